"""Writes tests/golden/halo_vectors.json from the reference's own halo-exchange test programs.

Runs ONLY in the build container, where the reference sources exist: oracle/ref/build_halo.sh
compiles src/OpenCL/test_halo_{read,write}{,_correct}.c with gcc, and this script runs each
program and stores what it prints (the array after the unpack, or the packed buffer) together
with the sizes and the inputs its main() sets up, restated here as data:

  test_halo_write          im=30 jm=40 km=2 h_w=3 v_dim=2; array = 1 (first half) / 2 (second
                           half), buffer = 5; exchange_2_halo_write; prints the array.
  test_halo_write_correct  the same with im=20 jm=20.
  test_halo_read           im=20 jm=20 km=2 h_w=3 v_dim=2; array = 1 / 2, buffer segments
                           north 6, south 7, west 8, east 9; write, buffer = 4, read; prints
                           the buffer.
  test_halo_read_correct   im=20 jm=20 km=2 h_w=1 v_dim=2; array[i] = i, buffer = -1;
                           the inner-ring read of halo_exchange_correct.cl; prints the array
                           and the buffer.

    python tests/golden/make_halo_vectors.py
"""
import json
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "halo")

CASES = {
    "test_halo_write": dict(im=30, jm=40, km=2, h_w=3, v_dim=2, prints=["array"]),
    "test_halo_write_correct": dict(im=20, jm=20, km=2, h_w=3, v_dim=2, prints=["array"]),
    "test_halo_read": dict(im=20, jm=20, km=2, h_w=3, v_dim=2, prints=["buffer"]),
    "test_halo_read_correct": dict(im=20, jm=20, km=2, h_w=1, v_dim=2, prints=["array", "buffer"]),
}


def numbers(text):
    return [float(t) for t in re.sub(r"==[a-z]+==", " ", text).split()]


def main():
    subprocess.run([os.path.join(ROOT, "oracle", "ref", "build_halo.sh")], check=True,
                   stdout=subprocess.DEVNULL)
    out = {}
    for name, c in CASES.items():
        text = subprocess.run([os.path.join(BIN, name)], check=True, capture_output=True,
                              text=True).stdout
        vals = numbers(text)
        im, jm, km, hw, vd = c["im"], c["jm"], c["km"], c["h_w"], c["v_dim"]
        arr_sz = vd * im * jm * km
        correct = name.endswith("_correct") and "read" in name
        buf_sz = vd * 2 * (im + jm - (4 if correct else 2) * hw) * hw * km
        ent = {k: c[k] for k in ("im", "jm", "km", "h_w", "v_dim")}
        pos = 0
        for what in c["prints"]:
            n = arr_sz if what == "array" else buf_sz
            ent[what] = vals[pos:pos + n]
            pos += n
        assert pos == len(vals), (name, pos, len(vals))
        out[name] = ent
    with open(os.path.join(HERE, "halo_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "halo_vectors.json"))


if __name__ == "__main__":
    main()
