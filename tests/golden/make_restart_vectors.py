"""Writes tests/golden/restart_data30.bin and restart_data31.bin: a restart pair in the Fortran
unformatted sequential layout, written by restart_records.f95 compiled with the image's Fortran
compiler (AMD flang), i.e. by the same runtime that reads and writes the reference's files.

    python tests/golden/make_restart_vectors.py
"""
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("FC", "/opt/rocm/lib/llvm/bin/amdflang")


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "restart_records")
        subprocess.run([FC, "-O0", "-o", exe, os.path.join(HERE, "restart_records.f95")],
                       check=True, cwd=tmp)
        subprocess.run([exe], check=True, cwd=tmp)
        for q in ("30", "31"):
            shutil.copy(os.path.join(tmp, "data%s.bin" % q), os.path.join(HERE, "restart_data%s.bin" % q))
    print("wrote restart_data30.bin, restart_data31.bin")


if __name__ == "__main__":
    main()
