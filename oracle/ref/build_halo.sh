#!/bin/bash
# Builds the reference's own C test programs of the halo-exchange pack/unpack
# (src/OpenCL/test_halo_{read,write}{,_correct}.c, plain C restatements of
# src/OpenCL/Kernels/halo_exchange{,_correct}.cl) from where they lie into oracle/_ref/halo/
# (never committed).  tests/golden/make_halo_vectors.py runs them and stores their output.
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
SRC=${LES_REFERENCE_OCL:-/root/reference/src/OpenCL}
OUT="$HERE/../_ref/halo"
mkdir -p "$OUT"
for t in test_halo_read test_halo_read_correct test_halo_write test_halo_write_correct; do
  gcc -O2 -std=gnu99 -o "$OUT/$t" "$SRC/$t.c"
done
echo "built $OUT"
