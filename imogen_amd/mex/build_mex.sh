#!/bin/sh
# Builds imogen_mex.mexa64 against the in-tree libimogen_hip.so where MATLAB is
# installed (mex on PATH); prints a notice and exits 0 elsewhere.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
if ! command -v mex >/dev/null 2>&1; then
    echo "mex not found: MATLAB is not installed here, the gateway is not built"
    exit 0
fi
mex -I"$ROOT/include" "$HERE/imogen_mex.cpp" -L"$ROOT/imogen_amd" -limogen_hip \
    LDFLAGS='$LDFLAGS -Wl,-rpath,'"$ROOT/imogen_amd" -outdir "$HERE"
