#!/usr/bin/env bash
# Copies the reference's committed sssp golden data (inputs and MOSP-OpenMP outputs) into
# tests/golden/. Source: cpp/tests/data/mosp_sssp (34 cases, cases.txt) and
# cpp/tests/data/mosp_graph_io (the inputs of 27 of them, plus the MOSP applyChangeBatch goldens
# in <case>/applied/), and cpp/tests/data/mosp_combined (8 preference cases of MOSP's combined
# solve: the combined distances, tree and path costs). Data only: no source file of the
# reference is copied.
set -euo pipefail
SRC=${1:-/root/reference/cpp/tests/data}
DST=$(cd "$(dirname "$0")" && pwd)
rm -rf "$DST/mosp_sssp" "$DST/mosp_graph_io" "$DST/mosp_combined"
mkdir -p "$DST/mosp_sssp" "$DST/mosp_graph_io" "$DST/mosp_combined"
cp "$SRC/mosp_sssp/cases.txt" "$DST/mosp_sssp/"
for d in "$SRC"/mosp_sssp/*/; do
  cp -r "$d" "$DST/mosp_sssp/"
done
cp "$SRC/mosp_graph_io/cases.txt" "$DST/mosp_graph_io/"
for d in "$SRC"/mosp_graph_io/*/; do
  name=$(basename "$d")
  [ "$name" = mtx ] && continue
  cp -r "$d" "$DST/mosp_graph_io/"
done
cp "$SRC/mosp_combined/cases.txt" "$DST/mosp_combined/"
for d in "$SRC"/mosp_combined/*/; do
  cp -r "$d" "$DST/mosp_combined/"
done
echo "copied $(ls "$DST/mosp_combined" | wc -l) mosp_combined entries"
echo "copied $(ls "$DST/mosp_sssp" | wc -l) mosp_sssp entries and $(ls "$DST/mosp_graph_io" | wc -l) mosp_graph_io entries"
