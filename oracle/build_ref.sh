#!/usr/bin/env bash
# Build the reference's own cutoff / macro-node / wiring / compaction / walk
# code from the reference source tree into oracle/_ref/pakref_k<K>, one
# binary per k (k is a compile-time constant in the reference, Makefile:5-20).
#
# Test infrastructure only.  Needs the reference tree (REF, default
# /root/reference) and MPICH (MPI_HOME, default /opt/conda).  The reference's
# counting TU src/kmer_counting.cpp is not built: it includes oneTBB headers
# that this image does not have; oracle/ref_harness.cpp supplies the counted
# table instead (see its header).  Usage: build_ref.sh [k ...]
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${REF:-/root/reference}"
MPI_HOME="${MPI_HOME:-/opt/conda}"
OUT="$HERE/_ref"
KS=("$@")
if [ ${#KS[@]} -eq 0 ]; then KS=(21 32 33 64); fi

if [ ! -f "$REF/src/mn_construct.cpp" ] || [ ! -f "$MPI_HOME/include/mpi.h" ]; then
  echo "build_ref.sh: reference tree or MPI not found; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
TUS=(distribute_kmers input_reading kmer_histogram mn_construct wire iterative_compaction
     merge_mn local_compaction serialize walk debug timers)
for K in "${KS[@]}"; do
  W=$((K - 1))
  DEFS="-DLMER_LENGTH=8 -DPERFORM_CONTIG_GENERATION -DDEBUG_P2TIME -DCOMPACT_PGRAGH -DWINDW_SIZE=$W"
  if [ "$K" -gt 32 ]; then DEFS="$DEFS -DEXTEND_KMER"; fi
  OBJ="$OUT/obj_k$K"
  mkdir -p "$OBJ"
  CXX="g++ -std=c++14 -O2 -fopenmp -w $DEFS -I$REF/includes -I$REF/src -I$MPI_HOME/include"
  # the harness also takes the library's node-record filler (include/pakhip_mnmap.hpp)
  HCXX="$CXX -I$HERE/../include"
  pids=()
  for tu in "${TUS[@]}"; do
    # distribute_kmers.cpp's main() is renamed so the harness can drive the
    # same globals and functions itself.
    $CXX -Dmain=pakref_unused_main -c "$REF/src/$tu.cpp" -o "$OBJ/$tu.o" &
    pids+=($!)
  done
  $HCXX -c "$HERE/ref_harness.cpp" -o "$OBJ/ref_harness.o" &
  pids+=($!)
  for p in "${pids[@]}"; do wait "$p"; done
  g++ -fopenmp -static-libstdc++ -static-libgcc "$OBJ"/*.o "$MPI_HOME/lib/libmpi.so.12" -Wl,-rpath,"$MPI_HOME/lib" \
      -o "$OUT/pakref_k$K"
  echo "built $OUT/pakref_k$K"
done
