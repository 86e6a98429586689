#!/bin/bash
# Host build of the small-cell device body (see emu_main.cpp): extracts the
# device code and the MixEngine's small-population builders from the
# product's engine_cell.hip into OUTDIR and compiles them with g++ and
# AddressSanitizer / UBSan.  usage: build.sh OUTDIR
set -euo pipefail
OUT=$1
HERE=$(cd "$(dirname "$0")" && pwd)
C=$(cd "$HERE/../../cerebellum-neuron-model-simulation-using-rkc_amd/csrc" && pwd)
S=$C/engine_cell.hip
mkdir -p "$OUT"
{ sed -n '/^enum Ctl {/,/^};/p' "$S"; grep -m1 '^constexpr double UROUND' "$S"; } > "$OUT/ctl.inc"
sed -n '/^constexpr int SMALL_NB/,/^\/\/ One persistent launch/p' "$S" | sed '$d' > "$OUT/small_body.inc"
sed -n '/^__global__ void __launch_bounds__(256) k_io_gj/,/^}/p' "$S" | sed 's/__launch_bounds__(256)//' > "$OUT/io_gj.inc"
sed -n '/  \/\/ stencil, RKC state and controller records common to the small types/,/  void set_lanes(int t, int G) {/p' "$S" | sed '$d' > "$OUT/mix_host.inc"
sed -n '/  static int nint_small(double x) {/,/^  }/p' "$S" > "$OUT/nint.inc"
for f in ctl small_body io_gj mix_host nint; do
  [ -s "$OUT/$f.inc" ] || { echo "extraction of $f from $S failed" >&2; exit 1; }
done
F="-std=c++17 -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=all -ffp-contract=off -I $HERE -I $OUT -I $C"
for f in gr_host go_host io_host morph; do
  g++ $F -DCB_HOST_EMU_HOST -c "$C/$f.cpp" -o "$OUT/$f.o"
done
g++ $F -DCB_HOST_EMU_DEVICE -c "$HERE/emu_main.cpp" -o "$OUT/emu_main.o"
g++ -fsanitize=address,undefined -pthread "$OUT/emu_main.o" "$OUT/gr_host.o" "$OUT/go_host.o" "$OUT/io_host.o" \
  "$OUT/morph.o" -o "$OUT/emu"
