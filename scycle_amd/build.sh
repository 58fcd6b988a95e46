#!/usr/bin/env bash
# Builds libscycle_hip.so for gfx950 in-tree (scycle_amd/libscycle_hip.so).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
src="$here/csrc"
out="$here/libscycle_hip.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
ARCH="${SC_OFFLOAD_ARCH:-gfx950}"
objdir="$here/_build"
mkdir -p "$objdir"
flags=(-O3 -std=c++17 -fPIC --offload-arch="$ARCH" -Wall -Wno-unused-function
       -Xclang -target-feature -Xclang -packed-fp32-ops -I"$here/../include")
objs=()
pids=()
for f in "$src"/*.hip "$src"/*.cpp; do
  o="$objdir/$(basename "$f").o"
  rm -f "$o"  # never link a stale object after a failed compile
  if [[ "$f" == *.hip ]]; then
    "$HIPCC" "${flags[@]}" -x hip -c "$f" -o "$o" &
  else
    "$HIPCC" "${flags[@]}" -c "$f" -o "$o" &
  fi
  objs+=("$o")
  pids+=("$!")
done
fail=0
for p in "${pids[@]}"; do
  wait "$p" || fail=1
done
if [[ $fail -ne 0 ]]; then
  echo "build failed" >&2
  exit 1
fi
"$HIPCC" -shared --offload-arch="$ARCH" -fPIC -o "$out.tmp" "${objs[@]}" -L/opt/rocm/lib -lrocblas -lrocsolver -Wl,-rpath,/opt/rocm/lib
# every symbol of the library's own namespace must resolve inside it
if nm -D --undefined-only "$out.tmp" | grep -q '_ZN2sc'; then
  nm -D --undefined-only "$out.tmp" | grep '_ZN2sc' >&2
  echo "unresolved library symbols" >&2
  rm -f "$out.tmp"
  exit 1
fi
mv "$out.tmp" "$out"
echo "built $out"
