#!/bin/sh
# Build the mock-MEX drivers of both gateways against libmesolf.so.
set -e
here=$(cd "$(dirname "$0")" && pwd)
root=$(cd "$here/../.." && pwd)
out=${1:-$here/_build}
mkdir -p "$out"
for gw in mex_lfm_convolution_vec recon_phase_space_mex; do
  g++ -O1 -std=c++17 -Wall -I"$here" -I"$root/include" \
      -DGATEWAY_FILE="\"$root/mesolf_amd/matlab/$gw.cpp\"" "$here/driver.cpp" \
      -L"$root/mesolf_amd" -lmesolf -Wl,-rpath,"$root/mesolf_amd" -o "$out/$gw"
done
