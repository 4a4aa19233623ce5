#!/bin/sh
# Regenerates tests/golden/std_kats.json with the system g++ / libstdc++.
set -e
cd "$(dirname "$0")"
g++ -O2 -std=c++11 -o /tmp/make_std_kats make_std_kats.cc
/tmp/make_std_kats > std_kats.json
echo "wrote $(pwd)/std_kats.json"
