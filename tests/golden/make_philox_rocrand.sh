#!/bin/sh
# Regenerates tests/golden/philox_rocrand.txt from rocRAND's host Philox.
set -e
cd "$(dirname "$0")"
hipcc --offload-arch=gfx950 -O1 -std=c++17 -o /tmp/philox_rocrand philox_rocrand.cpp
/tmp/philox_rocrand > philox_rocrand.txt
