#!/bin/bash
# builds tests/cxx/class_api (the class API test program) against the library
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
g++ -O2 -std=c++17 -Wall -I "$ROOT/jetshow_amd/cxx" "$HERE/class_api.cpp" "$ROOT/jetshow_amd/cxx/jetshow.cpp" \
    -L "$ROOT/jetshow_amd" -ljetshow_amd -Wl,-rpath,"$ROOT/jetshow_amd" -o "${1:-$HERE/class_api}"
