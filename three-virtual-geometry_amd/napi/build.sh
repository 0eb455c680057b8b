#!/bin/sh
# Builds vgcut.node without node-gyp: g++ against the Node headers, linked to lib/libvgcut.so and lib/libvgdag.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
NODE_INCLUDE=${NODE_INCLUDE:-/usr/include/node}
g++ -std=c++17 -O2 -shared -fPIC -DNODE_GYP_MODULE_NAME=vgcut -I"$NODE_INCLUDE" -I"$HERE/../../include" \
  "$HERE/vgcut_napi.cc" -L"$HERE/../lib" -lvgcut -lvgdag -Wl,-rpath,"$HERE/../lib" -o "$HERE/vgcut.node"
