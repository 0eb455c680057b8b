"""Locations of the in-tree native libraries (built by ``__graft_entry__.build()``)."""
from __future__ import annotations

import os

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PACKAGE_ROOT, "lib")


def library_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)
