"""Locates and loads the in-tree native libraries under graphitethree_amd/lib.

Nothing here falls back to Python or CPU code: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

# GRAPHITETHREE_AMD_LIB_DIR points at another build of the same libraries
# (the sanitizer build the CPU tests run against).
LIB_DIR = os.environ.get("GRAPHITETHREE_AMD_LIB_DIR") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "lib")
_loaded: dict[str, C.CDLL] = {}


def path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def load(name: str) -> C.CDLL:
    if name not in _loaded:
        p = path(name)
        if not os.path.exists(p):
            raise RuntimeError(
                f"{p} is missing: build it with `python -c 'import __graft_entry__ "
                f"as g; g.build()'` (or python -m graphitethree_amd.build)")
        _loaded[name] = C.CDLL(p, mode=C.RTLD_GLOBAL)
    return _loaded[name]
