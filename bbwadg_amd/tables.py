"""Interpolation nodes and cubature rules (data from ``cpp/include/NodeData.h``).

``nodedata.json`` holds the reference's Warburton nodes for orders 1..7 and its
tetrahedron/triangle cubature rules for degrees 1..15, regenerated by
``data/gen_nodedata.py``.  Rules are indexed by polynomial degree, as
``tet_cubature``/``tri_cubature`` in ``cpp/src/Basis.cpp:833-949`` index them.
"""
import json
import os
from functools import lru_cache

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "nodedata.json")


@lru_cache(maxsize=1)
def _load():
    with open(_PATH) as f:
        return json.load(f)


def _arr(d):
    return {k: np.asarray(v, dtype=np.float64) for k, v in d.items()}


def nodes(N):
    d = _load()["nodes"]
    if str(N) not in d:
        raise ValueError(f"no interpolation nodes tabulated for N={N}")
    return _arr(d[str(N)])


def tet_rule(deg):
    d = _load()["tet"]
    if str(deg) not in d:
        raise ValueError(f"no tetrahedral cubature of degree {deg}")
    return _arr(d[str(deg)])


def tri_rule(deg):
    d = _load()["tri"]
    if str(deg) not in d:
        raise ValueError(f"no triangle cubature of degree {deg}")
    return _arr(d[str(deg)])
