"""Regenerate ``nodedata.json``: interpolation nodes for N = 1..9 and cubature rules of degree 1..19.

The solver needs the same point sets as the reference so that the WADG
projection, the nodal initial-condition interpolation and the cubature-based
mass matrices agree with it.  The reference keeps them as generated C data
tables in ``cpp/include/NodeData.h`` (Warburton tetrahedral nodes ``p_r_N#``,
Xiao--Gimbutas tetrahedral rules ``p_rq_N#``/``p_wq_N#`` indexed by polynomial
degree, triangle rules ``p_rfq_N#``/``p_wfq_N#``; loaded by ``Nodes3D``,
``tet_cubature`` and ``tri_cubature`` at ``cpp/src/Basis.cpp:799-949``).

This script reads that header as text and writes the numbers only, as JSON,
for node orders 1..9 and cubature degrees 1..19 (the solver supports orders 1..9: the WADG
update uses the degree 2N+1 rule up to N = 7, the face and mass-matrix rules go to degree
2N+1 = 19 at N = 9).  It needs the reference checkout and
is only run when the table must be rebuilt; the JSON it writes is what the
package loads.  The tables are checked independently by ``tests/test_tables.py``
(cubature exactness on monomials, node symmetry and face counts).

Usage:  python bbwadg_amd/data/gen_nodedata.py /root/reference/cpp/include/NodeData.h
"""
import json
import os
import re
import sys

ARRAY_RE = re.compile(r"double\s+(p_\w+?)_N(\d+)\s*\[\s*(\d+)\s*\]\s*=\s*\{([^}]*)\}")
INT_RE = re.compile(r"int\s+(p_\w+?)_N(\d+)\s*=\s*(\d+)\s*;")


def parse(text):
    arrays = {}
    for name, idx, n, body in ARRAY_RE.findall(text):
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        assert len(vals) == int(n), (name, idx, n, len(vals))
        arrays[(name, int(idx))] = vals
    ints = {(name, int(idx)): int(v) for name, idx, v in INT_RE.findall(text)}
    return arrays, ints


def main(path):
    with open(path) as f:
        text = f.read()
    arrays, ints = parse(text)
    out = {"source": "cpp/include/NodeData.h (jlchan/bbwadg)", "nodes": {}, "tet": {}, "tri": {}}
    for N in range(1, 10):
        out["nodes"][str(N)] = {k: arrays[("p_" + k, N)] for k in ("r", "s", "t")}
    for deg in range(1, 20):
        nq = ints[("p_Nq", deg)]
        out["tet"][str(deg)] = {k: arrays[("p_" + k + "q", deg)] for k in ("r", "s", "t", "w")}
        assert all(len(v) == nq for v in out["tet"][str(deg)].values())
        nfq = ints[("p_Nfq", deg)]
        out["tri"][str(deg)] = {k: arrays[("p_" + k + "fq", deg)] for k in ("r", "s", "w")}
        assert all(len(v) == nfq for v in out["tri"][str(deg)].values())
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nodedata.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/cpp/include/NodeData.h")
