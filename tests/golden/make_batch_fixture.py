"""Oracle results for a size-stratified subset of the 2048-patch Hertzian batch (seed 2024),
as a fixture for the GPU batch tests.

The subset: the patches of the batch with N = mx*my <= 3000 elements (the CPU restatement
takes seconds on each), sorted by N and taken at 16 evenly spaced ranks, plus patches 381 and
1505, on which GDsteady stagnates and SteadyGS takes over.  Each is solved at the batch's
eps = 1e-8 (forces) and at eps = 1e-12 (forces, tractions and element states of every
element).

Usage: python tests/golden/make_batch_fixture.py    (writes tests/golden/batch_subset.npz)
"""
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def subset():
    from contact_amd import workloads as W
    cases = W.hertz_batch(n=2048, seed=2024)
    n = np.array([c["mx"] * c["my"] for c in cases])
    small = np.array([k for k in np.argsort(n, kind="stable") if n[k] <= 3000])
    pick = [int(small[i]) for i in np.linspace(0, len(small) - 1, 16).round().astype(int)]
    return sorted(set(pick) | {381, 1505})


def solve(k):
    from contact_amd import workloads as W
    from oracle import pyoracle as O
    c = W.hertz_batch(n=2048, seed=2024)[k]
    out = {}
    for tag, eps in (("e8", 1e-8), ("e12", 1e-12)):
        cc = dict(c)
        cc["eps"] = eps
        rc, f, P = O.solve_case(cc)
        out[tag] = (rc, np.array(f), P)
    pn, px, py = out["e12"][2].tractions()
    el = out["e12"][2].eldiv().astype(np.int8)
    return (k, out["e8"][0], out["e8"][1], out["e12"][0], out["e12"][1], pn, px, py, el,
            out["e8"][2].stagnated(), out["e12"][2].stagnated())


def main():
    ks = subset()
    with ProcessPoolExecutor(max_workers=6) as ex:
        res = list(ex.map(solve, ks))
    d = {"patches": np.array(ks)}
    for r in res:
        k = r[0]
        print(k, "rc", r[1], r[3], "forces e8", r[2], "e12", r[4], "stagnated", r[9], r[10], flush=True)
        d["rc_e8_%d" % k] = r[1]
        d["forces_e8_%d" % k] = r[2]
        d["rc_e12_%d" % k] = r[3]
        d["forces_e12_%d" % k] = r[4]
        d["pn_%d" % k], d["px_%d" % k], d["py_%d" % k], d["el_%d" % k] = r[5], r[6], r[7], r[8]
    np.savez_compressed(os.path.join(HERE, "batch_subset.npz"), **d)


if __name__ == "__main__":
    main()
