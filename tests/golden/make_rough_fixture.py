"""Oracle results of the rough patches at full size, as fixtures for the GPU parity tests.

The CPU restatement (oracle/cntc_oracle.c) needs minutes at 256^2 and most of an hour at
512^2 with eps = 1e-12, too long for the test suite; this script runs it once and stores the
forces, iteration counts and element-state counts, and the tractions and element states:
every element up to 256^2, a stride of 97 beyond.

Usage: python tests/golden/make_rough_fixture.py [--eps EPS] [n ...]   (default: 1e-12, 256 512)
Writes tests/golden/rough_patch_<n>_eps<k>.npz with k = -log10(eps).
"""
import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--eps", type=float, default=1e-12)
    ap.add_argument("sizes", type=int, nargs="*", default=[256, 512])
    a = ap.parse_args()
    from contact_amd import workloads as W
    from oracle import pyoracle as O
    for n in a.sizes:
        c = W.rough_patch(n=n, eps=a.eps)
        t0 = time.time()
        rc, f, P = O.solve_case(c)
        pn, px, py = P.tractions()
        el = P.eldiv()
        stride = 1 if n <= 256 else 97
        name = "rough_patch_%d_eps%d.npz" % (n, round(-math.log10(a.eps)))
        np.savez_compressed(os.path.join(HERE, name), rc=rc, eps=a.eps, forces=np.array(f),
                            iters=np.array(P.iters()), counts=np.bincount(el, minlength=4),
                            stride=stride, pn=pn[::stride], px=px[::stride], py=py[::stride],
                            el=el[::stride].astype(np.int8),
                            sums=np.array([pn.sum(), np.abs(px).sum(), np.abs(py).sum(), pn.max()]))
        print(n, a.eps, rc, f, P.iters(), "%.0f s" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
