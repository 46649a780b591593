"""Oracle results of the rough patches at full size, as a fixture for the GPU parity test.

The CPU restatement (oracle/cntc_oracle.c) needs minutes at 256^2 and about 13 minutes at
512^2, too long for the test suite; this script runs it once and stores the forces, iteration
counts, element-state counts and a strided sample of the tractions.

Usage: python tests/golden/make_rough_fixture.py [n ...]     (default: 256 512)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main(sizes):
    from contact_amd import workloads as W
    from oracle import pyoracle as O
    for n in sizes:
        c = W.rough_patch(n=n)
        rc, f, P = O.solve_case(c)
        pn, px, py = P.tractions()
        el = P.eldiv()
        stride = 97
        np.savez_compressed(os.path.join(HERE, "rough_patch_%d.npz" % n),
                            rc=rc, forces=np.array(f), iters=np.array(P.iters()),
                            counts=np.bincount(el, minlength=4), stride=stride,
                            pn=pn[::stride], px=px[::stride], py=py[::stride], el=el[::stride],
                            sums=np.array([pn.sum(), np.abs(px).sum(), np.abs(py).sum(), pn.max()]))
        print(n, rc, f, P.iters(), flush=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [256, 512])
