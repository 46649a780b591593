"""Extract the golden rows of the reference's test_table run into a small fixture.

Source: testbank/test_table_mx11.ref_fx of the reference (3220 steady-rolling Hertz cases of
src/test_table.f90, mx = 11, SteadyGS, eps = 1e-6): per case Fx/(mu Fn), Fy/(mu Fn),
Mz/(mu Fn cp) and the percentage of slip elements, as printed there.

Usage: python tests/golden/make_table_golden.py /path/to/reference
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    src = os.path.join(ref_root, "testbank", "test_table_mx11.ref_fx")
    rows = []
    for line in open(src):
        if line.startswith("Fx="):
            rows.append([x.split("=")[1].strip() for x in line.split(",")])
    with open(os.path.join(HERE, "test_table_mx11_fx.csv"), "w") as f:
        f.write("# fx_rel,fy_rel,mz_rel,pct_slip  (testbank/test_table_mx11.ref_fx)\n")
        for r in rows:
            f.write(",".join(r) + "\n")
    print(len(rows), "rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
