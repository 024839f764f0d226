"""Extract the first cases of the reference's dense-product fixtures.

Source: test/mat_prod/matlab/* of dmitryikh/nla3d (MIT licence), generated by
test_generator.m and read by test/mat_prod_test.cpp (100 cases each, compared
at 1e-6).  Only the numbers are copied: each output file holds NCASES cases,
one case per line, in the reference's read order (operands, then the expected
result).  Run in a container that has the reference checkout; the outputs are
committed next to this script.
"""
import os
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "test/mat_prod/matlab")
OUT = os.path.dirname(os.path.abspath(__file__))
NCASES = 25
# file -> values per case (operand sizes from mat_prod_test.cpp)
FILES = {
    "matBVprod_2409": 24 * 9 + 9 + 24,
    "matBTVprod_2409": 24 * 9 + 24 + 9,
    "matABprod_240912": 24 * 9 + 9 * 12 + 24 * 12,
    "matATBprod_240912": 24 * 9 + 24 * 12 + 9 * 12,
    "matBTDBprod_2409": 24 * 9 + 300 + 45,
}

for name, per in FILES.items():
    with open(os.path.join(SRC, name)) as f:
        tok = f.read().split()
    assert len(tok) >= per * 100, name
    with open(os.path.join(OUT, "mat_prod", name + ".txt"), "w") as g:
        for c in range(NCASES):
            g.write(" ".join(tok[c * per:(c + 1) * per]) + "\n")
print("ok")
