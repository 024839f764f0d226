"""The C-ABI library loads (no GPU needed) and exports every function include/*.h declares."""
import glob
import os
import re

import nla3d_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(nla3d_gpu_\w+)\s*\(", src, re.M))
    return names


def test_exports_match_header():
    L = nla3d_amd.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(nla3d_amd.ABI) | set(nla3d_amd.ABI_FREE) | {"nla3d_gpu_last_error"} >= names
