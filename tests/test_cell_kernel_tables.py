"""Constant tables of the cell kernel (kiwi_amd/csrc/kiwi_hip.hip), checked on the CPU.

kMagic16 turns the lane-strided box walk's divisions into multiplications: u / d == (u * m[d]) >> 16
with m[d] = ceil(2^16 / d), for every box face size d = n1 n2 or n2 (1..81, item boxes of at most 9
lattice points per axis) and every dividend the walk uses (a lane number 0..63, or the step 64).
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _magic_table():
    src = open(os.path.join(ROOT, "kiwi_amd", "csrc", "kiwi_hip.hip")).read()
    m = re.search(r"kMagic16\[(\d+)\]\s*=\s*\{([^}]*)\}", src)
    assert m, "kMagic16 not found"
    vals = [int(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
    assert len(vals) == int(m.group(1))
    return vals


def test_magic_table_is_ceil_of_two_to_the_16_over_d():
    vals = _magic_table()
    assert len(vals) == 82
    assert vals[1:] == [-(-65536 // d) for d in range(1, 82)]


def test_magic_division_is_exact_for_the_walks():
    vals = _magic_table()
    for d in range(1, 82):
        for u in range(0, 65):
            assert (u * vals[d]) >> 16 == u // d, (u, d)


def test_magic_division_bound_is_tight_enough():
    # the argument in the kernel's comment: u e < 2^16 with e = m d - 2^16 < d, u <= 64, d <= 81
    vals = _magic_table()
    for d in range(1, 82):
        e = vals[d] * d - 65536
        assert 0 <= e < d
        assert 64 * e < 65536
