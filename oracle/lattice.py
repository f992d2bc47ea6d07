"""Lattice tables for the oracle (D2Q9 and D3Q19).

The direction numbering and the per-wall direction groups are part of the
reference's data layout (``f[9, NX, NY]``, ``f[19, NX, NY, NZ]``), so they must
be identical to it:
  * D2Q9: vivsim/lbm/lattice.py:39-110
  * D3Q19: vivsim/lbm3d/lattice.py:39-126

Wall groups are derived here from the geometry (inward normal) and then
ordered as the reference orders them, because the NEBB / bounce-back formulas
of the reference pair ``in_dirs[k]`` with ``out_dirs[k]`` by position.
"""

import numpy as np

# ---------------------------------------------------------------- D2Q9
# lbm/lattice.py:44-57 : 0 rest, 1 +x, 2 +y, 3 -x, 4 -y, 5 (+x+y), 6 (-x+y), 7 (-x-y), 8 (+x-y)
C2 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
               [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int64)
W2 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float64)
Q2 = 9


def _opp(c):
    return np.array([int(np.flatnonzero((c == -ci).all(axis=1))[0]) for ci in c])


OPP2 = _opp(C2)

# Per-wall groups (lbm/lattice.py:64-109).  Fields:
#   axis     normal axis (0: x walls, 1: y walls)
#   sign     +1 if the inward normal points along +axis
#   idx      wall line index along the axis (0 or -1), nb = neighbour line (1 or -2)
#   in_dirs  unknown populations after streaming (c . n > 0), ordered [normal, diag, diag]
#   out_dirs opposite of in_dirs (same order)
#   tan      the two axis-aligned tangential dirs [+t, -t]
#   pos/neg  dirs with positive / negative tangential velocity component
WALLS2 = {
    "left":   dict(axis=0, sign=+1, idx=0,  nb=1,  in_dirs=[1, 5, 8], out_dirs=[3, 7, 6],
                   tan=[2, 4], pos=[2, 5, 6], neg=[4, 7, 8]),
    "right":  dict(axis=0, sign=-1, idx=-1, nb=-2, in_dirs=[3, 7, 6], out_dirs=[1, 5, 8],
                   tan=[2, 4], pos=[2, 5, 6], neg=[4, 7, 8]),
    "top":    dict(axis=1, sign=-1, idx=-1, nb=-2, in_dirs=[4, 7, 8], out_dirs=[2, 5, 6],
                   tan=[1, 3], pos=[1, 5, 8], neg=[3, 7, 6]),
    "bottom": dict(axis=1, sign=+1, idx=0,  nb=1,  in_dirs=[2, 5, 6], out_dirs=[4, 7, 8],
                   tan=[1, 3], pos=[1, 5, 8], neg=[3, 7, 6]),
}

# ---------------------------------------------------------------- D3Q19
# lbm3d/lattice.py:50-58
C3 = np.array([
    [0, 0, 0],
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0],
    [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1],
], dtype=np.int64)
W3 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
Q3 = 19
OPP3 = _opp(C3)

# lbm3d/lattice.py:64-125.  in_dirs / out_dirs are listed in the reference's
# order; note that out_dirs[k] is NOT opp(in_dirs[k]) for the diagonals (the
# reference pairs them by position in bounce-back, lbm3d/boundary/bb.py:22).
WALLS3 = {
    "left":   dict(axis=0, sign=+1, idx=0,  nb=1,  in_dirs=[1, 7, 9, 11, 13],
                   out_dirs=[2, 8, 10, 12, 14], zero=[0, 3, 4, 5, 6, 15, 16, 17, 18], tan_axes=(1, 2)),
    "right":  dict(axis=0, sign=-1, idx=-1, nb=-2, in_dirs=[2, 8, 10, 12, 14],
                   out_dirs=[1, 7, 9, 11, 13], zero=[0, 3, 4, 5, 6, 15, 16, 17, 18], tan_axes=(1, 2)),
    "bottom": dict(axis=1, sign=+1, idx=0,  nb=1,  in_dirs=[3, 7, 8, 15, 17],
                   out_dirs=[4, 9, 10, 16, 18], zero=[0, 1, 2, 5, 6, 11, 12, 13, 14], tan_axes=(0, 2)),
    "top":    dict(axis=1, sign=-1, idx=-1, nb=-2, in_dirs=[4, 9, 10, 16, 18],
                   out_dirs=[3, 7, 8, 15, 17], zero=[0, 1, 2, 5, 6, 11, 12, 13, 14], tan_axes=(0, 2)),
    "back":   dict(axis=2, sign=+1, idx=0,  nb=1,  in_dirs=[5, 11, 12, 15, 16],
                   out_dirs=[6, 13, 14, 17, 18], zero=[0, 1, 2, 3, 4, 7, 8, 9, 10], tan_axes=(0, 1)),
    "front":  dict(axis=2, sign=-1, idx=-1, nb=-2, in_dirs=[6, 13, 14, 17, 18],
                   out_dirs=[5, 11, 12, 15, 16], zero=[0, 1, 2, 3, 4, 7, 8, 9, 10], tan_axes=(0, 1)),
}


def check_tables():
    """Geometric consistency of the tables (used by the oracle tests)."""
    for c, opp, walls, d in ((C2, OPP2, WALLS2, 2), (C3, OPP3, WALLS3, 3)):
        assert (c[opp] == -c).all()
        for loc, s in walls.items():
            n = np.zeros(d, dtype=np.int64)
            n[s["axis"]] = s["sign"]
            cn = c @ n
            assert sorted(s["in_dirs"]) == sorted(np.flatnonzero(cn > 0).tolist()), loc
            assert sorted(s["out_dirs"]) == sorted(np.flatnonzero(cn < 0).tolist()), loc
            if d == 3:
                assert sorted(s["zero"]) == sorted(np.flatnonzero(cn == 0).tolist()), loc
    return True
