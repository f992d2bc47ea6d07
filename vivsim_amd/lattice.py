"""Lattice descriptors with the reference's field names (lbm/lattice.py:8-110,
lbm3d/lattice.py:13-126), so user code reading e.g. ``D2Q9.w`` or
``D2Q9.boundary_spec['left'].in_dirs`` keeps working.  The direction numbering
is the population layout of the HIP kernels (csrc/lattice.hpp)."""

from typing import NamedTuple

import numpy as np


class BoundarySpec(NamedTuple):
    wall: tuple
    neighbor: tuple
    in_dirs: np.ndarray
    out_dirs: np.ndarray
    tan_dirs: np.ndarray
    pos_side_dirs: np.ndarray
    neg_side_dirs: np.ndarray
    normal_sign: int
    normal_axis: int


class Lattice(NamedTuple):
    d: int
    q: int
    w: np.ndarray
    c: np.ndarray
    cs2: float
    right_dirs: np.ndarray
    left_dirs: np.ndarray
    up_dirs: np.ndarray
    down_dirs: np.ndarray
    all_dirs: np.ndarray
    opp_dirs: np.ndarray
    boundary_spec: dict


class BoundarySpec3D(NamedTuple):
    wall: tuple
    neighbor: tuple
    in_dirs: np.ndarray
    out_dirs: np.ndarray
    zero_dirs: np.ndarray
    normal_sign: int
    normal_axis: int
    tangential_axes: tuple


class Lattice3D(NamedTuple):
    d: int
    q: int
    w: np.ndarray
    c: np.ndarray
    cs2: float
    all_dirs: np.ndarray
    opp_dirs: np.ndarray
    boundary_spec: dict


def _i(*v):
    return np.array(v, dtype=np.int32)


_C2 = _i([0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1])


def _opp(c):
    return np.array([int(np.flatnonzero((c == -x).all(1))[0]) for x in c], dtype=np.int32)


D2Q9 = Lattice(
    d=2, q=9, cs2=1.0 / 3.0,
    w=np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float32),
    c=_C2,
    right_dirs=_i(1, 5, 8), left_dirs=_i(3, 7, 6), up_dirs=_i(2, 5, 6), down_dirs=_i(4, 7, 8),
    all_dirs=np.arange(9, dtype=np.int32), opp_dirs=_opp(_C2),
    boundary_spec={
        "left": BoundarySpec((0,), (1,), _i(1, 5, 8), _i(3, 7, 6), _i(2, 4), _i(2, 5, 6), _i(4, 7, 8), 1, 0),
        "right": BoundarySpec((-1,), (-2,), _i(3, 7, 6), _i(1, 5, 8), _i(2, 4), _i(2, 5, 6), _i(4, 7, 8), -1, 0),
        "top": BoundarySpec((slice(None), -1), (slice(None), -2), _i(4, 7, 8), _i(2, 5, 6), _i(1, 3),
                            _i(1, 5, 8), _i(3, 7, 6), -1, 1),
        "bottom": BoundarySpec((slice(None), 0), (slice(None), 1), _i(2, 5, 6), _i(4, 7, 8), _i(1, 3),
                               _i(1, 5, 8), _i(3, 7, 6), 1, 1),
    },
)

_C3 = _i([0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
         [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
         [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1])
_S = slice(None)
D3Q19 = Lattice3D(
    d=3, q=19, cs2=1.0 / 3.0,
    w=np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32),
    c=_C3, all_dirs=np.arange(19, dtype=np.int32), opp_dirs=_opp(_C3),
    boundary_spec={
        "left": BoundarySpec3D((0, _S, _S), (1, _S, _S), _i(1, 7, 9, 11, 13), _i(2, 8, 10, 12, 14),
                               _i(0, 3, 4, 5, 6, 15, 16, 17, 18), 1, 0, (1, 2)),
        "right": BoundarySpec3D((-1, _S, _S), (-2, _S, _S), _i(2, 8, 10, 12, 14), _i(1, 7, 9, 11, 13),
                                _i(0, 3, 4, 5, 6, 15, 16, 17, 18), -1, 0, (1, 2)),
        "bottom": BoundarySpec3D((_S, 0, _S), (_S, 1, _S), _i(3, 7, 8, 15, 17), _i(4, 9, 10, 16, 18),
                                 _i(0, 1, 2, 5, 6, 11, 12, 13, 14), 1, 1, (0, 2)),
        "top": BoundarySpec3D((_S, -1, _S), (_S, -2, _S), _i(4, 9, 10, 16, 18), _i(3, 7, 8, 15, 17),
                              _i(0, 1, 2, 5, 6, 11, 12, 13, 14), -1, 1, (0, 2)),
        "back": BoundarySpec3D((_S, _S, 0), (_S, _S, 1), _i(5, 11, 12, 15, 16), _i(6, 13, 14, 17, 18),
                               _i(0, 1, 2, 3, 4, 7, 8, 9, 10), 1, 2, (0, 1)),
        "front": BoundarySpec3D((_S, _S, -1), (_S, _S, -2), _i(6, 13, 14, 17, 18), _i(5, 11, 12, 15, 16),
                                _i(0, 1, 2, 3, 4, 7, 8, 9, 10), -1, 2, (0, 1)),
    },
)
