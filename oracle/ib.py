"""Oracle restatement of ``vivsim.ib`` / ``vivsim.ib3d``: delta kernels, stencils,
interpolate / spread, multi-direct forcing and marker geometry.

Out-of-range stencil indices follow JAX's documented indexing semantics (the
reference never relies on them; its examples keep markers inside the window):
negative indices wrap once, gathers clamp, scatter-adds drop.
"""

import numpy as np


def _abs(r):
    return np.abs(np.asarray(r))


# ------------------------------------------------------------------ ib/kernels.py

def kernel_peskin_3pt(r):
    """ib/kernels.py:4-22."""
    a = _abs(r)
    with np.errstate(invalid="ignore"):
        inner = (1 + np.sqrt(1 - 3 * a ** 2)) / 3
        outer = (5 - 3 * a - np.sqrt(-2 + 6 * a - 3 * a ** 2)) / 6
    return np.where(a > 1.5, 0, np.where(a < 0.5, inner, outer)).astype(a.dtype)


def kernel_peskin_4pt(r):
    """ib/kernels.py:25-43."""
    a = _abs(r)
    with np.errstate(invalid="ignore"):
        inner = (3 - 2 * a + np.sqrt(1 + 4 * a - 4 * a ** 2)) * 0.125
        outer = (5 - 2 * a - np.sqrt(-7 + 12 * a - 4 * a ** 2)) * 0.125
    return np.where(a > 2, 0, np.where(a < 1, inner, outer)).astype(a.dtype)


def kernel_cosine_4pt(r):
    """ib/kernels.py:46-61."""
    a = _abs(r)
    return np.where(a > 2, 0, (1 + np.cos(np.pi * a * 0.5)) * 0.25).astype(a.dtype)


def kernel_peskin_2pt(r):
    """Linear hat max(0, 1 - |r|).  Not in the reference (its README lists a
    2-point kernel that the code does not define); provided as an extension."""
    a = _abs(r)
    return np.where(a > 1, 0, 1 - a).astype(a.dtype)


KERNELS = {
    "peskin_2pt": kernel_peskin_2pt,
    "peskin_3pt": kernel_peskin_3pt,
    "peskin_4pt": kernel_peskin_4pt,
    "cosine_4pt": kernel_cosine_4pt,
}


# ------------------------------------------------------------------ ib/stencil.py, ib3d/stencil.py

def get_ib_stencil(marker_x, marker_y, ny, kernel=kernel_peskin_4pt, stencil_radius=2):
    """Tensor-product stencil, x-major / y-minor; idx = sx*ny + sy — ib/stencil.py:7-51."""
    mx = np.asarray(marker_x)
    my = np.asarray(marker_y)
    off = np.arange(-stencil_radius + 1, stencil_radius + 1, dtype=np.int32)
    ox, oy = np.meshgrid(off, off, indexing="ij")
    ox = ox.reshape(1, -1)
    oy = oy.reshape(1, -1)
    sx = np.floor(mx).astype(np.int32)[:, None] + ox
    sy = np.floor(my).astype(np.int32)[:, None] + oy
    dt = mx.dtype
    w = kernel(sx.astype(dt) - mx[:, None]) * kernel(sy.astype(dt) - my[:, None])
    idx = (sx * np.int32(ny) + sy).astype(np.int32)
    return w.astype(dt), idx


def get_ib_stencil3d(marker_coords, grid_shape, kernel=kernel_peskin_4pt, stencil_radius=2):
    """idx = sx*(ny*nz) + sy*nz + sz — ib3d/stencil.py:7-59 (same ValueErrors)."""
    if len(grid_shape) != 3:
        raise ValueError(f"grid_shape must be a 3-tuple, got {grid_shape}.")
    mc = np.asarray(marker_coords)
    if mc.ndim != 2 or mc.shape[1] != 3:
        raise ValueError(f"marker_coords must have shape (n_markers, 3), got {mc.shape}.")
    nx, ny, nz = grid_shape
    off = np.arange(-stencil_radius + 1, stencil_radius + 1, dtype=np.int32)
    ox, oy, oz = [a.reshape(1, -1) for a in np.meshgrid(off, off, off, indexing="ij")]
    ref = np.floor(mc).astype(np.int32)
    sx = ref[:, 0:1] + ox
    sy = ref[:, 1:2] + oy
    sz = ref[:, 2:3] + oz
    dt = mc.dtype
    w = (kernel(sx.astype(dt) - mc[:, 0:1]) * kernel(sy.astype(dt) - mc[:, 1:2])
         * kernel(sz.astype(dt) - mc[:, 2:3]))
    idx = (sx * np.int32(ny * nz) + sy * np.int32(nz) + sz).astype(np.int32)
    return w.astype(dt), idx


def _norm_idx(idx, n):
    idx = np.asarray(idx).astype(np.int64)
    return np.where(idx < 0, idx + n, idx)


def interpolate(grid_values, stencil_weights, stencil_indices):
    """marker[m, c] = sum_s W[m, s] grid[c, I[m, s]] — ib/stencil.py:54-78."""
    g = np.asarray(grid_values)
    flat = g.reshape(g.shape[0], -1)
    n = flat.shape[1]
    idx = np.clip(_norm_idx(stencil_indices, n), 0, n - 1)
    vals = flat[:, idx]                                   # (C, M, S)
    W = np.asarray(stencil_weights).astype(flat.dtype)
    return np.einsum("ms,cms->mc", W, vals)


def spread(marker_values, grid_values, stencil_weights, stencil_indices):
    """grid[c, I[m, s]] += marker[m, c] W[m, s] (scatter-add) — ib/stencil.py:81-110."""
    g = np.asarray(grid_values)
    shape = g.shape
    flat = g.reshape(shape[0], -1).copy()
    n = flat.shape[1]
    mv = np.asarray(marker_values).astype(flat.dtype)
    W = np.asarray(stencil_weights).astype(flat.dtype)
    vals = np.einsum("mc,ms->cms", mv, W)
    idx = _norm_idx(stencil_indices, n)
    keep = (idx >= 0) & (idx < n)
    for c in range(shape[0]):
        np.add.at(flat[c], idx[keep], vals[c][keep])
    return flat.reshape(shape)


# ------------------------------------------------------------------ ib/mdf.py

def multi_direct_forcing(grid_u, stencil_weights, stencil_indices, marker_u_target,
                         marker_ds, n_iter=5):
    """Multi-direct forcing — ib/mdf.py:10-64.

    ub = interp(u); repeat n_iter: dF = (U_t - ub) * 2 ds; F += dF;
    ub += interp(0.5 * spread(dF)).  Returns (spread(F), -F).
    """
    gu = np.asarray(grid_u)
    dt = gu.dtype
    ds = np.asarray(marker_ds, dtype=dt).reshape(-1, 1)
    ut = np.asarray(marker_u_target).astype(dt)
    zero = np.zeros_like(gu)
    ub = interpolate(gu, stencil_weights, stencil_indices)
    F = np.zeros_like(ut)
    for _ in range(n_iter):
        dF = (ut - ub) * (ds * 2.0)
        F = F + dF
        gstep = spread(dF, zero, stencil_weights, stencil_indices)
        ub = ub + interpolate(gstep * 0.5, stencil_weights, stencil_indices)
    return spread(F, zero, stencil_weights, stencil_indices), -F


# ------------------------------------------------------------------ ib/geometry.py

def get_area(marker_coords):
    """Shoelace area — ib/geometry.py:6-19."""
    p = np.asarray(marker_coords)
    x, y = p[:, 0], p[:, 1]
    return 0.5 * np.abs(np.sum(x * np.roll(y, 1) - y * np.roll(x, 1)))


def get_ds(marker_coords, closed=True):
    """Per-marker arc-length weights — ib/geometry.py:22-43."""
    p = np.asarray(marker_coords)
    if closed:
        seg = np.linalg.norm(p - np.roll(p, -1, axis=0), axis=1)
        return (seg + np.roll(seg, 1)) / 2.0
    seg = np.linalg.norm(p[1:] - p[:-1], axis=1)
    return np.pad(seg / 2, (1, 0)) + np.pad(seg / 2, (0, 1))


# ------------------------------------------------------------------ ib3d/geometry.py

def get_triangle_areas(vertex_coords, faces):
    """ib3d/geometry.py:6-18."""
    t = np.asarray(vertex_coords)[np.asarray(faces)]
    return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)


def get_surface_area(vertex_coords, faces):
    """ib3d/geometry.py:21-23."""
    return np.sum(get_triangle_areas(vertex_coords, faces))


def get_volume(vertex_coords, faces):
    """ib3d/geometry.py:26-45."""
    t = np.asarray(vertex_coords)[np.asarray(faces)]
    sv = np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])) / 6.0
    return np.abs(np.sum(sv))


def get_ds3d(vertex_coords, faces):
    """Vertex area weights (area/3 per incident triangle) — ib3d/geometry.py:48-67."""
    v = np.asarray(vertex_coords)
    fa = np.asarray(faces)
    a = get_triangle_areas(v, fa)
    out = np.zeros(v.shape[0], dtype=v.dtype)
    np.add.at(out, fa.reshape(-1), np.repeat(a / 3.0, 3))
    return out
