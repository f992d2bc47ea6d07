"""Oracle restatement of ``vivsim.dyn`` (rigid-body Newmark-beta and marker kinematics)."""

import numpy as np


def newmark(a, v, d, h, m, k, c, dt=1, gamma=0.5, beta=0.25):
    """Newmark-beta predictor / solve / corrector — dyn.py:5-51."""
    c1, c2 = gamma * dt, beta * dt ** 2
    v1 = v + dt * (1 - gamma) * a
    v2 = d + dt * v + dt ** 2 * (0.5 - beta) * a
    if np.ndim(m) > 0:
        an = np.linalg.solve(m + c1 * c + c2 * k, h - np.dot(c, v1) - np.dot(k, v2))
    else:
        an = (h - c * v1 - k * v2) / (m + c1 * c + c2 * k)
    return an, c1 * an + v1, c2 * an + v2


def newmark_2dof(a, v, d, h, m, k, c, dt=1, gamma=0.5, beta=0.25):
    """dyn.py:55-57."""
    return newmark(a, v, d, h, m, k, c, dt, gamma, beta)


def newmark_3dof(a, v, d, h, m, k, c, dt=1, gamma=0.5, beta=0.25):
    """dyn.py:60-62."""
    return newmark(a, v, d, h, m, k, c, dt, gamma, beta)


def get_markers_coords_2dof(x_markers_init, y_markers_init, d):
    """dyn.py:68-81."""
    return x_markers_init + d[0], y_markers_init + d[1]


def get_markers_coords_3dof(x_markers_init, y_markers_init, x_center_init, y_center_init, d):
    """Rigid rotation by d[2] about the displaced centre — dyn.py:84-100."""
    xr = x_markers_init - x_center_init
    yr = y_markers_init - y_center_init
    cs, sn = np.cos(d[2]), np.sin(d[2])
    return (x_center_init + d[0] + xr * cs - yr * sn,
            y_center_init + d[1] + xr * sn + yr * cs)


def get_markers_velocity_3dof(x_markers, y_markers, x_center_init, y_center_init, d, v):
    """v + omega x r — dyn.py:103-120."""
    xr = x_markers - x_center_init - d[0]
    yr = y_markers - y_center_init - d[1]
    return np.stack([v[0] - v[2] * yr, v[1] + v[2] * xr], axis=-1)


def get_force_to_obj(h_markers):
    """dyn.py:126-136."""
    return np.sum(h_markers, axis=0)


def get_torque_to_obj(x_markers, y_markers, x_center_init, y_center_init, d, h_markers):
    """dyn.py:139-154."""
    xr = x_markers - (x_center_init + d[0])
    yr = y_markers - (y_center_init + d[1])
    return np.sum(xr * h_markers[:, 1] - yr * h_markers[:, 0])
