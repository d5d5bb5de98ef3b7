"""Particle lattices of the reference's ``OpenONDA/utilities/vpm_solver_helper.py``.

Same function names (including the reference's spelling), arguments, point
order and volumes, so example scripts written against the reference run
unchanged.  Host NumPy, float64: these build the initial particle field once.
"""
from __future__ import annotations

import numpy as np

__all__ = ["gaussian", "compute_min_max", "get_rectangular_point_distribuition",
           "get_2D_rectangular_point_distribuition", "get_hexagonal_point_distribution",
           "get_cylindrical_point_distribuition"]


def gaussian(r, omega_0, a):
    """omega_0 exp(-r^2/a^2) (vpm_solver_helper.py:5-6)."""
    return omega_0 * np.exp(-(r ** 2) / a ** 2)


def compute_min_max(x, y, z):
    """(xmin, xmax, ymin, ymax, zmin, zmax) (vpm_solver_helper.py:10-16)."""
    return (np.min(x), np.max(x), np.min(y), np.max(y), np.min(z), np.max(z))


def _odd_count(lo, hi, spacing):
    # points per axis including both ends, bumped to an odd count
    n = int(np.ceil((hi - lo) / spacing)) + 1
    return n + 1 if n % 2 == 0 else n


def get_rectangular_point_distribuition(box_limits, spacing: float):
    """Box lattice, x-major order, equal volumes summing to the box volume
    (vpm_solver_helper.py:21-68)."""
    b = [float(v) for v in box_limits]
    axes = [np.linspace(b[2 * k], b[2 * k + 1], num=_odd_count(b[2 * k], b[2 * k + 1], spacing))
            for k in range(3)]
    grid = np.meshgrid(*axes, indexing="ij")
    points = np.stack([g.ravel() for g in grid], axis=1)
    box = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])
    return points, np.full(points.shape[0], box / points.shape[0])


def get_2D_rectangular_point_distribuition(box_limits, spacing: float):
    """Planar (z = 0) lattice over [xmin, xmax] x [ymin, ymax]; volumes are the
    box area shared equally (vpm_solver_helper.py:71-117)."""
    b = [float(v) for v in box_limits]
    xs = np.linspace(b[0], b[1], num=_odd_count(b[0], b[1], spacing))
    ys = np.linspace(b[2], b[3], num=_odd_count(b[2], b[3], spacing))
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    points = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    area = (b[1] - b[0]) * (b[3] - b[2])
    return points, np.full(points.shape[0], area / points.shape[0])


def get_hexagonal_point_distribution(box_limits, spacing: float):
    """Rows of spacing ``spacing`` along x, rows sqrt(3)/2 spacing apart in y with
    every other row shifted by half a spacing, planes ``spacing`` apart in z.
    Order: z, then row, then x.  Each point carries spacing^3 sqrt(3)/2
    (vpm_solver_helper.py:120-161)."""
    x_min, x_max, y_min, y_max, z_min, z_max = (float(v) for v in box_limits)
    dy = spacing * np.sqrt(3) / 2
    xs = np.arange(x_min, x_max + spacing, spacing)
    ys = np.arange(y_min, y_max + dy, dy)
    zs = np.arange(z_min, z_max + spacing, spacing)
    shift = np.where(np.arange(len(ys)) % 2 == 1, spacing / 2, 0.0)
    gz, gy, gx = np.meshgrid(zs, ys, xs, indexing="ij")
    gs = np.broadcast_to(shift[None, :, None], gx.shape)
    points = np.stack([(gx + gs).ravel(), gy.ravel(), gz.ravel()], axis=1)
    if points.shape[0] == 0:
        points = points.reshape(0, 3)
    return points, np.full(points.shape[0], spacing ** 3 * np.sqrt(3) / 2)


# coordinate slot of (r cos(theta), r sin(theta), axial) for each axis
_AXES = {"x": (1, 2, 0), "y": (0, 2, 1), "z": (0, 1, 2)}


def get_cylindrical_point_distribuition(particle_distance: float, radius=1.0, height=2.0,
                                        center=np.array([0, 0, 0]), axis="x"):
    """Polar lattice in a cylinder of ``radius`` and ``height`` along ``axis``
    (vpm_solver_helper.py:164-249).

    int(R/d) radii from dr to R, int(H/d)+1 axial stations over [-H/2, H/2],
    and int(pi R/d) - 1 angles: linspace(0, 2 pi, int(pi R/d)) without its
    last point -- spaced 2 pi/(int(pi R/d) - 1) although each volume
    r dtheta dr dz uses dtheta = 2 pi/int(pi R/d), as in the reference.
    Order: radius, station, angle; then one point per station on the axis
    with volume pi R^2 dz / (number of stations).
    """
    if axis not in _AXES:
        raise ValueError("Invalid axis. Choose from 'x', 'y', or 'z'.")
    n_r = int(radius / particle_distance)
    n_h = int(height / particle_distance)
    n_t = int(np.pi * radius / particle_distance)
    dtheta = 2 * np.pi / n_t
    dz = height / n_h
    dr = radius / n_r
    r = np.linspace(dr, radius, num=n_r)
    z = np.linspace(-height / 2, height / 2, num=n_h + 1)
    theta = np.linspace(0, 2 * np.pi, num=n_t)[:-1]
    c = np.asarray(center, dtype=np.float64)
    gr, gz, gt = np.meshgrid(r, z, theta, indexing="ij")
    local = (gr * np.cos(gt), gr * np.sin(gt), gz)       # (c1, c2, axial)
    slots = _AXES[axis]
    ring = np.empty((gr.size, 3))
    on_axis = np.empty((len(z), 3))
    for k in range(3):
        ring[:, slots[k]] = local[k].ravel() + c[slots[k]]
        on_axis[:, slots[k]] = (z if k == 2 else 0.0) + c[slots[k]]
    vol_ring = ((gr * dtheta) * dr * dz).ravel()
    vol_axis = np.full(len(z), np.pi * radius ** 2 * dz / len(z))
    return np.concatenate([ring, on_axis]), np.concatenate([vol_ring, vol_axis])
