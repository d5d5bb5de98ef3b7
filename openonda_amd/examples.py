"""The reference's Lagrangian example cases E1-E5 as data.

Each builder returns the initial particle field and the ParticleSystem
settings of one script of ``examples/lagrangian_module_examples`` (physical
parameters, lattice, generator, solver options, weak-particle removal and
step count).  ``coarsen`` multiplies the particle distance (and with it the
radius, the time step and the step count, through the scripts' own
formulas), so the same case can run at the scripts' resolution or small
enough for a CPU reference.

E4 calls ``vpm_fm.get_vortex_ring_vorticity``, which the reference does not
define (``OpenONDA/utilities/vpm_flow_models.py`` has ``vortex_ring_vpm``
with the same arguments); it is built with ``vortex_ring_vpm`` here.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import vpm_flow_models as fm
from . import vpm_solver_helper as sh


@dataclass
class Case:
    name: str
    groups: list                 # [(positions, velocities, strengths, radii, viscosities, group_id)]
    system: dict                 # ParticleSystem keyword arguments
    n_time_steps: int
    remove_weak: dict = field(default_factory=lambda: dict(
        mode="relative", threshold=1e-1, conserve_total_circulation=True))

    def arrays(self):
        """Concatenated (x, v, alpha, sigma, nu, group_id) in add order."""
        cols = list(zip(*self.groups))
        return tuple(np.concatenate(c) for c in cols[:5]) + (
            np.concatenate([np.full(len(g[0]), g[5], dtype=np.int64) for g in self.groups]),)


def _solver(dt, flow_model="LES", viscous_scheme="CoreSpreading", freq=10):
    return dict(flow_model=flow_model, time_step_size=dt, time_integration_method="RK2",
                viscous_scheme=viscous_scheme, processing_unit="GPU",
                monitor_variables=["Circulation", "Kinetic energy"], backup_frequency=freq)


def e1_lamb_oseen(coarsen=1.0) -> Case:
    """E1_Lamb_Oseen_Vortex.py:29-85."""
    core_radius, centre, gamma = 1.0, np.zeros(3), np.pi
    d = 0.25 * core_radius * coarsen
    radius, nu = 0.8 * d ** 0.5, np.pi * 1e-3
    dt = 3 * d ** 2 / gamma
    steps = int(100 / gamma / dt)
    box = np.array([-3.0, 3.0, -3.0, 3.0, -10.0, 10.0]) * core_radius
    pos, vol = sh.get_hexagonal_point_distribution(box, spacing=d)
    v, a, r, n = fm.lamb_oseen_vpm(pos, vol, radius, nu, centre, gamma, core_radius)
    return Case("E1", [(pos, v, a, r, n, 0)], _solver(dt, freq=5), steps)


def e2_vortex_ring(coarsen=1.0) -> Case:
    """E2_Vortex_Ring.py:33-76 (potential flow)."""
    centre, R, gamma = np.zeros(3), 1.0, 1.0
    thickness = 0.2 * R
    d = 0.25 * thickness * coarsen
    radius, nu = 0.8 * d ** 0.5, gamma / 750
    dt = 5 * d ** 2 / gamma
    steps = int(20 * R ** 2 / gamma / dt)
    box = [-2 * thickness, 2 * thickness, -2 * R, 2 * R, -2 * R, 2 * R]
    pos, vol = sh.get_hexagonal_point_distribution(box, spacing=d)
    v, a, r, n = fm.vortex_ring_vpm(pos, vol, radius, nu, centre, R, gamma, thickness / 1.28839)
    return Case("E2", [(pos, v, a, r, n, 0)],
                _solver(dt, flow_model="Potential", viscous_scheme=None, freq=25), steps)


def e3_ring_tube(coarsen=1.0) -> Case:
    """E3_Ring_Tube_Collision.py:32-99."""
    R, gamma = 1.0, 1.0
    thickness = 0.275 * R
    ring_centre, eps_w = np.array([-1.0 * R, 0.0, 0.0]), 0.02 * R
    tube_radius, tube_centre = 0.5 * thickness, np.zeros(3)
    d = 0.2 * thickness * coarsen
    radius, nu = 0.8 * d ** 0.5, gamma / 5000
    dt = 5 * d ** 2 / gamma
    steps = int(75 * R ** 2 / gamma / dt)
    box1 = [-2 * thickness + ring_centre[0], 2 * thickness + ring_centre[0],
            -2 * R, 2 * R, -2 * R, 2 * R]
    box2 = np.array([-2.0, 2.0, -2.0, 2.0, -25.0, 25.0]) * tube_radius
    p1, v1 = sh.get_rectangular_point_distribuition(box1, spacing=d)
    p2, v2 = sh.get_rectangular_point_distribuition(box2, spacing=d)
    g1 = fm.vortex_ring_vpm(p1, v1, radius, nu, ring_centre, R, gamma, thickness, eps_w)
    g2 = fm.lamb_oseen_vpm(p2, v2, radius, nu, tube_centre, gamma, tube_radius)
    return Case("E3", [(p1, *g1, 0), (p2, *g2, 1)], _solver(dt), steps)


def _two_rings(name, R, gammas, thickness, centres, eps, phases, d, re, t_end, corr):
    radius, nu = 0.8 * d ** 0.5, gammas[0] / re
    dt = 5 * d ** 2 / gammas[0]
    steps = int(t_end * R ** 2 / gammas[0] / dt)
    groups = []
    for k in range(2):
        c = centres[k]
        box = [-2 * thickness + c[0], 2 * thickness + c[0], -2 * R, 2 * R, -2 * R, 2 * R]
        p, vol = sh.get_rectangular_point_distribuition(box, spacing=d)
        g = fm.vortex_ring_vpm(p, vol, radius, nu, c, R, gammas[k], thickness / corr, eps,
                               phases[k])
        groups.append((p, *g, k))
    return Case(name, groups, _solver(dt), steps)


def e4_rings_collision(coarsen=1.0) -> Case:
    """E4_Rings_Collision.py:32-106 (head-on rings of opposite circulation)."""
    R, thickness = 1.0, 0.275
    return _two_rings("E4", R, (1.0, -1.0), thickness,
                      (np.array([-1.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0])), 0.04 * R,
                      (0.0, np.pi / 12), 0.2 * thickness * coarsen, 2200, 22, 1.0)


def e5_leapfrog(coarsen=1.0) -> Case:
    """E5_LeapFrogging_Rings.py:32-121 (coaxial rings of equal circulation)."""
    R, thickness = 1.0, 0.1
    return _two_rings("E5", R, (1.0, 1.0), thickness,
                      (np.array([-0.5, 0.0, 0.0]), np.array([0.5, 0.0, 0.0])), 0.04 * R,
                      (0.0, np.pi / 12), 0.25 * thickness * coarsen, 3000, 40, 1.28839)


CASES = {"E1": e1_lamb_oseen, "E2": e2_vortex_ring, "E3": e3_ring_tube,
         "E4": e4_rings_collision, "E5": e5_leapfrog}


def remove_weak(alpha, mode="relative", threshold=1e-1, conserve_total_circulation=True):
    """Indices kept and the strength scale of vpmModule.py:970-1037 (host)."""
    mags = np.linalg.norm(alpha, axis=1)
    if mode == "absolute":
        keep = ~(mags < threshold)
    else:
        keep = ~(mags / mags.max() < threshold)
    scale = 1.0
    if conserve_total_circulation:
        scale = mags.sum() / np.linalg.norm(alpha[keep], axis=1).sum()
    return np.nonzero(keep)[0], scale
