"""Vortex initial conditions of the reference's ``OpenONDA/utilities/vpm_flow_models.py``.

Same names, arguments, returns and values, so example scripts written against
the reference run unchanged.  Host NumPy, float64 (the field is built once).
"""
from __future__ import annotations

import numpy as np

__all__ = ["gaussian", "lamb_oseen_vpm", "vortex_ring_vpm", "isotropic_turbulence_vpm"]


def gaussian(r, omega_0, a):
    """omega_0 exp(-r^2/a^2) (vpm_flow_models.py:5-6)."""
    return omega_0 * np.exp(-(r ** 2) / a ** 2)


def vortex_ring_vpm(particle_positions, volumes, particle_radius, particle_viscosity,
                    ring_center, ring_radius, ring_strength, ring_thickness,
                    epsilon_W=0.0, phase_W=0.0, N_W=12):
    """Gaussian-core vortex ring with an azimuthal core perturbation.

    Restates vpm_flow_models.py:86-159 (arguments and returns identical):
    returns (velocities, strengths, radii, viscosities).
    """
    p = np.asarray(particle_positions, dtype=np.float64)
    vol = np.asarray(volumes, dtype=np.float64)
    X = p[:, 0] - ring_center[0]
    Y = p[:, 1] - ring_center[1]
    Z = p[:, 2] - ring_center[2]
    theta = np.arctan2(Z, Y)                                       # :127
    Y = Y - epsilon_W * np.cos(phase_W + N_W * theta)              # :128-131
    Z = Z - epsilon_W * np.sin(phase_W + N_W * theta)
    r_yz = np.sqrt(Y ** 2 + Z ** 2)                                # :134
    d = np.abs(r_yz - ring_radius)
    w = (ring_strength / (np.pi * ring_thickness ** 2)) * np.exp(-d ** 2 / ring_thickness ** 2)
    xd = np.exp(-X ** 2 / ring_thickness ** 2)                     # :137
    s = np.empty_like(p)
    s[:, 0] = 0.0                                                  # :140-148
    s[:, 1] = -w * np.sin(theta) * xd * vol
    s[:, 2] = w * np.cos(theta) * xd * vol
    g = np.sum(np.sqrt(np.sum(s ** 2, axis=1)))                    # :150-152
    s *= np.abs(2 * np.pi * ring_radius * ring_strength / g)
    n = len(p)
    return (np.zeros_like(s), s, np.full(n, particle_radius, dtype=np.float64),
            np.full(n, particle_viscosity, dtype=np.float64))


def lamb_oseen_vpm(particle_positions, volumes, particle_radius, particle_viscosity,
                   vortex_center, vortex_strength, core_radius, epsilon_W=0.0, phase_W=0.0,
                   N_W=12):
    """Perturbed Lamb-Oseen column (vpm_flow_models.py:13-82)."""
    p = np.asarray(particle_positions, dtype=np.float64)
    n = len(p)
    px = vortex_center[0] + epsilon_W * np.sin(phase_W + N_W * p[:, 2])
    py = vortex_center[1] + epsilon_W * np.cos(phase_W + N_W * p[:, 2])
    r = np.sqrt((p[:, 0] - px) ** 2 + (p[:, 1] - py) ** 2)
    s = np.zeros_like(p)
    s[:, 2] = (vortex_strength / (np.pi * core_radius ** 2)) * np.exp(-r ** 2 / core_radius ** 2) \
        * np.asarray(volumes, dtype=np.float64)
    return (np.zeros_like(s), s, particle_radius * np.ones(n), particle_viscosity * np.ones(n))


def _periodic_centred_difference(f, axis, h):
    # (f[i+1] - f[i-1]) / 2h with periodic wrap
    return (np.roll(f, -1, axis=axis) - np.roll(f, 1, axis=axis)) / (2 * h)


def isotropic_turbulence_vpm(box_size: float, particle_distance: float, particle_radius: float):
    """Particles on a periodic lattice carrying a random solenoidal velocity
    field with E(k) ~ k^-5/3 exp(-(k/k_d)^2) (vpm_flow_models.py:167-271).

    Returns (positions, velocities, strengths, radii).  The lattice is
    arange(-L/2, L/2, d) per axis (n points), x-major.  Spectrum: k_max =
    pi n/L, k_d = k_max/3, E(0) = 0; one set of random phases
    exp(2 pi i U[0,1)) drawn from the legacy NumPy generator seeded with 0
    (the reference calls np.random.seed(0); here a private RandomState(0)
    draws the same numbers without reseeding the global generator) is shared
    by the three components, which are then projected normal to k with the
    reference's 1e-10 guard, and transformed back (real part).

    Strengths are (4/3) d^3 times a periodic centred-difference vorticity in
    which, as in the reference, each component keeps only its first term:
    its second term stands on a line of its own and is discarded, so
    w_x = d_y u_z, w_y = d_z u_x, w_z = d_x u_y.
    """
    axis_pts = np.arange(-box_size / 2, box_size / 2, step=particle_distance)
    gx, gy, gz = np.meshgrid(axis_pts, axis_pts, axis_pts, indexing="ij")
    positions = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)

    n = len(axis_pts)
    k1 = np.fft.fftfreq(n, box_size / n) * 2 * np.pi
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    k_diss = (np.pi * n / box_size) / 3
    with np.errstate(divide="ignore"):
        E = K ** (-5 / 3) * np.exp(-(K / k_diss) ** 2)
    E[K == 0] = 0
    phases = np.exp(2j * np.pi * np.random.RandomState(0).rand(n, n, n))
    amp = np.sqrt(E) * phases
    Kg = K + 1e-10
    along = (KX * amp + KY * amp + KZ * amp) / Kg
    u = [np.real(np.fft.ifftn(amp - along * Kc / Kg)) for Kc in (KX, KY, KZ)]

    velocities = np.stack([c.reshape(-1) for c in u], axis=1)
    w = np.stack([_periodic_centred_difference(u[2], 1, particle_distance),
                  _periodic_centred_difference(u[0], 2, particle_distance),
                  _periodic_centred_difference(u[1], 0, particle_distance)], axis=-1)
    strengths = w.reshape(n * n * n, 3) * (4 / 3) * particle_distance ** 3
    radii = particle_radius * np.ones_like(positions[:, 0])
    return positions, velocities, strengths, radii
