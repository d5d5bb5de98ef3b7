"""Per-particle record, same fields and update rules as the reference's
``OpenONDA/solvers/VPM/Particle.py`` (``Particle``, :7-105).

In this package the particle state lives in SoA device tensors inside
:class:`~openonda_amd.vpmModule.ParticleSystem`; ``Particle`` objects are host
snapshots used by ``add_particle`` and by indexing the system.
"""
from __future__ import annotations

import numpy as np


class Particle:
    """position, velocity, strength (3-vectors), radius, viscosity,
    viscosity_t, viscosity_eff = viscosity + viscosity_t, group_id.

    Defaults as in Particle.py:43-64: random position in
    [-1, domain_size - 1)^3, zero velocity and strength, radius 0.01,
    viscosities 0.
    """

    def __init__(self, position=None, velocity=None, strength=None, radius=None, viscosity=None,
                 viscosity_t=None, domain_size=1.0, group_id=0):
        self.position = (np.array(position, dtype=np.float64) if position is not None
                         else domain_size * np.random.random(3) - 1.0)
        self.velocity = np.array(velocity if velocity is not None else np.zeros(3), dtype=np.float64)
        self.velocity_magnitude = np.linalg.norm(self.velocity)
        self.strength = np.array(strength if strength is not None else np.zeros(3), dtype=np.float64)
        self.strength_magnitude = np.linalg.norm(self.strength)
        self.radius = float(radius) if radius is not None else 0.01
        self.viscosity = float(viscosity) if viscosity is not None else 0.0
        self.viscosity_t = float(viscosity_t) if viscosity_t is not None else 0.0
        self.viscosity_eff = self.viscosity + self.viscosity_t
        self.group_id = group_id

    def update_state(self, position=None, velocity=None, strength=None, radius=None,
                     viscosity=None, viscosity_t=None):
        """Particle.py:67-105."""
        if position is not None:
            assert np.asarray(position).size == 3, "Position array must have 3 elements."
            self.position = np.array(position, dtype=np.float64)
        if velocity is not None:
            assert np.asarray(velocity).size == 3, "Velocity array must have 3 elements."
            self.velocity = np.array(velocity, dtype=np.float64)
            self.velocity_magnitude = np.linalg.norm(self.velocity)
        if strength is not None:
            assert np.asarray(strength).size == 3, "Strength array must have 3 elements."
            self.strength = np.array(strength, dtype=np.float64)
            self.strength_magnitude = np.linalg.norm(self.strength)
        if radius is not None:
            assert isinstance(radius, float), "Radius must be a float."
            self.radius = radius
        if viscosity is not None:
            assert isinstance(viscosity, float), "Viscosity must be a float."
            self.viscosity = viscosity
        if viscosity_t is not None:
            assert isinstance(viscosity_t, float), "Subgrid-scale viscosity must be a float."
            self.viscosity_t = viscosity_t
        if viscosity is not None or viscosity_t is not None:
            self.viscosity_eff = self.viscosity + self.viscosity_t

    def __str__(self):
        p, v, s = self.position, self.velocity, self.strength
        return (f"Location:({p[0]:.2E},{p[1]:.2E},{p[2]:.2E}) m, "
                f"Velocity:({v[0]:.2E},{v[1]:.2E},{v[2]:.2E}) m/s, "
                f"Strength:({s[0]:.2E},{s[1]:.2E},{s[2]:.2E}) m³/s, "
                f"Radius: {self.radius:.2E} m, "
                f"Viscosity: {self.viscosity:.2E} m²/s, "
                f"Subgrid-scale viscosity: {self.viscosity_t:.2E} m²/s")
