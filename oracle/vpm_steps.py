"""FP64 restatement of the reference ParticleSystem time step (test infrastructure).

TEST INFRASTRUCTURE ONLY -- imported by tests and by the bench's CPU leg, never
by the product package.

Restates the control flow of ``OpenONDA/solvers/VPM/vpmModule.py``:
``update_velocities`` (:699-710), ``update_strengths`` (:713-884, including its
strength-history quirk), the LES filter (:887-962) and ``update_state`` with the
Euler/RK2/RK3/RK4 position updates (:522-692), on plain arrays instead of a list
of Particle objects.  Pair sums come from :mod:`vpm_oracle`.
"""
from __future__ import annotations

import numpy as np

from . import vpm_oracle as O


class ReferenceStepper:
    """Array-state restatement of one reference driver step.

    A driver step is ``update_velocities(); update_strengths(); update_state()``
    (examples/lagrangian_module_examples/E5_LeapFrogging_Rings.py:118-121).
    """

    def __init__(self, x, alpha, sigma, nu, *, dt, method="RK4", flow_model="DNS",
                 viscous_scheme="PSE", variant="cuda", relax=False, nu_t=None,
                 time_step=0, flow_time=0.0, sums=None):
        self.x = np.array(x, dtype=np.float64)
        self.alpha = np.array(alpha, dtype=np.float64)
        self.sigma = np.array(sigma, dtype=np.float64)
        self.nu = np.array(nu, dtype=np.float64)
        self.nu_t = np.zeros_like(self.nu) if nu_t is None else np.array(nu_t, dtype=np.float64)
        self.velocity = np.zeros_like(self.x)
        self.dt = float(dt)
        self.method = method
        self.flow_model = flow_model
        self.viscous_scheme = viscous_scheme
        self.variant = variant
        self.relax = relax
        self.time_step = int(time_step)
        self.flow_time = float(flow_time)
        self.cnu = O.CNU_INIT                         # vpmModule.py:92
        self.enstrophy_ref = None                     # created lazily, :897-898
        self.prev = None                              # created lazily, :741-746
        self.prev2 = None
        self.prev3 = None
        # pair sums: the NumPy oracle, or an object with the same five
        # functions (oracle.cpu_baseline.CudaSums: the C port, GPU-plugin forms)
        self.sums = O if sums is None else sums

    @property
    def nu_eff(self):
        return self.nu + self.nu_t                    # Particle.py:61,104-105

    def _vel(self, x):
        return self.sums.selfinduced_velocity(x, self.alpha, self.sigma, self.variant)

    # vpmModule.py:699-710
    def update_velocities(self):
        self.velocity = self._vel(self.x)

    def _bdf(self, s, rhs):
        """The step-indexed combination used for both stretching and PSE."""
        dt = self.dt
        k = self.time_step
        if k == 0:                                            # :751-754
            return s + dt * rhs
        if k == 1:                                            # :757-760
            return (4 * s - self.prev) / 3 + (2 / 3) * dt * rhs
        if k == 2:                                            # :763-769
            return (18 * s - 9 * self.prev + 2 * self.prev2) / 11 + (6 / 11) * dt * rhs
        return ((48 * s - 36 * self.prev + 16 * self.prev2 - 3 * self.prev3) / 25
                + (12 / 25) * dt * rhs)                       # :772-779

    def _push_history(self, s):
        k = self.time_step                                    # :783-787
        if k >= 3:
            self.prev3 = self.prev2.copy()
        if k >= 2:
            self.prev2 = self.prev.copy()
        self.prev = s.copy()

    # vpmModule.py:713-884
    def update_strengths(self):
        if self.flow_model == "Potential":                    # :719-721
            return
        x, s, radii = self.x, self.alpha.copy(), self.sigma
        if self.flow_model == "LES":                          # :731-737
            vort = self.sums.vorticity_field_at(x, x, s, radii, self.variant)
            ens = self.sums.total_enstrophy(x, s, radii, self.variant)
            self.cnu, self.enstrophy_ref = O.les_filter_constant(ens, self.cnu, self.enstrophy_ref)
            nut, write = O.les_particle_viscosity(vort, radii, self.cnu)
            if write:
                self.nu_t = nut
        if self.prev is None:                                 # :741-746
            self.prev = np.zeros_like(s)
        if self.prev2 is None:
            self.prev2 = np.zeros_like(s)
        if self.prev3 is None:
            self.prev3 = np.zeros_like(s)
        if self.flow_model in ("LES", "DNS"):                 # :749-791
            rhs = self.sums.strength_gradients(x, s, radii)
            s = self._bdf(s, rhs)
            self._push_history(s)
            self.alpha = s.copy()
        if self.viscous_scheme == "PSE":                      # :797-845
            rhs = self.sums.strength_gradients_pse(x, s, radii, self.nu_eff, self.variant)
            s = self._bdf(s, rhs)
            self._push_history(s)
            self.alpha = s.copy()
        elif self.viscous_scheme == "CoreSpreading":          # :851-862
            self.sigma = O.core_spreading(radii, self.nu_eff, self.dt)
        if self.relax:                                        # :867-883
            vort = self.sums.vorticity_field_at(self.x, self.x, self.alpha, self.sigma, self.variant)
            vn = np.sqrt(np.sum(vort * vort, axis=1))
            sn = np.sqrt(np.sum(self.alpha * self.alpha, axis=1))
            self.alpha = self.alpha * 0.95 + 0.05 * (vort / vn[:, None]) * sn[:, None]

    # vpmModule.py:522-692
    def update_state(self):
        x, u1, dt = self.x, self.velocity, self.dt
        m = self.method.strip()
        if m == "Euler":                                      # :563-580
            xn = x + u1 * dt
        elif m == "RK2":                                      # :583-614
            u2 = self._vel(x + u1 * (dt * 0.5))
            xn = x + u2 * dt
        elif m == "RK3":                                      # :617-651
            u2 = self._vel(x + u1 * (dt / 2))
            u3 = self._vel(x + (dt / 2) * u2)
            xn = x + dt * u3
        elif m == "RK4":                                      # :653-692
            k1 = u1
            k2 = self._vel(x + (dt / 2) * k1)
            k3 = self._vel(x + (dt / 2) * k2)
            k4 = self._vel(x + dt * k3)
            xn = x + (dt / 6) * (k1 + 2 * k2 + 2 * k3 + k4)
        else:
            raise ValueError(f"Unknown time integration method: {m}")
        self.flow_time += dt                                  # :553-554
        self.time_step += 1
        self.x = xn

    def step(self):
        self.update_velocities()
        self.update_strengths()
        self.update_state()

    def remove(self, indices):
        """Drops the particles at ``indices`` (vpmModule.py:395-400 pops them
        from the particle list).  The reference leaves its strength-history
        arrays at the old size, so its next multistep update fails on the
        shape; here the history keeps the remaining particles' rows, which is
        the product's defined behaviour.  An empty index list changes nothing,
        as in the reference."""
        keep = np.ones(len(self.x), dtype=bool)
        keep[[i for i in indices if 0 <= i < len(self.x)]] = False
        for k in ("x", "alpha", "sigma", "nu", "nu_t", "velocity", "prev", "prev2", "prev3"):
            v = getattr(self, k)
            if v is not None:
                setattr(self, k, v[keep].copy())
