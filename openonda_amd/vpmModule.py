"""Device-resident vortex particle system with the reference's call surface.

``ParticleSystem`` keeps the constructor arguments, attributes and methods of
``OpenONDA/solvers/VPM/vpmModule.py`` (class at :27) so that a driver written
for the reference -- ``update_velocities(); update_strengths();
update_state()`` in a loop (examples/lagrangian_module_examples/
E5_LeapFrogging_Rings.py:118-121) -- runs unchanged.  The state lives in SoA
float64 tensors on the HIP device instead of a Python list of Particle
objects; every all-pairs sum runs in libvpm_hip.so and every O(N) update
(RK stages, the strength update, core spreading, the LES filter, relaxation,
impulses) in its device kernels, so a step transfers nothing to the host.

Deliberate differences (each also noted where it happens):
 * ``processing_unit`` selects whose arithmetic is reproduced -- "GPU": the
   reference's GPU plugin (ParticlesPhysicsCuda.py), "CPU": its Numba CPU
   plugin (ParticlesPhysics.py) -- and both run on the MI355X.  Without a
   HIP device the constructor raises instead of silently switching plugin
   (vpmModule.py:98-112).
 * ``precision`` chooses the pair-sum kernels: "plain" (default; the FP64
   restatement of the reference arithmetic), "f64" (fused FP64 pass) or
   "f32c" (fused FP32-compensated pass, GPU arithmetic only).
 * Messages are printed only with ``verbose=True``.
 * Backups are VTK XML PolyData written by :mod:`openonda_amd.vtp` (same
   arrays as vpmModule.py:498-515); ``write_backups=False`` disables them.
 * ``particles`` / indexing return host snapshots (Particle objects); mutate
   the system through its methods.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import ParticlesPhysicsHip
from . import _lib as L
from . import physics as P
from . import vtp
from .Particle import Particle

epsilon = 1e-9  # vpmModule.py:22


class ParticleSystem:
    """Manages particles on the HIP device and advances them in time.

    Arguments are those of vpmModule.py:49-61, plus the keyword-only
    ``precision``, ``device``, ``verbose``, ``write_backups`` and
    ``check_finite`` (scan the state after every update_state and raise
    NonFiniteError on NaN/Inf, see check_state).
    """

    def __init__(self, time_step_size: float = 1.0, time_integration_method: str = "RK2",
                 processing_unit: str = "CPU", flow_model: str = "DNS",
                 viscous_scheme: str = "CoreSpreading", monitor_variables: list = None,
                 flow_time: float = 0.0, time_step: int = 0, backup_frequency: int = 1,
                 backup_filename: str = "particle_data", relax_strength_solution: bool = False,
                 *, precision: str = "plain", device=None, verbose: bool = False,
                 write_backups: bool = True, check_finite: bool = False):
        if monitor_variables is None:                                 # :65-71
            monitor_variables = ["Circulation", "Linear Impulse", "Angular Impulse",
                                 "Kinetic Energy", "Enstrophy", "Helicity"]
        self._validate_float(time_step_size, "time_step_size")      # :74-78
        self._validate_integration_method(time_integration_method)
        self._validate_flow_model(flow_model)
        self._validate_viscous_scheme(viscous_scheme)
        self._validate_processing_unit(processing_unit)
        if precision not in ("plain", "f64", "f32c"):
            raise ValueError("precision must be 'plain', 'f64' or 'f32c'")
        if precision == "f32c" and processing_unit == "CPU":
            raise ValueError("precision 'f32c' reproduces the GPU plugin only (processing_unit='GPU')")
        P.require_gpu()
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.time_step_size = time_step_size                         # :81-93
        self.flow_time = flow_time
        self.time_step = time_step
        self.flow_model = flow_model
        self.viscous_scheme = viscous_scheme
        self.monitor_variables = monitor_variables
        self.backup_filename = backup_filename
        self.backup_frequency = backup_frequency
        self.time_integration_method = time_integration_method
        self.processing_unit = processing_unit
        self.relax_strength_solution = relax_strength_solution
        self.precision = precision
        self.verbose = verbose
        self.write_backups = write_backups
        self.check_finite = check_finite
        self.variant = "cuda" if processing_unit == "GPU" else "cpu"
        self.physics = ParticlesPhysicsHip                           # :141-154
        # C_nu, enstrophy_ref, has_ref: device state of the LES filter
        self._les = torch.tensor([0.02958, 0.0, 0.0], dtype=torch.float64, device=self.device)
        self._alloc(0)

        if self.flow_model == "Potential":                           # :117-137
            self.flow_model_description = "Potential flow"
            if self.viscous_scheme is not None:
                raise ValueError("For 'Potential' flow model, 'viscous_scheme' must be None.")
        elif self.flow_model == "pseudo2D":
            self.flow_model_description = "(ν)(∇²)ω"
            if self.viscous_scheme not in ["CoreSpreading", "PSE"]:
                raise ValueError("For 'pseudo2D' flow model, 'viscous_scheme' must be 'CoreSpreading' or 'PSE'.")
        elif self.flow_model == "LES":
            self.flow_model_description = "(ω.∇)u + (ν+νt)(∇²)ω"
            if self.viscous_scheme not in ["CoreSpreading", "PSE"]:
                raise ValueError("For 'LES' flow model, 'viscous_scheme' must be 'CoreSpreading' or 'PSE'.")
        elif self.flow_model == "DNS":
            self.flow_model_description = "(ω.∇)u + (ν)(∇²)ω"
            if self.viscous_scheme not in ["CoreSpreading", "PSE"]:
                raise ValueError("For 'DNS' flow model, 'viscous_scheme' must be 'CoreSpreading' or 'PSE'.")
        else:
            raise ValueError(f"Unsupported flow model: {self.flow_model}")

    # ------------------------------------------------------------ validation
    def _validate_float(self, value, name):                          # :158-160
        if not isinstance(value, (float, int, np.float64)):
            raise TypeError(f"{name} must be a float or int, got {type(value).__name__}")

    def _validate_integration_method(self, method):                  # :163-166
        valid = {"Euler", "RK2", "RK3", "RK4"}
        if method not in valid:
            raise ValueError(f"Invalid time integration method '{method}'. Choose from {valid}.")

    def _validate_viscous_scheme(self, scheme):                      # :169-172
        valid = {"PSE", "CoreSpreading", None}
        if scheme not in valid:
            raise ValueError(f"Invalid viscous scheme '{scheme}'. Choose from {valid}.")

    def _validate_processing_unit(self, unit):                       # :175-178
        valid = {"CPU", "GPU"}
        if unit not in valid:
            raise ValueError(f"Invalid processing unit '{unit}'. Choose from {valid}.")

    def _validate_flow_model(self, method):                          # :181-184
        valid = {"Potential", "pseudo2D", "DNS", "LES"}
        if method not in valid:
            raise ValueError(f"Invalid time integration method '{method}'. Choose from {valid}.")

    # ------------------------------------------------------------ state
    def _alloc(self, n):
        d, f = self.device, torch.float64
        self._pos = torch.zeros((3, n), dtype=f, device=d)
        self._vel = torch.zeros((3, n), dtype=f, device=d)
        self._str = torch.zeros((3, n), dtype=f, device=d)
        self._sigma = torch.zeros((n,), dtype=f, device=d)
        self._nu = torch.zeros((n,), dtype=f, device=d)
        self._nut = torch.zeros((n,), dtype=f, device=d)
        self._gid = torch.zeros((n,), dtype=torch.int64, device=d)
        self._hist = None
        self._unif = None

    def _set_state(self, pos, vel, st, sig, nu, nut, gid, hist=None):
        """Replaces the particle arrays; ``hist`` is the strength history of the
        multistep update for the new set (None: none yet)."""
        self._pos, self._vel, self._str = pos.contiguous(), vel.contiguous(), st.contiguous()
        self._sigma, self._nu, self._nut, self._gid = (sig.contiguous(), nu.contiguous(),
                                                       nut.contiguous(), gid.contiguous())
        self._hist = None if hist is None else [h.contiguous() for h in hist]
        self._unif = None

    @property
    def n(self) -> int:
        return int(self._pos.shape[1])

    @property
    def Cnu(self) -> float:
        return float(self._les[0].item())

    @Cnu.setter
    def Cnu(self, v):
        self._les[0] = float(v)

    @property
    def enstrophy_ref(self):
        return float(self._les[1].item()) if self._les[2].item() else None

    @property
    def previous_strengths(self):
        return None if self._hist is None else P.from_soa(self._hist[0])

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _ws(self, nbytes):
        return P.workspace(self.device, nbytes)

    # ------------------------------------------------------------ system info
    def __len__(self):
        return self.n

    def __getitem__(self, index):
        if isinstance(index, slice):
            return [self[i] for i in range(*index.indices(self.n))]
        if index < 0:
            index += self.n
        if not 0 <= index < self.n:
            raise IndexError("particle index out of range")
        return Particle(position=self._pos[:, index].cpu().numpy(),
                        velocity=self._vel[:, index].cpu().numpy(),
                        strength=self._str[:, index].cpu().numpy(),
                        radius=float(self._sigma[index]), viscosity=float(self._nu[index]),
                        viscosity_t=float(self._nut[index]), group_id=int(self._gid[index]))

    def __iter__(self):
        return iter(self.particles)

    @property
    def particles(self):
        """Host snapshots of all particles (the reference's list, vpmModule.py:81)."""
        pos, vel, st = (P.from_soa(t) for t in (self._pos, self._vel, self._str))
        sig, nu, nut = (t.cpu().numpy() for t in (self._sigma, self._nu, self._nut))
        gid = self._gid.cpu().numpy()
        return [Particle(position=pos[i], velocity=vel[i], strength=st[i], radius=float(sig[i]),
                         viscosity=float(nu[i]), viscosity_t=float(nut[i]), group_id=int(gid[i]))
                for i in range(self.n)]

    def __str__(self):
        return (f"\n{'=' * 60}\n"
                f"Particle System with {self.n} particles\n"
                f"{'-' * 60}\n"
                f"Starting from step:  {self.time_step}\n"
                f"Simulation Time:     {self.flow_time:.3E} s\n"
                f"Time Step Size:      {self.time_step_size:.3E} s\n"
                f"Viscous scheme:      {self.viscous_scheme}\n"
                f"Total strength:      {self.get_total_strength_magnitude():.3E} m³/s\n"
                f"Integration Method:  {self.time_integration_method}\n"
                f"Processing Unit:     {self.processing_unit}\n"
                f"Flow Model:          {self.flow_model}\n"
                f"Monitor Variables:   {', '.join(self.monitor_variables)}\n"
                f"Backup Frequency:    Every {self.backup_frequency} time steps\n"
                f"Backup Filename:     {self.backup_filename}\n"
                f"{'=' * 60}")

    def _log(self, msg):
        if self.verbose:
            print(msg)

    # ------------------------------------------------------------ getters (vpmModule.py:219-251)
    def get_particle_positions(self):
        return P.from_soa(self._pos)

    def get_particle_group_id(self):
        return self._gid.cpu().numpy()

    def get_particle_strengths(self):
        return P.from_soa(self._str)

    def get_particle_velocities(self):
        return P.from_soa(self._vel)

    def get_particle_radii(self):
        return self._sigma.cpu().numpy()

    def get_particle_viscosities(self):
        return self._nu.cpu().numpy()

    def get_particle_viscosities_t(self):
        return self._nut.cpu().numpy()

    def get_particle_viscosities_eff(self):
        return (self._nu + self._nut).cpu().numpy()

    def get_particle_strength_magnitudes(self) -> np.ndarray:
        return torch.linalg.vector_norm(self._str, dim=0).cpu().numpy()

    # device-tensor views (SoA) for callers that stay on the device
    @property
    def positions(self):
        return self._pos

    @property
    def strengths(self):
        return self._str

    @property
    def velocities(self):
        return self._vel

    @property
    def radii(self):
        return self._sigma

    # ------------------------------------------------------------ pair sums
    def _sigma_mode(self):
        """Uniform-radius kernels when all radii are equal (checked once per change of sigma)."""
        if self._unif is None:
            self._unif = P.sigma_is_uniform(self._sigma)
        return "uniform" if self._unif else "general"

    def _velocity(self, pos):
        if self.precision == "plain":
            return P.selfinduced_velocity(pos, self._str, self._sigma, self.variant)
        return P.forward(pos, self._str, self._sigma, outputs=("vel",), precision=self.precision,
                         sigma_mode=self._sigma_mode())["vel"]

    def _stretching(self, pos, st, sig):
        if self.precision == "plain":
            return P.strength_gradients(pos, st, sig)
        return P.forward(pos, st, sig, outputs=("stretch",), precision=self.precision,
                         sigma_mode=self._sigma_mode())["stretch"]

    def _pse(self, pos, st, sig, nu):
        if self.precision == "plain" or self.variant == "cpu":
            return P.strength_gradients_pse(pos, st, sig, nu, self.variant)
        return P.forward(pos, st, sig, nu, outputs=("pse",), precision=self.precision,
                         sigma_mode=self._sigma_mode())["pse"]

    def _vort_enstrophy(self, pos, st, sig):
        if self.precision == "plain" or self.variant == "cpu":
            return (P.vorticity_field_at(pos, pos, st, sig, self.variant),
                    P.total_enstrophy(pos, st, sig, self.variant).reshape(1))
        r = P.forward(pos, st, sig, outputs=("vort", "enstrophy"), precision=self.precision,
                      sigma_mode=self._sigma_mode())
        return r["vort"], r["enstrophy"].reshape(1)

    def _diag10(self):
        out = torch.empty(10, dtype=torch.float64, device=self.device)
        nb = L.load().vpm_state_workspace_bytes(self.n)
        ws = self._ws(nb)
        L.call("vpm_diagnostics", self.n, max(self.n, 1), self._pos.data_ptr(),
               self._str.data_ptr(), out.data_ptr(), ws.data_ptr(), ws.numel(), self._stream())
        return out

    def get_particle_vorticities(self) -> np.ndarray:               # :254-261
        return P.from_soa(P.vorticity_field_at(self._pos, self._pos, self._str, self._sigma,
                                               self.variant))

    def get_total_linear_impulse(self) -> np.ndarray:               # :264-267
        return self._diag10()[0:3].cpu().numpy()

    def get_total_angular_impulse(self) -> np.ndarray:              # :270-274
        return self._diag10()[3:6].cpu().numpy()

    def get_total_strength(self):                                   # :277-279
        return self._diag10()[6:9].cpu().numpy()

    def get_total_strength_magnitude(self) -> float:                # :282-284
        return float(self._diag10()[9].item())

    def get_total_enstrophy(self) -> float:                         # :287-295
        return float(P.total_enstrophy(self._pos, self._str, self._sigma, self.variant).item())

    def get_total_kinetic_energy(self) -> float:                    # :297-304
        return float(P.total_kinetic_energy(self._pos, self._str, self._sigma, self.variant).item())

    def get_total_helicity(self) -> float:                          # :307-322
        return float(P.total_helicity(self._pos, self._str, self._sigma, self.variant).item())

    def get_vorticity_field_at(self, grid_positions):               # :328-335
        y = P.to_soa(np.asarray(grid_positions, dtype=np.float64), self.device)
        return P.from_soa(P.vorticity_field_at(y, self._pos, self._str, self._sigma, self.variant))

    def get_induced_velocity_at(self, grid_positions):              # :338-345
        y = P.to_soa(np.asarray(grid_positions, dtype=np.float64), self.device)
        return P.from_soa(P.induced_velocity_at(y, self._pos, self._str, self._sigma))

    # ------------------------------------------------------------ diagnostics
    def log_diagnostics(self):                                      # :351-383
        """Prints the monitored quantities.  Like the reference, it tests the
        keys 'Kinetic energy', 'Linear impulse' and 'Angular impulse' (lower
        case), so the constructor's default names only print circulation,
        enstrophy and helicity."""
        print("\n" + "=" * 60)
        print(f"{'Flow Diagnostics':^5}")
        print("-" * 60)
        print(f"Flow time: {self.flow_time:0.3E} s")
        print(f"Time step: {self.time_step}\n")
        mv = self.monitor_variables
        if "Circulation" in mv:
            print(f"{'Circulation':<20}: {self.get_total_strength_magnitude():0.3E} m³/s")
        if "Kinetic energy" in mv:
            print(f"{'Kinetic energy':<20}: {self.get_total_kinetic_energy():0.3E} m⁵/s²")
        if "Linear impulse" in mv:
            I = self.get_total_linear_impulse()
            print(f"{'Linear Impulse':<20}: ({I[0]:0.3E}, {I[1]:0.3E}, {I[2]:0.3E}) m⁴/s")
        if "Angular impulse" in mv:
            A = self.get_total_angular_impulse()
            print(f"{'Angular Impulse':<20}: ({A[0]:0.3E}, {A[1]:0.3E}, {A[2]:0.3E}) m⁵/s")
        if "Enstrophy" in mv:
            print(f"{'Total enstrophy':<20}: {self.get_total_enstrophy():0.3E} m³/s²")
        if "Helicity" in mv:
            print(f"{'Total helicity':<20}: {self.get_total_helicity():0.3E} m²/s²")
        print("=" * 60 + "\n")

    # ------------------------------------------------------------ particle management
    def add_particle(self, particle: Particle):                     # :390-392
        self.add_particle_field(np.asarray(particle.position, dtype=np.float64)[None, :],
                                np.asarray(particle.velocity, dtype=np.float64)[None, :],
                                np.asarray(particle.strength, dtype=np.float64)[None, :],
                                np.array([particle.radius], dtype=np.float64),
                                np.array([particle.viscosity], dtype=np.float64),
                                np.array([particle.viscosity_t], dtype=np.float64),
                                np.array([int(particle.group_id)], dtype=np.int64))

    def remove_particles(self, particle_indices):                  # :395-400
        """Removes the particles at the given indices (out-of-range ones are
        ignored, as in the reference).  The strength history of the multistep
        update keeps the rows of the remaining particles, so the next
        update_strengths continues the same scheme; the reference keeps its
        history arrays at the old size, which fails on the next step after a
        real removal and continues unchanged when nothing was removed."""
        idx = sorted(set(int(i) for i in particle_indices))
        idx = [i for i in idx if 0 <= i < self.n]
        if idx:
            keep = torch.ones(self.n, dtype=torch.bool, device=self.device)
            keep[torch.tensor(idx, device=self.device)] = False
            hist = None if self._hist is None else [h[:, keep] for h in self._hist]
            self._set_state(self._pos[:, keep], self._vel[:, keep], self._str[:, keep],
                            self._sigma[keep], self._nu[keep], self._nut[keep], self._gid[keep],
                            hist)
        self._log(f"(info) Removed {len(particle_indices)} particles")

    def remove_all_particles(self):                                # :403-406
        self._alloc(0)
        self._log("(info) Removed all particles from particle system.")

    def print_particle_information(self, index=None):              # :409-419
        print("\nParticle system information ({} particles):".format(len(self)))
        if index is None:
            for p in self.particles:
                print(p)
        elif isinstance(index, int):
            print(self[index])

    def add_particle_field(self, positions, velocities, strengths, radii, viscosities,
                           viscosities_t=None, group_id=None):     # :422-461
        positions = np.asarray(positions, dtype=np.float64)
        velocities = np.asarray(velocities, dtype=np.float64)
        strengths = np.asarray(strengths, dtype=np.float64)
        radii = np.asarray(radii, dtype=np.float64).reshape(-1)
        viscosities = np.asarray(viscosities, dtype=np.float64).reshape(-1)
        N = positions.shape[0]
        if positions.shape[1] != 3 or velocities.shape[1] != 3 or strengths.shape[1] != 3:
            raise ValueError("Positions, velocities, and strengths must be of shape (N x 3).")
        if radii.shape[0] != N or viscosities.shape[0] != N:
            raise ValueError("Radii and viscosities must have the same number of elements as positions.")
        if viscosities_t is None:
            viscosities_t = np.zeros_like(viscosities, dtype=np.float64)
        if group_id is None:
            group_id = np.zeros(N, dtype=np.int64)
        elif isinstance(group_id, (int, np.integer)):
            group_id = np.full(N, int(group_id), dtype=np.int64)
        d = self.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(d)
        new_str = P.to_soa(strengths, d)
        # Particles added after the multistep strength update has started have
        # no history of their own: it is taken as their current strength (a
        # constant past), so their first multistep update is consistent
        # instead of scaling the strength by the scheme's leading coefficient.
        # The reference's history arrays keep the old size (it fails there).
        hist = None if self._hist is None else [torch.cat([h, new_str], 1) for h in self._hist]
        self._set_state(torch.cat([self._pos, P.to_soa(positions, d)], 1),
                        torch.cat([self._vel, P.to_soa(velocities, d)], 1),
                        torch.cat([self._str, new_str], 1),
                        torch.cat([self._sigma, t(radii)]), torch.cat([self._nu, t(viscosities)]),
                        torch.cat([self._nut, t(np.asarray(viscosities_t, dtype=np.float64).reshape(-1))]),
                        torch.cat([self._gid, t(np.asarray(group_id, dtype=np.int64).reshape(-1))]),
                        hist)

    def load_particle_field_from_backup(self, filename: str):      # :466-490
        pts, d = vtp.read_vtp(filename)
        self.add_particle_field(pts, d["Velocity"], d["Strength"], d["Radius"], d["Viscosity"],
                                d["Viscosity_t"], np.asarray(d["Group_ID"], dtype=np.int64))

    def backup_particle_field(self):                               # :492-516
        name = f"{self.backup_filename}_{str(self.time_step).zfill(7)}.vtp"
        vtp.write_vtp(name, self.get_particle_positions(), {
            "Velocity": self.get_particle_velocities(), "Strength": self.get_particle_strengths(),
            "Radius": self.get_particle_radii(), "Group_ID": self.get_particle_group_id(),
            "Viscosity": self.get_particle_viscosities(),
            "Viscosity_t": self.get_particle_viscosities_t()})
        self._log(f"• Particle data exported to {name}")

    # The VTP backup above holds the particle fields only, as the reference's
    # does.  A checkpoint adds what a bit-identical resume also needs: the
    # strength history of the multistep update (vpmModule.py:741-746,783-787),
    # the LES filter state (C_nu, enstrophy reference; :887-909) and the clock.
    _CKPT_VERSION = 1

    def save_checkpoint(self, filename: str):
        """Writes the full solver state to an .npz file (plain arrays, no pickles)."""
        arrays = {
            "version": np.array(self._CKPT_VERSION, dtype=np.int64),
            "positions": self._pos.cpu().numpy(), "velocities": self._vel.cpu().numpy(),
            "strengths": self._str.cpu().numpy(), "radii": self._sigma.cpu().numpy(),
            "viscosities": self._nu.cpu().numpy(), "viscosities_t": self._nut.cpu().numpy(),
            "group_id": self._gid.cpu().numpy(), "les": self._les.cpu().numpy(),
            "clock": np.array([self.flow_time, float(self.time_step), self.time_step_size]),
            "has_history": np.array(self._hist is not None),
        }
        if self._hist is not None:
            arrays["history"] = torch.stack(self._hist).cpu().numpy()
        with open(filename, "wb") as f:
            np.savez(f, **arrays)
        self._log(f"• Checkpoint written to {filename}")

    def load_checkpoint(self, filename: str):
        """Restores the state written by :meth:`save_checkpoint` (replacing the
        current particles); the solver settings of this instance are kept."""
        with np.load(filename, allow_pickle=False) as z:
            if int(z["version"]) != self._CKPT_VERSION:
                raise ValueError(f"unsupported checkpoint version {int(z['version'])}")
            d, f64 = self.device, torch.float64
            t = lambda k, dt=f64: torch.from_numpy(np.ascontiguousarray(z[k])).to(d, dt)
            self._set_state(t("positions"), t("velocities"), t("strengths"), t("radii"),
                            t("viscosities"), t("viscosities_t"), t("group_id", torch.int64))
            self._les = t("les")
            self.flow_time, step, self.time_step_size = (float(v) for v in z["clock"])
            self.time_step = int(step)
            if bool(z["has_history"]):
                h = t("history")
                self._hist = [h[k].contiguous() for k in range(h.shape[0])]

    # ------------------------------------------------------------ time integration
    def update_state(self):                                        # :522-560
        """Advances positions with the chosen RK method (stages use the
        strengths and radii produced by update_strengths)."""
        self._log("• Advecting particles and updating time-step.")
        self.time_integration_method = self.time_integration_method.strip()
        methods = {"Euler": self._euler_update_position, "RK2": self._rk2_update_position,
                   "RK3": self._rk3_update_position, "RK4": self._rk4_update_position}
        if self.time_integration_method not in methods:
            raise ValueError(f"Unknown time integration method: {self.time_integration_method}")
        new_positions = methods[self.time_integration_method](self._pos, self._vel)
        if self.write_backups and self.time_step % self.backup_frequency == 0:
            self.backup_particle_field()
            self.log_diagnostics()
        self.flow_time += self.time_step_size
        self.time_step += 1
        self._log(f"\n>>> Flow time: {self.flow_time:0.3E} s, Time-step: {self.time_step:d}\n")
        self._pos = new_positions
        if self.check_finite:
            self.check_state(f"state after time-step {self.time_step}")

    def check_state(self, where="particle state"):
        """Raises NonFiniteError (status VPM_ERR_NONFINITE) if any position or
        strength is NaN/Inf, a radius is not > 0 or a viscosity is negative
        (one device scan and a 4-byte read back).  No reference counterpart:
        the reference carries NaN on silently."""
        if self.n:
            P.check_particles(self._pos, self._str, self._sigma, self._nu, where=where)

    def _axpy(self, x, y, c):
        out = torch.empty_like(x)
        L.call("vpm_axpy3", self.n, max(self.n, 1), x.data_ptr(), y.data_ptr(), float(c),
               out.data_ptr(), self._stream())
        return out

    def _euler_update_position(self, positions, velocities_1):     # :563-580
        return self._axpy(positions, velocities_1, self.time_step_size)

    def _rk2_update_position(self, positions, velocities_1):       # :583-614
        dt = self.time_step_size
        p1 = self._axpy(positions, velocities_1, dt * 0.5)
        v2 = self._velocity(p1)
        return self._axpy(positions, v2, dt)

    def _rk3_update_position(self, positions, velocities_1):       # :617-651
        dt = self.time_step_size
        p1 = self._axpy(positions, velocities_1, dt / 2)
        v2 = self._velocity(p1)
        p2 = self._axpy(positions, v2, dt / 2)
        v3 = self._velocity(p2)
        return self._axpy(positions, v3, dt)

    def _rk4_update_position(self, positions, velocities_1):       # :653-692
        dt = self.time_step_size
        k1 = velocities_1
        k2 = self._velocity(self._axpy(positions, k1, dt / 2))
        k3 = self._velocity(self._axpy(positions, k2, dt / 2))
        k4 = self._velocity(self._axpy(positions, k3, dt))
        out = torch.empty_like(positions)
        L.call("vpm_rk4_combine", self.n, max(self.n, 1), positions.data_ptr(), k1.data_ptr(),
               k2.data_ptr(), k3.data_ptr(), k4.data_ptr(), float(dt), out.data_ptr(),
               self._stream())
        return out

    def update_dt(self, time_step_size: float):                    # :695-696
        self.time_step_size = time_step_size

    def update_velocities(self):                                   # :699-710
        self._vel = self._velocity(self._pos)
        self._log("• Updated particles' velocities")

    def _strength_update(self, st, rhs):
        h = self._hist
        L.call("vpm_strength_update", self.n, max(self.n, 1), int(self.time_step),
               float(self.time_step_size), st.data_ptr(), h[0].data_ptr(), h[1].data_ptr(),
               h[2].data_ptr(), rhs.data_ptr(), self._stream())

    def update_strengths(self):                                    # :713-884
        """Stretching, viscous diffusion (PSE or core spreading), the LES
        filter and relaxation, in the reference's order, including its
        strength-history bookkeeping (vpmModule.py:783-787)."""
        if self.flow_model == "Potential":
            self._log("• Update_strengths() ignored because flow model is set to potential flow.")
            return
        self._log(f"• Calculating the particles' strengths ::: {self.flow_model_description}")
        pos, st, radii = self._pos, self._str.clone(), self._sigma
        n = self.n
        if self.flow_model == "LES":                               # :731-737
            vort, ens = self._vort_enstrophy(pos, st, radii)
            self._les_update(vort, radii, ens)
        if self._hist is None:                                     # :741-746
            self._hist = [torch.zeros_like(st) for _ in range(3)]
        if self.flow_model in {"LES", "DNS"}:                      # :749-791
            rhs = self._stretching(pos, st, radii)
            self._strength_update(st, rhs)
            self._str = st
        if self.viscous_scheme == "PSE":                           # :797-845
            rhs = self._pse(pos, st, radii, self._nu + self._nut)
            self._strength_update(st, rhs)
            self._str = st
        elif self.viscous_scheme == "CoreSpreading":               # :851-862
            sig = self._sigma.clone()
            L.call("vpm_core_spreading", n, sig.data_ptr(), self._nu.data_ptr(),
                   self._nut.data_ptr(), float(self.time_step_size), self._stream())
            self._sigma = sig
            self._unif = None
        if self.relax_strength_solution:                           # :867-883
            vort = P.vorticity_field_at(self._pos, self._pos, self._str, self._sigma, self.variant)
            st2 = self._str.clone()
            L.call("vpm_relax_strengths", n, max(n, 1), st2.data_ptr(), vort.data_ptr(),
                   self._stream())
            self._str = st2

    def _les_update(self, vort, radii, ens, update_constant=True):
        nb = L.load().vpm_state_workspace_bytes(self.n)
        ws = self._ws(nb)
        nut = self._nut.clone()
        L.call("vpm_les_update", self.n, max(self.n, 1), vort.data_ptr(), radii.data_ptr(),
               ens.data_ptr() if ens is not None else None, self._les.data_ptr(),
               nut.data_ptr(), 1 if update_constant else 0, ws.data_ptr(), ws.numel(),
               self._stream())
        self._nut = nut

    def _update_LES_filter_constant(self, enstrophy: float):       # :887-909
        """Host entry of the C_nu recursion (a solver step runs it on the device)."""
        e = math.sqrt(enstrophy)
        ref = self.enstrophy_ref
        ref = e if ref is None else ref
        ref = 0.1 * e + 0.9 * ref
        c = self.Cnu * e / (ref + epsilon)
        self._les.copy_(torch.tensor([max(epsilon, min(c, 0.0761)), ref, 1.0], dtype=torch.float64))

    def _update_LES_particles_viscosity(self, vorticity, radii):   # :916-962
        """Host entry of the nu_t update for given vorticity (N,3) and radii (N,)."""
        w = P.to_soa(np.asarray(vorticity, dtype=np.float64), self.device)
        r = torch.from_numpy(np.ascontiguousarray(radii, dtype=np.float64)).to(self.device)
        self._les_update(w, r, None, update_constant=False)

    # ------------------------------------------------------------ particle field control
    def remove_weak_particles(self, mode: str, threshold: float,
                              conserve_total_circulation=False):   # :970-1037
        mags = torch.linalg.vector_norm(self._str, dim=0)
        total_before = float(mags.sum().item())
        if self.n == 0:
            print("(warning) No particles available to evaluate.")
            return
        if mode == "absolute":
            weak = mags < threshold
        elif mode == "relative":
            hi = float(mags.max().item())
            if hi == 0:
                print("(warning) All particle strengths are zero.")
                weak = torch.ones_like(mags, dtype=torch.bool)
            else:
                weak = mags / hi < threshold
        else:
            print(f"(error) Mode '{mode}' not recognized. Use 'absolute' or 'relative'.")
            return
        self.remove_particles(torch.nonzero(weak).flatten().tolist())
        if conserve_total_circulation:
            after = float(torch.linalg.vector_norm(self._str, dim=0).sum().item())
            self._str = self._str * (total_before / after)

    def subdivide_strong_particles(self, top_percentage=None, strength_threshold=None):
        """Splits the strongest particles in two along their strength direction
        (vpmModule.py:1040-1094).  The reference's version calls a method that
        does not exist (:1065); this one implements the documented behaviour:
        halves of the strength at +-radius/4 along alpha/|alpha|, new particles
        with zero velocity and the default radius (Particle.py:54)."""
        if top_percentage is None and strength_threshold is None:
            raise ValueError("Provide either top_percentage or strength_threshold.")
        mags = self.get_particle_strength_magnitudes()
        nz = np.nonzero(mags)[0]
        vals = mags[nz]
        order = np.argsort(vals)[::-1]
        if strength_threshold is not None:
            keep = vals > strength_threshold
            nz, vals = nz[keep], vals[keep]
            order = np.argsort(vals)[::-1]
            idx = nz[order]
        else:
            top_n = max(int(top_percentage * len(nz) / 100), 1)
            idx = nz[order[:top_n]]
        if len(idx) == 0:
            return
        pos, st, rad = self.get_particle_positions(), self.get_particle_strengths(), self.get_particle_radii()
        half = st[idx] / 2
        dx = rad[idx, None] * st[idx] / mags[idx, None]
        newp = np.concatenate([pos[idx] + dx * 0.25, pos[idx] - dx * 0.25])
        news = np.concatenate([half, half])
        m = len(newp)
        self.add_particle_field(newp, np.zeros((m, 3)), news, np.full(m, 0.01), np.zeros(m))
        self.remove_particles(idx.tolist())
        self._log("• Subdivided strongest particles")

    def remesh_solution(self, flow2D: bool = False):
        """The reference builds a lattice and stops (vpmModule.py:1097-1121);
        remeshing is not implemented there either."""
        raise NotImplementedError("remesh_solution is not implemented in the reference solver")
