"""GPU tests of the device-resident ParticleSystem against the reference-step oracle.

oracle/vpm_steps.py restates vpmModule.py's update_velocities /
update_strengths / update_state (including the strength-history bookkeeping
and the LES filter) in FP64 NumPy; the HIP system must follow it step by step.
"""
import math
import os

import numpy as np
import pytest
import torch

from oracle import vpm_oracle as O
from oracle.vpm_steps import ReferenceStepper
from openonda_amd import flow_models as FM
from openonda_amd import physics as P
from openonda_amd.Particle import Particle
from openonda_amd.vpmModule import ParticleSystem

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0)


def cloud_with_mean(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    a = rng.normal(0, 1, (n, 3)) / n + np.array([0.0, 0.0, 0.5]) / n
    s = np.full(n, 0.25)
    nu = np.full(n, 2e-3)
    return x, a, s, nu


CASES = [
    # flow_model, viscous, method, unit, relax, layout
    ("DNS", "PSE", "RK4", "GPU", False, "ring"),
    ("DNS", "CoreSpreading", "RK3", "GPU", True, "ring"),
    ("LES", "CoreSpreading", "RK2", "GPU", False, "cloud"),
    ("LES", "PSE", "RK4", "CPU", False, "cloud"),
    ("pseudo2D", "PSE", "Euler", "CPU", False, "ring"),
    ("Potential", None, "RK4", "GPU", False, "ring"),
    ("DNS", "PSE", "RK2", "CPU", False, "cloud"),
]


@pytest.mark.parametrize("flow,visc,method,unit,relax,layout", CASES)
def test_steps_follow_reference(gpu, flow, visc, method, unit, relax, layout):
    print(f"[system] {flow} {visc} {method} {unit} relax={relax} {layout}", flush=True)
    if layout == "ring":
        x, a, s, nu = FM.torus_ring(8, nu=1e-3)
    else:
        x, a, s, nu = cloud_with_mean(400, 7)
    dt = 0.02
    ref = ReferenceStepper(x, a, s, nu, dt=dt, method=method, flow_model=flow,
                           viscous_scheme=visc, variant="cuda" if unit == "GPU" else "cpu",
                           relax=relax)
    ps = ParticleSystem(time_step_size=dt, time_integration_method=method, processing_unit=unit,
                        flow_model=flow, viscous_scheme=visc, relax_strength_solution=relax,
                        write_backups=False)
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu)
    for k in range(5):
        ref.step()
        ps.update_velocities()
        ps.update_strengths()
        ps.update_state()
        tag = (flow, visc, method, unit, k)
        assert rel(ps.get_particle_positions(), ref.x) <= 1e-12, tag
        assert rel(ps.get_particle_strengths(), ref.alpha) <= 1e-11, tag
        assert rel(ps.get_particle_radii(), ref.sigma) <= 1e-14, tag
        assert rel(ps.get_particle_velocities(), ref.velocity) <= 1e-11, tag
        assert rel(ps.get_particle_viscosities_t(), ref.nu_t) <= 1e-10, tag
        if flow == "LES":
            assert ps.Cnu == pytest.approx(ref.cnu, rel=1e-12), tag
    assert ps.time_step == 5 and ps.flow_time == pytest.approx(5 * dt)


def test_fast_precisions_track_plain(gpu):
    # DNS: a ring's LES filter is decided by roundoff (its mean vorticity is ~0)
    x, a, s, nu = FM.torus_ring(16, nu=1e-3)
    runs = {}
    for prec in ("plain", "f64", "f32c"):
        ps = ParticleSystem(time_step_size=0.02, time_integration_method="RK4",
                            processing_unit="GPU", flow_model="DNS", viscous_scheme="PSE",
                            precision=prec, write_backups=False)
        ps.add_particle_field(x, np.zeros_like(x), a, s, nu)
        for _ in range(4):
            ps.update_velocities()
            ps.update_strengths()
            ps.update_state()
        runs[prec] = (ps.get_particle_positions(), ps.get_particle_strengths())
    assert rel(runs["f64"][0], runs["plain"][0]) <= 1e-12
    assert rel(runs["f64"][1], runs["plain"][1]) <= 1e-11
    assert rel(runs["f32c"][0], runs["plain"][0]) <= 1e-6
    assert rel(runs["f32c"][1], runs["plain"][1]) <= 1e-5


def test_diagnostics_match_oracle(gpu):
    x, a, s, nu = FM.torus_ring(16)
    ps = ParticleSystem(processing_unit="GPU", write_backups=False)
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu)
    assert rel(ps.get_total_linear_impulse(), O.total_linear_impulse(x, a)) <= 1e-13
    assert ps.get_total_strength_magnitude() == pytest.approx(np.sum(np.linalg.norm(a, axis=1)), rel=1e-13)
    assert ps.get_total_kinetic_energy() == pytest.approx(O.total_kinetic_energy(x, a, s), rel=1e-11)
    assert ps.get_total_enstrophy() == pytest.approx(O.total_enstrophy(x, a, s), rel=1e-11)
    y = FM.ring_probes(64)
    assert rel(ps.get_induced_velocity_at(y), O.induced_velocity_at(y, x, a, s)) <= 1e-12
    assert rel(ps.get_vorticity_field_at(y), O.vorticity_field_at(y, x, a, s)) <= 1e-12


def _tilted_offset_ring():
    """A ring whose impulses are O(1) in every component: the axis tilted off
    x, the centre moved off the origin."""
    x, a, s, nu = FM.torus_ring(16)
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0.0], [math.sin(th), math.cos(th), 0.0],
                   [0.0, 0.0, 1.0]])
    Rx = np.array([[1.0, 0.0, 0.0], [0.0, math.cos(0.4), -math.sin(0.4)],
                   [0.0, math.sin(0.4), math.cos(0.4)]])
    R = Rx @ Rz
    return x @ R.T + [0.8, -1.3, 2.1], a @ R.T, s, nu


@pytest.mark.parametrize("case", ["tilted_ring", "cloud_with_mean"])
def test_impulses_match_oracle_on_asymmetric_sets(gpu, case):
    """Linear and angular impulse (vpmModule.py:264-274) on sets where every
    component is far from zero, so a wrong sign or factor shows."""
    print(f"[system] impulses {case}", flush=True)
    if case == "tilted_ring":
        x, a, s, nu = _tilted_offset_ring()
    else:
        x, a, s, nu = cloud_with_mean(500, 11)
        x = x + [1.5, -0.5, 0.25]
    ps = ParticleSystem(processing_unit="GPU", write_backups=False)
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu)
    I_ref, A_ref = O.total_linear_impulse(x, a), O.total_angular_impulse(x, a)
    assert np.abs(A_ref).min() > 1e-3 * np.abs(A_ref).max() > 0, A_ref
    I, A = ps.get_total_linear_impulse(), ps.get_total_angular_impulse()
    assert np.allclose(I, I_ref, rtol=1e-12, atol=1e-14 * np.abs(I_ref).max())
    assert np.allclose(A, A_ref, rtol=1e-12, atol=1e-14 * np.abs(A_ref).max())
    mag = np.linalg.norm(a, axis=1).sum()
    assert np.allclose(ps.get_total_strength(), a.sum(0), rtol=1e-12, atol=1e-14 * mag)


@pytest.mark.parametrize("threshold", [2.1e-4, 1e-30])
def test_weak_particle_removal_mid_run_keeps_the_strength_history(gpu, threshold):
    """remove_weak_particles between steps 3 and 4 of a multistep strength
    update (vpmModule.py:970-1037): the scheme continues on the remaining
    particles' history.  threshold 1e-30 removes nothing (the reference then
    continues unchanged); 2.1e-4 removes the 80 weakest points of the core's
    outer ring (relative magnitudes after 3 steps run 1.9e-4 ... 2.09e-4,
    then 2.14e-4: a gap wide against any rounding difference)."""
    print(f"[system] removal mid-run threshold={threshold}", flush=True)
    x, a, s, nu = FM.torus_ring(8, nu=1e-3)
    dt = 0.02
    ref = ReferenceStepper(x, a, s, nu, dt=dt, method="RK2", flow_model="DNS",
                           viscous_scheme="PSE", variant="cuda")
    ps = ParticleSystem(time_step_size=dt, time_integration_method="RK2", processing_unit="GPU",
                        flow_model="DNS", viscous_scheme="PSE", write_backups=False)
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu)

    def both_step():
        ref.step()
        ps.update_velocities()
        ps.update_strengths()
        ps.update_state()

    for _ in range(3):
        both_step()
    mags = np.linalg.norm(ref.alpha, axis=1)
    weak = np.nonzero(mags / mags.max() < threshold)[0]
    assert (len(weak) > 0) == (threshold > 1e-20)
    ps.remove_weak_particles("relative", threshold)
    ref.remove(weak.tolist())
    assert len(ps) == len(ref.x)
    for k in range(3):
        both_step()
        tag = (threshold, k)
        assert rel(ps.get_particle_positions(), ref.x) <= 1e-12, tag
        assert rel(ps.get_particle_strengths(), ref.alpha) <= 1e-11, tag
        assert rel(ps.get_particle_velocities(), ref.velocity) <= 1e-11, tag


def test_subdivide_strong_particles(gpu):
    """subdivide_strong_particles (vpmModule.py:1040-1094, as documented): the
    strongest particles split into two halves at +-radius/4 along their
    strength direction."""
    rng = np.random.default_rng(5)
    n = 40
    x = rng.uniform(-1, 1, (n, 3))
    a = rng.normal(0, 1, (n, 3))
    s = rng.uniform(0.1, 0.3, n)
    ps = ParticleSystem(processing_unit="GPU", write_backups=False)
    ps.add_particle_field(x, np.zeros_like(x), a, s, np.zeros(n))
    mags = np.linalg.norm(a, axis=1)
    top = np.argsort(mags)[::-1][:4]  # 10 % of 40
    ps.subdivide_strong_particles(top_percentage=10)
    assert len(ps) == n + 4
    P, A, S = ps.get_particle_positions(), ps.get_particle_strengths(), ps.get_particle_radii()
    assert np.allclose(A.sum(0), a.sum(0), rtol=1e-14, atol=1e-14)
    rest = np.setdiff1d(np.arange(n), top)
    assert np.array_equal(P[:n - 4], x[rest]) and np.array_equal(A[:n - 4], a[rest])
    for k, i in enumerate(top):
        off = s[i] * a[i] / mags[i] * 0.25
        assert np.allclose(P[n - 4 + k], x[i] + off, rtol=0, atol=1e-15)
        assert np.allclose(P[n + k], x[i] - off, rtol=0, atol=1e-15)
        assert np.array_equal(A[n - 4 + k], a[i] / 2) and np.array_equal(A[n + k], a[i] / 2)
    assert np.all(S[n - 4:] == 0.01)
    # by threshold: only the particles stronger than it
    ps2 = ParticleSystem(processing_unit="GPU", write_backups=False)
    ps2.add_particle_field(x, np.zeros_like(x), a, s, np.zeros(n))
    thr = np.sort(mags)[-3] - 1e-12
    ps2.subdivide_strong_particles(strength_threshold=thr)
    assert len(ps2) == n + 3
    with pytest.raises(ValueError):
        ps2.subdivide_strong_particles()


def test_particle_management_and_backup(gpu, tmp_path):
    x, a, s, nu = FM.torus_ring(4, nu=1e-3)
    ps = ParticleSystem(processing_unit="GPU", write_backups=False,
                        backup_filename=str(tmp_path / "b"))
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu, group_id=3)
    ps.add_particle(Particle(position=[5.0, 0, 0], strength=[0, 0, 1e-9], radius=0.2))
    assert len(ps) == len(x) + 1
    p = ps[len(x)]
    assert p.radius == 0.2 and p.group_id == 0 and np.allclose(p.position, [5, 0, 0])
    ps.backup_particle_field()
    q = ParticleSystem(processing_unit="GPU", write_backups=False)
    q.load_particle_field_from_backup(str(tmp_path / "b_0000000.vtp"))
    assert np.array_equal(q.get_particle_positions(), ps.get_particle_positions())
    assert np.array_equal(q.get_particle_strengths(), ps.get_particle_strengths())
    assert np.array_equal(q.get_particle_group_id(), ps.get_particle_group_id())
    total = ps.get_total_strength_magnitude()
    # the Gaussian core's outer points are ~e^-9 of the peak; the added one 1e-9
    ps.remove_weak_particles("relative", 1e-7, conserve_total_circulation=True)
    assert len(ps) == len(x)
    assert ps.get_total_strength_magnitude() == pytest.approx(total, rel=1e-12)
    ps.remove_particles([0, 1, 1, 10 ** 6])
    assert len(ps) == len(x) - 2
    ps.remove_all_particles()
    assert len(ps) == 0


def test_constructor_surface(gpu):
    # the reference's only VPM test: tests/test1_import_modules.py:23-58
    ps = ParticleSystem(flow_model="LES", time_step_size=0.01, time_integration_method="RK2",
                        viscous_scheme="CoreSpreading", processing_unit="GPU",
                        monitor_variables=["Circulation", "Kinetic energy"],
                        backup_filename="particle_system_backup.dat", backup_frequency=10)
    assert ps.flow_model == "LES" and ps.time_step_size == 0.01
    with pytest.raises(ValueError):
        ParticleSystem(flow_model="Potential", viscous_scheme="PSE")
    with pytest.raises(ValueError):
        ParticleSystem(time_integration_method="RK5")
    with pytest.raises(TypeError):
        ParticleSystem(time_step_size="1")


@pytest.mark.parametrize("flow,visc", [("LES", "PSE"), ("DNS", "CoreSpreading")])
def test_checkpoint_resume_is_bit_identical(gpu, tmp_path, flow, visc):
    """Stepping on from a checkpoint reproduces the uninterrupted run exactly."""
    x, a, s, nu = cloud_with_mean(300, 9)

    def system():
        return ParticleSystem(time_step_size=0.02, time_integration_method="RK3",
                              processing_unit="GPU", flow_model=flow, viscous_scheme=visc,
                              write_backups=False)

    def step(ps, k):
        for _ in range(k):
            ps.update_velocities()
            ps.update_strengths()
            ps.update_state()

    ps = system()
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu, group_id=2)
    step(ps, 3)
    ps.save_checkpoint(str(tmp_path / "c.npz"))
    step(ps, 2)
    q = system()
    q.load_checkpoint(str(tmp_path / "c.npz"))
    assert q.time_step == 3 and q.flow_time == pytest.approx(0.06)
    step(q, 2)
    for get in ("get_particle_positions", "get_particle_strengths", "get_particle_radii",
                "get_particle_velocities", "get_particle_viscosities_t", "get_particle_group_id"):
        assert np.array_equal(getattr(q, get)(), getattr(ps, get)()), get
    assert q.Cnu == ps.Cnu and q.time_step == ps.time_step


def test_device_scan_flags_non_finite_state(gpu):
    """vpm_scan_particles through the C ABI: a clean set gives 0; a NaN
    position, an Inf strength, a zero radius and a negative viscosity each
    set their VPM_BAD_* bit; ParticleSystem(check_finite=True) raises
    NonFiniteError after the step that produced a NaN."""
    from openonda_amd import _lib as L

    x, a, s, nu = FM.torus_ring(8, nu=1e-3)
    pos, al = P.to_soa(x), P.to_soa(a)
    sig, vis = torch.from_numpy(s).cuda(), torch.from_numpy(nu).cuda()
    assert P.scan_particles(pos, al, sig, vis) == 0
    cases = [(pos, 1, 7, float("nan"), L.BAD_POSITION), (al, 2, 100, float("inf"), L.BAD_STRENGTH),
             (sig, None, 300, 0.0, L.BAD_RADIUS), (vis, None, 511, -1e-6, L.BAD_VISCOSITY)]
    for t, c, i, v, bit in cases:
        t2 = t.clone()
        if c is None:
            t2[i] = v
        else:
            t2[c, i] = v
        args = [pos, al, sig, vis]
        args[next(q for q, u in enumerate(args) if u is t)] = t2
        assert P.scan_particles(*args) == bit
        with pytest.raises(L.NonFiniteError) as e:
            P.check_particles(*args)
        assert e.value.code == L.VPM_ERR_NONFINITE and e.value.bits == bit
    ps = ParticleSystem(time_step_size=0.01, time_integration_method="RK2", processing_unit="GPU",
                        flow_model="DNS", viscous_scheme="CoreSpreading", write_backups=False,
                        check_finite=True)
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu)
    ps.update_velocities()
    ps.update_strengths()
    ps.update_state()                       # clean: no error
    ps._str[1, 5] = float("nan")            # a NaN strength reaches the next step
    ps.update_velocities()
    ps.update_strengths()
    with pytest.raises(L.NonFiniteError, match="time-step 2"):
        ps.update_state()
