"""CPU tests of the host-side logic (no kernel launches)."""
import base64
import math
import struct
import zlib

import numpy as np
import pytest
import torch

from openonda_amd import batch as B
from openonda_amd import flow_models as FM
from openonda_amd import physics as P
from openonda_amd import vtp


def test_work_list_orders_largest_set_first():
    sizes = [10, 1000, 0, 513, 512]
    item_set, item_block, first, count = B.work_list(sizes, 256)
    assert list(count) == [1, 4, 0, 3, 2]
    assert len(item_set) == sum(count)
    # sets in decreasing size, blocks of one set consecutive
    assert list(item_set[:4]) == [1] * 4 and list(item_block[:4]) == [0, 1, 2, 3]
    assert list(item_set[4:7]) == [3] * 3 and list(item_set[7:9]) == [4] * 2
    for s in range(len(sizes)):
        blk = item_block[first[s]:first[s] + count[s]]
        assert list(blk) == list(range(count[s]))
        assert np.all(item_set[first[s]:first[s] + count[s]] == s)


@pytest.mark.parametrize("binary", [True, False])
def test_vtp_round_trip(tmp_path, binary):
    rng = np.random.default_rng(0)
    pts = rng.normal(size=(17, 3))
    data = {"Velocity": rng.normal(size=(17, 3)), "Radius": rng.uniform(size=17),
            "Group_ID": np.arange(17, dtype=np.int64)}
    f = tmp_path / "p.vtp"
    vtp.write_vtp(f, pts, data, binary=binary)
    p2, d2 = vtp.read_vtp(f)
    assert np.array_equal(p2, pts)
    for k in data:
        assert np.array_equal(d2[k], data[k])


def test_vtp_reads_zlib_compressed_uint32_headers(tmp_path):
    # the layout VTK writes with vtkZLibDataCompressor and UInt32 headers
    pts = np.arange(12, dtype=np.float32).reshape(4, 3)
    raw = pts.tobytes()
    comp = zlib.compress(raw)
    head = struct.pack("<IIII", 1, len(raw), len(raw), len(comp))
    b64 = base64.b64encode(head).decode() + base64.b64encode(comp).decode()
    # VTK encodes header and data separately when compressed
    xml = f"""<?xml version="1.0"?>
<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt32" compressor="vtkZLibDataCompressor">
<PolyData><Piece NumberOfPoints="4"><PointData></PointData><Points>
<DataArray type="Float32" NumberOfComponents="3" format="binary">{base64.b64encode(head + comp).decode()}</DataArray>
</Points></Piece></PolyData></VTKFile>"""
    f = tmp_path / "c.vtp"
    f.write_text(xml)
    p, d = vtp.read_vtp(f)
    assert np.array_equal(p, pts.astype(np.float64)) and d == {}
    assert b64  # (documented alternative encoding, not needed by the reader)


def test_torus_ring_layout():
    x, a, s, nu = FM.torus_ring(64)
    assert x.shape == (4096, 3) and np.all(s == pytest.approx(0.13))
    # circulation-weighted: sum |alpha| approximates 2 pi R Gamma
    assert np.sum(np.linalg.norm(a, axis=1)) == pytest.approx(2 * math.pi, rel=0.07)
    x, a, s, nu = FM.torus_ring(64, renormalise=True)
    assert np.sum(np.linalg.norm(a, axis=1)) == pytest.approx(2 * math.pi, rel=1e-12)
    # vorticity is azimuthal: alpha . e_r = 0 and alpha_x = 0
    th = np.arctan2(x[:, 2], x[:, 1])
    assert np.allclose(a[:, 0], 0) and np.allclose(a[:, 1] * np.cos(th) + a[:, 2] * np.sin(th), 0)


def test_reference_ring_generator_renormalises():
    pts, vol = FM.rectangular_point_distribution([-0.3, 0.3, -1.3, 1.3, -1.3, 1.3], 0.05)
    v, s, r, nu = FM.vortex_ring_vpm(pts, vol, 0.1, 1e-3, np.zeros(3), 1.0, 1.0, 0.1, 0.02, 0.1)
    # vpm_flow_models.py:150-152: sum |alpha| = 2 pi R Gamma exactly
    assert np.sum(np.linalg.norm(s, axis=1)) == pytest.approx(2 * math.pi, rel=1e-12)
    assert np.all(s[:, 0] == 0) and np.all(v == 0) and np.all(r == 0.1) and np.all(nu == 1e-3)
    # odd lattice counts (vpm_solver_helper.py:40-48)
    n = round(len(pts) ** (1 / 3))
    assert len(pts) % 2 == 1


def test_cloud_and_batch_layouts():
    x, a, s, nu = FM.random_cloud(1 << 12)
    assert x.min() >= 0 and x.max() <= 2 * math.pi
    assert a.std() == pytest.approx(math.sqrt(1 / 4096), rel=0.05)
    assert s[0] == pytest.approx(1.3 * 2 * math.pi / 16)
    inst = FM.batch_instances()
    assert len(inst) == 96
    sizes = [len(i[0]) for i in inst]
    assert min(sizes) >= 2000 and max(sizes) <= 131072 + 64
    for i in inst:
        assert np.all(i[2] == i[2][0])  # uniform radius per set
    x, a, s, nu = FM.leapfrog_rings()
    assert x.shape == (65536, 3)
    assert FM.ring_probes().shape == (4096, 3)


def test_tensor_validation_without_gpu():
    cpu = torch.zeros((3, 4), dtype=torch.float64)
    with pytest.raises(ValueError, match="HIP device"):
        P._vec(cpu, "positions")
    with pytest.raises(TypeError):
        P._vec(cpu.float(), "positions")
    with pytest.raises(ValueError):
        P.to_soa(np.zeros((4, 2)))


def test_bench_refuses_several_gpus(capsys):
    import bench

    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device error path")
def test_no_silent_cpu_fallback():
    from openonda_amd.vpmModule import ParticleSystem
    from openonda_amd import ParticlesPhysicsHip as H

    with pytest.raises(RuntimeError, match="no CPU fallback"):
        ParticleSystem(processing_unit="GPU")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        H.get_selfinduced_velocity(np.zeros((2, 3)), np.zeros((2, 3)), np.ones(2))


def test_drop_in_precision_switch_checks():
    """The drop-in module's fused-pass switch rejects unknown values and the
    CPU plugin's arithmetic (the fused pass computes the GPU plugin's)."""
    from openonda_amd import ParticlesPhysicsHip as H

    try:
        H.PRECISION = "plain"
        assert H._fused() is None
        H.PRECISION = "f32c"
        assert H._fused() == "f32c"
        H.PRECISION = "f16"
        with pytest.raises(ValueError, match="PRECISION"):
            H._fused()
        H.PRECISION, H.VARIANT = "f64", "cpu"
        with pytest.raises(ValueError, match="VARIANT"):
            H._fused()
    finally:
        H.PRECISION, H.VARIANT = "plain", "cuda"


@pytest.mark.parametrize("which, bit", [("x", 1), ("a", 2), ("s", 4), ("nu", 8), ("y", 1)])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_drop_in_rejects_non_finite_inputs_before_any_gpu_call(which, bit, bad):
    """NaN/Inf inputs (and radii <= 0, viscosities < 0) raise NonFiniteError
    with status VPM_ERR_NONFINITE and the VPM_BAD_* bit of the array, on the
    host, before the library or a device is touched."""
    from openonda_amd import ParticlesPhysicsHip as H
    from openonda_amd import _lib as L

    x, a, s, nu = np.zeros((4, 3)), np.ones((4, 3)), np.full(4, 0.1), np.full(4, 1e-3)
    y = np.zeros((2, 3))
    arrs = dict(x=x, a=a, s=s, nu=nu, y=y)
    arrs[which] = arrs[which].copy()
    arrs[which].flat[1] = bad
    if which == "y":
        call = lambda: H.get_induced_velocity_at(arrs["y"], arrs["x"], arrs["a"], arrs["s"])
    elif which == "nu":
        call = lambda: H.get_strength_gradients_PSE(arrs["x"], arrs["a"], arrs["s"], arrs["nu"])
    else:
        call = lambda: H.get_selfinduced_velocity(arrs["x"], arrs["a"], arrs["s"])
    with pytest.raises(L.NonFiniteError) as e:
        call()
    assert e.value.code == L.VPM_ERR_NONFINITE == -6
    assert e.value.bits == bit


def test_drop_in_rejects_non_positive_radii_and_negative_viscosity():
    from openonda_amd import ParticlesPhysicsHip as H
    from openonda_amd import _lib as L

    x, a = np.zeros((3, 3)), np.ones((3, 3))
    with pytest.raises(L.NonFiniteError, match="radii"):
        H.get_strength_gradients(x, a, np.array([0.1, 0.0, 0.1]))
    with pytest.raises(L.NonFiniteError, match="viscosities"):
        H.get_strength_gradients_PSE(x, a, np.full(3, 0.1), np.array([1e-3, -1e-9, 0.0]))
    assert issubclass(L.NonFiniteError, ValueError)
