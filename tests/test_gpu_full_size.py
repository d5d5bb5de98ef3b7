"""The fused FP32 pass at the benchmark's full size (random cloud, N = 2^20).

The oracle cannot evaluate 1.1e12 pairs in a test, so the full size is held
to properties that do not depend on N, plus a strided sample against the
C/OpenMP FP64 port of the CPU plugin:
* velocity is linear and stretching quadratic in the strengths; scaling the
  strengths by 2 or -1 is exact in binary floating point, so the outputs
  must follow bit for bit;
* 256 strided targets against all 2^20 sources agree with the FP64 port to
  the rel-L2 bar of the FP32 pass.
"""
import numpy as np
import pytest
import torch

from oracle import cpu_baseline as CB
from openonda_amd import flow_models as FM
from openonda_amd import physics as P

pytestmark = pytest.mark.gpu

N = 1 << 20


@pytest.fixture(scope="module")
def cloud(gpu):
    x, a, s, _ = FM.random_cloud(N)
    return x, a, s, P.to_soa(x), P.to_soa(a), torch.from_numpy(s).cuda()


def run(pos, al, sig):
    r = P.forward(pos, al, sig, outputs=("vel", "stretch"), precision="f32c", sigma_mode="uniform")
    torch.cuda.synchronize()
    return r


def test_scaling_of_the_strengths_is_exact(cloud):
    _, _, _, pos, al, sig = cloud
    print("[full-size] N=2^20 scaling", flush=True)
    base = run(pos, al, sig)
    neg = run(pos, -al, sig)
    dbl = run(pos, 2.0 * al, sig)
    assert torch.equal(neg["vel"], -base["vel"])
    assert torch.equal(neg["stretch"], base["stretch"])
    assert torch.equal(dbl["vel"], 2.0 * base["vel"])
    assert torch.equal(dbl["stretch"], 4.0 * base["stretch"])


def test_strided_sample_against_fp64_port(cloud):
    x, a, s, pos, al, sig = cloud
    print("[full-size] N=2^20 sample vs FP64 port", flush=True)
    res = run(pos, al, sig)
    tidx = np.arange(0, N, N // 256, dtype=np.int64)
    u_ref = CB.selfinduced_velocity(x, a, s, tidx)
    s_ref = CB.strength_gradients(x, a, s, tidx)
    t = torch.from_numpy(tidx).cuda()
    u = res["vel"][:, t].t().cpu().numpy()
    st = res["stretch"][:, t].t().cpu().numpy()
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) <= 1e-6
    assert np.linalg.norm(st - s_ref) / np.linalg.norm(s_ref) <= 1e-6
