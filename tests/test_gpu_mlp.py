"""GPU parity of the BF16 MFMA MLP kernels (forward, backward, Adam) against
the oracle, through the C ABI.

Tolerances: the oracle's bf16 mode rounds at the same points as the kernels
(weights, encoded inputs, hidden activations, back-propagated pre-activation
gradients) and sums in FP64; the kernels sum in FP32 inside the MFMA, so a
value near a BF16 rounding boundary can round the other way (one BF16 ulp,
2^-8 relative).  Forward: |dy| <= 2e-2 max, 2e-3 mean (log radiance).
Gradient: relative L2 error per layer <= 2e-2.
"""
import numpy as np
import pytest

import althea_amd as aa
from althea_amd import mlp as M
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _batch(seed, n):
    dr = aa.synth.dir_rough(seed=seed, n=n)
    rng = np.random.default_rng(seed)
    t = np.concatenate([rng.lognormal(0.5, 1.0, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    return dr, t


SLICES = {"W1": (0, 1792), "b1": (1792, 1856), "W2": (1856, 5952), "b2": (5952, 6016), "W3": (6016, 10112),
          "b3": (10112, 10176), "W4": (10176, 10368), "b4": (10368, 10371)}


def test_mfma_lane_layout(torch):
    rng = np.random.default_rng(0)
    a = rng.integers(-4, 5, size=(16, 32)).astype(np.float32)
    b = rng.integers(-4, 5, size=(32, 16)).astype(np.float32)
    np.testing.assert_array_equal(M.mfma_probe(a, b), a @ b)


def test_forward_matches_oracle(torch):
    net = M.PrefilterMLP(seed=1)
    p = net.get_params()
    dr, _ = _batch(2, 4099)  # not a multiple of 16
    out = net.infer(torch.from_numpy(dr).cuda()).cpu().numpy()
    y = np.log(out[:, :3].astype(np.float64))
    ref = ol.mlp_forward(p, dr, bf16=True)
    err = np.abs(y - ref)
    assert err.max() < 2e-2 and err.mean() < 2e-3, (err.max(), err.mean())
    assert np.all(out[:, 3] == 1.0)


def test_gradients_match_oracle(torch):
    net = M.PrefilterMLP(seed=3)
    p = net.get_params()
    dr, t = _batch(4, 1000)
    g, loss = net.gradients(torch.from_numpy(dr).cuda(), torch.from_numpy(t).cuda())
    g = g.cpu().numpy().astype(np.float64)
    ref_loss, ref = ol.mlp_loss_grad(p, dr, t, bf16=True)
    assert abs(loss.item() - ref_loss) <= 1e-3 * ref_loss
    for name, (lo, hi) in SLICES.items():
        e = np.linalg.norm(g[lo:hi] - ref[lo:hi]) / max(np.linalg.norm(ref[lo:hi]), 1e-30)
        assert e < 2e-2, (name, e)
    # The BF16 kernels against exact FP64 arithmetic (reported scale of the
    # precision loss, not a parity claim).
    _, exact = ol.mlp_loss_grad(p, dr, t, bf16=False)
    e_exact = np.linalg.norm(g - exact) / np.linalg.norm(exact)
    assert e_exact < 0.1, e_exact


def test_gradient_is_batch_size_independent_in_scale(torch):
    net = M.PrefilterMLP(seed=5)
    dr, t = _batch(6, 2048)
    tdr, tt = torch.from_numpy(dr).cuda(), torch.from_numpy(t).cuda()
    g_full, _ = net.gradients(tdr, tt)
    g_a, _ = net.gradients(tdr[:1024].contiguous(), tt[:1024].contiguous())
    g_b, _ = net.gradients(tdr[1024:].contiguous(), tt[1024:].contiguous())
    torch.testing.assert_close(g_full, 0.5 * (g_a + g_b), rtol=2e-2, atol=1e-6)
    # A smaller batch after a larger one must not see stale activations.
    g_small, _ = net.gradients(tdr[:37].contiguous(), tt[:37].contiguous())
    p = net.get_params()
    _, ref = ol.mlp_loss_grad(p, dr[:37], t[:37], bf16=True)
    e = np.linalg.norm(g_small.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert e < 2e-2, e


def test_adam_step_matches_closed_form(torch):
    net = M.PrefilterMLP(seed=7, lr=1e-3)
    p0 = net.get_params().astype(np.float64)
    dr, t = _batch(8, 512)
    tdr, tt = torch.from_numpy(dr).cuda(), torch.from_numpy(t).cuda()
    g, _ = net.gradients(tdr, tt)
    g = g.cpu().numpy().astype(np.float64)
    net.train_step(tdr, tt)
    p1 = net.get_params().astype(np.float64)
    ref, _, _ = ol.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, 1e-3)
    np.testing.assert_allclose(p1, ref, rtol=0, atol=2e-7 + 1e-6 * 1e-3)


def test_training_reduces_loss(torch):
    env = aa.synth.env(seed=21, width=256, height=128, sigma=0.3, suns=2, peak_min=20.0, peak_max=20.0,
                       sun_sigma_deg=10.0)
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    dr = torch.from_numpy(aa.synth.dir_rough(seed=22, n=1 << 15, rmin=0.3, rmax=1.0)).cuda()
    labels = aa.prefilter_points(pyr, 256, 128, 256, dr)
    net = M.PrefilterMLP(seed=23, lr=3e-3, init_output_bias=float(np.log(labels[:, :3].mean().item())))
    loss = torch.zeros(1, device="cuda")
    losses = []
    for it in range(400):
        net.train_step(dr, labels, loss_out=loss)
        if it % 50 == 0 or it == 399:
            losses.append(loss.item())
    assert losses[-1] < losses[0] / 10, losses


def test_handles_alive_at_interpreter_exit_are_released_quietly(torch):
    """Handles still referenced by module globals when the interpreter exits
    are destroyed through the function bound at creation, with no exception
    reported from __del__."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import torch, althea_amd as aa\n"
            "from althea_amd import mlp as M\n"
            "ibl = aa.IBLResources(aa.IBLDesc(prefiltered_size=16, prefiltered_mips=3, irradiance_size=4,"
            " brdf_lut_size=8))\n"
            "net = M.PrefilterMLP(seed=3)\n"
            "torch.cuda.synchronize()\n"
            "print('handles made', flush=True)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "handles made" in r.stdout
    assert "Exception ignored" not in r.stderr and "TypeError" not in r.stderr, r.stderr[-2000:]
