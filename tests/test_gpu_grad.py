"""Gradients of the HIP sosfilt path (adjoint kernels) on the GPU.

Checked against (a) torch.autograd.gradcheck (FP64 central differences,
eps 1e-6, atol = rtol = 1e-4, the reference's settings in
test__sosfilt.py:264-375), (b) the reference's own backward: autograd of the
op-for-op PyTorch restatement of _lfilter.py / _sosfilt.py run on the CPU,
and (c) FP64 central-difference directional derivatives at larger sizes.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _F():
    from torchscience_amd.signal_processing import filter as F
    return F


def _ref():
    from oracle import torch_ref
    return torch_ref


def test_gradcheck_signal_and_coeffs():
    F, dev = _F(), _dev()
    import scipy.signal as ss
    sos_np = ss.butter(2, 0.3, output="sos")
    x = torch.randn(30, dtype=torch.float64, device=dev, requires_grad=True)
    sos = torch.tensor(sos_np, dtype=torch.float64, device=dev)
    torch.autograd.gradcheck(lambda x_: F.sosfilt(sos, x_), (x,), eps=1e-6, atol=1e-4, rtol=1e-4)
    n = sos_np.shape[0]

    def masked(part, x_):
        full = torch.zeros(n, 6, dtype=torch.float64, device=dev)
        full[:, 0], full[:, 1], full[:, 2] = part[:n], part[n:2 * n], part[2 * n:3 * n]
        full[:, 3] = 1.0
        full[:, 4], full[:, 5] = part[3 * n:4 * n], part[4 * n:]
        return F.sosfilt(full, x_)

    part = torch.cat([sos[:, 0], sos[:, 1], sos[:, 2], sos[:, 4], sos[:, 5]]).clone() \
        .requires_grad_(True)
    torch.autograd.gradcheck(masked, (part, x), eps=1e-6, atol=1e-4, rtol=1e-4)


def test_gradcheck_with_zi_and_zf():
    F, dev = _F(), _dev()
    import scipy.signal as ss
    sos = torch.tensor(ss.butter(4, 0.3, output="sos"), dtype=torch.float64, device=dev)
    x = torch.randn(3, 40, dtype=torch.float64, device=dev, requires_grad=True)
    zi = torch.randn(3, 2, 2, dtype=torch.float64, device=dev, requires_grad=True)

    def f(x_, zi_):
        y, zf = F.sosfilt(sos, x_, zi=zi_)
        return y, zf

    torch.autograd.gradcheck(f, (x, zi), eps=1e-6, atol=1e-4, rtol=1e-4)


@pytest.mark.parametrize("a0", [1.0, 2.0])
@pytest.mark.parametrize("T", [200, 2500])
def test_matches_reference_backward(a0, T):
    """Gradients equal the reference's autograd (restated op for op, CPU)."""
    F, dev, R = _F(), _dev(), _ref()
    import scipy.signal as ss
    torch.manual_seed(0)
    sos0 = torch.tensor(ss.cheby1(4, 0.5, 0.2, output="sos"), dtype=torch.float64)
    sos0[:, :3] *= a0
    sos0[:, 3:] *= a0
    B = 3
    x0 = torch.randn(B, T, dtype=torch.float64)
    zi0 = torch.randn(B, sos0.shape[0], 2, dtype=torch.float64)
    w = torch.randn(B, T, dtype=torch.float64)
    wz = torch.randn(B, sos0.shape[0], 2, dtype=torch.float64)

    def loss(fn, sos, x, zi):
        y, zf = fn(sos, x, zi)
        return (y * w.to(y.device)).sum() + (zf * wz.to(zf.device)).sum()

    s_c = sos0.clone().requires_grad_(True)
    x_c = x0.clone().requires_grad_(True)
    z_c = zi0.clone().requires_grad_(True)
    loss(R.sosfilt_loop, s_c, x_c, z_c).backward()

    s_g = sos0.to(dev).requires_grad_(True)
    x_g = x0.to(dev).requires_grad_(True)
    z_g = zi0.to(dev).requires_grad_(True)
    loss(lambda s, x, z: F.sosfilt(s, x, zi=z), s_g, x_g, z_g).backward()

    np.testing.assert_allclose(x_g.grad.cpu().numpy(), x_c.grad.numpy(), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(z_g.grad.cpu().numpy(), z_c.grad.numpy(), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(s_g.grad.cpu().numpy(), s_c.grad.numpy(), rtol=1e-8, atol=1e-8)
    if a0 == 1.0:
        assert torch.all(s_g.grad[:, 3] == 0)  # reference: untaken a0 branch


def test_zi_broadcast_and_batched_grads():
    F, dev = _F(), _dev()
    import scipy.signal as ss
    sos = torch.tensor(ss.butter(4, 0.25, output="sos"), dtype=torch.float64, device=dev)
    x = torch.randn(4, 3000, dtype=torch.float64, device=dev, requires_grad=True)
    zi = torch.randn(2, 2, dtype=torch.float64, device=dev, requires_grad=True)
    y, zf = F.sosfilt(sos, x, zi=zi)
    (y.square().mean() + zf.sum()).backward()
    g_b = zi.grad.clone()
    zi_full = zi.detach().expand(4, 2, 2).clone().requires_grad_(True)
    y2, zf2 = F.sosfilt(sos, x.detach(), zi=zi_full)
    (y2.square().mean() + zf2.sum()).backward()
    torch.testing.assert_close(g_b, zi_full.grad.sum(0), rtol=1e-10, atol=1e-12)
    # batched_filter_apply: gradient reaches filters and a broadcast signal
    filt = torch.tensor(np.stack([ss.butter(4, c, output="sos") for c in (0.2, 0.3)]),
                        device=dev, requires_grad=True)
    xs = torch.randn(700, dtype=torch.float64, device=dev, requires_grad=True)
    yb = F.batched_filter_apply(filt, xs)
    yb.sum().backward()
    ref_x = sum(torch.autograd.grad(F.sosfilt(filt[i].detach(), xs).sum(), xs)[0]
                for i in range(2))
    torch.testing.assert_close(xs.grad, ref_x, rtol=1e-10, atol=1e-10)
    assert not torch.isnan(filt.grad).any()


def test_fp32_directional_derivatives_config3_shape():
    """Differentiation workload shape (1024 x 16384, 4 sections): FP32 kernel
    gradients against FP64 central differences along random directions."""
    F, dev = _F(), _dev()
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    sos64 = torch.tensor(butterworth_sos(8, 0.2), dtype=torch.float64, device=dev)
    g = torch.Generator(device="cpu").manual_seed(3)
    B, T = 1024, 16384
    x = torch.randn(B, T, generator=g).to(dev)
    n = torch.arange(T, dtype=torch.float64)
    fr = torch.rand(3, generator=g, dtype=torch.float64) * 0.39 + 0.01
    tgt = (torch.sin(torch.pi * fr[:, None] * n).sum(0) + 0.1 * torch.randn(T, generator=g,
                                                                            dtype=torch.float64))
    tgt = tgt.to(dev)

    def loss64(sos, xx, zi):
        y, _ = F.sosfilt(sos, xx, zi=zi)
        return ((y - tgt) ** 2).mean()

    sos = sos64.float().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    zi = torch.zeros(B, 4, 2, device=dev, requires_grad=True)
    y, _ = F.sosfilt(sos, xg, zi=zi)
    ((y.double() - tgt) ** 2).mean().backward()
    h = 1e-6
    for d in range(4):
        gd = torch.Generator(device="cpu").manual_seed(100 + d)
        vs = torch.randn(4, 6, generator=gd, dtype=torch.float64).to(dev)
        vs[:, 3] = 0
        vx = torch.randn(B, T, generator=gd, dtype=torch.float64).to(dev)
        vz = torch.randn(B, 4, 2, generator=gd, dtype=torch.float64).to(dev)
        s64, x64 = sos.detach().double(), x.double()
        z64 = torch.zeros(B, 4, 2, dtype=torch.float64, device=dev)
        fd = (loss64(s64 + h * vs, x64 + h * vx, z64 + h * vz)
              - loss64(s64 - h * vs, x64 - h * vx, z64 - h * vz)) / (2 * h)
        an = (sos.grad.double() * vs).sum() + (xg.grad.double() * vx).sum() + \
            (zi.grad.double() * vz).sum()
        print("dir", d, float(fd), float(an), flush=True)
        assert abs(float(an) - float(fd)) <= 1e-3 * abs(float(fd)) + 1e-6


@pytest.mark.parametrize("T", [1000, 4096, 5000])
def test_fused_adjoint_matches_per_section(T):
    """At least 1024 channels run the fused adjoint (all sections in one pass,
    chained windows); fewer run one launch per section.  Per-channel filters
    with a0 != 1, zi and a zf term in the loss: the first 8 channels of a
    1024-channel problem get the same gradients (FP64) as those 8 channels
    alone, lengths covering a tail-only signal, whole windows and both."""
    F, dev = _F(), _dev()
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    rng = np.random.default_rng(T)
    B, S = 1024, 3
    filt0 = torch.tensor(batched_butterworth_sos(2 * S, rng.uniform(0.05, 0.45, B)))
    filt0[..., :] *= torch.tensor(rng.uniform(0.5, 2.0, (B, S, 1)))  # a0 != 1
    x0 = torch.tensor(rng.standard_normal((B, T)))
    zi0 = torch.tensor(rng.standard_normal((B, S, 2)))
    w = torch.tensor(rng.standard_normal((B, T)), device=dev)
    wz = torch.tensor(rng.standard_normal((B, S, 2)), device=dev)

    def run(n):
        f = filt0[:n].to(dev).requires_grad_(True)
        x = x0[:n].to(dev).requires_grad_(True)
        zi = zi0[:n].to(dev).requires_grad_(True)
        from torchscience_amd.signal_processing.filter._ops import SosfiltRows
        y, zf = SosfiltRows.apply(x, f, zi, n, 1, True, False)
        ((y * w[:n]).sum() + (zf * wz[:n]).sum()).backward()
        return x.grad, f.grad, zi.grad

    print("T", T, flush=True)
    gx, gf, gz = run(B)
    rx, rf, rz = run(8)
    torch.testing.assert_close(gx[:8], rx, rtol=1e-9, atol=1e-9)
    torch.testing.assert_close(gz[:8], rz, rtol=1e-9, atol=1e-9)
    torch.testing.assert_close(gf[:8], rf, rtol=1e-8, atol=1e-8)
    assert torch.isfinite(gx).all() and torch.isfinite(gf).all()
