"""Learned flux correction: the differentiable torch step, the BF16-MFMA network kernels, and the
corrected coarse solver on the HIP stage kernels.

CPU: the torch restatement of the step equals the oracle bit for bit; corrections keep the update
conservative; the float64 network oracle's gradients pass finite differences.
GPU: MFMA operand layout; network forward and all gradients against the float64 oracle (BF16
tolerances); directional finite differences of the kernels' own gradients; periodic translation
equivariance; corrected HIP stage kernels equal the torch step with the same increments; a zero head
reproduces the plain solver bit for bit.
"""
import numpy as np
import pytest
import torch

import oracle2d


def riemann_batch(n, idx):
    from npde.families import riemann
    return np.stack([riemann(k, n)[0] for k in idx])


# ------------------------------------------------------------------------------------------- CPU
@pytest.mark.parametrize("scheme,lim", [("hllc", "vanleer"), ("hllc", "minmod"), ("rusanov", "none"),
                                        ("hllc", "unlimited")])
def test_torch_step_equals_oracle(oracle, scheme, lim):
    from npde.learned import torch_solver as ts
    n = 40
    U0 = riemann_batch(n, [0, 3])
    g = oracle2d.Grid.make(n, n, bc_x="periodic", bc_y="periodic")
    dt = 0.0015
    out = ts.step(torch.from_numpy(U0), torch.full((2,), dt, dtype=torch.float64), 1 / n, 1 / n, scheme, lim)
    for k in range(2):
        a = U0[k].copy()
        oracle.step(g, scheme, lim, a, dt)
        assert np.array_equal(out[k].numpy(), a)


def test_torch_advance_lands_like_oracle(oracle):
    from npde.learned import torch_solver as ts
    n = 32
    U0 = riemann_batch(n, [5])
    g = oracle2d.Grid.make(n, n, bc_x="periodic", bc_y="periodic")
    a = U0[0].copy()
    r = oracle.run(g, "hllc", "vanleer", 0.5, a, 0.04)
    U, t, _ = ts.advance(torch.from_numpy(U0), torch.zeros(1, dtype=torch.float64),
                         torch.full((1,), 0.04, dtype=torch.float64), 1 / n, 1 / n)
    assert float(t[0]) == r["t"]
    assert np.array_equal(U[0].numpy(), a)


def test_corrections_are_conservative_and_differentiable():
    from npde.learned import torch_solver as ts
    n = 24
    U0 = torch.from_numpy(riemann_batch(n, [1]))
    cx = (0.01 * torch.randn(1, 4, n, n, dtype=torch.float64)).requires_grad_()
    cy = (0.01 * torch.randn(1, 4, n, n, dtype=torch.float64)).requires_grad_()
    U1 = ts.step(U0, torch.full((1,), 0.002, dtype=torch.float64), 1 / n, 1 / n, cx_corr=cx, cy_corr=cy)
    tot0, tot1 = U0.sum(dim=(-1, -2)), U1.sum(dim=(-1, -2))
    assert torch.allclose(tot0, tot1, rtol=0, atol=1e-12 * float(tot0.abs().max()))
    (U1[:, 0] ** 2).sum().backward()
    assert torch.isfinite(cx.grad).all() and cx.grad.abs().sum() > 0 and cy.grad.abs().sum() > 0


def test_network_oracle_gradients_pass_finite_differences():
    import nn_oracle
    g = torch.Generator().manual_seed(0)
    x = torch.randn(1, 6, 8, 3, generator=g, dtype=torch.float64)
    ws = [torch.randn(4, 3, 3, 3, generator=g, dtype=torch.float64, requires_grad=True),
          torch.randn(2, 4, 3, 3, generator=g, dtype=torch.float64, requires_grad=True)]
    bs = [torch.randn(4, generator=g, dtype=torch.float64, requires_grad=True),
          torch.randn(2, generator=g, dtype=torch.float64, requires_grad=True)]
    f = lambda w0, b0, w1, b1: nn_oracle.forward(x, [w0, w1], [b0, b1], round_bf16=False)
    assert torch.autograd.gradcheck(f, (ws[0], bs[0], ws[1], bs[1]), eps=1e-6, atol=1e-6)


def test_corrector_shape_and_identity_at_init():
    from npde.learned.model import FluxCorrector
    m = FluxCorrector()
    assert m.num_parameters() == 32392
    assert all(float(w.abs().max()) == 0.0 for w in [m.weights[-1]])


def test_dihedral_maps_commute_with_the_step():
    """The training augmentation's 8 symmetries of the periodic square are symmetries of the scheme: one
    SSP-RK2 step of a mapped state equals the mapped step to rounding (a wrong momentum sign or exchange
    fails by O(1)), and the maps form the group (each has an inverse in it)."""
    from npde.learned import torch_solver as ts
    from npde.learned.pipeline import dihedral
    n = 24
    U0 = torch.from_numpy(riemann_batch(n, [7])).expand(8, 4, n, n).contiguous()
    g = torch.arange(8)
    dt = torch.full((8,), 0.002, dtype=torch.float64)
    a = ts.step(dihedral(U0, g), dt, 1 / n, 1 / n)
    b = dihedral(ts.step(U0, dt, 1 / n, 1 / n), g)
    assert float((a - b).abs().max()) < 1e-12 * float(U0.abs().max())
    assert float((a - ts.step(U0, dt, 1 / n, 1 / n)).abs().max()) > 1e-3  # the maps do change the state
    inv = torch.tensor([0, 1, 2, 3, 4, 6, 5, 7])  # transpose after a single mirror: inverse mirrors the other axis
    assert torch.equal(dihedral(dihedral(U0, g), inv), U0)


def test_network_call_times_contain_the_frame_times_bitwise():
    """With S network calls per frame interval the solver lands on frame_times(50 S); every S-th of those is
    the frame time itself, bit for bit (0.5 S k and 0.5 k are exact, and both quotients are the correctly
    rounded value of the same real number), so the learned frames are taken at the truth's times."""
    from npde.learned.pipeline import frame_times
    base = frame_times(50)
    for S in (1, 2, 3, 4):
        assert frame_times(50 * S)[::S] == base


def test_fused_packing_places_every_weight():
    """pack_fused: fragment (layer, k-block b, n-tile t), lane l = 16 g + m, value j holds W[16 t + m][ci][tap]
    with (tap, ci) = (b, 8 g + j) for the 32-channel layers and (4 b + g, j) for the first layer; zero
    where the head has no channel or the first layer no tap; biases in layer order."""
    from npde.learned.model import FluxCorrector
    m = FluxCorrector(seed=3)
    g = torch.Generator().manual_seed(3)
    with torch.no_grad():
        for w in m.weights:
            w.copy_(torch.randn(w.shape, generator=g))
        for b in m.biases:
            b.copy_(torch.randn(b.shape, generator=g))
    bits, biases, scale = m.pack_fused()
    fr = bits.view(torch.bfloat16).float().reshape(69, 64, 8)
    ws = [w.detach().to(torch.bfloat16).float() for w in m.weights]
    starts, kbs, nts = [0, 6, 24, 42, 60], [3, 9, 9, 9, 9], [2, 2, 2, 2, 1]
    for L in range(5):
        W = ws[L]
        for b in range(kbs[L]):
            for t in range(nts[L]):
                for lane in range(64):
                    mm, gg = lane % 16, lane // 16
                    co = 16 * t + mm
                    for j in range(8):
                        tap, ci = (4 * b + gg, j) if L == 0 else (b, 8 * gg + j)
                        want = 0.0 if (tap >= 9 or co >= W.shape[0]) else float(W[co, ci, tap // 3, tap % 3])
                        assert float(fr[starts[L] + b * nts[L] + t, lane, j]) == want, (L, b, t, lane, j)
    assert torch.equal(biases, torch.cat([b.detach() for b in m.biases]))
    assert scale == m.out_scale


# ------------------------------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def cuda(npde_mod):
    return torch.device("cuda:0")


@pytest.mark.gpu
def test_mfma_operand_layout(cuda):
    import ctypes as C
    import npde._lib as L
    lib = L.load()
    a = torch.arange(256, dtype=torch.float32, device=cuda).reshape(16, 16) % 7 - 3
    b = torch.arange(256, dtype=torch.float32, device=cuda).reshape(16, 16) % 5 - 2
    d = torch.empty(16, 16, dtype=torch.float32, device=cuda)
    rc = lib.npde_mfma_probe(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(d.data_ptr()),
                             C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(d, a @ b)  # small integers: exact in bf16 and in fp32 accumulation


@pytest.mark.gpu
def test_mfma32_operand_layout(cuda):
    """The 16x16x32 BF16 MFMA operand layout the fused network kernel assumes (small integers: exact)."""
    import ctypes as C
    import npde._lib as L
    lib = L.load()
    a = torch.arange(512, dtype=torch.float32, device=cuda).reshape(16, 32) % 7 - 3
    b = torch.arange(512, dtype=torch.float32, device=cuda).reshape(32, 16) % 5 - 2
    d = torch.empty(16, 16, dtype=torch.float32, device=cuda)
    rc = lib.npde_mfma32_probe(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(d.data_ptr()),
                               C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(d, a @ b)


def _random_net(seed, cuda, head_scale=0.3):
    from npde.learned.model import FluxCorrector
    m = FluxCorrector(seed=seed).to(cuda)
    torch.manual_seed(seed)
    with torch.no_grad():
        m.weights[-1].copy_(torch.randn_like(m.weights[-1]) * head_scale)
        for b in m.biases:
            b.copy_(torch.randn_like(b) * 0.1)
    return m


@pytest.mark.gpu
def test_network_forward_and_gradients_match_oracle(cuda):
    import nn_oracle
    from npde.learned.conv import FluxCNN
    m = _random_net(1, cuda)
    x = torch.randn(2, 16, 32, 16, device=cuda, generator=torch.Generator(device=cuda).manual_seed(1))
    x[..., 8:] = 0
    x.requires_grad_(True)
    params = [p.detach().clone().requires_grad_(True) for p in m.params()]
    out = FluxCNN.apply(x, *params)
    gout = torch.randn_like(out)
    (out * gout).sum().backward()
    xo = x.detach().double().cpu().requires_grad_(True)
    po = [p.detach().double().cpu().requires_grad_(True) for p in params]
    ref = nn_oracle.forward(xo, po[0::2], po[1::2])
    (ref * gout.double().cpu()).sum().backward()
    err = (out.detach().double().cpu() - ref.detach()).abs().max() / ref.abs().max()
    assert err < 2e-3, err
    for got, want in [(x.grad, xo.grad)] + [(p.grad, q.grad) for p, q in zip(params, po)]:
        e = (got.double().cpu() - want).abs().max() / want.abs().max().clamp_min(1e-30)
        assert e < 3e-2, e


@pytest.mark.gpu
@pytest.mark.parametrize("B,H,W,cin,cout,fwd,wgrad", [(2, 16, 24, 32, 32, True, True), (1, 8, 40, 16, 48, True, True),
                                                       (2, 8, 24, 32, 96, True, True), (1, 16, 12, 16, 16, False, True),
                                                       (1, 12, 16, 16, 16, True, False), (1, 20, 8, 16, 32, True, False)])
def test_layer_kernels_match_float64_at_widths_not_a_multiple_of_16(cuda, B, H, W, cin, cout, fwd, wgrad):
    """npde_conv3x3 and npde_conv3x3_wgrad at image widths that are not multiples of 16 (8-column tiles; the
    weight gradient's K blocks of 16 pixels straddle image rows at varying offsets), through both weight-gradient
    kernels (LDS-staged up to 64 output channels, global loads beyond, and widths that are multiples of 4 only),
    and the forward at heights that are not multiples of 8 (tiles of 6 x 16 and 10 x 8 pixels): against float64
    periodic convolutions of the same bfloat16-rounded operands.  Products of two bfloat16 values are exact in
    float32, so only the float32 accumulation differs (relative 1e-5)."""
    import torch.nn.functional as Fn
    from npde.learned import conv as K
    g_ = torch.Generator().manual_seed(B * 1000 + W * 10 + cout)
    x = torch.randn(B, H, W, cin, generator=g_)
    w = torch.randn(cout, cin, 3, 3, generator=g_) * 0.2
    bias = torch.randn(cout, generator=g_)
    g = torch.randn(B, H, W, cout, generator=g_)
    bf = lambda t: t.to(torch.bfloat16).to(torch.float64)
    xpad = Fn.pad(bf(x).permute(0, 3, 1, 2), (1, 1, 1, 1), mode="circular")
    rel = lambda a, b: float((a.double().cpu() - b).abs().max() / b.abs().max())
    if fwd:
        y, _ = K.conv3x3(x.to(cuda), K.pack_weights(w.to(cuda)), bias.to(cuda), cout, cout, K.EPI_LINEAR)
        y_ref = (Fn.conv2d(xpad, bf(w)) + bias.double().view(1, -1, 1, 1)).permute(0, 2, 3, 1)
        assert rel(y, y_ref) < 1e-5
    if not wgrad:
        return
    wv = torch.zeros(cout, cin, 3, 3, dtype=torch.float64, requires_grad=True)
    (Fn.conv2d(xpad, wv) * bf(g).permute(0, 3, 1, 2)).sum().backward()
    dw, db = K.conv3x3_wgrad(x.to(cuda), g.to(cuda))
    assert rel(dw, wv.grad) < 1e-5
    assert rel(db, g.double().sum((0, 1, 2))) < 1e-5


@pytest.mark.gpu
def test_network_gradients_pass_directional_finite_differences(cuda):
    from npde.learned.conv import FluxCNN
    m = _random_net(2, cuda)
    x = torch.randn(1, 16, 16, 16, device=cuda, generator=torch.Generator(device=cuda).manual_seed(2))
    x[..., 8:] = 0
    params = [p.detach().clone().requires_grad_(True) for p in m.params()]
    gout = torch.randn(1, 16, 16, 8, device=cuda, generator=torch.Generator(device=cuda).manual_seed(4))
    (FluxCNN.apply(x, *params) * gout).sum().backward()
    gen = torch.Generator(device=cuda).manual_seed(3)
    for k in (0, 2, 8):
        # half along the gradient, half random: the difference quotient of bf16-rounded weights carries
        # rounding noise ~ |grad| * ulp, so the direction must not be nearly orthogonal to the gradient
        r = torch.randn(params[k].shape, device=cuda, generator=gen)
        v = params[k].grad / params[k].grad.norm() + r / r.norm()
        v /= v.norm()
        eps = 0.05 * float(params[k].detach().norm()) + 1e-3

        def f(s):
            ps = [p.detach().clone() for p in params]
            ps[k] += s * v
            return float((FluxCNN.apply(x, *ps) * gout).sum())

        fd = (f(eps) - f(-eps)) / (2 * eps)
        an = float((params[k].grad * v).sum())
        assert abs(fd - an) <= 0.05 * abs(an) + 1e-3, (k, fd, an)


@pytest.mark.gpu
def test_network_is_translation_equivariant(cuda):
    m = _random_net(4, cuda)
    U = torch.from_numpy(riemann_batch(32, [2])).to(cuda)
    cx, cy = m(U)
    cx2, cy2 = m(torch.roll(U, shifts=(16, 16), dims=(-2, -1)))
    assert torch.equal(torch.roll(cx, shifts=(16, 16), dims=(-2, -1)), cx2)
    cx3, _ = m(torch.roll(U, shifts=(1, 3), dims=(-2, -1)))
    assert torch.allclose(torch.roll(cx, shifts=(1, 3), dims=(-2, -1)), cx3, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("n,tuning,layout", [(64, {}, "row_groups"), (128, dict(row_groups=0, seams=2), "seams"),
                                            (128, dict(row_groups=0, seams=0), "strips")])
def test_corrected_hip_step_equals_torch_step(npde_mod, cuda, n, tuning, layout):
    """The learned increments enter the HIP step exactly as in the differentiable torch step, in every
    task layout (the seam pre-pass adds the increment of the faces it solves)."""
    from npde.learned import torch_solver as ts
    B = 3
    U0 = torch.from_numpy(riemann_batch(n, range(B))).to(cuda)
    cx = (0.02 * torch.randn(B, 4, n, n, dtype=torch.float64, device=cuda)).contiguous()
    cy = (0.02 * torch.randn(B, 4, n, n, dtype=torch.float64, device=cuda)).contiguous()
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B
    dt = torch.full((B,), 0.001, dtype=torch.float64, device=cuda)
    ref = ts.step(U0, dt, 1 / n, 1 / n, cx_corr=cx, cy_corr=cy)
    plain = ts.step(U0, dt, 1 / n, 1 / n)
    torch.cuda.synchronize()
    with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as b:
        if tuning:
            b.set_tuning(**tuning)
        assert b.task_layout == layout
        for k in range(B):
            b.set_state(k, U0[k].cpu().numpy())
            b.set_flux_correction(k, cx[k].data_ptr(), cy[k].data_ptr())
        b.step_all([0.001] * B)
        got = np.stack([b.get_state(k) for k in range(B)])
        zero = torch.zeros_like(cx)
        for k in range(B):
            b.set_state(k, U0[k].cpu().numpy())
            b.set_flux_correction(k, zero[k].data_ptr(), zero[k].data_ptr())
        torch.cuda.synchronize()
        b.step_all([0.001] * B)
        got0 = np.stack([b.get_state(k) for k in range(B)])
    assert np.array_equal(got, ref.cpu().numpy())
    assert np.array_equal(got0, plain.cpu().numpy())   # zero increments: the plain solver's bits
    tot0 = U0.sum(dim=(-1, -2)).cpu().numpy()
    assert np.allclose(got.sum(axis=(-1, -2)), tot0, rtol=0, atol=1e-11 * np.abs(tot0).max())


@pytest.mark.gpu
def test_learned_rollout_fallback_equals_plain_fine_solve(npde_mod, cuda):
    """Forced fallback: every learned instance is replaced by the plain binary64 512^2 solve, block
    averaged -- exactly the frames the fine solve produces on its own."""
    from npde.learned import pipeline as P
    from npde.learned.model import FluxCorrector
    idx = [100000, 100001]
    fr, _, fell = P.learned_rollout(FluxCorrector().to(cuda), idx, force_fallback=True, log=lambda *a: None)
    fine, _ = P.generate(idx, sizes=(64,), log=lambda *a: None)
    assert fell == idx
    assert np.array_equal(fr, fine[64])


@pytest.mark.gpu
def test_untrained_learned_rollout_is_the_plain_coarse_solver(npde_mod, cuda):
    """Zero-initialised head: the graph-captured learned rollout reproduces the plain 64^2 solver."""
    from npde.learned import pipeline as P
    from npde.learned.model import FluxCorrector
    idx = [100002, 100003]
    fr, _, fell = P.learned_rollout(FluxCorrector().to(cuda), idx, log=lambda *a: None)
    plain, _ = P.plain_rollout(idx, 64)
    assert fell == []
    assert np.array_equal(fr, plain)


@pytest.mark.gpu
def test_trained_corrector_reduces_coarse_error(npde_mod, cuda):
    """The committed trained weights lower the 64^2 rollout's density L1 against the block-averaged
    512^2 truth on held-out instances (the training indices are 0-447)."""
    from npde.learned import pipeline as P
    from npde.learned.model import load_trained
    idx = list(range(100100, 100116))
    truth, _ = P.generate(idx, sizes=(64,), log=lambda *a: None)
    plain, _ = P.plain_rollout(idx, 64)
    learned, _, fell = P.learned_rollout(load_trained(device=cuda), idx, log=lambda *a: None)
    # an instance that fell back carries the fine solve's frames: compare on the others only
    own = np.array([i not in fell for i in idx])
    assert own.sum() >= len(idx) - 2, fell
    e_plain = P.l1_density(plain[own], truth[64][own])[:, 1:].mean()
    e_learned = P.l1_density(learned[own], truth[64][own])[:, 1:].mean()
    assert e_learned < 0.8 * e_plain, (e_learned, e_plain)


@pytest.mark.gpu
def test_held_out_kh_evaluation_runs_and_charges_fallbacks(npde_mod, cuda):
    """The held-out Kelvin-Helmholtz protocol on a small setting (fine 256^2, coarse 64^2, two
    instances): every instance has a truth-relative density error for both solvers, an instance that
    fell back takes the fine frames (zero error) and is charged at least the fine solve's time."""
    from npde.learned.model import load_trained
    from npde.learned.pipeline import evaluate_kh
    res = evaluate_kh(load_trained(), [0, 1], n=64, n_fine=256, log=lambda *a: None)
    assert len(res["instances"]) == 2 and res["n"] == 64 and res["n_fine"] == 256
    for r in res["instances"]:
        assert 0.5 <= r["t_start"] <= 1.5
        assert np.isfinite(r["l1_rho_plain"]) and r["l1_rho_plain"] > 0.0
        if r["fallback"]:
            assert r["l1_rho_learned"] == 0.0 and r["learned_seconds"] >= r["fine_seconds"]
        else:
            assert np.isfinite(r["l1_rho_learned"]) and r["l1_rho_learned"] > 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 64, 96])
def test_fused_network_kernel_matches_oracle(npde_mod, cuda, n):
    """The fused inference kernel (features, four GELU layers and the head in one kernel, activations in
    LDS as bfloat16, 16x16x32 MFMAs) on solver states against the float64 network oracle with the same
    bfloat16 roundings, on grids of one, four and six tiles per side: the factors out_scale * tanh(head)
    agree to 2e-3 of their largest magnitude, and the instances of a batch are independent."""
    import nn_oracle
    from npde.learned.model import features
    m = _random_net(5, cuda, head_scale=0.05)
    m.out_scale = 1.0
    B = 3
    U0 = riemann_batch(n, range(10, 10 + B))
    with npde_mod.Batch([npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B, "hllc", "vanleer", 0.5) as b:
        for k in range(B):
            b.set_state(k, U0[k])
        b.set_flux_network(m)
        b.flux_network_eval()
        got = []
        for k in range(B):  # the call returned after the kernel completed: read the planes in place
            px, py = b.flux_network_planes(k)
            cx, cy = (torch.as_tensor(_DevView(p, (4, n, n)), device=cuda).clone() for p in (px, py))
            got.append(torch.cat([cx, cy]).permute(1, 2, 0).cpu())
        peak = b.flux_network_peak()
    x = features(torch.from_numpy(U0)).double()
    ws = [w.detach().double().cpu() for w in m.weights]
    bs = [bb.detach().double().cpu() for bb in m.biases]
    want = torch.tanh(nn_oracle.forward(x, ws, bs)) * m.out_scale  # (B, n, n, 8)
    for k in range(B):
        err = float((got[k] - want[k]).abs().max() / want[k].abs().max())
        assert err < 2e-3, (n, k, err)
    assert np.array_equal(peak, U0[:, 0].reshape(B, -1).max(axis=1).astype(np.float32))


class _DevView:
    """A (shape) float64 view of a device address, for torch.as_tensor."""

    def __init__(self, ptr, shape):
        strides = []
        acc = 8
        for d in reversed(shape):
            strides.insert(0, acc)
            acc *= d
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f8", "data": (ptr, False),
                                         "version": 3, "strides": tuple(strides)}


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["row_groups", "seams", "strips"])
def test_network_in_the_device_loop_equals_a_host_loop(npde_mod, cuda, layout):
    """npde_run_trajectory with the network set evaluates it at every frame landing of every instance: the
    frames and final states equal a host-driven loop that evaluates the network on every instance and then
    advances to the next frame time, bit for bit (the instances land at different step counts), in each
    task layout.  In the seam layout an instance's first step waits at the first frame for the network,
    which runs in the launch's fourth iteration by default: that step must not read seam edge columns no
    stage kernel has written yet (it once did, and the run went non-finite)."""
    m = _random_net(6, cuda, head_scale=0.3)
    m.out_scale = 0.1
    n, B, T, NOUT = 64, 4, 0.06, 6
    U0 = riemann_batch(n, range(20, 20 + B))
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B
    tuning = {"row_groups": {"row_groups": 1}, "seams": {"row_groups": 0, "seams": 2},
              "strips": {"row_groups": 0, "seams": 0}}[layout]
    with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as a, npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as h:
        a.set_tuning(**tuning)
        h.set_tuning(**tuning)
        assert a.task_layout == layout
        for k in range(B):
            a.set_state(k, U0[k])
            h.set_state(k, U0[k])
        a.set_flux_network(m)
        a.set_flux_network_limit(0)  # the host loop below has no switch-off
        h.set_flux_network(m)
        a.run_trajectory(T, NOUT, sizes=(n,))
        frames = {k: a.trajectory(k, n) for k in range(B)}
        for j in range(1, NOUT + 1):
            h.flux_network_eval()
            h.advance_to(0.0 + (T - 0.0) * j / NOUT)
            for k in range(B):
                assert np.array_equal(h.get_state(k).astype(np.float32), frames[k][j]), (j, k)
        steps = a.times()[1]
        assert len(set(steps.tolist())) > 1  # the instances took different numbers of steps
        for k in range(B):
            assert a.time(k) == h.time(k)
            assert np.array_equal(a.get_state(k), h.get_state(k))
        final = {k: a.get_state(k) for k in range(B)}
        # the network kernel's launch period and the graph length change only when a held batch is released,
        # never a result: every period (including ones longer than some launches) gives the same bits
        assert a.tuning["net_period"] == 4  # the default
        for period, graph in ((1, 32), (3, 32), (7, 5), (64, 8), (2, 1)):
            a.set_tuning(net_period=period, steps_per_graph=graph)
            for k in range(B):
                a.set_state(k, U0[k])
            a.run_trajectory(T, NOUT, sizes=(n,))
            for k in range(B):
                assert np.array_equal(a.trajectory(k, n), frames[k]), (period, graph, k)
                assert np.array_equal(a.get_state(k), final[k]), (period, graph, k)
            assert np.array_equal(a.times()[1], steps), (period, graph)
        for bad in (0, 65):
            with pytest.raises(npde_mod.NpdeError):
                a.set_tuning(net_period=bad)


@pytest.mark.gpu
def test_accuracy_cost_table(npde_mod, cuda):
    """The accuracy-versus-cost table on a small setting (8 held-out instances, plain 32^2 and 64^2 and the
    learned 64^2): every method has a positive cost and finite accuracy columns, conserves every total and has
    no bad cell, the finer plain solve is the more accurate, and the Pareto flags mark exactly the methods
    that no other method beats in both cost and last-frame error."""
    from npde.learned import pareto
    res = pareto.run(instances=8, ladder=(32, 64), log=lambda *a: None)
    rows = {r["method"]: r for r in res["rows"]}
    assert set(rows) == {"plain 32^2", "plain 64^2", "learned 64^2"}
    for r in rows.values():
        assert r["cost"] > 0 and np.isfinite(r["rmse_last"]) and np.isfinite(r["l1_rho"])
        assert r["bad_cells"] == 0 and abs(r["drift_max"]) < 1e-4  # periodic: every total is conserved
    assert rows["plain 64^2"]["rmse_last"] < rows["plain 32^2"]["rmse_last"]
    assert rows["plain 64^2"]["l1_rho"] < rows["plain 32^2"]["l1_rho"]
    assert any(r["pareto"] for r in rows.values())
    for r in rows.values():
        beaten = any(o is not r and o["cost"] <= r["cost"] and o["rmse_last"] <= r["rmse_last"]
                     and (o["cost"] < r["cost"] or o["rmse_last"] < r["rmse_last"]) for o in rows.values())
        assert r["pareto"] == (not beaten)


@pytest.mark.gpu
def test_network_switches_off_where_the_signal_speed_runs_away(npde_mod, cuda):
    """npde_set_flux_network_limit: at a frame landing where an instance's CFL signal-speed sum exceeds the
    limit times its first frame's, its factors become zero -- from then on it is the plain coarse solver.
    With a limit every instance exceeds at once, the in-loop run equals a host loop that evaluates the network
    at t0, advances to the first frame and then runs without it to every later frame time, bit for bit; with
    limit 0 nothing switches.
    The held-out instance 100511, whose corrected flow drains a cell toward vacuum (1,723 steps against 695
    plain without the switch), is caught by the default limit and finishes in far fewer steps (the drained
    cell, already there when the limit trips, keeps its steps above the plain solver's)."""
    from npde.learned import pipeline as P
    from npde.learned.model import load_trained
    m = _random_net(7, cuda, head_scale=0.3)
    m.out_scale = 0.1
    n, B, T, NOUT = 64, 3, 0.04, 4
    U0 = riemann_batch(n, range(30, 30 + B))
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=1)] * B
    with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as a, npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as h:
        for k in range(B):
            a.set_state(k, U0[k])
            h.set_state(k, U0[k])
        a.set_flux_network(m)
        a.set_flux_network_limit(1e-9)
        a.run_trajectory(T, NOUT, sizes=(n,))
        assert a.flux_network_switched_off().all()
        h.set_flux_network(m)
        h.flux_network_eval()
        h.advance_to(T * 1 / NOUT)
        h.set_flux_network(None)
        for j in range(2, NOUT + 1):  # landing on every frame time, as the recorded run does
            h.advance_to(T * j / NOUT)
        for k in range(B):
            assert a.time(k) == h.time(k)
            assert np.array_equal(a.get_state(k).view(np.uint64), h.get_state(k).view(np.uint64)), k
        a.set_flux_network_limit(0)
        for k in range(B):
            a.set_state(k, U0[k])
        a.run_trajectory(T, NOUT, sizes=(n,))
        assert not a.flux_network_switched_off().any()
    U = P.block_average_f64(P.riemann(100511, 512)[0], 64)[None]
    g = [npde_mod.make_grid(64, 64, bc_x=1, bc_y=1)]
    with npde_mod.Batch(g, "hllc", "vanleer", 0.5) as b:
        b.set_state(0, U[0])
        b.run_trajectory(0.5, 50, sizes=(64,))
        plain = int(b.times()[1][0])
        b.set_flux_network(load_trained(device=cuda))
        steps = {}
        for limit in (0.0, 2.5):
            b.set_flux_network_limit(limit)
            b.set_state(0, U[0])
            b.run_trajectory(0.5, 50, sizes=(64,))
            steps[limit] = (int(b.times()[1][0]), bool(b.flux_network_switched_off()[0]))
    print(f"instance 100511: plain {plain} steps, learned {steps}", flush=True)
    assert steps[0.0][0] > 1.5 * plain and not steps[0.0][1]
    assert steps[2.5][1] and steps[2.5][0] < 0.7 * steps[0.0][0]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_random_flux_network_runs_are_layout_and_launch_invariant(npde_mod, cuda, seed):
    """Seeded random flux-network runs: a batch of 2-5 instances of 64^2 or 128^2 (mixed, periodic; the network
    tiles are 16 x 16) recorded over a few frames with the network in the device loop, under a random task layout
    (row groups where every width is 64, seams, strips), band height, network launch period and graph length, give
    the bits of the same run in 60-column strips with the network launched in every iteration: every frame,
    final state, step count and switch-off flag.  Layouts, launch shapes and the network's launch period never
    change a result, so any difference is a bug in one of the paths (stale seam edge columns, a missed release,
    a frame written twice)."""
    rng = np.random.default_rng(1000 + seed)
    B = int(rng.integers(2, 6))
    sizes = [int(rng.choice([64, 128])) for _ in range(B)]
    if seed == 0:
        sizes = [64] * B  # row groups apply
    nout = int(rng.integers(3, 7))
    T = float(rng.uniform(0.02, 0.05))
    m = _random_net(10 + seed, cuda, head_scale=0.3)
    m.out_scale = 0.1
    from npde.families import riemann
    ics = [riemann(int(rng.integers(0, 10000)), n)[0] for n in sizes]
    grids = [npde_mod.make_grid(n, n, bc_x=1, bc_y=1) for n in sizes]
    layouts = ["seams", "strips"] + (["row_groups"] if all(n == 64 for n in sizes) else [])
    layout = layouts[seed % len(layouts)]
    tuning = {"row_groups": {"row_groups": 1, "seams": 0}, "seams": {"row_groups": 0, "seams": 2},
              "strips": {"row_groups": 0, "seams": 0}}[layout]
    tuning.update(net_period=int(rng.integers(1, 9)), steps_per_graph=int(rng.choice([1, 3, 8, 32])),
                  rows_per_task=int(rng.choice([0, 4, 16])))
    runs = {}
    for name, tu in (("reference", {"row_groups": 0, "seams": 0, "net_period": 1, "steps_per_graph": 32,
                                    "rows_per_task": 0}), ("random", tuning)):
        with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as b:
            b.set_tuning(**tu)
            if name == "random":
                assert b.task_layout == layout, (b.task_layout, layout)
            b.set_flux_network(m)
            for k, U in enumerate(ics):
                b.set_state(k, U)
            b.run_trajectory(T, nout, sizes=tuple(sorted(set(sizes))[:1]))
            m0 = min(sizes)
            runs[name] = ([b.trajectory(k, m0) for k in range(B)], [b.get_state(k) for k in range(B)],
                          b.times()[1].copy(), b.flux_network_switched_off().copy())
    (fa, sa, na, oa), (fb, sb, nb, ob) = runs["reference"], runs["random"]
    assert np.array_equal(na, nb), (seed, tuning, na, nb)
    assert np.array_equal(oa, ob), seed
    for k in range(B):
        assert np.all(np.isfinite(sa[k])), (seed, k)
        assert np.array_equal(fa[k], fb[k]), (seed, tuning, k)
        assert np.array_equal(sa[k], sb[k]), (seed, tuning, k)
