"""Full-state surrogates (npde.surrogates): the checks of the reference's tests/test_models.py,
test_unet.py, test_fno.py and test_positive.py that do not need its dataset.  Parity of the numbers is
unpinned (the reference's models are JAX/Equinox, absent here); what is pinned are the parameter counts
the reference's tests assert, the identity at initialisation, equivariance, the spectral convolution
against an explicit DFT, and trainability."""
import math

import numpy as np
import pytest
import torch

from npde import surrogates as M


def test_parameter_counts_match_reference():
    """test_models.py:39, test_unet.py:40, test_fno.py:94."""
    assert M.count_params(M.ResNet(width=64, depth=8)) == 591_428
    assert M.count_params(M.UNet(width=16, levels=3)) == 516_516
    assert M.count_params(M.FNO(width=32, depth=4, modes=16)) == 4_203_428


@pytest.mark.parametrize("name", ["resnet", "unet", "fno"])
def test_untrained_model_is_identity(name):
    x = torch.randn(2, 4, 32, 32, generator=torch.Generator().manual_seed(0))
    m = M.MODELS[name](**({"width": 16, "depth": 2} if name == "resnet" else
                         {"width": 8, "levels": 2} if name == "unet" else {"width": 8, "depth": 2, "modes": 6}))
    assert torch.equal(m(x), x)


def _noisy(m, seed=1):
    g = torch.Generator().manual_seed(seed)
    last = m.project2 if hasattr(m, "project2") else m.project
    with torch.no_grad():
        last.weight.copy_(0.1 * torch.randn(last.weight.shape, generator=g))
    return m


@pytest.mark.parametrize("name,shift", [("resnet", (7, -13)), ("fno", (5, 11)), ("unet", (8, -16))])
def test_translation_equivariance(name, shift):
    """Periodic padding: roll(model(x)) == model(roll(x)) (the U-Net for shifts that are multiples of
    2**levels, test_unet.py:50)."""
    torch.manual_seed(0)
    x = torch.randn(1, 4, 32, 32, dtype=torch.float64)
    kw = {"resnet": {"width": 16, "depth": 2}, "unet": {"width": 8, "levels": 3},
          "fno": {"width": 8, "depth": 2, "modes": 6}}[name]
    m = _noisy(M.MODELS[name](**kw)).double()
    r = lambda z: torch.roll(z, shift, dims=(-2, -1))
    assert torch.allclose(r(m(x)), m(r(x)), atol=1e-10)


def test_spectral_conv_equals_explicit_dft():
    """test_fno.py:40-58: the layer equals a direct sum over the kept modes."""
    torch.manual_seed(0)
    n, m, ci, co = 16, 4, 3, 2
    layer = M.SpectralConv2d(ci, co, m).double()
    x = torch.randn(1, ci, n, n, dtype=torch.float64)
    got = layer(x)[0]
    X = torch.fft.rfft2(x[0])
    W = torch.complex(layer.w_re, layer.w_im)
    Y = torch.zeros(co, n, n // 2 + 1, dtype=torch.complex128)
    for o in range(co):
        for kx in list(range(m)) + list(range(-m, 0)):
            for ky in range(m):
                blk, ix = (0, kx) if kx >= 0 else (1, kx + m)
                Y[o, kx % n, ky] = sum(X[i, kx % n, ky] * W[blk, i, o, ix, ky] for i in range(ci))
    ref = torch.fft.irfft2(Y, s=(n, n))
    assert torch.allclose(got, ref, atol=1e-12)


def test_spectral_conv_keeps_only_low_modes_and_is_linear():
    """test_fno.py:60-76: a single Fourier mode inside the kept range comes out as that mode only; one
    outside is removed; the layer is linear."""
    n, m = 32, 5
    layer = M.SpectralConv2d(1, 1, m).double()
    X, Y = torch.meshgrid(torch.arange(n, dtype=torch.float64), torch.arange(n, dtype=torch.float64), indexing="ij")
    for kx, ky, inside in [(2, 3, True), (-3, 1, True), (7, 2, False), (1, 9, False)]:
        z = torch.cos(2 * math.pi * (kx * X + ky * Y) / n)[None, None]
        out = layer(z)[0, 0]
        F_ = torch.fft.fft2(out)
        energy = F_.abs() ** 2
        keep = energy[kx % n, ky % n] + energy[-kx % n, -ky % n]
        if inside:
            assert (energy.sum() - keep) / energy.sum() < 1e-12
        else:
            assert out.abs().max() < 1e-12
    a, b = torch.randn(1, 1, n, n, dtype=torch.float64), torch.randn(1, 1, n, n, dtype=torch.float64)
    assert torch.allclose(layer(2 * a - 3 * b), 2 * layer(a) - 3 * layer(b), atol=1e-12)


@pytest.mark.parametrize("n,m", [(16, 4), (16, 8), (32, 12), (64, 16)])
def test_spectral_conv_dft_path_equals_rfft2(n, m):
    """The truncated-DFT path (backend "dft": matrix products over the kept modes only) equals the rfft2 /
    irfft2 path to binary64 rounding, in the output and in the gradients of the input and of both weight
    blocks; including 2 m = n, where the two kx blocks meet.  A sign, conjugation or normalisation slip in
    either transform, or a swapped weight block, fails by O(1)."""
    torch.manual_seed(n + m)
    layer = M.SpectralConv2d(3, 5, m, torch.Generator().manual_seed(1)).double()
    x = torch.randn(2, 3, n, n, dtype=torch.float64, requires_grad=True)
    res = {}
    for backend in ("torch", "dft"):
        layer.backend = backend
        y = layer(x)
        res[backend] = (y.detach(), torch.autograd.grad((y * torch.cos(y)).sum(), [x, layer.w_re, layer.w_im]))
    (ya, ga), (yb, gb) = res["torch"], res["dft"]
    assert float((ya - yb).abs().max()) < 1e-12 * float(ya.abs().max())
    for u, v in zip(ga, gb):  # sums over n^2 terms: 1e-11 of the largest entry
        assert float((u - v).abs().max()) < 1e-11 * float(u.abs().max())
    fno_t = _noisy(M.FNO(width=8, depth=2, modes=m // 2 or 1)).double()
    fno_d = M.FNO(width=8, depth=2, modes=m // 2 or 1, backend="dft").double()
    fno_d.load_state_dict(fno_t.state_dict())
    z = torch.randn(1, 4, n, n, dtype=torch.float64)
    assert torch.allclose(fno_t(z), fno_d(z), rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        M.FNO(width=8, depth=1, modes=4, backend="fft")


def test_fno_runs_at_any_resolution():
    m = _noisy(M.FNO(width=8, depth=2, modes=4)).double()
    n1 = torch.linspace(0, 2 * math.pi, 33, dtype=torch.float64)[:-1]
    n2 = torch.linspace(0, 2 * math.pi, 65, dtype=torch.float64)[:-1]
    f = lambda t: torch.stack([torch.sin(t)[:, None] * torch.cos(t)[None, :]] * 4)[None]
    a = m(f(n1))
    b = m(f(n2))[..., ::2, ::2]
    assert b.shape == a.shape and float((a - b).abs().max() / a.abs().max()) < 1e-3


def test_positive_wrapper_round_trip_and_positivity():
    """test_positive.py:30-58."""
    g = torch.Generator().manual_seed(0)
    rho = 0.1 + torch.rand(2, 16, 16, generator=g, dtype=torch.float64)
    u, v = torch.randn(2, 2, 16, 16, generator=g, dtype=torch.float64).unbind(1)
    p = 0.1 + torch.rand(2, 16, 16, generator=g, dtype=torch.float64)
    U = torch.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], dim=1)
    assert torch.allclose(M.logprim_to_conserved(M.conserved_to_logprim(U)), U, rtol=1e-12)
    mean, std = [1.0, 0.0, 0.0, 2.5], [0.5, 0.5, 0.5, 1.5]
    core = M.ResNet(width=16, depth=2).double()
    model = M.Positive(core, mean, std, [0.0, 0.0, 0.0, 0.5], [1.0, 1.0, 1.0, 1.0]).double()
    x = (U - torch.tensor(mean).view(1, 4, 1, 1)) / torch.tensor(std).view(1, 4, 1, 1)
    assert torch.allclose(model(x), x, atol=1e-10)
    with torch.no_grad():
        core.project.weight.copy_(5 * torch.randn(core.project.weight.shape, generator=g))
    out = model(x) * torch.tensor(std).view(1, 4, 1, 1) + torch.tensor(mean).view(1, 4, 1, 1)
    rho_o = out[:, 0]
    p_o = 0.4 * (out[:, 3] - 0.5 * (out[:, 1] ** 2 + out[:, 2] ** 2) / rho_o)
    assert torch.isfinite(out).all() and rho_o.min() > 0 and p_o.min() > -1e-9 * out[:, 3].abs().max()


@pytest.mark.parametrize("name", ["resnet", "unet", "fno"])
def test_gradients_reach_every_parameter_and_adam_fits_one_batch(name):
    """test_models.py:47-66: every parameter array gets a gradient once the last layer is non-zero, and
    Adam steps on one fixed batch at least halve the loss."""
    torch.manual_seed(0)
    kw = {"resnet": {"width": 16, "depth": 2}, "unet": {"width": 8, "levels": 2},
          "fno": {"width": 8, "depth": 2, "modes": 6}}[name]
    m = _noisy(M.MODELS[name](**kw))
    x = torch.randn(4, 4, 16, 16)
    y = x + 0.1 * torch.roll(x, 1, dims=-1)
    loss = torch.mean((m(x) - y) ** 2)
    loss.backward()
    assert all(p.grad is not None and float(p.grad.abs().max()) > 0 for p in m.parameters())
    opt = torch.optim.Adam(m.parameters(), lr=1e-3)
    l0 = float(loss)
    for _ in range(100):
        opt.zero_grad()
        loss = torch.mean((m(x) - y) ** 2)
        loss.backward()
        opt.step()
    assert float(loss) < 0.5 * l0


def test_downsample_upsample():
    """test_unet.py:33-34."""
    z = torch.randn(1, 3, 16, 16, dtype=torch.float64)
    assert torch.allclose(M.downsample(z).mean(), z.mean())
    assert torch.equal(M.downsample(M.upsample(z)), z)


def test_training_losses_and_rollout():
    """onestep / multistep / pushforward losses on a tiny set of trajectories; rollout shape."""
    torch.manual_seed(0)
    frames = torch.rand(3, 6, 4, 16, 16) + 0.5
    norm = M.Normaliser([1.0] * 4, [0.3] * 4, [0.1] * 4)
    for mode, h in (("onestep", 1), ("multistep", 2), ("pushforward", 3)):
        m = M.ResNet(width=8, depth=1)
        hist = M.train(m, frames, norm, steps=5, batch=2, horizon=h, mode=mode, log=lambda *a: None)
        assert len(hist) == 5 and all(np.isfinite(hist))
    out = M.rollout(m, norm.encode(frames[:, 0]), 4)
    assert out.shape == (3, 5, 4, 16, 16)
    m = M.rollout_metrics(norm.decode(out[0]), frames[0, :5], [0.3] * 4)
    rmse, l1 = m["rmse"], m["l1_rho"]
    assert rmse.shape == (5, 4) and l1.shape == (5,)


@pytest.mark.gpu
def test_resnet_learns_solver_frames_on_gpu(npde_mod):
    """End to end on the GPU: trajectories from the dataset pipeline's generator (GPU solver, block
    averaged to 32^2), a small ResNet trained for 150 steps with the one-step loss: its validation error
    drops below persistence (predicting "nothing changes"), the reference's yardstick (train.py:25-27)."""
    from npde import dataset as D
    items = [("riemann", i) for i in range(10)]
    trajs, _ = D.generate(items, (32,), riemann_n=64, batch=10, log=lambda *a: None)
    dev = torch.device("cuda")
    frames = torch.from_numpy(np.stack([trajs[k]["frames"][32] for k in items])).to(dev)
    train, val = frames[:8], frames[8:]
    flat = train.permute(2, 0, 1, 3, 4).reshape(4, -1)
    delta = (train[:, 1:] - train[:, :-1]).permute(2, 0, 1, 3, 4).reshape(4, -1)
    norm = M.Normaliser(flat.mean(1).cpu(), flat.std(1).cpu(), delta.std(1).cpu())
    model = M.ResNet(width=32, depth=4, delta_scale=(norm.delta_std / norm.std.flatten()).tolist()).to(dev)
    M.train(model, train, norm, steps=150, batch=16, lr=2e-3, log=lambda *a: None)
    with torch.no_grad():
        x, y = norm.encode(val[:, :-1].flatten(0, 1)), norm.encode(val[:, 1:].flatten(0, 1))
        err = float(torch.mean((model(x) - y) ** 2))
        persist = float(torch.mean((x - y) ** 2))
    assert err < 0.8 * persist, (err, persist)


@pytest.mark.gpu
def test_resnet_on_the_mfma_kernels_matches_the_torch_backend(npde_mod):
    """ResNet(backend="mfma") -- its 3x3 convolutions on the library's BF16-MFMA kernels, forward and backward --
    against the same parameters on torch's float32 convolutions: output change and every parameter gradient
    within BF16 tolerance (relative L2 error below 3e-2 and 5e-2), for widths 32 and 64; and a few Adam steps
    on the MFMA backend lower the one-batch loss."""
    dev = torch.device("cuda")
    for width in (32, 64):
        m = M.ResNet(width=width, depth=2, seed=3).to(dev)
        with torch.no_grad():  # a non-zero projection, so that the output depends on the trunk
            g = torch.Generator().manual_seed(1)
            m.project.weight.copy_(torch.randn(m.project.weight.shape, generator=g) * 0.1)
            m.project.bias.copy_(torch.randn(m.project.bias.shape, generator=g) * 0.1)
        x = torch.randn(4, 4, 64, 64, generator=torch.Generator().manual_seed(2)).to(dev)
        out, grads = {}, {}
        for backend in ("torch", "mfma"):
            m.backend = backend
            m.zero_grad()
            y = m(x)
            (y ** 2).mean().backward()
            out[backend] = (y - x).detach()
            grads[backend] = {k: p.grad.detach().clone() for k, p in m.named_parameters()}
        rel = lambda a, b: float((a - b).norm() / b.norm())
        assert rel(out["mfma"], out["torch"]) < 3e-2, width
        for k in grads["torch"]:
            assert rel(grads["mfma"][k], grads["torch"][k]) < 5e-2, (width, k)
    m.backend = "mfma"
    y_t = (x + 0.1 * torch.randn(x.shape, generator=torch.Generator().manual_seed(4)).to(dev)).detach()
    opt = torch.optim.Adam(m.parameters(), lr=1e-3)
    losses = []
    for _ in range(8):
        opt.zero_grad()
        loss = ((m(x) - y_t) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(float(loss))
    assert losses[-1] < 0.9 * losses[0], losses
    with pytest.raises(ValueError):
        M.ResNet(width=144, depth=1, backend="mfma").to(dev)(x)


@pytest.mark.gpu
def test_unet_on_the_mfma_kernels_matches_the_torch_backend(npde_mod):
    """UNet(backend="mfma") -- every 3x3 convolution on the library's BF16-MFMA kernels, forward and backward,
    with input channels 16 to 192 (skip concatenations), output channels 16 to 128 and images down to 8 x 8 --
    against the same parameters on torch's float32 convolutions: output change and every parameter gradient
    within BF16 tolerance (relative L2 error below 3e-2 and 6e-2), at the reference's default UNet(16, 3) and a
    two-level width-32 variant; and a few Adam steps on the MFMA backend lower the one-batch loss."""
    dev = torch.device("cuda")
    for width, levels in ((16, 3), (32, 2)):
        m = M.UNet(width=width, levels=levels, seed=5).to(dev)
        with torch.no_grad():
            g = torch.Generator().manual_seed(1)
            m.project.weight.copy_(torch.randn(m.project.weight.shape, generator=g) * 0.1)
            m.project.bias.copy_(torch.randn(m.project.bias.shape, generator=g) * 0.1)
        x = torch.randn(3, 4, 64, 64, generator=torch.Generator().manual_seed(2)).to(dev)
        out, grads = {}, {}
        for backend in ("torch", "mfma"):
            m.backend = backend
            m.zero_grad()
            y = m(x)
            (y ** 2).mean().backward()
            out[backend] = (y - x).detach()
            grads[backend] = {k: p.grad.detach().clone() for k, p in m.named_parameters()}
        rel = lambda a, b: float((a - b).norm() / b.norm())
        assert rel(out["mfma"], out["torch"]) < 3e-2, (width, levels)
        for k in grads["torch"]:
            assert rel(grads["mfma"][k], grads["torch"][k]) < 6e-2, (width, levels, k)
    m.backend = "mfma"
    y_t = (x + 0.1 * torch.randn(x.shape, generator=torch.Generator().manual_seed(4)).to(dev)).detach()
    opt = torch.optim.Adam(m.parameters(), lr=1e-3)
    losses = []
    for _ in range(8):
        opt.zero_grad()
        loss = ((m(x) - y_t) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(float(loss))
    assert losses[-1] < 0.9 * losses[0], losses
    with pytest.raises(ValueError):  # 96 x 96 does not halve to a multiple of 8 three times
        M.UNet(width=16, levels=3, backend="mfma").to(dev)(torch.zeros(1, 4, 96, 96, device=dev))


@pytest.mark.gpu
def test_fno_dft_path_on_the_gpu_matches_rfft2(npde_mod):
    """The FNO's truncated-DFT spectral layers on the GPU (float32 matrix products) against its rocFFT layers
    with the same parameters: outputs and gradients within float32 rounding (relative L2 below 1e-4)."""
    dev = torch.device("cuda")
    m = _noisy(M.FNO(width=32, depth=4, modes=16)).to(dev)
    x = torch.randn(4, 4, 64, 64, generator=torch.Generator().manual_seed(3)).to(dev)
    res = {}
    for backend in ("torch", "dft"):
        for b in m.blocks:
            b.spectral.backend = backend
        m.zero_grad()
        y = m(x)
        (y ** 2).mean().backward()
        res[backend] = (y.detach(), {k: p.grad.detach().clone() for k, p in m.named_parameters()})
    rel = lambda a, b: float((a - b).norm() / b.norm())
    assert rel(res["dft"][0], res["torch"][0]) < 1e-4
    for k in res["torch"][1]:
        assert rel(res["dft"][1][k], res["torch"][1][k]) < 1e-4, k


def test_rollout_drift_and_bad_cells_hand_computed():
    """surrogate/rollout.py:54-62: drift of each total from the truth's first frame over its total absolute
    amount, and the count of cells with rho <= 0 or p not > 0, on frames small enough to compute by hand."""
    true = torch.ones(2, 4, 2, 2, dtype=torch.float64)
    true[:, 1] = torch.tensor([[1.0, -1.0], [2.0, -2.0]])   # x momentum: total 0, absolute total 6
    true[:, 3] = 3.0                                          # E = 3: p = 0.4 (3 - 0.5 (mx^2 + 1) / 1) > 0
    pred = true.clone()
    pred[1, 0, 0, 0] = 1.5                                    # mass 4 -> 4.5: drift +0.5 / 4
    pred[1, 1, 1, 1] = -1.0                                   # x momentum 0 -> 1: drift 1 / 6
    pred[1, 3, 0, 1] = 0.5                                    # p = 0.4 (0.5 - 1) < 0: one bad cell
    pred[1, 0, 1, 0] = float("nan")                           # NaN density: bad (and mass drift NaN)
    m = M.rollout_metrics(pred, true, [1.0] * 4)
    d = m["drift"].numpy()
    assert np.allclose(d[0], 0.0) and np.isnan(d[1, 0])
    assert abs(d[1, 1] - 1.0 / 6.0) < 1e-15 and d[1, 2] == 0.0 and abs(d[1, 3] - (0.5 - 3.0) / 12.0) < 1e-15
    assert m["bad_cells"].tolist() == [0, 2]
    pred[1, 0, 1, 0] = 1.0
    m = M.rollout_metrics(pred, true, [1.0] * 4)
    assert abs(float(m["drift"][1, 0]) - 0.5 / 4.0) < 1e-15 and m["bad_cells"].tolist() == [0, 1]
