"""The flux-correction network's forward and backward on the project's BF16-MFMA kernels.

FluxCNN.apply(features, w1, b1, ..., wL, bL) is one torch.autograd.Function for the whole stack
  z_1 = conv(x) + b_1,  h_k = GELU(z_k),  z_{k+1} = conv(h_k) + b_{k+1},  out = z_L
(3x3 periodic convolutions, tanh-approximate GELU on hidden layers, linear head).  Its forward runs
npde_conv3x3 once per layer (bias + GELU fused in the epilogue, pre-activations kept for backward);
its backward runs, per layer from the top, npde_conv3x3_wgrad (weight and bias gradients) and the
data-gradient convolution with GELU'(z_{k-1}) fused into its epilogue.

Tensors: activations NHWC float32 on the GPU with channels padded to a multiple of 16 (values enter
the MFMAs rounded to bfloat16: the forward kernel stages each input tile in LDS as bfloat16 once);
weights (Cout, Cin, 3, 3) float32 master copies, packed to BF16 MFMA operand order on every call.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib

EPI_LINEAR, EPI_GELU, EPI_GELU_GRAD = 0, 1, 2
WGRAD_ROWS = 8


def _pad16(c: int) -> int:
    return (c + 15) // 16 * 16


def pack_weights(w: torch.Tensor, transpose_flip: bool = False) -> torch.Tensor:
    """(Cout, Cin, 3, 3) float32 -> bfloat16 A-operand fragments of the forward kernel,
    [tap][ci/32][co/16][lane][8]: lane = 16 g + m, element j holds W[co = 16 t + m][ci = 32 c + 8 g + j][tap]
    (tap = ky * 3 + kx), zero beyond Cin (padded to a multiple of 32) and Cout (to a multiple of 16).
    transpose_flip packs the data-gradient operator (taps reversed, ci and co exchanged)."""
    wt = w.permute(2, 3, 1, 0).reshape(9, w.shape[1], w.shape[0])  # [tap][ci][co]
    if transpose_flip:
        wt = wt.flip(0).transpose(1, 2)
    n_in, n_out = wt.shape[1], wt.shape[2]
    ci_p, co_p = (n_in + 31) // 32 * 32, _pad16(n_out)
    wp = wt.new_zeros(9, ci_p, co_p)
    wp[:, :n_in, :n_out] = wt
    wp = wp.reshape(9, ci_p // 32, 4, 8, co_p // 16, 16).permute(0, 1, 4, 2, 5, 3)  # [tap][c][t][g][m][j]
    return wp.contiguous().to(torch.bfloat16).contiguous()


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def conv3x3(x, wpack, bias, cout_p, cout_store, epilogue, want_pre=False, gpre=None):
    lib = _lib.load()
    B, H, W, cin = x.shape
    y = torch.empty(B, H, W, cout_store, device=x.device, dtype=torch.float32)
    pre = torch.empty(B, H, W, cout_p, device=x.device, dtype=torch.float32) if want_pre else None
    _check(lib.npde_conv3x3(_ptr(x), B, H, W, cin, _ptr(wpack), _ptr(bias), cout_p, cout_store, _ptr(y), _ptr(pre),
                            epilogue, _ptr(gpre), _stream()), "npde_conv3x3")
    return y, pre


def conv3x3_wgrad(x, g):
    """x: (B,H,W,Cin_p) layer input, g: (B,H,W,Cout_p) gradient of its pre-activation ->
    dW (Cout_p, Cin_p, 3, 3), db (Cout_p,)."""
    lib = _lib.load()
    B, H, W, cin = x.shape
    cout = g.shape[-1]
    nb = B * (H // WGRAD_ROWS)
    part = torch.empty(nb * 9 * cin * cout, device=x.device, dtype=torch.float32)
    bpart = torch.empty(nb * cout, device=x.device, dtype=torch.float32)
    dw = torch.empty(9, cin, cout, device=x.device, dtype=torch.float32)
    db = torch.empty(cout, device=x.device, dtype=torch.float32)
    _check(lib.npde_conv3x3_wgrad(_ptr(x), _ptr(g), B, H, W, cin, cout, WGRAD_ROWS, _ptr(part), _ptr(bpart), _ptr(dw),
                                  _ptr(db), _stream()), "npde_conv3x3_wgrad")
    return dw.view(3, 3, cin, cout).permute(3, 2, 0, 1), db


def pack_stack(params):
    """Per layer (packed BF16 weights, zero-padded float32 bias, padded Cout, stored Cout) for inference."""
    ws, bs = params[0::2], params[1::2]
    out = []
    for k, (w, b) in enumerate(zip(ws, bs)):
        cout = w.shape[0]
        cout_p = _pad16(cout)
        bias = torch.zeros(cout_p, device=w.device, dtype=torch.float32)
        bias[:cout] = b.detach()
        last = k == len(ws) - 1
        out.append((pack_weights(w.detach()), bias, cout_p, cout if last else cout_p))
    return out


def flux_cnn_infer(x, packed):
    """Forward of the stack without the pre-activations the backward needs (inference only)."""
    h = x
    for k, (wp, bias, cout_p, cout_store) in enumerate(packed):
        last = k == len(packed) - 1
        h, _ = conv3x3(h, wp, bias, cout_p, cout_store, EPI_LINEAR if last else EPI_GELU)
    return h


class FluxCNN(torch.autograd.Function):
    """Whole network as one autograd node; see the module docstring."""

    @staticmethod
    def forward(ctx, x, *params):
        ws, bs = params[0::2], params[1::2]
        L = len(ws)
        hs, zs = [x], []
        h = x
        for k in range(L):
            cout = ws[k].shape[0]
            cout_p = _pad16(cout)
            bias = torch.zeros(cout_p, device=x.device, dtype=torch.float32)
            bias[:cout] = bs[k]
            last = k == L - 1
            y, z = conv3x3(h, pack_weights(ws[k]), bias, cout_p, cout if last else cout_p,
                           EPI_LINEAR if last else EPI_GELU, want_pre=not last)
            if not last:
                zs.append(z)
                hs.append(y)
            h = y
        ctx.save_for_backward(*hs, *zs, *ws)
        ctx.L = L
        return h

    @staticmethod
    def backward(ctx, gout):
        L = ctx.L
        saved = ctx.saved_tensors
        hs, zs, ws = saved[:L], saved[L:2 * L - 1], saved[2 * L - 1:]
        grads_w, grads_b = [None] * L, [None] * L
        cout = ws[-1].shape[0]
        g = torch.zeros(*gout.shape[:-1], _pad16(cout), device=gout.device, dtype=torch.float32)
        g[..., :cout] = gout
        gx = None
        for k in range(L - 1, -1, -1):
            co, ci = ws[k].shape[:2]
            dw, db = conv3x3_wgrad(hs[k], g)
            grads_w[k] = dw[:co, :ci].contiguous()
            grads_b[k] = db[:co].contiguous()
            if k > 0 or ctx.needs_input_grad[0]:
                cin_p = hs[k].shape[-1]
                epi = EPI_GELU_GRAD if k > 0 else EPI_LINEAR
                g, _ = conv3x3(g, pack_weights(ws[k], transpose_flip=True), None, cin_p, cin_p, epi,
                               gpre=zs[k - 1] if k > 0 else None)
                if k == 0:
                    gx = g
        out = [gx]
        for k in range(L):
            out += [grads_w[k], grads_b[k]]
        return tuple(out)


def main(argv=None):
    """Time the corrector network at the training shape (layer kernels, forward + backward, batch 32 x 64^2)
    and at the rollout shape (the fused inference kernel on a batch of 64 x 64^2 solver states, and the
    layer kernels' forward for comparison):  python -m npde.learned.conv [--iters 20]"""
    import argparse
    import time
    from .. import solver as S
    from ..families import riemann
    from .model import FluxCorrector, features
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=20)
    args = ap.parse_args(argv)
    dev = torch.device("cuda")
    m = FluxCorrector().to(dev)
    with torch.no_grad():
        m.weights[-1].normal_(0.0, 0.1)
    g = torch.Generator(device=dev).manual_seed(0)

    def timed(f):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.iters):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.iters

    x = torch.randn(32, 64, 64, 16, device=dev, generator=g)
    x[..., 8:] = 0
    ms = timed(lambda: FluxCNN.apply(x, *m.params()).sum().backward())
    print(f"layer kernels, batch 32 x 64^2, forward + backward: {ms:.3f} ms per call", flush=True)
    B, n = 64, 64
    with S.Batch([S.make_grid(n, n, bc_x=1, bc_y=1)] * B, "hllc", "vanleer", 0.5) as b:
        for k in range(B):
            b.set_state(k, riemann(k, n)[0])
        b.set_flux_network(m)
        ms = timed(b.flux_network_eval)
        print(f"fused kernel, batch {B} x {n}^2 from the solver states: {ms:.3f} ms per call "
              f"(host-synchronised; the device time is in the kernel trace)", flush=True)
        U = b.state_view()
        packed = pack_stack(m.params())
        with torch.no_grad():
            ms = timed(lambda: flux_cnn_infer(features(U), packed))
        print(f"layer kernels, batch {B} x {n}^2 forward with features: {ms:.3f} ms per call", flush=True)


if __name__ == "__main__":
    main()
