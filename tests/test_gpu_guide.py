"""Guiding network on the GPU: BF16 MFMA forward / backward vs the FP64 oracle,
and online-style training that learns a conditional direction distribution."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, file=sys.stderr, flush=True)


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    inp = rng.random((n, 9)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.gamma(1.0, 1.0, n).astype(np.float32)
    return inp, d.astype(np.float32), w


def test_forward_matches_oracle(tt, ctx):
    from oracle import guide_ref as G
    g = tt.Guide(ctx, seed=3)
    p = g.params()
    inp, _, _ = _rows(1000, 1)                    # not a multiple of 16: ragged tail
    raw = g.forward(inp)
    ref, _ = G.forward(p, inp)
    err = np.abs(raw - ref)
    _say(f"forward: max abs err {err.max():.3g}, mean {err.mean():.3g}, ref scale {np.abs(ref).mean():.3g}")
    # features are bit-identical to the oracle's (same sincos and recurrence), so
    # only the FP32 accumulation order of the MFMA products differs
    assert err.max() < 1e-4 * max(1.0, np.abs(ref).max())
    assert err.mean() < 1e-5 * max(1.0, np.abs(ref).mean())
    g.close()


@pytest.mark.parametrize("n", [1, 16, 1000, 70001])
def test_fused_inference_matches_unfused_forward(tt, ctx, n):
    """The fused inference kernel (per-chain encoding into an LDS tile, 14-wave blocks)
    and the training path's k_encode + k_gemm chain perform the same arithmetic in the
    same order: outputs identical bit for bit, ragged tails included."""
    g = tt.Guide(ctx, seed=5)
    inp, _, _ = _rows(n, 7)
    a = g.forward(inp)
    b = g.forward_unfused(inp)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), np.abs(a - b).max()
    g.close()


@pytest.mark.parametrize("n,density,runs", [(1, 1.0, 1), (255, 0.5, 1), (257, 0.03, 1), (4096, 0.0, 1),
                                             (100003, 1.0, 1), (1000003, 0.4, 64), (3000017, 0.35, 64)])
def test_renderer_inference_rows_match_forward(tt, ctx, n, density, runs):
    """The renderer infers only its guided slots: ballot flags and block counts, the list
    of guided slots in slot order, the inference kernel over that list reading its inputs
    one tile ahead and writing each mixture back at its slot. Every guided slot gets the
    outputs the plain inference gives its inputs, bit for bit; no other slot is written.
    runs: the flag is drawn per run of that many slots (a pixel's samples share theirs);
    at 3M slots each block of the list kernels walks several 256-slot steps."""
    g = tt.Guide(ctx, seed=5)
    rng = np.random.default_rng(n)
    inp = rng.random((n, 9), dtype=np.float32)
    f = np.repeat(rng.random((n + runs - 1) // runs) < density, runs)[:n]
    if runs > 1:
        f ^= rng.random(n) < 0.02   # a few slots that differ from their run
    out, rows = g.infer_slots(inp, f)
    _say(f"infer_slots n={n}: {rows} guided rows")
    assert rows == int(f.sum())
    assert np.all(out[~f].view(np.uint32) == 0xFFFFFFFF)
    if rows:
        ref = g.forward(inp[f])
        assert np.array_equal(out[f].view(np.uint32), ref.view(np.uint32))
    g.close()


def test_gradient_is_bit_reproducible(tt, ctx):
    """The loss and every gradient element are the same bits on repeated evaluation of
    the same rows (the weight sum that scales the gradient and the loss are reduced in a
    fixed order, without float atomics), at a row count spanning many reduction blocks."""
    rng = np.random.default_rng(23)
    n = 200_003
    x = rng.random((n, 9), dtype=np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.gamma(1.0, 1.0, n).astype(np.float32)
    g = tt.Guide(ctx, seed=5, lr=1e-3)
    l0, g0 = g.grad(x, d, w)
    for _ in range(3):
        l1, g1 = g.grad(x, d, w)
        assert np.float32(l1).view(np.uint32) == np.float32(l0).view(np.uint32)
        assert np.array_equal(g1.view(np.uint32), g0.view(np.uint32))
    g.close()


@pytest.mark.parametrize("n", [1, 17, 1000, 70_001, 262_144])
def test_fused_training_kernels_match_unfused(tt, ctx, n):
    """The training step's fused forward (features and activations written by the
    inference kernel) and fused backward products give the loss and gradient of the
    separate encoding, layer and product kernels bit for bit."""
    rng = np.random.default_rng(n)
    x = rng.random((n, 9), dtype=np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.gamma(1.0, 1.0, n).astype(np.float32)
    g = tt.Guide(ctx, seed=7, lr=1e-3)
    g.train_step(x, d, w)                      # weights away from their initial values
    lf, gf = g.grad(x, d, w)
    lu, gu = g.grad_unfused(x, d, w)
    assert np.float32(lf).view(np.uint32) == np.float32(lu).view(np.uint32)
    assert np.array_equal(gf.view(np.uint32), gu.view(np.uint32)), \
        f"{np.count_nonzero(gf != gu)} gradient elements differ"
    g.close()


def test_gradient_matches_oracle(tt, ctx):
    from oracle import guide_ref as G
    g = tt.Guide(ctx, seed=5)
    p = g.params()
    inp, d, w = _rows(4096, 2)
    loss, grad = g.grad(inp, d, w)
    rloss, rgrad = G.grad(p, inp, d, w)
    _say(f"loss gpu {loss:.6f} oracle {rloss:.6f}")
    assert abs(loss - rloss) < 1e-3 * max(1.0, abs(rloss))
    o = 0
    for k, (i, n) in enumerate(G.LAYERS):
        for part, size in (("W", i * n), ("b", n)):
            a, b = grad[o:o + size], rgrad[o:o + size]
            o += size
            if np.linalg.norm(b) < 1e-12:
                assert np.linalg.norm(a) < 1e-6
                continue
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            _say(f"layer {k} {part}: cos {cos:.6f} rel {rel:.4f}")
            assert cos > 0.999 and rel < 0.04
    g.close()


def _vmf_sample(rng, mu, kap, n):
    u1, u2 = rng.random(n), rng.random(n)
    wz = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kap)) / kap
    s = np.sqrt(np.maximum(0, 1 - wz * wz))
    local = np.stack([s * np.cos(2 * np.pi * u2), s * np.sin(2 * np.pi * u2), wz], 1)
    mu = np.asarray(mu, float) / np.linalg.norm(mu)
    t = np.cross(mu, [0, 1, 0] if abs(mu[1]) < 0.9 else [1, 0, 0])
    t /= np.linalg.norm(t)
    b = np.cross(mu, t)
    return local[:, :1] * t + local[:, 1:2] * b + local[:, 2:] * mu


def test_training_learns_conditional_lobes(tt, ctx):
    """Inputs with x0 < 0.5 see radiance from a tight lobe about +z, the others
    from a lobe about +x; Adam steps on the GPU must move the mixture there."""
    from oracle import guide_ref as G
    rng = np.random.default_rng(7)
    g = tt.Guide(ctx, seed=11, lr=3e-3)
    first = None
    for step in range(300):
        n = 4096
        inp = rng.random((n, 9)).astype(np.float32)
        left = inp[:, 0] < 0.5
        d = np.where(left[:, None], _vmf_sample(rng, [0, 0, 1], 60.0, n), _vmf_sample(rng, [1, 0, 0], 20.0, n))
        loss = g.train_step(inp, d.astype(np.float32), np.ones(n, np.float32))
        first = loss if first is None else first
    _say(f"training loss {first:.3f} -> {loss:.3f}")
    assert loss < first - 2.0
    test = rng.random((512, 9)).astype(np.float32)
    raw = g.forward(test).astype(np.float64)
    left = test[:, 0] < 0.5
    tgt = np.where(left[:, None], [0, 0, 1.0], [1.0, 0, 0])
    other = np.where(left[:, None], [1.0, 0, 0], [0, 0, 1.0])
    lq_t, _, _ = G.mixture_logpdf(raw, tgt)
    lq_o, _, _ = G.mixture_logpdf(raw, other)
    _say(f"log q at target {np.median(lq_t):.2f}, at the other lobe {np.median(lq_o):.2f}, uniform {-np.log(4 * np.pi):.2f}")
    assert np.median(lq_t) > np.log(1 / (4 * np.pi)) + 2.0
    assert np.median(lq_t) > np.median(lq_o) + 2.0
    g.close()


def test_adam_step_matches_fp32_restatement(tt, ctx):
    """Two Adam steps (lr 1e-3, beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) of the
    device update equal a numpy float32 restatement of
    m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, p -= lr (m / c1) / (sqrt(v / c2) + eps)
    with c_t = 1 - b^t, using the gradients the device computes for the same rows
    (tt_guide_grad). A missing bias correction or a wrong moment update moves the
    parameters by a different amount (the update is lr at the first step)."""
    f = np.float32
    rng = np.random.default_rng(17)
    g = tt.Guide(ctx, seed=3, lr=1e-3)
    m = np.zeros(g.nparams, f)
    v = np.zeros(g.nparams, f)
    b1, b2, lr, eps = f(0.9), f(0.999), f(1e-3), f(1e-8)
    for step in (1, 2):
        n = 3000
        inp = rng.random((n, 9)).astype(f)
        d = rng.normal(size=(n, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f)
        w = rng.gamma(1.0, 1.0, n).astype(f)
        p0 = g.params()
        _, gr = g.grad(inp, d, w)
        g.train_step(inp, d, w)
        p1 = g.params()
        m = b1 * m + (f(1) - b1) * gr
        v = b2 * v + (f(1) - b2) * gr * gr
        c1 = f(1) - b1 ** f(step)
        c2 = f(1) - b2 ** f(step)
        want = p0 - lr * (m / c1) / (np.sqrt(v / c2) + eps)
        upd, upd_want = p1 - p0, want - p0
        moved = np.abs(upd_want) > 1e-6
        assert moved.mean() > 0.5
        rel = np.abs(upd[moved] - upd_want[moved]) / np.abs(upd_want[moved])
        _say(f"adam step {step}: max relative difference of the update {rel.max():.2e}")
        assert rel.max() < 1e-3
        assert np.abs(p1 - want).max() < 1e-6
    g.close()
