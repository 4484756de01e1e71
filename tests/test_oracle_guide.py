"""The FP64 guiding-network restatement pinned by finite differences and closed forms."""
import numpy as np


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    inp = rng.random((n, 9)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.gamma(1.0, 1.0, n)
    return inp, d, w


def test_vmf_density_normalised():
    from oracle import guide_ref as G
    rng = np.random.default_rng(0)
    d = rng.normal(size=(400000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for kap in (0.01, 1.0, 10.0):
        # E over uniform directions of pdf * 4 pi == 1
        est = np.exp(G.vmf_log(np.full(len(d), kap), d[:, 2])).mean() * 4 * np.pi
        assert abs(est - 1) < 0.03


def test_loss_gradient_matches_finite_differences():
    """d loss / d raw from the closed-form mixture gradient vs central differences
    (FP64, no rounding)."""
    from oracle import guide_ref as G
    rng = np.random.default_rng(1)
    n = 6
    raw = rng.normal(size=(n, 40)) * 0.7
    _, d, w = _rand(n, 2)
    _, g = G.loss_and_draw(raw, d, w)
    eps = 1e-6
    for j in [0, 5, 13, 24, 27, 31, 32, 36, 39]:
        for r in range(n):
            rp, rm = raw.copy(), raw.copy()
            rp[r, j] += eps
            rm[r, j] -= eps
            num = (G.loss_and_draw(rp, d, w)[0] - G.loss_and_draw(rm, d, w)[0]) / (2 * eps)
            assert abs(num - g[r, j]) < 1e-6 + 1e-5 * abs(num)


def test_network_gradient_matches_finite_differences():
    """Backward pass of the restatement vs finite differences of its loss, with the
    BF16 rounding switched off (identity) so the function is smooth."""
    from oracle import guide_ref as G
    saved = G.bf16
    G.bf16 = lambda x: np.asarray(x, np.float64)
    try:
        rng = np.random.default_rng(3)
        npar = sum(i * o + o for i, o in G.LAYERS)
        params = rng.normal(size=npar) * 0.1
        inp, d, w = _rand(16, 4)
        loss, g = G.grad(params, inp, d, w)

        def f(p):
            raw, _ = G.forward(p, inp)
            return G.loss_and_draw(raw, d, w)[0]
        for idx in rng.choice(npar, 25, replace=False):
            e = np.zeros(npar)
            e[idx] = 1e-6
            num = (f(params + e) - f(params - e)) / 2e-6
            assert abs(num - g[idx]) < 1e-7 + 1e-4 * abs(num)
    finally:
        G.bf16 = saved


def test_sincos_f32_bit_identical_to_c_oracle(oracle):
    """The numpy float32 restatement of the device sine / cosine equals the C
    oracle's o_sinf / o_cosf bit for bit (the C functions are pinned to the device
    in the GPU math test)."""
    from oracle import guide_ref as G
    rng = np.random.default_rng(3)
    x = np.concatenate([np.linspace(0, 3.2, 20001, dtype=np.float32), rng.uniform(0, 3.2, 20000).astype(np.float32)])
    s, c = G.sincos_f32(x)
    want_s = np.array([oracle.lib.o_sinf(float(v)) for v in x], np.float32)
    want_c = np.array([oracle.lib.o_cosf(float(v)) for v in x], np.float32)
    assert np.array_equal(s.view(np.int32), want_s.view(np.int32))
    assert np.array_equal(c.view(np.int32), want_c.view(np.int32))


def test_frequency_features_close_to_exact():
    """Features from the base sincos and the doubling recurrences stay within the
    error that doubling allows of sin / cos(2^f pi x) evaluated in FP64 (the input
    product x * pi is rounded to float32 first, which alone is worth up to
    2^f * 2.4e-7 at f)."""
    from oracle import guide_ref as G
    x = np.random.default_rng(4).uniform(0, 1, (4000, 9)).astype(np.float32)
    F = G.features_f32(x).reshape(4000, 9, 12, 2)
    for f in range(12):
        a = x.astype(np.float64) * np.pi * 2.0 ** f
        err = max(np.abs(F[:, :, f, 0] - np.sin(a)).max(), np.abs(F[:, :, f, 1] - np.cos(a)).max())
        assert err < 4e-7 * 2.0 ** f + 1e-6, (f, err)
