"""FP64 restatement of the guiding network (test infrastructure): frequency
encoding, MLP with BF16 rounding of weights and activations where the device
rounds them, the 8-lobe vMF mixture, its loss and the backward pass. Pinned in
tests by finite differences and closed forms (vMF normalisation)."""
import numpy as np

IN, FREQ, FEAT, K1, H, LOBES, NOUT = 9, 12, 216, 224, 64, 8, 40
LAYERS = [(K1, H), (H, H), (H, H), (H, 48)]
LK_MIN, LK_MAX = -5.0, 9.0


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    out, o = [], 0
    for i, n in LAYERS:
        W = params[o:o + i * n].reshape(n, i)
        o += i * n
        b = params[o:o + n]
        o += n
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def sincos_f32(x):
    """Sine and cosine in float32 arithmetic, step for step the device's
    tm_sincos (and the C oracle's o_sincos): Cody-Waite reduction by pi/2 with a
    three-part constant, Taylor polynomials of sin (to r^9) and cos (to r^10)."""
    f = np.float32
    x = np.asarray(x, np.float32)
    k = np.rint(x * f(float.fromhex("0x1.45f306p-1"))).astype(np.float32)
    r = x - k * f(float.fromhex("0x1.92p+0"))
    r = r - k * f(float.fromhex("0x1.fb4p-12"))
    r = r - k * f(float.fromhex("0x1.4442d2p-24"))
    r2 = r * r
    ps = f(-float.fromhex("0x1.555556p-3")) + r2 * (f(float.fromhex("0x1.111112p-7")) + r2 * (f(-float.fromhex("0x1.a01a02p-13")) + r2 * f(float.fromhex("0x1.71de3ap-19"))))
    sn = r + r * r2 * ps
    pc = f(float.fromhex("0x1.555556p-5")) + r2 * (f(-float.fromhex("0x1.6c16c2p-10")) + r2 * (f(float.fromhex("0x1.a01a02p-16")) + r2 * f(-float.fromhex("0x1.27e4fcp-22"))))
    cs = f(1.0) - f(0.5) * r2 + r2 * r2 * pc
    q = k.astype(np.int64) & 3
    s_out = np.select([q == 0, q == 1, q == 2], [sn, cs, -sn], -cs).astype(np.float32)
    c_out = np.select([q == 0, q == 1, q == 2], [cs, -sn, -cs], sn).astype(np.float32)
    return s_out, c_out


def features_f32(inp):
    """Frequency features before BF16 rounding, [n][216] float32: the base pair
    sincos(pi x) then the double-angle recurrences s' = (2 s) c, c' = (c - s)(c + s)
    in float32 (guide.hip freq_base / freq_next)."""
    f = np.float32
    x = np.asarray(inp, np.float32)
    out = np.zeros((len(x), FEAT), np.float32)
    for c in range(IN):
        s, co = sincos_f32(x[:, c] * f(3.14159265))
        for q in range(FREQ):
            out[:, c * 2 * FREQ + 2 * q] = s
            out[:, c * 2 * FREQ + 2 * q + 1] = co
            s, co = ((f(2.0) * s) * co).astype(np.float32), ((co - s) * (co + s)).astype(np.float32)
    return out


def encode(inp):
    feats = np.zeros((len(inp), K1))
    feats[:, :FEAT] = features_f32(inp)
    return bf16(feats)


def forward(params, inp):
    L = unpack(params)
    X = encode(inp)
    hs = [X]
    h = X
    for k, (W, b) in enumerate(L):
        z = h @ bf16(W).T + b.astype(np.float32)
        if k < 3:
            h = bf16(np.maximum(z, 0))
            hs.append(h)
        else:
            raw = z
    return raw[:, :NOUT], hs


def decode(raw):
    m = raw[:, :24].reshape(-1, LOBES, 3)
    ln = np.linalg.norm(m, axis=2) + 1e-6
    mu = m / ln[..., None]
    kap = np.exp(np.clip(raw[:, 24:32], LK_MIN, LK_MAX))
    lg = raw[:, 32:40]
    w = np.exp(lg - lg.max(1, keepdims=True))
    w /= w.sum(1, keepdims=True)
    return mu, kap, w, ln


def vmf_log(kap, c):
    return np.log(kap) - np.log(2 * np.pi) - np.log1p(-np.exp(-2 * kap)) + kap * (c - 1)


def mixture_logpdf(raw, d):
    mu, kap, w, _ = decode(raw)
    c = np.einsum("nkj,nj->nk", mu, d)
    lv = np.log(w) + vmf_log(kap, c)
    mx = lv.max(1, keepdims=True)
    return (mx + np.log(np.exp(lv - mx).sum(1, keepdims=True)))[:, 0], lv, c


def loss_and_draw(raw, d, wts):
    logq, lv, c = mixture_logpdf(raw, d)
    scale = wts / wts.sum()
    loss = float((scale * -logq).sum())
    mu, kap, w, ln = decode(raw)
    resp = np.exp(lv - logq[:, None])
    g = np.zeros((len(raw), 64))
    g[:, 32:40] = scale[:, None] * (w - resp)
    dlv = 1 / kap - 2 / np.expm1(2 * kap) + (c - 1)
    lk = raw[:, 24:32]
    inside = (lk > LK_MIN) & (lk < LK_MAX)
    g[:, 24:32] = np.where(inside, -scale[:, None] * resp * dlv * kap, 0.0)
    f = -scale[:, None] * resp * kap / ln
    g[:, :24] = (f[..., None] * (d[:, None, :] - c[..., None] * mu)).reshape(-1, 24)
    return loss, g


def grad(params, inp, d, wts):
    L = unpack(params)
    raw, hs = forward(params, inp)
    loss, g = loss_and_draw(raw, np.asarray(d, np.float64), np.asarray(wts, np.float64))
    dz = bf16(g)[:, :48]
    grads = [None] * 4
    for k in range(3, -1, -1):
        W, b = L[k]
        grads[k] = (hs[k].T @ dz, dz.sum(0))          # dW [in][out], db [out]
        if k > 0:
            dh = dz @ bf16(W)
            dz = bf16(np.where(hs[k] > 0, dh, 0.0))
    flat = []
    for (dW, db) in grads:
        flat.append(dW.T.reshape(-1))
        flat.append(db)
    return loss, np.concatenate(flat)
