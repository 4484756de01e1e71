"""FP64 restatement of the path-mode BSDF and environment densities (test
infrastructure). These parts have no reference counterpart; the formulas are the
published ones:
  - Lambert: f = albedo / pi, pdf = cos / pi
  - GGX (Walter et al. 2007) with alpha = roughness^2, height-correlated Smith
    masking (Heitz 2014), visible-normal sampling density G1(wo) D(h) / (4 wo.z)
    (Heitz 2018), Schlick Fresnel with F0 = albedo
  - smooth dielectric: exact unpolarised Fresnel, Snell refraction
  - equirectangular env: texel probability w_ij / sum w, solid-angle density
    p_ij W H / (2 pi^2 sin theta)
Pinned in tests/test_oracle_bsdf.py by closed forms (normalisation of D, G1 = 1 at
normal incidence, Fresnel at normal incidence and at the critical angle)."""
import numpy as np


def ggx_alpha(rough):
    return max(rough * rough, 1e-4)


def ggx_D(h, al):
    h = np.asarray(h, np.float64)
    x, y, z = h[..., 0] / al, h[..., 1] / al, h[..., 2]
    q = x * x + y * y + z * z
    return np.where(z > 0, 1.0 / (np.pi * al * al * q * q), 0.0)


def ggx_lambda(w, al):
    z = np.asarray(w, np.float64)[..., 2]
    z2 = z * z
    t2 = np.where(z2 > 0, np.maximum(0.0, 1 - z2) / np.maximum(z2, 1e-300), 0.0)
    return 0.5 * (-1 + np.sqrt(1 + al * al * t2))


def ggx_G1(w, al):
    return 1.0 / (1.0 + ggx_lambda(w, al))


def ggx_G2(wo, wi, al):
    return 1.0 / (1.0 + ggx_lambda(wo, al) + ggx_lambda(wi, al))


def schlick(f0, c):
    m = np.clip(1 - c, 0, 1)
    return np.asarray(f0)[None, :] + (1 - np.asarray(f0)[None, :]) * (m ** 5)[..., None]


def conductor_eval(wo, wi, rough, f0):
    """(f cos, pdf) for local wo (3,), wi (n, 3)."""
    al = ggx_alpha(rough)
    wo = np.asarray(wo, np.float64)
    wi = np.asarray(wi, np.float64)
    h = wo[None, :] + wi
    h = h / np.linalg.norm(h, axis=1, keepdims=True)
    D = ggx_D(h, al)
    G2 = ggx_G2(np.broadcast_to(wo, wi.shape), wi, al)
    G1 = ggx_G1(wo[None, :], al)
    F = schlick(f0, (h * wo[None, :]).sum(1))
    ok = (wi[:, 2] > 0) & (wo[2] > 0)
    fcos = np.where(ok[:, None], F * (D * G2 / (4 * wo[2]))[:, None], 0.0)
    pdf = np.where(ok, G1 * D / (4 * wo[2]), 0.0)
    return fcos, pdf


def fresnel_dielectric(cosi, eta):
    """Unpolarised reflectance, eta = eta_i / eta_t."""
    sin2t = eta * eta * max(0.0, 1 - cosi * cosi)
    if sin2t >= 1:
        return 1.0
    cost = np.sqrt(1 - sin2t)
    rs = (eta * cosi - cost) / (eta * cosi + cost)
    rp = (cosi - eta * cost) / (cosi + eta * cost)
    return 0.5 * (rs * rs + rp * rp)


def env_texel_probs(env):
    """Texel probabilities of the env distribution (luminance x sin theta)."""
    e = np.asarray(env, np.float64)
    h = e.shape[0]
    lum = 0.2126 * e[..., 0] + 0.7152 * e[..., 1] + 0.0722 * e[..., 2]
    w = np.maximum(lum, 0) * np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    return w / w.sum()


def sphere_bin_probs(pdf_fn, nth, nph, sub=24):
    """Probability of each (cos theta, phi) bin of the upper hemisphere under a
    solid-angle density, by midpoint integration with sub x sub points per bin."""
    out = np.zeros((nth, nph))
    for a in range(nth):
        c = 1 - (a + (np.arange(sub) + 0.5) / sub) / nth        # cos theta, equal-area bins
        for b in range(nph):
            ph = 2 * np.pi * (b + (np.arange(sub) + 0.5) / sub) / nph
            C, PH = np.meshgrid(c, ph, indexing="ij")
            S = np.sqrt(np.maximum(0, 1 - C * C))
            d = np.stack([S * np.cos(PH), S * np.sin(PH), C], -1).reshape(-1, 3)
            out[a, b] = pdf_fn(d).mean() * (2 * np.pi / nth / nph)
    return out
