"""The FP64 BSDF / env restatement pinned by closed forms."""
import numpy as np
import pytest


@pytest.mark.parametrize("rough", [0.05, 0.2, 0.5, 1.0])
def test_ggx_D_projected_area_is_one(oracle, rough):
    from oracle import bsdf_ref as B
    al = B.ggx_alpha(rough)
    # integral of D(h) cos(theta_h) over the hemisphere = 1; isotropic, so
    # 2 pi * int_0^1 D(c) c dc, midpoint rule in u with c = 1 - u^4 (dense near c = 1)
    n = 2_000_000
    u = (np.arange(n) + 0.5) / n
    c = 1 - u ** 4
    jac = 4 * u ** 3
    s = np.sqrt(1 - c * c)
    h = np.stack([s, np.zeros_like(s), c], 1)
    est = 2 * np.pi * (B.ggx_D(h, al) * c * jac).mean()
    assert abs(est - 1) < 1e-3


def test_ggx_masking_closed_forms(oracle):
    from oracle import bsdf_ref as B
    assert B.ggx_G1(np.array([0, 0, 1.0]), 0.3) == 1.0
    # Lambda of the GGX distribution: (-1 + sqrt(1 + a^2 tan^2)) / 2
    th = np.radians(60)
    w = np.array([np.sin(th), 0, np.cos(th)])
    assert abs(B.ggx_lambda(w, 0.5) - 0.5 * (-1 + np.sqrt(1 + 0.25 * np.tan(th) ** 2))) < 1e-15


def test_fresnel_closed_forms(oracle):
    from oracle import bsdf_ref as B
    for n in (1.33, 1.5, 2.4):
        assert abs(B.fresnel_dielectric(1.0, 1 / n) - ((n - 1) / (n + 1)) ** 2) < 1e-15
    crit = np.arcsin(1 / 1.5)
    assert B.fresnel_dielectric(np.cos(crit) - 1e-9, 1.5) == 1.0
    assert B.fresnel_dielectric(np.cos(crit) + 1e-6, 1.5) < 1.0


def test_env_probs_constant_map_uniform_solid_angle(oracle):
    from oracle import bsdf_ref as B
    h, w = 32, 64
    p = B.env_texel_probs(np.ones((h, w, 3), np.float32))
    th = np.pi * (np.arange(h) + 0.5) / h
    dens = p * w * h / (2 * np.pi ** 2 * np.sin(th)[:, None])   # solid-angle density
    assert np.allclose(dens, dens.mean(), rtol=0.02) and abs(dens.mean() - 1 / (4 * np.pi)) < 2e-3


@pytest.mark.parametrize("rough", [0.1, 0.4, 0.9])
def test_conductor_reciprocity_energy_and_pdf(oracle, rough):
    """The FP64 GGX conductor the GPU sampler is checked against: f(wo, wi) = f(wi, wo)
    (reciprocity), hemispherical albedo <= 1 and close to 1 for a smooth white
    conductor (single scattering loses energy only at high roughness), and the VNDF
    density integrating to at most 1 (the part below the horizon is lost)."""
    from oracle import bsdf_ref as B
    rng = np.random.default_rng(11)

    def hemi(k):
        v = rng.normal(size=(k, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        v[:, 2] = np.abs(v[:, 2]) + 1e-3
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    f0 = (0.9, 0.7, 0.4)
    a, b = hemi(300), hemi(300)
    for wo, wi in zip(a, b):
        f_ab = B.conductor_eval(wo, wi[None], rough, f0)[0][0] / wi[2]
        f_ba = B.conductor_eval(wi, wo[None], rough, f0)[0][0] / wo[2]
        assert np.allclose(f_ab, f_ba, rtol=1e-12, atol=0)
    # albedo and density integrals by uniform hemisphere sampling (pdf 1 / 2 pi)
    n = 400_000
    wi = hemi(n)
    for theta in (0.0, 45.0, 75.0):
        t = np.radians(theta)
        wo = np.array([np.sin(t), 0.0, np.cos(t)])
        fcos, pdf = B.conductor_eval(wo, wi, rough, (1.0, 1.0, 1.0))
        est = 2 * np.pi * fcos[:, 0]
        albedo, se = est.mean(), est.std() / np.sqrt(n)
        assert albedo <= 1 + 4 * se
        if rough <= 0.1 and theta <= 45:
            assert albedo > 0.97
        dens = 2 * np.pi * pdf
        assert dens.mean() <= 1 + 4 * dens.std() / np.sqrt(n)
