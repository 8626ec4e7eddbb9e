"""Surface BSDF of the oracle against what the mathematics fixes (no GPU).

The oracle's BSDF restates gpuupbp/kernel/bsdf.cuh:64-660 (component probabilities :143-177,
diffuse :227-249, Phong :251-284, mirror and glass sampling :455-559). The GPU's BSDF is tied to
the oracle's by the path-level render parity; these tests pin the oracle's own to closed forms:
* Phong normalisation: at normal incidence the lobe rho_s (n + 2) / (2 pi) cos^n integrates
  (against the cosine) to rho_s exactly, for any exponent; a diffuse lobe to rho_d;
* sample versus pdf: the density the sampler's directions follow is the pdf both report
  (E[cos / pdf] = pi over the hemisphere, a binned chi-square against the pdf), its weights
  f cos / pdf average to the directional albedo found by quadrature of the evaluation, and the
  reverse pdf of one direction pair is the forward pdf of the swapped pair;
* glass: Fresnel against the textbook (Hecht) amplitude coefficients, total internal reflection,
  refraction obeying Snell's law, event frequencies F and 1 - F, throughput weights 1 (mirror)
  and eta^2 (refraction on camera paths, radiance scaling) or 1 (light paths).
A dropped (n + 2), a pdf normalised over the wrong domain, a swapped forward / reverse pdf or a
Fresnel formula with the indices exchanged fails one of them.
"""
import numpy as np
import pytest

from oracle import oracle_py

DIFFUSE, PHONG, REFLECT, REFRACT = 1, 2, 4, 8


def _hemisphere_grid(n_mu=256, n_phi=512):
    """Gauss-Legendre nodes in mu = cos(theta) on [0, 1] times uniform phi: directions and weights
    (solid angle) of a quadrature over the upper hemisphere."""
    x, w = np.polynomial.legendre.leggauss(n_mu)
    mu = 0.5 * (x + 1.0)
    wm = 0.5 * w
    phi = (np.arange(n_phi) + 0.5) * (2 * np.pi / n_phi)
    M, P = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1.0 - M * M)
    d = np.stack([s * np.cos(P), s * np.sin(P), M], -1).reshape(-1, 3)
    wt = (wm[:, None] * np.full(n_phi, 2 * np.pi / n_phi)[None, :]).reshape(-1)
    return d, wt


def _bin_integrals(pdf, nb_mu, nb_phi, k=32):
    """Integral of pdf(directions) over each (cos theta, phi) cell of the upper hemisphere, with
    k x k Gauss-Legendre nodes inside every cell."""
    x, w = np.polynomial.legendre.leggauss(k)
    u, wu = 0.5 * (x + 1.0), 0.5 * w
    out = np.zeros((nb_mu, nb_phi))
    dmu, dphi = 1.0 / nb_mu, 2 * np.pi / nb_phi
    for i in range(nb_mu):
        mu = (i + u) * dmu
        for j in range(nb_phi):
            phi = (j + u) * dphi
            M, P = np.meshgrid(mu, phi, indexing="ij")
            s = np.sqrt(1 - M * M)
            d = np.stack([s * np.cos(P), s * np.sin(P), M], -1).reshape(-1, 3).astype(np.float32)
            out[i, j] = np.sum(pdf(d) * np.outer(wu, wu).reshape(-1)) * dmu * dphi
    return out


def _incoming(theta_deg, phi_deg=30.0):
    """Direction of travel of a ray arriving at the surface (normal +z) from polar angle theta."""
    t, p = np.radians(theta_deg), np.radians(phi_deg)
    return np.array([-np.sin(t) * np.cos(p), -np.sin(t) * np.sin(p), -np.cos(t)], np.float32)


@pytest.mark.parametrize("n", [1.0, 10.0, 100.0])
def test_phong_lobe_integrates_to_its_albedo_at_normal_incidence(n):
    d, wt = _hemisphere_grid()
    o = oracle_py.bsdf_eval(dict(phong=(0.5, 0.3, 0.7), exponent=n), _incoming(0.0), d)
    for c, rho in enumerate((0.5, 0.3, 0.7)):
        albedo = float(np.sum(o[:, c] * o[:, 3] * wt))
        assert abs(albedo - rho) < 2e-3 * rho, (n, c, albedo, rho)


def test_diffuse_lobe_integrates_to_its_albedo():
    d, wt = _hemisphere_grid()
    for theta in (0.0, 60.0):
        o = oracle_py.bsdf_eval(dict(diffuse=(0.6, 0.25, 0.9)), _incoming(theta), d)
        got = [float(np.sum(o[:, c] * o[:, 3] * wt)) for c in range(3)]
        assert np.allclose(got, [0.6, 0.25, 0.9], rtol=1e-4), got


@pytest.mark.parametrize("mat", [dict(diffuse=(0.3, 0.3, 0.3), phong=(0.4, 0.4, 0.4), exponent=20.0),
                                 dict(diffuse=(0.1, 0.1, 0.1), phong=(0.6, 0.6, 0.6), exponent=200.0)])
@pytest.mark.parametrize("theta", [0.0, 40.0, 75.0])
def test_sampler_follows_the_reported_pdf(mat, theta):
    rng = np.random.default_rng(7)
    n = 400000
    wi = _incoming(theta)
    s = oracle_py.bsdf_sample(mat, wi, rng.random((n, 3), np.float32))
    ok = (s[:, 5] > 0) & (s[:, 3] > 0)
    assert ok.mean() > 0.5
    # the pdf the sampler reports is the pdf the evaluation reports for that direction
    e = oracle_py.bsdf_eval(mat, wi, s[ok, :3])
    assert np.allclose(s[ok, 3], e[:, 4], rtol=1e-4)
    assert np.allclose(s[ok, 5:8], e[:, :3], rtol=1e-5)
    # E[cos / pdf] over the hemisphere is the integral of cos: pi (bounded integrand, the
    # diffuse component keeps pdf >= diff_prob cos / pi)
    est = np.where(ok, s[:, 4] / np.where(ok, s[:, 3], 1.0), 0.0)
    se = est.std() / np.sqrt(n)
    assert abs(est.mean() - np.pi) < 5 * se + 1e-3, (est.mean(), se)
    # weights f cos / pdf average to the directional albedo of the evaluation
    d, wt = _hemisphere_grid(512, 1024)
    q = oracle_py.bsdf_eval(mat, wi, d)
    albedo = float(np.sum(q[:, 0] * q[:, 3] * wt))
    wgt = np.where(ok, s[:, 5] * s[:, 4] / np.where(ok, s[:, 3], 1.0), 0.0)
    se = wgt.std() / np.sqrt(n)
    assert abs(wgt.mean() - albedo) < 5 * se + 2e-3 * albedo, (wgt.mean(), albedo, se)
    # binned chi-square of the directions against the pdf (cos theta x phi cells)
    nb_mu, nb_phi = 12, 16
    mu = s[ok, 2]
    phi = np.mod(np.arctan2(s[ok, 1], s[ok, 0]), 2 * np.pi)
    hist, _, _ = np.histogram2d(mu, phi, [nb_mu, nb_phi], [[0, 1], [0, 2 * np.pi]])
    expect = n * _bin_integrals(lambda dirs: oracle_py.bsdf_eval(mat, wi, dirs)[:, 4], nb_mu, nb_phi)
    m = expect > 20
    chi2 = float(np.sum((hist[m] - expect[m]) ** 2 / expect[m]))
    dof = int(m.sum()) - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


def test_reverse_pdf_is_the_forward_pdf_of_the_swapped_pair():
    mat = dict(diffuse=(0.3, 0.3, 0.3), phong=(0.4, 0.4, 0.4), exponent=15.0)
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[2], b[2] = abs(a[2]) + 0.1, abs(b[2]) + 0.1
        a, b = (a / np.linalg.norm(a)).astype(np.float32), (b / np.linalg.norm(b)).astype(np.float32)
        ab = oracle_py.bsdf_eval(mat, -a, b[None])[0]  # fixed a, generated b
        ba = oracle_py.bsdf_eval(mat, -b, a[None])[0]  # fixed b, generated a
        assert np.allclose(ab[:3], ba[:3], rtol=1e-5)  # reciprocity of diffuse + Phong
        assert abs(ab[5] - ba[4]) <= 1e-5 * max(ab[5], 1e-6)
        assert abs(ab[4] - ba[5]) <= 1e-5 * max(ab[4], 1e-6)


def _fresnel_textbook(cos_i, eta):
    """Unpolarised reflectance from the amplitude coefficients (Hecht, Optics, 4.34 / 4.40) with
    eta = n_i / n_t."""
    sin_t2 = eta * eta * (1 - cos_i * cos_i)
    if sin_t2 >= 1:
        return 1.0
    cos_t = np.sqrt(1 - sin_t2)
    rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("eta", [1 / 1.5, 1.5, 1 / 1.33])
def test_fresnel_against_the_textbook(eta):
    for c in np.linspace(0.02, 1.0, 40):
        assert abs(oracle_py.lib().oracle_fresnel(float(c), eta) - _fresnel_textbook(c, eta)) < 2e-6


@pytest.mark.parametrize("eta,theta", [(1 / 1.5, 30.0), (1 / 1.5, 70.0), (1.5, 20.0), (1.5, 50.0)])
@pytest.mark.parametrize("from_light", [False, True])
def test_glass_sampling(eta, theta, from_light):
    glass = dict(mirror=(1.0, 1.0, 1.0), ior=1.5)
    wi = _incoming(theta)
    rng = np.random.default_rng(11)
    n = 200000
    s = oracle_py.bsdf_sample(glass, wi, rng.random((n, 3), np.float32), ior=eta, from_light=from_light)
    F = _fresnel_textbook(np.cos(np.radians(theta)), eta)
    ev = s[:, 8].astype(int)
    refl = ev == REFLECT
    refr = (ev == REFRACT) & (s[:, 3] > 0)
    # reflection: the mirror direction, weight f cos / pdf = 1 (white mirror)
    r = s[refl]
    assert np.allclose(r[:, :3], [wi[0], wi[1], -wi[2]], atol=1e-6)
    assert np.allclose(r[:, 5] * r[:, 4] / r[:, 3], 1.0, rtol=1e-5)
    if F >= 1.0:  # total internal reflection
        assert refl.all()
        return
    # event frequencies F and 1 - F
    se = np.sqrt(F * (1 - F) / n)
    assert abs(refl.mean() - F) < 5 * se, (refl.mean(), F)
    assert abs(refr.mean() - (1 - F)) < 5 * se
    # refraction: Snell's law in the plane of incidence, weight eta^2 (camera) or 1 (light)
    t = s[refr]
    sin_i = np.sin(np.radians(theta))
    sin_t = np.linalg.norm(t[:, :2], axis=1)
    assert np.allclose(sin_t, eta * sin_i, rtol=1e-5)
    assert np.all(t[:, 2] < 0)  # through the surface
    assert np.allclose(t[:, :2] / sin_t[:, None], wi[:2] / np.linalg.norm(wi[:2]), atol=1e-5)
    w = t[:, 5] * t[:, 4] / t[:, 3]
    assert np.allclose(w, 1.0 if from_light else eta * eta, rtol=1e-5)
