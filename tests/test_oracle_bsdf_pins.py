"""Pins of the oracle's GGX, dielectric and Russian-roulette restatements by
relations the mathematics fixes, independent of how the oracle writes them:

  GlossyBxDF (glossy_bxdf.h:29-72): reciprocity of eval * cos(theta_o); eval and pdf
    equal closed forms built from the textbook GGX D and Smith Lambda
  MicrofacetDistribution::lambda (microfacet_distribution.h:137-145): it takes
    |cos/sin| where the textbook takes tan, i.e. it is the textbook Lambda at the
    complementary angle
  MicrofacetDistribution::sample (microfacet_distribution.h:13-58, 96-118) at oblique
    incidence: a chi-square test of the sampled half vectors against the visible
    normal density G1(wo) max(0, wo.wh) D(wh) / cos(theta_o) (Heitz and d'Eon 2014)
  RefractionBxDF::sample (refraction_bxdf.h:26-52): Snell's law on both sides,
    coplanarity, total internal reflection exactly past the critical angle
  russian_roulette (pathtracerMIS.h:24-27): survival values worked out by hand
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle


def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t = np.cross(n, rng.normal(size=3))
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    return t, b, n


def _upper(rng, n, t, b):
    """A random direction in the hemisphere of n, away from grazing."""
    ct = rng.uniform(0.1, 1.0)
    ph = rng.uniform(0, 2 * math.pi)
    st = math.sqrt(1 - ct * ct)
    return (math.cos(ph) * st) * t + (math.sin(ph) * st) * b + ct * n


def _ggx_d(alpha, cos_h):
    """Textbook isotropic GGX (Trowbridge-Reitz) normal distribution."""
    c2 = cos_h * cos_h
    t2 = (1 - c2) / c2
    return 1.0 / (math.pi * alpha * alpha * c2 * c2 * (1 + t2 / (alpha * alpha)) ** 2)


def _lambda_textbook(alpha, theta):
    """Smith Lambda of GGX: (-1 + sqrt(1 + alpha^2 tan^2 theta)) / 2."""
    return (-1 + math.sqrt(1 + (alpha * math.tan(theta)) ** 2)) / 2


@pytest.mark.parametrize("rough", [0.15, 0.3, 0.6, 0.9])
def test_glossy_eval_reciprocity(rough):
    """eval folds |n.wi| into the BRDF (eval = f |n.wi|), so eval(wo, wi) |n.wo|
    = f |n.wi| |n.wo| must be symmetric in wo and wi."""
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(300):
        t, b, n = _frame(rng)
        wo, wi = _upper(rng, n, t, b), _upper(rng, n, t, b)
        a = oracle.glossy_eval(t, b, n, rough, wo, wi) * float(np.dot(n, wo))
        c = oracle.glossy_eval(t, b, n, rough, wi, wo) * float(np.dot(n, wi))
        worst = max(worst, abs(a - c) / max(abs(a), 1e-20))
    assert worst < 2e-5, worst


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.7])
def test_lambda_is_textbook_at_complementary_angle(alpha):
    rng = np.random.default_rng(5)
    for theta in np.radians([3, 10, 25, 45, 60, 75, 87]):
        ph = rng.uniform(0, 2 * math.pi)
        w = (math.cos(ph) * math.sin(theta), math.sin(ph) * math.sin(theta), math.cos(theta))
        want = _lambda_textbook(alpha, math.pi / 2 - theta)
        got = oracle.ggx_lambda(alpha, w)
        # float directions: sin theta from a rounded cos theta near the pole
        assert abs(got - want) <= 3e-5 * max(1.0, want), (math.degrees(theta), got, want)
        assert abs(oracle.ggx_g1(alpha, w) - 1 / (1 + want)) < 3e-5
    # at normal incidence the reference's |cos/sin| is infinite and Lambda returns 0
    assert oracle.ggx_lambda(alpha, (0.0, 0.0, 1.0)) == 0.0


@pytest.mark.parametrize("rough", [0.2, 0.5, 0.8])
def test_glossy_eval_and_pdf_closed_forms(rough):
    """eval = G D / (4 cos_o) and pdf = D G1(wo) / (4 cos_o) with the textbook D, and
    G = 1 / (1 + Lambda(wo) + Lambda(wi)), G1 = 1 / (1 + Lambda(wo)), Lambda taken at
    the complementary angle (alpha = max(rough^2, 0.01), glossy_bxdf.h:77-79)."""
    alpha = max(rough * rough, 0.01)
    rng = np.random.default_rng(7)
    t, b, n = np.eye(3)[0], np.eye(3)[1], np.eye(3)[2]
    for _ in range(200):
        wo, wi = _upper(rng, n, t, b), _upper(rng, n, t, b)
        wh = (wo + wi) / np.linalg.norm(wo + wi)
        lo = _lambda_textbook(alpha, math.pi / 2 - math.acos(wo[2]))
        li = _lambda_textbook(alpha, math.pi / 2 - math.acos(wi[2]))
        d = _ggx_d(alpha, wh[2])
        ev = d / (1 + lo + li) / (4 * wo[2])
        pdf = d / (1 + lo) / (4 * wo[2])
        assert abs(oracle.glossy_eval(t, b, n, rough, wo, wi) - ev) <= 1e-4 * ev + 1e-7
        assert abs(oracle.glossy_pdf(t, b, n, rough, wo, wi) - pdf) <= 1e-4 * pdf + 1e-7


def _vndf_bin_probs(alpha, wo, nth, nphi, sub=16):
    """Probability of each (theta_h, phi_h) bin (theta_h in [0, pi/2) in nth equal
    steps) under the visible normal density Dwo(wh) = G1(wo) max(0, wo.wh) D(wh) /
    cos theta_o with the textbook G1; midpoint rule on sub x sub points per bin,
    d omega = sin theta d theta d phi."""
    th_o = math.acos(wo[2])
    g1 = 1 / (1 + _lambda_textbook(alpha, th_o))
    te = (np.arange(nth * sub) + 0.5) / (nth * sub) * (math.pi / 2)
    pe = (np.arange(nphi * sub) + 0.5) / (nphi * sub) * 2 * math.pi - math.pi
    T, P = np.meshgrid(te, pe, indexing="ij")
    C, S = np.cos(T), np.sin(T)
    dot = np.maximum(0, wo[0] * S * np.cos(P) + wo[1] * S * np.sin(P) + wo[2] * C)
    c2 = C * C
    D = 1.0 / (math.pi * alpha ** 2 * c2 * c2 * (1 + (1 - c2) / c2 / alpha ** 2) ** 2)
    dens = g1 * dot * D / wo[2] * S
    cell = (math.pi / 2 / (nth * sub)) * (2 * math.pi / (nphi * sub))
    return (dens * cell).reshape(nth, sub, nphi, sub).sum(axis=(1, 3))


@pytest.mark.parametrize("alpha", [0.2, 0.5])
@pytest.mark.parametrize("theta_o", [30, 60, 80])
def test_ggx_visible_normal_sampling_oblique(alpha, theta_o):
    th = math.radians(theta_o)
    wo = (math.sin(th) * math.cos(0.7), math.sin(th) * math.sin(0.7), math.cos(th))
    nth, nphi = 36, 24
    p = _vndf_bin_probs(alpha, wo, nth, nphi)
    assert abs(p.sum() - 1) < 2e-3  # the density integrates to 1 (a property of the VNDF)
    n = 400_000
    u = np.random.default_rng(theta_o + int(alpha * 100)).random((n, 2), dtype=np.float32)
    wh = oracle.ggx_sample_batch(alpha, wo, u).astype(np.float64)
    wh /= np.linalg.norm(wh, axis=1, keepdims=True)
    assert (wh[:, 2] > 0).all()
    ti = np.minimum((np.arccos(np.clip(wh[:, 2], -1, 1)) / (math.pi / 2) * nth).astype(int), nth - 1)
    pi = np.minimum(((np.arctan2(wh[:, 1], wh[:, 0]) + math.pi) / (2 * math.pi) * nphi).astype(int), nphi - 1)
    obs = np.bincount(ti * nphi + pi, minlength=nth * nphi).astype(float)
    exp = p.ravel() / p.sum() * n
    # bins of at least 200 expected samples: in the sparse far tail the PBRT rational
    # fit of the slope inverse (sample11's slope_y) departs by a few percent, which
    # the tail-mass check below bounds instead
    ok = exp >= 200
    chi2 = float((((obs - exp) ** 2) / exp)[ok].sum())
    dof = int(ok.sum()) - 1
    pval = stats.chi2.sf(chi2, dof)
    print(f"alpha {alpha} theta_o {theta_o}: chi2 {chi2:.1f} / {dof} dof, p {pval:.3g}, "
          f"mass outside the tested bins {obs[~ok].sum() / n:.2e} (expected {exp[~ok].sum() / n:.2e})")
    assert pval > 1e-4
    tail = exp[~ok].sum()
    assert abs(obs[~ok].sum() - tail) <= 5 * math.sqrt(tail + 1) + 0.03 * tail


def _tangent_dir(n, w):
    v = w - np.dot(w, n) * n
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("ior", [1.33, 1.5, 2.4])
def test_dielectric_snell_both_sides(ior):
    rng = np.random.default_rng(int(ior * 100))
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        if abs(np.dot(w, n)) < 0.05:
            continue
        wi = oracle.dielectric_sample(n, w, ior).astype(np.float64)
        assert abs(np.linalg.norm(wi) - 1) < 1e-5
        cos_i = float(np.dot(w, n))
        sin_i = math.sqrt(max(0.0, 1 - cos_i * cos_i))
        outside = cos_i > 0
        eta_i, eta_t = (1.0, ior) if outside else (ior, 1.0)
        if not outside and ior * sin_i > 1.0 + 1e-5:
            # total internal reflection: mirror about the normal, staying inside
            want = 2 * np.dot(n, w) * n - w
            assert np.abs(wi - want).max() < 1e-5
            continue
        if not outside and ior * sin_i > 1.0 - 1e-5:
            continue  # at the critical angle within rounding
        # transmitted to the other side, in the plane of n and wo, on the far side of n
        assert np.dot(wi, n) * cos_i < 0
        sin_t = math.sqrt(max(0.0, 1 - float(np.dot(wi, n)) ** 2))
        assert abs(eta_i * sin_i - eta_t * sin_t) < 2e-5 * max(1.0, ior)
        assert abs(np.dot(wi, np.cross(n, w))) < 2e-5
        if sin_i > 1e-3:
            assert np.dot(_tangent_dir(n, wi), _tangent_dir(n, w)) < -0.999


def test_dielectric_critical_angle_and_normal_incidence():
    ior = 1.5
    n = np.array([0.0, 0.0, 1.0])
    crit = math.asin(1 / ior)
    for d, tir in ((-2e-4, False), (2e-4, True)):
        th = crit + d
        w = np.array([math.sin(th), 0.0, -math.cos(th)])  # inside the medium (against n)
        wi = oracle.dielectric_sample(n, w, ior)
        if tir:
            assert np.abs(wi - (2 * np.dot(n, w) * n - w)).max() < 1e-5
        else:
            assert wi[2] > 0 and abs(wi[0]) > 0.99  # leaves close to grazing
    for w in (np.array([0.0, 0.0, 1.0]), np.array([0.0, 0.0, -1.0])):
        wi = oracle.dielectric_sample(n, w, ior)
        assert np.abs(wi + w).max() < 1e-6  # straight through


def test_russian_roulette_known_values():
    # q = |0.2 r + 0.5 g + 0.3 b| * 0.5 + 0.4, worked out by hand
    cases = [((1, 0, 0), 0.5), ((0, 1, 0), 0.65), ((0, 0, 1), 0.55), ((0, 0, 0), 0.4), ((1, 1, 1), 0.9),
             ((0, 4, 0), 1.4), ((-1, 0, 0), 0.5), ((2, -1, 0), 0.45), ((3, 3, 3), 1.9), ((0.5, 0.2, 0.1), 0.515)]
    for c, q in cases:
        assert abs(oracle.russian_roulette(c) - q) < 1e-6, (c, q)
        # the corrected mode clamps the survival probability at 1 and changes nothing else
        assert abs(oracle.russian_roulette(c, corrected=True) - min(q, 1.0)) < 1e-6, (c, q)
