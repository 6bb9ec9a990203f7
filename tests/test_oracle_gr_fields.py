"""Pins the oracle's GR field solver (CPU only).

* sqrt(det h) equals the square root of the determinant of the lower metric
  (pinned against the textbook Kerr-Schild line element in test_oracle_gr.py),
  sqrt(det h)~ is it over sin(theta), and the polar area is the integral of
  sqrt(det h) over the half cell next to the pole (exact for a = 0, where the
  metric factors do not depend on theta).
* Faraday leaves the discrete divergence of sqrt(det h) B unchanged, Ampere
  that of sqrt(det h) D (the discrete curl's divergence vanishes), whatever
  the fields; a Faraday with a wrong metric factor breaks it.
* Wald's solution (pgens/wald/pgen.hpp:41-223: B and D from the vector
  potential A_0, A_1, A_3 of a uniform field aligned with the spin) is a
  stationary vacuum solution: E and H from the auxiliary kernels give Faraday
  and Ampere rates that vanish with the grid at second order.  Dropping
  beta x B from E, or flipping the beta x D term of H, leaves an O(1) rate.
"""
import math

import numpy as np
import pytest

G = 2


def _prm(oracle, kind, n, a, rmax=12.0, h=0.0):
    rh = 1 + math.sqrt(1 - a * a) if kind != "kerr_schild_0" else 1.0
    rmin = 0.95 * rh if kind != "kerr_schild_0" else 1.0
    return oracle.gr_params(kind, n, ((rmin, rmax), (0.0, math.pi)), a=a, h=h)


def _metric(oracle, prm, x1, x2):
    x1, x2 = np.broadcast_arrays(np.asarray(x1, float), np.asarray(x2, float))
    q = oracle.gr_metric(prm, x1.ravel(), x2.ravel())
    return {k: v.reshape(x1.shape) for k, v in q.items()}


METRICS = [("kerr_schild", 0.9, 0.0), ("qkerr_schild", 0.9, 0.0), ("qkerr_schild", 0.6, 0.3),
           ("kerr_schild_0", 0.0, 0.0)]


@pytest.mark.parametrize("kind,a,h", METRICS)
def test_sqrt_det_h(oracle, kind, a, h):
    prm = _prm(oracle, kind, (40, 32), a, h=h)
    rng = np.random.default_rng(1)
    x1, x2 = rng.uniform(0, 40, 200), rng.uniform(0.5, 31.5, 200)
    q = _metric(oracle, prm, x1, x2)
    det = q["h_22"] * (q["h_11"] * q["h_33"] - q["h_13"] ** 2)
    np.testing.assert_allclose(q["sqrt_det_h"], np.sqrt(det), rtol=1e-12)
    np.testing.assert_allclose(q["sqrt_det_h_tilde"] * np.sin(q["theta"]), q["sqrt_det_h"],
                               rtol=1e-12)


@pytest.mark.parametrize("kind,h", [("kerr_schild", 0.0), ("qkerr_schild", 0.0),
                                    ("qkerr_schild", 0.3), ("kerr_schild_0", 0.0)])
def test_polar_area_is_the_polar_half_cell(oracle, kind, h):
    prm = _prm(oracle, kind, (40, 32), 0.0, h=h)
    x1 = np.array([0.5, 7.5, 33.5])
    # Simpson over x2 in [0, 1/2] of sqrt(det h)
    s = np.linspace(0.0, 0.5, 2001)
    X1, S = np.meshgrid(x1, s, indexing="ij")
    f = _metric(oracle, prm, X1, S)["sqrt_det_h"]
    w = np.ones(len(s))
    w[1:-1:2], w[2:-1:2] = 4, 2
    integral = (f * w).sum(1) * (s[1] - s[0]) / 3
    area = _metric(oracle, prm, x1, np.zeros(3))["polar_area"]
    np.testing.assert_allclose(area, integral, rtol=1e-10)


def _fields(n, ncomp, rng):
    return rng.standard_normal((ncomp, n[1] + 2 * G, n[0] + 2 * G))


def _div_B(oracle, prm, B, n):
    """div of sqrt(det h) B at cell centres (i + 1/2, j + 1/2), active cells."""
    i = np.arange(n[0] + 2 * G)[None, :] - G
    j = np.arange(n[1] + 2 * G)[:, None] - G
    s1 = _metric(oracle, prm, i + 0.0 * j, j + 0.5)["sqrt_det_h"] * B[3]      # B^1 at (i, j+1/2)
    s2 = _metric(oracle, prm, i + 0.5, j + 0.0 * i)["sqrt_det_h"] * B[4]      # B^2 at (i+1/2, j)
    return (s1[:-1, 1:] - s1[:-1, :-1]) + (s2[1:, :-1] - s2[:-1, :-1])


def _div_D(oracle, prm, D, n):
    """div of sqrt(det h) D at nodes (i, j)."""
    i = np.arange(n[0] + 2 * G)[None, :] - G
    j = np.arange(n[1] + 2 * G)[:, None] - G
    s1 = _metric(oracle, prm, i + 0.5, j + 0.0 * i)["sqrt_det_h"] * D[0]      # D^1 at (i+1/2, j)
    s2 = _metric(oracle, prm, i + 0.0 * j, j + 0.5)["sqrt_det_h"] * D[1]      # D^2 at (i, j+1/2)
    return (s1[1:, 1:] - s1[1:, :-1]) + (s2[1:, 1:] - s2[:-1, 1:])


@pytest.mark.parametrize("kind,a,h", METRICS)
def test_faraday_keeps_div_sqrt_h_B(oracle, kind, a, h):
    n = (24, 20)
    prm = _prm(oracle, kind, n, a, h=h)
    rng = np.random.default_rng(2)
    B, E = _fields(n, 6, rng), _fields(n, 6, rng)
    act = (slice(G + 1, G + n[1] - 1), slice(G + 1, G + n[0] - 1))
    d0 = _div_B(oracle, prm, B, n)[act]
    out = B.copy()
    assert oracle.gr_field("faraday", prm, B, out, E, coeff=0.3) == 0
    d1 = _div_B(oracle, prm, out, n)[act]
    assert np.abs(out - B).max() > 1e-2
    np.testing.assert_allclose(d1, d0, rtol=0, atol=1e-12 * np.abs(d0).max())
    bad = B.copy()
    assert oracle.gr_field("faraday", prm, B, bad, E, coeff=0.3, mutate=3) == 0
    assert np.abs(_div_B(oracle, prm, bad, n)[act] - d0).max() > 1e-3 * np.abs(d0).max()


@pytest.mark.parametrize("kind,a,h", METRICS)
def test_ampere_keeps_div_sqrt_h_D(oracle, kind, a, h):
    n = (24, 20)
    prm = _prm(oracle, kind, n, a, h=h)
    rng = np.random.default_rng(3)
    D, H = _fields(n, 6, rng), _fields(n, 6, rng)
    act = (slice(G + 1, G + n[1] - 1), slice(G + 1, G + n[0] - 1))
    d0 = _div_D(oracle, prm, D, n)[act]
    out = D.copy()
    assert oracle.gr_field("ampere", prm, D, out, H, coeff=0.3) == 0
    assert np.abs(out - D).max() > 1e-2
    np.testing.assert_allclose(_div_D(oracle, prm, out, n)[act], d0, rtol=0,
                               atol=1e-12 * np.abs(d0).max())


def wald_fields(oracle, prm, n, a):
    """Wald's stationary field (pgens/wald/pgen.hpp:41-223) on the staggered
    grid with ghosts: em array with D in 0..2, B in 3..5."""
    i = np.arange(n[0] + 2 * G, dtype=float)[None, :] - G
    j = np.arange(n[1] + 2 * G, dtype=float)[:, None] - G

    def q(x1, x2):
        return _metric(oracle, prm, x1 + 0 * x2, x2 + 0 * x1)

    def A3(x1, x2):
        m = q(x1, x2)
        return 0.5 * (m["h_33"] + 2 * a * m["h_13"] * m["beta1"])

    def A1(x1, x2):
        m = q(x1, x2)
        return 0.5 * (m["h_13"] + 2 * a * m["h_11"] * m["beta1"])

    def A0(x1, x2):
        m = q(x1, x2)
        g00 = -m["alpha"] ** 2 + m["h_11"] * m["beta1"] ** 2
        return 0.5 * (m["h_13"] * m["beta1"] + 2 * a * g00)

    em = np.zeros((6, n[1] + 2 * G, n[0] + 2 * G))
    with np.errstate(divide="ignore", invalid="ignore"):
        em[3] = (A3(i, j + 1) - A3(i, j)) / q(i, j + 0.5)["sqrt_det_h"]
        em[4] = np.where(j == 0, 0.0, -(A3(i + 1, j) - A3(i, j)) / q(i + 0.5, j)["sqrt_det_h"])
        em[5] = -(A1(i + 0.5, j + 1) - A1(i + 0.5, j)) / q(i + 0.5, j + 0.5)["sqrt_det_h"]
        mp0, m00, m0p = q(i + 0.5, j), q(i, j), q(i, j + 0.5)
        # covariant D_1 = d_1 A_0 / alpha, D_3 = d_1 A_3 beta / alpha and
        # D_2 = (d_2 A_0 - beta d_2 A_1) / alpha (the pgen's expressions),
        # each formed by centred differences at the point where the
        # contravariant component lives (the pgen takes D_3 for D^1, and D_1
        # for D^3, half a cell away, a first-order offset when a != 0)
        D1_p0 = (A0(i + 1, j) - A0(i, j)) / mp0["alpha"]
        D3_p0 = (A3(i + 1, j) - A3(i, j)) * mp0["beta1"] / mp0["alpha"]
        D1_00 = (A0(i + 0.5, j) - A0(i - 0.5, j)) / m00["alpha"]
        D3_00 = (A3(i + 0.5, j) - A3(i - 0.5, j)) * m00["beta1"] / m00["alpha"]
        em[0] = mp0["h11"] * D1_p0 + mp0["h13"] * D3_p0
        D2d = (A0(i, j + 1) - A0(i, j)) / m0p["alpha"] - \
            (A1(i, j + 1) - A1(i, j)) * m0p["beta1"] / m0p["alpha"]
        em[1] = m0p["h22"] * D2d
        em[2] = np.where(j == 0, m00["h13"] * D1_00, m00["h33"] * D3_00 + m00["h13"] * D1_00)
    return np.where(np.isfinite(em), em, 0.0)


def _wald_rates(oracle, kind, a, N, mutate=0):
    """Max |dB/dt| and |dD/dt| of Wald's field (Faraday and Ampere with the
    auxiliary E, H of the kernels), relative to max |B|, over r in [3, 8],
    theta in [0.6, pi - 0.6]."""
    n = (N, N)
    prm = _prm(oracle, kind, n, a)
    em = wald_fields(oracle, prm, n, a)
    aux = np.zeros_like(em)
    assert oracle.gr_field("aux_e", prm, em, em, aux, mutate=mutate) == 0
    assert oracle.gr_field("aux_h", prm, em, em, aux, mutate=mutate) == 0
    B1 = em.copy()
    assert oracle.gr_field("faraday", prm, em, B1, aux, coeff=1.0) == 0
    D1 = em.copy()
    assert oracle.gr_field("ampere", prm, em, D1, aux, coeff=1.0) == 0
    i = np.arange(N + 2 * G)[None, :] - G + 0.5
    j = np.arange(N + 2 * G)[:, None] - G + 0.5
    m = _metric(oracle, prm, i + 0 * j, j + 0 * i)
    sel = (m["r"] > 3) & (m["r"] < 8) & (m["theta"] > 0.6) & (m["theta"] < math.pi - 0.6)
    # rates per unit code time, made comparable across grids by the cell size
    scale = np.abs(em[3:6][:, sel]).max()
    rb = np.abs((B1 - em)[3:6][:, sel]).max() / scale
    rd = np.abs((D1 - em)[0:3][:, sel]).max() / scale
    return rb, rd


@pytest.mark.parametrize("kind,a", [("kerr_schild", 0.0), ("kerr_schild", 0.7),
                                    ("qkerr_schild", 0.9)])
def test_wald_field_is_stationary_at_second_order(oracle, kind, a):
    coarse = _wald_rates(oracle, kind, a, 48)
    fine = _wald_rates(oracle, kind, a, 96)
    # a per-step change of O(h^2): halving h cuts it about 4x (the rate per
    # unit code time carries one more factor of h from the cell size)
    for c, f in zip(coarse, fine):
        assert f < c / 2.5, (coarse, fine)
    assert max(fine) < 2e-3


@pytest.mark.parametrize("mutate", [1, 2])
def test_wald_check_sees_planted_mistakes(oracle, mutate):
    good = _wald_rates(oracle, "kerr_schild", 0.7, 48)
    bad = _wald_rates(oracle, "kerr_schild", 0.7, 48, mutate=mutate)
    assert max(bad) > 10 * max(good), (good, bad)
