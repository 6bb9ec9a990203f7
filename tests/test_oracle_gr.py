"""The oracle's GR pusher and Kerr-Schild metrics (src/kernels/pushers/gr.hpp,
src/metrics/kerr_schild.h, qkerr_schild.h, kerr_schild_0.h), pinned by:

* the textbook Kerr-Schild 3+1 split (spatial metric, lapse, shift), mapped to
  code coordinates by the Jacobian of each grid, independently of the
  restatement;
* h_ij h^jk = delta;
* every dr_/dt_ derivative against a central finite difference of the
  function it differentiates (the quasi-spherical h^22 derivatives omit the
  angle-stretching factor in the reference: pinned through that factor);
* the vector transformations: round trips, the tetrad norm h^ij u_i u_j;
* geodesic invariants of the pusher (u_phi exactly, energy -u_t to the
  scheme's accuracy), a gyration closing in flat space (KS0) with a uniform
  B, and the axis / absorbing boundaries.
"""
import math

import numpy as np
import pytest

EPS_FD = 1e-5


def _ks(a=0.9, n=(64, 64), rmax=20.0, **kw):
    import ntt_oracle as o
    rh = 1 + math.sqrt(1 - a * a)
    return o.gr_params("kerr_schild", n, ((0.9 * rh, rmax), (0.0, math.pi)), a=a, **kw)


def _qks(a=0.9, h=0.0, n=(64, 64), rmax=50.0, **kw):
    import ntt_oracle as o
    rh = 1 + math.sqrt(1 - a * a)
    return o.gr_params("qkerr_schild", n, ((0.9 * rh, rmax), (0.0, math.pi)), a=a, r0=0.0, h=h,
                       **kw)


def _ks0(n=(64, 64), **kw):
    import ntt_oracle as o
    return o.gr_params("kerr_schild_0", n, ((1.0, 20.0), (0.0, math.pi)), **kw)


METRICS = {"ks": lambda: _ks(), "qks": lambda: _qks(), "qks_h": lambda: _qks(h=0.25),
           "ks0": lambda: _ks0()}


def _points(rng, n=200, nx=64):
    x1 = rng.uniform(0.5, nx - 0.5, n)
    x2 = rng.uniform(1.0, nx - 1.0, n)   # away from the poles
    return x1, x2


def _jacobian(prm, x1, x2):
    """dr/dx1, dtheta/dx2 of the grid, independently of the oracle."""
    if prm.metric_kind == 1:
        chi_min = math.log(prm.x1_min - prm.r0)
        dchi = (math.log(prm.x1_max - prm.r0) - chi_min) / prm.nx1
        h = prm.h

        def e2t(e):
            return e + 2 * h * e * (math.pi - 2 * e) * (math.pi - e) / math.pi ** 2
        # eta grid: eta(theta_min), eta(theta_max) invert eta2theta numerically
        from scipy.optimize import brentq
        emin = brentq(lambda e: e2t(e) - prm.x2_min, -1, 4) if h else prm.x2_min
        emax = brentq(lambda e: e2t(e) - prm.x2_max, -1, 4) if h else prm.x2_max
        deta = (emax - emin) / prm.nx2
        eta = x2 * deta + emin
        dth_deta = 1 + 2 * h / math.pi ** 2 * (math.pi ** 2 - 6 * math.pi * eta + 6 * eta ** 2)
        return dchi * np.exp(x1 * dchi + chi_min), dth_deta * deta, dth_deta, deta
    dr = (prm.x1_max - prm.x1_min) / prm.nx1
    dth = (prm.x2_max - prm.x2_min) / prm.nx2
    return np.full_like(x1, dr), np.full_like(x2, dth), np.ones_like(x2), dth


@pytest.mark.parametrize("name", list(METRICS))
def test_textbook_kerr_schild(oracle, name):
    """gamma_ij, alpha, beta^r of Kerr in Kerr-Schild coordinates
    (Sigma = r^2 + a^2 cos^2, z = 2r/Sigma): gamma_rr = 1 + z, gamma_thth = Sigma,
    gamma_phph = A sin^2/Sigma, gamma_rph = -a (1 + z) sin^2, alpha = (1+z)^-1/2,
    beta^r = z/(1+z); mapped to code coordinates by the grid's Jacobian."""
    prm = METRICS[name]()
    rng = np.random.default_rng(1)
    x1, x2 = _points(rng)
    m = oracle.gr_metric(prm, x1, x2)
    r, th = m["r"], m["theta"]
    a = 0.0 if prm.metric_kind == 2 else prm.a
    S = r ** 2 + a ** 2 * np.cos(th) ** 2
    z = 2 * r / S if prm.metric_kind != 2 else np.zeros_like(r)
    D = r ** 2 - 2 * r + a ** 2
    A = (r ** 2 + a ** 2) ** 2 - a ** 2 * D * np.sin(th) ** 2
    if prm.metric_kind == 2:
        S, A = r ** 2, r ** 4
    Jr, Jt, _, _ = _jacobian(prm, x1, x2)
    np.testing.assert_allclose(m["h_11"], (1 + z) * Jr ** 2, rtol=1e-12)
    np.testing.assert_allclose(m["h_22"], S * Jt ** 2, rtol=1e-12)
    np.testing.assert_allclose(m["h_33"], A * np.sin(th) ** 2 / S, rtol=1e-12)
    np.testing.assert_allclose(m["h_13"], -a * (1 + z) * np.sin(th) ** 2 * Jr, rtol=1e-12,
                               atol=1e-300)
    np.testing.assert_allclose(m["alpha"], 1 / np.sqrt(1 + z), rtol=1e-13)
    np.testing.assert_allclose(m["beta1"] * Jr, z / (1 + z), rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("case", ["ks", "qks"])
def test_reference_ks_qks_cases(oracle, case):
    """The two grids of the reference's own metric test (tests/metrics/ks-qks.cpp:177-193):
    KerrSchild 64 x 54, r in [0.8, 50], a = 0.95, and QKerrSchild 32 x 42,
    r in [0.8, 10], r0 = -2, h = 0, a = 0.8; at every cell centre h_ij and
    h^ij equal the textbook Kerr-Schild expressions (the reference checks them
    in the physical basis to 10 ulp; here in code coordinates to 1e-12)."""
    import ntt_oracle as o
    if case == "ks":
        prm = o.gr_params("kerr_schild", (64, 54), ((0.8, 50.0), (0.0, math.pi)), a=0.95)
    else:
        prm = o.gr_params("qkerr_schild", (32, 42), ((0.8, 10.0), (0.0, math.pi)), a=0.8,
                          r0=-2.0, h=0.0)
    i, j = np.meshgrid(np.arange(prm.nx1), np.arange(prm.nx2), indexing="ij")
    x1, x2 = i.ravel() + 0.5, j.ravel() + 0.5
    m = oracle.gr_metric(prm, x1, x2)
    r, th, a = m["r"], m["theta"], prm.a
    S = r ** 2 + (a * np.cos(th)) ** 2
    z = 2 * r / S
    D = r ** 2 - 2 * r + a ** 2
    A = (r ** 2 + a ** 2) ** 2 - (a * np.sin(th)) ** 2 * D
    Jr, Jt, _, _ = _jacobian(prm, x1, x2)
    np.testing.assert_allclose(m["h_11"], (1 + z) * Jr ** 2, rtol=1e-12)
    np.testing.assert_allclose(m["h_22"], S * Jt ** 2, rtol=1e-12)
    np.testing.assert_allclose(m["h_33"], A * np.sin(th) ** 2 / S, rtol=1e-12)
    np.testing.assert_allclose(m["h_13"], -a * (1 + z) * np.sin(th) ** 2 * Jr, rtol=1e-12)
    np.testing.assert_allclose(m["h11"], A / (S * (S + 2 * r)) / Jr ** 2, rtol=1e-12)
    np.testing.assert_allclose(m["h22"], 1 / S / Jt ** 2, rtol=1e-12)
    np.testing.assert_allclose(m["h33"], 1 / (S * np.sin(th) ** 2), rtol=1e-12)
    np.testing.assert_allclose(m["h13"], a / S / Jr, rtol=1e-12)


@pytest.mark.parametrize("name", list(METRICS))
def test_upper_is_inverse_of_lower(oracle, name):
    prm = METRICS[name]()
    x1, x2 = _points(np.random.default_rng(2))
    m = oracle.gr_metric(prm, x1, x2)
    tol = 1e-12
    np.testing.assert_allclose(m["h_11"] * m["h11"] + m["h_13"] * m["h13"], 1, atol=tol)
    np.testing.assert_allclose(m["h_11"] * m["h13"] + m["h_13"] * m["h33"], 0,
                               atol=tol * np.abs(m["h_11"] * m["h13"]).max() + 1e-300)
    np.testing.assert_allclose(m["h_22"] * m["h22"], 1, atol=tol)
    np.testing.assert_allclose(m["h_13"] * m["h13"] + m["h_33"] * m["h33"], 1, atol=tol)


def _fd(oracle, prm, q, x1, x2, axis):
    e = EPS_FD
    if axis == 0:
        p, m = oracle.gr_metric(prm, x1 + e, x2), oracle.gr_metric(prm, x1 - e, x2)
    else:
        p, m = oracle.gr_metric(prm, x1, x2 + e), oracle.gr_metric(prm, x1, x2 - e)
    return (p[q] - m[q]) / (2 * e)


@pytest.mark.parametrize("name", list(METRICS))
@pytest.mark.parametrize("q", ["alpha", "beta1", "h11", "h22", "h33", "h13"])
def test_derivatives_match_finite_differences(oracle, name, q):
    prm = METRICS[name]()
    x1, x2 = _points(np.random.default_rng(3), n=100)
    m = oracle.gr_metric(prm, x1, x2)
    for axis, pre in ((0, "dr_"), (1, "dt_")):
        fd = _fd(oracle, prm, q, x1, x2, axis)
        got = m[pre + q]
        if name == "qks_h" and q == "h22":
            # qkerr_schild.h: the h^22 derivatives drop the angle-stretching
            # factor (dtheta/deta)^2 of h^22 = 1 / (Sigma (dtheta/deta)^2 deta^2)
            _, _, dtde, deta = _jacobian(prm, x1, x2)
            if axis == 0:
                fd = fd * dtde ** 2
            else:
                Sig = m["h_22"] / (deta ** 2 * dtde ** 2)
                e = EPS_FD
                Sp = oracle.gr_metric(prm, x1, x2 + e)["h_22"]
                Sm = oracle.gr_metric(prm, x1, x2 - e)["h_22"]
                _, _, dp, _ = _jacobian(prm, x1, x2 + e)
                _, _, dm, _ = _jacobian(prm, x1, x2 - e)
                dSig = (Sp / (deta ** 2 * dp ** 2) - Sm / (deta ** 2 * dm ** 2)) / (2 * e)
                fd = -dSig / Sig ** 2 / deta ** 2
        scale = np.abs(fd).max() + 1e-12
        np.testing.assert_allclose(got, fd, rtol=0, atol=2e-6 * scale,
                                   err_msg=f"{pre}{q} ({name})")


@pytest.mark.parametrize("name", list(METRICS))
def test_transform_round_trips_and_tetrad_norm(oracle, name):
    prm = METRICS[name]()
    rng = np.random.default_rng(4)
    x1, x2 = _points(rng, n=50)
    v = rng.standard_normal((50, 3))
    for fwd, back in (("D->U", "U->D"), ("D->T", "T->D"), ("U->T", "T->U")):
        w = oracle.gr_transform(prm, back, x1, x2, oracle.gr_transform(prm, fwd, x1, x2, v))
        np.testing.assert_allclose(w, v, rtol=1e-11, atol=1e-11, err_msg=fwd)
    # the tetrad is orthonormal: |v_hat|^2 = h^ij v_i v_j
    vu = oracle.gr_transform(prm, "D->U", x1, x2, v)
    vh = oracle.gr_transform(prm, "D->T", x1, x2, v)
    np.testing.assert_allclose((vh ** 2).sum(1), (v * vu).sum(1), rtol=1e-11)
    # the hatted vector does not depend on the route
    np.testing.assert_allclose(oracle.gr_transform(prm, "U->T", x1, x2, vu), vh, rtol=1e-11,
                               atol=1e-12)


def _gr_particles(i1, dx1, i2, dx2, u):
    n = len(i1)
    return {"i1": np.array(i1, np.int32), "i2": np.array(i2, np.int32),
            "dx1": np.array(dx1, np.float32), "dx2": np.array(dx2, np.float32),
            "i1_prev": np.zeros(n, np.int32), "i2_prev": np.zeros(n, np.int32),
            "dx1_prev": np.zeros(n, np.float32), "dx2_prev": np.zeros(n, np.float32),
            "ux1": np.array([x[0] for x in u]), "ux2": np.array([x[1] for x in u]),
            "ux3": np.array([x[2] for x in u]), "weight": np.ones(n), "phi": np.zeros(n),
            "tag": np.ones(n, np.int16)}


def _energy(oracle, prm, p, massless):
    """-u_t from g^{mu nu} u_mu u_nu = -m^2 in the 3+1 split: the future root
    is u_t = beta^i u_i - alpha Gamma with Gamma^2 = m^2 + h^ij u_i u_j."""
    x1 = p["i1"] + p["dx1"].astype(np.float64)
    x2 = p["i2"] + p["dx2"].astype(np.float64)
    m = oracle.gr_metric(prm, x1, x2)
    u1, u2, u3 = p["ux1"], p["ux2"], p["ux3"]
    hij = m["h11"] * u1 ** 2 + m["h22"] * u2 ** 2 + m["h33"] * u3 ** 2 + 2 * m["h13"] * u1 * u3
    Gam = np.sqrt(hij + (0.0 if massless else 1.0))
    return m["alpha"] * Gam - m["beta1"] * u1


@pytest.mark.parametrize("name", ["ks", "qks"])
@pytest.mark.parametrize("massless", [False, True])
def test_geodesic_invariants(oracle, name, massless):
    """Fields off: u_phi is carried unchanged (to round-off through the zero-
    field EM half pushes for massive particles) and -u_t is conserved up to the
    scheme's O(dt) error (x and u live half a step apart), around a = 0.9.
    (Quasi-spherical grids with angle stretching, h != 0, do not conserve it:
    their h^22 derivatives drop the stretching factor, see the derivative test.)"""
    prm = METRICS[name]()
    prm.dt = 0.01
    prm.massless = int(massless)
    prm.absorb_i1min = prm.absorb_i1max = 0
    rng = np.random.default_rng(5)
    n = 64
    i1 = rng.integers(20, 40, n)
    i2 = rng.integers(16, 48, n)
    u = [tuple(rng.normal(0, 0.5, 3) * (1, 4, 4)) for _ in range(n)]
    p = oracle.normalize_particles(_gr_particles(i1, rng.random(n), i2, rng.random(n), u), "f64")
    em = np.zeros((6, 64 + 4, 64 + 4))
    E0 = _energy(oracle, prm, p, massless)
    uphi0 = p["ux3"].copy()
    for _ in range(100):
        assert oracle.gr_push(prm, p, em, em, order=1) == 0
    E1 = _energy(oracle, prm, p, massless)
    if massless:
        np.testing.assert_array_equal(p["ux3"], uphi0)
    else:   # the zero-field EM half pushes round-trip u through the tetrad
        np.testing.assert_allclose(p["ux3"], uphi0, rtol=1e-12)
    assert (p["tag"] == 1).all()
    np.testing.assert_allclose(E1, E0, rtol=1e-3)
    assert np.abs(p["i1"] - i1).max() > 0  # the particles moved


def test_gyration_closes_in_flat_space(oracle):
    """KS0 is flat space in spherical coordinates.  A uniform B along z
    (B^r = B cos(th)/sqrt(h_11), B^th = -B sin(th)/sqrt(h_22) on the staggered
    grid) gyrates a massive particle in the equatorial plane on a circle of
    radius u/(omegaB0 B) with period 2 pi gamma/(omegaB0 B): after one period
    it is back where it started, with gamma unchanged."""
    import ntt_oracle as o
    n = (256, 256)
    prm = o.gr_params("kerr_schild_0", n, ((4.0, 16.0), (0.0, math.pi)))
    prm.absorb_i1min = prm.absorb_i1max = 0
    u, B = 0.5, 1.0
    gam = math.sqrt(1 + u * u)
    period = 2 * math.pi * gam / B
    nsteps = 400
    prm.dt = period / nsteps
    prm.charge, prm.mass, prm.omegaB0 = 1.0, 1.0, 1.0
    dr = 12.0 / n[0]
    dth = math.pi / n[1]
    G = 2
    em0 = np.zeros((6, n[1] + 2 * G, n[0] + 2 * G))
    ii = np.arange(n[0] + 2 * G) - G
    jj = np.arange(n[1] + 2 * G) - G
    # B^1 at (i, j + 1/2), B^2 at (i + 1/2, j), contravariant code components
    r1 = 4.0 + ii[None, :] * dr
    t1 = (jj[:, None] + 0.5) * dth
    em0[3] = B * np.cos(t1) / dr + 0 * r1
    r2 = 4.0 + (ii[None, :] + 0.5) * dr
    t2 = jj[:, None] * dth
    em0[4] = -B * np.sin(t2) / (dth * r2)
    em = np.zeros_like(em0)
    # start on the equator at r = 10 moving along +r (hatted u_r = u):
    # covariant u_1 = u * sqrt(h_11) = u * dr
    x1 = (10.0 - 4.0) / dr
    x2 = (math.pi / 2) / dth
    p = _gr_particles([int(x1)], [x1 - int(x1)], [int(x2)], [x2 - int(x2)], [(u * dr, 0.0, 0.0)])
    p = oracle.normalize_particles(p, "f64")
    for _ in range(nsteps):
        assert oracle.gr_push(prm, p, em, em0, order=1) == 0
    r = 4.0 + (p["i1"][0] + float(p["dx1"][0])) * dr
    th = (p["i2"][0] + float(p["dx2"][0])) * dth
    assert abs(r - 10.0) < 2e-3 and abs(th - math.pi / 2) < 2e-4
    uh = oracle.gr_transform(prm, "D->T", [p["i1"][0] + float(p["dx1"][0])],
                             [p["i2"][0] + float(p["dx2"][0])],
                             np.array([p["ux1"][0], p["ux2"][0], p["ux3"][0]]))
    assert abs(math.sqrt(1 + (uh ** 2).sum()) - gam) < 1e-4
    assert abs(uh[0, 0] - u) < 5e-3


def test_axis_reflects_and_horizon_absorbs(oracle):
    prm = _ks(a=0.5)
    prm.dt = 0.5
    prm.massless = 1
    # moving toward theta = 0 from the first cell: reflected off the axis
    p = oracle.normalize_particles(
        _gr_particles([40, 1], [0.5, 0.5], [0, 30], [0.05, 0.5], [(0.0, -3.0, 0.0), (-50.0, 0.0, 0.0)]),
        "f64")
    assert oracle.gr_push(prm, p, None, None) == 0
    assert p["i2"][0] == 0 and p["ux2"][0] > 0 and 0 <= p["dx2"][0] <= 1
    # falling through r_min with an absorbing inner boundary: dead
    assert p["tag"][1] == 0 and p["i1"][1] < 0


def test_invalid_tag_reported(oracle):
    prm = _ks()
    prm.massless = 1
    prm.dt = 0.1
    p = oracle.normalize_particles(_gr_particles([10], [0.5], [30], [0.5], [(0.1, 0.1, 0.1)]), "f64")
    p["tag"][0] = 7
    assert oracle.gr_push(prm, p, None, None) != 0
