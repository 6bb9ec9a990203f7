"""Pins the oracle's GRPIC boundaries, spherical current filter, GR deposit and
step (CPU only).

* Axis rows: the ghost rows the axis kernel writes are the values a field of
  definite parity across the pole takes there (D^theta and B^theta odd,
  the other components even, B^theta and D^phi zero on the pole), for
  fields sampled from closed-form functions of theta; a flipped sign fails.
* Horizon rows copy the row above them; the match layer blends with
  s = tanh(4 |r - r_edge| / ds) where r comes from the grid's own definition
  (r = r0 + exp(chi) for QKS), at each component's staggered x1; B^3 taken
  at the wrong x1 fails; a field equal to its target is left unchanged.
* The absorbing layer of J scales by tanh(|r - r_edge| / (ds / 4)).
* The spherical filter is the separable binomial filter off the axis, and on
  the two rows next to each pole it is that filter applied to the current
  continued across the pole with the parities above (jx1 even, jx2 odd,
  jx3 zero on the pole).
* The GR deposit's in-plane current is the SR Esirkepov current of the same
  moves; its out-of-plane current sums to q w alpha h^3j u_j / sqrt(1 + u_i
  h^ij u_j) per particle, with alpha, h^ij from the textbook Kerr-Schild
  metric.
* The step: Wald's field with match layers at both r boundaries stays
  stationary, with a deviation that falls at second order with the grid;
  with particles, the discrete Gauss law div(sqrt(h) D) - (q0/B0) rho holds
  cell by cell to round-off over the steps (charge conservation through the
  deposit, the current's placement and the Ampere updates).
"""
import math

import numpy as np
import pytest

from test_oracle_gr_fields import G, _metric, _prm, wald_fields


def _ks0(n):
    import ntt_oracle as o
    return o.gr_params("kerr_schild_0", n, ((1.0, 20.0), (0.0, math.pi)))


def _qks(n, a=0.9, rmax=12.0, h=0.0):
    import ntt_oracle as o
    rh = 1 + math.sqrt(1 - a * a)
    return o.gr_params("qkerr_schild", n, ((0.95 * rh, rmax), (0.0, math.pi)), a=a, h=h)


def _theta(prm, x2):
    if prm.metric_kind == 1:
        return x2 * math.pi / prm.nx2   # h = 0: theta = eta
    return prm.x2_min + x2 * (prm.x2_max - prm.x2_min) / prm.nx2


def _r(prm, x1):
    """The grid's radius of code x1, from its definition."""
    if prm.metric_kind == 1:
        c0 = math.log(prm.x1_min - prm.r0)
        dchi = (math.log(prm.x1_max - prm.r0) - c0) / prm.nx1
        return prm.r0 + np.exp(c0 + x1 * dchi)
    return prm.x1_min + x1 * (prm.x1_max - prm.x1_min) / prm.nx1


def _grid(n):
    i = np.arange(n[0] + 2 * G, dtype=float)[None, :] - G
    j = np.arange(n[1] + 2 * G, dtype=float)[:, None] - G
    return i, j


# ---- axis ------------------------------------------------------------------
def _parity_fields(prm, n):
    """Fields with definite parity across both poles, sampled at each
    component's staggered point (ghost rows included): odd ones ~ sin theta,
    even ones ~ cos theta / cos 2 theta, D^phi ~ sin^2 theta (zero on the
    poles)."""
    i, j = _grid(n)
    f = np.zeros((6, n[1] + 2 * G, n[0] + 2 * G))
    g = 1.0 + 0.1 * i
    f[0] = g * np.cos(_theta(prm, j))                 # D^1 (i+1/2, j), unused by the BC
    f[1] = g * np.sin(_theta(prm, j + 0.5))           # D^2 (i, j+1/2): odd
    f[2] = g * np.sin(_theta(prm, j)) ** 2 + 0.0 * i  # D^3 (i, j): even, zero on the pole
    f[3] = g * np.cos(_theta(prm, j + 0.5))           # B^1 (i, j+1/2): even
    f[4] = g * np.sin(_theta(prm, j))                 # B^2 (i+1/2, j): odd, zero on the pole
    f[5] = g * np.cos(2 * _theta(prm, j + 0.5))       # B^3 (i+1/2, j+1/2): even
    return f


@pytest.mark.parametrize("which", ["ks0", "qks"])
@pytest.mark.parametrize("upper", [False, True])
def test_axis_ghosts_follow_the_parities(oracle, which, upper):
    n = (12, 16)
    prm = _ks0(n) if which == "ks0" else _qks(n)
    exact = _parity_fields(prm, n)
    e = G + n[1] if upper else G
    # rows the kernel writes: lower pole e-1 (and e for D^3, B^2); upper e, e+1
    rows = {1: [e - 1] if not upper else [e], 2: [e, e - 1] if not upper else [e, e + 1],
            3: [e - 1] if not upper else [e], 4: [e, e - 1] if not upper else [e, e + 1],
            5: [e - 1] if not upper else [e]}
    F = exact.copy()
    for c, rr in rows.items():
        F[c, rr, :] = 123.0
    assert oracle.gr_bc("axis", prm, F, tags=oracle.BC_D | oracle.BC_B, upper=upper) == 0
    for c, rr in rows.items():
        np.testing.assert_allclose(F[c, rr, :], exact[c, rr, :], rtol=0, atol=1e-12,
                                   err_msg=f"component {c}")
    # D^1 is not touched
    np.testing.assert_array_equal(F[0], exact[0])
    bad = exact.copy()
    for c, rr in rows.items():
        bad[c, rr, :] = 123.0
    assert oracle.gr_bc("axis", prm, bad, tags=oracle.BC_D, upper=upper, mutate=2) == 0
    assert np.abs(bad[1, rows[1], :] - exact[1, rows[1], :]).max() > 0.1
    # B rows left alone when only D is tagged
    np.testing.assert_array_equal(bad[3, rows[3], :], 123.0)


# ---- horizon -----------------------------------------------------------------
@pytest.mark.parametrize("nfilter", [0, 2])
def test_horizon_rows_copy_the_row_above(oracle, nfilter):
    n = (12, 10)
    prm = _qks(n)
    rng = np.random.default_rng(1)
    F0 = rng.standard_normal((6, n[1] + 2 * G, n[0] + 2 * G))
    F = F0.copy()
    assert oracle.gr_bc("horizon", prm, F, tags=oracle.BC_B, nfilter=nfilter) == 0
    src = G + 1 + nfilter
    rows = slice(G, G + n[1] + 1)
    for c in range(6):
        for i in range(0, 3 + nfilter):
            want = F0[c, rows, src] if c >= 3 else F0[c, rows, i]
            np.testing.assert_array_equal(F[c, rows, i], want)
    # everything else unchanged
    mask = np.ones_like(F, bool)
    mask[3:, rows, : 3 + nfilter] = False
    np.testing.assert_array_equal(F[mask], F0[mask])


# ---- match layer and absorbing layer ---------------------------------------------
@pytest.mark.parametrize("which", ["ks0", "qks"])
def test_match_layer_closed_form(oracle, which):
    n = (40, 16)
    prm = _ks0(n) if which == "ks0" else _qks(n)
    ds = 0.15 * (prm.x1_max - prm.x1_min)
    lo, hi, edge = oracle.gr_match_range(prm, True, ds)
    assert edge == prm.x1_max and hi == n[0] + 2 * G
    # the layer's first cell holds r >= r_max - ds, the one before does not
    assert _r(prm, lo - G + 1.0) > prm.x1_max - ds >= _r(prm, lo - G - 1.0)
    shape = (6, n[1] + 2 * G, n[0] + 2 * G)
    F = np.zeros(shape)
    T = np.ones(shape)
    axis = (True, True)
    assert oracle.gr_bc("match", prm, F, T, xg_edge=edge, ds=ds, i1_range=(lo, hi),
                        tags=oracle.BC_D | oracle.BC_B, axis=axis) == 0
    i, j = _grid(n)
    s_half = np.tanh(4 * np.abs(_r(prm, i + 0.5) - edge) / ds) + 0 * j
    s_int = np.tanh(4 * np.abs(_r(prm, i) - edge) / ds) + 0 * j
    inside = np.zeros(shape[1:], bool)
    inside[:, lo:hi] = True
    for c, s in ((0, s_half), (1, s_int), (3, s_int), (4, s_half), (5, s_half)):
        want = np.where(inside, 1.0 - s, 0.0)
        np.testing.assert_allclose(F[c], want, rtol=0, atol=1e-13, err_msg=f"comp {c}")
    # D^3: only damped on and beyond the axis rows
    pole = np.zeros(shape[1:], bool)
    pole[: G + 1, :] = True
    pole[G + n[1]:, :] = True
    np.testing.assert_allclose(F[2], np.where(inside & ~pole, 1.0 - s_int, 0.0), rtol=0,
                               atol=1e-13)
    # a field equal to its target is left unchanged (up to rounding)
    W = T.copy()
    assert oracle.gr_bc("match", prm, W, T, xg_edge=edge, ds=ds, i1_range=(lo, hi)) == 0
    np.testing.assert_allclose(W[:, G + 1:G + n[1]], 1.0, rtol=0, atol=1e-15)
    # only the components the target defines, only the tagged ones
    P = np.zeros(shape)
    assert oracle.gr_bc("match", prm, P, T, xg_edge=edge, ds=ds, i1_range=(lo, hi),
                        tags=oracle.BC_B, comps=0b100000) == 0
    assert np.abs(P[:5]).max() == 0 and np.abs(P[5]).max() > 0.1
    # B^3 at the wrong x1 is caught
    M = np.zeros(shape)
    assert oracle.gr_bc("match", prm, M, T, xg_edge=edge, ds=ds, i1_range=(lo, hi),
                        mutate=1) == 0
    assert np.abs(M[5] - np.where(inside, 1.0 - s_half, 0.0)).max() > 1e-2


def test_match_range_lower_side(oracle):
    """At r_min the box has ghosts excluded below and included above, so the
    range runs from the first active cell to ceil(x1(r_min + ds)) + 2 ghosts."""
    n = (40, 16)
    prm = _qks(n)
    ds = 0.5
    lo, hi, edge = oracle.gr_match_range(prm, False, ds)
    assert edge == prm.x1_min and lo == G
    x1 = oracle.gr_x1_to_code(prm, prm.x1_min + ds)
    assert hi == math.ceil(x1) + 2 * G
    assert abs(_r(prm, x1) - (prm.x1_min + ds)) < 1e-12


def test_absorb_currents_closed_form(oracle):
    n = (40, 16)
    prm = _qks(n)
    ds = 2.0
    lo, hi, edge = oracle.gr_match_range(prm, True, ds)
    J = np.ones((3, n[1] + 2 * G, n[0] + 2 * G))
    assert oracle.gr_bc("absorb_currents", prm, J, xg_edge=edge, ds=ds, i1_range=(lo, hi)) == 0
    i, j = _grid(n)
    f = np.tanh(np.abs(_r(prm, i) - edge) / (ds / 4)) + 0 * j
    want = np.ones_like(f)
    want[:, lo:hi] = f[:, lo:hi]
    for c in range(3):
        np.testing.assert_allclose(J[c], want, rtol=0, atol=1e-13)


# ---- spherical filter --------------------------------------------------------------
def _binomial(B):
    """The separable binomial filter (1/4, 1/2, 1/4) x (1/4, 1/2, 1/4) of an
    array (rows j, columns i), interior points."""
    w = np.array([0.25, 0.5, 0.25])
    out = np.zeros_like(B)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            out[1:-1, 1:-1] += w[dj + 1] * w[di + 1] * B[1 + dj:B.shape[0] - 1 + dj,
                                                          1 + di:B.shape[1] - 1 + di]
    return out


def test_spherical_filter(oracle):
    n = (12, 14)
    prm = _qks(n)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((3, n[1] + 2 * G, n[0] + 2 * G))
    A = rng.standard_normal(B.shape)
    A0 = A.copy()
    assert oracle.gr_bc("filter", prm, A, B.copy(), axis=(True, True)) == 0
    act1 = slice(G, G + n[0])
    e0, e1 = G, G + n[1]            # the two pole rows
    # off the poles: the binomial filter
    rows = slice(e0 + 2, e1 - 1)
    for c in range(3):
        np.testing.assert_allclose(A[c, rows, act1], _binomial(B[c])[rows, act1], rtol=0,
                                   atol=1e-14)
    # next to the poles: the binomial filter of the current continued across
    # the pole (jx1 at integer theta rows even, jx2 at half rows odd, jx3 zero
    # on the pole)
    ext = B.copy()
    ext[0, e0 - 1] = B[0, e0 + 1]
    ext[1, e0 - 1] = -B[1, e0]
    ext[2, e0] = 0.0
    ext[0, e1 + 1] = B[0, e1 - 1]
    ext[1, e1] = -B[1, e1 - 1]
    ext[2, e1] = 0.0
    want = np.stack([_binomial(ext[c]) for c in range(3)])
    for c, rr in ((0, [e0, e0 + 1, e1 - 1, e1]), (1, [e0, e0 + 1, e1 - 1]),
                  (2, [e0 + 1, e1 - 1])):
        np.testing.assert_allclose(A[c, rr, act1], want[c, rr, act1], rtol=0, atol=1e-14,
                                   err_msg=f"component {c}")
    np.testing.assert_array_equal(A[2, [e0, e1], act1], 0.0)
    # jx2 on the upper pole row is not written
    np.testing.assert_array_equal(A[1, e1, act1], A0[1, e1, act1])


# ---- GR deposit ----------------------------------------------------------------------
def _particles(rng, n, npart, prm):
    i1 = rng.integers(3, n[0] - 3, npart)
    i2 = rng.integers(3, n[1] - 3, npart)
    dx1, dx2 = rng.random(npart).astype(np.float32), rng.random(npart).astype(np.float32)
    step1 = rng.integers(-1, 2, npart)
    step2 = rng.integers(-1, 2, npart)
    dxp1 = np.clip(dx1 - 0.3 * step1 + 0.1 * rng.standard_normal(npart), 0.01, 0.99)
    dxp2 = np.clip(dx2 - 0.3 * step2 + 0.1 * rng.standard_normal(npart), 0.01, 0.99)
    return dict(i1=i1, i2=i2, dx1=dx1, dx2=dx2, i1_prev=i1 - (step1 * (rng.random(npart) < 0.3)),
                i2_prev=i2 - (step2 * (rng.random(npart) < 0.3)),
                dx1_prev=dxp1.astype(np.float32), dx2_prev=dxp2.astype(np.float32),
                ux1=rng.standard_normal(npart), ux2=rng.standard_normal(npart),
                ux3=rng.standard_normal(npart) * 3, weight=rng.uniform(0.5, 2, npart),
                tag=np.ones(npart, np.int16), phi=np.zeros(npart))


def test_gr_deposit(oracle):
    n = (24, 20)
    prm = _qks(n, a=0.9)
    rng = np.random.default_rng(7)
    p = _particles(rng, n, 300, prm)
    charge, dt = -1.0, 0.3
    J = np.zeros((3, n[1] + 2 * G, n[0] + 2 * G))
    assert oracle.gr_deposit(prm, p, J, charge=charge, dt=dt) == 0
    # in-plane: the SR Esirkepov current of the same moves
    Jsr = np.zeros_like(J)
    assert oracle.deposit(p, Jsr, dim=2, charge=charge, dt=dt, order=1, prec="f64") == 0
    np.testing.assert_allclose(J[:2], Jsr[:2], rtol=0, atol=1e-13)
    # out of plane: its total is sum q w v^3 with the textbook Kerr-Schild
    # alpha and h^ij at the particle (code coordinates via the grid Jacobian)
    x1 = p["i1"] + p["dx1"].astype(float)
    x2 = p["i2"] + p["dx2"].astype(float)
    m = oracle.gr_metric(prm, x1, x2)
    r, th, a = m["r"], m["theta"], prm.a
    S = r ** 2 + a ** 2 * np.cos(th) ** 2
    z = 2 * r / S
    Dl = r ** 2 - 2 * r + a ** 2
    Av = (r ** 2 + a ** 2) ** 2 - a ** 2 * Dl * np.sin(th) ** 2
    # contravariant spatial metric of Kerr-Schild (the inverse of gamma_ij)
    g11 = Av / (S * (S + 2 * r))
    g33 = 1 / (S * np.sin(th) ** 2)
    g13 = a / S
    g22 = 1 / S
    c0 = math.log(prm.x1_min - prm.r0)
    dchi = (math.log(prm.x1_max - prm.r0) - c0) / prm.nx1
    Jr = dchi * (r - prm.r0)                  # dr/dx1
    Jt = math.pi / prm.nx2                    # dtheta/dx2 (h = 0)
    u1, u2, u3 = p["ux1"], p["ux2"], p["ux3"]   # covariant code components
    # u_r = u1 / Jr, u_th = u2 / Jt; v^3 = g^{3r} u_r + g^{33} u_3
    ur, uth = u1 / Jr, u2 / Jt
    v3 = g13 * ur + g33 * u3
    usq = g11 * ur ** 2 + g22 * uth ** 2 + g33 * u3 ** 2 + 2 * g13 * ur * u3
    alpha = 1 / np.sqrt(1 + z)
    want = np.sum(charge * p["weight"] * v3 * alpha / np.sqrt(1 + usq))
    assert J[2].sum() == pytest.approx(want, rel=1e-11)


# ---- the step ----------------------------------------------------------------------
def _wald_step(oracle, N, nsteps, a=0.9):
    n = (N, N)
    prm = _prm(oracle, "qkerr_schild", n, a, rmax=12.0)
    rh = 1 + math.sqrt(1 - a * a)
    prm.dt = 0.3 * (math.log(12.0) - math.log(0.95 * rh)) / N * 0.95 * rh
    em = wald_fields(oracle, prm, n, a)
    st = oracle.GRStep(prm, em, [], flds_bc=((oracle.FLDS_MATCH, oracle.FLDS_MATCH),
                                             (oracle.FLDS_AXIS, oracle.FLDS_AXIS)),
                       prtl_bc=((oracle.PRTL_HORIZON, oracle.PRTL_ABSORB),
                                (oracle.PRTL_AXIS, oracle.PRTL_AXIS)),
                       ds=((0.3, 1.5),), target=em, deposit=False)
    for _ in range(nsteps):
        st.step_forward()
    return prm, em, st


def test_step_keeps_wald_stationary(oracle):
    """Two resolutions over the same physical time: the deviation from Wald's
    field, relative to |B|, falls by about 4 (second order), in r in [2.5, 8],
    theta in [0.4, pi - 0.4]."""
    errs = []
    for N in (32, 64):
        prm, em, st = _wald_step(oracle, N, 4 if N == 32 else 8)
        m = _metric(oracle, prm, *np.meshgrid(np.arange(N + 2 * G) - G,
                                              np.arange(N + 2 * G) - G))
        sel = (m["r"] > 2.5) & (m["r"] < 8) & (m["theta"] > 0.4) & (m["theta"] < math.pi - 0.4)
        scale = np.abs(em[3:6][:, sel]).max()
        errs.append(max(np.abs(st.em[c][sel] - em[c][sel]).max() for c in range(6)) / scale)
    assert errs[0] < 2e-2
    assert errs[0] / errs[1] > 3.0, errs


def test_step_gauss_law(oracle):
    """div(sqrt(h) D) - (q0 / B0) rho per cell, rho the order-1 charge of the
    particles at their positions, is unchanged by the steps to round-off
    (cells two away from the poles and the r layers)."""
    n = (40, 32)
    a = 0.6
    prm = _prm(oracle, "qkerr_schild", n, a, rmax=12.0)
    prm.dt = 0.01
    rng = np.random.default_rng(3)
    em = rng.standard_normal((6, n[1] + 2 * G, n[0] + 2 * G)) * 0.01
    species = []
    for q in (-1.0, 1.0):
        npart = 600
        sp = oracle.gr_params("qkerr_schild", n, ((prm.x1_min, prm.x1_max), (0.0, math.pi)),
                              a=a, dt=prm.dt, mass=1.0, charge=q, omegaB0=1.0,
                              absorb=(True, True))
        p = dict(i1=rng.integers(12, 28, npart), i2=rng.integers(10, 22, npart),
                 dx1=rng.random(npart).astype(np.float32), dx2=rng.random(npart).astype(np.float32),
                 ux1=rng.standard_normal(npart), ux2=rng.standard_normal(npart),
                 ux3=rng.standard_normal(npart), weight=rng.uniform(0.5, 1.5, npart),
                 tag=np.ones(npart, np.int16), phi=np.zeros(npart))
        for k in ("i1", "i2", "dx1", "dx2"):
            p[k + "_prev"] = p[k].copy()
        species.append((sp, p))
    q0, B0 = 0.05, 1.0
    st = oracle.GRStep(prm, em, species, flds_bc=((oracle.FLDS_MATCH, oracle.FLDS_MATCH),
                                                  (oracle.FLDS_AXIS, oracle.FLDS_AXIS)),
                       prtl_bc=((oracle.PRTL_HORIZON, oracle.PRTL_ABSORB),
                                (oracle.PRTL_AXIS, oracle.PRTL_AXIS)),
                       ds=((0.2, 0.5),), target=np.zeros_like(em), q0=q0, B0=B0)

    def invariant():
        i, j = _grid(n)
        sd = lambda x1, x2: _metric(oracle, prm, x1 + 0 * x2, x2 + 0 * x1)["sqrt_det_h"]
        D = st.em
        div = (sd(i + 0.5, j) * D[0] - sd(i - 0.5, j) * np.roll(D[0], 1, axis=1) +
               sd(i, j + 0.5) * D[1] - sd(i, j - 0.5) * np.roll(D[1], 1, axis=0))
        rho = np.zeros_like(div)
        for sp, p in st.species:
            m = p["_n"]
            x1 = p["i1"][:m] + G + p["dx1"][:m].astype(float)
            x2 = p["i2"][:m] + G + p["dx2"][:m].astype(float)
            c1, c2 = np.floor(x1).astype(int), np.floor(x2).astype(int)
            f1, f2 = x1 - c1, x2 - c2
            w = sp.charge * p["weight"][:m]
            for di, wi in ((0, 1 - f1), (1, f1)):
                for dj, wj in ((0, 1 - f2), (1, f2)):
                    np.add.at(rho, (c2 + dj, c1 + di), w * wi * wj)
        return div - (q0 / B0) * rho

    sel = (slice(G + 3, G + n[1] - 3), slice(G + 6, G + n[0] - 8))
    I0 = None
    for k in range(6):
        st.step_forward()
        if k == 0:
            I0 = invariant()[sel]   # after the first step: D and x at the same time
    I1 = invariant()[sel]
    moved = sum(np.abs(p["i1"][: p["_n"]] - p["i1_prev"][: p["_n"]]).sum() for _, p in st.species)
    assert moved > 0
    scale = np.abs(I0).max()
    np.testing.assert_allclose(I1, I0, rtol=0, atol=1e-10 * scale)
