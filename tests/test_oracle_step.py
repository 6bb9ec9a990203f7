"""Pins the oracle's step composition and the per-cell deposit values (CPU only).

* Vacuum Yee eigenmodes through ntt_oracle_sr_step (no particles): the step
  of srpic.hpp:65-188 (Faraday 1/2, Faraday 1/2, Ampere 1) is the leapfrog
  B^{n+1/2} = B^{n-1/2} - dt curl E^n, E^{n+1} = E^n + dt curl B^{n+1/2}, whose
  discrete plane waves E = e cos(k.x - w t), B = (kt x e)/|kt| cos(k.x - w t)
  (kt_d = 2 sin(k_d dx / 2) / dx, B stored half a step behind E) obey
  sin(w dt / 2) = (dt / 2)|kt|.  Sampled at every component's staggered
  position, the mode must come back to FP64 round-off after many steps.
  The same check applied to wrong compositions built from the oracle's own
  Faraday/Ampere primitives (Ampere with dt/2, Ampere between the Faraday
  halves) fails, which shows the check sees those mistakes.
* Push placement: a zero-weight particle in the wave is pushed by the fields
  after the first Faraday half exactly (sr_push on Faraday(1/2)(em)), and
  not by those of the step's start or end.
* 2D out-of-plane current per cell: Esirkepov's J_z at node (a, b) is
  q w v_z * int_0^1 prod_d [(1 - t) S(a_d - x_d^prev) + t S(a_d - x_d)] dt
  with textbook B-splines S (currents_deposit.hpp:476-555), and zig-zag's J
  (all three components) is the two-segment relay-point split of Umeda et al.
  (2003) with area weights at each segment's midpoint
  (currents_deposit.hpp:171-305).  A variant with the half-weights of the
  Esirkepov J_z swapped keeps every sum but fails the per-cell comparison.
"""
import math

import numpy as np
import pytest

from _helpers import G, bspline, field_array, move_particles, random_particles

# staggered positions (in cells) of em components 0..5 (Ex, Ey, Ez, Bx, By, Bz)
STAGGER = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5), (0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0)]


def _mode(dim, n, m, pol):
    """Wave vector k, discrete kt, E amplitude e (kt . e = 0), B amplitude b, frequency."""
    k = np.array([2 * math.pi * m[d] / n[d] if d < dim else 0.0 for d in range(3)])
    kt = 2 * np.sin(k / 2)
    e = np.asarray(pol, dtype=np.float64)
    e = e - kt * (e @ kt) / (kt @ kt)   # discrete transversality
    e /= np.linalg.norm(e)
    b = np.cross(kt, e) / np.linalg.norm(kt)
    return k, kt, e, b


def _sample(dim, n, k, e, b, omega, tE, tB):
    """em array (ghosts included) holding the mode with E at time tE, B at tB."""
    em = field_array(dim, n, 6, dtype=np.float64)
    ext = [x + 2 * G for x in n]
    for c in range(6):
        amp, t = (e[c], tE) if c < 3 else (b[c - 3], tB)
        axes = [np.arange(ext[d], dtype=np.float64) - G + STAGGER[c][d] for d in range(dim)]
        grids = np.meshgrid(*axes, indexing="ij")
        ph = sum(k[d] * grids[d] for d in range(dim)) - omega * t
        em[c] = (amp * np.cos(ph)).T
    return em


def _vacuum_case(dim, n, m, pol, dt):
    k, kt, e, b = _mode(dim, n, m, pol)
    omega = 2 / dt * math.asin(dt * np.linalg.norm(kt) / 2)
    return k, e, b, omega


def _step_params(dt, dim):
    return dict(dt=dt, correction=1.0, omegaB0=1.0, q0=1.0, B0=1.0, V0=1.0, ppc0=1.0, dx=1.0,
                shape_order=1, n_filters=0, fieldsolver_enable=1, deposit_enable=1, step=0)


VACUUM = [(2, (24, 20), (1, 0), (0, 0, 1)), (2, (24, 20), (2, 3), (0, 0, 1)),
          (2, (24, 20), (2, -1), (1, 0, 0)), (3, (12, 10, 8), (1, 2, -1), (1, 0.3, -0.2)),
          (3, (12, 10, 8), (0, 0, 2), (0, 1, 0))]


@pytest.mark.parametrize("case", VACUUM, ids=lambda c: f"{c[0]}d-m{c[2]}")
def test_vacuum_yee_mode_through_step(oracle, case):
    dim, n, m, pol = case
    dt = 0.5 / math.sqrt(dim)
    k, e, b, omega = _vacuum_case(dim, n, m, pol, dt)
    em = _sample(dim, n, k, e, b, omega, 0.0, -0.5 * dt)
    cur = field_array(dim, n, 3, dtype=np.float64)
    buff = np.zeros_like(cur)
    nsteps = 40
    prm = _step_params(dt, dim)
    for s in range(nsteps):
        prm["step"] = s
        assert oracle.sr_step(em, cur, buff, [], dim=dim, prm=prm, bc=((1, 1),) * 3,
                              prec="f64") == 0
    want = _sample(dim, n, k, e, b, omega, nsteps * dt, (nsteps - 0.5) * dt)
    err = np.abs(em - want).max()
    assert err < 1e-12, err
    # the wave moved: after 40 steps its phase advanced by omega * 40 dt
    assert omega * nsteps * dt > 1.0
    assert not cur.any()


def _wrong_step(oracle, em, dim, dt, variant):
    c = lambda f: (f * dt, f * dt) if dim == 2 else (f * dt, 0.0)   # noqa: E731
    if variant == "ampere_half":
        seq = [("F", 0.5), ("F", 0.5), ("A", 0.5)]
    else:  # Ampere between the two Faraday halves
        seq = [("F", 0.5), ("A", 1.0), ("F", 0.5)]
    for op, f in seq:
        c1, c2 = c(f)
        if op == "F":
            oracle.faraday(em, dim=dim, coeff1=c1, coeff2=c2, prec="f64")
            oracle.communicate(em, 3, 6, dim=dim, prec="f64")
        else:
            oracle.ampere(em, dim=dim, coeff1=c1, coeff2=c2, prec="f64")
            oracle.communicate(em, 0, 3, dim=dim, prec="f64")


@pytest.mark.parametrize("variant", ["ampere_half", "ampere_between_halves"])
def test_vacuum_check_rejects_wrong_compositions(oracle, variant):
    dim, n, m, pol = VACUUM[1]
    dt = 0.5 / math.sqrt(dim)
    k, e, b, omega = _vacuum_case(dim, n, m, pol, dt)
    em = _sample(dim, n, k, e, b, omega, 0.0, -0.5 * dt)
    nsteps = 40
    for _ in range(nsteps):
        _wrong_step(oracle, em, dim, dt, variant)
    want = _sample(dim, n, k, e, b, omega, nsteps * dt, (nsteps - 0.5) * dt)
    assert np.abs(em - want).max() > 1e-3


@pytest.mark.parametrize("dim", [2, 3])
def test_push_sees_fields_after_first_faraday_half(oracle, dim):
    n = (16, 12, 10)[:dim]
    dt = 0.5 / math.sqrt(dim)
    k, e, b, omega = _vacuum_case(dim, n, (1, 2, 1)[:dim], (0.3, -0.2, 1.0), dt)
    em0 = _sample(dim, n, k, e, b, omega, 0.0, -0.5 * dt)
    # add a uniform B so the rotation is not small
    em0[3:6] += np.array([0.2, -0.1, 0.7]).reshape((3,) + (1,) * dim)
    rng = np.random.default_rng(7)
    p = random_particles(rng, dim, n, 64, margin=3)
    p["weight"] = np.zeros(64)
    order, pusher = 2, 2
    sp = [{"p": oracle.normalize_particles(p, "f64"), "mass": 1.0, "charge": -1.0,
           "pusher": pusher}]
    em = em0.copy()
    cur = field_array(dim, n, 3, dtype=np.float64)
    prm = _step_params(dt, dim)
    prm["shape_order"] = order
    assert oracle.sr_step(em, cur, np.zeros_like(cur), sp, dim=dim, prm=prm,
                          bc=((1, 1),) * 3, prec="f64") == 0

    def pushed(fields):
        q, rc = oracle.sr_push(p, fields, dim=dim, mass=1.0, charge=-1.0, dt=dt, omegaB0=1.0,
                               pusher=pusher, order=order, prec="f64")
        assert rc == 0
        return q

    half = em0.copy()
    c1, c2 = (0.5 * dt, 0.5 * dt) if dim == 2 else (0.5 * dt, 0.0)
    oracle.faraday(half, dim=dim, coeff1=c1, coeff2=c2, prec="f64")
    oracle.communicate(half, 3, 6, dim=dim, prec="f64")
    ref = pushed(half)
    got = sp[0]["p"]
    for key in ("ux1", "ux2", "ux3", "dx1", "dx2"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
    # the fields at the step's start or end give measurably different momenta
    for other in (em0, em):
        q = pushed(other.copy())
        assert np.abs(q["ux1"] - ref["ux1"]).max() > 1e-6
    # zero weight deposits nothing: the fields evolve as in vacuum
    vac = em0.copy()
    prm0 = dict(prm)
    cur0 = field_array(dim, n, 3, dtype=np.float64)
    assert oracle.sr_step(vac, cur0, np.zeros_like(cur0), [], dim=dim, prm=prm0,
                          bc=((1, 1),) * 3, prec="f64") == 0
    np.testing.assert_array_equal(em, vac)


def _moved_2d(n, npart, seed, maxdisp):
    rng = np.random.default_rng(seed)
    p = random_particles(rng, 2, n, npart, margin=3)
    disp = [rng.uniform(-maxdisp, maxdisp, npart) for _ in range(2)]
    q = move_particles(p, 2, disp)
    q["weight"] = rng.uniform(0.5, 2.0, npart)
    q["ux3"] = rng.uniform(-3, 3, npart)
    return q


def _positions(q, prev):
    suf = "_prev" if prev else ""
    return [q[f"i{d}{suf}"].astype(np.float64) + q[f"dx{d}{suf}"].astype(np.float64)
            for d in (1, 2)]


def _vz(q):
    u = np.stack([q["ux1"], q["ux2"], q["ux3"]]).astype(np.float64)
    return u[2] / np.sqrt(1 + (u ** 2).sum(0))


def _jz_time_integral(q, n, order, charge, swapped=False):
    """q w v_z * int_0^1 prod_d [(1-t) iS_d + t fS_d] dt per node, from textbook
    B-splines; swapped=True builds the half-weight-swapped variant."""
    Jz = np.zeros(tuple(s + 2 * G for s in reversed(n)))
    xi, xf = _positions(q, True), _positions(q, False)
    vz = _vz(q)
    w = q["weight"].astype(np.float64) * charge
    a_ii, a_ff, a_if = (1 / 3, 1 / 3, 1 / 6) if not swapped else (1 / 6, 1 / 6, 1 / 3)
    for p in range(len(w)):
        base = [int(np.floor(min(xi[d][p], xf[d][p]))) for d in range(2)]
        for a in range(base[0] - 2, base[0] + 4):
            iS1, fS1 = bspline(order, a - xi[0][p]), bspline(order, a - xf[0][p])
            for bb in range(base[1] - 2, base[1] + 4):
                iS2, fS2 = bspline(order, bb - xi[1][p]), bspline(order, bb - xf[1][p])
                val = a_ii * iS1 * iS2 + a_ff * fS1 * fS2 + a_if * (iS1 * fS2 + fS1 * iS2)
                Jz[bb + G, a + G] += w[p] * vz[p] * val
    return Jz


@pytest.mark.parametrize("order", [1, 2])
def test_esirkepov_2d_jz_per_cell(oracle, order):
    n = (12, 11)
    q = _moved_2d(n, 80, seed=30 + order, maxdisp=0.9)
    charge, dt = -1.3, 0.37
    J = field_array(2, n, 3, dtype=np.float64)
    assert oracle.deposit(q, J, dim=2, charge=charge, dt=dt, order=order, prec="f64") == 0
    want = _jz_time_integral(q, n, order, charge)
    scale = np.abs(want).max()
    np.testing.assert_allclose(J[2], want, rtol=0, atol=1e-13 * scale)
    # the half-weight swap keeps the sum but moves the per-cell values
    bad = _jz_time_integral(q, n, order, charge, swapped=True)
    assert abs(bad.sum() - want.sum()) < 1e-10 * np.abs(want).sum()
    assert np.abs(bad - J[2]).max() > 1e-3 * scale


def _zigzag_expected(q, n, charge, dt, relay_at_face=True):
    """Umeda et al. (2003) zig-zag: relay point at the mid-point when the cell
    does not change and at the crossed face otherwise; each segment deposits
    its flux with area weights at its own mid-point, J_z with half of q w v_z
    per segment."""
    J = np.zeros((3,) + tuple(s + 2 * G for s in reversed(n)))
    xi, xf = _positions(q, True), _positions(q, False)
    ci = [q[f"i{d}_prev"].astype(np.int64) for d in (1, 2)]
    cf = [q[f"i{d}"].astype(np.int64) for d in (1, 2)]
    vz = _vz(q)
    w = q["weight"].astype(np.float64) * charge
    for p in range(len(w)):
        xr = []
        for d in range(2):
            if ci[d][p] == cf[d][p] or not relay_at_face:
                xr.append(0.5 * (xi[d][p] + xf[d][p]))
            else:
                xr.append(float(max(ci[d][p], cf[d][p])))
        segs = [(ci[0][p], ci[1][p], [xi[0][p], xi[1][p]]),
                (cf[0][p], cf[1][p], [xf[0][p], xf[1][p]])]
        for s, (c1, c2, xe) in enumerate(segs):
            # segment 0 runs x_prev -> x_r, segment 1 runs x_r -> x
            flux = [(xr[d] - xe[d]) if s == 0 else (xe[d] - xr[d]) for d in range(2)]
            mid = [0.5 * (xe[d] + xr[d]) - (c1 if d == 0 else c2) for d in range(2)]
            Fx, Fy = (f * w[p] / dt for f in flux)
            Fz = 0.5 * w[p] * vz[p]
            J[0][c2 + G, c1 + G] += Fx * (1 - mid[1])
            J[0][c2 + 1 + G, c1 + G] += Fx * mid[1]
            J[1][c2 + G, c1 + G] += Fy * (1 - mid[0])
            J[1][c2 + G, c1 + 1 + G] += Fy * mid[0]
            for da, wa in ((0, 1 - mid[0]), (1, mid[0])):
                for db, wb in ((0, 1 - mid[1]), (1, mid[1])):
                    J[2][c2 + db + G, c1 + da + G] += Fz * wa * wb
    return J


def test_zigzag_2d_per_cell(oracle):
    n = (12, 11)
    q = _moved_2d(n, 120, seed=41, maxdisp=0.6)
    charge, dt = 0.8, 0.41
    J = field_array(2, n, 3, dtype=np.float64)
    assert oracle.deposit(q, J, dim=2, charge=charge, dt=dt, order=0, prec="f64") == 0
    want = _zigzag_expected(q, n, charge, dt)
    # the relay point is formed in single precision even in the double build
    # (prtldx_t arithmetic, currents_deposit.hpp:175-181)
    for c in range(3):
        scale = np.abs(want[c]).max()
        np.testing.assert_allclose(J[c], want[c], rtol=0, atol=2e-6 * scale, err_msg=f"J{c}")
    # a relay point always at the mid-point (no face split) is caught
    crossing = (q["i1"] != q["i1_prev"]) | (q["i2"] != q["i2_prev"])
    assert crossing.sum() > 20
    bad = _zigzag_expected(q, n, charge, dt, relay_at_face=False)
    assert np.abs(bad[2] - J[2]).max() > 1e-3 * np.abs(want[2]).max()
