"""Pins the oracle's field kernels (CPU only).

* Faraday and Ampere reproduce analytic curls of trigonometric fields to
  O(dx^2), as in tests/kernels/faraday_mink.cpp:111-329 and
  tests/kernels/ampere_mink.cpp (same fields and tolerances dx^2/3).
* Standard-Yee Faraday leaves the discrete div B unchanged (any E).
* Extended-stencil Faraday with all coefficients zero equals standard Yee;
  with non-zero coefficients it still differentiates linear fields exactly
  (alpha = 1 - 2 beta - 3 delta makes every stencil consistent).
* 2D filter impulse response and sum = 1 (tests/kernels/digital_filter.cpp:31-128);
  the 3D stencil's weights, read off its impulse response, are those written
  in src/kernels/digital_filter.hpp:349-378.
* Periodic ghost fill is the periodic extension; the current fold conserves
  the total current and equals a brute-force periodic accumulation.
"""
import math

import numpy as np
import pytest

from _helpers import G, field_array, node_coords


def _coords(dim, n, ext):
    dx = ext[0] / n[0]
    return dx


def test_faraday_2d_analytic(oracle):
    # tests/kernels/faraday_mink.cpp:111-206
    n = (64, 128)
    sx, sy = 2 * math.pi, 4 * math.pi
    dx = sx / n[0]
    em = field_array(2, n, 6, dtype=np.float64)
    X00 = [c * dx for c in node_coords(2, n)]
    Xp0 = [c * dx for c in node_coords(2, n, (0.5, 0, 0))]
    X0p = [c * dx for c in node_coords(2, n, (0, 0.5, 0))]
    Xpp = [c * dx for c in node_coords(2, n, (0.5, 0.5, 0))]
    em[0] = np.sin(2 * Xp0[0]) * np.cos(Xp0[1]) / dx   # transform<1,T,U>: / sqrt(h_11)
    em[1] = np.cos(2 * X0p[0]) * np.cos(X0p[1]) / dx
    em[2] = np.cos(2 * X00[0]) * np.sin(X00[1])
    assert oracle.faraday(em, dim=2, coeff1=1.0 / dx**2, coeff2=1.0, prec="f64") == 0
    a = (slice(G, G + n[1]), slice(G, G + n[0]))
    bx1 = em[3][a] * dx
    bx2 = em[4][a] * dx
    bx3 = em[5][a]
    e1 = -np.cos(2 * X0p[0]) * np.cos(X0p[1])
    e2 = -2 * np.sin(2 * Xp0[0]) * np.sin(Xp0[1])
    e3 = 4 * np.cos(Xpp[0]) * np.cos(Xpp[1]) * np.sin(Xpp[0]) - \
        2 * np.cos(Xpp[0]) * np.sin(Xpp[0]) * np.sin(Xpp[1])
    max_err = dx**2 / 3
    assert np.abs(bx1 - e1[a]).max() < max_err
    assert np.abs(bx2 - e2[a]).max() < max_err
    assert np.abs(bx3 - e3[a]).max() < 3 * max_err


def test_faraday_3d_analytic(oracle):
    n = (32, 64, 32)
    sx = 2 * math.pi
    dx = sx / n[0]
    em = field_array(3, n, 6, dtype=np.float64)
    P = {c: [v * dx for v in node_coords(3, n, sh)] for c, sh in
         {0: (0.5, 0, 0), 1: (0, 0.5, 0), 2: (0, 0, 0.5),
          3: (0, 0.5, 0.5), 4: (0.5, 0, 0.5), 5: (0.5, 0.5, 0)}.items()}
    f = [lambda x, y, z: np.sin(2 * x) * np.cos(y) * np.cos(z),
         lambda x, y, z: np.cos(x) * np.sin(y) * np.cos(2 * z),
         lambda x, y, z: np.cos(x) * np.cos(2 * y) * np.sin(z)]
    for c in range(3):
        em[c] = f[c](*P[c]) / dx
    assert oracle.faraday(em, dim=3, coeff1=1.0 / dx, coeff2=0.0, prec="f64") == 0
    a = (slice(G, G + n[2]), slice(G, G + n[1]), slice(G, G + n[0]))
    h = 1e-6

    def d(fn, k, x, y, z):
        e = [0, 0, 0]
        e[k] = h
        return (fn(x + e[0], y + e[1], z + e[2]) - fn(x - e[0], y - e[1], z - e[2])) / (2 * h)
    # B_new = -curl E (stored contravariant: B^i = B_phys / dx)
    curl = [
        lambda x, y, z: d(f[2], 1, x, y, z) - d(f[1], 2, x, y, z),
        lambda x, y, z: d(f[0], 2, x, y, z) - d(f[2], 0, x, y, z),
        lambda x, y, z: d(f[1], 0, x, y, z) - d(f[0], 1, x, y, z),
    ]
    for c in range(3):
        got = em[3 + c][a] * dx
        exp = -curl[c](*P[3 + c])[a]
        assert np.abs(got - exp).max() < dx**2, c


def _div_B(em, dim):
    # Bx1 at (i, j+1/2, k+1/2), Bx2 at (i+1/2, j, k+1/2), Bx3 at (i+1/2, j+1/2, k):
    # divergence at cell centres (i+1/2, j+1/2, k+1/2)
    B = em[3:6]
    if dim == 2:
        return (B[0][:, 1:] - B[0][:, :-1])[:-1, :] + (B[1][1:, :] - B[1][:-1, :])[:, :-1]
    return ((B[0][:, :, 1:] - B[0][:, :, :-1])[:-1, :-1, :] +
            (B[1][:, 1:, :] - B[1][:, :-1, :])[:-1, :, :-1] +
            (B[2][1:, :, :] - B[2][:-1, :, :])[:, :-1, :-1])


@pytest.mark.parametrize("dim", [2, 3])
def test_faraday_extended_exact_for_linear_E(oracle, dim):
    n = (10, 9, 8)[:dim]
    fd = dict(delta_x=0.05, delta_y=-0.03, delta_z=0.02, beta_xy=0.01, beta_yx=0.04,
              beta_xz=-0.02, beta_zx=0.03, beta_yz=0.015, beta_zy=-0.01)
    em = field_array(dim, n, 6, dtype=np.float64)
    g = np.array([[0.3, -0.7, 0.4], [1.1, 0.2, -0.5], [-0.6, 0.8, 0.9]])  # dE_c/dx_d
    for c in range(3):
        X = node_coords(dim, n)
        em[c] = sum(g[c, d] * X[d] for d in range(dim))
    assert oracle.faraday(em, dim=dim, coeff1=1.0, coeff2=1.0 if dim == 2 else 0.0, fd=fd,
                          prec="f64") == 0
    gg = g.copy()
    if dim == 2:
        gg[:, 2] = 0.0
    curl = [gg[2, 1] - gg[1, 2], gg[0, 2] - gg[2, 0], gg[1, 0] - gg[0, 1]]
    core = tuple(slice(G + 1, G + x - 2) for x in reversed(n))
    for c in range(3):
        np.testing.assert_allclose(em[3 + c][core], -curl[c], atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_faraday_preserves_divB(oracle, dim):
    n = (10, 9, 8)[:dim]
    rng = np.random.default_rng(dim)
    em = rng.standard_normal(field_array(dim, n, 6).shape)
    fd = None
    # discrete div B is a face-to-cell operator; on interior cells it is
    # unchanged by Faraday since div curl = 0
    before = _div_B(em, dim).copy()
    c2 = 0.3 if dim == 2 else 0.0
    assert oracle.faraday(em, dim=dim, coeff1=0.3, coeff2=c2, fd=fd, prec="f64") == 0
    after = _div_B(em, dim)
    if dim == 2:
        core = (slice(G + 2, G + n[1] - 3), slice(G + 2, G + n[0] - 3))
    else:
        core = (slice(G + 2, G + n[2] - 3), slice(G + 2, G + n[1] - 3), slice(G + 2, G + n[0] - 3))
    np.testing.assert_allclose(after[core], before[core], atol=1e-12)


def test_faraday_zero_extended_equals_yee(oracle):
    rng = np.random.default_rng(0)
    n = (8, 8, 8)
    em = rng.standard_normal(field_array(3, n, 6).shape)
    em2 = em.copy()
    oracle.faraday(em, dim=3, coeff1=0.4, coeff2=0.0, prec="f64")
    oracle.faraday(em2, dim=3, coeff1=0.4, coeff2=0.0, fd=dict(), prec="f64")
    np.testing.assert_array_equal(em, em2)


def test_ampere_2d_analytic(oracle):
    n = (64, 128)
    dx = 2 * math.pi / n[0]
    em = field_array(2, n, 6, dtype=np.float64)
    P0p = [c * dx for c in node_coords(2, n, (0, 0.5, 0))]
    Pp0 = [c * dx for c in node_coords(2, n, (0.5, 0, 0))]
    Ppp = [c * dx for c in node_coords(2, n, (0.5, 0.5, 0))]
    P00 = [c * dx for c in node_coords(2, n)]
    em[3] = np.sin(2 * P0p[0]) * np.cos(P0p[1]) / dx
    em[4] = np.cos(2 * Pp0[0]) * np.cos(Pp0[1]) / dx
    em[5] = np.cos(2 * Ppp[0]) * np.sin(Ppp[1])
    assert oracle.ampere(em, dim=2, coeff1=1.0 / dx**2, coeff2=1.0, prec="f64") == 0
    a = (slice(G, G + n[1]), slice(G, G + n[0]))
    # E += curl B
    ex = em[0][a] * dx
    ey = em[1][a] * dx
    ez = em[2][a]
    exp_x = np.cos(2 * Pp0[0]) * np.cos(Pp0[1])                     # dBz/dy
    exp_y = 2 * np.sin(2 * P0p[0]) * np.sin(P0p[1])                 # -dBz/dx
    exp_z = -2 * np.sin(2 * P00[0]) * np.cos(P00[1]) + np.sin(2 * P00[0]) * np.sin(P00[1])
    max_err = dx**2
    assert np.abs(ex - exp_x[a]).max() < max_err
    assert np.abs(ey - exp_y[a]).max() < max_err
    assert np.abs(ez - exp_z[a]).max() < 3 * max_err


def test_currents_ampere(oracle):
    n = (6, 5)
    rng = np.random.default_rng(1)
    em = rng.standard_normal(field_array(2, n, 6).shape).astype(np.float32)
    J = rng.standard_normal(field_array(2, n, 3).shape).astype(np.float32)
    em0, J0 = em.copy(), J.copy()
    assert oracle.currents_ampere(em, J, dim=2, coeff=-0.25, ppc0=16.0) == 0
    a = (slice(G, G + n[1]), slice(G, G + n[0]))
    for c in range(3):
        np.testing.assert_array_equal(em[c][a], em0[c][a] + J0[c][a] * np.float32(-0.25))
        np.testing.assert_array_equal(J[c][a], J0[c][a] / np.float32(16.0))
    np.testing.assert_array_equal(em[3:], em0[3:])


def test_filter_2d_impulse(oracle):
    # tests/kernels/digital_filter.cpp:57-128
    n = (10, 10)
    J = field_array(2, n, 3, dtype=np.float64)
    J[0][5 + G, 5 + G] = 1.0
    J[1][5 + G, 4 + G] = 1.0
    J[2][4 + G, 5 + G] = 1.0
    buf = J.copy()
    assert oracle.filter_pass(J, buf, dim=2, prec="f64") == 0
    for c in range(3):
        assert abs(J[c].sum() - 1.0) < 1e-14
    for i1 in range(4, 7):
        for i2 in range(4, 7):
            expect = 0.5 ** (abs(i1 - 5) + 1) * 0.5 ** (abs(i2 - 5) + 1)
            assert J[0][i2 + G, i1 + G] == pytest.approx(expect)
            assert J[1][i2 + G, i1 + G - 1] == pytest.approx(expect)
            assert J[2][i2 + G - 1, i1 + G] == pytest.approx(expect)


def test_filter_3d_weights_as_written(oracle):
    n = (8, 8, 8)
    J = field_array(3, n, 3, dtype=np.float64)
    c = 4 + G
    J[:, c, c, c] = 1.0
    buf = J.copy()
    assert oracle.filter_pass(J, buf, dim=3, prec="f64") == 0
    w = J[0][c - 1:c + 2, c - 1:c + 2, c - 1:c + 2]  # [dz, dy, dx] offsets -1..1
    assert abs(J[0].sum() - 1.0) < 1e-14
    # the output at offset r receives the weight of input offset -r
    def weight(dx, dy, dz):
        return w[-dz + 1, -dy + 1, -dx + 1]
    assert weight(0, 0, 0) == 1 / 8
    for f in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]:
        assert weight(*f) == 1 / 16
    # (0,0,+-1) appear both as faces (1/16) and in the 1/32 group
    assert weight(0, 0, 1) == 1 / 16 + 1 / 32
    assert weight(0, 0, -1) == 1 / 16 + 1 / 32
    # the edge pair missing from the 1/32 group
    assert weight(0, -1, 1) == 0.0 and weight(0, 1, -1) == 0.0
    for e in [(-1, -1, 0), (1, 1, 0), (-1, 1, 0), (1, -1, 0), (0, -1, -1), (0, 1, 1),
              (-1, 0, -1), (1, 0, 1), (-1, 0, 1), (1, 0, -1)]:
        assert weight(*e) == 1 / 32
    for s in [(a, b, cc) for a in (-1, 1) for b in (-1, 1) for cc in (-1, 1)]:
        assert weight(*s) == 1 / 64


@pytest.mark.parametrize("dim", [2, 3])
def test_ghost_fill_is_periodic_extension(oracle, dim):
    n = (6, 5, 4)[:dim]
    rng = np.random.default_rng(2)
    F = rng.standard_normal(field_array(dim, n, 6).shape)
    assert oracle.communicate(F, 0, 6, dim=dim, prec="f64") == 0
    act = F[(slice(None),) + tuple(slice(G, G + x) for x in reversed(n))]
    ext = np.pad(act, [(0, 0)] + [(G, G)] * dim, mode="wrap")
    np.testing.assert_array_equal(F, ext)


@pytest.mark.parametrize("dim", [2, 3])
def test_fold_conserves_and_wraps(oracle, dim):
    n = (6, 5, 4)[:dim]
    rng = np.random.default_rng(4)
    J = rng.standard_normal(field_array(dim, n, 3).shape)
    buf = np.zeros_like(J)
    total = J.sum(axis=tuple(range(1, dim + 1)))
    # brute force: every extended entry lands on its periodic image
    expect = np.zeros((3,) + tuple(reversed(n)))
    idx = np.indices(J.shape[1:])
    src = J.reshape(3, -1)
    img = [((idx[k].ravel() - G) % list(reversed(n))[k]) for k in range(dim)]
    flat = np.ravel_multi_index(img, tuple(reversed(n)))
    for c in range(3):
        np.add.at(expect[c].reshape(-1), flat, src[c])
    assert oracle.synchronize(J, buf, dim=dim, prec="f64") == 0
    act = J[(slice(None),) + tuple(slice(G, G + x) for x in reversed(n))]
    np.testing.assert_allclose(act, expect, atol=1e-12)
    np.testing.assert_allclose(act.sum(axis=tuple(range(1, dim + 1))), total, atol=1e-10)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_threaded_step_matches_sequential_step(oracle, prec):
    """The OpenMP step (the CPU baseline in FP32, the finite-difference
    reference of the gradient checks in FP64) pushes the same particles and,
    up to the current's summation order, gives the same fields."""
    from _helpers import field_array, random_particles
    dt_ = np.float32 if prec == "f32" else np.float64
    tol = 1e-5 if prec == "f32" else 1e-12
    rng = np.random.default_rng(4)
    dim, n = 3, (12, 10, 8)
    em = (rng.standard_normal(field_array(dim, n, 6).shape) * 0.05).astype(dt_)
    out = []
    for threads in (False, True):
        sp = []
        for q in (-1.0, 1.0):
            p = random_particles(np.random.default_rng(5), dim, n, 4000, umax=0.5)
            sp.append({"p": oracle.normalize_particles(p, prec), "mass": 1.0, "charge": q,
                       "pusher": 4})
        e = em.copy()
        cur = field_array(dim, n, 3).astype(dt_)
        prm = dict(dt=0.28, correction=1.0, omegaB0=1.0, q0=0.01, B0=1.0, V0=1.0, ppc0=8.0,
                   dx=1.0, shape_order=2, n_filters=2, fieldsolver_enable=1, deposit_enable=1,
                   step=0)
        for s in range(3):
            prm["step"] = s
            assert oracle.sr_step(e, cur, np.zeros_like(cur), sp, dim=dim, prm=prm,
                                  bc=((1, 1),) * 3, prec=prec, threads=threads) == 0
        out.append((sp, e, cur))
    (sa, ea, ja), (sb, eb, jb) = out
    for a, b in zip(sa, sb):
        for k in ("i1", "i2", "i3", "dx1", "dx2", "dx3", "ux1", "ux2", "ux3"):
            np.testing.assert_allclose(a["p"][k], b["p"][k], rtol=tol, atol=tol / 10,
                                       err_msg=k)
    np.testing.assert_allclose(ja, jb, rtol=0, atol=tol * np.abs(ja).max())
    np.testing.assert_allclose(ea, eb, rtol=0, atol=tol * np.abs(ea).max())
    assert oracle.omp_threads() >= 1
