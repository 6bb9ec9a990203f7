"""Pins the oracle's current deposit (CPU only).

* Pointwise charge continuity rho^{n+1} - rho^n + dt div J = 0 on every node,
  with rho built from textbook B-splines of the deposit order (independent of
  the oracle's shape code) -- zig-zag (O = 0, linear charge shape) and
  Esirkepov O = 1, 2 in 2D and 3D.  The 3D zig-zag weights each current
  component by the product of the other two midpoint weights and carries no
  dy*dz/12 correction, so it conserves charge exactly only for moves in a
  coordinate plane; general 3D moves are checked against the size of that
  missing term instead.
* The closed-form zig-zag values of tests/kernels/deposit.cpp:85-127 (same
  particle, same expected per-cell sums), including the out-of-plane Jz.
* sum(div J) == 0 (tests/kernels/deposit.cpp:163-196).
* Out-of-plane Esirkepov current sums to q w v_z * 3 * THIRD (THIRD = 0.333333
  in single precision, src/global/utils/numeric.h:41).
"""
import numpy as np
import pytest

from _helpers import G, charge_density, divergence_J, field_array, move_particles, random_particles


def _moved(dim, n, npart, seed, maxdisp=0.45, planar=False):
    rng = np.random.default_rng(seed)
    p = random_particles(rng, dim, n, npart, margin=3)
    disp = [rng.uniform(-maxdisp, maxdisp, npart) for _ in range(dim)]
    if planar:
        zero = rng.integers(0, dim, npart)
        for d in range(dim):
            disp[d][zero == d] = 0.0
    return move_particles(p, dim, disp), rng


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_continuity_pointwise(oracle, dim, order):
    n = (12, 11, 10)[:dim]
    planar = dim == 3 and order == 0
    q, rng = _moved(dim, n, 60 if dim == 3 else 200, seed=10 + order + dim, planar=planar)
    q["weight"] = rng.uniform(0.5, 2.0, len(q["weight"]))
    J = field_array(dim, n, 3, dtype=np.float64)
    dt, charge = 0.37, -1.3
    rc = oracle.deposit(q, J, dim=dim, charge=charge, dt=dt, order=order, prec="f64")
    assert rc == 0
    shape_order = 1 if order == 0 else order
    rho0 = charge_density(q, dim, n, shape_order, prev=True, charge=charge)
    rho1 = charge_density(q, dim, n, shape_order, prev=False, charge=charge)
    resid = rho1 - rho0 + dt * divergence_J(J, dim)
    # the zig-zag relay point and its sums with dx_prev are formed in single
    # precision even in the double build (prtldx_t arithmetic,
    # currents_deposit.hpp:175-181), which bounds its continuity to ~1e-7
    tol = 1e-6 if order == 0 else 1e-12
    assert np.abs(resid).max() < tol * max(1.0, np.abs(rho1).max())
    # and the divergence is not trivially zero
    assert np.abs(rho1 - rho0).max() > 1e-3


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_continuity_single_precision(oracle, dim, order):
    n = (12, 11, 10)[:dim]
    q, _ = _moved(dim, n, 40, seed=5 + order, planar=(dim == 3 and order == 0))
    J = field_array(dim, n, 3, dtype=np.float32)
    dt = 0.35
    assert oracle.deposit(q, J, dim=dim, charge=1.0, dt=dt, order=order, prec="f32") == 0
    so = 1 if order == 0 else order
    resid = charge_density(q, dim, n, so) - charge_density(q, dim, n, so, prev=True) \
        + dt * divergence_J(J, dim)
    assert np.abs(resid).max() < 2e-5


def test_zigzag_3d_general_moves_miss_only_the_cross_term(oracle):
    # per node, the continuity residual is bounded by the omitted
    # sum_c q w |d_a d_b| / 12 terms of the particles whose stencil covers it
    n = (12, 11, 10)
    q, _ = _moved(3, n, 1, seed=21)
    J = field_array(3, n, 3, dtype=np.float64)
    assert oracle.deposit(q, J, dim=3, charge=1.0, dt=1.0, order=0, prec="f64") == 0
    resid = charge_density(q, 3, n, 1) - charge_density(q, 3, n, 1, prev=True) \
        + divergence_J(J, 3)
    d = [abs((q[f"i{k}"] + q[f"dx{k}"].astype(np.float64)) -
             (q[f"i{k}_prev"] + q[f"dx{k}_prev"].astype(np.float64)))[0] for k in (1, 2, 3)]
    bound = (d[0] * d[1] + d[1] * d[2] + d[0] * d[2]) / 12 * 2 + 1e-9
    assert 0 < np.abs(resid).max() <= bound


def test_zigzag_closed_form_reference_case(oracle):
    # tests/kernels/deposit.cpp:85-127: particle from (25.65, 21.65) to (24.99, 20.80)
    n = (50, 50)
    i0, j0, i0f, j0f = 25, 21, 24, 20
    dxi, dxf, dyi, dyf = np.float32(0.65), np.float32(0.99), np.float32(0.65), np.float32(0.80)
    uz = 2.5
    p = {"i1": np.array([i0f], np.int32), "i2": np.array([j0f], np.int32),
         "i1_prev": np.array([i0], np.int32), "i2_prev": np.array([j0], np.int32),
         "dx1": np.array([dxf]), "dx2": np.array([dyf]),
         "dx1_prev": np.array([dxi]), "dx2_prev": np.array([dyi]),
         "ux1": np.zeros(1), "ux2": np.zeros(1), "ux3": np.array([uz]),
         "weight": np.ones(1), "tag": np.ones(1, np.int16)}
    J = field_array(2, n, 3, dtype=np.float64)
    assert oracle.deposit(p, J, dim=2, charge=1.0, dt=1.0, order=0, prec="f64") == 0
    xi, xf = i0 + float(dxi), i0f + float(dxf)
    yi, yf = j0 + float(dyi), j0f + float(dyf)
    xr, yr = 0.5 * (xi + xf), 0.5 * (yi + yf)
    Wx1, Wx2 = 0.5 * (xi + xr) - i0, 0.5 * (xf + xr) - i0
    Wy1, Wy2 = 0.5 * (yi + yr) - j0, 0.5 * (yf + yr) - j0
    Fx1, Fx2 = xr - xi, xf - xr
    Fy1, Fy2 = yr - yi, yf - yr
    Fz = 0.5 * uz / np.sqrt(1 + uz * uz)
    Jx = Fx1 * (1 - Wy1) + Fx2 * (1 - Wy2) + Fx1 * Wy1 + Fx2 * Wy2
    Jy = Fy1 * (1 - Wx1) + Fy2 * (1 - Wx2) + Fy1 * Wx1 + Fy2 * Wx2
    Jz = 2 * Fz
    act = (slice(G, G + n[1]), slice(G, G + n[0]))
    np.testing.assert_allclose(J[0][act].sum(), Jx, rtol=1e-12)
    np.testing.assert_allclose(J[1][act].sum(), Jy, rtol=1e-12)
    np.testing.assert_allclose(J[2][act].sum(), Jz, rtol=1e-12)
    # sum of div J over the active region is zero (deposit.cpp:163-196)
    sl = J[:, G:G + n[1], G:G + n[0]]
    prev1 = J[0][G:G + n[1], G - 1:G + n[0] - 1]
    prev2 = J[1][G - 1:G + n[1] - 1, G:G + n[0]]
    assert abs((sl[0] - prev1 + sl[1] - prev2).sum()) < 1e-12


@pytest.mark.parametrize("order", [1, 2])
def test_esirkepov_out_of_plane_sum(oracle, order):
    n = (10, 10)
    q, _ = _moved(2, n, 20, seed=order)
    q["ux3"] = np.linspace(-2, 2, 20)
    J = field_array(2, n, 3, dtype=np.float32)
    assert oracle.deposit(q, J, dim=2, charge=1.0, dt=0.4, order=order, prec="f32") == 0
    u = np.stack([q["ux1"], q["ux2"], q["ux3"]]).astype(np.float64)
    vz = u[2] / np.sqrt(1 + (u ** 2).sum(0))
    expect = (vz * 3 * np.float64(np.float32(0.333333))).sum()
    np.testing.assert_allclose(J[2].astype(np.float64).sum(), expect, rtol=1e-5, atol=1e-6)


def test_deposit_skips_dead(oracle):
    n = (8, 8)
    q, _ = _moved(2, n, 10, seed=3)
    q["tag"] = np.zeros(10, np.int16)
    J = field_array(2, n, 3, dtype=np.float32)
    assert oracle.deposit(q, J, dim=2, charge=1.0, dt=0.4, order=1) == 0
    assert not J.any()
