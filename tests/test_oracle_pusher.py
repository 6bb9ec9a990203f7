"""Pins the oracle's SR pusher against closed forms (CPU only).

* Gyration in a uniform B field for Boris and Vay, 2000 steps, as in the
  reference's tests/kernels/pusher.cpp:98-272 (same field, initial state,
  omegaB0, dt and tolerance 1e-2 (FP32) / 1e-3 (FP64)), for every shape order.
* |u| is conserved by Boris in a pure B field to round-off.
* Free streaming in zero field with periodic wrap (tests/kernels/prtl_bc.cpp).
* Interpolation reproduces constant and linear fields exactly on the Yee
  staggering, for every order (catches a primal/dual swap).
* Absorbing and reflecting boundaries (sr.hpp:659-790).
"""
import math

import numpy as np
import pytest

from _helpers import G, field_array, node_coords

BORIS, VAY, PHOTON = 2, 4, 1


def check_value(target, value, eps):
    # tests/kernels/pusher.cpp:28-42
    diff = abs(target - value)
    s = 0.5 * (abs(target) + abs(value))
    assert not (((s > eps) and (diff / s > eps)) or ((s <= eps) and (diff > eps / 10.0))), \
        (target, value)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_gyration_boris_vay(oracle, prec, order):
    n = (10, 10, 10)
    dt_ = np.float32 if prec == "f32" else np.float64
    em = field_array(3, n, 6, dtype=dt_)
    bx1, bx2, bx3 = 0.66, 0.55, 0.44
    em[3], em[4], em[5] = bx1, bx2, bx3
    b_mag = math.sqrt(bx1**2 + bx2**2 + bx3**2)
    x0 = (1.15, 1.85, 1.25)
    u0 = (1.0, -2.0, 0.1)
    gamma0 = math.sqrt(1 + sum(u * u for u in u0))
    omegaB0, dt = 0.2, 0.01
    eps = 1e-2 if prec == "f32" else 1e-3
    p = {}
    for d in range(3):
        p[f"i{d+1}"] = np.full(1, int(x0[d]), np.int32)
        p[f"dx{d+1}"] = np.full(1, x0[d] - math.floor(x0[d]), np.float32)
        p[f"i{d+1}_prev"] = np.zeros(1, np.int32)
        p[f"dx{d+1}_prev"] = np.zeros(1, np.float32)
        p[f"ux{d+1}"] = np.full(1, u0[d])
    p["weight"] = np.ones(1)
    p["tag"] = np.ones(1, np.int16)
    qb = oracle.normalize_particles(p, prec)
    qv = oracle.normalize_particles(p, prec)
    u0b = sum(a * b for a, b in zip(u0, (bx1, bx2, bx3))) / b_mag
    cr = ((u0[1] * bx3 - u0[2] * bx2) / b_mag, (u0[2] * bx1 - u0[0] * bx3) / b_mag,
          (u0[0] * bx2 - u0[1] * bx1) / b_mag)
    bb = (bx1, bx2, bx3)
    for t in range(2000):
        time = t * dt
        for q, alg in ((qb, BORIS), (qv, VAY)):
            _, rc = oracle.sr_push(q, em, dim=3, mass=1.0, charge=1.0, dt=dt, omegaB0=omegaB0,
                                   pusher=alg, order=order, prec=prec)
            assert rc == 0
        arg = (b_mag * omegaB0 * (time + dt)) / gamma0
        for q in (qb, qv):
            u = [float(q[f"ux{d+1}"][0]) for d in range(3)]
            check_value(math.sqrt(1 + sum(x * x for x in u)), gamma0, eps)
            for d in range(3):
                expect = (bb[d] / b_mag) * u0b + (-(bb[d] / b_mag) * u0b + u0[d]) * math.cos(arg) \
                    + cr[d] * math.sin(arg)
                check_value(u[d], expect, eps)


def test_boris_conserves_u_in_pure_B(oracle):
    n = (8, 8, 8)
    em = field_array(3, n, 6, dtype=np.float64)
    em[3], em[4], em[5] = 0.3, -1.2, 2.0
    rng = np.random.default_rng(7)
    from _helpers import random_particles
    p = random_particles(rng, 3, n, 64, umax=3.0, dtype=np.float64)
    q = oracle.normalize_particles(p, "f64")
    u2 = q["ux1"] ** 2 + q["ux2"] ** 2 + q["ux3"] ** 2
    for _ in range(50):
        _, rc = oracle.sr_push(q, em, dim=3, mass=1.0, charge=-1.0, dt=0.2, omegaB0=1.0,
                               pusher=BORIS, order=1, prec="f64")
        assert rc == 0
    u2n = q["ux1"] ** 2 + q["ux2"] ** 2 + q["ux3"] ** 2
    np.testing.assert_allclose(u2n, u2, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_free_streaming_periodic(oracle, dim, prec):
    n = (12, 9, 7)[:dim]
    em = field_array(dim, n, 6, dtype=np.float32 if prec == "f32" else np.float64)
    rng = np.random.default_rng(3)
    from _helpers import random_particles
    p = random_particles(rng, dim, n, 100, umax=2.0, dtype=np.float64)
    q = oracle.normalize_particles(p, prec)
    x0 = [q[f"i{d+1}"].astype(np.float64) + q[f"dx{d+1}"] for d in range(dim)]
    u = np.stack([q[f"ux{d+1}"].astype(np.float64) for d in range(3)])
    gam = np.sqrt(1 + (u ** 2).sum(0))
    dt, steps = 0.3, 60
    for _ in range(steps):
        _, rc = oracle.sr_push(q, em, dim=dim, mass=1.0, charge=1.0, dt=dt, omegaB0=1.0,
                               pusher=BORIS, order=1, prec=prec)
        assert rc == 0
    # dx is stored in single precision in both builds (prtldx_t, global.h:347)
    tol = 2e-3 if prec == "f32" else 1e-5
    for d in range(dim):
        x = q[f"i{d+1}"].astype(np.float64) + q[f"dx{d+1}"]
        xe = np.mod(x0[d] + u[d] / gam * dt * steps, n[d])
        dist = np.abs((x - xe + n[d] / 2) % n[d] - n[d] / 2)
        assert dist.max() < tol
        assert (q[f"i{d+1}"] >= 0).all() and (q[f"i{d+1}"] < n[d]).all()
        assert (q[f"dx{d+1}"] >= 0).all() and (q[f"dx{d+1}"] <= 1).all()
    np.testing.assert_array_equal(q["ux1"], p["ux1"].astype(q["ux1"].dtype))


STAG = {  # Yee staggering (fields.h:40-52): half-cell shift per direction
    0: (0.5, 0.0, 0.0), 1: (0.0, 0.5, 0.0), 2: (0.0, 0.0, 0.5),
    3: (0.0, 0.5, 0.5), 4: (0.5, 0.0, 0.5), 5: (0.5, 0.5, 0.0),
}


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_interpolation_exact_for_linear_fields(oracle, dim, order):
    n = (8, 7, 6)[:dim]
    em = field_array(dim, n, 6, dtype=np.float64)
    coef = np.array([[0.3, -0.7, 0.2, 0.5], [1.1, 0.4, -0.6, 0.1], [-0.2, 0.9, 0.35, 2.0],
                     [0.8, -0.3, 0.45, -1.0], [0.25, 0.6, -0.15, 0.7], [-0.9, 0.05, 0.3, 0.4]])
    for c in range(6):
        sh = STAG[c] if dim == 3 else (STAG[c][0], STAG[c][1], 0.0)
        X = node_coords(dim, n, sh)
        val = coef[c, 3] + sum(coef[c, d] * X[d] for d in range(dim))
        em[c] = val
    rng = np.random.default_rng(11)
    from _helpers import random_particles
    p = random_particles(rng, dim, n, 200, dtype=np.float64)
    e, b = oracle.interpolate(p, em, dim=dim, order=order, prec="f64")
    x = [p[f"i{d+1}"].astype(np.float64) + p[f"dx{d+1}"].astype(np.float64) for d in range(dim)]
    for c in range(6):
        expect = coef[c, 3] + sum(coef[c, d] * x[d] for d in range(dim))
        got = e[:, c] if c < 3 else b[:, c - 3]
        np.testing.assert_allclose(got, expect, rtol=0, atol=1e-12)


@pytest.mark.parametrize("bc_kind", ["absorb", "reflect"])
def test_absorb_and_reflect(oracle, bc_kind):
    n = (6, 6)
    em = field_array(2, n, 6, dtype=np.float32)
    p = {"i1": np.array([5, 0], np.int32), "dx1": np.array([0.95, 0.05], np.float32),
         "i2": np.array([2, 3], np.int32), "dx2": np.array([0.5, 0.5], np.float32),
         "i1_prev": np.zeros(2, np.int32), "i2_prev": np.zeros(2, np.int32),
         "dx1_prev": np.zeros(2, np.float32), "dx2_prev": np.zeros(2, np.float32),
         "ux1": np.array([3.0, -3.0]), "ux2": np.zeros(2), "ux3": np.zeros(2),
         "weight": np.ones(2), "tag": np.ones(2, np.int16)}
    code = 2 if bc_kind == "absorb" else 5
    q, rc = oracle.sr_push(p, em, dim=2, mass=1.0, charge=1.0, dt=0.5, omegaB0=1.0,
                           pusher=BORIS, order=1, bc=((code, code), (1, 1), (1, 1)))
    assert rc == 0
    if bc_kind == "absorb":
        assert list(q["tag"]) == [0, 0]
    else:
        assert list(q["tag"]) == [1, 1]
        assert list(q["i1"]) == [5, 0]
        assert q["ux1"][0] < 0 and q["ux1"][1] > 0
        # mirrored offsets: dx -> 1 - dx (sr.hpp:673-676)
        assert 0.0 <= q["dx1"][0] <= 1.0


def test_invalid_tag_reports(oracle):
    n = (4, 4)
    em = field_array(2, n, 6)
    p = {"i1": np.array([1], np.int32), "dx1": np.array([0.5], np.float32),
         "i2": np.array([1], np.int32), "dx2": np.array([0.5], np.float32),
         "i1_prev": np.zeros(1, np.int32), "i2_prev": np.zeros(1, np.int32),
         "dx1_prev": np.zeros(1, np.float32), "dx2_prev": np.zeros(1, np.float32),
         "ux1": np.zeros(1), "ux2": np.zeros(1), "ux3": np.zeros(1), "weight": np.ones(1),
         "tag": np.array([7], np.int16)}
    _, rc = oracle.sr_push(p, em, dim=2, mass=1.0, charge=1.0, dt=0.5, omegaB0=1.0,
                           pusher=BORIS, order=1)
    assert rc & 1
