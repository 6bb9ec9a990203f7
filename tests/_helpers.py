"""Shared helpers for the tests: textbook B-splines, field/particle builders.

The B-splines here are written from their closed forms (the same closed
forms quoted in the comments of src/kernels/particle_shapes.hpp:20-45), not
from the oracle's code, so that checks built on them are independent.
"""
from __future__ import annotations

import numpy as np

G = 2  # ghost layers


def bspline(order: int, x):
    """Centered B-spline of the given order evaluated at x (vectorized)."""
    a = np.abs(np.asarray(x, dtype=np.float64))
    if order in (0, 1):
        return np.where(a < 1.0, 1.0 - a, 0.0)
    if order == 2:
        return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    raise ValueError(order)


def field_array(dim, n, ncomp, dtype=np.float32):
    ext = [x + 2 * G for x in n]
    return np.zeros((ncomp,) + tuple(reversed(ext)), dtype=dtype)


def node_coords(dim, n, shift=(0.0, 0.0, 0.0)):
    """Code-unit coordinates (ghost-extended grid) of every array entry: entry
    g holds position g - G + shift."""
    ext = [x + 2 * G for x in n]
    axes = [np.arange(e, dtype=np.float64) - G + shift[d] for d, e in enumerate(ext)]
    grids = np.meshgrid(*axes, indexing="ij")
    # transpose to (z, y, x) / (y, x) storage order
    return [g.T for g in grids]


def random_particles(rng, dim, n, npart, *, umax=1.0, margin=0, dtype=np.float32):
    lo = margin
    i = [rng.integers(lo, n[d] - lo, npart).astype(np.int32) for d in range(dim)]
    dx = [rng.random(npart, dtype=np.float32) for _ in range(dim)]
    u = [(rng.standard_normal(npart) * umax).astype(dtype) for _ in range(3)]
    p = {}
    for d in range(dim):
        p[f"i{d+1}"] = i[d]
        p[f"dx{d+1}"] = dx[d]
        p[f"i{d+1}_prev"] = i[d].copy()
        p[f"dx{d+1}_prev"] = dx[d].copy()
    for d in range(3):
        p[f"ux{d+1}"] = u[d]
    p["weight"] = np.ones(npart, dtype=dtype)
    p["tag"] = np.ones(npart, dtype=np.int16)
    return p


def move_particles(p, dim, disp):
    """Applies in-cell displacements the way the reference's position update
    does (sr.hpp:540-547), keeping *_prev as the starting state."""
    q = {k: v.copy() for k, v in p.items()}
    for d in range(dim):
        k = d + 1
        q[f"i{k}_prev"] = p[f"i{k}"].copy()
        q[f"dx{k}_prev"] = p[f"dx{k}"].copy()
        dx = (p[f"dx{k}"] + disp[d].astype(np.float32)).astype(np.float32)
        i = p[f"i{k}"] + (dx >= 1).astype(np.int32) - (dx < 0).astype(np.int32)
        dx = (dx - (dx >= 1)).astype(np.float32)
        dx = (dx + (dx < 0)).astype(np.float32)
        q[f"i{k}"] = i.astype(np.int32)
        q[f"dx{k}"] = dx
    return q


def charge_density(p, dim, n, order, *, prev=False, charge=1.0):
    """rho on primal nodes of the ghost-extended grid: sum_p q w S(node - x_p)."""
    rho = np.zeros(tuple(reversed([x + 2 * G for x in n])), dtype=np.float64)
    suf = "_prev" if prev else ""
    x = [p[f"i{d+1}{suf}"].astype(np.float64) + p[f"dx{d+1}{suf}"].astype(np.float64)
         for d in range(dim)]
    w = p["weight"].astype(np.float64) * charge
    reach = 2
    for idx in range(len(w)):
        base = [int(np.floor(x[d][idx])) for d in range(dim)]
        ranges = [range(base[d] - reach, base[d] + reach + 1) for d in range(dim)]
        if dim == 2:
            for a in ranges[0]:
                sa = bspline(order, a - x[0][idx])
                if sa == 0:
                    continue
                for b in ranges[1]:
                    sb = bspline(order, b - x[1][idx])
                    if sb:
                        rho[b + G, a + G] += w[idx] * sa * sb
        else:
            for a in ranges[0]:
                sa = bspline(order, a - x[0][idx])
                if sa == 0:
                    continue
                for b in ranges[1]:
                    sb = bspline(order, b - x[1][idx])
                    if sb == 0:
                        continue
                    for c in ranges[2]:
                        sc = bspline(order, c - x[2][idx])
                        if sc:
                            rho[c + G, b + G, a + G] += w[idx] * sa * sb * sc
    return rho


def divergence_J(J, dim):
    """Backward-difference divergence at primal nodes (dx = 1), as in
    tests/kernels/deposit.cpp:163-173."""
    J = J.astype(np.float64)
    div = np.zeros(J.shape[1:], dtype=np.float64)
    if dim == 2:
        div[:, 1:] += J[0][:, 1:] - J[0][:, :-1]
        div[:, 0] += J[0][:, 0]
        div[1:, :] += J[1][1:, :] - J[1][:-1, :]
        div[0, :] += J[1][0, :]
    else:
        div[:, :, 1:] += J[0][:, :, 1:] - J[0][:, :, :-1]
        div[:, :, 0] += J[0][:, :, 0]
        div[:, 1:, :] += J[1][:, 1:, :] - J[1][:, :-1, :]
        div[:, 0, :] += J[1][:, 0, :]
        div[1:, :, :] += J[2][1:, :, :] - J[2][:-1, :, :]
        div[0, :, :] += J[2][0, :, :]
    return div


def oracle_field_energy(eng, nsteps, *, dux=(), dbz=0.0):
    """L = sum |B|^2 / 2 over the active cells after `nsteps` steps of the
    oracle in FP64 (OpenMP threads), started from the 2D engine's current
    state with species s's u_x shifted by dux[s] and the active B_z shifted
    by dbz: the finite-difference reference of the adjoint checks, with
    momenta and fields in double (offsets in float, as in the reference's
    FP64 build)."""
    import ntt_oracle
    p = eng.params
    n1, n2 = p.n
    em = eng.fields.em.detach().cpu().numpy().astype(np.float64)
    em[5, G:G + n2, G:G + n1] += dbz
    cur = np.zeros((3,) + em.shape[1:], np.float64)
    sp = []
    for k, s in enumerate(eng.species):
        q = ntt_oracle.normalize_particles(s.to_host(), "f64")
        if k < len(dux):
            q["ux1"] += dux[k]
        sp.append({"p": q, "mass": s.mass, "charge": s.charge, "pusher": s.pusher})
    for k in range(nsteps):
        prm = dict(dt=p.dt, correction=p.correction, omegaB0=p.omegaB0, q0=p.q0, B0=p.B0,
                   V0=p.V0, ppc0=p.ppc0, dx=p.dx, shape_order=p.shape_order,
                   n_filters=p.n_filters, fieldsolver_enable=1, deposit_enable=1,
                   step=eng.step + k)
        assert ntt_oracle.sr_step(em, cur, np.zeros_like(cur), sp, dim=2, prm=prm,
                                  bc=((1, 1),) * 3, prec="f64", threads=True) == 0
    return 0.5 * float((em[3:6, G:G + n2, G:G + n1] ** 2).sum())
