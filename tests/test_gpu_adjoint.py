"""Adjoint of the 2D periodic SR step (ntt_sr_push_vjp, ntt_deposit_vjp,
ntt_yee_vjp and the checkpointed driver entity_amd.adjoint).

The reference has no adjoint, so these checks are fixed by mathematics:
* the field VJPs are the transposes of the forward operators run by the
  library: <lam, A x> = <A^T lam, x> for random x, lam (dot-product test);
* the particle VJPs match central finite differences of the forward kernels
  (ntt_sr_push, ntt_deposit_currents + fold) along random directions, with
  particles kept away from the kinks of the shape functions and from cell
  faces so that the finite difference stays on one smooth piece;
* the end-to-end gradient of L = sum |B|^2 / 2 after several steps of a
  counter-streaming (Weibel) run, with respect to the drift u_d and the
  amplitude b of a seed B_z = b cos(2 pi y / Ly), matches central finite
  differences of the forward engine; float and double adjoints agree, and
  the checkpoint interval does not change the result.
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

G = 2  # ghost cells


def _active(t, n):
    return t[:, G:G + n[1], G:G + n[0]]


def _stream():
    return torch.cuda.current_stream().cuda_stream


# ---------------------------------------------------------------- field transposes
@pytest.mark.parametrize("prec", [32, 64])
def test_yee_vjp_transposes(hip, prec):
    from entity_amd import _abi
    from entity_amd.containers import GridShape
    from entity_amd.kernels import communicate_fields, mink, DigitalFilter_kernel
    lib = _abi.load(False)
    n = (24, 20)
    shape = GridShape(2, n)
    nn = (C.c_int32 * 2)(*n)
    rdt = torch.float64 if prec == 64 else torch.float32
    tol = 1e-5 if prec == 64 else 2e-5
    g = torch.Generator().manual_seed(3)
    c1, c2 = 0.3, 0.7

    def rnd(*s):
        return torch.randn(*s, generator=g, dtype=torch.float64)

    def vjp(op, lin, a=c1, b=c2):
        lout = torch.zeros_like(lin)
        _abi.check(lib.ntt_yee_vjp(op, nn, a, b, prec, lin.data_ptr(), lout.data_ptr(),
                                   _stream()), "ntt_yee_vjp")
        return lout

    em0 = torch.zeros(shape.tensor_shape(6), dtype=torch.float32, device="cuda")
    # Faraday: B = M E
    x = rnd(3, n[1], n[0])
    em = em0.clone()
    _active(em, n)[0:3] = x.float().cuda()
    communicate_fields(em, shape, 0, 3)
    mink.Faraday_kernel(em, shape, c1, c2)()
    lam = rnd(3, n[1], n[0])
    lhs = float((lam * _active(em, n)[3:6].double().cpu()).sum())
    rhs = float((vjp(_abi.VJP_FARADAY, lam.to(rdt).cuda()).double().cpu() * x.float().double()).sum())
    assert abs(lhs - rhs) <= tol * (abs(lhs) + 1.0), (lhs, rhs)
    # Ampere: E = N B
    x = rnd(3, n[1], n[0])
    em = em0.clone()
    _active(em, n)[3:6] = x.float().cuda()
    communicate_fields(em, shape, 3, 6)
    mink.Ampere_kernel(em, shape, c1, c2)()
    lam = rnd(3, n[1], n[0])
    lhs = float((lam * _active(em, n)[0:3].double().cpu()).sum())
    rhs = float((vjp(_abi.VJP_AMPERE, lam.to(rdt).cuda()).double().cpu() * x.float().double()).sum())
    assert abs(lhs - rhs) <= tol * (abs(lhs) + 1.0), (lhs, rhs)
    # filter: cur = F buff
    x = rnd(3, n[1], n[0])
    buf = torch.zeros(shape.tensor_shape(3), dtype=torch.float32, device="cuda")
    cur = torch.zeros_like(buf)
    _active(buf, n)[:] = x.float().cuda()
    communicate_fields(buf, shape, 0, 3)
    DigitalFilter_kernel(cur, buf, shape)()
    lam = rnd(3, n[1], n[0])
    lhs = float((lam * _active(cur, n).double().cpu()).sum())
    rhs = float((vjp(_abi.VJP_FILTER, lam.to(rdt).cuda()).double().cpu() * x.float().double()).sum())
    assert abs(lhs - rhs) <= tol * (abs(lhs) + 1.0), (lhs, rhs)


def test_yee_vjp_rejects_bad_arguments(hip):
    from entity_amd import _abi
    lib = _abi.load(False)
    nn = (C.c_int32 * 2)(8, 8)
    a = torch.zeros(3 * 64, device="cuda")
    b = torch.zeros_like(a)
    assert lib.ntt_yee_vjp(7, nn, 1.0, 1.0, 32, a.data_ptr(), b.data_ptr(), None) == _abi.NTT_ERR_ARG
    assert lib.ntt_yee_vjp(0, nn, 1.0, 1.0, 16, a.data_ptr(), b.data_ptr(), None) == _abi.NTT_ERR_ARG
    assert lib.ntt_yee_vjp(0, nn, 1.0, 1.0, 32, a.data_ptr(), a.data_ptr(), None) == _abi.NTT_ERR_ARG


# ---------------------------------------------------------------- particle VJPs
def _smooth_mask(dx, h, order, staggered_too=True):
    """Offsets at least 4h from the cell faces and from 1/2 (the joints of the
    order-1 and order-2 shape pieces)."""
    m = (dx > 4 * h) & (dx < 1 - 4 * h)
    if staggered_too:
        m &= np.abs(dx - 0.5) > 4 * h
    return m


def _particles(n, npart, rng, spread=0.45):
    from entity_amd.containers import Particles
    p = Particles(2, npart, device="cuda")
    i1 = rng.integers(0, n[0], npart)
    i2 = rng.integers(0, n[1], npart)
    dx1 = rng.uniform(0, 1, npart)
    dx2 = rng.uniform(0, 1, npart)
    u = rng.standard_normal((3, npart)) * spread
    p.load(i1=i1, i2=i2, dx1=dx1.astype(np.float32), dx2=dx2.astype(np.float32),
           i1_prev=i1, i2_prev=i2, dx1_prev=dx1.astype(np.float32),
           dx2_prev=dx2.astype(np.float32), ux1=u[0].astype(np.float32),
           ux2=u[1].astype(np.float32), ux3=u[2].astype(np.float32),
           weight=rng.uniform(0.5, 1.5, npart).astype(np.float32),
           tag=np.ones(npart, np.int16))
    return p


def _padj(t5):
    from entity_amd import _abi
    return _abi.NttParticleAdjoint(*(x.data_ptr() for x in t5), t5[0].numel())


@pytest.mark.parametrize("order", [1, 2])
def test_push_vjp_matches_finite_differences(hip, order):
    """d/deps of sum lam_out . (x', u') along a random direction in (dx, u, E, B)."""
    from entity_amd import _abi
    from entity_amd.containers import GridShape, field_abi
    from entity_amd.kernels import Minkowski, PusherBoundaries, PusherContext, communicate_fields, sr
    n = (20, 16)
    shape = GridShape(2, n)
    rng = np.random.default_rng(10 + order)
    npart = 4000
    h = 1e-3
    base = _particles(n, npart, rng)
    st = base.to_host()
    ok = (_smooth_mask(st["dx1"], h, order) & _smooth_mask(st["dx2"], h, order))
    # smooth fields (the gather is linear in them, so any values work)
    yy, xx = np.meshgrid(np.arange(n[1] + 2 * G), np.arange(n[0] + 2 * G), indexing="ij")
    em_np = np.stack([0.3 * np.sin(2 * np.pi * (xx * (c + 1) + yy * (2 - c % 2)) / 16 + c)
                      for c in range(6)]).astype(np.float32)
    em0 = torch.from_numpy(em_np).cuda()
    communicate_fields(em0, shape, 0, 6)
    ctx = PusherContext(_abi.PUSHER_BORIS, 1.0, -1.0, 0.35, 2.0, n[0], n[1])
    metric = Minkowski(2, 1.0)
    v_dx = [np.where(ok, rng.uniform(-1, 1, npart), 0.0) for _ in range(2)]
    v_u = [rng.uniform(-1, 1, npart) for _ in range(3)]
    v_em = rng.uniform(-1, 1, (6, n[1], n[0]))
    lam = [rng.standard_normal(npart) for _ in range(5)]

    def run(eps):
        p = _particles(n, npart, np.random.default_rng(0))  # container, reloaded below
        p.load(i1=st["i1"], i2=st["i2"], i1_prev=st["i1"], i2_prev=st["i2"],
               dx1=(st["dx1"] + eps * v_dx[0]).astype(np.float32),
               dx2=(st["dx2"] + eps * v_dx[1]).astype(np.float32),
               dx1_prev=st["dx1"], dx2_prev=st["dx2"],
               ux1=(st["ux1"] + eps * v_u[0]).astype(np.float32),
               ux2=(st["ux2"] + eps * v_u[1]).astype(np.float32),
               ux3=(st["ux3"] + eps * v_u[2]).astype(np.float32),
               weight=st["weight"], tag=st["tag"])
        em = em0.clone()
        _active(em, n).add_(torch.from_numpy(eps * v_em).float().cuda())
        communicate_fields(em, shape, 0, 6)
        sr.Pusher_kernel(ctx, PusherBoundaries(), p, em, shape, metric, order)()
        o = p.to_host()
        x1 = o["i1"] + o["dx1"].astype(np.float64)
        x2 = o["i2"] + o["dx2"].astype(np.float64)
        return x1, x2, o["ux1"].astype(np.float64), o["ux2"].astype(np.float64), \
            o["ux3"].astype(np.float64)

    plus, minus = run(h), run(-h)
    fd = 0.0
    for k, (a, b) in enumerate(zip(plus, minus)):
        d = a - b
        if k < 2:  # unwrap periodic crossings
            L = n[k]
            d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
        fd += float((lam[k] * d).sum()) / (2 * h)
    # adjoint in double
    dev = "cuda"
    post = [torch.from_numpy(x).to(dev) for x in lam]
    pre = [torch.zeros(npart, dtype=torch.float64, device=dev) for _ in range(5)]
    lem = torch.zeros((6, n[1], n[0]), dtype=torch.float64, device=dev)
    p0 = _particles(n, npart, np.random.default_rng(0))
    p0.load(**{k: st[k] for k in st})
    lib = _abi.load(False)
    _abi.check(lib.ntt_sr_push_vjp(C.byref(ctx.abi()), C.byref(p0.abi()),
                                   C.byref(field_abi(em0, shape)), C.byref(metric.abi()), order,
                                   64, C.byref(_padj(post)), C.byref(_padj(pre)), lem.data_ptr(),
                                   _stream()), "ntt_sr_push_vjp")
    pre = [x.cpu().numpy() for x in pre]
    ad = float((pre[0] * v_dx[0]).sum() + (pre[1] * v_dx[1]).sum())
    ad += sum(float((pre[2 + c] * v_u[c]).sum()) for c in range(3))
    ad += float((lem.cpu().numpy() * v_em).sum())
    assert abs(ad - fd) <= 2e-3 * max(abs(fd), 1.0), (ad, fd)


@pytest.mark.parametrize("order", [1, 2])
def test_deposit_vjp_matches_finite_differences(hip, order):
    """d/deps of sum lam_J . J (folded) along a random direction in
    (dx_prev, dx, u) of particles that moved less than half a cell."""
    from entity_amd import _abi
    from entity_amd.containers import GridShape, Particles
    from entity_amd.kernels import DepositCurrents_kernel, Minkowski, synchronize_currents
    n = (18, 14)
    shape = GridShape(2, n)
    rng = np.random.default_rng(20 + order)
    npart = 3000
    h = 1e-3
    i0 = [rng.integers(0, n[0], npart), rng.integers(0, n[1], npart)]
    x0 = [rng.uniform(0, 1, npart), rng.uniform(0, 1, npart)]
    step = [rng.uniform(-0.45, 0.45, npart), rng.uniform(-0.45, 0.45, npart)]
    i1, x1 = [], []
    for d in range(2):
        x = x0[d] + step[d]
        fl = np.floor(x).astype(np.int64)
        i1.append(i0[d] + fl)
        x1.append(x - fl)
    # periodic wrap of i and i_prev together (as the pusher's boundary step does)
    for d in range(2):
        sh = np.where(i1[d] < 0, n[d], np.where(i1[d] >= n[d], -n[d], 0))
        i1[d] = i1[d] + sh
        i0[d] = i0[d] + sh
    u = rng.standard_normal((3, npart)) * 0.5
    w = rng.uniform(0.5, 1.5, npart)
    ok = np.ones(npart, bool)
    for d in range(2):
        ok &= _smooth_mask(x0[d], h, order) & _smooth_mask(x1[d], h, order)
    v = [np.where(ok, rng.uniform(-1, 1, npart), 0.0) for _ in range(4)]
    v_u = [rng.uniform(-1, 1, npart) for _ in range(3)]
    lamJ = rng.standard_normal((3, n[1], n[0]))
    charge, dt = -1.0, 0.35

    def fill(p, eps):
        p.load(i1=i1[0], i2=i1[1], i1_prev=i0[0], i2_prev=i0[1],
               dx1_prev=(x0[0] + eps * v[0]).astype(np.float32),
               dx2_prev=(x0[1] + eps * v[1]).astype(np.float32),
               dx1=(x1[0] + eps * v[2]).astype(np.float32),
               dx2=(x1[1] + eps * v[3]).astype(np.float32),
               ux1=(u[0] + eps * v_u[0]).astype(np.float32),
               ux2=(u[1] + eps * v_u[1]).astype(np.float32),
               ux3=(u[2] + eps * v_u[2]).astype(np.float32),
               weight=w.astype(np.float32), tag=np.ones(npart, np.int16))
        return p

    def run(eps):
        p = fill(Particles(2, npart, device="cuda"), eps)
        J = torch.zeros(shape.tensor_shape(3), dtype=torch.float32, device="cuda")
        B = torch.zeros_like(J)
        DepositCurrents_kernel(J, shape, p, Minkowski(2, 1.0), charge, dt, order)()
        synchronize_currents(J, B, shape)
        return _active(J, n).double().cpu().numpy()

    fd = float((lamJ * (run(h) - run(-h))).sum()) / (2 * h)
    p = fill(Particles(2, npart, device="cuda"), 0.0)
    post = [torch.zeros(npart, dtype=torch.float64, device="cuda") for _ in range(5)]
    pre = [torch.zeros(npart, dtype=torch.float64, device="cuda") for _ in range(5)]
    lj = torch.from_numpy(lamJ).cuda()
    lib = _abi.load(False)
    nn = (C.c_int32 * 2)(*n)
    _abi.check(lib.ntt_deposit_vjp(C.byref(p.abi()), nn, charge, dt, order, 64, lj.data_ptr(),
                                   C.byref(_padj(post)), C.byref(_padj(pre)), _stream()),
               "ntt_deposit_vjp")
    post = [x.cpu().numpy() for x in post]
    lxp = [x.cpu().numpy() for x in pre[:2]]
    ad = float((lxp[0] * v[0]).sum() + (lxp[1] * v[1]).sum() + (post[0] * v[2]).sum() +
               (post[1] * v[3]).sum())
    ad += sum(float((post[2 + c] * v_u[c]).sum()) for c in range(3))
    assert abs(ad - fd) <= 2e-3 * max(abs(fd), 1.0), (ad, fd)


@pytest.mark.parametrize("n", [(24, 16), (48, 40)])
@pytest.mark.parametrize("order", [1, 2])
def test_fused_particle_vjp_and_ids(hip, order, n):
    """ntt_sr_particle_vjp equals ntt_deposit_vjp followed by ntt_sr_push_vjp;
    with an ID payload the adjoint arrays are indexed by ID (a permutation of
    the particles gives the permuted adjoint); an out-of-range ID is reported."""
    from entity_amd import _abi
    from entity_amd.containers import GridShape, Particles, field_abi
    from entity_amd.kernels import (Minkowski, PusherBoundaries, PusherContext,
                                    communicate_fields, device_errors, sr)
    shape = GridShape(2, n)
    rng = np.random.default_rng(40 + order)
    npart = 3000
    pre = _particles(n, npart, rng)
    em = torch.from_numpy((rng.standard_normal((6, n[1] + 2 * G, n[0] + 2 * G)) * 0.2)
                          .astype(np.float32)).cuda()
    communicate_fields(em, shape, 0, 6)
    ctx = PusherContext(_abi.PUSHER_BORIS, 1.0, -1.0, 0.35, 2.0, n[0], n[1])
    metric = Minkowski(2, 1.0)
    post = Particles(2, npart, device="cuda").load(**{k: v for k, v in pre.to_host().items()})
    sr.Pusher_kernel(ctx, PusherBoundaries(), post, em, shape, metric, order)()
    lam_j = torch.from_numpy(rng.standard_normal((3, n[1], n[0]))).cuda()
    lam_post = [torch.from_numpy(rng.standard_normal(npart)).cuda() for _ in range(5)]
    lib = _abi.load(False)
    nn = (C.c_int32 * 2)(*n)
    fe = field_abi(em, shape)

    def zeros(k):
        return [torch.zeros(npart, dtype=torch.float64, device="cuda") for _ in range(k)]

    # separate: deposit VJP (adds to a copy of lam_post), then push VJP
    post_sep = [x.clone() for x in lam_post]
    pre_sep, lem_sep = zeros(5), torch.zeros((6, n[1], n[0]), dtype=torch.float64, device="cuda")
    _abi.check(lib.ntt_deposit_vjp(C.byref(post.abi()), nn, -1.0, 0.35, order, 64,
                                   lam_j.data_ptr(), C.byref(_padj(post_sep)),
                                   C.byref(_padj(pre_sep)), _stream()), "ntt_deposit_vjp")
    _abi.check(lib.ntt_sr_push_vjp(C.byref(ctx.abi()), C.byref(pre.abi()), C.byref(fe),
                                   C.byref(metric.abi()), order, 64, C.byref(_padj(post_sep)),
                                   C.byref(_padj(pre_sep)), lem_sep.data_ptr(), _stream()),
               "ntt_sr_push_vjp")
    # fused
    pre_fus, lem_fus = zeros(5), torch.zeros_like(lem_sep)
    _abi.check(lib.ntt_sr_particle_vjp(C.byref(ctx.abi()), C.byref(pre.abi()), C.byref(post.abi()),
                                       C.byref(fe), C.byref(metric.abi()), order, 64,
                                       lam_j.data_ptr(), C.byref(_padj(lam_post)),
                                       C.byref(_padj(pre_fus)), lem_fus.data_ptr(), _stream()),
               "ntt_sr_particle_vjp")
    for a, b in zip(pre_sep, pre_fus):
        np.testing.assert_allclose(b.cpu().numpy(), a.cpu().numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(lem_fus.cpu().numpy(), lem_sep.cpu().numpy(), rtol=1e-11,
                               atol=1e-11 * float(lem_sep.abs().max()))
    # IDs: particle k carries ID perm[k]; adjoints live at their IDs
    perm = torch.from_numpy(rng.permutation(npart)).cuda()
    pre.pld_i = perm.clone()
    post.pld_i = perm.clone()
    lam_post_id = [torch.empty_like(x) for x in lam_post]
    for a, b in zip(lam_post_id, lam_post):
        a[perm] = b
    pre_id, lem_id = zeros(5), torch.zeros_like(lem_sep)
    _abi.check(lib.ntt_sr_particle_vjp(C.byref(ctx.abi()), C.byref(pre.abi()), C.byref(post.abi()),
                                       C.byref(fe), C.byref(metric.abi()), order, 64,
                                       lam_j.data_ptr(), C.byref(_padj(lam_post_id)),
                                       C.byref(_padj(pre_id)), lem_id.data_ptr(), _stream()),
               "ntt_sr_particle_vjp")
    for a, b in zip(pre_fus, pre_id):
        np.testing.assert_allclose(b[perm].cpu().numpy(), a.cpu().numpy(), rtol=1e-12, atol=1e-12)
    assert device_errors() == 0
    # an ID beyond the adjoint arrays: that particle is skipped and reported
    pre.pld_i[5] = npart + 3
    pre_bad = zeros(5)
    _abi.check(lib.ntt_sr_particle_vjp(C.byref(ctx.abi()), C.byref(pre.abi()), C.byref(post.abi()),
                                       C.byref(fe), C.byref(metric.abi()), order, 64,
                                       lam_j.data_ptr(), C.byref(_padj(lam_post_id)),
                                       C.byref(_padj(pre_bad)), lem_id.data_ptr(), _stream()),
               "ntt_sr_particle_vjp")
    assert device_errors() & _abi.DEVERR_OUT_OF_BOUNDS


def test_particle_vjps_reject_bad_arguments(hip):
    from entity_amd import _abi
    from entity_amd.containers import Particles
    lib = _abi.load(False)
    p = Particles(2, 8, device="cuda")
    p.npart = 8
    a = [torch.zeros(8, device="cuda") for _ in range(5)]
    b = [torch.zeros(8, device="cuda") for _ in range(5)]
    nn = (C.c_int32 * 2)(8, 8)
    lj = torch.zeros(3 * 64, device="cuda")
    st = lib.ntt_deposit_vjp(C.byref(p.abi()), nn, 1.0, 0.1, 3, 32, lj.data_ptr(),
                             C.byref(_padj(a)), C.byref(_padj(b)), None)
    assert st == _abi.NTT_ERR_UNSUPPORTED
    st = lib.ntt_deposit_vjp(C.byref(p.abi()), nn, 1.0, 0.1, 1, 32, None,
                             C.byref(_padj(a)), C.byref(_padj(b)), None)
    assert st == _abi.NTT_ERR_ARG
    # adjoint arrays indexed by position must hold every particle
    short = _abi.NttParticleAdjoint(*(x.data_ptr() for x in a), 4)
    st = lib.ntt_deposit_vjp(C.byref(p.abi()), nn, 1.0, 0.1, 1, 32, lj.data_ptr(),
                             C.byref(short), C.byref(_padj(b)), None)
    assert st == _abi.NTT_ERR_ARG


# ---------------------------------------------------------------- end to end
N_BOX = (32, 24)
SIGN = (1.0, -1.0, 0.0)


def _weibel(ud, b, order=2, n_filters=1, sort_interval=0):
    """Two counter-streaming electron beams (u_x shifted by +-u_d) over a
    positron background, with a seed B_z = b cos(2 pi y / Ly).  With a
    re-sort interval the particles carry IDs."""
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    prm = SRParams(dim=2, n=N_BOX, shape_order=order, n_filters=n_filters, ppc0=8.0,
                   larmor0=1.0, skindepth0=3.0, sort_interval=sort_interval, tile=(8, 8))
    sp = [SpeciesSpec("e1", 1.0, -1.0), SpeciesSpec("e2", 1.0, -1.0), SpeciesSpec("p", 1.0, 1.0)]
    ppc = 4
    eng = SRPICEngine(prm, sp, N_BOX[0] * N_BOX[1] * ppc, with_payload=sort_interval > 0)
    eng.inject_uniform(ppc, temperature=0.01, seed=5)
    for s, sp in zip(SIGN, eng.species):   # u_x += sign * u_d on a thermal plasma
        sp.ux1[: sp.npart] += s * ud
    j = torch.arange(N_BOX[1], dtype=torch.float64) + 0.5
    bz = (b * torch.cos(2 * math.pi * j / N_BOX[1])).float().cuda()
    _active(eng.fields.em, N_BOX)[5] = bz[:, None].expand(N_BOX[1], N_BOX[0])
    return eng


def _loss(ud, b, nsteps, **kw):
    from entity_amd.adjoint import SRAdjoint
    eng = _weibel(ud, b, **kw)
    return SRAdjoint(eng, nsteps).forward()


def _grads(ud, b, nsteps, checkpoint, precision, per_particle=False, **kw):
    from entity_amd.adjoint import SRAdjoint
    eng = _weibel(ud, b, **kw)
    loss, lam, lam_u = SRAdjoint(eng, nsteps, checkpoint, precision).gradient()
    j = torch.arange(N_BOX[1], dtype=torch.float64) + 0.5
    seed = torch.cos(2 * math.pi * j / N_BOX[1])[:, None].expand(N_BOX[1], N_BOX[0])
    g_b = float((lam[5].double().cpu() * seed).sum())
    g_ud = sum(s * float(lu[0].double().sum()) for s, lu in zip(SIGN, lam_u))
    if per_particle:
        return loss, g_ud, g_b, [torch.stack(lu).cpu().numpy() for lu in lam_u]
    return loss, g_ud, g_b


@pytest.mark.parametrize("order", [1, 2])
def test_weibel_gradient_matches_finite_differences(hip, order):
    ud, b, nsteps, h = 0.4, 0.05, 10, 1e-3
    loss, g_ud, g_b = _grads(ud, b, nsteps, 4, 64, order=order)
    assert loss == pytest.approx(_loss(ud, b, nsteps, order=order), rel=1e-12)
    fd_ud = (_loss(ud + h, b, nsteps, order=order) - _loss(ud - h, b, nsteps, order=order)) / (2 * h)
    fd_b = (_loss(ud, b + h, nsteps, order=order) - _loss(ud, b - h, nsteps, order=order)) / (2 * h)
    print(f"order {order}: L={loss:.6e} dL/dud adj={g_ud:.6e} fd={fd_ud:.6e} "
          f"dL/db adj={g_b:.6e} fd={fd_b:.6e}", flush=True)
    assert g_b == pytest.approx(fd_b, rel=2e-3)
    assert g_ud == pytest.approx(fd_ud, rel=1e-2, abs=1e-6 * abs(g_b))


def test_adjoint_precisions_and_checkpoints_agree(hip):
    ud, b, nsteps = 0.4, 0.05, 9
    ref = _grads(ud, b, nsteps, 4, 64)
    one = _grads(ud, b, nsteps, 1, 64)
    full = _grads(ud, b, nsteps, nsteps, 64)
    f32 = _grads(ud, b, nsteps, 4, 32)
    for other in (one, full):
        assert other[0] == ref[0]
        assert other[1] == pytest.approx(ref[1], rel=1e-9)
        assert other[2] == pytest.approx(ref[2], rel=1e-9)
    assert f32[1] == pytest.approx(ref[1], rel=1e-3)
    assert f32[2] == pytest.approx(ref[2], rel=1e-3)


def test_gradient_through_resorts(hip):
    """Re-sorting every 2 steps (inside segments and at their starts)
    permutes the particles; with IDs the adjoint follows them, and since the
    forward does not depend on particle order
    (integer current sums) the loss is identical and the gradients agree to
    round-off with the run that never re-sorts."""
    from entity_amd.kernels import device_errors
    ud, b, nsteps = 0.4, 0.05, 10
    ref = _grads(ud, b, nsteps, 4, 64, per_particle=True)
    srt = _grads(ud, b, nsteps, 4, 64, per_particle=True, sort_interval=2)
    assert device_errors() == 0
    assert srt[0] == ref[0]
    assert srt[1] == pytest.approx(ref[1], rel=1e-9)
    assert srt[2] == pytest.approx(ref[2], rel=1e-9)
    # every particle's initial-momentum adjoint, indexed by ID in both runs
    for a, r in zip(srt[3], ref[3]):
        np.testing.assert_allclose(a, r, rtol=1e-9, atol=1e-12 * float(np.abs(r).max()))


def test_autograd_function(hip):
    from entity_amd.adjoint import FieldEnergyLoss

    def setup(ud, b):
        return _weibel(ud, b, n_filters=0)

    j = torch.arange(N_BOX[1], dtype=torch.float64) + 0.5
    profile = torch.cos(2 * math.pi * j / N_BOX[1])[:, None].expand(N_BOX[1], N_BOX[0])
    ud = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(0.04, dtype=torch.float64, requires_grad=True)
    L = FieldEnergyLoss.apply(ud, b, setup, SIGN, profile, 6, 3, 64)
    (2.0 * L).backward()
    ref = _grads(0.3, 0.04, 6, 3, 64, n_filters=0)
    assert float(L.detach()) == pytest.approx(ref[0], rel=1e-12)
    assert float(ud.grad) == pytest.approx(2.0 * ref[1], rel=1e-9)
    assert float(b.grad) == pytest.approx(2.0 * ref[2], rel=1e-9)


def test_adjoint_rejects_unsupported_setups(hip):
    from entity_amd.adjoint import SRAdjoint
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    prm = SRParams(dim=2, n=(16, 16), shape_order=0, sort_interval=0)
    eng = SRPICEngine(prm, [SpeciesSpec("e", 1.0, -1.0)], 16)
    with pytest.raises(ValueError):
        SRAdjoint(eng, 2)
    prm = SRParams(dim=2, n=(16, 16), shape_order=1, sort_interval=4)
    eng = SRPICEngine(prm, [SpeciesSpec("e", 1.0, -1.0)], 16)
    with pytest.raises(ValueError):   # re-sorting without particle IDs
        SRAdjoint(eng, 2)
