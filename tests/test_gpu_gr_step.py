"""GPU parity of the full GRPIC step's pieces and of the step itself.

* Match layer, horizon and axis rows, the current's absorbing layer and the
  spherical filter (ntt_gr_match_boundaries, _horizon_, _axis_,
  ntt_gr_absorb_currents, ntt_gr_digital_filter) against the oracle on random
  FP32 fields: the same values to FP32 round-off (the copies exactly).
* The GR deposit (ntt_gr_deposit_currents) against the oracle's FP64
  deposit with the same moves: in-plane currents to FP32 round-off of the
  SR scheme, the out-of-plane one to the metric's round-off in the
  precision asked for.
* GRPICEngine against the oracle's GRStep: Wald's field on a QKS grid with a
  pair plasma, deposit, one filter pass, match layers and the absorbing
  layer, 1 and 6 steps: fields within 2e-5 of their maximum after 1 step and
  2e-4 after 6 (plus a few ulp of |B| for D, which grows from curls of
  B-sized terms), particles as a multiset with momenta to 1e-4.
* The discrete Gauss law holds on the GPU over the steps to FP32 round-off.
"""
import math

import numpy as np
import pytest
import torch

from test_oracle_gr_fields import G, _metric, wald_fields

pytestmark = pytest.mark.gpu

KINDS = {"kerr_schild": 0, "qkerr_schild": 1, "kerr_schild_0": 2}


def _setup(oracle, kind, n, a=0.9, rmax=12.0, h=0.0):
    from entity_amd.containers import GridShape
    from entity_amd.kernels import gr
    rh = 1 + math.sqrt(1 - a * a) if kind != "kerr_schild_0" else 1.0
    rmin = 0.95 * rh if kind != "kerr_schild_0" else 1.0
    prm = oracle.gr_params(kind, n, ((rmin, rmax), (0.0, math.pi)), a=a, h=h)
    met = gr.KerrSchild(KINDS[kind], n, (rmin, rmax), (0.0, math.pi), a=a, h=h)
    return prm, met, GridShape(2, n)


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def _close(got, ref, what, rtol=2e-6):
    scale = max(np.abs(ref).max(), 1e-30)
    np.testing.assert_allclose(got.astype(np.float64), ref, rtol=rtol, atol=rtol * scale,
                               err_msg=what)


AXIS, MATCH = 7, 2
BCS = [(MATCH, MATCH), (AXIS, AXIS)]


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("upper", [True, False])
def test_match_layer(hip, oracle, kind, upper):
    from entity_amd.grpic import match_range
    from entity_amd.kernels import device_errors, gr
    n = (48, 24)
    prm, met, shape = _setup(oracle, kind, n)
    ds = 0.2 * (prm.x1_max - prm.x1_min)
    lo, hi, edge = match_range(met, upper, ds)
    assert (lo, hi, edge) == oracle.gr_match_range(prm, upper, ds)
    rng = np.random.default_rng(21)
    F = rng.standard_normal((6, n[1] + 2 * G, n[0] + 2 * G)).astype(np.float32)
    T = rng.standard_normal(F.shape).astype(np.float32)
    for tags, comps in ((0x7, 0x3F), (0x38, 0x3F), (0x3F, 0b101101)):
        f = _dev(F)
        gr.MatchBoundaries_kernel(f, _dev(T), shape, met, edge, ds, tags, (lo, hi), comps,
                                  BCS)()
        torch.cuda.synchronize()
        assert device_errors() == 0
        ref = F.astype(np.float64)
        assert oracle.gr_bc("match", prm, ref, T.astype(np.float64), xg_edge=edge, ds=ds,
                            i1_range=(lo, hi), tags=tags, comps=comps) == 0
        _close(f.cpu().numpy(), ref, f"match tags={tags} comps={comps}")


@pytest.mark.parametrize("nfilter", [0, 1, 3])
def test_horizon_and_axis_rows(hip, oracle, nfilter):
    from entity_amd.kernels import device_errors, gr
    n = (20, 16)
    prm, met, shape = _setup(oracle, "qkerr_schild", n)
    rng = np.random.default_rng(22)
    F = rng.standard_normal((6, n[1] + 2 * G, n[0] + 2 * G)).astype(np.float32)
    for tags in (0x7, 0x38, 0x3F):
        f = _dev(F)
        gr.HorizonBoundaries_kernel(f, shape, tags, nfilter)()
        gr.AxisBoundaries_kernel(f, shape, False, tags)()
        gr.AxisBoundaries_kernel(f, shape, True, tags)()
        torch.cuda.synchronize()
        assert device_errors() == 0
        ref = F.astype(np.float64)
        assert oracle.gr_bc("horizon", prm, ref, tags=tags, nfilter=nfilter) == 0
        assert oracle.gr_bc("axis", prm, ref, tags=tags, upper=False) == 0
        assert oracle.gr_bc("axis", prm, ref, tags=tags, upper=True) == 0
        np.testing.assert_array_equal(f.cpu().numpy().astype(np.float64), ref)


def test_absorb_currents_and_filter(hip, oracle):
    from entity_amd.grpic import match_range
    from entity_amd.kernels import device_errors, gr
    n = (40, 24)
    prm, met, shape = _setup(oracle, "qkerr_schild", n, a=0.6)
    rng = np.random.default_rng(23)
    J = rng.standard_normal((3, n[1] + 2 * G, n[0] + 2 * G)).astype(np.float32)
    lo, hi, edge = match_range(met, True, 3.0)
    j = _dev(J)
    gr.AbsorbCurrents_kernel(j, shape, met, edge, 3.0, (lo, hi))()
    torch.cuda.synchronize()
    ref = J.astype(np.float64)
    assert oracle.gr_bc("absorb_currents", prm, ref, xg_edge=edge, ds=3.0, i1_range=(lo, hi)) == 0
    _close(j.cpu().numpy(), ref, "absorb")
    for axis in ((True, True), (False, False)):
        A0 = rng.standard_normal(J.shape).astype(np.float32)
        a = _dev(A0)
        bcs = [(MATCH, MATCH), (AXIS if axis[0] else MATCH, AXIS if axis[1] else MATCH)]
        gr.DigitalFilter_kernel(a, _dev(J), shape, bcs)()
        torch.cuda.synchronize()
        assert device_errors() == 0
        ref = A0.astype(np.float64)
        assert oracle.gr_bc("filter", prm, ref, J.astype(np.float64), axis=axis) == 0
        _close(a.cpu().numpy(), ref, f"filter axis={axis}")


def _gr_particles(rng, n, npart):
    i1 = rng.integers(3, n[0] - 3, npart)
    i2 = rng.integers(2, n[1] - 2, npart)
    dx1, dx2 = rng.random(npart).astype(np.float32), rng.random(npart).astype(np.float32)
    s1, s2 = rng.integers(-1, 2, npart), rng.integers(-1, 2, npart)
    return dict(i1=i1, i2=i2, dx1=dx1, dx2=dx2,
                i1_prev=i1 - s1 * (rng.random(npart) < 0.3),
                i2_prev=i2 - s2 * (rng.random(npart) < 0.3),
                dx1_prev=np.clip(dx1 + 0.2 * rng.standard_normal(npart), 0.01, 0.99).astype(np.float32),
                dx2_prev=np.clip(dx2 + 0.2 * rng.standard_normal(npart), 0.01, 0.99).astype(np.float32),
                ux1=rng.standard_normal(npart).astype(np.float32),
                ux2=rng.standard_normal(npart).astype(np.float32),
                ux3=(3 * rng.standard_normal(npart)).astype(np.float32),
                weight=rng.uniform(0.5, 2, npart).astype(np.float32),
                tag=np.where(rng.random(npart) < 0.1, 0, 1).astype(np.int16),
                phi=np.zeros(npart, np.float32))


def _load(sp, h):
    n = len(h["i1"])
    for k, v in h.items():
        t = getattr(sp, k, None)
        if t is not None and k != "phi":
            t[:n] = torch.from_numpy(np.ascontiguousarray(v)).to(t.dtype).cuda()
    sp.npart = n


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("ordered", [False, True])
def test_gr_deposit(hip, oracle, kind, precision, ordered):
    """Random order: each workgroup's particles span the grid and deposit
    with global atomics; cell order: they fit the LDS window."""
    from entity_amd.containers import Particles
    from entity_amd.kernels import device_errors, gr
    n = (32, 24)
    prm, met, shape = _setup(oracle, kind, n, a=0.9 if kind != "kerr_schild_0" else 0.0)
    rng = np.random.default_rng(24)
    h = _gr_particles(rng, n, 4000)
    if ordered:
        order = np.lexsort((h["i1"], h["i2"]))
        h = {k: v[order] for k, v in h.items()}
    sp = Particles(2, 4096, device="cuda", with_phi=True)
    _load(sp, h)
    J = torch.zeros(shape.tensor_shape(3), device="cuda")
    gr.DepositCurrents_kernel(J, shape, sp, met, -1.0, 0.37, precision)()
    torch.cuda.synchronize()
    assert device_errors() == 0
    ref = np.zeros((3, n[1] + 2 * G, n[0] + 2 * G))
    assert oracle.gr_deposit(prm, h, ref, charge=-1.0, dt=float(np.float32(0.37))) == 0
    got = J.cpu().numpy()
    _close(got[:2], ref[:2], "in-plane", rtol=4e-6)
    _close(got[2], ref[2], "out-of-plane", rtol=4e-6 if precision == 64 else 2e-5)


def _engine_case(oracle, n=(48, 40), a=0.9, ppc=2, seed=31, nfilter=1, cfl=2.0):
    """Wald's field on a QKS grid with an e-/e+ plasma: the engine and the
    oracle's step with the same start."""
    from bench import gr_run
    from entity_amd.grpic import GRPICEngine, GRSpeciesSpec
    from entity_amd import _abi
    run = gr_run(n, a, ppc=ppc, rmax=12.0, seed=seed, cfl=cfl)
    prm, met = run_prm(oracle, run, a)
    em = run.em.cpu().numpy().astype(np.float64)
    flds = ((_abi.FLDS_BC_MATCH, _abi.FLDS_BC_MATCH), (_abi.FLDS_BC_AXIS, _abi.FLDS_BC_AXIS))
    prtl = ((_abi.PRTL_BC_HORIZON, _abi.PRTL_BC_ABSORB), (_abi.PRTL_BC_AXIS, _abi.PRTL_BC_AXIS))
    ds = ((0.2, 1.5),)
    specs = [GRSpeciesSpec(sp.label, sp.mass, sp.charge) for sp in run.species]
    eng = GRPICEngine(met, specs, run.species[0].maxnpart, dt=run.dt, flds_bc=flds, prtl_bc=prtl,
                      ds=ds, target=run.em.clone(), nfilter=nfilter, q0=0.02, B0=1.0,
                      clear_interval=4)
    eng.em.copy_(run.em)
    hosts = []
    for src, dst in zip(run.species, eng.species):
        h = src.to_host()
        # keep particles off the axis rows, where the geodesic push's 1/sin
        # terms amplify round-off
        h["i2"] = np.clip(h["i2"], 2, n[1] - 3)
        _load(dst, h)
        hosts.append(h)
    oprm = []
    for sp in eng.species:
        p = oracle.gr_params("qkerr_schild", n, ((met.r_extent[0], met.r_extent[1]), (0, math.pi)),
                             a=a, dt=float(np.float32(run.dt)), mass=sp.mass, charge=sp.charge,
                             omegaB0=1.0, niter=10, absorb=(True, True), axis=(True, True))
        oprm.append(p)
    st = oracle.GRStep(prm, em, list(zip(oprm, hosts)),
                       flds_bc=((oracle.FLDS_MATCH, oracle.FLDS_MATCH),
                                (oracle.FLDS_AXIS, oracle.FLDS_AXIS)),
                       prtl_bc=((oracle.PRTL_HORIZON, oracle.PRTL_ABSORB),
                                (oracle.PRTL_AXIS, oracle.PRTL_AXIS)),
                       ds=ds, target=em, nfilter=nfilter, q0=0.02, B0=1.0, clear_interval=4)
    return eng, st


def run_prm(oracle, run, a):
    met = run.metric
    prm = oracle.gr_params("qkerr_schild", tuple(met.n), (tuple(met.r_extent), (0.0, math.pi)),
                           a=a, dt=float(np.float32(run.dt)))
    return prm, met


@pytest.mark.parametrize("nsteps,tol", [(1, 2e-5), (6, 2e-4)])
def test_engine_matches_oracle_step(hip, oracle, nsteps, tol):
    from entity_amd.kernels import device_errors
    eng, st = _engine_case(oracle)
    for _ in range(nsteps):
        eng.step_forward()
        st.step_forward()
    torch.cuda.synchronize()
    assert device_errors() == 0
    for name in ("em", "em0"):
        got = getattr(eng, name).cpu().numpy().astype(np.float64)
        ref = getattr(st, name)
        # D starts at zero and grows from curls of H ~ B: its FP32 round-off
        # is that of the B-sized terms, a floor of 2e-7 |B| (a few ulp of B)
        floor = 2e-7 * (10 if nsteps > 1 else 1) * np.abs(ref).max()
        for c in range(6):
            sc = max(np.abs(ref[c]).max(), 1e-30)
            np.testing.assert_allclose(got[c], ref[c], rtol=0, atol=tol * sc + floor,
                                       err_msg=f"{name}[{c}] after {nsteps} steps")
    for name in ("cur", "cur0"):
        got = getattr(eng, name).cpu().numpy().astype(np.float64)
        ref = getattr(st, name)
        sc = max(np.abs(ref).max(), 1e-30)
        np.testing.assert_allclose(got, ref, rtol=0, atol=tol * sc * 10, err_msg=name)
    for sp, (_, q) in zip(eng.species, st.species):
        h = sp.to_host()
        m = q["_n"]
        assert sp.npart == m
        for k in ("ux1", "ux2", "ux3"):
            a = np.sort(h[k][:m].astype(np.float64))
            b = np.sort(q[k][:m])
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-4 * max(np.abs(b).max(), 1))
        assert np.array_equal(np.sort(h["tag"][:m]), np.sort(q["tag"][:m]))


def test_engine_gauss_law(hip, oracle):
    """div(sqrt(h) D) - (q0 / B0) rho per interior cell, on the GPU engine,
    stays within FP32 round-off of its value after the first step (CFL 0.5:
    no particle moves a cell or more per step, as the deposit requires)."""
    eng, _ = _engine_case(oracle, nfilter=0, cfl=0.5)
    prm, _ = run_prm(oracle, eng, eng.metric.a)
    n = tuple(eng.metric.n)
    i = np.arange(n[0] + 2 * G, dtype=float)[None, :] - G
    j = np.arange(n[1] + 2 * G, dtype=float)[:, None] - G

    def sd(x1, x2):
        return _metric(oracle, prm, x1 + 0 * x2, x2 + 0 * x1)["sqrt_det_h"]

    w1, w1m, w2, w2m = sd(i + 0.5, j), sd(i - 0.5, j), sd(i, j + 0.5), sd(i, j - 0.5)

    def invariant():
        D = eng.em.cpu().numpy().astype(np.float64)
        div = (w1 * D[0] - w1m * np.roll(D[0], 1, axis=1) +
               w2 * D[1] - w2m * np.roll(D[1], 1, axis=0))
        rho = np.zeros_like(div)
        for sp in eng.species:
            h = sp.to_host()
            m = sp.npart
            live = h["tag"][:m] == 1
            x1 = h["i1"][:m][live] + G + h["dx1"][:m][live].astype(float)
            x2 = h["i2"][:m][live] + G + h["dx2"][:m][live].astype(float)
            c1, c2 = np.floor(x1).astype(int), np.floor(x2).astype(int)
            f1, f2 = x1 - c1, x2 - c2
            wq = sp.charge * h["weight"][:m][live].astype(float)
            for di, wi in ((0, 1 - f1), (1, f1)):
                for dj, wj in ((0, 1 - f2), (1, f2)):
                    np.add.at(rho, (c2 + dj, c1 + di), wq * wi * wj)
        return div - (0.02 / 1.0) * rho

    def max_move():
        m = 0.0
        for sp in eng.species:
            h = sp.to_host()
            k = sp.npart
            for d in ("1", "2"):
                x = h["i" + d][:k] + h["dx" + d][:k].astype(float)
                xp = h["i" + d + "_prev"][:k] + h["dx" + d + "_prev"][:k].astype(float)
                m = max(m, float(np.abs(x - xp).max()))
        return m

    # interior cells: off the poles' rows and outside both match layers
    # (which blend D towards the target) and the absorbing current layer
    lo_x1 = eng._layers[0][1] + 1
    hi_x1 = eng._layers[1][0] - 1
    assert hi_x1 - lo_x1 > 20
    sel = (slice(G + 3, G + n[1] - 3), slice(lo_x1, hi_x1))
    eng.step_forward()
    I0 = invariant()[sel]
    for _ in range(5):
        eng.step_forward()
        assert max_move() < 1.0
    I1 = invariant()[sel]
    scale = np.abs(I0).max()
    np.testing.assert_allclose(I1, I0, rtol=0, atol=2e-5 * scale)


@pytest.mark.parametrize("kind", ["kerr_schild", "kerr_schild_0"])
def test_engine_other_metrics_and_horizon(hip, oracle, kind):
    """The engine against the oracle's step on the uniform-r grids (KS at
    a = 0.6 with a horizon boundary at r_min inside the horizon; KS0, flat
    space), Wald's field built from its vector potential, a random pair
    plasma kept off the poles: 3 steps, fields and particles as in
    test_engine_matches_oracle_step."""
    from entity_amd import _abi
    from entity_amd.grpic import GRPICEngine, GRSpeciesSpec
    from entity_amd.kernels import device_errors
    n, a = (40, 32), (0.6 if kind == "kerr_schild" else 0.0)
    prm, met, shape = _setup(oracle, kind, n, a=a, rmax=10.0)
    em = wald_fields(oracle, prm, n, a).astype(np.float32)
    dr = (prm.x1_max - prm.x1_min) / n[0]
    dt = float(np.float32(0.25 * min(dr, prm.x1_min * math.pi / n[1])))
    prm.dt = dt
    lo_bc = _abi.FLDS_BC_HORIZON if kind == "kerr_schild" else _abi.FLDS_BC_MATCH
    flds = ((lo_bc, _abi.FLDS_BC_MATCH), (_abi.FLDS_BC_AXIS, _abi.FLDS_BC_AXIS))
    prtl = ((_abi.PRTL_BC_HORIZON, _abi.PRTL_BC_ABSORB), (_abi.PRTL_BC_AXIS, _abi.PRTL_BC_AXIS))
    specs = [GRSpeciesSpec("e-", 1.0, -1.0), GRSpeciesSpec("e+", 1.0, 1.0)]
    ds = ((0.5, 1.5),)
    eng = GRPICEngine(met, specs, 4096, dt=dt, flds_bc=flds, prtl_bc=prtl, ds=ds,
                      target=_dev(em), nfilter=1, q0=0.05, B0=1.0, clear_interval=2)
    eng.em.copy_(_dev(em))
    rng = np.random.default_rng(41)
    hosts, oprm = [], []
    for sp in eng.species:
        m = 3000
        cell = np.sort(rng.integers(0, n[0] * (n[1] - 6), m))
        h = dict(i1=(cell % n[0]).astype(np.int32), i2=(3 + cell // n[0]).astype(np.int32),
                 dx1=rng.random(m).astype(np.float32), dx2=rng.random(m).astype(np.float32),
                 ux1=(0.3 * rng.standard_normal(m)).astype(np.float32),
                 ux2=(0.3 * rng.standard_normal(m)).astype(np.float32),
                 ux3=(0.3 * rng.standard_normal(m)).astype(np.float32),
                 weight=np.ones(m, np.float32), tag=np.ones(m, np.int16))
        for k in ("i1", "i2", "dx1", "dx2"):
            h[k + "_prev"] = h[k].copy()
        _load(sp, h)
        hosts.append(h)
        oprm.append(oracle.gr_params(kind, n, ((prm.x1_min, prm.x1_max), (0.0, math.pi)), a=a,
                                     dt=dt, mass=sp.mass, charge=sp.charge, omegaB0=1.0,
                                     niter=10, absorb=(True, True), axis=(True, True)))
    bc_o = {_abi.FLDS_BC_HORIZON: oracle.FLDS_HORIZON, _abi.FLDS_BC_MATCH: oracle.FLDS_MATCH}
    st = oracle.GRStep(prm, em.astype(np.float64), list(zip(oprm, hosts)),
                       flds_bc=((bc_o[lo_bc], oracle.FLDS_MATCH),
                                (oracle.FLDS_AXIS, oracle.FLDS_AXIS)),
                       prtl_bc=((oracle.PRTL_HORIZON, oracle.PRTL_ABSORB),
                                (oracle.PRTL_AXIS, oracle.PRTL_AXIS)),
                       ds=ds, target=em.astype(np.float64), nfilter=1, q0=0.05, B0=1.0,
                       clear_interval=2)
    for _ in range(3):
        eng.step_forward()
        st.step_forward()
    torch.cuda.synchronize()
    assert device_errors() == 0
    for name in ("em", "em0"):
        got = getattr(eng, name).cpu().numpy().astype(np.float64)
        ref = getattr(st, name)
        floor = 2e-6 * np.abs(ref).max()
        for c in range(6):
            sc = max(np.abs(ref[c]).max(), 1e-30)
            np.testing.assert_allclose(got[c], ref[c], rtol=0, atol=1e-4 * sc + floor,
                                       err_msg=f"{kind} {name}[{c}]")
    for sp, (_, q) in zip(eng.species, st.species):
        h = sp.to_host()
        m = q["_n"]
        assert sp.npart == m
        for k in ("ux1", "ux2", "ux3"):
            np.testing.assert_allclose(np.sort(h[k][:m].astype(np.float64)), np.sort(q[k][:m]),
                                       rtol=0, atol=1e-4 * max(np.abs(q[k][:m]).max(), 1))
