"""GPU parity of the kernel-level push and deposit off the unit-cell,
all-periodic case, and of the step's chunking past the last sorted tile.

* Cell size dx != 1 (metric::Minkowski, minkowski.h:263-299): the gathered
  in-plane E, B are multiplied by dx and the velocity divided by dx in the
  position update.  Strict build: bit-identical to the oracle; product build:
  the near-face rule of the other kernel tests.
* Absorbing and reflecting particle boundaries (sr.hpp:659-790) through
  ntt_sr_push: absorbed particles are tagged dead, reflected ones mirror
  their offset and momentum; bit-identical to the oracle in the strict build.
* ntt_sr_step after particles are appended past the last tile of the sort
  whose offsets the engine keeps: the appended particles are pushed and
  deposited (the fused kernel's tail chunk), as in the oracle's step.
"""
import numpy as np
import pytest
import torch

from test_gpu_kernels import ALGS, make_fields, make_particles, near_face, to_device

pytestmark = pytest.mark.gpu

ABSORB, REFLECT, PERIODIC = 2, 5, 1


def _compare(out, q, dim, strict):
    ints = [f"i{d+1}" for d in range(dim)] + [f"i{d+1}_prev" for d in range(dim)] + ["tag"]
    floats = [f"dx{d+1}" for d in range(dim)] + [f"dx{d+1}_prev" for d in range(dim)] + \
        ["ux1", "ux2", "ux3"]
    if strict:
        for k in ints + floats:
            np.testing.assert_array_equal(out[k], q[k], err_msg=k)
        return
    flagged = np.zeros(len(q["tag"]), bool)
    for d in range(dim):
        flagged |= near_face(q[f"dx{d+1}"]) | near_face(out[f"dx{d+1}"])
    mism = np.zeros_like(flagged)
    for k in ints:
        mism |= out[k] != q[k]
    assert not (mism & ~flagged).any(), "integer mismatch away from cell faces"
    assert mism.sum() <= max(8, len(mism) // 200)
    ok = ~mism
    for k in floats:
        np.testing.assert_allclose(out[k][ok], q[k][ok], rtol=2e-6, atol=2e-6, err_msg=k)


def _push(oracle, p, em_h, dim, n, order, alg, strict, dx, bc):
    from entity_amd.containers import GridShape
    from entity_amd.kernels import Minkowski, PusherBoundaries, PusherContext, device_errors, sr
    dt = float(np.float32(0.45 * dx / np.sqrt(dim)))
    mass, charge, omegaB0 = 1.0, -1.0, 0.7
    prt = to_device(p, dim)
    em = torch.from_numpy(em_h).cuda()
    ctx = PusherContext(ALGS[alg], mass, charge, dt, omegaB0, *n, *([1] * (3 - dim)))
    sr.Pusher_kernel(ctx, PusherBoundaries(list(bc)), prt, em, GridShape(dim, n),
                     Minkowski(dim, dx), order, strict=strict)()
    torch.cuda.synchronize()
    assert device_errors(strict) == 0
    q, rc = oracle.sr_push(p, em_h.copy(), dim=dim, mass=mass, charge=charge, dt=dt,
                           omegaB0=omegaB0, pusher=ALGS[alg], order=order, dx=dx,
                           bc=tuple(bc) + ((PERIODIC, PERIODIC),) * (3 - len(bc)))
    assert rc == 0
    return prt.to_host(), q


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("alg", ["boris", "vay"])
@pytest.mark.parametrize("strict", [True, False])
def test_push_parity_cell_size(hip, oracle, dim, order, alg, strict):
    n = (13, 10, 9)[:dim]
    dx = 0.37
    em_h = make_fields(oracle, dim, n, seed=40 + dim * 10 + order)
    p = make_particles(dim, n, 4096 if dim == 2 else 2048, seed=order + 9 * dim)
    out, q = _push(oracle, p, em_h, dim, n, order, alg, strict, dx, [(PERIODIC,) * 2] * dim)
    _compare(out, q, dim, strict)
    # dx enters: the same push with dx = 1 moves the particles differently
    out1, _ = _push(oracle, p, em_h, dim, n, order, alg, strict, 1.0, [(PERIODIC,) * 2] * dim)
    assert np.abs(out1["dx1"].astype(np.float64) - out["dx1"]).max() > 1e-3


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_deposit_parity_cell_size(hip, oracle, dim, order):
    from entity_amd.containers import GridShape
    from entity_amd.kernels import DepositCurrents_kernel, Minkowski, device_errors
    from _helpers import field_array
    n = (13, 10, 9)[:dim]
    dx = 0.37
    rng = np.random.default_rng(200 + order + dim)
    p = make_particles(dim, n, 4096, seed=5 * order + dim)
    for d in range(dim):
        disp = rng.uniform(-0.45, 0.45, len(p["tag"])).astype(np.float32)
        x = p[f"dx{d+1}"] + disp
        p[f"i{d+1}_prev"] = (p[f"i{d+1}"] + np.floor(x).astype(np.int32)).astype(np.int32)
        p[f"dx{d+1}_prev"] = (x - np.floor(x)).astype(np.float32)
    dt = float(np.float32(0.4 * dx))
    shape = GridShape(dim, n)
    J = torch.zeros(shape.tensor_shape(3), device="cuda")
    DepositCurrents_kernel(J, shape, to_device(p, dim), Minkowski(dim, dx), -1.0, dt, order,
                           strict=True)()
    torch.cuda.synchronize()
    assert device_errors(True) == 0
    J_ref = field_array(dim, n, 3)
    assert oracle.deposit(p, J_ref, dim=dim, charge=-1.0, dt=dt, order=order, dx=dx) == 0
    scale = np.abs(J_ref).max()
    np.testing.assert_allclose(J.cpu().numpy(), J_ref, rtol=0,
                               atol=2e-6 * scale * np.sqrt(len(p["tag"])) / 10)


BCS = {
    "absorb-x_reflect-y": [(ABSORB, ABSORB), (REFLECT, REFLECT), (PERIODIC, PERIODIC)],
    "reflect-x_absorb-y": [(REFLECT, ABSORB), (ABSORB, REFLECT), (REFLECT, ABSORB)],
}


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("bcname", list(BCS))
@pytest.mark.parametrize("strict", [True, False])
def test_push_absorb_reflect(hip, oracle, dim, order, bcname, strict):
    n = (13, 10, 9)[:dim]
    bc = BCS[bcname][:dim]
    em_h = make_fields(oracle, dim, n, seed=70 + order + dim)
    npart = 4096 if dim == 2 else 2048
    p = make_particles(dim, n, npart, seed=11 * order + dim, dead_frac=0.02)
    rng = np.random.default_rng(3 + order)
    # a third of the particles in the edge cells, moving outward fast
    k = npart // 3
    for d in range(dim):
        lo = rng.random(k) < 0.5
        idx = rng.choice(npart, k, replace=False)
        p[f"i{d+1}"][idx] = np.where(lo, 0, n[d] - 1).astype(np.int32)
        p[f"dx{d+1}"][idx] = np.where(lo, rng.uniform(0, 0.3, k),
                                      rng.uniform(0.7, 1.0, k)).astype(np.float32)
        p[f"ux{d+1}"][idx] = np.where(lo, -1, 1) * rng.uniform(1.0, 4.0, k).astype(np.float32)
        p[f"i{d+1}_prev"] = p[f"i{d+1}"].copy()
        p[f"dx{d+1}_prev"] = p[f"dx{d+1}"].copy()
    out, q = _push(oracle, p, em_h, dim, n, order, "boris", strict, 1.0, bc)
    _compare(out, q, dim, strict)
    # both boundary kinds acted: particles died and momenta flipped at a wall
    died = (p["tag"] == 1) & (q["tag"] == 0)
    assert died.sum() > 20
    flipped = np.zeros(npart, bool)
    for d in range(dim):
        if REFLECT in bc[d]:
            flipped |= (np.sign(q[f"ux{d+1}"]) != np.sign(p[f"ux{d+1}"])) & \
                (q[f"i{d+1}"] == np.where(p[f"ux{d+1}"] < 0, 0, n[d] - 1))
    assert flipped.sum() > 20
    for d in range(dim):
        alive = q["tag"] == 1
        assert ((q[f"i{d+1}"][alive] >= 0) & (q[f"i{d+1}"][alive] < n[d])).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_step_pushes_particles_appended_after_sort(hip, oracle, dim):
    """The engine keeps the tile offsets of its last sort; particles loaded
    past them (a larger npart, no re-sort) are still pushed and deposited."""
    from test_gpu_engine import make_engine, oracle_step, snapshot
    n = (32, 24) if dim == 2 else (16, 12, 8)
    order = 1
    eng = make_engine(dim, n, order, 2, 1, strict=True, ppc=2,
                      maxnpart=int(np.prod(n)) * 3)
    # grow every species by 50% with a copy of its first particles moved by
    # half a cell (appended after the tiles the offsets describe)
    for sp in eng.species:
        h = sp.to_host()
        m = sp.npart // 2
        grown = {}
        for key, v in h.items():
            if v is None or not hasattr(v, "shape") or v.shape[:1] != (sp.npart,):
                continue
            extra = v[:m].copy()
            if key.startswith("dx"):
                extra = ((extra + np.float32(0.5)) % np.float32(1.0)).astype(v.dtype)
            grown[key] = np.concatenate([v, extra])
        assert sp.maxnpart >= len(grown["tag"])
        sp.load(**{k: torch.from_numpy(np.ascontiguousarray(v)) for k, v in grown.items()})
    torch.cuda.synchronize()
    parts0, flds0 = snapshot(eng)
    eng.advance()       # no re-sort: the offsets of the generator's layout stay
    from entity_amd.kernels import device_errors
    assert device_errors(True) == 0
    parts1, flds1 = snapshot(eng)
    sp_ref, em_ref, cur_ref = oracle_step(oracle, eng, parts0, flds0)
    for got, ref in zip(parts1, sp_ref):
        assert len(got["tag"]) == len(ref["p"]["tag"]) > 0
        keys = [f"i{d+1}" for d in range(dim)] + [f"dx{d+1}" for d in range(dim)]
        for k in keys + ["ux1", "ux2", "ux3"]:
            np.testing.assert_array_equal(got[k], ref["p"][k], err_msg=k)
    np.testing.assert_array_equal(flds1["em"][3:], em_ref[3:])
    jscale = np.abs(cur_ref).max()
    np.testing.assert_allclose(flds1["cur"], cur_ref, rtol=0, atol=4e-6 * jscale)


@pytest.mark.parametrize("case", [(2, (32, 24), 1, 2, 1), (3, (16, 12, 8), 2, 4, 2)],
                         ids=["2d-o1-boris", "3d-o2-vay"])
def test_step_parity_cell_size(hip, oracle, case):
    """ntt_sr_step with dx = 0.5 (extent n / 2): particles bit-identical to
    the oracle's step in the strict build, J and E within the atomic-order
    tolerance of the unit-cell test, B exactly."""
    from test_gpu_engine import make_engine, oracle_step, snapshot
    from entity_amd.kernels import device_errors
    dim, n, order, alg, nf = case
    extent = tuple((0.0, 0.5 * x) for x in n)
    eng = make_engine(dim, n, order, alg, nf, strict=True, extent=extent)
    assert eng.params.dx == 0.5
    parts0, flds0 = snapshot(eng)
    eng.step_forward()
    assert device_errors(True) == 0
    parts1, flds1 = snapshot(eng)
    sp_ref, em_ref, cur_ref = oracle_step(oracle, eng, parts0, flds0)
    keys = [f"i{d+1}" for d in range(dim)] + [f"dx{d+1}" for d in range(dim)]
    for got, ref in zip(parts1, sp_ref):
        for k in keys + ["ux1", "ux2", "ux3", "tag"]:
            np.testing.assert_array_equal(got[k], ref["p"][k], err_msg=k)
    np.testing.assert_array_equal(flds1["em"][3:], em_ref[3:])
    jscale = np.abs(cur_ref).max()
    escale = np.abs(em_ref).max()
    np.testing.assert_allclose(flds1["cur"], cur_ref, rtol=0, atol=4e-6 * jscale)
    np.testing.assert_allclose(flds1["em"][:3], em_ref[:3], rtol=0, atol=4e-6 * escale)
