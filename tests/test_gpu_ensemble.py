"""Batched ensembles (ntt_ensemble_step): boxes of different sizes, shape
orders, pushers, filter counts, particle loads, temperatures and drifts
advance together, and every box ends bit-identical (fields, and particles as
a multiset) to the same box stepped alone with ntt_sr_step; a captured HIP
graph replays the same steps, with eager work between the replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

BORIS, VAY = 2, 4

BOXES = [  # (n, order, pusher, filters, ppc, theta, drift, seed)
    ((32, 32), 1, BORIS, 1, 4, 0.1, 0.5, 1),
    ((64, 32), 2, VAY, 2, 6, 0.3, 0.2, 2),
    ((48, 48), 1, VAY, 0, 8, 0.02, 0.8, 3),
    ((24, 40), 0, BORIS, 1, 5, 0.5, 0.0, 4),
    ((32, 64), 2, BORIS, 3, 3, 1.0, 0.3, 5),
]


BOXES_3D = [
    ((16, 12, 8), 2, VAY, 2, 4, 0.3, 0.3, 11),
    ((12, 8, 16), 1, BORIS, 1, 3, 0.1, 0.6, 12),
]


def make_box(spec, sort_interval=4):
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    n, order, alg, nf, ppc, theta, drift, seed = spec
    dim = len(n)
    prm = SRParams(dim=dim, n=n, shape_order=order, n_filters=nf, ppc0=2.0 * ppc,
                   tile=(8, 8) if dim == 2 else (4, 4, 4),
                   larmor0=0.5, skindepth0=2.0, sort_interval=sort_interval)
    eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0, alg), SpeciesSpec("e+", 1.0, 1.0, alg)],
                      int(np.prod(n)) * ppc)
    eng.inject_uniform(ppc, temperature=theta, drifts=[(drift, 0, 0), (-drift, 0, 0)], seed=seed)
    rng = np.random.default_rng(seed)
    em = (rng.standard_normal(eng.fields.em.shape) * 0.05).astype(np.float32)
    eng.fields.em.copy_(torch.from_numpy(em))
    return eng


def rows(host):
    keys = ["i1", "i2", "dx1", "dx2", "i1_prev", "i2_prev", "dx1_prev", "dx2_prev",
            "ux1", "ux2", "ux3", "weight", "tag"]
    if "i3" in host:
        keys += ["i3", "dx3", "i3_prev", "dx3_prev"]
    m = np.stack([np.asarray(host[k]).astype(np.float64) for k in keys], 1)
    return m[np.lexsort(m.T[::-1])]


def assert_same(a, b):
    torch.cuda.synchronize()
    for k in ("em", "cur"):
        np.testing.assert_array_equal(getattr(a.fields, k).cpu().numpy(),
                                      getattr(b.fields, k).cpu().numpy(), err_msg=k)
    for sa, sb in zip(a.species, b.species):
        np.testing.assert_array_equal(rows(sa.to_host()), rows(sb.to_host()))


def test_ensemble_equals_single_boxes(hip):
    from entity_amd.ensemble import SREnsemble
    from entity_amd.kernels import device_errors
    ens_boxes = [make_box(s) for s in BOXES]
    ref_boxes = [make_box(s) for s in BOXES]
    ens = SREnsemble(ens_boxes)
    for _ in range(10):
        ens.step_forward()
        for b in ref_boxes:
            b.step_forward()
    assert device_errors() == 0
    for a, b in zip(ens_boxes, ref_boxes):
        assert_same(a, b)


def test_ensemble_graph_replays(hip):
    from entity_amd.ensemble import SREnsemble
    from entity_amd.kernels import device_errors
    ens_boxes = [make_box(s, sort_interval=0) for s in BOXES[:3]]
    ref_boxes = [make_box(s, sort_interval=0) for s in BOXES[:3]]
    ens = SREnsemble(ens_boxes)
    ens.step_forward()
    for b in ref_boxes:
        b.step_forward()
    ens.capture()
    for r in range(4):
        ens.step_graph()
        for b in ref_boxes:
            b.step_forward()
        # eager work and a synchronisation between the replays
        _ = (torch.ones(1000, device="cuda") * r).sum().item()
        for a, b in zip(ens_boxes, ref_boxes):
            assert_same(a, b)
    assert device_errors() == 0


def test_ensemble_sort_orders_every_box(hip):
    """ntt_ensemble_sort over boxes of different sizes, dimensions of tiles and
    species: every species ends grouped by tile with dead particles last,
    every particle carried with all its arrays, and tile offsets equal to the
    counts of the tiles."""
    from entity_amd.ensemble import SREnsemble
    specs = [BOXES[0], BOXES[1], BOXES[3]]
    boxes = [make_box(s, sort_interval=4) for s in specs]
    rng = np.random.default_rng(9)
    befores = []
    for eng in boxes:
        for sp in eng.species:
            n = sp.npart
            sp.pld_i = torch.arange(sp.maxnpart, device="cuda")
            perm = torch.from_numpy(rng.permutation(n)).cuda()
            for v in sp.arrays().values():
                v.copy_(v[perm].clone())
            dead = torch.from_numpy(rng.random(n) < 0.05).cuda()
            sp.tag[:n][dead] = 0
            befores.append(sp.to_host())
    ens = SREnsemble(boxes)
    _abi_check = __import__("entity_amd._abi", fromlist=["check"]).check
    _abi_check(ens.lib.ntt_ensemble_sort(ens._plan, torch.cuda.current_stream().cuda_stream),
               "ntt_ensemble_sort")
    torch.cuda.synchronize()
    k = 0
    for eng in boxes:
        t1, t2 = eng.params.tile
        nt1 = (eng.params.n[0] + t1 - 1) // t1
        for s, sp in enumerate(eng.species):
            before, after = befores[k], sp.to_host()
            k += 1
            inv = np.empty(len(before["pld_i"]), dtype=np.int64)
            inv[before["pld_i"]] = np.arange(len(before["pld_i"]))
            src = inv[after["pld_i"]]
            for name, arr in before.items():
                if name.endswith("_prev"):   # left in place: the next push rewrites them
                    np.testing.assert_array_equal(after[name], arr, err_msg=name)
                else:
                    np.testing.assert_array_equal(after[name], arr[src], err_msg=name)
            alive = int((after["tag"] == 1).sum())
            assert (after["tag"][:alive] == 1).all() and (after["tag"][alive:] == 0).all()
            key = after["i1"][:alive] // t1 + nt1 * (after["i2"][:alive] // t2)
            assert (np.diff(key) >= 0).all()
            off = eng.tile_offsets[s].cpu().numpy()
            counts = np.bincount(key, minlength=len(off) - 1)
            np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(counts)]))


def test_ensemble_sort_with_arrays_held_by_some_boxes(hip):
    """One batched re-sort over boxes that hold different arrays: a 2D box
    whose species carry an ID payload (8-byte pld_i), a 2D box without one and
    a 3D box (i3, dx3).  The sort moves an array in the boxes that hold it and
    leaves the others alone: every species ends grouped by tile with dead
    particles last, carrying each particle's arrays together (exactly, through
    the IDs, where there are IDs; as a multiset of rows elsewhere), with its
    prev arrays untouched and tile offsets equal to the tile counts."""
    from entity_amd.ensemble import SREnsemble
    specs = [BOXES[1], BOXES[0], BOXES_3D[0]]
    boxes = [make_box(s, sort_interval=4) for s in specs]
    rng = np.random.default_rng(17)
    befores = []
    for b, eng in enumerate(boxes):
        for sp in eng.species:
            n = sp.npart
            if b == 0:
                sp.pld_i = torch.arange(sp.maxnpart, device="cuda") * 3 + 1
            perm = torch.from_numpy(rng.permutation(n)).cuda()
            for v in sp.arrays().values():
                v.copy_(v[perm].clone())
            dead = torch.from_numpy(rng.random(n) < 0.05).cuda()
            sp.tag[:n][dead] = 0
            befores.append(sp.to_host())
    ens = SREnsemble(boxes)
    _abi_check = __import__("entity_amd._abi", fromlist=["check"]).check
    _abi_check(ens.lib.ntt_ensemble_sort(ens._plan, torch.cuda.current_stream().cuda_stream),
               "ntt_ensemble_sort")
    torch.cuda.synchronize()

    def moved_rows(h):
        names = sorted(k for k in h if not k.endswith("_prev"))
        m = np.stack([np.asarray(h[k]).astype(np.float64) for k in names], 1)
        return m[np.lexsort(m.T[::-1])]

    k = 0
    for b, eng in enumerate(boxes):
        t = eng.params.tile
        nt = [(eng.params.n[d] + t[d] - 1) // t[d] for d in range(len(t))]
        for s, sp in enumerate(eng.species):
            before, after = befores[k], sp.to_host()
            k += 1
            assert ("pld_i" in after) == (b == 0)
            for name, arr in before.items():
                if name.endswith("_prev"):
                    np.testing.assert_array_equal(after[name], arr, err_msg=name)
            if b == 0:
                inv = np.full(int(before["pld_i"].max()) + 1, -1, dtype=np.int64)
                inv[before["pld_i"]] = np.arange(len(before["pld_i"]))
                src = inv[after["pld_i"]]
                assert (src >= 0).all()
                for name, arr in before.items():
                    if not name.endswith("_prev"):
                        np.testing.assert_array_equal(after[name], arr[src], err_msg=name)
            else:
                np.testing.assert_array_equal(moved_rows(after), moved_rows(before))
            alive = int((after["tag"] == 1).sum())
            assert (after["tag"][:alive] == 1).all() and (after["tag"][alive:] == 0).all()
            key = np.zeros(alive, dtype=np.int64)
            for d in reversed(range(len(t))):
                key = key * nt[d] + after["i%d" % (d + 1)][:alive] // t[d]
            assert (np.diff(key) >= 0).all()
            off = eng.tile_offsets[s].cpu().numpy()
            counts = np.bincount(key, minlength=len(off) - 1)
            np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(counts)]))


def test_ensemble_graph_replays_with_resorts(hip):
    """Re-sorting every 2 steps inside the replayed graphs: fields bit-identical
    to the boxes stepped alone, particles equal as sets."""
    from entity_amd.ensemble import SREnsemble
    from entity_amd.kernels import device_errors
    ens_boxes = [make_box(s, sort_interval=2) for s in BOXES[:4]]
    ref_boxes = [make_box(s, sort_interval=2) for s in BOXES[:4]]
    ens = SREnsemble(ens_boxes)
    ens.step_forward()
    for b in ref_boxes:
        b.step_forward()
    ens.capture()
    for r in range(6):
        ens.step_graph()
        for b in ref_boxes:
            b.step_forward()
        _ = (torch.ones(1000, device="cuda") * r).sum().item()
        for a, b in zip(ens_boxes, ref_boxes):
            assert_same(a, b)
    assert device_errors() == 0


def test_ensemble_with_an_empty_species(hip):
    """A box whose second species holds no particles: the batched step skips
    it, the batched re-sort writes its tile offsets as zeros, and every box
    stays bit-identical to itself stepped alone, across graph replays with
    re-sorts every 2 steps."""
    from entity_amd.ensemble import SREnsemble
    from entity_amd.kernels import device_errors

    def boxes():
        out = [make_box(s, sort_interval=2) for s in BOXES[:3]]
        out[1].species[1].npart = 0
        out[1].sort()
        return out

    ens_boxes, ref_boxes = boxes(), boxes()
    ens = SREnsemble(ens_boxes)
    ens.step_forward()
    for b in ref_boxes:
        b.step_forward()
    ens.capture()
    for r in range(5):
        ens.step_graph()
        for b in ref_boxes:
            b.step_forward()
        _ = (torch.ones(1000, device="cuda") * r).sum().item()
        for a, b in zip(ens_boxes, ref_boxes):
            assert_same(a, b)
    assert ens_boxes[1].species[1].npart == 0
    assert int(ens_boxes[1].tile_offsets[1].abs().sum().item()) == 0
    assert device_errors() == 0


def test_ensemble_mixes_2d_and_3d_boxes(hip):
    """2D and 3D boxes (different dimension groups of the batched fused
    kernel, and 3D tiles in the batched re-sort) in one ensemble, replayed
    from graphs with re-sorts every 2 steps: every box equals itself stepped
    alone."""
    from entity_amd.ensemble import SREnsemble
    from entity_amd.kernels import device_errors
    specs = BOXES[:2] + BOXES_3D
    ens_boxes = [make_box(s, sort_interval=2) for s in specs]
    ref_boxes = [make_box(s, sort_interval=2) for s in specs]
    ens = SREnsemble(ens_boxes)
    ens.step_forward()
    for b in ref_boxes:
        b.step_forward()
    ens.capture()
    for _ in range(5):
        ens.step_graph()
        for b in ref_boxes:
            b.step_forward()
    assert device_errors() == 0
    for a, b in zip(ens_boxes, ref_boxes):
        assert_same(a, b)


def test_ensemble_detects_boxes_changed_behind_its_back(hip):
    """A box re-sorted out of place on its own (its arrays swapped) no longer
    matches the plan's pointers: the ensemble refuses to step until rebuilt."""
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    from entity_amd.ensemble import SREnsemble
    prm = SRParams(dim=2, n=(32, 32), sort_interval=4, tile=(8, 8))
    eng = SRPICEngine(prm, [SpeciesSpec("e", 1.0, -1.0)], 32 * 32 * 2, double_buffer=True)
    eng.inject_uniform(2, temperature=0.1)
    ens = SREnsemble([eng])
    ens.step_forward()
    eng.sort()
    with pytest.raises(RuntimeError):
        ens.step_forward()
    ens._build()
    ens.step_forward()
