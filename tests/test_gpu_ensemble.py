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


def make_box(spec, sort_interval=4):
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    n, order, alg, nf, ppc, theta, drift, seed = spec
    prm = SRParams(dim=2, n=n, shape_order=order, n_filters=nf, ppc0=2.0 * ppc, tile=(8, 8),
                   larmor0=0.5, skindepth0=2.0, sort_interval=sort_interval)
    eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0, alg), SpeciesSpec("e+", 1.0, 1.0, alg)],
                      n[0] * n[1] * ppc)
    eng.inject_uniform(ppc, temperature=theta, drifts=[(drift, 0, 0), (-drift, 0, 0)], seed=seed)
    rng = np.random.default_rng(seed)
    em = (rng.standard_normal(eng.fields.em.shape) * 0.05).astype(np.float32)
    eng.fields.em.copy_(torch.from_numpy(em))
    return eng


def rows(host):
    keys = ["i1", "i2", "dx1", "dx2", "i1_prev", "i2_prev", "dx1_prev", "dx2_prev",
            "ux1", "ux2", "ux3", "weight", "tag"]
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
