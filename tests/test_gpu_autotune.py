"""The autotuner on a real box: every candidate of the 2D search space runs,
the fastest is returned, and -- since the current is summed in integers --
the box's fields after several steps are bit-identical, and its particles
equal as a set, whichever tile shape, re-sort cadence or chunk size runs it."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N = (96, 64)
PPC = 6


def _snapshot():
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    prm = SRParams(dim=2, n=N, shape_order=2, n_filters=1, ppc0=2.0 * PPC, larmor0=2.0,
                   skindepth0=3.0, tile=(8, 8))
    eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0), SpeciesSpec("e+", 1.0, 1.0)],
                      N[0] * N[1] * PPC)
    eng.inject_uniform(PPC, temperature=0.2, drifts=[(0.3, 0, 0), (-0.3, 0, 0)], seed=3)
    rng = np.random.default_rng(3)
    em = (rng.standard_normal(eng.fields.em.shape) * 0.05).astype(np.float32)
    return [sp.to_host() for sp in eng.species], em


def _factory(snap):
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    prts, em = snap

    def make(c):
        prm = SRParams(dim=2, n=N, shape_order=2, n_filters=1, ppc0=2.0 * PPC, larmor0=2.0,
                       skindepth0=3.0, tile=c.tile, sort_interval=c.sort_interval,
                       particles_per_thread=c.particles_per_thread)
        eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0), SpeciesSpec("e+", 1.0, 1.0)],
                          N[0] * N[1] * PPC)
        for sp, h in zip(eng.species, prts):
            sp.load(**h)
        eng.fields.em.copy_(torch.from_numpy(em))
        if c.sort_interval:
            eng.sort()
        return eng
    return make


def test_autotune_picks_a_candidate_and_results_do_not_depend_on_it(hip):
    from entity_amd.autotune import Candidate, candidates_for, engine_runner, tune
    from entity_amd.kernels import device_errors
    snap = _snapshot()
    make = _factory(snap)
    cands = candidates_for(2)
    res = tune(engine_runner(make, steps=4, warmup=1), cands)
    assert res.best in cands and all(t > 0 for t in res.ms_per_step.values())
    assert device_errors() == 0

    def run(c, steps=8):
        eng = make(c)
        for _ in range(steps):
            eng.step_forward()
        torch.cuda.synchronize()
        rows = []
        for sp in eng.species:
            h = sp.to_host()
            m = np.stack([np.asarray(h[k]).astype(np.float64) for k in
                          ("i1", "i2", "dx1", "dx2", "ux1", "ux2", "ux3")], 1)
            rows.append(m[np.lexsort(m.T[::-1])])
        return eng.fields.em.cpu().numpy(), rows

    ref_em, ref_rows = run(Candidate((8, 8), 4))
    for c in {res.best, Candidate((16, 16), 2), Candidate((8, 8), 0, 1)}:
        em, rows = run(c)
        np.testing.assert_array_equal(em, ref_em, err_msg=str(c))
        for a, b in zip(rows, ref_rows):
            np.testing.assert_array_equal(a, b, err_msg=str(c))
    assert device_errors() == 0
