"""GPU parity of the GR field solver kernels against the oracle.

Random D, B, E, H, J (FP32) on 40 x 32 (r, theta) grids of every metric of the
Kerr-Schild family, with axis boundaries at both theta ends and without, and
Wald's stationary field on a QKS grid.  The device evaluates the metric in
FP64 and rounds the FP64 result to the FP32 field; the oracle runs in FP64 on
the same FP32 inputs.  Outputs agree to FP32 round-off (2e-6 relative, plus
2e-6 of the largest value for results of cancelling sums), every cell of the
kernel's range, ghost rows included.  Ampere and Faraday are also run in
place (Din = Dout, Bin = Bout), where the axis rows read the row next to them
as it was before the update.
"""
import math

import numpy as np
import pytest
import torch

from test_oracle_gr_fields import G, _prm, wald_fields

pytestmark = pytest.mark.gpu

KINDS = {"kerr_schild": 0, "qkerr_schild": 1, "kerr_schild_0": 2}
METRICS = [("kerr_schild", 0.9, 0.0), ("qkerr_schild", 0.9, 0.0), ("qkerr_schild", 0.6, 0.3),
           ("kerr_schild_0", 0.0, 0.0)]
N = (40, 32)
AXIS, MATCH = 7, 2


def _setup(oracle, kind, a, h, axis=True):
    """Grid of the metric: theta over [0, pi] with axis rows, else over
    [0.2, pi - 0.2] (sqrt(det h) vanishes on the axis)."""
    from entity_amd.containers import GridShape
    from entity_amd.kernels import gr
    prm = _prm(oracle, kind, N, a, h=h)
    if not axis:
        prm = oracle.gr_params(kind, N, ((prm.x1_min, prm.x1_max), (0.2, math.pi - 0.2)),
                               a=a, h=h)
    met = gr.KerrSchild(KINDS[kind], N, (prm.x1_min, prm.x1_max), (prm.x2_min, prm.x2_max),
                        a=a, h=h)
    return prm, met, GridShape(2, N)


def _rand(rng, ncomp, scale=1.0):
    return (rng.standard_normal((ncomp, N[1] + 2 * G, N[0] + 2 * G)) * scale).astype(np.float32)


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _close(got, ref, what):
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got.astype(np.float64), ref, rtol=2e-6, atol=2e-6 * scale,
                               err_msg=what)


def _bcs(axis):
    return [(MATCH, MATCH), (AXIS, AXIS) if axis else (MATCH, MATCH)]


@pytest.mark.parametrize("kind,a,h", METRICS)
@pytest.mark.parametrize("axis", [True, False])
def test_aux_fields(hip, oracle, kind, a, h, axis):
    from entity_amd.kernels import device_errors, gr
    prm, met, shape = _setup(oracle, kind, a, h, axis)
    rng = np.random.default_rng(11)
    D, B = _rand(rng, 6), _rand(rng, 6)
    for op, K in (("aux_e", gr.ComputeAuxE_kernel), ("aux_h", gr.ComputeAuxH_kernel)):
        out0 = _rand(rng, 6)
        out = _dev(out0)
        K(_dev(D), _dev(B), out, shape, met, _bcs(axis))()
        torch.cuda.synchronize()
        assert device_errors() == 0
        ref = out0.astype(np.float64)
        assert oracle.gr_field(op, prm, D.astype(np.float64), B.astype(np.float64), ref,
                               axis=(axis, axis)) == 0
        comps = slice(0, 3) if op == "aux_e" else slice(3, 6)
        got = out.cpu().numpy()
        _close(got[comps], ref[comps], op)
        # the other three components are left as they were
        np.testing.assert_array_equal(got[3:6] if op == "aux_e" else got[0:3],
                                      out0[3:6] if op == "aux_e" else out0[0:3])


@pytest.mark.parametrize("kind,a,h", METRICS)
@pytest.mark.parametrize("axis", [True, False])
@pytest.mark.parametrize("inplace", [False, True])
def test_faraday_ampere(hip, oracle, kind, a, h, axis, inplace):
    from entity_amd.kernels import device_errors, gr
    prm, met, shape = _setup(oracle, kind, a, h, axis)
    rng = np.random.default_rng(12)
    X, Y, F = _rand(rng, 6), _rand(rng, 6), _rand(rng, 6)
    for op, K, coeff in (("faraday", gr.Faraday_kernel, 0.37), ("ampere", gr.Ampere_kernel, -0.21)):
        xin = _dev(X)
        xout = xin if inplace else _dev(Y)
        K(xin, xout, _dev(F), shape, met, coeff, _bcs(axis))()
        torch.cuda.synchronize()
        assert device_errors() == 0
        a_in = X.astype(np.float64)
        a_out = a_in if inplace else Y.astype(np.float64)
        assert oracle.gr_field(op, prm, a_in, a_out, F.astype(np.float64),
                               coeff=float(np.float32(coeff)), axis=(axis, axis)) == 0
        comps = slice(3, 6) if op == "faraday" else slice(0, 3)
        _close(xout.cpu().numpy()[comps], a_out[comps], op)


@pytest.mark.parametrize("kind,a,h", METRICS)
@pytest.mark.parametrize("axis", [True, False])
def test_currents_ampere_and_time_averages(hip, oracle, kind, a, h, axis):
    from entity_amd.kernels import device_errors, gr
    prm, met, shape = _setup(oracle, kind, a, h, axis)
    rng = np.random.default_rng(13)
    D, J = _rand(rng, 6), _rand(rng, 3)
    d = _dev(D)
    gr.CurrentsAmpere_kernel(d, _dev(J), shape, met, -0.3, _bcs(axis))()
    torch.cuda.synchronize()
    assert device_errors() == 0
    ref = D.astype(np.float64)
    assert oracle.gr_field("currents_ampere", prm, ref, J.astype(np.float64),
                           coeff=float(np.float32(-0.3)), axis=(axis, axis)) == 0
    _close(d.cpu().numpy()[0:3], ref[0:3], "currents_ampere")
    # time averages: bit for bit ((x + y) * 0.5 in FP32 on both sides)
    BD, BD0 = _rand(rng, 6), _rand(rng, 6)
    b0 = _dev(BD0)
    gr.TimeAverageDB_kernel(_dev(BD), b0, shape)()
    J1, J0 = _rand(rng, 3), _rand(rng, 3)
    j1 = _dev(J1)
    gr.TimeAverageJ_kernel(j1, _dev(J0), shape)()
    torch.cuda.synchronize()
    act = (slice(None), slice(G, G + N[1]), slice(G, G + N[0]))
    want = BD0.copy()
    want[act] = np.float32(0.5) * (BD0[act] + BD[act])
    np.testing.assert_array_equal(b0.cpu().numpy(), want)
    want = J1.copy()
    want[act] = np.float32(0.5) * (J1[act] + J0[act])
    np.testing.assert_array_equal(j1.cpu().numpy(), want)


def test_wald_field_rates(hip, oracle):
    """Wald's field on a QKS grid through aux E/H, Faraday and Ampere on the
    GPU: the same rates as the oracle's (which vanish at second order)."""
    from entity_amd.kernels import device_errors, gr
    kind, a = "qkerr_schild", 0.9
    prm, met, shape = _setup(oracle, kind, a, 0.0)
    em = wald_fields(oracle, prm, N, a).astype(np.float32)
    bcs = _bcs(True)
    e = _dev(em)
    aux = torch.zeros_like(e)
    gr.ComputeAuxE_kernel(e, e, aux, shape, met, bcs)()
    gr.ComputeAuxH_kernel(e, e, aux, shape, met, bcs)()
    b1, d1 = e.clone(), e.clone()
    gr.Faraday_kernel(e, b1, aux, shape, met, 1.0, bcs)()
    gr.Ampere_kernel(e, d1, aux, shape, met, 1.0, bcs)()
    torch.cuda.synchronize()
    assert device_errors() == 0
    ref_aux = np.zeros_like(em, dtype=np.float64)
    e64 = em.astype(np.float64)
    assert oracle.gr_field("aux_e", prm, e64, e64, ref_aux) == 0
    assert oracle.gr_field("aux_h", prm, e64, e64, ref_aux) == 0
    _close(aux.cpu().numpy(), ref_aux, "aux")
    rb, rd = e64.copy(), e64.copy()
    assert oracle.gr_field("faraday", prm, e64, rb, ref_aux, coeff=1.0) == 0
    assert oracle.gr_field("ampere", prm, e64, rd, ref_aux, coeff=1.0) == 0
    # the rates themselves are small differences of O(1) fields: compare them
    # against the scale of the fields
    scale = np.abs(e64).max()
    np.testing.assert_allclose(b1.cpu().numpy()[3:6] - em[3:6], rb[3:6] - e64[3:6], rtol=0,
                               atol=4e-6 * scale)
    np.testing.assert_allclose(d1.cpu().numpy()[0:3] - em[0:3], rd[0:3] - e64[0:3], rtol=0,
                               atol=4e-6 * scale)
