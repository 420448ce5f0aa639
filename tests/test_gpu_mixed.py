"""Affine physical derivative on matrix cores (dgk_phys_deriv_affine):
FP64 against the oracle's calcPhysicalDerivative
(tetnudg/tet_nudg_mesh_physical_derivative_test.go:344-367), FP32 and split
BF16x3 against the FP64 result with stated error bounds."""
import numpy as np
import pytest

from dgkernel_amd import _lib, dg
from dgkernel_amd import mesh as M
from oracle import oracle as O
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

METRICS = ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")
PRECS = {"f64": _lib.DGK_PREC_F64, "f32": _lib.DGK_PREC_F32, "bf16x3": _lib.DGK_PREC_BF16X3,
         "bf16x6": _lib.DGK_PREC_BF16X6, "i8x6": _lib.DGK_PREC_I8X6}
# Bounds on the error against FP64 (relative L2, max error scaled by max |du|)
# for FP32-accumulated products of length Np <= 84 on O(1) data. FP32
# operands carry 24 bits (measured ~1e-7); the two-part bf16 split carries
# 16 bits per operand and drops lo*lo, so BF16x3 sits near 2^-16 / sqrt(Np)
# (measured ~2e-6); the three-part split (six products) is at FP32 level.
# The exact-integer split (I8X6) carries 46 bits per operand and exact int32
# sums: FP64-level, bounded here at 1e-11 (L2) and 1e-10 (max).
BOUND = {"f32": (1e-6, 1e-5), "bf16x3": (1e-5, 5e-5), "bf16x6": (1e-6, 1e-5),
         "i8x6": (1e-11, 1e-10)}


def _run(device, N, D, U, geo, prec, rst):
    K, Np = U.shape
    plan = dg.DerivPlan(device, N, D, prec, rst)
    du = device.upload(np.ascontiguousarray(U))
    dg_ = device.upload(np.ascontiguousarray(geo))
    outs = [device.zeros(K * Np, plan.out_dtype) for _ in range(3)]
    plan.apply(K, du, dg_.ptr, geo.shape[1], outs)
    device.sync()
    res = [o.to_host().reshape(K, Np).astype(np.float64) for o in outs]
    plan.free()
    return res


def _rst(el):
    return (el["r"], el["s"], el["t"])


def _mesh_case(N, seed=21):
    m = M.cube_mesh(6, 6, 5, periodic=True, jitter=0.1, seed=seed)  # 1080 tets, K % 16 = 8
    o = O.Setup.from_mesh(N, m)
    assert o.status == 0
    el = O.element(N)
    U = M.uniform(seed, 5, m.K * el["Np"], -1.0, 1.0).reshape(m.K, el["Np"])
    mets = [o.field(k) for k in METRICS]
    geo = np.stack([mm[:, 0] for mm in mets])  # affine: one value per element
    for q in range(9):  # per-node values are Dr X etc.: constant up to round-off ~ cond(V)
        assert np.max(np.abs(mets[q] - geo[q][:, None])) <= 1e-9 * np.max(np.abs(geo[q]))
    # The oracle's per-node metrics of an affine element agree to round-off;
    # the reference value is the chain rule with the element's one metric value.
    mets_e = [np.repeat(geo[q][:, None], el["Np"], axis=1) for q in range(9)]
    host = O.phys_deriv(el["Dr"], el["Ds"], el["Dt"], U, mets_e)
    return el, U, geo, host


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_affine_fp64_vs_oracle(device, N):
    el, U, geo, host = _mesh_case(N)
    got = _run(device, N, (el["Dr"], el["Ds"], el["Dt"]), U, geo, PRECS["f64"], _rst(el))
    for d in range(3):
        scale = np.max(np.abs(host[d]))
        assert np.max(np.abs(got[d] - host[d])) <= 1e-13 * scale


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16x6", "i8x6"])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_affine_mixed_error_bounds(device, N, prec):
    el, U, geo, host = _mesh_case(N, seed=22)
    got = _run(device, N, (el["Dr"], el["Ds"], el["Dt"]), U, geo, PRECS[prec], _rst(el))
    l2b, maxb = BOUND[prec]
    for d in range(3):
        e2 = rel_l2(got[d], host[d])
        emax = np.max(np.abs(got[d] - host[d])) / np.max(np.abs(host[d]))
        print(f"N={N} {prec} d{d}: rel L2 {e2:.3e}, max {emax:.3e}")
        assert e2 <= l2b and emax <= maxb
        # the error is rounding, not a wrong entry: far above FP32-operand level it would not be
        assert e2 > 0.0


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3", "bf16x6", "i8x6"])
@pytest.mark.parametrize("N", [4, 6])
def test_affine_grid_stride_and_ragged(device, prec, N):
    """K above one sweep of the persistent grid, K % 16 != 0 (K Np odd at N=4),
    random metrics."""
    el = O.element(N)
    Np = el["Np"]
    K = 70001
    U = M.uniform(3, 1, K * Np, -1.0, 1.0).reshape(K, Np)
    geo = M.uniform(3, 2, 9 * K, -2.0, 2.0).reshape(9, K)
    stride = K + 37
    geo_s = np.zeros((9, stride))
    geo_s[:, :K] = geo
    got = _run(device, N, (el["Dr"], el["Ds"], el["Dt"]), U, geo_s, PRECS[prec], _rst(el))
    mets = [np.repeat(geo[q][:, None], Np, axis=1) for q in range(9)]
    host = O.phys_deriv(el["Dr"], el["Ds"], el["Dt"], U, mets)
    for d in range(3):
        scale = np.max(np.abs(host[d]))
        err = np.max(np.abs(got[d] - host[d])) / scale
        assert err <= (1e-13 if prec == "f64" else BOUND[prec][1])


def test_affine_empty_and_errors(device):
    el = O.element(2)
    plan = dg.DerivPlan(device, 2, (el["Dr"], el["Ds"], el["Dt"]))
    with pytest.raises(_lib.DgkError):  # node coordinates required below FP64
        dg.DerivPlan(device, 2, (el["Dr"], el["Ds"], el["Dt"]), _lib.DGK_PREC_F32)
    buf = device.zeros(16)
    plan.apply(0, buf, buf.ptr, 0, [buf, buf, buf])  # K = 0 is a no-op
    with pytest.raises(_lib.DgkError):
        plan.apply(4, buf, buf.ptr, 2, [buf, buf, buf])  # stride < K
    with pytest.raises(_lib.DgkError):
        dg.DerivPlan(device, 7, (np.eye(120),) * 3)
    plan.free()
