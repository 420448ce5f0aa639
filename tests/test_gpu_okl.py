"""Source-built kernels on the GPU, through the C ABI the Go runner binds
(dgk_kernel_build_from_source + dgk_kernel_launch): the reference's own
runner and derivative kernels, in INT64 and INT32, against their known
answers and the CPU oracle."""
import numpy as np
import pytest

from dgkernel_amd import _lib
from dgkernel_amd import mesh as M
from oracle import oracle as O
from tests import okl_kernels as KS
from tests.helpers import neu_mesh
from tests.runner_text import DeviceRun

pytestmark = pytest.mark.gpu

INTS = ["INT64", "INT32"]


def _run(device, case, scalars=None, host_K=False):
    rt, src, name, order = case[:4]
    run = DeviceRun(rt, device)
    k = _lib.Kernel(device, src, name)
    hk = np.asarray(rt.K, dtype=rt.int_np) if host_K else None
    k.launch(run.args(order, scalars), host_K=hk)
    device.sync()
    return run


@pytest.mark.parametrize("int_type", INTS)
@pytest.mark.parametrize("maker", [KS.matmul_known_answer, KS.differentiate_device_loop,
                                   KS.static_matmul, KS.device_matmul, KS.chain])
def test_runner_matrix_kernels(device, maker, int_type):
    case = maker(int_type)
    run = _run(device, case)
    for name, want in case[4].items():
        got = run.result(name)
        assert np.array_equal(got, want) or np.max(np.abs(got - want)) <= 1e-13 * max(1, np.abs(want).max()), name


def test_matmul_known_answer_exact(device):
    """runner_matrix_test.go:100-124: [2x+y, y+z, x+2z] of [e+1, e+2, e+3]."""
    rt, src, name, order, _ = KS.matmul_known_answer()
    run = _run(device, (rt, src, name, order), host_K=True)
    V = run.result("V").reshape(-1, 3)
    for e in range(5):
        x, y, z = e + 1.0, e + 2.0, e + 3.0
        assert list(V[e]) == [2 * x + y, y + z, x + 2 * z]


def _neu_derivative_inputs(N, parts=None):
    m = neu_mesh()
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    e = O.element(N)
    X, Y, Z = o.field("X"), o.field("Y"), o.field("Z")
    U = X ** N + Y ** N + Z ** N
    mets = {nm: o.field(nm.lower()).reshape(-1) for nm in KS.METRICS}
    K = parts or [m.K]
    return e, U, mets, K, (X, Y, Z)


@pytest.mark.parametrize("int_type", INTS)
@pytest.mark.parametrize("parts", [None, [200, 165, 200]])
def test_tetnudg_physical_derivative(device, int_type, parts):
    """tet_nudg_mesh_physical_derivative_test.go:747-919 on cube-partitioned.neu,
    P=2: the device result equals the closed form d(x^N+y^N+z^N) to 1e-8 and
    the oracle's calcPhysicalDerivative to 1e-13; the temporaries hold Dr u etc."""
    N = 2
    e, U, mets, K, (X, Y, Z) = _neu_derivative_inputs(N, parts)
    D = [e["Dr"], e["Ds"], e["Dt"]]
    case = KS.differentiate_full(D, U.reshape(-1), mets, K, int_type)
    assert "physical derivative" in _lib.describe_kernel(case[1], case[2])
    run = _run(device, case)
    mo = [mets[k].reshape(-1, e["Np"]) for k in ("Rx", "Ry", "Rz", "Sx", "Sy", "Sz", "Tx", "Ty", "Tz")]
    want = O.phys_deriv(e["Dr"], e["Ds"], e["Dt"], U, mo)
    closed = [N * X ** (N - 1), N * Y ** (N - 1), N * Z ** (N - 1)]
    for q, nm in enumerate(("DuDx", "DuDy", "DuDz")):
        got = run.result(nm).reshape(U.shape)
        assert np.max(np.abs(got - closed[q])) < 1e-8
        assert np.max(np.abs(got - want[q])) <= 1e-13 * np.abs(want[q]).max()
    for nm, Dq in (("Ur", e["Dr"]), ("Us", e["Ds"]), ("Ut", e["Dt"])):
        t = run.result(nm).reshape(U.shape)
        assert np.max(np.abs(t - U @ Dq.T)) <= 1e-13 * np.abs(U @ Dq.T).max()


@pytest.mark.parametrize("int_type", INTS)
def test_physical_derivative_fp32_metrics(device, int_type):
    """:240-278 (FP32 metrics and outputs, flat loop), with the outputs declared
    float: FP32 rounding of the metrics and of the result only."""
    N = 3
    m = M.cube_mesh(3, 3, 3, jitter=0.1, seed=5)
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    e = O.element(N)
    U = np.sin(o.field("X")) * np.cos(o.field("Y")) + o.field("Z") ** 2
    mets = {nm: o.field(nm.lower()).reshape(-1).astype(np.float32) for nm in KS.METRICS}
    case = KS.differentiate_full([e["Dr"], e["Ds"], e["Dt"]], U.reshape(-1), mets, [60, 102], int_type,
                                 short="NudgTet3", fp32=True)
    run = _run(device, case)
    mo = [mets[k].astype(np.float64).reshape(-1, e["Np"])
          for k in ("Rx", "Ry", "Rz", "Sx", "Sy", "Sz", "Tx", "Ty", "Tz")]
    want = O.phys_deriv(e["Dr"], e["Ds"], e["Dt"], U, mo)
    for q, nm in enumerate(("DuDx", "DuDy", "DuDz")):
        got = run.result(nm).astype(np.float64).reshape(U.shape)
        assert np.max(np.abs(got - want[q])) <= 2 ** -23 * np.abs(want[q]).max() * 2


@pytest.mark.parametrize("int_type", INTS)
def test_dx_kernel(device, int_type):
    """physical_derivative_test.go:92-113: Dx = Rx .* (Dr U), four partitions."""
    N = 2
    e, U, mets, _, _ = _neu_derivative_inputs(N)
    K = [141, 141, 141, 142]
    case = KS.differentiate_dx(e["Dr"], U.reshape(-1), mets["Rx"], K, int_type)
    run = _run(device, case)
    for nm, want in case[4].items():
        got = run.result(nm)
        assert np.max(np.abs(got - want)) <= 1e-13 * np.abs(want).max(), nm


def test_many_partitions(device):
    """300 partitions (above the 256 of the by-value tile table): the fused
    derivative, a MATMUL and a pointwise loop, K read from the device."""
    N = 2
    e, U, mets, _, _ = _neu_derivative_inputs(N)
    rng = np.random.default_rng(3)
    K = np.full(300, 1)
    K[: 565 - 300] += 1
    rng.shuffle(K)
    K = list(K)
    case = KS.differentiate_full([e["Dr"], e["Ds"], e["Dt"]], U.reshape(-1), mets, K)
    run = _run(device, case)
    mo = [mets[k].reshape(-1, e["Np"]) for k in ("Rx", "Ry", "Rz", "Sx", "Sy", "Sz", "Tx", "Ty", "Tz")]
    want = O.phys_deriv(e["Dr"], e["Ds"], e["Dt"], U, mo)
    got = run.result("DuDy").reshape(U.shape)
    assert np.max(np.abs(got - want[1])) <= 1e-13 * np.abs(want[1]).max()
    case = KS.differentiate_dx(e["Dr"], U.reshape(-1), mets["Rx"], K)
    run = _run(device, case, host_K=True)
    assert np.max(np.abs(run.result("Dx") - case[4]["Dx"])) <= 1e-13 * np.abs(case[4]["Dx"]).max()


def test_pointwise_scalars_and_float_rounding(device):
    """A flat loop with a double scalar, float arrays and a float literal: C's
    types (float op float is an FP32 operation) against a numpy restatement."""
    from tests.runner_text import RunnerText
    K = [33, 64, 7]
    rng = np.random.default_rng(1)
    n = sum(K) * 4
    A = rng.standard_normal(n).astype(np.float32)
    B = rng.standard_normal(n).astype(np.float32)
    rt = RunnerText(K)
    rt.array("A", 4, np.float32, data=A)
    rt.array("B", 4, np.float32, data=B)
    rt.array("C", 4, np.float32, output=True)
    rt.array("D", 4, np.float64, output=True)
    rt.scalar("dt")
    order = ["A", "B", "C", "D", "dt"]
    src = rt.source(f"""
#define NP 4
@kernel void pw({rt.signature(order)}) {{
  for (int part = 0; part < NPART; ++part; @outer) {{
    const float* A = A_PART(part);
    const float* B = B_PART(part);
    float* C = C_PART(part);
    double* D = D_PART(part);
    for (int i = 0; i < NP*KpartMax; ++i; @inner) {{
      C[i] = A[i]*B[i] + 0.5f*A[i];
      D[i] = -A[i]*dt + B[i] / 3;
      C[i] += 1.0f;
    }}
  }}
}}""")
    run = _run(device, (rt, src, "pw", order), scalars={"dt": 0.1})
    f = np.float32
    wantC = (A * B + f(0.5) * A) + f(1.0)
    wantD = (-A.astype(np.float64)) * 0.1 + (B / f(3)).astype(np.float64)
    assert np.array_equal(run.result("C"), wantC)
    assert np.array_equal(run.result("D"), wantD)


def test_launch_rejects_bad_arguments(device):
    rt, src, name, order, _ = KS.matmul_known_answer()
    run = DeviceRun(rt, device)
    k = _lib.Kernel(device, src, name)
    args = run.args(order)
    with pytest.raises(_lib.DgkError, match="argument 1 .* must be non-NULL"):
        k.launch([args[0], 0] + args[2:])
    with pytest.raises(_lib.DgkError, match="arguments, the signature has"):
        k.launch(args[:-1])
    with pytest.raises(_lib.DgkError, match="outside 0..KpartMax"):
        k.launch(args, host_K=np.array([2, 9], dtype=np.int64))
