"""The OKL front end of the source-built kernels, on the CPU: what program
libdgk reads from the runner's text (dgk_kernel_describe needs no device),
and which texts it rejects."""
import numpy as np
import pytest

from dgkernel_amd import _lib
from tests import okl_kernels as KS
from tests.runner_text import RunnerText


def _prog(src, name):
    return _lib.describe_kernel(src, name).strip().splitlines()


def _fail(src, name):
    with pytest.raises(_lib.DgkError) as ei:
        _lib.describe_kernel(src, name)
    return ei.value


@pytest.mark.parametrize("int_type", ["INT64", "INT32"])
def test_runner_matrix_kernels_are_read(int_type):
    rt, src, name, _, _ = KS.matmul_known_answer(int_type)
    p = _prog(src, name)
    assert p[0] == f"kernel matmul: NPART 2, KpartMax 3, int_t {'long' if int_type == 'INT64' else 'int'}"
    assert p[1:] == ["matmul TestMat [3x3 static double, column-major, sum double] U -> V"]
    _, src, name, _, _ = KS.differentiate_device_loop(int_type)
    assert _prog(src, name)[1:] == ["matmul Dr [4x4 device double, row-major, sum double] U -> Ur"]
    _, src, name, _, _ = KS.static_matmul(int_type)
    assert _prog(src, name)[1:] == ["matmul M [10x10 static double, column-major, sum double] U -> V"]
    _, src, name, _, _ = KS.device_matmul(int_type)
    assert _prog(src, name)[1:] == ["matmul M [10x10 device double, row-major, sum double] U -> V"]
    _, src, name, _, _ = KS.chain(int_type)
    assert _prog(src, name)[1:] == ["matmul A [5x5 static double, column-major, sum double] U -> V",
                                    "matmul B [5x5 static double, column-major, sum double] V -> W"]


def _deriv_inputs(np_=10, K=(7, 5)):
    rng = np.random.default_rng(0)
    D = [rng.standard_normal((np_, np_)) for _ in range(3)]
    n = sum(K) * np_
    mets = {m: rng.standard_normal(n) for m in KS.METRICS}
    return D, rng.standard_normal(n), mets, list(K)


@pytest.mark.parametrize("fp32", [False, True])
def test_physical_derivative_body_is_fused(fp32):
    D, U, mets, K = _deriv_inputs()
    _, src, name, _ = KS.differentiate_full(D, U, mets, K, fp32=fp32)
    p = _prog(src, name)
    assert len(p) == 2
    assert p[1] == ("physical derivative (Dr_NudgTet2, Ds_NudgTet2, Dt_NudgTet2) of U into DuDx, "
                    "DuDy, DuDz; temporaries Ur, Us, Ut")


def test_dx_body_is_matmul_then_pointwise():
    D, U, mets, K = _deriv_inputs()
    _, src, name, _, _ = KS.differentiate_dx(D[0], U, mets["Rx"], K)
    p = _prog(src, name)
    assert p[1] == "matmul Dr_NudgTet2 [10x10 device double, column-major, sum double] U -> Ur"
    assert p[2] == "pointwise element-node loop x10: Dx = ..."


def test_swapped_chain_rule_runs_unfused():
    """Products written temp-first are not the recognised form: they run as
    three MATMULs and a pointwise loop (same result, more launches)."""
    D, U, mets, K = _deriv_inputs()
    _, src, name, _ = KS.differentiate_full(D, U, mets, K)
    src = src.replace("Rx[i]*Ur[i]", "Ur[i]*Rx[i]")
    p = _prog(src, name)
    assert [l.split()[0] for l in p[1:]] == ["matmul", "matmul", "matmul", "pointwise"]


def test_reference_fp32_kernel_as_written_is_rejected():
    """physical_derivative_test.go:253-255 declares double* for the float
    outputs; that text does not compile under OCCA either."""
    D, U, mets, K = _deriv_inputs()
    _, src, name, _ = KS.differentiate_full(D, U, mets, K, fp32=True)
    bad = src.replace("float* DuDx = DuDx_PART", "double* DuDx = DuDx_PART")
    e = _fail(bad, name)
    assert e.code == 1 and "DuDx declared double but the array holds float" in str(e)


def _one_array_kernel(body, scalars=(), K=(3, 4)):
    rt = RunnerText(list(K))
    rt.array("U", 4)
    rt.array("V", 4, output=True)
    for s in scalars:
        rt.scalar(s)
    order = ["U", "V", *scalars]
    src = rt.source(f"""
#define NP 4
@kernel void k1({rt.signature(order)}) {{
  for (int part = 0; part < NPART; ++part; @outer) {{
    const double* U = U_PART(part);
    double* V = V_PART(part);
{body}
  }}
}}""")
    return src


def test_scalars_and_flat_loop():
    src = _one_array_kernel("""    for (int i = 0; i < NP*KpartMax; ++i; @inner) {
      V[i] = -U[i]*dt + 2.0f*V[i] / (1 + U[i]);
    }""", scalars=("dt",))
    assert _prog(src, "k1")[1] == "pointwise flat-node loop x4: V = ..."


@pytest.mark.parametrize("body,code,needle", [
    ("    V[0] = U[0];", 4, "statement in the partition loop"),
    ("    for (int i = 0; i < NP*KpartMax; ++i; @inner) { V[i] = U[i+1]; }", 4,
     "array read away from the loop's node index"),
    ("    for (int i = 0; i < NP*KpartMax; ++i; @inner) { U[i] = 1.0; }", 1,
     "assignment to const array U"),
    ("    for (int i = 0; i < NP*KpartMax; ++i; @inner) { V[i] = sin(U[i]); }", 4,
     "expression form"),
    ("    while (1) {}", 4, "'while' statement"),
    ("    MATMUL_Q(U, V, K[part]);", 1, "no macro MATMUL_Q"),
])
def test_unsupported_statements_name_their_line(body, code, needle):
    src = _one_array_kernel(body)
    e = _fail(src, "k1")
    assert e.code == code, str(e)
    assert needle in str(e) and "line " in str(e)


def test_missing_kernel_and_bad_signature():
    src = _one_array_kernel("")
    assert _fail(src, "nope").code == 3
    bad = src.replace("const int_t* K,", "")
    assert "first parameter must be the K array" in str(_fail(bad, "k1"))
    nodef = src.replace("#define NPART 2", "")
    assert "NPART" in str(_fail(nodef, "k1"))


def test_user_macros_are_expanded():
    src = _one_array_kernel("""    for (int i = 0; i < NP*KpartMax; ++i; @inner) {
      V[i] = SQ(U[i]) + HALF;
    }""").replace("#define NP 4", "#define NP 4\n#define SQ(x) ((x)*(x))\n#define HALF 0.5")
    assert _prog(src, "k1")[1] == "pointwise flat-node loop x4: V = ..."
