"""Kernel sources in the shapes the reference's callers write, each with its
runner configuration (tests/runner_text.RunnerText) and expected answer.

- matmul:        runner/runner_matrix_test.go:13-125 (static TestMat, K=[2,3])
- differentiate: runner/runner_matrix_test.go:128-212 (device Dr, explicit
                 loop over M[i*NP + j] -> Ur)
- static_matmul, device_matmul: runner/runner_matrix_test.go:215-373
- chain:         runner/runner_matrix_test.go:376-480 (W = B (A U))
- differentiate (Dx = Rx .* Dr U):
                 tetnudg/tet_nudg_mesh_physical_derivative_test.go:92-113
- differentiate (three MATMULs + chain rule, element-node loop):
                 tetnudg/tet_nudg_mesh_physical_derivative_test.go:850-892
- differentiate (FP32 metrics and outputs, flat loop):
                 tetnudg/tet_nudg_mesh_physical_derivative_test.go:240-278
"""
from __future__ import annotations

import numpy as np

from tests.runner_text import RunnerText

OUTER_OPEN = "\tfor (int part = 0; part < NPART; ++part; @outer) {\n"


def matmul_known_answer(int_type="INT64"):
    K = [2, 3]
    np_ = 3
    A = np.array([[2.0, 1.0, 0.0], [0.0, 1.0, 1.0], [1.0, 0.0, 2.0]])
    U = np.array([float(e + i + 1) for e in range(sum(K)) for i in range(np_)])
    rt = RunnerText(K, int_type)
    rt.static_matrix("TestMat", A)
    rt.array("U", np_, data=U)
    rt.array("V", np_, output=True)
    order = ["TestMat", "U", "V"]
    src = rt.source(f"""
#define NP {np_}

@kernel void matmul(
\t{rt.signature(order)}
) {{
{OUTER_OPEN}\t\tconst double* U = U_PART(part);
\t\tdouble* V = V_PART(part);
\t\t
\t\tMATMUL_TestMat(U, V, K[part]);
\t}}
}}""")
    want = {"V": (U.reshape(-1, np_) @ A.T).reshape(-1)}
    return rt, src, "matmul", order, want


def differentiate_device_loop(int_type="INT64"):
    K = [10]
    np_ = 4
    Dr = np.array([[-2.0, 3.0, -1.0, 0.0], [-1.0, 0.0, 1.0, 0.0], [0.0, -1.0, 0.0, 1.0],
                   [0.0, 1.0, -3.0, 2.0]])
    U = np.array([(e / (K[0] - 1)) ** 2 for e in range(K[0]) for _ in range(np_)])
    rt = RunnerText(K, int_type)
    rt.device_matrix("Dr", Dr)
    rt.array("U", np_, data=U)
    rt.array("Ur", np_, output=True)
    order = ["Dr", "U", "Ur"]
    src = rt.source(f"""
#define NP {np_}

@kernel void differentiate(
\t{rt.signature(order)}
) {{
\tfor (int part = 0; part < NPART; ++part; @outer) {{
\t\tconst double* U = U_PART(part);
\t\tdouble* Ur = Ur_PART(part);
\t\t
\t\t// Matrix multiply with runtime matrix
\t\tfor (int elem = 0; elem < K[part]; ++elem) {{
\t\t\tfor (int i = 0; i < NP; ++i; @inner) {{
\t\t\t\tdouble sum = 0.0;
\t\t\t\tfor (int j = 0; j < NP; ++j) {{
\t\t\t\t\tsum += Dr[i*NP + j] * U[elem*NP + j];
\t\t\t\t}}
\t\t\t\tUr[elem*NP + i] = sum;
\t\t\t}}
\t\t}}
\t}}
}}""")
    # the device copy is column-major (runner.go:375-415), so Dr[i*NP + j]
    # reads Dr(j, i): the loop applies Dr transposed
    mem = Dr.T.reshape(-1)
    want = {"Ur": (U.reshape(-1, np_) @ mem.reshape(np_, np_).T).reshape(-1)}
    return rt, src, "differentiate", order, want


def _tridiag(n):
    M = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                M[i, j] = 2.0
            elif abs(i - j) == 1:
                M[i, j] = -1.0
    return M


def static_matmul(int_type="INT64"):
    K = [50]
    np_ = 10
    M = _tridiag(np_)
    U = np.array([float(i % 10) for i in range(K[0] * np_)])
    rt = RunnerText(K, int_type)
    rt.static_matrix("M", M)
    rt.array("U", np_, data=U)
    rt.array("V", np_, output=True)
    order = ["M", "U", "V"]
    src = rt.source(f"""
#define NP {np_}
@kernel void static_matmul({rt.signature(order)}) {{
\tfor (int part = 0; part < NPART; ++part; @outer) {{
\t\tconst double* U = U_PART(part);
\t\tdouble* V = V_PART(part);
\t\tMATMUL_M(U, V, K[part]);
\t}}
}}""")
    return rt, src, "static_matmul", order, {"V": (U.reshape(-1, np_) @ M.T).reshape(-1)}


def device_matmul(int_type="INT64"):
    K = [50]
    np_ = 10
    M = _tridiag(np_)
    M[0, 1] = -0.5  # not symmetric: the row-major read of column-major memory shows
    U = np.array([float(i % 10) for i in range(K[0] * np_)])
    rt = RunnerText(K, int_type)
    rt.device_matrix("M", M)
    rt.array("U", np_, data=U)
    rt.array("V", np_, output=True)
    order = ["M", "U", "V"]
    src = rt.source(f"""
#define NP {np_}
@kernel void device_matmul({rt.signature(order)}) {{
\tfor (int part = 0; part < NPART; ++part; @outer) {{
\t\tconst double* U = U_PART(part);
\t\tdouble* V = V_PART(part);
\t\t
\t\tfor (int elem = 0; elem < K[part]; ++elem) {{
\t\t\tfor (int i = 0; i < NP; ++i; @inner) {{
\t\t\t\tdouble sum = 0.0;
\t\t\t\tfor (int j = 0; j < NP; ++j) {{
\t\t\t\t\tsum += M[i*NP + j] * U[elem*NP + j];
\t\t\t\t}}
\t\t\t\tV[elem*NP + i] = sum;
\t\t\t}}
\t\t}}
\t}}
}}""")
    return rt, src, "device_matmul", order, {"V": (U.reshape(-1, np_) @ M).reshape(-1)}


def chain(int_type="INT64"):
    K = [20]
    np_ = 5
    A = np.array([[(i + j) / 10.0 for j in range(np_)] for i in range(np_)])
    B = np.array([[abs(i - j) / 5.0 for j in range(np_)] for i in range(np_)])
    U = np.array([float(i % np_) for i in range(K[0] * np_)])
    rt = RunnerText(K, int_type)
    rt.static_matrix("A", A)
    rt.static_matrix("B", B)
    rt.array("U", np_, data=U)
    rt.array("V", np_, output=True)
    rt.array("W", np_, output=True)
    order = ["A", "B", "U", "V", "W"]
    src = rt.source(f"""
#define NP {np_}

@kernel void chain(
\t{rt.signature(order)}
) {{
\tfor (int part = 0; part < NPART; ++part; @outer) {{
\t\tconst double* U = U_PART(part);
\t\tdouble* V = V_PART(part);
\t\tdouble* W = W_PART(part);
\t\t
\t\t// First operation: V = A * U
\t\tMATMUL_A(U, V, K[part]);
\t\t
\t\t// Second operation: W = B * V
\t\tMATMUL_B(V, W, K[part]);
\t}}
}}""")
    V = U.reshape(-1, np_) @ A.T
    return rt, src, "chain", order, {"V": V.reshape(-1), "W": (V @ B.T).reshape(-1)}


METRICS = ("Rx", "Sx", "Tx", "Ry", "Sy", "Ty", "Rz", "Sz", "Tz")


def differentiate_dx(Dr, U, Rx, K, int_type="INT64", short="NudgTet2"):
    """Dx = Rx .* (Dr U) (physical_derivative_test.go:92-113); Dr a device matrix."""
    np_ = Dr.shape[0]
    rt = RunnerText(K, int_type)
    rt.device_matrix("Dr_" + short, Dr)
    rt.array("U", np_, data=U)
    rt.array("Rx", np_, data=Rx)
    rt.array("Dx", np_, output=True)
    rt.array("Ur", np_, output=True)
    order = ["U", "Rx", "Dx", "Dr_" + short, "Ur"]
    src = rt.source(f"""
#define NP {np_}

@kernel void differentiate({rt.signature(order)}) {{
    for (int part = 0; part < NPART; ++part; @outer) {{
        const double* U = U_PART(part);
        double* Ur = Ur_PART(part);
        MATMUL_Dr_{short}(U, Ur, K[part]);

        double* Dx = Dx_PART(part);
        const double* Rx = Rx_PART(part);
		for (int k = 0; k < KpartMax; ++k; @inner) {{
			if (k < K[part]) {{
				for (int n = 0; n < NP; ++n) {{
					int i = n + k*NP;
					Dx[i] = Rx[i]*Ur[i];
				}}
			}}
		}}
    }}
}}
""")
    Ur = (U.reshape(-1, np_) @ Dr.T).reshape(-1)
    return rt, src, "differentiate", order, {"Ur": Ur, "Dx": Rx * Ur}


def differentiate_full(D, U, mets, K, int_type="INT64", short="NudgTet2", fp32=False):
    """The physical derivative kernel: FP64 form with the element-node loop
    (:850-892), or FP32 metrics/outputs with the flat loop (:240-278).
    mets: dict name -> flat array (element-major per partition)."""
    np_ = D[0].shape[0]
    rt = RunnerText(K, int_type)
    for q, nm in enumerate(("Dr_", "Ds_", "Dt_")):
        rt.device_matrix(nm + short, D[q])
    mt = np.float32 if fp32 else np.float64
    rt.array("U", np_, data=U)
    for nm in METRICS:
        rt.array(nm, np_, dtype=mt, data=mets[nm])
    for nm in ("DuDx", "DuDy", "DuDz"):
        rt.array(nm, np_, dtype=mt, output=True)
    for nm in ("Ur", "Us", "Ut"):
        rt.array(nm, np_, output=True)
    order = (["U", *METRICS, "DuDx", "DuDy", "DuDz"] + ["Dr_" + short, "Ds_" + short, "Dt_" + short]
             + ["Ur", "Us", "Ut"])
    ty = "float" if fp32 else "double"
    decl = "\n".join(f"        const {ty}* {nm} = {nm}_PART(part);" for nm in METRICS)
    if fp32:
        loop = """		for (int i = 0; i < NP*KpartMax; ++i; @inner) {
			DuDx[i] = Rx[i]*Ur[i] + Sx[i]*Us[i] + Tx[i]*Ut[i];
			DuDy[i] = Ry[i]*Ur[i] + Sy[i]*Us[i] + Ty[i]*Ut[i];
			DuDz[i] = Rz[i]*Ur[i] + Sz[i]*Us[i] + Tz[i]*Ut[i];
		}"""
    else:
        loop = """		for (int k = 0; k < KpartMax; ++k; @inner) {
			if (k < K[part]) {
				for (int n = 0; n < NP; ++n) {
					int i = n + k*NP;
					DuDx[i] = Rx[i]*Ur[i] + Sx[i]*Us[i] + Tx[i]*Ut[i];
					DuDy[i] = Ry[i]*Ur[i] + Sy[i]*Us[i] + Ty[i]*Ut[i];
					DuDz[i] = Rz[i]*Ur[i] + Sz[i]*Us[i] + Tz[i]*Ut[i];
				}
			}
		}"""
    src = rt.source(f"""
#define NP {np_}

@kernel void differentiate({rt.signature(order)}) {{
    for (int part = 0; part < NPART; ++part; @outer) {{
        const double* U = U_PART(part);
        double* Ur = Ur_PART(part);
        double* Us = Us_PART(part);
        double* Ut = Ut_PART(part);
        MATMUL_Dr_{short}(U, Ur, K[part]);
        MATMUL_Ds_{short}(U, Us, K[part]);
        MATMUL_Dt_{short}(U, Ut, K[part]);

        double* DuDx = DuDx_PART(part);
        double* DuDy = DuDy_PART(part);
        double* DuDz = DuDz_PART(part);
{decl}
        // ************************************************************
		// *** This computation is happening in column-major format ***
        // ************************************************************
{loop}
    }}
}}
""".replace("double* DuDx", f"{ty}* DuDx").replace("double* DuDy", f"{ty}* DuDy")
       .replace("double* DuDz", f"{ty}* DuDz"))
    return rt, src, "differentiate", order
