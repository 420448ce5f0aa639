"""Reference tetrahedron of the NUDG family, built on the host with numpy.

This is the host-side setup a caller runs once before launching the device
kernels: warp & blend nodes, the orthonormal PKD basis, the Vandermonde matrix
and its inverse, Dr/Ds/Dt, the face masks and LIFT. It mirrors what the
reference's Go caller gets from ``gonudg.NewDG3D``/``StartUp3D``
(element/library/gonudg/startup3d.go:73-118) and hands to the runner as device
matrices (tetnudg/tet_nudg_mesh.go:370-389, names ``Dr_NudgTetN`` ...).

The node construction follows element/library/gonudg/nodes3d.go literally,
including its two departures from the Hesthaven-Warburton text for P >= 3
(0-based ``alphastore`` lookup, nodes3d.go:19-28; ``evalwarp`` pairing the
negated GL nodes with ascending equispaced nodes and multiplying by 1-x^2,
nodes3d.go:227-231,290-323): the reference's r,s,t -- and hence every operator
-- at P >= 3 are defined by that code.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# nodes3d.go:19-23
_ALPHASTORE = (0.0000, 0.0000, 0.0000, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
               1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655)
NODETOL = 1e-7  # startup3d.go:81


def jacobi_p(x, alpha: float, beta: float, n: int) -> np.ndarray:
    """Orthonormal Jacobi polynomial P_n^(alpha,beta)(x); gonudg/jacobip.go:30-97."""
    x = np.asarray(x, dtype=np.float64)
    ab, ab1, a1, b1 = alpha + beta, alpha + beta + 1.0, alpha + 1.0, beta + 1.0
    g0 = 2.0 ** ab1 / ab1 * math.gamma(a1) * math.gamma(b1) / math.gamma(ab1)
    p_prev = np.full_like(x, 1.0 / math.sqrt(g0))
    if n == 0:
        return p_prev
    g1 = a1 * b1 / (ab + 3.0) * g0
    p = ((ab + 2.0) * x / 2.0 + (alpha - beta) / 2.0) * (1.0 / math.sqrt(g1))
    if n == 1:
        return p
    aold = 2.0 / (2.0 + ab) * math.sqrt(a1 * b1 / (ab + 3.0))
    for i in range(1, n):
        h1 = 2.0 * i + ab
        anew = 2.0 / (h1 + 2.0) * math.sqrt((i + 1.0) * (i + ab1) * (i + a1) * (i + b1)
                                            / (h1 + 1.0) / (h1 + 3.0))
        bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0)
        p_prev, p = p, 1.0 / anew * (-aold * p_prev + (x - bnew) * p)
        aold = anew
    return p


def grad_jacobi_p(x, alpha: float, beta: float, n: int) -> np.ndarray:
    """d/dx of the orthonormal Jacobi polynomial; gonudg/jacobip.go:107-121."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.zeros_like(x)
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def jacobi_gq(alpha: float, beta: float, n: int):
    """Gauss-Jacobi nodes/weights by Golub-Welsch; gonudg/jacobi_quadrature.go:39-101."""
    if n == 0:
        return np.array([-(alpha - beta) / (alpha + beta + 2.0)]), np.array([2.0])
    h1 = 2.0 * np.arange(n + 1) + alpha + beta
    with np.errstate(divide="ignore", invalid="ignore"):
        d0 = (beta * beta - alpha * alpha) / (h1 * (h1 + 2.0))
    if alpha + beta < 10 * 1e-16:
        d0[0] = 0.0
    i1 = np.arange(1, n + 1, dtype=np.float64)
    d1 = 2.0 / (h1[:n] + 2.0) * np.sqrt(i1 * (i1 + alpha + beta) * (i1 + alpha) * (i1 + beta)
                                        / (h1[:n] + 1.0) / (h1[:n] + 3.0))
    jj = np.diag(d0) + np.diag(d1, 1) + np.diag(d1, -1)
    x, v = np.linalg.eigh(jj)
    g0 = (math.gamma(alpha + 1) * math.gamma(beta + 1) * 2.0 ** (alpha + beta + 1)
          / (alpha + beta + 1) / math.gamma(alpha + beta + 1))
    return x, v[0, :] ** 2 * g0


def jacobi_gl(alpha: float, beta: float, n: int) -> np.ndarray:
    """Gauss-Lobatto nodes; gonudg/jacobi_quadrature.go:16-37."""
    if n == 0:
        return np.array([0.0])
    if n == 1:
        return np.array([-1.0, 1.0])
    xint, _ = jacobi_gq(alpha + 1, beta + 1, n - 2)
    return np.concatenate(([-1.0], xint, [1.0]))


def equi_nodes3d(p: int):
    """Equispaced lattice in r,s,t order l fastest; gonudg/nodes3d.go:177-204."""
    if p == 0:
        return np.array([-1.0]), np.array([-1.0]), np.array([-1.0])
    lmn = [(l, m, n) for n in range(p + 1) for m in range(p + 1 - n) for l in range(p + 1 - n - m)]
    a = np.array(lmn, dtype=np.float64)
    return -1.0 + 2.0 * a[:, 0] / p, -1.0 + 2.0 * a[:, 1] / p, -1.0 + 2.0 * a[:, 2] / p


def _evalwarp(p: int, gauss_x: np.ndarray, xs: np.ndarray) -> np.ndarray:
    """gonudg/nodes3d.go:290-323 (Lagrange interpolant of gauss_x - xeq, times 1-x^2)."""
    xeq = -1.0 + 2.0 * np.arange(p + 1) / p
    warp = np.zeros_like(xs)
    for j in range(p + 1):
        lag = np.ones_like(xs)
        for k in range(p + 1):
            if k != j:
                lag = lag * ((xs - xeq[k]) / (xeq[j] - xeq[k]))
        warp = warp + lag * (gauss_x[j] - xeq[j])
    return warp * (1.0 - xs * xs)


def _evalshift(p: int, pval: float, l1, l2, l3):
    """gonudg/nodes3d.go:221-287."""
    gauss_x = -jacobi_gl(0.0, 0.0, p)
    w1 = _evalwarp(p, gauss_x, l3 - l2) * 4.0
    w2 = _evalwarp(p, gauss_x, l1 - l3) * 4.0
    w3 = _evalwarp(p, gauss_x, l2 - l1) * 4.0
    warp1 = l2 * l3 * w1 * (1.0 + pval * pval * l1 * l1)
    warp2 = l1 * l3 * w2 * (1.0 + pval * pval * l2 * l2)
    warp3 = l1 * l2 * w3 * (1.0 + pval * pval * l3 * l3)
    c2, c4 = math.cos(2.0 * math.pi / 3.0), math.cos(4.0 * math.pi / 3.0)
    s2, s4 = math.sin(2.0 * math.pi / 3.0), math.sin(4.0 * math.pi / 3.0)
    return (1.0 * warp1 + c2 * warp2 + c4 * warp3,
            0.0 * warp1 + s2 * warp2 + s4 * warp3)


_SQ3, _SQ6 = math.sqrt(3.0), math.sqrt(6.0)
_V1 = np.array([-1.0, -1.0 / _SQ3, -1.0 / _SQ6])
_V2 = np.array([1.0, -1.0 / _SQ3, -1.0 / _SQ6])
_V3 = np.array([0.0, 2.0 / _SQ3, -1.0 / _SQ6])
_V4 = np.array([0.0, 0.0, 3.0 / _SQ6])


def nodes3d(p: int):
    """Warp & blend nodes in the equilateral tetrahedron; gonudg/nodes3d.go:17-174."""
    alpha = _ALPHASTORE[p] if 0 <= p < 15 else 1.0
    tol = 1e-10
    r, s, t = equi_nodes3d(p)
    L1, L2, L3, L4 = (1.0 + t) / 2.0, (1.0 + s) / 2.0, -(1.0 + r + s + t) / 2.0, (1.0 + r) / 2.0
    t1 = [_V2 - _V1, _V2 - _V1, _V3 - _V2, _V3 - _V1]
    t2 = [_V3 - 0.5 * (_V1 + _V2), _V4 - 0.5 * (_V1 + _V2),
          _V4 - 0.5 * (_V2 + _V3), _V4 - 0.5 * (_V1 + _V3)]
    t1 = [v / math.sqrt(float(v @ v)) for v in t1]
    t2 = [v / math.sqrt(float(v @ v)) for v in t2]
    X = L3 * _V1[0] + L4 * _V2[0] + L2 * _V3[0] + L1 * _V4[0]
    Y = L3 * _V1[1] + L4 * _V2[1] + L2 * _V3[1] + L1 * _V4[1]
    Z = L3 * _V1[2] + L4 * _V2[2] + L2 * _V3[2] + L1 * _V4[2]
    shift = np.zeros((3, X.size))
    orders = [(L1, L2, L3, L4), (L2, L1, L3, L4), (L3, L1, L4, L2), (L4, L1, L3, L2)]
    for face, (La, Lb, Lc, Ld) in enumerate(orders):
        warp1, warp2 = _evalshift(p, alpha, Lb, Lc, Ld)
        blend = Lb * Lc * Ld
        denom = (Lb + 0.5 * La) * (Lc + 0.5 * La) * (Ld + 0.5 * La)
        big = denom > tol
        blend = np.where(big, (1.0 + alpha * alpha * La * La) * blend / np.where(big, denom, 1.0), blend)
        for d in range(3):
            shift[d] += blend * warp1 * t1[face][d] + blend * warp2 * t2[face][d]
        cnt = (Lb > tol).astype(int) + (Lc > tol) + (Ld > tol)
        onface = (La < tol) & (cnt < 3)
        for d in range(3):
            shift[d] = np.where(onface, warp1 * t1[face][d] + warp2 * t2[face][d], shift[d])
    return X + shift[0], Y + shift[1], Z + shift[2]


def xyz_to_rst(X, Y, Z):
    """Equilateral -> reference tetrahedron; gonudg/xyztorst.go:8-69."""
    A = 0.5 * np.column_stack((_V2 - _V1, _V3 - _V1, _V4 - _V1))
    off = 0.5 * (_V2 + _V3 + _V4 - _V1)
    rst = np.linalg.solve(A, np.vstack((X - off[0], Y - off[1], Z - off[2])))
    return rst[0], rst[1], rst[2]


def rst_to_abc(r, s, t):
    """Collapsed coordinates; gonudg/rsttoabc.go:7-34."""
    tol = 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(np.abs(s + t) < tol, -1.0, 2.0 * (1.0 + r) / (-s - t) - 1.0)
        b = np.where(np.abs(t - 1.0) < tol, -1.0, 2.0 * (1.0 + s) / (1.0 - t) - 1.0)
    return a, b, np.array(t, dtype=np.float64)


def simplex3dp(a, b, c, i: int, j: int, k: int) -> np.ndarray:
    """Orthonormal PKD mode (i,j,k); gonudg/simplex3dp.go:15-36."""
    h1 = jacobi_p(a, 0.0, 0.0, i)
    h2 = jacobi_p(b, 2.0 * i + 1, 0.0, j)
    h3 = jacobi_p(c, 2.0 * (i + j) + 2, 0.0, k)
    return 2.0 * math.sqrt(2.0) * h1 * h2 * (1.0 - b) ** i * h3 * (1.0 - c) ** (i + j)


def grad_simplex3dp(r, s, t, i: int, j: int, k: int):
    """Gradient of mode (i,j,k) in r,s,t; gonudg/simplex3dp.go:56-121."""
    a, b, c = rst_to_abc(r, s, t)
    fa, dfa = jacobi_p(a, 0, 0, i), grad_jacobi_p(a, 0, 0, i)
    gb, dgb = jacobi_p(b, 2 * i + 1, 0, j), grad_jacobi_p(b, 2 * i + 1, 0, j)
    hc, dhc = jacobi_p(c, 2 * (i + j) + 2, 0, k), grad_jacobi_p(c, 2 * (i + j) + 2, 0, k)
    hb, hcc = 0.5 * (1.0 - b), 0.5 * (1.0 - c)
    vr = dfa * gb * hc
    if i > 0:
        vr = vr * hb ** (i - 1)
    if i + j > 0:
        vr = vr * hcc ** (i + j - 1)
    tmp = dgb * hb ** i
    if i > 0:
        tmp = tmp - (0.5 * i) * gb * hb ** (i - 1)
    if i + j > 0:
        tmp = tmp * hcc ** (i + j - 1)
    tmp = fa * tmp * hc
    vs = 0.5 * (1.0 + a) * vr + tmp
    vt = 0.5 * (1.0 + a) * vr + 0.5 * (1.0 + b) * tmp
    tmp2 = dhc * hcc ** (i + j)
    if i + j > 0:
        tmp2 = tmp2 - (0.5 * (i + j)) * hc * hcc ** (i + j - 1)
    vt = vt + fa * gb * tmp2 * hb ** i
    nf = 2.0 ** (2 * i + j + 1.5)
    return vr * nf, vs * nf, vt * nf


def _modes3d(n: int):
    return [(i, j, k) for i in range(n + 1) for j in range(n + 1 - i) for k in range(n + 1 - i - j)]


def vandermonde3d(n: int, r, s, t) -> np.ndarray:
    """V_ij = phi_j(r_i); gonudg/vandermonde3d.go:15-41."""
    a, b, c = rst_to_abc(r, s, t)
    return np.column_stack([simplex3dp(a, b, c, i, j, k) for (i, j, k) in _modes3d(n)])


def grad_vandermonde3d(n: int, r, s, t):
    """gonudg/vandermonde3d.go:45-72."""
    cols = [grad_simplex3dp(r, s, t, i, j, k) for (i, j, k) in _modes3d(n)]
    return tuple(np.column_stack([c[d] for c in cols]) for d in range(3))


def vandermonde2d(n: int, R, S) -> np.ndarray:
    """2D face Vandermonde; gonudg/vandermonde2D.go:9-72."""
    R, S = np.asarray(R, np.float64), np.asarray(S, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(S != 1, 2 * (1 + R) / (1 - S) - 1, -1.0)
    b = S
    cols = []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            cols.append(1.4142135623730951 * jacobi_p(a, 0, 0, i) * jacobi_p(b, 2 * i + 1, 0, j)
                        * (1 - b) ** i)
    return np.column_stack(cols)


@dataclass
class RefTet:
    """Operators of the order-N NUDG tetrahedron (the reference's NUDGTet fields)."""

    N: int
    Np: int
    Nfp: int
    r: np.ndarray
    s: np.ndarray
    t: np.ndarray
    V: np.ndarray
    Vinv: np.ndarray
    Mass: np.ndarray
    Dr: np.ndarray
    Ds: np.ndarray
    Dt: np.ndarray
    LIFT: np.ndarray
    Fmask: np.ndarray  # [4, Nfp]
    extra: dict = field(default_factory=dict)

    @property
    def short_name(self) -> str:
        """``GetProperties().ShortName``; tetnudg/tet_nudg_mesh.go:230-246."""
        return "NudgTet" + chr(ord("0") + self.N)

    def ref_matrices(self) -> dict:
        """``GetRefMatrices()`` names; tetnudg/tet_nudg_mesh.go:370-389.

        ``M_`` is V*V^T exactly as the reference computes it
        (tet_nudg_mesh.go:338-349); that product is the inverse mass matrix.
        """
        sn = self.short_name
        M = self.V @ self.V.T
        return {"V_" + sn: self.V, "Vinv_" + sn: self.Vinv, "M_" + sn: M,
                "Minv_" + sn: np.linalg.inv(M), "Dr_" + sn: self.Dr, "Ds_" + sn: self.Ds,
                "Dt_" + sn: self.Dt, "LIFT_" + sn: self.LIFT}


def build_ref_tet(N: int) -> RefTet:
    """``StartUp3D`` reference-element part; gonudg/startup3d.go:94-118,179-209, lift3d.go."""
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    Nfp = (N + 1) * (N + 2) // 2
    r, s, t = xyz_to_rst(*nodes3d(N))
    V = vandermonde3d(N, r, s, t)
    Vinv = np.linalg.inv(V)
    Vr, Vs, Vt = grad_vandermonde3d(N, r, s, t)
    Dr, Ds, Dt = Vr @ Vinv, Vs @ Vinv, Vt @ Vinv
    fm = [np.nonzero(np.abs(1.0 + t) < NODETOL)[0], np.nonzero(np.abs(1.0 + s) < NODETOL)[0],
          np.nonzero(np.abs(1.0 + r + s + t) < NODETOL)[0], np.nonzero(np.abs(1.0 + r) < NODETOL)[0]]
    if any(f.size != Nfp for f in fm):
        raise ValueError(f"face mask construction failed at N={N}")
    Fmask = np.stack(fm).astype(np.int32)
    face_coords = [(r, s), (r, t), (s, t), (s, t)]  # lift3d.go:251-283
    E = np.zeros((Np, 4 * Nfp))
    for f in range(4):
        fr, fs = face_coords[f]
        VF = vandermonde2d(N, fr[Fmask[f]], fs[Fmask[f]])
        E[np.ix_(Fmask[f], np.arange(f * Nfp, (f + 1) * Nfp))] = np.linalg.inv(VF @ VF.T)
    LIFT = V @ (V.T @ E)
    return RefTet(N=N, Np=Np, Nfp=Nfp, r=r, s=s, t=t, V=V, Vinv=Vinv, Mass=Vinv.T @ Vinv,
                  Dr=Dr, Ds=Ds, Dt=Dt, LIFT=LIFT, Fmask=Fmask)


def mode_degrees(N: int) -> np.ndarray:
    """Total degree i+j+k of each modal column of V, in Vandermonde3D order."""
    return np.array([i + j + k for (i, j, k) in _modes3d(N)], dtype=np.int32)
