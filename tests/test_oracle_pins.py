"""Pins of the CPU oracle against the reference's own known-answer tests,
closed forms and invariants (no GPU). Each test names the reference test or
the mathematical fact it restates."""
import itertools
import math

import numpy as np
import pytest

from dgkernel_amd import dg
from dgkernel_amd import mesh as M
from dgkernel_amd import nudg
from oracle import oracle as O
from tests.helpers import neu_mesh, rel_l2


# element/library/gonudg/jacobi_quadrature_test.go:9-130
@pytest.mark.parametrize("alpha,beta,N,expected", [
    (0.0, 0.0, 0, [0.0]),
    (0.0, 0.0, 1, [-1.0, 1.0]),
    (0.0, 0.0, 2, [-1.0, 0.0, 1.0]),
    (0.0, 0.0, 3, [-1.0, -0.4472135954999579, 0.4472135954999579, 1.0]),
    (0.0, 0.0, 4, [-1.0, -0.6546536707079771, 0.0, 0.6546536707079771, 1.0]),
    (0.0, 0.0, 5, [-1.0, -0.7650553239294647, -0.2852315164806451, 0.2852315164806451,
                   0.7650553239294647, 1.0]),
    (1.0, 1.0, 2, [-1.0, 0.0, 1.0]),
])
def test_jacobi_gl_known_values(alpha, beta, N, expected):
    got = O.jacobi_gl(alpha, beta, N)
    assert np.max(np.abs(got - np.array(expected))) <= 1e-14
    assert np.max(np.abs(nudg.jacobi_gl(alpha, beta, N) - np.array(expected))) <= 1e-14


@pytest.mark.parametrize("N", [1, 2, 3, 5, 8, 12])
def test_jacobi_gq_is_gauss_legendre(N):
    """alpha=beta=0 Gauss-Jacobi = Gauss-Legendre (numpy's leggauss), weights included."""
    x, w = O.jacobi_gq(0.0, 0.0, N)
    xl, wl = np.polynomial.legendre.leggauss(N + 1)
    assert np.max(np.abs(x - xl)) <= 1e-14
    assert np.max(np.abs(w - wl)) <= 1e-14


def test_jacobi_p_normalization_constants():
    """element/library/gonudg/jacobip_test.go:8-93: P_0 values and P000 = sqrt(3)/2."""
    x = np.array([-1.0, -0.3, 0.0, 0.7, 1.0])
    p00 = O.jacobi_p(x, 0, 0, 0)
    p10 = O.jacobi_p(x, 1, 0, 0)
    p20 = O.jacobi_p(x, 2, 0, 0)
    assert np.allclose(p00, 1 / math.sqrt(2), atol=1e-14, rtol=0)
    assert np.allclose(p10, 1 / math.sqrt(2), atol=1e-14, rtol=0)
    assert np.allclose(p20, math.sqrt(3 / 8), atol=1e-14, rtol=0)
    r = np.array([-1.0, -0.5, 0.1]); s = np.array([-1.0, -0.2, -0.5]); t = np.array([-1.0, 0.0, -0.9])
    P, *_ = O.simplex3dp(r, s, t, 0, 0, 0)
    assert np.allclose(P, math.sqrt(3) / 2, atol=1e-14, rtol=0)


@pytest.mark.parametrize("a,b,N", [(0, 0, 3), (1, 0, 4), (3, 0, 2), (0, 0, 7)])
def test_jacobi_p_orthonormal(a, b, N):
    """Orthonormal w.r.t. (1-x)^a (1+x)^b on [-1,1], checked with Gauss-Jacobi quadrature."""
    xq, wq = O.jacobi_gq(float(a), float(b), N + 2)
    for m in range(N + 1):
        for n in range(N + 1):
            v = np.sum(wq * O.jacobi_p(xq, a, b, m) * O.jacobi_p(xq, a, b, n))
            assert abs(v - (1.0 if m == n else 0.0)) <= 1e-13


def test_grad_jacobi_p_finite_difference():
    """simplex3dp_test.go:123-194 style: derivative against a central difference."""
    x = np.linspace(-0.9, 0.9, 7)
    h = 1e-6
    for (a, b, N) in [(0, 0, 4), (3, 0, 3), (6, 0, 2)]:
        g = O.jacobi_p(x, a, b, N, grad=True)
        fd = (O.jacobi_p(x + h, a, b, N) - O.jacobi_p(x - h, a, b, N)) / (2 * h)
        assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.abs(g).max())


def test_nodes_p1_are_vertices_and_p2_equispaced():
    """nodes3d_test.go:101-128 (P=1 vertices); P=2 nodes equispaced (alpha=0 and zero warp)."""
    e1 = O.element(1)
    want = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)
    got = np.stack([e1["r"], e1["s"], e1["t"]], axis=1)
    assert np.max(np.abs(got - want)) <= 1e-14
    e2 = O.element(2)
    r, s, t = nudg.equi_nodes3d(2)
    assert np.max(np.abs(e2["r"] - r)) <= 1e-14
    assert np.max(np.abs(e2["s"] - s)) <= 1e-14
    assert np.max(np.abs(e2["t"] - t)) <= 1e-14


@pytest.mark.parametrize("N", range(1, 9))
def test_nodes_two_restatements_agree(N):
    """nodes3d.go transcribed twice (C++ oracle, numpy product setup)."""
    X, Y, Z = O.nodes3d(N)
    x, y, z = nudg.nodes3d(N)
    assert max(np.abs(X - x).max(), np.abs(Y - y).max(), np.abs(Z - z).max()) <= 1e-15


# --- Nodes3D at P >= 3, pinned by what the reference's code reduces to -------
#
# The reference's warp & blend nodes (element/library/gonudg/nodes3d.go:17-323)
# depart from the Hesthaven-Warburton text: alphastore is read 0-based
# (:19-28) and evalwarp pairs the negated (descending) GL points with the
# ascending equispaced points and multiplies by 1 - x^2 (:227-231, :295-319).
# The closed forms below are derived by hand from that code; they are
# evaluated here through numpy's Legendre class and a polynomial fit, not by
# the oracle's Golub-Welsch / Lagrange-product route.
#
# Edge s = t = -1 (L1 = L2 = 0): the blends of faces 2 and 3 vanish
# (Lb = L1 = 0), faces 0 and 1 overwrite the shift (La = 0, count = 2), so
# only face 1's warp survives. In its evalshift call (L1, L2, L3) =
# (0, (1-r)/2, (1+r)/2), so warp1 = (1-r)(1+r)/4 * 4 evalwarp(r), warp2 =
# warp3 = 0 and the node moves along t1[1] = (1,0,0):
#     r' = r + (1 - r^2)^2 sum_j l_j(r) (-x_j^GL - r_j),
# l_j the Lagrange basis on the equispaced r_j, x^GL ascending. At the
# equispaced nodes themselves l_j(r_k) = delta_jk. At P=3 this puts the edge
# nodes at +-0.283395 where the textbook puts the GL points +-0.447214.
#
# Face t = -1 interior node at P=4, barycentrics (a, b, c) = (1/2, 1/4, 1/4)
# of (v3, v1, v2): faces 1-3 have zero blend, face 0 has blend 1 and no
# overwrite (count = 3). In evalshift w1 = 0 (c - b = 0, and the warp is odd),
# w3 = -w2, so dx = 0 and dy = sqrt(3) w2 along t2[0] = (0,1,0), which moves
# the node along the median to v3: lambda(v3) = 1/2 + w2 with
#     w2 = a c 4 evalwarp(1/4) (1 + alpha^2 b^2) = (1/2)(1 + alpha^2/16)(15/16) W(1/4),
# W(x) = sum_j l_j(x)(-x_j^GL - r_j), alpha = alphastore[4] = 1.1332 (0-based
# read; the textbook's alpha(4) would be 0.1002).


def _gl_points(P):
    """Gauss-Lobatto-Legendre points, ascending: the ends and the roots of P_N'."""
    inner = np.polynomial.legendre.Legendre.basis(P).deriv().roots()
    return np.concatenate(([-1.0], np.sort(inner.real), [1.0]))


def _warp_interp(P, x):
    """W(x): the degree-P interpolant through (r_j, -x_j^GL - r_j), by polyfit."""
    r = np.linspace(-1.0, 1.0, P + 1)
    c = np.polynomial.polynomial.polyfit(r, -_gl_points(P) - r, P)
    return np.polynomial.polynomial.polyval(x, c)


def _bary(r, s, t):
    """Barycentrics of the vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1)."""
    return np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2], axis=1)


def _edge_error(lam, N):
    """Max deviation of the nodes on the six edges from the closed form."""
    req = np.linspace(-1.0, 1.0, N + 1)
    want = np.sort((1 + req + (1 - req ** 2) ** 2 * (-_gl_points(N) - req)) / 2)
    worst = 0.0
    for i, j in itertools.combinations(range(4), 2):
        others = [k for k in range(4) if k not in (i, j)]
        on = np.all(np.abs(lam[:, others]) <= 1e-12, axis=1)
        got = np.sort(lam[on, j])
        if got.size != N + 1:
            return np.inf
        worst = max(worst, float(np.abs(got - want).max()))
    return worst


def _face_p4_error(lam):
    """Max deviation of the 12 face-interior nodes at P=4 from the closed form."""
    w2 = 0.5 * (1 + 1.1332 ** 2 / 16) * (15 / 16) * _warp_interp(4, 0.25)
    worst = 0.0
    for f in range(4):
        face = [k for k in range(4) if k != f]
        on = (np.abs(lam[:, f]) <= 1e-12) & np.all(lam[:, face] > 1e-12, axis=1)
        if on.sum() != 3:
            return np.inf
        for row in lam[on][:, face]:
            want = np.sort([0.5 + w2, (0.5 - w2) / 2, (0.5 - w2) / 2])
            worst = max(worst, float(np.abs(np.sort(row) - want).max()))
    return worst


def _symmetry_error(lam):
    """Max distance from each vertex-permuted node to its nearest node."""
    worst = 0.0
    for p in itertools.permutations(range(4)):
        d = np.abs(lam[:, p][:, None, :] - lam[None, :, :]).max(axis=2).min(axis=1).max()
        worst = max(worst, float(d))
    return worst


def _node_sets(N):
    e = O.element(N)
    r, s, t = nudg.xyz_to_rst(*nudg.nodes3d(N))
    return {"oracle": (e["r"], e["s"], e["t"]), "nudg": (r, s, t)}


@pytest.mark.parametrize("N", range(3, 7))
def test_nodes_edge_closed_form(N):
    """Edge nodes on all six edges equal the reduction of nodes3d.go above."""
    for name, (r, s, t) in _node_sets(N).items():
        assert _edge_error(_bary(r, s, t), N) <= 2e-15, name
    if N == 3:  # the textbook's edge nodes (the GL points) are not these
        r, s, t = _node_sets(3)["oracle"]
        edge = np.sort(r[(np.abs(s + 1) < 1e-12) & (np.abs(t + 1) < 1e-12)])
        assert np.allclose(edge, [-1.0, -0.283395, 0.283395, 1.0], atol=1e-6)


def test_nodes_face_closed_form_p4():
    """Face-interior nodes at P=4 (they carry the 0-based alpha lookup)."""
    for name, (r, s, t) in _node_sets(4).items():
        assert _face_p4_error(_bary(r, s, t)) <= 2e-15, name


@pytest.mark.parametrize("N", range(1, 7))
def test_nodes_symmetric_under_vertex_permutations(N):
    """The barycentric node set is invariant under all 24 vertex permutations;
    at P=3 the face-interior nodes sit at the face centroids and at P=4 the
    interior node at the tet centroid (the odd warp vanishes at 0)."""
    for name, (r, s, t) in _node_sets(N).items():
        lam = _bary(r, s, t)
        assert _symmetry_error(lam) <= 2e-15, name
        assert np.all(np.abs(lam.sum(axis=1) - 1) <= 1e-15)
        if N == 3:
            cen = np.all(np.abs(np.sort(lam, axis=1) - [0, 1 / 3, 1 / 3, 1 / 3]) <= 2e-15, axis=1)
            assert cen.sum() == 4, name
        if N == 4:
            assert np.any(np.all(np.abs(lam - 0.25) <= 2e-15, axis=1)), name


@pytest.mark.parametrize("N", range(3, 7))
def test_nodes_in_tet_and_on_faces(N):
    """nodes3d_test.go:68-130 (TestNodes3D: XYZtoRST of the nodes lies in the
    reference tet, IsValidRST tolerance 1e-10, rsttoabc.go:104-107), extended to
    P=3..6; and the warp is tangential: nodes of the equispaced set on a face
    (a zero barycentric) stay on it."""
    req, seq, teq = nudg.equi_nodes3d(N)
    lam_eq = _bary(req, seq, teq)
    for name, (r, s, t) in _node_sets(N).items():
        tol = 1e-10
        assert np.all((r >= -1 - tol) & (s >= -1 - tol) & (t >= -1 - tol) & (r + s + t <= -1 + tol)), name
        lam = _bary(r, s, t)
        zero = np.abs(lam_eq) <= 1e-15
        assert np.all(np.abs(lam[zero]) <= 1e-14), name


def test_evalwarp_zero_at_endpoints():
    """nodes3d_test.go:176-206 (TestEvalwarp) at P=3..6."""
    for P in range(3, 7):
        xeq = np.linspace(-1.0, 1.0, P + 1)
        w = nudg._evalwarp(P, -nudg.jacobi_gl(0.0, 0.0, P), xeq)
        assert abs(w[0]) <= 1e-14 and abs(w[P]) <= 1e-14


def _textbook_pairing(p, gauss_x, xs):
    xeq = -1.0 + 2.0 * np.arange(p + 1) / p
    c = np.polynomial.polynomial.polyfit(xeq, -gauss_x - xeq, p)  # ascending GL
    return np.polynomial.polynomial.polyval(xs, c) * (1.0 - xs * xs)


def _textbook_divide(p, gauss_x, xs):
    xeq = -1.0 + 2.0 * np.arange(p + 1) / p
    c = np.polynomial.polynomial.polyfit(xeq, gauss_x - xeq, p)
    inner = np.abs(xs) < 1 - 1e-10
    return np.where(inner, np.polynomial.polynomial.polyval(xs, c) / np.where(inner, 1 - xs * xs, 1), 0.0)


@pytest.mark.parametrize("mutant", ["pairing", "divide", "alpha_1based"])
def test_node_pins_reject_textbook_corrections(monkeypatch, mutant):
    """Planted on the CPU: each plausible 'correction' of nodes3d.go toward the
    textbook fails at least one of the pins above."""
    if mutant == "pairing":
        monkeypatch.setattr(nudg, "_evalwarp", _textbook_pairing)
    elif mutant == "divide":
        monkeypatch.setattr(nudg, "_evalwarp", _textbook_divide)
    else:
        monkeypatch.setattr(nudg, "_ALPHASTORE", (0.0,) + nudg._ALPHASTORE[:-1])
    failures = 0
    for N in (3, 4):
        lam = _bary(*nudg.xyz_to_rst(*nudg.nodes3d(N)))
        failures += _edge_error(lam, N) > 2e-15
        if N == 4:
            failures += _face_p4_error(lam) > 2e-15
    assert failures >= 1


@pytest.mark.parametrize("N", range(1, 7))
def test_operators_exact_on_linears(N):
    """vandermonde3d_test.go:115-159: Dr,Ds,Dt of a r + b s + c t + d; V Vinv = I."""
    e = O.element(N)
    r, s, t = e["r"], e["s"], e["t"]
    f = 0.7 * r - 1.3 * s + 2.1 * t + 0.4
    assert np.max(np.abs(e["Dr"] @ f - 0.7)) <= 1e-10
    assert np.max(np.abs(e["Ds"] @ f + 1.3)) <= 1e-10
    assert np.max(np.abs(e["Dt"] @ f - 2.1)) <= 1e-10
    assert np.max(np.abs(e["V"] @ e["Vinv"] - np.eye(e["Np"]))) <= 1e-10
    # exact on degree-N monomials: Dr (r^N) = N r^(N-1)
    g = r ** N
    assert np.max(np.abs(e["Dr"] @ g - N * r ** (N - 1))) <= 1e-9 * max(1, N ** 2)


@pytest.mark.parametrize("N", range(1, 7))
def test_mass_and_lift_invariants(N):
    """M integrates exactly: 1'M1 = |T| = 4/3, r'M1 = -2/3; M LIFT = E (face mass blocks,
    each face integrated in its own 2D reference coordinates of area 2, lift3d.go:251-308),
    so 1' M LIFT 1 = 4 faces x 2 = 8."""
    e = O.element(N)
    M_ = e["Mass"]
    one = np.ones(e["Np"])
    tol = 1e-14 * max(1.0, np.linalg.cond(e["V"]))  # M = Vinv' Vinv inherits cond(V)
    assert abs(one @ M_ @ one - 4 / 3) <= tol
    assert abs(e["r"] @ M_ @ one + 2 / 3) <= tol
    ML = M_ @ e["LIFT"]
    area = one @ ML @ np.ones(4 * e["Nfp"])
    assert abs(area - 8.0) <= 1e-9
    # M LIFT vanishes off the face nodes
    for f in range(4):
        off = np.setdiff1d(np.arange(e["Np"]), e["Fmask"][f])
        blk = ML[off][:, f * e["Nfp"]:(f + 1) * e["Nfp"]]
        assert np.max(np.abs(blk)) <= 1e-9 * np.abs(ML).max()


@pytest.mark.parametrize("N", range(1, 7))
def test_operator_restatements_agree(N):
    """Dr/Ds/Dt/LIFT/Mass of the oracle vs the numpy product setup, relative to the
    operator size times cond(V) (both invert V; the reference's nodes make V
    ill-conditioned from P=5)."""
    e = O.element(N)
    r = nudg.build_ref_tet(N)
    kappa = np.linalg.cond(r.V)
    for k in ("Dr", "Ds", "Dt", "LIFT", "Mass", "V", "Vinv"):
        a, b = e[k], getattr(r, k)
        assert np.max(np.abs(a - b)) <= 1e-15 * kappa * max(1.0, np.abs(b).max()) * 50, k
    assert np.array_equal(e["Fmask"], r.Fmask)


def test_single_tet_geometry_closed_form():
    """normals3d_test.go: unit right tet J = V/V_ref = 1/8, outward normals, sJ = area ratio."""
    VX, VY, VZ = np.array([0.0, 1, 0, 0]), np.array([0.0, 0, 1, 0]), np.array([0.0, 0, 0, 1])
    o = O.Setup(3, VX, VY, VZ, np.array([[0, 1, 2, 3]]))
    assert o.status == 0
    assert np.allclose(o.field("J"), 1 / 8, atol=1e-15)
    nx, ny, nz = (o.field(k).reshape(4, -1) for k in ("nx", "ny", "nz"))
    want = [(0, 0, -1), (0, -1, 0), (1 / math.sqrt(3),) * 3, (-1, 0, 0)]
    for f in range(4):
        assert np.allclose(nx[f], want[f][0], atol=1e-14)
        assert np.allclose(ny[f], want[f][1], atol=1e-14)
        assert np.allclose(nz[f], want[f][2], atol=1e-14)
    # sJ = physical face area / 2 (face reference triangles have area 2)
    sJ = o.field("sJ").reshape(4, -1)
    for f, area in enumerate((0.5, 0.5, math.sqrt(3) / 2, 0.5)):
        assert np.allclose(sJ[f], area / 2, atol=1e-14)
    # metric identity: [rx ry rz; sx sy sz; tx ty tz] = inverse of d(x,y,z)/d(r,s,t)
    Jm = 0.5 * np.array([[1.0, 0, 0], [0, 1, 0], [0, 0, 1]])  # columns x_r, x_s, x_t
    inv = np.linalg.inv(Jm)
    for (q, k) in enumerate(("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")):
        assert np.allclose(o.field(k), inv[q // 3, q % 3], atol=1e-13)


def test_metric_identities_on_jittered_mesh():
    """normals3d_test.go:106-196: rx*xr + sx*xs + tx*xt = 1 etc., per node."""
    N = 3
    m = M.cube_mesh(3, 3, 3, jitter=0.15, seed=9)
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    e = O.element(N)
    X, Y, Z = o.field("X"), o.field("Y"), o.field("Z")
    xr, xs, xt = X @ e["Dr"].T, X @ e["Ds"].T, X @ e["Dt"].T
    yr, ys, yt = Y @ e["Dr"].T, Y @ e["Ds"].T, Y @ e["Dt"].T
    rx, sx, tx = o.field("rx"), o.field("sx"), o.field("tx")
    ry, sy, ty = o.field("ry"), o.field("sy"), o.field("ty")
    assert np.max(np.abs(rx * xr + sx * xs + tx * xt - 1)) <= 1e-10
    assert np.max(np.abs(ry * xr + sy * xs + ty * xt)) <= 1e-10
    assert np.max(np.abs(rx * yr + sx * ys + tx * yt)) <= 1e-10
    # total volume from J and the mass matrix
    vol = (o.field("J") @ e["Mass"] @ np.ones(e["Np"])).sum()
    assert abs(vol - 1.0) <= 1e-12


def test_cube_six_tets_twelve_boundary_faces():
    """ticonnect3d_test.go:110-154: the 6-tet cube has 12 boundary faces; reciprocity."""
    m = M.cube_mesh(1, 1, 1)
    o = O.Setup(1, m.VX, m.VY, m.VZ, m.EToV)
    E, F = o.imap("EToE"), o.imap("EToF")
    bnd = sum(1 for k in range(6) for f in range(4) if E[k, f] == k)
    assert bnd == 12
    for k in range(6):
        for f in range(4):
            k2, f2 = E[k, f], F[k, f]
            assert E[k2, f2] == k and F[k2, f2] == f
    assert np.all(o.field("J") > 0)


def test_buildmaps_reciprocity_and_continuity():
    """buildmaps3d_test.go:138-165: vmapP(mapP) = vmapM, matched nodes coincide."""
    N = 3
    m = M.cube_mesh(2, 3, 2, jitter=0.1, seed=2)
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    vM, vP, mP = (o.imap(k).reshape(-1) for k in ("vmapM", "vmapP", "mapP"))
    assert np.array_equal(vM[mP], vP)
    X, Y, Z = (o.field(k).reshape(-1) for k in ("X", "Y", "Z"))
    d = np.sqrt((X[vM] - X[vP]) ** 2 + (Y[vM] - Y[vP]) ** 2 + (Z[vM] - Z[vP]) ** 2)
    assert d.max() <= 1e-12
    interior = mP != np.arange(mP.size)
    assert interior.sum() > 0 and np.array_equal(mP[mP], np.arange(mP.size))


def test_periodic_vmapP_matches_shifted_coordinates():
    """Periodic extension: the matched node sits one period away (new in this build)."""
    N = 2
    m = M.cube_mesh(3, 3, 3, periodic=True, jitter=0.1, seed=4)
    o = O.Setup.from_mesh(N, m)
    vM, vP = o.imap("vmapM").reshape(-1), o.imap("vmapP").reshape(-1)
    assert np.all(vM != vP)  # no boundary faces left
    X, Y, Z = (o.field(k).reshape(-1) for k in ("X", "Y", "Z"))
    for A in (X, Y, Z):
        d = A[vP] - A[vM]
        assert np.max(np.abs(d - np.round(d))) <= 1e-12


def test_physical_derivative_closed_form_neu_p2():
    """tetnudg/tet_nudg_mesh_physical_derivative_test.go:747-919 (host half)."""
    N = 2
    m = neu_mesh()
    assert m.K == 565
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    assert o.status == 0
    e = O.element(N)
    X, Y, Z = o.field("X"), o.field("Y"), o.field("Z")
    U = X ** N + Y ** N + Z ** N
    mets = [o.field(k) for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")]
    dx, dy, dz = O.phys_deriv(e["Dr"], e["Ds"], e["Dt"], U, mets)
    assert np.max(np.abs(dx - N * X ** (N - 1))) < 1e-8
    assert np.max(np.abs(dy - N * Y ** (N - 1))) < 1e-8
    assert np.max(np.abs(dz - N * Z ** (N - 1))) < 1e-8


def test_matmul_known_answer():
    """runner/runner_matrix_test.go:13-125: [2,1,0;0,1,1;1,0,2] times sequential columns."""
    A = np.array([[2.0, 1.0, 0.0], [0.0, 1.0, 1.0], [1.0, 0.0, 2.0]])
    IN = np.arange(1.0, 13.0).reshape(4, 3)
    out = O.matmul(A, IN)
    assert np.array_equal(out, IN @ A.T)
    assert np.array_equal(O.matmul(A, IN, add_to=out), 2 * out)


def test_aligned_offsets_reference_cases():
    """runner/runner_test.go:474-535 (K=[10,15,20], one value per element, no alignment)
    and aligned cases; two implementations (oracle, product) agree."""
    offs, total = O.aligned_offsets([10, 15, 20], 45 * 8, 2, 1)
    assert list(offs) == [0, 10, 25, 45] and total == 360
    offs2, total2 = O.aligned_offsets([3, 5, 7], 15 * 20 * 8, 2, 128)
    mine, mbytes = dg.aligned_offsets(np.array([3, 5, 7]), 20, 8, 128)
    assert np.array_equal(offs2, mine) and total2 == mbytes
    assert all((o * 8) % 128 == 0 for o in offs2)
    offs3, _ = O.aligned_offsets([3, 5], 8 * 10 * 4, 1, 64)  # Float32
    mine3, _ = dg.aligned_offsets(np.array([3, 5]), 10, 4, 64)
    assert np.array_equal(offs3, mine3)


def test_face_connector_invariants_and_element_id_exchange():
    """utils/face_connector_test.go:179-397: exchanging element ids delivers, at every
    face point, the id of the neighbour element; Verify invariants hold."""
    N = 1
    m = neu_mesh()
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    e = O.element(N)
    vP = o.imap("vmapP").reshape(-1)
    NP, counts, pick, place = O.face_connector(m.K, 4, e["Nfp"], e["Np"], vP, m.EToP)
    assert NP == 5 and counts[0].sum() == 0 and counts[:, 0].sum() == 0  # empty p0
    fc = M.new_face_connector(m.K, 4, e["Nfp"], e["Np"], vP, m.EToP)
    fc.verify()
    for k in pick:
        assert np.array_equal(pick[k], fc.pick[k])
        assert np.array_equal(place[(k[1], k[0])], fc.place[(k[1], k[0])])
    ids = [np.repeat(fc.local_to_global[p], e["Np"]).astype(float) for p in range(NP)]
    P = [np.full(int(k) * 4 * e["Nfp"], -1.0) for k in fc.elems_per_partition]
    for (s, d), pk in fc.pick.items():
        P[d][fc.place[(d, s)]] = ids[s][pk]
    E = o.imap("EToE")
    for p in range(NP):
        for le, g in enumerate(fc.local_to_global[p]):
            got = P[p][le * 4 * e["Nfp"]:(le + 1) * 4 * e["Nfp"]].reshape(4, e["Nfp"])
            assert np.all(got == E[g][:, None])


def test_face_connector_rejects_bad_input():
    with pytest.raises(ValueError):
        M.new_face_connector(0, 4, 3, 4, np.zeros(0), np.zeros(0))
    with pytest.raises(ValueError):
        M.new_face_connector(2, 4, 3, 4, np.zeros(5), np.zeros(2))
    with pytest.raises(ValueError):
        O.face_connector(0, 4, 3, 4, np.zeros(0), np.zeros(0))


def _burgers_setup(n, N):
    m = M.cube_mesh(n, n, n, periodic=True)
    return m, O.Setup.from_mesh(N, m), dg.make_setup(m, N)


def test_burgers_constant_state_and_conservation():
    m, o, st = _burgers_setup(4, 3)
    e = O.element(3)
    assert np.max(np.abs(o.burgers_rhs(np.full((m.K, e["Np"]), 0.7)))) <= 1e-13
    x, y, z = st.node_coords()
    u = 0.5 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
    rhs = o.burgers_rhs(u)
    J = o.field("J")
    total = ((rhs @ e["Mass"].T) * J).sum()
    assert abs(total) <= 1e-13 * (np.abs(rhs @ e["Mass"].T) * J).sum()


def test_partition_schedule_equals_element_schedule():
    """The reference's OCCA-OpenMP schedule (parallel over partitions, serial
    element loop) computes the same operator as the element-parallel one, in
    the same summation order: results are bit-identical, also over SSP-RK3."""
    m, o, st = _burgers_setup(4, 2)
    x, y, z = st.node_coords()
    u = 0.5 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
    Kp = M.slab_partition(m.K, 5, 0.4, seed=3)
    assert np.array_equal(o.burgers_rhs_partitions(u, Kp), o.burgers_rhs(u))
    hmin = o.min_node_distance()
    a, da = o.ssprk3(u, 3, 0.3, hmin)
    b, db = o.ssprk3_partitions(u, 3, 0.3, hmin, Kp)
    assert np.array_equal(a, b) and np.array_equal(da, db)


def _exact_burgers(x, y, z, t, u0):
    """Characteristic solution u = u0(x - u t (1,1,1)) (before shocks), by Newton."""
    u = u0(x, y, z)
    for _ in range(60):
        s = u * t
        f = u - u0(x - s, y - s, z - s)
        h = 1e-7
        df = 1 - (u0(x - s - h * t, y - s - h * t, z - s - h * t) - u0(x - s, y - s, z - s)) / h
        u = u - f / df
    return u


def test_burgers_spatial_convergence():
    """Smooth data, short time: the nodal error after T is T times the local
    truncation error of the nodal derivative, O(h^P); P=3 must converge at
    about third order (a dropped or sign-flipped flux/LIFT term gives O(1))."""
    N = 3
    u0 = lambda x, y, z: 0.5 + 0.1 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)  # noqa
    errs = []
    T = 0.02
    for n in (4, 8):
        m, o, st = _burgers_setup(n, N)
        x, y, z = st.node_coords()
        steps = 10 * n
        u, _ = o.ssprk3(u0(x, y, z), steps, fixed_dt=T / steps)
        errs.append(rel_l2(u, _exact_burgers(x, y, z, T, u0)))
    rate = math.log2(errs[0] / errs[1])
    assert rate > N - 0.5, (errs, rate)


def test_ssprk3_third_order_in_time():
    N = 2
    m, o, st = _burgers_setup(3, N)
    x, y, z = st.node_coords()
    u0 = 0.5 + 0.2 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    T = 0.05
    sols = [o.ssprk3(u0, n, fixed_dt=T / n)[0] for n in (5, 10, 20, 160)]
    e1, e2, e3 = (rel_l2(s, sols[-1]) for s in sols[:3])
    assert 2.6 < math.log2(e1 / e2) < 3.4 and 2.6 < math.log2(e2 / e3) < 3.4, (e1, e2, e3)
