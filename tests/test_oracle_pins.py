"""Pins of the CPU oracle against the reference's own known-answer tests,
closed forms and invariants (no GPU). Each test names the reference test or
the mathematical fact it restates."""
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
