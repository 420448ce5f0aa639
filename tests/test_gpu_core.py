"""GPU parity of the hot path against the CPU oracle (calls through the C ABI)."""
import ctypes as C

import numpy as np
import pytest

from dgkernel_amd import _lib, dg
from dgkernel_amd import mesh as M
from dgkernel_amd.nudg import build_ref_tet
from oracle import oracle as O
from tests.helpers import neu_mesh, pooled, rel_l2, smooth_u0

pytestmark = pytest.mark.gpu

METRICS = ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")


def _pool(parts, Kp, per_elem, dtype):
    """Runner pooled layout (WarpAlign) of per-partition arrays of one dtype."""
    vs = np.dtype(dtype).itemsize
    offs, nbytes = dg.aligned_offsets(np.asarray(Kp), per_elem, vs)
    out = np.zeros(nbytes // vs, dtype=dtype)
    for p, v in enumerate(parts):
        out[offs[p]:offs[p] + v.size] = v.reshape(-1)
    return out, offs


def _deriv_gpu(device, N, Kp, U_parts, met_parts, mtype=_lib.DGK_F64, otype=_lib.DGK_F64, el=None):
    """Run dgk_phys_deriv_nodal on per-partition [K_p, Np] arrays."""
    el = el or O.element(N)
    Np = el["Np"]
    mdt = np.float64 if mtype == _lib.DGK_F64 else np.float32
    odt = np.float64 if otype == _lib.DGK_F64 else np.float32
    u_pool, uoff = _pool(U_parts, Kp, Np, np.float64)
    du, duo = device.upload(u_pool), device.upload(uoff)
    mbufs, moff = [], None
    for q in range(9):
        mp, moff = _pool([m[q] for m in met_parts], Kp, Np, mdt)
        mbufs.append(device.upload(mp))
    dmo = device.upload(moff)
    o_pool, ooff = _pool([np.zeros((k, Np)) for k in Kp], Kp, Np, odt)
    outs = [device.upload(o_pool) for _ in range(3)]
    doo = device.upload(ooff)
    Dm = [device.upload(np.ascontiguousarray(el[k])) for k in ("Dr", "Ds", "Dt")]
    Karr = np.asarray(Kp, dtype=np.int64)
    mptrs = (C.c_void_p * 9)(*[b.ptr.value for b in mbufs])
    _lib.call("dgk_phys_deriv_nodal", device.h, N, len(Kp), Karr.ctypes.data_as(C.POINTER(C.c_int64)),
              Dm[0].ptr, Dm[1].ptr, Dm[2].ptr, du.ptr, duo.ptr, mptrs, mtype, dmo.ptr,
              outs[0].ptr, outs[1].ptr, outs[2].ptr, otype, doo.ptr)
    device.sync()
    res = []
    for o in outs:
        h = o.to_host()
        res.append([h[ooff[p]:ooff[p] + Kp[p] * Np].reshape(Kp[p], Np) for p in range(len(Kp))])
    return res


def test_physical_derivative_neu_p2_closed_form(device):
    """tetnudg/tet_nudg_mesh_physical_derivative_test.go:747-919 on the GPU."""
    N = 2
    m = neu_mesh()
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    assert o.status == 0
    el = O.element(N)
    X, Y, Z = o.field("X"), o.field("Y"), o.field("Z")
    U = X ** N + Y ** N + Z ** N
    mets = [o.field(k) for k in METRICS]
    host = O.phys_deriv(el["Dr"], el["Ds"], el["Dt"], U, mets)
    out = _deriv_gpu(device, N, [m.K], [U], [mets], el=el)
    exact = [N * X ** (N - 1), N * Y ** (N - 1), N * Z ** (N - 1)]
    for d in range(3):
        g = out[d][0]
        assert np.max(np.abs(g - exact[d])) < 1e-8
        assert np.max(np.abs(g - host[d])) <= 1e-13 * np.max(np.abs(host[d]))


@pytest.mark.parametrize("N,parts", [(4, [700, 1100, 813]), (3, [1500, 1113]), (1, [2613]),
                                     (5, [900, 1713]), (6, [2613])])
def test_physical_derivative_partitioned_fp64(device, N, parts):
    m = M.cube_mesh(8, 8, 7, jitter=0.1, seed=3)  # 2688 tets, jittered, non-periodic
    K = sum(parts)
    m.EToV, m.EToVc = m.EToV[:K], m.EToVc[:K]
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    assert o.status == 0
    el = O.element(N)
    U = M.uniform(11, 3, K * el["Np"], -1.0, 1.0).reshape(K, el["Np"])
    mets = [o.field(k) for k in METRICS]
    host = O.phys_deriv(el["Dr"], el["Ds"], el["Dt"], U, mets)
    starts = np.concatenate(([0], np.cumsum(parts)))
    Up = [U[starts[p]:starts[p + 1]] for p in range(len(parts))]
    Mp = [[mm[starts[p]:starts[p + 1]] for mm in mets] for p in range(len(parts))]
    out = _deriv_gpu(device, N, parts, Up, Mp, el=el)
    for d in range(3):
        g = np.concatenate(out[d])
        scale = np.max(np.abs(host[d]))
        assert np.max(np.abs(g - host[d])) <= 1e-13 * scale


def test_physical_derivative_fp32_metrics(device):
    """FP32 metrics and outputs (the reference's CalculatePhysicalDerivative32)."""
    N = 4
    m = M.cube_mesh(6, 6, 6, jitter=0.1, seed=4)
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    el = O.element(N)
    U = M.uniform(12, 3, m.K * el["Np"], -1.0, 1.0).reshape(m.K, el["Np"])
    mets = [o.field(k) for k in METRICS]
    mets32 = [mm.astype(np.float32).astype(np.float64) for mm in mets]
    host = O.phys_deriv(el["Dr"], el["Ds"], el["Dt"], U, mets32)
    out = _deriv_gpu(device, N, [m.K], [U], [mets], mtype=_lib.DGK_F32, otype=_lib.DGK_F32, el=el)
    for d in range(3):
        g = out[d][0].astype(np.float64)
        # FP64 arithmetic on FP32 metrics, rounded once to FP32 on store.
        assert np.max(np.abs(g - host[d]) / (np.abs(host[d]) + 1e-3 * np.abs(host[d]).max())) < 2e-7


def test_matmul_known_answer_and_random(device):
    """runner/runner_matrix_test.go:13-125 known answer, plus MATMUL_ADD."""
    A = np.array([[2.0, 1.0, 0.0], [0.0, 1.0, 1.0], [1.0, 0.0, 2.0]])
    Kp = [3, 5]
    IN = np.arange(1.0, 25.0).reshape(8, 3)
    want = O.matmul(A, IN)
    np.testing.assert_allclose(want[0], A @ IN[0])
    inp, ioff = pooled([IN[:3], IN[3:]], Kp, 3)
    outp, ooff = pooled([np.zeros((3, 3)), np.zeros((5, 3))], Kp, 3)
    dA = device.upload(np.ascontiguousarray(A.T))  # column-major device matrix
    din, dout = device.upload(inp), device.upload(outp)
    dio, doo = device.upload(ioff), device.upload(ooff)
    Karr = np.array(Kp, dtype=np.int64)
    _lib.call("dgk_matmul", device.h, 3, 3, dA.ptr, 2, Karr.ctypes.data_as(C.POINTER(C.c_int64)),
              din.ptr, dio.ptr, dout.ptr, doo.ptr, 0)
    _lib.call("dgk_matmul", device.h, 3, 3, dA.ptr, 2, Karr.ctypes.data_as(C.POINTER(C.c_int64)),
              din.ptr, dio.ptr, dout.ptr, doo.ptr, 1)
    h = dout.to_host()
    got = np.concatenate([h[ooff[0]:ooff[0] + 9].reshape(3, 3), h[ooff[1]:ooff[1] + 15].reshape(5, 3)])
    np.testing.assert_array_equal(got, O.matmul(A, IN, add_to=want))


def _exchange_gpu(device, fc, u_parts, Np, Nfp):
    """dgk_face_exchange of per-partition u into per-partition P buffers."""
    NP = fc.num_partitions
    Kp = fc.elems_per_partition
    u_pool, uoff = pooled(u_parts, Kp, Np)
    P_pool, Poff = pooled([np.full(int(k) * 4 * Nfp, np.nan) for k in Kp], Kp, 4 * Nfp)
    counts = np.array([[fc.pick[(s, d)].size for d in range(NP)] for s in range(NP)], dtype=np.int64)
    pick = np.concatenate([fc.pick[(s, d)] for s in range(NP) for d in range(NP)]).astype(np.int32)
    place = np.concatenate([fc.place[(d, s)] for d in range(NP) for s in range(NP)]).astype(np.int32)
    plan = C.c_void_p()
    _lib.call("dgk_face_plan_create", device.h, NP, counts.ctypes.data_as(C.POINTER(C.c_int64)),
              pick.ctypes.data_as(C.POINTER(C.c_int32)), place.ctypes.data_as(C.POINTER(C.c_int32)),
              C.byref(plan))
    du, dP = device.upload(u_pool), device.upload(P_pool)
    duo, dPo = device.upload(uoff), device.upload(Poff)
    _lib.call("dgk_face_exchange", plan, du.ptr, duo.ptr, dP.ptr, dPo.ptr)
    device.sync()
    h = dP.to_host()
    _lib.call("dgk_face_plan_destroy", plan)
    return [h[Poff[p]:Poff[p] + int(Kp[p]) * 4 * Nfp] for p in range(NP)]


def _simulate_exchange(fc, u_parts, Nfp):
    """utils/face_connector_test.go:104-156 semantics on the host."""
    out = [np.full(int(k) * 4 * Nfp, np.nan) for k in fc.elems_per_partition]
    for (s, d), pk in fc.pick.items():
        out[d][fc.place[(d, s)]] = u_parts[s].reshape(-1)[pk]
    return out


def test_face_exchange_partial_lists_keep_untouched_slots(device):
    """Lists that leave a P slot unfilled (one entry dropped from one pair)
    take the pair-ordered path; the filled slots match the host simulation and
    the dropped slot keeps its old content."""
    import copy
    m = M.cube_mesh(3, 3, 3, periodic=True)
    st = dg.make_setup(m, 1, Kp=M.slab_partition(m.K, 3))
    _, vp = st.vmaps()
    fc = M.new_face_connector(m.K, 4, st.ref.Nfp, st.ref.Np, vp, M.etop_from_counts(st.Kp))
    part = copy.copy(fc)
    part.pick, part.place = dict(fc.pick), dict(fc.place)
    part.pick[(0, 1)] = fc.pick[(0, 1)][:-1]
    part.place[(1, 0)] = fc.place[(1, 0)][:-1]
    u = M.uniform(8, 3, m.K * st.ref.Np, -1, 1).reshape(m.K, st.ref.Np)
    u_parts = [u[fc.local_to_global[q]] for q in range(3)]
    got = _exchange_gpu(device, part, u_parts, st.ref.Np, st.ref.Nfp)
    want = _simulate_exchange(part, u_parts, st.ref.Nfp)
    for q in range(3):
        assert np.array_equal(got[q], want[q], equal_nan=True)
    assert np.isnan(got[1][fc.place[(1, 0)][-1]])
    full = _exchange_gpu(device, fc, u_parts, st.ref.Np, st.ref.Nfp)
    assert not any(np.isnan(f).any() for f in full)


def test_face_exchange_neu_five_partitions_bit_exact(device):
    """cube-partitioned.neu: groups p1..p4 give 5 partitions, p0 empty (face_connector.go:59-64)."""
    N = 2
    m = neu_mesh()
    o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
    el = O.element(N)
    vP = o.imap("vmapP").reshape(-1)
    fc = M.new_face_connector(m.K, 4, el["Nfp"], el["Np"], vP, m.EToP)
    assert fc.num_partitions == 5 and fc.elems_per_partition[0] == 0
    fc.verify()
    NPo, counts, pk, pl = O.face_connector(m.K, 4, el["Nfp"], el["Np"], vP, m.EToP)
    assert all(np.array_equal(pk[k], fc.pick[k]) for k in pk)
    u = M.uniform(5, 9, m.K * el["Np"], -1, 1).reshape(m.K, el["Np"])
    u_parts = [u[fc.local_to_global[p]] for p in range(fc.num_partitions)]
    got = _exchange_gpu(device, fc, u_parts, el["Np"], el["Nfp"])
    want = _simulate_exchange(fc, u_parts, el["Nfp"])
    for p in range(fc.num_partitions):
        assert np.array_equal(got[p], want[p])  # bit-exact copy
        # and the value placed is u at VmapP (element-ID style check)
    flat_P = np.zeros((m.K, 4 * el["Nfp"]))
    for p in range(fc.num_partitions):
        flat_P[fc.local_to_global[p]] = got[p].reshape(-1, 4 * el["Nfp"])
    assert np.array_equal(flat_P.reshape(-1), u.reshape(-1)[vP])


@pytest.mark.parametrize("N,n,per,parts", [(2, 6, True, [300, 500, 496]), (3, 5, True, [750]),
                                           (3, 4, False, [100, 284])])
def test_trace_plus_matches_face_exchange(device, N, n, per, parts):
    """The fused stage's neighbour gather equals the FaceConnector exchange, bit for bit."""
    m = M.cube_mesh(n, n, n, periodic=per)
    st = dg.make_setup(m, N, Kp=parts)
    mesh = dg.DGMesh(device, st)
    ref = st.ref
    u = M.uniform(6, 1, m.K * ref.Np, -1, 1).reshape(m.K, ref.Np)
    du = device.upload(st.to_pooled(u))
    dP = device.zeros(m.K * 4 * ref.Nfp)
    mesh.trace_plus(du, dP)
    device.sync()
    got = dP.to_host()
    vm, vp = st.vmaps()
    EToP = M.etop_from_counts(st.Kp)
    fc = M.new_face_connector(m.K, 4, ref.Nfp, ref.Np, vp, EToP)
    u_parts = [u[fc.local_to_global[p]] for p in range(fc.num_partitions)]
    ex = _simulate_exchange(fc, u_parts, ref.Nfp)
    assert np.array_equal(got, np.concatenate(ex))
    # and the oracle's coordinate-matched VmapP agrees
    o = O.Setup.from_mesh(N, m)
    assert np.array_equal(o.imap("vmapP").reshape(-1), vp)


def test_geometry_and_hmin_vs_oracle(device):
    N = 3
    m = M.cube_mesh(5, 4, 3, periodic=True, jitter=0.1, seed=8)
    st = dg.make_setup(m, N, Kp=[100, 260])
    mesh = dg.DGMesh(device, st)
    geo = mesh.geometry()
    o = O.Setup.from_mesh(N, m)
    for q, k in enumerate(METRICS):
        ref = o.field(k)
        assert np.max(np.abs(geo[q][:, None] - ref)) <= 1e-12 * np.max(np.abs(ref))
    J = o.field("J")
    assert np.max(np.abs(geo[9][:, None] - J)) <= 1e-12 * J.max()
    Fs = o.field("Fscale").reshape(m.K, 4, -1)
    for f in range(4):
        assert np.max(np.abs(geo[13 + 4 * f][:, None] - Fs[:, f, :])) <= 1e-12 * Fs.max()
        for d, nm in enumerate(("nx", "ny", "nz")):
            nn = o.field(nm).reshape(m.K, 4, -1)[:, f, :]
            assert np.max(np.abs(geo[10 + 4 * f + d][:, None] - nn)) <= 1e-13
    assert abs(mesh.min_node_distance() - o.min_node_distance()) <= 1e-15


@pytest.mark.parametrize("variant", [1, 2, 4, 5, 6])
@pytest.mark.parametrize("N,n,parts", [(2, 8, [768] * 4), (3, 5, [200, 350, 200]), (1, 6, [1296]),
                                       (4, 3, [100, 62]), (3, 4, [1, 63, 64, 65, 191])])
def test_burgers_rhs_vs_oracle(device, N, n, parts, variant):
    """Stage kernels against the oracle's L(u): VALU (1), pipelined FP64 MFMA
    (2), FP64 MFMA without the load pipeline (4), deep pipeline (5, N <= 3)."""
    if variant in (5, 6) and N > 3:
        pytest.skip("the deep-pipeline and register-operator kernels are built for N <= 3")
    m = M.cube_mesh(n, n, n, periodic=True)
    st = dg.make_setup(m, N, Kp=parts)
    u = smooth_u0(st, noise=1e-3)
    o = O.Setup.from_mesh(N, m)
    want = o.burgers_rhs(u)
    solver = dg.BurgersSolver(device, st)
    solver.mesh.set_variant(variant)
    got = solver.rhs(u)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("variant", [1, 2, 4, 5, 6])
@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_burgers_rhs_boundary_faces_vs_oracle(device, N, variant):
    """Non-periodic jittered cube: a boundary face is its own neighbour, so its
    LF jump is zero (VmapP = VmapM there, BuildMaps3D)."""
    if variant in (5, 6) and N > 3:
        pytest.skip("the deep-pipeline and register-operator kernels are built for N <= 3")
    m = M.cube_mesh(4, 4, 4, periodic=False, jitter=0.05, seed=40 + N)
    st = dg.make_setup(m, N, Kp=[150, 1, 233])
    o = O.Setup.from_mesh(N, m)
    assert int(np.sum(o.imap("EToE") == np.arange(m.K)[:, None])) == 6 * 2 * 4 * 4
    u = smooth_u0(st, amp=0.8, noise=1e-2)
    want = o.burgers_rhs(u)
    solver = dg.BurgersSolver(device, st)
    solver.mesh.set_variant(variant)
    got = solver.rhs(u)
    spread = max(float(np.max(np.abs(mm - mm[:, :1])) / np.max(np.abs(mm)))
                 for mm in (o.field(k) for k in METRICS))
    assert np.max(np.abs(got - want)) <= (1e-12 + 4.0 * spread) * np.max(np.abs(want))


@pytest.mark.parametrize("N", [3, 5])
def test_burgers_rhs_rcb_partitions_vs_oracle(device, N):
    """Partitions from recursive coordinate bisection (uneven counts), elements
    renumbered partition-major: every stage form against the oracle on the
    same renumbered mesh."""
    m = M.cube_mesh(5, 5, 5, periodic=True, jitter=0.05, seed=12)
    kp = M.slab_partition(m.K, 6, 0.4, seed=3)
    mo, kp = M.order_by_partition(m, M.rcb_partition(m, kp))
    st = dg.make_setup(mo, N, Kp=kp)
    u = smooth_u0(st, noise=1e-3)
    o = O.Setup.from_mesh(N, mo)
    want = o.burgers_rhs(u)
    spread = max(float(np.max(np.abs(mm - mm[:, :1])) / np.max(np.abs(mm)))
                 for mm in (o.field(k) for k in METRICS))
    for variant in ((1, 2, 5) if N <= 3 else (0,)):
        solver = dg.BurgersSolver(device, st)
        if variant:
            solver.mesh.set_variant(variant)
        got = solver.rhs(u)
        assert np.max(np.abs(got - want)) <= (1e-12 + 4.0 * spread) * np.max(np.abs(want)), variant


@pytest.mark.parametrize("N", [2, 3])
def test_burgers_kway_partitions_vs_oracle(device, N):
    """Partitions from the multilevel k-way graph partitioner (7 parts, volume
    objective), elements renumbered partition-major: the stage against the
    oracle's L(u), and 10 SSP-RK3 steps against the oracle's."""
    from dgkernel_amd import partitioner as P
    m = M.cube_mesh(6, 6, 6, periodic=True, jitter=0.05, seed=4)
    mo, kp = M.order_by_partition(m, P.kway_partition(m, 7, "vol", seed=2))
    assert kp.size == 7 and kp.min() > 0
    st = dg.make_setup(mo, N, Kp=kp)
    u = smooth_u0(st, noise=1e-3)
    o = O.Setup.from_mesh(N, mo)
    want = o.burgers_rhs(u)
    spread = max(float(np.max(np.abs(mm - mm[:, :1])) / np.max(np.abs(mm)))
                 for mm in (o.field(k) for k in METRICS))
    solver = dg.BurgersSolver(device, st)
    got = solver.rhs(u)
    assert np.max(np.abs(got - want)) <= (1e-12 + 4.0 * spread) * np.max(np.abs(want))
    hmin = o.min_node_distance()
    want, _ = o.ssprk3(u, 10, 0.3, hmin)
    solver = dg.BurgersSolver(device, st, cfl=0.3, hmin=hmin)
    solver.set_state(u)
    solver.step(10)
    assert rel_l2(solver.state(), want) <= 1e-11


@pytest.mark.parametrize("variant", [1, 2, 5, 6])
def test_config1_burgers_100_steps_vs_oracle(device, variant):
    """Config 1: 8^3 hexes x 6, P=2, 4 slabs, LF + SSP-RK3, CFL 0.3, 100 steps."""
    m = M.cube_mesh(8, 8, 8, periodic=True)
    st = dg.make_setup(m, 2, Kp=M.slab_partition(m.K, 4))
    u0 = smooth_u0(st, amp=0.3, noise=1e-3, seed=1)
    o = O.Setup.from_mesh(2, m)
    hmin = o.min_node_distance()
    want, dts = o.ssprk3(u0, 100, 0.3, hmin)
    solver = dg.BurgersSolver(device, st, cfl=0.3, hmin=hmin)
    solver.mesh.set_variant(variant)
    solver.set_state(u0)
    solver.step(100)
    got = solver.state()
    assert rel_l2(got, want) <= 1e-12
    assert np.max(np.abs(got - want)) <= 1e-11
    # dt follows max|u|, which agrees to the state's round-off
    assert abs(solver.dt() - dts[-1]) <= 1e-12 * dts[-1]


def test_graph_replay_matches_eager(device):
    m = M.cube_mesh(6, 6, 6, periodic=True)
    st = dg.make_setup(m, 3, Kp=[400, 896])
    u0 = smooth_u0(st)
    a = dg.BurgersSolver(device, st)
    a.set_state(u0)
    b = dg.BurgersSolver(device, st, hmin=a.hmin)
    b.set_state(u0)
    for rep in range(3):
        a.step(2)
        b.step(2, graph=True)
        device.sync()
        assert np.array_equal(a.state(), b.state()), f"replay {rep}"


@pytest.mark.parametrize("variant,tb,bpc,chunk", [(1, 64, 0, 1), (1, 128, 0, 1), (1, 256, 0, 0),
                                                  (1, 64, 3, 1), (2, 64, 1, 1), (2, 64, 0, 0),
                                                  (2, 16, 5, 1), (2, 32, 0, 0), (2, 16, 0, 0),
                                                  (5, 64, 1, 1), (5, 64, 0, 0), (5, 32, 0, 1),
                                                  (5, 32, 3, 0), (4, 64, 0, 0), (4, 16, 2, 1),
                                                  (6, 64, 0, 1), (6, 32, 0, 0), (6, 64, 1, 0)])
def test_tile_shapes_agree(device, variant, tb, bpc, chunk):
    """Tile shape and grid size do not change the result of a kernel variant."""
    m = M.cube_mesh(6, 5, 4, periodic=True)
    st = dg.make_setup(m, 3, Kp=[333, 387])
    u0 = smooth_u0(st)
    ref = dg.BurgersSolver(device, st)
    ref.mesh.set_variant(variant)
    ref.set_state(u0)
    ref.step(3)
    s = dg.BurgersSolver(device, st, hmin=ref.hmin)
    s.mesh.set_variant(variant)
    s.mesh.set_tile(tb, bpc)
    s.mesh.set_order(bool(chunk))
    s.set_state(u0)
    s.step(3)
    assert np.array_equal(ref.state(), s.state())


@pytest.mark.parametrize("N,variant,prec", [(3, 1, 0), (3, 2, 0), (3, 5, 0), (3, 6, 0), (4, 2, 0),
                                            (5, 3, 0), (5, 3, 2), (5, 3, 1 | 16)])
def test_spatial_tile_order_agrees(device, N, variant, prec):
    """Walking the tiles along the Morton curve of their centroids, across
    partitions, gives the same state bit for bit as the partition order."""
    m = M.cube_mesh(7, 6, 5, periodic=True)
    st = dg.make_setup(m, N, Kp=M.slab_partition(m.K, 4, 0.4, seed=N))
    u0 = smooth_u0(st)
    out = []
    for spatial in (False, True):
        s = dg.BurgersSolver(device, st)
        s.mesh.set_variant(variant)
        s.mesh.set_precision(prec)
        s.mesh.set_order(True)
        s.mesh.set_tile_order(spatial)
        s.set_state(u0)
        s.step(3)
        out.append(s.state())
    assert np.array_equal(out[0], out[1])


def test_mfma_and_valu_stage_agree(device):
    m = M.cube_mesh(6, 6, 6, periodic=True)
    st = dg.make_setup(m, 3, Kp=M.slab_partition(m.K, 5, 0.4, seed=3))
    u0 = smooth_u0(st)
    out = []
    for v in (1, 2, 4, 5, 6):
        s = dg.BurgersSolver(device, st)
        s.mesh.set_variant(v)
        s.set_state(u0)
        s.step(10)
        out.append(s.state())
    for o in out[1:]:
        assert rel_l2(o, out[0]) <= 1e-14


def test_face_exchange_c2_size_four_slabs_bit_exact(device):
    """2^20 jittered tets, P=4, 4 slabs: the device exchange equals the
    host simulation value for value and places u at VmapP."""
    m = M.cube_mesh(64, 64, 43, jitter=0.1, seed=5, kmax=2 ** 20)
    st = dg.make_setup(m, 4, Kp=[m.K])
    Np, Nfp = st.ref.Np, st.ref.Nfp
    _, vp = st.vmaps()
    etop = M.etop_from_counts(M.slab_partition(m.K, 4, 0.0, seed=0))
    fc = M.new_face_connector(m.K, 4, Nfp, Np, vp, etop)
    u = M.uniform(6, 9, m.K * Np, -1, 1).reshape(m.K, Np)
    u_parts = [u[fc.local_to_global[p]] for p in range(4)]
    got = _exchange_gpu(device, fc, u_parts, Np, Nfp)
    want = _simulate_exchange(fc, u_parts, Nfp)
    flat = np.empty((m.K, 4 * Nfp))
    for p in range(4):
        assert np.array_equal(got[p], want[p])
        flat[fc.local_to_global[p]] = got[p].reshape(-1, 4 * Nfp)
    assert np.array_equal(flat.reshape(-1), u.reshape(-1)[vp])
