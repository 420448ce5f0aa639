"""Host-side setup and the C ABI surface, without a GPU."""
import re
from pathlib import Path

import numpy as np
import pytest

from dgkernel_amd import _lib, dg
from dgkernel_amd import mesh as M
from dgkernel_amd import nudg
from oracle import oracle as O
from tests.helpers import neu_mesh

ROOT = Path(__file__).resolve().parent.parent


def _declared_symbols():
    text = "\n".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dgk_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) > 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.dgk_version().decode().startswith("libdgk")


def test_device_create_rejects_cpu_modes():
    import ctypes as C
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.dgk_device_create(b'{"mode": "OpenMP"}', C.byref(h))
    assert rc == 4 and "not available" in lib.dgk_last_error().decode()


def test_splitmix_uniform_deterministic():
    a = M.uniform(7, 3, 1000, -1, 1)
    b = M.uniform(7, 3, 1000, -1, 1)
    c = M.uniform(7, 3, 500, -1, 1, start=500)
    assert np.array_equal(a, b) and np.array_equal(a[500:], c)
    assert -1 <= a.min() and a.max() < 1 and abs(a.mean()) < 0.1
    assert not np.array_equal(a, M.uniform(8, 3, 1000, -1, 1))


@pytest.mark.parametrize("n,per,jit", [(3, False, 0.0), (4, True, 0.0), (3, True, 0.1), (4, False, 0.1)])
def test_cube_mesh_positive_volume_and_conforming(n, per, jit):
    m = M.cube_mesh(n, n, n, periodic=per, jitter=jit, seed=3)
    v = np.stack([m.VX[m.EToV], m.VY[m.EToV], m.VZ[m.EToV]], axis=2)
    det = np.linalg.det(np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2))
    assert np.all(det > 0)
    assert abs(det.sum() / 6 - 1.0) < 1e-12
    E, F = M.connect(m)
    nb = (E == np.arange(m.K)[:, None]).sum()
    assert nb == (0 if per else 12 * n * n)
    k = np.arange(m.K)[:, None]
    assert np.all(E[E, F] == k) and np.all(F[E, F] == np.arange(4)[None, :])


def test_kmax_truncation_keeps_generation_order():
    full = M.cube_mesh(4, 4, 3)
    part = M.cube_mesh(4, 4, 3, kmax=100)
    assert np.array_equal(part.EToV, full.EToV[:100])


def test_slab_partition_counts():
    kp = M.slab_partition(5308416, 64, 0.4, seed=2024)
    assert kp.sum() == 5308416 and kp.size == 64
    assert kp.min() > 0.6 * 5308416 / 64 * 0.9 and kp.max() < 1.4 * 5308416 / 64 * 1.1
    assert np.array_equal(M.slab_partition(10, 3), [4, 3, 3])


@pytest.mark.parametrize("N", [1, 2, 3, 4, 6])
def test_vmapP_orientation_codes_equal_coordinate_match(N):
    """VmapP from orientation codes equals buildmaps3d.go's coordinate match, bit for bit."""
    for (n, per, jit) in [(3, True, 0.0), (2, False, 0.1), (3, True, 0.1)]:
        m = M.cube_mesh(n, n, n, periodic=per, jitter=jit, seed=N)
        st = dg.make_setup(m, N)
        vm, vp = st.vmaps()
        o = O.Setup.from_mesh(N, m)
        assert np.array_equal(vp, o.imap("vmapP").reshape(-1))
        assert np.array_equal(vm, o.imap("vmapM").reshape(-1))
        assert np.array_equal(st.EToE, o.imap("EToE")) and np.array_equal(st.EToF, o.imap("EToF"))


def test_vmapP_on_gambit_mesh():
    m = neu_mesh()
    for N in (1, 2, 3):
        st = dg.make_setup(m, N)
        o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
        assert np.array_equal(st.vmaps()[1], o.imap("vmapP").reshape(-1))


def test_gambit_reader_partitions():
    """mesh/readers/gambit_reader_file_test.go:36-38: p1..p4 groups, 565 tets."""
    m = neu_mesh()
    assert m.K == 565 and m.VX.size == 175
    assert sorted(np.bincount(m.EToP).tolist()) == [0, 141, 141, 141, 142]


@pytest.mark.parametrize("N,n", [(2, 4), (3, 3)])
def test_face_connector_numpy_equals_oracle(N, n):
    """Includes an empty partition in the middle."""
    m = M.cube_mesh(n, n, n, periodic=True)
    parts = [m.K // 4, 0, m.K // 2, m.K - m.K // 4 - m.K // 2]
    st = dg.make_setup(m, N, Kp=parts)
    vm, vp = st.vmaps()
    etop = M.etop_from_counts(st.Kp)
    fc = M.new_face_connector(m.K, 4, st.ref.Nfp, st.ref.Np, vp, etop)
    fc.verify()
    NP, counts, pk, pl = O.face_connector(m.K, 4, st.ref.Nfp, st.ref.Np, vp, etop)
    assert NP == fc.num_partitions
    for k in pk:
        assert np.array_equal(pk[k], fc.pick[k])
    for k in pl:
        assert np.array_equal(pl[k], fc.place[k])


def test_face_connector_verify_catches_corruption():
    m = M.cube_mesh(2, 2, 2, periodic=True)
    st = dg.make_setup(m, 1, Kp=[20, 28])
    fc = M.new_face_connector(m.K, 4, 3, 4, st.vmaps()[1], M.etop_from_counts(st.Kp))
    fc.verify()
    fc.pick[(0, 1)] = fc.pick[(0, 1)].copy()
    fc.pick[(0, 1)][0] = 10 ** 6
    with pytest.raises(ValueError, match="invalid pick index"):
        fc.verify()


def test_ref_matrices_names():
    """tetnudg/tet_nudg_mesh.go:370-389 names; M_ is V V^T (= M^-1, as the reference computes)."""
    r = nudg.build_ref_tet(2)
    mats = r.ref_matrices()
    assert set(mats) == {f"{k}_NudgTet2" for k in ("V", "Vinv", "M", "Minv", "Dr", "Ds", "Dt", "LIFT")}
    assert np.allclose(mats["M_NudgTet2"] @ r.Mass, np.eye(r.Np), atol=1e-12)


def test_pooled_layout_roundtrip():
    m = M.cube_mesh(3, 3, 3)
    st = dg.make_setup(m, 2, Kp=[50, 60, 52])
    u = np.arange(m.K * st.ref.Np, dtype=float).reshape(m.K, -1)
    pool = st.to_pooled(u)
    assert pool.size * 8 == st.state_bytes
    assert np.array_equal(st.from_pooled(pool).reshape(u.shape), u)
    assert all((o * 8) % dg.WARP_ALIGN == 0 for o in st.state_off)


def test_face_connector_c2_size_one_and_four_slabs():
    """2^20 jittered tets at P=4 (first 2^20 tets of a 64x64x43-hex cube):
    pick/place maps for 1 partition and for 4 slabs equal the oracle's."""
    m = M.cube_mesh(64, 64, 43, jitter=0.1, seed=5, kmax=2 ** 20)
    st = dg.make_setup(m, 4, Kp=[m.K])
    _, vp = st.vmaps()
    for parts in (1, 4):
        Kp = M.slab_partition(m.K, parts, 0.0, seed=0) if parts > 1 else np.array([m.K])
        etop = M.etop_from_counts(Kp)
        fc = M.new_face_connector(m.K, 4, st.ref.Nfp, st.ref.Np, vp, etop)
        NP, counts, pk, pl = O.face_connector(m.K, 4, st.ref.Nfp, st.ref.Np, vp, etop)
        assert NP == parts == fc.num_partitions
        assert sum(v.size for v in fc.pick.values()) == m.K * 4 * st.ref.Nfp
        for k in pk:
            assert np.array_equal(pk[k], fc.pick[k])
            assert np.array_equal(pl[(k[1], k[0])], fc.place[(k[1], k[0])])
