"""Mesh files: the Gambit neutral writer against the reader (the reference's
format, mesh/readers/gambit_reader.go:14-338) and the .npz store for generated
periodic meshes, so that CPU and GPU runs read identical inputs."""
import numpy as np
import pytest

from dgkernel_amd import mesh as M
from oracle import oracle as O
from tests.helpers import neu_mesh


def _same(a, b):
    assert np.array_equal(a.VX, b.VX) and np.array_equal(a.VY, b.VY) and np.array_equal(a.VZ, b.VZ)
    assert np.array_equal(a.EToV, b.EToV)
    assert (a.EToP is None) == (b.EToP is None)
    if a.EToP is not None:
        assert np.array_equal(a.EToP, b.EToP)


def test_gambit_round_trip_of_the_reference_fixture(tmp_path):
    a = neu_mesh()
    f = tmp_path / "cube.neu"
    M.write_gambit_neutral(a, str(f))
    b = M.read_gambit_neutral(str(f))
    _same(a, b)
    # the connectivity the oracle builds from it is the same, face for face
    ea, fa = M.connect(a)
    eb, fb = M.connect(b)
    assert np.array_equal(ea, eb) and np.array_equal(fa, fb)


@pytest.mark.parametrize("parts", [None, [0] * 162, "slabs"])
def test_gambit_round_trip_of_a_generated_mesh(tmp_path, parts):
    m = M.cube_mesh(3, 3, 3, periodic=False, jitter=0.1, seed=5)
    if parts == "slabs":
        m.EToP = M.etop_from_counts(M.slab_partition(m.K, 3, imbalance=0.3, seed=2))
    elif parts is not None:
        m.EToP = np.array(parts, dtype=np.int64)
    f = tmp_path / "gen.neu"
    M.write_gambit_neutral(m, str(f))
    _same(m, M.read_gambit_neutral(str(f)))


def test_gambit_writer_refuses_periodic_meshes(tmp_path):
    with pytest.raises(ValueError):
        M.write_gambit_neutral(M.cube_mesh(2, 2, 2, periodic=True), str(tmp_path / "p.neu"))


def test_npz_round_trip_of_a_periodic_mesh(tmp_path):
    m = M.cube_mesh(4, 3, 2, periodic=True, jitter=0.05, seed=9)
    m.EToP = M.etop_from_counts(M.slab_partition(m.K, 2))
    f = tmp_path / "per.npz"
    M.save_mesh(m, str(f))
    b = M.load_mesh(str(f))
    _same(m, b)
    assert np.array_equal(m.EToVc, b.EToVc) and b.period == m.period and b.info == m.info
    # the periodic connectivity and the oracle's geometry are identical
    assert all(np.array_equal(x, y) for x, y in zip(M.connect(m), M.connect(b)))
    oa, ob = O.Setup.from_mesh(2, m), O.Setup.from_mesh(2, b)
    assert np.array_equal(oa.imap("vmapP"), ob.imap("vmapP"))
    assert np.array_equal(oa.field("rx"), ob.field("rx"))
