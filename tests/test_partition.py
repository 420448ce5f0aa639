"""Geometric partitioning (recursive coordinate bisection) and partition-major
renumbering: exact counts, fewer faces between partitions than slabs, and a
FaceConnector that verifies on the result."""
import numpy as np
import pytest

from dgkernel_amd import dg
from dgkernel_amd import mesh as M
from oracle import oracle as O


def _cut_faces(m, etop):
    EToE, _ = M.connect(m)
    own = np.repeat(np.arange(m.K), 4).reshape(m.K, 4)
    return int(np.sum((EToE != own) & (etop[EToE] != etop[own]))) // 2


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("nparts,imb", [(8, 0.0), (7, 0.4), (64, 0.4)])
def test_rcb_counts_and_cut(periodic, nparts, imb):
    m = M.cube_mesh(8, 8, 8, periodic=periodic, jitter=0.05, seed=3)
    kp = M.slab_partition(m.K, nparts, imb, seed=11)
    etop = M.rcb_partition(m, kp)
    assert np.array_equal(np.bincount(etop, minlength=nparts), kp)
    slabs = M.etop_from_counts(kp)
    cut_rcb, cut_slab = _cut_faces(m, etop), _cut_faces(m, slabs)
    print(f"{nparts} parts (imbalance {imb}, periodic {periodic}): cut faces rcb {cut_rcb}, slabs {cut_slab}")
    assert cut_rcb < cut_slab
    if nparts == 8 and not periodic:
        # eight octants against eight x-slabs: 3 cut planes of 2*8*8 faces against 7
        assert cut_rcb == 3 * 2 * 8 * 8 and cut_slab == 7 * 2 * 8 * 8


def test_rcb_rejects_bad_counts():
    m = M.cube_mesh(2, 2, 2)
    with pytest.raises(ValueError):
        M.rcb_partition(m, [10, 10])


def test_order_by_partition_and_face_connector():
    m = M.cube_mesh(5, 5, 5, periodic=True)
    kp = M.slab_partition(m.K, 5, 0.3, seed=4)
    etop = M.rcb_partition(m, kp)
    mo, kp2 = M.order_by_partition(m, etop)
    assert np.array_equal(kp2, kp)
    assert np.array_equal(mo.EToP, M.etop_from_counts(kp))
    # same element set, each element with its own vertices
    assert sorted(map(tuple, mo.EToV)) == sorted(map(tuple, m.EToV))
    st = dg.make_setup(mo, 2, Kp=kp2)
    _, vp = st.vmaps()
    fc = M.new_face_connector(mo.K, 4, st.ref.Nfp, st.ref.Np, vp, mo.EToP)
    fc.verify()
    NP, _, pk, pl = O.face_connector(mo.K, 4, st.ref.Nfp, st.ref.Np, vp, mo.EToP)
    assert NP == 5
    for k in pk:
        assert np.array_equal(pk[k], fc.pick[k])
    for k in pl:
        assert np.array_equal(pl[k], fc.place[k])
