"""dgk_phys_deriv_nodal_noalloc (the allocation-free physical derivative)
against dgk_phys_deriv_nodal, through the C ABI: two calls in a row, the
second on a smaller input, each bit-identical to the existing entry point and
within 1e-13 of the CPU oracle; a graph capture of it replayed three times;
argument checks on the host."""
import ctypes as C

import numpy as np
import pytest

from dgkernel_amd import _lib, dg
from dgkernel_amd import mesh as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

METRICS = ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")
P64 = C.POINTER(C.c_int64)


class _Case:
    """Pooled inputs and outputs for one partitioning of the first K tets."""

    def __init__(self, device, N, K, nparts, seed):
        m = M.cube_mesh(8, 8, 7, jitter=0.1, seed=3)
        m.EToV, m.EToVc = m.EToV[:K], m.EToVc[:K]
        o = O.Setup(N, m.VX, m.VY, m.VZ, m.EToV)
        assert o.status == 0
        el = O.element(N)
        Np = el["Np"]
        rng = np.random.default_rng(seed)
        Kp = np.ones(nparts, dtype=np.int64)
        Kp += np.bincount(rng.integers(0, nparts, K - nparts), minlength=nparts)
        assert Kp.sum() == K
        self.Kp, self.Np, self.N = Kp, Np, N
        U = M.uniform(seed, 3, K * Np, -1.0, 1.0).reshape(K, Np)
        mets = [o.field(k) for k in METRICS]
        self.want = O.phys_deriv(el["Dr"], el["Ds"], el["Dt"], U, mets)
        offs, nbytes = dg.aligned_offsets(Kp, Np, 8)
        self.offs = offs
        starts = np.concatenate(([0], np.cumsum(Kp)))

        def pool(a):
            out = np.zeros(nbytes // 8)
            for p in range(nparts):
                out[offs[p]:offs[p] + Kp[p] * Np] = a[starts[p]:starts[p + 1]].reshape(-1)
            return device.upload(out)

        self.u = pool(U)
        self.m = [pool(mm) for mm in mets]
        self.off = device.upload(np.asarray(offs, dtype=np.int64))
        self.D = [device.upload(np.ascontiguousarray(el[k])) for k in ("Dr", "Ds", "Dt")]
        self.outs = {name: [device.zeros(nbytes // 8) for _ in range(3)] for name in ("ref", "na")}
        self.starts = starts
        self.mptrs = (C.c_void_p * 9)(*[b.ptr.value for b in self.m])

    def call(self, device, fn, which):
        o = self.outs[which]
        _lib.call(fn, device.h, self.N, len(self.Kp), self.Kp.ctypes.data_as(P64), self.D[0].ptr,
                  self.D[1].ptr, self.D[2].ptr, self.u.ptr, self.off.ptr, self.mptrs, _lib.DGK_F64,
                  self.off.ptr, o[0].ptr, o[1].ptr, o[2].ptr, _lib.DGK_F64, self.off.ptr)

    def result(self, which):
        out = []
        for b in self.outs[which]:
            h = b.to_host()
            out.append(np.concatenate([h[self.offs[p]:self.offs[p] + self.Kp[p] * self.Np]
                                       for p in range(len(self.Kp))]).reshape(-1, self.Np))
        return out


def _check(device, case):
    case.call(device, "dgk_phys_deriv_nodal", "ref")
    case.call(device, "dgk_phys_deriv_nodal_noalloc", "na")
    device.sync()
    ref, na = case.result("ref"), case.result("na")
    for d in range(3):
        assert np.array_equal(na[d], ref[d]), d
        assert np.max(np.abs(na[d] - case.want[d])) <= 1e-13 * np.abs(case.want[d]).max()


@pytest.mark.parametrize("N", [3, 6])
def test_noalloc_two_calls_second_smaller(device, N):
    """600 partitions over 2688 tets, then 300 partitions over 1500 tets (the
    table keeps the first call's tail), then 5 partitions (below the by-value
    limit of the other entry point)."""
    big = _Case(device, N, 2688, 600, seed=1)
    _check(device, big)
    small = _Case(device, N, 1500, 300, seed=2)
    _check(device, small)
    few = _Case(device, N, 900, 5, seed=4)
    _check(device, few)


def test_noalloc_graph_replay(device):
    """Captured once (kernels only, no copies or fills), replayed three times
    with eager work between replays; every replay equals the eager result."""
    case = _Case(device, 3, 2688, 400, seed=5)
    case.call(device, "dgk_phys_deriv_nodal", "ref")
    device.sync()
    ref = case.result("ref")
    g = _lib.Graph(device, lambda: case.call(device, "dgk_phys_deriv_nodal_noalloc", "na"))
    other = _Case(device, 3, 1200, 260, seed=6)
    for rep in range(3):
        for b in case.outs["na"]:
            b.fill0()
        g.launch()
        device.sync()
        got = case.result("na")
        for d in range(3):
            assert np.array_equal(got[d], ref[d]), (rep, d)
        _check(device, other)  # eager calls in between rewrite the shared table


def test_noalloc_rejects_bad_sizes(device):
    case = _Case(device, 2, 600, 3, seed=7)
    Kbig = np.ones(65536, dtype=np.int64)
    o = case.outs["na"]
    with pytest.raises(_lib.DgkError, match="outside 1..65535"):
        _lib.call("dgk_phys_deriv_nodal_noalloc", device.h, 2, 65536, Kbig.ctypes.data_as(P64),
                  case.D[0].ptr, case.D[1].ptr, case.D[2].ptr, case.u.ptr, case.off.ptr, case.mptrs,
                  _lib.DGK_F64, case.off.ptr, o[0].ptr, o[1].ptr, o[2].ptr, _lib.DGK_F64, case.off.ptr)
    Kneg = np.array([200, -1, 401], dtype=np.int64)
    with pytest.raises(_lib.DgkError, match="out of range"):
        _lib.call("dgk_phys_deriv_nodal_noalloc", device.h, 2, 3, Kneg.ctypes.data_as(P64),
                  case.D[0].ptr, case.D[1].ptr, case.D[2].ptr, case.u.ptr, case.off.ptr, case.mptrs,
                  _lib.DGK_F64, case.off.ptr, o[0].ptr, o[1].ptr, o[2].ptr, _lib.DGK_F64, case.off.ptr)
