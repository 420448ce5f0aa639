"""Autotuner host logic (CPU) and a small tuning run on the GPU."""
import json

import numpy as np
import pytest

from dgkernel_amd import autotune


class _FakeMesh:
    def __init__(self, N):
        self.calls = []

        class _Ref:
            pass

        class _St:
            pass

        self.st = _St()
        self.st.ref = _Ref()
        self.st.ref.N = N

    def set_variant(self, v):
        self.calls.append(("variant", v))

    def set_order(self, x):
        self.calls.append(("order", x))

    def set_tile_order(self, x):
        self.calls.append(("walk", x))

    def set_tile(self, t, b):
        self.calls.append(("tile", t, b))


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_candidates_cover_the_kernel_families(N):
    cs = autotune.candidates(N)
    assert cs and all(set(c) == {"variant", "tile", "bpc", "xcd", "spatial"} for c in cs)
    vs = {c["variant"] for c in cs}
    if N <= 3:
        assert vs == {1, 2, 4, 5, 6}
    elif N == 4:
        assert vs == {1, 2, 4}
    else:
        assert vs == {0}
    assert len(autotune.candidates(N, quick=True)) <= len(cs)


def test_configure_sets_variant_before_tile_and_apply_uses_table():
    m = _FakeMesh(3)
    autotune.configure(m, {"variant": 5, "tile": 32, "bpc": 2, "xcd": 1, "spatial": 1})
    assert m.calls == [("variant", 5), ("order", True), ("walk", True), ("tile", 32, 2)]
    # entries written before the tile walk was tuned keep the partition order
    table = {"entries": {"3": {"variant": 2, "tile": 64, "bpc": 0, "xcd": 0, "ms": 1.0}}}
    m2 = _FakeMesh(3)
    assert autotune.apply(m2, table)["variant"] == 2
    assert m2.calls == [("variant", 2), ("order", False), ("walk", False), ("tile", 64, 0)]
    m3 = _FakeMesh(4)
    assert autotune.apply(m3, table) is None and m3.calls == []
    assert autotune.entry_for({"entries": {"3": {"variant": 2}}}, 3) is None  # incomplete entry


def test_load_table_tolerates_missing_or_bad_files(tmp_path):
    assert autotune.load_table(tmp_path / "none.json") is None
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    assert autotune.load_table(bad) is None
    good = tmp_path / "good.json"
    good.write_text(json.dumps({"entries": {}}))
    assert autotune.load_table(good) == {"entries": {}}


@pytest.mark.gpu
def test_tune_small_mesh_and_apply(device):
    from dgkernel_amd import dg
    from dgkernel_amd import mesh as M
    from tests.helpers import rel_l2, smooth_u0
    table = autotune.tune(device, [2], tets=6000, reps=2, quick=True, log=lambda s: None)
    e = table["entries"]["2"]
    assert e["ms"] > 0 and e["ms"] == min(r["ms"] for r in table["measured"]["2"])
    m = M.cube_mesh(5, 5, 5, periodic=True)
    st = dg.make_setup(m, 2, Kp=[300, 450])
    u0 = smooth_u0(st)
    a = dg.BurgersSolver(device, st, tuned=False)
    a.set_state(u0)
    a.step(4)
    b = dg.BurgersSolver(device, st, hmin=a.hmin, tuned=False)
    assert autotune.apply(b.mesh, table) == e
    b.set_state(u0)
    b.step(4)
    assert rel_l2(b.state(), a.state()) <= 1e-13
