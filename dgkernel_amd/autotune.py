"""Tile-shape autotuner for the Burgers stage kernel.

For each polynomial order it times every stage-kernel variant and launch
shape on a representative periodic mesh (HIP events on the library's stream,
warm-up launches first) and records the fastest in a JSON table,
``dgkernel_amd/tuned_stage.json``. ``apply(mesh)`` sets a device mesh to the
tuned entry for its order; BurgersSolver applies it when the table exists.

Usage: python -m dgkernel_amd.autotune [--orders 1,2,3,4] [--tets 600000]
                                        [--reps 12] [--parts 8] [--out PATH] [--morton]
                                        [--merge]

Candidates (variant numbers of dgk_burgers_set_variant):
  1 VALU stage (one element per lane; 64/128/256 elements per block),
  2 pipelined FP64-MFMA stage, 4 MFMA stage without the load pipeline,
  5 deep-pipeline MFMA stage and 6 register-operator MFMA stage (N <= 3),
  each MFMA kernel with 2 or 4 waves per block (32/64 elements), 1-4
  blocks per CU, and three tile walks: grid stride, XCD-swept in partition
  order, XCD-swept along the Morton curve of the tile centroids of all
  partitions. Orders 5 and 6 run the split stage only (one candidate).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path

import numpy as np

TABLE = Path(__file__).resolve().parent / "tuned_stage.json"


def candidates(N: int, quick: bool = False) -> list[dict]:
    if N > 4:
        return [{"variant": 0, "tile": 0, "bpc": 0, "xcd": 0, "spatial": 0}]
    out = []
    mfma = [2, 4] + ([5, 6] if N <= 3 else [])
    # tile walks: grid stride, XCD-swept partition order, XCD-swept spatial
    # (Morton) order of the tiles of all partitions
    walks = ((0, 0),) if quick else ((0, 0), (1, 0), (1, 1))
    for v in mfma:
        for tile in (32, 64):
            for bpc in ((0,) if quick else (1, 2, 3, 4)):
                for xcd, spatial in walks:
                    out.append({"variant": v, "tile": tile, "bpc": bpc, "xcd": xcd,
                                "spatial": spatial})
    for tile in (64, 128, 256):
        for bpc in ((0,) if quick else (0, 2)):
            out.append({"variant": 1, "tile": tile, "bpc": bpc, "xcd": 0, "spatial": 0})
    return out


def configure(mesh, cfg: dict) -> None:
    """Apply one candidate to a DGMesh (variant first: it selects the kernel
    family the tile shape refers to)."""
    mesh.set_variant(int(cfg["variant"]))
    mesh.set_order(bool(cfg.get("xcd", 0)))
    mesh.set_tile_order(bool(cfg.get("spatial", 0)))
    if cfg.get("tile", 0) or cfg.get("bpc", 0):
        mesh.set_tile(int(cfg.get("tile", 0)), int(cfg.get("bpc", 0)))


def load_table(path: Path = TABLE) -> dict | None:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return None


def entry_for(table: dict | None, N: int) -> dict | None:
    if not table:
        return None
    e = table.get("entries", {}).get(str(N))
    if not e or any(k not in e for k in ("variant", "tile", "bpc")):
        return None
    return e


def apply(mesh, table: dict | None = None) -> dict | None:
    """Set `mesh` to the tuned entry of its order (no-op without a table
    entry); returns the entry applied."""
    table = load_table() if table is None else table
    e = entry_for(table, mesh.st.ref.N)
    if e is not None:
        configure(mesh, e)
    return e


def time_stage(solver, reps: int) -> float:
    """Mean stage duration (ms) of the solver's current configuration."""
    dev = solver.dev
    solver.step(1)
    dev.sync()
    t = dev.timer()
    t.start()
    solver.step(reps)
    t.stop()
    return t.ms() / (3 * reps)


def tune(dev, orders, tets: int = 600_000, reps: int = 12, quick: bool = False, log=print,
         morton: bool = False, parts: int = 8) -> dict:
    from . import dg
    from . import mesh as M
    table = {"device": dev.name, "tets_target": tets, "reps": reps, "entries": {}, "measured": {}}
    for N in orders:
        n = max(2, int(round((tets / 6) ** (1 / 3))))
        m = M.cube_mesh(n, n, n, periodic=True)
        Kp = M.slab_partition(m.K, parts, 0.4, seed=1)
        if morton:
            m = M.morton_within_partitions(m, Kp)
        st = dg.make_setup(m, N, Kp=Kp)
        x, y, z = st.node_coords()
        u0 = 0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
        solver = dg.BurgersSolver(dev, st, cfl=0.3, tuned=False)
        # clocks up before the first candidate: at least ~100 ms of stages
        solver.set_state(u0)
        warm = time_stage(solver, 2)
        solver.step(max(1, int(100.0 / max(3 * warm, 1e-3))))
        res = []
        for cfg in candidates(N, quick):
            configure(solver.mesh, cfg)
            solver.set_state(u0)
            ms = time_stage(solver, reps)
            res.append(dict(cfg, ms=ms))
            log(f"N={N} K={m.K} {cfg}: {ms:.4f} ms/stage")
        best = min(res, key=lambda r: r["ms"])
        table["entries"][str(N)] = dict(best, K=m.K, morton=morton, parts=parts,
                                        dof_stage_per_s=m.K * st.ref.Np / (best["ms"] * 1e-3))
        table["measured"][str(N)] = res
        del solver
    return table


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--orders", default="1,2,3,4,5,6")
    ap.add_argument("--tets", type=int, default=600_000)
    ap.add_argument("--reps", type=int, default=12)
    ap.add_argument("--quick", action="store_true", help="default launch shapes only")
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--morton", action="store_true",
                    help="tune on a mesh renumbered in Morton order within partitions")
    ap.add_argument("--parts", type=int, default=8,
                    help="slab partitions of the tuning mesh (+-40%% imbalance)")
    ap.add_argument("--merge", action="store_true",
                    help="update only the tuned orders of an existing table at --out")
    a = ap.parse_args(argv)
    from . import _lib
    dev = _lib.Device()
    orders = [int(v) for v in a.orders.split(",") if v]
    table = tune(dev, orders, a.tets, a.reps, a.quick,
                 log=lambda s: print(s, file=sys.stderr, flush=True), morton=a.morton,
                 parts=a.parts)
    old = load_table(Path(a.out)) if a.merge else None
    if old:
        for key in ("entries", "measured"):
            old.setdefault(key, {}).update(table[key])
        old["device"] = table["device"]
        table = old
    Path(a.out).write_text(json.dumps(table, indent=1) + "\n")
    print(json.dumps({k: v for k, v in table["entries"].items()}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
