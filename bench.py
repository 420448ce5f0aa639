"""Benchmark: DOF-stage-updates/s of the fused tet DG Burgers stage on one MI355X.

Workload (default): 96^3 hexes x 6 = 5,308,416 tets on the periodic unit
cube, P=3 (Np=20), FP64 state, u0 = 0.5 + 0.4 sin(2 pi x) cos(2 pi y)
sin(2 pi z), 64 contiguous x-slab partitions with seeded +-40% imbalance,
strong-form Lax-Friedrichs flux + LIFT, SSP-RK3 with dt = 0.3 hmin/max|u|
recomputed on the device every step. One persistent launch per stage covers
all partitions.

A step of the timed region is one SSP-RK3 step (three fused stage launches
plus the one-thread dt kernel); inputs are resident in HBM before timing
starts. Each stage launch is bracketed by HIP events on the library's
stream, which gives the per-launch duration used for the roofline figure.

Prints one JSON line. ``--gpus`` above 1 is rejected: this build runs on a
single GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes_per_elem(Np: int) -> float:
    """Algorithmic HBM bytes per element per SSP-RK3 stage, FP64 state, affine tets:
    (8/3)*8*Np state (2, 3, 3 arrays over the three stages) + 25 geometry doubles
    + 4 neighbour ids + 4 orientation codes (int32)."""
    return (8.0 / 3.0) * 8.0 * Np + 8.0 * 25 + 32.0


def stage_flops_per_elem(Np: int, Nfp: int) -> float:
    return 6.0 * Np * Np + 8.0 * Np * Nfp + 12.0 * Np + 72.0 * Nfp


def stage_kernel_name(N: int, solver) -> str:
    """The stage kernel the solver's configuration launches (variant numbers of
    dgk_burgers_set_variant; the tuned entry when the autotuner table is used)."""
    e = solver.tuned or {}
    v = e.get("variant", 0)
    if N > 4 or v == 3:
        return f"split stage (deriv_affine_kernel<{N}, 0, true> + surface_update_kernel<{N}, 0>)"
    if v == 1:
        return f"burgers_stage_kernel<{N}, {e.get('tile') or 64}>"
    if v == 4:
        return f"burgers_stage_mfma_kernel<{N}, {max(1, (e.get('tile') or 64) // 16)}>"
    kind = {5: 1, 6: 2}.get(v, 0)
    return f"burgers_stage_pipe_kernel<{N}, {max(1, (e.get('tile') or 64) // 16)}, MODE, {kind}>"


def read_traffic(name: str, key: str = "traffic_bytes_per_launch"):
    """HBM bytes per launch recorded from rocprofv3 --pmc runs (profiles/<name>)."""
    v = read_profile_key(name, key)
    try:
        return None if v is None else float(v.split()[0])
    except ValueError:
        return None


def read_profile_key(name: str, key: str):
    """The text after `key:` on its own line of profiles/<name>, or None."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.strip().startswith(key + ":"):
            return line.split(":", 1)[1].strip()
    return None


def stage_source_hash() -> str:
    """SHA-256 (first 16 hex digits) of the fused stage kernel's source: a
    profile records it beside the traffic it measured, so a line can say
    whether that figure belongs to the kernel it timed."""
    import hashlib
    return hashlib.sha256((ROOT / "dgkernel_amd" / "csrc" / "burgers.hip").read_bytes()).hexdigest()[:16]


def build_c3(n: int, N: int, nparts: int, seed: int, morton: bool = False):
    from dgkernel_amd import dg
    from dgkernel_amd import mesh as M
    m = M.cube_mesh(n, n, n, periodic=True)
    Kp = M.slab_partition(m.K, nparts, 0.4, seed=seed) if nparts > 1 else np.array([m.K])
    if morton:
        m = M.morton_within_partitions(m, Kp)
    st = dg.make_setup(m, N, Kp=Kp)
    x, y, z = st.node_coords()
    u0 = 0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
    del x, y, z
    return m, st, u0


def cpu_baseline(N: int, n: int, steps: int, threads: int | None, parts: int = 64,
                 seed: int = 2024):
    """The CPU oracle (C++ restatement) on a bounded sample of the same problem,
    in two schedules: all threads element-parallel (the reported value), and
    the reference's OCCA-OpenMP schedule -- parallel over the (+-40% uneven)
    slab partitions with a serial element loop inside, so at most
    min(partitions, threads) threads are busy."""
    from oracle import oracle as O
    if threads:
        O.set_threads(threads)
    m, st, u0 = build_c3(n, N, parts, seed)
    o = O.Setup.from_mesh(N, m)
    hmin = o.min_node_distance()
    dof_stage = m.K * st.ref.Np * 3 * steps
    t0 = time.perf_counter()
    o.ssprk3(u0, steps, 0.3, hmin)
    dt_e = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.ssprk3_partitions(u0, steps, 0.3, hmin, st.Kp)
    dt_p = time.perf_counter() - t0
    nthr = O.num_threads()
    return {"value": dof_stage / dt_p, "unit": "DOF-stage-updates/s",
            "cores": min(st.npart, nthr), "kind": "port",
            "sample": f"{n}^3 hexes x 6 = {m.K} tets in {st.npart} slab partitions, P={N}, "
                      f"{steps} SSP-RK3 steps ({3 * steps} stages) of the same periodic Burgers "
                      f"problem; C++ oracle with the reference's OCCA-OpenMP schedule (omp "
                      f"parallel for over partitions, serial element loop), {dt_p:.2f} s",
            "element_parallel": {"value": dof_stage / dt_e, "threads": nthr, "seconds": dt_e,
                                 "schedule": "every operator pass omp-parallel over all "
                                             "elements"},
            "cpu": _cpu_model()}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Sustained-load warm-up: on a GPU that was idle the stage kernels run up to
# ~25 % slower during the first ~25 ms of load (clocks ramp up), so every timed
# region is preceded by at least this much of the same work.
WARM_MS = 60.0
# The headline stage mode warms up for longer: its timed region is short
# (~110 ms) and any residue of the ramp shows in its mean.
WARM_STAGE_MS = 150.0


def _warm(dev, fn, min_calls: int = 2, min_ms: float = WARM_MS) -> int:
    """Call ``fn`` at least ``min_calls`` times and until ``min_ms`` of wall time
    has passed with the device busy; returns the number of calls."""
    dev.sync()
    t0 = time.perf_counter()
    n = 0
    while n < min_calls or (time.perf_counter() - t0) * 1e3 < min_ms:
        fn()
        n += 1
        dev.sync()
    return n


def _time_ms(dev, fn, reps: int) -> float:
    """Mean duration of ``fn`` (enqueued on the library stream) over ``reps`` calls, HIP events,
    after the sustained-load warm-up."""
    _warm(dev, fn)
    t = dev.timer()
    t.start()
    for _ in range(reps):
        fn()
    t.stop()
    return t.ms() / reps


def _errs(x, ref):
    x = np.asarray(x, dtype=np.float64)
    return {"rel_l2": float(np.linalg.norm(x - ref) / np.linalg.norm(ref)),
            "max_abs": float(np.max(np.abs(x - ref))),
            "max_rel": float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))}


def bench_deriv(a, dev) -> dict:
    """Physical derivative: per-node (reference layout) and affine MFMA forms.

    c2: first 2^20 tets of a 64x64x43-hex cube, vertices jittered +-10% of h,
        P=4, u ~ U[-1,1]; FP64 and FP32.
    c4: 64^3 x 6 = 1,572,864 periodic tets, P=6, smooth field + U(+-1e-6);
        FP64, FP32, split BF16x3 and BF16x6 (FP32 outputs below FP64).
    """
    import ctypes as C
    from dgkernel_amd import _lib, dg
    from dgkernel_amd import mesh as M
    case = a.case
    if case == "c2":
        m = M.cube_mesh(64, 64, 43, jitter=0.1, seed=a.seed, kmax=2 ** 20)
        N = 4
    else:
        m = M.cube_mesh(64, 64, 64, periodic=True)
        N = 6
    st = dg.make_setup(m, N, Kp=[m.K])
    K, Np = m.K, st.ref.Np
    if case == "c2":
        u = M.uniform(a.seed, 1, K * Np, -1.0, 1.0)
    else:
        x, y, z = st.node_coords()
        u = (0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
             ).reshape(-1) + M.uniform(a.seed, 7, K * Np, -1e-6, 1e-6)
        del x, y, z
    mesh = dg.DGMesh(dev, st)
    geo_ptr = mesh.geometry_dev_ptr()
    du = dev.upload(u)
    D = (st.ref.Dr, st.ref.Ds, st.ref.Dt)
    flops = (6.0 * Np * Np + 15.0 * Np) * K
    out = {"metric": f"physical derivative ({case})", "tets": K, "order": N, "Np": Np,
           "reps": a.reps, "kernels": {}}
    ref64 = None
    res = {}
    precs = [("affine_f64", _lib.DGK_PREC_F64), ("affine_f32", _lib.DGK_PREC_F32),
             ("affine_bf16x3", _lib.DGK_PREC_BF16X3), ("affine_bf16x6", _lib.DGK_PREC_BF16X6),
             ("affine_i8x6", _lib.DGK_PREC_I8X6)]
    rst = (st.ref.r, st.ref.s, st.ref.t)
    if a.precs:
        keep = set(a.precs.split(","))
        precs = [pp for pp in precs if pp[0].split("_", 1)[1] in keep or pp[0] == "affine_f64"]
    for name, prec in precs:
        plan = dg.DerivPlan(dev, N, D, prec, rst)
        outs = [dev.empty(K * Np, plan.out_dtype) for _ in range(3)]
        ms = _time_ms(dev, lambda: plan.apply(K, du, geo_ptr, K, outs), a.reps)
        ob = 8 if prec in (_lib.DGK_PREC_F64, _lib.DGK_PREC_I8X6) else 4
        nbytes = (Np * 8 + 3 * Np * ob + 72) * K
        got = [o.to_host().astype(np.float64) for o in outs]
        if ref64 is None:
            ref64 = got
        e = [_errs(g, r) for g, r in zip(got, ref64)]
        res[name] = ms
        out["kernels"][name] = {
            "ms": ms, "bytes_per_elem": nbytes / K, "GBps": nbytes / ms * 1e-6,
            "hbm_frac": nbytes / ms * 1e-6 / HBM_PEAK_GBS, "TFLOPs": flops / ms * 1e-9,
            "err_vs_f64": {"rel_l2": max(q["rel_l2"] for q in e),
                           "max_abs": max(q["max_abs"] for q in e),
                           "max_rel": max(q["max_rel"] for q in e)}}
        log(f"{name}: {ms:.4f} ms, {nbytes / ms * 1e-6:.0f} GB/s, {flops / ms * 1e-9:.1f} TF/s, "
            f"err {out['kernels'][name]['err_vs_f64']}")
        del outs, got
        plan.free()
    if case == "c2" and not a.precs:
        # Reference layout: per-node metrics and outputs, 13 Np K values per call.
        geo = mesh.geometry()[:9]
        Dm = [dev.upload(np.ascontiguousarray(d)) for d in D]
        off = dev.upload(np.array([0, K * Np], dtype=np.int64))
        Karr = np.array([K], dtype=np.int64)
        for name, mt in (("nodal_f64", _lib.DGK_F64), ("nodal_f32", _lib.DGK_F32)):
            mdt = np.float64 if mt == _lib.DGK_F64 else np.float32
            mb = [dev.upload(np.repeat(geo[q], Np).astype(mdt)) for q in range(9)]
            mptrs = (C.c_void_p * 9)(*[b.ptr.value for b in mb])
            outs = [dev.empty(K * Np, mdt) for _ in range(3)]
            fn = lambda: _lib.call(  # noqa: E731
                "dgk_phys_deriv_nodal", dev.h, N, 1, Karr.ctypes.data_as(C.POINTER(C.c_int64)),
                Dm[0].ptr, Dm[1].ptr, Dm[2].ptr, du.ptr, off.ptr, mptrs, mt, off.ptr,
                outs[0].ptr, outs[1].ptr, outs[2].ptr, mt, off.ptr)
            ms = _time_ms(dev, fn, a.reps)
            vb = 8 if mt == _lib.DGK_F64 else 4
            nbytes = (Np * 8 + 12 * Np * vb) * K
            got = [o.to_host().astype(np.float64) for o in outs]
            e = [_errs(g, r) for g, r in zip(got, ref64)]
            out["kernels"][name] = {
                "ms": ms, "bytes_per_elem": nbytes / K, "GBps": nbytes / ms * 1e-6,
                "hbm_frac": nbytes / ms * 1e-6 / HBM_PEAK_GBS, "TFLOPs": flops / ms * 1e-9,
                "err_vs_affine_f64": {"rel_l2": max(q["rel_l2"] for q in e),
                                      "max_rel": max(q["max_rel"] for q in e)}}
            log(f"{name}: {ms:.4f} ms, {nbytes / ms * 1e-6:.0f} GB/s")
            del mb, outs, got
        # FaceConnector exchange for 4 slabs, bit-exact maps; bytes: 8 read + 8 write + 2x4 index.
        _, vp = st.vmaps()
        Kp = M.slab_partition(K, 4, 0.0, seed=0)
        etop = M.etop_from_counts(Kp)
        fc = M.new_face_connector(K, 4, st.ref.Nfp, Np, vp, etop)
        uoff, _ = dg.aligned_offsets(Kp, Np)
        Poff, Pbytes = dg.aligned_offsets(Kp, 4 * st.ref.Nfp)
        counts = np.array([[fc.pick[(s_, d_)].size for d_ in range(4)] for s_ in range(4)],
                          dtype=np.int64)
        pick = np.concatenate([fc.pick[(s_, d_)] for s_ in range(4) for d_ in range(4)]).astype(np.int32)
        place = np.concatenate([fc.place[(d_, s_)] for d_ in range(4) for s_ in range(4)]).astype(np.int32)
        plan = C.c_void_p()
        _lib.call("dgk_face_plan_create", dev.h, 4, counts.ctypes.data_as(C.POINTER(C.c_int64)),
                  pick.ctypes.data_as(C.POINTER(C.c_int32)), place.ctypes.data_as(C.POINTER(C.c_int32)),
                  C.byref(plan))
        up = dev.empty(int(uoff[-1]))
        dP = dev.empty(Pbytes // 8)
        duo, dPo = dev.upload(uoff), dev.upload(Poff)
        ms = _time_ms(dev, lambda: _lib.call("dgk_face_exchange", plan, up.ptr, duo.ptr, dP.ptr,
                                             dPo.ptr), a.reps)
        nbytes = pick.size * 24
        out["kernels"]["face_exchange_4slabs"] = {"ms": ms, "GBps": nbytes / ms * 1e-6,
                                                  "hbm_frac": nbytes / ms * 1e-6 / HBM_PEAK_GBS,
                                                  "points": int(pick.size)}
        log(f"face exchange: {ms:.4f} ms, {nbytes / ms * 1e-6:.0f} GB/s")
        _lib.call("dgk_face_plan_destroy", plan)
    else:
        out["speedup_vs_f64"] = {k: res["affine_f64"] / v for k, v in res.items()}
    out["device"] = dev.name
    return out


# Stage precision settings: Dr/Ds/Dt (and LIFT) precision; "_lift64" keeps
# the LIFT product in FP64.
STAGE_PRECS = {"f64": 0, "f32": 1, "bf16x3": 2, "bf16x6": 3,
               "f32_lift64": 1 | 16, "bf16x3_lift64": 2 | 16, "bf16x6_lift64": 3 | 16,
               "i8x6": 4}


def bench_c4_stage(a, dev) -> dict:
    """Mixed precision at high order: 64^3 x 6 = 1,572,864 periodic tets, P=6,
    u0 = 0.5 + 0.4 sin(2 pi x) cos(2 pi y) sin(2 pi z) + U(+-1e-6), 2 slab
    partitions, ``a.steps`` SSP-RK3 steps per precision of the D and LIFT
    contractions. The state, geometry, flux and dt stay FP64; each precision's
    final state is compared with the FP64 GPU run."""
    from dgkernel_amd import _lib, dg
    from dgkernel_amd import mesh as M
    n, N = a.n, a.order
    m = M.cube_mesh(n, n, n, periodic=True)
    Kp = M.slab_partition(m.K, max(2, a.parts), 0.0, seed=a.seed)
    if a.morton:
        m = M.morton_within_partitions(m, Kp)
    st = dg.make_setup(m, N, Kp=Kp)
    x, y, z = st.node_coords()
    u0 = (0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
          + M.uniform(a.seed, 7, x.size, -1e-6, 1e-6).reshape(x.shape))
    del x, y, z
    K, Np = m.K, st.ref.Np
    ref = None
    out = {"metric": "DOF-stage-updates/s per contraction precision (P=6 high-order case)",
           "config": {"workload": f"Burgers SSP-RK3, {n}^3x6={K} periodic tets, P={N}, "
                                  f"{st.npart} slab partitions, {a.steps} steps"
                                  + (", Morton order within partitions" if a.morton else ""),
                      "tets": K, "order": N, "Np": Np, "steps": a.steps},
           "precisions": {}}
    solver = None
    hmin = None
    precs = [p for p in STAGE_PRECS if not a.precs or p == "f64" or p in a.precs.split(",")]
    for name in precs:
        solver = dg.BurgersSolver(dev, st, cfl=0.3, hmin=hmin)
        hmin = solver.hmin
        if a.tile_order >= 0:
            solver.mesh.set_tile_order(bool(a.tile_order))
        solver.mesh.set_precision(STAGE_PRECS[name])
        if a.variant >= 0:
            solver.mesh.set_variant(a.variant)
        solver.set_state(u0)
        _warm(dev, lambda: solver.step(1), min_calls=1)  # outside the timed run, restarted below
        solver.set_state(u0)
        dev.sync()
        t = dev.timer()
        t.start()
        solver.step(a.steps, graph=not a.no_graph)
        t.stop()
        ms = t.ms()
        u = solver.state()
        rate = K * Np * 3 * a.steps / (ms * 1e-3)
        rec = {"ms_per_step": ms / a.steps, "value": rate, "final_dt": solver.dt()}
        if ref is None:
            ref = u
        else:
            rec["err_vs_f64"] = _errs(u, ref)
        out["precisions"][name] = rec
        log(f"c4 stage {name}: {ms / a.steps:.3f} ms/step, {rate:.3e} DOF-stage/s, "
            f"{rec.get('err_vs_f64', '')}")
        del solver
    out["device"] = dev.name
    return out


def p6_section(dev, seed: int, steps: int = 5, reps: int = 10) -> dict:
    """The P=6 half of the metric, for the default bench line: the 64^3 x 6 =
    1,572,864-tet periodic P=6 problem (2 slab partitions, u0 = 0.5 + 0.4
    sin(2 pi x) cos(2 pi y) sin(2 pi z) + U(+-1e-6)). SSP-RK3 steps with the
    Dr/Ds/Dt contractions in FP64, in the exact-integer split (I8X6) and in
    BF16x3 (LIFT in FP64 for both), and with D and LIFT both in BF16x3 or both
    in BF16x6 (the fused stage: one launch per stage), each after a sustained
    warm-up and compared
    with the FP64 run after the same steps; and the physical derivative alone
    (three outputs) in the FP64, I8X6 and BF16x3 plans."""
    from dgkernel_amd import _lib, dg
    from dgkernel_amd import mesh as M
    N = 6
    m = M.cube_mesh(64, 64, 64, periodic=True)
    st = dg.make_setup(m, N, Kp=M.slab_partition(m.K, 2, 0.0, seed=seed))
    x, y, z = st.node_coords()
    u0 = (0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
          + M.uniform(seed, 7, x.size, -1e-6, 1e-6).reshape(x.shape))
    del x, y, z
    K, Np = m.K, st.ref.Np
    out = {"workload": f"Burgers SSP-RK3, 64^3x6={K} periodic tets, P=6, 2 slab partitions, "
                       f"{steps} steps per plan", "stage": {}, "derivative": {}}
    ref, hmin = None, None
    for name, prec in (("f64", 0), ("i8x6", 4), ("bf16x3_lift64", 2 | 16), ("bf16x3", 2), ("bf16x6", 3)):
        solver = dg.BurgersSolver(dev, st, cfl=0.3, hmin=hmin)
        hmin = solver.hmin
        solver.mesh.set_precision(prec)
        solver.set_state(u0)
        _warm(dev, lambda: solver.step(1), min_calls=1, min_ms=WARM_STAGE_MS)
        solver.set_state(u0)
        dev.sync()
        t = dev.timer()
        t.start()
        solver.step(steps)
        t.stop()
        ms = t.ms() / steps
        u = solver.state()
        rec = {"ms_per_step": ms, "value": K * Np * 3 / (ms * 1e-3)}
        if ref is None:
            ref = u
        else:
            rec["err_vs_f64"] = _errs(u, ref)
            rec["speedup_vs_f64"] = out["stage"]["f64"]["ms_per_step"] / ms
        out["stage"][name] = rec
        log(f"p6 stage {name}: {ms:.3f} ms/step {rec.get('err_vs_f64', '')}")
        del solver
    mesh = dg.DGMesh(dev, st)
    geo_ptr = mesh.geometry_dev_ptr()
    du = dev.upload(u0.reshape(-1))
    D, rst = (st.ref.Dr, st.ref.Ds, st.ref.Dt), (st.ref.r, st.ref.s, st.ref.t)
    ref = None
    for name, prec in (("f64", 0), ("i8x6", 4), ("bf16x3", 2)):
        plan = dg.DerivPlan(dev, N, D, prec, rst)
        outs = [dev.empty(K * Np, plan.out_dtype) for _ in range(3)]
        ms = _time_ms(dev, lambda: plan.apply(K, du, geo_ptr, K, outs), reps)
        got = [o.to_host().astype(np.float64) for o in outs]
        rec = {"ms": ms}
        if ref is None:
            ref = got
        else:
            e = [_errs(g, r) for g, r in zip(got, ref)]
            rec["err_vs_f64"] = {"rel_l2": max(q["rel_l2"] for q in e),
                                 "max_rel": max(q["max_rel"] for q in e)}
            rec["speedup_vs_f64"] = out["derivative"]["f64"]["ms"] / ms
        out["derivative"][name] = rec
        log(f"p6 derivative {name}: {ms:.4f} ms {rec.get('err_vs_f64', '')}")
        del outs, got
        plan.free()
    return out


def bench_partition(a) -> dict:
    """Partition quality on the n^3 x 6 periodic cube (host only): faces
    between partitions (the exchange's face count), communication volume
    (per element, the other partitions among its face neighbours), load
    imbalance and wall time, for the stage bench's slabs (+-40 %), equal
    slabs, coordinate bisection and the k-way graph partitioner (cut and
    volume objectives, 5 % imbalance)."""
    from dgkernel_amd import mesh as M
    from dgkernel_amd import partitioner as P
    m = M.cube_mesh(a.n, a.n, a.n, periodic=True)
    k = a.parts
    EToE, _ = M.connect(m)
    own = np.repeat(np.arange(m.K), 4).reshape(m.K, 4)

    def quality(etop, secs):
        c = (EToE != own) & (etop[EToE] != etop[own])
        pe = np.sort(np.where(c, etop[EToE], -1), axis=1)
        first = np.concatenate([np.ones((m.K, 1), bool), pe[:, 1:] != pe[:, :-1]], axis=1)
        cnt = np.bincount(etop, minlength=k)
        return {"cut_faces": int(c.sum()) // 2, "volume": int(((pe >= 0) & first).sum()),
                "imbalance": float(cnt.max() / cnt.mean() - 1.0), "seconds": secs}

    out = {"metric": "partition quality", "workload": f"{a.n}^3x6={m.K} periodic tets, {k} partitions",
           "partitions": {}}
    kp40 = M.slab_partition(m.K, k, 0.4, seed=a.seed)
    kp = M.slab_partition(m.K, k, 0.0)
    out["partitions"]["slabs_imbalanced_40pct"] = quality(M.etop_from_counts(kp40), 0.0)
    out["partitions"]["slabs"] = quality(M.etop_from_counts(kp), 0.0)
    t = time.perf_counter()
    e = M.rcb_partition(m, kp)
    out["partitions"]["rcb"] = quality(e, time.perf_counter() - t)
    for obj in ("cut", "vol"):
        t = time.perf_counter()
        e = P.kway_partition(m, k, obj, seed=a.seed)
        out["partitions"][f"kway_{obj}"] = quality(e, time.perf_counter() - t)
    for name, q in out["partitions"].items():
        log(f"{name}: {q}")
    return out


def bench_c5(a, dev) -> dict:
    """Learned artificial viscosity: 64x64x85 hexes x 6 = 2,088,960 periodic
    tets (2^21 is not a multiple of 6), P=4, 2 slab partitions,
    u0 = 0.5 + sin(2 pi x) sin(2 pi y) sin(2 pi z), SSP-RK3 past shock formation.

    The sensor network is trained on the GPU (hand-written kernels) on rows
    (normalised modal coefficients, Persson-Peraire fraction) from
    Persson-Peraire runs of 32 seeded fields on a 16^3 x 6 mesh; then the
    Persson-Peraire and learned-viscosity runs are timed on the full mesh and
    judged against the exact entropy solution on a node sample."""
    from dgkernel_amd import _lib
    from dgkernel_amd import av as AV
    from dgkernel_amd import dg
    from dgkernel_amd import mesh as M
    N = 4
    out = {"metric": "learned vs Persson-Peraire artificial viscosity (P=4 shock-forming Burgers)"}
    # 1. training data and training
    t0 = time.perf_counter()
    # training mesh with the target mesh's hex aspect ratio (nz / n)
    mt = M.cube_mesh(a.train_n, a.train_n, max(2, round(a.train_n * a.nz / a.c5_n)), periodic=True)
    stt = dg.make_setup(mt, N, Kp=[mt.K])
    p = AV.AVParams.for_order(N)
    X, T, nrows = AV.generate_training_data(dev, stt, p, seeds=range(1000, 1000 + a.train_seeds),
                                            t_end=a.t_end, every=a.train_every)
    t_gen = time.perf_counter() - t0
    log(f"training data: {nrows} rows from {a.train_seeds} fields on {mt.K} tets, {t_gen:.1f} s")
    net = AV.SensorNet(dev, stt.ref.Np, seed=1)
    dev.sync()
    t0 = time.perf_counter()
    hist = net.train(X, T, nrows, steps=a.train_steps, batch=a.train_batch, lr=3e-3, lr_final=5e-5,
                     report_every=max(1, a.train_steps // 5))
    dev.sync()
    t_train = time.perf_counter() - t0
    log(f"training: {a.train_steps} Adam steps x {a.train_batch}, {t_train:.1f} s, loss {hist}")
    out["training"] = {"rows": nrows, "fields": a.train_seeds, "mesh_tets": mt.K,
                       "generate_s": t_gen, "steps": a.train_steps, "batch": a.train_batch,
                       "train_s": t_train, "samples_per_s": a.train_steps * a.train_batch / t_train,
                       "loss_history": hist}
    # The gradient over the whole training set (one dgk_mlp_grad call on the
    # BF16 matrix cores): rows/s and the algorithmic rate against the dense
    # BF16 peak. Flops per row: forward 2 (64 nin + 64 x 64), backward
    # W2^T dz2, dz2^T h1 (2 x 64 x 64 each) and dz1^T x (2 x 64 nin).
    nin = stt.ref.Np
    gbuf = dev.zeros(net.nparams, np.float32)
    lbuf = dev.zeros(4, np.float32)

    def full_grad():
        _lib.call("dgk_mlp_grad", net.h, X.ptr, T.ptr, nrows, gbuf.ptr, lbuf.ptr)

    _warm(dev, full_grad)
    reps = 10
    dev.sync()
    tm = dev.timer()
    tm.start()
    for _ in range(reps):
        full_grad()
    tm.stop()
    g_ms = tm.ms() / reps
    flops_row = 4 * 64 * nin + 6 * 64 * 64
    out["training"]["grad_full"] = {
        "rows": nrows, "ms": g_ms, "rows_per_s": nrows / (g_ms * 1e-3),
        "tflops": flops_row * nrows / (g_ms * 1e-3) / 1e12,
        "frac_bf16_peak": flops_row * nrows / (g_ms * 1e-3) / 2.5e15,
        "flops_per_row": flops_row}
    log(f"gradient over {nrows} rows: {g_ms:.3f} ms, {nrows / (g_ms * 1e-3):.3e} rows/s, "
        f"{out['training']['grad_full']['tflops']:.1f} TFLOP/s")
    del X, T, gbuf, lbuf
    # 2. the full mesh
    n = a.c5_n
    m = M.cube_mesh(n, n, a.nz, periodic=True)
    st = dg.make_setup(m, N, Kp=M.slab_partition(m.K, 2))
    x, y, z = st.node_coords()
    u0 = 0.5 + np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    out["config"] = {"workload": f"{n}x{n}x{a.nz}x6={m.K} periodic tets, P={N}, 2 slab "
                                 f"partitions, SSP-RK3 to t={a.t_end}", "tets": m.K, "order": N,
                     "Np": st.ref.Np, "t_end": a.t_end, "cfl": 0.3, "cfl_v": p.cfl_v,
                     "ceps": p.ceps, "kappa": p.kappa, "s0": p.s0}
    rng = np.random.default_rng(a.seed)
    sample = rng.choice(x.size, size=min(a.exact_nodes, x.size), replace=False)
    runs = {}
    for sensor in ("pp", "mlp"):
        s = AV.AVSolver(dev, st, p, sensor, net=net if sensor == "mlp" else None)
        s.set_state(u0)
        # sensor cost: one evaluation on the initial state, HIP events
        ms_sensor = _time_ms(dev, s.sense, a.reps)
        s.set_state(u0)
        dev.sync()
        tm = dev.timer()
        tm.start()
        last = [time.perf_counter()]

        def progress(k, t, name=sensor):
            if time.perf_counter() - last[0] > 20.0:
                last[0] = time.perf_counter()
                log(f"{name}: {k} steps, t = {t:.4f}")

        steps = (s.run(a.t_end, chunk=64, progress=progress) if a.c5_steps == 0
                 else (s.step(a.c5_steps, a.t_end) or a.c5_steps))
        tm.stop()
        ms = tm.ms()
        u = s.state()
        # sensor cost on the final (shocked) state
        ms_sensor_end = _time_ms(dev, s.sense, a.reps)
        runs[sensor] = {"u": u.reshape(-1)[sample], "u_full": u, "t": s.time()}
        out[sensor] = {"run_ms": ms, "steps": steps, "t_final": s.time(),
                       "sensor_ms_smooth": ms_sensor, "sensor_ms_shocked": ms_sensor_end,
                       "finite": bool(np.isfinite(u).all())}
        log(f"{sensor}: {steps} steps to t={s.time():.4f} in {ms / 1e3:.1f} s; sensor "
            f"{ms_sensor:.4f} ms (smooth) / {ms_sensor_end:.4f} ms (shocked)")
        del s
    from oracle import av_oracle as A
    ue = A.burgers_exact(lambda X_, Y_, Z_: 0.5 + np.sin(2 * np.pi * X_) * np.sin(2 * np.pi * Y_)
                         * np.sin(2 * np.pi * Z_), x.reshape(-1)[sample], y.reshape(-1)[sample],
                         z.reshape(-1)[sample], runs["pp"]["t"], umax=1.5)
    for sensor in ("pp", "mlp"):
        d = runs[sensor]["u"] - ue
        out[sensor]["err_vs_exact"] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ue)),
                                       "nodes": int(sample.size)}
    e_pp, e_ml = out["pp"]["err_vs_exact"]["rel_l2"], out["mlp"]["err_vs_exact"]["rel_l2"]
    out["mlp_vs_pp"] = {"err_ratio": e_ml / e_pp,
                        "state_rel_l2": _errs(runs["mlp"]["u_full"], runs["pp"]["u_full"])["rel_l2"],
                        "sensor_speedup_smooth": out["pp"]["sensor_ms_smooth"] / out["mlp"]["sensor_ms_smooth"],
                        "sensor_speedup_shocked": out["pp"]["sensor_ms_shocked"] / out["mlp"]["sensor_ms_shocked"]}
    out["device"] = dev.name
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=25,
                    help="untimed SSP-RK3 steps first; at least WARM_STAGE_MS of them run "
                         "whatever this says (the clocks ramp over the first ~25 ms of load)")
    ap.add_argument("--n", type=int, default=96, help="hexes per side")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--parts", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tile", type=int, default=0, help="elements per block (0 = tuned)")
    ap.add_argument("--bpc", type=int, default=0, help="blocks per CU (0 = tuned)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="stage kernel (0 auto, 1 VALU, 2 MFMA pipelined, 3 split, 4 MFMA)")
    ap.add_argument("--xcd", type=int, default=-1, help="1: XCD-swept tile order")
    ap.add_argument("--tile-order", type=int, default=-1,
                    help="1: tiles of all partitions in the Morton order of their centroids, "
                         "0: partition by partition (library default)")
    ap.add_argument("--morton", dest="morton", action="store_true", default=None,
                    help="renumber elements in Morton order within each partition "
                         "(default for the P=3 stage, off for --mode c4)")
    ap.add_argument("--no-morton", dest="morton", action="store_false",
                    help="keep the generator's element order")
    ap.add_argument("--cpu-n", type=int, default=48, help="hexes per side of the CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-p6", action="store_true",
                    help="leave the P=6 section (C4 stage and derivative) out of the stage line")
    ap.add_argument("--mode", choices=["stage", "deriv", "c4", "c5", "partition"], default="stage")
    ap.add_argument("--c5-n", type=int, default=64, help="hexes in x and y (c5)")
    ap.add_argument("--nz", type=int, default=85, help="hexes in z (c5)")
    ap.add_argument("--t-end", type=float, default=0.4, help="final time (c5)")
    ap.add_argument("--c5-steps", type=int, default=0, help="c5: fixed step count (0: run to t-end)")
    ap.add_argument("--train-n", type=int, default=16, help="c5: hexes per side of the training mesh")
    ap.add_argument("--train-seeds", type=int, default=32)
    ap.add_argument("--train-every", type=int, default=100)
    ap.add_argument("--train-steps", type=int, default=60000)
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--exact-nodes", type=int, default=20000)
    ap.add_argument("--case", choices=["c2", "c4"], default="c4", help="derivative workload")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--precs", default="", help="derivative precisions to run (f32,bf16x3,...)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (PMC calibration)")
    a = ap.parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}, WORLD_SIZE={world}); "
              "multi-GPU sharding is described in DESIGN.md but not built.", file=sys.stderr)
        return 2

    from dgkernel_amd import _lib, dg

    if a.mode == "partition":
        print(json.dumps(bench_partition(a)), flush=True)
        return 0
    if a.copy_calib:
        dev = _lib.Device()
        src, dst = dev.zeros(a.copy_calib // 8), dev.zeros(a.copy_calib // 8)
        for _ in range(3):
            src.copy_to(dst)
        dev.sync()
        print(json.dumps({"copy_calib_bytes": a.copy_calib}), flush=True)
        return 0

    if a.mode == "deriv":
        print(json.dumps(bench_deriv(a, _lib.Device())), flush=True)
        return 0
    if a.mode == "c4":
        print(json.dumps(bench_c4_stage(a, _lib.Device())), flush=True)
        return 0
    if a.mode == "c5":
        print(json.dumps(bench_c5(a, _lib.Device())), flush=True)
        return 0

    t_setup = time.perf_counter()
    morton = a.morton is None or a.morton
    m, st, u0 = build_c3(a.n, a.order, a.parts, a.seed, morton=morton)
    log(f"host setup {time.perf_counter() - t_setup:.1f} s, K={m.K}, parts={st.npart}")
    dev = _lib.Device()
    solver = dg.BurgersSolver(dev, st, cfl=0.3)
    if a.variant >= 0:
        solver.mesh.set_variant(a.variant)
    if a.xcd >= 0:
        solver.mesh.set_order(bool(a.xcd))
    if a.tile_order >= 0:
        solver.mesh.set_tile_order(bool(a.tile_order))
    if a.tile or a.bpc:
        solver.mesh.set_tile(a.tile or 64, a.bpc)
    if a.variant >= 0 or a.tile or a.bpc or a.xcd >= 0 or a.tile_order >= 0:
        e = dict(solver.tuned or {})
        e.update({k: v for k, v in (("variant", a.variant), ("tile", a.tile), ("bpc", a.bpc),
                                     ("xcd", a.xcd)) if v not in (-1, 0) or k == "variant" and v == 0})
        if a.tile_order >= 0:
            e["spatial"] = a.tile_order
        solver.tuned = e
    solver.set_state(u0)
    del u0
    dev.sync()
    log(f"device setup done {time.perf_counter() - t_setup:.1f} s, hmin={solver.hmin:.6g}")
    K, Np, Nfp = m.K, st.ref.Np, st.ref.Nfp

    t_w = time.perf_counter()
    warm_steps = _warm(dev, lambda: solver.step(1, graph=False), min_calls=max(1, a.warmup),
                       min_ms=WARM_STAGE_MS)
    warm_ms = (time.perf_counter() - t_w) * 1e3
    log(f"warm-up: {warm_steps} steps, {warm_ms:.1f} ms")

    # Timed region: eager launches, one event pair per stage launch.
    timers = [dev.timer() for _ in range(3 * a.steps)]
    region = dev.timer()
    import ctypes as C
    L = _lib
    msh = solver.mesh.h
    dev.sync()
    region.start()
    for s in range(a.steps):
        L.call("dgk_cfl_dt", dev.h, solver.maxabs_ptr, solver.cfl, solver.hmin, solver.dt_ptr)
        tm = timers[3 * s:3 * s + 3]
        tm[0].start()
        L.call("dgk_burgers_stage", msh, solver.u.ptr, None, solver.u1.ptr, 0.0, 1.0, solver.dt_ptr,
               None)
        tm[0].stop()
        tm[1].start()
        L.call("dgk_burgers_stage", msh, solver.u1.ptr, solver.u.ptr, solver.u2.ptr, 0.75, 0.25,
               solver.dt_ptr, None)
        tm[1].stop()
        tm[2].start()
        L.call("dgk_burgers_stage", msh, solver.u2.ptr, solver.u.ptr, solver.u.ptr, 1.0 / 3.0,
               2.0 / 3.0, solver.dt_ptr, solver.maxabs_ptr)
        tm[2].stop()
    region.stop()
    dev.sync()
    total_ms = region.ms()
    stage_ms = np.array([t.ms() for t in timers])
    log(f"timed {a.steps} steps: {total_ms:.3f} ms, stage mean {stage_ms.mean():.4f} ms "
        f"(min {stage_ms.min():.4f}, max {stage_ms.max():.4f})")

    graph_ms = None
    if not a.no_graph:
        solver.step(1, graph=True)
        dev.sync()
        g = dev.timer()
        g.start()
        solver.step(a.steps, graph=True)
        g.stop()
        graph_ms = g.ms()
        log(f"graph replay {a.steps} steps: {graph_ms:.3f} ms")

    u_end = solver.state()
    finite = bool(np.isfinite(u_end).all())
    dof_stage = K * Np * 3 * a.steps
    value = dof_stage / (total_ms * 1e-3)
    bpe = stage_bytes_per_elem(Np)
    achieved = bpe * K / (stage_ms.mean() * 1e-3) / 1e9
    flops = stage_flops_per_elem(Np, Nfp) * K / (stage_ms.mean() * 1e-3) / 1e12
    prof = f"burgers_stage_p{a.order}_{a.n}cube_pmc.md"
    traffic = read_traffic(prof, "traffic_bytes_per_launch" if morton
                           else "traffic_bytes_per_launch_generator_order")
    # The traffic is not counted in this run (hardware counters need their own
    # rocprofv3 --pmc runs); it comes from the profile, which records the
    # hash of the stage source it was measured on.
    src = stage_source_hash()
    traffic_src = read_profile_key(prof, "stage_source_sha256")

    kname = stage_kernel_name(a.order, solver)
    walk = "morton across partitions" if (solver.tuned or {}).get("spatial") else "partition order"
    del solver
    p6 = None
    if not a.no_p6:
        p6 = p6_section(dev, a.seed)

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(a.order, a.cpu_n, a.cpu_steps, a.cpu_threads or None)
            log(f"cpu baseline {cpu['value']:.3e} DOF-stage/s on {cpu['cores']} threads "
                f"(partition schedule); element-parallel {cpu['element_parallel']['value']:.3e}")
        except Exception as e:  # the oracle library may be absent on a bare install
            log(f"cpu baseline skipped: {e}")

    out = {
        "metric": "DOF-stage-updates/s at P=3 and P=6, % HBM/MFMA roofline, L2 err vs FP64",
        "value": value,
        "unit": "DOF-stage-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "warmup_run": {"steps": warm_steps, "ms": warm_ms},
        "ms_per_step": total_ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (smooth initial field on a generated periodic mesh)",
        "config": {"workload": f"batched Burgers SSP-RK3 stages, {a.n}^3x6={K} tets, P={a.order}, "
                               f"{st.npart} slab partitions (+-40% imbalance), one persistent launch"
                               + (", elements in Morton order within each partition" if morton
                                  else ""),
                   "tets": K, "order": a.order, "Np": Np, "partitions": st.npart,
                   "kpart_max": int(st.Kp.max()), "stages_timed": 3 * a.steps,
                   "element_order": "morton" if morton else "generator",
                   "tile_walk": walk},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": {"profile": f"profiles/{prof}",
                                        "measured_on_this_source": traffic_src == src,
                                        "stage_source_sha256": src},
                     "kernel": kname,
                     "bytes_per_elem": bpe,
                     "stage_ms_mean": float(stage_ms.mean()),
                     "fp64_tflops": flops, "fp64_frac": flops / FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "graph_ms_per_step": None if graph_ms is None else graph_ms / a.steps,
        "graph_value": None if graph_ms is None else dof_stage / (graph_ms * 1e-3),
        "state_finite": finite,
        "p6": p6,
        "device": dev.name,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
