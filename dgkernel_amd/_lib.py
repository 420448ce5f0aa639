"""ctypes binding of libdgk.so (the C ABI declared in include/dgk.h).

The shared library is built in-tree by ``python -m dgkernel_amd.build`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: the
package has no CPU fallback for any device entry point.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libdgk.so"

_vp = C.c_void_p
_i32, _i64, _f64, _sz = C.c_int, C.c_int64, C.c_double, C.c_size_t
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)
_pf64 = C.POINTER(C.c_double)
_pf32 = C.POINTER(C.c_float)

DGK_OK = 0
DGK_F64, DGK_F32 = 0, 1
DGK_PREC_F64, DGK_PREC_F32, DGK_PREC_BF16X3, DGK_PREC_BF16X6, DGK_PREC_I8X6 = 0, 1, 2, 3, 4
DGK_PREC_LIFT_F64 = 16  # flag: LIFT product in FP64 whatever the D precision
ARG_MEM, ARG_F64, ARG_F32, ARG_I64, ARG_I32 = 0, 1, 2, 3, 4


class DgkError(RuntimeError):
    """A libdgk call returned a nonzero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed (status {code}): {msg}")
        self.code = code


class _ArgVal(C.Union):
    _fields_ = [("mem", C.c_void_p), ("f64", C.c_double), ("f32", C.c_float), ("i64", C.c_int64),
                ("i32", C.c_int32)]


class Arg(C.Structure):
    _fields_ = [("kind", C.c_int), ("pad", C.c_int), ("v", _ArgVal)]


class AVParams(C.Structure):
    _fields_ = [("s0", C.c_double), ("kappa", C.c_double), ("ceps", C.c_double),
                ("tau0", C.c_double)]


class DGMeshDesc(C.Structure):
    _fields_ = [("N", C.c_int), ("npart", C.c_int), ("K", _pi64), ("state_off", _pi64),
                ("Dr", _pf64), ("Ds", _pf64), ("Dt", _pf64), ("LIFT", _pf64),
                ("Fmask", C.POINTER(C.c_int)), ("r", _pf64), ("s", _pf64), ("t", _pf64),
                ("face_nodes", C.POINTER(C.c_int)), ("EToE", _pi64),
                ("codes", C.POINTER(C.c_int)), ("vtx", _pf64)]


# name -> (restype, argtypes)
_SIGS = {
    "dgk_last_error": (C.c_char_p, []),
    "dgk_version": (C.c_char_p, []),
    "dgk_device_create": (_i32, [C.c_char_p, C.POINTER(_vp)]),
    "dgk_device_destroy": (_i32, [_vp]),
    "dgk_device_sync": (_i32, [_vp]),
    "dgk_device_mode": (C.c_char_p, [_vp]),
    "dgk_device_info": (_i32, [_vp, C.c_char_p, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32),
                               C.POINTER(_i32)]),
    "dgk_malloc": (_i32, [_vp, _sz, _vp, C.POINTER(_vp)]),
    "dgk_free": (_i32, [_vp, _vp]),
    "dgk_copy_h2d": (_i32, [_vp, _vp, _vp, _sz, _sz]),
    "dgk_copy_d2h": (_i32, [_vp, _vp, _vp, _sz, _sz]),
    "dgk_copy_d2d": (_i32, [_vp, _vp, _vp, _sz]),
    "dgk_fill_bytes": (_i32, [_vp, _vp, _i32, _sz]),
    "dgk_kernel_list": (_i32, [C.c_char_p, _sz]),
    "dgk_kernel_build_from_source": (_i32, [_vp, C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "dgk_kernel_describe": (_i32, [C.c_char_p, C.c_char_p, C.c_char_p, _sz]),
    "dgk_kernel_launch": (_i32, [_vp, _vp, _i32, C.POINTER(Arg)]),
    "dgk_kernel_free": (_i32, [_vp]),
    "dgk_graph_begin": (_i32, [_vp]),
    "dgk_graph_end": (_i32, [_vp, C.POINTER(_vp)]),
    "dgk_graph_launch": (_i32, [_vp]),
    "dgk_graph_destroy": (_i32, [_vp]),
    "dgk_timer_create": (_i32, [_vp, C.POINTER(_vp)]),
    "dgk_timer_start": (_i32, [_vp]),
    "dgk_timer_stop": (_i32, [_vp]),
    "dgk_timer_elapsed_ms": (_i32, [_vp, C.POINTER(_f64)]),
    "dgk_timer_destroy": (_i32, [_vp]),
    "dgk_phys_deriv_nodal": (_i32, [_vp, _i32, _i32, _pi64, _vp, _vp, _vp, _vp, _vp,
                                    C.POINTER(_vp), _i32, _vp, _vp, _vp, _vp, _i32, _vp]),
    "dgk_phys_deriv_nodal_noalloc": (_i32, [_vp, _i32, _i32, _pi64, _vp, _vp, _vp, _vp, _vp,
                                            C.POINTER(_vp), _i32, _vp, _vp, _vp, _vp, _i32, _vp]),
    "dgk_deriv_plan_create": (_i32, [_vp, _i32, _pf64, _pf64, _pf64, _pf64, _pf64, _pf64, _i32,
                                     C.POINTER(_vp)]),
    "dgk_deriv_plan_destroy": (_i32, [_vp]),
    "dgk_phys_deriv_affine": (_i32, [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp]),
    "dgk_matmul": (_i32, [_vp, _i32, _i32, _vp, _i32, _pi64, _vp, _vp, _vp, _vp, _i32]),
    "dgk_face_plan_create": (_i32, [_vp, _i32, _pi64, _pi32, _pi32, C.POINTER(_vp)]),
    "dgk_partition_kway": (_i32, [_i64, _pi64, _pi64, _pi64, _pi64, _i32, _f64, _i32, C.c_uint64,
                                  _pi32, _pi64]),
    "dgk_face_plan_destroy": (_i32, [_vp]),
    "dgk_face_exchange": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "dgk_dgmesh_create": (_i32, [_vp, C.POINTER(DGMeshDesc), C.POINTER(_vp)]),
    "dgk_dgmesh_destroy": (_i32, [_vp]),
    "dgk_dgmesh_geometry": (_i32, [_vp, _pf64]),
    "dgk_dgmesh_geometry_dev": (_i32, [_vp, C.POINTER(_vp)]),
    "dgk_dgmesh_min_node_distance": (_i32, [_vp, _pf64]),
    "dgk_dgmesh_trace_plus": (_i32, [_vp, _vp, _vp]),
    "dgk_burgers_stage": (_i32, [_vp, _vp, _vp, _vp, _f64, _f64, _vp, _vp]),
    "dgk_burgers_rhs": (_i32, [_vp, _vp, _vp]),
    "dgk_max_abs": (_i32, [_vp, _vp, _vp]),
    "dgk_cfl_dt": (_i32, [_vp, _vp, _f64, _f64, _vp]),
    "dgk_burgers_set_tile": (_i32, [_vp, _i32, _i32]),
    "dgk_burgers_set_variant": (_i32, [_vp, _i32]),
    "dgk_burgers_set_precision": (_i32, [_vp, _i32]),
    "dgk_burgers_set_order": (_i32, [_vp, _i32]),
    "dgk_dgmesh_set_tile_order": (_i32, [_vp, _i32]),
    "dgk_av_create": (_i32, [_vp, _pf64, C.POINTER(AVParams), C.POINTER(_vp)]),
    "dgk_av_destroy": (_i32, [_vp]),
    "dgk_av_eps_dev": (_i32, [_vp, C.POINTER(_vp)]),
    "dgk_av_sensor_pp": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "dgk_av_sensor_mlp": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "dgk_av_rhs": (_i32, [_vp, _vp, _vp]),
    "dgk_av_rhs_total": (_i32, [_vp, _vp, _vp]),
    "dgk_av_stage": (_i32, [_vp, _vp, _vp, _vp, _f64, _f64, _vp, _vp]),
    "dgk_av_cfl_dt": (_i32, [_vp, _vp, _f64, _f64, _f64, _f64]),
    "dgk_mlp_create": (_i32, [_vp, _i32, _pf32, C.POINTER(_vp)]),
    "dgk_mlp_destroy": (_i32, [_vp]),
    "dgk_mlp_num_params": (_i32, [_vp, C.POINTER(_i32)]),
    "dgk_mlp_get": (_i32, [_vp, _pf32]),
    "dgk_mlp_set": (_i32, [_vp, _pf32]),
    "dgk_mlp_forward": (_i32, [_vp, _vp, _i64, _vp]),
    "dgk_mlp_grad": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "dgk_mlp_train_step": (_i32, [_vp, _vp, _vp, _i64, _i32, C.c_uint64, C.c_float, _vp]),
}

_lib = None
_finalizing = False


def _mark_finalizing() -> None:
    global _finalizing
    _finalizing = True


atexit.register(_mark_finalizing)


def alive(dev: "Device | None" = None) -> bool:
    """True while device calls are allowed (not at interpreter exit, device open)."""
    return (not _finalizing) and (dev is None or bool(dev.h))


def load() -> C.CDLL:
    """Load libdgk.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m dgkernel_amd.build` "
                "(the package has no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    return [n for n in _SIGS if hasattr(load(), n)]


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != DGK_OK:
        raise DgkError(name, rc, lib.dgk_last_error().decode())


def last_error() -> str:
    return load().dgk_last_error().decode()


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Device:
    """A GPU device handle with one HIP stream (``gocca.NewDevice`` replacement)."""

    def __init__(self, props: str = '{"mode": "HIP", "device_id": 0}'):
        h = C.c_void_p()
        call("dgk_device_create", props.encode(), C.byref(h))
        self.h = h
        name = C.create_string_buffer(256)
        cu, lds, ma, mi = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        call("dgk_device_info", self.h, name, C.byref(cu), C.byref(lds), C.byref(ma), C.byref(mi))
        self.name = name.value.decode()
        self.cu_count = cu.value
        self.lds_bytes = lds.value

    def mode(self) -> str:
        return load().dgk_device_mode(self.h).decode()

    def sync(self) -> None:
        call("dgk_device_sync", self.h)

    def free(self) -> None:
        if self.h and alive():
            call("dgk_device_destroy", self.h)
        self.h = None

    def malloc(self, nbytes: int, init: np.ndarray | None = None) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes, init)

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(self, a.nbytes, a)
        b.dtype, b.shape = a.dtype, a.shape
        return b

    def empty(self, shape, dtype=np.float64) -> "DeviceBuffer":
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        b = DeviceBuffer(self, n, None)
        b.dtype, b.shape = np.dtype(dtype), tuple(np.atleast_1d(shape))
        return b

    def zeros(self, shape, dtype=np.float64) -> "DeviceBuffer":
        b = self.empty(shape, dtype)
        b.fill0()
        return b

    def timer(self) -> "Timer":
        return Timer(self)


class DeviceBuffer:
    """Owned device allocation (``OCCAMemory`` replacement)."""

    def __init__(self, dev: Device, nbytes: int, init: np.ndarray | None):
        self.dev = dev
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(np.float64) if init is None else init.dtype
        self.shape = (self.nbytes // self.dtype.itemsize,) if init is None else init.shape
        p = C.c_void_p()
        call("dgk_malloc", dev.h, C.c_size_t(self.nbytes),
             ptr(np.ascontiguousarray(init)) if init is not None else None, C.byref(p))
        self.p = p

    @property
    def ptr(self):
        return self.p

    def copy_from(self, a: np.ndarray, offset_bytes: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes + offset_bytes > self.nbytes:
            raise ValueError("copy exceeds allocation")
        call("dgk_copy_h2d", self.dev.h, self.p, ptr(a), C.c_size_t(a.nbytes), C.c_size_t(offset_bytes))

    def to_host(self, dtype=None, shape=None, offset_bytes: int = 0, count: int | None = None):
        dtype = np.dtype(dtype or self.dtype)
        if count is None:
            count = (self.nbytes - offset_bytes) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        call("dgk_copy_d2h", self.dev.h, ptr(out), self.p, C.c_size_t(out.nbytes),
             C.c_size_t(offset_bytes))
        if shape is not None:
            out = out.reshape(shape)
        elif self.shape and int(np.prod(self.shape)) == count and dtype == self.dtype:
            out = out.reshape(self.shape)
        return out

    def fill0(self) -> None:
        call("dgk_fill_bytes", self.dev.h, self.p, 0, C.c_size_t(self.nbytes))

    def copy_to(self, other: "DeviceBuffer") -> None:
        call("dgk_copy_d2d", self.dev.h, other.p, self.p, C.c_size_t(min(self.nbytes, other.nbytes)))

    def free(self) -> None:
        if self.p:
            call("dgk_free", self.dev.h, self.p)
            self.p = None

    def __del__(self):
        try:
            if getattr(self, "p", None) and alive(self.dev):
                load().dgk_free(self.dev.h, self.p)
        except Exception:
            pass


class Timer:
    """HIP-event timer on the device stream."""

    def __init__(self, dev: Device):
        h = C.c_void_p()
        call("dgk_timer_create", dev.h, C.byref(h))
        self.h = h
        self.dev = dev

    def start(self):
        call("dgk_timer_start", self.h)

    def stop(self):
        call("dgk_timer_stop", self.h)

    def ms(self) -> float:
        v = C.c_double()
        call("dgk_timer_elapsed_ms", self.h, C.byref(v))
        return v.value

    def __del__(self):
        try:
            if alive(self.dev):
                load().dgk_timer_destroy(self.h)
        except Exception:
            pass


class Graph:
    """A captured HIP graph replayed on the device stream."""

    def __init__(self, dev: Device, fn):
        self.dev = dev
        call("dgk_graph_begin", dev.h)
        try:
            fn()
        finally:
            h = C.c_void_p()
            call("dgk_graph_end", dev.h, C.byref(h))
        self.h = h

    def launch(self):
        call("dgk_graph_launch", self.h)

    def __del__(self):
        try:
            if alive(self.dev):
                load().dgk_graph_destroy(self.h)
        except Exception:
            pass


def describe_kernel(source: str, name: str) -> str:
    """The statement program libdgk would run for kernel ``name`` of ``source``
    (parse only; needs no device)."""
    buf = C.create_string_buffer(1 << 16)
    call("dgk_kernel_describe", source.encode(), name.encode(), buf, C.c_size_t(len(buf)))
    return buf.value.decode()


class Kernel:
    """A kernel built from the runner's full OKL text (``BuildKernelFromString``
    replacement); ``launch`` is ``RunWithArgs`` with one value per signature
    parameter: DeviceBuffer (or a raw device pointer) for pointers, and
    (kind, value) pairs for scalars, e.g. (ARG_F64, 0.5)."""

    def __init__(self, dev: Device, source: str, name: str):
        h = C.c_void_p()
        call("dgk_kernel_build_from_source", dev.h, source.encode(), name.encode(), C.byref(h))
        self.h = h
        self.dev = dev
        self.name = name

    def launch(self, args, host_K: np.ndarray | None = None) -> None:
        arr = (Arg * len(args))()
        for i, a in enumerate(args):
            if isinstance(a, tuple):
                kind, val = a
                arr[i].kind = kind
                field = {ARG_F64: "f64", ARG_F32: "f32", ARG_I64: "i64", ARG_I32: "i32",
                         ARG_MEM: "mem"}[kind]
                setattr(arr[i].v, field, val)
            else:
                arr[i].kind = ARG_MEM
                arr[i].v.mem = a.ptr.value if isinstance(a, DeviceBuffer) else a
        hk = None
        if host_K is not None:
            host_K = np.ascontiguousarray(host_K)
            hk = ptr(host_K)
        call("dgk_kernel_launch", self.h, hk, len(args), arr)

    def free(self) -> None:
        if self.h and alive(self.dev):
            call("dgk_kernel_free", self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
