"""Shock capturing on the device: the Persson-Peraire modal-decay sensor,
BR1 artificial viscosity, and the learned sensor (a 64-64 MLP over element
modal coefficients) with its GPU training loop, all through libdgk's C ABI.

The reference has no shock capturing; this module implements the
learned-viscosity workload around the tet DG Burgers path:
    u_t + div(u^2/2 (1,1,1)) = div(eps grad u),  eps constant per element,
with eps from the sensor of u at the start of each SSP-RK3 step (frozen over
the step) and dt = min(cfl hmin / max|u|, cfl_v hmin^2 / max eps).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .dg import DGMesh, DGSetup
from .mesh import uniform


# Viscous step factors: about 0.8 of SSP-RK3's real-axis stability limit
# (2.51) over the spectral radius of L_v / eps in units of 1/hmin^2, measured
# by power iteration on periodic Kuhn meshes with tau0 = 1:
# rho hmin^2 = 130, 82, 61, 8.2 for N = 1..4 (hmin is the minimum node
# distance, whose ratio to h drops sharply at N = 4 for the NUDG nodes).
CFL_V = {1: 0.015, 2: 0.025, 3: 0.033, 4: 0.25}


@dataclass
class AVParams:
    """Sensor and viscosity parameters (Persson & Peraire, AIAA 2006-112)."""

    s0: float            # log10 threshold of the top-mode energy fraction
    kappa: float = 1.0   # half width of the sine ramp (decades)
    ceps: float = 0.5    # eps0 = ceps * h * sqrt(3) max|u| / N
    tau0: float = 1.0    # jump penalty tau = tau0 * max(eps-, eps+) (per unit Fscale)
    cfl_v: float = 0.02  # viscous step limit cfl_v * hmin^2 / max eps

    @classmethod
    def for_order(cls, N: int, **kw) -> "AVParams":
        kw.setdefault("cfl_v", CFL_V[N])
        return cls(s0=-4.0 * math.log10(N), **kw)

    def c_struct(self) -> _lib.AVParams:
        return _lib.AVParams(self.s0, self.kappa, self.ceps, self.tau0)


def seeded_field(st: DGSetup, seed: int, amp: tuple[float, float] = (0.7, 1.1),
                 mean: tuple[float, float] = (0.3, 0.7)) -> np.ndarray:
    """Smooth periodic initial field of a seeded family:
    m + a sin(2 pi (x + px)) sin(2 pi (y + py)) sin(2 pi (z + pz)),
    m ~ U(mean), a ~ U(amp), phases ~ U(0, 1) from (seed, stream 11)."""
    r = uniform(seed, 11, 5)
    m = mean[0] + (mean[1] - mean[0]) * r[0]
    a = amp[0] + (amp[1] - amp[0]) * r[1]
    x, y, z = st.node_coords()
    return m + a * (np.sin(2 * np.pi * (x + r[2])) * np.sin(2 * np.pi * (y + r[3]))
                    * np.sin(2 * np.pi * (z + r[4])))


class SensorNet:
    """The sensor network nin -> 64 -> 64 -> 1 (tanh, tanh, sigmoid) on the
    device (dgk_mlp). Parameters are FP32; inference runs on BF16 matrix
    cores, training in FP32 with Adam."""

    H = 64

    def __init__(self, dev: _lib.Device, nin: int, params: np.ndarray | None = None,
                 seed: int = 0):
        self.dev, self.nin = dev, nin
        if params is None:
            params = self.init_params(nin, seed)
        params = np.ascontiguousarray(params, dtype=np.float32)
        h = C.c_void_p()
        _lib.call("dgk_mlp_create", dev.h, nin, params.ctypes.data_as(_lib._pf32), C.byref(h))
        self.h = h
        n = C.c_int()
        _lib.call("dgk_mlp_num_params", h, C.byref(n))
        self.nparams = n.value
        self._loss = dev.zeros(4, np.float32)

    @staticmethod
    def layout(nin: int) -> dict:
        H = SensorNet.H
        o = {"W1": (0, (H, nin))}
        o["b1"] = (H * nin, (H,))
        o["W2"] = (H * nin + H, (H, H))
        o["b2"] = (H * nin + H + H * H, (H,))
        o["w3"] = (H * nin + 2 * H + H * H, (H,))
        o["b3"] = (H * nin + 3 * H + H * H, (1,))
        return o

    @classmethod
    def init_params(cls, nin: int, seed: int) -> np.ndarray:
        """Glorot-uniform weights, zero biases, from (seed, streams 21-23)."""
        H = cls.H
        L = cls.layout(nin)
        p = np.zeros(L["b3"][0] + 1, dtype=np.float32)

        def glorot(stream, fan_in, fan_out, n):
            lim = math.sqrt(6.0 / (fan_in + fan_out))
            return uniform(seed, stream, n, -lim, lim).astype(np.float32)

        p[0:H * nin] = glorot(21, nin, H, H * nin)
        o = L["W2"][0]
        p[o:o + H * H] = glorot(22, H, H, H * H)
        o = L["w3"][0]
        p[o:o + H] = glorot(23, H, 1, H)
        return p

    def get(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _lib.call("dgk_mlp_get", self.h, out.ctypes.data_as(_lib._pf32))
        return out

    def set(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        _lib.call("dgk_mlp_set", self.h, params.ctypes.data_as(_lib._pf32))

    def forward(self, x: _lib.DeviceBuffer, n: int, y: _lib.DeviceBuffer) -> None:
        _lib.call("dgk_mlp_forward", self.h, x.ptr, n, y.ptr)

    def grad(self, x: _lib.DeviceBuffer, t: _lib.DeviceBuffer, n: int) -> tuple[np.ndarray, float]:
        g = self.dev.zeros(self.nparams, np.float32)
        _lib.call("dgk_mlp_grad", self.h, x.ptr, t.ptr, n, g.ptr, self._loss.ptr)
        return g.to_host(), float(self._loss.to_host()[0])

    def train(self, x: _lib.DeviceBuffer, t: _lib.DeviceBuffer, ndata: int, steps: int,
              batch: int = 4096, lr: float = 2e-3, seed: int = 1, report_every: int = 0,
              lr_final: float | None = None) -> list[float]:
        """Adam steps on batches drawn from the ndata rows; returns the mean
        batch loss of each report window (or of the whole run)."""
        hist = []
        self._loss.fill0()
        window = report_every or steps
        for k in range(steps):
            lr_k = lr if lr_final is None else lr * (lr_final / lr) ** (k / max(1, steps - 1))
            _lib.call("dgk_mlp_train_step", self.h, x.ptr, t.ptr, ndata, batch, seed, lr_k,
                      self._loss.ptr)
            if (k + 1) % window == 0:
                hist.append(float(self._loss.to_host()[0]) / window)
                self._loss.fill0()
        return hist

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib.alive(self.dev):
                _lib.load().dgk_mlp_destroy(self.h)
        except Exception:
            pass


class AVSolver:
    """SSP-RK3 for Burgers with artificial viscosity on the device.

    sensor: "pp" (Persson-Peraire, FP64), "mlp" (SensorNet), or None (no
    viscosity; the plain stage kernels). Device scalars: [max|u|, dt, t, max eps].
    """

    def __init__(self, dev: _lib.Device, st: DGSetup, params: AVParams, sensor: str | None = "pp",
                 net: SensorNet | None = None, cfl: float = 0.3, hmin: float | None = None):
        if sensor not in ("pp", "mlp", None):
            raise ValueError(f"unknown sensor {sensor!r}")
        if sensor == "mlp" and net is None:
            raise ValueError("the learned sensor needs a SensorNet")
        self.dev, self.st, self.params, self.sensor, self.net = dev, st, params, sensor, net
        self.cfl = cfl
        self.mesh = DGMesh(dev, st)
        n = st.state_bytes // 8
        self.u, self.u1, self.u2 = dev.zeros(n), dev.zeros(n), dev.zeros(n)
        self.scal = dev.zeros(4)
        self.hmin = self.mesh.min_node_distance() if hmin is None else hmin
        self._Vinv = np.ascontiguousarray(st.ref.Vinv, dtype=np.float64)
        self._cp = params.c_struct()
        h = C.c_void_p()
        _lib.call("dgk_av_create", self.mesh.h, self._Vinv.ctypes.data_as(_lib._pf64),
                  C.byref(self._cp), C.byref(h))
        self.av = h
        e = C.c_void_p()
        _lib.call("dgk_av_eps_dev", h, C.byref(e))
        self._eps_ptr = e

    def _sp(self, i: int):
        return C.c_void_p(self.scal.ptr.value + 8 * i)

    def set_state(self, u_dense: np.ndarray) -> None:
        self.u.copy_from(self.st.to_pooled(u_dense))
        self.scal.fill0()
        _lib.call("dgk_max_abs", self.mesh.h, self.u.ptr, self._sp(0))

    def state(self) -> np.ndarray:
        return self.st.from_pooled(self.u.to_host()).reshape(self.st.K, self.st.ref.Np)

    def time(self) -> float:
        return float(self.scal.to_host()[2])

    def eps(self) -> np.ndarray:
        out = np.empty(self.st.K)
        _lib.call("dgk_copy_d2h", self.dev.h, out.ctypes.data_as(C.c_void_p), self._eps_ptr,
                  C.c_size_t(out.nbytes), C.c_size_t(0))
        return out

    def sense(self, u=None, frac=None, frac32=None, feats=None) -> None:
        """Sensor of u (default: the state) into eps, max eps into scal[3]."""
        u = self.u if u is None else u
        if self.sensor == "pp":
            _lib.call("dgk_av_sensor_pp", self.av, u.ptr, self._sp(3),
                      None if frac is None else frac.ptr, None if frac32 is None else frac32.ptr,
                      None if feats is None else feats.ptr)
        elif self.sensor == "mlp":
            _lib.call("dgk_av_sensor_mlp", self.av, self.net.h, u.ptr, self._sp(3),
                      None if frac is None else frac.ptr)

    def _enqueue_step(self, t_end: float) -> None:
        self.sense()
        p = self.params
        _lib.call("dgk_av_cfl_dt", self.dev.h, self.scal.ptr, self.cfl, p.cfl_v, self.hmin, t_end)
        dt = self._sp(1)
        if self.sensor is None:
            m = self.mesh.h
            _lib.call("dgk_burgers_stage", m, self.u.ptr, None, self.u1.ptr, 0.0, 1.0, dt, None)
            _lib.call("dgk_burgers_stage", m, self.u1.ptr, self.u.ptr, self.u2.ptr, 0.75, 0.25, dt,
                      None)
            _lib.call("dgk_burgers_stage", m, self.u2.ptr, self.u.ptr, self.u.ptr, 1.0 / 3.0,
                      2.0 / 3.0, dt, self._sp(0))
            return
        a = self.av
        _lib.call("dgk_av_stage", a, self.u.ptr, None, self.u1.ptr, 0.0, 1.0, dt, None)
        _lib.call("dgk_av_stage", a, self.u1.ptr, self.u.ptr, self.u2.ptr, 0.75, 0.25, dt, None)
        _lib.call("dgk_av_stage", a, self.u2.ptr, self.u.ptr, self.u.ptr, 1.0 / 3.0, 2.0 / 3.0, dt,
                  self._sp(0))

    def step(self, n: int = 1, t_end: float = 0.0) -> None:
        for _ in range(n):
            self._enqueue_step(t_end)

    def run(self, t_end: float, chunk: int = 32, max_steps: int = 1_000_000,
            progress=None) -> int:
        """Steps until t reaches t_end (the last dt is clipped); returns the
        steps enqueued (whole chunks; steps after t_end are no-ops, dt = 0).
        progress(steps, t) is called after every chunk if given."""
        steps = 0
        while steps < max_steps:
            self.step(chunk, t_end)
            steps += chunk
            t = self.time()
            if progress is not None:
                progress(steps, t)
            if t >= t_end * (1 - 1e-15):
                break
        return steps

    def rhs_total(self, u_dense: np.ndarray) -> np.ndarray:
        """L(u) + L_v(u; eps) with the current eps (testing aid)."""
        self.u1.copy_from(self.st.to_pooled(u_dense))
        _lib.call("dgk_av_rhs_total", self.av, self.u1.ptr, self.u2.ptr)
        return self.st.from_pooled(self.u2.to_host()).reshape(self.st.K, self.st.ref.Np)

    def rhs_visc(self, u_dense: np.ndarray) -> np.ndarray:
        """L_v(u; eps) alone (testing aid)."""
        self.u1.copy_from(self.st.to_pooled(u_dense))
        _lib.call("dgk_av_rhs", self.av, self.u1.ptr, self.u2.ptr)
        return self.st.from_pooled(self.u2.to_host()).reshape(self.st.K, self.st.ref.Np)

    def set_eps(self, eps: np.ndarray) -> None:
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        _lib.call("dgk_copy_h2d", self.dev.h, self._eps_ptr, eps.ctypes.data_as(C.c_void_p),
                  C.c_size_t(eps.nbytes), C.c_size_t(0))

    def __del__(self):
        try:
            if getattr(self, "av", None) and _lib.alive(self.dev):
                _lib.load().dgk_av_destroy(self.av)
        except Exception:
            pass


def generate_training_data(dev: _lib.Device, st: DGSetup, params: AVParams, seeds, t_end: float,
                           every: int, cfl: float = 0.3, max_rows: int | None = None):
    """Persson-Peraire runs from seeded fields; every `every` steps each
    element's normalised modal coefficients and sensor fraction become one
    (input, target) row. Returns (X [n, Np] FP32, T [n] FP32) device buffers
    and n."""
    K, Np = st.K, st.ref.Np
    solver = AVSolver(dev, st, params, "pp", cfl=cfl)
    # rows per run: samples at steps 0, every, 2 every, ... until t_end
    chunks = []
    for sd in seeds:
        solver.set_state(seeded_field(st, sd))
        feats = dev.zeros(K * Np, np.float32)
        frac = dev.zeros(K, np.float32)
        k = 0
        while solver.time() < t_end * (1 - 1e-15):
            if k % every == 0:
                f = dev.zeros(K * Np, np.float32)
                fr = dev.zeros(K, np.float32)
                solver.sense(frac32=fr, feats=f)
                chunks.append((f, fr))
            solver.step(1, t_end)
            k += 1
        del feats, frac
    n = K * len(chunks)
    if max_rows is not None:
        n = min(n, max_rows)
    X = dev.empty(n * Np, np.float32)
    T = dev.empty(n, np.float32)
    o = 0
    for f, fr in chunks:
        take = min(K, n - o)
        if take <= 0:
            break
        _lib.call("dgk_copy_d2d", dev.h, C.c_void_p(X.ptr.value + 4 * o * Np), f.ptr,
                  C.c_size_t(4 * take * Np))
        _lib.call("dgk_copy_d2d", dev.h, C.c_void_p(T.ptr.value + 4 * o), fr.ptr,
                  C.c_size_t(4 * take))
        o += take
    dev.sync()
    return X, T, n
