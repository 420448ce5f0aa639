"""Shared builders for the tests: meshes, initial data, pooled layouts."""
from __future__ import annotations

import numpy as np

from dgkernel_amd import dg
from dgkernel_amd import mesh as M
from tests.conftest import GOLDEN


def smooth_u0(st: dg.DGSetup, amp: float = 0.3, noise: float = 1e-3, seed: int = 1):
    """0.5 + amp sin(2 pi x) sin(2 pi y) sin(2 pi z) + U(-noise, noise) per node."""
    x, y, z = st.node_coords()
    u = 0.5 + amp * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    if noise:
        u = u + M.uniform(seed, 7, x.size, -noise, noise).reshape(x.shape)
    return u


def neu_mesh():
    return M.read_gambit_neutral(str(GOLDEN / "cube-partitioned.neu"))


def pooled(values_by_part, K, per_elem, alignment=dg.WARP_ALIGN):
    """Pack per-partition [K_p, per_elem] arrays into the runner pooled layout."""
    offs, nbytes = dg.aligned_offsets(np.asarray(K), per_elem, 8, alignment)
    out = np.zeros(nbytes // 8)
    for p, v in enumerate(values_by_part):
        out[offs[p]:offs[p] + v.size] = v.reshape(-1)
    return out, offs


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
