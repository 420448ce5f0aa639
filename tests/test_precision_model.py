"""CPU model of the below-FP64 contractions (no GPU): what a split that lost
one of its parts would do to the error.

The device forms w = f - Lf per element (Lf: the linear interpolant through
the four vertex nodes, removed in FP64), rounds w to FP32 and splits it and
the operator into bf16 parts (hi = bf16(x), lo = bf16(x - hi)); BF16x3 sums
hi*hi + hi*lo + lo*hi with FP32 accumulation, and the exact slopes of Lf are
added back. This model does the same products in FP64 (the FP32
accumulation adds ~1e-7, far below the BF16x3 error), so it isolates operand
rounding. Planted mistakes -- a dropped lo part of either operand, or plain
BF16 -- must raise the error by well over the 10x margin the GPU stage bounds
keep above the measured BF16x3 error (tests/test_gpu_split.py)."""
import numpy as np
import pytest

from dgkernel_amd import dg
from dgkernel_amd import mesh as M


def bf16(x):
    """Round to bfloat16 (round to nearest even on the FP32 bit pattern)."""
    b = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return b.view(np.float32).astype(np.float64)


def test_bf16_rounding_known_values():
    """1 + 2^-8 is a tie (rounds to even 1), 1 + 3*2^-8 rounds up to 1 + 2^-6."""
    assert bf16(1.0 + 2.0 ** -8)[()] == 1.0
    assert bf16(1.0 + 3 * 2.0 ** -8)[()] == 1.0 + 2.0 ** -6
    x = np.linspace(-3, 3, 101)
    assert np.all(np.abs(bf16(x) - x) <= 2.0 ** -8 * np.abs(x) + 1e-300)


def _model_errors(N):
    m = M.cube_mesh(4, 4, 4, periodic=True)
    st = dg.make_setup(m, N)
    x, y, z = st.node_coords()
    u = 0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
    f = 0.5 * u * u
    ref = st.ref
    lam = np.stack([-(1 + ref.r + ref.s + ref.t) / 2, (1 + ref.r) / 2, (1 + ref.s) / 2,
                    (1 + ref.t) / 2], axis=1)
    vtx = [int(np.argmax(lam[:, k])) for k in range(4)]
    Lf = f[:, vtx] @ lam.T
    w32 = (f - Lf).astype(np.float32).astype(np.float64)
    out = {}
    for name, D in (("Dr", ref.Dr), ("Ds", ref.Ds), ("Dt", ref.Dt)):
        exact = f @ D.T
        slope = Lf @ D.T
        Dh, wh = bf16(D), bf16(w32)
        Dl, wl = bf16(D - Dh), bf16(w32 - wh)
        models = {
            "bf16x3": wh @ Dh.T + wl @ Dh.T + wh @ Dl.T,
            "no_w_lo": wh @ Dh.T + wh @ Dl.T,
            "no_D_lo": wh @ Dh.T + wl @ Dh.T,
            "bf16x1": wh @ Dh.T,
        }
        for k, v in models.items():
            e = float(np.linalg.norm(v + slope - exact) / np.linalg.norm(exact))
            out[k] = max(out.get(k, 0.0), e)
    return out


@pytest.mark.parametrize("N", [3, 5, 6])
def test_dropped_split_part_exceeds_margin(N):
    e = _model_errors(N)
    assert e["bf16x3"] > 0.0
    for mutant in ("no_w_lo", "no_D_lo", "bf16x1"):
        assert e[mutant] >= 30.0 * e["bf16x3"], (mutant, e)
