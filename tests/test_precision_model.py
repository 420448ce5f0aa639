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


# --- exact-integer split (the I8X6 plan) ------------------------------------

def _digits(X, T):
    """Balanced base-256 digits (int8 range) of int64 X, least significant first."""
    X = X.copy()
    out = []
    for _ in range(T - 1):
        d = X & 255
        d = np.where(d >= 128, d - 256, d)
        out.append(d)
        X = (X - d) >> 8
    out.append(X)
    return out


def _split(V, T):
    """Rows of V scaled by 2^-E (|v| < 2^E), rounded to 8T-2 bits, as digits."""
    m = np.max(np.abs(V), axis=1, keepdims=True)
    E = np.frexp(np.where(m > 0, m, 1.0))[1]
    X = np.rint(np.ldexp(V, 8 * T - 2 - E)).astype(np.int64)
    return _digits(X, T), E


def test_balanced_digits_exact_and_offset_form():
    """Digits reassemble exactly and lie in [-128, 127]; the device's offset
    form (bytes of X + 0x808080808080, each XOR 0x80) gives the same digits."""
    rng = np.random.default_rng(3)
    X = rng.integers(-(2 ** 46), 2 ** 46, 10000, dtype=np.int64)
    d = _digits(X, 6)
    assert all(np.all((x >= -128) & (x <= 127)) for x in d)
    assert np.array_equal(sum(x * 256 ** k for k, x in enumerate(d)), X)
    Z = X + 0x808080808080
    for k in range(6):
        byte = ((Z >> (8 * k)) & 255) ^ 0x80
        assert np.array_equal(np.where(byte >= 128, byte - 256, byte), d[k])


def _ozaki(A, W, thr=5, drop=None, SA=6, SW=6):
    Ad, ea = _split(A, SA)
    Wd, ew = _split(W.T, SW)
    Y = np.zeros((A.shape[0], W.shape[1]))
    for a in range(SA):
        for b in range(SW):
            if a + b < thr or (a, b) == drop:
                continue
            Y += (Ad[a] @ Wd[b].T).astype(np.float64) * 256.0 ** (a + b)
    return np.ldexp(Y, ea + ew.T - (8 * SA - 2) - (8 * SW - 2))


def test_integer_split_reaches_fp64_level_and_planted_drops_fail():
    """P=6 operator on the C4-like field (h = 1/64, U(+-1e-6) noise): the 21
    pairs of weight >= 5 reproduce D w to ~2e-11 of itself; dropping one kept pair,
    or cutting one weight lower, raises the error 20x to 1e4x."""
    N = 6
    m = M.cube_mesh(4, 4, 4, periodic=True, L=1 / 16)
    st = dg.make_setup(m, N)
    x, y, z = st.node_coords()
    rng = np.random.default_rng(0)
    u = (0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
         + rng.uniform(-1e-6, 1e-6, x.shape))
    ref = st.ref
    A = np.vstack([ref.Dr, ref.Ds, ref.Dt])
    lam = np.stack([-(1 + ref.r + ref.s + ref.t) / 2, (1 + ref.r) / 2, (1 + ref.s) / 2,
                    (1 + ref.t) / 2], axis=1)
    vtx = [int(np.argmax(lam[:, k])) for k in range(4)]
    W = (u - u[:, vtx] @ lam.T).T
    exact = A.astype(np.longdouble) @ W.astype(np.longdouble)
    def err(Y):
        return float(np.linalg.norm((Y - exact).astype(np.float64)) / np.linalg.norm(exact.astype(np.float64)))
    e = err(_ozaki(A, W))
    # relative to D w itself (the linear part is removed, so D w is a small
    # part of the derivative): ~2e-11
    assert e <= 1e-10, e
    assert err(_ozaki(A, W, drop=(5, 0))) >= 10 * e   # a lowest-weight pair
    assert err(_ozaki(A, W, drop=(3, 3))) >= 1e3 * e
    assert err(_ozaki(A, W, thr=6)) >= 30 * e
