"""Pins of the shock-capturing oracle (oracle/dgk_oracle.cpp ppSensor, avRHS,
orc_ssprk3_av; oracle/av_oracle.py). The reference has no shock capturing,
so each function is pinned by a closed form or an invariant it must satisfy:

* sensor: single-mode and two-mode fields, whose top-mode energy fraction S
  and ramp value follow from the orthonormal modal basis by hand;
* viscous operator: a constant state gives 0; discrete conservation over a
  periodic mesh; energy dissipation u'MJ L_v(u) <= 0; convergence to
  eps Laplacian(u) for a smooth field and constant eps;
* exact solution: agrees with the characteristic solution before the shock;
* network gradient: central finite differences;
* the viscous SSP-RK3 run: stays bounded past shock formation and is closer
  to the exact entropy solution than the run without viscosity.
"""
import math

import numpy as np
import pytest

from dgkernel_amd import dg
from dgkernel_amd import mesh as M
from dgkernel_amd.av import AVParams
from oracle import av_oracle as A
from oracle import oracle as O


def _mass_weights(o, el):
    J = o.field("J")
    return (el["Mass"] @ np.ones(el["Np"]))[None, :] * J  # 1'M per node times J


def _ramp(S, s0, kappa):
    if S <= 0:
        return 0.0
    s = math.log10(S)
    if s < s0 - kappa:
        return 0.0
    if s > s0 + kappa:
        return 1.0
    return 0.5 * (1 + math.sin(math.pi * (s - s0) / (2 * kappa)))


@pytest.mark.parametrize("N", [2, 3, 4])
def test_pp_sensor_closed_form_single_and_two_modes(N):
    m = M.cube_mesh(2, 2, 2, periodic=True)  # 48 tets
    o = O.Setup.from_mesh(N, m)
    el = O.element(N)
    Np, K = el["Np"], m.K
    deg = [i + j + k for i in range(N + 1) for j in range(N + 1 - i) for k in range(N + 1 - i - j)]
    top = [q for q in range(Np) if deg[q] == N]
    s0, kappa, ceps = -4 * math.log10(N), 1.0, 0.7
    rng = np.random.default_rng(N)
    uh = np.zeros((K, Np))
    uh[:, 0] = rng.uniform(0.5, 2.0, K)                 # mean mode
    b = 10.0 ** rng.uniform(-4, 0, K)                    # top mode, S spans 1e-8 .. 0.5
    uh[np.arange(K), np.array(top)[rng.integers(0, len(top), K)]] = b * uh[:, 0]
    u = uh @ el["V"].T                                  # nodal values, u = V uh per element
    eps, frac, e0, feats = o.pp_sensor(u, s0, kappa, ceps)
    J = o.field("J")[:, 0]
    for k in range(K):
        S = b[k] ** 2 / (1 + b[k] ** 2)
        assert abs(frac[k] - _ramp(S, s0, kappa)) <= 1e-9
        h = (6 * 4.0 / 3.0 * J[k]) ** (1 / 3)
        assert abs(e0[k] - ceps * h * math.sqrt(3) * np.max(np.abs(u[k])) / N) <= 1e-12
        nrm = np.linalg.norm(uh[k])
        assert np.max(np.abs(feats[k] - uh[k] / nrm)) <= 1e-6
    assert np.allclose(eps, e0 * frac, rtol=1e-14, atol=0)
    # u = 0 and a constant give no viscosity
    eps0, fr0, _, f0 = o.pp_sensor(np.zeros_like(u), s0, kappa, ceps)
    assert not eps0.any() and not fr0.any() and not f0.any()
    _, frc, _, _ = o.pp_sensor(np.full_like(u, 0.3), s0, kappa, ceps)
    assert np.max(frc) == 0.0


def test_viscous_operator_invariants():
    N = 3
    m = M.cube_mesh(3, 3, 3, periodic=True, jitter=0.08, seed=5)
    o = O.Setup.from_mesh(N, m)
    el = O.element(N)
    rng = np.random.default_rng(7)
    K, Np = m.K, el["Np"]
    eps = rng.uniform(0.0, 0.05, K)
    # constant state
    assert np.max(np.abs(o.av_rhs(np.full((K, Np), 0.37), eps, 1.0))) <= 1e-12
    w = _mass_weights(o, el)
    for trial in range(3):
        u = rng.uniform(-1, 1, (K, Np))
        L = o.av_rhs(u, eps, 1.0)
        scale = np.sum(np.abs(w * L))
        assert abs(np.sum(w * L)) <= 1e-12 * scale          # conservation
        JM = o.field("J")[:, :1] * (u @ el["Mass"])          # J u'M per element
        energy = np.sum(JM * L)
        assert energy <= 1e-12 * np.sum(np.abs(JM * L))      # dissipation


@pytest.mark.parametrize("N,rate_min", [(3, 1.4), (4, 2.5)])
def test_viscous_operator_converges_to_eps_laplacian(N, rate_min):
    """Pointwise consistency of a DG second-derivative operator on
    interpolated data is O(h^(N-1)); measured rates 1.5-1.8 (N=3), 2.9-3.1
    (N=4) between 3^3 and 9^3 hexes."""
    errs = []
    for n in (3, 6):
        m = M.cube_mesh(n, n, n, periodic=True)
        o = O.Setup.from_mesh(N, m)
        st = dg.make_setup(m, N)
        x, y, z = st.node_coords()
        u = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
        lap = -2 * (2 * np.pi) ** 2 * u
        eps = np.full(m.K, 0.01)
        L = o.av_rhs(u, eps, 1.0)
        errs.append(np.sqrt(np.mean((L - 0.01 * lap) ** 2)) / np.sqrt(np.mean((0.01 * lap) ** 2)))
    rate = math.log2(errs[0] / errs[1])
    assert errs[1] < errs[0] and rate > rate_min, (errs, rate)


def _u0(x, y, z):
    return 0.5 + np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)


def test_exact_solution_matches_characteristics_before_shock():
    rng = np.random.default_rng(3)
    X, Y, Z = (rng.uniform(0, 1, 400) for _ in range(3))
    t = 0.05
    v = A.burgers_exact(_u0, X, Y, Z, t, umax=1.5, n_grid=8001)
    w = _u0(X, Y, Z)
    for _ in range(200):  # v = u0(x - v t (1,1,1)), a contraction for t max|grad u0| < 1
        w = _u0(X - w * t, Y - w * t, Z - w * t)
    assert np.max(np.abs(v - w)) <= 1e-5


def test_mlp_gradient_finite_differences():
    nin = 10
    rng = np.random.default_rng(2)
    P = rng.normal(0, 0.3, A.layout(nin)["b3"][0] + 1)
    x = rng.normal(0, 1, (13, nin))
    t = rng.uniform(0, 1, 13)
    g, _ = A.mlp_grad(P, x, t, nin)

    def loss(p):
        return float(np.mean((A.mlp_forward(p, x, nin) - t) ** 2))

    for i in rng.choice(P.size, 40, replace=False):
        e = np.zeros_like(P)
        e[i] = 1e-6
        fd = (loss(P + e) - loss(P - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd)


def test_mlp_gradient_bf16_path():
    """The BF16 gradient path: with its rounding replaced by the identity it is
    the exact (finite-difference-pinned) gradient; with bfloat16 rounding it
    stays within the rounding's reach of the exact gradient (a few 1e-3
    relative), while a dropped tanh derivative in the backward pass lands
    far outside it."""
    nin = 35
    rng = np.random.default_rng(7)
    P = rng.normal(0, 0.25, A.layout(nin)["b3"][0] + 1)
    x = rng.uniform(-1, 1, (200, nin))
    t = rng.uniform(0, 1, 200)
    g, loss = A.mlp_grad(P, x, t, nin)
    gi, li = A.mlp_grad(P, x, t, nin, rnd=lambda v: np.asarray(v, np.float64))
    assert np.array_equal(g, gi) and loss == li
    gb, lb = A.mlp_grad(P, x, t, nin, bf16_path=True)
    e = np.linalg.norm(gb - g) / np.linalg.norm(g)
    assert 1e-4 < e < 2e-2
    assert abs(lb - loss) <= 2e-2 * loss
    # planted mistake: tanh' of the first layer omitted from the backward pass
    w = A.unpack(P, nin)
    h1 = np.tanh(x @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    y = 1.0 / (1.0 + np.exp(-(h2 @ w["w3"] + w["b3"][0])))
    dz2 = np.outer(2.0 * (y - t) / len(t) * y * (1 - y), w["w3"]) * (1 - h2 * h2)
    bad = (dz2 @ w["W2"]).T @ x
    o = A.layout(nin)["W1"][0]
    good = g[o:o + 64 * nin].reshape(64, nin)
    assert np.linalg.norm(bad - good) / np.linalg.norm(good) > 10 * e


def test_viscous_run_bounded_and_closer_to_entropy_solution():
    N, n, t_end = 3, 5, 0.3
    m = M.cube_mesh(n, n, n, periodic=True)
    o = O.Setup.from_mesh(N, m)
    st = dg.make_setup(m, N)
    x, y, z = st.node_coords()
    u0 = _u0(x, y, z)
    hmin = o.min_node_distance()
    p = AVParams.for_order(N)
    uv, dts, t = o.ssprk3_av(u0, t_end, 0.3, p.cfl_v, hmin, p.s0, p.kappa, p.ceps, p.tau0, mode=1)
    assert abs(t - t_end) <= 1e-14 and np.isfinite(uv).all()
    assert np.max(np.abs(uv)) <= 2.0  # exact max is 1.5; the coarse mesh overshoots at the shock
    ex = A.burgers_exact(_u0, x, y, z, t_end, umax=1.5).reshape(uv.shape)
    e_av = np.linalg.norm(uv - ex) / np.linalg.norm(ex)
    un, _, tn = o.ssprk3_av(u0, t_end, 0.3, p.cfl_v, hmin, p.s0, p.kappa, p.ceps, p.tau0, mode=0)
    e_no = np.linalg.norm(un - ex) / np.linalg.norm(ex) if np.isfinite(un).all() else np.inf
    # measured: 0.19 with viscosity, 0.66 without (h = 1/5, P = 3)
    assert e_av < 0.25 and e_av < 0.5 * e_no, (e_av, e_no)
