"""Shock capturing on the GPU against the oracle: Persson-Peraire sensor,
BR1 viscous operator, the viscous SSP-RK3 stepper, the sensor network's
forward pass and gradients (BF16 matrix cores), and the learned sensor
trained on the GPU from Persson-Peraire runs."""
import math

import numpy as np
import pytest

from dgkernel_amd import av as AV
from dgkernel_amd import dg
from dgkernel_amd import mesh as M
from oracle import av_oracle as A
from oracle import oracle as O
from tests.helpers import rel_l2, smooth_u0

pytestmark = pytest.mark.gpu

METRICS = ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")


def _spread(o):
    """Relative round-off spread of the oracle's per-node metrics (the device
    uses one exact value per affine element)."""
    return max(float(np.max(np.abs(mm - mm[:, :1])) / np.max(np.abs(mm)))
               for mm in (o.field(k) for k in METRICS))


def _u0(x, y, z):
    return 0.5 + np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)


@pytest.mark.parametrize("N,n,parts", [(2, 4, [200, 184]), (3, 4, [1, 383]), (4, 3, [80, 82])])
def test_pp_sensor_vs_oracle(device, N, n, parts):
    m = M.cube_mesh(n, n, n, periodic=True, jitter=0.06, seed=N)
    st = dg.make_setup(m, N, Kp=parts)
    o = O.Setup.from_mesh(N, m)
    p = AV.AVParams.for_order(N, kappa=1.5)
    u = smooth_u0(st, amp=1.0, noise=1e-3)
    eps, frac, e0, feats = o.pp_sensor(u, p.s0, p.kappa, p.ceps)
    s = AV.AVSolver(device, st, p, "pp")
    s.u.copy_from(st.to_pooled(u))
    fr = device.zeros(m.K)
    f32 = device.zeros(m.K * st.ref.Np, np.float32)
    s.scal.fill0()
    s.sense(frac=fr, feats=f32)
    geps, gfr = s.eps(), fr.to_host()
    assert (frac > 0).any() and (frac < 1).any()
    assert np.max(np.abs(gfr - frac)) <= 1e-10
    assert np.max(np.abs(geps - eps)) <= 1e-10 * max(eps.max(), 1e-30)
    assert np.max(np.abs(f32.to_host().reshape(m.K, -1) - feats)) <= 1e-6
    assert abs(s.scal.to_host()[3] - eps.max()) <= 1e-12 * eps.max()


@pytest.mark.parametrize("N,n,parts", [(1, 4, [384]), (2, 4, [100, 284]), (3, 3, [50, 1, 111]),
                                       (4, 3, [162])])
def test_viscous_rhs_vs_oracle(device, N, n, parts):
    m = M.cube_mesh(n, n, n, periodic=True, jitter=0.06, seed=10 + N)
    st = dg.make_setup(m, N, Kp=parts)
    o = O.Setup.from_mesh(N, m)
    u = smooth_u0(st, amp=0.8, noise=1e-2)
    eps = M.uniform(N, 3, m.K, 0.0, 0.02)
    want = o.av_rhs(u, eps, 1.0)
    s = AV.AVSolver(device, st, AV.AVParams.for_order(N), "pp")
    s.set_eps(eps)
    got = s.rhs_visc(u)
    tol = 1e-11 + 50 * _spread(o)
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))
    tot = s.rhs_total(u)
    want_t = want + o.burgers_rhs(u)
    assert np.max(np.abs(tot - want_t)) <= tol * np.max(np.abs(want_t))


@pytest.mark.parametrize("N", [2, 4])
def test_viscous_operator_and_sensor_in_spatial_tile_order(device, N):
    """The viscous operator and the sensor walk the same tile list as the
    stage: in the Morton order of the tile centroids they still match the
    oracle."""
    m = M.cube_mesh(4, 4, 3, periodic=True, jitter=0.06, seed=40 + N)
    st = dg.make_setup(m, N, Kp=[100, 1, m.K - 101])
    o = O.Setup.from_mesh(N, m)
    p = AV.AVParams.for_order(N, kappa=1.5)
    u = smooth_u0(st, amp=0.8, noise=1e-2)
    eps = M.uniform(N, 5, m.K, 0.0, 0.02)
    want = o.av_rhs(u, eps, 1.0) + o.burgers_rhs(u)
    s = AV.AVSolver(device, st, p, "pp")
    s.mesh.set_tile_order(True)
    s.set_eps(eps)
    got = s.rhs_total(u)
    tol = 1e-11 + 50 * _spread(o)
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))
    eps_pp = o.pp_sensor(u, p.s0, p.kappa, p.ceps)[0]
    s.u.copy_from(st.to_pooled(u))
    s.sense()
    assert np.max(np.abs(s.eps() - eps_pp)) <= 1e-10 * max(eps_pp.max(), 1e-30)


@pytest.mark.parametrize("N,parts", [(2, [162]), (4, [60, 102])])
def test_viscous_rhs_with_boundary_faces_vs_oracle(device, N, parts):
    """Non-periodic cube: boundary faces are their own neighbours (zero jump
    in u and in the viscous flux), as in the oracle's VmapP."""
    m = M.cube_mesh(3, 3, 3, periodic=False, jitter=0.05, seed=30 + N)
    st = dg.make_setup(m, N, Kp=parts)
    o = O.Setup.from_mesh(N, m)
    u = smooth_u0(st, amp=0.8, noise=1e-2)
    eps = M.uniform(N, 4, m.K, 0.0, 0.02)
    want = o.av_rhs(u, eps, 1.0)
    s = AV.AVSolver(device, st, AV.AVParams.for_order(N), "pp")
    s.set_eps(eps)
    got = s.rhs_visc(u)
    tol = 1e-11 + 50 * _spread(o)
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))


@pytest.mark.parametrize("N", [3, 4])
def test_viscous_ssprk3_vs_oracle(device, N):
    m = M.cube_mesh(4, 4, 4, periodic=True)
    st = dg.make_setup(m, N, Kp=M.slab_partition(m.K, 3))
    o = O.Setup.from_mesh(N, m)
    p = AV.AVParams.for_order(N)
    x, y, z = st.node_coords()
    u0 = _u0(x, y, z)
    hmin = o.min_node_distance()
    t_end = 0.12
    want, dts, t = o.ssprk3_av(u0, t_end, 0.3, p.cfl_v, hmin, p.s0, p.kappa, p.ceps, p.tau0)
    s = AV.AVSolver(device, st, p, "pp", hmin=hmin)
    s.set_state(u0)
    steps = s.run(t_end, chunk=8)
    assert steps >= len(dts) and abs(s.time() - t) <= 1e-13
    assert rel_l2(s.state(), want) <= 1e-10


def _random_params(nin, seed):
    return AV.SensorNet.init_params(nin, seed) * 1.5 + np.float32(0.05) * (
        M.uniform(seed, 99, A.layout(nin)["b3"][0] + 1, -1, 1).astype(np.float32))


@pytest.mark.parametrize("nin,n", [(35, 1000), (20, 37), (10, 16)])
def test_mlp_forward_vs_numpy(device, nin, n):
    P = _random_params(nin, 3)
    x = M.uniform(4, 1, n * nin, -1, 1).astype(np.float32).reshape(n, nin)
    net = AV.SensorNet(device, nin, P)
    xd = device.upload(x)
    yd = device.zeros(n, np.float32)
    net.forward(xd, n, yd)
    got = yd.to_host()
    want = A.mlp_forward(P, x, nin, bf16_path=True)
    assert np.max(np.abs(got - want)) <= 2e-3 and np.mean(np.abs(got - want)) <= 2e-4
    full = A.mlp_forward(P, x, nin)  # without bf16 rounding: the rounding is what differs
    assert np.max(np.abs(got - full)) <= 2e-2


@pytest.mark.parametrize("nin,n", [(35, 500), (12, 64), (35, 1), (35, 70001)])
def test_mlp_gradient_vs_numpy(device, nin, n):
    """The gradient kernel (BF16 matrix cores, FP32 accumulation) against the
    oracle's restatement of the same rounding points, and against the exact
    gradient within the rounding's reach; 70001 rows give each wave several
    32-row chunks and a partial last one. Two calls agree bit for bit (the
    per-wave sums are added in a fixed order)."""
    P = _random_params(nin, 5)
    x = M.uniform(6, 1, n * nin, -1, 1).astype(np.float32).reshape(n, nin)
    t = M.uniform(6, 2, n, 0, 1).astype(np.float32)
    net = AV.SensorNet(device, nin, P)
    xd, td = device.upload(x), device.upload(t)
    g, loss = net.grad(xd, td, n)
    g2, loss2 = net.grad(xd, td, n)
    assert np.array_equal(g, g2) and loss == loss2
    gq, lq = A.mlp_grad(P, x, t, nin, bf16_path=True)
    gw, lw = A.mlp_grad(P, x, t, nin)
    # Every gradient entry carries the factor y - t, so FP32 rounding of y
    # (~1e-7 y) is amplified by y / |y - t| (large when a single sample's
    # output is close to its target); batch means average this out.
    y = A.mlp_forward(P, x, nin, bf16_path=True)
    amp = float(np.max(np.abs(y) / np.maximum(np.abs(y - t), 1e-12)))
    eq, ew = rel_l2(g, gq), rel_l2(g, gw)
    print(f"nin={nin} n={n}: rel L2 vs BF16-path oracle {eq:.3e}, vs exact {ew:.3e}, amp {amp:.1f}")
    assert eq <= 1e-4 + 2e-8 * amp
    if n >= 64:  # a single sample with y close to t amplifies the BF16 forward error
        assert ew <= 3e-2
    assert abs(loss - lq) <= 1e-5 * lq + 6e-7 * amp * lq  # (y - t)^2: twice the amplification


def test_mlp_sensor_matches_network_on_modal_features(device):
    N = 4
    m = M.cube_mesh(3, 3, 3, periodic=True, jitter=0.05, seed=2)
    st = dg.make_setup(m, N, Kp=[100, 62])
    o = O.Setup.from_mesh(N, m)
    p = AV.AVParams.for_order(N)
    u = smooth_u0(st, amp=1.0, noise=1e-3)
    _, _, e0, feats = o.pp_sensor(u, p.s0, p.kappa, p.ceps)
    P = _random_params(st.ref.Np, 8)
    net = AV.SensorNet(device, st.ref.Np, P)
    s = AV.AVSolver(device, st, p, "mlp", net=net)
    s.u.copy_from(st.to_pooled(u))
    fr = device.zeros(m.K)
    s.scal.fill0()
    s.sense(frac=fr)
    want = A.mlp_forward(P, feats, st.ref.Np, bf16_path=True)
    got = fr.to_host()
    assert np.max(np.abs(got - want)) <= 3e-3
    assert np.max(np.abs(s.eps() - e0 * got)) <= 1e-5 * e0.max()


def test_learned_viscosity_trained_on_gpu(device):
    """Train the sensor network on Persson-Peraire data from 8 seeded fields,
    then run the learned-viscosity solver past shock formation: its error
    against the exact entropy solution is within 5% of the Persson-Peraire
    run's, and both beat the run without viscosity by far."""
    N, n, t_end = 4, 4, 0.3
    m = M.cube_mesh(n, n, n, periodic=True)
    st = dg.make_setup(m, N, Kp=M.slab_partition(m.K, 2))
    p = AV.AVParams.for_order(N)
    X, T, nrows = AV.generate_training_data(device, st, p, seeds=range(100, 108), t_end=t_end,
                                            every=2)
    frac = T.to_host()
    assert nrows > 100000 and (frac > 0.5).sum() > 1000
    net = AV.SensorNet(device, st.ref.Np, seed=1)
    # measured: loss 5.6e-3 -> 7.7e-4, learned run within 2.2% of the Persson-Peraire run
    hist = net.train(X, T, nrows, steps=10000, batch=2048, lr=3e-3, lr_final=1e-4,
                     report_every=2000)
    assert hist[-1] < 0.25 * hist[0] and hist[-1] < 2e-3, hist
    x, y, z = st.node_coords()
    ex = A.burgers_exact(_u0, x, y, z, t_end, umax=1.5).reshape(m.K, -1)
    errs = {}
    for sensor in ("pp", "mlp", None):
        s = AV.AVSolver(device, st, p, sensor, net=net if sensor == "mlp" else None)
        s.set_state(_u0(x, y, z))
        s.run(t_end, chunk=16, max_steps=4000)
        u = s.state()
        errs[sensor] = rel_l2(u, ex) if np.isfinite(u).all() else np.inf
    print("rel L2 vs exact:", errs)
    assert abs(errs["mlp"] - errs["pp"]) <= 0.05 * errs["pp"], errs
    assert errs["pp"] < 0.5 * errs[None], errs
