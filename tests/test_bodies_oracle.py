"""The oracle's gas/body couplings and diagnostics against the reference's own
tests (test/test_sinks.jl, test/test_diagnostics.jl) and against what the
physics fixes independently of the formulas:

- gas_force_on_bh (src/gravity_bh.jl:91-113): Newton's third law against the
  separately written point-mass source add_bh_gravity_source!, and the force of
  a uniform ball on a distant body (shell theorem);
- accrete! (src/sinks.jl:73-109): the reference's mass and threshold tests, and
  the momentum the body gains equals what the plain (not torque-free) sink
  drains from the gas over dt;
- the diagnostics (src/diagnostics.jl): the reference's tests, and closed forms
  for uniform flow, solid-body rotation and the bound radius of a point mass.
"""
import math

import numpy as np

from oracle import oracle as O

NG = 3


def empty(n):
    return np.zeros((n + 2 * NG, n + 2 * NG, n + 2 * NG, 5))


def act(U):
    return U[NG:-NG, NG:-NG, NG:-NG]


def centres(n, d, x0):
    return x0 + (np.arange(1, n + 1) - 0.5) * d


# --------------------------------------------------------------------------
# gas_force_on_bh

def test_force_is_minus_the_point_mass_momentum_source():
    rng = np.random.default_rng(3)
    n, d, x0 = 12, 0.1, -0.6
    U = empty(n)
    act(U)[..., 0] = rng.uniform(0.2, 2.0, (n, n, n))
    act(U)[..., 4] = 1.0
    pos, M, eps = (0.07, -0.13, 0.21), 2.5, 0.05
    F = O.gas_force_on_bh(U, d, d, d, pos, M, eps, x0, x0, x0)
    dU = np.zeros_like(U)
    O.bh_gravity_source(dU, U, d, d, d, x0, x0, x0, [pos], [M], [eps])
    src = act(dU)[..., 1:4] * d ** 3
    scale = np.abs(src).sum((0, 1, 2))
    assert np.all(np.abs(F + src.sum((0, 1, 2))) <= 1e-13 * scale)
    assert np.all(np.abs(F) > 1e-3 * scale)         # not a trivial zero


def test_force_of_a_uniform_ball_on_a_distant_body():
    # shell theorem: F = M m_ball / D^2 toward the ball, here along -x
    n, x0 = 48, -1.0
    d = 2.0 / n
    U = empty(n)
    c = centres(n, d, x0)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    act(U)[..., 0] = np.where(X * X + Y * Y + Z * Z < 0.6 ** 2, 1.0, 0.0)
    m_ball = act(U)[..., 0].sum() * d ** 3
    D, M = 25.0, 3.0
    F = O.gas_force_on_bh(U, d, d, d, (D, 0.0, 0.0), M, 1e-8, x0, x0, x0)
    assert abs(F[0] / (-M * m_ball / D ** 2) - 1) < 1e-3
    assert abs(F[1]) < 1e-9 * abs(F[0]) and abs(F[2]) < 1e-9 * abs(F[0])


# --------------------------------------------------------------------------
# accrete!

def _sink_case(rho, P0=1e-2, g=5.0 / 3.0, n=8, d=0.1):
    U = empty(n)
    act(U)[..., 0] = rho
    act(U)[..., 4] = P0 / (g - 1)
    return U


def test_accrete_mass_matches_the_gas_inside_the_sink():
    # test/test_sinks.jl:140-187: dM = M_gas(r < r_sink) / t_sink * dt, rtol 1e-14
    n, d, dt = 8, 0.1, 0.01
    U = _sink_case(1.5)
    pos = (0.5 * n * d,) * 3
    M0, c_code, r_floor = 1.0, 1e6, 2.5 * d
    rs = max(6.0 * M0 / c_code ** 2, r_floor)
    ts = 1.0 * rs / math.sqrt(2.0 * M0 / rs)
    c = centres(n, d, 0.0)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    inside = np.sqrt((X - pos[0]) ** 2 + (Y - pos[1]) ** 2 + (Z - pos[2]) ** 2) < rs
    M_gas = (act(U)[..., 0] * inside).sum() * d ** 3
    s = O.accrete_sums(U, d, d, d, 0.0, 0.0, 0.0, pos, rs, ts, dt)
    mass, vel = O.accrete(M0, [0.0, 0.0, 0.0], s)
    assert abs((mass - M0) - M_gas / ts * dt) <= 1e-14 * (M_gas / ts * dt)
    assert inside.sum() > 8


def test_accrete_threshold_skips_low_density_cells():
    # test/test_sinks.jl:191-249: only the planted cell above rho_sink_min counts
    n, d, dt = 8, 0.1, 0.01
    U = _sink_case(1e-3)
    ic = NG + n // 2
    U[ic, ic, ic, 0] = 1.0
    pos = (0.5 * n * d,) * 3
    rs = 3.0 * d
    ts = rs / math.sqrt(2.0 * 1.0 / rs)
    s = O.accrete_sums(U, d, d, d, 0.0, 0.0, 0.0, pos, rs, ts, dt, rho_sink_min=2e-3)
    assert abs(s[0] - (1.0 * d ** 3 / ts) * dt) <= 1e-10 * s[0]
    s0 = O.accrete_sums(U, d, d, d, 0.0, 0.0, 0.0, pos, rs, ts, dt)
    assert s0[0] > s[0]


def test_accreted_momentum_is_what_the_plain_sink_drains():
    rng = np.random.default_rng(8)
    n, d, dt = 10, 0.1, 0.02
    U = empty(n)
    act(U)[..., 0] = rng.uniform(0.5, 2.0, (n, n, n))
    act(U)[..., 1:4] = rng.normal(0.0, 1.0, (n, n, n, 3))
    act(U)[..., 4] = 5.0
    pos, vel, rs, ts = (0.43, 0.52, 0.47), (0.1, -0.2, 0.3), 0.31, 0.7
    s = O.accrete_sums(U, d, d, d, 0.0, 0.0, 0.0, pos, rs, ts, dt)
    dU = np.zeros_like(U)
    O.sink_sources(dU, U, d, d, d, 0.0, 0.0, 0.0, [pos], [vel], [rs], [ts], torque_free=False)
    drained = -act(dU)[..., :4].sum((0, 1, 2)) * d ** 3 * dt
    assert np.allclose(s, drained, rtol=1e-13, atol=0)
    # the body update conserves momentum: M' v' = M v + dP
    M0 = 2.0
    mass, v = O.accrete(M0, vel, s)
    assert mass == M0 + s[0]
    assert np.allclose(mass * np.array(v), M0 * np.array(vel) + s[1:], rtol=1e-15, atol=1e-16)
    # nothing accreted: the body is unchanged
    assert O.accrete(M0, vel, [0.0, 1.0, 1.0, 1.0]) == (M0, list(vel))


# --------------------------------------------------------------------------
# diagnostics

def test_energy_momentum_and_bound_mass_reference_checks():
    # test/test_diagnostics.jl:13-50 and :78-102
    n, d, x0, g = 8, 0.1, -0.4, 5.0 / 3.0
    U = empty(n)
    act(U)[..., 0] = 1.0
    act(U)[..., 4] = 2.0
    out = O.gas_diagnostics(U, d, d, d, x0, x0, x0)
    E_ref = 2.0 * n ** 3 * d ** 3
    assert abs(out[0] - E_ref) <= 1e-14 * E_ref
    act(U)[..., 4] = 0.01
    assert np.abs(O.gas_diagnostics(U, d, d, d, x0, x0, x0)[2:5]).max() < 1e-14
    act(U)[..., 4] = 1e-6 / (g - 1)
    out = O.gas_diagnostics(U, d, d, d, x0, x0, x0, [0.0, 0.0, 0.0], [10.0], [0.01], g)
    assert out[8] / (n ** 3 * d ** 3) > 0.9


def test_uniform_flow_and_solid_rotation_closed_forms():
    n, rho = 10, 1.3
    d = 1.0 / n
    x0 = -0.5
    c = centres(n, d, x0)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    U = empty(n)
    act(U)[..., 0] = rho
    v = np.array([0.3, -0.2, 0.5])
    act(U)[..., 1:4] = rho * v
    act(U)[..., 4] = 1.0
    out = O.gas_diagnostics(U, d, d, d, x0, x0, x0)
    V = 1.0
    assert np.allclose(out[1], 0.5 * rho * (v @ v) * V, rtol=1e-13)
    assert np.allclose(out[2:5], rho * v * V, rtol=1e-13)
    # a uniform flow through a centred box carries no angular momentum
    assert np.abs(out[5:8]).max() < 1e-14
    # solid rotation about z, omega: L_z = rho omega sum (x^2 + y^2) dV with
    # sum_i x_i^2 = d^2 n (n^2 - 1) / 12 on a centred axis of n cells
    w = 0.7
    act(U)[..., 1] = -rho * w * Y
    act(U)[..., 2] = rho * w * X
    act(U)[..., 3] = 0.0
    out = O.gas_diagnostics(U, d, d, d, x0, x0, x0)
    sx2 = d * d * n * (n * n - 1) / 12.0
    Lz = rho * w * 2 * (n * n) * sx2 * d ** 3
    assert abs(out[7] / Lz - 1) < 1e-13
    assert abs(out[5]) < 1e-14 * Lz and abs(out[6]) < 1e-14 * Lz


def test_bound_mass_is_the_ball_inside_the_bound_radius():
    # gas at rest, e_int = e0: bound exactly where M / sqrt(r^2 + eps^2) >= e0
    n, x0, g = 16, -0.8, 5.0 / 3.0
    d = 1.6 / n
    c = centres(n, d, x0)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    U = empty(n)
    rho, e0, M, eps = 0.9, 2.0, 1.0, 0.01
    act(U)[..., 0] = rho
    act(U)[..., 4] = rho * e0
    R_b = math.sqrt((M / e0) ** 2 - eps ** 2)
    r = np.sqrt(X * X + Y * Y + Z * Z)
    assert np.abs(r - R_b).min() > 1e-6              # no cell on the edge
    expect = rho * (r <= R_b).sum() * d ** 3
    out = O.gas_diagnostics(U, d, d, d, x0, x0, x0, [0.0, 0.0, 0.0], [M], [eps], g)
    assert abs(out[8] - expect) <= 1e-13 * expect
    assert 0 < (r <= R_b).sum() < n ** 3
