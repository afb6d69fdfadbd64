"""Pin the oracle's ferroelectric stage (ferroelectric_solver.cpp:4-131,
534-842, 1786-1999) by closed forms: the q increment is the negative
gradient of the 8th-order Landau energy (finite differences, any
rotation), minus the fields, minus damping; a small transverse wave on a
polarised chain oscillates at omega^2 = (f''(Ps) + G11 lambda(k)) / m with
the discrete Laplacian eigenvalue lambda, advanced by the exact RK4
propagator of that linear system."""
import numpy as np
import pytest

from goferro import FIELD_ESTAT, FIELD_P, FIELD_Q, FIELD_P_STORE, Material, Problem
from goferro.workloads import pto_material


def landau_energy(M, p):
    x, y, z = p
    x2, y2, z2 = x * x, y * y, z * z
    return (M.a1 * (x2 + y2 + z2) + M.a11 * (x2 * x2 + y2 * y2 + z2 * z2)
            + M.a12 * (x2 * y2 + y2 * z2 + z2 * x2)
            + M.a111 * (x2 ** 3 + y2 ** 3 + z2 ** 3)
            + M.a112 * (x2 * x2 * (y2 + z2) + y2 * y2 * (x2 + z2) + z2 * z2 * (x2 + y2))
            + M.a123 * x2 * y2 * z2
            + M.a1111 * (x2 ** 4 + y2 ** 4 + z2 ** 4)
            + M.a1112 * (x2 ** 3 * (y2 + z2) + y2 ** 3 * (x2 + z2) + z2 ** 3 * (x2 + y2))
            + M.a1122 * (x2 * x2 * y2 * y2 + y2 * y2 * z2 * z2 + z2 * z2 * x2 * x2)
            + M.a1123 * (x2 * x2 * y2 * z2 + y2 * y2 * z2 * x2 + z2 * z2 * x2 * y2))


def full_landau_material():
    # every coefficient non-zero so each term of the force is exercised
    return Material.make(if_FE=1, FE_mass=2e-16, FE_damping=37.0, a1=-1.7e8, a11=-7.3e7, a12=7.5e8, a111=2.6e8,
                         a112=6.1e8, a123=-3.7e9, a1111=4.1e8, a1112=-2.2e8, a1122=9.3e8, a1123=1.7e9, G11=1e-10)


@pytest.mark.parametrize("rotate", [False, True])
def test_dq_is_landau_gradient(oracle_lib, rotate):
    M = full_landau_material()
    R = None
    if rotate:
        c, s = np.cos(0.6), np.sin(0.6)
        R = np.array([[1, 0, 0], [0, c, -s], [0, s, c]]) @ np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    dt = 1e-18
    prob = Problem(shape=(1, 1, 1), spacing=(1e-9,) * 3, materials=[M], material_cell=np.ones((1, 1, 1), np.uint32),
                   dt=dt, flags=(1 << 4) if rotate else 0, Tg2c=R, Eext_stat=(2e6, -1e6, 3e6))
    p0 = np.array([0.31, -0.22, 0.58])
    q0 = np.array([1e5, -3e5, 2e5])
    E = np.array([4e6, 1e6, -2e6])
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_P, p0.reshape(3, 1, 1, 1))
    o.set(FIELD_Q, q0.reshape(3, 1, 1, 1))
    o.set(FIELD_ESTAT, E.reshape(3, 1, 1, 1))
    o.time_advance()
    o.fe_stage(1)
    dq = 2 * (o.get(oracle_lib.FIELD_Q_STORE).reshape(3) - q0)
    T = np.eye(3) if R is None else R
    h = 1e-6
    grad = np.array([(landau_energy(M, T @ (p0 + h * e)) - landau_energy(M, T @ (p0 - h * e))) / (2 * h)
                     for e in np.eye(3)])
    Etot = E + np.array([2e6, -1e6, 3e6])
    ref = -dt / M.FE_mass * (grad - Etot) - dt * M.FE_damping / M.FE_mass * q0
    np.testing.assert_allclose(dq, ref, rtol=0, atol=1e-8 * np.abs(ref).max())
    # dp = q_store * dt
    p_store = o.get(FIELD_P_STORE).reshape(3)
    np.testing.assert_allclose(p_store, p0 + 0.5 * q0 * dt, rtol=1e-15, atol=0)


@pytest.mark.parametrize("q", [1, 2, 5])
def test_transverse_soft_mode_dispersion(oracle_lib, q):
    M = pto_material(FE_mass=1e-16, FE_damping=0.0)
    nx, d, dt, N, eps = 16, 1e-9, 2e-15, 400, 1e-6
    a1, a11, a111 = M.a1, M.a11, M.a111
    Ps = np.sqrt((-a11 + np.sqrt(a11 * a11 - 3 * a1 * a111)) / (3 * a111))
    prob = Problem(shape=(nx, 1, 1), spacing=(d, d, d), materials=[M], material_cell=np.ones((nx, 1, 1), np.uint32),
                   dt=dt)
    k = 2 * np.pi * q / (nx * d)
    x = np.arange(nx) * d
    p0 = np.zeros((3, nx, 1, 1))
    p0[2] = Ps
    p0[0, :, 0, 0] = eps * np.cos(k * x)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_P, p0)
    o.set(FIELD_Q, np.zeros_like(p0))
    o.step(N)
    px = o.get(FIELD_P)[0, :, 0, 0]
    lam = (2 - 2 * np.cos(k * d)) / d ** 2
    f2 = 2 * a1 + 2 * M.a12 * Ps ** 2 + 2 * M.a112 * Ps ** 4
    omega = np.sqrt((f2 + M.G11 * lam) / M.FE_mass)
    # exact RK4 propagator of x'' = -omega^2 x
    A = np.array([[0.0, 1.0], [-omega ** 2, 0.0]]) * dt
    R = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    y = np.linalg.matrix_power(R, N) @ np.array([1.0, 0.0])
    np.testing.assert_allclose(px, eps * y[0] * np.cos(k * x), rtol=0, atol=1e-6 * eps)


def test_spontaneous_polarisation_relaxes(oracle_lib):
    M = pto_material(FE_mass=1e-16, FE_damping=0.0)
    a1, a11, a111 = M.a1, M.a11, M.a111
    Ps = np.sqrt((-a11 + np.sqrt(a11 * a11 - 3 * a1 * a111)) / (3 * a111))
    assert abs(Ps - 0.754) < 2e-3   # the configured PbTiO3-like film
    f2 = 2 * a1 + 12 * a11 * Ps ** 2 + 30 * a111 * Ps ** 4
    M.FE_damping = 2 * np.sqrt(M.FE_mass * f2)       # critical damping at Ps
    prob = Problem(shape=(1, 1, 1), spacing=(1e-9,) * 3, materials=[M], material_cell=np.ones((1, 1, 1), np.uint32),
                   dt=1e-15)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_P, np.array([0.0, 0.0, 0.6]).reshape(3, 1, 1, 1))
    o.set(FIELD_Q, np.zeros((3, 1, 1, 1)))
    o.step(8000)
    assert abs(o.get(FIELD_P).reshape(3)[2] - Ps) < 1e-9
