"""Pin the oracle's electrodynamics (EMdynamic_solver.cpp:337-4556,
EMdynamic_AbsorbingBoundary.cpp:3-874) by closed forms:

- a vacuum Yee mode of wavenumber q along each axis and polarisation
  advances by the exact RK4 propagator of its 2x2 semi-discrete system,
  [a, b]' = [[0, -s/eps0], [s/mu0, 0]] [a, b] with s = 2 sin(q d/2)/d,
  which fixes the curl stencils, their staggering, signs and eps0/mu0;
- a uniform field in a conductor decays by RK4 of E' = -sigma/(eps0 er) E;
- a uniform Drude current and field follow RK4 of the coupled linear ODE
  E' = -J/(eps0 er), J' = eps0 wp^2 cn E - J/tau;
- a uniform dP/dt gives E = -(P - P0)/(eps0 er) (Ampere with a
  polarisation current), a uniform dm/dt gives H = -Ms (m - m0);
- a first-order Mur wall absorbs a normally incident pulse (reflection
  below 3 %), a PEC wall reflects it with the tangential E held at zero;
- a current sheet K = Jf dz radiates plane waves of amplitude eta0 K / 2.

The four-cell averaging at material interfaces is pinned in
test_oracle_pins.py; the half/full-stage Mur coefficients are pinned in
test_oracle_boundary_pins.py (exact for outgoing linear waves).
"""
import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_H_EM, FIELD_JP, FIELD_M, FIELD_P, FIELD_Q, FLAG_EM_FROM_M, FLAG_EM_FROM_P,
                     FLAG_EMDYNAMIC, FLAG_JF_INPUT, FLAG_PEC_Z, MU0, Material, Problem, field_size, split_yee)

EPS0 = 8.854187817e-12


def vacuum(shape, d=(1e-9, 1e-9, 1e-9), dt=1e-18, periodic=(True, True, True), flags=0, mats=(), cell=None, **kw):
    if cell is None:
        cell = np.zeros(shape, np.uint32)
    return Problem(shape=shape, spacing=d, materials=list(mats), material_cell=cell, dt=dt, periodic=periodic,
                   flags=FLAG_EMDYNAMIC | flags, **kw)


def rk4_matrix(M, h):
    A = M * h
    I = np.eye(A.shape[0])
    return I + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24


def edge_coords(shape, comp, axis, d):
    """Coordinate along `axis` of each Yee E edge of component comp (edges
    sit on cell corners along the axes other than comp)."""
    from goferro import FIELD_E_EM, yee_shapes
    s = yee_shapes(shape, FIELD_E_EM)[comp]
    idx = np.indices(s)[axis].astype(float)
    return (idx + (0.5 if axis == comp else 0.0)) * d[axis]


def face_coords(shape, comp, axis, d):
    from goferro import FIELD_H_EM, yee_shapes
    s = yee_shapes(shape, FIELD_H_EM)[comp]
    idx = np.indices(s)[axis].astype(float)
    return (idx + (0.0 if axis == comp else 0.5)) * d[axis]


# (propagation axis, E component, H component, sign of the H partner)
MODES = [(2, 0, 1, +1), (2, 1, 0, -1), (0, 1, 2, +1), (0, 2, 1, -1), (1, 2, 0, +1), (1, 0, 2, -1)]


@pytest.mark.parametrize("axis,ec,hc,sgn", MODES)
@pytest.mark.parametrize("q", [1, 3])
def test_vacuum_mode_follows_rk4_propagator(oracle_lib, axis, ec, hc, sgn, q):
    n, d, N = 16, 1e-9, 40
    shape = [2, 2, 2]
    shape[axis] = n
    shape = tuple(shape)
    sp = (d, d, d)
    dt = 0.3 * d / 3e8
    prob = vacuum(shape, sp, dt)
    kq = 2 * np.pi * q / (n * d)
    from goferro import yee_shapes
    Es = [np.zeros(s) for s in yee_shapes(shape, FIELD_E_EM)]
    Es[ec] = np.cos(kq * edge_coords(shape, ec, axis, sp))
    Hs = [np.zeros(s) for s in yee_shapes(shape, FIELD_H_EM)]
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_E_EM, Es)
    o.set(FIELD_H_EM, Hs)
    o.step(N)
    s_ = 2 * np.sin(kq * d / 2) / d
    # a' = -sgn s/eps0 b, b' = sgn s/mu0 a for E = a cos(qx), H = b sin(q x')
    M = np.array([[0.0, -sgn * s_ / EPS0], [sgn * s_ / MU0, 0.0]])
    a, b = np.linalg.matrix_power(rk4_matrix(M, dt), N) @ np.array([1.0, 0.0])
    Eg = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)
    Hg = split_yee(o.get(FIELD_H_EM), shape, FIELD_H_EM)
    np.testing.assert_allclose(Eg[ec], a * np.cos(kq * edge_coords(shape, ec, axis, sp)), rtol=0, atol=1e-12)
    hwant = b * np.sin(kq * face_coords(shape, hc, axis, sp))
    np.testing.assert_allclose(Hg[hc], hwant, rtol=0, atol=1e-12 * max(abs(b), 1e-300))
    for c in range(3):
        if c != ec:
            assert np.abs(Eg[c]).max() <= 1e-12
        if c != hc:
            assert np.abs(Hg[c]).max() <= 1e-12 * abs(b)


def uniform_E(shape, comp, value):
    from goferro import yee_shapes
    Es = [np.zeros(s) for s in yee_shapes(shape, FIELD_E_EM)]
    Es[comp][...] = value
    return Es


def test_conductor_decay(oracle_lib):
    sigma, er, dt, N = 3e4, 4.0, 2e-17, 25
    mat = Material.make(r_permittivity=er, conductivity=sigma, comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0,
                        tao_e_n3=1.0)
    shape = (2, 2, 2)
    prob = vacuum(shape, dt=dt, mats=[mat], cell=np.ones(shape, np.uint32))
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_E_EM, uniform_E(shape, 1, 7.0))
    o.step(N)
    r = rk4_matrix(np.array([[-sigma / (EPS0 * er)]]), dt)[0, 0]
    Eg = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)
    np.testing.assert_allclose(Eg[1], 7.0 * r ** N, rtol=1e-12)


def test_drude_current(oracle_lib):
    wp, tau, er, dt, N = 2e15, 3e-15, 2.0, 1e-17, 60
    mat = Material.make(r_permittivity=er, comp_n1=1.0, omega_plasma_n1=wp, tao_e_n1=tau, tao_e_n2=1.0,
                        tao_e_n3=1.0)
    shape = (2, 2, 2)
    prob = vacuum(shape, dt=dt, mats=[mat], cell=np.ones(shape, np.uint32))
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_E_EM, uniform_E(shape, 2, 1e6))
    o.step(N)
    M = np.array([[0.0, -1.0 / (EPS0 * er)], [EPS0 * wp * wp, -1.0 / tau]])
    E, J = np.linalg.matrix_power(rk4_matrix(M, dt), N) @ np.array([1e6, 0.0])
    Eg = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)
    np.testing.assert_allclose(Eg[2], E, rtol=1e-11)
    Jp = o.get(FIELD_JP)
    np.testing.assert_allclose(Jp[0, 2], J, rtol=1e-11)
    assert np.abs(Jp[1:]).max() == 0.0


def test_polarisation_current(oracle_lib):
    """A ferroelectric with no restoring force keeps q: p = p0 + q t and the
    uniform field follows E = -(p - p0)/(eps0 er)."""
    er, dt, N = 10.0, 1e-18, 30
    fe = Material.make(if_FE=1, FE_mass=1e-16, r_permittivity=er, comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0,
                       tao_e_n3=1.0)
    shape = (2, 2, 2)
    prob = vacuum(shape, dt=dt, mats=[fe], cell=np.ones(shape, np.uint32), flags=FLAG_EM_FROM_P)
    o = oracle_lib.Oracle(prob)
    p0 = np.zeros((3,) + shape)
    p0[2] = 0.5
    q = np.zeros((3,) + shape)
    q[2] = 3e12
    o.set(FIELD_P, p0)
    o.set(FIELD_Q, q)
    o.step(N)
    t = N * dt
    np.testing.assert_allclose(o.get(FIELD_P)[2], 0.5 + 3e12 * t, rtol=1e-13)
    Eg = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)
    np.testing.assert_allclose(Eg[2], -3e12 * t / (EPS0 * er), rtol=1e-12)
    assert np.abs(Eg[0]).max() == 0.0 and np.abs(Eg[1]).max() == 0.0


def test_magnetisation_current(oracle_lib):
    """Uniform precession in a periodic box: Faraday's law with dM gives
    H = -Ms (m - m0) at every face."""
    from goferro.workloads import GYRO
    Ms, dt, N = 1.4e5, 1e-14, 50
    fm = Material.make(if_FM=1, Ms=Ms, gyro=GYRO, FM_damping=0.05, Aex=3.7e-12, comp_n1=1.0, tao_e_n1=1.0,
                       tao_e_n2=1.0, tao_e_n3=1.0, r_permittivity=1.0)
    shape = (2, 2, 2)
    prob = vacuum(shape, dt=dt, mats=[fm], cell=np.ones(shape, np.uint32), flags=FLAG_EM_FROM_M,
                  Hext_stat=(0.0, 0.0, 8e4))
    m0 = np.zeros((3,) + shape)
    m0[0], m0[2] = np.sin(0.4), np.cos(0.4)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m0)
    o.step(N)
    m = o.get(FIELD_M)
    Hg = split_yee(o.get(FIELD_H_EM), shape, FIELD_H_EM)
    for c in range(3):
        np.testing.assert_allclose(Hg[c], -Ms * (m[c, 0, 0, 0] - m0[c, 0, 0, 0]), rtol=0, atol=1e-9 * Ms)


def pulse_run(oracle_lib, flags, N=900):
    n, d = 240, 1e-9
    shape = (2, 2, n)
    dt = 0.5 * d / 3e8
    prob = vacuum(shape, (d, d, d), dt, periodic=(True, True, False), flags=flags)
    from goferro import yee_shapes
    z_e = edge_coords(shape, 0, 2, (d, d, d))
    z_h = face_coords(shape, 1, 2, (d, d, d))
    z0, w = 120 * d, 8 * d
    Es = [np.zeros(s) for s in yee_shapes(shape, FIELD_E_EM)]
    Hs = [np.zeros(s) for s in yee_shapes(shape, FIELD_H_EM)]
    eta0 = np.sqrt(MU0 / EPS0)
    Es[0] = np.exp(-((z_e - z0) / w) ** 2)
    Hs[1] = np.exp(-((z_h - z0) / w) ** 2) / eta0          # travels towards +z
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_E_EM, Es)
    o.set(FIELD_H_EM, Hs)
    e0 = (Es[0] ** 2).sum()
    o.step(N)
    Eg = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)
    return Eg, e0


def test_mur_wall_absorbs(oracle_lib):
    Eg, e0 = pulse_run(oracle_lib, 0)
    assert (Eg[0] ** 2).sum() < (0.03 ** 2) * e0


def test_pec_wall_reflects(oracle_lib):
    Eg, e0 = pulse_run(oracle_lib, FLAG_PEC_Z)
    assert np.abs(Eg[0][:, :, [0, -1]]).max() == 0.0
    # the reflected pulse keeps (nearly) all its energy
    assert (Eg[0] ** 2).sum() > 0.9 * e0


def test_current_sheet_radiates_half_eta_K(oracle_lib):
    # a 1D sheet radiates E = -eta0 K/2 sin(2 pi f (t - |z|/c)) both ways;
    # the probe 40 cells away sees the first quarter period's crest before
    # anything reflected from the walls 200 cells away returns
    n, d = 400, 1e-9
    shape = (2, 2, n)
    dt = 0.5 * d / 3e8
    f, J0 = 5e14, 1e15
    k0, probe = n // 2, n // 2 + 40
    prob = vacuum(shape, (d, d, d), dt, periodic=(True, True, False), flags=FLAG_JF_INPUT, Jf_input_type=0,
                  Jf_input_component="x", Jf_input_amp=J0, Jf_input_freq=f, Jfin_lo=(0, 0, k0), Jfin_hi=(1, 1, k0))
    o = oracle_lib.Oracle(prob)
    peak = 0.0
    for i in range(700):
        o.step(1)
        Eg = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)
        peak = max(peak, np.abs(Eg[0][0, 0, probe]))
    eta0 = np.sqrt(MU0 / EPS0)
    assert abs(peak / (eta0 * J0 * d / 2) - 1) < 0.02
