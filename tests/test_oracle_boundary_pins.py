"""Pin the oracle's boundary restatements that the bulk closed forms leave
open, each by a check the mathematics fixes, and show that a planted
mistake in the oracle fails the check (the oracle is rebuilt with
GFO_PLANT = k, a compile-time switch that is 0 in the shipped build):

- the interfacial-DMI surface slope dm/dn = D/(2A) (-m_z, 0, m_x) on x
  faces and (0, -m_z, m_y) on y faces (magnetic_solver.cpp:1667-1672,
  :1704-1720): a chain in the rotation plane is torque free exactly when it
  is the uniform spiral with tan(dtheta) = -D dx/(2A) (the discrete
  Rohart-Thiaville condition, derived below), and a relaxed chain with an
  easy axis cants at its edges by sin(theta0) = D/(2 sqrt(A K)) with
  opposite senses on the two edges (Rohart and Thiaville, Phys. Rev. B 88,
  184422 (2013));
- the Liao multi-transmitting extrapolation at z = 0
  (elastodynamics_AbsorbingBoundary.cpp:3-230): along the characteristic
  (k, n - k) the formula is a fourth difference (weight w = 1) blended
  with a second difference (1 - w), so it reproduces any displacement
  history that is cubic in (k, n) at w = 1 and any linear one at every w;
  this fixes the coefficients and which step each history slot holds;
- the first-order Mur walls (EMdynamic_initialize.cpp:48-56,
  EMdynamic_AbsorbingBoundary.cpp:58-874): E_b' = E_i + C (E_i' - E_b)
  with C = (v dt_s - d)/(v dt_s + d) is exact for an outgoing wave linear in
  (x +- v t); the half stages use dt_s = dt/2 (coefficient with 2d), the
  full stage and the step dt_s = dt.  Checked on the stage stores and the
  state, on both walls of every axis, with a dielectric on the z walls.
  The elastic Mur walls (elastodynamics_AbsorbingBoundary.cpp:92-230 with
  if_1D_ABC) use the same formula on the displacement with the shear and
  longitudinal speeds of the bottom and top materials: a linear shear or
  compression wave leaving through either z wall is reproduced exactly.

Planted mistakes (gf_oracle.c, GFO_PLANT): 1 DMI slope sign, 2 DMI slope
D/A, 3 Liao history one step too old, 4 Liao t1/t2 coefficients swapped,
5 Mur half/full coefficients swapped, 6 Mur reading the base state for the
stage store, 7 elastic Mur half/full coefficients swapped.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_EM, FIELD_M, FIELD_U, FIELD_U_STORE, FIELD_V,
                     FLAG_1D_ABC, FLAG_ELASTODYNAMICS, FLAG_EMDYNAMIC, MU0, Material, Problem, split_yee, yee_shapes)
from goferro.workloads import GYRO, set_elastic
from test_oracle_em import EPS0, edge_coords, face_coords


@pytest.fixture(scope="module")
def planted(oracle_lib, tmp_path_factory):
    """Variant oracles, one per planted mistake (built in parallel)."""
    out = tmp_path_factory.mktemp("planted")
    with ThreadPoolExecutor(7) as ex:
        paths = list(ex.map(lambda k: oracle_lib.build_planted(k, out), range(1, 8)))
    return {k: p for k, p in zip(range(1, 8), paths)}


# ------------------------------------------------------------- DMI slope

def dmi_chain(axis, n, A=1e-11, D=3e-3, Ms=8e5, d=1e-9, K=0.0, alpha=0.5, dt=1e-14):
    """A chain of n ferromagnetic cells along x (axis 0) or y (axis 1) with
    free ends, interfacial DMI, an optional easy axis along z, no field."""
    shape = [1, 1, 1]
    shape[axis] = n
    mat = Material.make(if_FM=1, anisotropy_type=2, Ms=Ms, gyro=GYRO, FM_damping=alpha, K1=K, Aex=A, iDMI=D)
    per = [True, True, True]
    per[axis] = False
    return Problem(shape=tuple(shape), spacing=(d, d, d), materials=[mat], material_cell=np.ones(shape, np.uint32),
                   dt=dt, periodic=tuple(per), flags=0)


def spiral(axis, n, dtheta, theta0=0.3):
    th = theta0 + dtheta * np.arange(n)
    m = np.zeros((3, n))
    m[axis] = np.sin(th)
    m[2] = np.cos(th)
    shape = [1, 1, 1]
    shape[axis] = n
    return m.reshape((3,) + tuple(shape))


def spiral_torque(oracle_lib, axis, lib_path=None, n=24, A=1e-11, D=3e-3, Ms=8e5, d=1e-9, dt=1e-14):
    """Largest stage-1 increment of the discrete equilibrium spiral, in units
    of the increment a spiral angle step of D d/(2A) produces at an edge.

    With m_i = (sin t_i, 0, cos t_i) along x, the interior exchange and DMI
    fields are parallel to m_i for any uniform step dt_; at the first cell
    the surface slope g = D/(2A) (-cos t0, 0, sin t0) enters both, and the
    field along e_t = (cos t0, 0, -sin t0) is
    (2/(mu0 Ms)) [A (sin dt_/d + g_t)/d + D (cos dt_ - 1)/(2 d)] with
    g_t = -D/(2A), which vanishes iff A sin dt_/d = -(D/2) cos dt_, i.e.
    tan dt_ = -D d/(2A); the last cell gives the same condition."""
    prob = dmi_chain(axis, n, A=A, D=D, Ms=Ms, d=d, dt=dt)
    o = oracle_lib.Oracle(prob, lib_path)
    o.set(FIELD_M, spiral(axis, n, -np.arctan(D * d / (2 * A))))
    o.time_advance()
    o.L.gfo_mag_get_dm(o._h, 1)
    k1 = o.get(oracle_lib.FIELD_K1)
    o.close()
    alpha = 0.5
    scale = GYRO / (1 + alpha ** 2) * dt * 2 * A / (MU0 * Ms) * (D / (2 * A)) / d
    return float(np.abs(k1).max() / scale)


@pytest.mark.parametrize("axis", [0, 1])
def test_dmi_spiral_is_torque_free(oracle_lib, axis):
    assert spiral_torque(oracle_lib, axis) < 1e-9


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("plant", [1, 2])
def test_dmi_spiral_fails_planted_slope(oracle_lib, planted, axis, plant):
    assert spiral_torque(oracle_lib, axis, planted[plant]) > 1e-2


def edge_canting(oracle_lib, axis, lib_path=None):
    """Relax a 100-cell chain (d = 1 nm, wall width sqrt(A/K) = 10 nm) with an
    easy axis z at damping 1 and return (m_axis at the first cell, at the
    last cell, the continuum prediction at the cell centre d/2)."""
    A, K, D, Ms, d, n = 1e-11, 1e5, 0.6e-3, 1e6, 1e-9, 100
    prob = dmi_chain(axis, n, A=A, D=D, Ms=Ms, d=d, K=K, alpha=1.0, dt=1e-13)
    o = oracle_lib.Oracle(prob, lib_path)
    m0 = np.zeros((3,) + prob.shape)
    m0[2] = 1.0
    o.set(FIELD_M, m0)
    o.step(8000)
    m = o.get(FIELD_M).reshape(3, n)
    o.close()
    lw = np.sqrt(A / K)
    th0 = np.arcsin(D / (2 * np.sqrt(A * K)))
    th_c = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-(d / 2) / lw))
    return float(m[axis, 0]), float(m[axis, -1]), float(np.sin(th_c))


@pytest.mark.parametrize("axis", [0, 1])
def test_dmi_edge_canting_matches_rohart_thiaville(oracle_lib, axis):
    first, last, want = edge_canting(oracle_lib, axis)
    # discretisation error O((d / wall width)^2) = 1 %
    assert abs(first / want - 1) < 0.03, (first, want)
    assert abs(-last / want - 1) < 0.03, (last, want)


@pytest.mark.parametrize("plant", [1, 2])
def test_dmi_edge_canting_fails_planted_slope(oracle_lib, planted, plant):
    first, last, want = edge_canting(oracle_lib, 0, planted[plant])
    assert abs(first / want - 1) > 0.2 or abs(-last / want - 1) > 0.2


# ------------------------------------------------------------ Liao MTF

def liao_run(oracle_lib, w, cubic, lib_path=None, nsteps=12):
    """A stiffness-free column (the force vanishes, v is constant, so the
    displacement is u0(k) + n dt v(k) exactly); returns the worst error of
    the Liao plane u(k = 0) against that polynomial at steps n >= 4, in
    units of the largest displacement increment per step."""
    nz, d, dt = 8, 2e-9, 1e-13
    mat = set_elastic(Material.make(), 0.0, 0.0, 0.0, 5000.0)
    prob = Problem(shape=(1, 1, nz), spacing=(d, d, d), materials=[mat], material_cell=np.ones((1, 1, nz), np.uint32),
                   dt=dt, periodic=(True, True, False), flags=FLAG_ELASTODYNAMICS, weighting_factor_fourth_Liao=w)
    k = np.arange(nz, dtype=float)
    u0, v = np.zeros((3, 1, 1, nz)), np.zeros((3, 1, 1, nz))
    coef_u = [(1e-12, 3e-13, -2e-14, 5e-15), (-2e-12, 1e-13, 4e-14, -3e-15), (5e-13, -4e-13, 1e-14, 2e-15)]
    coef_v = [(3.0, -0.5, 0.04), (-1.0, 0.7, -0.03), (2.0, 0.2, 0.05)]
    for a in range(3):
        cu, cv = coef_u[a], coef_v[a]
        if cubic:
            u0[a, 0, 0] = cu[0] + cu[1] * k + cu[2] * k ** 2 + cu[3] * k ** 3
            v[a, 0, 0] = cv[0] + cv[1] * k + cv[2] * k ** 2
        else:
            u0[a, 0, 0] = cu[0] + cu[1] * k
            v[a, 0, 0] = cv[0] + 0.0 * k
    o = oracle_lib.Oracle(prob, lib_path)
    o.set(FIELD_U, u0)
    o.set(FIELD_V, v)
    worst = 0.0
    for n in range(1, nsteps + 1):
        o.step(1)
        u = o.get(FIELD_U)
        exact = u0 + n * dt * v
        # interior planes follow the polynomial (the force is zero)
        assert np.abs(u[..., 1:] - exact[..., 1:]).max() <= 1e-12 * np.abs(exact).max()
        if n >= 4:
            worst = max(worst, float(np.abs(u[..., 0] - exact[..., 0]).max() / (dt * np.abs(v).max())))
    o.close()
    return worst


def test_liao_exact_for_cubic_history_at_unit_weight(oracle_lib):
    assert liao_run(oracle_lib, 1.0, cubic=True) < 1e-9


@pytest.mark.parametrize("w", [0.0, 0.3, 0.5, 1.0])
def test_liao_exact_for_linear_history_at_any_weight(oracle_lib, w):
    assert liao_run(oracle_lib, w, cubic=False) < 1e-9


@pytest.mark.parametrize("plant", [3, 4])
def test_liao_fails_planted_mistake(oracle_lib, planted, plant):
    assert max(liao_run(oracle_lib, 1.0, True, planted[plant]), liao_run(oracle_lib, 0.5, False, planted[plant])) > 1e-3


# ---------------------------------------------------------------- Mur

# (axis, E component, H component, sign of H for a wave towards -axis);
# H = sign * f / eta for E = f(x_axis + v t), and -sign * f / eta for
# E = f(x_axis - v t) (derived from the two curl equations)
MUR_WAVES = [(2, 0, 1, -1), (0, 2, 1, +1), (1, 0, 2, +1)]


def mur_run(oracle_lib, axis, ec, hc, sgn, towards, er=1.0, lib_path=None, nsteps=6):
    """A linear wave E = a + b (s + towards... ) leaving through the wall
    at the low (towards = -1) or high (+1) end of `axis`; returns the worst
    error of that wall's E (stage stores after stages 1-3, the state after
    the step), in units of the change of E per step."""
    n, d = 48, 1e-9
    shape = [1, 1, 1]
    shape[axis] = n
    shape = tuple(shape)
    sp = (d, d, d)
    per = [True, True, True]
    per[axis] = False
    v = 1.0 / np.sqrt(MU0 * EPS0 * er)       # phase speed of the code's constants
    eta = np.sqrt(MU0 / (EPS0 * er))
    dt = 0.5 * d / v
    mats, cell = [], np.zeros(shape, np.uint32)
    if er != 1.0:
        mats = [Material.make(r_permittivity=er)]
        cell[:] = 1
    prob = Problem(shape=shape, spacing=sp, materials=mats, material_cell=cell, dt=dt, periodic=tuple(per),
                   flags=FLAG_EMDYNAMIC)
    E0, L = 1.0, n * d
    a, b = 0.3 * E0, E0 / L
    phase = -towards      # f(s - towards v t)

    def fields(t):
        Es = [np.zeros(s) for s in yee_shapes(shape, FIELD_E_EM)]
        Hs = [np.zeros(s) for s in yee_shapes(shape, FIELD_H_EM)]
        Es[ec] = a + b * (edge_coords(shape, ec, axis, sp) + phase * v * t)
        Hs[hc] = sgn * phase * (a + b * (face_coords(shape, hc, axis, sp) + phase * v * t)) / eta
        return Es, Hs

    Es, Hs = fields(0.0)
    o = oracle_lib.Oracle(prob, lib_path)
    o.set(FIELD_E_EM, Es)
    o.set(FIELD_H_EM, Hs)
    unit = b * v * dt
    worst = 0.0
    for st in range(nsteps):
        o.time_advance()
        for s in range(1, 5):
            o.em_stage(s)
            t = (st + (0.5 if s < 3 else 1.0)) * dt
            got = split_yee(o.get(FIELD_E_EM_STORE if s < 4 else FIELD_E_EM), shape, FIELD_E_EM)[ec]
            want = fields(t)[0][ec]
            idx = [slice(None)] * 3
            idx[axis] = 0 if towards < 0 else -1
            worst = max(worst, float(np.abs(got[tuple(idx)] - want[tuple(idx)]).max() / unit))
    o.close()
    return worst


@pytest.mark.parametrize("axis,ec,hc,sgn", MUR_WAVES)
@pytest.mark.parametrize("towards", [-1, +1])
def test_mur_exact_for_outgoing_linear_wave(oracle_lib, axis, ec, hc, sgn, towards):
    assert mur_run(oracle_lib, axis, ec, hc, sgn, towards) < 1e-7


@pytest.mark.parametrize("towards", [-1, +1])
def test_mur_exact_in_dielectric_on_z_walls(oracle_lib, towards):
    # the z-wall coefficients use c / sqrt(er) of the material on the wall plane
    assert mur_run(oracle_lib, 2, 0, 1, -1, towards, er=4.0) < 1e-7


@pytest.mark.parametrize("plant", [5, 6])
@pytest.mark.parametrize("axis,ec,hc,sgn", MUR_WAVES)
def test_mur_fails_planted_mistake(oracle_lib, planted, plant, axis, ec, hc, sgn):
    assert mur_run(oracle_lib, axis, ec, hc, sgn, -1, lib_path=planted[plant]) > 1e-2


# ------------------------------------------------------ elastic Mur walls

def elastic_mur_run(oracle_lib, comp, towards, lib_path=None, nsteps=6):
    """A uniform-strain wave u_comp = a + b (z - towards v t) (v the shear
    speed for x, y and the longitudinal one for z, of the wall's material)
    leaving through the bottom (towards = -1) or top (+1) z wall of a
    column under Mur walls; interior forces vanish (uniform strain), so
    the wall plane must follow the wave.  Returns the worst error of that
    plane (stage stores after stages 1-3, state after the step) in units of
    the change per step."""
    nz, d = 24, 2e-9
    c11, c12, c44, rho = 2.69e11, 1.08e11, 0.76e11, 5170.0
    mat = set_elastic(Material.make(), c11, c12, c44, rho)
    v = np.sqrt((c11 if comp == 2 else c44) / rho)
    dt = 0.4 * d / np.sqrt(c11 / rho)
    prob = Problem(shape=(1, 1, nz), spacing=(d, d, d), materials=[mat], material_cell=np.ones((1, 1, nz), np.uint32),
                   dt=dt, periodic=(True, True, False), flags=FLAG_ELASTODYNAMICS | FLAG_1D_ABC)
    z = (np.arange(nz) + 0.5) * d
    a, b = 1e-12, 1e-4
    phase = -towards

    def u_at(t):
        u = np.zeros((3, 1, 1, nz))
        u[comp, 0, 0] = a + b * (z + phase * v * t)
        return u

    vel = np.zeros((3, 1, 1, nz))
    vel[comp] = b * phase * v
    o = oracle_lib.Oracle(prob, lib_path)
    o.set(FIELD_U, u_at(0.0))
    o.set(FIELD_V, vel)
    unit = b * v * dt
    plane = 0 if towards < 0 else nz - 1
    worst = 0.0
    for st in range(nsteps):
        o.time_advance()
        for s in range(1, 5):
            o.elastic_stage(s)
            t = (st + (0.5 if s < 3 else 1.0)) * dt
            got = o.get(FIELD_U_STORE if s < 4 else FIELD_U)
            worst = max(worst, float(np.abs(got[comp, 0, 0, plane] - u_at(t)[comp, 0, 0, plane]) / unit))
    o.close()
    return worst


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("towards", [-1, +1])
def test_elastic_mur_exact_for_outgoing_linear_wave(oracle_lib, comp, towards):
    assert elastic_mur_run(oracle_lib, comp, towards) < 1e-9


@pytest.mark.parametrize("comp", [0, 2])
def test_elastic_mur_fails_planted_mistake(oracle_lib, planted, comp):
    assert elastic_mur_run(oracle_lib, comp, -1, planted[7]) > 1e-2
