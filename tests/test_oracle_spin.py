"""Pin the oracle's spin pumping (GF_FLAG_SPIN_PUMPING; get_J_ISHE_RK*,
magnetic_solver.cpp:231-283, set-up global.cpp:169-235):

* the planes: a pumping normal metal with a magnet on exactly one side
  carries the current on every plane, sourced from the magnet plane next
  to it; none when both or neither side is magnetic; the layers come from
  nz_layer (two adjacent layers of one material stay two layers);
* the current is the inverse spin Hall relation J = c(k) z x (m x dm) / dt
  of the source plane (numpy cross products), with
  c(k) = theta_SH e g_mix / (2 pi) * s sinh((t - d) / l_sd) / sinh(t / l_sd),
  d = (|k - sz| - 1/2) dz, s = +1 above the source and -1 below: the
  spin-diffusion profile that vanishes at the far surface of the layer;
* in Ampere's law it is a current density: starting from E = H = 0 with a
  tiny dt, one step changes an edge's E by -dt J / (eps0 er) (the mean of
  the four cells around it).
* at a step boundary the current is get_J_ISHE's (magnetic_solver.cpp:
  1460-1511, called from update_m before m += dm): the state before the
  update crossed with the RK-combined increment; with an FM and an AFM the
  AFM update's call stands, which sees the FM state already updated.
"""
import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_J_ISHE, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FLAG_EMDYNAMIC,
                     FLAG_SPIN_PUMPING, Material, Problem, split_yee, workloads)
from goferro.workloads import GYRO

E_CHARGE = 1.602176634e-19
EPS0 = 8.854187817e-12
EM_MAT = dict(comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0, tao_e_n3=1.0)
SHA, GMIX, LSD = 0.08, 5e18, 3e-9


def metal():
    return Material.make(if_spin_pump=1, spin_hall_angle=SHA, spin_mix_cond=GMIX, spin_diffus_length=LSD,
                         r_permittivity=2.0, **EM_MAT)


def magnet():
    return Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, K1=-610.0,
                         Aex=3.7e-12, r_permittivity=5.0, **EM_MAT)


def stack(column, shape_xy=(4, 3), dz=1e-9, flags=FLAG_SPIN_PUMPING, nz_layer=None, mats=None):
    """A film whose z column follows `column` (material id per plane)."""
    col = np.asarray(column, np.uint32)
    cell = np.broadcast_to(col, shape_xy + (col.size,)).copy()
    mats = mats or [magnet(), metal()]
    return Problem(shape=shape_xy + (col.size,), spacing=(1.5e-9, 2e-9, dz), materials=mats, material_cell=cell,
                   dt=1e-15, periodic=(True, True, False), flags=flags, Hext_stat=(3e4, -1e4, 8e4),
                   nz_layer=nz_layer)


def coef(k, sz, t, dz):
    d = (abs(k - sz) - 0.5) * dz
    s = 1.0 if k > sz else -1.0
    return SHA * E_CHARGE * GMIX / (2 * np.pi) * s * np.sinh((t - d) / LSD) / np.sinh(t / LSD)


def stage1_current(oracle_lib, prob, m, afm=None):
    o = oracle_lib.Oracle(prob)
    if m is not None:
        o.set(FIELD_M, m)
    if afm is not None:
        o.set(FIELD_M_AFM1, afm[0])
        o.set(FIELD_M_AFM2, afm[1])
    o.L.gfo_time_advance(o._h)
    dm = None
    if m is not None:
        o.L.gfo_mag_get_dm(o._h, 1)
        dm = o.get(103)                       # the stage-1 increment of m
    if afm is not None:
        o.L.gfo_afm_get_dm(o._h, 1)
    J = o.get(FIELD_J_ISHE)
    o.close()
    return J, dm


def expected(m, dm, planes, dt):
    """planes: {k: (sz, c)}; J = c z x (m x dm) / dt of the source plane."""
    J = np.zeros_like(m)
    for k, (sz, c) in planes.items():
        s = np.cross(m[:, :, :, sz], dm[:, :, :, sz], axis=0)
        J[0, :, :, k] = -c * s[1] / dt
        J[1, :, :, k] = c * s[0] / dt
    return J


@pytest.mark.parametrize("below", [True, False])
def test_current_is_ishe_of_the_source_plane(oracle_lib, below):
    # vacuum 2 | magnet 3 | metal 4 | vacuum 2, and the mirror image
    col = [0, 0, 1, 1, 1, 2, 2, 2, 2, 0, 0]
    if not below:
        col = col[::-1]
    prob = stack(col)
    m = workloads.tilted_m(prob.shape, 3, theta_max=0.7)
    J, dm = stage1_current(oracle_lib, prob, m)
    dz = prob.spacing[2]
    if below:
        planes = {k: (4, coef(k, 4, 4 * dz, dz)) for k in range(5, 9)}
    else:
        planes = {k: (6, coef(k, 6, 4 * dz, dz)) for k in range(2, 6)}
    want = expected(m, dm, planes, prob.dt)
    scale = np.abs(want).max()
    assert scale > 0
    np.testing.assert_allclose(J, want, rtol=0, atol=1e-13 * scale)
    # the profile: largest next to the magnet, sinh(dz / 2 / l) / sinh(t / l) of it at the far plane
    near, far = (5, 8) if below else (5, 2)
    r = np.abs(J[:2, :, :, far]).max() / np.abs(J[:2, :, :, near]).max()
    assert r == pytest.approx(np.sinh(0.5 * dz / LSD) / np.sinh(3.5 * dz / LSD), rel=1e-9)


def test_no_current_between_two_magnets_or_none(oracle_lib):
    for col in ([1, 1, 2, 2, 1, 1], [0, 2, 2, 2, 0, 1]):
        prob = stack(col)
        m = workloads.tilted_m(prob.shape, 5, theta_max=0.7)
        J, _ = stage1_current(oracle_lib, prob, m)
        assert not J.any(), col


def test_layers_from_nz_layer(oracle_lib):
    # magnet 2 | metal 2 | metal 3: given as three layers, only the first
    # metal layer touches the magnet; as runs, the metal is one layer of 5
    col = [1, 1, 2, 2, 2, 2, 2]
    dz = 1e-9
    m = workloads.tilted_m((4, 3, 7), 6, theta_max=0.7)
    J3, dm = stage1_current(oracle_lib, stack(col, nz_layer=[2, 2, 3]), m)
    J1, _ = stage1_current(oracle_lib, stack(col), m)
    np.testing.assert_allclose(J3, expected(m, dm, {k: (1, coef(k, 1, 2 * dz, dz)) for k in (2, 3)}, 1e-15),
                               rtol=0, atol=1e-13 * np.abs(J3).max())
    np.testing.assert_allclose(J1, expected(m, dm, {k: (1, coef(k, 1, 5 * dz, dz)) for k in range(2, 7)}, 1e-15),
                               rtol=0, atol=1e-13 * np.abs(J1).max())


def test_afm_sublattices_add(oracle_lib):
    """Both sublattices pump: J = c z x (m1 x dm1 + m2 x dm2) / dt."""
    prob0, a1, a2 = workloads.afm_block(shape=(4, 3, 6))
    mats = [prob0.materials[0], metal()]
    prob = stack([1, 1, 1, 2, 2, 0], mats=mats)
    cell = np.asarray(prob.material_cell)
    m1, m2 = a1.copy(), a2.copy()
    m1[:, cell != 1] = 0.0
    m2[:, cell != 1] = 0.0
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M_AFM1, m1)
    o.set(FIELD_M_AFM2, m2)
    o.L.gfo_time_advance(o._h)
    o.L.gfo_afm_get_dm(o._h, 1)
    J = o.get(FIELD_J_ISHE)
    d1, d2 = o.get(104), o.get(105)
    o.close()
    dz = prob.spacing[2]
    planes = {k: (2, coef(k, 2, 2 * dz, dz)) for k in (3, 4)}
    want = expected(m1, d1, planes, prob.dt) + expected(m2, d2, planes, prob.dt)
    np.testing.assert_allclose(J, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_ampere_source_sign_and_mean(oracle_lib):
    """E = H = 0, one tiny step: dE_x = -dt <J_x> / (eps0 <er>) per edge."""
    col = [1, 1, 1, 2, 2, 2, 0, 0]
    prob = stack(col, flags=FLAG_SPIN_PUMPING | FLAG_EMDYNAMIC)
    prob.periodic = (True, True, True)
    prob.dt = 1e-21
    m = workloads.tilted_m(prob.shape, 8, theta_max=0.7)
    m[:, np.asarray(prob.material_cell) != 1] = 0.0
    J, _ = stage1_current(oracle_lib, prob, m)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m)
    o.step(1)
    Ex = split_yee(o.get(FIELD_E_EM), prob.shape, FIELD_E_EM)[0]
    o.close()
    nx, ny, nz = prob.shape
    er_of = {0: 1.0, 1: 5.0, 2: 2.0}
    cell = np.asarray(prob.material_cell)
    worst = 0.0
    for j in range(ny + 1):
        for k in range(nz + 1):
            y1, y3 = (j - 1) % ny, j % ny
            z1, z2 = (k - 1) % nz, k % nz
            cells = [(0, y1, z1), (0, y1, z2), (0, y3, z1), (0, y3, z2)]
            jm = np.mean([J[0][c] for c in cells])
            er = np.mean([er_of[int(cell[c])] for c in cells])
            want = -prob.dt * jm / (EPS0 * er)
            worst = max(worst, abs(Ex[0, j, k] - want))
    scale = prob.dt * np.abs(J[0]).max() / EPS0
    assert scale > 0 and worst <= 1e-6 * scale



def step_and_stage4_current(oracle_lib, prob, fields):
    """One step stage by stage: the current after stage 4's get phase (the
    stage current) and after its update phase (the step current)."""
    o = oracle_lib.Oracle(prob)
    for fid, v in fields.items():
        o.set(fid, v)
    fm, afm = FIELD_M in fields, FIELD_M_AFM1 in fields
    o.time_advance()
    for s in range(1, 5):
        if fm:
            o.L.gfo_mag_get_dm(o._h, s)
        if afm:
            o.L.gfo_afm_get_dm(o._h, s)
        if s == 4:
            J4 = o.get(FIELD_J_ISHE)
        if fm:
            o.L.gfo_mag_update(o._h, s)
        if afm:
            o.L.gfo_afm_update(o._h, s)
    out = {"J4": J4, "J": o.get(FIELD_J_ISHE), "dm": o.get(100), "dm1": o.get(109), "dm2": o.get(110)}
    o.close()
    return out


def test_step_boundary_current_uses_base_state_and_combined_increment(oracle_lib):
    col = [0, 1, 1, 1, 2, 2, 2, 0]
    prob = stack(col)
    m0 = workloads.tilted_m(prob.shape, 11, theta_max=0.7)
    m0[:, np.asarray(prob.material_cell) != 1] = 0.0
    r = step_and_stage4_current(oracle_lib, prob, {FIELD_M: m0})
    dz = prob.spacing[2]
    planes = {k: (3, coef(k, 3, 3 * dz, dz)) for k in (4, 5, 6)}
    want = expected(m0, r["dm"], planes, prob.dt)
    scale = np.abs(want).max()
    assert scale > 0
    np.testing.assert_allclose(r["J"], want, rtol=0, atol=1e-13 * scale)
    # the stage-4 current (stage-3 store and k4) is a different quantity
    assert np.abs(r["J4"] - want).max() > 1e-3 * scale


def test_step_boundary_current_of_afm_sublattices(oracle_lib):
    """Metal under an AFM, an FM elsewhere in the stack: the current is the
    AFM states before the update crossed with their combined increments
    (the FM update's call is overwritten by the AFM update's)."""
    prob0, a1, a2 = workloads.afm_block(shape=(4, 3, 8))
    mats = [magnet(), metal(), prob0.materials[0]]
    col = [0, 2, 2, 3, 3, 0, 1, 1]
    prob = stack(col, mats=mats)
    cell = np.asarray(prob.material_cell)
    m = workloads.tilted_m(prob.shape, 13, theta_max=0.7)
    m[:, cell != 1] = 0.0
    s1, s2 = a1.copy(), a2.copy()
    s1[:, cell != 3] = 0.0
    s2[:, cell != 3] = 0.0
    r = step_and_stage4_current(oracle_lib, prob, {FIELD_M: m, FIELD_M_AFM1: s1, FIELD_M_AFM2: s2})
    dz = prob.spacing[2]
    planes = {k: (3, coef(k, 3, 2 * dz, dz)) for k in (1, 2)}
    want = expected(s1, r["dm1"], planes, prob.dt) + expected(s2, r["dm2"], planes, prob.dt)
    scale = np.abs(want).max()
    assert scale > 0
    np.testing.assert_allclose(r["J"], want, rtol=0, atol=1e-13 * scale)
    assert np.abs(r["J4"] - want).max() > 1e-3 * scale
