"""GPU parity of the elastodynamics stage and its couplings to m and p,
through the C ABI, against the CPU oracle on the same seeded inputs.

Displacement and velocity are compared relative to their peak (1e-12 per
stage, 1e-9 after 1000 steps, the parity bounds); m absolutely (|m| ~ 1).
Every case prints its name and shape first so a failing GPU step names
itself.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE, FIELD_P, FIELD_P_STORE, FIELD_Q,
                     FIELD_STRAIN_DIAG, FIELD_STRAIN_OFF, FIELD_U, FIELD_U_STORE, FIELD_V, FLAG_1D_ABC,
                     FLAG_1D_ABC_ELAST_ONLYTOP, FLAG_ELASTO_BACKACTION, FLAG_ELASTO_ON_PM, FLAG_ELASTODYNAMICS,
                     FLAG_GAUSSIAN_PULSE, FLAG_ORIENT_ROTATE, Context, Material, Problem, workloads)
from goferro.workloads import set_elastic

pytestmark = pytest.mark.gpu

C11, C12, C44, RHO = 2.69e11, 1.08e11, 0.76e11, 5170.0


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rot():
    c, s = np.cos(0.4), np.sin(0.4)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    c2, s2 = np.cos(0.9), np.sin(0.9)
    Rx = np.array([[1, 0, 0], [0, c2, -s2], [0, s2, c2]])
    return Rx @ Rz


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def random_uv(shape, seed):
    rng = np.random.default_rng(seed)
    return 1e-11 * rng.standard_normal((3,) + tuple(shape)), 2.0 * rng.standard_normal((3,) + tuple(shape))


def elastic_block(shape=(12, 10, 16), rotate=False, flags=0, source=None, periodic=(True, True, False)):
    """Two elastic materials with damping, vacuum pockets away from the
    k = 0 plane, an optional rotation and displacement source."""
    rng = np.random.default_rng(21)
    T = rot() if rotate else None
    a = set_elastic(Material.make(), C11, C12, C44, RHO, Tg2c=T, mass_damping=2e9, stiff_damping=3e-15)
    b = set_elastic(Material.make(), 1.6e11, 0.7e11, 0.5e11, 7800.0, Tg2c=T, stiff_damping=1e-15)
    cell = rng.choice(np.array([0, 1, 2], np.uint32), size=shape, p=[0.08, 0.6, 0.32])
    cell[:, :, 0] = 1
    kw = {}
    if source is not None:
        kind, plane, comp = source
        flags |= FLAG_GAUSSIAN_PULSE
        kw = dict(input_strain_type=kind, input_strain_sourcez=plane, input_strain_component=comp,
                  sigma_gauss=(4e-13 if kind != 2 else 3e11), amplitude_gauss=(1e-3 if kind != 3 else 2e8),
                  num_strain_cycles=2)
    return Problem(shape=shape, spacing=(2e-9, 2.5e-9, 2e-9), materials=[a, b], material_cell=cell, dt=1e-14,
                   periodic=periodic, flags=FLAG_ELASTODYNAMICS | flags | (FLAG_ORIENT_ROTATE if rotate else 0),
                   Tg2c=T, weighting_factor_fourth_Liao=0.5, **kw)


def run_stages(prob, fields, nsteps, phys=("el",), check=()):
    """Stage by stage on both builds in the reference's order (all get
    phases, then all updates); compares the u store and the listed stores
    after every stage."""
    from oracle.oracle import Oracle
    say("case", phys, prob.shape, hex(prob.flags))
    g, o = Context(prob), Oracle(prob)
    for fid, val in fields.items():
        g.upload(fid, val)
        o.set(fid, val)
    L, h = o.L, o._h
    gets = {"el": L.gfo_elastic_get, "fm": L.gfo_mag_get_dm, "afm": L.gfo_afm_get_dm, "fe": L.gfo_fe_get}
    upds = {"el": L.gfo_elastic_update, "fm": L.gfo_mag_update, "afm": L.gfo_afm_update, "fe": L.gfo_fe_update}
    calls = {"el": g.elastic_stage, "fm": g.mag_stage, "afm": g.afm_stage, "fe": g.fe_stage}
    for step in range(nsteps):
        g.time_advance()
        o.time_advance()
        for s in range(1, 5):
            for p in phys:
                calls[p](s)
            for p in phys:
                gets[p](h, s)
            for p in phys:
                upds[p](h, s)
            d = rel(g.download(FIELD_U_STORE), o.get(FIELD_U_STORE))
            assert d <= 1e-12, ("u store", step, s, d)
            for fid in check:
                d = np.abs(g.download(fid) - o.get(fid)).max()
                assert d <= 1e-12, (fid, step, s, d)
    for fid, tol in ((FIELD_U, 1e-12), (FIELD_V, 1e-12), (FIELD_STRAIN_DIAG, 1e-12), (FIELD_STRAIN_OFF, 1e-12)):
        d = rel(g.download(fid), o.get(fid))
        assert d <= tol, (fid, d)
    assert abs(g.time - o.time) <= 1e-30
    return g, o


@pytest.mark.parametrize("rotate", [False, True])
def test_elastic_stages_liao_with_source(rotate):
    prob = elastic_block(rotate=rotate, source=(1, 15, "x"))
    u, v = random_uv(prob.shape, 3)
    g, o = run_stages(prob, {FIELD_U: u, FIELD_V: v}, nsteps=3)
    g.close()
    o.close()


@pytest.mark.parametrize("flags,source", [(FLAG_1D_ABC, (2, 0, "z")),
                                          (FLAG_1D_ABC | FLAG_1D_ABC_ELAST_ONLYTOP, (3, 11, "y")),
                                          (FLAG_1D_ABC, (3, 11, "z"))])
def test_elastic_stages_mur(flags, source):
    prob = elastic_block(shape=(8, 8, 12), flags=flags, source=source)
    u, v = random_uv(prob.shape, 4)
    g, o = run_stages(prob, {FIELD_U: u, FIELD_V: v}, nsteps=3)
    g.close()
    o.close()


def test_elastic_periodic_z_cube():
    # periodic z is supported where the reference's wrap test (k == ny-1)
    # is the periodic one: ny == nz
    prob = elastic_block(shape=(6, 8, 8), periodic=(True, True, True), rotate=True)
    u, v = random_uv(prob.shape, 5)
    g, o = run_stages(prob, {FIELD_U: u, FIELD_V: v}, nsteps=2)
    g.close()
    o.close()


def magnetoelastic(shape=(16, 16, 12), layers=4, rotate=False):
    prob, m = workloads.magnetoelastic_film(shape=shape, magnet_layers=layers)
    if rotate:
        T = rot()
        for mat in prob.materials:
            set_elastic(mat, mat.c11, mat.c12, mat.c44, mat.density, Tg2c=T)
        prob.materials[0].anisotropy_type = 1
        prob.materials[0].K1 = -610.0
        prob.flags |= FLAG_ORIENT_ROTATE
        prob.Tg2c = T
    return prob, m


@pytest.mark.parametrize("rotate", [False, True])
def test_magnetoelastic_stages(rotate):
    prob, m = magnetoelastic(rotate=rotate)
    u, v = random_uv(prob.shape, 6)
    g, o = run_stages(prob, {FIELD_M: m, FIELD_U: u, FIELD_V: v}, nsteps=2, phys=("el", "fm"),
                      check=(FIELD_M_STORE,))
    np.testing.assert_allclose(g.download(FIELD_M), o.get(FIELD_M), rtol=0, atol=1e-12)
    g.close()
    o.close()


def test_magnetoelastic_1000_steps():
    """The parity bounds after 1000 coupled steps through gf_step."""
    from oracle.oracle import Oracle
    prob, m = magnetoelastic(shape=(8, 8, 8), layers=3)
    u, v = random_uv(prob.shape, 7)
    u *= 0.1
    say("case magnetoelastic 1000 steps", prob.shape)
    g, o = Context(prob), Oracle(prob)
    for fid, val in ((FIELD_M, m), (FIELD_U, u), (FIELD_V, v)):
        g.upload(fid, val)
        o.set(fid, val)
    g.step(1000)
    o.step(1000)
    a, b = g.download(FIELD_M), o.get(FIELD_M)
    assert np.abs(a - b).max() <= 1e-9
    assert np.abs(np.sqrt((a * a).sum(0)) - np.sqrt((b * b).sum(0))).max() <= 1e-12
    assert rel(g.download(FIELD_U), o.get(FIELD_U)) <= 1e-9
    assert rel(g.download(FIELD_V), o.get(FIELD_V)) <= 1e-9


def fe_elastic(shape=(8, 8, 16), rotate=False):
    from goferro.workloads import pto_material
    T = rot() if rotate else None
    fe = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    fe.Q11, fe.Q12, fe.Q44 = 0.089, -0.026, 0.0675
    fe.tv1, fe.tv2, fe.tv3 = 3.05e10, 6.4e8, 2.97e10
    set_elastic(fe, 2.3e11, 1.0e11, 1.1e11, 7900.0, Tg2c=T)
    sub = set_elastic(Material.make(), 3.2e11, 1.0e11, 1.2e11, 5100.0, Tg2c=T, mass_damping=1e9)
    cell = np.zeros(shape, np.uint32)
    cell[:, :, :5] = 2
    cell[:, :, 5:12] = 1
    prob = Problem(shape=shape, spacing=(1e-9, 1e-9, 1e-9), materials=[fe, sub], material_cell=cell, dt=2e-15,
                   periodic=(True, True, False),
                   flags=FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM |
                   (FLAG_ORIENT_ROTATE if rotate else 0), Tg2c=T, Eext_stat=(0, 0, 1e6))
    rng = np.random.default_rng(31)
    p = 0.05 * rng.standard_normal((3,) + shape)
    p[2] += 0.7
    p[:, cell != 1] = 0.0
    q = 1e8 * rng.standard_normal((3,) + shape)
    q[:, cell != 1] = 0.0
    return prob, p, q


@pytest.mark.parametrize("rotate", [False, True])
def test_ferroelectric_elastic_stages(rotate):
    prob, p, q = fe_elastic(rotate=rotate)
    u, v = random_uv(prob.shape, 8)
    g, o = run_stages(prob, {FIELD_P: p, FIELD_Q: q, FIELD_U: u, FIELD_V: v}, nsteps=2, phys=("el", "fe"))
    for fid in (FIELD_P_STORE, FIELD_P, FIELD_Q):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.close()
    o.close()


def test_afm_elastic_stages():
    prob, m1, m2 = workloads.afm_block(shape=(8, 8, 8))
    mat = prob.materials[0]
    set_elastic(mat, C11, C12, C44, RHO)
    mat.B1_AFM1 = mat.B1_AFM2 = 6e5
    mat.B2_AFM1 = mat.B2_AFM2 = 4e5
    mat.lamda100_AFM1 = mat.lamda100_AFM2 = -2e-5
    mat.lamda111_AFM1 = mat.lamda111_AFM2 = 1e-5
    prob.flags |= FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM
    u, v = random_uv(prob.shape, 9)
    g, o = run_stages(prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2, FIELD_U: u, FIELD_V: v}, nsteps=2,
                      phys=("el", "afm"))
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
        np.testing.assert_allclose(g.download(fid), o.get(fid), rtol=0, atol=1e-12)
    g.close()
    o.close()


def random_e0t0(shape, seed, scale=2e-5):
    rng = np.random.default_rng(seed)
    e = scale * rng.standard_normal((6,) + tuple(shape))
    return {FIELD_EIGENSTRAIN_T0_DIAG: e[:3], FIELD_EIGENSTRAIN_T0_OFF: e[3:]}


@pytest.mark.parametrize("rotate", [False, True])
def test_magnetoelastic_stages_with_static_eigenstrain(rotate):
    """The static eigenstrain at t0 (ex*0t0_crt) subtracted from the
    store eigenstrain (elastic_dynamic_solver.cpp:39-65), stage by stage."""
    prob, m = magnetoelastic(rotate=rotate)
    u, v = random_uv(prob.shape, 16)
    fields = {FIELD_M: m, FIELD_U: u, FIELD_V: v, **random_e0t0(prob.shape, 17)}
    g, o = run_stages(prob, fields, nsteps=2, phys=("el", "fm"), check=(FIELD_M_STORE,))
    np.testing.assert_allclose(g.download(FIELD_M), o.get(FIELD_M), rtol=0, atol=1e-12)
    for fid in (FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF):
        np.testing.assert_array_equal(g.download(fid), fields[fid])
    g.close()
    o.close()


@pytest.mark.parametrize("rotate", [False, True])
def test_magnetoelastic_stages_with_static_strain(rotate):
    """A static strain uploaded beside the elastodynamics: the magnets read
    exxt0 + Dexx (magnetic_solver.cpp:89-96), stage by stage."""
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF
    prob, m = magnetoelastic(rotate=rotate)
    u, v = random_uv(prob.shape, 36)
    rng = np.random.default_rng(37)
    e = 3e-5 * rng.standard_normal((6,) + prob.shape)
    fields = {FIELD_M: m, FIELD_U: u, FIELD_V: v, FIELD_STRAIN_T0_DIAG: e[:3], FIELD_STRAIN_T0_OFF: e[3:]}
    g, o = run_stages(prob, fields, nsteps=2, phys=("el", "fm"), check=(FIELD_M_STORE,))
    np.testing.assert_allclose(g.download(FIELD_M), o.get(FIELD_M), rtol=0, atol=1e-12)
    g.close()
    o.close()


def test_afm_elastic_stages_with_static_strain():
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF
    prob, m1, m2 = workloads.afm_block(shape=(8, 8, 8))
    mat = prob.materials[0]
    set_elastic(mat, C11, C12, C44, RHO)
    mat.B1_AFM1 = mat.B1_AFM2 = 6e5
    mat.B2_AFM1 = mat.B2_AFM2 = 4e5
    mat.lamda100_AFM1 = mat.lamda100_AFM2 = -2e-5
    mat.lamda111_AFM1 = mat.lamda111_AFM2 = 1e-5
    prob.flags |= FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM
    u, v = random_uv(prob.shape, 38)
    e = 3e-5 * np.random.default_rng(39).standard_normal((6,) + prob.shape)
    fields = {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2, FIELD_U: u, FIELD_V: v, FIELD_STRAIN_T0_DIAG: e[:3],
              FIELD_STRAIN_T0_OFF: e[3:]}
    g, o = run_stages(prob, fields, nsteps=2, phys=("el", "afm"))
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
        np.testing.assert_allclose(g.download(fid), o.get(fid), rtol=0, atol=1e-12)
    g.close()
    o.close()


def test_ferroelectric_elastic_stages_with_static_eigenstrain():
    """The ferroelectric reads the same eigenstrain back (its elastic driving
    force, ferroelectric_solver.cpp:630-640)."""
    prob, p, q = fe_elastic()
    u, v = random_uv(prob.shape, 18)
    fields = {FIELD_P: p, FIELD_Q: q, FIELD_U: u, FIELD_V: v, **random_e0t0(prob.shape, 19, scale=1e-3)}
    g, o = run_stages(prob, fields, nsteps=2, phys=("el", "fe"))
    for fid in (FIELD_P_STORE, FIELD_P, FIELD_Q):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.close()
    o.close()


def test_static_eigenstrain_cancels_a_frozen_magnet():
    """With the eigenstrain of a frozen uniform m as the static one the film
    stays at rest through gf_step; without it the top surface moves."""
    from goferro import FLAG_GAUSSIAN_PULSE
    say("case static eigenstrain, frozen magnet")
    prob, _ = workloads.magnetoelastic_film(shape=(8, 8, 12), magnet_layers=4)
    prob.flags &= ~FLAG_GAUSSIAN_PULSE
    mag = prob.materials[0]
    mag.gyro = 0.0
    mdir = np.array([0.36, -0.48, 0.8])
    magnet = np.asarray(prob.material_cell) == 1
    m = np.zeros((3,) + prob.shape)
    m[:, magnet] = mdir[:, None]
    mx, my, mz = mdir
    l100, l111 = mag.lamda100, mag.lamda111
    e0 = np.zeros((6,) + prob.shape)
    e0[:, magnet] = np.array([1.5 * l100 * (mx * mx - 1.0 / 3.0), 1.5 * l100 * (my * my - 1.0 / 3.0),
                              1.5 * l100 * (mz * mz - 1.0 / 3.0), 1.5 * l111 * my * mz, 1.5 * l111 * mx * mz,
                              1.5 * l111 * my * mx])[:, None]
    peaks = []
    for t0 in (None, e0):
        g = Context(prob)
        g.upload(FIELD_M, m)
        if t0 is not None:
            g.upload(FIELD_EIGENSTRAIN_T0_DIAG, t0[:3])
            g.upload(FIELD_EIGENSTRAIN_T0_OFF, t0[3:])
        g.step(20)
        peaks.append(np.abs(g.download(FIELD_V)).max())
        g.close()
    say("  peak |v| without, with", peaks)
    assert peaks[0] > 1e-6
    assert peaks[1] <= 1e-12 * peaks[0]


def test_stage_order_is_enforced():
    prob, m = magnetoelastic(shape=(8, 8, 8), layers=3)
    say("case stage order", prob.shape)
    g = Context(prob)
    g.upload(FIELD_M, m)
    g.time_advance()
    with pytest.raises(Exception, match="gf_elastic_stage"):
        g.mag_stage(1)
    g.elastic_stage(1)
    g.mag_stage(1)
    with pytest.raises(Exception, match="gf_elastic_stage"):
        g.mag_stage(2)
    g.close()
