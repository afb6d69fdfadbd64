"""GPU parity of the electrodynamics stage (Yee E/H, Mur/PEC walls, plane
and current sources, Drude currents) and its couplings to p and m, through
the C ABI, against the CPU oracle on the same seeded inputs.

E and H are compared relative to their peak, 1e-12 per stage and 1e-9
after 1000 steps (the parity bounds); p and m likewise.  Every case prints
its name first so a failing GPU step names itself.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_E_CELL, FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_CELL, FIELD_H_EM, FIELD_H_EM_STORE,
                     FIELD_JP, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE, FIELD_P, FIELD_P_STORE, FIELD_Q,
                     FLAG_1D_ABC, FLAG_1D_ABC_EM_ONLYTOP, FLAG_EM_BACKACTION_M, FLAG_EM_BACKACTION_P, FLAG_EM_FROM_M,
                     FLAG_EM_FROM_P, FLAG_EMDYNAMIC, FLAG_JF_INPUT, FLAG_MAG_1D, FLAG_MAGNETOSTATICS, FLAG_PEC_XY,
                     FLAG_PEC_Z, FLAG_PLANE_EM, Context, Material, Problem, workloads, yee_shapes)
from goferro.workloads import GYRO

pytestmark = pytest.mark.gpu

EM_MAT = dict(comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0, tao_e_n3=1.0)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def random_yee(shape, field, scale, seed):
    rng = np.random.default_rng(seed)
    return [scale * rng.standard_normal(s) for s in yee_shapes(shape, field)]


def run_stages(prob, fields, nsteps, phys, check=(), tol=1e-12):
    from oracle.oracle import Oracle
    say("case", phys, prob.shape, hex(prob.flags))
    g, o = Context(prob), Oracle(prob)
    for fid, val in fields.items():
        g.upload(fid, val)
        o.set(fid, val)
    L, h = o.L, o._h
    gets = {"fm": L.gfo_mag_get_dm, "afm": L.gfo_afm_get_dm, "fe": L.gfo_fe_get, "em": L.gfo_em_get}
    upds = {"fm": L.gfo_mag_update, "afm": L.gfo_afm_update, "fe": L.gfo_fe_update, "em": L.gfo_em_update}
    calls = {"fm": g.mag_stage, "afm": g.afm_stage, "fe": g.fe_stage, "em": g.em_stage}
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
            for fid in (FIELD_E_EM_STORE, FIELD_H_EM_STORE, FIELD_E_CELL, FIELD_H_CELL) + tuple(check):
                d = rel(g.download(fid), o.get(fid))
                assert d <= tol, (fid, step, s, d)
    for fid in (FIELD_E_EM, FIELD_H_EM):
        assert rel(g.download(fid), o.get(fid)) <= tol, fid
    assert abs(g.time - o.time) <= 1e-30
    return g, o


def fe_film(shape=(4, 4, 96), flags=0, source=True):
    """Config-3-like column: vacuum / PbTiO3-like film / vacuum, Mur walls in
    z, a Gaussian-modulated THz current sheet in the lower vacuum."""
    from goferro.workloads import pto_material
    fe = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    fe.r_permittivity = 10.0
    for k, v in EM_MAT.items():
        setattr(fe, k, v)
    nz = shape[2]
    cell = np.zeros(shape, np.uint32)
    cell[:, :, nz // 3: 2 * nz // 3] = 1
    kw = {}
    if source:
        flags |= FLAG_JF_INPUT
        kw = dict(Jf_input_type=6, Jf_input_component="x", Jf_input_amp=5.3e16, Jf_input_freq=1e12,
                  Jf_input_sigma=2e-17, Jfin_lo=(0, 0, nz // 6), Jfin_hi=(shape[0] - 1, shape[1] - 1, nz // 6))
    prob = Problem(shape=shape, spacing=(1e-9, 1e-9, 1e-9), materials=[fe], material_cell=cell, dt=1e-18,
                   periodic=(True, True, False),
                   flags=FLAG_EMDYNAMIC | FLAG_EM_FROM_P | FLAG_EM_BACKACTION_P | flags, **kw)
    rng = np.random.default_rng(3)
    p = np.zeros((3,) + shape)
    p[2] = 0.754 * (1 + 0.01 * rng.standard_normal(shape))
    p[:, cell == 0] = 0.0
    q = 1e9 * rng.standard_normal((3,) + shape)
    q[:, cell == 0] = 0.0
    return prob, p, q


@pytest.mark.parametrize("flags", [0, FLAG_1D_ABC | FLAG_1D_ABC_EM_ONLYTOP, FLAG_PEC_Z])
def test_fe_em_stages(flags):
    prob, p, q = fe_film(flags=flags)
    E = random_yee(prob.shape, FIELD_E_EM, 1e6, 4)
    H = random_yee(prob.shape, FIELD_H_EM, 3e3, 5)
    g, o = run_stages(prob, {FIELD_P: p, FIELD_Q: q, FIELD_E_EM: E, FIELD_H_EM: H}, nsteps=3, phys=("fe", "em"),
                      check=(FIELD_P_STORE,))
    for fid in (FIELD_P, FIELD_Q):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.close()
    o.close()


def boxed_media(shape=(8, 10, 12), flags=0, jf_type=1, plane=True):
    """Non-periodic in every direction: Mur (or PEC) walls on all six faces,
    a conductor, a Drude metal with two poles, a PEC inclusion, a plane
    source and a current box with a spatial profile."""
    rng = np.random.default_rng(11)
    cond = Material.make(r_permittivity=3.0, conductivity=2e5, **EM_MAT)
    drude = Material.make(r_permittivity=2.0, comp_n1=0.6, comp_n2=0.4, omega_plasma_n1=3e15, omega_plasma_n2=1e15,
                          tao_e_n1=4e-15, tao_e_n2=9e-15, tao_e_n3=1.0)
    pec = Material.make(if_PEC=1, r_permittivity=1.0, **EM_MAT)
    cell = rng.choice(np.array([0, 1, 2], np.uint32), size=shape, p=[0.5, 0.25, 0.25])
    cell[3:5, 4:6, 5:7] = 3
    f = FLAG_EMDYNAMIC | FLAG_JF_INPUT | flags | (FLAG_PLANE_EM if plane else 0)
    return Problem(shape=shape, spacing=(1e-9, 1.5e-9, 1.2e-9), materials=[cond, drude, pec], material_cell=cell,
                   dt=1e-18, periodic=(False, False, False), flags=f, planeEM_E=(3e6, -2e6, 1e6),
                   planeEM_E_freq=(1e14, 2e14, 3e14), planeEM_source_z=2, Jf_input_type=jf_type,
                   Jf_input_component="y", Jf_input_amp=4e15, Jf_input_freq=3e14, Jf_input_sigma=5e-16,
                   num_Jf_cycles=3, Jfin_lo=(1, 2, 3), Jfin_hi=(6, 7, 9))


@pytest.mark.parametrize("flags,jf_type", [(0, 1), (0, 2), (FLAG_PEC_XY, 6), (FLAG_PEC_XY | FLAG_PEC_Z, 3)])
def test_boxed_media_stages(flags, jf_type):
    prob = boxed_media(flags=flags, jf_type=jf_type)
    E = random_yee(prob.shape, FIELD_E_EM, 1e6, 6)
    H = random_yee(prob.shape, FIELD_H_EM, 3e3, 7)
    Jp = 1e9 * np.random.default_rng(8).standard_normal((3, 3) + prob.shape)
    g, o = run_stages(prob, {FIELD_E_EM: E, FIELD_H_EM: H, FIELD_JP: Jp}, nsteps=2, phys=("em",), check=(FIELD_JP,))
    g.close()
    o.close()


@pytest.mark.parametrize("jf_type", [0, 4, 5])
def test_periodic_box_sources(jf_type):
    # all walls periodic: no boundary planes, no history
    prob = boxed_media(shape=(6, 6, 8), jf_type=jf_type, plane=False)
    prob.periodic = (True, True, True)
    prob.Jf_input_freq = 2e-16 if jf_type in (4, 5) else 3e14
    prob.Jfin_lo, prob.Jfin_hi = (0, 1, 2), (5, 4, 6)
    E = random_yee(prob.shape, FIELD_E_EM, 1e6, 9)
    H = random_yee(prob.shape, FIELD_H_EM, 3e3, 10)
    g, o = run_stages(prob, {FIELD_E_EM: E, FIELD_H_EM: H}, nsteps=2, phys=("em",))
    g.close()
    o.close()


@pytest.mark.parametrize("one_d", [False, True])
def test_fm_em_stages(one_d):
    """Faraday's law with dM from the magnet and the H back-action on m."""
    shape = (8, 8, 8)
    fm = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.02, K1=-610.0, Aex=3.7e-12,
                       r_permittivity=5.0, **EM_MAT)
    cell = np.zeros(shape, np.uint32)
    cell[:, :, 2:6] = 1
    flags = FLAG_EMDYNAMIC | FLAG_EM_FROM_M | FLAG_EM_BACKACTION_M
    if one_d:
        flags |= FLAG_MAG_1D | FLAG_MAGNETOSTATICS
    # FDTD needs c dt / dx well below 1: dt = 1e-18 s at 2 nm
    prob = Problem(shape=shape, spacing=(2e-9, 2e-9, 2e-9), materials=[fm], material_cell=cell, dt=1e-18,
                   periodic=(True, False, False), flags=flags, Hext_stat=(0, 0, 8e4))
    m = workloads.tilted_m(shape, 12, theta_max=0.8)
    m[:, cell == 0] = 0.0
    E = random_yee(shape, FIELD_E_EM, 1e5, 13)
    H = random_yee(shape, FIELD_H_EM, 1e3, 14)
    g, o = run_stages(prob, {FIELD_M: m, FIELD_E_EM: E, FIELD_H_EM: H}, nsteps=2, phys=("fm", "em"),
                      check=(FIELD_M_STORE,))
    assert np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max() <= 1e-12
    g.close()
    o.close()


def test_afm_em_stages():
    prob, m1, m2 = workloads.afm_block(shape=(6, 6, 8), dt=1e-18)
    mat = prob.materials[0]
    mat.r_permittivity = 4.0
    for k, v in EM_MAT.items():
        setattr(mat, k, v)
    prob.flags |= FLAG_EMDYNAMIC | FLAG_EM_FROM_M | FLAG_EM_BACKACTION_M
    E = random_yee(prob.shape, FIELD_E_EM, 1e5, 15)
    H = random_yee(prob.shape, FIELD_H_EM, 1e3, 16)
    g, o = run_stages(prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2, FIELD_E_EM: E, FIELD_H_EM: H}, nsteps=2,
                      phys=("afm", "em"))
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
        assert np.abs(g.download(fid) - o.get(fid)).max() <= 1e-12
    g.close()
    o.close()


def test_fe_em_1000_steps():
    """The parity bounds after 1000 coupled steps through gf_step."""
    from oracle.oracle import Oracle
    prob, p, q = fe_film(shape=(2, 2, 64))
    say("case fe+em 1000 steps", prob.shape)
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_P, p), (FIELD_Q, q)):
        g.upload(fid, v)
        o.set(fid, v)
    g.step(1000)
    o.step(1000)
    for fid in (FIELD_P, FIELD_E_EM, FIELD_H_EM):
        assert rel(g.download(fid), o.get(fid)) <= 1e-9, fid


def test_em_stage_order_is_enforced():
    prob, p, q = fe_film(shape=(2, 2, 16), source=False)
    say("case em stage order", prob.shape)
    g = Context(prob)
    g.upload(FIELD_P, p)
    g.time_advance()
    with pytest.raises(Exception, match="gf_em_stage comes after"):
        g.em_stage(1)
    g.fe_stage(1)
    g.em_stage(1)
    with pytest.raises(Exception, match="gf_em_stage comes after"):
        g.em_stage(2)
    g.fe_stage(2)
    g.em_stage(2)
    g.close()
