"""GPU parity: the gfx950 library against the CPU oracle, through the C ABI.

Tolerances are the parity bounds: 1e-12 per stage/step on m (|m| ~ 1, so
absolute), 1e-9 on m after 1000 steps; the magnetostatic field is compared
relative to its peak (1e-12).  Every case prints its name and shape first so
a failing GPU step names itself.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_HEXT_NONUNIF, FIELD_HSTAT, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE,
                     FLAG_DEMAG_FACTOR, FLAG_MAG_1D, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE, FLAG_PRESCRIBE_HEXT,
                     MU0, Context, Material, Problem, workloads)
from goferro.workloads import GYRO

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rot():
    c, s = np.cos(0.3), np.sin(0.3)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    c2, s2 = np.cos(0.7), np.sin(0.7)
    Ry = np.array([[c2, 0, s2], [0, 1, 0], [-s2, 0, c2]])
    return Ry @ Rz


def mixed_fm_problem(shape=(16, 16, 8), flags=FLAG_MAGNETOSTATICS, periodic=(False, True, False), rotate=True):
    """Two ferromagnets with different Ms/Aex/anisotropy/DMI, a non-magnetic
    material and vacuum, a non-periodic axis, a rotation and an AC drive."""
    rng = np.random.default_rng(7)
    m1 = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.02, K1=-2e4, K2=5e3,
                       Aex=3.7e-12, iDMI=1.5e-3)
    m2 = Material.make(if_FM=1, anisotropy_type=2, Ms=8e5, gyro=GYRO, FM_damping=0.1, K1=4e5, Aex=1.3e-11,
                       iDMI=0.0)
    m3 = Material.make(if_FM=0)
    cell = rng.choice(np.array([0, 1, 2, 3], np.uint32), size=shape, p=[0.1, 0.45, 0.35, 0.1])
    prob = Problem(shape=shape, spacing=(2e-9, 2.5e-9, 1.5e-9), materials=[m1, m2, m3], material_cell=cell,
                   dt=2e-14, periodic=periodic, flags=flags | (FLAG_ORIENT_ROTATE if rotate else 0),
                   Hext_stat=(1e4, -2e4, 5e4), Hext_altr=(3e4, 0, 1e4), H_altr_freq=(5e9, 0, 2e10),
                   demag_fac=(0.1, 0.2, 0.7), Tg2c=rot() if rotate else None)
    m = workloads.tilted_m(shape, 9, theta_max=1.2)
    return prob, m


def compare_stages(prob, fields, nsteps=2, hnu=None, check_h=True, afm=False, fm=True):
    from oracle.oracle import Oracle
    say("case", prob.shape, hex(prob.flags))
    g = Context(prob)
    o = Oracle(prob)
    for fid, val in fields.items():
        g.upload(fid, val)
        o.set(fid, val)
    if hnu is not None:
        g.upload(FIELD_HEXT_NONUNIF, hnu)
        o.set(FIELD_HEXT_NONUNIF, hnu)
    if prob.flags & FLAG_MAGNETOSTATICS:
        g.demag()
        o.demag()
    worst = 0.0
    for step in range(nsteps):
        g.time_advance()
        o.time_advance()
        for s in range(1, 5):
            if fm:
                g.mag_stage(s)
                o.L.gfo_mag_get_dm(o._h, s)
            if afm:
                g.afm_stage(s)
                o.L.gfo_afm_get_dm(o._h, s)
            if fm:
                o.L.gfo_mag_update(o._h, s)
            if afm:
                o.L.gfo_afm_update(o._h, s)
            if fm:
                d = np.abs(g.download(FIELD_M_STORE) - o.get(FIELD_M_STORE)).max()
                worst = max(worst, d)
                assert d <= 1e-12, (step, s, d)
            if prob.flags & FLAG_MAGNETOSTATICS:
                g.demag()
                o.demag()
                if check_h:
                    hg, ho = g.download(FIELD_HSTAT), o.get(FIELD_HSTAT)
                    dh = np.abs(hg - ho).max() / max(np.abs(ho).max(), 1e-300)
                    assert dh <= 1e-12, (step, s, dh)
    if fm:
        np.testing.assert_allclose(g.download(FIELD_M), o.get(FIELD_M), rtol=0, atol=1e-12)
        np.testing.assert_allclose(g.average(FIELD_M), o.average(FIELD_M), rtol=0, atol=1e-13)
    if afm:
        for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
            np.testing.assert_allclose(g.download(fid), o.get(fid), rtol=0, atol=1e-12)
            np.testing.assert_allclose(g.average(fid), o.average(fid), rtol=0, atol=1e-13)
    assert abs(g.time - o.time) <= 1e-30
    g.close()
    o.close()
    return worst


@pytest.mark.parametrize("shape", [(8, 8, 4), (16, 16, 8), (64, 64, 16), (4, 8, 32)])
def test_fm_film_stages(shape):
    prob, m = workloads.fm_film(shape)
    compare_stages(prob, {FIELD_M: m}, nsteps=2)


def test_fm_mixed_materials_rotation_dmi_drive():
    prob, m = mixed_fm_problem()
    hnu = np.random.default_rng(3).standard_normal((3,) + prob.shape) * 3e4
    compare_stages(prob, {FIELD_M: m}, nsteps=2, hnu=hnu)


def test_fm_mixed_materials_lean_fused_stage():
    """Two ferromagnets (one with DMI), a non-magnet and vacuum, FFT
    magnetostatics, no rotation and no optional fields: the fused stage's
    lean form, which reads the material records from its LDS copy, with
    interfaces between every pair of materials."""
    prob, m = mixed_fm_problem(rotate=False)
    compare_stages(prob, {FIELD_M: m}, nsteps=3)


@pytest.mark.parametrize("mode", [FLAG_DEMAG_FACTOR | FLAG_MAGNETOSTATICS, FLAG_MAG_1D | FLAG_MAGNETOSTATICS, 0])
def test_fm_local_statics(mode):
    prob, m = mixed_fm_problem(shape=(8, 16, 4), flags=mode, periodic=(True, False, True))
    compare_stages(prob, {FIELD_M: m}, nsteps=2)


def test_fm_prescribed_drive_table():
    prob, m = workloads.fm_film((8, 8, 4))
    prob.flags |= FLAG_PRESCRIBE_HEXT
    prob.prescribe_Hext = np.array([[0.0, 1e4, 0, 2e5], [2.5e-14, -3e4, 1e4, 1e5], [5.5e-14, 0, 0, 3e5]])
    compare_stages(prob, {FIELD_M: m}, nsteps=6)


def test_afm_stages_with_demag():
    prob, m1, m2 = workloads.afm_block((8, 8, 4))
    prob.flags |= FLAG_MAGNETOSTATICS
    compare_stages(prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}, nsteps=2, afm=True, fm=False)


def test_fm_and_afm_together():
    shape = (8, 8, 8)
    prob_fm, m = workloads.fm_film(shape)
    prob_afm, m1, m2 = workloads.afm_block(shape)
    cell = np.ones(shape, np.uint32)
    cell[:, :, 4:] = 2
    cell[0, 0, :] = 0
    prob = Problem(shape=shape, spacing=(2e-9,) * 3, materials=[prob_fm.materials[0], prob_afm.materials[0]],
                   material_cell=cell, dt=1e-15, periodic=(True, True, False), flags=FLAG_MAGNETOSTATICS,
                   Hext_stat=(0, 0, 0.1 / MU0), Hext_altr=(0.05 / MU0, 0, 0), H_altr_freq=(1e12, 0, 0))
    compare_stages(prob, {FIELD_M: m, FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}, nsteps=2, afm=True, fm=True)


def fe_problem(shape=(8, 8, 32), rotate=True, periodic=(True, True, False), prescribe=False):
    """FE film between vacuum layers, a second non-FE material, random
    p, q and static field, AC drive, rotation."""
    from goferro.workloads import pto_material
    rng = np.random.default_rng(12)
    fe = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    fe2 = Material.make(if_FE=1, FE_mass=3e-16, FE_damping=1e-3, a1=-1.1e8, a11=-2.0e7, a12=3.5e8, a111=1.6e8,
                        a112=2.1e8, a123=-1.7e9, a1111=4e7, a1112=-1e7, a1122=3e7, a1123=2e7, G11=3e-10)
    other = Material.make(if_FE=0)
    cell = np.zeros(shape, np.uint32)
    nz = shape[2]
    cell[:, :, nz // 4: 3 * nz // 4] = 1
    cell[:, :, nz // 2: nz // 2 + 2] = 2
    cell[0, 0, :] = 3
    prob = Problem(shape=shape, spacing=(1e-9, 1.5e-9, 1e-9), materials=[fe, fe2, other], material_cell=cell,
                   dt=2e-15, periodic=periodic, flags=FLAG_ORIENT_ROTATE if rotate else 0,
                   Tg2c=rot() if rotate else None, Eext_stat=(1e6, 0, -2e6), Eext_altr=(0, 3e6, 1e7),
                   E_altr_freq=(0, 2e11, 5e11))
    if prescribe:
        prob.flags |= 1 << 7
        prob.prescribe_Eext = np.array([[0.0, 1e6, 0, 2e7], [3e-15, -3e6, 1e6, 1e7], [7e-15, 0, 0, 3e7]])
    p = 0.1 * rng.standard_normal((3,) + shape)
    p[2] += 0.7
    q = 1e9 * rng.standard_normal((3,) + shape)
    E = 1e7 * rng.standard_normal((3,) + shape)
    return prob, p, q, E


@pytest.mark.parametrize("rotate,prescribe", [(False, False), (True, False), (True, True)])
def test_fe_stages(rotate, prescribe):
    from goferro import FIELD_ESTAT, FIELD_P, FIELD_P_STORE, FIELD_Q
    from oracle.oracle import Oracle
    prob, p, q, E = fe_problem(rotate=rotate, prescribe=prescribe)
    say("case fe", prob.shape, rotate, prescribe)
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_P, p), (FIELD_Q, q), (FIELD_ESTAT, E)):
        g.upload(fid, v)
        o.set(fid, v)
    pscale = np.abs(p).max()
    for step in range(3):
        g.time_advance()
        o.time_advance()
        for s in range(1, 5):
            g.fe_stage(s)
            o.fe_stage(s)
            d = np.abs(g.download(FIELD_P_STORE) - o.get(FIELD_P_STORE)).max()
            assert d <= 1e-12 * pscale, (step, s, d)
    np.testing.assert_allclose(g.download(FIELD_P), o.get(FIELD_P), rtol=0, atol=1e-12 * pscale)
    qo = o.get(FIELD_Q)
    np.testing.assert_allclose(g.download(FIELD_Q), qo, rtol=0, atol=1e-12 * np.abs(qo).max())
    np.testing.assert_allclose(g.average(FIELD_P), o.average(FIELD_P), rtol=0, atol=1e-13 * pscale)


def test_fe_and_fm_step_together():
    from goferro import FIELD_P, FIELD_Q
    from oracle.oracle import Oracle
    shape = (8, 8, 16)
    prob_fe, p, q, _ = fe_problem(shape=shape, rotate=False)
    prob_fm, m = workloads.fm_film(shape)
    cell = prob_fe.material_cell.copy()
    cell[cell == 0] = 4
    prob = Problem(shape=shape, spacing=prob_fe.spacing, materials=list(prob_fe.materials) + [prob_fm.materials[0]],
                   material_cell=cell, dt=1e-15, periodic=(True, True, False), flags=FLAG_MAGNETOSTATICS,
                   Hext_stat=(0, 0, 8e4), Eext_altr=(0, 0, 1e7), E_altr_freq=(0, 0, 1e12))
    say("case fe+fm", shape)
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_M, m), (FIELD_P, p), (FIELD_Q, q)):
        g.upload(fid, v)
        o.set(fid, v)
    g.demag()
    o.demag()
    g.step(20)
    o.step(20)
    np.testing.assert_allclose(g.download(FIELD_M), o.get(FIELD_M), rtol=0, atol=1e-12)
    np.testing.assert_allclose(g.download(FIELD_P), o.get(FIELD_P), rtol=0, atol=1e-12)


def test_fm_film_1000_steps():
    """The parity bound after 1000 steps, at a size the oracle runs in seconds."""
    from oracle.oracle import Oracle
    shape = (16, 16, 8)
    say("case 1000 steps", shape)
    prob, m = workloads.fm_film(shape)
    g, o = Context(prob), Oracle(prob)
    g.upload(FIELD_M, m)
    o.set(FIELD_M, m)
    g.demag()
    o.demag()
    g.step(1000)
    o.step(1000)
    a, b = g.download(FIELD_M), o.get(FIELD_M)
    assert np.abs(a - b).max() <= 1e-9
    # |m| is not renormalised by the reference: the builds must drift alike
    na, nb = np.sqrt((a * a).sum(0)), np.sqrt((b * b).sum(0))
    assert np.abs(na - nb).max() <= 1e-12


def test_uniform_film_full_size_equals_column():
    """Full-size stage kernel check: a laterally uniform film without demag
    evolves exactly like a 1x1 periodic column (translation symmetry)."""
    from oracle.oracle import Oracle
    nz = 32
    shape = (256, 256, nz)
    say("case uniform film", shape)
    prob, _ = workloads.fm_film(shape, demag=False)
    prob_c, _ = workloads.fm_film((1, 1, nz), demag=False)
    col = workloads.tilted_m((1, 1, nz), 4, theta_max=0.5)
    m = np.broadcast_to(col, (3,) + shape).copy()
    g, o = Context(prob), Oracle(prob_c)
    g.upload(FIELD_M, m)
    o.set(FIELD_M, col)
    g.step(50)
    o.step(50)
    a = g.download(FIELD_M)
    ref = o.get(FIELD_M)
    assert np.abs(a - ref).max() <= 1e-12


def test_demag_full_size_shift_and_linearity():
    """Circular convolution commutes with circular shifts and is linear;
    checked at the benchmark grid where the oracle is too slow."""
    shape = (256, 256, 128)
    say("case demag properties", shape)
    prob, m = workloads.fm_film(shape)
    g = Context(prob)
    g.upload(FIELD_M, m)
    g.demag()
    h0 = g.download(FIELD_HSTAT)
    sh = (37, -101, 5)
    g.upload(FIELD_M, np.roll(m, sh, axis=(1, 2, 3)))
    g.demag()
    h1 = g.download(FIELD_HSTAT)
    scale = np.abs(h0).max()
    assert np.abs(np.roll(h0, sh, axis=(1, 2, 3)) - h1).max() <= 1e-12 * scale
    m2 = workloads.tilted_m(shape, 77, theta_max=2.0)
    g.upload(FIELD_M, m2)
    g.demag()
    h2 = g.download(FIELD_HSTAT)
    g.upload(FIELD_M, 0.25 * m - 1.5 * m2)
    g.demag()
    h3 = g.download(FIELD_HSTAT)
    assert np.abs(h3 - (0.25 * h0 - 1.5 * h2)).max() <= 1e-12 * np.abs(h3).max()


def test_fm_shaped_drive_and_static_strain():
    """Shaped AC drive profile plus a static crystal-frame strain acting
    through B1/B2, with rotation, two materials and FFT magnetostatics."""
    from goferro import FIELD_HDRIVE, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTO_ON_PM, FLAG_HDRIVE
    prob, m = mixed_fm_problem(shape=(8, 8, 8))
    mats = list(prob.materials)
    mats[0] = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.02, K1=-2e4, K2=5e3,
                            Aex=3.7e-12, iDMI=1.5e-3, B1=3.5e5, B2=-2e5)
    prob.materials = mats
    prob.flags |= FLAG_HDRIVE | FLAG_ELASTO_ON_PM
    prob.Hdrive_freq = 3e10
    rng = np.random.default_rng(12)
    fields = {FIELD_M: m,
              FIELD_HDRIVE: rng.standard_normal((3,) + prob.shape) * 4e4,
              FIELD_STRAIN_T0_DIAG: rng.standard_normal((3,) + prob.shape) * 1e-3,
              FIELD_STRAIN_T0_OFF: rng.standard_normal((3,) + prob.shape) * 1e-3}
    compare_stages(prob, fields, nsteps=3)


def test_afm_static_strain():
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTO_ON_PM
    prob, m1, m2 = workloads.afm_block((8, 8, 4))
    mat = prob.materials[0]
    mat.B1_AFM1, mat.B2_AFM1, mat.B1_AFM2, mat.B2_AFM2 = 2e6, -1e6, 1.5e6, 5e5
    prob.flags |= FLAG_ELASTO_ON_PM
    rng = np.random.default_rng(13)
    fields = {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2,
              FIELD_STRAIN_T0_DIAG: rng.standard_normal((3,) + prob.shape) * 1e-3,
              FIELD_STRAIN_T0_OFF: rng.standard_normal((3,) + prob.shape) * 1e-3}
    compare_stages(prob, fields, nsteps=2, afm=True, fm=False)


def test_set_material_matches_fresh_context():
    """gf_set_material: changing alpha, Aex, K1 and B1 of a material (a) before
    any step equals a context created with the new numbers, bit for bit;
    (b) mid-run equals a fresh context started from the state at the switch.
    Both runs use the oracle-pinned stage kernels, so they are checked
    against each other."""
    from goferro import FIELD_STRAIN_T0_DIAG, FLAG_ELASTO_ON_PM
    say("case set_material")
    shape = (8, 8, 4)
    prob, m = workloads.fm_film(shape)
    prob.flags |= FLAG_ELASTO_ON_PM
    diag = np.zeros((3,) + shape)
    diag[0] = 1e-3
    new = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.03, K1=-900.0,
                        Aex=4.1e-12, B1=2e5)

    def ctx(p, mm):
        c = Context(p)
        c.upload(FIELD_M, mm)
        c.upload(FIELD_STRAIN_T0_DIAG, diag)
        c.demag()
        return c

    import copy
    p_new = copy.deepcopy(prob)
    p_new.materials = [new]
    a = ctx(prob, m)
    a.set_material(1, new)
    b = ctx(p_new, m)
    a.step(3)
    b.step(3)
    assert np.array_equal(a.download(FIELD_M), b.download(FIELD_M))
    # mid-run: no AC drive in this film, so only the state matters
    c = ctx(prob, m)
    c.step(2)
    mid = c.download(FIELD_M)
    c.set_material(1, new)
    c.step(2)
    d = ctx(p_new, mid)
    d.step(2)
    # the fresh context builds its first field with a separate z pass rather
    # than inside the stage kernel: equal up to rounding
    np.testing.assert_allclose(c.download(FIELD_M), d.download(FIELD_M), rtol=0, atol=1e-13)
    bad = Material.make(if_AFM=1)
    with pytest.raises(Exception, match="kind flags"):
        a.set_material(1, bad)
