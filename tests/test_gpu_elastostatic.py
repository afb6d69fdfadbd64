"""GPU parity of the elastostatics (GF_FLAG_ELASTOSTATIC) through the C ABI
against the oracle (pinned by closed forms in test_oracle_elastostatic.py):
the static strain of random magnetisations on mixed blocks (rotated
crystal, two magnets and a non-magnetic material, an antiferromagnet), and
magnets stepping with the static strain recomputed after every step and
acting back through the magnetoelastic field; a ferroelectric layer whose
electrostrictive eigenstrain enters the solve and which reads the static
strain and eigenstrain back, with its restart; the same stack with
elastodynamics, the static solve once at start-up.  1e-12 of the peak.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF,
                     FLAG_ELASTO_ON_PM, FLAG_ELASTODYNAMICS, FLAG_ELASTOSTATIC, FLAG_ORIENT_ROTATE, Context,
                     GoFerroError, Material, Problem, workloads)
from goferro.workloads import GYRO, set_elastic, set_magnetostriction

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rot():
    c, s = np.cos(0.3), np.sin(0.3)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    c2, s2 = np.cos(0.7), np.sin(0.7)
    Ry = np.array([[c2, 0, s2], [0, 1, 0], [-s2, 0, c2]])
    return Ry @ Rz


def mixed(shape=(16, 8, 32), rotate=True, flags=0):
    T = rot() if rotate else None
    a = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, K1=-610.0, Aex=3.7e-12)
    set_elastic(a, 2.69e11, 1.08e11, 0.76e11, 5170.0, Tg2c=T)
    set_magnetostriction(a, 3.5e5, 2.0e5)
    b = Material.make(if_FM=1, anisotropy_type=2, Ms=8e5, gyro=GYRO, FM_damping=0.1, K1=2e4, Aex=1.3e-11)
    set_elastic(b, 2.2e11, 1.2e11, 0.9e11, 8900.0, Tg2c=T)
    set_magnetostriction(b, -6e5, 4e5)
    sub = set_elastic(Material.make(), 3.0e11, 1.0e11, 1.1e11, 5100.0, Tg2c=T)
    rng = np.random.default_rng(3)
    cell = rng.choice(np.array([0, 1, 2, 3], np.uint32), size=shape, p=[0.1, 0.45, 0.3, 0.15])
    prob = Problem(shape=shape, spacing=(2e-9, 2.5e-9, 1.5e-9), materials=[a, b, sub], material_cell=cell,
                   dt=5e-15, periodic=(True, True, False), Tg2c=T, Hext_stat=(1e4, 0, 6e4),
                   flags=FLAG_ELASTOSTATIC | flags | (FLAG_ORIENT_ROTATE if rotate else 0), scale_elasto=1.3)
    m = workloads.tilted_m(shape, 5, theta_max=0.6)
    return prob, {FIELD_M: m}


def strain(obj):
    return np.concatenate([obj.get(FIELD_STRAIN_T0_DIAG) if hasattr(obj, "get") else obj.download(FIELD_STRAIN_T0_DIAG),
                           obj.get(FIELD_STRAIN_T0_OFF) if hasattr(obj, "get") else obj.download(FIELD_STRAIN_T0_OFF)])


def pair(prob, fields):
    from oracle.oracle import Oracle
    g, o = Context(prob), Oracle(prob)
    for fid, v in fields.items():
        g.upload(fid, v)
        o.set(fid, v)
    return g, o


@pytest.mark.parametrize("rotate", [False, True])
def test_static_strain_matches_oracle(rotate):
    say("case elastostatic strain, rotate", rotate)
    prob, f = mixed(rotate=rotate)
    g, o = pair(prob, f)
    g.demag()
    o.demag()
    d = rel(strain(g), strain(o))
    say("  rel", d)
    assert d <= 1e-12
    g.close()
    o.close()


def test_afm_static_strain():
    say("case elastostatic strain, AFM")
    prob, m1, m2 = workloads.afm_block(shape=(8, 16, 16))
    mat = prob.materials[0]
    set_elastic(mat, 2.5e11, 1.3e11, 0.8e11, 6800.0)
    mat.lamda100_AFM1, mat.lamda100_AFM2, mat.lamda111_AFM1, mat.lamda111_AFM2 = -2e-5, -1e-5, 1e-5, 3e-5
    prob.flags |= FLAG_ELASTOSTATIC
    g, o = pair(prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2})
    g.demag()
    o.demag()
    assert rel(strain(g), strain(o)) <= 1e-12
    g.close()
    o.close()


def test_magnets_step_with_the_static_strain():
    say("case elastostatic coupled steps")
    prob, f = mixed(shape=(8, 8, 16), flags=FLAG_ELASTO_ON_PM)
    g, o = pair(prob, f)
    g.demag()
    o.demag()
    g.step(3)
    o.step(3)
    assert np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max() <= 1e-12
    assert rel(strain(g), strain(o)) <= 1e-12
    g.close()
    o.close()


def fe_mixed(shape=(8, 8, 16), rotate=True):
    """A ferroelectric layer (PbTiO3-like, electrostrictive), a magnet and a
    substrate, random p and m, the strain acting on both."""
    from goferro.workloads import pto_material
    T = rot() if rotate else None
    fe = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    fe.Q11, fe.Q12, fe.Q44 = 0.089, -0.026, 0.0675
    fe.tv1, fe.tv2, fe.tv3 = 3.05e10, 6.4e8, 2.97e10
    set_elastic(fe, 2.3e11, 1.0e11, 1.1e11, 7900.0, Tg2c=T)
    mag = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, K1=-610.0, Aex=3.7e-12)
    set_elastic(mag, 2.69e11, 1.08e11, 0.76e11, 5170.0, Tg2c=T)
    set_magnetostriction(mag, 3.5e5, 2.0e5)
    sub = set_elastic(Material.make(), 3.2e11, 1.0e11, 1.2e11, 5100.0, Tg2c=T)
    cell = np.full(shape, 3, np.uint32)
    cell[:, :, 4:10] = 1
    cell[:, :, 10:] = 2
    prob = Problem(shape=shape, spacing=(1e-9, 1e-9, 1e-9), materials=[fe, mag, sub], material_cell=cell, dt=2e-15,
                   periodic=(True, True, False), Tg2c=T, Hext_stat=(0, 0, 5e4), Eext_stat=(0, 0, 1e6),
                   flags=FLAG_ELASTOSTATIC | FLAG_ELASTO_ON_PM | (FLAG_ORIENT_ROTATE if rotate else 0))
    rng = np.random.default_rng(41)
    p = 0.05 * rng.standard_normal((3,) + shape)
    p[2] += 0.6
    p[:, cell != 1] = 0.0
    q = 1e8 * rng.standard_normal((3,) + shape)
    q[:, cell != 1] = 0.0
    m = workloads.tilted_m(shape, 9, theta_max=0.5)
    m[:, cell != 2] = 0.0
    from goferro import FIELD_P, FIELD_Q
    return prob, {FIELD_P: p, FIELD_Q: q, FIELD_M: m}


@pytest.mark.parametrize("rotate", [False, True])
def test_ferroelectric_static_strain_and_steps(rotate):
    """The static strain and eigenstrain of a ferroelectric and a magnet, and
    three gf_steps with both reading them back."""
    from goferro import FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF, FIELD_P, FIELD_Q
    say("case elastostatic ferroelectric, rotate", rotate)
    prob, f = fe_mixed(rotate=rotate)
    g, o = pair(prob, f)
    g.demag()
    o.demag()
    assert rel(strain(g), strain(o)) <= 1e-12
    for fid in (FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.step(3)
    o.step(3)
    for fid in (FIELD_P, FIELD_Q):
        d = rel(g.download(fid), o.get(fid))
        say("  rel", fid, d)
        assert d <= 1e-12, fid
    assert np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max() <= 1e-12
    assert rel(strain(g), strain(o)) <= 1e-12
    g.close()
    o.close()


def test_ferroelectric_elastostatic_restart():
    """The state image carries the static eigenstrain the ferroelectric
    reads: n1 + n2 steps equal n1, image, fresh context, n2, bit for bit."""
    from goferro import FIELD_P, FIELD_Q
    say("case elastostatic ferroelectric restart")
    prob, f = fe_mixed(rotate=False)
    a = Context(prob)
    for fid, v in f.items():
        a.upload(fid, v)
    a.demag()
    a.step(5)
    want = a.download(FIELD_P), a.download(FIELD_Q)
    a.close()
    b = Context(prob)
    for fid, v in f.items():
        b.upload(fid, v)
    b.demag()
    b.step(2)
    image = b.state_image()
    b.close()
    c = Context(prob)
    c.restore_state(image)
    c.step(3)
    assert np.array_equal(c.download(FIELD_P), want[0]) and np.array_equal(c.download(FIELD_Q), want[1])
    c.close()


def beside(rotate):
    """fe_mixed with elastodynamics acting back, random u and v."""
    from goferro import FIELD_U, FIELD_V, FLAG_ELASTO_BACKACTION
    prob, f = fe_mixed(rotate=rotate)
    prob.flags |= FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION
    rng = np.random.default_rng(47)
    f[FIELD_U] = 1e-12 * rng.standard_normal((3,) + prob.shape)
    f[FIELD_V] = 0.5 * rng.standard_normal((3,) + prob.shape)
    return prob, f


@pytest.mark.parametrize("rotate", [False, True])
def test_elastostatics_beside_elastodynamics(rotate):
    """The static strain and eigenstrain solved once at start-up
    (main.cpp:148-158); three steps of the magnet, the ferroelectric and the
    elastodynamics reading them beside the dynamic ones; the static strain
    kept through the steps; an uploaded static strain replaces the solve."""
    from goferro import FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF, FIELD_P, FIELD_Q, FIELD_U, FIELD_V
    say("case elastostatics beside elastodynamics, rotate", rotate)
    prob, f = beside(rotate)
    g, o = pair(prob, f)
    g.demag()
    o.demag()
    assert rel(strain(g), strain(o)) <= 1e-12
    for fid in (FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    s0 = strain(g)
    g.step(3)
    o.step(3)
    for fid in (FIELD_P, FIELD_Q, FIELD_U, FIELD_V):
        d = rel(g.download(fid), o.get(fid))
        say("  rel", fid, d)
        assert d <= 1e-12, fid
    assert np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max() <= 1e-12
    g.demag()
    assert np.array_equal(strain(g), s0)
    g.close()
    o.close()
    g = Context(prob)
    for fid, v in f.items():
        g.upload(fid, v)
    up = 1e-5 * np.random.default_rng(5).standard_normal((3,) + prob.shape)
    g.upload(FIELD_STRAIN_T0_DIAG, up)
    g.demag()
    assert np.array_equal(g.download(FIELD_STRAIN_T0_DIAG), up)
    g.close()


def test_elastostatics_beside_elastodynamics_restart():
    """The image carries the start-up static strain and that it was
    solved: n1 + n2 steps equal n1, image, fresh context, gf_demag, n2."""
    from goferro import FIELD_P, FIELD_U
    say("case elastostatics beside elastodynamics, restart")
    prob, f = beside(False)
    a = Context(prob)
    for fid, v in f.items():
        a.upload(fid, v)
    a.demag()
    a.step(5)
    want = a.download(FIELD_P), a.download(FIELD_U), a.download(FIELD_M)
    a.close()
    b = Context(prob)
    for fid, v in f.items():
        b.upload(fid, v)
    b.demag()
    b.step(2)
    image = b.state_image()
    b.close()
    c = Context(prob)
    c.restore_state(image)
    c.demag()   # does not solve again: the image holds the start-up solve
    c.step(3)
    got = c.download(FIELD_P), c.download(FIELD_U), c.download(FIELD_M)
    assert all(np.array_equal(x, y) for x, y in zip(got, want))
    c.close()


def test_afm_elastostatics_beside_elastodynamics():
    """An antiferromagnet whose static strain is solved once at start-up and
    stepped with elastodynamics acting back: three steps against the oracle,
    the static strain kept."""
    from goferro import FIELD_U, FIELD_V, FLAG_ELASTO_BACKACTION
    say("case AFM elastostatics beside elastodynamics")
    prob, m1, m2 = workloads.afm_block(shape=(8, 8, 8))
    mat = prob.materials[0]
    set_elastic(mat, 2.5e11, 1.3e11, 0.8e11, 6800.0)
    mat.lamda100_AFM1, mat.lamda100_AFM2, mat.lamda111_AFM1, mat.lamda111_AFM2 = -2e-5, -1e-5, 1e-5, 3e-5
    mat.B1_AFM1 = mat.B1_AFM2 = 6e5
    mat.B2_AFM1 = mat.B2_AFM2 = 4e5
    prob.flags |= FLAG_ELASTOSTATIC | FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM
    rng = np.random.default_rng(53)
    f = {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2, FIELD_U: 1e-12 * rng.standard_normal((3,) + prob.shape),
         FIELD_V: 0.5 * rng.standard_normal((3,) + prob.shape)}
    g, o = pair(prob, f)
    g.demag()
    o.demag()
    assert rel(strain(g), strain(o)) <= 1e-12
    s0 = strain(g)
    g.step(3)
    o.step(3)
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
        assert np.abs(g.download(fid) - o.get(fid)).max() <= 1e-12, fid
    for fid in (FIELD_U, FIELD_V):
        d = rel(g.download(fid), o.get(fid))
        say("  rel", fid, d)
        assert d <= 1e-12, fid
    assert np.array_equal(strain(g), s0)
    g.close()
    o.close()


def test_elastostatic_rejections():
    say("case elastostatic rejections")
    from goferro import FLAG_FLEXO
    prob, _ = fe_mixed()
    prob.flags |= FLAG_FLEXO | FLAG_ELASTODYNAMICS
    with pytest.raises(GoFerroError, match="flexoelectric"):
        Context(prob)
    prob, _ = mixed(shape=(8, 8, 12))
    with pytest.raises(GoFerroError, match="power-of-two"):
        Context(prob)
