"""GPU parity of the spin pumping (GF_FLAG_SPIN_PUMPING) through the C ABI
against the oracle (pinned in test_oracle_spin.py): the inverse spin Hall
current after every magnet stage, alone and as a source in Ampere's law,
for a ferromagnet and an antiferromagnet, with the layer stack given and
derived.  Tolerance 1e-12 relative to each field's peak per stage.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_EM, FIELD_H_EM_STORE, FIELD_J_ISHE, FIELD_M,
                     FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE, FLAG_1D_ABC, FLAG_EM_BACKACTION_M, FLAG_EM_FROM_M,
                     FLAG_EMDYNAMIC, FLAG_SPIN_PUMPING, Context, Material, Problem, workloads, yee_shapes)
from goferro.workloads import GYRO

pytestmark = pytest.mark.gpu

EM_MAT = dict(comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0, tao_e_n3=1.0)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def metal():
    return Material.make(if_spin_pump=1, spin_hall_angle=0.08, spin_mix_cond=5e18, spin_diffus_length=3e-9,
                         r_permittivity=2.0, conductivity=1e5, **EM_MAT)


def film(shape=(8, 6, 12), flags=0, afm=False, nz_layer=None):
    """vacuum | metal 1 | magnet 3 | metal 4 | vacuum: the lower metal is
    sourced from the magnet above it (the negative branch), the upper one
    from the magnet below."""
    nz = shape[2]
    if afm:
        prob0, a1, a2 = workloads.afm_block(shape=shape, dt=1e-18)
        mag = prob0.materials[0]
        mag.r_permittivity = 4.0
        for k, v in EM_MAT.items():
            setattr(mag, k, v)
    else:
        mag = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, K1=-610.0,
                            Aex=3.7e-12, r_permittivity=5.0, **EM_MAT)
    col = np.zeros(nz, np.uint32)
    col[1] = 2
    col[2:5] = 1
    col[5:9] = 2
    cell = np.broadcast_to(col, shape).copy()
    prob = Problem(shape=shape, spacing=(2e-9, 2e-9, 1e-9), materials=[mag, metal()], material_cell=cell,
                   dt=1e-18, periodic=(True, True, False), flags=FLAG_SPIN_PUMPING | flags,
                   Hext_stat=(2e4, 0, 8e4), Hext_altr=(3e4, 1e4, 0), H_altr_freq=(5e11, 7e11, 0), nz_layer=nz_layer)
    if afm:
        m1, m2 = a1.copy(), a2.copy()
        m1[:, cell != 1] = 0.0
        m2[:, cell != 1] = 0.0
        return prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}
    m = workloads.tilted_m(shape, 21, theta_max=0.8)
    m[:, cell != 1] = 0.0
    return prob, {FIELD_M: m}


def run(prob, fields, phys, nsteps=2, em_fields=True):
    from oracle.oracle import Oracle
    say("case spin pumping", phys, prob.shape, hex(prob.flags))
    g, o = Context(prob), Oracle(prob)
    if em_fields and "em" in phys:
        rng = np.random.default_rng(3)
        fields = dict(fields)
        fields[FIELD_E_EM] = [1e4 * rng.standard_normal(s) for s in yee_shapes(prob.shape, FIELD_E_EM)]
        fields[FIELD_H_EM] = [1e2 * rng.standard_normal(s) for s in yee_shapes(prob.shape, FIELD_H_EM)]
    for fid, v in fields.items():
        g.upload(fid, v)
        o.set(fid, v)
    L, h = o.L, o._h
    gets = {"fm": L.gfo_mag_get_dm, "afm": L.gfo_afm_get_dm, "em": L.gfo_em_get}
    upds = {"fm": L.gfo_mag_update, "afm": L.gfo_afm_update, "em": L.gfo_em_update}
    calls = {"fm": g.mag_stage, "afm": g.afm_stage, "em": g.em_stage}
    checks = [FIELD_J_ISHE] + ([FIELD_E_EM_STORE, FIELD_H_EM_STORE] if "em" in phys else [])
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
            for fid in checks:
                d = rel(g.download(fid), o.get(fid))
                assert d <= 1e-12, (fid, step, s, d)
    J = o.get(FIELD_J_ISHE)
    if prob.nz_layer is None:
        on = [1, 5, 6, 7, 8]
        assert all(np.abs(J[:2, :, :, k]).max() > 0 for k in on)
        assert not J[:, :, :, [k for k in range(prob.shape[2]) if k not in on]].any()
    return g, o


@pytest.mark.parametrize("em", [False, True])
def test_fm_spin_pumping(em):
    flags = (FLAG_EMDYNAMIC | FLAG_1D_ABC) if em else 0
    prob, f = film(flags=flags)
    g, o = run(prob, f, ("fm", "em") if em else ("fm",))
    assert np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max() <= 1e-12
    if em:
        for fid in (FIELD_E_EM, FIELD_H_EM):
            assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.close()
    o.close()


def test_afm_spin_pumping_with_em_coupling():
    prob, f = film(flags=FLAG_EMDYNAMIC | FLAG_EM_FROM_M | FLAG_EM_BACKACTION_M, afm=True)
    g, o = run(prob, f, ("afm", "em"))
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2, FIELD_E_EM, FIELD_H_EM):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.close()
    o.close()


def test_layer_stack_given():
    # the metal above the magnet as two layers of 2: only the lower one pumps
    prob, f = film(nz_layer=[1, 1, 3, 2, 2, 3])
    from oracle.oracle import Oracle
    say("case spin pumping layers given")
    g, o = Context(prob), Oracle(prob)
    for fid, v in f.items():
        g.upload(fid, v)
        o.set(fid, v)
    g.step(2)
    o.step(2)
    J = o.get(FIELD_J_ISHE)
    assert np.abs(J[:, :, :, 5:7]).max() > 0 and not J[:, :, :, 7:].any()
    assert rel(g.download(FIELD_J_ISHE), J) <= 1e-12
    assert np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max() <= 1e-12
    g.close()
    o.close()


def test_fm_and_afm_pump_into_two_metals():
    """vacuum | metal 2 | FM 3 | vacuum | AFM 3 | metal 2 | vacuum: the lower
    metal pumps from the FM, the upper from the AFM; at the step boundary
    the AFM update's current stands (the FM state already updated)."""
    shape = (6, 5, 14)
    prob0, a1, a2 = workloads.afm_block(shape=shape, dt=1e-18)
    afm = prob0.materials[0]
    fm = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, K1=-610.0, Aex=3.7e-12)
    col = np.array([0, 2, 2, 1, 1, 1, 0, 3, 3, 3, 2, 2, 0, 0], np.uint32)
    cell = np.broadcast_to(col, shape).copy()
    prob = Problem(shape=shape, spacing=(2e-9, 2e-9, 1e-9), materials=[fm, metal(), afm], material_cell=cell,
                   dt=1e-18, periodic=(True, True, False), flags=FLAG_SPIN_PUMPING, Hext_stat=(2e4, 0, 8e4),
                   Hext_altr=(3e4, 1e4, 0), H_altr_freq=(5e11, 7e11, 0))
    m = workloads.tilted_m(shape, 23, theta_max=0.8)
    m[:, cell != 1] = 0.0
    m1, m2 = a1.copy(), a2.copy()
    m1[:, cell != 3] = 0.0
    m2[:, cell != 3] = 0.0
    fields = {FIELD_M: m, FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}
    from oracle.oracle import Oracle
    say("case spin pumping fm + afm", shape)
    g, o = Context(prob), Oracle(prob)
    for fid, v in fields.items():
        g.upload(fid, v)
        o.set(fid, v)
    L, h = o.L, o._h
    for step in range(2):
        g.time_advance()
        o.time_advance()
        for s in range(1, 5):
            g.mag_stage(s)
            g.afm_stage(s)
            L.gfo_mag_get_dm(h, s)
            L.gfo_afm_get_dm(h, s)
            L.gfo_mag_update(h, s)
            L.gfo_afm_update(h, s)
            d = rel(g.download(FIELD_J_ISHE), o.get(FIELD_J_ISHE))
            assert d <= 1e-12, (step, s, d)
    J = o.get(FIELD_J_ISHE)
    assert (np.abs(J[:2][..., [1, 2, 10, 11]]).max(axis=(0, 1, 2)) > 0).all()
    g.step(3)
    o.step(3)
    assert rel(g.download(FIELD_J_ISHE), o.get(FIELD_J_ISHE)) <= 1e-12
    for fid in fields:
        assert np.abs(g.download(fid) - o.get(fid)).max() <= 1e-12, fid
    g.close()
    o.close()
