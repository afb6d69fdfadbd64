"""GPU tests of the FFT electrostatics (GF_FLAG_ELECTROSTATICS) through the
C ABI against the oracle (itself pinned by closed forms in
test_oracle_estatic.py): the static field of a random polarisation on a
block of ferroelectric, non-polar material and vacuum, and coupled
ferroelectric steps whose field is recomputed after every update phase.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_ESTAT, FIELD_HSTAT, FIELD_M, FIELD_P, FIELD_Q, FLAG_DEMAG_FACTOR, FLAG_ELECTROSTATICS,
                     FLAG_MAG_1D, FLAG_MAGNETOSTATICS, Context, GoFerroError, Material, Problem, workloads)
from goferro.workloads import GYRO, pto_material

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def block(shape=(8, 8, 16), periodic=(True, True, False), flags=FLAG_ELECTROSTATICS, scale=1.0):
    fe = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    fe.r_permittivity = 12.0
    sub = Material.make(r_permittivity=25.0)
    cell = np.zeros(shape, np.uint32)
    cell[:, :, : shape[2] // 4] = 2
    cell[:, :, shape[2] // 4: 3 * shape[2] // 4] = 1
    prob = Problem(shape=shape, spacing=(1e-9, 1.2e-9, 0.8e-9), materials=[fe, sub], material_cell=cell,
                   dt=2e-15, periodic=periodic, flags=flags, Eext_stat=(0.0, 0.0, 2e6), scale_elec=scale)
    rng = np.random.default_rng(17)
    p = 0.05 * rng.standard_normal((3,) + shape)
    p[2] += 0.7
    p[:, cell != 1] = 0.0
    q = 1e8 * rng.standard_normal((3,) + shape)
    q[:, cell != 1] = 0.0
    return prob, p, q


def pair(prob, p, q):
    from oracle.oracle import Oracle
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_P, p), (FIELD_Q, q)):
        g.upload(fid, v)
        o.set(fid, v)
    return g, o


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("shape,scale", [((8, 8, 16), 1.0), ((16, 4, 32), 1.7), ((1, 1, 64), 1.0), ((6, 10, 15), 1.0),
                                         ((1, 1, 60), 1.3), ((7, 3, 20), 1.0)])
def test_static_field_matches_oracle(shape, scale):
    say("case electrostatic field", shape, scale)
    prob, p, q = block(shape, scale=scale)
    g, o = pair(prob, p, q)
    g.demag()
    o.demag()
    d = rel(g.download(FIELD_ESTAT), o.get(FIELD_ESTAT))
    say("  rel", d)
    assert d <= 1e-12
    g.close()
    o.close()


def test_ferroelectric_steps_with_electrostatics():
    say("case FE steps with electrostatics")
    prob, p, q = block()
    g, o = pair(prob, p, q)
    g.demag()
    o.demag()
    g.step(6)
    o.step(6)
    for fid in (FIELD_P, FIELD_Q, FIELD_ESTAT):
        d = rel(g.download(fid), o.get(fid))
        say("  field", fid, "rel", d)
        assert d <= 1e-11
    assert abs(g.time - o.time) <= 1e-30
    g.close()
    o.close()


def with_magnet(flags):
    """The block with its non-polar layer made a ferromagnet."""
    prob, p, q = block(flags=flags)
    prob.materials[1] = Material.make(if_FM=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, Aex=3.7e-12,
                                      r_permittivity=25.0)
    prob.demag_fac = (-0.2, -0.3, -0.5)
    m = workloads.tilted_m(prob.shape, 4, theta_max=0.4)
    m[:, np.asarray(prob.material_cell) != 2] = 0.0
    return prob, p, q, m


@pytest.mark.parametrize("mode", [FLAG_DEMAG_FACTOR, FLAG_MAG_1D])
def test_electrostatics_with_local_magnetostatics(mode):
    """The demag-factor and 1D magnetostatics need no FFT plan and run
    beside the FFT electrostatics."""
    from oracle.oracle import Oracle
    say("case electrostatics with local magnetostatics", hex(mode))
    prob, p, q, m = with_magnet(FLAG_ELECTROSTATICS | FLAG_MAGNETOSTATICS | mode)
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_P, p), (FIELD_Q, q), (FIELD_M, m)):
        g.upload(fid, v)
        o.set(fid, v)
    g.demag()
    o.demag()
    g.step(4)
    o.step(4)
    for fid in (FIELD_P, FIELD_Q, FIELD_ESTAT, FIELD_M, FIELD_HSTAT):
        d = rel(g.download(fid), o.get(fid))
        say("  field", fid, "rel", d)
        assert d <= 1e-11
    g.close()
    o.close()


@pytest.mark.parametrize("shape", [(8, 8, 16), (6, 5, 12)])
def test_electrostatics_with_fft_magnetostatics(shape):
    """A multiferroic block: FFT electrostatics of the polar layer and FFT
    magnetostatics of the magnetic one in one context, each on its own
    plan, after every update phase (main.cpp:299-306)."""
    from oracle.oracle import Oracle
    say("case electrostatics with fft magnetostatics", shape)
    prob, p, q = block(shape=shape, flags=FLAG_ELECTROSTATICS | FLAG_MAGNETOSTATICS)
    prob.materials[1] = Material.make(if_FM=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, Aex=3.7e-12,
                                      r_permittivity=25.0)
    m = workloads.tilted_m(prob.shape, 4, theta_max=0.4)
    m[:, np.asarray(prob.material_cell) != 2] = 0.0
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_P, p), (FIELD_Q, q), (FIELD_M, m)):
        g.upload(fid, v)
        o.set(fid, v)
    g.demag()
    o.demag()
    for fid in (FIELD_ESTAT, FIELD_HSTAT):
        d = rel(g.download(fid), o.get(fid))
        say("  initial field", fid, "rel", d)
        assert d <= 1e-12
    g.step(4)
    o.step(4)
    for fid in (FIELD_P, FIELD_Q, FIELD_ESTAT, FIELD_M, FIELD_HSTAT):
        d = rel(g.download(fid), o.get(fid))
        say("  field", fid, "rel", d)
        assert d <= 1e-11
    g.close()
    o.close()


def test_electrostatics_rejections():
    say("case electrostatics rejections")
    # without magnets the magnetostatics flag sets nothing up
    prob, _, _ = block(flags=FLAG_ELECTROSTATICS | FLAG_MAGNETOSTATICS)
    Context(prob).close()
    prob, _, _ = block(shape=(8, 8, 22))
    with pytest.raises(GoFerroError, match="prime factors"):
        Context(prob)
    shape = (4, 4, 4)
    prob = Problem(shape=shape, spacing=(1e-9,) * 3, materials=[Material.make(r_permittivity=3.0)],
                   material_cell=np.ones(shape, np.uint32), dt=1e-15, flags=FLAG_ELECTROSTATICS)
    with pytest.raises(GoFerroError, match="ferroelectric"):
        Context(prob)
