"""1000 coupled steps through gf_step for the couplings that act inside the
stages (spin pumping into the electrodynamics, flexoelectricity with
elastodynamics, the elastostatics acting on the magnets), against the
oracle at the parity bound of 1e-9 of each field's peak.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_H_EM, FIELD_J_ISHE, FIELD_M, FIELD_P, FIELD_Q, FIELD_STRAIN_T0_DIAG, FIELD_U,
                     FIELD_V, FLAG_1D_ABC, FLAG_ELASTO_ON_PM, FLAG_EMDYNAMIC, Context)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run(name, prob, fields, ids, nsteps=1000, statics=False):
    from oracle.oracle import Oracle
    say("case", name, prob.shape, nsteps, "steps")
    g, o = Context(prob), Oracle(prob)
    for fid, v in fields.items():
        g.upload(fid, v)
        o.set(fid, v)
    if statics:
        g.demag()
        o.demag()
    g.step(nsteps)
    o.step(nsteps)
    for fid in ids:
        d = rel(g.download(fid), o.get(fid))
        say("  field", fid, "rel", d)
        assert d <= 1e-9, (name, fid, d)
    g.close()
    o.close()


def test_spin_pumping_em_1000_steps():
    from test_gpu_spin import film
    prob, f = film(shape=(2, 2, 12), flags=FLAG_EMDYNAMIC | FLAG_1D_ABC)
    run("spin pumping + em", prob, f, (FIELD_M, FIELD_E_EM, FIELD_H_EM, FIELD_J_ISHE))


def test_flexo_elastic_1000_steps():
    from test_gpu_flexo import flexo_film
    prob, p = flexo_film(shape=(2, 2, 16))
    z = np.zeros((3,) + prob.shape)
    run("flexo + elastic", prob, {FIELD_P: p, FIELD_Q: z, FIELD_U: z, FIELD_V: z}, (FIELD_P, FIELD_Q, FIELD_U, FIELD_V))


def test_elastostatic_magnets_1000_steps():
    from test_gpu_elastostatic import mixed
    prob, f = mixed(shape=(4, 4, 8), flags=FLAG_ELASTO_ON_PM)
    run("elastostatics + magnets", prob, f, (FIELD_M, FIELD_STRAIN_T0_DIAG), statics=True)
