"""GPU parity of the flexoelectric coupling (GF_FLAG_FLEXO) through the C ABI
against the oracle (pinned in test_oracle_flexo.py), stage by stage, with
and without a rotated crystal.

The ferroelectric has no Landau energy and no gradient energy, and its
f11, f12, f44 are large, so that the flexoelectric force dominates its q
increment and the flexoelectric eigenstrain dominates the stress that
drives u (u and v start at zero): a wrong term shows at the 1e-12 bar.
The film is substrate | ferroelectric | vacuum along z with Liao at the
bottom, so FE surfaces and stress-free faces are both exercised.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_P, FIELD_P_STORE, FIELD_Q, FIELD_U, FIELD_V, FLAG_ELASTO_BACKACTION, FLAG_ELASTO_ON_PM,
                     FLAG_ELASTODYNAMICS, FLAG_FLEXO, FLAG_ORIENT_ROTATE, Context, GoFerroError, Material, Problem)
from goferro.workloads import set_elastic

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


def flexo_film(shape=(8, 8, 16), rotate=False):
    T = rot() if rotate else None
    fe = Material.make(if_FE=1, FE_mass=1e-16, FE_damping=2e-4)
    fe.Q11, fe.Q12, fe.Q44 = 2e-3, -1e-3, 3e-3
    set_elastic(fe, 2.3e11, 1.0e11, 1.1e11, 7900.0, Tg2c=T)
    fe.f11, fe.f12, fe.f44 = 3.0, -1.0, 2.0
    sub = set_elastic(Material.make(), 3.2e11, 1.0e11, 1.2e11, 5100.0, Tg2c=T)
    cell = np.zeros(shape, np.uint32)
    cell[:, :, :4] = 2
    cell[:, :, 4:12] = 1
    prob = Problem(shape=shape, spacing=(1e-9, 1.2e-9, 1e-9), materials=[fe, sub], material_cell=cell, dt=2e-15,
                   periodic=(True, True, False),
                   flags=FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM | FLAG_FLEXO |
                   (FLAG_ORIENT_ROTATE if rotate else 0), Tg2c=T, weighting_factor_fourth_Liao=0.5)
    rng = np.random.default_rng(41)
    p = 0.05 * rng.standard_normal((3,) + shape)
    p[2] += 0.3
    p[:, cell != 1] = 0.0
    return prob, p


@pytest.mark.parametrize("rotate", [False, True])
def test_flexo_stages(rotate):
    from test_gpu_elastic import run_stages
    prob, p = flexo_film(rotate=rotate)
    z = np.zeros((3,) + prob.shape)
    g, o = run_stages(prob, {FIELD_P: p, FIELD_Q: z, FIELD_U: z, FIELD_V: z}, nsteps=2, phys=("el", "fe"))
    for fid in (FIELD_P_STORE, FIELD_P, FIELD_Q, FIELD_U, FIELD_V):
        d = rel(g.download(fid), o.get(fid))
        say("  field", fid, "rel", d)
        assert d <= 1e-12, fid
    assert np.abs(o.get(FIELD_U)).max() > 0
    g.close()
    o.close()


def test_flexo_without_elastodynamics_and_rejections():
    from oracle.oracle import Oracle
    say("case flexo without elastodynamics")
    prob, p = flexo_film()
    prob.flags = FLAG_FLEXO
    g, o = Context(prob), Oracle(prob)
    for fid, v in ((FIELD_P, p),):
        g.upload(fid, v)
        o.set(fid, v)
    g.step(3)
    o.step(3)
    for fid in (FIELD_P, FIELD_Q):
        assert rel(g.download(fid), o.get(fid)) <= 1e-12, fid
    g.close()
    o.close()
    bad, _ = flexo_film()
    bad.materials[0].c12 = bad.materials[0].c11
    with pytest.raises(GoFerroError, match="c11 != c12"):
        Context(bad)
    nofe = Problem(shape=(4, 4, 8), spacing=(1e-9,) * 3, materials=[Material.make()],
                   material_cell=np.ones((4, 4, 8), np.uint32), dt=1e-15, flags=FLAG_FLEXO)
    with pytest.raises(GoFerroError, match="ferroelectric"):
        Context(nofe)
