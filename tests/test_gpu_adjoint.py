"""GPU tests of the adjoint sweep (gf_adjoint_*): the per-stage vector-Jacobian
product against central differences of the oracle's stage increment, the
reverse-sweep gradients against the oracle's tangent-linear reference
(itself pinned by finite differences on the CPU) and against central finite
differences of the GPU forward, checkpoint invariance, state restoration and
the torch.autograd binding.  The bar is 1e-6 relative (the stated
gradient agreement); most checks land far below it.
"""
import dataclasses
import sys

import numpy as np
import pytest

from goferro import (FIELD_HDRIVE, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF,
                     FLAG_ELASTO_ON_PM, FLAG_HDRIVE, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE, MU0, Context,
                     GoFerroError, Material, Problem, workloads)
from goferro.workloads import GYRO

pytestmark = pytest.mark.gpu

FREQ = 3e10


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def clone(mat):
    return type(mat).from_buffer_copy(mat)


def rot():
    c, s = np.cos(0.35), np.sin(0.35)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]]) @ np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def dmi_cells(shape):
    """Vacuum cells that give the magnet x and y surfaces away from the probe."""
    nx, ny, _ = shape
    cell = np.ones(shape, np.uint32)
    cell[nx - 2, ny // 2:ny // 2 + 2, :] = 0
    cell[nx // 2 + 1, 0, 1:] = 0
    return cell


def problem(shape=(16, 16, 8), alpha=0.02, aex=3.7e-12, b1=3.5e5, rotate=True, demag=True, anis=1, dmi=0.0):
    """With DMI: free x ends and vacuum cells, so that its surface slope acts."""
    mat = Material.make(if_FM=1, anisotropy_type=anis, Ms=1.4e5, gyro=GYRO, FM_damping=alpha, K1=-2e4, K2=3e3,
                        Aex=aex, B1=b1, B2=-2e5, iDMI=dmi)
    flags = FLAG_HDRIVE | FLAG_ELASTO_ON_PM | (FLAG_MAGNETOSTATICS if demag else 0)
    if rotate:
        flags |= FLAG_ORIENT_ROTATE
    cell = dmi_cells(shape) if dmi else np.ones(shape, np.uint32)
    return Problem(shape=shape, spacing=(2e-9,) * 3, materials=[mat], material_cell=cell, dt=2e-14,
                   periodic=(not dmi, True, False), flags=flags, Hext_stat=(0, 0, 0.1 / MU0), Hext_altr=(2e3, 0, 0),
                   H_altr_freq=(2e10, 0, 0), Hdrive_freq=FREQ, Tg2c=rot() if rotate else None)


def inputs(shape, seed=21, dmi=False):
    rng = np.random.default_rng(seed)
    m0 = workloads.tilted_m(shape, seed, theta_max=0.3)
    if dmi:
        m0[:, dmi_cells(shape) == 0] = 0.0
    hd = np.zeros((3,) + shape)
    hd[0, :, :, -1] = rng.uniform(0.5, 1.0, shape[:2]) * 4e4
    diag = rng.standard_normal((3,) + shape) * 1e-3
    off = rng.standard_normal((3,) + shape) * 1e-3
    return {FIELD_M: m0, FIELD_HDRIVE: hd, FIELD_STRAIN_T0_DIAG: diag, FIELD_STRAIN_T0_OFF: off}


def gpu(prob, fields):
    c = Context(prob)
    for k, v in fields.items():
        c.upload(k, v)
    return c


def oracle(prob, fields):
    from oracle.oracle import Oracle
    o = Oracle(prob)
    for k, v in fields.items():
        o.set(k, v)
    return o


def probe(shape):
    nx, ny, nz = shape
    return (nx // 4, ny // 4, 0), (nx // 4 + 3, ny // 4 + 3, 1)


@pytest.mark.parametrize("nadv,rotate,demag,dmi", [(1, True, True, 0.0), (3, False, True, 0.0), (5, True, False, 0.0),
                                                   (2, True, True, 1.2e-3), (1, False, False, -8e-4)])
def test_stage_vjp_against_oracle_differences(nadv, rotate, demag, dmi):
    """<ybar, v> = <kbar, dk/dy v> for the stage increment k = f(y), the
    right side from central differences of the oracle's k (a polynomial of
    low degree in y, so the difference is exact up to rounding), at the
    stage-1 time after nadv time steps; with interfacial DMI and its
    surface slopes at free x ends and vacuum cells as well."""
    from oracle.oracle import FIELD_K1
    say("case stage vjp", nadv, rotate, demag, dmi)
    prob = problem(shape=(8, 8, 4), rotate=rotate, demag=demag, dmi=dmi)
    f = inputs(prob.shape, dmi=bool(dmi))
    rng = np.random.default_rng(40 + nadv)
    y = f[FIELD_M] + 0.05 * rng.standard_normal((3,) + prob.shape)
    if dmi:
        y[:, dmi_cells(prob.shape) == 0] = 0.0
    kbar = rng.standard_normal((3,) + prob.shape)
    v = rng.standard_normal((3,) + prob.shape)
    g = gpu(prob, f)
    for _ in range(nadv):
        g.time_advance()
    ybar, grad, ghd = g.adjoint_stage(1, 1, y, kbar)

    def k_of(yy, mat=None):
        p = prob if mat is None else dataclasses.replace(prob, materials=[mat])
        o = oracle(p, {**f, FIELD_M: yy})
        if demag:
            o.demag()
        for _ in range(nadv):
            o.time_advance()
        o.L.gfo_mag_get_dm(o._h, 1)
        k = o.get(FIELD_K1)
        o.close()
        return k

    eps = 1e-4
    jv = (k_of(y + eps * v) - k_of(y - eps * v)) / (2 * eps)
    lhs, rhs = float((ybar * v).sum()), float((kbar * jv).sum())
    say("  <ybar,v>", lhs, "<kbar,Jv>", rhs)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-7)
    # parameter gradients of the same stage
    base = prob.materials[0]
    for idx, name, rel in ((0, "FM_damping", 1e-4), (1, "Aex", 1e-4), (2, "B1", 1e-4)):
        up, dn = clone(base), clone(base)
        h = rel * getattr(base, name)
        setattr(up, name, getattr(base, name) + h)
        setattr(dn, name, getattr(base, name) - h)
        fd = float((kbar * (k_of(y, up) - k_of(y, dn)) / (2 * h)).sum())
        np.testing.assert_allclose(grad[idx], fd, rtol=1e-6, err_msg=name)
    # drive profile: k is linear in it, so a large step (1e3 A/m per unit
    # of v) is exact and keeps the difference far above rounding
    sv = 1e3
    o_up = {**f, FIELD_HDRIVE: f[FIELD_HDRIVE] + sv * v}
    o_dn = {**f, FIELD_HDRIVE: f[FIELD_HDRIVE] - sv * v}

    def k_drive(ff):
        o = oracle(prob, {**ff, FIELD_M: y})
        if demag:
            o.demag()
        for _ in range(nadv):
            o.time_advance()
        o.L.gfo_mag_get_dm(o._h, 1)
        k = o.get(FIELD_K1)
        o.close()
        return k

    fd = float((kbar * (k_drive(o_up) - k_drive(o_dn)) / (2 * sv)).sum())
    np.testing.assert_allclose(float((ghd * v).sum()), fd, rtol=1e-7)
    g.close()


def two_material_problem(shape=(8, 8, 4), rotate=True, demag=True, dmi=False):
    """Two ferromagnets (different Ms, Aex, damping, B1, anisotropy and,
    with dmi, DMI of opposite signs) in stripes along x, a non-magnetic
    pocket, the same drives as problem()."""
    prob = problem(shape=shape, rotate=rotate, demag=demag)
    a = prob.materials[0]
    b = Material.make(if_FM=1, anisotropy_type=2, Ms=8e5, gyro=GYRO, FM_damping=0.05, K1=4e4, K2=0.0, Aex=1.1e-11,
                      B1=-6e5, B2=3e5)
    if dmi:
        a.iDMI, b.iDMI = 1.2e-3, -7e-4
    vac = Material.make()
    cell = np.ones(shape, np.uint32)
    cell[shape[0] // 2:, :, :] = 2
    cell[0, 0, 0] = 3
    cell[shape[0] - 1, shape[1] - 1, 1] = 3
    return dataclasses.replace(prob, materials=[a, b, vac], material_cell=cell)


@pytest.mark.parametrize("material,rotate,demag,dmi", [(1, True, True, False), (2, False, True, False),
                                                      (2, True, False, False), (1, True, True, True),
                                                      (2, False, False, True)])
def test_stage_vjp_two_materials(material, rotate, demag, dmi):
    """The stage VJP with two ferromagnets: <ybar, v> against the oracle's
    central differences, and the gradients of each material's alpha, Aex
    (its own faces and the harmonic-mean interface) and B1; with DMI of
    opposite signs in the two, the neighbours' DMI terms across the
    interface carry each neighbour's own D."""
    from oracle.oracle import FIELD_K1
    say("case stage vjp two materials", material, rotate, demag, dmi)
    prob = two_material_problem(rotate=rotate, demag=demag, dmi=dmi)
    f = inputs(prob.shape)
    magnet = np.asarray(prob.material_cell) < 3
    f[FIELD_M] = np.where(magnet[None], f[FIELD_M], 0.0)
    rng = np.random.default_rng(60 + material)
    y = f[FIELD_M] + 0.05 * rng.standard_normal((3,) + prob.shape) * magnet[None]
    kbar = rng.standard_normal((3,) + prob.shape)
    v = rng.standard_normal((3,) + prob.shape) * magnet[None]
    g = gpu(prob, f)
    g.time_advance()
    ybar, grad, _ = g.adjoint_stage(material, 1, y, kbar)

    def k_of(yy, mats=None):
        p = prob if mats is None else dataclasses.replace(prob, materials=mats)
        o = oracle(p, {**f, FIELD_M: yy})
        if demag:
            o.demag()
        o.time_advance()
        o.L.gfo_mag_get_dm(o._h, 1)
        k = o.get(FIELD_K1)
        o.close()
        return k

    eps = 1e-4
    jv = (k_of(y + eps * v) - k_of(y - eps * v)) / (2 * eps)
    lhs, rhs = float((ybar * v).sum()), float((kbar * jv).sum())
    say("  <ybar,v>", lhs, "<kbar,Jv>", rhs)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-7)
    for idx, name in ((0, "FM_damping"), (1, "Aex"), (2, "B1")):
        mats_up = [clone(m) for m in prob.materials]
        mats_dn = [clone(m) for m in prob.materials]
        base = getattr(prob.materials[material - 1], name)
        h = 1e-4 * base
        setattr(mats_up[material - 1], name, base + h)
        setattr(mats_dn[material - 1], name, base - h)
        fd = float((kbar * (k_of(y, mats_up) - k_of(y, mats_dn)) / (2 * h)).sum())
        say("  ", name, "adjoint", grad[idx], "fd", fd)
        np.testing.assert_allclose(grad[idx], fd, rtol=1e-6, err_msg=name)
    g.close()


def test_two_material_sweep_against_gpu_differences():
    """The reverse sweep with two ferromagnets: each material's gradients
    against central differences of the GPU forward loss, 20 steps."""
    say("case adjoint two materials vs gpu finite differences")
    prob = two_material_problem(shape=(16, 16, 8))
    f = inputs(prob.shape)
    magnet = np.asarray(prob.material_cell) < 3
    f[FIELD_M] = np.where(magnet[None], f[FIELD_M], 0.0)
    N = 20
    lo, hi = probe(prob.shape)
    g = gpu(prob, f)
    for material in (1, 2):
        g.upload(FIELD_M, f[FIELD_M])
        g.time = 0.0
        g.adjoint_forward(material, lo, hi, FREQ, N)
        grad, _ = g.adjoint_backward()
        base = clone(prob.materials[material - 1])

        def loss(p):
            g.set_material(material, p)
            g.upload(FIELD_M, f[FIELD_M])
            g.time = 0.0
            return g.adjoint_forward(material, lo, hi, FREQ, N)

        for idx, name in ((0, "FM_damping"), (1, "Aex"), (2, "B1")):
            h = 1e-5 * getattr(base, name)
            up, dn = clone(base), clone(base)
            setattr(up, name, getattr(base, name) + h)
            setattr(dn, name, getattr(base, name) - h)
            fd = (loss(up) - loss(dn)) / (2 * h)
            g.set_material(material, base)
            say("  material", material, name, "adjoint", grad[idx], "fd", fd)
            np.testing.assert_allclose(grad[idx], fd, rtol=1e-6)
    g.close()


def run_adjoint(prob, fields, nsteps, K=0):
    lo, hi = probe(prob.shape)
    g = gpu(prob, fields)
    L = g.adjoint_forward(1, lo, hi, FREQ, nsteps, checkpoint_every=K)
    grad, ghd = g.adjoint_backward()
    return g, L, grad, ghd


@pytest.mark.parametrize("dmi", [0.0, 1.2e-3])
def test_gradients_match_tangent_linear_reference(dmi):
    say("case adjoint vs tangent linear 16x16x8, 30 steps, K=7, DMI", dmi)
    prob = problem(dmi=dmi)
    f = inputs(prob.shape, dmi=bool(dmi))
    N = 30
    g, L, grad, ghd = run_adjoint(prob, f, N, K=7)
    lo, hi = probe(prob.shape)
    o = oracle(prob, f)
    L_o, _ = o.tl_probe_loss(1, lo, hi, FREQ, N)
    o.close()
    say("  loss gpu", L, "oracle", L_o)
    np.testing.assert_allclose(L, L_o, rtol=1e-9)
    for i in range(3):
        dpar = [0.0, 0.0, 0.0]
        dpar[i] = 1.0
        o = oracle(prob, f)
        _, dL = o.tl_probe_loss(1, lo, hi, FREQ, N, dpar=dpar)
        o.close()
        say("  param", i, "adjoint", grad[i], "tangent", dL)
        np.testing.assert_allclose(grad[i], dL, rtol=1e-6)
    v = np.random.default_rng(9).standard_normal((3,) + prob.shape)
    o = oracle(prob, f)
    _, dL = o.tl_probe_loss(1, lo, hi, FREQ, N, dhdrive=v)
    o.close()
    say("  drive <g,v>", float((ghd * v).sum()), "tangent", dL)
    np.testing.assert_allclose(float((ghd * v).sum()), dL, rtol=1e-6)
    g.close()


def test_checkpoint_spacing_does_not_change_gradients():
    say("case checkpoint invariance")
    prob = problem(shape=(8, 8, 4))
    f = inputs(prob.shape)
    out = []
    for K in (1, 4, 10, 0):
        g, L, grad, ghd = run_adjoint(prob, f, 10, K=K)
        out.append((L, grad, ghd))
        g.close()
    for L, grad, ghd in out[1:]:
        assert L == out[0][0]
        assert np.array_equal(grad, out[0][1])
        assert np.array_equal(ghd, out[0][2])


@pytest.mark.parametrize("demag", [True, False])
def test_resident_steps_do_not_change_gradients(demag):
    """Stage stores and fields kept by the forward pass for the last M steps
    (none, some, all of them) give the same loss, gradients, final state and
    time, bit for bit, as checkpointing every step."""
    say("case resident steps, demag", demag)
    prob = problem(shape=(16, 16, 8), demag=demag)
    f = inputs(prob.shape)
    lo, hi = probe(prob.shape)
    out = []
    for mib in (0, 1, 3, 100):
        g = gpu(prob, f)
        L = g.adjoint_forward(1, lo, hi, FREQ, 12, checkpoint_every=5 if mib == 1 else 0, resident_mib=mib)
        m_fwd, t_fwd = g.download(FIELD_M), g.time
        grad, ghd = g.adjoint_backward()
        assert np.array_equal(g.download(FIELD_M), m_fwd) and g.time == t_fwd
        g.step(2)
        out.append((L, grad, ghd, m_fwd, g.download(FIELD_M)))
        g.close()
    for L, grad, ghd, m_fwd, m_after in out[1:]:
        assert L == out[0][0]
        assert np.array_equal(grad, out[0][1]) and np.array_equal(ghd, out[0][2])
        assert np.array_equal(m_fwd, out[0][3]) and np.array_equal(m_after, out[0][4])


def test_gradients_match_gpu_finite_differences():
    """Central differences of the GPU forward loss at 32x32x8, 60 steps:
    the three scalars and four random drive cells of the top layer."""
    say("case adjoint vs gpu finite differences")
    shape = (32, 32, 8)
    prob = problem(shape=shape)
    f = inputs(shape)
    N = 60
    lo, hi = probe(shape)
    g = gpu(prob, f)
    L0 = g.adjoint_forward(1, lo, hi, FREQ, N)
    grad, ghd = g.adjoint_backward()

    def loss(p=None, hd=None):
        c = g
        if p is not None:
            c.set_material(1, p)
        c.upload(FIELD_M, f[FIELD_M])
        c.upload(FIELD_HDRIVE, f[FIELD_HDRIVE] if hd is None else hd)
        c.time = 0.0
        return c.adjoint_forward(1, lo, hi, FREQ, N)

    base = clone(prob.materials[0])
    # relative step 1e-5: the truncation error of the central difference
    # (h^2/6 times the third derivative) stays near 1e-8 even where the loss
    # bends sharply with Aex; rounding adds ~1e-15/1e-5
    for idx, name in ((0, "FM_damping"), (1, "Aex"), (2, "B1")):
        h = 1e-5 * getattr(base, name)
        up, dn = clone(base), clone(base)
        setattr(up, name, getattr(base, name) + h)
        setattr(dn, name, getattr(base, name) - h)
        fd = (loss(up) - loss(dn)) / (2 * h)
        loss(base)
        say("  ", name, "adjoint", grad[idx], "fd", fd)
        np.testing.assert_allclose(grad[idx], fd, rtol=1e-6)
    rng = np.random.default_rng(3)
    for _ in range(4):
        i, j = int(rng.integers(shape[0])), int(rng.integers(shape[1]))
        # one cell's amplitude moves the loss by ~1e-11 of itself per percent,
        # near its rounding noise: a step of twice the amplitude keeps the
        # difference well above it, and the loss is quadratic in the drive to
        # far below 1e-6 (linear response), so the central difference stays
        # exact (a half-amplitude step left it within 5% of the 1e-6 bar)
        h = 2.0 * f[FIELD_HDRIVE][0, i, j, -1]
        up, dn = f[FIELD_HDRIVE].copy(), f[FIELD_HDRIVE].copy()
        up[0, i, j, -1] += h
        dn[0, i, j, -1] -= h
        fd = (loss(hd=up) - loss(hd=dn)) / (2 * h)
        say("  drive cell", i, j, "adjoint", ghd[0, i, j, -1], "fd", fd)
        np.testing.assert_allclose(ghd[0, i, j, -1], fd, rtol=1e-6)
    assert L0 > 0
    g.close()


def test_dmi_sweep_against_gpu_differences():
    """Interfacial DMI with its surface slopes (free x ends, vacuum cells):
    the FP64 sweep's alpha and Aex gradients and the drive gradient of two
    top-layer cells against central differences of the GPU forward loss,
    16x16x8, 40 steps."""
    say("case adjoint with DMI vs gpu finite differences")
    shape = (16, 16, 8)
    prob = problem(shape=shape, dmi=1.2e-3)
    f = inputs(shape, dmi=True)
    N = 40
    lo, hi = probe(shape)
    g = gpu(prob, f)
    L0 = g.adjoint_forward(1, lo, hi, FREQ, N, resident_mib=64)
    grad, ghd = g.adjoint_backward()

    def loss(p=None, hd=None):
        if p is not None:
            g.set_material(1, p)
        g.upload(FIELD_M, f[FIELD_M])
        g.upload(FIELD_HDRIVE, f[FIELD_HDRIVE] if hd is None else hd)
        g.time = 0.0
        return g.adjoint_forward(1, lo, hi, FREQ, N)

    base = clone(prob.materials[0])
    for idx, name in ((0, "FM_damping"), (1, "Aex")):
        h = 1e-5 * getattr(base, name)
        up, dn = clone(base), clone(base)
        setattr(up, name, getattr(base, name) + h)
        setattr(dn, name, getattr(base, name) - h)
        fd = (loss(up) - loss(dn)) / (2 * h)
        loss(base)
        say("  ", name, "adjoint", grad[idx], "fd", fd)
        np.testing.assert_allclose(grad[idx], fd, rtol=1e-6)
    for i, j in ((3, 5), (13, 9)):
        h = 2.0 * f[FIELD_HDRIVE][0, i, j, -1]
        up, dn = f[FIELD_HDRIVE].copy(), f[FIELD_HDRIVE].copy()
        up[0, i, j, -1] += h
        dn[0, i, j, -1] -= h
        fd = (loss(hd=up) - loss(hd=dn)) / (2 * h)
        say("  drive cell", i, j, "adjoint", ghd[0, i, j, -1], "fd", fd)
        np.testing.assert_allclose(ghd[0, i, j, -1], fd, rtol=1e-6)
    assert L0 > 0
    g.close()


def test_backward_restores_the_forward_final_state():
    say("case state restore")
    prob = problem(shape=(8, 8, 4))
    f = inputs(prob.shape)
    g = gpu(prob, f)
    lo, hi = probe(prob.shape)
    g.adjoint_forward(1, lo, hi, FREQ, 12, checkpoint_every=5)
    m_fwd, t_fwd = g.download(FIELD_M), g.time
    g.adjoint_backward()
    assert np.array_equal(g.download(FIELD_M), m_fwd)
    assert g.time == t_fwd
    # the context keeps stepping like a plain run of the same length
    g.step(3)
    h = gpu(prob, f)
    h.demag()
    h.step(15)
    np.testing.assert_allclose(g.download(FIELD_M), h.download(FIELD_M), rtol=0, atol=1e-13)
    g.close()
    h.close()


def test_adjoint_rejections():
    say("case adjoint rejections")
    prob = problem(shape=(4, 4, 4))
    f = inputs(prob.shape)
    g = gpu(prob, f)
    with pytest.raises(GoFerroError, match="gf_adjoint_forward first"):
        g.adjoint_backward()
    with pytest.raises(GoFerroError, match="probe"):
        g.adjoint_forward(1, (0, 0, 0), (9, 0, 0), FREQ, 2)
    with pytest.raises(GoFerroError, match="resident_mib"):
        g.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2, resident_mib=-1)
    g.close()
    pa, m1, m2 = workloads.afm_block((4, 4, 4))
    a = Context(pa)
    a.upload(FIELD_M_AFM1, m1)
    a.upload(FIELD_M_AFM2, m2)
    with pytest.raises(GoFerroError, match="ferromagnet-only"):
        a.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2)
    a.close()
    # DMI: the FP64 sweep takes it, the FP32 variant refuses it
    pd = problem(shape=(4, 4, 4), rotate=False)
    mat = pd.materials[0]
    mat.iDMI = 1e-3
    d = gpu(pd, f)
    d.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2)
    with pytest.raises(GoFerroError, match="DMI"):
        d.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2, precision=32)
    d.close()
    # the elastostatic solver recomputes the static strain from m after
    # every step, which the sweep would not differentiate: refused, in FP64
    # and FP32
    from goferro import FLAG_ELASTOSTATIC
    pe = problem(shape=(4, 4, 4), rotate=False)
    pe.flags |= FLAG_ELASTOSTATIC
    from goferro.workloads import set_elastic
    set_elastic(pe.materials[0], 2.69e11, 1.08e11, 0.76e11, 5170.0)
    e = gpu(pe, {FIELD_M: f[FIELD_M], FIELD_HDRIVE: f[FIELD_HDRIVE]})
    for prec in (64, 32):
        with pytest.raises(GoFerroError, match="elastostatic"):
            e.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2, precision=prec)
    e.close()


def test_torch_autograd_function():
    import torch

    from goferro.autograd import ProbeLossSim, probe_loss
    say("case torch autograd")
    prob = problem(shape=(16, 16, 8))
    f = inputs(prob.shape)
    lo, hi = probe(prob.shape)
    sim = ProbeLossSim(prob, f[FIELD_M], material=1, probe_lo=lo, probe_hi=hi, freq=FREQ, nsteps=20,
                       drive_vector=(4e4, 0.0, 0.0), drive_layer=prob.shape[2] - 1,
                       static_fields={FIELD_STRAIN_T0_DIAG: f[FIELD_STRAIN_T0_DIAG],
                                      FIELD_STRAIN_T0_OFF: f[FIELD_STRAIN_T0_OFF]})
    amp = torch.tensor(f[FIELD_HDRIVE][0, :, :, -1] / 4e4, dtype=torch.float64, requires_grad=True)
    alpha = torch.tensor(0.02, dtype=torch.float64, requires_grad=True)
    aex = torch.tensor(3.7e-12, dtype=torch.float64, requires_grad=True)
    b1 = torch.tensor(3.5e5, dtype=torch.float64, requires_grad=True)
    L = probe_loss(sim, alpha, aex, b1, amp)
    L.backward()
    g, L2, grad, ghd = run_adjoint(prob, f, 20)
    np.testing.assert_allclose(float(L.detach()), L2, rtol=1e-12)
    np.testing.assert_allclose([alpha.grad.item(), aex.grad.item(), b1.grad.item()], grad, rtol=1e-12)
    np.testing.assert_allclose(amp.grad.numpy(), 4e4 * ghd[0, :, :, -1], rtol=1e-12, atol=0)
    g.close()
    sim.close()


# --- FP32 variant (gf_adjoint_desc.precision = 32) ------------------------
# The FP32 sweep runs the same forward and discrete adjoint in FP32.  Its
# bar is agreement with the FP64 sweep to FP32 accuracy over the trajectory:
# loss 1e-5, scalar gradients 1e-4, drive gradient 1e-5 relative (observed
# 4e-7, 1e-5 for alpha and below 1e-6 for the rest, 3e-7).

def run_adjoint32(prob, fields, nsteps, K=0):
    lo, hi = probe(prob.shape)
    g = gpu(prob, fields)
    L = g.adjoint_forward(1, lo, hi, FREQ, nsteps, checkpoint_every=K, precision=32)
    grad, ghd = g.adjoint_backward()
    return g, L, grad, ghd


@pytest.mark.parametrize("demag,anis", [(True, 1), (False, 2)])
def test_fp32_sweep_tracks_fp64(demag, anis):
    say("case FP32 adjoint vs FP64, 16x16x8, 40 steps, demag", demag, "anisotropy", anis)
    prob = problem(rotate=False, demag=demag, anis=anis)
    f = inputs(prob.shape)
    N = 40
    g64, L64, grad64, ghd64 = run_adjoint(prob, f, N)
    g32, L32, grad32, ghd32 = run_adjoint32(prob, f, N)
    say("  loss", L64, L32, "rel", abs(L32 - L64) / L64)
    assert abs(L32 - L64) <= 1e-5 * L64
    for i in range(3):
        rel = abs(grad32[i] - grad64[i]) / abs(grad64[i])
        say("  param", i, grad64[i], grad32[i], "rel", rel)
        assert rel <= 1e-4
    rel = np.linalg.norm(ghd32 - ghd64) / np.linalg.norm(ghd64)
    say("  drive gradient rel", rel)
    assert rel <= 1e-5
    g64.close()
    g32.close()


def test_fp32_checkpoints_and_state():
    """The FP32 gradients do not depend on the checkpoint spacing (the
    recomputed FP32 trajectory is bit-identical), and the sweep leaves the
    context's FP64 state and time unchanged."""
    say("case FP32 checkpoints and state")
    prob = problem(shape=(8, 8, 4), rotate=False)
    f = inputs(prob.shape)
    out = []
    for K in (1, 4, 0):
        g = gpu(prob, f)
        m0, t0 = g.download(FIELD_M), g.time
        lo, hi = probe(prob.shape)
        L = g.adjoint_forward(1, lo, hi, FREQ, 10, checkpoint_every=K, precision=32)
        assert np.array_equal(g.download(FIELD_M), m0) and g.time == t0
        grad, ghd = g.adjoint_backward()
        assert np.array_equal(g.download(FIELD_M), m0) and g.time == t0
        out.append((L, grad, ghd))
        g.close()
    for L, grad, ghd in out[1:]:
        assert L == out[0][0]
        assert np.array_equal(grad, out[0][1])
        assert np.array_equal(ghd, out[0][2])


def test_fp32_rejections():
    say("case FP32 rejections")
    prob = problem(shape=(4, 4, 4), rotate=True)
    f = inputs(prob.shape)
    g = gpu(prob, f)
    with pytest.raises(GoFerroError, match="rotation"):
        g.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2, precision=32)
    with pytest.raises(GoFerroError, match="precision"):
        g.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2, precision=16)
    g.close()
    p2 = two_material_problem(shape=(4, 4, 4), rotate=False)
    g = gpu(p2, {FIELD_M: f[FIELD_M], FIELD_HDRIVE: f[FIELD_HDRIVE]})
    with pytest.raises(GoFerroError, match="one FM material"):
        g.adjoint_forward(1, (0, 0, 0), (1, 1, 1), FREQ, 2, precision=32)
    g.close()


def test_fp32_torch_autograd_function():
    import torch

    from goferro.autograd import ProbeLossSim, probe_loss
    say("case torch autograd FP32")
    prob = problem(shape=(8, 8, 4), rotate=False)
    f = inputs(prob.shape)
    lo, hi = probe(prob.shape)
    sim = ProbeLossSim(prob, f[FIELD_M], material=1, probe_lo=lo, probe_hi=hi, freq=FREQ, nsteps=12,
                       drive_vector=(4e4, 0.0, 0.0), drive_layer=prob.shape[2] - 1,
                       static_fields={FIELD_STRAIN_T0_DIAG: f[FIELD_STRAIN_T0_DIAG],
                                      FIELD_STRAIN_T0_OFF: f[FIELD_STRAIN_T0_OFF]}, precision=32)
    amp = torch.tensor(f[FIELD_HDRIVE][0, :, :, -1] / 4e4, dtype=torch.float64, requires_grad=True)
    alpha = torch.tensor(0.02, dtype=torch.float64, requires_grad=True)
    aex = torch.tensor(3.7e-12, dtype=torch.float64, requires_grad=True)
    b1 = torch.tensor(3.5e5, dtype=torch.float64, requires_grad=True)
    L = probe_loss(sim, alpha, aex, b1, amp)
    L.backward()
    g, L2, grad, ghd = run_adjoint32(prob, f, 12)
    np.testing.assert_allclose(float(L.detach()), L2, rtol=1e-12)
    np.testing.assert_allclose([alpha.grad.item(), aex.grad.item(), b1.grad.item()], grad, rtol=1e-12)
    np.testing.assert_allclose(amp.grad.numpy(), 4e4 * ghd[0, :, :, -1], rtol=1e-12, atol=0)
    g.close()
    sim.close()
