"""Pin the oracle's elastostatics (GF_FLAG_ELASTOSTATIC; get_strain_static
is declared at elastic_system.h:190 but not defined in the reference, so
its set-up, elastic_initialize.cpp:22-40 and :340-389, and closed forms fix
it).  The eigenstrain is the magnetostrictive one of a cubic crystal,
e0_ii = 3/2 lambda100 (m_i^2 - 1/3), e0_ij = 3/2 lambda111 m_i m_j
(textbook form, restated here in numpy):

* uniform m: the strain is the eigenstrain, everywhere;
* m varying along z only: the laterally periodic solution
  e_xx, e_yy, e_xy = their means, e_xz = e0_xz, e_yz = e0_yz,
  e_zz = e0_zz + (c12 / c11) (e0_xx + e0_yy - mean), from sigma_iz = 0;
* random m: at every non-Nyquist k != 0 the stress C (e - e0) is
  divergence-free (k . sigma = 0) and the strain is compatible
  (Saint-Venant: k x (k x e)^T = 0), and the mean strain is the mean
  eigenstrain; together these fix the solution;
* the homogeneous stiffness is the component-wise largest c11, c12, c44
  times scale_elasto.
"""
import numpy as np
import pytest

from goferro import (FIELD_M, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTOSTATIC, Material, Problem,
                     workloads)
from goferro.workloads import GYRO, set_elastic

C11, C12, C44 = 2.69e11, 1.08e11, 0.76e11
L100, L111 = -1.4e-5, 2.2e-5


def magnet(c=(C11, C12, C44)):
    m = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.1, Aex=3.7e-12,
                      lamda100=L100, lamda111=L111)
    return set_elastic(m, c[0], c[1], c[2], 5170.0)


def problem(shape, mats=None, cell=None, scale=1.0, spacing=(1e-9, 1.5e-9, 2e-9)):
    mats = mats or [magnet()]
    cell = np.ones(shape, np.uint32) if cell is None else cell
    return Problem(shape=shape, spacing=spacing, materials=mats, material_cell=cell, dt=1e-14,
                   periodic=(True, True, True), flags=FLAG_ELASTOSTATIC, scale_elasto=scale)


def static_strain(oracle_lib, prob, m):
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m)
    o.demag()
    e = np.concatenate([o.get(FIELD_STRAIN_T0_DIAG), o.get(FIELD_STRAIN_T0_OFF)])
    o.close()
    return e


def eigen(m, l100=L100, l111=L111):
    e = np.empty((6,) + m.shape[1:])
    for a in range(3):
        e[a] = 1.5 * l100 * (m[a] ** 2 - 1. / 3.)
    e[3] = 1.5 * l111 * m[1] * m[2]
    e[4] = 1.5 * l111 * m[0] * m[2]
    e[5] = 1.5 * l111 * m[0] * m[1]
    return e


def test_uniform_magnetisation_gives_its_eigenstrain(oracle_lib):
    shape = (4, 4, 8)
    m = np.zeros((3,) + shape)
    m[:] = (np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8]))[:, None, None, None]
    e = static_strain(oracle_lib, problem(shape), m)
    want = eigen(m)
    np.testing.assert_allclose(e, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_profile_along_z(oracle_lib):
    shape = (2, 2, 16)
    z = np.arange(16)
    th = 0.9 * np.sin(2 * np.pi * z / 16) + 0.4 * np.cos(6 * np.pi * z / 16)
    ph = 0.3 + 0.5 * np.sin(4 * np.pi * z / 16)
    m = np.zeros((3,) + shape)
    m[0] = (np.sin(th) * np.cos(ph))[None, None, :]
    m[1] = (np.sin(th) * np.sin(ph))[None, None, :]
    m[2] = np.cos(th)[None, None, :]
    e = static_strain(oracle_lib, problem(shape), m)
    e0 = eigen(m)
    want = np.empty_like(e0)
    for a in (0, 1, 5):
        want[a] = e0[a].mean()
    want[3], want[4] = e0[3], e0[4]
    s = e0[0] + e0[1]
    want[2] = e0[2] + C12 / C11 * (s - s.mean())
    np.testing.assert_allclose(e, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("rotate", [False, True])
def test_random_field_is_in_equilibrium_and_compatible(oracle_lib, rotate):
    shape = (8, 8, 8)
    spacing = (1e-9, 1.5e-9, 2e-9)
    prob = problem(shape, spacing=spacing)
    T = np.eye(3)
    if rotate:
        c, s = np.cos(0.4), np.sin(0.4)
        T = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, np.cos(0.3), np.sin(0.3)],
                                                                       [0, -np.sin(0.3), np.cos(0.3)]])
        from goferro import FLAG_ORIENT_ROTATE
        prob.flags |= FLAG_ORIENT_ROTATE
        prob.Tg2c = T
    m = workloads.tilted_m(shape, 17, theta_max=1.4)
    e = static_strain(oracle_lib, prob, m)
    mc = np.einsum("ab,b...->a...", T, m)              # crystal frame
    e0 = eigen(mc)
    assert np.abs(e.reshape(6, -1).mean(axis=1) - e0.reshape(6, -1).mean(axis=1)).max() <= 1e-12 * np.abs(e0).max()
    d = e - e0
    sig = np.empty_like(d)
    sig[0] = C11 * d[0] + C12 * (d[1] + d[2])
    sig[1] = C11 * d[1] + C12 * (d[0] + d[2])
    sig[2] = C11 * d[2] + C12 * (d[0] + d[1])
    sig[3:] = 2 * C44 * d[3:]
    Sk = np.array([np.fft.fftn(x) for x in sig])
    Ek = np.array([np.fft.fftn(x) for x in e])
    kv = [2 * np.pi * np.fft.fftfreq(n, h) for n, h in zip(shape, spacing)]
    nyq = [np.isclose(np.abs(np.fft.fftfreq(n)), 0.5) for n in shape]
    K = np.array(np.meshgrid(*kv, indexing="ij"))
    Kc = np.einsum("ab,b...->a...", T, K)              # crystal-frame wave vectors
    N = np.meshgrid(*nyq, indexing="ij")
    ok = ~(N[0] | N[1] | N[2]) & ((K ** 2).sum(0) > 0)
    full = lambda v: np.array([[v[0], v[5], v[4]], [v[5], v[1], v[3]], [v[4], v[3], v[2]]])
    S, E = full(Sk), full(Ek)
    div = np.einsum("ij...,j...->i...", S, Kc)
    smax = np.abs(np.einsum("ij...,j...->i...", np.abs(S), np.abs(Kc))).max()
    assert np.abs(div[:, ok]).max() <= 1e-10 * smax
    # Saint-Venant: eps_ijk eps_lmn k_j k_m E_kn = 0
    lc = np.zeros((3, 3, 3))
    lc[0, 1, 2] = lc[1, 2, 0] = lc[2, 0, 1] = 1
    lc[0, 2, 1] = lc[2, 1, 0] = lc[1, 0, 2] = -1
    inc = np.einsum("ijk,lmn,j...,m...,kn...->il...", lc, lc, Kc, Kc, E)
    emax = np.abs(np.einsum("j...,m...,kn...->...", np.abs(Kc), np.abs(Kc), np.abs(E))).max()
    assert np.abs(inc[:, :, ok]).max() <= 1e-10 * emax


def test_homogeneous_stiffness_is_the_largest_scaled(oracle_lib):
    # two magnets with different constants: the z-profile closed form
    # holds with the component-wise largest c11 and c12 (times the scale)
    shape = (2, 2, 8)
    mats = [magnet((2.0e11, 1.3e11, 0.5e11)), magnet((2.5e11, 0.9e11, 0.7e11))]
    cell = np.ones(shape, np.uint32)
    cell[:, :, 4:] = 2
    z = np.arange(8)
    th = 0.7 * np.sin(2 * np.pi * z / 8)
    m = np.zeros((3,) + shape)
    m[0] = np.sin(th)[None, None, :]
    m[2] = np.cos(th)[None, None, :]
    e = static_strain(oracle_lib, problem(shape, mats=mats, cell=cell, scale=1.7), m)
    e0 = eigen(m)
    s = e0[0] + e0[1]
    want = e0[2] + 1.3e11 / 2.5e11 * (s - s.mean())
    np.testing.assert_allclose(e[2], want, rtol=0, atol=1e-12 * np.abs(want).max())


# ---- ferroelectrics: the electrostrictive eigenstrain e0_ii = Q11 p_i^2 +
# Q12 (p_j^2 + p_k^2), e0_ij = Q44 p_i p_j (elastic_dynamic_solver.cpp:
# 39-65, textbook form restated in numpy), and the ferroelectric reading the
# static strain and eigenstrain (ferroelectric_solver.cpp:636-650).

Q11, Q12, Q44 = 0.089, -0.026, 0.0675
TV1, TV2, TV3 = 3.05e10, 6.4e8, 2.97e10


def ferroelectric(c=(2.3e11, 1.0e11, 1.1e11)):
    from goferro.workloads import pto_material
    f = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    f.Q11, f.Q12, f.Q44 = Q11, Q12, Q44
    f.tv1, f.tv2, f.tv3 = TV1, TV2, TV3
    return set_elastic(f, c[0], c[1], c[2], 7900.0)


def fe_eigen(p):
    e = np.empty((6,) + p.shape[1:])
    for a in range(3):
        b, c = [x for x in range(3) if x != a]
        e[a] = Q11 * p[a] ** 2 + Q12 * (p[b] ** 2 + p[c] ** 2)
    e[3] = Q44 * p[1] * p[2]
    e[4] = Q44 * p[0] * p[2]
    e[5] = Q44 * p[0] * p[1]
    return e


def fe_profile(shape):
    z = np.arange(shape[2])
    p = np.zeros((3,) + shape)
    p[0] = (0.1 * np.sin(2 * np.pi * z / shape[2]))[None, None, :]
    p[1] = (-0.05 + 0.08 * np.cos(4 * np.pi * z / shape[2]))[None, None, :]
    p[2] = (0.5 + 0.2 * np.sin(2 * np.pi * z / shape[2] + 0.3))[None, None, :]
    return p


def fe_problem(shape, flags=0):
    from goferro import FLAG_ELASTO_ON_PM
    cell = np.ones(shape, np.uint32)
    return Problem(shape=shape, spacing=(1e-9, 1e-9, 1e-9), materials=[ferroelectric()], material_cell=cell,
                   dt=2e-15, periodic=(True, True, True), flags=FLAG_ELASTOSTATIC | flags)


def test_ferroelectric_uniform_and_profile(oracle_lib):
    """Uniform p: the strain and the kept static eigenstrain are the
    eigenstrain; p along z: the laterally periodic closed form, with the
    component-wise largest stiffness (here the one material's)."""
    from goferro import FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF, FIELD_P
    c11, c12 = 2.3e11, 1.0e11
    for uniform in (True, False):
        shape = (2, 2, 16)
        p = fe_profile(shape)
        if uniform:
            p[:] = p[:, :, :, :1]
        o = oracle_lib.Oracle(fe_problem(shape))
        o.set(FIELD_P, p)
        o.demag()
        e = np.concatenate([o.get(FIELD_STRAIN_T0_DIAG), o.get(FIELD_STRAIN_T0_OFF)])
        e0k = np.concatenate([o.get(FIELD_EIGENSTRAIN_T0_DIAG), o.get(FIELD_EIGENSTRAIN_T0_OFF)])
        o.close()
        e0 = fe_eigen(p)
        np.testing.assert_allclose(e0k, e0, rtol=0, atol=1e-15 * np.abs(e0).max())
        if uniform:
            want = e0
        else:
            want = np.empty_like(e0)
            for a in (0, 1, 5):
                want[a] = e0[a].mean()
            want[3], want[4] = e0[3], e0[4]
            s = e0[0] + e0[1]
            want[2] = e0[2] + c12 / c11 * (s - s.mean())
        np.testing.assert_allclose(e, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_ferroelectric_reads_static_strain_and_eigenstrain(oracle_lib):
    """The stage-1 q increment with the coupling on, less the one without,
    is -dt/m [(tv1 (e0xx - exx) + tv2 (e0yy - eyy + e0zz - ezz)) px
    + tv3 ((e0xy - exy) py + (e0xz - exz) pz)] and its y, z analogues, with
    the closed-form static strain of a p profile along z: a wrong array, sign
    or frame in the wiring breaks it."""
    from goferro import FIELD_P, FLAG_ELASTO_ON_PM
    shape = (2, 2, 16)
    p = fe_profile(shape)
    c11, c12 = 2.3e11, 1.0e11
    kq = []
    for flags in (FLAG_ELASTO_ON_PM, 0):
        o = oracle_lib.Oracle(fe_problem(shape, flags))
        o.set(FIELD_P, p)
        o.demag()
        o.L.gfo_fe_get(o._h, 1)
        kq.append(o.get(106))   # the stage-1 q increment
        o.close()
    e0 = fe_eigen(p)
    e = np.empty_like(e0)
    for a in (0, 1, 5):
        e[a] = e0[a].mean()
    e[3], e[4] = e0[3], e0[4]
    s = e0[0] + e0[1]
    e[2] = e0[2] + c12 / c11 * (s - s.mean())
    d = e0 - e     # e0 - e, Voigt
    pre = -2e-15 / 1e-16
    want = np.empty((3,) + shape)
    want[0] = pre * ((TV1 * d[0] + TV2 * (d[1] + d[2])) * p[0] + TV3 * (d[5] * p[1] + d[4] * p[2]))
    want[1] = pre * ((TV1 * d[1] + TV2 * (d[0] + d[2])) * p[1] + TV3 * (d[5] * p[0] + d[3] * p[2]))
    want[2] = pre * ((TV1 * d[2] + TV2 * (d[0] + d[1])) * p[2] + TV3 * (d[4] * p[0] + d[3] * p[1]))
    got = kq[0] - kq[1]
    assert np.abs(want).max() > 1e-3 * np.abs(kq[1]).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * np.abs(kq[1]).max())


def beside_problem(flags, shape=(8, 8, 16)):
    """A ferroelectric layer, a magnet and a substrate (the reference's
    multiferroic stack) with random p and m, the strain acting on both."""
    from goferro import FIELD_M, FIELD_P, FLAG_ELASTO_ON_PM
    from goferro.workloads import GYRO, set_magnetostriction
    fe = ferroelectric()
    mag = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.05, K1=-610.0, Aex=3.7e-12)
    set_elastic(mag, 2.69e11, 1.08e11, 0.76e11, 5170.0)
    set_magnetostriction(mag, 3.5e5, 2.0e5)
    sub = set_elastic(Material.make(), 3.2e11, 1.0e11, 1.2e11, 5100.0)
    cell = np.full(shape, 3, np.uint32)
    cell[:, :, 4:10] = 1
    cell[:, :, 10:] = 2
    prob = Problem(shape=shape, spacing=(1e-9, 1e-9, 1e-9), materials=[fe, mag, sub], material_cell=cell,
                   dt=2e-15, periodic=(True, True, False), Hext_stat=(0, 0, 5e4),
                   flags=FLAG_ELASTO_ON_PM | flags)
    rng = np.random.default_rng(43)
    p = 0.05 * rng.standard_normal((3,) + shape)
    p[2] += 0.6
    p[:, cell != 1] = 0.0
    m = np.zeros((3,) + shape)
    m[:, cell == 2] = np.array([0.3, 0.2, np.sqrt(1 - 0.13)])[:, None]
    m[0] += 0.05 * rng.standard_normal(shape) * (cell == 2)
    m /= np.maximum(np.linalg.norm(m, axis=0), 1e-300)
    return prob, {FIELD_P: p, FIELD_M: m}


def test_start_up_solve_beside_elastodynamics(oracle_lib):
    """With elastodynamics the static strain is solved once at start-up
    (main.cpp:148-158) and the dynamics carries the change from t0: the
    eigenstrain of the elastic stage less the static one is zero at t0, so
    with u = 0 the stage-1 velocity increment vanishes; the magnet reads
    exxt0 + Dexx and the ferroelectric Dexx + exxt0, Dexx0 + exx0t0, so their
    stage-1 increments equal those of the elastostatics alone.  After steps
    the static strain is still the start-up one (without elastodynamics it
    follows the stores), and an uploaded static strain replaces the solve."""
    from goferro import (FIELD_STRAIN_T0_DIAG, FLAG_ELASTO_BACKACTION, FLAG_ELASTODYNAMICS)
    dyn = FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION
    out = {}
    for name, flags in (("static", FLAG_ELASTOSTATIC), ("both", FLAG_ELASTOSTATIC | dyn), ("dyn", dyn)):
        prob, f = beside_problem(flags)
        o = oracle_lib.Oracle(prob)
        for fid, v in f.items():
            o.set(fid, v)
        o.demag()
        e_start = o.get(FIELD_STRAIN_T0_DIAG)
        if flags & FLAG_ELASTODYNAMICS:
            o.L.gfo_elastic_get(o._h, 1)
        o.L.gfo_mag_get_dm(o._h, 1)
        o.L.gfo_fe_get(o._h, 1)
        out[name] = {"k1": o.get(103), "kq": o.get(106), "e": e_start,
                     "dv": o.get(102) if flags & FLAG_ELASTODYNAMICS else None}
        o.close()
    s, b, d = out["static"], out["both"], out["dyn"]
    assert np.array_equal(b["e"], s["e"]) and np.abs(s["e"]).max() > 1e-6
    assert np.abs(d["dv"]).max() > 0.
    assert np.abs(b["dv"]).max() <= 1e-12 * np.abs(d["dv"]).max()
    for key in ("k1", "kq"):
        scale = np.abs(s[key]).max()
        assert np.abs(s[key] - d[key]).max() > 1e-6 * scale, key   # the static strain matters
        np.testing.assert_allclose(b[key], s[key], rtol=0, atol=1e-13 * scale)
    # steps: the start-up solve is kept with elastodynamics, redone without
    for flags, kept in ((FLAG_ELASTOSTATIC | dyn, True), (FLAG_ELASTOSTATIC, False)):
        prob, f = beside_problem(flags)
        o = oracle_lib.Oracle(prob)
        for fid, v in f.items():
            o.set(fid, v)
        o.demag()
        e0 = o.get(FIELD_STRAIN_T0_DIAG)
        o.step(3)
        o.demag()
        assert np.array_equal(o.get(FIELD_STRAIN_T0_DIAG), e0) == kept
        o.close()
    prob, f = beside_problem(FLAG_ELASTOSTATIC | dyn)
    o = oracle_lib.Oracle(prob)
    for fid, v in f.items():
        o.set(fid, v)
    up = 1e-5 * np.random.default_rng(5).standard_normal((3,) + prob.shape)
    o.set(FIELD_STRAIN_T0_DIAG, up)
    o.demag()
    assert np.array_equal(o.get(FIELD_STRAIN_T0_DIAG), up)
    o.close()


def test_start_up_solve_cancels_an_antiferromagnets_eigenstrain(oracle_lib):
    """The same start-up cancellation for the two sublattices' eigenstrain
    (0.5 * 1.5 lambda terms of each, elastic_dynamic_solver.cpp:4-104): with
    u = 0 the first stage-1 velocity increment is zero with the solve, and
    not without it."""
    from goferro import (FIELD_M_AFM1, FIELD_M_AFM2, FLAG_ELASTO_BACKACTION, FLAG_ELASTO_ON_PM,
                         FLAG_ELASTODYNAMICS, workloads)
    dv = {}
    for solve in (True, False):
        prob, m1, m2 = workloads.afm_block(shape=(8, 8, 8))
        mat = prob.materials[0]
        set_elastic(mat, 2.5e11, 1.3e11, 0.8e11, 6800.0)
        mat.lamda100_AFM1, mat.lamda100_AFM2, mat.lamda111_AFM1, mat.lamda111_AFM2 = -2e-5, -1e-5, 1e-5, 3e-5
        prob.flags |= FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM
        if solve:
            prob.flags |= FLAG_ELASTOSTATIC
        o = oracle_lib.Oracle(prob)
        o.set(FIELD_M_AFM1, m1)
        o.set(FIELD_M_AFM2, m2)
        o.demag()
        o.L.gfo_elastic_get(o._h, 1)
        dv[solve] = np.abs(o.get(102)).max()
        o.close()
    assert dv[False] > 0.
    assert dv[True] <= 1e-12 * dv[False]
