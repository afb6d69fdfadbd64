"""Pin the oracle's tangent-linear reference of the adjoint sweep's probe
spectral loss by central finite differences of the oracle's own forward
(the reference's stages, magnetic_solver.cpp:3-229, main.cpp:243-483), for
each differentiated parameter: damping, exchange constant, B1, the shaped
drive profile and the initial magnetisation.  Also pins the loss itself
against a direct evaluation from the trajectory."""
import copy

import numpy as np
import pytest

from goferro import (FIELD_HDRIVE, FIELD_M, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTO_ON_PM,
                     FLAG_HDRIVE, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE, MU0, Material, Problem, workloads)
from goferro.workloads import GYRO

LO, HI, FREQ = (1, 2, 0), (3, 4, 1), 3e10
N = 24


def rot():
    c, s = np.cos(0.35), np.sin(0.35)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]]) @ np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def dmi_cells(shape):
    """Vacuum cells that give the magnet x and y surfaces away from the probe."""
    cell = np.ones(shape, np.uint32)
    cell[6, 5:7, :] = 0
    cell[5, 0, 1:] = 0
    return cell


def problem(alpha=0.02, aex=3.7e-12, b1=3.5e5, rotate=True, demag=True, dmi=0.0):
    shape = (8, 8, 4)
    mat = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=alpha, K1=-2e4, K2=3e3,
                        Aex=aex, B1=b1, B2=-2e5, iDMI=dmi)
    flags = FLAG_HDRIVE | FLAG_ELASTO_ON_PM | (FLAG_MAGNETOSTATICS if demag else 0)
    if rotate:
        flags |= FLAG_ORIENT_ROTATE
    # with DMI: free x ends and vacuum cells, so that its surface slope acts
    cell = dmi_cells(shape) if dmi else np.ones(shape, np.uint32)
    return Problem(shape=shape, spacing=(2e-9,) * 3, materials=[mat], material_cell=cell,
                   dt=2e-14, periodic=(not dmi, True, False), flags=flags, Hext_stat=(0, 0, 0.1 / MU0),
                   Hext_altr=(2e3, 0, 0), H_altr_freq=(2e10, 0, 0), Hdrive_freq=FREQ, Tg2c=rot() if rotate else None)


def inputs(shape, dmi=False):
    rng = np.random.default_rng(21)
    m0 = workloads.tilted_m(shape, 21, theta_max=0.3)
    if dmi:
        m0[:, dmi_cells(shape) == 0] = 0.0
    hd = np.zeros((3,) + shape)
    hd[0, :, :, -1] = rng.uniform(0.5, 1.0, shape[:2]) * 4e4
    diag = rng.standard_normal((3,) + shape) * 1e-3
    off = rng.standard_normal((3,) + shape) * 1e-3
    return m0, hd, diag, off


def oracle_for(oracle_lib, prob, m0, hd, diag, off):
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m0)
    o.set(FIELD_HDRIVE, hd)
    o.set(FIELD_STRAIN_T0_DIAG, diag)
    o.set(FIELD_STRAIN_T0_OFF, off)
    return o


def loss_of(oracle_lib, prob, m0, hd, diag, off):
    o = oracle_for(oracle_lib, prob, m0, hd, diag, off)
    L, _ = o.tl_probe_loss(1, LO, HI, FREQ, N)
    o.close()
    return L


def test_loss_matches_direct_evaluation(oracle_lib):
    prob = problem()
    m0, hd, diag, off = inputs(prob.shape)
    L = loss_of(oracle_lib, prob, m0, hd, diag, off)
    o = oracle_for(oracle_lib, prob, m0, hd, diag, off)
    o.demag()
    S = 0j
    for _ in range(N):
        o.step(1)
        m = o.get(FIELD_M)
        X = m[0, LO[0]:HI[0] + 1, LO[1]:HI[1] + 1, LO[2]:HI[2] + 1].mean()
        S += X * prob.dt * np.exp(-2j * np.pi * FREQ * o.time)
    np.testing.assert_allclose(L, abs(S) ** 2, rtol=1e-12)
    assert L > 0


@pytest.mark.parametrize("which", ["alpha", "aex", "b1"])
def test_tangent_linear_matches_central_differences_scalars(oracle_lib, which):
    base = dict(alpha=0.02, aex=3.7e-12, b1=3.5e5)
    prob = problem(**base)
    m0, hd, diag, off = inputs(prob.shape)
    dpar = {"alpha": (1.0, 0, 0), "aex": (0, 1.0, 0), "b1": (0, 0, 1.0)}[which]
    o = oracle_for(oracle_lib, prob, m0, hd, diag, off)
    L0, dL = o.tl_probe_loss(1, LO, HI, FREQ, N, dpar=dpar)
    o.close()
    eps = 1e-4 * base[which]
    up, dn = dict(base), dict(base)
    up[which] += eps
    dn[which] -= eps
    fd = (loss_of(oracle_lib, problem(**up), m0, hd, diag, off)
          - loss_of(oracle_lib, problem(**dn), m0, hd, diag, off)) / (2 * eps)
    assert abs(dL) > 0
    np.testing.assert_allclose(dL, fd, rtol=1e-6)


@pytest.mark.parametrize("what", ["drive", "m0"])
def test_tangent_linear_matches_central_differences_fields(oracle_lib, what):
    prob = problem()
    m0, hd, diag, off = inputs(prob.shape)
    v = np.random.default_rng(5).standard_normal((3,) + prob.shape)
    if what == "drive":
        v[:, :, :, :-1] = 0.0
        v *= 4e4
    o = oracle_for(oracle_lib, prob, m0, hd, diag, off)
    kw = {"dhdrive": v} if what == "drive" else {"dm0": v}
    _, dL = o.tl_probe_loss(1, LO, HI, FREQ, N, **kw)
    o.close()
    eps = 1e-5 if what == "drive" else 1e-6
    if what == "drive":
        fd = (loss_of(oracle_lib, prob, m0, hd + eps * v, diag, off)
              - loss_of(oracle_lib, prob, m0, hd - eps * v, diag, off)) / (2 * eps)
    else:
        fd = (loss_of(oracle_lib, prob, m0 + eps * v, hd, diag, off)
              - loss_of(oracle_lib, prob, m0 - eps * v, hd, diag, off)) / (2 * eps)
    np.testing.assert_allclose(dL, fd, rtol=1e-6)


def test_tangent_linear_without_rotation_or_magnetostatics(oracle_lib):
    prob = problem(rotate=False, demag=False)
    m0, hd, diag, off = inputs(prob.shape)
    o = oracle_for(oracle_lib, prob, m0, hd, diag, off)
    _, dL = o.tl_probe_loss(1, LO, HI, FREQ, N, dpar=(1.0, 0, 0))
    o.close()
    eps = 2e-6
    p_up, p_dn = problem(alpha=0.02 + eps, rotate=False, demag=False), problem(alpha=0.02 - eps, rotate=False,
                                                                              demag=False)
    fd = (loss_of(oracle_lib, p_up, m0, hd, diag, off) - loss_of(oracle_lib, p_dn, m0, hd, diag, off)) / (2 * eps)
    np.testing.assert_allclose(dL, fd, rtol=1e-6)


@pytest.mark.parametrize("which", ["alpha", "aex", "m0"])
def test_tangent_linear_with_dmi_surfaces(oracle_lib, which, tmp_path):
    """Interfacial DMI with its surface slope -D/(2A) at free x ends and
    vacuum cells (magnetic_solver.cpp:1639-1831): the tangent of the stencil
    (linear in m for given A and D, the slope's own derivative in A) against
    central differences, for alpha, Aex and the initial state."""
    base = dict(alpha=0.02, aex=3.7e-12, b1=3.5e5, dmi=1.2e-3)
    prob = problem(**base)
    m0, hd, diag, off = inputs(prob.shape, dmi=True)
    o = oracle_for(oracle_lib, prob, m0, hd, diag, off)
    if which == "m0":
        v = np.random.default_rng(6).standard_normal((3,) + prob.shape)
        v[:, dmi_cells(prob.shape) == 0] = 0.0
        _, dL = o.tl_probe_loss(1, LO, HI, FREQ, N, dm0=v)
        o.close()
        eps = 1e-6
        fd = (loss_of(oracle_lib, prob, m0 + eps * v, hd, diag, off)
              - loss_of(oracle_lib, prob, m0 - eps * v, hd, diag, off)) / (2 * eps)
    else:
        dpar = {"alpha": (1.0, 0, 0), "aex": (0, 1.0, 0)}[which]
        _, dL = o.tl_probe_loss(1, LO, HI, FREQ, N, dpar=dpar)
        o.close()
        eps = 1e-4 * base[which]
        up, dn = dict(base), dict(base)
        up[which] += eps
        dn[which] -= eps
        fd = (loss_of(oracle_lib, problem(**up), m0, hd, diag, off)
              - loss_of(oracle_lib, problem(**dn), m0, hd, diag, off)) / (2 * eps)
    assert abs(dL) > 0
    np.testing.assert_allclose(dL, fd, rtol=1e-6)
    if which == "aex":
        # the slope's own dependence on A matters: a reference that leaves it
        # out (planted mistake 9) fails the same comparison
        path = oracle_lib.build_planted(9, tmp_path)
        o = oracle_lib.Oracle(prob, lib_path=path)
        o.set(FIELD_M, m0)
        o.set(FIELD_HDRIVE, hd)
        o.set(FIELD_STRAIN_T0_DIAG, diag)
        o.set(FIELD_STRAIN_T0_OFF, off)
        _, dL9 = o.tl_probe_loss(1, LO, HI, FREQ, N, dpar=(0, 1.0, 0))
        o.close()
        assert abs(dL9 - fd) > 1e-4 * abs(fd)
