"""Pin the oracle's flexoelectric coupling (GF_FLAG_FLEXO; the terms of
get_dq_RK*, ferroelectric_solver.cpp:800-832, with the gradients of
get_p_gradient, :134-281, and get_Dstrain_gradient,
elastic_dynamic_solver.cpp:230-399, and the eigenstrain term :82-89).

The flexoelectric part of the q increment is isolated as the difference of
two oracles that differ only in the flag, divided by -dt/m; the material
has no Landau, gradient or electrostrictive energy of its own beyond
small Q's, so the difference is the flexoelectric force alone, and the
Q's are chosen so that every term is visible in it.  On periodic
blocks with sinusoidal profiles every discrete derivative has a closed form
(central differences halved: sin(kh)/h; second differences:
(2 cos kh - 2)/h^2), and with c12 = Q12 = 0 the derived constants of
read.cpp:181-199 reduce to F11 = f11/c11, F12 = f12/c11, F44 = f44/c44,
tv4 = 2 c11 Q11 F11, tv5 = 2 c11 Q11 F12, tv6 = 2 c44 Q44 F44,
tv7 = c11 (F11^2 + 2 F12^2), tv8 = c44 F44^2.
"""
import numpy as np
import pytest

from goferro import (FIELD_P, FIELD_Q, FIELD_U, FLAG_ELASTO_BACKACTION, FLAG_ELASTODYNAMICS, FLAG_FLEXO, Material,
                     Problem)
from goferro.workloads import set_elastic

C11, C44 = 2.3e11, 1.1e11
Q11, Q44 = 2e-12, 3e-12
F11C, F12C, F44C = 1e-10, -3e-11, 2e-11      # f11, f12, f44
KQ1, E0D, E0O = 106, 107, 108


def block(shape, flags, spacing=(1e-9, 1.5e-9, 2e-9), q=(Q11, Q44)):
    fe = Material.make(if_FE=1, FE_mass=1e-16, FE_damping=0.0)
    fe.Q11, fe.Q12, fe.Q44 = q[0], 0.0, q[1]
    set_elastic(fe, C11, 0.0, C44, 7900.0)
    fe.f11, fe.f12, fe.f44 = F11C, F12C, F44C
    return Problem(shape=shape, spacing=spacing, materials=[fe], material_cell=np.ones(shape, np.uint32), dt=1e-15,
                   periodic=(True, True, True), flags=flags)


def flexo_force(oracle_lib, shape, flags, fields, spacing=(1e-9, 1.5e-9, 2e-9)):
    """(kq with the flag - kq without) / (-dt / m) at stage 1."""
    out = []
    for f in (flags | FLAG_FLEXO, flags):
        prob = block(shape, f, spacing)
        o = oracle_lib.Oracle(prob)
        for fid, v in fields.items():
            o.set(fid, v)
        o.L.gfo_time_advance(o._h)
        o.L.gfo_fe_get(o._h, 1)
        out.append(o.get(KQ1))
        o.close()
    return (out[0] - out[1]) / (-prob.dt / 1e-16), prob


def constants():
    F11, F12, F44 = F11C / C11, F12C / C11, F44C / C44
    return dict(F11=F11, F12=F12, F44=F44, tv4=2 * C11 * Q11 * F11, tv5=2 * C11 * Q11 * F12, tv6=2 * C44 * Q44 * F44,
                tv7=C11 * (F11 ** 2 + 2 * F12 ** 2), tv8=C44 * F44 ** 2)


def test_polarisation_gradient_terms(oracle_lib):
    # p = (A sin kx, B sin kx, 0): dpx/dx, dpy/dx and their second
    # differences are the only non-zero derivatives
    shape = (16, 2, 2)
    h = 1e-9
    kx = 2 * np.pi / (shape[0] * h)
    x = np.arange(shape[0]) * h
    A, B = 0.6, 0.25
    p = np.zeros((3,) + shape)
    p[0] = (A * np.sin(kx * x))[:, None, None]
    p[1] = (B * np.sin(kx * x))[:, None, None]
    F, prob = flexo_force(oracle_lib, shape, 0, {FIELD_P: p})
    c = constants()
    s, co = np.sin(kx * x), np.cos(kx * x)
    g = np.sin(kx * h) / h * co                      # d sin / dx, central difference halved
    d2 = (2 * np.cos(kx * h) - 2) / h ** 2 * s       # second difference
    px, py = A * s, B * s
    want_x = c["tv4"] * px * (A * g) + c["tv5"] * py * (B * g) - c["tv7"] * (A * d2)
    want_y = c["tv6"] * (py * (A * g) + px * (B * g)) - c["tv8"] * (B * d2)
    scale = np.abs(want_x).max()
    # every term carries weight in the sums checked
    for t in (c["tv4"] * px * (A * g), c["tv5"] * py * (B * g), c["tv7"] * (A * d2)):
        assert np.abs(t).max() >= 1e-3 * scale
    for t in (c["tv6"] * (py * (A * g) + px * (B * g)), c["tv8"] * (B * d2)):
        assert np.abs(t).max() >= 1e-3 * np.abs(want_y).max()
    np.testing.assert_allclose(F[0, :, 0, 0], want_x, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(F[1, :, 0, 0], want_y, rtol=0, atol=1e-12 * scale)
    assert np.abs(F[2]).max() <= 1e-12 * scale


@pytest.mark.parametrize("mode", ["normal", "shear"])
def test_strain_gradient_terms(oracle_lib, mode):
    # u_x = U sin(k x): e_xx = u_x', de_xx/dx the second difference;
    # u_x = U sin(k z): e_xz = U cos(kz) sin(kh) / (2h), de_xz/dz its
    # central difference
    shape = (8, 8, 8)
    h = 1e-9
    k = 2 * np.pi / (8 * h)
    r = np.arange(8) * h
    U = 1e-11
    u = np.zeros((3,) + shape)
    if mode == "normal":
        u[0] = (U * np.sin(k * r))[:, None, None]
    else:
        u[0] = (U * np.sin(k * r))[None, None, :]
    F, prob = flexo_force(oracle_lib, shape, FLAG_ELASTODYNAMICS, {FIELD_U: u}, spacing=(h, h, h))
    if mode == "normal":
        dexxdx = U * np.sin(k * r) * (2 * np.cos(k * h) - 2) / h ** 2
        want = -F11C * dexxdx
        got = F[0, :, 0, 0]
    else:
        dexzdz = -U * np.sin(k * r) * np.sin(k * h) ** 2 / (2 * h ** 2)
        want = -2 * F44C * dexzdz
        got = F[0, 0, 0, :]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())
    for comp in (1, 2):
        assert np.abs(F[comp]).max() <= 1e-11 * np.abs(want).max()


def test_eigenstrain_term(oracle_lib):
    # -F grad p: with p = (A sin kx, 0, 0), e0_xx -= F11 dpx/dx and
    # e0_yy, e0_zz -= F12 dpx/dx (elastodynamics needs nz >= 5 and, periodic
    # in z, ny = nz)
    shape = (16, 6, 6)
    h = 1e-9
    kx = 2 * np.pi / (shape[0] * h)
    x = np.arange(shape[0]) * h
    A = 0.4
    p = np.zeros((3,) + shape)
    p[0] = (A * np.sin(kx * x))[:, None, None]
    e0 = []
    for f in (FLAG_FLEXO, 0):
        prob = block(shape, FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | f, q=(0.0, 0.0))
        o = oracle_lib.Oracle(prob)
        o.set(FIELD_P, p)
        o.L.gfo_time_advance(o._h)
        o.L.gfo_elastic_get(o._h, 1)
        e0.append((o.get(E0D), o.get(E0O)))
        o.close()
    c = constants()
    g = A * np.sin(kx * h) / h * np.cos(kx * x)
    d = e0[0][0] - e0[1][0]
    np.testing.assert_allclose(d[0, :, 0, 0], -c["F11"] * g, rtol=1e-12, atol=1e-25)
    np.testing.assert_allclose(d[1, :, 0, 0], -c["F12"] * g, rtol=1e-12, atol=1e-25)
    np.testing.assert_allclose(d[2, :, 0, 0], -c["F12"] * g, rtol=1e-12, atol=1e-25)
    assert np.abs(e0[0][1] - e0[1][1]).max() == 0.0


def test_surface_halves_the_one_sided_gradient(oracle_lib):
    # an FE layer between non-FE layers along z: at the surface cell the
    # neighbour across the face is replaced by the cell itself, and the
    # difference is still halved (ferroelectric_solver.cpp:236-270)
    shape = (2, 2, 8)
    prob = block(shape, FLAG_FLEXO)
    prob.materials = [prob.materials[0], Material.make()]
    cell = np.ones(shape, np.uint32)
    cell[:, :, :2] = 2
    cell[:, :, 6:] = 2
    prob.material_cell = cell
    h = prob.spacing[2]
    z = np.arange(8.0)
    p = np.zeros((3,) + shape)
    p[2] = (0.1 * z)[None, None, :]
    out = []
    for f in (FLAG_FLEXO, 0):
        prob.flags = f
        o = oracle_lib.Oracle(prob)
        o.set(FIELD_P, p)
        o.L.gfo_time_advance(o._h)
        o.L.gfo_fe_get(o._h, 1)
        out.append(o.get(KQ1))
        o.close()
    F = (out[0] - out[1]) / (-prob.dt / 1e-16)
    c = constants()
    # Fz = tv4 pz dpz/dz - tv7 d2pz/dz2; the slope is 0.1 per cell
    g = np.array([0, 0, 0.05, 0.1, 0.1, 0.05, 0, 0]) / h
    d2 = np.array([0, 0, 0.1, 0, 0, -0.1, 0, 0]) / h ** 2
    want = c["tv4"] * p[2, 0, 0] * g - c["tv7"] * d2
    assert np.abs(c["tv7"] * d2).max() >= 1e-3 * np.abs(want).max()
    np.testing.assert_allclose(F[2, 0, 0], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
