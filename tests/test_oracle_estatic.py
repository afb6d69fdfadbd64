"""Pin the oracle's FFT electrostatics (GF_FLAG_ELECTROSTATICS; the
reference declares get_E_static at ferroelectric_system.h:103 without a
definition, so only its set-up, ferroelectric_initialize.cpp:310-338 and
global.cpp:270-318, and closed forms fix it):

* a polarisation that varies along one axis only gives the exact 1D
  depolarisation field E = -(P_par - <P_par>) / (eps0 er) along that axis
  and no field from the transverse components;
* a uniform polarisation gives no field (k = 0 carries none);
* for a general polarisation the field is curl-free and satisfies Gauss's
  law div(eps0 er E + P) = 0 in Fourier space (every non-Nyquist k);
* er is the largest relative permittivity times scale_elec.
"""
import numpy as np
import pytest

from goferro import FIELD_ESTAT, FIELD_P, FLAG_ELECTROSTATICS, Material, Problem
from goferro.workloads import pto_material

EPS0 = 8.854187817e-12


def problem(shape, spacing=(1e-9, 1.5e-9, 2e-9), er=(10.0, 4.0), scale=1.0, cell=None):
    fe = pto_material()
    fe.r_permittivity = er[0]
    other = Material.make(r_permittivity=er[1])
    if cell is None:
        cell = np.ones(shape, np.uint32)
    return Problem(shape=shape, spacing=spacing, materials=[fe, other], material_cell=cell, dt=1e-15,
                   periodic=(True, True, True), flags=FLAG_ELECTROSTATICS, scale_elec=scale)


def efield(oracle_lib, prob, p):
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_P, p)
    o.demag()
    E = o.get(FIELD_ESTAT)
    o.close()
    return E


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_one_dimensional_profile_gives_the_exact_depolarisation_field(oracle_lib, axis):
    shape = (8, 4, 16)
    prob = problem(shape)
    rng = np.random.default_rng(axis)
    prof = rng.standard_normal(shape[axis])
    trans = rng.standard_normal(shape[axis])
    p = np.zeros((3,) + shape)
    idx = [None, None, None]
    idx[axis] = slice(None)
    p[axis] += prof[tuple(idx)]
    other = (axis + 1) % 3
    p[other] += trans[tuple(idx)]          # transverse component: no field
    E = efield(oracle_lib, prob, p)
    er = 10.0
    want = -(p[axis] - p[axis].mean()) / (EPS0 * er)
    scale = np.abs(want).max()
    np.testing.assert_allclose(E[axis], want, rtol=0, atol=1e-12 * scale)
    for a in range(3):
        if a != axis:
            assert np.abs(E[a]).max() <= 1e-12 * scale


def test_uniform_polarisation_has_no_field_and_er_scales(oracle_lib):
    shape = (4, 4, 8)
    p = np.ones((3,) + shape) * np.array([0.1, -0.2, 0.7])[:, None, None, None]
    E = efield(oracle_lib, problem(shape), p)
    assert np.abs(E).max() <= 1e-12 * 0.7 / (EPS0 * 10.0)
    rng = np.random.default_rng(4)
    p = np.zeros((3,) + shape)
    p[2] = rng.standard_normal(shape[2])[None, None, :]
    E1 = efield(oracle_lib, problem(shape, er=(10.0, 4.0)), p)
    E2 = efield(oracle_lib, problem(shape, er=(4.0, 20.0)), p)      # er = max = 20
    E3 = efield(oracle_lib, problem(shape, er=(10.0, 4.0), scale=3.0), p)
    np.testing.assert_allclose(E2, E1 * 10.0 / 20.0, rtol=1e-12, atol=1e-12 * np.abs(E1).max())
    np.testing.assert_allclose(E3, E1 / 3.0, rtol=1e-12, atol=1e-12 * np.abs(E1).max())


def test_general_field_is_curl_free_and_satisfies_gauss(oracle_lib):
    shape = (8, 8, 16)
    spacing = (1e-9, 1.5e-9, 2e-9)
    prob = problem(shape, spacing=spacing)
    rng = np.random.default_rng(11)
    p = rng.standard_normal((3,) + shape)
    E = efield(oracle_lib, prob, p)
    er = 10.0
    kv = [2 * np.pi * np.fft.fftfreq(n, d) for n, d in zip(shape, spacing)]
    nyq = [np.isclose(np.abs(np.fft.fftfreq(n)), 0.5) for n in shape]
    K = np.meshgrid(*kv, indexing="ij")
    N = np.meshgrid(*nyq, indexing="ij")
    Ek = np.array([np.fft.fftn(E[a]) for a in range(3)])
    Pk = np.array([np.fft.fftn(p[a]) for a in range(3)])
    ok = ~(N[0] | N[1] | N[2])
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    ok &= k2 > 0
    div = sum(K[a] * (EPS0 * er * Ek[a] + Pk[a]) for a in range(3))
    ref = sum(np.abs(K[a] * Pk[a]) for a in range(3))
    assert np.abs(div[ok]).max() <= 1e-12 * ref[ok].max()
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        curl = K[b] * Ek[c] - K[c] * Ek[b]
        assert np.abs(curl[ok]).max() <= 1e-12 * np.abs(K[b] * Ek[c]).max()
    assert np.abs(Ek[:, 0, 0, 0]).max() <= 1e-12 * np.abs(Ek).max()
