"""Pin the oracle's real 3D transforms (stand-ins for cuFFT D2Z/Z2D,
magnetostatic_solver.cpp:149-185) against a naive DFT evaluated in long
double and against numpy.fft, which share no code with it."""
import numpy as np
import pytest


@pytest.mark.parametrize("shape", [(8, 4, 8), (4, 8, 2), (2, 2, 16), (1, 4, 4), (4, 1, 1),
                                   (6, 5, 10), (3, 7, 9), (12, 1, 15), (5, 4, 7), (14, 3, 6)])
def test_rfft3_matches_naive_dft(oracle_lib, shape):
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape)
    got = oracle_lib.rfft3(x)
    ref = oracle_lib.rdft3_naive(x)
    assert got.shape == (shape[0], shape[1], shape[2] // 2 + 1)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", [(16, 8, 32), (32, 32, 4), (8, 16, 1), (20, 12, 10), (48, 9, 25), (7, 49, 6),
                                   (100, 6, 20)])
def test_rfft3_matches_numpy(oracle_lib, shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    got = oracle_lib.rfft3(x)
    ref = np.fft.rfftn(x)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", [(16, 8, 32), (8, 8, 2), (4, 4, 1), (6, 10, 15), (12, 7, 5), (3, 5, 14)])
def test_irfft3_is_unnormalised_inverse(oracle_lib, shape):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(shape)
    n = x.size
    X = oracle_lib.rfft3(x)
    back = oracle_lib.irfft3(X, shape[2])
    # cuFFT semantics: unnormalised, so the round trip scales by n
    assert np.max(np.abs(back / n - x)) <= 1e-13 * np.max(np.abs(x))
    np.testing.assert_allclose(back, np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * n, rtol=0, atol=1e-12 * np.max(np.abs(back)))


def test_parseval(oracle_lib):
    rng = np.random.default_rng(8)
    x = rng.standard_normal((8, 16, 8))
    X = oracle_lib.rfft3(x)
    nz = 8
    w = np.ones(nz // 2 + 1)
    w[1:nz // 2] = 2.0  # bins with a conjugate partner outside the half spectrum
    energy_k = np.sum(w * np.abs(X) ** 2) / x.size
    assert abs(energy_k - np.sum(x * x)) <= 1e-12 * np.sum(x * x)
