"""Pin the oracle's magnetostatics (Newell kernel, magnetic_initialize.cpp:
1176-1557, and its FFT convolution, magnetostatic_solver.cpp:114-219) by
what the mathematics fixes: the demagnetising-tensor trace, Aharoni's closed
form for a prism, the point-dipole far field, and a real-space circular
convolution that shares no code with the FFT path."""
import numpy as np
import pytest

from goferro import FIELD_HSTAT, FIELD_M, FLAG_MAGNETOSTATICS, Material, Problem
from goferro.workloads import GYRO


def aharoni_dz(a, b, c):
    """Demagnetising factor D_z of a 2a x 2b x 2c prism (A. Aharoni,
    J. Appl. Phys. 83, 3432 (1998), eq. 1)."""
    r = np.sqrt(a * a + b * b + c * c)
    ab = np.sqrt(a * a + b * b)
    bc = np.sqrt(b * b + c * c)
    ac = np.sqrt(a * a + c * c)
    pd = ((b * b - c * c) / (2 * b * c) * np.log((r - a) / (r + a))
          + (a * a - c * c) / (2 * a * c) * np.log((r - b) / (r + b))
          + b / (2 * c) * np.log((ab + a) / (ab - a))
          + a / (2 * c) * np.log((ab + b) / (ab - b))
          + c / (2 * a) * np.log((bc - b) / (bc + b))
          + c / (2 * b) * np.log((ac - a) / (ac + a))
          + 2 * np.arctan(a * b / (c * r))
          + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c)
          + (a * a + b * b - 2 * c * c) / (3 * a * b * c) * r
          + c / (a * b) * (ac + bc)
          - (ab ** 3 + bc ** 3 + ac ** 3) / (3 * a * b * c))
    return pd / np.pi


def test_aharoni_reference_itself():
    # the closed form's own invariants: cube 1/3, factors sum to one
    assert abs(aharoni_dz(1, 1, 1) - 1 / 3) < 1e-14
    a, b, c = 1.0, 1.7, 0.4
    s = aharoni_dz(a, b, c) + aharoni_dz(b, c, a) + aharoni_dz(c, a, b)
    assert abs(s - 1) < 1e-13


@pytest.mark.parametrize("cell", [(2e-9, 2e-9, 2e-9), (2e-9, 3e-9, 5e-9), (4e-9, 1e-9, 1.5e-9)])
def test_self_term_matches_prism_factors(oracle_lib, cell):
    dx, dy, dz = cell
    shape = (8, 8, 8)
    k = oracle_lib.newell_real(shape, cell) * np.prod(shape)   # undo the /n
    V = dx * dy * dz
    nxx, nyy, nzz = k[0, 0, 0, 0] / V, k[1, 0, 0, 0] / V, k[2, 0, 0, 0] / V
    assert abs(nxx + nyy + nzz - 1) < 1e-12
    assert abs(nzz - aharoni_dz(dx / 2, dy / 2, dz / 2)) < 1e-12
    assert abs(nxx - aharoni_dz(dy / 2, dz / 2, dx / 2)) < 1e-12
    assert abs(nyy - aharoni_dz(dz / 2, dx / 2, dy / 2)) < 1e-12
    # off-diagonal self terms vanish by symmetry
    for s in range(3, 9):
        assert abs(k[s, 0, 0, 0]) < 1e-12 * V


def _single_cell_problem(m_dir, shape=(16, 16, 16), d=2e-9, Ms=8e5):
    mat = Material.make(if_FM=1, anisotropy_type=0, Ms=Ms, gyro=GYRO, FM_damping=0.0, Aex=1e-11)
    cell = np.zeros(shape, dtype=np.uint32)
    cell[0, 0, 0] = 1
    prob = Problem(shape=shape, spacing=(d, d, d), materials=[mat], material_cell=cell, dt=1e-14,
                   periodic=(True, True, True), flags=FLAG_MAGNETOSTATICS)
    m = np.zeros((3,) + shape)
    m[:, 0, 0, 0] = m_dir
    return prob, m


def test_isolated_cube_self_and_dipole_field(oracle_lib):
    m_dir = np.array([0.3, -0.5, 0.8])
    m_dir /= np.linalg.norm(m_dir)
    Ms, d = 8e5, 2e-9
    prob, m = _single_cell_problem(m_dir, Ms=Ms, d=d)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m)
    o.demag()
    H = o.get(FIELD_HSTAT)
    # uniformly magnetised cube: H = -M/3 inside
    np.testing.assert_allclose(H[:, 0, 0, 0], -Ms * m_dir / 3, rtol=0, atol=1e-10 * Ms)
    # far field of a cube equals the point dipole up to O((a/r)^4)
    V = d ** 3
    for off in [(6, 0, 0), (0, 6, 0), (0, 0, -6), (4, 5, 0), (-3, 4, 5), (5, -5, 5)]:
        r = np.array(off, dtype=float) * d
        rn = np.linalg.norm(r)
        rh = r / rn
        Hd = Ms * V / (4 * np.pi) * (3 * np.dot(m_dir, rh) * rh - m_dir) / rn ** 3
        got = H[:, off[0] % 16, off[1] % 16, off[2] % 16]
        assert np.linalg.norm(got - Hd) < 3e-3 * np.linalg.norm(Hd), off


@pytest.mark.parametrize("shape", [(8, 8, 4), (6, 5, 3), (5, 6, 7)])
def test_fft_convolution_equals_real_space_sum(oracle_lib, shape):
    d = (2e-9, 2.5e-9, 3e-9)
    rng = np.random.default_rng(21)
    mat = Material.make(if_FM=1, anisotropy_type=0, Ms=1.4e5, gyro=GYRO, FM_damping=0.0, Aex=1e-11)
    cell = (rng.random(shape) < 0.7).astype(np.uint32)
    prob = Problem(shape=shape, spacing=d, materials=[mat], material_cell=cell, dt=1e-14,
                   periodic=(True, True, True), flags=FLAG_MAGNETOSTATICS)
    m = rng.standard_normal((3,) + shape)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m)
    o.demag()
    H = o.get(FIELD_HSTAT)
    n = int(np.prod(shape))
    K = oracle_lib.newell_real(shape, d) * n          # placed kernel without the /n
    M = m * 1.4e5 * cell[None]
    V = d[0] * d[1] * d[2]
    # pairing of magnetostatic_solver.cpp:163-177: spectra order
    # Axyz Ayzx Azxy Bxyz Byzx Bzxy Byxz Bzyx Bxzy
    pair = {0: [(0, 0), (1, 3), (2, 8)], 1: [(1, 1), (2, 4), (0, 6)], 2: [(2, 2), (0, 5), (1, 7)]}
    idx = np.indices(shape).reshape(3, -1).T
    Hd = np.zeros((3, n))
    for a in range(3):
        for (b, s) in pair[a]:
            Ks = K[s]
            Mb = M[b].reshape(-1)
            for t, (i, j, k) in enumerate(idx):
                di = (i - idx[:, 0]) % shape[0]
                dj = (j - idx[:, 1]) % shape[1]
                dk = (k - idx[:, 2]) % shape[2]
                Hd[a, t] += np.dot(Ks[di, dj, dk], Mb)
    Hd = -Hd / V
    np.testing.assert_allclose(H.reshape(3, -1), Hd, rtol=0, atol=1e-11 * np.abs(Hd).max())


def test_thin_film_limit(oracle_lib):
    # a film infinite in x, y (periodic) magnetised along z has Hz = -Ms;
    # the kernel truncation at |d| <= n/2-1 leaves an error O(t / L)
    shape = (64, 64, 4)
    mat = Material.make(if_FM=1, anisotropy_type=0, Ms=1.4e5, gyro=GYRO, FM_damping=0.0, Aex=1e-11)
    cell = np.zeros(shape, dtype=np.uint32)
    cell[:, :, :1] = 1
    prob = Problem(shape=shape, spacing=(2e-9,) * 3, materials=[mat], material_cell=cell, dt=1e-14,
                   periodic=(True, True, True), flags=FLAG_MAGNETOSTATICS)
    m = np.zeros((3,) + shape)
    m[2] = 1.0
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m)
    o.demag()
    H = o.get(FIELD_HSTAT)
    hz = H[2, :, :, 0]
    assert np.all(np.abs(hz / -1.4e5 - 1) < 0.04)
    assert np.abs(H[0, :, :, 0]).max() < 1e-6 * 1.4e5


@pytest.mark.parametrize("shape", [(6, 5, 7), (9, 10, 3), (4, 12, 15)])
def test_kernel_placement_on_any_grid(oracle_lib, shape):
    """magnetic_initialize.cpp:1204-1357 places the kernel of offset d at
    index d (d >= 0) or n + d (d < 0) for 1 - n/2 <= d <= n/2 - 1 (integer
    halves) and leaves the other indices zero, for odd and even n alike:
    check the zero pattern and that the placed values are those of the
    offsets, through the evenness of the diagonal kernels."""
    d = (2e-9, 2.5e-9, 3e-9)
    K = oracle_lib.newell_real(shape, d)
    for ax, n in enumerate(shape):
        half = n // 2
        kept = set(range(0, half)) | set(range(n - half + 1, n))
        for i in range(n):
            sl = [slice(None)] * 3
            sl[ax] = i
            plane = K[(slice(None),) + tuple(sl)]
            if i in kept:
                assert np.abs(plane[0]).max() > 0, (ax, i)
            else:
                assert np.abs(plane).max() == 0, (ax, i)
        # the xx kernel is even in each offset: K(d) at index d equals K(-d)
        # at index n - d, up to the evaluation's round-off
        for dd in range(1, half):
            a = np.take(K[0], dd, axis=ax)
            b = np.take(K[0], n - dd, axis=ax)
            np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9 * np.abs(K[0]).max())
