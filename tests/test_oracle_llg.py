"""Pin the oracle's LLG stages (magnetic_solver.cpp:3-229, 1515-1831;
magnetic_AFM_solver.cpp:3-430, 1715-2219) and its RK4 protocol
(main.cpp:243-483) by closed forms: macrospin precession and damping,
fourth-order convergence, drive-time placement, the discrete exchange
spin-wave dispersion, anisotropy fields as energy gradients, and the
antiferromagnetic resonance frequency."""
import numpy as np
import pytest

from goferro import FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE, MU0, Material, Problem
from goferro.workloads import GYRO


def macrospin(Ms=1.4e5, alpha=0.0, anis=0, K1=0.0, K2=0.0, H=(0, 0, 0), dt=1e-13, altr=(0, 0, 0),
              freq=(0, 0, 0), Tg2c=None):
    mat = Material.make(if_FM=1, anisotropy_type=anis, Ms=Ms, gyro=GYRO, FM_damping=alpha, K1=K1, K2=K2,
                        Aex=1e-11)
    flags = 0
    if Tg2c is not None:
        flags |= 1 << 4
    return Problem(shape=(1, 1, 1), spacing=(2e-9,) * 3, materials=[mat], material_cell=np.ones((1, 1, 1), np.uint32),
                   dt=dt, periodic=(True, True, True), flags=flags, Hext_stat=H, Hext_altr=altr, H_altr_freq=freq,
                   Tg2c=Tg2c)


def sph(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


def run_macrospin(oracle_lib, prob, m0, nsteps):
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m0.reshape(3, 1, 1, 1))
    o.step(nsteps)
    return o.get(FIELD_M).reshape(3)


def test_precession_and_damping_closed_form(oracle_lib):
    H0, alpha, dt, N = 8e4, 0.05, 1e-13, 400
    th0, ph0 = 0.7, 0.3
    m = run_macrospin(oracle_lib, macrospin(alpha=alpha, H=(0, 0, H0), dt=dt), sph(th0, ph0), N)
    g = GYRO / (1 + alpha * alpha)
    t = N * dt
    theta = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-alpha * g * H0 * t))
    phi = ph0 + g * H0 * t
    # the reference does not renormalise |m|; with omega*dt ~ 2e-3 the RK4
    # truncation is far below the tolerance
    np.testing.assert_allclose(m, sph(theta, phi), rtol=0, atol=1e-10)


def test_rk4_is_fourth_order(oracle_lib):
    H0, alpha = 1e6, 0.02
    g = GYRO / (1 + alpha * alpha)
    T = 40 * 0.25 / (g * H0)          # 40 steps at omega*dt = 0.25
    errs = []
    for N in (40, 80):
        m = run_macrospin(oracle_lib, macrospin(alpha=alpha, H=(0, 0, H0), dt=T / N), sph(0.9, 0.0), N)
        theta = 2 * np.arctan(np.tan(0.45) * np.exp(-alpha * g * H0 * T))
        errs.append(np.linalg.norm(m - sph(theta, g * H0 * T)))
    ratio = errs[0] / errs[1]
    assert 14.0 < ratio < 18.0, ratio


def test_drive_time_placement(oracle_lib):
    # H_z(t) = H0 + H1 sin(2 pi f t): stages must sample t-dt, t-dt/2, t-dt/2, t
    # (magnetic_solver.cpp:1834-1861); any other placement drops the order
    # of the time-dependent problem below four
    H0, H1, f = 2e5, 1.5e5, 7e9
    errs = []
    for N, dt in ((100, 1e-12), (200, 5e-13)):
        m = run_macrospin(oracle_lib, macrospin(H=(0, 0, H0), altr=(0, 0, H1), freq=(0, 0, f), dt=dt),
                          sph(0.5, 0.0), N)
        T = N * dt
        phase = GYRO * (H0 * T + H1 * (1 - np.cos(2 * np.pi * f * T)) / (2 * np.pi * f))
        got = np.arctan2(m[1], m[0])
        errs.append(abs((got - phase + np.pi) % (2 * np.pi) - np.pi))
    assert errs[1] < 1e-6
    assert 13.0 < errs[0] / errs[1] < 19.0, errs


def num_grad(E, m, h=1e-6):
    g = np.zeros(3)
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        g[a] = (E(m + e) - E(m - e)) / (2 * h)
    return g


@pytest.mark.parametrize("anis", [1, 2])
@pytest.mark.parametrize("rotate", [False, True])
def test_anisotropy_field_is_energy_gradient(oracle_lib, anis, rotate):
    Ms, K1, K2, alpha, dt = 1.4e5, -610.0 * 50, 2.3e3, 0.1, 1e-13
    R = None
    if rotate:
        c, s = np.cos(0.4), np.sin(0.4)
        Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
        c2, s2 = np.cos(0.9), np.sin(0.9)
        Rx = np.array([[1, 0, 0], [0, c2, -s2], [0, s2, c2]])
        R = Rx @ Rz
    prob = macrospin(Ms=Ms, alpha=alpha, anis=anis, K1=K1, K2=K2, dt=dt, Tg2c=R)
    m0 = sph(1.1, 0.7)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m0.reshape(3, 1, 1, 1))
    o.time_advance()
    o.mag_stage(1)
    k1 = 2 * (o.get(FIELD_M_STORE).reshape(3) - m0)
    T = np.eye(3) if R is None else R

    def E(m):
        c = T @ m
        if anis == 1:
            return K1 * (c[0] ** 2 * c[1] ** 2 + c[1] ** 2 * c[2] ** 2 + c[2] ** 2 * c[0] ** 2) \
                + K2 * c[0] ** 2 * c[1] ** 2 * c[2] ** 2
        return -K1 * c[2] ** 2

    H = -num_grad(E, m0) / (MU0 * Ms)
    precess = -GYRO / (1 + alpha ** 2) * dt
    k_ref = precess * np.cross(m0, H) + alpha * precess * np.cross(m0, np.cross(m0, H))
    np.testing.assert_allclose(k1, k_ref, rtol=0, atol=1e-8 * np.abs(k_ref).max())


@pytest.mark.parametrize("q", [1, 3, 8])
def test_exchange_spin_wave_dispersion(oracle_lib, q):
    nx, d, A, Ms, H0 = 32, 2e-9, 3.7e-12, 1.4e5, 8e4
    eps, dt, N = 0.02, 2e-14, 300
    mat = Material.make(if_FM=1, anisotropy_type=0, Ms=Ms, gyro=GYRO, FM_damping=0.0, Aex=A)
    prob = Problem(shape=(nx, 1, 1), spacing=(d, d, d), materials=[mat], material_cell=np.ones((nx, 1, 1), np.uint32),
                   dt=dt, periodic=(True, True, True), Hext_stat=(0, 0, H0))
    k = 2 * np.pi * q / (nx * d)
    x = np.arange(nx) * d
    m0 = np.stack([eps * np.cos(k * x), eps * np.sin(k * x), np.full(nx, np.sqrt(1 - eps ** 2))]).reshape(3, nx, 1, 1)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m0)
    o.step(N)
    m = o.get(FIELD_M).reshape(3, nx)
    lam = (2 - 2 * np.cos(k * d)) / d ** 2
    omega = GYRO * (H0 + 2 * A / (MU0 * Ms) * lam * np.sqrt(1 - eps ** 2))
    # m_z is conserved, so m_x + i m_y obeys z' = i omega z exactly and RK4
    # advances it by R(i omega dt) = 1 + x + x^2/2 + x^3/6 + x^4/24 per step
    x1 = 1j * omega * dt
    R = 1 + x1 + x1 ** 2 / 2 + x1 ** 3 / 6 + x1 ** 4 / 24
    want = N * np.angle(R)
    d = (np.arctan2(m[1], m[0]) - k * x - want + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() < 1e-9 * want
    np.testing.assert_allclose(np.hypot(m[0], m[1]), eps * np.abs(R) ** N, rtol=1e-9)


def test_afm_resonance_frequency(oracle_lib):
    Ms, HE_T, HK_T, dt, N = 3.5e5, 900.0, 0.6, 1e-15, 4000
    mat = Material.make(if_AFM=1, Ms_AFM1=Ms, Ms_AFM2=Ms, gyro_AFM1=GYRO, gyro_AFM2=GYRO, damping_AFM1=0.0,
                        damping_AFM2=0.0, K1_AFM1=HK_T / 2 * Ms, K1_AFM2=HK_T / 2 * Ms, uniaxis_AFM=(0, 0, 1),
                        Aex_AFM1=1e-12, Aex_AFM2=1e-12, J_AFM=HE_T * Ms)
    prob = Problem(shape=(1, 1, 1), spacing=(2e-9,) * 3, materials=[mat], material_cell=np.ones((1, 1, 1), np.uint32),
                   dt=dt, periodic=(True, True, True))
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M_AFM1, sph(0.002, 0.0).reshape(3, 1, 1, 1))
    # l = m1 - m2 tilted, net moment zero: the linear regime of both
    # (degenerate) resonance modes; the amplitude shift ~ tilt^2/4 stays at 1e-6
    o.set(FIELD_M_AFM2, -sph(0.002, 0.0).reshape(3, 1, 1, 1))
    xs = []
    for _ in range(N):
        o.step(1)
        xs.append(o.get(FIELD_M_AFM1).reshape(3)[0])
    xs = np.array(xs)
    t = dt * np.arange(1, N + 1)
    HE, HK = HE_T / MU0, HK_T / MU0
    omega = GYRO * np.sqrt(2 * HE * HK + HK * HK)
    # least-squares fit of a cos + b sin around the predicted frequency
    best = None
    for w in omega * (1 + np.linspace(-2e-3, 2e-3, 4001)):
        Aw = np.stack([np.cos(w * t), np.sin(w * t)], 1)
        coef, res, *_ = np.linalg.lstsq(Aw, xs, rcond=None)
        r = np.sum((Aw @ coef - xs) ** 2)
        if best is None or r < best[0]:
            best = (r, w)
    assert abs(best[1] / omega - 1) < 5e-6, best[1] / omega - 1


def test_shaped_drive_reduces_to_uniform_drive(oracle_lib):
    """A shaped drive whose profile is the same vector in every cell is the
    uniform AC drive (magnetic_solver.cpp:1834-1861): same stage times, same
    sums, so the trajectories agree bit for bit."""
    from goferro import FIELD_HDRIVE, FLAG_HDRIVE
    shape = (4, 3, 2)
    h, f = np.array([3e4, -1e4, 2e4]), 9e9
    base = dict(shape=shape, spacing=(2e-9,) * 3, material_cell=np.ones(shape, np.uint32), dt=2e-14,
                periodic=(True, True, False), Hext_stat=(0, 0, 5e4))
    mat = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.02, K1=-2e4, Aex=3.7e-12)
    m0 = np.random.default_rng(4).standard_normal((3,) + shape)
    m0 /= np.linalg.norm(m0, axis=0)
    a = oracle_lib.Oracle(Problem(materials=[mat], Hext_altr=tuple(h), H_altr_freq=(f, f, f), **base))
    b = oracle_lib.Oracle(Problem(materials=[mat], flags=FLAG_HDRIVE, Hdrive_freq=f, **base))
    b.set(FIELD_HDRIVE, np.broadcast_to(h.reshape(3, 1, 1, 1), (3,) + shape).copy())
    for o in (a, b):
        o.set(FIELD_M, m0)
        o.step(7)
    assert np.array_equal(a.get(FIELD_M), b.get(FIELD_M))
    assert np.abs(a.get(FIELD_M) - m0).max() > 1e-4   # the drive did act


def test_shaped_drive_is_local_and_linear(oracle_lib):
    """A profile in one cell only: the stage increment of that cell is the
    drive's precession term at the stage time, other
    cells are untouched after one stage (uniform m: no exchange field)."""
    from goferro import FIELD_HDRIVE, FLAG_HDRIVE
    shape = (3, 3, 1)
    mat = Material.make(if_FM=1, anisotropy_type=0, Ms=1.4e5, gyro=GYRO, FM_damping=0.0, Aex=1e-11)
    prob = Problem(shape=shape, spacing=(2e-9,) * 3, materials=[mat], material_cell=np.ones(shape, np.uint32),
                   dt=1e-13, periodic=(False, False, False), flags=FLAG_HDRIVE, Hdrive_freq=2e10)
    m0 = np.zeros((3,) + shape)
    m0[0] = 1.0
    prof = np.zeros((3,) + shape)
    prof[2, 1, 1, 0] = 1e5
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m0)
    o.set(FIELD_HDRIVE, prof)
    o.time_advance()
    o.L.gfo_mag_get_dm(o._h, 1)
    o.L.gfo_mag_update(o._h, 1)
    ms = o.get(FIELD_M_STORE)
    # stage 1 samples t - dt = 0: sin(0) = 0, nothing moves yet
    np.testing.assert_array_equal(ms, m0)
    o.L.gfo_mag_get_dm(o._h, 2)
    o.L.gfo_mag_update(o._h, 2)
    ms = o.get(FIELD_M_STORE)
    # stage 2 samples t - dt/2: k2 = -gyro dt (m x H), m = x, H = h z  ->  k_y = +gyro dt h
    s = np.sin(2 * np.pi * 2e10 * 0.5e-13)
    ky = GYRO * 1e-13 * 1e5 * s
    np.testing.assert_allclose(ms[1, 1, 1, 0], 0.5 * ky, rtol=1e-14)
    ms[1, 1, 1, 0] = 0.0
    np.testing.assert_array_equal(ms, m0)


def test_static_strain_equals_equivalent_anisotropy(oracle_lib):
    """Static strain ezz with B1 acts as a uniaxial anisotropy along z with
    K1 = -B1 ezz (magnetic_solver.cpp:124-150): H_z = -2 B1 ezz mz / (mu0 Ms)
    against 2 K1 mz / (mu0 Ms).  Equal up to rounding of the two forms."""
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTO_ON_PM
    shape = (4, 4, 2)
    B1, ezz = 3.5e5, 2e-3
    base = dict(shape=shape, spacing=(2e-9,) * 3, material_cell=np.ones(shape, np.uint32), dt=1e-14,
                periodic=(True, True, False), Hext_stat=(0, 0, 1e4))
    m_me = Material.make(if_FM=1, anisotropy_type=0, Ms=1.4e5, gyro=GYRO, FM_damping=0.01, Aex=3.7e-12, B1=B1, B2=7e5)
    m_an = Material.make(if_FM=1, anisotropy_type=2, Ms=1.4e5, gyro=GYRO, FM_damping=0.01, Aex=3.7e-12,
                         K1=-B1 * ezz)
    m0 = np.random.default_rng(5).standard_normal((3,) + shape)
    m0 /= np.linalg.norm(m0, axis=0)
    a = oracle_lib.Oracle(Problem(materials=[m_me], flags=FLAG_ELASTO_ON_PM, **base))
    diag = np.zeros((3,) + shape)
    diag[2] = ezz
    a.set(FIELD_STRAIN_T0_DIAG, diag)
    a.set(FIELD_STRAIN_T0_OFF, np.zeros((3,) + shape))
    b = oracle_lib.Oracle(Problem(materials=[m_an], **base))
    for o in (a, b):
        o.set(FIELD_M, m0)
        o.step(20)
    d = np.abs(a.get(FIELD_M) - b.get(FIELD_M)).max()
    assert d <= 1e-13, d
    assert np.abs(a.get(FIELD_M) - m0).max() > 1e-3
