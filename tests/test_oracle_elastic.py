"""Pin the oracle's elastodynamics (elastic_dynamic_solver.cpp:4-227,
841-3284; elastodynamics_AbsorbingBoundary.cpp:3-230) by closed forms:
the stress divergence of a plane wave along each axis and component equals
-(C/rho) (2 - 2 cos kd)/d^2 u (C = c11 longitudinal, c44 transverse), a
pulse advances by the exact RK4 propagator of that discrete operator, the
reference's mass damping acts on the base velocity (v_N = v_0 (1 - eta
dt)^N), and the cell strain is the average of face gradients.

The Liao boundary (elastodynamics_AbsorbingBoundary.cpp:92-230) is pinned
in test_oracle_boundary_pins.py: it reproduces displacement histories that
are cubic (weight 1) or linear (any weight) along its characteristic.

The reference applies the z = 0 absorbing boundary whatever the
periodicity (update_Du_Boundary* run in every update_u*), so waves that
are time-integrated are kept away from that plane."""
import numpy as np
import pytest

from goferro import (FIELD_STRAIN_DIAG, FIELD_STRAIN_OFF, FIELD_U, FIELD_V, FLAG_ELASTODYNAMICS, Material,
                     Problem)
from goferro.workloads import set_elastic

C11, C12, C44, RHO = 2.69e11, 1.08e11, 0.76e11, 5170.0


def chain(axis, n, d=2e-9, dt=2e-14, periodic=(True, True, True), c=(C11, C12, C44), **kw):
    shape = [1, 1, 5]          # the z = 0 Liao plane reads k = 1..4
    shape[axis] = n
    mat = set_elastic(Material.make(), c[0], c[1], c[2], RHO, **kw)
    return Problem(shape=tuple(shape), spacing=(d, d, d), materials=[mat], material_cell=np.ones(shape, np.uint32),
                   dt=dt, periodic=periodic, flags=FLAG_ELASTODYNAMICS)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("q", [1, 3, 8])
def test_stress_divergence_of_plane_wave(oracle_lib, axis, comp, q):
    # z is never periodic here: the reference's periodic-z stencil reads
    # outside its arrays unless ny == nz (elastic_dynamic_solver.cpp:1247);
    # along z the free-ended chain's eigenmodes cos(k (z + d/2)) are used
    n, d, dt, eps = 16, 2e-9, 2e-14, 1e-10
    prob = chain(axis, n, d, dt, periodic=(True, True, False))
    x = np.arange(n) * d
    if axis == 2:
        k = np.pi * q / (n * d)
        mode = np.cos(k * (x + d / 2))
    else:
        k = 2 * np.pi * q / (n * d)
        mode = np.cos(k * x)
    u0 = np.zeros((3,) + prob.shape)
    line = [1, 1, 1]
    line[axis] = n
    u0[comp] = np.broadcast_to((eps * mode).reshape(line), prob.shape)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_U, u0)
    o.set(FIELD_V, np.zeros_like(u0))
    o.time_advance()
    o.L.gfo_elastic_get(o._h, 1)
    dv = o.get(oracle_lib.FIELD_DV1)
    C = C11 if comp == axis else C44
    lam = (2 - 2 * np.cos(k * d)) / d ** 2
    want = -dt / RHO * C * lam * u0[comp].reshape(-1)
    np.testing.assert_allclose(dv[comp].reshape(-1), want, rtol=0, atol=1e-12 * np.abs(want).max())
    # other components vanish away from the free z walls (a longitudinal wave
    # along x or y pushes the walls through c12: the Poisson effect)
    for a in range(3):
        if a != comp:
            inner = dv[a] if axis == 2 else dv[a][:, :, 1:-1]
            assert np.abs(inner).max() <= 1e-12 * np.abs(want).max()


def test_pulse_follows_rk4_propagator(oracle_lib):
    n, d, N = 160, 2e-9, 60
    vt = np.sqrt(C44 / RHO)
    dt = 0.4 * d / vt
    prob = chain(2, n, d, dt, periodic=(True, True, False))
    z = np.arange(n) * d
    u0 = np.zeros((3, 1, 1, n))
    u0[0, 0, 0] = 1e-10 * np.exp(-((z - 100 * d) / (6 * d)) ** 2)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_U, u0)
    o.set(FIELD_V, np.zeros_like(u0))
    o.step(N)
    got = o.get(FIELD_U)[0, 0, 0]
    # discrete operator of a free-ended chain (stress-free walls)
    L = np.zeros((n, n))
    for i in range(n):
        if i > 0:
            L[i, i - 1] += 1
            L[i, i] -= 1
        if i < n - 1:
            L[i, i + 1] += 1
            L[i, i] -= 1
    A = np.zeros((2 * n, 2 * n))
    A[:n, n:] = np.eye(n)
    A[n:, :n] = C44 / RHO / d ** 2 * L
    A *= dt
    R = np.eye(2 * n) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    y = np.linalg.matrix_power(R, N) @ np.concatenate([u0[0, 0, 0], np.zeros(n)])
    assert np.abs(y[:20]).max() < 1e-25          # never reached the Liao plane
    np.testing.assert_allclose(got, y[:n], rtol=0, atol=1e-12 * 1e-10)


def test_strain_of_a_linear_shear(oracle_lib):
    # cell strain averages face gradients: u_x = g z gives e_xz = g/2 inside
    n, d, g = 12, 2e-9, 1e-3
    prob = chain(2, n, d, periodic=(True, True, False))
    z = (np.arange(n) + 0.5) * d
    u0 = np.zeros((3, 1, 1, n))
    u0[0, 0, 0] = g * z
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_U, u0)
    off = o.get(FIELD_STRAIN_OFF)[1].reshape(-1)        # e_xz
    np.testing.assert_allclose(off[1:-1], g / 2, rtol=1e-12)
    np.testing.assert_allclose(off[[0, -1]], g / 4, rtol=1e-12)   # stress-free wall faces
    assert np.abs(o.get(FIELD_STRAIN_DIAG)).max() == 0.0


def test_mass_damping_acts_on_base_velocity(oracle_lib):
    eta, dt, N = 3e10, 1e-13, 50
    prob = chain(2, 8, dt=dt, periodic=(True, True, False), mass_damping=eta, c=(0.0, 0.0, 0.0))
    v0 = np.zeros((3,) + prob.shape)
    v0[1] = 2.5
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_U, np.zeros_like(v0))
    o.set(FIELD_V, v0)
    o.step(N)
    v = o.get(FIELD_V)[1].reshape(-1)
    # the same base-state derivative in all four stages makes RK4 forward Euler
    np.testing.assert_allclose(v, 2.5 * (1 - eta * dt) ** N, rtol=1e-12)


def _store_eigenstrain_crt(mat, m_crt):
    """The magnetostrictive eigenstrain of a uniform crystal-frame m
    (elastic_dynamic_solver.cpp:39-65 with p = 0 and no AFM), in the
    reference's evaluation order: xx yy zz | yz xz xy."""
    mx, my, mz = (float(v) for v in m_crt)
    l100, l111 = mat.lamda100, mat.lamda111
    return [1.5 * l100 * (mx * mx - 1.0 / 3.0), 1.5 * l100 * (my * my - 1.0 / 3.0), 1.5 * l100 * (mz * mz - 1.0 / 3.0),
            1.5 * l111 * my * mz, 1.5 * l111 * mx * mz, 1.5 * l111 * my * mx]


@pytest.mark.parametrize("rotate", [False, True])
def test_static_eigenstrain_cancels_a_frozen_magnet(oracle_lib, rotate):
    """A static eigenstrain at t0 equal to the eigenstrain of the magnet's
    m (elastic_dynamic_solver.cpp:39-65 subtracts ex*0t0_crt) leaves no
    stress: with m frozen (gyro = 0) and u = v = 0 the film stays at rest,
    while without it, or with its shear entries permuted, the free top
    surface moves."""
    from goferro import FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF, FIELD_M, FLAG_GAUSSIAN_PULSE, \
        FLAG_ORIENT_ROTATE
    from goferro.workloads import magnetoelastic_film
    prob, _ = magnetoelastic_film(shape=(4, 4, 8), magnet_layers=3)
    prob.flags &= ~FLAG_GAUSSIAN_PULSE
    mag = prob.materials[0]
    mag.gyro = 0.0
    T = np.eye(3)
    if rotate:
        c, s = np.cos(0.7), np.sin(0.7)
        T = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]]) @ np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
        for mat in prob.materials:
            set_elastic(mat, mat.c11, mat.c12, mat.c44, mat.density, Tg2c=T)
        prob.flags |= FLAG_ORIENT_ROTATE
        prob.Tg2c = T
    mdir = np.array([0.36, -0.48, 0.8])
    magnet = np.asarray(prob.material_cell) == 1
    m = np.zeros((3,) + prob.shape)
    m[:, magnet] = mdir[:, None]
    e0 = np.zeros((6,) + prob.shape)
    e0[:, magnet] = np.asarray(_store_eigenstrain_crt(mag, T @ mdir))[:, None]

    def run(t0):
        o = oracle_lib.Oracle(prob)
        o.set(FIELD_M, m)
        if t0 is not None:
            o.set(FIELD_EIGENSTRAIN_T0_DIAG, t0[:3])
            o.set(FIELD_EIGENSTRAIN_T0_OFF, t0[3:])
        o.step(20)
        v = o.get(FIELD_V)
        o.close()
        return np.abs(v).max()

    free = run(None)
    assert free > 1e-6
    assert run(e0) <= 1e-12 * free
    swapped = e0.copy()
    swapped[[3, 5]] = e0[[5, 3]]
    assert run(swapped) > 1e-3 * free


def test_static_strain_adds_to_the_dynamic_one_for_magnets(oracle_lib):
    """magnetic_solver.cpp:89-96: the magnet reads exx = exxt0_crt + Dexx_crt.
    The stage-1 increment is linear in H for a fixed m, and H linear in the
    strain, so k(S + D) - k(D) equals k(S) - k(0), where S is an uploaded
    static strain and D the dynamic strain of a displacement field; and
    with u = 0 the increment equals that of a context without
    elastodynamics given the same static strain."""
    from goferro import (FIELD_M, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTO_BACKACTION,
                         FLAG_ELASTODYNAMICS, FLAG_GAUSSIAN_PULSE)
    from goferro.workloads import magnetoelastic_film
    from oracle.oracle import FIELD_K1
    prob, m = magnetoelastic_film(shape=(6, 6, 8), magnet_layers=3)
    prob.flags &= ~(FLAG_GAUSSIAN_PULSE | FLAG_ELASTO_BACKACTION)
    rng = np.random.default_rng(77)
    S = 1e-4 * rng.standard_normal((6,) + prob.shape)
    u = 1e-11 * rng.standard_normal((3,) + prob.shape)

    def k1(problem, static, disp):
        o = oracle_lib.Oracle(problem)
        o.set(FIELD_M, m)
        if static is not None:
            o.set(FIELD_STRAIN_T0_DIAG, static[:3])
            o.set(FIELD_STRAIN_T0_OFF, static[3:])
        if disp is not None:
            o.set(FIELD_U, disp)
        o.time_advance()
        if problem.flags & FLAG_ELASTODYNAMICS:
            o.L.gfo_elastic_get(o._h, 1)
        o.L.gfo_mag_get_dm(o._h, 1)
        k = o.get(FIELD_K1)
        o.close()
        return k

    zero = np.zeros_like(S)
    a = k1(prob, S, u) - k1(prob, zero, u)
    b = k1(prob, S, None) - k1(prob, zero, None)
    assert np.abs(b).max() > 0
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())
    import dataclasses
    plain = dataclasses.replace(prob, flags=prob.flags & ~FLAG_ELASTODYNAMICS)
    np.testing.assert_array_equal(k1(prob, S, None), k1(plain, S, None))
