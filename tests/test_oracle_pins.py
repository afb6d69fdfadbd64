"""Further pins of the oracle against textbook forms.

* The dynamic eigenstrain (elastic_dynamic_solver.cpp:4-104) of random
  magnetisations and polarisations in a rotated crystal equals the
  magnetostrictive (3/2 lambda100 (a_i^2 - 1/3), 3/2 lambda111 a_i a_j;
  half weight per AFM sublattice) plus electrostrictive (Q11 p_i^2 +
  Q12 (p_j^2 + p_k^2), Q44 p_i p_j) tensor, built in the crystal frame
  and rotated back with the tensor law (numpy here).
* Ampere's law with the four-cell means of permittivity and conductivity
  at material interfaces (EMdynamic_solver.cpp:392-499): from a uniform
  E_x and H = 0, one tiny step of a layered conductor changes every x
  edge by -dt <sigma> E / (eps0 <er>), the means over the four cells
  around the edge.
"""
import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_P, FLAG_ELASTO_BACKACTION,
                     FLAG_ELASTODYNAMICS, FLAG_EMDYNAMIC, FLAG_ORIENT_ROTATE, Material, Problem, split_yee,
                     workloads)
from goferro.workloads import GYRO, set_elastic

EPS0 = 8.854187817e-12
E0D, E0O = 107, 108


def rot():
    c, s = np.cos(0.5), np.sin(0.5)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    c2, s2 = np.cos(0.8), np.sin(0.8)
    Rx = np.array([[1, 0, 0], [0, c2, -s2], [0, s2, c2]])
    return Rx @ Rz


def tensor(v):
    return np.array([[v[0], v[5], v[4]], [v[5], v[1], v[3]], [v[4], v[3], v[2]]])


@pytest.mark.parametrize("kind", ["fm", "afm", "fe"])
def test_eigenstrain_is_the_textbook_tensor(oracle_lib, kind):
    shape = (6, 6, 6)
    T = rot()
    if kind == "fm":
        mat = Material.make(if_FM=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.1, Aex=3.7e-12, lamda100=-2e-5, lamda111=3e-5)
    elif kind == "afm":
        mat = Material.make(if_AFM=1, Ms_AFM1=3e5, Ms_AFM2=3e5, gyro_AFM1=GYRO, gyro_AFM2=GYRO, damping_AFM1=0.01,
                            damping_AFM2=0.01, uniaxis_AFM=(0, 0, 1), lamda100_AFM1=-2e-5, lamda100_AFM2=1e-5,
                            lamda111_AFM1=3e-5, lamda111_AFM2=-4e-5)
    else:
        mat = Material.make(if_FE=1, FE_mass=1e-16, Q11=0.089, Q12=-0.026, Q44=0.0675)
    set_elastic(mat, 2.6e11, 1.1e11, 0.8e11, 6000.0, Tg2c=T)
    prob = Problem(shape=shape, spacing=(1e-9,) * 3, materials=[mat], material_cell=np.ones(shape, np.uint32),
                   dt=1e-15, periodic=(True, True, True), Tg2c=T,
                   flags=FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ORIENT_ROTATE)
    o = oracle_lib.Oracle(prob)
    rng = np.random.default_rng(9)
    if kind == "fm":
        a = workloads.tilted_m(shape, 3, theta_max=2.5)
        o.set(FIELD_M, a)
        parts = [(a, 1.0, mat.lamda100, mat.lamda111)]
    elif kind == "afm":
        a1 = workloads.tilted_m(shape, 4, theta_max=2.5)
        a2 = workloads.tilted_m(shape, 5, theta_max=2.5)
        o.set(FIELD_M_AFM1, a1)
        o.set(FIELD_M_AFM2, a2)
        parts = [(a1, 0.5, mat.lamda100_AFM1, mat.lamda111_AFM1), (a2, 0.5, mat.lamda100_AFM2, mat.lamda111_AFM2)]
    else:
        p = 0.4 * rng.standard_normal((3,) + shape)
        o.set(FIELD_P, p)
    o.L.gfo_time_advance(o._h)
    o.L.gfo_elastic_get(o._h, 1)
    got = np.concatenate([o.get(E0D), o.get(E0O)])
    o.close()
    want = np.zeros((6,) + shape)
    for idx in np.ndindex(*shape):
        ec = np.zeros((3, 3))
        if kind == "fe":
            pc = T @ p[(slice(None),) + idx]
            for i in range(3):
                j, k = (i + 1) % 3, (i + 2) % 3
                ec[i, i] = mat.Q11 * pc[i] ** 2 + mat.Q12 * (pc[j] ** 2 + pc[k] ** 2)
            for i, j in ((1, 2), (0, 2), (0, 1)):
                ec[i, j] = ec[j, i] = mat.Q44 * pc[i] * pc[j]
        else:
            for vec, w, l100, l111 in parts:
                ac = T @ vec[(slice(None),) + idx]
                for i in range(3):
                    ec[i, i] += w * 1.5 * l100 * (ac[i] ** 2 - 1. / 3.)
                for i, j in ((1, 2), (0, 2), (0, 1)):
                    ec[i, j] += w * 1.5 * l111 * ac[i] * ac[j]
                    ec[j, i] = ec[i, j]
        eg = T.T @ ec @ T                                 # crystal -> global
        want[(slice(None),) + idx] = [eg[0, 0], eg[1, 1], eg[2, 2], eg[1, 2], eg[0, 2], eg[0, 1]]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_ampere_interface_means(oracle_lib):
    shape = (2, 4, 8)
    cond = Material.make(r_permittivity=3.0, conductivity=2e5, comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0, tao_e_n3=1.0)
    diel = Material.make(r_permittivity=7.0, conductivity=5e4, comp_n1=1.0, tao_e_n1=1.0, tao_e_n2=1.0, tao_e_n3=1.0)
    cell = np.zeros(shape, np.uint32)
    cell[:, :, 0:3] = 1
    cell[:, :, 3:5] = 2
    cell[:, 1, :] = np.where(cell[:, 1, :] == 1, 2, cell[:, 1, :])     # an interface along y too
    prob = Problem(shape=shape, spacing=(1e-9, 1.2e-9, 0.8e-9), materials=[cond, diel], material_cell=cell,
                   dt=1e-21, periodic=(True, True, True), flags=FLAG_EMDYNAMIC)
    E0 = 3e5
    nx, ny, nz = shape
    Ex = np.full((nx, ny + 1, nz + 1), E0)
    Ey = np.zeros((nx + 1, ny, nz + 1))
    Ez = np.zeros((nx + 1, ny + 1, nz))
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_E_EM, np.concatenate([Ex.ravel(), Ey.ravel(), Ez.ravel()]))
    o.step(1)
    got = split_yee(o.get(FIELD_E_EM), shape, FIELD_E_EM)[0]
    o.close()
    er = {0: 1.0, 1: 3.0, 2: 7.0}
    sg = {0: 0.0, 1: 2e5, 2: 5e4}
    want = np.empty((ny + 1, nz + 1))
    rates = np.empty((ny + 1, nz + 1))
    for j in range(ny + 1):
        for k in range(nz + 1):
            y1, y3, z1, z2 = (j - 1) % ny, j % ny, (k - 1) % nz, k % nz
            cells = [cell[0, y1, z1], cell[0, y1, z2], cell[0, y3, z1], cell[0, y3, z2]]
            rates[j, k] = np.mean([sg[int(c)] for c in cells]) / (EPS0 * np.mean([er[int(c)] for c in cells]))
            want[j, k] = E0 * np.exp(-rates[j, k] * prob.dt)
    assert len(np.unique(np.round(rates / rates.max(), 9))) >= 5       # many distinct means checked
    assert np.abs(got[0] - want).max() <= 1e-6 * E0 * rates.max() * prob.dt


def test_interface_exchange_is_the_harmonic_mean(oracle_lib):
    """magnetic_initialize.cpp:381-524: across a face between two
    ferromagnets the exchange stiffness is 2 A1 A2 / (A1 + A2); the
    exchange field of cell i is 2 / (mu0 Ms_i) sum_faces A_f (m_nb - m_i) /
    h^2, and the stage-1 increment is the LLG step
    -gyro dt / (1 + a^2) [m x H + a m x (m x H)] (textbook forms)."""
    MU0 = 1.256637061e-6
    shape = (8, 1, 1)
    A = {1: 3e-12, 2: 1.2e-11}
    MS = {1: 1.4e5, 2: 8e5}
    AL = {1: 0.02, 2: 0.3}
    mats = [Material.make(if_FM=1, anisotropy_type=1, K1=0.0, Ms=MS[t], gyro=GYRO, FM_damping=AL[t], Aex=A[t])
            for t in (1, 2)]
    cell = np.array([1, 1, 1, 1, 2, 2, 2, 2], np.uint32).reshape(shape)
    h = 2e-9
    prob = Problem(shape=shape, spacing=(h, 2e-9, 2e-9), materials=mats, material_cell=cell, dt=1e-14,
                   periodic=(True, True, True))
    m = workloads.tilted_m(shape, 6, theta_max=1.5)
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_M, m)
    o.L.gfo_time_advance(o._h)
    o.L.gfo_mag_get_dm(o._h, 1)
    k1 = o.get(103)
    o.close()
    want = np.zeros_like(m)
    for i in range(8):
        ti = int(cell[i, 0, 0])
        H = np.zeros(3)
        for nb in ((i - 1) % 8, (i + 1) % 8):
            tn = int(cell[nb, 0, 0])
            Af = 2 * A[ti] * A[tn] / (A[ti] + A[tn])
            H += Af * (m[:, nb, 0, 0] - m[:, i, 0, 0]) / h ** 2
        H *= 2 / (MU0 * MS[ti])
        mi = m[:, i, 0, 0]
        pre = -GYRO / (1 + AL[ti] ** 2) * prob.dt
        mxh = np.cross(mi, H)
        want[:, i, 0, 0] = pre * mxh + AL[ti] * pre * np.cross(mi, mxh)
    np.testing.assert_allclose(k1, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_gaussian_displacement_source_emits_its_amplitude(oracle_lib):
    """The type-1 source (elastic_dynamic_solver.cpp:3211-3232) holds the
    source plane at u_x = A sqrt(e) sigma v_t exp(-(t - 5 sigma)^2 / (2
    sigma^2)), v_t = sqrt(c44 / rho): the plane follows it after every
    step, and since its slowness-scaled rate peaks at A, the shear wave it
    launches into the column carries du_x/dz of peak A (2 e_xz = A),
    the textbook transmitted-strain amplitude."""
    from goferro import FIELD_STRAIN_OFF, FIELD_U, FLAG_GAUSSIAN_PULSE
    c11, c12, c44, rho = 2.69e11, 1.08e11, 0.76e11, 5170.0
    mat = set_elastic(Material.make(), c11, c12, c44, rho)
    nz = 256
    shape = (1, 1, nz)
    src = nz - 1
    sigma, amp = 5e-13, 1e-4
    prob = Problem(shape=shape, spacing=(1e-9, 1e-9, 0.25e-9), materials=[mat], material_cell=np.ones(shape, np.uint32),
                   dt=1e-14, periodic=(True, True, False), flags=FLAG_ELASTODYNAMICS | FLAG_GAUSSIAN_PULSE,
                   input_strain_type=1, input_strain_sourcez=src, input_strain_component="x", sigma_gauss=sigma,
                   amplitude_gauss=amp, weighting_factor_fourth_Liao=0.5)
    vt = np.sqrt(c44 / rho)
    o = oracle_lib.Oracle(prob)
    peak = 0.0
    for step in range(1, 1301):
        o.step(1)
        t = step * prob.dt
        if step % 100 == 0 and t < 12.5 * sigma:
            ux = o.get(FIELD_U)[0, 0, 0, src]
            want = amp * 1.6487210 * sigma * vt * np.exp(-(t - 5 * sigma) ** 2 / (2 * sigma ** 2))
            assert ux == pytest.approx(want, rel=1e-12, abs=1e-30), (step, ux, want)
        if step in (1000, 1100, 1200, 1300):
            exz = o.get(FIELD_STRAIN_OFF)[1, 0, 0, : src - 40]          # the pulse away from the source
            peak = max(peak, 2 * np.abs(exz).max())
    o.close()
    assert peak == pytest.approx(amp, rel=0.03)


def shear_column(src_type, sigma, amp, nz=256, cycles=0):
    from goferro import FLAG_GAUSSIAN_PULSE
    c11, c12, c44, rho = 2.69e11, 1.08e11, 0.76e11, 5170.0
    mat = set_elastic(Material.make(), c11, c12, c44, rho)
    shape = (1, 1, nz)
    prob = Problem(shape=shape, spacing=(1e-9, 1e-9, 0.25e-9), materials=[mat], material_cell=np.ones(shape, np.uint32),
                   dt=1e-14, periodic=(True, True, False), flags=FLAG_ELASTODYNAMICS | FLAG_GAUSSIAN_PULSE,
                   input_strain_type=src_type, input_strain_sourcez=nz - 1, input_strain_component="x",
                   sigma_gauss=sigma, amplitude_gauss=amp, num_strain_cycles=cycles, weighting_factor_fourth_Liao=0.5)
    return prob, c44, np.sqrt(c44 / rho)


def test_sinusoidal_displacement_source_emits_its_amplitude(oracle_lib):
    """The type-2 source (elastic_dynamic_solver.cpp:3211-3232; sigma_gauss
    is its frequency f) holds the source plane at u_x = A v_t / (2 pi f)
    (cos(2 pi f t) - 1) for num_cycles periods and at 0 after them: the
    plane follows that after every step, and since its slowness-scaled rate
    -A sin(2 pi f t) peaks at A, the one-period shear wave it launches
    carries du_x/dz of peak A, the textbook transmitted-strain amplitude."""
    from goferro import FIELD_STRAIN_OFF, FIELD_U
    f, amp = 1e11, 1e-4
    prob, _, vt = shear_column(2, f, amp, cycles=1)
    src = prob.shape[2] - 1
    o = oracle_lib.Oracle(prob)
    peak = 0.0
    for step in range(1, 1301):
        o.step(1)
        t = step * prob.dt
        if step % 50 == 0 and step != 1000:
            ux = o.get(FIELD_U)[0, 0, 0, src]
            want = amp * vt / (2 * np.pi * f) * (np.cos(2 * np.pi * f * t) - 1) if t < 1 / f else 0.0
            assert ux == pytest.approx(want, rel=1e-12, abs=1e-12 * amp * vt / f), (step, ux, want)
        if step in (1100, 1200, 1300):
            exz = o.get(FIELD_STRAIN_OFF)[1, 0, 0, : src - 40]
            peak = max(peak, 2 * np.abs(exz).max())
    o.close()
    assert peak == pytest.approx(amp, rel=0.01)


def test_gaussian_traction_source_emits_its_stress(oracle_lib):
    """The type-3 source (elastic_dynamic_solver.cpp:1249-1262) replaces the
    stress on the stress-free top face of the source plane by a traction
    sigma_xz = A exp(-(t - 5 sigma)^2 / (2 sigma^2)).  A traction T(t) on a
    free face launches the shear wave whose stress is T, so the pulse in
    the column carries du_x/dz = sigma_xz / c44 of peak A / c44 (textbook
    1D wave equation); with no traction the column stays at rest."""
    from goferro import FIELD_STRAIN_OFF, FIELD_U
    sigma = 5e-13
    prob, c44, _ = shear_column(3, sigma, 1e-4 * 0.76e11)
    src = prob.shape[2] - 1
    o = oracle_lib.Oracle(prob)
    peak = 0.0
    for step in range(1, 1301):
        o.step(1)
        if step in (1000, 1100, 1200, 1300):
            exz = o.get(FIELD_STRAIN_OFF)[1, 0, 0, : src - 40]
            peak = max(peak, 2 * np.abs(exz).max())
    o.close()
    assert peak == pytest.approx(1e-4, rel=0.01)
    quiet, _, _ = shear_column(3, sigma, 0.0)
    o = oracle_lib.Oracle(quiet)
    o.step(200)
    assert np.abs(o.get(FIELD_U)).max() == 0.0
    o.close()


@pytest.mark.parametrize("eta,beta", [(2.5e-13, 0.0), (0.0, 3e10)])
def test_staged_damping_is_kelvin_voigt(oracle_lib, eta, beta):
    """The stiffness damping (elastic_dynamic_solver.cpp:1270-1290) adds
    (eta / rho)(f - F_store) to the velocity increment of RK stages 1 and 4
    and twice that in stages 2 and 3, F_store being the stage-1 force (of
    the previous step in stage 1); each stage then approximates
    (eta / rho) dt df/dt, the Kelvin-Voigt stress C (e + eta de/dt).  With
    the mass damping -beta v, a standing shear mode u_y = a(t) cos(k x)
    obeys a'' + (beta + eta w^2) a' + w^2 a = 0, w^2 = c44 k^2 / rho, whose
    velocity extrema shrink by exp(-g T_d / 2) per half period with
    g = (beta + eta w^2) / 2 (textbook damped oscillator).  Starting from
    u = 0 keeps the zero F_store of the first step exact.  The mode is
    uniform in z inside a block with stress-free top and bottom faces
    (which it satisfies), under a vacuum plane at z = 0: the Liao plane is
    written there on every step whatever the periodicity
    (elastic_dynamic_solver.cpp:2916-3029), from a displacement history
    that starts at zero."""
    from goferro import FIELD_V
    c44, rho = 0.76e11, 5170.0
    mat = set_elastic(Material.make(), 2.69e11, 1.08e11, c44, rho, mass_damping=beta, stiff_damping=eta)
    shape = (64, 6, 6)
    h = 1e-9
    cell = np.ones(shape, np.uint32)
    cell[:, :, 0] = 0
    prob = Problem(shape=shape, spacing=(h, h, h), materials=[mat], material_cell=cell,
                   dt=2e-14, periodic=(True, True, False), flags=FLAG_ELASTODYNAMICS)
    k = 2 * np.pi / (shape[0] * h)
    w2 = c44 * k * k / rho
    mode = np.cos(k * h * np.arange(shape[0]))
    v0 = np.zeros((3,) + shape)
    v0[1] = mode[:, None, None] * cell
    o = oracle_lib.Oracle(prob)
    o.set(FIELD_V, v0)
    b = [1.0]
    for _ in range(2100):
        o.step(1)
        vy = o.get(FIELD_V)[1]
        assert np.abs(vy[:, :, 1:] - vy[:, :1, 3:4]).max() <= 1e-12
        b.append(float(vy[:, 0, 3] @ mode / (mode @ mode)))
    o.close()
    b = np.abs(np.array(b))
    peaks = [i for i in range(1, len(b) - 1) if b[i] >= b[i - 1] and b[i] > b[i + 1]]
    peaks = [0] + peaks
    g = -np.log(b[peaks[-1]] / b[peaks[0]]) / ((peaks[-1] - peaks[0]) * prob.dt)
    assert len(peaks) >= 4
    assert g == pytest.approx((beta + eta * w2) / 2, rel=0.02)
