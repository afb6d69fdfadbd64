"""Synthetic workloads of the benchmark and the parity tests.

Parameters are stated in the reference's input units (A/m, J/m^3, gyro =
gamma*mu0), following read.cpp:113-215.  Fields quoted in tesla are divided
by mu0 = 1.256637061e-6.  Initial magnetisations are seeded so every run of
a workload starts from the same state.
"""
from __future__ import annotations

import numpy as np

from . import (FLAG_MAGNETOSTATICS, MU0, Material, Problem)

GAMMA = 1.76e11                 # rad/(s T)
GYRO = GAMMA * MU0              # the reference's gyro is gamma*mu0 (magnetic_solver.cpp:207-208)


def yig_material() -> Material:
    """YIG-like ferromagnet: Ms 1.4e5 A/m, Aex 3.7e-12 J/m, cubic K1 -610 J/m^3, alpha 0.01."""
    return Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.01,
                         K1=-610.0, K2=0.0, Aex=3.7e-12, iDMI=0.0, B1=0.0, B2=0.0)


def tilted_m(shape, seed: int, theta_max: float = 0.05) -> np.ndarray:
    """m along +z tilted by theta = theta_max*U[0,1) at azimuth 2*pi*U[0,1) per cell."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    theta = theta_max * rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    m = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    return m.reshape((3,) + tuple(shape))


def fm_film(shape=(64, 64, 16), dx=2e-9, dt=1e-14, seed=1, demag=True, bias_T=0.1):
    """Plain FP64 ferromagnet LLG + RK4 film: exchange, cubic anisotropy, FFT
    demag, uniform bias along z; periodic in x/y, free surfaces in z.
    Returns (Problem, m0)."""
    cell = np.ones(shape, dtype=np.uint32)
    prob = Problem(shape=tuple(shape), spacing=(dx, dx, dx), materials=[yig_material()], material_cell=cell,
                   dt=dt, periodic=(True, True, False), flags=FLAG_MAGNETOSTATICS if demag else 0,
                   Hext_stat=(0.0, 0.0, bias_T / MU0))
    return prob, tilted_m(shape, seed)


def cijkl_glb(c11: float, c12: float, c44: float, Tg2c=None) -> np.ndarray:
    """6x6 global-frame stiffness of a cubic crystal, Voigt order
    xx yy zz yz xz xy, rotated as global_parameters::get_cijkl does
    (global.cpp:365-441): C_wxyz = T_wi T_xj T_yk T_zl C_ijkl with T = Tc2g."""
    Cc = np.zeros((6, 6))
    Cc[:3, :3] = c12
    for a in range(3):
        Cc[a, a] = c11
        Cc[3 + a, 3 + a] = c44
    voigt = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    C4 = np.zeros((3, 3, 3, 3))
    for I, (w, x) in enumerate(voigt):
        for J, (y, z) in enumerate(voigt):
            for (a, b) in ((w, x), (x, w)):
                for (c, d) in ((y, z), (z, y)):
                    C4[a, b, c, d] = Cc[I, J]
    T = np.eye(3) if Tg2c is None else np.asarray(Tg2c, dtype=float).T
    Cg = np.einsum("wi,xj,yk,zl,ijkl->wxyz", T, T, T, T, C4)
    out = np.zeros((6, 6))
    for I, (w, x) in enumerate(voigt):
        for J, (y, z) in enumerate(voigt):
            out[I, J] = Cg[w, x, y, z]
    return out


def set_elastic(mat: Material, c11: float, c12: float, c44: float, density: float, Tg2c=None,
                mass_damping: float = 0.0, stiff_damping: float = 0.0) -> Material:
    """Fill the elastic record of a material (read.cpp:128-130, global.cpp:365-441)."""
    mat.c11, mat.c12, mat.c44, mat.density = c11, c12, c44, density
    mat.elast_mass_damping, mat.elast_stiff_damping = mass_damping, stiff_damping
    C = cijkl_glb(c11, c12, c44, Tg2c)
    for a in range(6):
        for b in range(6):
            mat.cijkl_glb[a][b] = float(C[a, b])
    return mat


def set_magnetostriction(mat: Material, B1: float, B2: float) -> Material:
    """Magnetostriction constants from the magnetoelastic coefficients,
    inverting read.cpp:141-142 (B1 = -3/2 lambda100 (c11-c12), B2 = -3 lambda111 c44)."""
    mat.lamda100 = -2.0 * B1 / (3.0 * (mat.c11 - mat.c12))
    mat.lamda111 = -B2 / (3.0 * mat.c44)
    mat.B1 = -3.0 / 2.0 * mat.lamda100 * (mat.c11 - mat.c12)
    mat.B2 = -3.0 * mat.lamda111 * mat.c44
    return mat


def magnetoelastic_film(shape=(128, 128, 64), magnet_layers=16, dx=2e-9, dt=5e-15, seed=2):
    """Magnon-phonon film: a YIG-like magnet (alpha 1e-3, B1 = B2 = 3.5e5
    J/m^3) on a non-magnetic substrate (c11 2.69e11, c12 1.08e11, c44
    0.76e11 Pa, rho 5170 kg/m^3) with the same stiffness; elastodynamics
    coupled both ways; fourth-order Liao absorbing bottom; a Gaussian shear
    displacement pulse (amplitude 1e-4, width 1 ps) launched at the top
    layer.  Returns (Problem, m0)."""
    from . import FLAG_ELASTO_BACKACTION, FLAG_ELASTO_ON_PM, FLAG_ELASTODYNAMICS, FLAG_GAUSSIAN_PULSE
    c11, c12, c44, rho = 2.69e11, 1.08e11, 0.76e11, 5170.0
    mag = yig_material()
    mag.FM_damping = 1e-3
    set_elastic(mag, c11, c12, c44, rho)
    set_magnetostriction(mag, 3.5e5, 3.5e5)
    sub = set_elastic(Material.make(if_FM=0), c11, c12, c44, rho)
    cell = np.full(shape, 2, dtype=np.uint32)
    cell[:, :, shape[2] - magnet_layers:] = 1
    prob = Problem(shape=tuple(shape), spacing=(dx, dx, dx), materials=[mag, sub], material_cell=cell, dt=dt,
                   periodic=(True, True, False),
                   flags=FLAG_ELASTODYNAMICS | FLAG_ELASTO_BACKACTION | FLAG_ELASTO_ON_PM | FLAG_GAUSSIAN_PULSE,
                   input_strain_type=1, input_strain_sourcez=shape[2] - 1, input_strain_component="x",
                   sigma_gauss=1e-12, amplitude_gauss=1e-4, weighting_factor_fourth_Liao=0.5)
    m = tilted_m(shape, seed)
    m[:, :, :, : shape[2] - magnet_layers] = 0.0
    return prob, m


def pto_material(T=300.0, FE_mass=1e-16, FE_damping=0.0, G11=1e-10, eps_r=10.0) -> Material:
    """PbTiO3-like ferroelectric with a 6th-order Landau potential:
    a1 = 3.74e5 (T - 752), a11 = -7.3e7, a12 = 7.5e8, a111 = 2.6e8,
    a112 = 6.1e8, a123 = -3.7e9 (SI).  The reference's dynamics is
    second order, m q' + gamma q = -dF/dp (ferroelectric_solver.cpp:740-792),
    so a first-order mobility L maps to gamma = 1/L with a small mass."""
    return Material.make(if_FE=1, FE_mass=FE_mass, FE_damping=FE_damping, a1=3.74e5 * (T - 752.0), a11=-7.3e7,
                         a12=7.5e8, a111=2.6e8, a112=6.1e8, a123=-3.7e9, G11=G11)


def afm_block(shape=(32, 32, 16), dt=1e-15, seed=5000, alpha=1e-3, drive_T=0.05, freq=1e12, Aex=1e-12):
    """Two-sublattice NiO-like antiferromagnet: Ms 3.5e5 A/m per sublattice,
    exchange field 900 T (J_AFM = 900*Ms), easy-axis field 0.6 T
    (K1 = 0.3*Ms) along z, AC drive along x.  Aex_AFM is not given by the
    workload and zero would make the face harmonic mean 0/0, so a finite
    1e-12 J/m is used.  Returns (Problem, m1, m2)."""
    Ms = 3.5e5
    mat = Material.make(if_AFM=1, Ms_AFM1=Ms, Ms_AFM2=Ms, gyro_AFM1=GYRO, gyro_AFM2=GYRO,
                        damping_AFM1=alpha, damping_AFM2=alpha, K1_AFM1=0.3 * Ms, K1_AFM2=0.3 * Ms,
                        uniaxis_AFM=(0.0, 0.0, 1.0), Aex_AFM1=Aex, Aex_AFM2=Aex, J_AFM=900.0 * Ms)
    cell = np.ones(shape, dtype=np.uint32)
    prob = Problem(shape=tuple(shape), spacing=(2e-9, 2e-9, 2e-9), materials=[mat], material_cell=cell, dt=dt,
                   periodic=(True, True, False), flags=0, Hext_altr=(drive_T / MU0, 0.0, 0.0),
                   H_altr_freq=(freq, 0.0, 0.0))
    m1 = tilted_m(shape, seed)
    m2 = -tilted_m(shape, seed + 100000)
    return prob, m1, m2


SWEEP_SIZES = ((16, 16, 16), (32, 32, 16), (32, 32, 32), (64, 64, 16), (64, 64, 32))


def afm_sweep(count=256, seed=5, sizes=SWEEP_SIZES, dt=1e-15):
    """Batched-sweep instances: NiO-like two-sublattice antiferromagnets
    (afm_block) with a grid drawn per instance from `sizes`, damping
    log-uniform in [1e-4, 1e-2], AC drive along x with amplitude U[0.01, 0.1] T
    and frequency U[0.5, 2] THz, all drawn from one seed; instance i starts
    from seeds 5000 + i.  Returns a list of (Problem, m1, m2)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        shape = sizes[int(rng.integers(len(sizes)))]
        alpha = float(10.0 ** rng.uniform(-4.0, -2.0))
        amp = float(rng.uniform(0.01, 0.1))
        freq = float(rng.uniform(0.5e12, 2e12))
        out.append(afm_block(shape, dt=dt, seed=5000 + i, alpha=alpha, drive_T=amp, freq=freq))
    return out


def adjoint_film(shape=(128, 128, 32), seed=4, drive_T=1e-3, drive_freq=3e10, strain=1e-4, dt=1e-14):
    """The adjoint workload: the plain ferromagnet film physics (fm_film) with
    a shaped AC drive along x on the top layer k = nz-1, amplitude
    a(x, y) * drive_T with a ~ U[0.5, 1] (seeded), at drive_freq, and a
    static biaxial strain exx = eyy = `strain` so that B1 acts.  Returns
    (Problem, fields, amp) with fields the uploads (m, drive profile,
    static strain) and amp the (nx, ny) amplitude map."""
    from . import FIELD_HDRIVE, FIELD_M, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTO_ON_PM, FLAG_HDRIVE
    prob, m = fm_film(shape, seed=seed, dt=dt)
    mat = prob.materials[0]
    mat.B1, mat.B2 = 3.5e5, 3.5e5
    prob.flags |= FLAG_HDRIVE | FLAG_ELASTO_ON_PM
    prob.Hdrive_freq = drive_freq
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.5, 1.0, shape[:2])
    hd = np.zeros((3,) + tuple(shape))
    hd[0, :, :, -1] = amp * drive_T / MU0
    diag = np.zeros((3,) + tuple(shape))
    diag[0] = strain
    diag[1] = strain
    fields = {FIELD_M: m, FIELD_HDRIVE: hd, FIELD_STRAIN_T0_DIAG: diag,
              FIELD_STRAIN_T0_OFF: np.zeros((3,) + tuple(shape))}
    return prob, fields, amp


def adjoint_probe(shape=(128, 128, 32), size=(16, 16, 4)):
    """Probe box of the adjoint workload: centred in x and y, the bottom
    `size[2]` layers (away from the driven top layer)."""
    lo = ((shape[0] - size[0]) // 2, (shape[1] - size[1]) // 2, 0)
    hi = (lo[0] + size[0] - 1, lo[1] + size[1] - 1, size[2] - 1)
    return lo, hi


def fe_thz_column(shape=(8, 8, 1024), fe_layers=200, dz=1e-9, dt=1e-18, seed=3, sigma=2e-17):
    """Ferroelectric + electrodynamics THz column: vacuum / PbTiO3-like film
    (fe_layers cells, relative permittivity 10) / vacuum, Mur walls in z,
    periodic x and y, a Gaussian-modulated 1 THz free-current sheet (type 6,
    amplitude 2 E0 / (eta0 dz) for E0 = 1e7 V/m) in the lower vacuum; dP/dt
    drives the field and the field acts back on P.  p starts at
    (0, 0, 0.754)(1 + 0.01 N(0, 1)) in the film, q = 0.
    Returns (Problem, p0, q0)."""
    from . import FLAG_EM_BACKACTION_P, FLAG_EM_FROM_P, FLAG_EMDYNAMIC, FLAG_JF_INPUT
    fe = pto_material(FE_mass=1e-16, FE_damping=4e-4)
    fe.r_permittivity = 10.0
    fe.comp_n1, fe.tao_e_n1, fe.tao_e_n2, fe.tao_e_n3 = 1.0, 1.0, 1.0, 1.0
    nz = shape[2]
    lo = (nz - fe_layers) // 2
    cell = np.zeros(shape, np.uint32)
    cell[:, :, lo:lo + fe_layers] = 1
    src = lo // 2
    prob = Problem(shape=tuple(shape), spacing=(dz, dz, dz), materials=[fe], material_cell=cell, dt=dt,
                   periodic=(True, True, False), flags=FLAG_EMDYNAMIC | FLAG_EM_FROM_P | FLAG_EM_BACKACTION_P |
                   FLAG_JF_INPUT, Jf_input_type=6, Jf_input_component="x", Jf_input_amp=2e7 / (376.73 * dz),
                   Jf_input_freq=1e12, Jf_input_sigma=sigma, Jfin_lo=(0, 0, src),
                   Jfin_hi=(shape[0] - 1, shape[1] - 1, src))
    rng = np.random.default_rng(seed)
    p = np.zeros((3,) + tuple(shape))
    p[2] = 0.754 * (1 + 0.01 * rng.standard_normal(shape))
    p[:, cell == 0] = 0.0
    return prob, p, np.zeros_like(p)
