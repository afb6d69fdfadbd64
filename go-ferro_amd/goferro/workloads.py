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
