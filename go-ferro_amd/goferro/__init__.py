"""goferro — Python binding of the gfx950 coupled-RK4 library (libgoferro.so).

The binding is a thin ctypes layer over the C ABI declared in
``include/goferro.h``; no torch types cross it.  ``Problem`` describes one
simulation the way the reference's input files do (grid, material table,
per-cell material ids, drive, flags); ``Context`` owns the device state.

The library is built in-tree by ``goferro._build.build()``.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import dataclasses
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "Material", "Problem", "Context", "GoFerroError", "load_library", "LIB_PATH",
    "MU0", "FLAG_MAGNETOSTATICS", "FLAG_MAG_1D", "FLAG_DEMAG_FACTOR", "FLAG_PRESCRIBE_HEXT",
    "FLAG_ORIENT_ROTATE", "FLAG_HEXT_NONUNIF", "FLAG_EMDYNAMIC", "FLAG_PRESCRIBE_EEXT",
    "FIELD_M", "FIELD_M_STORE", "FIELD_HSTAT", "FIELD_HEXT_NONUNIF", "FIELD_M_AFM1", "FIELD_M_AFM2",
    "FIELD_P", "FIELD_Q", "FIELD_P_STORE", "FIELD_ESTAT",
    "FIELD_U", "FIELD_V", "FIELD_U_STORE", "FIELD_STRAIN_DIAG", "FIELD_STRAIN_OFF",
    "FLAG_ELASTODYNAMICS", "FLAG_ELASTO_BACKACTION", "FLAG_ELASTO_ON_PM", "FLAG_GAUSSIAN_PULSE", "FLAG_1D_ABC",
    "FLAG_1D_ABC_ELAST_ONLYTOP",
    "FLAG_EM_FROM_P", "FLAG_EM_FROM_M", "FLAG_EM_BACKACTION_P", "FLAG_EM_BACKACTION_M", "FLAG_PLANE_EM",
    "FLAG_JF_INPUT", "FLAG_PEC_XY", "FLAG_PEC_Z", "FLAG_1D_ABC_EM_ONLYTOP",
    "FIELD_E_EM", "FIELD_H_EM", "FIELD_E_EM_STORE", "FIELD_H_EM_STORE", "FIELD_E_CELL", "FIELD_H_CELL", "FIELD_JP",
    "FLAG_HDRIVE", "FLAG_ELECTROSTATICS", "FLAG_SPIN_PUMPING", "FLAG_FLEXO", "FLAG_ELASTOSTATIC", "FIELD_HDRIVE", "FIELD_STRAIN_T0_DIAG",
    "FIELD_STRAIN_T0_OFF", "FIELD_J_ISHE", "FIELD_ELASTOFORCE", "FIELD_EIGENSTRAIN_T0_DIAG",
    "FIELD_EIGENSTRAIN_T0_OFF",
    "field_size", "yee_shapes", "split_yee", "join_yee",
    "KERNEL_CLASSES", "EXPORTED_SYMBOLS", "Batch",
]

LIB_PATH = Path(__file__).resolve().parent / "libgoferro.so"
MU0 = 1.256637061e-6  # magnetic_system.h:14

ABI_VERSION = 4
FLAG_MAGNETOSTATICS = 1 << 0
FLAG_MAG_1D = 1 << 1
FLAG_DEMAG_FACTOR = 1 << 2
FLAG_PRESCRIBE_HEXT = 1 << 3
FLAG_ORIENT_ROTATE = 1 << 4
FLAG_HEXT_NONUNIF = 1 << 5
FLAG_EMDYNAMIC = 1 << 6
FLAG_PRESCRIBE_EEXT = 1 << 7
FLAG_ELASTODYNAMICS = 1 << 8
FLAG_ELASTO_BACKACTION = 1 << 9
FLAG_ELASTO_ON_PM = 1 << 10
FLAG_GAUSSIAN_PULSE = 1 << 11
FLAG_1D_ABC = 1 << 12
FLAG_1D_ABC_ELAST_ONLYTOP = 1 << 13
FLAG_EM_FROM_P = 1 << 14
FLAG_EM_FROM_M = 1 << 15
FLAG_EM_BACKACTION_P = 1 << 16
FLAG_EM_BACKACTION_M = 1 << 17
FLAG_PLANE_EM = 1 << 18
FLAG_JF_INPUT = 1 << 19
FLAG_PEC_XY = 1 << 20
FLAG_PEC_Z = 1 << 21
FLAG_1D_ABC_EM_ONLYTOP = 1 << 22
FLAG_HDRIVE = 1 << 23
FLAG_ELECTROSTATICS = 1 << 24
FLAG_SPIN_PUMPING = 1 << 25
FLAG_FLEXO = 1 << 26
FLAG_ELASTOSTATIC = 1 << 27
FLAG_PRESCRIBED_M = 1 << 28

(FIELD_M, FIELD_M_STORE, FIELD_HSTAT, FIELD_HEXT_NONUNIF, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_P, FIELD_Q, FIELD_P_STORE,
 FIELD_ESTAT, FIELD_U, FIELD_V, FIELD_U_STORE, FIELD_STRAIN_DIAG, FIELD_STRAIN_OFF, FIELD_E_EM, FIELD_H_EM,
 FIELD_E_EM_STORE, FIELD_H_EM_STORE, FIELD_E_CELL, FIELD_H_CELL, FIELD_JP, FIELD_HDRIVE, FIELD_STRAIN_T0_DIAG,
 FIELD_STRAIN_T0_OFF, FIELD_J_ISHE, FIELD_ELASTOFORCE, FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF) = range(29)


def yee_shapes(shape, field):
    """Shapes of the three component arrays of a Yee field
    (EMdynamic_initialize.cpp:68-95): E on edges, H on faces."""
    nx, ny, nz = (int(v) for v in shape)
    if field in (FIELD_E_EM, FIELD_E_EM_STORE):
        return [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    if field in (FIELD_H_EM, FIELD_H_EM_STORE):
        return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    raise ValueError("not a Yee field")


def field_size(shape, field) -> int:
    """Number of doubles a field occupies in gf_upload / gf_download."""
    n = int(shape[0] * shape[1] * shape[2])
    if field in (FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_EM, FIELD_H_EM_STORE):
        return sum(a * b * c for a, b, c in yee_shapes(shape, field))
    if field == FIELD_JP:
        return 9 * n
    return 3 * n


def split_yee(flat, shape, field):
    """Concatenated x|y|z Yee array -> list of three component arrays."""
    out, o = [], 0
    for s in yee_shapes(shape, field):
        k = s[0] * s[1] * s[2]
        out.append(np.asarray(flat[o:o + k]).reshape(s))
        o += k
    return out


def join_yee(parts):
    return np.concatenate([np.ascontiguousarray(p, dtype=np.float64).reshape(-1) for p in parts])

KERNEL_CLASSES = ["mag_stage", "afm_stage", "fft_z_fwd", "fft_y_fwd", "fft_x_conv", "fft_y_inv",
                  "fft_z_inv", "other", "fe_stage", "elastic_stage", "elastic_strain", "em_e", "em_h"]

# every entry point declared in include/goferro.h
EXPORTED_SYMBOLS = [
    "gf_create", "gf_destroy", "gf_last_error", "gf_upload", "gf_download", "gf_time_advance",
    "gf_mag_stage", "gf_afm_stage", "gf_fe_stage", "gf_elastic_stage", "gf_em_stage", "gf_field_size", "gf_demag",
    "gf_prescribe_m", "gf_step", "gf_average", "gf_average_noalloc", "gf_get_time", "gf_sync",
    "gf_get_stream", "gf_set_kernel_timing", "gf_kernel_times", "gf_set_material", "gf_set_time",
    "gf_adjoint_forward", "gf_adjoint_backward", "gf_adjoint_stage",
    "gf_batch_create", "gf_batch_destroy", "gf_batch_last_error", "gf_batch_step", "gf_batch_sync",
    "gf_batch_get_stream", "gf_batch_config", "gf_batch_set_config", "gf_batch_tune",
    "gf_state_size", "gf_state_save", "gf_state_load", "gf_state_write", "gf_state_read",
    "gf_write_cells", "gf_read_cells", "gf_selftest_stage_io",
]


class GoFerroError(RuntimeError):
    """A gf_* call returned a non-zero status."""


class Material(ctypes.Structure):
    """gf_material (include/goferro.h); field meaning per material.h:3-58."""

    _fields_ = [
        ("if_FM", ctypes.c_int32), ("anisotropy_type", ctypes.c_int32),
        ("Ms", ctypes.c_double), ("gyro", ctypes.c_double), ("FM_damping", ctypes.c_double),
        ("K1", ctypes.c_double), ("K2", ctypes.c_double), ("Aex", ctypes.c_double), ("iDMI", ctypes.c_double),
        ("B1", ctypes.c_double), ("B2", ctypes.c_double),
        ("if_AFM", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("Ms_AFM1", ctypes.c_double), ("Ms_AFM2", ctypes.c_double),
        ("gyro_AFM1", ctypes.c_double), ("gyro_AFM2", ctypes.c_double),
        ("damping_AFM1", ctypes.c_double), ("damping_AFM2", ctypes.c_double),
        ("K1_AFM1", ctypes.c_double), ("K1_AFM2", ctypes.c_double),
        ("uniaxis_AFM", ctypes.c_double * 3),
        ("Aex_AFM1", ctypes.c_double), ("Aex_AFM2", ctypes.c_double),
        ("J_AFM", ctypes.c_double),
        ("B1_AFM1", ctypes.c_double), ("B1_AFM2", ctypes.c_double),
        ("B2_AFM1", ctypes.c_double), ("B2_AFM2", ctypes.c_double),
        ("if_FE", ctypes.c_int32), ("reserved1", ctypes.c_int32),
        ("FE_mass", ctypes.c_double), ("FE_damping", ctypes.c_double),
        ("a1", ctypes.c_double), ("a11", ctypes.c_double), ("a12", ctypes.c_double), ("a111", ctypes.c_double),
        ("a112", ctypes.c_double), ("a123", ctypes.c_double), ("a1111", ctypes.c_double), ("a1112", ctypes.c_double),
        ("a1122", ctypes.c_double), ("a1123", ctypes.c_double),
        ("G11", ctypes.c_double),
        ("Q11", ctypes.c_double), ("Q12", ctypes.c_double), ("Q44", ctypes.c_double),
        ("tv1", ctypes.c_double), ("tv2", ctypes.c_double), ("tv3", ctypes.c_double),
        ("c11", ctypes.c_double), ("c12", ctypes.c_double), ("c44", ctypes.c_double), ("density", ctypes.c_double),
        ("elast_mass_damping", ctypes.c_double), ("elast_stiff_damping", ctypes.c_double),
        ("cijkl_glb", (ctypes.c_double * 6) * 6),
        ("lamda100", ctypes.c_double), ("lamda111", ctypes.c_double),
        ("lamda100_AFM1", ctypes.c_double), ("lamda100_AFM2", ctypes.c_double),
        ("lamda111_AFM1", ctypes.c_double), ("lamda111_AFM2", ctypes.c_double),
        ("if_PEC", ctypes.c_int32), ("reserved2", ctypes.c_int32),
        ("r_permittivity", ctypes.c_double), ("conductivity", ctypes.c_double),
        ("comp_n1", ctypes.c_double), ("omega_plasma_n1", ctypes.c_double), ("tao_e_n1", ctypes.c_double),
        ("comp_n2", ctypes.c_double), ("omega_plasma_n2", ctypes.c_double), ("tao_e_n2", ctypes.c_double),
        ("comp_n3", ctypes.c_double), ("omega_plasma_n3", ctypes.c_double), ("tao_e_n3", ctypes.c_double),
        ("if_spin_pump", ctypes.c_int32), ("reserved3", ctypes.c_int32),
        ("spin_hall_angle", ctypes.c_double), ("spin_mix_cond", ctypes.c_double),
        ("spin_diffus_length", ctypes.c_double),
        ("f11", ctypes.c_double), ("f12", ctypes.c_double), ("f44", ctypes.c_double),
    ]

    @classmethod
    def make(cls, **kw) -> "Material":
        m = cls()
        for k, v in kw.items():
            if k == "uniaxis_AFM":
                u = np.asarray(v, dtype=float)
                u = u / np.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])  # read.cpp:151-155
                m.uniaxis_AFM[:] = list(u)
            else:
                setattr(m, k, v)
        return m


class _Desc(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("periodic", ctypes.c_int32 * 3), ("flags", ctypes.c_uint32),
        ("dt", ctypes.c_double),
        ("n_materials", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("materials", ctypes.POINTER(Material)),
        ("material_cell", ctypes.POINTER(ctypes.c_uint32)),
        ("Tg2c", (ctypes.c_double * 3) * 3),
        ("Hext_stat", ctypes.c_double * 3), ("Hext_altr", ctypes.c_double * 3),
        ("H_altr_freq", ctypes.c_double * 3), ("demag_fac", ctypes.c_double * 3),
        ("n_prescribe_Hext", ctypes.c_int32), ("reserved1", ctypes.c_int32),
        ("prescribe_Hext", ctypes.POINTER(ctypes.c_double)),
        ("Eext_stat", ctypes.c_double * 3), ("Eext_altr", ctypes.c_double * 3),
        ("E_altr_freq", ctypes.c_double * 3),
        ("n_prescribe_Eext", ctypes.c_int32), ("reserved2", ctypes.c_int32),
        ("prescribe_Eext", ctypes.POINTER(ctypes.c_double)),
        ("input_strain_type", ctypes.c_int32), ("input_strain_sourcez", ctypes.c_int32),
        ("input_strain_component", ctypes.c_int32), ("num_strain_cycles", ctypes.c_int32),
        ("sigma_gauss", ctypes.c_double), ("amplitude_gauss", ctypes.c_double),
        ("weighting_factor_fourth_Liao", ctypes.c_double),
        ("planeEM_E", ctypes.c_double * 3), ("planeEM_E_freq", ctypes.c_double * 3),
        ("planeEM_source_z", ctypes.c_int32), ("Jf_input_type", ctypes.c_int32),
        ("Jf_input_component", ctypes.c_int32), ("num_Jf_cycles", ctypes.c_int32),
        ("Jfin_lo", ctypes.c_int32 * 3), ("Jfin_hi", ctypes.c_int32 * 3),
        ("Jf_input_amp", ctypes.c_double), ("Jf_input_freq", ctypes.c_double), ("Jf_input_sigma", ctypes.c_double),
        ("Hdrive_freq", ctypes.c_double),
        ("scale_elec", ctypes.c_double),
        ("n_layers", ctypes.c_int32), ("reserved3", ctypes.c_int32),
        ("nz_layer", ctypes.POINTER(ctypes.c_int32)),
        ("scale_elasto", ctypes.c_double),
        ("precess_angle", ctypes.c_double), ("precess_frequency", ctypes.c_double),
    ]


class _AdjDesc(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("material", ctypes.c_int32),
        ("probe_lo", ctypes.c_int32 * 3), ("probe_hi", ctypes.c_int32 * 3),
        ("probe_freq", ctypes.c_double), ("nsteps", ctypes.c_int64), ("checkpoint_every", ctypes.c_int64),
        ("precision", ctypes.c_int32), ("resident_mib", ctypes.c_int32),
    ]


@dataclasses.dataclass
class Problem:
    """One simulation, in the reference's terms (read.cpp:4-237)."""

    shape: tuple[int, int, int]
    spacing: tuple[float, float, float]
    materials: Sequence[Material]
    material_cell: np.ndarray                    # uint32, shape `shape`, 0 = vacuum
    dt: float
    periodic: tuple[bool, bool, bool] = (True, True, True)
    flags: int = 0
    Hext_stat: tuple[float, float, float] = (0.0, 0.0, 0.0)
    Hext_altr: tuple[float, float, float] = (0.0, 0.0, 0.0)
    H_altr_freq: tuple[float, float, float] = (0.0, 0.0, 0.0)
    demag_fac: tuple[float, float, float] = (0.0, 0.0, 0.0)
    Tg2c: Optional[np.ndarray] = None
    prescribe_Hext: Optional[np.ndarray] = None  # rows (t, Hx, Hy, Hz)
    Eext_stat: tuple[float, float, float] = (0.0, 0.0, 0.0)
    Eext_altr: tuple[float, float, float] = (0.0, 0.0, 0.0)
    E_altr_freq: tuple[float, float, float] = (0.0, 0.0, 0.0)
    prescribe_Eext: Optional[np.ndarray] = None  # rows (t, Ex, Ey, Ez)
    input_strain_type: int = 0
    input_strain_sourcez: int = 0
    input_strain_component: str = "x"
    num_strain_cycles: int = 0
    sigma_gauss: float = 0.0
    amplitude_gauss: float = 0.0
    weighting_factor_fourth_Liao: float = 0.5
    planeEM_E: tuple[float, float, float] = (0.0, 0.0, 0.0)
    planeEM_E_freq: tuple[float, float, float] = (0.0, 0.0, 0.0)
    planeEM_source_z: int = 0
    Jf_input_type: int = 0
    Jf_input_component: str = "x"
    num_Jf_cycles: int = 0
    Jfin_lo: tuple[int, int, int] = (0, 0, 0)
    Jfin_hi: tuple[int, int, int] = (0, 0, 0)
    Jf_input_amp: float = 0.0
    Jf_input_freq: float = 0.0
    Jf_input_sigma: float = 0.0
    Hdrive_freq: float = 0.0
    scale_elec: float = 1.0
    nz_layer: Optional[Sequence[int]] = None       # layer stack bottom to top; None: runs of column (0, 0)
    scale_elasto: float = 1.0
    precess_angle: float = 0.0                     # degrees (FLAG_PRESCRIBED_M)
    precess_frequency: float = 0.0                 # Hz

    @property
    def n(self) -> int:
        return int(self.shape[0] * self.shape[1] * self.shape[2])

    def desc(self):
        """Return (gf_desc, keep-alive list) for this problem."""
        cell = np.ascontiguousarray(self.material_cell, dtype=np.uint32).reshape(-1)
        if cell.size != self.n:
            raise ValueError("material_cell has the wrong size")
        mats = (Material * max(1, len(self.materials)))(*self.materials)
        d = _Desc()
        d.struct_size = ctypes.sizeof(_Desc)
        d.abi_version = ABI_VERSION
        d.nx, d.ny, d.nz = (int(v) for v in self.shape)
        d.dx, d.dy, d.dz = (float(v) for v in self.spacing)
        d.periodic[:] = [int(bool(p)) for p in self.periodic]
        d.flags = int(self.flags)
        d.dt = float(self.dt)
        d.n_materials = len(self.materials)
        d.materials = ctypes.cast(mats, ctypes.POINTER(Material))
        d.material_cell = cell.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        T = np.eye(3) if self.Tg2c is None else np.asarray(self.Tg2c, dtype=float)
        for a in range(3):
            for b in range(3):
                d.Tg2c[a][b] = float(T[a, b])
        d.Hext_stat[:] = [float(v) for v in self.Hext_stat]
        d.Hext_altr[:] = [float(v) for v in self.Hext_altr]
        d.H_altr_freq[:] = [float(v) for v in self.H_altr_freq]
        d.demag_fac[:] = [float(v) for v in self.demag_fac]
        d.Eext_stat[:] = [float(v) for v in self.Eext_stat]
        d.Eext_altr[:] = [float(v) for v in self.Eext_altr]
        d.E_altr_freq[:] = [float(v) for v in self.E_altr_freq]
        d.input_strain_type = int(self.input_strain_type)
        d.input_strain_sourcez = int(self.input_strain_sourcez)
        d.input_strain_component = ord(self.input_strain_component[0]) if self.input_strain_component else 0
        d.num_strain_cycles = int(self.num_strain_cycles)
        d.sigma_gauss = float(self.sigma_gauss)
        d.amplitude_gauss = float(self.amplitude_gauss)
        d.weighting_factor_fourth_Liao = float(self.weighting_factor_fourth_Liao)
        d.planeEM_E[:] = [float(v) for v in self.planeEM_E]
        d.planeEM_E_freq[:] = [float(v) for v in self.planeEM_E_freq]
        d.planeEM_source_z = int(self.planeEM_source_z)
        d.Jf_input_type = int(self.Jf_input_type)
        d.Jf_input_component = ord(self.Jf_input_component[0]) if self.Jf_input_component else 0
        d.num_Jf_cycles = int(self.num_Jf_cycles)
        d.Jfin_lo[:] = [int(v) for v in self.Jfin_lo]
        d.Jfin_hi[:] = [int(v) for v in self.Jfin_hi]
        d.Jf_input_amp = float(self.Jf_input_amp)
        d.Jf_input_freq = float(self.Jf_input_freq)
        d.Jf_input_sigma = float(self.Jf_input_sigma)
        d.Hdrive_freq = float(self.Hdrive_freq)
        d.scale_elec = float(self.scale_elec)
        d.scale_elasto = float(self.scale_elasto)
        d.precess_angle = float(self.precess_angle)
        d.precess_frequency = float(self.precess_frequency)
        keep = [cell, mats]
        if self.nz_layer is not None:
            lay = np.ascontiguousarray(self.nz_layer, dtype=np.int32)
            d.n_layers = lay.size
            d.nz_layer = lay.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            keep.append(lay)
        if self.prescribe_Hext is not None:
            pre = np.ascontiguousarray(self.prescribe_Hext, dtype=np.float64).reshape(-1, 4)
            d.n_prescribe_Hext = pre.shape[0]
            d.prescribe_Hext = pre.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            keep.append(pre)
        if self.prescribe_Eext is not None:
            pre = np.ascontiguousarray(self.prescribe_Eext, dtype=np.float64).reshape(-1, 4)
            d.n_prescribe_Eext = pre.shape[0]
            d.prescribe_Eext = pre.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            keep.append(pre)
        return d, keep


_lib = None


def load_library(path: Path = LIB_PATH) -> ctypes.CDLL:
    """Load libgoferro.so (built in-tree); raise if it is missing.  The
    environment variable GOFERRO_LIB names another build of the same
    library (a tuning variant) to load instead."""
    global _lib
    if _lib is not None:
        return _lib
    if path == LIB_PATH and os.environ.get("GOFERRO_LIB"):
        path = Path(os.environ["GOFERRO_LIB"])
    if not Path(path).exists():
        raise GoFerroError(f"{path} is missing: build it with goferro._build.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(str(path))
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "gf_create": ([ctypes.POINTER(_Desc), ctypes.POINTER(P)], ctypes.c_int),
        "gf_destroy": ([P], ctypes.c_int),
        "gf_last_error": ([P], ctypes.c_char_p),
        "gf_upload": ([P, ctypes.c_int, dp, ctypes.c_size_t], ctypes.c_int),
        "gf_download": ([P, ctypes.c_int, dp, ctypes.c_size_t], ctypes.c_int),
        "gf_time_advance": ([P], ctypes.c_int),
        "gf_mag_stage": ([P, ctypes.c_int], ctypes.c_int),
        "gf_afm_stage": ([P, ctypes.c_int], ctypes.c_int),
        "gf_fe_stage": ([P, ctypes.c_int], ctypes.c_int),
        "gf_elastic_stage": ([P, ctypes.c_int], ctypes.c_int),
        "gf_em_stage": ([P, ctypes.c_int], ctypes.c_int),
        "gf_field_size": ([P, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "gf_demag": ([P], ctypes.c_int),
        "gf_prescribe_m": ([P], ctypes.c_int),
        "gf_step": ([P, ctypes.c_int64], ctypes.c_int),
        "gf_average": ([P, ctypes.c_int, dp], ctypes.c_int),
        "gf_average_noalloc": ([P, ctypes.c_int, P], ctypes.c_int),
        "gf_get_time": ([P, dp], ctypes.c_int),
        "gf_sync": ([P], ctypes.c_int),
        "gf_get_stream": ([P, ctypes.POINTER(P)], ctypes.c_int),
        "gf_set_kernel_timing": ([P, ctypes.c_int], ctypes.c_int),
        "gf_kernel_times": ([P, dp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gf_set_material": ([P, ctypes.c_int, ctypes.POINTER(Material)], ctypes.c_int),
        "gf_set_time": ([P, ctypes.c_double], ctypes.c_int),
        "gf_adjoint_forward": ([P, ctypes.POINTER(_AdjDesc), dp], ctypes.c_int),
        "gf_adjoint_backward": ([P, dp, dp, ctypes.c_size_t], ctypes.c_int),
        "gf_adjoint_stage": ([P, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_size_t], ctypes.c_int),
        "gf_batch_create": ([ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
        "gf_batch_destroy": ([P], ctypes.c_int),
        "gf_batch_last_error": ([P], ctypes.c_char_p),
        "gf_batch_step": ([P, ctypes.c_int64], ctypes.c_int),
        "gf_batch_sync": ([P], ctypes.c_int),
        "gf_batch_get_stream": ([P, ctypes.POINTER(P)], ctypes.c_int),
        "gf_batch_config": ([P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "gf_batch_set_config": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "gf_batch_tune": ([P, ctypes.c_int, dp], ctypes.c_int),
        "gf_state_size": ([P, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "gf_state_save": ([P, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
        "gf_state_load": ([P, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
        "gf_state_write": ([P, ctypes.c_char_p], ctypes.c_int),
        "gf_state_read": ([P, ctypes.c_char_p], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """Device-resident simulation (one gf_ctx)."""

    def __init__(self, problem: Problem):
        self.lib = load_library()
        self.problem = problem
        self.shape = tuple(int(v) for v in problem.shape)
        self.n = problem.n
        d, self._keep = problem.desc()
        h = ctypes.c_void_p()
        st = self.lib.gf_create(ctypes.byref(d), ctypes.byref(h))
        if st != 0:
            raise GoFerroError(f"gf_create failed ({st}): {self.lib.gf_last_error(None).decode()}")
        self._h = h

    # -- plumbing
    def _check(self, st: int, what: str) -> None:
        if st != 0:
            raise GoFerroError(f"{what} failed ({st}): {self.lib.gf_last_error(self._h).decode()}")

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self.lib.gf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- fields
    def upload(self, field: int, value) -> None:
        """Upload a field: (3, nx, ny, nz) for cell vectors, (3, 3, nx, ny, nz)
        for Jp, the concatenated x|y|z array (or a list of the three
        component arrays) for the Yee E and H fields."""
        if isinstance(value, (list, tuple)):
            value = join_yee(value)
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
        if a.size != field_size(self.shape, field):
            raise ValueError(f"field {field} must hold {field_size(self.shape, field)} values")
        self._check(self.lib.gf_upload(self._h, field, _dptr(a), a.size), "gf_upload")

    def download(self, field: int) -> np.ndarray:
        """Download a field; Yee fields come back concatenated x|y|z
        (see split_yee)."""
        a = np.empty(field_size(self.shape, field), dtype=np.float64)
        self._check(self.lib.gf_download(self._h, field, _dptr(a), a.size), "gf_download")
        if field in (FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_EM, FIELD_H_EM_STORE):
            return a
        if field == FIELD_JP:
            return a.reshape((3, 3) + self.shape)
        return a.reshape((3,) + self.shape)

    def set_material(self, index: int, material: "Material") -> None:
        """Replace the numbers of material `index` (1-based); kinds stay."""
        self._check(self.lib.gf_set_material(self._h, int(index), ctypes.byref(material)), "gf_set_material")
        mats = list(self.problem.materials)
        mats[index - 1] = material
        self.problem.materials = mats

    def field_size(self, field: int) -> int:
        n = ctypes.c_size_t()
        self._check(self.lib.gf_field_size(self._h, field, ctypes.byref(n)), "gf_field_size")
        return int(n.value)

    # -- protocol
    def time_advance(self) -> None:
        self._check(self.lib.gf_time_advance(self._h), "gf_time_advance")

    def mag_stage(self, stage: int) -> None:
        self._check(self.lib.gf_mag_stage(self._h, stage), "gf_mag_stage")

    def afm_stage(self, stage: int) -> None:
        self._check(self.lib.gf_afm_stage(self._h, stage), "gf_afm_stage")

    def fe_stage(self, stage: int) -> None:
        self._check(self.lib.gf_fe_stage(self._h, stage), "gf_fe_stage")

    def elastic_stage(self, stage: int) -> None:
        self._check(self.lib.gf_elastic_stage(self._h, stage), "gf_elastic_stage")

    def em_stage(self, stage: int) -> None:
        self._check(self.lib.gf_em_stage(self._h, stage), "gf_em_stage")

    def demag(self) -> None:
        self._check(self.lib.gf_demag(self._h), "gf_demag")

    def prescribe_m(self) -> None:
        """Set the prescribed bottom layer at the current time (start-up)."""
        self._check(self.lib.gf_prescribe_m(self._h), "gf_prescribe_m")

    def step(self, nsteps: int = 1) -> None:
        self._check(self.lib.gf_step(self._h, int(nsteps)), "gf_step")

    def average(self, field: int = FIELD_M) -> np.ndarray:
        out = np.zeros(3)
        self._check(self.lib.gf_average(self._h, field, _dptr(out)), "gf_average")
        return out

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        self._check(self.lib.gf_get_time(self._h, ctypes.byref(t)), "gf_get_time")
        return t.value

    @time.setter
    def time(self, t: float) -> None:
        self._check(self.lib.gf_set_time(self._h, float(t)), "gf_set_time")

    def sync(self) -> None:
        self._check(self.lib.gf_sync(self._h), "gf_sync")

    # -- restart image (gf_state_*)
    def state_image(self) -> bytes:
        """Everything the next step reads, at full precision, as bytes."""
        n = ctypes.c_size_t()
        self._check(self.lib.gf_state_size(self._h, ctypes.byref(n)), "gf_state_size")
        buf = ctypes.create_string_buffer(n.value)
        self._check(self.lib.gf_state_save(self._h, buf, n.value), "gf_state_save")
        return buf.raw

    def restore_state(self, image: bytes) -> None:
        """Continue from a state_image() of a context of the same problem."""
        buf = ctypes.create_string_buffer(bytes(image), len(image))
        self._check(self.lib.gf_state_load(self._h, buf, len(image)), "gf_state_load")
        self._sync_materials(bytes(image))

    def save_state(self, path) -> None:
        self._check(self.lib.gf_state_write(self._h, str(path).encode()), "gf_state_write")

    def load_state(self, path) -> None:
        self._check(self.lib.gf_state_read(self._h, str(path).encode()), "gf_state_read")
        self._sync_materials(Path(path).read_bytes())

    def _sync_materials(self, image: bytes) -> None:
        # the image's material table replaces the problem's (gf_state_load
        # applied any record that differs)
        from .io import _StateHeader
        hs = ctypes.sizeof(_StateHeader)
        h = _StateHeader.from_buffer_copy(image[:hs])
        sz = ctypes.sizeof(Material)
        self.problem.materials = [Material.from_buffer_copy(image[hs + i * sz:hs + (i + 1) * sz])
                                  for i in range(h.n_materials)]

    # -- adjoint sweep
    def adjoint_forward(self, material: int, probe_lo, probe_hi, freq: float, nsteps: int,
                        checkpoint_every: int = 0, precision: int = 64, resident_mib: int = 0) -> float:
        """N steps keeping checkpoints; returns the probe spectral loss.
        precision=32 runs the FP32 variant on copies of the state (the
        context's state and time are left unchanged).  resident_mib: device
        memory for the stage stores and fields of the last steps, which the
        reverse sweep then walks without recomputing them."""
        d = _AdjDesc()
        d.precision = int(precision)
        d.resident_mib = int(resident_mib)
        d.struct_size = ctypes.sizeof(_AdjDesc)
        d.material = int(material)
        d.probe_lo[:] = [int(v) for v in probe_lo]
        d.probe_hi[:] = [int(v) for v in probe_hi]
        d.probe_freq = float(freq)
        d.nsteps = int(nsteps)
        d.checkpoint_every = int(checkpoint_every)
        L = ctypes.c_double()
        self._check(self.lib.gf_adjoint_forward(self._h, ctypes.byref(d), ctypes.byref(L)), "gf_adjoint_forward")
        return L.value

    def adjoint_backward(self, drive: bool = True):
        """Reverse sweep: (dL/d(alpha, Aex, B1), dL/dHdrive as (3,)+shape or None)."""
        g = np.zeros(3)
        gh = np.zeros(3 * self.n) if drive else None
        null = ctypes.POINTER(ctypes.c_double)()
        self._check(self.lib.gf_adjoint_backward(self._h, _dptr(g), null if gh is None else _dptr(gh),
                                                 0 if gh is None else gh.size), "gf_adjoint_backward")
        return g, (None if gh is None else gh.reshape((3,) + self.shape))

    def adjoint_stage(self, material: int, stage: int, store, kbar, drive: bool = True):
        """Vector-Jacobian product of one stage increment: (ybar, grad3, grad_hdrive)."""
        y = np.ascontiguousarray(store, dtype=np.float64).reshape(-1)
        kb = np.ascontiguousarray(kbar, dtype=np.float64).reshape(-1)
        yb = np.zeros(3 * self.n)
        g = np.zeros(3)
        gh = np.zeros(3 * self.n) if drive else None
        null = ctypes.POINTER(ctypes.c_double)()
        self._check(self.lib.gf_adjoint_stage(self._h, int(material), int(stage), _dptr(y), _dptr(kb), _dptr(yb),
                                              _dptr(g), null if gh is None else _dptr(gh), yb.size),
                    "gf_adjoint_stage")
        return yb.reshape((3,) + self.shape), g, (None if gh is None else gh.reshape((3,) + self.shape))

    def set_kernel_timing(self, enable: bool) -> None:
        self._check(self.lib.gf_set_kernel_timing(self._h, int(bool(enable))), "gf_set_kernel_timing")

    def kernel_times(self) -> dict:
        ms = (ctypes.c_double * len(KERNEL_CLASSES))()
        cnt = (ctypes.c_int64 * len(KERNEL_CLASSES))()
        self._check(self.lib.gf_kernel_times(self._h, ms, cnt), "gf_kernel_times")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNEL_CLASSES)}


class Batch:
    """Independent contexts stepped together (gf_batch): one launch per RK4
    stage covers every instance and each step replays a HIP graph.  The
    contexts are borrowed and stay usable between batch steps."""

    def __init__(self, contexts):
        self.lib = load_library()
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c._h.value for c in self.contexts])
        h = ctypes.c_void_p()
        st = self.lib.gf_batch_create(arr, len(self.contexts), ctypes.byref(h))
        if st != 0:
            raise GoFerroError(f"gf_batch_create failed ({st}): {self.lib.gf_batch_last_error(None).decode()}")
        self._h = h

    def _check(self, st: int, what: str) -> None:
        if st != 0:
            raise GoFerroError(f"{what} failed ({st}): {self.lib.gf_batch_last_error(self._h).decode()}")

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self.lib.gf_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, nsteps: int = 1) -> None:
        self._check(self.lib.gf_batch_step(self._h, int(nsteps)), "gf_batch_step")

    def sync(self) -> None:
        self._check(self.lib.gf_batch_sync(self._h), "gf_batch_sync")

    @property
    def config(self) -> tuple:
        """(workgroup size, cells per thread, grid cap)."""
        t, p, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.gf_batch_config(self._h, ctypes.byref(t), ctypes.byref(p), ctypes.byref(g)),
                    "gf_batch_config")
        return t.value, p.value, g.value

    def set_config(self, threads: int, cells_per_thread: int = 1, grid_cap: int = 0) -> None:
        self._check(self.lib.gf_batch_set_config(self._h, int(threads), int(cells_per_thread), int(grid_cap)),
                    "gf_batch_set_config")

    def tune(self, reps: int = 3) -> float:
        """Pick the fastest launch shape; returns its ms per stage."""
        ms = ctypes.c_double()
        self._check(self.lib.gf_batch_tune(self._h, int(reps), ctypes.byref(ms)), "gf_batch_tune")
        return ms.value
