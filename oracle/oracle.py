"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product.  It consumes the same
``goferro.Problem`` description as the GPU binding (types only).
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
FIELD_DM = 100   # oracle-only: combined RK4 increment of the last step
FIELD_Q_STORE = 101   # oracle-only: ferroelectric q store
FIELD_DV1 = 102       # oracle-only: elastic velocity increment of stage 1
FIELD_K1 = 103        # oracle-only: FM increment of the last stage-1 get_dm

_libs = {}


def build() -> Path:
    """Compile liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def build_planted(plant: int, outdir) -> Path:
    """Compile a variant of the oracle with the deliberate mistake number
    `plant` (GFO_PLANT in gf_oracle.c) into outdir, for the tests that show
    a pin fails on it.  Same flags as the Makefile."""
    out = Path(outdir) / f"liboracle_plant{int(plant)}.so"
    subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-fopenmp", "-ffp-contract=off", f"-DGFO_PLANT={int(plant)}",
                    "-shared", "-o", str(out), str(HERE / "gf_oracle.c"), f"-I{HERE.parent / 'include'}", "-lm"],
                   check=True)
    return out


def lib(path=None) -> ctypes.CDLL:
    path = str(LIB if path is None else path)
    if path in _libs:
        return _libs[path]
    if path == str(LIB) and not LIB.exists():
        build()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    i64 = ctypes.c_int64
    sig = {
        "gfo_create": ([P], P), "gfo_destroy": ([P], None),
        "gfo_set_field": ([P, ctypes.c_int, dp, ctypes.c_size_t], ctypes.c_int),
        "gfo_get_field": ([P, ctypes.c_int, dp, ctypes.c_size_t], ctypes.c_int),
        "gfo_time_advance": ([P], None),
        "gfo_mag_get_dm": ([P, ctypes.c_int], None), "gfo_mag_update": ([P, ctypes.c_int], None),
        "gfo_afm_get_dm": ([P, ctypes.c_int], None), "gfo_afm_update": ([P, ctypes.c_int], None),
        "gfo_fe_get": ([P, ctypes.c_int], None), "gfo_fe_update": ([P, ctypes.c_int], None),
        "gfo_elastic_get": ([P, ctypes.c_int], None), "gfo_elastic_update": ([P, ctypes.c_int], None),
        "gfo_em_get": ([P, ctypes.c_int], None), "gfo_em_update": ([P, ctypes.c_int], None),
        "gfo_demag": ([P], None), "gfo_step": ([P, i64], None), "gfo_prescribe_m": ([P], None),
        "gfo_average": ([P, ctypes.c_int, dp], None), "gfo_time": ([P], ctypes.c_double),
        "gfo_newell_real": ([i64, i64, i64, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp], None),
        "gfo_newell_F2": ([ctypes.c_double] * 3, ctypes.c_double),
        "gfo_newell_G2": ([ctypes.c_double] * 3, ctypes.c_double),
        "gfo_newell_A": ([ctypes.c_double] * 6, ctypes.c_double),
        "gfo_newell_B": ([ctypes.c_double] * 6, ctypes.c_double),
        "gfo_rfft3": ([i64, i64, i64, dp, dp], None),
        "gfo_irfft3": ([i64, i64, i64, dp, dp], None),
        "gfo_rdft3_naive": ([i64, i64, i64, dp, dp], None),
        "gfo_threads": ([], ctypes.c_int),
        "gfo_tl_probe_loss": ([P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                               ctypes.c_double, i64, dp, dp, dp, dp, dp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _libs[path] = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """The reference's protocol on the CPU for one Problem."""

    def __init__(self, problem, lib_path=None):
        self.L = lib(lib_path)
        self.problem = problem
        self.shape = tuple(int(v) for v in problem.shape)
        self.n = int(np.prod(self.shape))
        d, self._keep = problem.desc()
        self._h = self.L.gfo_create(ctypes.byref(d))
        if not self._h:
            raise RuntimeError("gfo_create rejected the problem")

    def close(self):
        if self._h:
            self.L.gfo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, field, value):
        from goferro import join_yee
        if isinstance(value, (list, tuple)):
            value = join_yee(value)
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
        if self.L.gfo_set_field(self._h, field, _dp(a), a.size) != 0:
            raise ValueError("bad field or size")

    def get(self, field):
        from goferro import FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_EM, FIELD_H_EM_STORE, FIELD_JP, field_size
        size = field_size(self.shape, field) if field < 100 else 3 * self.n
        a = np.empty(size)
        if self.L.gfo_get_field(self._h, field, _dp(a), a.size) != 0:
            raise ValueError("bad field or size")
        if field in (FIELD_E_EM, FIELD_E_EM_STORE, FIELD_H_EM, FIELD_H_EM_STORE):
            return a
        if field == FIELD_JP:
            return a.reshape((3, 3) + self.shape)
        return a.reshape((3,) + self.shape)

    def time_advance(self):
        self.L.gfo_time_advance(self._h)

    def mag_stage(self, s):
        """get_dm_RKs then update_m_RKs (one reference stage, FM only)."""
        self.L.gfo_mag_get_dm(self._h, s)
        self.L.gfo_mag_update(self._h, s)

    def afm_stage(self, s):
        self.L.gfo_afm_get_dm(self._h, s)
        self.L.gfo_afm_update(self._h, s)

    def fe_stage(self, s):
        self.L.gfo_fe_get(self._h, s)
        self.L.gfo_fe_update(self._h, s)

    def elastic_stage(self, s):
        self.L.gfo_elastic_get(self._h, s)
        self.L.gfo_elastic_update(self._h, s)

    def em_stage(self, s):
        self.L.gfo_em_get(self._h, s)
        self.L.gfo_em_update(self._h, s)

    def demag(self):
        self.L.gfo_demag(self._h)

    def prescribe_m(self):
        self.L.gfo_prescribe_m(self._h)

    def step(self, n=1):
        self.L.gfo_step(self._h, int(n))

    def average(self, field=0):
        out = np.zeros(3)
        self.L.gfo_average(self._h, field, _dp(out))
        return out

    def tl_probe_loss(self, material, lo, hi, freq, nsteps, dpar=(0.0, 0.0, 0.0), dhdrive=None, dm0=None):
        """Tangent-linear reference: (L, dL) of the probe spectral loss over
        nsteps steps from the current state along (dpar, dhdrive, dm0)."""
        lo_a = (ctypes.c_int * 3)(*[int(v) for v in lo])
        hi_a = (ctypes.c_int * 3)(*[int(v) for v in hi])
        dp3 = np.ascontiguousarray(dpar, dtype=np.float64)
        dh = None if dhdrive is None else np.ascontiguousarray(dhdrive, dtype=np.float64).reshape(-1)
        d0 = None if dm0 is None else np.ascontiguousarray(dm0, dtype=np.float64).reshape(-1)
        L, dL = np.zeros(1), np.zeros(1)
        null = ctypes.POINTER(ctypes.c_double)()
        st = self.L.gfo_tl_probe_loss(self._h, int(material), lo_a, hi_a, float(freq), int(nsteps), _dp(dp3),
                                      null if dh is None else _dp(dh), null if d0 is None else _dp(d0), _dp(L), _dp(dL))
        if st != 0:
            raise ValueError("unsupported context for the tangent-linear reference")
        return float(L[0]), float(dL[0])

    @property
    def time(self):
        return self.L.gfo_time(self._h)


def threads() -> int:
    """Host threads the oracle's parallel loops use."""
    return int(lib().gfo_threads())


def newell_real(shape, spacing):
    nx, ny, nz = shape
    out = np.empty(9 * nx * ny * nz)
    lib().gfo_newell_real(nx, ny, nz, *spacing, _dp(out))
    return out.reshape((9, nx, ny, nz))


def rfft3(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nx, ny, nz = x.shape
    out = np.empty(2 * nx * ny * (nz // 2 + 1))
    lib().gfo_rfft3(nx, ny, nz, _dp(x), _dp(out))
    o = out.reshape(nx, ny, nz // 2 + 1, 2)
    return o[..., 0] + 1j * o[..., 1]


def irfft3(X, nz):
    nx, ny, nzh = X.shape
    a = np.empty((nx, ny, nzh, 2))
    a[..., 0] = X.real
    a[..., 1] = X.imag
    a = np.ascontiguousarray(a)
    out = np.empty(nx * ny * nz)
    lib().gfo_irfft3(nx, ny, nz, _dp(a), _dp(out))
    return out.reshape(nx, ny, nz)


def rdft3_naive(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nx, ny, nz = x.shape
    out = np.empty(2 * nx * ny * (nz // 2 + 1))
    lib().gfo_rdft3_naive(nx, ny, nz, _dp(x), _dp(out))
    o = out.reshape(nx, ny, nz // 2 + 1, 2)
    return o[..., 0] + 1j * o[..., 1]
