"""The reference's on-disk contract: its input deck, its per-cell text files,
its output schedule, and a full-precision restart image.

* ``read_deck(directory)`` reads ``system_setting.in`` and ``materials.in``
  token for token as ``geometry_parameters::readgeo`` and
  ``global_parameters::read_global`` / ``read_materials`` do (read.cpp:4-237),
  applies the derivations of ``set_geometry`` (geometry.cpp:3-43),
  ``set_global`` (global.cpp:14-318), ``check_material`` and ``get_cijkl``
  (global.cpp:320-441), and returns a ``Deck``: the ``Problem`` plus every
  raw setting.  Options the library does not implement are reported, not
  ignored.
* ``write_cells`` / ``read_cells`` are the per-cell text files of
  inoutput.cpp (``m.<step>.dat``, ``m.in`` ...), through the library's
  ``gf_write_cells`` / ``gf_read_cells``.
* ``Outputs`` writes what main.cpp:169-197 and :486-650 write, with the
  reference's file names and formats, and ``run`` is main.cpp's time loop.
* ``Context.save_state`` / ``load_state`` (``gf_state_write`` /
  ``gf_state_read``) hold the restart image; ``state_info`` reads its header
  without a GPU.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import re
import struct
import warnings
from pathlib import Path
from typing import Optional

import numpy as np

from . import (
    FIELD_E_CELL, FIELD_E_EM, FIELD_ESTAT, FIELD_H_CELL, FIELD_H_EM, FIELD_HEXT_NONUNIF, FIELD_HSTAT, FIELD_JP,
    FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_P, FIELD_Q, FIELD_STRAIN_DIAG, FIELD_STRAIN_OFF,
    FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FIELD_U, FIELD_V, FIELD_ELASTOFORCE, FIELD_EIGENSTRAIN_T0_DIAG,
    FIELD_EIGENSTRAIN_T0_OFF,
    FLAG_1D_ABC, FLAG_1D_ABC_ELAST_ONLYTOP, FLAG_1D_ABC_EM_ONLYTOP, FLAG_DEMAG_FACTOR, FLAG_ELASTO_BACKACTION,
    FLAG_ELASTO_ON_PM, FLAG_ELASTODYNAMICS, FLAG_ELECTROSTATICS, FLAG_EM_BACKACTION_M, FLAG_EM_BACKACTION_P,
    FLAG_EM_FROM_M, FLAG_EM_FROM_P, FLAG_EMDYNAMIC, FLAG_GAUSSIAN_PULSE, FLAG_HEXT_NONUNIF, FLAG_JF_INPUT,
    FLAG_MAG_1D, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE, FLAG_PEC_XY, FLAG_PEC_Z, FLAG_PLANE_EM,
    FLAG_PRESCRIBE_EEXT, FLAG_PRESCRIBE_HEXT, FLAG_SPIN_PUMPING, FLAG_FLEXO, FLAG_ELASTOSTATIC, FIELD_J_ISHE,
    FLAG_PRESCRIBED_M,
    Context, GoFerroError, Material, Problem, load_library, split_yee, yee_shapes,
)

__all__ = [
    "Deck", "UnsupportedInput", "read_deck", "read_system_setting", "read_materials", "layer_stack",
    "write_cells", "read_cells", "initial_fields", "load_inputs", "Outputs", "run", "state_info",
]

PI = 3.14159265358979323846  # mathlib.h


class UnsupportedInput(GoFerroError):
    """The deck asks for an option the library does not implement."""


# ---------------------------------------------------------------------------
# std::ifstream semantics over the text of a file

_INT = re.compile(r"[+-]?\d+")
_FLOAT = re.compile(r"[+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)")
_BOOL = re.compile(r"true|false")


class _Stream:
    """`file >> x` (skip white space, then the longest prefix of the type),
    `file.ignore(1000, '\\n')` and `std::getline`.  Bools are read under
    std::boolalpha, so only `true` / `false` parse."""

    def __init__(self, text: str, name: str):
        self.s = text
        self.i = 0
        self.name = name

    def _where(self) -> str:
        return f"{self.name}, line {self.s.count(chr(10), 0, self.i) + 1}"

    def _skip(self) -> None:
        while self.i < len(self.s) and self.s[self.i] in " \t\n\r\v\f":
            self.i += 1

    def _take(self, rx: re.Pattern, what: str) -> str:
        self._skip()
        m = rx.match(self.s, self.i)
        if not m:
            got = self.s[self.i:self.i + 20].split("\n")[0]
            raise GoFerroError(f"{self._where()}: expected {what}, found {got!r}")
        self.i = m.end()
        return m.group(0)

    def int(self) -> int:
        return int(self._take(_INT, "an integer"))

    def uint(self) -> int:
        v = self.int()
        if v < 0:
            raise GoFerroError(f"{self._where()}: expected a non-negative integer")
        return v

    def float(self) -> float:
        return float(self._take(_FLOAT, "a number"))

    def bool(self) -> bool:
        return self._take(_BOOL, "true or false") == "true"

    def char(self) -> str:
        self._skip()
        if self.i >= len(self.s):
            raise GoFerroError(f"{self._where()}: expected a character")
        self.i += 1
        return self.s[self.i - 1]

    def ignore(self) -> None:
        """file.ignore(1000, '\\n')"""
        end = min(len(self.s), self.i + 1000)
        j = self.s.find("\n", self.i, end)
        self.i = end if j < 0 else j + 1

    def getline(self) -> str:
        j = self.s.find("\n", self.i)
        if j < 0:
            line, self.i = self.s[self.i:], len(self.s)
        else:
            line, self.i = self.s[self.i:j], j + 1
        return line


def _read_text(path: Path) -> str:
    try:
        return Path(path).read_text()
    except OSError as e:
        raise GoFerroError(f"cannot read {path}: {e}") from None


# ---------------------------------------------------------------------------
# system_setting.in

def read_system_setting(path) -> dict:
    """The raw settings of system_setting.in, keyed by the reference's
    variable names: readgeo (read.cpp:217-237) and read_global (:4-111),
    which re-reads the file after skipping its first 8 non-empty lines."""
    text = _read_text(path)
    f = _Stream(text, str(path))
    g: dict = {}
    g["nx_system"], g["ny_system"], g["nz_system"] = f.int(), f.int(), f.int(); f.ignore()
    g["dx"], g["dy"], g["dz"] = f.float(), f.float(), f.float(); f.ignore()
    g["periodicX"], g["periodicY"], g["periodicZ"] = f.bool(), f.bool(), f.bool(); f.ignore()
    g["if_read_struct"] = f.bool(); f.ignore()
    g["nx_work"], g["ny_work"] = f.int(), f.int(); f.ignore()
    g["num_layer_unit"] = f.int(); f.ignore()
    g["num_periods"], g["id_firstlayer"], g["id_lastlayer_unit"] = f.int(), f.int(), f.int(); f.ignore()
    nlu = g["num_layer_unit"]
    if nlu < 1:
        raise GoFerroError(f"{path}: num_layer_unit must be at least 1")
    g["pt_nz_layer_unit"] = [f.uint() for _ in range(nlu)]

    f = _Stream(text, str(path))
    done = 0
    while done < 8:
        if f.i >= len(f.s):
            raise GoFerroError(f"{path}: fewer than 8 non-empty lines before the global settings")
        if len(f.getline()) > 0:
            done += 1

    def row(*names, kind="float"):
        for nm in names:
            g[nm] = getattr(f, kind)()
        f.ignore()

    row("step", kind="uint")
    row("dt")
    row("num_materials", kind="uint")
    g["pt_material_layer_unit"] = [f.uint() for _ in range(nlu)]
    f.ignore()
    row("if_OrientRotate", kind="bool")
    g["Tg2c"] = []
    for _ in range(3):
        g["Tg2c"].append([f.float(), f.float(), f.float()])
        f.ignore()
    row("theta_mag", "phi_mag")
    row("px_init", "py_init", "pz_init")
    row("if_read_Hext_nonunif", kind="bool")
    g["Hext_stat"] = [f.float() for _ in range(3)]; f.ignore()
    g["Hext_altr"] = [f.float() for _ in range(3)]; f.ignore()
    g["H_altr_freq"] = [f.float() for _ in range(3)]; f.ignore()
    g["if_prescribe_Hext"] = f.bool(); g["num_prescribe_Hext"] = f.int(); f.ignore()
    g["Eext_stat"] = [f.float() for _ in range(3)]; f.ignore()
    g["Eext_altr"] = [f.float() for _ in range(3)]; f.ignore()
    g["E_altr_freq"] = [f.float() for _ in range(3)]; f.ignore()
    g["if_prescribe_Eext"] = f.bool(); g["num_prescribe_Eext"] = f.int(); f.ignore()
    row("if_magnetostatics", "if_mag_1Dmodel", kind="bool")
    g["if_demag_factor"] = f.bool()
    g["demag_fac_x"], g["demag_fac_y"], g["demag_fac_z"] = f.float(), f.float(), f.float(); f.ignore()
    row("if_elec_1Dmodel", "if_elec_1D_compensate", kind="bool")
    row("if_EMdynamic", "if_input_planeEM_E", kind="bool")
    g["planeEM_E"] = [f.float() for _ in range(3)]; f.ignore()
    g["planeEM_E_freq"] = [f.float() for _ in range(3)]; f.ignore()
    row("if_EM_fromP", "if_EM_fromM", kind="bool")
    row("if_EM_backaction_on_P", "if_EM_backaction_on_M", kind="bool")
    g["free_charge_nzi"], g["free_charge_nzf"], g["free_charge"] = f.int(), f.int(), f.float(); f.ignore()
    g["scale_elec"], g["elec_solver_limit"], g["elec_solver_tol"] = f.float(), f.uint(), f.float(); f.ignore()
    row("if_elastodynamics", "if_elasto_backaction_from_pORm", "if_elasto_on_pORm", kind="bool")
    g["if_gaussian_strain_pulse"] = f.bool()
    g["input_strain_type"], g["input_strain_sourcez"] = f.uint(), f.uint(); f.ignore()
    g["input_strain_component"] = f.char()
    g["sigma_gauss"], g["amplitude_gauss"], g["num_strain_cycles"] = f.float(), f.float(), f.uint(); f.ignore()
    row("if_elastostatic", "if_elastostatic_1D", kind="bool")
    row("if_read_strain_ext", "if_elastostatic_film", kind="bool")
    g["strain_ext_nzi"], g["strain_ext_nzf"] = f.int(), f.int()
    g["exx_ext_glb"], g["eyy_ext_glb"], g["exy_ext_glb"] = f.float(), f.float(), f.float(); f.ignore()
    g["scale_elasto"], g["elasto_solver_limit"], g["elasto_solver_tol"] = f.float(), f.uint(), f.float(); f.ignore()
    g["if_1D_ABC"], g["if_1D_ABC_ELAST_onlytop"], g["if_1D_ABC_EM_onlytop"] = f.bool(), f.bool(), f.bool()
    g["weighting_factor_fourth_Liao"] = f.float(); f.ignore()
    row("if_PEC_XY", "if_PEC_Z", kind="bool")
    row("if_flexo", kind="bool")
    row("if_spin_pumping", kind="bool")
    row("if_input_m", "if_input_AFMm", "if_input_Hstat", kind="bool")
    row("if_input_pandq", "if_input_Estat", kind="bool")
    row("if_input_em", "if_input_Jp", kind="bool")
    row("if_input_uandv", "if_input_straint0", "if_input_eigenstraint0_crt", kind="bool")
    row("if_input_elastoforce", kind="bool")

    def out(*names):
        for nm in names:
            g[nm] = f.bool() if nm.startswith("if_") else f.uint()
        f.ignore()

    out("if_output_ave", "output_step_ave")
    out("if_output_m", "output_step_m", "if_output_only_magcell", "output_step_magcell")
    out("if_output_AFMm", "output_step_AFMm")
    out("if_output_p", "output_step_p")
    out("if_output_em", "output_step_em", "if_output_em_onecell", "output_step_em_onecell", "em_onecell_index")
    out("if_output_strain", "output_step_strain")
    out("if_output_Hstat", "output_step_Hstat")
    out("if_output_Estat", "output_step_Estat")
    out("if_output_eigenstraint0_crt", "output_step_eigenstraint0_crt")
    out("if_output_q", "output_step_q")
    out("if_output_uandv", "output_step_uandv")
    out("if_output_elastoforce", "output_step_elastoforce")
    out("if_output_emYee", "output_step_emYee")
    out("if_output_Jp", "output_step_Jp")
    out("if_output_Jishe", "output_step_Jishe")
    row("if_Jf_input", kind="bool")
    g["Jfin_xi"], g["Jfin_yi"], g["Jfin_zi"] = f.uint(), f.uint(), f.uint(); f.ignore()
    g["Jfin_xf"], g["Jfin_yf"], g["Jfin_zf"] = f.uint(), f.uint(), f.uint(); f.ignore()
    g["Jf_input_type"] = f.uint()
    g["Jf_input_component"] = f.char()
    g["Jf_input_amp"], g["Jf_input_freq"], g["Jf_input_sigma"] = f.float(), f.float(), f.float()
    g["num_Jf_cycles"] = f.uint(); f.ignore()
    g["if_prescribed_m"] = f.bool(); g["precess_angle"], g["precess_frequency"] = f.float(), f.float(); f.ignore()
    return g


# ---------------------------------------------------------------------------
# materials.in

_MAT_ROWS = [
    ("c11", "c12", "c44"), ("density",), ("elast_mass_damping", "elast_stiff_damping"),
    ("?if_FM",), ("Ms", "gyro", "FM_damping"), ("#anisotropy_type",), ("K1", "K2"), ("Aex",), ("iDMI",),
    ("lamda100", "lamda111"),
    ("?if_AFM",), ("Ms_AFM1", "Ms_AFM2"), ("gyro_AFM1", "gyro_AFM2"), ("damping_AFM1", "damping_AFM2"),
    ("K1_AFM1", "K1_AFM2"), ("uniaxis_AFM0", "uniaxis_AFM1", "uniaxis_AFM2"), ("Aex_AFM1", "Aex_AFM2"), ("J_AFM",),
    ("lamda100_AFM1", "lamda100_AFM2"), ("lamda111_AFM1", "lamda111_AFM2"),
    ("?if_spin_pump",), ("spin_hall_angle", "spin_mix_cond", "spin_diffus_length"),
    ("?if_FE",), ("FE_mass", "FE_damping"), ("a1",), ("a11", "a12"), ("a111", "a112", "a123"),
    ("a1111", "a1112", "a1122", "a1123"), ("G11",), ("Q11", "Q12", "Q44"), ("f11", "f12", "f44"),
    ("?if_PEC",), ("r_permittivity",), ("conductivity",),
    ("comp_n1", "omega_plasma_n1", "tao_e_n1"), ("comp_n2", "omega_plasma_n2", "tao_e_n2"),
    ("comp_n3", "omega_plasma_n3", "tao_e_n3"),
]


def read_materials(path, num_materials: int) -> list[dict]:
    """The raw records of materials.in (read_materials, read.cpp:113-215):
    per material one title line, then one row per group of constants."""
    f = _Stream(_read_text(path), str(path))
    out = []
    for _ in range(num_materials):
        rec = {"title": f.getline().strip()}
        for names in _MAT_ROWS:
            for nm in names:
                if nm[0] == "?":
                    rec[nm[1:]] = f.bool()
                elif nm[0] == "#":
                    rec[nm[1:]] = f.int()
                else:
                    rec[nm] = f.float()
            f.ignore()
        out.append(rec)
    return out


def _cijkl_glb(c11: float, c12: float, c44: float, Tc2g) -> list[list[float]]:
    """get_cijkl (global.cpp:365-441), with its summation order."""
    Ccrt = [[0.0] * 6 for _ in range(6)]
    for a in range(3):
        Ccrt[a][a] = c11
        Ccrt[3 + a][3 + a] = c44
    for a, b in ((0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)):
        Ccrt[a][b] = c12
    voigt = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    C = [[[[0.0] * 3 for _ in range(3)] for _ in range(3)] for _ in range(3)]
    for I, (w, x) in enumerate(voigt):
        for J, (y, z) in enumerate(voigt):
            v = Ccrt[I][J]
            C[w][x][y][z] = v
            C[x][w][y][z] = v
            C[w][x][z][y] = v
            C[x][w][z][y] = v
    T = Tc2g
    out = [[0.0] * 6 for _ in range(6)]
    for I, (w, x) in enumerate(voigt):
        for J, (y, z) in enumerate(voigt):
            s = 0.0
            for i in range(3):
                for j in range(3):
                    for k in range(3):
                        for l in range(3):
                            s = s + T[w][i] * T[x][j] * T[y][k] * T[z][l] * C[i][j][k][l]
            out[I][J] = s
    return out


def _material(rec: dict, Tc2g) -> Material:
    """One gf_material from a raw record: the derived constants of
    read_materials (read.cpp:141-142, 162-165, 181-199), check_material's
    zeroing and the sequential normalisation of comp_n1..3 (global.cpp:
    320-363), and get_cijkl."""
    r = dict(rec)
    c11, c12, c44 = r["c11"], r["c12"], r["c44"]
    r["B1"] = -3. / 2. * r["lamda100"] * (c11 - c12)
    r["B2"] = -3. * r["lamda111"] * c44
    u = [r["uniaxis_AFM0"], r["uniaxis_AFM1"], r["uniaxis_AFM2"]]
    norm = math.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    r["uniaxis_AFM"] = [v / norm for v in u] if norm > 0 else [math.nan] * 3
    for s in ("1", "2"):
        r["B1_AFM" + s] = -3. / 2. * r["lamda100_AFM" + s] * (c11 - c12)
        r["B2_AFM" + s] = -3. * r["lamda111_AFM" + s] * c44
    r["tv1"] = 2. * (c11 * r["Q11"] + 2. * c12 * r["Q12"])
    r["tv2"] = 2. * (c11 * r["Q12"] + c12 * r["Q11"] + c12 * r["Q12"])
    r["tv3"] = 4. * c44 * r["Q44"]
    if not r["if_FM"]:
        for k in ("Ms", "gyro", "FM_damping", "K1", "K2", "Aex", "iDMI", "lamda100", "lamda111", "B1", "B2"):
            r[k] = 0.0
        r["anisotropy_type"] = 0
    if not r["if_AFM"]:
        for k in ("Ms_AFM1", "Ms_AFM2", "gyro_AFM1", "gyro_AFM2", "damping_AFM1", "damping_AFM2", "K1_AFM1",
                  "K1_AFM2", "Aex_AFM1", "Aex_AFM2", "J_AFM", "lamda100_AFM1", "lamda100_AFM2", "lamda111_AFM1",
                  "lamda111_AFM2", "B1_AFM1", "B1_AFM2", "B2_AFM1", "B2_AFM2"):
            r[k] = 0.0
    if not r["if_FE"]:
        for k in ("FE_mass", "FE_damping", "a1", "a11", "a12", "a111", "a112", "a123", "a1111", "a1112", "a1122",
                  "a1123", "G11", "Q11", "Q12", "Q44", "tv1", "tv2", "tv3", "f11", "f12", "f44"):
            r[k] = 0.0
    # global.cpp:360-362 divides each share by a sum that already holds the
    # shares normalised before it
    r["comp_n1"] = r["comp_n1"] / (r["comp_n1"] + r["comp_n2"] + r["comp_n3"])
    r["comp_n2"] = r["comp_n2"] / (r["comp_n1"] + r["comp_n2"] + r["comp_n3"])
    r["comp_n3"] = r["comp_n3"] / (r["comp_n1"] + r["comp_n2"] + r["comp_n3"])
    m = Material()
    for name, ctype in Material._fields_:
        if name.startswith("reserved") or name in ("uniaxis_AFM", "cijkl_glb"):
            continue
        if name in r:
            setattr(m, name, int(r[name]) if ctype is ctypes.c_int32 else float(r[name]))
    m.uniaxis_AFM[:] = [float(v) for v in r["uniaxis_AFM"]]
    C = _cijkl_glb(c11, c12, c44, Tc2g)
    for a in range(6):
        for b in range(6):
            m.cijkl_glb[a][b] = C[a][b]
    return m


# ---------------------------------------------------------------------------
# geometry

def _expand_layers(unit: list, g: dict) -> list:
    """set_geometry's / set_global's layer list (geometry.cpp:6-25,
    global.cpp:44-60): the layers from id_firstlayer to id_lastlayer_unit
    (1-based) repeat num_periods times; those before and after stay."""
    nlu, first, last_unit = g["num_layer_unit"], g["id_firstlayer"], g["id_lastlayer_unit"]
    if not (1 <= first <= last_unit <= nlu) or g["num_periods"] < 1:
        raise GoFerroError("the layer period (id_firstlayer, id_lastlayer_unit, num_periods) is inconsistent")
    n_unit = last_unit - first + 1
    id_last = first - 1 + g["num_periods"] * n_unit
    num_layer = nlu + (g["num_periods"] - 1) * n_unit
    out, j = [], first
    for i in range(num_layer):
        if i < first - 1:
            out.append(unit[i])
        elif i > id_last - 1:
            out.append(unit[i - num_layer + nlu])
        else:
            out.append(unit[j - 1])
            j += 1
            if j > last_unit:
                j = first
    return out


def layer_stack(g: dict) -> tuple[list, list]:
    """(nz per layer, material id per layer) bottom to top."""
    return _expand_layers(g["pt_nz_layer_unit"], g), _expand_layers(g["pt_material_layer_unit"], g)


def _material_cell(g: dict, directory: Path) -> np.ndarray:
    nx, ny, nz = g["nx_system"], g["ny_system"], g["nz_system"]
    if g["if_read_struct"]:
        ids = read_cells(directory / "struct.in", (nx, ny, nz), 1)[0]
        if not np.all(ids == np.round(ids)) or ids.min() < 0 or ids.max() > g["num_materials"]:
            raise GoFerroError("struct.in holds material ids outside 0..num_materials")
        return ids.astype(np.uint32)
    nzl, mats = layer_stack(g)
    nxw, nyw = g["nx_work"], g["ny_work"]
    if not (0 < nxw <= nx and 0 < nyw <= ny):
        raise GoFerroError("the work region (nx_work, ny_work) does not fit the grid")
    if sum(nzl) > nz:
        raise GoFerroError(f"the layers hold {sum(nzl)} cells in z, more than nz_system = {nz}")
    for t in mats:
        if t > g["num_materials"]:
            raise GoFerroError(f"layer material {t} exceeds num_materials = {g['num_materials']}")
    xi = (nx - nxw) // 2            # geometry.cpp:33-38, 0-based
    yi = (ny - nyw) // 2
    cell = np.zeros((nx, ny, nz), dtype=np.uint32)
    z = 0
    for h, t in zip(nzl, mats):
        cell[xi:xi + nxw, yi:yi + nyw, z:z + h] = t
        z += h
    return cell


# ---------------------------------------------------------------------------
# the deck

@dataclasses.dataclass
class Deck:
    """An input deck: the Problem the library runs, the raw settings, the
    raw material records, and the fields to upload after create."""

    problem: Problem
    settings: dict
    materials: list
    directory: Path
    steps: int
    fields: dict = dataclasses.field(default_factory=dict)       # field id -> array, uploaded after create
    skipped_outputs: list = dataclasses.field(default_factory=list)

    @property
    def any_fm(self) -> bool:
        return any(m["if_FM"] for m in self.materials)

    @property
    def any_afm(self) -> bool:
        return any(m["if_AFM"] for m in self.materials)

    @property
    def any_fe(self) -> bool:
        return any(m["if_FE"] for m in self.materials)


def _unsupported(g: dict, mats: list) -> list[str]:
    any_fe = any(m["if_FE"] for m in mats)
    out = []
    if g["if_elastostatic_1D"] or g["if_read_strain_ext"] or g["if_elastostatic_film"]:
        out.append("the 1D and film elastostatics and external strain (if_elastostatic_1D, if_elastostatic_film, "
                   "if_read_strain_ext)")
    if g["if_elastostatic"] and any_fe and g["if_flexo"]:
        out.append("if_elastostatic with the flexoelectric coupling")
    if any_fe and g["if_elec_1Dmodel"]:
        out.append("if_elec_1Dmodel (get_E_static_1D is declared but not defined in the reference)")
    return out


def read_deck(directory, allow_unsupported: bool = False) -> Deck:
    """Read system_setting.in and materials.in (and struct.in, Hext.in,
    Eext.in, Hext_nonunif.in when the deck asks for them) from `directory`.

    Raises UnsupportedInput naming every option the library does not
    implement, unless allow_unsupported (then they are left out)."""
    d = Path(directory)
    g = read_system_setting(d / "system_setting.in")
    raw = read_materials(d / "materials.in", g["num_materials"])
    Tg2c = g["Tg2c"]
    Tc2g = [[Tg2c[b][a] for b in range(3)] for a in range(3)]      # read.cpp:104-107
    mats = [_material(r, Tc2g) for r in raw]
    bad = _unsupported(g, raw)
    if bad and not allow_unsupported:
        raise UnsupportedInput("the deck asks for options this library does not implement: " + "; ".join(bad))
    nx, ny, nz = g["nx_system"], g["ny_system"], g["nz_system"]
    cell = _material_cell(g, d)
    any_fe = any(r["if_FE"] for r in raw)
    flag_of = [
        ("if_magnetostatics", FLAG_MAGNETOSTATICS), ("if_mag_1Dmodel", FLAG_MAG_1D),
        ("if_demag_factor", FLAG_DEMAG_FACTOR), ("if_prescribe_Hext", FLAG_PRESCRIBE_HEXT),
        ("if_OrientRotate", FLAG_ORIENT_ROTATE), ("if_read_Hext_nonunif", FLAG_HEXT_NONUNIF),
        ("if_EMdynamic", FLAG_EMDYNAMIC), ("if_prescribe_Eext", FLAG_PRESCRIBE_EEXT),
        ("if_elastodynamics", FLAG_ELASTODYNAMICS), ("if_elasto_backaction_from_pORm", FLAG_ELASTO_BACKACTION),
        ("if_elasto_on_pORm", FLAG_ELASTO_ON_PM), ("if_gaussian_strain_pulse", FLAG_GAUSSIAN_PULSE),
        ("if_1D_ABC", FLAG_1D_ABC), ("if_1D_ABC_ELAST_onlytop", FLAG_1D_ABC_ELAST_ONLYTOP),
        ("if_EM_fromP", FLAG_EM_FROM_P), ("if_EM_fromM", FLAG_EM_FROM_M),
        ("if_EM_backaction_on_P", FLAG_EM_BACKACTION_P), ("if_EM_backaction_on_M", FLAG_EM_BACKACTION_M),
        ("if_input_planeEM_E", FLAG_PLANE_EM), ("if_Jf_input", FLAG_JF_INPUT), ("if_PEC_XY", FLAG_PEC_XY),
        ("if_PEC_Z", FLAG_PEC_Z), ("if_1D_ABC_EM_onlytop", FLAG_1D_ABC_EM_ONLYTOP),
        ("if_spin_pumping", FLAG_SPIN_PUMPING),
    ]
    flags = 0
    for key, bit in flag_of:
        if g[key]:
            flags |= bit
    # update_prescribed_m* act inside the FM updates (magnetic_solver.cpp:
    # 1526-1609), which run only with a ferromagnet
    if g["if_prescribed_m"] and any(r["if_FM"] for r in raw):
        flags |= FLAG_PRESCRIBED_M
    # the flexoelectric terms act only with a ferroelectric (if_FE_all &&
    # if_flexo, elastic_dynamic_solver.cpp:63, main.cpp:161-166)
    if g["if_flexo"] and any_fe:
        flags |= FLAG_FLEXO
    # get_strain_static before the loop and after every step without
    # elastodynamics (main.cpp:214, 472-474), once at start-up with it
    # (main.cpp:148-158); without a magnet or ferroelectric the eigenstrain,
    # and so the static strain, is zero, and nothing is solved
    any_mag = any(r["if_FM"] or r["if_AFM"] for r in raw)
    if g["if_elastostatic"] and (any_mag or any_fe):
        flags |= FLAG_ELASTOSTATIC
    # get_E_static runs whenever there is a ferroelectric (main.cpp:120-131,
    # 214-226, 299-302)
    if any_fe and not g["if_elec_1Dmodel"]:
        flags |= FLAG_ELECTROSTATICS
    demag = (0.0, 0.0, 0.0)
    if g["if_demag_factor"]:
        s = g["demag_fac_x"] + g["demag_fac_y"] + g["demag_fac_z"]     # global.cpp:271-275
        demag = tuple(g[k] / s * -1. for k in ("demag_fac_x", "demag_fac_y", "demag_fac_z"))
    nzl, _ = layer_stack(g)
    prob = Problem(
        shape=(nx, ny, nz), spacing=(g["dx"], g["dy"], g["dz"]), materials=mats, material_cell=cell, dt=g["dt"],
        periodic=(g["periodicX"], g["periodicY"], g["periodicZ"]), flags=flags,
        Hext_stat=tuple(g["Hext_stat"]), Hext_altr=tuple(g["Hext_altr"]), H_altr_freq=tuple(g["H_altr_freq"]),
        demag_fac=demag, Tg2c=np.array(Tg2c, dtype=float),
        Eext_stat=tuple(g["Eext_stat"]), Eext_altr=tuple(g["Eext_altr"]), E_altr_freq=tuple(g["E_altr_freq"]),
        input_strain_type=g["input_strain_type"], input_strain_sourcez=g["input_strain_sourcez"],
        input_strain_component=g["input_strain_component"], num_strain_cycles=g["num_strain_cycles"],
        sigma_gauss=g["sigma_gauss"], amplitude_gauss=g["amplitude_gauss"],
        weighting_factor_fourth_Liao=g["weighting_factor_fourth_Liao"],
        planeEM_E=tuple(g["planeEM_E"]), planeEM_E_freq=tuple(g["planeEM_E_freq"]),
        planeEM_source_z=nzl[0] // 5,                                    # EMdynamic_initialize.cpp:18
        Jf_input_type=g["Jf_input_type"], Jf_input_component=g["Jf_input_component"],
        num_Jf_cycles=g["num_Jf_cycles"],
        Jfin_lo=(g["Jfin_xi"], g["Jfin_yi"], g["Jfin_zi"]), Jfin_hi=(g["Jfin_xf"], g["Jfin_yf"], g["Jfin_zf"]),
        Jf_input_amp=g["Jf_input_amp"], Jf_input_freq=g["Jf_input_freq"], Jf_input_sigma=g["Jf_input_sigma"],
        scale_elec=g["scale_elec"],
        nz_layer=None if g["if_read_struct"] else list(nzl),
        scale_elasto=g["scale_elasto"],
        precess_angle=g["precess_angle"], precess_frequency=g["precess_frequency"],
    )
    if g["if_prescribe_Hext"]:
        prob.prescribe_Hext = _table(d / "Hext.in", g["num_prescribe_Hext"])
    if g["if_prescribe_Eext"]:
        prob.prescribe_Eext = _table(d / "Eext.in", g["num_prescribe_Eext"])
    deck = Deck(problem=prob, settings=g, materials=raw, directory=d, steps=g["step"])
    if g["if_read_Hext_nonunif"]:
        deck.fields[FIELD_HEXT_NONUNIF] = read_cells(d / "Hext_nonunif.in", (nx, ny, nz), 3)
    if g["if_output_Jishe"] and not g["if_spin_pumping"]:
        deck.skipped_outputs.append("J_ISHE without spin pumping")
    if g["if_output_elastoforce"] and not g["if_elastodynamics"]:
        deck.skipped_outputs.append("elastoforce without elastodynamics (no force store is kept)")
    if g["if_output_strain"] and not g["if_elastodynamics"] and not g["if_elastostatic"]:
        # main.cpp:582-594 writes nothing then (neither branch copies a strain)
        deck.skipped_outputs.append("strain without elastodynamics or elastostatics (the reference copies none)")
    return deck


def _table(path: Path, rows: int) -> np.ndarray:
    """read_prescribe_Hext / _Eext (read.cpp:239-257): rows of (t, x, y, z)."""
    f = _Stream(_read_text(path), str(path))
    return np.array([[f.float() for _ in range(4)] for _ in range(rows)], dtype=np.float64).reshape(rows, 4)


# ---------------------------------------------------------------------------
# per-cell text files

def _lib():
    lib = load_library()
    if not getattr(lib, "_textio_ready", False):
        i64x3 = ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        for name in ("gf_write_cells", "gf_read_cells"):
            fn = getattr(lib, name)
            fn.argtypes = [ctypes.c_char_p, i64x3, ctypes.c_int32, dp]
            fn.restype = ctypes.c_int
        lib._textio_ready = True
    return lib


def write_cells(path, arrays) -> None:
    """Write ncomp arrays of one grid shape (an array (ncomp, nx, ny, nz) or
    a list of (nx, ny, nz) arrays) as the reference's per-cell text file."""
    a = np.ascontiguousarray(np.asarray(arrays, dtype=np.float64))
    if a.ndim == 3:
        a = a[None]
    if a.ndim != 4:
        raise ValueError("write_cells takes (ncomp, nx, ny, nz) values")
    dims = (ctypes.c_int64 * 3)(*a.shape[1:])
    lib = _lib()
    st = lib.gf_write_cells(str(path).encode(), dims, a.shape[0], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if st != 0:
        raise GoFerroError(lib.gf_last_error(None).decode())


def read_cells(path, dims, ncomp: int) -> np.ndarray:
    """Read a per-cell text file of grid `dims` into (ncomp, *dims)."""
    out = np.empty((ncomp,) + tuple(int(v) for v in dims), dtype=np.float64)
    cd = (ctypes.c_int64 * 3)(*out.shape[1:])
    lib = _lib()
    st = lib.gf_read_cells(str(path).encode(), cd, ncomp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if st != 0:
        raise GoFerroError(lib.gf_last_error(None).decode())
    return out


# ---------------------------------------------------------------------------
# initial state and inputs

def initial_fields(deck: Deck) -> dict:
    """The state before any input file: m along (theta_mag, phi_mag) in FM
    cells, the AFM sublattices along +/- that direction
    (magnetic_initialize.cpp:1131-1173), p = p_init in FE cells
    (ferroelectric_initialize.cpp:381-395); everything else zero."""
    g = deck.settings
    cell = np.asarray(deck.problem.material_cell)
    kind = {k: np.zeros(len(deck.materials) + 1, dtype=bool) for k in ("if_FM", "if_AFM", "if_FE")}
    for i, r in enumerate(deck.materials):
        for k in kind:
            kind[k][i + 1] = bool(r[k])
    theta = g["theta_mag"] / 180. * PI
    phi = g["phi_mag"] / 180 * PI
    mdir = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
    out = {}
    if deck.any_fm:
        fm = kind["if_FM"][cell]
        out[FIELD_M] = mdir[:, None, None, None] * fm[None]
    if deck.any_afm:
        afm = kind["if_AFM"][cell]
        out[FIELD_M_AFM1] = mdir[:, None, None, None] * afm[None]
        out[FIELD_M_AFM2] = -mdir[:, None, None, None] * afm[None]
    if deck.any_fe:
        fe = kind["if_FE"][cell]
        p0 = np.array([g["px_init"], g["py_init"], g["pz_init"]])
        out[FIELD_P] = p0[:, None, None, None] * fe[None]
    return out


def load_inputs(ctx: Context, deck: Deck, directory=None) -> list[str]:
    """Upload the initial state and the input files the deck names
    (main.cpp:44-89, inoutput.cpp:11-327); returns the files read."""
    d = Path(directory) if directory is not None else deck.directory
    g = deck.settings
    shape = deck.problem.shape
    for fid, val in initial_fields(deck).items():
        ctx.upload(fid, val)
    for fid, val in deck.fields.items():
        ctx.upload(fid, val)
    read = []

    def cells(name, ncomp, dims=shape):
        read.append(name)
        return read_cells(d / name, dims, ncomp)

    if g["if_input_m"] and deck.any_fm:
        ctx.upload(FIELD_M, cells("m.in", 3))
    if g["if_input_AFMm"] and deck.any_afm:
        a = cells("m_AFM.in", 6)
        ctx.upload(FIELD_M_AFM1, a[:3])
        ctx.upload(FIELD_M_AFM2, a[3:])
    if g["if_input_Hstat"]:
        ctx.upload(FIELD_HSTAT, cells("Hstat.in", 3))
    if g["if_input_pandq"] and deck.any_fe:
        ctx.upload(FIELD_P, cells("p.in", 3))
        ctx.upload(FIELD_Q, cells("q.in", 3))
    if g["if_input_Estat"] and deck.any_fe:
        ctx.upload(FIELD_ESTAT, cells("Estat.in", 3))
    if g["if_input_em"] and g["if_EMdynamic"]:
        ctx.upload(FIELD_E_EM, [cells(f"E{a}_Yee.in", 1, s)[0]
                                for a, s in zip("xyz", yee_shapes(shape, FIELD_E_EM))])
        ctx.upload(FIELD_H_EM, [cells(f"H{a}_Yee.in", 1, s)[0]
                                for a, s in zip("xyz", yee_shapes(shape, FIELD_H_EM))])
    if g["if_input_Jp"] and g["if_EMdynamic"]:
        ctx.upload(FIELD_JP, np.stack([cells(f"Jp{p}.in", 3) for p in (1, 2, 3)]))
    if g["if_input_uandv"] and g["if_elastodynamics"]:
        ctx.upload(FIELD_U, cells("u.in", 3))
        ctx.upload(FIELD_V, cells("v.in", 3))
    if g["if_input_elastoforce"] and g["if_elastodynamics"]:
        # input_elastoforce (inoutput.cpp:220-237): the stage-1 force that the
        # stiffness damping of the first step differences against
        ctx.upload(FIELD_ELASTOFORCE, cells("elastoforce.in", 3))
    if g["if_input_eigenstraint0_crt"] and g["if_elastodynamics"]:
        # input_eigenstraint0 (inoutput.cpp:259-280): the static eigenstrain
        # the elastic stage subtracts from that of the stores
        e = cells("eigenstraint0_crt.in", 6)
        ctx.upload(FIELD_EIGENSTRAIN_T0_DIAG, e[:3])
        ctx.upload(FIELD_EIGENSTRAIN_T0_OFF, e[3:])
    if g["if_input_straint0"] and g["if_elasto_on_pORm"]:
        # input_straint0 (inoutput.cpp:238-257) reads the global-frame static
        # strain; set-up turns it into the crystal frame the magnets read
        # (elastic_initialize.cpp:430-460), beside the dynamic strain when
        # there is one (magnetic_solver.cpp:89-96)
        e = cells("strain_stat.in", 6)
        if deck.problem.flags & FLAG_ORIENT_ROTATE:
            e = glb2crt_voigt(e, deck.problem.Tg2c)
        ctx.upload(FIELD_STRAIN_T0_DIAG, e[:3])
        ctx.upload(FIELD_STRAIN_T0_OFF, e[3:])
    return read


def glb2crt_voigt(e, Tg2c) -> np.ndarray:
    """transform_matrix_glb2crt (mathlib.cpp:78-100) of a Voigt field
    (xx yy zz yz xz xy, leading axis): out_ij = sum_k sum_l T_ik T_jl in_kl,
    summed k-major as there."""
    T = np.asarray(Tg2c, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    full = [[e[0], e[5], e[4]], [e[5], e[1], e[3]], [e[4], e[3], e[2]]]
    out = {}
    for (i, j) in ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)):
        acc = np.zeros_like(e[0])
        for k in range(3):
            for l in range(3):
                acc = acc + T[i][k] * T[j][l] * full[k][l]
        out[(i, j)] = acc
    return np.stack([out[(0, 0)], out[(1, 1)], out[(2, 2)], out[(1, 2)], out[(0, 2)], out[(0, 1)]])


# ---------------------------------------------------------------------------
# outputs

def _fmt_line(nstep: int, values, gaps) -> str:
    s = f"{nstep}"
    for v, gap in zip(values, gaps):
        s += gap + f"{v:.7e}"
    return s + "\n"


class Outputs:
    """The files main.cpp writes, under `directory`: the initial ones
    (:169-197) and those of step nstep (:486-650)."""

    def __init__(self, deck: Deck, directory=None):
        self.deck = deck
        self.g = deck.settings
        self.dir = Path(directory) if directory is not None else deck.directory
        self.dir.mkdir(parents=True, exist_ok=True)
        self.written: list[str] = []
        for what in deck.skipped_outputs:
            warnings.warn(f"output skipped: {what}")

    def _due(self, key: str, step_key: str, nstep: int) -> bool:
        return bool(self.g[key]) and self.g[step_key] > 0 and nstep % self.g[step_key] == 0

    def next_due(self, nstep: int) -> Optional[int]:
        """The first step after nstep at which anything is written."""
        best = None
        for key, step_key in self._schedule():
            if self.g[key] and self.g[step_key] > 0:
                s = (nstep // self.g[step_key] + 1) * self.g[step_key]
                best = s if best is None else min(best, s)
        return best

    def _schedule(self):
        return [("if_output_ave", "output_step_ave"), ("if_output_m", "output_step_m"),
                ("if_output_only_magcell", "output_step_magcell"), ("if_output_AFMm", "output_step_AFMm"),
                ("if_output_p", "output_step_p"), ("if_output_q", "output_step_q"),
                ("if_output_em", "output_step_em"), ("if_output_em_onecell", "output_step_em_onecell"),
                ("if_output_emYee", "output_step_emYee"), ("if_output_Estat", "output_step_Estat"),
                ("if_output_Hstat", "output_step_Hstat"), ("if_output_strain", "output_step_strain"),
                ("if_output_uandv", "output_step_uandv"), ("if_output_Jp", "output_step_Jp"),
                ("if_output_Jishe", "output_step_Jishe"), ("if_output_elastoforce", "output_step_elastoforce"),
                ("if_output_eigenstraint0_crt", "output_step_eigenstraint0_crt")]

    def _cells(self, name: str, arrays) -> None:
        write_cells(self.dir / name, arrays)
        self.written.append(name)

    def _append(self, name: str, line: str) -> None:
        with open(self.dir / name, "a") as f:
            f.write(line)
        if name not in self.written:
            self.written.append(name)

    def _averages(self, ctx: Context, nstep: int) -> None:
        deck = self.deck
        if deck.any_fm:
            self._append("average_m.dat", _fmt_line(nstep, ctx.average(FIELD_M), (" ", "  ", "  ")))
        if deck.any_afm:
            v = list(ctx.average(FIELD_M_AFM1)) + list(ctx.average(FIELD_M_AFM2))
            self._append("average_AFMm.dat", _fmt_line(nstep, v, (" ", "  ", "  ", "  ", "  ", "  ")))
        if deck.any_fe:
            self._append("average_p.dat", _fmt_line(nstep, ctx.average(FIELD_P), (" ", "  ", "  ")))

    def initial(self, ctx: Context) -> None:
        """main.cpp:169-197 (nstep = 0)."""
        g = self.g
        if g["if_output_ave"]:
            self._averages(ctx, 0)
        if g["if_output_Estat"] and self.deck.any_fe:
            self._cells("Estat.0.dat", ctx.download(FIELD_ESTAT))
        if g["if_output_Hstat"]:
            self._cells("Hstat.0.dat", ctx.download(FIELD_HSTAT))
        if g["if_elastodynamics"] and g["if_output_strain"]:
            self._cells("strain_stat.0.dat", np.zeros((6,) + tuple(self.deck.problem.shape)))
        if g["if_output_eigenstraint0_crt"]:
            self._eigenstrain_t0(ctx, 0)

    def step(self, ctx: Context, nstep: int) -> None:
        """The files of step nstep, in main.cpp's order (:601-650)."""
        g, deck = self.g, self.deck
        shape = deck.problem.shape
        if self._due("if_output_ave", "output_step_ave", nstep):
            self._averages(ctx, nstep)
        if self._due("if_output_m", "output_step_m", nstep) and deck.any_fm:
            self._cells(f"m.{nstep}.dat", ctx.download(FIELD_M))
        if self._due("if_output_only_magcell", "output_step_magcell", nstep) and deck.any_fm:
            self._magcell(ctx, nstep)
        if self._due("if_output_AFMm", "output_step_AFMm", nstep) and deck.any_afm:
            self._cells(f"m_AFM.{nstep}.dat",
                        np.concatenate([ctx.download(FIELD_M_AFM1), ctx.download(FIELD_M_AFM2)]))
        if self._due("if_output_p", "output_step_p", nstep) and deck.any_fe:
            self._cells(f"p.{nstep}.dat", ctx.download(FIELD_P))
        if self._due("if_output_q", "output_step_q", nstep) and deck.any_fe:
            self._cells(f"q.{nstep}.dat", ctx.download(FIELD_Q))
        em = g["if_EMdynamic"]
        if em and self._due("if_output_em", "output_step_em", nstep):
            self._cells(f"Eem.{nstep}.dat", ctx.download(FIELD_E_CELL))
            self._cells(f"Hem.{nstep}.dat", ctx.download(FIELD_H_CELL))
        if em and self._due("if_output_em_onecell", "output_step_em_onecell", nstep):
            idx = int(g["em_onecell_index"]) - 1
            e = ctx.download(FIELD_E_CELL).reshape(3, -1)[:, idx]
            h = ctx.download(FIELD_H_CELL).reshape(3, -1)[:, idx]
            self._append("EM_onecell.dat", _fmt_line(nstep, list(e) + list(h), (" ", "  ", "  ", " ", "  ", "  ")))
        if em and self._due("if_output_emYee", "output_step_emYee", nstep):
            for fid, lead in ((FIELD_E_EM, "E"), (FIELD_H_EM, "H")):
                for a, part in zip("xyz", split_yee(ctx.download(fid), shape, fid)):
                    self._cells(f"{lead}{a}_Yee.{nstep}.dat", part)
        if self._due("if_output_Estat", "output_step_Estat", nstep) and deck.any_fe:
            self._cells(f"Estat.{nstep}.dat", ctx.download(FIELD_ESTAT))
        if self._due("if_output_Hstat", "output_step_Hstat", nstep):
            self._cells(f"Hstat.{nstep}.dat", ctx.download(FIELD_HSTAT))
        if self._due("if_output_strain", "output_step_strain", nstep) and g["if_elastodynamics"]:
            # output_strain (inoutput.cpp:648-672): exxt0_glb + Dexx_glb, the
            # static part as read (input_straint0) or solved at start-up
            dyn = np.concatenate([ctx.download(FIELD_STRAIN_DIAG), ctx.download(FIELD_STRAIN_OFF)])
            if g["if_input_straint0"]:
                t0 = read_cells(self.deck.directory / "strain_stat.in", self.deck.problem.shape, 6)
            else:
                t0 = self._static_strain_glb(ctx)
            self._cells(f"strain.{nstep}.dat", t0 + dyn)
        elif self._due("if_output_strain", "output_step_strain", nstep) and g["if_elastostatic"]:
            # main.cpp:587-593 and output_strain (inoutput.cpp:648-672): the
            # global-frame static strain plus a dynamic strain that stays zero
            self._cells(f"strain.{nstep}.dat", self._static_strain_glb(ctx))
        if self._due("if_output_uandv", "output_step_uandv", nstep) and g["if_elastodynamics"]:
            self._cells(f"u.{nstep}.dat", ctx.download(FIELD_U))
            self._cells(f"v.{nstep}.dat", ctx.download(FIELD_V))
        if self._due("if_output_elastoforce", "output_step_elastoforce", nstep) and g["if_elastodynamics"]:
            self._cells(f"elastoforce.{nstep}.dat", ctx.download(FIELD_ELASTOFORCE))   # inoutput.cpp:742-761
        if self._due("if_output_eigenstraint0_crt", "output_step_eigenstraint0_crt", nstep):
            self._eigenstrain_t0(ctx, nstep)
        if em and self._due("if_output_Jp", "output_step_Jp", nstep):
            jp = ctx.download(FIELD_JP)
            for p in range(3):
                self._cells(f"Jp{p + 1}.{nstep}.dat", jp[p])
        if g["if_spin_pumping"] and self._due("if_output_Jishe", "output_step_Jishe", nstep):
            self._cells(f"J_ISHE.{nstep}.dat", ctx.download(FIELD_J_ISHE))   # Jx, Jy, 0 (inoutput.cpp:846-864)

    def _eigenstrain_t0(self, ctx: Context, nstep: int) -> None:
        """output_eigenstraint0_crt (inoutput.cpp:762-789): the static
        eigenstrain at t0 (with elastodynamics, as read; with elastostatics,
        that of the last static solve); otherwise the context holds none and
        the file repeats what was read (or zeros)."""
        g = self.g
        if g["if_elastodynamics"] or g["if_elastostatic"]:
            e = np.concatenate([ctx.download(FIELD_EIGENSTRAIN_T0_DIAG), ctx.download(FIELD_EIGENSTRAIN_T0_OFF)])
        elif g["if_input_eigenstraint0_crt"]:
            e = read_cells(self.deck.directory / "eigenstraint0_crt.in", self.deck.problem.shape, 6)
        else:
            e = np.zeros((6,) + tuple(self.deck.problem.shape))
        self._cells(f"eigenstraint0_crt.{nstep}.dat", e)

    def _static_strain_glb(self, ctx: Context) -> np.ndarray:
        """get_strain_static_glb (elastic_system.h:196): the crystal-frame
        static strain rotated to the global frame, e_g = Tc2g e_c Tc2g^T, in
        the order xx yy zz yz xz xy; zero when nothing is solved."""
        prob = self.deck.problem
        shape = tuple(prob.shape)
        if not prob.flags & FLAG_ELASTOSTATIC:
            return np.zeros((6,) + shape)
        d, o = ctx.download(FIELD_STRAIN_T0_DIAG), ctx.download(FIELD_STRAIN_T0_OFF)
        if not prob.flags & FLAG_ORIENT_ROTATE:
            return np.concatenate([d, o])
        e = np.empty((3, 3) + shape)
        for a in range(3):
            e[a, a] = d[a]
        e[1, 2] = e[2, 1] = o[0]
        e[0, 2] = e[2, 0] = o[1]
        e[0, 1] = e[1, 0] = o[2]
        T = np.asarray(prob.Tg2c, dtype=np.float64).T          # Tc2g
        g = np.einsum("ik,jl,kl...->ij...", T, T, e)
        return np.stack([g[0, 0], g[1, 1], g[2, 2], g[1, 2], g[0, 2], g[0, 1]])

    def _magcell(self, ctx: Context, nstep: int) -> None:
        """output_magcell (inoutput.cpp:349-366): FM cells only, in index order."""
        cell = np.asarray(self.deck.problem.material_cell).reshape(-1)
        fm = np.zeros(len(self.deck.materials) + 1, dtype=bool)
        for i, r in enumerate(self.deck.materials):
            fm[i + 1] = bool(r["if_FM"])
        ids = np.nonzero(fm[cell])[0]
        m = ctx.download(FIELD_M).reshape(3, -1)[:, ids]
        ny, nz = self.deck.problem.shape[1:]
        i, j, k = ids // (ny * nz), (ids // nz) % ny, ids % nz
        name = f"m_magcell.{nstep}.dat"
        with open(self.dir / name, "w") as f:
            for q in range(ids.size):
                f.write(f"{i[q] + 1}  {j[q] + 1}  {k[q] + 1}  {m[0, q]:.7e}  {m[1, q]:.7e}  {m[2, q]:.7e}\n")
        self.written.append(name)


def _statics_needed(deck: Deck) -> tuple[bool, bool]:
    g = deck.settings
    need_e = deck.any_fe and not g["if_elec_1Dmodel"]
    need_h = (deck.any_fm or deck.any_afm) and bool(g["if_magnetostatics"])
    return need_e, need_h


def run(directory, steps: Optional[int] = None, out_dir=None, restart: Optional[str] = None,
        checkpoint: Optional[str] = None, checkpoint_every: int = 0, progress=None) -> Context:
    """main.cpp's program on the library: read the deck, set the initial
    state and the input files, the initial statics and outputs, then the
    time loop with the outputs of every step.  `restart` names a state image
    to continue from instead of the initial state (its t and step count
    carry on); `checkpoint` is where the image is written every
    `checkpoint_every` steps and at the end.  Returns the open Context."""
    deck = read_deck(directory)
    total = deck.steps if steps is None else int(steps)
    ctx = Context(deck.problem)
    outs = Outputs(deck, out_dir)
    need_e, need_h = _statics_needed(deck)
    need_s = bool(deck.problem.flags & FLAG_ELASTOSTATIC)
    g = deck.settings
    if restart is not None:
        ctx.load_state(restart)
        nstep = int(round(ctx.time / deck.problem.dt))
    else:
        load_inputs(ctx, deck)
        nstep = 0
        # initialize_device sets the prescribed layer before the initial
        # statics (main.cpp:113, magnetic_initialize.cpp:1074-1077); a
        # restart image already holds it at its time
        if deck.problem.flags & FLAG_PRESCRIBED_M:
            ctx.prescribe_m()
        # initial statics (main.cpp:120-137): each static field is computed
        # unless it was read; gf_demag computes both, so a field that was
        # read is uploaded again afterwards
        # gf_demag also solves the static strain: at start-up with
        # elastodynamics unless a static strain or eigenstrain was read
        # (main.cpp:148-158; load_inputs uploads those first)
        if (need_e and not g["if_input_Estat"]) or (need_h and not g["if_input_Hstat"]) or need_s:
            ctx.demag()
            if need_e and g["if_input_Estat"]:
                ctx.upload(FIELD_ESTAT, read_cells(deck.directory / "Estat.in", deck.problem.shape, 3))
            if need_h and g["if_input_Hstat"]:
                ctx.upload(FIELD_HSTAT, read_cells(deck.directory / "Hstat.in", deck.problem.shape, 3))
        outs.initial(ctx)
        # statics before the loop without electrodynamics (main.cpp:214-238)
        if not g["if_EMdynamic"] and (need_e or need_h or need_s):
            ctx.demag()
    while nstep < total:
        nxt = outs.next_due(nstep)
        stop = total if nxt is None else min(nxt, total)
        if checkpoint and checkpoint_every > 0:
            stop = min(stop, (nstep // checkpoint_every + 1) * checkpoint_every)
        ctx.step(stop - nstep)
        nstep = stop
        outs.step(ctx, nstep)
        if checkpoint and checkpoint_every > 0 and nstep % checkpoint_every == 0:
            ctx.save_state(checkpoint)
        if progress is not None:
            progress(nstep, total)
    if checkpoint:
        ctx.save_state(checkpoint)
    return ctx


# ---------------------------------------------------------------------------
# state image header (gf_state_*), readable without a GPU

class _StateHeader(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_char * 8), ("version", ctypes.c_uint32), ("header_bytes", ctypes.c_uint32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("per", ctypes.c_int32 * 3), ("flags", ctypes.c_uint32), ("cell_hash", ctypes.c_uint64),
        ("time", ctypes.c_double), ("prescribe_idx", ctypes.c_int64), ("prescribe_e_idx", ctypes.c_int64),
        ("hist_head", ctypes.c_int32), ("hist_slots", ctypes.c_int32),
        ("mag_sel", ctypes.c_int32), ("afm_sel", ctypes.c_int32), ("fe_sel", ctypes.c_int32),
        ("e_sel", ctypes.c_int32), ("h_sel", ctypes.c_int32), ("bits", ctypes.c_uint32),
        ("n_materials", ctypes.c_int32), ("n_arrays", ctypes.c_int32), ("material_bytes", ctypes.c_uint64),
    ]


def _fnv(data: bytes) -> int:
    """The image checksum: FNV-1a over little-endian 64-bit words, then the
    tail bytes."""
    h = 14695981039346656037
    prime = 1099511628211
    mask = (1 << 64) - 1
    nw = len(data) // 8
    words = np.frombuffer(data, dtype="<u8", count=nw)
    for w in words.tolist():
        h = ((h ^ w) * prime) & mask
    for b in data[nw * 8:]:
        h = ((h ^ b) * prime) & mask
    return h


def state_info(path, verify: bool = True) -> dict:
    """Header, material table and array list of a state image, checked
    against its checksum when `verify` (a pure-Python pass: slow for large
    images)."""
    data = Path(path).read_bytes()
    hs = ctypes.sizeof(_StateHeader)
    if len(data) < hs + 8:
        raise GoFerroError(f"{path}: truncated state image")
    h = _StateHeader.from_buffer_copy(data[:hs])
    if bytes(h.magic) != b"GFSTATE1" or h.header_bytes != hs:
        raise GoFerroError(f"{path}: not a goferro state image of this version")
    if verify and _fnv(data[:-8]) != struct.unpack("<Q", data[-8:])[0]:
        raise GoFerroError(f"{path}: checksum mismatch")
    o = hs
    mats = []
    for _ in range(h.n_materials):
        mats.append(Material.from_buffer_copy(data[o:o + ctypes.sizeof(Material)]))
        o += ctypes.sizeof(Material)
    arrays = []
    for _ in range(h.n_arrays):
        tag, _res, count = struct.unpack("<IIq", data[o:o + 16])
        arrays.append((tag >> 8, tag & 0xFF, count))
        o += 16 + 8 * count
    if o + 8 != len(data):
        raise GoFerroError(f"{path}: array list does not match the image size")
    info = {name: getattr(h, name) for name, _ in _StateHeader._fields_ if name not in ("magic", "per")}
    info["per"] = tuple(h.per)
    info["materials"] = mats
    info["arrays"] = arrays
    return info
