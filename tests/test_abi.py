"""The C-ABI library loads on a CPU-only host and exports exactly what
include/goferro.h declares (no compute calls: no GPU here)."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "goferro.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:gf_status|const char\*)\s+(gf_[a-z_]+)\s*\(", text, re.M)))


def test_header_declares_the_binding_list():
    import goferro
    assert declared_functions() == sorted(goferro.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    import goferro
    if not goferro.LIB_PATH.exists():
        pytest.skip("libgoferro.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", str(goferro.LIB_PATH)], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared_functions() if s not in exported]
    assert not missing, missing


def test_library_loads_and_binds_without_gpu():
    import goferro
    if not goferro.LIB_PATH.exists():
        pytest.skip("libgoferro.so not built (run __graft_entry__.build())")
    lib = goferro.load_library()
    for name in goferro.EXPORTED_SYMBOLS:
        assert getattr(lib, name) is not None
    # the create-error string is readable without a context
    assert isinstance(lib.gf_last_error(None), bytes)


def test_desc_layout_matches_header():
    """ctypes mirror of gf_desc / gf_material has the C layout (compiled probe)."""
    import ctypes

    import goferro
    probe = ROOT / "_build_probe_sizes.c"
    exe = ROOT / "_build_probe_sizes"
    probe.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "goferro.h"\n'
                     'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(gf_desc), sizeof(gf_material),'
                     ' offsetof(gf_desc, prescribe_Hext), offsetof(gf_material, B2_AFM2));return 0;}\n')
    try:
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(probe), "-o", str(exe)], check=True)
        got = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    finally:
        probe.unlink(missing_ok=True)
        exe.unlink(missing_ok=True)
    want = [ctypes.sizeof(goferro._Desc), ctypes.sizeof(goferro.Material),
            goferro._Desc.prescribe_Hext.offset, goferro.Material.B2_AFM2.offset]
    assert [int(v) for v in got] == want


def test_missing_library_fails_loudly(tmp_path):
    import goferro
    with pytest.raises(goferro.GoFerroError):
        _load_fresh(goferro, tmp_path / "nope.so")


def _load_fresh(goferro, path):
    saved = goferro._lib
    goferro._lib = None
    try:
        return goferro.load_library(path)
    finally:
        goferro._lib = saved


@pytest.mark.parametrize("case", ["periodic_z", "liao_thin", "mur_single", "vacuum_wrap"])
def test_create_rejects_grids_the_reference_cannot_run(case, oracle_lib):
    """Elastic grids on which the reference indexes outside its arrays or
    reads an undefined material (elastic_dynamic_solver.cpp:1247,
    elastodynamics_AbsorbingBoundary.cpp:92-230) are refused by gf_create
    before any device call, and by the oracle alike."""
    import numpy as np

    import goferro
    from goferro import FLAG_1D_ABC, FLAG_ELASTODYNAMICS, Material, Problem
    from goferro.workloads import set_elastic
    if not goferro.LIB_PATH.exists():
        pytest.skip("libgoferro.so not built (run __graft_entry__.build())")
    shape, periodic, flags = (4, 6, 8), (True, True, False), FLAG_ELASTODYNAMICS
    cell = np.ones(shape, np.uint32)
    if case == "periodic_z":
        periodic = (True, True, True)
    elif case == "liao_thin":
        shape = (4, 6, 4)
        cell = np.ones(shape, np.uint32)
    elif case == "mur_single":
        shape, flags = (4, 4, 1), flags | FLAG_1D_ABC
        cell = np.ones(shape, np.uint32)
    else:
        cell[1, 2, 0] = 0
    mat = set_elastic(Material.make(), 2.69e11, 1.08e11, 0.76e11, 5170.0)
    prob = Problem(shape=shape, spacing=(2e-9, 2e-9, 2e-9), materials=[mat], material_cell=cell, dt=1e-14,
                   periodic=periodic, flags=flags)
    with pytest.raises(goferro.GoFerroError, match="gf_create failed \\(-2\\)"):
        goferro.Context(prob)
    with pytest.raises(RuntimeError):
        oracle_lib.Oracle(prob)


@pytest.mark.parametrize("plant,where", [(0, None), (1, "FE stage"), (2, "FM stage"), (3, "AFM stage"),
                                         (4, "FE stage")])
def test_stage_arguments_never_alias(plant, where):
    """The stage kernels load every input of a cell before its code is
    known, so no array a launch writes may be read by it at neighbouring
    cells or as another quantity.  The library checks this before every
    stage launch; gf_selftest_stage_io runs the check over the arguments
    its builders produce for every physics/coupling/optional-field
    combination, stage and stage-1 store on host placeholder contexts (no
    device work), and shows that a planted alias is reported."""
    import ctypes

    import goferro
    if not goferro.LIB_PATH.exists():
        pytest.skip("libgoferro.so not built (run __graft_entry__.build())")
    lib = goferro.load_library()
    fn = lib.gf_selftest_stage_io
    fn.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    fn.restype = ctypes.c_int
    buf = ctypes.create_string_buffer(512)
    assert fn(plant, buf, len(buf)) == 0, buf.value.decode()
    text = buf.value.decode()
    if where is None:
        assert text.startswith("no aliasing in") and int(text.split()[3]) > 10000
    else:
        assert text.startswith(where), text


def build_binding_driver(out_dir) -> Path:
    """tests/cpp/binding_driver.cpp (a C++ caller in the pattern of the
    reference-side binding) compiled with warnings as errors against
    include/goferro.h and linked to the built library."""
    import goferro
    exe = Path(out_dir) / "binding_driver"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "cpp" / "binding_driver.cpp"), f"-L{goferro.LIB_PATH.parent}", "-lgoferro",
                    f"-Wl,-rpath,{goferro.LIB_PATH.parent}", "-o", str(exe)], check=True)
    return exe


def test_cpp_binding_compiles_and_reports_create_errors(tmp_path):
    """The C++ caller builds against the header and the library; creating a
    context either succeeds (a GPU is present) or is refused with the
    message gf_last_error(NULL) returns (no GPU here)."""
    import goferro
    if not goferro.LIB_PATH.exists():
        pytest.skip("libgoferro.so not built (run __graft_entry__.build())")
    exe = build_binding_driver(tmp_path)
    r = subprocess.run([str(exe), "probe"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("probe: context created") or "create refused: " in r.stdout
