"""The C++ caller of tests/cpp/binding_driver.cpp, which drives the library
as the reference-side binding does (descriptor from plain arrays, upload
of three component arrays, main.cpp's loop body as member calls
gf_time_advance / gf_mag_stage / gf_demag, then gf_average and
gf_download), against gf_step of the same film through ctypes: bit for
bit, since both run the same kernels in the same order."""
import subprocess
import sys

import numpy as np
import pytest

from goferro import FIELD_M, Context, workloads
from test_abi import build_binding_driver

pytestmark = pytest.mark.gpu


def test_cpp_binding_steps_like_gf_step(tmp_path):
    shape, steps = (12, 8, 16), 5
    print("case c++ binding driver", shape, steps, file=sys.stderr, flush=True)
    prob, m = workloads.fm_film(shape, seed=6)
    mat = prob.materials[0]
    params = [*shape, steps, prob.spacing[0], prob.dt, mat.Ms, mat.gyro, mat.FM_damping, mat.K1, mat.Aex,
              prob.Hext_stat[2]]
    (tmp_path / "params.txt").write_text(" ".join(repr(float(v)) if isinstance(v, float) else str(v) for v in params))
    np.ascontiguousarray(m, dtype="<f8").tofile(tmp_path / "m_in.bin")
    exe = build_binding_driver(tmp_path)
    r = subprocess.run([str(exe), "run", str(tmp_path / "params.txt"), str(tmp_path / "m_in.bin"),
                        str(tmp_path / "m_out.bin")], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "m_out.bin", dtype="<f8").reshape((3,) + shape)
    g = Context(prob)
    g.upload(FIELD_M, m)
    g.demag()
    g.step(steps)
    ref = g.download(FIELD_M)
    avg = g.average(FIELD_M)
    g.close()
    assert np.array_equal(got, ref)
    printed = [float(v) for v in r.stdout.split()[2:5]]
    assert printed == list(avg)
