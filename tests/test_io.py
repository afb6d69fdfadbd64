"""The input deck, the per-cell text files and the state-image header, on
the CPU (the library's text I/O is host code; nothing here touches a GPU).

The golden deck under tests/golden/deck_fm was typed line by line from the
reference's reader (read.cpp:4-237); the values checked below come from it,
from closed forms (a cubic crystal rotated 45 degrees about z), from hand
arithmetic (the sequential comp_n normalisation of global.cpp:360-362, the
layer period of geometry.cpp:3-26) and from the magnetoelastic constants of
the magnon-phonon film (B1 = 3.5e5 J/m^3 for lambda100 = -1.4493e-6).
"""
import ctypes
import math
import random
from pathlib import Path

import numpy as np
import pytest

from goferro import (FIELD_HEXT_NONUNIF, FLAG_DEMAG_FACTOR, FLAG_ELASTO_ON_PM, FLAG_ELASTODYNAMICS,
                     FLAG_ELECTROSTATICS, FLAG_EMDYNAMIC, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE,
                     FLAG_PRESCRIBE_HEXT, FLAG_PRESCRIBED_M, FLAG_SPIN_PUMPING, FLAG_FLEXO, GoFerroError, Material,
                     workloads)
from goferro import io as gio

DECK = Path(__file__).resolve().parent / "golden" / "deck_fm"


def _line_of(lines, key, start=0):
    """Index of the first line from `start` whose trailing comment begins
    with `key` (the comments name the reference's variables)."""
    for i in range(start, len(lines)):
        parts = lines[i].split("  ")
        comment = parts[-1].strip() if len(parts) > 1 else ""
        if comment.startswith(key):
            return i
    raise KeyError(key)


def deck_copy(tmp_path, edits=(), mat_edits=()):
    """The golden deck in tmp_path.  edits: (comment key, new line) of
    system_setting.in; mat_edits: (material number, comment key, new line)
    of materials.in."""
    d = tmp_path / "deck"
    d.mkdir()
    lines = (DECK / "system_setting.in").read_text().split("\n")
    for key, text in edits:
        lines[_line_of(lines, key)] = text
    (d / "system_setting.in").write_text("\n".join(lines))
    lines = (DECK / "materials.in").read_text().split("\n")
    starts = [i for i, ln in enumerate(lines) if ln.startswith("Material ")]
    for mat, key, text in mat_edits:
        lines[_line_of(lines, key, starts[mat - 1])] = text
    (d / "materials.in").write_text("\n".join(lines))
    (d / "m.in").write_text((DECK / "m.in").read_text())
    return d


def test_golden_deck_values():
    deck = gio.read_deck(DECK)
    p = deck.problem
    assert p.shape == (8, 8, 4) and p.spacing == (2e-9, 2e-9, 2e-9) and p.dt == 1e-14
    assert p.periodic == (True, True, False)
    assert deck.steps == 20
    assert p.flags == FLAG_MAGNETOSTATICS
    assert p.Hext_stat == (0.0, 0.0, 7.9577e4) and p.Hext_altr == (2e3, 0.0, 0.0) and p.H_altr_freq == (5e9, 0, 0)
    cell = np.asarray(p.material_cell)
    assert (cell[:, :, :3] == 1).all() and (cell[:, :, 3] == 2).all()
    fm, sub = p.materials
    assert fm.if_FM == 1 and fm.Ms == 1.4e5 and fm.gyro == 2.2116812e5 and fm.FM_damping == 0.01
    assert fm.anisotropy_type == 1 and fm.K1 == -610.0 and fm.Aex == 3.7e-12
    # B1 = -3/2 lambda100 (c11 - c12) recovers the film's 3.5e5 J/m^3
    assert fm.B1 == pytest.approx(3.5e5, rel=1e-4) and fm.B2 == pytest.approx(3.5e5, rel=1e-4)
    # check_material zeroes the magnetic constants of a non-ferromagnet
    assert sub.if_FM == 0 and sub.Ms == 0.0 and sub.gyro == 0.0 and sub.Aex == 0.0 and sub.anisotropy_type == 0
    # uniaxis (3, 4, 0) normalised
    assert list(fm.uniaxis_AFM) == [0.6, 0.8, 0.0]
    # comp (1, 1, 1): 1/3, then 1 / (1/3 + 1 + 1) = 3/7, then 1 / (1/3 + 3/7 + 1) = 21/37
    assert fm.comp_n1 == pytest.approx(1 / 3, rel=1e-15)
    assert fm.comp_n2 == pytest.approx(3 / 7, rel=1e-15)
    assert fm.comp_n3 == pytest.approx(21 / 37, rel=1e-15)
    assert (sub.comp_n1, sub.comp_n2, sub.comp_n3) == (1.0, 0.0, 0.0)
    # unrotated cubic stiffness
    C = np.array([[fm.cijkl_glb[a][b] for b in range(6)] for a in range(6)])
    assert C[0, 0] == 2.69e11 and C[0, 1] == 1.08e11 and C[3, 3] == 0.76e11 and C[0, 3] == 0.0
    assert deck.settings["output_step_m"] == 10 and deck.settings["if_output_only_magcell"]
    assert deck.skipped_outputs == []


def test_initial_fields_follow_theta_phi(tmp_path):
    deck = gio.read_deck(DECK)
    m = gio.initial_fields(deck)[0]
    th, ph = math.radians(10), math.radians(30)
    want = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
    np.testing.assert_allclose(m[:, 0, 0, 0], want, rtol=0, atol=1e-15)
    assert (m[:, :, :, 3] == 0).all()      # the substrate is not magnetic


def test_layer_period_and_work_region(tmp_path):
    # 4 unit layers (nz 2, 1, 3, 1; materials 1, 2, 1, 2); layers 2..3
    # repeat 3 times: 2 | 1 3 | 1 3 | 1 3 | 1 in z, work region 4 x 2
    # centred in 8 x 6
    d = deck_copy(tmp_path, edits=[("nx_system", "8 6 16"), ("nx_work", "4 2"), ("num_layer_unit", "4"),
                                   ("num_periods", "3 2 3"), ("nz of each", "2 1 3 1"),
                                   ("material of each", "1 2 1 2")])
    g = gio.read_system_setting(d / "system_setting.in")
    nz, mats = gio.layer_stack(g)
    assert nz == [2, 1, 3, 1, 3, 1, 3, 1]
    assert mats == [1, 2, 1, 2, 1, 2, 1, 2]
    deck = gio.read_deck(d)
    cell = np.asarray(deck.problem.material_cell)
    assert cell.shape == (8, 6, 16)
    col = [1, 1, 2, 1, 1, 1, 2, 1, 1, 1, 2, 1, 1, 1, 2, 0]
    assert list(cell[2, 2]) == col and list(cell[5, 3]) == col
    assert (cell[:2] == 0).all() and (cell[6:] == 0).all() and (cell[:, :2] == 0).all() and (cell[:, 4:] == 0).all()
    assert deck.problem.planeEM_source_z == 0     # nz of the first layer (2) / 5


def test_rotated_stiffness_closed_form(tmp_path):
    s = math.sqrt(0.5)
    d = deck_copy(tmp_path, edits=[("if_OrientRotate", "true"), ("Tg2c row 1", f"{s!r} {s!r} 0"),
                                   ("Tg2c row 2", f"{-s!r} {s!r} 0"), ("Tg2c row 3", "0 0 1")])
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_ORIENT_ROTATE
    fm = deck.problem.materials[0]
    c11, c12, c44 = 2.69e11, 1.08e11, 0.76e11
    C = np.array([[fm.cijkl_glb[a][b] for b in range(6)] for a in range(6)])
    # a cubic crystal turned 45 degrees about z
    assert C[0, 0] == pytest.approx((c11 + c12 + 2 * c44) / 2, rel=1e-14)
    assert C[0, 1] == pytest.approx((c11 + c12 - 2 * c44) / 2, rel=1e-14)
    assert C[5, 5] == pytest.approx((c11 - c12) / 2, rel=1e-14)
    assert C[2, 2] == pytest.approx(c11, rel=1e-14) and C[3, 3] == pytest.approx(c44, rel=1e-14)
    # and the independent einsum form of the same rotation
    T = np.array([[s, s, 0], [-s, s, 0], [0, 0, 1]])
    np.testing.assert_allclose(C, workloads.cijkl_glb(c11, c12, c44, T), rtol=0, atol=1e-4)


def test_flags_demag_factor_and_drive_table(tmp_path):
    d = deck_copy(tmp_path, edits=[("if_prescribe_Hext", "true 3"), ("if_demag_factor", "true 1 1 2  demag factors"),
                                   ("if_read_Hext_nonunif", "true")])
    (d / "Hext.in").write_text("1e-13 1 2 3\n2e-13 4 5 6\n3e-13 7 8 9\n")
    rng = np.random.default_rng(1)
    h = rng.standard_normal((3, 8, 8, 4))
    gio.write_cells(d / "Hext_nonunif.in", h)
    deck = gio.read_deck(d)
    p = deck.problem
    assert p.flags & FLAG_DEMAG_FACTOR and p.flags & FLAG_PRESCRIBE_HEXT
    assert p.demag_fac == (-0.25, -0.25, -0.5)
    np.testing.assert_array_equal(p.prescribe_Hext, [[1e-13, 1, 2, 3], [2e-13, 4, 5, 6], [3e-13, 7, 8, 9]])
    np.testing.assert_allclose(deck.fields[FIELD_HEXT_NONUNIF], h, rtol=5e-8, atol=0)


def test_ferroelectric_sets_electrostatics_and_elastic_flags(tmp_path):
    d = deck_copy(tmp_path, edits=[("px_init", "0 0 0.5"), ("if_magnetostatics", "false false"),
                         ("if_elastodynamics", "true false true")],
                  mat_edits=[(2, "if_FE", "true"), (2, "FE_mass", "1e-16 4e-4"), (2, "a1", "-1.69e8")])
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_ELECTROSTATICS
    assert deck.problem.flags & FLAG_ELASTODYNAMICS and deck.problem.flags & FLAG_ELASTO_ON_PM
    sub = deck.problem.materials[1]
    assert sub.if_FE == 1 and sub.FE_mass == 1e-16 and sub.a1 == -1.69e8
    p = gio.initial_fields(deck)[6]
    assert (p[2, :, :, 3] == 0.5).all() and (p[:, :, :, :3] == 0).all()


def test_spin_pumping_deck(tmp_path):
    d = deck_copy(tmp_path, edits=[("if_spin_pumping", "true")],
                  mat_edits=[(2, "if_spin_pump", "true"), (2, "spin_hall_angle", "0.08 5e18 3e-9")])
    deck = gio.read_deck(d)
    p = deck.problem
    assert p.flags & FLAG_SPIN_PUMPING
    sub = p.materials[1]
    assert sub.if_spin_pump == 1 and (sub.spin_hall_angle, sub.spin_mix_cond, sub.spin_diffus_length) == (0.08, 5e18,
                                                                                                         3e-9)
    assert list(p.nz_layer) == [3, 1]


def test_flexo_deck(tmp_path):
    # flexo acts only with a ferroelectric (if_FE_all && if_flexo)
    d = deck_copy(tmp_path, edits=[("if_flexo", "true"), ("if_magnetostatics", "false false")],
                  mat_edits=[(2, "if_FE", "true"), (2, "f11 f12 f44", "1e-10 -2e-11 3e-11")])
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_FLEXO
    sub = deck.problem.materials[1]
    assert (sub.f11, sub.f12, sub.f44) == (1e-10, -2e-11, 3e-11)
    assert deck.problem.materials[0].f11 == 0.0     # not a ferroelectric: zeroed
    (tmp_path / "b").mkdir()
    d2 = deck_copy(tmp_path / "b", edits=[("if_flexo", "true")])
    assert not gio.read_deck(d2).problem.flags & FLAG_FLEXO


def test_prescribed_m_deck(tmp_path):
    d = deck_copy(tmp_path, edits=[("if_prescribed_m", "true 30 2.5e10")])
    p = gio.read_deck(d).problem
    assert p.flags & FLAG_PRESCRIBED_M
    assert (p.precess_angle, p.precess_frequency) == (30.0, 2.5e10)


@pytest.mark.parametrize("key,text,what", [
    ("if_elastostatic ", "true true", "elastostatics"),
    ("if_elastostatic ", "false true", "1D")])
def test_unsupported_options_are_named(tmp_path, key, text, what):
    d = deck_copy(tmp_path, edits=[(key, text)])
    with pytest.raises(gio.UnsupportedInput, match=what):
        gio.read_deck(d)
    gio.read_deck(d, allow_unsupported=True)


class _RecordingContext:
    """Stands in for a Context in the deck runner's host logic: records
    uploads and serves them (zeros otherwise) to the output writer."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.uploads = {}

    def upload(self, field, value):
        self.uploads[field] = np.asarray(value, dtype=np.float64)

    def download(self, field):
        return self.uploads.get(field, np.zeros((3,) + self.shape))

    def average(self, field):
        return np.zeros(3)


def test_elastoforce_input_and_output(tmp_path):
    """if_input_elastoforce reads elastoforce.in into the force store
    (main.cpp:74-76, inoutput.cpp:220-237) and if_output_elastoforce writes
    elastoforce.<n>.dat in the reference's cell format (inoutput.cpp:742-761)
    at its own interval."""
    from goferro import FIELD_ELASTOFORCE
    d = deck_copy(tmp_path, edits=[
        ("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "true true true"),
        ("if_input_elastoforce", "true"), ("if_output_elastoforce", "true 4")])
    rng = np.random.default_rng(11)
    f = rng.standard_normal((3, 8, 8, 4)) * 1e13
    gio.write_cells(d / "elastoforce.in", f)
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_ELASTODYNAMICS
    assert not any("elastoforce" in s for s in deck.skipped_outputs)
    ctx = _RecordingContext(deck.problem.shape)
    read = gio.load_inputs(ctx, deck)
    assert "elastoforce.in" in read
    np.testing.assert_array_equal(ctx.uploads[FIELD_ELASTOFORCE], gio.read_cells(d / "elastoforce.in", (8, 8, 4), 3))
    assert np.abs(ctx.uploads[FIELD_ELASTOFORCE] - f).max() <= 5e-8 * np.abs(f).max()
    out = gio.Outputs(deck, tmp_path / "out")
    assert out.next_due(0) == 4
    for n in (3, 4):
        out.step(ctx, n)
    assert not (tmp_path / "out" / "elastoforce.3.dat").exists()
    assert (tmp_path / "out" / "elastoforce.4.dat").read_text() == reference_lines(ctx.uploads[FIELD_ELASTOFORCE])
    # without elastodynamics nothing is read and the output is reported skipped
    (tmp_path / "b").mkdir()
    d2 = deck_copy(tmp_path / "b", edits=[("if_input_elastoforce", "true"), ("if_output_elastoforce", "true 4")])
    deck2 = gio.read_deck(d2)
    assert any("elastoforce" in s for s in deck2.skipped_outputs)
    assert "elastoforce.in" not in gio.load_inputs(_RecordingContext(deck2.problem.shape), deck2)


def test_eigenstrain_t0_input_and_output(tmp_path):
    """if_input_eigenstraint0_crt reads eigenstraint0_crt.in (six values a
    cell, inoutput.cpp:259-280) into the static eigenstrain, and
    if_output_eigenstraint0_crt writes it at step 0 and at its interval
    (main.cpp:196-198, 696-700; inoutput.cpp:762-789)."""
    from goferro import FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF
    d = deck_copy(tmp_path, edits=[
        ("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "true true true"),
        ("if_input_uandv", "false false true"), ("if_output_elastoforce", "false 1"),
        ("if_output_eigenstraint0_crt", "true 5")])
    rng = np.random.default_rng(12)
    e = rng.standard_normal((6, 8, 8, 4)) * 1e-5
    gio.write_cells(d / "eigenstraint0_crt.in", e)
    deck = gio.read_deck(d)
    ctx = _RecordingContext(deck.problem.shape)
    assert "eigenstraint0_crt.in" in gio.load_inputs(ctx, deck)
    back = gio.read_cells(d / "eigenstraint0_crt.in", (8, 8, 4), 6)
    np.testing.assert_array_equal(ctx.uploads[FIELD_EIGENSTRAIN_T0_DIAG], back[:3])
    np.testing.assert_array_equal(ctx.uploads[FIELD_EIGENSTRAIN_T0_OFF], back[3:])
    out = gio.Outputs(deck, tmp_path / "out")
    out.initial(ctx)
    out.step(ctx, 5)
    for n in (0, 5):
        assert (tmp_path / "out" / f"eigenstraint0_crt.{n}.dat").read_text() == reference_lines(back)
    # without elastodynamics the file repeats what was read
    (tmp_path / "b").mkdir()
    d2 = deck_copy(tmp_path / "b", edits=[("if_input_uandv", "false false true"),
                                          ("if_output_eigenstraint0_crt", "true 5")])
    gio.write_cells(d2 / "eigenstraint0_crt.in", e)
    deck2 = gio.read_deck(d2)
    ctx2 = _RecordingContext(deck2.problem.shape)
    assert "eigenstraint0_crt.in" not in gio.load_inputs(ctx2, deck2)
    out2 = gio.Outputs(deck2, tmp_path / "out2")
    out2.initial(ctx2)
    assert (tmp_path / "out2" / "eigenstraint0_crt.0.dat").read_text() == reference_lines(back)


def test_input_static_strain_rotated_and_with_elastodynamics(tmp_path):
    """input_straint0 reads the global-frame static strain; with a rotated
    crystal the upload is Tg2c e Tg2c^T (transform_matrix_glb2crt,
    mathlib.cpp:78-100), also beside elastodynamics."""
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF
    c, s = np.cos(0.6), np.sin(0.6)
    d = deck_copy(tmp_path, edits=[
        ("if_OrientRotate", "true"), ("Tg2c row 1", f"{c:.17e} {-s:.17e} 0"), ("Tg2c row 2", f"{s:.17e} {c:.17e} 0"),
        ("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "true true true"),
        ("if_input_uandv", "false true false")])
    rng = np.random.default_rng(13)
    e = rng.standard_normal((6, 8, 8, 4)) * 1e-4
    gio.write_cells(d / "strain_stat.in", e)
    deck = gio.read_deck(d)
    ctx = _RecordingContext(deck.problem.shape)
    assert "strain_stat.in" in gio.load_inputs(ctx, deck)
    back = gio.read_cells(d / "strain_stat.in", (8, 8, 4), 6)
    full = np.empty((3, 3, 8, 8, 4))
    for a in range(3):
        full[a, a] = back[a]
    full[1, 2] = full[2, 1] = back[3]
    full[0, 2] = full[2, 0] = back[4]
    full[0, 1] = full[1, 0] = back[5]
    T = np.asarray(deck.problem.Tg2c)
    want = np.einsum("ik,jl,kl...->ij...", T, T, full)
    got = np.concatenate([ctx.uploads[FIELD_STRAIN_T0_DIAG], ctx.uploads[FIELD_STRAIN_T0_OFF]])
    for q, (i, j) in enumerate(((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))):
        np.testing.assert_allclose(got[q], want[i, j], rtol=0, atol=1e-19)
    # a rotation about z mixes xx and yy
    assert np.abs(got[0] - back[0]).max() > 1e-6


def test_elastostatics_beside_elastodynamics_deck(tmp_path):
    """if_elastostatic with elastodynamics is read (the static strain solved
    at start-up, main.cpp:148-158); a deck whose only polar material is a
    ferroelectric sets the elastostatics as well; with elastodynamics the
    strain output is exxt0_glb + Dexx_glb (inoutput.cpp:648-672), the
    static part as read by input_straint0."""
    from goferro import FLAG_ELASTOSTATIC, FIELD_STRAIN_DIAG, FIELD_STRAIN_OFF
    d = deck_copy(tmp_path, edits=[
        ("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "true true true"),
        ("if_elastostatic if_elastostatic_1D", "true false"), ("if_input_uandv", "false true false"),
        ("if_output_strain", "true 5")])
    rng = np.random.default_rng(17)
    gio.write_cells(d / "strain_stat.in", rng.standard_normal((6, 8, 8, 4)) * 1e-4)
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_ELASTOSTATIC and deck.problem.flags & FLAG_ELASTODYNAMICS
    ctx = _RecordingContext(deck.problem.shape)
    assert "strain_stat.in" in gio.load_inputs(ctx, deck)
    dyn = rng.standard_normal((6, 8, 8, 4)) * 1e-5
    ctx.uploads[FIELD_STRAIN_DIAG], ctx.uploads[FIELD_STRAIN_OFF] = dyn[:3], dyn[3:]
    out = gio.Outputs(deck, tmp_path / "out")
    out.step(ctx, 5)
    t0 = gio.read_cells(d / "strain_stat.in", (8, 8, 4), 6)
    assert (tmp_path / "out" / "strain.5.dat").read_text() == reference_lines(t0 + dyn)
    (tmp_path / "b").mkdir()
    d2 = deck_copy(tmp_path / "b", edits=[("if_elastostatic if_elastostatic_1D", "true false")],
                   mat_edits=[(1, "if_FM", "false"), (1, "if_FE", "true")])
    deck2 = gio.read_deck(d2)
    assert deck2.any_fe and not deck2.any_fm
    assert deck2.problem.flags & FLAG_ELASTOSTATIC


def test_fft_electrostatics_with_fft_magnetostatics_is_read(tmp_path):
    """A multiferroic deck with both FFT statics (main.cpp:299-306) maps to
    both flags."""
    d = deck_copy(tmp_path, mat_edits=[(2, "if_FE", "true")])
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_ELECTROSTATICS
    assert deck.problem.flags & FLAG_MAGNETOSTATICS


def test_malformed_token_names_its_line(tmp_path):
    d = deck_copy(tmp_path, edits=[("periodicX", "1 1 0")])
    with pytest.raises(GoFerroError, match=r"system_setting.in, line 3: expected true or false"):
        gio.read_deck(d)


def test_comments_after_values_and_ignore(tmp_path):
    # values end where their type ends (std::istream >>); the rest of the
    # line is ignored, also text glued to the last value
    d = deck_copy(tmp_path, edits=[("dt (s)", "1e-14s  time step"), ("dx dy dz", "2e-9 2e-9 2e-9! cells")])
    deck = gio.read_deck(d)
    assert deck.problem.dt == 1e-14 and deck.problem.spacing == (2e-9, 2e-9, 2e-9)


# ---- per-cell text files ---------------------------------------------------

def reference_lines(a):
    """output_m's format string (inoutput.cpp:340-342) applied in Python."""
    out = []
    nc, nx, ny, nz = a.shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                out.append("%d  %d  %d" % (i + 1, j + 1, k + 1) + "".join("  %.7e" % a[c, i, j, k] for c in range(nc))
                           + "\n")
    return "".join(out)


@pytest.mark.parametrize("ncomp,dims", [(3, (3, 4, 5)), (1, (2, 3, 4)), (6, (1, 1, 7))])
def test_write_cells_matches_the_reference_format(tmp_path, ncomp, dims):
    rng = np.random.default_rng(ncomp)
    a = rng.standard_normal((ncomp,) + dims) * 10.0 ** rng.integers(-20, 20, size=(ncomp,) + dims)
    a.reshape(-1)[0] = -0.0
    a.reshape(-1)[-1] = 0.0
    gio.write_cells(tmp_path / "f.dat", a)
    assert (tmp_path / "f.dat").read_text() == reference_lines(a)
    back = gio.read_cells(tmp_path / "f.dat", dims, ncomp)
    np.testing.assert_array_equal(back, np.array([[float("%.7e" % v) for v in row] for row in a.reshape(ncomp, -1)])
                                  .reshape(a.shape))


def test_read_cells_any_order_and_full_precision(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((3, 4, 3, 2))
    lines = []
    for i in range(4):
        for j in range(3):
            for k in range(2):
                lines.append("%d %d %d %r %r %r\n" % (i + 1, j + 1, k + 1, *map(float, a[:, i, j, k])))
    random.Random(5).shuffle(lines)
    (tmp_path / "m.in").write_text("".join(lines))
    np.testing.assert_array_equal(gio.read_cells(tmp_path / "m.in", (4, 3, 2), 3), a)


@pytest.mark.parametrize("body,msg", [
    ("1 1 1 0 0 0\n1 1 1 0 0 0\n", "cell given twice"),
    ("1 1 1 0 0 0\n", "1 records for 2 cells"),
    ("1 1 1 0 0 0\n1 1 3 0 0 0\n", "outside the grid"),
    ("1 1 1 0 0 0\n1 1 2 0 0\n", "expected 3 values"),
    ("1 1 1 0 0 0\n1 1 2 0 0 0\n1 1 1 0 0 0\n", "more records"),
    ("1 1 1 0 0 0\n1 1.5 2 0 0 0\n", "integer cell index"),
])
def test_read_cells_rejects_malformed_files(tmp_path, body, msg):
    (tmp_path / "bad.in").write_text(body)
    with pytest.raises(GoFerroError, match=msg):
        gio.read_cells(tmp_path / "bad.in", (1, 1, 2), 3)


def test_read_cells_missing_file(tmp_path):
    with pytest.raises(GoFerroError, match="cannot open"):
        gio.read_cells(tmp_path / "none.in", (1, 1, 1), 3)


def test_golden_m_in_reads_at_full_precision():
    m = gio.read_cells(DECK / "m.in", (8, 8, 4), 3)
    np.testing.assert_allclose((m * m).sum(axis=0), 1.0, rtol=0, atol=1e-15)


# ---- state image -----------------------------------------------------------

def test_image_checksum_is_fnv1a_on_short_inputs():
    # on fewer than 8 bytes the word loop is empty and the sum is plain
    # 64-bit FNV-1a (published vectors)
    assert gio._fnv(b"") == 0xcbf29ce484222325
    assert gio._fnv(b"a") == 0xaf63dc4c8601ec8c
    assert gio._fnv(b"foobar") == 0x85944171f73967e8


def test_state_header_layout():
    # the C header is 168 bytes with no padding holes
    assert ctypes.sizeof(gio._StateHeader) == 168
    assert ctypes.sizeof(Material) % 8 == 0


def test_state_info_rejects_foreign_files(tmp_path):
    (tmp_path / "x.bin").write_bytes(b"not an image" * 20)
    with pytest.raises(GoFerroError, match="not a goferro state image"):
        gio.state_info(tmp_path / "x.bin")


class _CallLog(_RecordingContext):
    """A Context stand-in for run(): records the order of uploads, statics
    and steps (the deck runner's host logic only)."""

    def __init__(self, problem):
        super().__init__(problem.shape)
        self.calls = []
        self.time = 0.0

    def upload(self, field, value):
        super().upload(field, value)
        self.calls.append(("upload", field))

    def demag(self):
        self.calls.append(("demag",))

    def step(self, n):
        self.calls.append(("step", n))

    def prescribe_m(self):
        self.calls.append(("prescribe",))


@pytest.mark.parametrize("with_input", [False, True])
def test_run_orders_the_start_up_static_strain(tmp_path, monkeypatch, with_input):
    """With elastodynamics and if_elastostatic, run() calls gf_demag (which
    solves the static strain once) after load_inputs has uploaded any static
    strain read (main.cpp:78-80, 148-158): the upload comes first, so the
    library skips the solve; without the input the first gf_demag solves."""
    from goferro import FIELD_STRAIN_T0_DIAG
    edits = [("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "true true true"),
             ("if_elastostatic if_elastostatic_1D", "true false")]
    if with_input:
        edits.append(("if_input_uandv", "false true false"))
    d = deck_copy(tmp_path, edits=edits)
    if with_input:
        gio.write_cells(d / "strain_stat.in", np.full((6, 8, 8, 4), 1e-5))
    made = []
    monkeypatch.setattr(gio, "Context", lambda prob: made.append(_CallLog(prob)) or made[-1])
    gio.run(d, steps=2, out_dir=tmp_path / "out")
    calls = made[0].calls
    first_demag = calls.index(("demag",))
    uploads = [i for i, c in enumerate(calls) if c == ("upload", FIELD_STRAIN_T0_DIAG)]
    assert bool(uploads) == with_input
    if with_input:
        assert uploads[0] < first_demag
    assert any(c[0] == "step" for c in calls[first_demag:])
