"""Restart through the state image, and the reference's deck driven end to
end, on the GPU through the C ABI.

A run of n1 + n2 steps and a run of n1 steps, saved, loaded into a fresh
context and continued for n2 steps must agree bit for bit (np.array_equal
on every field, and t): the image has to carry every array the next step
reads (stage stores, the elastic force store, the Liao and Mur histories,
the fused-magnetostatics H spectrum, the drive-table cursor).  The cases
cover each of those.  The deck run is checked against the oracle on the
same deck (8 written digits) and its restart against the uninterrupted run
(identical files).  Every case prints its name first.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from goferro import (FIELD_E_CELL, FIELD_E_EM, FIELD_ESTAT, FIELD_H_CELL, FIELD_H_EM, FIELD_HSTAT, FIELD_JP,
                     FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE, FIELD_P, FIELD_Q, FIELD_STRAIN_DIAG,
                     FIELD_STRAIN_OFF, FIELD_U, FIELD_V, FLAG_1D_ABC, FLAG_MAGNETOSTATICS, FLAG_PRESCRIBE_HEXT,
                     Context, GoFerroError, workloads)
from goferro import io as gio

pytestmark = pytest.mark.gpu

DECK = Path(__file__).resolve().parent / "golden" / "deck_fm"


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def fields_of(ctx, ids):
    return {f: ctx.download(f) for f in ids}


def split_run(name, prob, fields, ids, n1, n2, statics=True):
    """Continuous n1+n2 steps against n1, image, fresh context, n2."""
    say("case", name, prob.shape, hex(prob.flags))
    a = Context(prob)
    for f, v in fields.items():
        a.upload(f, v)
    if statics:
        a.demag()
    a.step(n1 + n2)
    want = fields_of(a, ids)
    t_want = a.time
    a.close()

    b = Context(prob)
    for f, v in fields.items():
        b.upload(f, v)
    if statics:
        b.demag()
    b.step(n1)
    image = b.state_image()
    b.close()
    c = Context(prob)
    c.restore_state(image)
    c.step(n2)
    got = fields_of(c, ids)
    assert c.time == t_want
    c.close()
    for f in ids:
        assert np.array_equal(got[f], want[f]), (name, f, np.abs(got[f] - want[f]).max())


def test_restart_fm_fused_magnetostatics():
    prob, m = workloads.fm_film(shape=(32, 32, 16))
    prob.Hext_altr = (4e3, 0.0, 0.0)
    prob.H_altr_freq = (7e9, 0.0, 0.0)
    split_run("fm fused", prob, {FIELD_M: m}, (FIELD_M, FIELD_HSTAT), 7, 6)


def test_restart_mixed_fm_with_drive_table():
    from test_gpu_parity import mixed_fm_problem
    prob, m = mixed_fm_problem()
    prob.flags |= FLAG_PRESCRIBE_HEXT
    prob.prescribe_Hext = np.array([[5e-14, 1e4, 0, 0], [1.5e-13, 0, 2e4, 0], [2.5e-13, 0, 0, 3e4]])
    split_run("mixed fm + table", prob, {FIELD_M: m}, (FIELD_M, FIELD_HSTAT), 6, 9)


def test_restart_uploaded_stage_store():
    # a stage store that differs from m (uploaded) travels in the image
    prob, m = workloads.fm_film(shape=(16, 16, 8))
    ms = m.copy()
    ms[0] += 1e-3
    split_run("fm store", prob, {FIELD_M: m, FIELD_M_STORE: ms}, (FIELD_M,), 0, 3, statics=False)


def test_restart_afm():
    prob, m1, m2 = workloads.afm_block(shape=(16, 16, 8))
    split_run("afm", prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}, (FIELD_M_AFM1, FIELD_M_AFM2), 5, 5, statics=False)


@pytest.mark.parametrize("mur", [False, True])
def test_restart_magnetoelastic(mur):
    from test_gpu_elastic import magnetoelastic, random_uv
    prob, m = magnetoelastic(shape=(16, 16, 12))
    if mur:
        prob.flags |= FLAG_1D_ABC
    u, v = random_uv(prob.shape, 6)
    split_run("magnetoelastic" + (" mur" if mur else " liao"), prob, {FIELD_M: m, FIELD_U: u, FIELD_V: v},
              (FIELD_M, FIELD_U, FIELD_V, FIELD_STRAIN_DIAG, FIELD_STRAIN_OFF), 6, 7, statics=False)


def test_restart_carries_the_static_strain_beside_elastodynamics():
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF
    from test_gpu_elastic import magnetoelastic, random_uv
    prob, m = magnetoelastic(shape=(8, 8, 12))
    u, v = random_uv(prob.shape, 28)
    e = 3e-5 * np.random.default_rng(29).standard_normal((6,) + prob.shape)
    f = {FIELD_M: m, FIELD_U: u, FIELD_V: v, FIELD_STRAIN_T0_DIAG: e[:3], FIELD_STRAIN_T0_OFF: e[3:]}
    split_run("magnetoelastic, static strain", prob, f, (FIELD_M, FIELD_U, FIELD_V), 3, 4, statics=False)


def test_restart_carries_the_static_eigenstrain():
    from goferro import FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF
    from test_gpu_elastic import magnetoelastic, random_e0t0, random_uv
    prob, m = magnetoelastic(shape=(8, 8, 12))
    u, v = random_uv(prob.shape, 26)
    f = {FIELD_M: m, FIELD_U: u, FIELD_V: v, **random_e0t0(prob.shape, 27)}
    split_run("magnetoelastic, static eigenstrain", prob, f,
              (FIELD_M, FIELD_U, FIELD_V, FIELD_EIGENSTRAIN_T0_DIAG, FIELD_EIGENSTRAIN_T0_OFF), 3, 4, statics=False)


def test_restart_ferroelectric_electrostatics():
    from test_gpu_estatic import block
    prob, p, q = block()
    split_run("fe estatic", prob, {FIELD_P: p, FIELD_Q: q}, (FIELD_P, FIELD_Q, FIELD_ESTAT), 4, 5)


def test_restart_em_walls_sources_drude():
    from test_gpu_em import boxed_media, random_yee
    prob = boxed_media(jf_type=6)
    E = random_yee(prob.shape, FIELD_E_EM, 1e6, 4)
    H = random_yee(prob.shape, FIELD_H_EM, 3e3, 5)
    split_run("em boxed", prob, {FIELD_E_EM: E, FIELD_H_EM: H},
              (FIELD_E_EM, FIELD_H_EM, FIELD_JP, FIELD_E_CELL, FIELD_H_CELL), 5, 6, statics=False)


def test_restart_fe_em_column():
    from test_gpu_em import fe_film
    prob, p, q = fe_film(flags=FLAG_1D_ABC)
    split_run("fe em", prob, {FIELD_P: p, FIELD_Q: q}, (FIELD_P, FIELD_Q, FIELD_E_EM, FIELD_H_EM), 4, 4,
              statics=False)


def test_state_file_and_rejections(tmp_path):
    say("case state file and rejections")
    prob, m = workloads.fm_film(shape=(16, 16, 8))
    a = Context(prob)
    a.upload(FIELD_M, m)
    a.demag()
    a.step(3)
    path = tmp_path / "state.gfs"
    a.save_state(path)
    info = gio.state_info(path)          # pure-Python checksum agrees with the library's
    assert info["time"] == a.time and (info["nx"], info["ny"], info["nz"]) == (16, 16, 8)
    assert (1, 0, 16 * 16 * 8) in info["arrays"]
    b = Context(prob)
    b.load_state(path)
    assert np.array_equal(b.download(FIELD_M), a.download(FIELD_M)) and b.time == a.time
    # another grid
    other, m2 = workloads.fm_film(shape=(16, 16, 4))
    c = Context(other)
    with pytest.raises(GoFerroError, match="another grid"):
        c.load_state(path)
    c.close()
    # a flipped byte
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 1
    with pytest.raises(GoFerroError, match="checksum"):
        b.restore_state(bytes(raw))
    with pytest.raises(GoFerroError, match="truncated|checksum"):
        b.restore_state(bytes(raw[:100]))
    # only at a step boundary
    b.time_advance()
    b.mag_stage(1)
    with pytest.raises(GoFerroError, match="step boundary"):
        b.state_image()
    a.close()
    b.close()


def test_material_update_travels_in_the_image():
    say("case material in image")
    prob, m = workloads.fm_film(shape=(16, 16, 8))
    a = Context(prob)
    a.upload(FIELD_M, m)
    mat = prob.materials[0]
    changed = type(mat).from_buffer_copy(mat)
    changed.FM_damping = 0.2
    a.set_material(1, changed)
    image = a.state_image()
    a.step(4)
    fresh = Context(workloads.fm_film(shape=(16, 16, 8))[0])
    fresh.restore_state(image)
    assert fresh.problem.materials[0].FM_damping == 0.2
    fresh.step(4)
    assert np.array_equal(fresh.download(FIELD_M), a.download(FIELD_M))
    a.close()
    fresh.close()


def test_deck_run_matches_oracle(tmp_path):
    """The golden deck through io.run against the oracle on the same deck."""
    from oracle.oracle import Oracle
    say("case deck run")
    ctx = gio.run(DECK, out_dir=tmp_path)
    deck = gio.read_deck(DECK)
    o = Oracle(deck.problem)
    o.set(FIELD_M, gio.read_cells(DECK / "m.in", deck.problem.shape, 3))
    o.demag()
    got10 = gio.read_cells(tmp_path / "m.10.dat", deck.problem.shape, 3)
    o.step(10)
    want10 = o.get(FIELD_M)
    o.step(10)
    want20 = o.get(FIELD_M)
    got20 = gio.read_cells(tmp_path / "m.20.dat", deck.problem.shape, 3)
    # 8 significant digits written; the GPU state agrees with the oracle to 1e-12
    for got, want in ((got10, want10), (got20, want20)):
        assert np.abs(got - want).max() <= 5e-8 * np.abs(want).max() + 1e-12
    np.testing.assert_allclose(ctx.download(FIELD_M), want20, rtol=0, atol=1e-12)
    ave = np.loadtxt(tmp_path / "average_m.dat")
    assert ave.shape == (5, 4) and list(ave[:, 0]) == [0, 5, 10, 15, 20]
    np.testing.assert_allclose(ave[-1, 1:], want20.sum(axis=(1, 2, 3)) / (8 * 8 * 3), rtol=0, atol=1e-7)
    assert (tmp_path / "Hstat.0.dat").exists() and (tmp_path / "Hstat.20.dat").exists()
    mag = np.loadtxt(tmp_path / "m_magcell.20.dat")
    assert mag.shape == (8 * 8 * 3, 6) and mag[:, 2].max() == 3
    ctx.close()
    o.close()


def test_deck_restart_reproduces_the_run(tmp_path):
    say("case deck restart")
    full, part = tmp_path / "full", tmp_path / "part"
    gio.run(DECK, out_dir=full).close()
    ck = tmp_path / "ck.gfs"
    gio.run(DECK, steps=10, out_dir=part, checkpoint=str(ck)).close()
    gio.run(DECK, out_dir=part, restart=str(ck)).close()
    for name in ("m.20.dat", "m_magcell.20.dat", "Hstat.20.dat"):
        assert (part / name).read_bytes() == (full / name).read_bytes(), name
    assert (part / "average_m.dat").read_text() == (full / "average_m.dat").read_text()


def test_restart_spin_pumping_flexo_elastostatics():
    """The couplings added to the stages carry through the image too."""
    from test_gpu_elastostatic import mixed
    from test_gpu_flexo import flexo_film
    from test_gpu_spin import film
    from goferro import FIELD_J_ISHE, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_EMDYNAMIC, FLAG_ELASTO_ON_PM
    prob, f = film(flags=FLAG_EMDYNAMIC | FLAG_1D_ABC)
    split_run("spin + em", prob, f, (FIELD_M, FIELD_E_EM, FIELD_H_EM, FIELD_J_ISHE), 3, 4, statics=False)
    prob, p = flexo_film()
    z = np.zeros((3,) + prob.shape)
    split_run("flexo", prob, {FIELD_P: p, FIELD_Q: z, FIELD_U: z, FIELD_V: z}, (FIELD_P, FIELD_Q, FIELD_U, FIELD_V),
              3, 3, statics=False)
    prob, f = mixed(shape=(8, 8, 16), flags=FLAG_ELASTO_ON_PM)
    split_run("elastostatics", prob, f, (FIELD_M, FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF), 2, 3)


def test_fe_em_deck_run_matches_oracle(tmp_path):
    """A ferroelectric/electrodynamics deck (the golden deck edited) through
    io.run: p, the Yee files, the one-cell EM record and the averages."""
    from oracle.oracle import Oracle
    from test_io import deck_copy
    say("case fe + em deck run")
    fe_rows = [(1, "if_FM", "false"), (1, "if_FE", "true"), (1, "FE_mass", "1e-16 4e-4"), (1, "a1", "-1.69e8"),
               (1, "a11 a12", "-7.3e7 7.5e8"), (1, "a111", "2.6e8 6.1e8 -3.7e9"), (1, "G11", "1e-10"),
               (1, "r_permittivity", "10"), (2, "r_permittivity", "4")]
    d = deck_copy(tmp_path, mat_edits=fe_rows, edits=[
        ("dt (s)", "1e-18"), ("px_init", "0 0 0.5"), ("if_magnetostatics", "false false"),
        ("if_EMdynamic", "true false"), ("if_EM_fromP", "true false"), ("if_EM_backaction_on_P", "true false"),
        ("if_input_m", "false false false"), ("if_output_m", "false 1 false 1"), ("if_output_Hstat", "false 1"),
        ("if_output_p", "true 10"), ("if_output_em ", "true 10 true 5 3"), ("if_output_emYee", "true 20"),
        ("if_output_Estat", "true 20"), ("if_Jf_input", "true"), ("Jfin_xi", "0 0 0"), ("Jfin_xf", "7 7 0"),
        ("Jf_input_type", "6 x 5e15 1e12 2e-17 1")])
    out = tmp_path / "out"
    ctx = gio.run(d, out_dir=out)
    deck = gio.read_deck(d)
    o = Oracle(deck.problem)
    for fid, v in gio.initial_fields(deck).items():
        o.set(fid, v)
    o.demag()
    o.step(20)
    want = o.get(FIELD_P)
    got = gio.read_cells(out / "p.20.dat", deck.problem.shape, 3)
    assert np.abs(got - want).max() <= 5e-8 * np.abs(want).max()
    np.testing.assert_allclose(ctx.download(FIELD_P), want, rtol=0, atol=1e-12 * np.abs(want).max())
    ex = gio.read_cells(out / "Ex_Yee.20.dat", (8, 9, 5), 1)[0]
    ex_want = np.asarray(o.get(FIELD_E_EM)[:8 * 9 * 5]).reshape(8, 9, 5)
    assert np.abs(ex - ex_want).max() <= 5e-8 * np.abs(ex_want).max() + 1e-300
    one = np.loadtxt(out / "EM_onecell.dat")
    assert one.shape == (4, 7) and list(one[:, 0]) == [5, 10, 15, 20]
    ave = np.loadtxt(out / "average_p.dat")
    assert list(ave[:, 0]) == [0, 5, 10, 15, 20]
    assert (out / "Estat.0.dat").exists() and (out / "Eem.20.dat").exists() and (out / "Hz_Yee.20.dat").exists()
    ctx.close()
    o.close()


def test_elastostatic_deck_writes_the_global_static_strain(tmp_path):
    """if_elastostatic without elastodynamics: strain.<n>.dat holds the
    static strain in the global frame (main.cpp:587-593, output_strain), the
    crystal-frame solution rotated by Tc2g = Tg2c^T."""
    from test_io import deck_copy
    from goferro import FIELD_STRAIN_T0_DIAG, FIELD_STRAIN_T0_OFF, FLAG_ELASTOSTATIC
    say("case elastostatic deck, rotated crystal")
    c, s = np.cos(0.5), np.sin(0.5)
    d = deck_copy(tmp_path, edits=[
        ("if_OrientRotate", "true"), ("Tg2c row 1", f"{c:.17e} {-s:.17e} 0"), ("Tg2c row 2", f"{s:.17e} {c:.17e} 0"),
        ("if_elastostatic if_elastostatic_1D", "true false"),
        ("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "false false true"),
        ("if_output_strain", "true 10")])
    deck = gio.read_deck(d)
    assert deck.problem.flags & FLAG_ELASTOSTATIC
    out = tmp_path / "out"
    ctx = gio.run(d, out_dir=out)
    got = gio.read_cells(out / "strain.20.dat", deck.problem.shape, 6)
    dg, og = ctx.download(FIELD_STRAIN_T0_DIAG), ctx.download(FIELD_STRAIN_T0_OFF)
    ctx.close()
    ec = np.zeros(deck.problem.shape + (3, 3))
    for a in range(3):
        ec[..., a, a] = dg[a]
    ec[..., 1, 2] = ec[..., 2, 1] = og[0]
    ec[..., 0, 2] = ec[..., 2, 0] = og[1]
    ec[..., 0, 1] = ec[..., 1, 0] = og[2]
    R = np.asarray(deck.problem.Tg2c).T
    eg = R @ ec @ R.T
    want = np.stack([eg[..., 0, 0], eg[..., 1, 1], eg[..., 2, 2], eg[..., 1, 2], eg[..., 0, 2], eg[..., 0, 1]])
    assert np.abs(want).max() > 0
    assert np.abs(got - want).max() <= 1e-7 * np.abs(want).max()
    # a rotation about z mixes xx and yy: the file is not the crystal frame
    assert np.abs(got[0] - dg[0]).max() > 1e-3 * np.abs(want).max()


def test_elastoforce_deck_run_matches_oracle(tmp_path):
    """A magnetoelastic deck with stiffness damping whose force store is read
    from elastoforce.in (main.cpp:74-76): the first stage-1 velocity update
    differences against it (elastic_dynamic_solver.cpp:1313-1323), so v and
    the store written to elastoforce.<n>.dat follow the oracle given the
    same store, and differ from a run without it."""
    from oracle.oracle import Oracle
    from test_io import deck_copy
    from goferro import FIELD_ELASTOFORCE
    say("case elastoforce deck run")
    # dv gains (eta/rho)(f - f_store) with no dt (elastic_dynamic_solver.cpp:1310-1313):
    # eta = 1e-13 Pa s keeps that jump the size of the magnetostrictive response
    damping = "0 1e-13                     elast_mass_damping elast_stiff_damping"
    d = deck_copy(tmp_path, edits=[
        ("if_elastodynamics if_elasto_backaction_from_pORm if_elasto_on_pORm", "true true true"),
        ("if_1D_ABC", "true false false 0.5"), ("if_input_elastoforce", "true"),
        ("if_output_elastoforce", "true 10"), ("if_output_uandv", "true 20")],
        mat_edits=[(1, "elast_mass_damping", damping), (2, "elast_mass_damping", damping)])
    rng = np.random.default_rng(5)
    gio.write_cells(d / "elastoforce.in", rng.standard_normal((3, 8, 8, 4)) * 1e13)
    out = tmp_path / "out"
    ctx = gio.run(d, out_dir=out)
    deck = gio.read_deck(d)
    shape = deck.problem.shape
    store0 = gio.read_cells(d / "elastoforce.in", shape, 3)

    def oracle_run(with_store):
        o = Oracle(deck.problem)
        for fid, v in gio.initial_fields(deck).items():
            o.set(fid, v)
        o.set(FIELD_M, gio.read_cells(d / "m.in", shape, 3))
        if with_store:
            o.set(FIELD_ELASTOFORCE, store0)
        o.demag()
        o.step(10)
        f10 = o.get(FIELD_ELASTOFORCE)
        o.step(10)
        res = f10, o.get(FIELD_ELASTOFORCE), o.get(FIELD_V), o.get(FIELD_U)
        o.close()
        return res

    f10, f20, v20, u20 = oracle_run(True)
    scale_v = np.abs(v20).max()
    assert scale_v > 0
    assert np.abs(ctx.download(FIELD_V) - v20).max() <= 1e-11 * scale_v
    assert np.abs(ctx.download(FIELD_ELASTOFORCE) - f20).max() <= 1e-11 * np.abs(f20).max()
    for n, want in ((10, f10), (20, f20)):
        got = gio.read_cells(out / f"elastoforce.{n}.dat", shape, 3)
        assert np.abs(got - want).max() <= 5e-8 * np.abs(want).max(), n
    got_v = gio.read_cells(out / "v.20.dat", shape, 3)
    assert np.abs(got_v - v20).max() <= 5e-8 * scale_v
    ctx.close()
    _, _, v_free, _ = oracle_run(False)
    assert np.abs(v_free - v20).max() > 1e-3 * scale_v
