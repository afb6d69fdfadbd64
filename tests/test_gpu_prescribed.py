"""GPU parity of the prescribed bottom layer (if_prescribed_m,
magnetic_solver.cpp:1526-1609, 1891-1957) against the oracle: every stage
store of a step, then many steps, on the fused FFT-magnetostatics stage and
on the unfused stage (frame rotation), and a deck with the option through
the run driver (io.run)."""
import shutil
import sys
from pathlib import Path

import numpy as np
import pytest

from goferro import (FIELD_M, FIELD_M_STORE, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE, FLAG_PRESCRIBED_M, Context,
                     io as gio, workloads)

pytestmark = pytest.mark.gpu
DECK = Path(__file__).resolve().parent / "golden" / "deck_fm"


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def problem(shape=(16, 16, 8), flags=FLAG_MAGNETOSTATICS, rotate=False):
    prob, m = workloads.fm_film(shape, seed=11)
    prob.flags = flags | FLAG_PRESCRIBED_M
    prob.precess_angle, prob.precess_frequency = 15.0, 4e10
    # a vacuum cell in the bottom layer stays untouched
    cell = prob.material_cell.copy()
    cell[3, 5, 0] = 0
    prob.material_cell = cell
    if rotate:
        c, s = np.cos(0.4), np.sin(0.4)
        prob.Tg2c = np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])
        prob.flags |= FLAG_ORIENT_ROTATE
    return prob, m


@pytest.mark.parametrize("rotate", [False, True], ids=["fused", "rotated"])
def test_prescribed_layer_matches_oracle(rotate):
    from oracle.oracle import Oracle
    prob, m = problem(rotate=rotate)
    say("case prescribed layer", prob.shape, "rotate", rotate)
    g, o = Context(prob), Oracle(prob)
    g.upload(FIELD_M, m)
    o.set(FIELD_M, m)
    g.prescribe_m()
    o.prescribe_m()
    np.testing.assert_array_equal(g.download(FIELD_M), o.get(FIELD_M))
    g.demag()
    o.demag()
    g.time_advance()
    o.time_advance()
    for s in (1, 2, 3, 4):
        g.mag_stage(s)
        o.mag_stage(s)
        g.demag()
        o.demag()
        d = float(np.abs(g.download(FIELD_M_STORE) - o.get(FIELD_M_STORE)).max())
        say("  stage", s, "store max abs", d)
        assert d <= 1e-12
    g.step(40)
    o.step(40)
    a, b = g.download(FIELD_M), o.get(FIELD_M)
    d = float(np.abs(a - b).max())
    say("  after 41 steps max abs", d)
    assert d <= 1e-11
    # the layer sits on the cone at the last step's time
    np.testing.assert_array_equal(a[:, :, :, 0], b[:, :, :, 0])
    g.close()
    o.close()


def test_prescribed_deck_run_matches_oracle(tmp_path):
    from oracle.oracle import Oracle
    deck_dir = tmp_path / "deck"
    shutil.copytree(DECK, deck_dir)
    setting = deck_dir / "system_setting.in"
    lines = setting.read_text().splitlines()
    for i, line in enumerate(lines):
        if "if_prescribed_m" in line:
            lines[i] = "true 25 3e10                 if_prescribed_m precess_angle precess_frequency"
    setting.write_text("\n".join(lines) + "\n")
    say("case prescribed deck run")
    ctx = gio.run(deck_dir, out_dir=tmp_path / "out")
    deck = gio.read_deck(deck_dir)
    assert deck.problem.flags & FLAG_PRESCRIBED_M
    o = Oracle(deck.problem)
    o.set(FIELD_M, gio.read_cells(deck_dir / "m.in", deck.problem.shape, 3))
    o.prescribe_m()
    o.demag()
    o.step(deck.steps)
    np.testing.assert_allclose(ctx.download(FIELD_M), o.get(FIELD_M), rtol=0, atol=1e-12)
    ctx.close()
    o.close()
