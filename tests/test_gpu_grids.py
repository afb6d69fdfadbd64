"""GPU parity of the FFT magnetostatics on grids whose sizes are not all
powers of two (the reference plans cuFFT for any grid,
magnetic_initialize.cpp:114-115): the mixed-radix passes against the
oracle's mixed-radix transform (itself pinned against a long-double DFT and
numpy in test_oracle_fft.py), through the unfused path (nz not a power of
two), the fused stage with mixed-radix y and x passes (nz a power of two),
the AFM with magnetostatics, a 100x100x20 film over 20 steps, and the
shift and linearity properties on a large odd grid.  The FP32 adjoint
sweep, whose transforms are radix-2, refuses such grids by name.
"""
import sys

import numpy as np
import pytest

from goferro import FIELD_HSTAT, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FLAG_MAGNETOSTATICS, Context, GoFerroError, workloads
from test_gpu_parity import compare_stages, mixed_fm_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("shape,rotate", [((12, 10, 6), True), ((10, 9, 7), True), ((24, 20, 16), False),
                                          ((6, 14, 32), False), ((5, 1, 15), True)])
def test_fm_mixed_materials_general_grid(shape, rotate):
    prob, m = mixed_fm_problem(shape=shape, rotate=rotate)
    worst = compare_stages(prob, {FIELD_M: m}, nsteps=2)
    say("  worst stage store difference", worst)


def test_fm_film_general_grid_fused_and_periodic():
    """The config-1 film physics on 20x12x16: periodic in x and y, the fused
    stage (nz = 16) with P = 4 columns dividing ny = 12."""
    prob, m = workloads.fm_film((20, 12, 16))
    compare_stages(prob, {FIELD_M: m}, nsteps=3)


def test_afm_with_demag_general_grid():
    prob, m1, m2 = workloads.afm_block((10, 6, 5))
    prob.flags |= FLAG_MAGNETOSTATICS
    compare_stages(prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}, nsteps=2, afm=True, fm=False)


def test_fm_film_100x100x20_steps():
    """A deck-sized film that is not a power of two (100 = 4*5*5, 20 = 4*5):
    20 gf_steps against the oracle."""
    from oracle.oracle import Oracle
    shape = (100, 100, 20)
    say("case fm film", shape, "20 steps")
    prob, m = workloads.fm_film(shape)
    g, o = Context(prob), Oracle(prob)
    g.upload(FIELD_M, m)
    o.set(FIELD_M, m)
    g.demag()
    o.demag()
    hg, ho = g.download(FIELD_HSTAT), o.get(FIELD_HSTAT)
    assert np.abs(hg - ho).max() <= 1e-12 * np.abs(ho).max()
    g.step(20)
    o.step(20)
    d = float(np.abs(g.download(FIELD_M) - o.get(FIELD_M)).max())
    say("  max |dm| after 20 steps", d)
    assert d <= 1e-11
    g.close()
    o.close()


def test_demag_shift_and_linearity_large_odd_grid():
    """Circular convolution commutes with circular shifts and is linear, on a
    grid where the oracle is slow: 210 x 150 x 49 (2*3*5*7, 2*3*5*5, 7*7)."""
    shape = (210, 150, 49)
    say("case demag properties", shape)
    prob, m = workloads.fm_film(shape)
    g = Context(prob)
    g.upload(FIELD_M, m)
    g.demag()
    h0 = g.download(FIELD_HSTAT)
    sh = (37, -101, 5)
    g.upload(FIELD_M, np.roll(m, sh, axis=(1, 2, 3)))
    g.demag()
    h1 = g.download(FIELD_HSTAT)
    scale = np.abs(h0).max()
    assert np.abs(np.roll(h0, sh, axis=(1, 2, 3)) - h1).max() <= 1e-12 * scale
    m2 = workloads.tilted_m(shape, 77, theta_max=2.0)
    g.upload(FIELD_M, m2)
    g.demag()
    h2 = g.download(FIELD_HSTAT)
    g.upload(FIELD_M, 0.25 * m - 1.5 * m2)
    g.demag()
    h3 = g.download(FIELD_HSTAT)
    assert np.abs(h3 - (0.25 * h0 - 1.5 * h2)).max() <= 1e-12 * np.abs(h3).max()
    g.close()


def test_general_grid_rejections():
    say("case general grid rejections")
    prob, m = workloads.fm_film((8, 8, 11))
    with pytest.raises(GoFerroError, match="prime factors"):
        Context(prob)
    from goferro.workloads import adjoint_film
    prob, fields, _ = adjoint_film((12, 10, 6))
    g = Context(prob)
    for fid, v in fields.items():
        g.upload(fid, v)
    g.demag()
    lo, hi = (2, 2, 1), (5, 5, 2)
    with pytest.raises(GoFerroError, match="power-of-two"):
        g.adjoint_forward(1, lo, hi, 3e10, 4, precision=32)
    g.close()
