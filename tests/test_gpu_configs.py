"""GPU parity at the benchmark workloads' own sizes, against the oracle
(its cell loops on all host cores) or, where the oracle would take hours,
against GPU central differences and single-context stepping:

- the plain ferromagnet film 64x64x16 with FFT magnetostatics, 1000 RK4
  steps: m within 1e-9 and |m| drifting alike to 1e-12;
- the magnetoelastic film 128x128x64 (16 magnet / 48 substrate layers, Liao
  bottom, shear pulse): every stage store of u and m within 1e-12 over two
  steps, then 10 more steps within 1e-9;
- the ferroelectric THz column 8x8x1024 (200-cell PbTiO3-like film,
  electrodynamics with Mur walls, type-6 current sheet): P, Q, E and H
  within 1e-9 of their peaks after 1000 steps;
- the NiO-like antiferromagnet block 32x32x16 with its AC drive: both
  sublattices within 1e-9 after 1000 steps;
- the adjoint film 128x128x32, 2000 steps: the gradients with respect to
  alpha, Aex and B1 and 32 seeded top-layer drive cells against central
  differences of the GPU forward at 1e-6;
- the full 256-instance AFM sweep: every instance of a batched run equal to
  its own context stepped alone after 20 steps.

Each case prints its name and size before it touches the GPU.
"""
import sys
import time

import numpy as np
import pytest

from goferro import (FIELD_E_EM, FIELD_H_EM, FIELD_HDRIVE, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_M_STORE,
                     FIELD_P, FIELD_Q, FIELD_U, FIELD_V, Batch, Context, workloads)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def pair(prob, fields):
    from oracle.oracle import Oracle
    g, o = Context(prob), Oracle(prob)
    for fid, v in fields.items():
        g.upload(fid, v)
        o.set(fid, v)
    return g, o


def test_fm_film_64x64x16_1000_steps():
    shape = (64, 64, 16)
    say("case fm film", shape, "1000 steps")
    prob, m = workloads.fm_film(shape)
    g, o = pair(prob, {FIELD_M: m})
    g.demag()
    o.demag()
    t0 = time.perf_counter()
    g.step(1000)
    g.sync()
    t1 = time.perf_counter()
    o.step(1000)
    say(f"  gpu {t1 - t0:.2f} s, oracle {time.perf_counter() - t1:.2f} s")
    a, b = g.download(FIELD_M), o.get(FIELD_M)
    d = float(np.abs(a - b).max())
    na, nb = np.sqrt((a * a).sum(0)), np.sqrt((b * b).sum(0))
    dn = float(np.abs(na - nb).max())
    say("  max |dm|", d, "max ||m| - |m_oracle||", dn)
    assert d <= 1e-9
    assert dn <= 1e-12
    np.testing.assert_allclose(g.average(FIELD_M), o.average(FIELD_M), rtol=0, atol=1e-9)
    g.close()
    o.close()


def test_magnetoelastic_film_128x128x64():
    from test_gpu_elastic import run_stages
    say("case magnetoelastic film (128, 128, 64)")
    prob, m = workloads.magnetoelastic_film()
    z = np.zeros((3,) + prob.shape)
    g, o = run_stages(prob, {FIELD_M: m, FIELD_U: z, FIELD_V: z}, nsteps=2, phys=("el", "fm"),
                      check=(FIELD_M_STORE,))
    g.step(10)
    o.step(10)
    for fid in (FIELD_M, FIELD_U, FIELD_V):
        d = rel(g.download(fid), o.get(fid))
        say("  after 12 steps field", fid, "rel", d)
        assert d <= 1e-9, (fid, d)
    g.close()
    o.close()


def test_fe_thz_column_8x8x1024_1000_steps():
    say("case fe thz column (8, 8, 1024) 1000 steps")
    prob, p, q = workloads.fe_thz_column()
    g, o = pair(prob, {FIELD_P: p, FIELD_Q: q})
    g.step(1000)
    o.step(1000)
    for fid in (FIELD_P, FIELD_Q, FIELD_E_EM, FIELD_H_EM):
        d = rel(g.download(fid), o.get(fid))
        say("  field", fid, "rel", d)
        assert d <= 1e-9, (fid, d)
    # the sheet has radiated by now, so the check is not on a quiet field:
    # a current sheet K(t) radiates eta0 K / 2 each way in vacuum (less
    # where the wave has entered the er = 10 film), and the type-6 sheet's
    # K peaks where exp(-(t - 5 sigma)^2 / 2 sigma^2) sin(2 pi f t) does
    t = prob.dt * np.arange(1, 1001)
    sg, f = prob.Jf_input_sigma, prob.Jf_input_freq
    k_peak = prob.Jf_input_amp * prob.spacing[2] * np.abs(np.exp(-(t - 5 * sg) ** 2 / (2 * sg ** 2))
                                                           * np.sin(2 * np.pi * f * t)).max()
    e_free = 376.73 * k_peak / 2
    e_max = float(np.abs(g.download(FIELD_E_EM)).max())
    say("  max |E|", e_max, "free-space sheet field", e_free)
    assert 0.2 * e_free < e_max < 1.05 * e_free
    g.close()
    o.close()


def test_afm_block_1000_steps():
    say("case afm block (32, 32, 16) 1000 steps")
    prob, m1, m2 = workloads.afm_block((32, 32, 16))
    g, o = pair(prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2})
    g.step(1000)
    o.step(1000)
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
        d = float(np.abs(g.download(fid) - o.get(fid)).max())
        say("  field", fid, "max abs", d)
        assert d <= 1e-9, (fid, d)
    g.close()
    o.close()


def test_adjoint_film_128x128x32_against_gpu_differences():
    """The adjoint workload at its size and length: gradients of the probe
    spectral loss against central differences of the same GPU forward."""
    shape, N, freq = (128, 128, 32), 2000, 3e10
    say("case adjoint film", shape, N, "steps")
    prob, fields, amp = workloads.adjoint_film(shape)
    lo, hi = workloads.adjoint_probe(shape)
    g = Context(prob)

    def loss(mat=None, hd=None):
        if mat is not None:
            g.set_material(1, mat)
        for fid, v in fields.items():
            g.upload(fid, v)
        if hd is not None:
            g.upload(FIELD_HDRIVE, hd)
        g.time = 0.0
        g.demag()
        return g.adjoint_forward(1, lo, hi, freq, N)

    t0 = time.perf_counter()
    L0 = loss()
    grad, ghd = g.adjoint_backward()
    say(f"  loss {L0:.6e}, forward + reverse {time.perf_counter() - t0:.2f} s")
    base = type(prob.materials[0]).from_buffer_copy(prob.materials[0])
    for idx, name in ((0, "FM_damping"), (1, "Aex"), (2, "B1")):
        h = 1e-5 * getattr(base, name)
        up = type(base).from_buffer_copy(base)
        dn = type(base).from_buffer_copy(base)
        setattr(up, name, getattr(base, name) + h)
        setattr(dn, name, getattr(base, name) - h)
        fd = (loss(up) - loss(dn)) / (2 * h)
        say("  ", name, "adjoint", grad[idx], "fd", fd)
        np.testing.assert_allclose(grad[idx], fd, rtol=1e-6)
    g.set_material(1, base)
    rng = np.random.default_rng(4)
    cells = set()
    while len(cells) < 32:
        cells.add((int(rng.integers(shape[0])), int(rng.integers(shape[1]))))
    hd0 = fields[FIELD_HDRIVE]
    worst = 0.0
    for i, j in sorted(cells):
        # the loss is quadratic in the drive to far below 1e-6 (linear
        # response at 1 mT), so a step of twice the cell's amplitude keeps
        # the central difference exact while lifting it above the rounding
        # of the loss
        h = 2.0 * hd0[0, i, j, -1]
        up, dn = hd0.copy(), hd0.copy()
        up[0, i, j, -1] += h
        dn[0, i, j, -1] -= h
        fd = (loss(hd=up) - loss(hd=dn)) / (2 * h)
        r = abs(ghd[0, i, j, -1] - fd) / abs(fd)
        worst = max(worst, r)
        say("  drive cell", i, j, "adjoint", ghd[0, i, j, -1], "fd", fd, "rel", r)
    say("  worst drive-cell relative difference", worst)
    assert worst <= 1e-6
    g.close()


def test_full_afm_sweep_batched_equals_single_contexts():
    say("case full afm sweep, 256 instances, 20 steps")
    inst = workloads.afm_sweep(count=256)

    def make():
        out = []
        for prob, m1, m2 in inst:
            c = Context(prob)
            c.upload(FIELD_M_AFM1, m1)
            c.upload(FIELD_M_AFM2, m2)
            out.append(c)
        return out

    A, B = make(), make()
    batch = Batch(A)
    batch.tune(reps=1)
    batch.step(20)
    for c in B:
        c.step(20)
    batch.sync()
    worst = 0.0
    for a, b in zip(A, B):
        for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
            worst = max(worst, float(np.abs(a.download(fid) - b.download(fid)).max()))
        assert a.time == b.time
    say("  worst |batched - single| over 256 instances", worst)
    # the drive's device sin() may differ from the host's by an ulp
    assert worst <= 1e-13
    batch.close()
    for c in A + B:
        c.close()
