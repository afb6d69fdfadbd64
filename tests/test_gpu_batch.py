"""GPU tests of the batched sweep (gf_batch): instances of uneven grid sizes
stepped together by one launch per stage and a replayed step graph.

A batch step must equal gf_step of each context on its own (the stage
arithmetic is shared; the drive's device sin() may differ from the host's
by an ulp, so the bar is 1e-13 on |m| ~ 1 rather than bit equality), and
one instance is also checked against the CPU oracle.  Graph replays are
interleaved with eager single-context work and every replay is checked.
"""
import sys

import numpy as np
import pytest

from goferro import (FIELD_HSTAT, FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FLAG_MAGNETOSTATICS, FLAG_ORIENT_ROTATE,
                     Batch, Context, GoFerroError, Material, Problem, workloads)
from goferro.workloads import GYRO

pytestmark = pytest.mark.gpu

ATOL = 1e-13


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rot():
    c, s = np.cos(0.4), np.sin(0.4)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]) @ np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def fm_instance(shape, seed):
    """FM instance without magnetostatics: two materials, DMI, rotation, a
    static field uploaded into Hstat and an AC drive."""
    rng = np.random.default_rng(seed)
    m1 = Material.make(if_FM=1, anisotropy_type=1, Ms=1.4e5, gyro=GYRO, FM_damping=0.02, K1=-2e4, K2=5e3,
                       Aex=3.7e-12, iDMI=1.5e-3)
    m2 = Material.make(if_FM=1, anisotropy_type=2, Ms=8e5, gyro=GYRO, FM_damping=0.1, K1=4e5, Aex=1.3e-11)
    cell = rng.choice(np.array([0, 1, 2], np.uint32), size=shape, p=[0.1, 0.5, 0.4])
    prob = Problem(shape=shape, spacing=(2e-9, 2.5e-9, 1.5e-9), materials=[m1, m2], material_cell=cell, dt=2e-14,
                   periodic=(True, False, True), flags=FLAG_ORIENT_ROTATE, Hext_stat=(1e4, -2e4, 5e4),
                   Hext_altr=(3e4, 0, 1e4), H_altr_freq=(5e9, 0, 2e10), Tg2c=rot())
    m = workloads.tilted_m(shape, seed, theta_max=1.0)
    hstat = rng.standard_normal((3,) + shape) * 2e4
    return prob, {FIELD_M: m, FIELD_HSTAT: hstat}


def afm_instance(shape, seed, alpha=1e-3, amp=0.05, freq=1e12):
    prob, m1, m2 = workloads.afm_block(shape, seed=seed, alpha=alpha, drive_T=amp, freq=freq)
    return prob, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}


def mixed_instances():
    """AFM instances of uneven sizes (including sizes that are not a
    multiple of any workgroup size) and two FM instances."""
    out = []
    shapes = [(8, 8, 4), (4, 4, 4), (16, 8, 4), (3, 5, 7), (8, 8, 8), (16, 16, 4)]
    for i, sh in enumerate(shapes):
        out.append(afm_instance(sh, 700 + i, alpha=10.0 ** (-4 + 0.4 * i), amp=0.01 + 0.015 * i,
                                freq=(0.5 + 0.3 * i) * 1e12))
    out.append(fm_instance((8, 6, 4), 31))
    out.append(fm_instance((5, 9, 3), 32))
    return out


def make(instances):
    ctxs = []
    for prob, fields in instances:
        c = Context(prob)
        for fid, v in fields.items():
            c.upload(fid, v)
        ctxs.append(c)
    return ctxs


def fields_of(prob):
    mats = prob.materials
    out = []
    if any(m.if_FM for m in mats):
        out.append(FIELD_M)
    if any(m.if_AFM for m in mats):
        out += [FIELD_M_AFM1, FIELD_M_AFM2]
    return out


def compare(batch_ctxs, solo_ctxs, instances, what):
    worst = 0.0
    for i, (a, b) in enumerate(zip(batch_ctxs, solo_ctxs)):
        for fid in fields_of(instances[i][0]):
            d = float(np.abs(a.download(fid) - b.download(fid)).max())
            worst = max(worst, d)
            assert d <= ATOL, (what, i, fid, d)
        assert a.time == b.time, (what, i, a.time, b.time)
    say(what, "worst", worst)
    return worst


def test_batch_equals_single_context_steps():
    say("case batch vs single contexts, mixed AFM/FM, 8 instances")
    inst = mixed_instances()
    A, B = make(inst), make(inst)
    batch = Batch(A)
    done = 0
    # 3 single-step replays, then 16+1 (the multi-step graph and a single),
    # then eager single-context steps on the batched contexts in between
    for k, n in enumerate((3, 17, 2)):
        batch.step(n)
        for c in B:
            c.step(n)
        done += n
        batch.sync()
        compare(A, B, inst, f"after {done} steps")
        if k == 1:
            for a, b in zip(A, B):
                a.step(1)
                b.step(1)
            done += 1
            compare(A, B, inst, f"after eager step {done}")
    batch.close()


@pytest.mark.parametrize("threads,cpt,cap", [(64, 1, 0), (128, 2, 7), (256, 4, 3), (64, 8, 0)])
def test_batch_launch_shapes(threads, cpt, cap):
    say("case batch launch shape", threads, cpt, cap)
    inst = mixed_instances()
    A, B = make(inst), make(inst)
    batch = Batch(A)
    batch.set_config(threads, cpt, cap)
    assert batch.config == (threads, cpt, cap)
    batch.step(5)
    for c in B:
        c.step(5)
    batch.sync()
    compare(A, B, inst, f"shape {threads}/{cap}")


def test_batch_tune_leaves_state():
    say("case batch tune")
    inst = mixed_instances()
    A, B = make(inst), make(inst)
    batch = Batch(A)
    batch.step(2)
    for c in B:
        c.step(2)
    ms = batch.tune(reps=2)
    assert ms > 0
    t, cpt, cap = batch.config
    assert t in (64, 128, 256) and cpt in (1, 2, 4, 8) and cap >= 0
    batch.step(4)
    for c in B:
        c.step(4)
    batch.sync()
    compare(A, B, inst, "after tune")


def test_batch_instance_against_oracle():
    from oracle.oracle import Oracle
    say("case batch instance vs oracle")
    inst = [afm_instance((8, 8, 4), 900, alpha=3e-3, amp=0.08, freq=1.7e12), fm_instance((6, 4, 4), 33)]
    A = make(inst)
    batch = Batch(A)
    batch.step(6)
    batch.sync()
    for (prob, fields), c in zip(inst, A):
        o = Oracle(prob)
        for fid, v in fields.items():
            o.set(fid, v)
        o.step(6)
        for fid in fields_of(prob):
            np.testing.assert_allclose(c.download(fid), o.get(fid), rtol=0, atol=1e-12)
        assert abs(c.time - o.time) <= 1e-30
        o.close()


def test_batch_sweep_instances_step():
    """A slice of the sweep workload (sizes drawn from the five sweep grids)
    against single-context stepping."""
    say("case sweep slice")
    inst = [(p, {FIELD_M_AFM1: m1, FIELD_M_AFM2: m2}) for p, m1, m2 in workloads.afm_sweep(count=6)]
    A, B = make(inst), make(inst)
    batch = Batch(A)
    batch.tune(reps=1)
    batch.step(20)
    for c in B:
        c.step(20)
    batch.sync()
    compare(A, B, inst, "sweep slice")
    m1 = A[0].download(FIELD_M_AFM1)
    assert np.isfinite(m1).all()


def test_batch_of_one_magnet_kind():
    """Batches with only FM or only AFM instances have no stage table of the
    other kind; the clock kernel must leave the missing table alone."""
    say("case fm-only and afm-only batches")
    for inst in ([fm_instance((8, 6, 4), 51), fm_instance((5, 9, 3), 52)],
                 [afm_instance((8, 8, 4), 53), afm_instance((3, 5, 7), 54)]):
        A, B = make(inst), make(inst)
        batch = Batch(A)
        batch.step(3)
        for c in B:
            c.step(3)
        batch.sync()
        compare(A, B, inst, "one kind")
        batch.close()


def test_batch_follows_uploads_made_after_create():
    """Hext_nonunif and Hstat uploaded after the batch exists, and a stage-1
    store that is not m (an uploaded M_STORE), act on the next batch step as
    on gf_step; afterwards batch and single-context steps alternate."""
    from goferro import FIELD_HEXT_NONUNIF, FIELD_M_STORE, FLAG_HEXT_NONUNIF
    say("case batch after uploads")
    inst = []
    for i, sh in enumerate([(8, 6, 4), (5, 9, 3)]):
        prob, f = fm_instance(sh, 40 + i)
        prob.flags |= FLAG_HEXT_NONUNIF
        inst.append((prob, {FIELD_M: f[FIELD_M]}))
    inst.append(afm_instance((8, 8, 4), 41))
    A, B = make(inst), make(inst)
    batch = Batch(A)
    rng = np.random.default_rng(5)
    for i in range(2):
        shape = inst[i][0].shape
        hnu = rng.standard_normal((3,) + shape) * 3e4
        hst = rng.standard_normal((3,) + shape) * 2e4
        mst = workloads.tilted_m(shape, 90 + i, theta_max=0.9)
        for c in (A[i], B[i]):
            c.upload(FIELD_HEXT_NONUNIF, hnu)
            c.upload(FIELD_HSTAT, hst)
            c.upload(FIELD_M_STORE, mst)
    batch.step(3)
    for c in B:
        c.step(3)
    batch.sync()
    compare(A, B, inst, "after uploads")
    for c in A:
        c.step(2)
    for c in B:
        c.step(2)
    batch.step(2)
    for c in B:
        c.step(2)
    batch.sync()
    compare(A, B, inst, "alternating")
    batch.close()


def test_batch_rejects_unbatchable_and_mid_step():
    say("case batch rejections")
    prob, m = workloads.fm_film((8, 8, 4))          # FFT magnetostatics on
    c = Context(prob)
    c.upload(FIELD_M, m)
    with pytest.raises(GoFerroError, match="magnetostatics"):
        Batch([c])
    c.close()
    inst = [afm_instance((4, 4, 4), 5)]
    A = make(inst)
    batch = Batch(A)
    A[0].afm_stage(1)                                # now mid-step
    with pytest.raises(GoFerroError, match="middle of a step"):
        batch.step(1)
    with pytest.raises(GoFerroError, match="middle of a step"):
        batch.tune(1)
    with pytest.raises(GoFerroError, match="threads"):
        batch.set_config(96, 1, 0)
    with pytest.raises(GoFerroError, match="cells_per_thread"):
        batch.set_config(64, 0, 0)
    for s in (2, 3, 4):
        A[0].afm_stage(s)
    batch.step(1)
    batch.sync()
    with pytest.raises(GoFerroError, match="twice"):
        Batch([A[0], A[0]])
