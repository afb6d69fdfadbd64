"""Benchmark of the coupled FP64 RK4 hot path on one MI355X.

Default workload: ferromagnet LLG + RK4 with exchange, cubic anisotropy,
uniform bias and zero-padding-free FFT magnetostatics on a 256x256x128 grid
(YIG-like parameters, dt 1e-14 s), state resident in HBM.  A step is one
full RK4 step (4 fused stage kernels + 4 magnetostatic convolutions).

Prints one JSON line: throughput in cell-RK4-steps/s, the roofline of the
dominant kernel from HIP events recorded around its launches on the
library's stream, the whole-step fraction of the HBM roofline at the
algorithmic bytes per cell-step, and the CPU oracle timed on a bounded
sample of the same physics.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "go-ferro_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "coupled RK4 cell-steps/s + % HBM roofline; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(cls: str, n: int, nh: int) -> float:
    """Algorithmic bytes of one launch of a kernel class (each input read
    once, each output written once).  n cells, nh = nx*ny*(nz/2+1) complex
    points per component.

    The magnetic stage counts the stage share of the fused-RK4 definition:
    per cell 102 B of state (136 B per state component and step over four
    stages, three components), 24 B of magnetostatic field and 4 B of
    material id, 130 B in all.  With the z transforms fused into it, the
    stage reads the H z-spectrum instead of the real field and writes the M
    z-spectrum; those are FFT intermediates, outside the definition, and
    stage_bytes_with_spectra() states them separately."""
    c16 = 16.0
    if cls == "mag_stage":
        return 130.0 * n
    if cls == "fft_z_fwd":
        return 28.0 * n + 3 * nh * c16
    if cls in ("fft_y_fwd", "fft_y_inv"):
        return 6 * nh * c16
    if cls == "fft_x_conv":
        # 3 M spectra in, 3 H spectra out, 9 tensor spectra at fp64 real + fp32 imaginary
        return (6 * c16 + 9 * 12.0) * nh
    if cls == "fft_z_inv":
        return 3 * nh * c16 + 24.0 * n
    return 0.0


def stage_bytes_with_spectra(n: int, nh: int) -> float:
    """Bytes the fused magnetic stage moves if its z-spectrum input and
    output are counted once each: 106 B per cell of state and id plus the
    H z-spectrum read and the M z-spectrum written (3 components each)."""
    return 106.0 * n + 6 * nh * 16.0


def measured_traffic(kernel_cls: str):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (already scaled by the
    1 GiB copy calibration), or None when the summary has no such row."""
    names = {"mag_stage": "mag_stage_fused_kernel<true", "fft_z_fwd": "zfwd_kernel<false>", "fft_y_fwd": "yfwd_kernel",
             "fft_x_conv": "xconv_kernel<true>", "fft_y_inv": "yinv_kernel", "fft_z_inv": "zinv_kernel"}
    path = ROOT / "profiles" / "fm256_demag_kernel_stats.md"
    if kernel_cls not in names or not path.exists():
        return None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0].strip("`").startswith(names[kernel_cls]):
            return float(cells[6]) * 1e6
    return None


def b_alg_fm(shape) -> float:
    """Algorithmic FP64 bytes per cell-step of FM LLG + FFT magnetostatics:
    fused-RK4 state traffic 136*3, magnetostatic field reads 4*24, material
    id 4*4, and per convolution 24 in + 24 out + 4 id + 6 real tensor
    spectra at 48*(nz/2+1)/nz."""
    nz = shape[2]
    conv = 24 + 24 + 4 + 48.0 * (nz // 2 + 1) / nz
    return 136 * 3 + 4 * 24 + 4 * 4 + 4 * conv


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(shape, seconds_hint: float = 12.0, max_steps: int = 8) -> dict:
    """The oracle (the C port of the reference algorithm, its cell and
    FFT-line loops on all host cores with OpenMP) on the bench workload
    itself: the same grid and physics, a few RK4 steps after one untimed
    step.  Returns the rate with the thread count, nproc and CPU model."""
    from goferro import FIELD_M, workloads
    from oracle.oracle import Oracle, threads

    prob, m = workloads.fm_film(shape)
    t_set = time.perf_counter()
    o = Oracle(prob)
    o.set(FIELD_M, m)
    o.demag()
    o.step(1)
    t_set = time.perf_counter() - t_set
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds_hint or steps >= max_steps:
            break
    o.close()
    n = shape[0] * shape[1] * shape[2]
    nthr = threads()
    model = cpu_model()
    print(f"bench: cpu baseline {nthr} threads (nproc {os.cpu_count()}, {model}): {steps} steps in {el:.1f} s "
          f"(+{t_set:.1f} s set-up and one untimed step)", file=sys.stderr, flush=True)
    return {"value": n * steps / el, "unit": "cell-steps/s", "cores": nthr, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": model,
            "sample": f"FM LLG + FFT magnetostatics {shape[0]}x{shape[1]}x{shape[2]} (the bench workload), "
                      f"{steps} RK4 steps in {el:.1f} s, oracle/gf_oracle.c with OpenMP on {nthr} threads"}


def bench_batch(count: int, steps_batch: int, steps_solo: int) -> dict:
    """Batched sweep: `count` NiO-like AFM instances of grids drawn from
    16^3 .. 64x64x32, one launch per stage for all of them with each RK4 step
    replayed from a HIP graph (launch shape autotuned), against the same
    contexts stepped one by one with gf_step.  Rates are cell-steps/s over
    all instances."""
    import goferro
    from goferro import FIELD_M_AFM1, FIELD_M_AFM2, workloads

    t_setup = time.perf_counter()
    inst = workloads.afm_sweep(count=count)
    ctxs = []
    for prob, m1, m2 in inst:
        c = goferro.Context(prob)
        c.upload(FIELD_M_AFM1, m1)
        c.upload(FIELD_M_AFM2, m2)
        ctxs.append(c)
    cells = sum(c.n for c in ctxs)
    batch = goferro.Batch(ctxs)
    tune_ms = batch.tune(reps=3)
    threads, cpt, cap = batch.config
    print(f"bench: batch of {count} instances, {cells} cells, set-up {time.perf_counter() - t_setup:.1f} s, "
          f"tuned to {threads} threads x {cpt} cells / grid cap {cap}", file=sys.stderr, flush=True)
    batch.step(4)
    batch.sync()
    t0 = time.perf_counter()
    batch.step(steps_batch)
    batch.sync()
    el_b = time.perf_counter() - t0
    for c in ctxs:
        c.step(1)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(steps_solo)
    for c in ctxs:
        c.sync()
    el_s = time.perf_counter() - t0
    # one simulation after the other, each run to completion before the next
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(steps_solo)
        c.sync()
    el_q = time.perf_counter() - t0
    rate_b = cells * steps_batch / el_b
    rate_s = cells * steps_solo / el_s
    rate_q = cells * steps_solo / el_q
    b_alg = 136.0 * 6 + 16          # AFM: two sublattice state vectors + material id per cell-step
    batch.close()
    for c in ctxs:
        c.close()
    print(f"bench: batched {1e3 * el_b / steps_batch:.3f} ms/step, one by one on own streams "
          f"{1e3 * el_s / steps_solo:.3f} ms/step, one after the other {1e3 * el_q / steps_solo:.3f} ms/step",
          file=sys.stderr, flush=True)
    return {"workload": f"{count} AFM instances (NiO-like, grids 16^3..64x64x32, sweep seed 5), dt 1e-15 s",
            "cells": cells, "unit": "cell-steps/s",
            "batched": {"value": rate_b, "ms_per_step": 1e3 * el_b / steps_batch, "steps": steps_batch,
                        "threads": threads, "cells_per_thread": cpt, "grid_cap": cap, "tuned_stage_ms": tune_ms,
                        "achieved_GBs": b_alg * rate_b / 1e9, "hbm_frac": b_alg * rate_b / (HBM_PEAK_GBS * 1e9)},
            "one_by_one": {"value": rate_s, "ms_per_step": 1e3 * el_s / steps_solo, "steps": steps_solo,
                           "note": "gf_step per context, each context on its own HIP stream (overlapping)"},
            "one_after_another": {"value": rate_q, "ms_per_step": 1e3 * el_q / steps_solo, "steps": steps_solo,
                                  "note": "gf_step per context, each run to completion before the next"},
            "speedup": rate_b / rate_s, "speedup_vs_one_after_another": rate_b / rate_q}


def bench_adjoint(shape, nsteps: int) -> dict:
    """Adjoint sweep of the probe spectral loss (30 GHz, 16x16x4 probe) on
    the FM film with a shaped top-layer drive and static strain: wall time
    of the reverse sweep (segment recomputation included) against the plain
    forward (gf_step) of the same steps, plus the checkpointed forward."""
    import goferro
    from goferro import workloads

    prob, fields, _ = workloads.adjoint_film(shape)
    lo, hi = workloads.adjoint_probe(shape)
    ctx = goferro.Context(prob)

    def reset():
        for fid, v in fields.items():
            ctx.upload(fid, v)
        ctx.time = 0.0
        ctx.demag()
        ctx.sync()

    reset()
    ctx.step(2)                      # warm-up (kernels, allocations)
    ctx.sync()
    reset()
    t0 = time.perf_counter()
    ctx.step(nsteps)
    ctx.sync()
    t_plain = time.perf_counter() - t0
    reset()
    t0 = time.perf_counter()
    L = ctx.adjoint_forward(1, lo, hi, 3e10, nsteps)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad, ghd = ctx.adjoint_backward()
    t_bwd = time.perf_counter() - t0
    # the same sweep with the stage stores and fields of as many steps as fit
    # in 80% of the free HBM kept by the forward pass (all of them when they
    # fit), so that the reverse sweep recomputes none of those steps
    import ctypes

    import numpy as np
    n = shape[0] * shape[1] * shape[2]
    per_step = 8 * 3 * n * 8
    hip = ctypes.CDLL("libamdhip64.so")
    free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free_b), ctypes.byref(total_b)):
        raise RuntimeError("hipMemGetInfo failed")
    free = free_b.value
    mib = int(-(-min(nsteps * per_step, 0.8 * free) // 1048576))
    resident = {"resident_mib": mib, "steps_resident": min(nsteps, (mib * 1048576) // per_step)}
    L_r = None
    if mib > 0:
        reset()
        t0 = time.perf_counter()
        try:
            L_r = ctx.adjoint_forward(1, lo, hi, 3e10, nsteps, resident_mib=mib)
        except goferro.GoFerroError as e:   # the budget could not be allocated after all
            resident["error"] = str(e)
        t_fwd_r = time.perf_counter() - t0
    if L_r is not None:
        t0 = time.perf_counter()
        grad_r, ghd_r = ctx.adjoint_backward()
        t_bwd_r = time.perf_counter() - t0
        ctx.adjoint_forward(1, lo, hi, 3e10, 1)   # frees the resident buffers
        resident.update({"forward_s": t_fwd_r, "reverse_s": t_bwd_r,
                         "adjoint_over_forward": t_bwd_r / t_plain,
                         "gradient_over_forward": (t_fwd_r + t_bwd_r) / t_plain,
                         "identical_to_checkpointed": bool(L_r == L and np.array_equal(grad_r, grad)
                                                           and np.array_equal(ghd_r, ghd))})
        print(f"bench: adjoint with {resident['steps_resident']} resident steps ({mib} MiB): forward "
              f"{t_fwd_r:.3f} s, reverse {t_bwd_r:.3f} s", file=sys.stderr, flush=True)
    # the FP32 variant of the same sweep (its state is a copy; the context is
    # reset anyway so that it starts from the same m and time)
    reset()
    ctx.adjoint_forward(1, lo, hi, 3e10, 4, precision=32)   # warm-up
    ctx.adjoint_backward()
    reset()
    t0 = time.perf_counter()
    L32 = ctx.adjoint_forward(1, lo, hi, 3e10, nsteps, precision=32)
    t_fwd32 = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad32, ghd32 = ctx.adjoint_backward()
    t_bwd32 = time.perf_counter() - t0
    ctx.close()
    print(f"bench: adjoint {shape}, {nsteps} steps: plain forward {t_plain:.3f} s, checkpointed forward "
          f"{t_fwd:.3f} s, reverse sweep {t_bwd:.3f} s; FP32 forward {t_fwd32:.3f} s, reverse {t_bwd32:.3f} s",
          file=sys.stderr, flush=True)
    fp32 = {"dtype": "f32", "checkpointed_forward_s": t_fwd32, "reverse_s": t_bwd32,
            "reverse_over_forward": t_bwd32 / t_fwd32, "forward_cell_steps_per_s": n * nsteps / t_fwd32,
            "speedup_vs_f64": {"forward": t_fwd / t_fwd32, "reverse": t_bwd / t_bwd32},
            "loss": L32, "grad": {"alpha": grad32[0], "Aex": grad32[1], "B1": grad32[2]},
            "rel_to_f64": {"loss": abs(L32 - L) / abs(L),
                           "alpha": abs(grad32[0] - grad[0]) / abs(grad[0]),
                           "Aex": abs(grad32[1] - grad[1]) / abs(grad[1]),
                           "B1": abs(grad32[2] - grad[2]) / abs(grad[2]),
                           "drive": float(((ghd32 - ghd) ** 2).sum() ** 0.5 / (ghd ** 2).sum() ** 0.5)}}
    return {"workload": f"FM film {shape[0]}x{shape[1]}x{shape[2]} + shaped top-layer drive (30 GHz) + static strain, "
                        f"probe 16x16x4 spectral power at 30 GHz, {nsteps} steps, sqrt(N) checkpoints "
                        f"(and, under 'resident', the stage data of every step that fits kept in HBM)",
            "dtype": "f64", "nsteps": nsteps, "forward_s": t_plain, "checkpointed_forward_s": t_fwd,
            "reverse_s": t_bwd, "adjoint_over_forward": t_bwd / t_plain,
            "gradient_over_forward": (t_fwd + t_bwd) / t_plain,
            "ratio_note": "adjoint_over_forward = reverse sweep / plain forward; gradient_over_forward = "
                          "(checkpointed forward + reverse sweep) / plain forward, the whole gradient's cost",
            "forward_cell_steps_per_s": n * nsteps / t_plain, "loss": L,
            "grad": {"alpha": grad[0], "Aex": grad[1], "B1": grad[2]},
            "grad_drive_norm": float((ghd ** 2).sum() ** 0.5),
            "gradient_over_forward_resident": resident.get("gradient_over_forward"),
            "resident": resident, "fp32": fp32}


def bench_coupled(which: str, steps: int) -> dict:
    """The coupled workloads of the other physics rows, timed on the GPU with
    per-kernel-class HIP-event times: 'magnetoelastic' (128x128x64 magnet on
    substrate, elastodynamics both ways, Liao bottom, shear pulse) and
    'fe_thz' (8x8x1024 ferroelectric film + electrodynamics, Mur walls,
    THz current sheet).  Per-class achieved bandwidth uses the compulsory
    state bytes of each class per launch (each state array read and written
    once per stage as in the fused RK4 accounting)."""
    import goferro
    from goferro import FIELD_M, FIELD_P, FIELD_Q, workloads

    if which == "magnetoelastic":
        prob, m = workloads.magnetoelastic_film()
        ctx = goferro.Context(prob)
        ctx.upload(FIELD_M, m)
        name = "magnetoelastic film 128x128x64 (16 magnet / 48 substrate), dt 5e-15 s"
    else:
        prob, p, q = workloads.fe_thz_column()
        ctx = goferro.Context(prob)
        ctx.upload(FIELD_P, p)
        ctx.upload(FIELD_Q, q)
        name = "FE + FDTD column 8x8x1024 (vacuum / PbTiO3-like 200 cells / vacuum), dt 1e-18 s"
    n = prob.n
    ctx.step(3)
    ctx.sync()
    ctx.set_kernel_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    el = time.perf_counter() - t0
    kt = ctx.kernel_times()
    ctx.set_kernel_timing(False)
    t1 = time.perf_counter()
    ctx.step(steps)             # untimed per kernel: the event bracketing adds launch overhead
    ctx.sync()
    el_plain = time.perf_counter() - t1
    ctx.close()
    # compulsory bytes per stage of the state each class owns (fused RK4: 17 double
    # accesses per state component per step = 4.25 per stage) plus its cross reads.
    # The elastic class is its two launches (strain, eigenstrain and stress parts; then
    # the stress divergence and updates): u, v state, the strain written for the
    # magnets, the m store read for the eigenstrain.  The EM class is one launch
    # for Ampere, Faraday and the ferroelectric stage (gf_step runs them together).
    if which == "magnetoelastic":
        nmag = n * 16 // 64
        per = {"elastic_stage": n * (6 * 4.25 * 8 + 4 + 6 * 8) + nmag * 3 * 8,
               "mag_stage": nmag * (3 * 4.25 * 8 + 6 * 8 + 4 + 24)}
    else:
        nfe = n * 200 // 1024
        ne = 6.77 * n
        per = {"em_e": ne * 4.25 * 8 + 3 * 8 * n + nfe * (6 * 4.25 * 8 + 3 * 8 + 4)}
    kernels = {}
    for cls, (ms, cnt) in kt.items():
        if cnt:
            b = per.get(cls)
            kernels[cls] = {"launches": cnt, "avg_us": 1e3 * ms / cnt,
                            "achieved_GBs": (b / (ms / cnt * 1e-3) / 1e9) if b else None}
    print(f"bench: {which}: {1e3 * el_plain / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return {"workload": name, "cells": n, "steps": steps, "ms_per_step": 1e3 * el_plain / steps,
            "value": n * steps / el_plain, "unit": "cell-steps/s", "kernels": kernels,
            "ms_per_step_with_kernel_events": 1e3 * el / steps}


def copy_probe(nbytes: int) -> int:
    """One HIP device-to-device copy of nbytes (runtime blit kernel): the
    reference point that scales FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed for the copy probe")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", default="256x256x128")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=256,
                    help="instances of the batched AFM sweep measured after the main workload (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--solo-steps", type=int, default=10)
    ap.add_argument("--coupled-steps", type=int, default=200,
                    help="steps of the magnetoelastic and FE+FDTD measurements (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="steps of the adjoint measurement on the 128x128x32 film (0 = skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import goferro
    from goferro import FIELD_M, workloads

    shape = tuple(int(v) for v in args.grid.split("x"))
    n = shape[0] * shape[1] * shape[2]
    nh = shape[0] * shape[1] * (shape[2] // 2 + 1)
    prob, m = workloads.fm_film(shape)
    t_setup = time.perf_counter()
    ctx = goferro.Context(prob)
    ctx.upload(FIELD_M, m)
    ctx.demag()
    ctx.sync()
    t_setup = time.perf_counter() - t_setup
    print(f"bench: setup {t_setup:.1f} s for {args.grid}", file=sys.stderr, flush=True)

    ctx.step(args.warmup)
    ctx.sync()
    # the timed steps run without per-kernel events (they add launch
    # overhead); the kernel roofline comes from a second, event-timed run of
    # the same number of steps
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    el = time.perf_counter() - t0
    print(f"bench: {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)
    kt = {}
    if not args.no_kernel_timing:
        ctx.set_kernel_timing(True)
        ctx.sync()
        t1 = time.perf_counter()
        ctx.step(args.steps)
        ctx.sync()
        el_ev = time.perf_counter() - t1
        kt = ctx.kernel_times()
        ctx.set_kernel_timing(False)
        print(f"bench: {args.steps} steps with per-kernel events in {el_ev:.3f} s", file=sys.stderr, flush=True)

    value = n * args.steps / el
    balg = b_alg_fm(shape)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FM LLG+RK4, exchange + cubic anisotropy + bias + FFT magnetostatics, "
                               f"{args.grid}, dx 2 nm, YIG-like, dt 1e-14 s",
                   "grid": list(shape), "cells": n},
        "step_roofline": {"bytes_per_cell_step": balg,
                          "achieved_GBs": balg * value / 1e9,
                          "frac": balg * value / (HBM_PEAK_GBS * 1e9)},
    }
    if kt:
        per = {}
        fused = kt.get("fft_z_fwd", (0.0, 0))[1] == 0
        for cls, (ms, cnt) in kt.items():
            if cnt:
                b = kernel_bytes(cls, n, nh)
                per[cls] = {"launches": cnt, "avg_us": 1e3 * ms / cnt,
                            "share": ms / max(1e-12, sum(v[0] for v in kt.values())),
                            "achieved_GBs": (b / (ms / cnt * 1e-3)) / 1e9 if b else None}
        dom = max(per, key=lambda k: per[k]["avg_us"] * per[k]["launches"])
        d = per[dom]
        traffic = measured_traffic(dom) if shape == (256, 256, 128) else None
        out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": d["achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                           "bytes_per_launch": kernel_bytes(dom, n, nh), "avg_launch_us": d["avg_us"]}
        if dom == "mag_stage" and fused:
            bs = stage_bytes_with_spectra(n, nh)
            out["roofline"]["with_z_spectra"] = {
                "bytes_per_launch": bs, "achieved": bs / (d["avg_us"] * 1e-6) / 1e9,
                "frac": bs / (d["avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS,
                "note": "the fused stage also reads the H z-spectrum and writes the M z-spectrum "
                        "(FFT intermediates, not in the algorithmic bytes)"}
        out["kernels"] = per
        out["ms_per_step_with_kernel_events"] = 1e3 * el_ev / args.steps
    ctx.close()
    if args.copy_probe_mib:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe_mib << 20)
    if args.batch_instances > 0:
        out["batched_sweep"] = bench_batch(args.batch_instances, args.batch_steps, args.solo_steps)
    if args.adjoint_steps > 0:
        out["adjoint"] = bench_adjoint((128, 128, 32), args.adjoint_steps)
    if args.coupled_steps > 0:
        out["magnetoelastic"] = bench_coupled("magnetoelastic", args.coupled_steps)
        out["fe_thz"] = bench_coupled("fe_thz", 10 * args.coupled_steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(shape)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
