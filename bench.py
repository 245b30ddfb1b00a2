#!/usr/bin/env python3
"""Benchmark of the force path on one MI355X.

Workload (the GPU configuration): N = 128^3 = 2,097,152 particles on a
lattice displaced by Gaussian noise of 0.2 cell (seed 1), Plummer softening
eps = 1/(50*128), GM = 1/N.  One step is one all-pairs FP32 compensated kick
(N^2 = 4.398e12 interactions, with the half-kick/rung epilogue) over inputs
already resident in HBM.  The same run times the 3-D complex FP32 FFT
(256^3 round trip of that configuration and a 512^3 forward transform), the
drift, the host upload and download, PM alone and with the learned
correction (also on a held-out seed), PM + the direct near field in three
settings, rung-subset kicks, the clustered precision study, training steps
of the learned correction, and the stepping configuration (2^24 particles,
10 KDK steps) with each force method, every force judged against the FP64
direct sum; and the CPU baseline on the host's cores.

Prints one JSON line.  value = interactions/s of the kick.
"""
from __future__ import annotations

import argparse
import subprocess
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ct2_amd.timing import HipEvents, time_calls  # noqa: E402

METRIC = "kick interactions/s + 3D FFT GB/s; step time PM+MLP vs direct at ≤1e-3 force err"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOP_PER_INTERACTION = 20
# HBM bytes per kick_sym_pairs launch (64 per kick) at N = 128^3 from rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs scaled by a plain 1 GiB copy (profiles/kick_sym_n128cubed.md); None at other sizes
KICK_TRAFFIC_BYTES = {128: 1.82e9}
# FP32 flop per unordered pair evaluation counted by SQ_INSTS_VALU_FLOPS_FP32 over kick_sym_pairs,
# and VALU instructions per pair evaluation (profiles/kick_sym_valu_counters_n128cubed.md)
KICK_HW_FLOP_PER_EVAL = 26.19
KICK_VALU_PER_EVAL = 19.07


def rel_err(F, J):
    return np.linalg.norm(F - J, axis=0) / np.linalg.norm(J, axis=0)


def pm_leg(api, ev: HipEvents, n: int, side: int, kick_ms: float, idx, J, ng: int, reps: int) -> dict:
    """Times the PM force and the PM + learned-correction force of the current
    particles and measures both against the periodic FP64 judge."""
    from ct2_amd import _lib
    from ct2_amd.train_mlp import DEFAULT_OUT, load_weights

    lib = _lib.load()
    trained = load_weights(DEFAULT_OUT)
    meta = trained[1] if trained is not None else {}
    order, rs = int(meta.get("order", 1)), float(meta.get("rs_cells", 1.25))  # the mesh the model corrects
    api.pm_assignment(order)
    api.pm_init(ng, rs)
    out = {"pm_grid": ng, "assignment": {1: "cic", 2: "tsc"}[order], "rs_cells": rs,
           "direct_kick_ms": round(kick_ms, 3)}
    _lib.check(lib.ct2_pm_forces())
    pm_ms = float(np.median(time_calls(ev, lambda: _lib.check(lib.ct2_pm_forces()), reps)))
    fx, fy, fz, _ = api.forces_download()
    e_pm = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J[:3])
    out.update({"pm_ms": round(pm_ms, 3), "pm_median_rel_err": float(np.median(e_pm)),
                "pm_p90_rel_err": float(np.percentile(e_pm, 90))})
    if trained is None:
        out["pm_mlp"] = "no trained weights (run ct2_amd/train_mlp.py)"
        api.pm_assignment(1)
        return out
    w = trained[0]
    api.mlp_init(1)
    api.mlp_set_weights(w)
    _lib.check(lib.ct2_pm_mlp_forces())
    ms = float(np.median(time_calls(ev, lambda: _lib.check(lib.ct2_pm_mlp_forces()), reps)))
    fx, fy, fz, _ = api.forces_download()
    e_ml = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J[:3])
    out.update({"pm_mlp_ms": round(ms, 3), "pm_mlp_median_rel_err": float(np.median(e_ml)),
                "pm_mlp_p90_rel_err": float(np.percentile(e_ml, 90)),
                "speedup_pm_mlp_vs_direct": round(kick_ms / ms, 1),
                "mlp_flops_per_particle": 189440, "mlp_tflops": round(189440 * n / (ms - pm_ms) / 1e9, 2)
                if ms > pm_ms else None,
                "target_median_1e-3_met": bool(np.median(e_ml) <= 1e-3),
                "model": {k: meta.get(k) for k in ("ng", "rs_cells", "order", "kinds", "seeds", "samples_per_set",
                                                   "steps", "batch")}})
    api.pm_assignment(1)
    return out


def near_leg(api, ev: HipEvents, side: int, kick_ms: float, idx, J, reps: int, ng: int, rs: float,
             rcut_rs: float, order: int) -> dict:
    """PM + direct near field (ct2's kick() structure: neighbour-cell pairs
    directly, the rest from the mesh) timed and judged like the PM leg."""
    from ct2_amd import _lib
    lib = _lib.load()
    api.pm_assignment(order)
    api.pm_init(ng, rs)
    api.pm_near(rcut_rs)
    _lib.check(lib.ct2_pm_forces())
    ms = float(np.median(time_calls(ev, lambda: _lib.check(lib.ct2_pm_forces()), reps)))
    fx, fy, fz, pot = api.forces_download()
    e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J[:3])
    # the near field's work: pairs closer than rcut, N^2 (4/3) pi rcut^3 on a near-uniform set
    rcut = rcut_rs * rs / ng
    n = side ** 3
    pairs = float(n) * n * 4.0 / 3.0 * np.pi * rcut ** 3
    out = {"assignment": {1: "cic", 2: "tsc"}[order], "pm_grid": ng, "rs_cells": rs, "rcut_rs": rcut_rs,
           "cells_per_side": api.pm_near_cells(), "ms": round(ms, 3), "near_pairs_within_rcut": pairs,
           "near_pairs_per_s_incl_mesh": round(pairs / (ms * 1e-3), 1), "median_rel_err": float(np.median(e)),
           "p90_rel_err": float(np.percentile(e, 90)), "p99_rel_err": float(np.percentile(e, 99)),
           "pot_max_rel_err": float(np.max(np.abs(pot[idx] - J[3])) / np.abs(J[3]).max()),
           "speedup_vs_direct": round(kick_ms / ms, 1),
           "target_median_1e-3_and_10x_met": bool(np.median(e) <= 1e-3 and kick_ms / ms >= 10.0)}
    api.pm_near(0.0)
    api.pm_assignment(1)
    return out


def mm_leg(api, ev: HipEvents, n: int, kick_ms: float, idx, J, reps: int, order: int) -> dict:
    """ct2's multipole mesh (near-set direct pairs + Cartesian expansions of
    the given order through FFTs on about 8 particles per cell), timed and
    judged like the PM legs; init_s is the one-time table build."""
    nc = api.mm_cells_for(n)
    t0 = time.perf_counter()
    api.mm_init(nc, order)
    init_s = time.perf_counter() - t0
    from ct2_amd import _lib
    lib = _lib.load()
    _lib.check(lib.ct2_mm_forces())
    ms = float(np.median(time_calls(ev, lambda: _lib.check(lib.ct2_mm_forces()), reps)))
    fx, fy, fz, pot = api.forces_download()
    e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J[:3])
    d = pot[idx] - J[3]
    return {"cells": nc, "order": order, "components": api.mm_ncomp(order), "ms": round(ms, 3),
            "init_s": round(init_s, 2), "median_rel_err": float(np.median(e)),
            "p90_rel_err": float(np.percentile(e, 90)), "p99_rel_err": float(np.percentile(e, 99)),
            # the judge softens every pair's nearest image, the mesh only the direct pairs:
            # a near-constant offset of the potential (GM sum eps^2 / 2r^3 over far pairs)
            "pot_offset_gm": float(d.mean() * n), "pot_spread_rel": float(d.std() / np.abs(J[3]).mean()),
            "speedup_vs_direct": round(kick_ms / ms, 1),
            "target_median_1e-3_and_10x_met": bool(np.median(e) <= 1e-3 and kick_ms / ms >= 10.0)}


def stepping_leg(api, ev: HipEvents, rate: float, side: int, steps: int, judge: int, method: str) -> dict:
    """The stepping configuration: side^3 particles from a Gaussian random
    field, 512^3 grid, KDK leapfrog a = 0.02 -> 0.03 with FP64 factors, the
    periodic FP64 direct sum judging `judge` particles each step.  method
    "pm_mlp": PM (the assignment and split the model was trained on: TSC,
    1.0 cells) + the learned correction trained on this configuration's
    distribution (the force the configuration names);
    "pm_near": TSC PM (split 2 cells) + the near field to 5.5 split radii;
    "mm": ct2's multipole mesh (8 particles per cell, order 6).
    The direct kick at this N is not run (N^2 / rate at the measured rate
    would take minutes per step); its time is derived."""
    from ct2_amd import stepping
    from ct2_amd.train_mlp import load_weights
    kw = dict(steps=steps, judge=judge, timer=ev.time, log=lambda m: print(m, file=sys.stderr, flush=True))
    if method == "pm_mlp":
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ct2_amd", "weights", "mlp_grf_ng2x.npz")
        trained = load_weights(path)
        if trained is None:
            return {"method": method, "skipped": f"no trained weights at {path}"}
        meta = trained[1]
        rec = stepping.run(side, 2 * side, method="pm_mlp", weights=trained[0],
                           rs_cells=float(meta.get("rs_cells", 1.25)), near=0.0, order=int(meta.get("order", 1)),
                           meta=meta, **kw)
    elif method == "mm":
        rec = stepping.run(side, 2 * side, method="mm", mm_order=6, **kw)
    else:
        rec = stepping.run(side, 2 * side, method="pm", rs_cells=2.0, near=5.5, order=2, **kw)
    n = side ** 3
    ms = [r["force_kick_ms"] for r in rec["records"]][1:]  # the first call includes one-time set-up
    med = [r["median_rel_err"] for r in rec["records"]]
    direct_ms = n * float(n) / rate * 1e3
    out = {"method": method, "n": n, "pm_grid": rec["ng"], "assignment": rec["assignment"],
           **({"mm_cells": rec["mm_cells"], "mm_order": rec["mm_order"]} if method == "mm" else {}),
           "rs_cells": rec["rs_cells"], "rcut_rs": rec["near_rcut_rs"], "steps": steps,
           "a": [rec["a_start"], rec["a_end"]], "force_kick_ms_median": round(float(np.median(ms)), 2),
           "drift_ms_median": round(float(np.median([r["drift_ms"] for r in rec["records"][:-1]])), 3),
           "median_rel_err_per_step": [float(f"{v:.3e}") for v in med],
           "p90_rel_err_max": float(max(r["p90_rel_err"] for r in rec["records"])),
           "judge_sinks_per_step": judge, "direct_kick_ms_derived": round(direct_ms, 1),
           "speedup_vs_direct": round(direct_ms / float(np.median(ms)), 1),
           "target_median_1e-3_and_10x_met": bool(max(med) <= 1e-3 and direct_ms / float(np.median(ms)) >= 10)}
    if method == "pm_mlp":
        out["pm_alone_median_rel_err_per_step"] = [float(f"{r['no_mlp_median_rel_err']:.3e}") for r in rec["records"]]
    return out


def fft_leg(api, ev: HipEvents, reps: int) -> dict:
    """FP32 3-D FFT: 256^3 round trip (96 n^3 B) and 512^3 forward (48 n^3 B)
    on device-resident random data, median of `reps` HIP-event timings."""
    fft = {}
    for nf, mode in ((256, "round_trip"), (512, "forward")):
        buf = api.DeviceBuffer(nf ** 3 * 8)
        rng = np.random.default_rng(nf)
        buf.upload((rng.standard_normal(2 * nf ** 3, dtype=np.float32)).view(np.complex64))
        if mode == "round_trip":
            def op():
                api.fft3d_device(buf, nf, 1)
                api.fft3d_device(buf, nf, -1)
            alg = 96 * nf ** 3
        else:
            def op():
                api.fft3d_device(buf, nf, 1)
            alg = 48 * nf ** 3
        op()
        ms = float(np.median(time_calls(ev, op, reps)))
        fft[f"{nf}^3_{mode}"] = {"ms": round(ms, 4), "GB/s": round(alg / ms / 1e6, 1),
                                 "frac_hbm": round(alg / ms / 1e6 / HBM_PEAK_GBS, 4),
                                 "compulsory_GB/s": round(16 * nf ** 3 * (2 if mode == "round_trip" else 1) / ms / 1e6, 1)}
        buf.close()
        print(f"[bench] fft {nf}^3 {mode}: {ms:.3f} ms", file=sys.stderr, flush=True)
    return fft


def train_leg(api, ev: HipEvents, n: int) -> list:
    """Training throughput of the learned correction: Adam steps of the
    library's own forward and backward kernels on 32,768 judged samples of the
    current particles (PM grid of the PM + MLP leg), 568,320 flop per sample
    (forward + backward)."""
    api.mlp_init(7)
    api.mlp_data_reset(32768)
    api.mlp_data_append(np.random.default_rng(11).choice(n, 32768, replace=False))
    out = []
    for batch in (16384, 65536):
        api.mlp_train(2, batch)
        steps = 10
        ms = float(np.median(time_calls(ev, lambda: api.mlp_train(steps, batch), 3))) / steps
        out.append({"batch": batch, "ms_per_step": round(ms, 4), "TFLOP/s": round(568320 * batch / ms / 1e9, 1),
                    "frac_bf16_mfma": round(568320 * batch / ms / 1e9 / 2500.0, 4)})
    return out


def heldout_mlp_leg(api, side: int, ng: int, seed: int = 6) -> dict:
    """The trained correction on particles it never saw: a lattice +
    Gaussian(0.2 cell) set of a seed outside the training seeds (1-5), PM and
    PM + correction against the periodic FP64 direct sum on 4,096 sinks.
    Replaces the particle container."""
    from ct2_amd import _lib
    from ct2_amd.train_mlp import DEFAULT_OUT, load_weights

    trained = load_weights(DEFAULT_OUT)
    if trained is None:
        return {"skipped": "no trained weights"}
    lib = _lib.load()
    n = side ** 3
    x, y, z = api.ic_lattice(side, 0.2, seed)
    api.particle_allocate(n)
    api.particles_upload(x, y, z)
    api.set_force_options(1.0 / (50 * side), 1.0 / n, api.KERNEL_PLUMMER)
    idx = np.random.default_rng(seed).choice(n, 4096, replace=False)
    J = np.stack(api.kick_sample_fp64(idx, periodic=True))
    api.pm_assignment(int(trained[1].get("order", 1)))
    api.pm_init(ng, float(trained[1].get("rs_cells", 1.25)))
    api.mlp_init(1)
    api.mlp_set_weights(trained[0])
    out = {"seed": seed, "training_seeds": trained[1]["seeds"], "pm_grid": ng}
    for name, fn in (("pm", lib.ct2_pm_forces), ("pm_mlp", lib.ct2_pm_mlp_forces)):
        _lib.check(fn())
        fx, fy, fz, _ = api.forces_download()
        e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J[:3])
        out[name + "_median_rel_err"] = float(np.median(e))
        out[name + "_p90_rel_err"] = float(np.percentile(e, 90))
    api.pm_assignment(1)
    return out


def pct(e) -> dict:
    return {"p50": float(np.percentile(e, 50)), "p90": float(np.percentile(e, 90)), "p99": float(np.percentile(e, 99))}


def precision_leg(api, ev: HipEvents, side: int, reps: int) -> dict:
    """The clustered precision study: side^3 particles, half drawn from 64
    Plummer spheres (scale radius 0.01, truncated at 0.1, seeded centres) and
    half uniform, eps = 2e-4.  The FP32 kick with compensated and with naive
    accumulation against the FP64 direct sum (minimum image) on 4,096 sampled
    sinks; the PM force on a 512^3 grid (FP64 Green's table, FP32 FFTs) alone
    and with the near field against the periodic FP64 direct sum on the same
    sinks.  Replaces the particle container."""
    from ct2_amd import _lib
    lib = _lib.load()
    n = side ** 3
    x, y, z = api.ic_plummer_mix(n, n // 2, 64, 0.01, 0.1, 3)
    api.particle_allocate(n)
    api.particles_upload(x, y, z)
    api.set_force_options(2e-4, 1.0 / n, api.KERNEL_PLUMMER)
    idx = np.random.default_rng(7).choice(n, 4096, replace=False)
    Jm = np.stack(api.kick_sample_fp64(idx))
    out = {"n": n, "spheres": 64, "scale_radius": 0.01, "truncation_radius": 0.1, "eps": 2e-4, "sinks": 4096}
    for name, flag in (("fp32_compensated", 0), ("fp32_naive", api.KICK_NAIVE)):
        ms = float(time_calls(ev, lambda: api.kick(None, flag | api.KICK_STORE), 1)[0])
        fx, fy, fz, pot = api.forces_download()
        e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), Jm[:3])
        out[name] = {"ms": round(ms, 1), **pct(e), "pot_p50": float(np.median(np.abs(pot[idx] / Jm[3] - 1)))}
    J = np.stack(api.kick_sample_fp64(idx, periodic=True))
    for name, order, rs, rc in (("pm_cic_512", 1, 1.25, 0.0), ("pm_tsc_512_near", 2, 2.0, 5.5)):
        api.pm_assignment(order)
        api.pm_init(4 * side, rs)
        api.pm_near(rc)
        _lib.check(lib.ct2_pm_forces())
        ms = float(np.median(time_calls(ev, lambda: _lib.check(lib.ct2_pm_forces()), reps)))
        fx, fy, fz, _ = api.forces_download()
        e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J[:3])
        out[name] = {"pm_grid": 4 * side, "rs_cells": rs, "rcut_rs": rc, "ms": round(ms, 3), **pct(e)}
        api.pm_near(0.0)
        api.pm_assignment(1)
    return out


def subset_leg(api, ev: HipEvents, n: int, params, reps: int) -> list:
    """Kicks of a rung subset, as kick_step makes them (main.cpp:50-68):
    particles on rung >= 1 are partitioned to the end (particle_sort_by_rung)
    and kicked against every source: pairs within the active set once
    (symmetric core) when it holds >= 65,536 particles, the rest by the
    ordered kernel.  Leaves the container permuted, so it runs after the legs
    that judge fixed indices."""
    out = []
    for frac in (8, 64):
        rung = np.where(np.arange(n) % frac == 0, 1, 0).astype(np.int8)
        api.particles_upload(rung=rung)
        api.particle_sort_by_rung(1)
        first, last = api.particle_current_range()
        sub = api.kick_params(ascending=True, descending=True, top=False, first=False, rung=1, t0=1e-3,
                              eta=0.15, max_dt=1e-3, scale=1.0)
        api.kick(sub)
        ms = float(np.median(time_calls(ev, lambda: api.kick(sub), reps)))
        m = last - first
        out.append({"active": m, "sources": n, "ms": round(ms, 3), "interactions_per_s": m * float(n) / (ms * 1e-3),
                    "kernel": "kick_sym_pairs core + kick_direct" if m >= 65536 else "kick_direct (ordered pairs)"})
        api.particle_sort_by_rung(0)
    return out


def cpu_baseline(x, y, z, soft, gm, target_s: float):
    """FP64 oracle kick timed on the host cores on a bounded sample of sinks
    (every source), scaled to interactions/s."""
    from oracle import oracle as O
    n = x.size
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    rng = np.random.default_rng(7)
    m = 32
    O.kick(x, y, z, gm, soft, O.KERNEL_PLUMMER, sinks=rng.choice(n, m, replace=False))  # warm-up
    t0 = time.perf_counter()
    O.kick(x, y, z, gm, soft, O.KERNEL_PLUMMER, sinks=rng.choice(n, m, replace=False))
    dt = time.perf_counter() - t0
    m = int(min(n, max(m, m * target_s / max(dt, 1e-6))))
    sinks = rng.choice(n, m, replace=False)
    t0 = time.perf_counter()
    O.kick(x, y, z, gm, soft, O.KERNEL_PLUMMER, sinks=sinks)
    dt = time.perf_counter() - t0
    return {"value": m * n / dt, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"FP64 all-pairs kick of {m} random sinks against all {n} sources "
                      f"({m * n:.3e} interactions, {dt:.1f} s, OpenMP threads={threads})",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "cpus_usable": len(os.sched_getaffinity(0)),
            "cpu_case": cpu_case(threads)}


def cpu_model() -> str:
    """The host CPU's model name as `lscpu` reports it (else /proc/cpuinfo),
    with the family/model/stepping numbers where the name is generic."""
    name, fms = "unknown", []
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                name = v
            elif k in ("CPU family", "Model", "Stepping"):
                fms.append(f"{k.lower()} {v}")
    except (OSError, subprocess.SubprocessError):
        pass
    if name == "unknown":
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        name = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return f"{name} ({', '.join(fms)})" if fms else name


def cpu_case(threads: int) -> dict:
    """The CPU-runnable case (32^3 particles, lattice + Gaussian 0.2 cell,
    eps = 1/1600): one FP64 all-pairs kick (N^2 = 1.07e9 interactions) on all
    threads and on one, one drift, and a 32^3 complex FP64 forward + inverse
    FFT of the CIC density with the radix-2 transform in ct2's bit-reversed
    order; medians of 5 runs after a warm-up."""
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    side = 32
    n = side ** 3
    q = (np.indices((side,) * 3).reshape(3, -1).T + 0.5) / side + rng.normal(0.0, 0.2 / side, (n, 3))
    x, y, z = (np.round((q[:, d] % 1.0) * 2.0 ** 32).astype(np.uint64).astype(np.uint32) for d in range(3))
    soft, gm = 1.0 / (50 * side), 1.0 / n
    v = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]

    def med(fn, reps=5):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    out = {}
    for t in (threads, 1):
        O.set_threads(t)
        s = med(lambda: O.kick(x, y, z, gm, soft, O.KERNEL_PLUMMER), reps=5 if t > 1 else 3)
        out[f"kick_fp64_{t}_threads"] = {"s": round(s, 4), "interactions_per_s": float(n) * n / s}
    O.set_threads(threads)
    s = med(lambda: O.drift(x, y, z, *v, 1e-3, 1.0))
    out["drift"] = {"s": s, "GB/s": round(36 * n / s / 1e9, 2), "note": "includes the wrapper's copies of x, y, z"}
    rho = O.cic_deposit(x, y, z, 5).astype(np.float64) * 2.0 ** -32
    s = med(lambda: O.fft3d(O.fft3d(rho.astype(np.complex128), 1), -1))
    out["fft3d_32_fp64_round_trip"] = {"s": s, "GB/s": round(2 * 96 * 32 ** 3 / s / 1e9, 3)}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=128, help="lattice side (N = side^3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--fft-reps", type=int, default=10)
    ap.add_argument("--reps", type=int, default=10, help="repetitions of the PM timings")
    ap.add_argument("--pm-ng", type=int, default=256, help="PM grid of the PM + MLP leg")
    ap.add_argument("--stepping-side", type=int, default=256, help="particles^(1/3) of the stepping leg (0: skip)")
    ap.add_argument("--stepping-steps", type=int, default=10)
    ap.add_argument("--kick-only", action="store_true",
                    help="time the kick alone (for counter runs under rocprofv3 --pmc)")
    ap.add_argument("--only", choices=["fft", "near", "mm", "mlp", "train"], default=None,
                    help="time one leg alone (for counter runs under rocprofv3 --pmc): the FP32 FFTs, the "
                         "PM + near-field force (TSC, grid 2 side, split 1.5 cells, cutoff 5.5 split radii), or "
                         "the multipole mesh (about 8 particles per cell, order --mm-order), or PM + the "
                         "learned correction (the committed model, grid --pm-ng), or its training steps")
    ap.add_argument("--mm-order", type=int, default=8, help="expansion order of --only mm")
    ap.add_argument("--mm-cells", type=int, default=0, help="cell grid of --only mm (0: about 8 particles per cell)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB three times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the kick would shard by sinks "
              "after one all_gather of positions, the FFT by slabs with one all_to_all)", file=sys.stderr)
        return 2

    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: 3 x {args.copy_probe} MiB read + write", file=sys.stderr, flush=True)

    from ct2_amd import api
    api.init(0)
    ev = HipEvents(api.stream())

    side = args.side
    n = side ** 3
    soft, gm = 1.0 / (50 * side), 1.0 / n
    x, y, z = api.ic_lattice(side, 0.2, 1)
    api.particle_allocate(n)
    api.particles_upload(x, y, z, np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32),
                         np.zeros(n, np.int8))
    api.set_force_options(soft, gm, api.KERNEL_PLUMMER)
    if args.only == "fft":
        line = {"leg": "fft", "fft_fp32": fft_leg(api, ev, args.fft_reps)}
        api.finalize()
        print(json.dumps(line), flush=True)
        return 0
    if args.only == "near":
        from ct2_amd import _lib
        lib = _lib.load()
        api.pm_assignment(2)
        api.pm_init(2 * side, 1.5)
        api.pm_near(5.5)
        _lib.check(lib.ct2_pm_forces())
        ms = [float(v) for v in time_calls(ev, lambda: _lib.check(lib.ct2_pm_forces()), args.reps)]
        line = {"leg": "pm_near", "n": n, "pm_grid": 2 * side, "rs_cells": 1.5, "rcut_rs": 5.5,
                "cells_per_side": api.pm_near_cells(), "ms": ms}
        api.finalize()
        print(json.dumps(line), flush=True)
        return 0
    if args.only == "mlp":
        # PM and PM + learned correction (the committed model's mesh) on the
        # lattice, timed alone: the command the inference counter profile runs
        from ct2_amd import _lib
        from ct2_amd.train_mlp import DEFAULT_OUT, load_weights
        lib = _lib.load()
        w, meta = load_weights(DEFAULT_OUT)
        api.pm_assignment(int(meta.get("order", 1)))
        api.pm_init(args.pm_ng, float(meta.get("rs_cells", 1.25)))
        api.mlp_init(1)
        api.mlp_set_weights(w)
        _lib.check(lib.ct2_pm_forces())
        pm_ms = [float(v) for v in time_calls(ev, lambda: _lib.check(lib.ct2_pm_forces()), args.reps)]
        _lib.check(lib.ct2_pm_mlp_forces())
        ms = [float(v) for v in time_calls(ev, lambda: _lib.check(lib.ct2_pm_mlp_forces()), args.reps)]
        line = {"leg": "pm_mlp", "n": n, "pm_grid": args.pm_ng, "pm_ms": pm_ms, "pm_mlp_ms": ms,
                "mlp_tflops": round(189440 * n / (float(np.median(ms)) - float(np.median(pm_ms))) / 1e9, 2)}
        api.finalize()
        print(json.dumps(line), flush=True)
        return 0
    if args.only == "train":
        api.pm_assignment(2)
        api.pm_init(args.pm_ng, 1.0)
        line = {"leg": "mlp_training", "n": n, "pm_grid": args.pm_ng, "steps": train_leg(api, ev, n)}
        api.finalize()
        print(json.dumps(line), flush=True)
        return 0
    if args.only == "mm":
        nc = args.mm_cells or api.mm_cells_for(n)
        api.mm_init(nc, args.mm_order)
        from ct2_amd import _lib
        lib = _lib.load()
        _lib.check(lib.ct2_mm_forces())
        ms = [float(v) for v in time_calls(ev, lambda: _lib.check(lib.ct2_mm_forces()), args.reps)]
        line = {"leg": "mm", "n": n, "cells": nc, "order": args.mm_order, "ms": ms}
        api.finalize()
        print(json.dumps(line), flush=True)
        return 0
    # top-level kick of kick_step (main.cpp:50-68): ascending, descending, top
    params = api.kick_params(ascending=True, descending=True, top=True, first=False, rung=0, t0=1e-3,
                             eta=0.15, max_dt=1e-3, scale=1.0)
    print(f"[bench] kick N={n} warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        api.kick(params)
    api.synchronize()
    ev.device_sync()
    t0 = time.perf_counter()
    kick_ms = time_calls(ev, lambda: api.kick(params), args.steps)
    api.synchronize()
    ev.device_sync()
    wall = time.perf_counter() - t0
    inter = float(n) * n
    kick_avg = sum(kick_ms) / len(kick_ms)
    value = inter * args.steps / wall
    achieved_tflops = FLOP_PER_INTERACTION * inter / (kick_avg * 1e-3) / 1e12
    print(f"[bench] kick {kick_avg:.1f} ms/step, {inter / kick_avg * 1e3:.3e} int/s", file=sys.stderr, flush=True)
    if args.kick_only:
        api.finalize()
        print(json.dumps({"metric": METRIC, "value": value, "unit": "interactions/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
                          "note": "kick only"}), flush=True)
        return 0

    # the host boundary: the container upload (25 B per particle, pageable host
    # arrays) and the force download (16 B per particle, into new arrays),
    # synchronous, wall clock
    zf, zr = np.zeros(n, np.float32), np.zeros(n, np.int8)
    t_up = []
    for _ in range(3):
        t1 = time.perf_counter()
        api.particles_upload(x, y, z, zf, zf, zf, zr)
        t_up.append((time.perf_counter() - t1) * 1e3)
    t_down = []
    for _ in range(3):
        t1 = time.perf_counter()
        api.forces_download()
        t_down.append((time.perf_counter() - t1) * 1e3)
    # the same copies from and into page-locked arrays (ct2_host_alloc, as the
    # ct2 binding keeps its host arrays)
    pin = [api.pinned_empty(n, dt) for dt in (np.uint32, np.uint32, np.uint32, np.float32, np.float32, np.float32,
                                             np.int8)]
    for dst, src in zip(pin, (x, y, z, zf, zf, zf, zr)):
        dst[:] = src
    t_upp = []
    for _ in range(3):
        t1 = time.perf_counter()
        api.particles_upload(*pin)
        t_upp.append((time.perf_counter() - t1) * 1e3)
    fout = [api.pinned_empty(n, np.float32) for _ in range(4)]
    t_downp = []
    for _ in range(3):
        t1 = time.perf_counter()
        api.forces_download(fout)
        t_downp.append((time.perf_counter() - t1) * 1e3)
    del pin, fout
    host_io = {"upload_ms": round(float(np.median(t_up)), 3), "upload_bytes": 25 * n,
               "download_forces_ms": round(float(np.median(t_down)), 3), "download_bytes": 16 * n,
               "upload_pinned_ms": round(float(np.median(t_upp)), 3),
               "download_forces_pinned_ms": round(float(np.median(t_downp)), 3)}

    # drift: 36 B/particle
    drift_ms = time_calls(ev, lambda: api.drift(1e-9, 1.0), 20)
    drift_avg = float(np.median(drift_ms))

    fft = fft_leg(api, ev, args.fft_reps)

    # force accuracy of the direct kick against the FP64 direct sum (same
    # minimum-image force) on sampled sinks
    api.kick(None, api.KICK_STORE)
    fx, fy, fz, _ = api.forces_download()
    idx = np.random.default_rng(42).choice(n, 4096, replace=False)
    jx, jy, jz, _ = api.kick_sample_fp64(idx)
    e_direct = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), np.stack([jx, jy, jz]))

    # PM and PM + learned correction on the same particles, against the
    # periodic FP64 direct sum (Ewald) on the same sinks
    J = np.stack(api.kick_sample_fp64(idx, periodic=True))
    pmres = pm_leg(api, ev, n, side, kick_avg, idx, J, args.pm_ng, args.reps)
    print("[bench] PM + near field ...", file=sys.stderr, flush=True)
    nearres = [near_leg(api, ev, side, kick_avg, idx, J, args.reps, ng, rs, rc, order)
               for ng, rs, rc, order in ((2 * side, 1.5, 5.5, 2), (2 * side, 2.0, 5.5, 2), (4 * side, 4.0, 5.5, 1))]
    print("[bench] multipole mesh ...", file=sys.stderr, flush=True)
    mmres = [mm_leg(api, ev, n, kick_avg, idx, J, args.reps, order) for order in (4, 5, 6, 8)]
    print("[bench] learned-correction training ...", file=sys.stderr, flush=True)
    train = train_leg(api, ev, n)
    print("[bench] rung-subset kicks ...", file=sys.stderr, flush=True)
    subset = subset_leg(api, ev, n, params, args.reps)
    print("[bench] clustered precision study ...", file=sys.stderr, flush=True)
    precision = precision_leg(api, ev, side, args.reps)
    pmres["held_out"] = heldout_mlp_leg(api, side, args.pm_ng)
    steps_res = None
    if args.stepping_side:
        print("[bench] stepping configuration ...", file=sys.stderr, flush=True)
        steps_res = [stepping_leg(api, ev, inter / (kick_avg * 1e-3), args.stepping_side, args.stepping_steps, 16384,
                                  method) for method in ("pm_near", "pm_mlp", "mm")]
    pmres["direct_fp32_vs_fp64_min_image"] = {"median_rel_err": float(np.median(e_direct)),
                                              "p90": float(np.percentile(e_direct, 90)),
                                              "p99": float(np.percentile(e_direct, 99))}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"all-pairs direct-sum kick, N={side}^3 lattice+Gaussian(0.2 cell), "
                               f"Plummer eps=1/(50*{side}), FP32 compensated, with half-kick/rung epilogue",
                   "n_particles": n, "interactions_per_step": inter},
        "roofline": {"bound": "valu", "achieved": round(achieved_tflops, 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4), "traffic": KICK_TRAFFIC_BYTES.get(side),
                     "kernel": "kick_sym_pairs (+kick_sym_diag, kick_finalize): each pair evaluated once",
                     "flop_per_interaction": FLOP_PER_INTERACTION, "kick_ms": round(kick_avg, 3),
                     "traffic_per": "kick_sym_pairs launch (64 per kick); achieved = 20 N^2 / kick time",
                     # information only: the same kick at 38 flop per interaction (the
                     # Warren et al. 1997 / Hamada et al. 2009 count)
                     "info_38flop_tflops": round(achieved_tflops * 38 / FLOP_PER_INTERACTION, 3),
                     # information only: the work the hardware executes, each unordered pair once
                     # (N^2 / 2 evaluations) at the counted flop and VALU per evaluation
                     "info_executed": {"pair_evaluations": inter / 2,
                                       "hw_flop_per_evaluation": KICK_HW_FLOP_PER_EVAL,
                                       "valu_per_evaluation": KICK_VALU_PER_EVAL,
                                       "hw_tflops": round(KICK_HW_FLOP_PER_EVAL * inter / 2 / (kick_avg * 1e-3) / 1e12, 3),
                                       "hw_frac": round(KICK_HW_FLOP_PER_EVAL * inter / 2 / (kick_avg * 1e-3) / 1e12
                                                        / FP32_PEAK_TFLOPS, 4)}},
        "fft_fp32": fft,
        "drift": {"ms": round(drift_avg, 4), "GB/s": round(36 * n / drift_avg / 1e6, 1),
                  "frac_hbm": round(36 * n / drift_avg / 1e6 / HBM_PEAK_GBS, 4)},
        "pm_mlp_vs_direct": pmres,
        "pm_near_vs_direct": nearres,
        "multipole_mesh_vs_direct": mmres,
        "rung_subset_kicks": subset,
        "host_io": host_io,
        "clustered_precision": precision,
        "mlp_training": train,
        "stepping": steps_res,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(x, y, z, soft, gm, args.cpu_seconds)
    api.finalize()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
