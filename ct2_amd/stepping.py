#!/usr/bin/env python3
"""The global-time-step leapfrog of the stepping configuration, on the device.

ct2's step loop is `kick_step` (main.cpp:8-92): kicks by `kick()` with
half-kicks of 0.5 t0 / 2^rung (kick.cpp:80,95) and drifts by `drift(dt, scale)`
(main.cpp:80-81).  With every particle on rung 0 that is kick-drift-kick
leapfrog.  Here it runs in comoving coordinates x and canonical momentum
u = a^2 dx/dt (G = 1, mean comoving density 1, Omega_m = 1 so
H0 = sqrt(8 pi / 3)):
    kick   u += F * int dt/a      (ct2_kick_* with t0 = 2 * factor, rung 0)
    drift  x += u * int dt/a^2    (ct2_drift(factor, 1))
with both factors integrated in FP64 (ct2_cosmos_kick_factor /
ct2_cosmos_drift_factor).  The second half-kick of a step and the first of the
next share one force evaluation.  Forces: PM + learned correction (or PM
alone, or the direct sum), optionally with the direct near-field sum of
pairs closer than `near` split radii (ct2's neighbour-cell structure), or
ct2's multipole mesh (method mm); each step the periodic FP64 direct sum
judges the force on sampled particles.

Initial conditions: a side^3 lattice displaced by a Gaussian random field,
density P(k) ~ k^-2 up to Nyquist, rms displacement `rms` cells per
component (Zel'dovich), u = H0 a^1/2 psi (growing mode, D = a).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def rel_err(F, J):
    return np.linalg.norm(F - J, axis=0) / np.linalg.norm(J, axis=0)


def run(side=256, ng=512, a_start=0.02, a_end=0.03, steps=10, rms=0.3, seed=1, judge=16384, method="pm_mlp",
        weights=None, timer=None, log=print, rs_cells=1.25, near=0.0, order=1, meta=None, mm_order=6, mm_cells=0):
    """Runs the loop; returns a record with per-step force/step times and the
    judged force errors.  `meta` (pm_mlp) is the trained model's metadata: the
    mesh it records is handed to the library, and the run stops if the kick
    does not take the learned path (ct2_kick_pm_mlp falls back to the direct
    sum on another mesh)."""
    from ct2_amd import _lib, api

    lib = _lib.load()
    n = side ** 3
    H = float(np.sqrt(8 * np.pi / 3))
    cosmo = api.cosmology(H)
    api.particle_allocate(n)
    api.ic_grf(side, rms, -2.0, seed, H * a_start ** 0.5)
    api.set_force_options(1.0 / (50 * side), 1.0 / n, api.KERNEL_PLUMMER)
    if method in ("pm", "pm_mlp"):
        api.pm_assignment(order)
        api.pm_init(ng, rs_cells)
        api.pm_near(near)
    if method == "pm_mlp":
        if weights is None:
            raise ValueError("pm_mlp needs trained weights")
        api.mlp_init(1)
        api.mlp_set_weights(weights)
        if meta is not None:
            api.mlp_set_mesh(*api.mlp_mesh_of(meta))
    if method == "mm":  # ct2's multipole mesh: direct pairs over the near cells + expansions of mm_order
        mm_cells = mm_cells or api.mm_cells_for(n)
        api.mm_init(mm_cells, mm_order)
    kick_fn = {"pm_mlp": lib.ct2_kick_pm_mlp, "pm": lib.ct2_kick_pm, "direct": lib.ct2_kick, "mm": lib.ct2_kick_mm}[method]

    a = np.linspace(a_start, a_end, steps + 1)
    ah = 0.5 * (a[:-1] + a[1:])
    rng = np.random.default_rng(seed + 7)
    records = []
    pending = api.cosmos_kick_factor(cosmo, a[0], ah[0])
    for i in range(steps + 1):
        params = api.kick_params(ascending=True, descending=False, top=True, first=False, rung=0,
                                 t0=float(2.0 * pending))
        t0 = time.perf_counter()
        ms_force = timer(lambda: _lib.check(kick_fn(params, api.KICK_STORE, None))) if timer else None
        if timer is None:
            _lib.check(kick_fn(params, api.KICK_STORE, None))
        if method == "pm_mlp" and api.last_force_path() != api.PATH_PM_MLP:
            raise RuntimeError("the learned correction does not fit this mesh; the kick fell back to the direct sum")
        rec = {"step": i, "a": float(a[i]), "kick_factor": float(pending), "force_kick_ms": ms_force}
        if judge:
            idx = rng.choice(n, judge, replace=False)
            fx, fy, fz, _ = api.forces_download()
            jx, jy, jz, _ = api.kick_sample_fp64(idx, periodic=True)
            # the force of the kick just made, at the positions it was computed on
            J = np.stack([jx, jy, jz])
            e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J)
            rec.update({"median_rel_err": float(np.median(e)), "p90_rel_err": float(np.percentile(e, 90)),
                        "p99_rel_err": float(np.percentile(e, 99))})
            if method == "pm_mlp":  # the same force without the learned correction, for comparison
                _lib.check(lib.ct2_pm_forces())
                fx, fy, fz, _ = api.forces_download()
                e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J)
                rec.update({"no_mlp_median_rel_err": float(np.median(e)),
                            "no_mlp_p90_rel_err": float(np.percentile(e, 90))})
        if i == steps:
            records.append(rec)
            break
        D = api.cosmos_drift_factor(cosmo, a[i], a[i + 1])
        ms_drift = timer(lambda: api.drift(float(D), 1.0)) if timer else None
        if timer is None:
            api.drift(float(D), 1.0)
        rec.update({"drift_factor": float(D), "drift_ms": ms_drift, "wall_s": time.perf_counter() - t0})
        records.append(rec)
        log(f"[steps] step {i}: a {a[i]:.4f} -> {a[i + 1]:.4f} force+kick {ms_force} ms "
            f"median err {rec.get('median_rel_err')}")
        pending = api.cosmos_kick_factor(cosmo, ah[i], a[i + 1])
        if i + 1 < steps:
            pending += api.cosmos_kick_factor(cosmo, a[i + 1], ah[i + 1])
    p = api.particles_download()
    if method in ("pm", "pm_mlp"):
        api.pm_near(0.0)
        api.pm_assignment(1)
    mesh = method in ("pm", "pm_mlp")
    out = {"mm_cells": mm_cells, "mm_order": mm_order} if method == "mm" else {}
    return {**out, "side": side, "n": n, "ng": ng if mesh else None, "method": method,
            "rs_cells": rs_cells if mesh else None, "near_rcut_rs": near if mesh else None,
            "assignment": {1: "cic", 2: "tsc"}[order] if mesh else None,
            "steps": steps, "a_start": a_start, "a_end": a_end,
            "records": records, "v_rms": float(np.sqrt(np.mean(p.vx.astype(np.float64) ** 2)))}


def main() -> int:
    ap = argparse.ArgumentParser(description="global-step leapfrog on one GPU")
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--ng", type=int, default=None,
                    help="PM grid side (pm_mlp: the weights' grid-to-lattice ratio times --side; otherwise 2 side)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--judge", type=int, default=16384)
    ap.add_argument("--method", default="pm_mlp", choices=["pm_mlp", "pm", "direct", "mm"])
    ap.add_argument("--mm-order", type=int, default=6, help="expansion order of the multipole mesh (method mm)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "ct2_amd", "weights", "mlp_grf_ng2x.npz"))
    ap.add_argument("--rs", type=float, default=None,
                    help="PM split radius in cells (pm_mlp: the weights' own; otherwise 1.25)")
    ap.add_argument("--near", type=float, default=0.0, help="near-field cutoff in units of rs (0: off)")
    ap.add_argument("--assign", type=int, default=None, choices=[1, 2],
                    help="mass assignment: 1 CIC, 2 TSC (pm_mlp: the weights' own; otherwise CIC)")
    args = ap.parse_args()
    from ct2_amd import api
    from ct2_amd.timing import HipEvents
    from ct2_amd.train_mlp import load_weights

    api.init(0)
    w = meta = None
    if args.method == "pm_mlp":
        t = load_weights(args.weights)
        if t is None:
            print(f"no weights at {args.weights}", file=sys.stderr)
            return 2
        w, meta = t
        # the mesh the model was trained to correct; another one is refused
        ratio = meta["ng"] / meta["side"]
        if args.ng is None:
            args.ng = int(round(args.side * ratio))
        if args.rs is None:
            args.rs = float(meta.get("rs_cells", 1.25))
        if args.assign is None:
            args.assign = int(meta.get("order", 1))
        want = (ratio, float(meta.get("rs_cells", 1.25)), int(meta.get("order", 1)),
                float(meta.get("near_rcut_rs", 0.0)))
        have = (args.ng / args.side, args.rs, args.assign, args.near)
        if any(abs(a - b) > 1e-9 * max(abs(a), abs(b), 1.0) for a, b in zip(want, have)):
            print(f"the weights were trained for (grid/lattice, rs, assignment, near) = {want}, not {have}",
                  file=sys.stderr)
            return 2
    if args.ng is None:
        args.ng = 2 * args.side
    if args.rs is None:
        args.rs = 1.25
    if args.assign is None:
        args.assign = 1
    ev = HipEvents(api.stream())
    rec = run(args.side, args.ng, steps=args.steps, judge=args.judge, method=args.method, weights=w, timer=ev.time,
              log=lambda m: print(m, file=sys.stderr, flush=True), rs_cells=args.rs, near=args.near, order=args.assign,
              meta=meta, mm_order=args.mm_order)
    print(json.dumps(rec))
    api.finalize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
