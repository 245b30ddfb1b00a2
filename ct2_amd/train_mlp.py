#!/usr/bin/env python3
"""Trains the learned short-range correction on the GPU.

Labels come from the direct-sum method itself: for sampled particles the
library computes the periodic FP64 direct sum (Ewald) and stores
(judge - PM) / F_unit with the PM-grid features; training runs the
library's BF16-MFMA forward/backward kernels and Adam.

Default: N = 128^3, TSC PM grid 256^3 (split 1.0 cells), no near field
(`--near` adds the direct near-field sum to the base force, and the model
then learns the mesh's remaining error); training sets from the
lattice+Gaussian(0.2 cell) configuration (eps = 1/(50*128)) and the
clustered configuration (2^21 particles, half in 64 Plummer spheres of scale
radius 0.01, eps = 2e-4), 5 seeds each, `--samples` labels per set.
Writes the FP32 weights and the training record to ``--out`` (.npz).
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

DEFAULT_OUT = os.path.join(ROOT, "ct2_amd", "weights", "mlp_n128_ng256.npz")


def particle_set(api, kind: str, side: int, seed: int):
    n = side ** 3
    if kind == "lattice":
        x, y, z = api.ic_lattice(side, 0.2, seed)
        soft = 1.0 / (50 * side)
    elif kind == "clustered":
        x, y, z = api.ic_plummer_mix(n, n // 2, 64, 0.01, 0.1, seed)
        soft = 2e-4
    else:
        raise ValueError(kind)
    return x, y, z, soft


def build_training_set(api, kinds, seeds, side, samples, log):
    n = side ** 3
    for seed in seeds:
        for kind in kinds:
            t0 = time.perf_counter()
            api.particle_allocate(n)
            if kind == "grf":  # the stepping configuration's initial conditions, built on the device
                api.ic_grf(side, 0.3, -2.0, seed, 0.0)
                soft = 1.0 / (50 * side)
            else:
                x, y, z, soft = particle_set(api, kind, side, seed)
                api.particles_upload(x, y, z)
            api.set_force_options(soft, 1.0 / n, api.KERNEL_PLUMMER)
            idx = np.random.default_rng(1000 + seed).choice(n, min(samples, n), replace=False)
            api.mlp_data_append(idx)
            api.synchronize()
            log(f"labels: {kind} seed {seed}: {idx.size} samples in {time.perf_counter() - t0:.1f} s "
                f"(total {api.mlp_data_size()})")


def train(api, steps, batch, lr, seed, log):
    losses = []
    chunk = 500
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        k = min(chunk, steps - done)
        frac = done / steps
        cur = lr * (0.5 * (1 + np.cos(np.pi * frac)) * 0.95 + 0.05)  # cosine decay to 5 %
        l = api.mlp_train(k, batch, float(cur), seed + done)
        losses.append(l)
        done += k
        log(f"train: step {done}/{steps} lr {cur:.2e} loss {float(np.mean(l[-50:])):.4e} "
            f"({time.perf_counter() - t0:.1f} s)")
    return np.concatenate(losses)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--side", type=int, default=128)
    ap.add_argument("--ng", type=int, default=256)
    ap.add_argument("--rs", type=float, default=1.0)
    ap.add_argument("--order", type=int, default=2, choices=[1, 2], help="mass assignment: 1 CIC, 2 TSC")
    ap.add_argument("--near", type=float, default=0.0,
                    help="near-field cutoff in units of rs (0: PM alone); labels are then judge - (PM + near field)")
    ap.add_argument("--samples", type=int, default=2_000_000, help="labels per particle set")
    ap.add_argument("--seeds", type=int, default=5)
    ap.add_argument("--kinds", default="lattice,clustered")
    ap.add_argument("--steps", type=int, default=150000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()

    from ct2_amd import api

    def log(msg):
        print(f"[train_mlp] {msg}", flush=True)

    api.init(0)
    kinds = args.kinds.split(",")
    n = args.side ** 3
    api.pm_assignment(args.order)
    api.pm_init(args.ng, args.rs)
    api.pm_near(args.near)
    api.mlp_init(args.init_seed)
    api.mlp_data_reset(args.samples * args.seeds * len(kinds))
    build_training_set(api, kinds, range(1, args.seeds + 1), args.side, args.samples, log)
    losses = train(api, args.steps, args.batch, args.lr, args.init_seed, log)
    w = api.mlp_get_weights()
    meta = {"side": args.side, "n": n, "ng": args.ng, "rs_cells": args.rs, "order": args.order,
            "near_rcut_rs": args.near, "kinds": kinds, "seeds": args.seeds,
            "samples_per_set": args.samples, "steps": args.steps, "batch": args.batch, "lr": args.lr,
            "final_loss": float(np.mean(losses[-200:]))}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez_compressed(args.out, weights=w, losses=losses[:: max(1, losses.size // 2000)],
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    log(f"wrote {args.out}: {json.dumps(meta)}")
    api.finalize()
    return 0


def load_weights(path: str = DEFAULT_OUT):
    """(weights, meta) of a trained model, or None when the file is absent."""
    if not os.path.exists(path):
        return None
    d = np.load(path, allow_pickle=False)
    meta = json.loads(bytes(d["meta"]).decode())
    return d["weights"], meta


if __name__ == "__main__":
    sys.exit(main())
