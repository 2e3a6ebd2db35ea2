#!/usr/bin/env python3
"""Benchmark: matrix-free FP64 Hubbard H.v on one MI355X.

Workload (one step = one H.v over the whole vector, inputs resident in HBM):
2D Hubbard, 4 x 4 periodic square lattice, t = 1, U = 8, N_up = N_dn = 7,
D = 11,440^2 = 130,873,600, complex128 vectors with seeded N(0,1) entries.

value = algorithmic GB/s = 32 * D bytes (read x 16 B + write y 16 B per
element) / mean time per H.v.  The roofline object prices the same kernel
against MI355X's 8 TB/s HBM; `traffic` is the measured HBM bytes per launch
(rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) when
profiles/hv_c128_4x4_7p7_traffic.json holds it.  cpu_baseline times the
oracle's OpenMP matrix-free H.v on the host cores on the same vector (the
reference's sparse-matrix path needs ~109 GB of CSC here and cannot run).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hv_c128_4x4_7p7_traffic.json")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import manybody_amd as mb

    torch.cuda.set_device(0)
    Lx, Ly, nu, nd, t, U = 4, 4, 7, 7, 1.0, 8.0
    basis = mb.FermionBasis.sector(Lx * Ly, nu, nd)
    op = mb.HubbardModel2D(t, U, Lx, Ly).operator(basis)
    D = op.dimension()
    x = op.gaussian(seed=20260101)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        op.hv(x, y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        op.hv(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    alg_bytes = 32.0 * D
    gbs = alg_bytes / (ms * 1e-3) / 1e9

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        P = po.ProductOperator([Lx, Ly], nu, nd, t, U, threads=threads)
        xh = mb.from_planar(x)
        P.matvec(xh)  # warm-up (page faults)
        runs = []
        for _ in range(3):
            c0 = time.perf_counter()
            P.matvec(xh)
            runs.append(time.perf_counter() - c0)
        tc = float(np.median(runs))
        cpu = {"value": round(alg_bytes / tc / 1e9, 3), "unit": "GB/s", "cores": threads,
               "kind": "port",
               "sample": f"median of 3 full-vector H.v (D={D}) by the oracle's OpenMP "
                         f"matrix-free restatement, {tc * 1e3:.0f} ms each"}
        del xh

    line = {
        "metric": "Hubbard H·v GB/s (% HBM peak); batched solves/s + p99 latency at |ΔE|<1e-8",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded N(0,1) complex128 vector (SplitMix64 + Box-Muller per reference index)",
        "config": {"workload": "hubbard2d_4x4_pbc_t1_U8_7up7dn_hv_c128", "D": D,
                   "algorithmic_bytes_per_step": int(alg_bytes), "parallelism": "single GPU",
                   "wall_s": round(wall, 3)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
