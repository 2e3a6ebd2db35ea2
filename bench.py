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
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-n", type=int, default=512)
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-c2-cpu-energy", action="store_true",
                    help="skip the oracle's 300-step OpenMP Lanczos on the host (about 2 minutes) "
                         "that checks the 4x4 ground energy")
    ap.add_argument("--no-csc-baseline", action="store_true")
    ap.add_argument("--no-kpm", action="store_true")
    ap.add_argument("--kpm-moments", type=int, default=1024,
                    help="Chebyshev moments of the L=14 KPM leg (1024; fewer for profiling runs)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the one-at-a-time comparison of the batched leg (profiling runs)")
    ap.add_argument("--no-momentum", action="store_true")
    ap.add_argument("--no-conductivity", action="store_true")
    ap.add_argument("--no-expression", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-n", type=int, default=200)
    ap.add_argument("--serving-rate", type=float, default=40.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain device copy of one full vector (2 D doubles) instead; "
                         "run under rocprofv3 --pmc to calibrate FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--traffic", nargs=4, metavar=("HV_FETCH", "HV_WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="(no GPU) turn four rocprofv3 --pmc output folders into the traffic "
                         "record read by the roofline object, then exit")
    args = ap.parse_args()
    if args.traffic:
        write_traffic(*args.traffic)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs on a single MI355X and has no "
              "multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.steps, args.warmup)
        return

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

    copy_gbs = device_copy_rate(x)

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")

    host = host_info()
    threads = args.cpu_threads or min(16, host["affinity"])
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po
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
                         f"matrix-free restatement, {tc * 1e3:.0f} ms each",
               "host": host}
        del xh

    def leg(name, skip, fn, *fargs):
        if skip:
            return None
        progress(f"leg {name}")
        return fn(*fargs)

    lanczos = leg("lanczos_4x4", args.no_lanczos, bench_lanczos_c2, op, x, basis,
                  (Lx, Ly, nu, nd, t, U), not args.no_c2_cpu_energy, threads)
    observables = leg("observables_4x4", args.no_lanczos, bench_observables, op, x)
    del y
    momentum = leg("momentum_4x4_K00", args.no_momentum, bench_momentum)
    conductivity = leg("conductivity_4x4", args.no_conductivity, bench_conductivity)
    expression = leg("expression_4x4", args.no_expression, bench_expression)
    batched = leg("batched", args.no_batched, bench_batched, args.batched_n, not args.no_single)
    kpm = leg("kpm_l14", args.no_kpm, bench_kpm_c3, args.kpm_moments)
    krylov = leg("krylov_l14", args.no_kpm, bench_krylov_fp32)
    serving = leg("serving", args.no_serving, bench_serving, args.serving_n, args.serving_rate)
    csc = leg("cpu_csc_1d_l8", args.no_csc_baseline, bench_cpu_csc_c1, threads)

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
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "copy_GBps": round(copy_gbs, 1), "frac_vs_copy": round(gbs / copy_gbs, 4)},
        "cpu_baseline": cpu,
        "lanczos_4x4": lanczos,
        "observables_4x4": observables,
        "momentum_4x4_K00": momentum,
        "conductivity_4x4": conductivity,
        "expression_4x4": expression,
        "batched": batched,
        "kpm_l14": kpm,
        "krylov_l14": krylov,
        "serving": serving,
        "cpu_csc_1d_l8": csc,
    }
    print(json.dumps(line))


D_4X4_7P7 = 11440 * 11440


def progress(msg):
    """One progress line on stderr (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_info():
    """The host the CPU legs run on: logical CPUs, the ones this process may use,
    the CPU model (/proc/cpuinfo) and OMP_NUM_THREADS."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def device_copy_rate(x, reps=10):
    """GB/s of a plain device copy of x (read + write bytes / time): the
    achievable HBM rate the H.v is also priced against."""
    import torch
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del y
    return 2.0 * x.numel() * x.element_size() / (ms * 1e-3) / 1e9


def bench_cpu_csc_c1(threads, reps=5, k=200):
    """The reference's cost structure on the 1D ring L = 8, 4 + 4, U = 4
    (D = 4,900), restated by the oracle: OpenMP column-parallel CSC build with
    a hash lookup per nonzero (matrix_elements.h:74-105) on `threads` threads,
    then a 200-step Lanczos (lanczos.h:130-175) whose CSC product runs on one
    thread (linear_operator.h:87; ARMA_OPENMP_THREADS=1, src/CMakeLists.txt:20),
    vector passes on one thread too.  Median of `reps` runs after a warm-up,
    beside the device's 200-step Lanczos from the same start vector."""
    import numpy as np
    import torch

    import manybody_amd as mb
    from oracle import pyoracle as po
    L, nu, nd, t, U = 8, 4, 4, 1.0, 4.0
    op = mb.HubbardModel(t, U, L).operator(mb.FermionBasis.sector(L, nu, nd))
    b = op.gaussian(1)
    b_ref = op.to_reference_order(b).cpu().numpy()
    builds, solves, gpu = [], [], []
    A = po.build_csc([L], nu, nd, t, U, threads=threads)
    for _ in range(reps):
        c0 = time.perf_counter()
        A = po.build_csc([L], nu, nd, t, U, threads=threads)
        builds.append(time.perf_counter() - c0)
    po.lib().orc_set_threads(1)
    po.lanczos_csc(A, b_ref, k)
    for _ in range(reps):
        c0 = time.perf_counter()
        al, be, _ = po.lanczos_csc(A, b_ref, k)
        e_cpu = float(po.tridiag_eigvals(al, be).min())
        solves.append(time.perf_counter() - c0)
    po.lib().orc_set_threads(threads)
    mb.lanczos_pass_one(op, b, k)
    for _ in range(reps):
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        dec = mb.lanczos_pass_one(op, b, k)
        w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
        gpu.append(time.perf_counter() - c0)
    tb, ts, tg = (float(np.median(v)) for v in (builds, solves, gpu))
    return {"workload": "1D ring L=8 4+4 U=4 (D=4900): CSC build + 200-step Lanczos, "
                        "oracle restatement of the reference's cost structure",
            "nnz": int(A.col_ptr[-1]), "build_threads": threads, "build_ms": round(tb * 1e3, 3),
            "solve_threads": 1, "solve_ms": round(ts * 1e3, 3),
            "ms_per_step_cpu": round(ts / k * 1e3, 4), "E0_cpu": e_cpu,
            "gpu_solve_ms": round(tg * 1e3, 3), "E0_gpu": float(w[0]),
            "rel_dE": abs(float(w[0]) - e_cpu) / abs(e_cpu)}


def copy_calibration(steps, warmup):
    """Plain device copy of 2 D doubles (one complex128 vector of the 4x4 7+7
    sector, 2.09 GB): the known-size kernel that calibrates the HBM counters."""
    import torch
    torch.cuda.set_device(0)
    n = 2 * D_4X4_7P7
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        y.copy_(x)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "ms": round(ms, 5),
                      "GB/s": round(16 * n / (ms * 1e-3) / 1e9, 1)}))


def _pmc_per_kernel(folder, counter):
    """Mean counter value (KiB per dispatch) per kernel name from a rocprofv3
    --pmc --output-format csv folder."""
    import csv
    import glob
    vals = {}
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] == counter:
                    vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}


def write_traffic(hv_fetch, hv_write, copy_fetch, copy_write):
    """HBM bytes per H.v launch pair (up kernel + row kernel) from FETCH_SIZE
    and WRITE_SIZE, each scaled by the factor that makes the same counter read
    the known byte count of a plain copy of one vector."""
    known = 16.0 * D_4X4_7P7  # bytes read (and written) by the calibration copy

    def copy_kernel(d):
        big = {k: v for k, v in d.items() if "opy" in k or "elementwise" in k}
        return max(big.values())

    f_copy = copy_kernel(_pmc_per_kernel(copy_fetch, "FETCH_SIZE")) * 1024.0
    w_copy = copy_kernel(_pmc_per_kernel(copy_write, "WRITE_SIZE")) * 1024.0
    ff, wf = known / f_copy, known / w_copy
    fetch = _pmc_per_kernel(hv_fetch, "FETCH_SIZE")
    write = _pmc_per_kernel(hv_write, "WRITE_SIZE")
    kernels = {}
    for name in fetch:
        for short in ("up_block_kernel", "up_strip_kernel", "row_kernel_sorted", "row_kernel_compact",
                      "row_kernel_ell"):
            if short + "<" in name:
                kernels[short] = {"fetch_bytes": round(fetch[name] * 1024.0 * ff),
                                  "write_bytes": round(write.get(name, 0.0) * 1024.0 * wf)}
    total = sum(v["fetch_bytes"] + v["write_bytes"] for v in kernels.values())
    out = {"workload": "hubbard2d_4x4_pbc_t1_U8_7up7dn_hv_c128", "D": D_4X4_7P7,
           "bytes_per_launch": round(total), "algorithmic_bytes": 32 * D_4X4_7P7,
           "fetch_factor": round(ff, 4), "write_factor": round(wf, 4), "kernels": kernels}
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def bench_lanczos_c2(op, x, basis, model, cpu_energy, threads, k=300):
    """300-step FP64 Lanczos on the 4x4 7+7 sector from the bench vector: lowest
    Ritz value and time per step; unless switched off, the oracle's OpenMP
    Lanczos on the host from the same vector (about 2 minutes) for the energy
    comparison (bar: 1e-10 relative)."""
    import numpy as np
    import torch

    import manybody_amd as mb
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dec = mb.lanczos_pass_one(op, x, k)
    t_gpu = time.perf_counter() - t0
    w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
    out = {"steps": dec.steps_taken, "E0": float(w[0]), "s": round(t_gpu, 3),
           "ms_per_step": round(t_gpu / k * 1e3, 3),
           "algorithmic_GBps": round(96.0 * op.dimension() * k / t_gpu / 1e9, 1)}
    if cpu_energy:
        import scipy.linalg as sla
        from oracle import pyoracle as po
        Lx, Ly, nu, nd, t, U = model
        P = po.ProductOperator([Lx, Ly], nu, nd, t, U, threads=threads)
        xh = mb.from_planar(x)
        po.lib().orc_set_progress(25)  # progress lines on stderr during the host run
        c0 = time.perf_counter()
        al, be, _ = P.lanczos(xh, k)
        t_cpu = time.perf_counter() - c0
        po.lib().orc_set_progress(0)
        e_cpu = float(sla.eigvalsh_tridiagonal(np.asarray(al), np.asarray(be))[0])
        rel = abs(out["E0"] - e_cpu) / abs(e_cpu)
        out.update({"E0_cpu_oracle": e_cpu, "cpu_s": round(t_cpu, 1), "cpu_threads": threads,
                    "rel_dE": rel, "rel_dE_below_1e-10": bool(rel < 1e-10)})
    return out


def bench_observables(op, x, reps=5):
    """Host-synchronous observables of the bench vector on 4x4 7+7: all 16
    opposite-spin correlations G(r) in one pass (16 D algorithmic bytes: x read
    once) and the (energy, residual) pair (one H.v plus two passes over x and
    Hx).  Median wall time of `reps` calls each, tables and allocation included."""
    import numpy as np
    import torch
    torch.cuda.synchronize()
    D = op.dimension()
    op.opposite_spin_correlation(x)
    op.energy_residual(x)
    tc, te = [], []
    for _ in range(reps):
        c0 = time.perf_counter()
        g = op.opposite_spin_correlation(x)
        tc.append(time.perf_counter() - c0)
        c0 = time.perf_counter()
        e, r = op.energy_residual(x)
        te.append(time.perf_counter() - c0)
    tc, te = float(np.median(tc)), float(np.median(te))
    return {"correlation_ms": round(tc * 1e3, 3),
            "correlation_GBps": round(16.0 * D / tc / 1e9, 1),
            "sum_rule_error": abs(float(g.sum()) - 49 / 16),
            "energy_residual_ms": round(te * 1e3, 3), "energy": e, "residual": r}


def bench_momentum(steps=20, k=300):
    """Fixed-momentum sector K = (0, 0) of the 4x4 7+7 model (D_K ~ D / 16): host
    table build, FP64 H.v (32 D_K algorithmic bytes) and a 300-step Lanczos."""
    import torch

    import manybody_amd as mb
    t0 = time.perf_counter()
    basis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(16, 14, 0, (4, 4), (0, 0))
    op = mb.HubbardMomentumFactorizedFixedK(basis, (4, 4), 1.0, 8.0, (0, 0))
    t_build = time.perf_counter() - t0
    D = op.dimension()
    x = op.gaussian(seed=20260101)
    y = torch.empty_like(x)
    for _ in range(2):
        op.hv(x, y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        op.hv(x, y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    c0 = time.perf_counter()
    dec = mb.lanczos_pass_one(op, x, k)
    t_l = time.perf_counter() - c0
    w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
    return {"workload": "4x4 7+7 U=8, total momentum K=(0,0), factorized momentum-space H",
            "D_K": D, "build_s": round(t_build, 2), "hv_ms": round(ms, 4),
            "hv_algorithmic_GBps": round(32.0 * D / ms / 1e6, 1),
            "lanczos_steps": dec.steps_taken, "lanczos_s": round(t_l, 3), "E0_K00": float(w[0])}


def bench_conductivity(steps=10, samples=1, beta=4.0, dt=0.1, krylov=30, reps=20):
    """Optical conductivity of the 4x4 7+7 model at U = 8: sectors K = (0,0) and
    K + Q with Q = (1,0), the x current J_x(Q) (one transfer kernel, 16 D_K read +
    16 D_KQ written algorithmically), and the stochastic correlator: one sample,
    exp(-beta H_K / 2) then `steps` time points of two 30-step Krylov evolutions."""
    import torch

    import manybody_amd as mb
    t, U = 1.0, 8.0
    t0 = time.perf_counter()
    bk = mb.FermionBasis.with_fixed_particle_number_spin_momentum(16, 14, 0, (4, 4), (0, 0))
    bq = mb.FermionBasis.with_fixed_particle_number_spin_momentum(16, 14, 0, (4, 4), (1, 0))
    hk = mb.HubbardMomentumFactorizedFixedK(bk, (4, 4), t, U, (0, 0))
    hq = mb.HubbardMomentumFactorizedFixedK(bq, (4, 4), t, U, (1, 0))
    J = mb.CurrentOperatorFixedK(bk, bq, (4, 4), t, 1, 0)
    t_build = time.perf_counter() - t0
    DK, DQ = hk.dimension(), hq.dimension()
    x = hk.gaussian(seed=5)
    y = J.apply(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        J.apply(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del x, y
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    C = mb.optical_correlator(hk, hq, J, samples, beta, dt, steps, seed=20260101,
                              opts=mb.EvolutionOptions(krylov))
    t_c = time.perf_counter() - c0
    r = mb.compute_optical_conductivity(C, dt, beta, 16.0)
    return {"workload": "4x4 7+7 U=8, J_x(Q=(1,0)) from K=(0,0), beta=4, dt=0.1",
            "D_K": DK, "D_KQ": DQ, "build_s": round(t_build, 2), "current_ms": round(ms, 4),
            "current_algorithmic_GBps": round(16.0 * (DK + DQ) / ms / 1e6, 1),
            "correlator_samples": samples, "correlator_steps": steps, "krylov_steps": krylov,
            "correlator_s": round(t_c, 3), "C0": [float(C[0].real), float(C[0].imag)],
            "sum_rule": r.sum_rule}


def bench_expression(steps=10):
    """HubbardModel2D(1, 8, 4, 4).hamiltonian() as a generic FermionExpression
    compiled for the device, on the full 7 + 7 sector: host compile time, H.v time
    against the native two-pass H.v on the same vector (32 D algorithmic bytes).
    The expression has Hubbard form, so by default it runs on Hubbard tables
    built from its own hop lists; generic_hv_ms is the same expression on the
    generic expression kernels (MANYBODY_AMD_EXPR_NATIVE=0)."""
    import os

    import torch

    import manybody_amd as mb
    basis = mb.FermionBasis.sector(16, 7, 7)
    model = mb.HubbardModel2D(1.0, 8.0, 4, 4)
    t0 = time.perf_counter()
    expr = model.hamiltonian()
    op = mb.FermionExpressionOperator(basis, expr)
    t_build = time.perf_counter() - t0
    D = op.dimension()
    x = op.gaussian(seed=20260101)
    y = torch.empty_like(x)

    def timed(o):
        o.hv(x, y)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            o.hv(x, y)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms = timed(op)
    y1 = y.clone()
    del op
    os.environ["MANYBODY_AMD_EXPR_NATIVE"] = "0"
    try:
        gen = mb.FermionExpressionOperator(basis, expr)
    finally:
        del os.environ["MANYBODY_AMD_EXPR_NATIVE"]
    ms_generic = timed(gen)
    err_generic = (torch.abs(y1 - y).max() / torch.abs(y).max()).item()
    del gen
    native = model.operator(basis)
    ms_native = timed(native)
    err = (torch.abs(y1 - y).max() / torch.abs(y).max()).item()
    return {"workload": "HubbardModel2D(t=1, U=8, 4x4).hamiltonian() as a compiled expression, 7+7",
            "terms": len(expr), "D": D, "compile_s": round(t_build, 2), "hv_ms": round(ms, 3),
            "hv_algorithmic_GBps": round(32.0 * D / ms / 1e6, 1),
            "native_hv_ms": round(ms_native, 3), "max_rel_diff_vs_native": err,
            "generic_hv_ms": round(ms_generic, 3),
            "max_rel_diff_generic_vs_hubbard_form": err_generic}


def bench_kpm_c3(M=1024, R=16, seed=7):
    """KPM density of states, 1D ring L = 14, 7 + 7, t = 1, U = 4
    (D = 11,778,624): bounds from a 100-step FP64 Lanczos, spectrum mapped to
    [-0.99, 0.99], 1024 moments over 16 random-phase vectors, H.v in FP32 and in
    FP64; the FP32 error against the FP64 run is stated."""
    import numpy as np
    import torch

    import manybody_amd as mb
    L, nu, nd, t, U = 14, 7, 7, 1.0, 4.0
    op = mb.HubbardModel(t, U, L).operator(mb.FermionBasis.sector(L, nu, nd))
    D = op.dimension()
    b = op.gaussian(11)
    dec = mb.lanczos_pass_one(op, b, 100)
    w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
    a, c = mb.rescaling_from_bounds(float(w[0]), float(w[-1]), 1.0 / 0.99 - 1.0)
    del b
    res = {}
    for fp32 in (True, False):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mu = np.array(mb.kpm_dos_moments(op, a, c, M, R, seed, fp32=fp32))
        res[fp32] = (mu, time.perf_counter() - t0)
    mu32, t32 = res[True]
    mu64, t64 = res[False]
    g = np.array(mb.jackson_kernel(M - 1))
    xs = np.linspace(-0.99, 0.99, 4001)
    Tn = np.cos(np.outer(np.arange(M), np.arccos(xs)))

    def dos(mu):
        c_n = g * mu
        c_n[1:] *= 2.0
        return (c_n @ Tn) / (np.pi * np.sqrt(1.0 - xs ** 2))

    r32, r64 = dos(mu32), dos(mu64)
    l1 = float(np.trapezoid(np.abs(r32 - r64), xs) / np.trapezoid(np.abs(r64), xs))
    steps = M // 2
    return {"workload": "KPM DOS, 1D ring L=14 7+7 U=4, 1024 moments x 16 random-phase vectors",
            "D": D, "bounds": [float(w[0]), float(w[-1])], "a": a, "b": c,
            "fp32_s": round(t32, 3), "fp64_s": round(t64, 3),
            "fp32_GBps": round(24.0 * R * D * steps / t32 / 1e9, 1),  # 24 B per complex element
            "max_abs_dmu_fp32_vs_fp64": float(np.abs(mu32 - mu64).max()),
            "dos_rel_L1_fp32_vs_fp64": l1, "mu0": float(mu64[0]),
            "max_abs_mu": float(np.abs(mu64).max())}


def bench_krylov_fp32(t1=1.0, dt=0.05, k=30):
    """Krylov exp(-iHt) on the 1D ring L = 14, 7 + 7, U = 4 (D = 11,778,624):
    normalised seeded start, t = 0 -> 1 in steps of 0.05, Krylov dimension 30,
    Krylov vectors and H.v in FP64 and in FP32 (state, reductions and the
    exponential stay FP64); the FP32 error against the FP64 run is stated."""
    import torch

    import manybody_amd as mb
    op = mb.HubbardModel(1.0, 4.0, 14).operator(mb.FermionBasis.sector(14, 7, 7))
    psi = op.gaussian(21)
    psi = psi / torch.linalg.vector_norm(psi)
    opts = mb.EvolutionOptions(k)
    out = {}
    for fp32 in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out[fp32] = (mb.time_evolve_state_steps(op, psi, 0.0, t1, dt, opts, fp32=fp32),
                     time.perf_counter() - t0)
    err = (torch.linalg.vector_norm(out[True][0] - out[False][0]) /
           torch.linalg.vector_norm(out[False][0])).item()
    ov = torch.sum(psi[0] * out[False][0][0] + psi[1] * out[False][0][1]).item()
    return {"workload": "Krylov exp(-iHt), 1D ring L=14 7+7 U=4, t=1, dt=0.05, k=30",
            "D": op.dimension(), "fp64_s": round(out[False][1], 3), "fp32_s": round(out[True][1], 3),
            "rel_state_err_fp32_vs_fp64": err, "re_overlap_psi0_psit": ov,
            "norm_fp32": torch.linalg.vector_norm(out[True][0]).item()}


def bench_serving(n, rate, slots=4, arena=64 << 30, lane=8 << 30):
    """Poisson request stream (manybody_amd.workloads.serving_requests): 70 %
    ground state, 20 % KPM DOS (FP32 H.v), 10 % Krylov exp(-iHt) on 1D L = 8..14
    and 2D 3x3 / 4x3 rings, U ~ U[0, 12].  Served with `slots` stream slots and
    the Lanczos lane: p50 / p99 latency (arrival -> result) at `rate`; then the
    same requests all at once (saturated) against one request at a time without
    the lane."""
    import numpy as np
    from dataclasses import replace

    import manybody_amd as mb
    from manybody_amd.workloads import serving_requests
    reqs = serving_requests(n, rate=rate, seed=5)

    def run(rs, n_slots, lane_bytes):
        with mb.Server(n_slots, arena, lane_bytes=lane_bytes) as S:
            ids = [S.submit(r) for r in rs]
            out = [S.wait(i) for i in ids]
        bad = [x for x in out if x["status"] != 0]
        if bad:
            raise RuntimeError("served request failed: " + bad[0]["error"])
        span = max(x["finished"] for x in out) - min(x["arrival"] for x in out)
        lat = np.array([x["finished"] - x["arrival"] for x in out])
        return out, span, lat

    out, span, lat = run(reqs, slots, lane)
    sat = [replace(r, arrival=-1.0) for r in reqs]
    _, span_sat, _ = run(sat, slots, lane)
    _, span_one, _ = run(sat, 1, 0)
    kinds = np.bincount([r.kind for r in reqs], minlength=3)
    return {"workload": f"{n} requests, Poisson {rate}/s: {kinds[0]} ground, {kinds[1]} KPM, "
                        f"{kinds[2]} Krylov; 1D L 8-14, 2D 3x3/4x3",
            "slots": slots, "lane_bytes": lane,
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "solves_per_s": round(n / span, 1),
            "saturated_solves_per_s": round(n / span_sat, 1),
            "one_at_a_time_solves_per_s": round(n / span_one, 1),
            "speedup_vs_one_at_a_time": round(span_one / span_sat, 2),
            "in_lane": int(sum(1 for x in out if x["slot"] >= slots))}


def bench_batched(n, single=True):
    """512 independent 1D ring ground states (L ~ U{8..12}, N_up, N_dn ~
    U{1..L/2}, U ~ U[0,12], t = 1), 150 FP64 Lanczos steps each from seeded
    N(0,1) start vectors: batched engine (autotuned mapping, one hipGraph)
    against the plain one-at-a-time path (find_min_eigenpair per instance)."""
    import numpy as np
    import torch

    import manybody_amd as mb
    from manybody_amd.workloads import batch_1d_instances

    k = 150
    inst = batch_1d_instances(n, seed=4)
    ops = [mb.HubbardModel(1.0, U, L).operator(mb.FermionBasis.sector(L, nu, nd))
           for (L, nu, nd, U) in inst]
    seeds = np.arange(1000, 1000 + n, dtype=np.uint64)
    B = mb.Batch(ops)
    B.fill_gaussian(seeds)
    tile, block = B.autotune(3)
    B.ground_energies(k, use_graph=True)  # capture + warm
    # five timed replays of the captured graph: median, spread, and whether the
    # replays give the same energies bit for bit
    times, e_batch, identical = [], None, True
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = np.asarray(B.ground_energies(k, use_graph=True))
        times.append(time.perf_counter() - t0)
        if e_batch is None:
            e_batch = e
        else:
            identical = identical and bool(np.array_equal(e, e_batch))
    t_batch = float(np.median(times))
    dims = np.array([op.dimension() for op in ops])
    if not single:
        return {"workload": f"{n} x 1D Hubbard ring ground states, L 8-12, 150 FP64 Lanczos steps",
                "solves_per_s": round(n / t_batch, 1), "batch_s": round(t_batch, 4),
                "replays": len(times), "batch_s_min": round(min(times), 4),
                "batch_s_max": round(max(times), 4), "replays_bit_identical": identical,
                "mapping": {"tile": tile, "block": block}}
    # plain path: one find_min_eigenpair per instance (start vectors filled beforehand)
    starts = [op.gaussian(int(sd)) for op, sd in zip(ops, seeds)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_single = np.array([mb.find_min_eigenpair(op, k, b, want_vector=False).value
                         for op, b in zip(ops, starts)])
    t_single = time.perf_counter() - t0
    return {"workload": f"{n} x 1D Hubbard ring ground states, L 8-12, 150 FP64 Lanczos steps",
            "solves_per_s": round(n / t_batch, 1), "batch_s": round(t_batch, 4),
            "replays": len(times), "batch_s_min": round(min(times), 4),
            "batch_s_max": round(max(times), 4), "replays_bit_identical": identical,
            "p50_ms": round(float(np.percentile(times, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(times, 99)) * 1e3, 2),
            "single_solves_per_s": round(n / t_single, 1),
            "speedup_vs_single": round(t_single / t_batch, 2),
            "max_abs_dE_vs_single": float(np.abs(e_batch - e_single).max()),
            "mapping": {"tile": tile, "block": block},
            "dims": {"min": int(dims.min()), "mean": float(dims.mean()), "max": int(dims.max())}}


if __name__ == "__main__":
    main()
