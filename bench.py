"""Benchmark of the FP64 drop-in path on one MI355X.

Workload (one step = one pass of the hot path over one block, inputs already
resident in HBM): the non-grey drop-in block, 64 x 64 columns, nz = 256,
nw = 128 log-spaced 100 A - 10 um, seeded disk profiles
(ocotillo_amd.workloads.config2) -> 1.342e8 (mesh cell, wavelength) points.

Prints one JSON line: points/s, the roofline of the dominant kernel (device
time from HIP events on the library's stream) and a CPU baseline (the FP64
CPU oracle, OpenMP over the host cores, on the same block).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 (vector and matrix): 256 CUs x 64 lanes x 2 x 2.4 GHz
FP64_LANE_OPS = 256 * 64 * 2.4e9  # VALU FP64 instruction-lanes per second (one op per lane per cycle)

# VALU instructions per call of each elementary function / FP64 divide, counted
# in this build's gfx950 ISA (oco_math.h, -O3 -ffp-contract=off), frozen here.
W_EXP, W_EXP10, W_LOG, W_DIV = 37, 40, 74, 11
# Operations per point after hoisting the temperature-only transcendentals
# into the per-cell kernel: the reference's 138 (neutral) / 143 (ionised)
# arithmetic operations with 7 divides, plus the wavelength-dependent
# transcendentals (neutral: 2 exp10 + 1 exp; ionised adds 3 log + 3 exp).
W_NEUTRAL = 138 + 7 * (W_DIV - 1) + 2 * W_EXP10 + W_EXP
W_IONISED = 143 + 7 * (W_DIV - 1) + 2 * W_EXP10 + W_EXP + 3 * W_LOG + 3 * W_EXP
# HBM bytes per sweep launch on this block from the rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes (x2.0 / x1.0 copy-kernel calibration); None if not measured
SWEEP_TRAFFIC_BYTES = None


def ionised_fraction(rho, T, fully_ionized_T=2e4):
    """Share of cells on the bremsstrahlung branch ((1-x) <= 0.01,
    source/continuous_opacity.f90:262), from the Saha root of the inputs."""
    k_ev, mp = 8.617333262145179e-05, 1.67262192369e-24
    with np.errstate(over="ignore", under="ignore"):
        C = 2414683039571967.0 * T ** 1.5 * np.exp(-13.6 / (k_ev * T)) * mp / rho
        x = 0.5 * (np.sqrt(C * C + 4 * C) - C)
    x = np.where(T > fully_ionized_T, 1.0, x)
    return float(((1.0 - x) <= 0.01).mean())


def algorithmic_bytes(nz, ncol, nw):
    """Compulsory HBM bytes of one block: U with ghosts, V and alpha written
    once (FP64), rho and T read once."""
    return 8.0 * ((nz + 6) + 2 * nz) * ncol * nw + 16.0 * nz * ncol


def sweep_bytes(nz, ncol, nw):
    """Algorithmic bytes of the sweep kernel per launch: its outputs U, V, alpha."""
    return 8.0 * ((nz + 6) + 2 * nz) * ncol * nw


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--nw", type=int, default=128)
    ap.add_argument("--config", default="c2", choices=["c2", "c3"],
                    help="c2: the 64x64x256 x 128-wavelength drop-in block (default); "
                         "c3: 4096 uneven columns through the batched persistent path")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (3 times): the reference point "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the path shards by columns/wavelengths with no "
              "exchange (see DESIGN.md), multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "copies": 3}))
        return
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config2
    if args.config == "c3":
        return bench_batch(args)

    rho, T, dz, waves = config2(args.nx, args.ny, args.nz, args.nw)
    nx, ny, nz = rho.shape
    nw = waves.size
    ncol = nx * ny
    points = float(nz) * ncol * nw

    dev = torch.device("cuda:0")
    d_rho = torch.from_numpy(rho).to(dev)
    d_T = torch.from_numpy(T).to(dev)
    d_U = torch.empty((nw, nx, ny, nz + 6), dtype=torch.float64, device=dev)
    d_V = torch.empty((nw, nx, ny, nz), dtype=torch.float64, device=dev)
    d_A = torch.empty((nw, nx, ny, nz), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    ctx = oco.Context(0)
    ctx.set_wavelengths(waves)
    ctx.set_physics(2e4, False, 0.4)

    def step():
        return ctx.solve_block_device(nz, ny, nx, dz, d_rho.data_ptr(), d_T.data_ptr(), d_U.data_ptr(),
                                      d_V.data_ptr(), d_A.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    cells_ms, sweep_ms = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        d = step()
        cells_ms.append(d.cells_ms)
        sweep_ms.append(d.sweep_ms)
        print("step %d: cells %.3f ms sweep %.3f ms" % (i, d.cells_ms, d.sweep_ms), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = points / (ms_per_step * 1e-3)

    sw = float(np.mean(sweep_ms))
    fion = ionised_fraction(rho, T)
    w_pt = (1.0 - fion) * W_NEUTRAL + fion * W_IONISED
    lane_ops = points * w_pt / (sw * 1e-3)
    # the binding roof of the sweep is FP64 VALU issue (no MFMA: nothing here is
    # a dense contraction); TFLOP/s counts each FP64 lane-op at the FMA rate
    roofline = {"kernel": "sweep_fp64", "bound": "valu", "achieved": round(2 * lane_ops / 1e12, 3),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(lane_ops / FP64_LANE_OPS, 4),
                "traffic": SWEEP_TRAFFIC_BYTES, "kernel_ms": round(sw, 4),
                "ops_per_point": round(w_pt, 1), "ionised_fraction": round(fion, 4),
                "weights": {"exp": W_EXP, "exp10": W_EXP10, "log": W_LOG, "div": W_DIV},
                "cells_kernel_ms": round(float(np.mean(cells_ms)), 4)}
    achieved_gbs = sweep_bytes(nz, ncol, nw) / (sw * 1e-3) / 1e9
    roofline_hbm = {"kernel": "sweep_fp64", "bound": "hbm", "achieved": round(achieved_gbs, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
                    "bytes_per_point": round(sweep_bytes(nz, ncol, nw) / points, 3)}

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle
        o = Oracle(waves)
        nthreads = min(os.cpu_count() or 1, 16)
        tc0 = time.perf_counter()
        o.solve_block(rho, T, dz, nthreads=nthreads)
        tc = time.perf_counter() - tc0
        cpu = {"value": round(points / tc, 1), "unit": "points/s", "cores": nthreads, "kind": "port",
               "sample": "the full 64x64x256 x 128-wavelength block, one pass (%.2f s wall)" % tc}

    out = {
        "metric": "mesh*lambda points/s mixed vs FP64 at 1e-6 flux err; serve p50/p99 latency",
        "value": round(value, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "FP64 drop-in block: %dx%d columns, nz=%d, nw=%d log-spaced 100A-10um, "
                               "seeded disk profiles (rho floor 1e-15)" % (nx, ny, nz, nw),
                   "points_per_step": points, "mode": "fp64"},
        "roofline": roofline,
        "roofline_hbm": roofline_hbm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_batch(args):
    """Config 3: 4096 independent columns, nz ~ U{32..1024}, nw ~ U{16..256};
    one persistent launch per step (autotuned, replayed as a hipGraph)."""
    import torch
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config3
    nz, nw, waves, dz, rho, T = config3()
    points = float((nz.astype(np.int64) * nw).sum())
    ctx = oco.Context(0)
    ctx.set_physics(2e4, False, 0.4)
    b = oco.Batch(ctx, nz, nw, waves, dz)
    d_rho, d_T, d_U, d_V, d_A = b.buffers()
    t_rho = torch.from_numpy(rho).to("cuda:0")
    t_T = torch.from_numpy(T).to("cuda:0")
    torch.cuda.synchronize()
    ctypes_copy(d_rho, t_rho.data_ptr(), rho.nbytes)
    ctypes_copy(d_T, t_T.data_ptr(), T.nbytes)
    for _ in range(max(args.warmup, 1)):   # the first run autotunes and captures the graph
        b.run()
    torch.cuda.synchronize()
    ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        d = b.run()
        ms.append(d.total_ms)
        print("step %d: graph %.3f ms" % (i, d.total_ms), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    variant, nblocks, _ = b.tuning()
    fion = ionised_fraction(rho, T)
    w_pt = (1.0 - fion) * W_NEUTRAL + fion * W_IONISED
    kern = float(np.mean(ms))
    lane_ops = points * w_pt / (kern * 1e-3)
    out = {
        "metric": "mesh*lambda points/s mixed vs FP64 at 1e-6 flux err; serve p50/p99 latency",
        "value": round(points / (ms_per_step * 1e-3), 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched uneven FP64: 4096 columns, nz~U{32..1024}, nw~U{16..256} log-spaced "
                               "100A-10um, disk profiles x lognormal(0.5 dex) rho", "points_per_step": points,
                   "mode": "fp64", "autotuned_variant": variant, "workgroups": nblocks},
        "roofline": {"kernel": "batch_fp64 (graph)", "bound": "valu", "achieved": round(2 * lane_ops / 1e12, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(lane_ops / FP64_LANE_OPS, 4),
                     "traffic": None, "kernel_ms": round(kern, 4), "ops_per_point": round(w_pt, 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    b.close()
    ctx.close()


def ctypes_copy(dst, src, nbytes):
    """Device-to-device copy between raw pointers (torch's allocator and the plan's)."""
    import ctypes
    import torch
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    if hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3) != 0:
        raise RuntimeError("hipMemcpy failed")


if __name__ == "__main__":
    main()
