"""Gravity forward + adjoint throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (pm_blobs_1m): 128x128x64 cells of 500 m (1,048,576 cells), 16,384
stations on a 128x128 grid with 0-2,000 m topography, densities from 8 seeded
Gaussian blobs + noise. One step = one FP32 forward g = G rho plus one FP32
adjoint G^T r over the whole grid and station set (2 x 1.718e10 station-cell
pairs), through libztm_gfm's plan API, inputs resident in HBM.

Prints one JSON line: value = station x cell pairs per second (forward +
adjoint), the FP64 GPU path's rate on the same step and the speedup over it,
the FP32 path's max-normalised error against the FP64 GPU path, a roofline
entry for the dominant kernel (timed with HIP events on its launch stream),
and the CPU baseline: the FP64 restatement of the reference CPU gravity
(TFM.cpp:95-99 arithmetic, pow(r^2,1.5)) timed on this host on a bounded
sample, which also checks the GPU result on that sample.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "station×cell pairs/s fwd+adjoint, FP32/BF16 vs FP64 on GPU at ≤1e-5 rel err"
FLOP_PER_PAIR = 13            # 3 sub, r^2 (1 mul + 2 fma), 2 mul r^-3, 1 mul rho*dz, 1 fma
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector, published
PEAK_HBM_GBS = 8000.0
PEAK_FP64_MFMA_TFLOPS = 78.6  # MI355X dense FP64 matrix, published


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def relerr(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def cpu_baseline(w, r, g32, a32, n_st=512, n_cells=32768):
    """Time the oracle's FP64 restatement of the reference CPU path on a sample:
    the forward for n_st stations over all cells, the adjoint for n_cells cells
    over all stations. Returns the baseline dict and the GPU errors on it."""
    import oracle
    threads = oracle.DEFAULT_THREADS
    st_idx = np.unique((np.arange(n_st) * (w.n_stations / n_st)).astype(np.int64))
    cell_idx = np.unique((np.arange(n_cells) * (w.grid.n_cells / n_cells)).astype(np.int64))
    oracle.lib()
    t0 = time.perf_counter()
    g_cpu = oracle.pm_forward(w.grid, w.kappa, w.rho, w.stations[st_idx], threads=threads)
    t1 = time.perf_counter()
    a_cpu = oracle.pm_adjoint_cells(w.grid, w.kappa, w.stations, r, cell_idx, threads=threads)
    t2 = time.perf_counter()
    pairs = st_idx.size * w.grid.n_cells + cell_idx.size * w.n_stations
    base = {
        "value": pairs / (t2 - t0),
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP64 pow(r^2,1.5) point-mass sum restating TFM.cpp:95-99 on {threads} threads "
                   f"({cpu_model()}, nproc={os.cpu_count()}): forward for {st_idx.size} of "
                   f"{w.n_stations} stations over all {w.grid.n_cells} cells + adjoint for "
                   f"{cell_idx.size} cells over all stations = {pairs:.3e} pairs in "
                   f"{t2 - t0:.2f} s"),
    }
    errs = {"forward": relerr(g32[st_idx], g_cpu), "adjoint": relerr(a32[cell_idx], a_cpu)}
    return base, errs


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the stored-matrix and prism legs")
    ap.add_argument("--leg", choices=["prism", "batched", "uneven", "tfm", "accumulation", "gn",
                                      "gn_lib", "lm", "mfma"],
                    default=None,
                    help="run only this extra leg and print its JSON object (for profiling)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run 5 device-to-device copies of this many MiB "
                         "(calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE against known bytes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    from ztm_amd import gfm, workloads as W

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    if args.copy_calib_mib:
        n = args.copy_calib_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n, "copies": 5}))
        return 0
    if args.leg in ("batched", "uneven", "tfm", "accumulation", "gn", "gn_lib", "lm", "mfma"):
        leg = {"batched": batched_leg, "uneven": uneven_leg, "tfm": tfm_leg,
               "accumulation": accumulation_leg, "gn": gn_leg, "gn_lib": gn_lib_leg,
               "lm": lm_leg, "mfma": mfma_leg}[args.leg]
        print(json.dumps({args.leg: leg(torch.cuda.current_stream(), args)}))
        return 0
    w = W.pm_blobs_1m()
    r = W.normal(W.SEED0 + 2, 40, w.n_stations, 0.5)
    pairs = w.pairs
    if args.leg == "prism":
        print(json.dumps({"prism": prism_leg(w, r, torch.cuda.current_stream(), args)}))
        return 0

    plan = gfm.Plan(w.grid, w.stations, kappa=w.kappa)
    rho32 = torch.from_numpy(w.rho.astype(np.float32)).to(dev)
    r32 = torch.from_numpy(r.astype(np.float32)).to(dev)
    rho64 = torch.from_numpy(w.rho).to(dev)
    r64 = torch.from_numpy(r).to(dev)
    g32 = torch.empty(w.n_stations, dtype=torch.float64, device=dev)
    a32 = torch.empty(w.grid.n_cells, dtype=torch.float32, device=dev)
    g64 = torch.empty(w.n_stations, dtype=torch.float64, device=dev)
    a64 = torch.empty(w.grid.n_cells, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    def step32():
        plan.forward(rho32, out=g32, stream=stream)
        plan.adjoint(r32, out=a32, stream=stream)

    def step64():
        plan.forward(rho64, out=g64, stream=stream)
        plan.adjoint(r64, out=a64, stream=stream)

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()

    gfm.profile_reset()
    gfm.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step32()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gfm.profile_enable(False)
    fwd_ms, fwd_n = gfm.profile_read("pm_fwd")
    adj_ms, adj_n = gfm.profile_read("pm_adj")
    gfm.profile_reset()

    ms_step = (t1 - t0) * 1e3 / args.steps
    value = 2.0 * pairs * args.steps / (t1 - t0)

    # FP64 GPU path, same step
    value64 = None
    err = None
    g32h, a32h = g32.cpu().numpy(), a32.cpu().numpy()
    if args.fp64_steps > 0:
        step64()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.fp64_steps):
            step64()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        value64 = 2.0 * pairs * args.fp64_steps / (t3 - t2)
        g64h, a64h = g64.cpu().numpy(), a64.cpu().numpy()
        err = {"forward": relerr(g32h, g64h), "adjoint": relerr(a32h, a64h)}

    fwd_avg, adj_avg = fwd_ms / max(fwd_n, 1), adj_ms / max(adj_n, 1)
    dom, dom_ms = ("pm_adj", adj_avg) if adj_avg >= fwd_avg else ("pm_fwd", fwd_avg)
    achieved = FLOP_PER_PAIR * pairs / (dom_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom, {}).get("hbm_bytes_per_launch")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Gaussian-blob densities, Gaussian-hill station heights)",
        "config": {"workload": "pm_blobs_1m: FP32 forward + adjoint, 128x128x64 cells of 500 m "
                               "(1,048,576) x 16,384 stations, point-mass cells",
                   "cells": w.grid.n_cells, "stations": w.n_stations,
                   "pairs_per_step": 2 * pairs},
        "fp64_gpu": {"value": value64, "unit": "pairs/s", "steps": args.fp64_steps},
        "speedup_vs_fp64_gpu": value / value64 if value64 else None,
        "rel_err_vs_fp64": err,
        "kernels_ms": {"pm_fwd": fwd_avg, "pm_adj": adj_avg},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic, "kernel": dom,
                     "flop_per_pair": FLOP_PER_PAIR, "pairs_per_launch": pairs},
    }
    if not args.no_cpu_baseline:
        base, cerr = cpu_baseline(w, r, g32h, a32h)
        line["cpu_baseline"] = base
        line["rel_err_vs_oracle_sample"] = cerr
        line["speedup_vs_cpu"] = value / base["value"]
    if not args.no_extras:
        line["stored"] = stored_leg(plan, rho32, r32, g64h if err else None, stream, args)
    plan.close()
    del g64, a64, rho64, r64
    if not args.no_extras:
        line["prism"] = prism_leg(w, r, stream, args)
        line["batched"] = batched_leg(stream, args)
        line["uneven_batch"] = uneven_leg(stream, args)
        line["tfm_fm"] = tfm_leg(stream, args)
        line["gn_step"] = gn_leg(stream, args)
    line["summary"] = summary(line)
    print(json.dumps(line))
    return 0


def summary(line):
    """The speedups of every forward + adjoint path over its FP64 GPU
    counterpart and over the CPU baseline, with the largest relative error of
    each, gathered from the legs above (None where a leg did not run)."""
    def get(d, *keys):
        for k in keys:
            if not isinstance(d, dict) or k not in d:
                return None
            d = d[k]
        return d

    def worst(e):
        return max(e.values()) if isinstance(e, dict) else e

    out = {"fwd_adj_vs_fp64_gpu": {
               "point_mass_fp32": line.get("speedup_vs_fp64_gpu"),
               "prism_fp32": get(line, "prism", "speedup_vs_fp64_gpu"),
               "prism_fp32_vs_fp64_same_algorithm": get(line, "prism",
                                                        "speedup_vs_fp64_same_algorithm"),
               "ensemble_split3": get(line, "batched", "split3", "forward_plus_adjoint",
                                      "speedup_vs_fp64_gpu"),
               "ensemble_split2": get(line, "batched", "split2", "forward_plus_adjoint",
                                      "speedup_vs_fp64_gpu"),
               "ensemble_fp16x2": get(line, "batched", "fp16x2", "forward_plus_adjoint",
                                      "speedup_vs_fp64_gpu"),
               "ensemble_fp16x2_vs_fp64_same_algorithm": get(
                   line, "batched", "fp16x2", "forward_plus_adjoint",
                   "speedup_vs_fp64_same_algorithm")},
           "point_mass_vs_cpu_baseline": line.get("speedup_vs_cpu"),
           "gn_step_vs_cpu_baseline": get(line, "gn_step", "speedup_vs_cpu"),
           "gn_step_normal_equations_mfma_frac": get(line, "gn_step", "roofline", "frac"),
           "max_rel_err_vs_fp64": {
               "point_mass_fp32": worst(line.get("rel_err_vs_fp64")),
               "prism_fp32": worst(get(line, "prism", "rel_err_vs_fp64")),
               "ensemble_split3": max((v for v in (get(line, "batched", "split3", "rel_err_vs_fp64"),
                                                   get(line, "batched", "split3", "adjoint",
                                                       "rel_err_vs_fp64")) if v is not None),
                                      default=None),
               "ensemble_split2": max((v for v in (get(line, "batched", "split2", "rel_err_vs_fp64"),
                                                   get(line, "batched", "split2", "adjoint",
                                                       "rel_err_vs_fp64")) if v is not None),
                                      default=None),
               "ensemble_fp16x2": max((v for v in (get(line, "batched", "fp16x2", "rel_err_vs_fp64"),
                                                   get(line, "batched", "fp16x2", "adjoint",
                                                       "rel_err_vs_fp64")) if v is not None),
                                      default=None)}}
    return out


def _progress(msg):
    """One line on stderr as each part of a leg starts (the JSON line stays on
    stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _timed(fn, reps, sync):
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return (time.perf_counter() - t0) / reps


def stored_leg(plan, rho32, r32, g64h, stream, args):
    """Stored-matrix variant on the same plan: G (FP32, FP64-evaluated) is
    materialised in HBM once, then forward/adjoint stream it (HBM-bound)."""
    import torch
    from ztm_amd import gfm
    nbytes = plan.store_bytes()
    free, _ = torch.cuda.mem_get_info()
    if nbytes > 0.9 * free:
        return {"skipped": f"needs {nbytes / 1e9:.1f} GB, {free / 1e9:.1f} GB free"}
    t0 = time.perf_counter()
    plan.store(stream=stream)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    g = torch.empty(plan.n_stations, dtype=torch.float64, device=rho32.device)
    a = torch.empty(plan.n_cells, dtype=torch.float32, device=rho32.device)
    gfm.profile_reset()
    gfm.profile_enable(True)
    t_step = _timed(lambda: (plan.forward_stored(rho32, out=g, stream=stream),
                             plan.adjoint_stored(r32, out=a, stream=stream)),
                    max(2, args.steps // 4), torch.cuda.synchronize)
    gfm.profile_enable(False)
    f_ms, f_n = gfm.profile_read("stored_fwd")
    a_ms, a_n = gfm.profile_read("stored_adj")
    gfm.profile_reset()
    pairs = plan.n_stations * plan.n_cells
    f_avg, a_avg = f_ms / max(f_n, 1), a_ms / max(a_n, 1)
    out = {"value": 2.0 * pairs / t_step, "unit": "pairs/s", "G_bytes": nbytes,
           "build_s": build_s, "kernels_ms": {"stored_fwd": f_avg, "stored_adj": a_avg},
           "roofline": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                        "achieved": {"stored_fwd": nbytes / (f_avg * 1e-3) / 1e9,
                                     "stored_adj": nbytes / (a_avg * 1e-3) / 1e9},
                        "algorithmic_bytes_per_launch": nbytes}}
    out["roofline"]["frac"] = {k: v / PEAK_HBM_GBS for k, v in out["roofline"]["achieved"].items()}
    if g64h is not None:
        out["rel_err_vs_fp64"] = relerr(g.cpu().numpy(), g64h)
    plan.store_release()
    return out


PEAK_BF16_TFLOPS = 2500.0     # MI355X dense BF16 (and FP16) MFMA, published
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 (matrix and vector), published


def batched_leg(stream, args):
    """Ensemble forward and adjoint on split-BF16 MFMA (ensemble_256: 96x96x48
    cells of 1 km, 9,216 stations at z = +10 m, 256 smoothed models, 256
    residual vectors N(0, 0.5)), BF16 split 3 and 2 and the FP16 hi/lo split,
    against the FP64 single-model
    kernels on models / vectors 0, 85, 170, 255 (extrapolated to 256 for the
    FP64 rates) and against the FP64 ensemble (the same algorithm in FP64)."""
    _progress("batched leg")
    import torch
    from ztm_amd import gfm, workloads as W
    dev = torch.device("cuda:0")
    w = W.ensemble_256()
    B = w.rho.shape[0]
    P = torch.from_numpy(w.rho).to(dev)
    pm = w.n_stations * w.grid.n_cells * B                  # pair-models per call
    out = {"workload": "ensemble_256: 96x96x48 cells (442,368) x 9,216 stations x 256 models",
           "unit": "pair-models/s", "pair_models_per_call": pm}
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        Y = torch.empty(B, w.n_stations, dtype=torch.float64, device=dev)
        check = (0, 85, 170, 255)
        # FP64 single-model forward and adjoint on 4 models / vectors: inputs and
        # outputs resident, the calls timed back to back (no copies inside)
        x64 = {b: torch.from_numpy(w.rho[b].astype(np.float64)).to(dev) for b in check}
        y64 = {b: torch.empty(w.n_stations, dtype=torch.float64, device=dev) for b in check}
        t64 = _timed(lambda: [p.forward(x64[b], out=y64[b], stream=stream) for b in check], 1,
                     torch.cuda.synchronize)
        g64 = {b: y64[b].cpu().numpy() for b in check}
        out["fp64_gpu"] = {"value": pm / (t64 / len(check) * B), "unit": "pair-models/s",
                           "kind": "FP64 single-model kernel, timed on 4 models (resident, "
                                   "no copies in the timed region)"}
        R = np.stack([W.normal(W.SEED0 + 40, b, w.n_stations, 0.5) for b in range(B)]
                     ).astype(np.float32)
        Rt = torch.from_numpy(R).to(dev)
        X = torch.empty(B, w.grid.n_cells, dtype=torch.float32, device=dev)
        r64 = {b: torch.from_numpy(R[b].astype(np.float64)).to(dev) for b in check}
        z64 = {b: torch.empty(w.grid.n_cells, dtype=torch.float64, device=dev) for b in check}
        ta64 = _timed(lambda: [p.adjoint(r64[b], out=z64[b], stream=stream) for b in check], 1,
                      torch.cuda.synchronize)
        a64 = {b: z64[b].cpu().numpy() for b in check}
        out["fp64_gpu_adjoint"] = {"value": pm / (ta64 / len(check) * B),
                                   "unit": "pair-models/s",
                                   "kind": "FP64 single-vector adjoint, timed on 4 vectors "
                                           "(resident, no copies in the timed region)"}
        fp64_fa = 2.0 * pm / ((t64 + ta64) / len(check) * B)
        # the same algorithm in FP64: G once per tile in FP64, reused across the
        # 256 models on FP64 MFMA (inputs are the same values in double)
        _progress("ensemble FP64 (same algorithm)")
        P64 = P.double()
        R64 = Rt.double()
        Y64 = torch.empty(B, w.n_stations, dtype=torch.float64, device=dev)
        X64 = torch.empty(B, w.grid.n_cells, dtype=torch.float64, device=dev)
        same = {}
        for name, fn, kname in (
                ("forward", lambda: p.forward_batched64(P64, out=Y64, stream=stream),
                 "batched64_fwd"),
                ("adjoint", lambda: p.adjoint_batched64(R64, out=X64, stream=stream),
                 "batched64_adj")):
            gfm.profile_reset()
            gfm.profile_enable(True)
            tt = _timed(fn, 2, torch.cuda.synchronize)
            gfm.profile_enable(False)
            ms, n = gfm.profile_read(kname)
            gfm.profile_reset()
            km = ms / max(n, 1)
            flop = 2.0 * pm
            ref = g64 if name == "forward" else a64
            got = (Y64 if name == "forward" else X64).cpu().numpy()
            same[name] = {"value": pm / tt, "ms_per_call": tt * 1e3, "kernel_ms": km,
                          "rel_err_vs_fp64_single": max(relerr(got[b], ref[b]) for b in check),
                          "roofline": {"bound": "mfma", "unit": "TFLOP/s",
                                       "peak": PEAK_FP64_TFLOPS,
                                       "achieved": flop / (km * 1e-3) / 1e12,
                                       "frac": flop / (km * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                                       "flop": "2 x pair-models on v_mfma_f64_16x16x4_f64"}}
        del P64, R64, Y64, X64
        t_same = same["forward"]["ms_per_call"] + same["adjoint"]["ms_per_call"]
        same["forward_plus_adjoint"] = {"value": 2.0 * pm / (t_same * 1e-3),
                                        "ms_per_pair_of_calls": t_same}
        same["kind"] = ("FP64 ensemble: G generated once per tile in FP64 and reused across "
                        "the batch on FP64 MFMA (ztm_plan_forward/adjoint_batched64)")
        out["fp64_gpu_same_algorithm"] = same
        for split in (3, 2, 4):
            _progress(f"ensemble split {split}")
            p.set_batched_tuning(split=split)
            t0 = time.perf_counter()
            p.autotune_batched(B, reps=3, stream=stream)
            t_tune = time.perf_counter() - t0
            products = 6 if split == 3 else 3
            gfm.profile_reset()
            gfm.profile_enable(True)
            t = _timed(lambda: p.forward_batched(P, out=Y, stream=stream), 3,
                       torch.cuda.synchronize)
            gfm.profile_enable(False)
            ms, n = gfm.profile_read("batched_fwd")
            gfm.profile_reset()
            kms = ms / max(n, 1)
            mfma_flop = 2.0 * pm * products
            Yh = Y.cpu().numpy()
            leg = {"value": pm / t, "ms_per_call": t * 1e3, "kernel_ms": kms,
                   "rel_err_vs_fp64": max(relerr(Yh[b], g64[b]) for b in check),
                   "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": PEAK_BF16_TFLOPS,
                                "achieved": mfma_flop / (kms * 1e-3) / 1e12,
                                "frac": mfma_flop / (kms * 1e-3) / 1e12 / PEAK_BF16_TFLOPS,
                                "flop": "issued MFMA: 2 x pair-models x products",
                                "products": products}}
            leg["speedup_vs_fp64_gpu"] = leg["value"] / out["fp64_gpu"]["value"]
            gfm.profile_reset()
            gfm.profile_enable(True)
            ta = _timed(lambda: p.adjoint_batched(Rt, out=X, stream=stream), 3,
                        torch.cuda.synchronize)
            gfm.profile_enable(False)
            ms, n = gfm.profile_read("batched_adj")
            gfm.profile_reset()
            kma = ms / max(n, 1)
            Xh = X.cpu().numpy()
            leg["adjoint"] = {
                "value": pm / ta, "ms_per_call": ta * 1e3, "kernel_ms": kma,
                "rel_err_vs_fp64": max(relerr(Xh[b], a64[b]) for b in check),
                "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": PEAK_BF16_TFLOPS,
                             "achieved": mfma_flop / (kma * 1e-3) / 1e12,
                             "frac": mfma_flop / (kma * 1e-3) / 1e12 / PEAK_BF16_TFLOPS},
                "speedup_vs_fp64_gpu": (pm / ta) / out["fp64_gpu_adjoint"]["value"]}
            fa = 2.0 * pm / (t + ta)
            leg["forward_plus_adjoint"] = {
                "value": fa, "unit": "pair-models/s", "ms_per_pair_of_calls": (t + ta) * 1e3,
                "speedup_vs_fp64_gpu": fa / fp64_fa,
                "speedup_vs_fp64_same_algorithm":
                    fa / same["forward_plus_adjoint"]["value"]}
            leg["autotune"] = {"s": t_tune,
                               "forward": dict(zip(("variant", "waves", "slices"),
                                                   p.batched_kernel(0, B))),
                               "adjoint": dict(zip(("variant", "waves", "slices"),
                                                   p.batched_kernel(1, B)))}
            out["fp16x2" if split == 4 else f"split{split}"] = leg
        out["default_split"] = "fp16x2"
    return out


def uneven_leg(stream, args, iters=10, solves=2):
    """uneven_48: 48 independent problems (cells 2e4-2e6, stations 1e3-4e4,
    edges 250-2,000 m) solved together: 10 CG iterations of the regularised
    normal equations per problem (one G^T d pass + 10 x (forward + adjoint)),
    every pass one persistent launch over all 48, the whole solve one HIP
    graph, tiles autotuned per size class. Compared with the same passes run
    problem by problem (FP32 plans and FP64 plans, eager)."""
    _progress("uneven leg")
    import torch
    from ztm_amd import gfm, workloads as W
    dev = torch.device("cuda:0")
    ps = W.uneven_48()
    pairs = sum(p.pairs for p in ps)
    passes = 1 + 2 * iters
    out = {"workload": "uneven_48: 48 problems, cells log-U[2e4,2e6], stations log-U[1e3,4e4], "
                       f"edge log-U[250,2000] m; {iters} CG iterations (fwd + adj) each",
           "unit": "pairs/s", "pairs_per_pass": pairs, "passes_per_solve": passes,
           "cells": int(sum(p.grid.n_cells for p in ps)),
           "stations": int(sum(p.stations.shape[0] for p in ps))}
    lam = np.array([p.lam for p in ps])
    with gfm.Batch([(p.grid, p.stations, p.kappa) for p in ps]) as b:
        t0 = time.perf_counter()
        b.autotune(reps=2, stream=stream)
        torch.cuda.synchronize()
        out["autotune_s"] = time.perf_counter() - t0
        rho_t = b.pack_cells([p.rho for p in ps], device=dev)
        g = b.forward(rho_t, stream=stream)
        d = (g + b.pack_stations([p.noise for p in ps], dtype=torch.float64, device=dev)).float()
        x = torch.zeros(b.total_cells, dtype=torch.float32, device=dev)
        rr = torch.zeros(iters + 1, b.n_problems, dtype=torch.float64, device=dev)
        b.solve(d, lam, iters, x=x, rr=rr, use_graph=True, stream=stream)   # capture + run
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(solves):
            b.solve(d, lam, iters, x=x, rr=rr, use_graph=True, stream=stream)
        torch.cuda.synchronize()
        t_graph = (time.perf_counter() - t0) / solves
        t_eager = _timed(lambda: b.solve(d, lam, iters, x=x, rr=rr, use_graph=False,
                                         stream=stream), 1, torch.cuda.synchronize)
        gfm.profile_reset()
        gfm.profile_enable(True)
        r_one = torch.ones(b.total_stations, dtype=torch.float32, device=dev)
        gb = torch.empty(b.total_stations, dtype=torch.float64, device=dev)
        ab = torch.empty(b.total_cells, dtype=torch.float32, device=dev)
        t_pass = _timed(lambda: (b.forward(x, out=gb, stream=stream),
                                 b.adjoint(r_one, out=ab, stream=stream)), 3,
                        torch.cuda.synchronize)
        gfm.profile_enable(False)
        fms, fn = gfm.profile_read("pm_fwd_grouped")
        ams, an = gfm.profile_read("pm_adj_grouped")
        gfm.profile_reset()
        # solution quality
        gx = b.forward(x, stream=stream)
        rrh = rr.cpu().numpy()
        res = b.unpack_stations((gx - d.double()).cpu().numpy())
        dn = b.unpack_stations(d.cpu().numpy())
        xs = b.unpack_cells(x.cpu().numpy())
        gbh = b.unpack_stations(gb.cpu().numpy())
        xh = xs
        tunings = [b.tuning(i) for i in range(len(ps))]

    rec = [W.recovery_stats(xs[i], p) for i, p in enumerate(ps)]
    corr = [r["corr"] for r in rec]
    chi2 = [float(res[i] @ res[i] / (len(res[i]) * 0.1 ** 2)) for i in range(len(ps))]
    # problem by problem, eager: FP32 and FP64 plans, forward + adjoint of the
    # same vectors, and the FP32 batch result against FP64
    t32 = t64 = 0.0
    err = 0.0
    for i, p in enumerate(ps):
        with gfm.Plan(p.grid, p.stations, kappa=p.kappa) as pl:
            x32 = torch.from_numpy(xh[i]).to(dev)
            x64 = x32.double()
            r32 = torch.ones(p.stations.shape[0], dtype=torch.float32, device=dev)
            r64 = r32.double()
            t32 += _timed(lambda: (pl.forward(x32, stream=stream), pl.adjoint(r32, stream=stream)),
                          1, torch.cuda.synchronize)
            t64 += _timed(lambda: (pl.forward(x64, stream=stream), pl.adjoint(r64, stream=stream)),
                          1, torch.cuda.synchronize)
            err = max(err, relerr(gbh[i], pl.forward(x64, stream=stream).cpu().numpy()))
    fwd_k, adj_k = fms / max(fn, 1), ams / max(an, 1)
    out.update({
        "value": passes * pairs / t_graph,
        "ms_per_solve_graph": t_graph * 1e3,
        "ms_per_solve_eager": t_eager * 1e3,
        "ms_per_pass_batched": t_pass * 1e3,
        "ms_per_pass_per_problem_fp32": t32 * 1e3,
        "ms_per_pass_per_problem_fp64": t64 * 1e3,
        "speedup_vs_fp64_gpu": t64 / t_pass,
        "speedup_vs_per_problem_fp32": t32 / t_pass,
        "rel_err_vs_fp64": err,
        "kernels_ms": {"pm_fwd_grouped": fwd_k, "pm_adj_grouped": adj_k},
        "roofline": {"bound": "valu", "unit": "TFLOP/s", "peak": PEAK_FP32_TFLOPS,
                     "achieved": FLOP_PER_PAIR * pairs / (max(fwd_k, adj_k) * 1e-3) / 1e12,
                     "frac": FLOP_PER_PAIR * pairs / (max(fwd_k, adj_k) * 1e-3) / 1e12
                     / PEAK_FP32_TFLOPS,
                     "kernel": "pm_adj_grouped" if adj_k >= fwd_k else "pm_fwd_grouped"},
        "solution": {"rr_ratio_max": float(np.max(rrh[-1] / rrh[0])),
                     "misfit_ratio_median": float(np.median(
                         [np.linalg.norm(res[i]) / np.linalg.norm(dn[i]) for i in range(len(ps))])),
                     "chi2_per_station_median": float(np.median(chi2)),
                     "column_mass_corr_median": float(np.median(corr)),
                     "recovery": {k: {"median": float(np.median([r[k] for r in rec])),
                                      "min": float(np.min([r[k] for r in rec])),
                                      "max": float(np.max([r[k] for r in rec]))}
                                  for k in ("mass", "contrast", "offset")}},
        "tiles": {"st_per_lane": sorted({t.st_per_lane for t in tunings}),
                  "cells_per_lane": sorted({t.cells_per_lane for t in tunings}),
                  "size_classes": sorted({t.size_class for t in tunings})},
    })
    return out


def tfm_leg(stream, args, reps=20):
    """One gravity forward model of a ZTM FM call at the reference's own grid
    (TFM.cpp:83-94; tfm_fm: 61x61x33 coarse, fh=1, fv=2 -> 241,865 fine cells,
    1,681 interior stations): the density producer and GF on the device with
    inputs resident in HBM, the same through the host entry point (PCIe in
    and out), the GFM program behind the reference's file contract (process
    start, files, PCIe), and the reference's CPU path restated (producer +
    TFM.cpp:95-99 gravity with one thread per station row, and on 16
    threads) on this host."""
    _progress("tfm leg")
    import subprocess
    import tempfile
    import torch
    import oracle
    from ztm_amd import gfm, workloads as W
    dev = torch.device("cuda:0")
    w = W.tfm_fm()
    M = w.di * w.dj
    ni, nj = w.di - 2 * w.bfr, w.dj - 2 * w.bfr
    pairs = w.n_fine * ni * nj
    t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    sea, moho, vel = t64(w.sea), t64(w.moho), t64(w.vel)
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    drho = torch.empty(w.n_fine, dtype=torch.float32, device=dev)
    g = torch.zeros(M, dtype=torch.float32, device=dev)

    def fm_device():
        gfm.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, sea, moho, vel, w.vm, w.ro, w.dr,
                     out=drho, stream=stream)
        g.zero_()
        gfm.GF(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, g, drho, s32, stream=stream)

    gfm.profile_reset()
    gfm.profile_enable(True)
    t_dev = _timed(fm_device, reps, torch.cuda.synchronize)
    gfm.profile_enable(False)
    p_ms, p_n = gfm.profile_read("tfm_drho")
    f_ms, f_n = gfm.profile_read("pm_fwd")
    gfm.profile_reset()
    g_dev = g.cpu().numpy().astype(np.float64)
    host_args = (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel, w.surf, w.vm,
                 w.ro, w.dr)
    t_host = _timed(lambda: gfm.tfm_gravity_host(*host_args), reps, lambda: None)
    G_host = gfm.tfm_gravity_host(*host_args)
    # the GFM program on the reference's files (R.dat / S.dat -> G.dat)
    exe = os.path.join(ROOT, "ztm_amd", "bin", "GFM")
    t_prog = None
    with tempfile.TemporaryDirectory() as tmp:
        drho.cpu().numpy().tofile(os.path.join(tmp, "R.dat"))
        w.surf.astype(np.float32).tofile(os.path.join(tmp, "S.dat"))
        env = dict(os.environ, ZTM_GFM_DIR=tmp)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            ok = subprocess.run([exe], env=env, capture_output=True).returncode == 0
            runs.append(time.perf_counter() - t0)
            if not ok:
                break
        if ok:
            t_prog = min(runs)
            G_prog = np.fromfile(os.path.join(tmp, "G.dat"), dtype=np.float32)
    # the reference's CPU path, restated (producer + gravity lambda)
    t0 = time.perf_counter()
    d64, nn = oracle.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel, w.vm,
                              w.ro, w.dr)
    t_prod_cpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    G_cpu = oracle.tfm_gravity(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, d64, w.surf, nn, threads=0)
    t_grav_rows = time.perf_counter() - t0
    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    oracle.tfm_gravity(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, d64, w.surf, nn, threads=cores)
    t_grav_cores = time.perf_counter() - t0
    interior = np.zeros((w.dj, w.di), bool)
    interior[w.bfr:w.dj - w.bfr, w.bfr:w.di - w.bfr] = True
    interior = interior.ravel()
    err = relerr(g_dev[interior], G_cpu[interior])
    # finite-difference Jacobian columns: 256 variants (gravity parameters and
    # Moho perturbed per variant) in one batched call
    B = 256
    prm = np.array([[w.vm + 1e-3 * (b % 7), w.ro - 1e-3 * (b % 5), w.dr + 1e-3 * (b % 3)]
                    for b in range(B)])
    bump = np.sin(np.arange(M) * 0.37)
    mohoB = t64(np.stack([w.moho + 0.01 * (b % 11) * bump for b in range(B)]))
    prmB = t64(prm)
    gB = torch.empty(B, M, dtype=torch.float64, device=dev)
    dB = torch.empty(B, w.n_fine, dtype=torch.float32, device=dev)
    t_batch = _timed(lambda: gfm.tfm_gravity_batched(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc,
                                                     sea, mohoB, vel, prmB, s32, drho=dB, g=gB,
                                                     stream=stream), 3, torch.cuda.synchronize)
    # exact gravity columns for (vm, ro, dr): derivative fields + one forward
    ddB = torch.empty(3, w.n_fine, dtype=torch.float32, device=dev)
    dgB = torch.empty(3, M, dtype=torch.float64, device=dev)
    t_jvp = _timed(lambda: gfm.tfm_gravity_jvp(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, sea,
                                               t64(w.moho), vel, w.vm, w.ro, w.dr, s32,
                                               ddrho=ddB, dg=dgB, stream=stream), 5,
                   torch.cuda.synchronize)
    rB = t64(W.normal(W.SEED0 + 31, 1, M, 1.0))
    vjB = torch.empty(3, dtype=torch.float64, device=dev)
    t_vjp = _timed(lambda: gfm.tfm_gravity_vjp(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, sea,
                                               t64(w.moho), vel, w.vm, w.ro, w.dr, s32, rB,
                                               ddrho=ddB, out=vjB, stream=stream), 5,
                   torch.cuda.synchronize)
    # the Moho and velocity grids: the gradient of r . G over every node in one
    # adjoint pass, and 256 node columns in one tangent call (TFM.cpp:134 runs
    # one forward model per node)
    n_nodes = M + M * w.dk
    t_vjp_grids = _timed(lambda: gfm.tfm_gravity_vjp_grids(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv,
                                                           w.rc, sea, moho, vel, w.vm, w.ro,
                                                           w.dr, s32, rB, stream=stream), 5,
                         torch.cuda.synchronize)
    dmo = torch.zeros(B, M, dtype=torch.float64, device=dev)
    dve = torch.zeros(B, M * w.dk, dtype=torch.float64, device=dev)
    for b in range(B):   # unit directions on 128 Moho and 128 velocity nodes
        if b < 128:
            dmo[b, (b * 29) % M] = 1.0
        else:
            dve[b, (b * 977) % (M * w.dk)] = 1.0
    ddG = torch.empty(B, w.n_fine, dtype=torch.float32, device=dev)
    dgG = torch.empty(B, M, dtype=torch.float64, device=dev)
    t_jvp_grids = _timed(lambda: gfm.tfm_gravity_jvp_grids(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv,
                                                           w.rc, sea, moho, vel, w.vm, w.ro,
                                                           w.dr, s32, dmoho=dmo, dvel=dve,
                                                           ddrho=ddG, dg=dgG, stream=stream), 3,
                         torch.cuda.synchronize)
    out = {
        "workload": "tfm_fm: TFM.cpp:83-94 at the reference grid, 241,865 fine cells x 1,681 "
                    "interior stations, producer + gravity per FM call",
        "unit": "ms per FM call", "pairs_per_call": pairs, "kept_cells": int(nn.size),
        "ms_device_resident": t_dev * 1e3,
        "kernels_ms": {"tfm_drho": p_ms / max(p_n, 1), "pm_fwd": f_ms / max(f_n, 1)},
        "ms_host_api": t_host * 1e3,
        "jacobian_256_variants": {"ms_per_call": t_batch * 1e3,
                                  "ms_per_variant": t_batch * 1e3 / B,
                                  "vs_single_device_calls": t_dev * B / t_batch},
        "gravity_jvp_3_params": {"ms_per_call": t_jvp * 1e3,
                                 "note": "exact dG/d(vm, ro, dr): derivative fields + forward"},
        "gravity_vjp_3_params": {"ms_per_call": t_vjp * 1e3,
                                 "note": "r . dG/d(vm, ro, dr): derivative fields + adjoint"},
        "gravity_vjp_grids": {"ms_per_call": t_vjp_grids * 1e3, "nodes": n_nodes,
                              "fd_equivalent_ms": n_nodes * t_dev * 1e3,
                              "note": "d(r . G)/d(Moho, velocity) over every node: one adjoint "
                                      "+ the transpose of the tangent map; the FD Jacobian "
                                      "needs one FM call per node"},
        "gravity_jvp_grids_256": {"ms_per_call": t_jvp_grids * 1e3,
                                  "ms_per_column": t_jvp_grids * 1e3 / B,
                                  "note": "256 exact node columns (128 Moho, 128 velocity): "
                                          "tangent fields + split MFMA ensemble forward"},
        "ms_gfm_program": None if t_prog is None else t_prog * 1e3,
        "cpu_port": {"ms_producer": t_prod_cpu * 1e3,
                     "ms_gravity_thread_per_row": t_grav_rows * 1e3, "rows": nj,
                     "ms_gravity_cores": t_grav_cores * 1e3, "cores": cores,
                     "kind": "port", "sample": "full call (producer + gravity), FP64 pow"},
        "speedup_vs_cpu_port": (t_prod_cpu + t_grav_rows) / t_dev,
        "rel_err_vs_cpu_path": err,
        "host_api_equals_device": bool(np.array_equal(G_host, g_dev)),
    }
    if t_prog is not None:
        out["gfm_program_equals_device"] = bool(np.array_equal(G_prog.astype(np.float64), g_dev))
    return out


def accumulation_leg(stream, args):
    """Error of the ensemble forward against the K-slice count: each slice is
    one FP32 accumulation run on the MFMAs (N / 16 / slices chunk steps),
    summed across slices in FP64. ensemble_256 at half size (48x48x24 cells,
    2,304 stations), 64 models; error = max over models 0, 21, 63 of
    max |Y - Y64| / max |Y64| against the FP64 single-model forward."""
    _progress("accumulation leg")
    import torch
    from ztm_amd import gfm, workloads as W
    dev = torch.device("cuda:0")
    w = W.ensemble_256(n=48, nz=24)
    Pt = torch.from_numpy(w.rho[:64].copy()).to(dev)
    n_chunks = (w.grid.n_cells + 15) // 16
    out = {"workload": "ensemble_256 half size: 48x48x24 cells, 2,304 stations, 64 models",
           "chunks": n_chunks, "rows": []}
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        ref = {b: p.forward(torch.from_numpy(w.rho[b].astype(np.float64)).to(dev),
                            stream=stream).cpu().numpy() for b in (0, 21, 63)}
        for ks in (1, 2, 4, 9, 27, 81, 243, n_chunks):
            row = {"slices": ks, "chunks_per_run": -(-n_chunks // ks)}
            for split, name in ((3, "split3"), (4, "fp16x2"), (2, "split2")):
                p.set_batched_tuning(slices=ks, split=split)
                Y = p.forward_batched(Pt, stream=stream).cpu().numpy()
                row[name] = max(relerr(Y[b], r) for b, r in ref.items())
            out["rows"].append(row)
    return out


def mfma_leg(stream, args):
    """The MFMA pipe's sustained rate per shape (ztm_calib_mfma): every wave
    streams MFMAs on register operands, nothing else in the loop, on random
    and on low-entropy operands. Under the chip's power limit the rate with
    random operands is the ceiling of a data-dependent MFMA kernel."""
    _progress("mfma calibration leg")
    from ztm_amd import gfm
    out = {}
    for op in gfm.CALIB_OPS:
        peak = 2500.0   # dense FP16 / BF16, published
        for low in (False, True):
            v = gfm.calib_mfma(op, low_entropy=low, iters=2000, stream=stream)
            out[f"{op}_{'low_entropy' if low else 'random'}"] = {
                "tflops": v, "peak": peak, "frac": v / peak}
    return out


def lm_leg(stream, args, reps=10):
    """One Levenberg-Marquardt iteration of ZTM's gravity fit on the device
    (ztm_amd.lm.GravityLM) at the reference grid (tfm_fm) with 9 parameters
    (vm, ro, dr and 6 Moho basis coefficients) and the reference's 12
    damping conditions: the forward model, the exact Jacobian, the normal
    equations, the damped steps and the 12 candidates' forward models, with
    inputs resident. Against the reference's CPU iteration restated: its
    forward models, 2 + n + nd = 23 of them (LMA::fit's residual, Jac's one
    per parameter, one per damping condition; TFM.cpp:124-135,157-161),
    each timed as the oracle's producer + TFM.cpp:95-99 gravity on 16
    threads (the dense LA of a 9-parameter step is negligible beside them).
    Then one whole fit, LMA::fit's loop of up to 6 iterations
    (GravityLM.fit, TFM.cpp:157-166), timed once with its host reads."""
    _progress("lm leg")
    import torch
    import oracle
    from ztm_amd import lm, workloads as W
    dev = torch.device("cuda:0")
    w = W.tfm_fm()
    basis = W.tfm_moho_basis(w)
    model = lm.GravityLM(w, basis, dev)
    b0 = np.array([w.vm, w.ro, w.dr] + [0.0] * basis.shape[0])
    d = model.forward(b0 + W.TFM_LM_STEP, stream)[0]
    sync = torch.cuda.synchronize
    out = {}

    def it():
        out["r"] = model.iteration(b0, d, W.TFM_LM_SIGMA, W.GN_DAMPING, stream)

    t_dev = _timed(it, reps, sync)
    _, misfits, best, misfit0 = out["r"]
    n, nd = model.n, W.GN_DAMPING.shape[0]
    c0 = time.perf_counter()
    drho, nn = oracle.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel,
                               w.vm, w.ro, w.dr)
    oracle.tfm_gravity(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, drho, w.surf, nn, threads=16)
    t_fm = time.perf_counter() - c0
    fms = 2 + n + nd
    sync()
    c0 = time.perf_counter()
    bf, Q, nb = model.fit(b0, d, W.TFM_LM_SIGMA, W.GN_DAMPING, iters=6, stream=stream)
    sync()
    t_fit = time.perf_counter() - c0
    return {
        "workload": f"tfm_fm (61x61x33, fh=1, fv=2), {n} parameters, {nd} damping conditions",
        "ms_per_iteration": t_dev * 1e3,
        "misfit": {"start": float(misfit0), "best": float(misfits[best]), "condition": best},
        "cpu_baseline": {"ms_per_iteration": t_fm * fms * 1e3, "forward_models": fms,
                         "ms_per_forward_model": t_fm * 1e3, "cores": 16, "kind": "port",
                         "sample": "one oracle forward model (producer + TFM.cpp:95-99 gravity) "
                                   f"timed, times the iteration's {fms}"},
        "speedup_vs_cpu": t_fm * fms / t_dev,
        "fit": {"iterations": len(Q) - 1, "ms": t_fit * 1e3, "misfit_log": Q,
                "last_condition": nb,
                "max_error_vs_truth": float(np.max(np.abs(bf - b0 - W.TFM_LM_STEP)))},
    }


def gn_lib_leg(stream, args, reps=10):
    """The vendor libraries on the gn_step shapes, through torch (rocBLAS /
    hipBLASLt for the product, rocSOLVER for the factorisation): J^T J as a
    full FP64 product and the Cholesky factors of the 12 damped systems. A
    comparison point for the gn leg, not part of it."""
    _progress("gn library leg")
    import torch
    from ztm_amd import gfm, workloads as W
    dev = torch.device("cuda:0")
    w = W.gn_step()
    rows, n = w.J.shape
    t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    J, r, sig = t64(w.J), t64(w.r), t64(w.sig)
    Jt = J.T.contiguous()
    JJ, _ = gfm.gn_normal(J, r, stream=stream)
    dJ = torch.diagonal(JJ)
    K = torch.stack([JJ + torch.diag(d0 * dJ + d1 / sig ** 2) for d0, d1 in w.dm.tolist()])
    sync = torch.cuda.synchronize
    t_mm = _timed(lambda: torch.mm(Jt, J), reps, sync)
    t_chol = _timed(lambda: torch.linalg.cholesky(K), reps, sync)
    return {
        "workload": f"gn_step: J {rows:,} x {n:,}, {K.shape[0]} damped systems",
        "torch_mm_JtJ_ms": t_mm * 1e3,
        "torch_mm_JtJ_tflops": 2.0 * rows * n * n / t_mm / 1e12,
        "torch_linalg_cholesky_ms": t_chol * 1e3,
        "kind": "torch.mm (FP64, full product) and torch.linalg.cholesky (batched FP64)",
    }


def gn_leg(stream, args, reps=10, cpu_rows=2048):
    """Dense LA of one damped Gauss-Newton step (LMA::fit, TFM.cpp:157-161) on
    gn_step (J 32,768 x 1,024, the reference's 12 damping conditions), FP64 on
    the device: J^T J with J^T r (one symmetric product on FP64 MFMA), the 12
    damped steps (batched blocked Cholesky and substitutions), and
    diag(inv(J^T J + I d)) (create / annihilate, TFM.cpp:150,173). Compared
    with the oracle's FP64 restatement of zmath's MTM and Gauss-Jordan inv on
    the host: MTM timed on the first cpu_rows rows and scaled to all rows
    (its cost is linear in rows), the damped steps timed whole on the GPU's
    J^T J, whose result also checks the device steps."""
    _progress("gn leg")
    import torch
    import oracle
    from ztm_amd import gfm, workloads as W
    dev = torch.device("cuda:0")
    w = W.gn_step()
    rows, n = w.J.shape
    nd = w.dm.shape[0]
    t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    J, r, sig, b, dm = t64(w.J), t64(w.r), t64(w.sig), t64(w.b), t64(w.dm)
    JJ = torch.empty(n, n, dtype=torch.float64, device=dev)
    jtr = torch.empty(n, dtype=torch.float64, device=dev)
    B = torch.empty(nd, n, dtype=torch.float64, device=dev)
    V = torch.empty(n, dtype=torch.float64, device=dev)
    work = torch.empty(gfm.gn_workspace_bytes(rows, n, nd), dtype=torch.uint8, device=dev)
    sync = torch.cuda.synchronize

    def normal():
        gfm.gn_normal(J, r, stream=stream, out=(JJ, jtr), work=work)

    def steps():
        gfm.gn_damped_steps(JJ, jtr, sig, dm, b, stream=stream, out=B, work=work)

    def invdiag():
        gfm.gn_inverse_diag(JJ, 1e-3, stream=stream, out=V, work=work)

    normal()
    steps()
    invdiag()
    sync()
    gfm.profile_reset()
    gfm.profile_enable(True)
    t_normal = _timed(normal, reps, sync)
    t_steps = _timed(steps, reps, sync)
    gfm.profile_enable(False)
    t_inv = _timed(invdiag, reps, sync)
    prof = {k: gfm.profile_read(k) for k in ("gn_syrk", "gn_cholesky", "gn_solve")}
    gfm.profile_reset()
    kms = {k: ms / max(cnt, 1) for k, (ms, cnt) in prof.items()}
    syrk_flop = float(rows) * n * (n + 3)     # 2 per row for each of the n (n + 1) / 2 entries + J^T r
    chol_flop = nd * n ** 3 / 3.0
    JJh, jtrh, Bh, Vh = JJ.cpu().numpy(), jtr.cpu().numpy(), B.cpu().numpy(), V.cpu().numpy()
    err_normal = max(relerr(JJh, w.J.T @ w.J), relerr(jtrh, w.J.T @ w.r))
    # CPU: the oracle's MTM on a row sample, scaled; the damped steps on the
    # device's J^T J (Gauss-Jordan per condition, conditions on threads)
    Js = np.ascontiguousarray(w.J[:cpu_rows])
    c0 = time.perf_counter()
    oracle.mtm(Js)
    c1 = time.perf_counter()
    Bc = oracle.gn_steps(JJh, jtrh, w.sig, w.dm, w.b)
    c2 = time.perf_counter()
    cpu_mtm = (c1 - c0) * rows / cpu_rows
    cpu_ms = (cpu_mtm + (c2 - c1)) * 1e3
    steps_err = max(float(np.max(np.abs(Bh[k] - Bc[k])) / np.max(np.abs(Bc[k] - w.b)))
                    for k in range(nd))
    inv_err = relerr(Vh, oracle.inv_diag(JJh, 1e-3))
    ms_gpu = (t_normal + t_steps) * 1e3
    return {
        "workload": f"gn_step: J {rows:,} data x {n:,} parameters, {nd} damping conditions "
                    "(TFM.cpp:157-161)",
        "unit": "ms per GN step (J^T J, J^T r and the damped steps of every condition)",
        "ms_per_step": ms_gpu,
        "ms_normal": t_normal * 1e3, "ms_damped_steps": t_steps * 1e3,
        "ms_inverse_diag": t_inv * 1e3,
        "kernels_ms": kms,
        "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": PEAK_FP64_MFMA_TFLOPS,
                     "kernel": "gn_syrk",
                     "achieved": syrk_flop / (kms["gn_syrk"] * 1e-3) / 1e12,
                     "frac": syrk_flop / (kms["gn_syrk"] * 1e-3) / 1e12 / PEAK_FP64_MFMA_TFLOPS,
                     "flop": "rows x n x (n+3): the upper triangle of J^T J and J^T r"},
        "cholesky_tflops": chol_flop / (kms["gn_cholesky"] * 1e-3) / 1e12,
        "rel_err": {"normal_vs_host_blas": err_normal, "steps_vs_oracle": steps_err,
                    "inverse_diag_vs_oracle": inv_err},
        "cpu_baseline": {"ms_per_step": cpu_ms, "ms_mtm": cpu_mtm * 1e3,
                         "ms_damped_steps": (c2 - c1) * 1e3,
                         "cores": oracle.DEFAULT_THREADS, "kind": "port",
                         "sample": f"oracle MTM (zmath.cpp:260-262 loop order) on {cpu_rows} of "
                                   f"{rows} rows scaled to all rows, + Gauss-Jordan damped steps "
                                   f"(zmath.cpp:129-140) for {nd} conditions on "
                                   f"{oracle.DEFAULT_THREADS} threads ({cpu_model()})"},
        "speedup_vs_cpu": cpu_ms / ms_gpu,
    }


def prism_leg(w, r, stream, args):
    """Prism cell formula on the same workload: FP32 hybrid (point-mass far
    field, Gauss-Legendre mid zone, FP64 closed-form near field) against the
    FP64 closed form on every pair, forward + adjoint."""
    _progress("prism leg")
    import torch
    from ztm_amd import gfm
    dev = torch.device("cuda:0")
    pairs = w.pairs
    names = ("prism_near_fwd", "pm_fwd", "pm_fwd_reduce", "pm_adj", "pm_adj_reduce")
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=gfm.PRISM) as p:
        rho32 = torch.from_numpy(w.rho.astype(np.float32)).to(dev)
        r32 = torch.from_numpy(r.astype(np.float32)).to(dev)
        g32 = torch.empty(w.n_stations, dtype=torch.float64, device=dev)
        a32 = torch.empty(w.grid.n_cells, dtype=torch.float32, device=dev)
        reps = max(2, args.steps // 4)
        gfm.profile_reset()
        gfm.profile_enable(True)
        t32 = _timed(lambda: (p.forward(rho32, out=g32, stream=stream),
                              p.adjoint(r32, out=a32, stream=stream)),
                     reps, torch.cuda.synchronize)
        gfm.profile_enable(False)
        kms = {}
        for name in names:
            tot, cnt = gfm.profile_read(name)
            kms[name] = tot / max(cnt, 1)
        gfm.profile_reset()
        dom = "pm_adj" if kms["pm_adj"] >= kms["pm_fwd"] else "pm_fwd"
        achieved = FLOP_PER_PAIR * pairs / (kms[dom] * 1e-3) / 1e12
        out = {"value": 2.0 * pairs / t32, "unit": "pairs/s", "ms_per_step": t32 * 1e3,
               "kernels_ms": kms,
               "kernels_share_of_step": {k: v / (t32 * 1e3) for k, v in kms.items()},
               "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                            "traffic": None, "kernel": dom, "flop_per_pair": FLOP_PER_PAIR,
                            "pairs_per_launch": pairs,
                            "note": "far-field kernel (point mass on cubic cells) over every "
                                    "pair; the mid-zone and near-field corrections run in "
                                    "prism_near_fwd and pm_adj_reduce"}}
        if args.fp64_steps > 0:
            rho64 = torch.from_numpy(w.rho).to(dev)
            r64 = torch.from_numpy(r).to(dev)
            g64 = torch.empty(w.n_stations, dtype=torch.float64, device=dev)
            a64 = torch.empty(w.grid.n_cells, dtype=torch.float64, device=dev)
            t64 = _timed(lambda: (p.forward(rho64, out=g64, stream=stream),
                                  p.adjoint(r64, out=a64, stream=stream)),
                         1, torch.cuda.synchronize)
            out["fp64_gpu"] = {"value": 2.0 * pairs / t64, "unit": "pairs/s",
                               "ms_per_step": t64 * 1e3, "kind": "exact closed form, FP64"}
            out["speedup_vs_fp64_gpu"] = t64 / t32
            out["rel_err_vs_fp64"] = {"forward": relerr(g32.cpu().numpy(), g64.cpu().numpy()),
                                      "adjoint": relerr(a32.cpu().numpy(), a64.cpu().numpy())}
            # the same zoned algorithm in FP64: what FP32 itself buys
            gz = torch.empty(w.n_stations, dtype=torch.float64, device=dev)
            az = torch.empty(w.grid.n_cells, dtype=torch.float64, device=dev)
            tz = _timed(lambda: (p.forward(rho64, out=gz, stream=stream, zoned=True),
                                 p.adjoint(r64, out=az, stream=stream, zoned=True)),
                        reps, torch.cuda.synchronize)
            out["fp64_gpu_same_algorithm"] = {
                "value": 2.0 * pairs / tz, "unit": "pairs/s", "ms_per_step": tz * 1e3,
                "kind": "FP64 with the FP32 path's zones (ZTM_FP64_ZONED)",
                "rel_err_vs_fp64": {"forward": relerr(gz.cpu().numpy(), g64.cpu().numpy()),
                                    "adjoint": relerr(az.cpu().numpy(), a64.cpu().numpy())}}
            out["speedup_vs_fp64_same_algorithm"] = tz / t32
    return out


if __name__ == "__main__":
    sys.exit(main())
