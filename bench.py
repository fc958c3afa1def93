"""Benchmark: (k, omega) surface Green's function solves per second at Ndim = 72.

Workload (one step = one pass of the hot path over the whole grid): the
36-orbital synthetic model (R in [-3,3]^3, 343 R vectors, seed 2), Np = 2
principal layers (Ndim = 72), 512 k points on the (-0.5,0)->(0.5,0) path x 512
energies on [-1, 1) eV = 262,144 instances, eta = 1e-3 eV, the reference
decimation stop rule (sum |alpha| <= 1e-16 Ha, itermax 100).  Inputs (model,
k points, energies) are resident on the device before timing starts; a step
is the Fourier sum plus the batched decimation and traces.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32_refine|auto]
    python bench.py --summarize-rocprof TRACE_DIR FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR

Prints one JSON line.  The decimation runs as rounds of two kernels on two
streams (gj_round_kernel and gemm_round_kernel); the roofline figures refer to
that whole decimation stage, timed with HIP events on the first stream, which
waits for the second at the end of the stage.  --summarize-rocprof turns the
rocprofv3 runs of the same command (kernel trace; FETCH_SIZE and WRITE_SIZE
counters; the same counters on a plain 1 GiB copy for calibration) into the
summary kept under profiles/.  The FP64 path always runs (it defines the per-instance
reference iteration count I_ref used for the algorithmic flop count
F = (56 I_ref + 24) N^3 per solve).  The reported `value` is the solves/s of
the fastest path whose A(k, omega) error against FP64 is <= 1e-6.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from wannier_tools_amd import model as M  # noqa: E402

TRAFFIC_JSON = "decimation_fp64_ndim72_512x512.traffic.json"
DECIMATION_KERNELS = ("gj_round_kernel", "gemm_round_kernel")
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
NW, NP_ = 36, 2
NDIM = NW * NP_


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(nk: int, nomega: int):
    mdl = M.model_ndim72()
    irvec, hr = mdl.surface_cell()
    kpts = M.kpath_slab([((-0.5, 0.0), (0.5, 0.0))], nk)
    omegas = M.omega_grid(-1.0, 1.0, nomega) * M.EV2HARTREE
    eta = 1e-3 * M.EV2HARTREE
    return irvec, hr, kpts, omegas, eta


def time_path(gf, precision: str, steps: int, warmup: int):
    for i in range(warmup):
        gf.run_async(gf.eta, 1.0, precision)
        gf.sync()
        log(f"[bench] {precision} warmup {i + 1}/{warmup}: decimation {gf.timing()[1]:.1f} ms")
    dec_ms = []
    t0 = time.perf_counter()
    for i in range(steps):
        gf.run_async(gf.eta, 1.0, precision)
        gf.sync()
        dec_ms.append(gf.timing()[1])
        log(f"[bench] {precision} step {i + 1}/{steps}: decimation {dec_ms[-1]:.1f} ms")
    t1 = time.perf_counter()
    out = gf.download()
    return (t1 - t0) / steps, float(np.mean(dec_ms)), out


# ---- CPU baseline: the reference's own Fortran (oracle/_ref) or the C port --
def _cpu_worker(args):
    kind, idx, kpts, omegas, eta, irvec, hr = args
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    import oracle
    n_done = 0
    if kind == "reference":
        ref = oracle.RefHarness()
        ref.set_model(NW, NP_, irvec, hr)
        cache = {}
        for ik, iw in idx:
            if ik not in cache:
                cache = {ik: ref.ham_qlayer2qlayer(kpts[ik])}
            h00, h01 = cache[ik]
            ref.surfgreen_1985(omegas[iw], eta, h00, h01)
            n_done += 1
    else:
        o = oracle.Oracle()
        cache = {}
        for ik, iw in idx:
            if ik not in cache:
                cache = {ik: o.ham_qlayer2qlayer(NW, NP_, irvec, hr, kpts[ik])}
            h00, h01 = cache[ik]
            o.surfgreen_1985(omegas[iw], eta, h00, h01)
            n_done += 1
    return n_done


def cpu_baseline(kpts, omegas, eta, irvec, hr, per_worker: int, workers: int):
    """Time the reference CPU path on a bounded, evenly strided sample of the grid:
    `workers` processes (the reference's MPI k round-robin), one OpenBLAS thread each."""
    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    nk, nom = kpts.shape[0], omegas.size
    total = per_worker * workers
    flat = np.linspace(0, nk * nom - 1, total).astype(np.int64)
    pairs = [(int(t // nom), int(t % nom)) for t in flat]
    chunks = [pairs[w::workers] for w in range(workers)]
    ctx = mp.get_context("spawn")
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    with ctx.Pool(workers) as pool:
        t0 = time.perf_counter()
        done = sum(pool.map(_cpu_worker, [(kind, c, kpts, omegas, eta, irvec, hr) for c in chunks]))
        dt = time.perf_counter() - t0
    return {
        "value": done / dt,
        "unit": "solves/s",
        "cores": workers,
        "kind": kind,
        "sample": (f"{done} (k, omega) instances strided over the 512x512 grid, "
                   f"{'reference ham_qlayer2qlayer + surfgreen_1985 (Fortran, OpenBLAS)' if kind == 'reference' else 'C FP64 restatement'}"
                   f", {workers} processes x 1 thread, {dt:.1f} s"),
    }


def _csv_rows(d: str, suffix: str):
    import csv
    import glob
    paths = glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)
    if not paths:
        raise SystemExit(f"no *{suffix} under {d}")
    with open(paths[0]) as f:
        return list(csv.DictReader(f))


def _counter_kib(d: str, counter: str, kernels) -> float:
    """Sum of a rocprofv3 counter over the dispatches of the named kernels (KiB)."""
    return sum(float(r["Counter_Value"]) for r in _csv_rows(d, "counter_collection.csv")
               if r["Counter_Name"] == counter and any(k in r["Kernel_Name"] for k in kernels))


def summarize_rocprof(trace, fetch, write, cfetch, cwrite) -> int:
    """profiles/ summary of the decimation stage from rocprofv3 runs of
    `bench.py --steps 1 --warmup 0 --no-cpu-baseline --precision fp64`."""
    stats = _csv_rows(trace, "kernel_stats.csv")
    rows = _csv_rows(trace, "kernel_trace.csv")
    dec = [r for r in rows if any(k in r["Kernel_Name"] for k in DECIMATION_KERNELS)]
    span_ns = max(int(r["End_Timestamp"]) for r in dec) - min(int(r["Start_Timestamp"]) for r in dec)
    busy = {k: sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in dec
                   if k in r["Kernel_Name"]) for k in DECIMATION_KERNELS}
    # counter scale: KiB reported for a copy of known size
    copy_bytes = 1 << 30
    f_scale = copy_bytes / (_counter_kib(cfetch, "FETCH_SIZE", ("copyBuffer",)) * 1024)
    w_scale = copy_bytes / (_counter_kib(cwrite, "WRITE_SIZE", ("copyBuffer",)) * 1024)
    fetch_b = _counter_kib(fetch, "FETCH_SIZE", DECIMATION_KERNELS) * 1024 * f_scale
    write_b = _counter_kib(write, "WRITE_SIZE", DECIMATION_KERNELS) * 1024 * w_scale
    json.dump({"stage": "decimation rounds (gj_round_kernel + gemm_round_kernel)",
               "workload": "ndim72 512x512 eta=1e-3 eV", "fetch_bytes": fetch_b,
               "write_bytes": write_b, "fetch_scale": f_scale, "write_scale": w_scale,
               "span_ns": span_ns}, open(os.path.join(ROOT, "profiles", TRAFFIC_JSON), "w"),
              indent=1)
    print("| Kernel | Calls | Total ms | Average us |\n|---|---|---|---|")
    for r in stats[:8]:
        print(f"| `{r['Name'][:90]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.1f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} |")
    print(f"\nDecimation stage span (first round start to last round end): {span_ns / 1e6:.1f} ms; "
          + "; ".join(f"{k} busy {v / 1e6:.1f} ms" for k, v in busy.items()))
    print(f"\nHBM traffic of the stage: fetch {fetch_b:.4e} B (scale {f_scale:.3f}), "
          f"write {write_b:.4e} B (scale {w_scale:.3f}), {(fetch_b + write_b) / span_ns:.1f} GB/s "
          f"over the span")
    return 0


def main() -> int:
    if len(sys.argv) == 7 and sys.argv[1] == "--summarize-rocprof":
        return summarize_rocprof(*sys.argv[2:])
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "fp32_refine"])
    ap.add_argument("--nk", type=int, default=512)
    ap.add_argument("--nomega", type=int, default=512)
    ap.add_argument("--cpu-workers", type=int, default=16)
    ap.add_argument("--cpu-per-worker", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the surface Green's function path is not "
              "sharded across GPUs in this build (run with --gpus 1)", file=sys.stderr)
        return 2

    from wannier_tools_amd.surface import SurfaceGF

    irvec, hr, kpts, omegas, eta = workload(args.nk, args.nomega)
    ninst = kpts.shape[0] * omegas.size
    gf = SurfaceGF(0)
    gf.set_model_arrays(NW, NP_, irvec, hr)
    gf.upload_grid(kpts, omegas)
    gf.eta = eta

    log(f"[bench] FP64 path: {ninst} instances, Ndim={NDIM}")
    s64, dec64, out64 = time_path(gf, "fp64", args.steps, args.warmup)
    dl64, dr64, db64, it64, st64 = out64
    flops = float(np.sum(56.0 * it64.astype(np.float64) + 24.0)) * NDIM ** 3
    res = {
        "fp64": {"s_per_step": s64, "solves_per_s": ninst / s64, "decimate_ms": dec64,
                 "mean_iters": float(it64.mean()), "max_iters": int(it64.max()),
                 "not_converged": int(np.count_nonzero(st64 & 2)),
                 "singular": int(np.count_nonzero(st64 & 1)),
                 "mfma_tflops_algorithmic": flops / (dec64 * 1e-3) / 1e12},
    }
    log(f"[bench] fp64: {ninst / s64:.1f} solves/s, {s64:.3f} s/step, decimation {dec64:.1f} ms")

    best = ("fp64", s64, dec64, FP64_MFMA_PEAK_TFLOPS)
    if args.precision in ("auto", "fp32_refine"):
        try:
            s32, dec32, out32 = time_path(gf, "fp32_refine", args.steps, args.warmup)
        except Exception as e:  # path not available in this build
            log(f"[bench] fp32_refine unavailable: {e}")
        else:
            dl, dr, db = out32[:3]
            err = max(float(np.max(np.abs(x - y) / np.abs(y))) for x, y in
                      ((dl, dl64), (dr, dr64), (db, db64)))
            med = float(np.median(np.concatenate([np.abs(x - y).ravel() / np.abs(y).ravel() for x, y in
                                                  ((dl, dl64), (dr, dr64), (db, db64))])))
            rel = np.concatenate([np.abs(x - y).ravel() / np.abs(y).ravel() for x, y in
                                  ((dl, dl64), (dr, dr64), (db, db64))])
            res["fp32_refine"] = {"s_per_step": s32, "solves_per_s": ninst / s32, "decimate_ms": dec32,
                                  "max_rel_err_A": err, "median_rel_err_A": med,
                                  "frac_within_1e-6": float(np.mean(rel <= 1e-6)),
                                  "speedup_vs_fp64": s64 / s32}
            log(f"[bench] fp32_refine: {ninst / s32:.1f} solves/s, max rel err {err:.2e}")
            if err <= 1e-6 and s32 < best[1]:
                best = ("fp32_refine", s32, dec32, FP32_MFMA_PEAK_TFLOPS)

    name, s_step, dec_ms, peak = best
    achieved = flops / (dec_ms * 1e-3) / 1e12
    # HBM bytes per launch of this kernel on this workload, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a 1 GiB copy calibration)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", TRAFFIC_JSON)
    if name == "fp64" and (args.nk, args.nomega) == (512, 512) and os.path.exists(tpath):
        t = json.load(open(tpath))
        traffic = t["fetch_bytes"] + t["write_bytes"]
    out = {
        "metric": "G(k,w) solves/s @Ndim=72: FP32+refine vs FP64 GPU, err<=1e-6; %MFMA peak",
        "value": ninst / s_step,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if name == "fp64" else "c64+c128-refine",
        "data": "synthetic",
        "config": {"workload": "surface GF k-path x energy grid, 36-orbital synthetic model, Np=2",
                   "ndim": NDIM, "num_wann": NW, "np": NP_, "nr": int(irvec.shape[0]),
                   "nk": int(kpts.shape[0]), "nomega": int(omegas.size), "instances": ninst,
                   "eta_ev": 1e-3, "stop_rule": "sum|alpha|<=1e-16 Ha, itermax 100",
                   "path": name},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "kernel": ("decimation stage: gj_round_kernel<5, %s> + gemm_round_kernel<5, %s> "
                                "rounds on two streams" % ((("double",) * 2) if name == "fp64" else
                                                           (("float",) * 2))),
                     "flops_per_launch": flops},
        "paths": res,
    }
    gf.close()
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(kpts, omegas, eta, irvec, hr, args.cpu_per_worker,
                                           args.cpu_workers)
        log(f"[bench] cpu: {out['cpu_baseline']['value']:.1f} solves/s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
