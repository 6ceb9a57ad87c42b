#!/usr/bin/env python3
"""Benchmark: GP iterations/s of the electrostatic placement loop on one MI355X.

A step is one Placement::step (placement.rs:180-228) over the 1M-cell workload
(1,000,000 movable cells, 1,050,000 nets, degree cap 256, 300 macros,
utilisation 0.8, 1024x1024 bins), synthetic netlist from the seeded generator.
Inputs are resident in HBM before timing; W warmup iterations run on a scratch
context state which is then reset, and K iterations are timed from iteration 0
(the reference's own loop starts there, main.rs:78).

Prints one JSON line (rank 0) with the metric, a roofline object for the
dominant kernel (its average launch time measured with HIP events on the
context's stream) and the CPU baseline (the oracle's f32 OpenMP path on a
bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "reda-erplacer_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = 2500.0     # dense bf16 (no sparsity)
METRIC = "GP iters/s @1M cells & placements/min + p50/p99 latency @ HPWL within 0.5%"




def kernel_bytes(name, d, grid, cells_pins):
    """Algorithmic HBM bytes of one launch of `name`: each logical array the
    kernel needs is read or written once (gathers count the gathered element)."""
    import numpy as np
    C = d.num_movable
    deg = np.diff(d.net_start)
    B = grid * grid
    n2 = int(d.n_2nets)
    multi = deg[n2:]
    P2 = n2 * 2
    narrow = multi[multi <= 256]
    wide = multi[multi > 256]
    table = {
        # per pin: slot node (4) + pin offset (8) + instance position gather (8) + grad
        # write (8); per multi-pin net its CSR start (8)
        "wa_narrow": (P2 + int(narrow.sum())) * 28 + int(narrow.size) * 8,
        "wa_wide": int(wide.sum()) * 28 + int(wide.size) * 8,
        # sorted record (16) + ratio (4) + offset (8) + new order (4) read; order (4),
        # constants (16) and ratio (4) written back for the next sort; int64 map
        # written once
        "scatter": C * 56 + B * 8,
        # order (4), position gather (8), constants (16); key (4) + corner (8) written
        "sort_keys": C * 40,
        # key, order, corner, constants, ratio read (36); order, record, ratio,
        # offset written in sorted order (32)
        "sort_place": C * 68,
        # read q_mov, q_fix (16), clear q_mov (8), write density (4) and row transform (4)
        "pass1_row_dct": B * 32,
        # pass 1 and the WA net groups up to degree 256 in one launch
        "pass1_wa": B * 32 + (P2 + int(narrow.sum())) * 28 + int(narrow.size) * 8,
        # MFMA form: read q_mov, q_fix (16), clear q_mov (8), write density (4) and
        # the split row operand (four bf16 planes of half the row: 4)
        "mfma_finalize": B * 32,
        "mfma_finalize_wa": B * 32 + (P2 + int(narrow.sum())) * 28 + int(narrow.size) * 8,
        # read row transform (4), write three column-inverse maps (12)
        "pass2_col_spectral": B * 16,
        # read three maps (12) + density (4), write field x/y and potential (12)
        "pass3_row_final": B * 28,
        # per sorted cell: record (16) and ratio/id record (16) read, field force
        # written to the cell's id (8); field maps read once (8 per bin)
        "field_sorted": C * 40 + B * 8,
        # per cell: CSR start (4), position (8), ratio/pins/area (16), previous position
        # and gradient (16), field force (8), gradient write (8); per pin: slot id (4) +
        # pin gradient (8)
        "cell_fused": C * (4 + 8 + 16 + 16 + 8 + 8) + cells_pins * 12,
        # position, previous position, gradient, clamp bound read; three writes
        "update": C * 56,
    }
    return table.get(name, 0)


def kernel_flops(name, grid, potential):
    """MFMA flops one launch of the split-BF16 spectral form executes: every
    product is two half-size products (even / odd frequencies, K = L/2) of three
    bf16 terms each, 3 m n^2 for a row pass and 3 m^2 n for a column pass, times
    the products the launch carries (inverse columns: U and V, 2; inverse rows:
    3, or 2 without the potential map).  The useful FP32-equivalent work is a
    third of it."""
    m = n = grid
    nz = 3 if potential else 2
    return {"mfma_fwd_row": 3 * m * n * n, "mfma_fwd_col": 3 * m * m * n,
            "mfma_inv_col": 2 * 3 * m * m * n, "mfma_inv_row": nz * 3 * m * n * n}.get(name, 0)


# timer name -> kernel function (without the k_ prefix) in the rocprofv3 summaries
KERNEL_OF_TIMER = {"pass1_row_dct": "pass1", "pass2_col_spectral": "col_spectral",
                   "pass3_row_final": "row_final"}


def measured_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)["bytes_per_dispatch"]
    v = table.get(KERNEL_OF_TIMER.get(kernel, kernel))
    return int(round(v)) if v is not None else None


# the reference scatters sequentially below this iteration and in parallel from
# it on (density.rs:102-237, epotential.rs:245)
SEQUENTIAL_SCATTER_ITERS = 200


def cpu_baseline(d, grid, td, workload_iters, par_iters, threads, repeats=3):
    """The oracle's f32 OpenMP path on the same workload, `repeats` runs from
    iteration 0 (the median is reported): every iteration of the
    sequential-scatter phase is timed, then `par_iters` of the parallel phase;
    a run's rate is the workload's iteration count over the time the two
    phases would take at those per-iteration costs.  Returns (median rate,
    all rates, seconds timed, iterations timed per run)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    n_seq = min(SEQUENTIAL_SCATTER_ITERS, workload_iters)
    n_par = workload_iters - n_seq
    k = min(par_iters, n_par)
    rates, timed = [], 0.0
    for _ in range(repeats):
        P = O.Placer(d, grid, target_density=td, sfx="_f32", nthreads=threads)
        t0 = time.perf_counter()
        for _ in range(n_seq):
            P.step()
        t_seq = time.perf_counter() - t0
        t1 = time.perf_counter()
        for _ in range(k):
            P.step()
        t_par = time.perf_counter() - t1
        projected = t_seq + (n_par * t_par / k if k else 0.0)
        rates.append(workload_iters / projected)
        timed += t_seq + t_par
        del P
    return float(np.median(rates)), rates, timed, n_seq + k


def cpu_serving_baseline(reqs, check_every, max_iterations, threads, budget_s):
    """Sequential plain solves on the CPU: the oracle's f32 OpenMP loop places the
    serving stream's requests one after another (stop at overflow < 0.1 checked
    every `check_every` iterations, as the GPU server does) until `budget_s`
    seconds have passed; returns placements/min over the requests finished."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from erplacer import workloads as W
    stop = W.SERVING["stop_overflow"]
    td = W.SERVING["target_density"]
    done, cells, t0 = 0, 0, time.perf_counter()
    for _, d, g in reqs:
        P = O.Placer(d, g, target_density=td, sfx="_f32", nthreads=threads)
        it = 0
        while it < max_iterations:
            for _ in range(min(check_every, max_iterations - it)):
                r = P.step()
                it += 1
            if r["overflow"] < stop:
                break
        del P
        done += 1
        cells += d.num_movable
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"placements_per_min": round(60 * done / dt, 2), "requests": done,
            "cells": cells, "seconds": round(dt, 2), "cores": threads, "kind": "port",
            "sample": f"the first {done} requests of the stream placed one after another "
                      f"(f32 OpenMP restatement, context setup included)"}


def serving_requests(n_requests, rate):
    from erplacer import workloads as W
    t0 = time.perf_counter()
    reqs = W.serving_requests(n_requests, rate_per_s=rate)
    print(f"[serving] generated {len(reqs)} requests in {time.perf_counter() - t0:.1f}s "
          f"({sum(d.num_movable for _, d, _ in reqs)} cells)", file=sys.stderr, flush=True)
    return reqs


def serving_profile(reqs, check_every, max_active, max_iterations):
    """The saturated serving stream on an eager server (the batched kernels of
    each round launched directly; results identical to the graph replay)."""
    from erplacer import workloads as W
    from erplacer.server import Server
    stop = W.SERVING["stop_overflow"]
    td = W.SERVING["target_density"]
    t0 = time.perf_counter()
    with Server(max_active=max_active, check_every=check_every, eager=True) as srv:
        ids = [srv.submit(d, g, max_iterations, stop, target_density=td) for _, d, g in reqs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    conv = sum(1 for r in res if r["converged"])
    print(f"[serving] eager stream: {len(res)} requests, {conv} converged, "
          f"{time.perf_counter() - t0:.2f}s, stats {st}", file=sys.stderr, flush=True)


def serving_bench(reqs, rates, check_every, max_active, max_iterations):
    """Serving form: the request stream (log-uniform 2k-150k cells, grids 64-512,
    stop at overflow < 0.1) placed (1) one after another alone, (2) all submitted at
    once to the continuous-batching server, (3) with Poisson arrivals at each rate
    (the request sizes are the same; only the arrival times scale).
    Returns the serving object of the JSON line."""
    import numpy as np
    from erplacer import workloads as W
    from erplacer.placement import Placement
    from erplacer.server import Server
    stop = W.SERVING["stop_overflow"]
    td = W.SERVING["target_density"]
    # (1) sequential plain solves (context creation included)
    seq_t, seq_res = [], []
    for _, d, g in reqs:
        t = time.perf_counter()
        with Placement(d, g, target_density=td) as P:
            P.set_store_gradients(False)
            info = P.run_until(max_iterations, stop, check_every)
            dbu = P.final_hpwl_dbu()
        seq_t.append(time.perf_counter() - t)
        seq_res.append((info["iteration"], dbu))
    seq_total = sum(seq_t)
    print(f"[serving] sequential: {seq_total:.2f}s", file=sys.stderr, flush=True)

    def served(arrivals):
        with Server(max_active=max_active, check_every=check_every) as srv:
            t_start = time.perf_counter()
            ids = []
            for (arr, d, g) in zip(arrivals, [r[1] for r in reqs], [r[2] for r in reqs]):
                if arr > 0:
                    dt = arr - (time.perf_counter() - t_start)
                    if dt > 0:
                        time.sleep(dt)
                ids.append(srv.submit(d, g, max_iterations, stop, target_density=td))
            res = [srv.wait(i) for i in ids]
            makespan = time.perf_counter() - t_start
            st = srv.stats()
        return res, makespan, st

    res_sat, mk_sat, st_sat = served([0.0] * len(reqs))
    print(f"[serving] saturated: {mk_sat:.2f}s stats {st_sat}", file=sys.stderr, flush=True)
    same = all(r["status"] == 0 and (r["iterations"], r["hpwl_dbu"]) == s for r, s in zip(res_sat, seq_res))
    base_arrivals = np.array([a for a, _, _ in reqs])
    latency = []
    for rate in rates:
        arrivals = list(base_arrivals * (rates[0] / rate))
        res_poi, mk_poi, st_poi = served(arrivals)
        same = same and all(r["status"] == 0 and (r["iterations"], r["hpwl_dbu"]) == s
                            for r, s in zip(res_poi, seq_res))
        lat = np.array([r["t_done"] - r["t_submit"] for r in res_poi])
        # a sequential server under the same arrivals: FIFO, one alone solve at a time
        done, seq_lat = 0.0, []
        for a, t in zip(arrivals, seq_t):
            done = max(done, a) + t
            seq_lat.append(done - a)
        seq_lat = np.array(seq_lat)
        row = {"arrival_rate_per_s": rate,
               "p50_s": round(float(np.percentile(lat, 50)), 4),
               "p99_s": round(float(np.percentile(lat, 99)), 4),
               "sequential_p50_s": round(float(np.percentile(seq_lat, 50)), 4),
               "sequential_p99_s": round(float(np.percentile(seq_lat, 99)), 4),
               "makespan_s": round(mk_poi, 3), "max_batch": st_poi["max_batch"]}
        print(f"[serving] poisson {rate}/s: {row}", file=sys.stderr, flush=True)
        latency.append(row)
    conv = sum(1 for r in res_sat if r["converged"])
    return {
        "requests": len(reqs), "check_every": check_every,
        "max_active": max_active, "max_iterations": max_iterations, "stop_overflow": stop,
        "converged": conv,
        "placements_per_min": round(60 * len(reqs) / mk_sat, 1),
        "sequential_placements_per_min": round(60 * len(reqs) / seq_total, 1),
        "speedup_vs_sequential": round(seq_total / mk_sat, 3),
        "p50_s": latency[-1]["p50_s"], "p99_s": latency[-1]["p99_s"],
        "arrival_rate_per_s": latency[-1]["arrival_rate_per_s"],
        "latency": latency, "max_batch": st_sat["max_batch"],
        "results_identical_to_alone": bool(same),
    }


def traced_run(P, iterations, every=50):
    """The workload from iteration 0 in chunks of `every` (bit-identical to one
    run): (iteration, hpwl, overflow) after each chunk."""
    P.reset()
    out, done = [], 0
    while done < iterations:
        k = min(every, iterations - done)
        r = P.run(k)
        done += k
        out.append((done, r["hpwl"], r["overflow"]))
    return out


def end_state(workload, info, hpwl_dbu, trace=None):
    """The end state of the whole workload against the FP64 oracle's runs of the
    same workload (tests/golden/<workload>_band.json, written by the oracle on
    the CPU): the unperturbed FP64 run and the spread of its one-ulp
    perturbations.  The metric's qualifier ('HPWL within 0.5 %') is reported
    as measured, beside the band the oracle itself produces."""
    path = os.path.join(ROOT, "tests", "golden", f"{workload}_band.json")
    out = {"hpwl": info["hpwl"], "overflow": info["overflow"], "hpwl_dbu": hpwl_dbu,
           "iteration": info["iteration"]}
    if not os.path.exists(path):
        out["fp64_reference"] = None
        return out
    with open(path) as f:
        band = json.load(f)
    ref = band["runs"][0]                       # unperturbed, float map (the faithful oracle)
    s = band["fp64"]
    rel = info["hpwl"] / ref["hpwl"] - 1
    z = (info["hpwl"] - s["hpwl_mean"]) / s["hpwl_std"]
    zo = (info["overflow"] - s["overflow_mean"]) / s["overflow_std"]
    out.update({
        "fp64_run_hpwl": ref["hpwl"], "fp64_run_overflow": ref["overflow"],
        "hpwl_rel_to_fp64_run": round(rel, 6),
        "overflow_diff_to_fp64_run": round(info["overflow"] - ref["overflow"], 6),
        "within_0.5pct_and_0.005": bool(abs(rel) <= 0.005 and abs(info["overflow"] - ref["overflow"]) <= 0.005),
        "fp64_band": {"runs": s["n"], "hpwl_mean": s["hpwl_mean"],
                      "hpwl_std_rel": round(s["hpwl_std"] / s["hpwl_mean"], 6),
                      "hpwl_range_rel": round(s["hpwl_max"] / s["hpwl_min"] - 1, 6),
                      "overflow_mean": s["overflow_mean"], "overflow_std": s["overflow_std"]},
        "hpwl_z_in_band": round(z, 3), "overflow_z_in_band": round(zo, 3),
        "within_band_3sigma": bool(abs(z) <= 3 and abs(zo) <= 3),
    })
    f64 = [r for r in band["runs"] if r["sfx"] == "_f64"]
    if trace and all("trace" in r for r in f64):
        # the GPU's HPWL every 50 iterations against the FP64 runs' envelope
        env = {}
        for r in f64:
            for it, h, _ in r["trace"]:
                lo, hi = env.get(it, (h, h))
                env[it] = (min(lo, h), max(hi, h))
        # the oracle at the GPU's precision split (f32 arrays, FP64 scalars) and
        # in plain f32 at the same iterations: while the one-ulp FP64 ensemble is
        # still tight, the precisions have already parted
        other = {}
        for r in band["runs"]:
            if r["sfx"] in ("_f32d", "_f32") and "trace" in r:
                for it, h, _ in r["trace"]:
                    other.setdefault((r["sfx"], it), []).append(h)
        rows = []
        for it, h, o in trace:
            if it in env:
                lo, hi = env[it]
                mid = 0.5 * (lo + hi)
                row = {"iteration": it, "hpwl": h, "overflow": round(o, 5),
                       "fp64_min": lo, "fp64_max": hi,
                       "rel_to_fp64_mid": round(h / mid - 1, 5),
                       "inside": bool(lo <= h <= hi)}
                if ("_f32d", it) in other:
                    row["oracle_f32_fp64_scalars"] = other[("_f32d", it)][0]
                if ("_f32", it) in other:
                    row["oracle_f32_min"] = min(other[("_f32", it)])
                    row["oracle_f32_max"] = max(other[("_f32", it)])
                rows.append(row)
        out["trace_vs_fp64"] = rows
    return out


def copy_calibration(nbytes, reps):
    """Device-to-device hipMemcpy (the runtime's blit kernel) of a known size,
    the reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"[bench] copy calibration: {reps} x {nbytes} bytes read + written", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=800)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="large")
    ap.add_argument("--spectral", default="fft", choices=["fft", "mfma"])
    ap.add_argument("--skip-potential", action="store_true",
                    help="trajectory only: skip the potential map and energy, which "
                         "Placement::step discards (placement.rs:185); not the default")
    ap.add_argument("--timing-steps", type=int, default=0,
                    help="iterations re-run with per-kernel HIP events for the roofline "
                         "(default: the workload's full iteration count, whatever --steps is)")
    ap.add_argument("--cpu-iters", type=int, default=20,
                    help="timed CPU iterations of the parallel-scatter phase (the "
                         "sequential phase before it is timed whole)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-repeats", type=int, default=3,
                    help="CPU baseline runs (the median is reported)")
    ap.add_argument("--cpu-serving-seconds", type=float, default=15.0,
                    help="time budget of the CPU sequential-solve leg of the serving form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only the per-kernel timing loop, the launches whose average the "
                         "roofline quotes (for rocprofv3 runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device-to-device hipMemcpy after the loop, "
                         "the reference dispatch for rocprofv3 FETCH_SIZE / WRITE_SIZE scaling")
    ap.add_argument("--serving-profile", action="store_true",
                    help="only the serving stream, all requests at once, on a server that "
                         "launches the batched kernels directly instead of replaying graphs "
                         "(for rocprofv3 kernel traces of the batched kernels)")
    ap.add_argument("--serving-requests", type=int, default=512,
                    help="requests of the serving measurement (0 skips it)")
    ap.add_argument("--serving-rate", default="20,60",
                    help="Poisson arrival rates (requests/s), comma separated; the last is the "
                         "headline p50/p99 (60/s exceeds one-at-a-time capacity)")
    ap.add_argument("--serving-check-every", type=int, default=10)
    ap.add_argument("--serving-max-active", type=int, default=64)
    ap.add_argument("--serving-max-iterations", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(placement jobs shard across GPUs with no exchange: run one server per GPU)", file=sys.stderr)
        sys.exit(2)

    from erplacer import lib as L
    from erplacer import workloads as W
    from erplacer.placement import Placement

    if args.serving_profile:
        serving_profile(serving_requests(args.serving_requests, float(args.serving_rate.split(",")[0])),
                        args.serving_check_every, args.serving_max_active, args.serving_max_iterations)
        return

    t0 = time.perf_counter()
    d, grid, iters_cfg, td = W.make(args.workload)
    gen_s = time.perf_counter() - t0
    print(f"[bench] workload {args.workload}: {d.summary()} grid {grid} (generated in {gen_s:.1f}s)",
          file=sys.stderr, flush=True)
    mode = L.SPECTRAL_FFT if args.spectral == "fft" else L.SPECTRAL_MFMA_BF16X3
    P = Placement(d, grid, target_density=td, spectral_mode=mode)
    P.set_store_gradients(False)
    if args.skip_potential:
        P.set_compute_potential(False)
    if args.warmup:
        P.run(args.warmup)
    P.reset()
    P.info()  # synchronise
    window = args.timing_steps or iters_cfg
    if args.profile:
        P.reset()
        P.set_timing(True)
        P.run(window)
        P.set_timing(False)
        t = P.timing()
        print(f"[bench] profile loop per-kernel avg ms: { {k: round(v, 5) for k, v in t.items()} }",
              file=sys.stderr, flush=True)
        if args.copy_calibration:
            copy_calibration(256 << 20, 3)
        return
    t0 = time.perf_counter()
    info = P.run(args.steps)   # returns after a stream synchronize
    elapsed = time.perf_counter() - t0
    ips = args.steps / elapsed
    print(f"[bench] timed {args.steps} iterations in {elapsed:.3f}s -> {ips:.1f} it/s; "
          f"hpwl {info['hpwl']:.6e} overflow {info['overflow']:.5f}", file=sys.stderr, flush=True)
    hpwl_dbu = P.final_hpwl_dbu()

    # the whole workload (its own iteration count from iteration 0), for the
    # end state and the rate over every iteration of it
    if args.steps == iters_cfg:
        full_ips, full_info, full_dbu = ips, info, hpwl_dbu
    else:
        P.reset()
        P.info()
        t0 = time.perf_counter()
        full_info = P.run(iters_cfg)
        full_ips = iters_cfg / (time.perf_counter() - t0)
        full_dbu = P.final_hpwl_dbu()

    trace = traced_run(P, iters_cfg)
    if trace[-1][1] != full_info["hpwl"]:
        raise RuntimeError("the chunked run of the workload differs from the single run")

    # per-kernel device time with HIP events over a second timed region: the
    # workload's full iteration count, so the averages cover the clustered start
    # and the spread end alike (what a kernel trace of the whole run averages)
    P.reset()
    P.set_timing(True)
    P.run(window)
    P.set_timing(False)
    times = P.timing()
    if not times:
        raise RuntimeError("no kernel timings were recorded")
    pins_cells = int(d.node_pin_start[d.num_movable])
    per_iter = {}
    for k, ms in times.items():
        per_iter[k] = ms
    dom = max((k for k in per_iter if kernel_bytes(k, d, grid, pins_cells) > 0),
              key=lambda k: per_iter[k])
    dom_ms = per_iter[dom]
    nbytes = kernel_bytes(dom, d, grid, pins_cells)
    achieved = nbytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.workload, dom) if args.spectral == "fft" else None,
            "bytes_per_launch": nbytes, "avg_launch_ms": round(dom_ms, 5),
            "window_iterations": window}
    kernels = {}
    for k, v in per_iter.items():
        e = {"avg_ms": round(v, 5)}
        nb, nf = kernel_bytes(k, d, grid, pins_cells), kernel_flops(k, grid, not args.skip_potential)
        if nb > 0 and v > 0:
            gbs = nb / (v * 1e-3) / 1e9
            e.update(gbs=round(gbs, 1), hbm_frac=round(gbs / HBM_PEAK_GBS, 4))
        if nf > 0 and v > 0:
            tf = nf / (v * 1e-3) / 1e12
            e.update(mfma_tflops=round(tf, 1), mfma_frac=round(tf / MFMA_PEAK_TFLOPS, 4))
        kernels[k] = e
    print(f"[bench] per-kernel avg ms: {kernels}", file=sys.stderr, flush=True)

    rates = [float(r) for r in str(args.serving_rate).split(",") if r]
    reqs = serving_requests(args.serving_requests, rates[0]) if args.serving_requests > 0 else None
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cps, crates, cdt, cn = cpu_baseline(d, grid, td, iters_cfg, args.cpu_iters, threads,
                                            args.cpu_repeats)
        cpu = {"value": round(cps, 4), "unit": "GP iterations/s", "cores": threads, "kind": "port",
               "runs": [round(r, 4) for r in crates],
               "sample": f"median of {len(crates)} runs, each the first {cn} iterations of the "
                         f"same workload ({cdt:.1f}s in all; the sequential-scatter phase whole, "
                         f"then {cn - min(SEQUENTIAL_SCATTER_ITERS, iters_cfg)} of the parallel "
                         f"one, projected onto its {iters_cfg}), f32, OpenMP restatement of the "
                         f"Rayon path"}
        print(f"[bench] cpu baseline {cps:.4f} it/s on {threads} threads (runs {crates})",
              file=sys.stderr, flush=True)
        if args.serving_requests > 0 and args.cpu_serving_seconds > 0:
            cpu["serving"] = cpu_serving_baseline(reqs, args.serving_check_every,
                                                  args.serving_max_iterations, threads,
                                                  args.cpu_serving_seconds)
            print(f"[bench] cpu serving baseline {cpu['serving']}", file=sys.stderr, flush=True)

    serving = None
    if reqs:
        serving = serving_bench(reqs, rates, args.serving_check_every,
                                args.serving_max_active, args.serving_max_iterations)
        print(f"[serving] {serving}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(ips, 2), "unit": "GP iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator, reda-db contract)",
        "config": {"workload": f"{args.workload}: {d.num_movable} cells, {d.num_nets} nets, "
                               f"{d.num_pin_total} pins, {grid}x{grid} bins",
                   "iterations_of_workload": iters_cfg, "spectral": args.spectral,
                   "potential": "skipped (trajectory only)" if args.skip_potential else "computed",
                   "parallelism": "single GPU"},
        "final": {"hpwl": info["hpwl"], "overflow": info["overflow"], "hpwl_dbu": hpwl_dbu,
                  "iteration": info["iteration"]},
        "full_workload": {"iterations": iters_cfg, "it_per_s": round(full_ips, 2),
                          "end_state": end_state(args.workload, full_info, full_dbu, trace)},
        "roofline": roof, "kernels": kernels, "cpu_baseline": cpu, "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
