"""Benchmark of the GeodesicFDTD step loop on one MI355X.

Workload (default): s7 icosahedral mesh (163842 dual cells) x 64 radial cells
to 100 km, FP32 fields, exponential conductive update with
sigma(z) = 1e-6 exp((z-70 km)/5 km) S/m, surface-impedance ground
(sigma_g = 1e-2 S/m), electron Drude plasma above 60 km, 20 Hz Gaussian
vertical current, 32 receivers sampled every step.  A "step" is one full
leapfrog step (H kernel + E kernel) of this grid; inputs are resident in HBM.

Prints one JSON line: Mcell-updates/s (steps/s x V x R), the roofline of the
dominant kernel (algorithmic bytes / measured kernel time vs 8 TB/s) and the
CPU oracle (the NumPy restatement of the reference path) timed on a bounded
sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))

HBM_PEAK_GBS = 8000.0
MAPPING_NAMES = {1: "rows", 2: "flat"}
C2_RUN_STEPS = 4000  # length of one ensemble run (16 receiver traces per run)
METRIC = "Mcell-updates/s + % HBM roofline (1 GPU); ensemble trace queries/s at ≤2% err"


def algorithmic_bytes(sim) -> dict:
    """Compulsory bytes per launch of each kernel (the per-unit figures of the
    step's byte model, split between the two half-step kernels).

    Fields: every value is read once by the half-step that uses it as a source
    and read + written once by the half-step that updates it (3 accesses per
    DOF and step).  Plasma current: read + written once per step, 3 components
    per species on the layers with plasma drive.  Surface impedance: memory
    read + written.  Tables: index / metric rows read once per entity.
    """
    V, E, F = sim.mesh.n_vertices, sim.mesh.n_edges, sim.mesh.n_faces
    R = len(sim.radial_midpoints_m)
    sf = sim._er_buf.element_size()
    sc = 8 if sim.backend.dtype_name == "float64" else 4
    band = 0
    species = 0
    if sim._plasma_buf is not None:
        lo, hi = sim._plasma_info["band"]
        band = hi - lo
        species = sim._plasma_buf.shape[0]
    si_terms = sim._si_memory.shape[1] if sim._si_memory is not None else 0
    h_fields = sf * (V * (R + 1) + E * R + 2 * E * (R + 1) + 2 * F * R)
    h_plasma = sc * 6 * F * band * species
    h_si = sc * 2 * si_terms * E
    h_geom = 12 * E + 24 * F + (48 * F if species else 0)
    e_fields = sf * (E * (R + 1) + F * R + 2 * V * (R + 1) + 2 * E * R)
    e_geom = 48 * V + 12 * E + ((48 * V + 8 * E) if species else 0)
    return {"h": h_fields + h_plasma + h_si + h_geom, "e": e_fields + e_geom,
            "h_fields": h_fields, "h_plasma": h_plasma, "h_si": h_si, "h_geom": h_geom,
            "e_fields": e_fields, "e_geom": e_geom}


def graph_chunk(steps: int, limit: int) -> int:
    """Largest divisor of ``steps`` that is at most ``limit``: the timed steps
    are then whole graph replays only."""
    for c in range(max(1, min(limit, steps)), 0, -1):
        if steps % c == 0:
            return c
    return 1


def read_traffic(kernel: str, workload: str) -> float | None:
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), from the committed summary of that
    workload (None when there is none)."""
    path = REPO / "profiles" / f"{kernel}_{workload}_traffic.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("bytes_per_launch="):
            return float(line.split("=", 1)[1])
    return None


# The reference's published CPU rates for its own benchmark workload (vacuum,
# PEC, seeded 1e-6 fields, s7 x 80, FP32; 32 warm-up steps, median of 3
# repeats of 32 steps): artifacts/benchmarks/backend-scaling-eager-rtx3060.json,
# docs/benchmarks/backend-comparison.md:268.
PUBLISHED_S7_R80_F32_STEPS_PER_S = {"numpy": 0.8190335876595383, "torch-cpu": 0.9233466471837272}


def _torch_cpu_module():
    import importlib.util

    spec = importlib.util.spec_from_file_location("torch_cpu", REPO / "oracle" / "torch_cpu.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _time_cpu_leg(sim, warmup: int, steps: int, repeats: int) -> list[float]:
    """The reference benchmark's method (benchmarks/backend_matrix.py:144-156):
    warm-up steps, then `repeats` timings of `steps` steps."""
    sim.step(warmup)
    out = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        sim.step(steps)
        out.append(time.perf_counter() - t0)
    return out


def cpu_legs(workload: str, warmup: int, steps: int, repeats: int) -> dict:
    """The oracle (NumPy restatement of the reference array path) and its
    torch-CPU restatement (the reference's backend="torch", device="cpu"
    path) on the workload's grid, FP32, timed on this host.  NumPy ufuncs and
    fancy indexing run on one core; torch uses every core."""
    import numpy as np
    import scenarios as S

    import ionosphere_fdtd_amd as g

    sc = S.scenario(workload, dtype="float32")
    mesh = g.build_geodesic_mesh(sc["subdivision"])  # bit-identical to the oracle's builder
    cells = mesh.n_vertices * sc["radial_cells"]
    legs = {}
    threads = os.cpu_count() or 1
    for leg in ("numpy", "torch-cpu"):
        ref = S.build_oracle(sc, dtype="float32", mesh=mesh)
        if sc["seed_fields"]:
            S.apply_seed_fields(ref, lambda k, v: getattr(ref, k).__setitem__(slice(None), v))
        sim = ref if leg == "numpy" else _torch_cpu_module().TorchCpuSim(ref, threads=threads)
        times = _time_cpu_leg(sim, warmup, steps, repeats)
        median = float(np.median(times))
        legs[leg] = {"steps_per_s": steps / median, "mcell_updates_per_s": steps * cells / median / 1e6,
                     "cores": 1 if leg == "numpy" else threads,
                     "seconds": [round(t, 3) for t in times]}
        print(f"cpu {workload} {leg}: {legs[leg]}", file=sys.stderr, flush=True)
        del sim, ref
    return {"workload": workload, "warmup_steps": warmup, "steps": steps, "repeats": repeats,
            "host_logical_cpus": threads, "legs": legs}


def cpu_baseline() -> dict:
    """Bounded CPU sample of the benchmark workload (s7 x 64, plasma + SI,
    FP32): one warm-up step, then the median of 3 timings of 2 steps, for the
    NumPy and torch-CPU legs.  `value` is the faster leg."""
    res = cpu_legs("c1", warmup=1, steps=2, repeats=3)
    best = max(res["legs"], key=lambda k: res["legs"][k]["mcell_updates_per_s"])
    leg = res["legs"][best]
    return {"value": round(leg["mcell_updates_per_s"], 4), "unit": "Mcell-updates/s",
            "cores": leg["cores"], "kind": "port",
            "sample": (f"s7 x 64 plasma+SI FP32 workload, 1 warm-up step then median of 3 x 2 "
                       f"steps per leg (the reference benchmark's method with fewer steps); "
                       f"faster leg: {best}; NumPy leg single-threaded ufuncs, torch-CPU leg "
                       f"{res['host_logical_cpus']} threads"),
            "legs": {k: {"mcell_updates_per_s": round(v["mcell_updates_per_s"], 4),
                         "steps_per_s": round(v["steps_per_s"], 4), "cores": v["cores"]}
                     for k, v in res["legs"].items()}}


def precision_sweep(args) -> dict:
    """C3 precision sweep: the same grid and physics with FP64, FP32, FP16
    (dynamic power-of-two scales) and BF16 field storage, FP32 update
    arithmetic for the last three.  Errors against the FP64 run on the GPU:
    receiver-trace relative L2, per-field relative L2 against the field
    family's norm (Er, Et) / (Hr, Ht), and the discrete div-B residual
    max |D| / flux scale.  Two initial states: zero fields driven by the
    20 Hz source, and the reference benchmark's seeded random fields
    (rng(20260814).standard_normal * 1e-6, source off) so that every cell is
    active from the first step."""
    import numpy as np
    import torch

    import scenarios as S

    out = {}
    for init in ("source", "seeded"):
        sc = S.scenario(args.workload)
        if init == "seeded":
            sc.update(source=None, seed_fields=True)
        ref_fields = ref_traces = None
        res = {}
        for variant in ["float64"] + [d for d in args.sweep_dtypes.split(",") if d != "float64"]:
            # "float16-nearest" / "bfloat16-nearest": narrow storage rounded to
            # nearest instead of stochastically
            dtype, _, rounding = variant.partition("-")
            t0 = time.perf_counter()
            sim = S.build_sim(sc, dtype=dtype, compile_step=True, compile_chunk_size=args.chunk,
                              storage_rounding=rounding or "rotation")
            if sc["seed_fields"]:
                S.apply_seed_fields(sim, lambda k, v: setattr(sim, k, v))
            pts = S.receiver_points(sc)
            v, l, w = S.receiver_tables(sim.mesh, sim.altitudes_m, pts)
            setup = time.perf_counter() - t0
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            traces = sim.record_er_observations(v, l, w, args.steps)
            torch.cuda.synchronize()
            run = time.perf_counter() - t1
            diag = sim.field_diagnostics()
            fields = {k: getattr(sim, k).to(torch.float64) for k in ("er", "et", "hr", "ht")}
            entry = {"ms_per_step": round(run * 1e3 / args.steps, 4),
                     "divb_residual": diag["divb_max_abs"] / max(diag["divb_flux_scale"], 1e-300),
                     "finite": bool(diag["nonfinite"] == 0 and diag["storage_overflow"] == 0),
                     "trace_norm": float(np.linalg.norm(traces))}
            if dtype == "float64":
                ref_fields, ref_traces = fields, traces
            else:
                entry["trace_rel_l2"] = S.rel_l2(traces, ref_traces)
                for fam in (("er", "et"), ("hr", "ht")):
                    norm = torch.sqrt(sum(torch.sum(ref_fields[k] ** 2) for k in fam))
                    for k in fam:
                        entry[f"{k}_rel_l2"] = float(torch.linalg.norm(fields[k] - ref_fields[k])
                                                     / norm)
                if dtype in ("float16", "bfloat16"):
                    entry["final_scales"] = sim.storage_scales
            res[variant] = entry
            print(f"precision {init} {variant}: setup {setup:.1f} s, run {run:.1f} s, {entry}",
                  file=sys.stderr, flush=True)
            del sim, fields
            torch.cuda.empty_cache()
        out[init] = res
    return out


def ensemble_bench(args) -> dict:
    """Batched ensemble: 64 seeded runs on the s5 mesh (R in {32, 40, 48},
    own h', beta, ground sigma, source, carrier, 16 receivers each), FP32,
    advanced together in one H and one E launch per step (HIP graphs of
    --chunk steps) with the autotuned column-block shape.  The same runs
    stepped one after another (each with its own graphs) are timed beside it."""
    import numpy as np
    import torch

    import scenarios as S

    import ionosphere_fdtd_amd as g

    t0 = time.perf_counter()
    scs = S.c2_runs(n_runs=args.runs)
    mesh = g.build_geodesic_mesh(scs[0]["subdivision"])
    runs = [S.build_sim(sc, mesh=mesh, compile_step=True, compile_chunk_size=args.chunk)
            for sc in scs]
    rx = [S.receiver_tables(mesh, r.altitudes_m, S.receiver_points(sc)) for r, sc in zip(runs, scs)]
    print(f"ensemble setup {time.perf_counter() - t0:.1f} s, {len(runs)} runs", file=sys.stderr,
          flush=True)
    ens = g.EnsembleFDTD(runs, compile_chunk_size=args.chunk, autotune=False)
    shape = ens.tune(10)
    print(f"autotuned shape {shape}", file=sys.stderr, flush=True)
    ens.record_er_observations(rx, args.warmup)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    traces = ens.record_er_observations(rx, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    assert np.all(np.isfinite(traces))
    cells = ens.cells
    step_bytes = sum(algorithmic_bytes(r)["h"] + algorithmic_bytes(r)["e"] for r in runs)
    # the same runs, one after another
    solo_steps = max(args.steps // 10, args.chunk)
    for r, x in zip(runs, rx):
        r.record_er_observations(*x, args.chunk)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for r, x in zip(runs, rx):
        r.record_er_observations(*x, solo_steps)
    torch.cuda.synchronize()
    solo = time.perf_counter() - t2
    n_traces = len(runs) * traces.shape[2]
    return {
        "metric": "ensemble Mcell-updates/s (64 uneven runs, one launch per half-step)",
        "value": round(args.steps * cells / elapsed / 1e6, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(runs)} runs, s5 mesh, R in {{32,40,48}} (seeded), Drude "
                               "plasma h'~U(68,80) km beta~U(0.25,0.55)/km, SI ground "
                               "sigma~logU(1e-4,1e-1), carrier~U(10,40) Hz, 16 receivers each",
                   "cells": cells,
                   "radial_cells": [len(r.radial_midpoints_m) for r in runs],
                   "graph_chunk": args.chunk, "column_block": shape},
        # one query = one receiver trace of a full 4000-step run
        "trace_queries_per_s": round(n_traces / (elapsed * C2_RUN_STEPS / args.steps), 2),
        "traces_per_ensemble_run": n_traces,
        "steps_per_run": C2_RUN_STEPS,
        "roofline": {"bound": "hbm", "kernel": "h_ensemble_kernel+e_ensemble_kernel",
                     "achieved": round(step_bytes / (elapsed / args.steps) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes_per_step": step_bytes},
        "runs_one_after_another": {
            "ms_per_step": round(solo * 1e3 / solo_steps, 5),
            "value": round(solo_steps * cells / solo / 1e6, 1),
            "ensemble_speedup": round((solo / solo_steps) / (elapsed / args.steps), 2)},
    }


# (label, M, N, K, A layout, B layout): "k" = K index contiguous, "r" = row
# (M for A, N for B) contiguous
GEMM_SHAPES = (
    ("trunk layer, batch 2048", 2048, 256, 256, "k", "r"),
    ("decoder convolution as im2col", 131072, 32, 160, "k", "r"),
    ("decoder data gradient", 131072, 160, 32, "k", "k"),
    ("weight gradient, split K", 256, 256, 131072, "r", "r"),
    ("square 4096", 4096, 4096, 4096, "k", "r"),
    ("square 4096, B stored transposed", 4096, 4096, 4096, "k", "k"),
    ("square 8192", 8192, 8192, 8192, "k", "r"),
)


def gemm_bench(args) -> dict:
    """geo_gemm_bf16 (BF16 in, FP32 out) beside torch's matmul (hipBLASLt, BF16
    out) on the surrogate's own products and on large squares: HIP events on
    the current stream, median of 20 launches after 3 warm-up launches."""
    import numpy as np
    import torch
    from ionosphere_fdtd_amd import _native

    lib = _native.load_library()
    stream = torch.cuda.current_stream()
    out = []
    for label, M, N, K, al, bl in GEMM_SHAPES:
        g = torch.Generator(device="cuda").manual_seed(0)
        a = torch.randn((M, K) if al == "k" else (K, M), device="cuda", generator=g).bfloat16()
        b = torch.randn((N, K) if bl == "k" else (K, N), device="cuda", generator=g).bfloat16()
        a_rs, a_cs = (K, 1) if al == "k" else (1, M)
        b_rs, b_cs = (1, K) if bl == "k" else (N, 1)
        c = torch.empty((M, N), device="cuda", dtype=torch.float32)
        at = a if al == "k" else a.t()
        bt = b.t() if bl == "k" else b

        def ours():
            _native.check(lib.geo_gemm_bf16(M, N, K, a.data_ptr(), a_rs, a_cs, b.data_ptr(),
                                            b_rs, b_cs, None, c.data_ptr(), None, N, 0,
                                            stream.cuda_stream), lib)

        def theirs():
            torch.matmul(at, bt)

        row = {"shape": label, "M": M, "N": N, "K": K}
        for name, fn in (("geo_gemm_bf16", ours), ("hipblaslt", theirs)):
            for _ in range(3):
                fn()
            times = []
            for _ in range(20):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                fn()
                e1.record(stream)
                e1.synchronize()
                times.append(e0.elapsed_time(e1))
            ms = float(np.median(times))
            row[name + "_tflops"] = round(2.0 * M * N * K / (ms * 1e-3) / 1e12, 1)
            row[name + "_ms"] = round(ms, 4)
        ref = (at.float() @ bt.float())
        row["max_rel_err_vs_fp32"] = float((c - ref).abs().max() / ref.abs().max())
        out.append(row)
        print(json.dumps(row), file=sys.stderr, flush=True)
    return {"metric": "BF16 GEMM TFLOP/s (FP32 accumulation)", "peak_dense_bf16_tflops": 2500.0,
            "results": out}


def surrogate_bench(args) -> dict:
    """Learned surrogate with FDTD fallback on held-out random-geometry queries.

    --queries held-out queries (one receiver per configuration by default:
    every query its own physics, its source anywhere on the sphere, its
    receiver within 250 km; seeds disjoint from training) are answered (a) by
    batched plain FDTD -- the ground truth and the time to beat -- and (b) by
    the query service: the trained committee for in-range queries it agrees
    on, FDTD for the rest, one spot-checked configuration per batch (any error
    > 2 % sends the batch to FDTD).  A further 16 out-of-range queries
    exercise the fallback."""
    import numpy as np
    import torch

    import ionosphere_fdtd_amd as g
    from ionosphere_fdtd_amd import trace_queries as tq

    paths = sorted((REPO / "ionosphere-fdtd_amd" / "data").glob("trace_surrogate_*.npz"))
    committee = tq.SurrogateCommittee.load(paths)
    mesh = g.build_geodesic_mesh(tq.SUBDIVISION)
    per = max(1, args.receivers_per_config)
    queries = tq.random_queries(args.queries // per, per, seed=20261206)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    truth = tq.fdtd_traces(queries, mesh)
    torch.cuda.synchronize()
    t_fdtd = time.perf_counter() - t0
    print(f"batched FDTD: {len(queries)} queries in {t_fdtd:.1f} s", file=sys.stderr, flush=True)
    thresholds = [float(x) for x in str(args.disagreement).split(",")]
    runs = {}
    for tau in thresholds:
        service = tq.TraceQueryService(committee, mesh, tolerance=0.02, disagreement=tau,
                                       spot_checks=1, seed=7)
        service.answer(queries[:8])  # warm-up (graphs, kernels)
        tq._RECEIVER_TABLES.clear()  # every timed run finds its faces afresh
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        answers, served, report = service.answer(queries)
        torch.cuda.synchronize()
        runs[tau] = (time.perf_counter() - t1, answers, served, report)
        e = tq.trace_rel_l2(answers, truth)
        print(f"threshold {tau}: {runs[tau][0]:.3f} s, {report.get('uncertain')} to FDTD, "
              f"within 2 % {np.mean(e <= 0.02):.4f}, max {e.max():.4f}", file=sys.stderr,
              flush=True)
    tau = thresholds[0]
    t_service, answers, served, report = runs[tau]
    err = tq.trace_rel_l2(answers, truth)
    pred, spread = committee.predict(queries, mesh)
    raw = tq.trace_rel_l2(pred, truth)
    by_net = served == "surrogate"
    # fallback: out-of-range physics and a receiver beyond 250 km
    from ionosphere_fdtd_amd.sources import destination_point

    odd = queries[:16].copy()
    odd[:8, 1] = 0.7
    for k in range(8, 16):
        odd[k, 5], odd[k, 6] = destination_point(odd[k, 3], odd[k, 4], 45.0 * k, 400e3,
                                                 tq.EARTH_RADIUS_M)
    o_ans, o_served, o_report = service.answer(odd)
    o_truth = tq.fdtd_traces(odd, mesh)
    return {
        "metric": "ensemble trace queries/s at <=2% err (surrogate + FDTD fallback)",
        "value": round(len(queries) / t_service, 1), "unit": "queries/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 network / f32 FDTD",
        "data": "synthetic", "vs_baseline": None,
        "config": {"workload": f"{len(queries)} held-out random-geometry queries "
                               f"({len(queries) // per} configurations x {per} receivers; s5, "
                               "R=32, 4096 steps, 512-sample Er traces)",
                   "models": [p.name for p in paths],
                   "disagreement_threshold": tau},
        "other_thresholds": {
            f"{t:g}": {"queries_per_s": round(len(queries) / r[0], 1),
                       "speedup_vs_batched_fdtd": round(t_fdtd / r[0], 1),
                       "to_fdtd": r[3].get("uncertain"),
                       "within_2pct": float(np.mean(tq.trace_rel_l2(r[1], truth) <= 0.02)),
                       "max_error": float(tq.trace_rel_l2(r[1], truth).max())}
            for t, r in runs.items() if t != tau},
        "batched_fdtd_queries_per_s": round(len(queries) / t_fdtd, 2),
        "speedup_vs_batched_fdtd": round(t_fdtd / t_service, 1),
        "served_by_surrogate": int(np.sum(by_net)),
        "served_by_fdtd": int(np.sum(~by_net)),
        "uncertain_to_fdtd": report.get("uncertain"),
        "spot_check_errors": report.get("spot_check_errors"),
        "service_seconds": {"total": round(t_service, 4),
                            "surrogate": round(report.get("surrogate_seconds", 0.0), 4),
                            "fdtd_spot_check_and_fallback": round(report.get("fdtd_seconds", 0.0),
                                                                   4)},
        "error_vs_fdtd": {"median": float(np.median(err)), "p90": float(np.quantile(err, 0.9)),
                          "p99": float(np.quantile(err, 0.99)), "max": float(err.max()),
                          "within_2pct": float(np.mean(err <= 0.02))},
        "committee_alone": {"median": float(np.median(raw)), "p90": float(np.quantile(raw, 0.9)),
                            "p99": float(np.quantile(raw, 0.99)), "max": float(raw.max()),
                            "within_2pct": float(np.mean(raw <= 0.02)),
                            "flagged": int(np.sum(spread > tau)),
                            "above_2pct_not_flagged": int(np.sum((raw > 0.02)
                                                                 & (spread <= tau)))},
        "fallback_check": {"queries": len(odd), "served_by_fdtd": int(np.sum(o_served == "fdtd")),
                           "max_error_vs_fdtd": float(tq.trace_rel_l2(o_ans, o_truth).max())},
        # the router's trade-off: queries sent to FDTD and the errors left among
        # the rest, for a range of disagreement thresholds
        "threshold_sweep": {f"{t:g}": {"to_fdtd": int(np.sum(spread > t)),
                                       "served_max_error": float(raw[spread <= t].max())
                                       if np.any(spread <= t) else None,
                                       "served_above_2pct": int(np.sum((raw > 0.02)
                                                                       & (spread <= t)))}
                            for t in (0.002, 0.004, 0.006, 0.008, 0.01, 0.015, 0.02, 0.03)},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--kernel-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-full", action="store_true",
                    help="only time the CPU legs with the reference benchmark's own method (32 "
                         "warm-up steps, median of 3 x 32 steps) on its published s7 x 80 "
                         "vacuum workload, beside the published rates")
    ap.add_argument("--workload", default="c1",
                    help="c1 (default), c1_noplasma, c1_pec, c2 (64-run ensemble), c3, "
                         "c4 (surrogate + FDTD fallback), vacuum_s7_r64, vacuum_s7_r80, "
                         "gemm (the surrogate's BF16 GEMM beside hipBLASLt)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="also time a 1 GiB device copy (read + write GB/s)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 4 torch copies of 512 MiB (counter calibration)")
    ap.add_argument("--precision-sweep", action="store_true",
                    help="run the FP64/FP32/FP16/BF16 storage sweep on --workload (c3: s8 x 96, "
                         "refined 60-90 km band) for --steps steps instead of the throughput "
                         "bench")
    ap.add_argument("--runs", type=int, default=64, help="ensemble runs (--workload c2)")
    ap.add_argument("--queries", type=int, default=2000, help="held-out queries (--workload c4)")
    ap.add_argument("--receivers-per-config", type=int, default=1,
                    help="queries sharing one configuration (--workload c4)")
    ap.add_argument("--disagreement", default="0.01,0.015,0.02",
                    help="committee disagreement above which a query goes to FDTD "
                         "(--workload c4); a comma-separated list times the service at "
                         "each, the first one reported as the result")
    ap.add_argument("--sweep-dtypes", default="float32,float16,bfloat16",
                    help="storage types compared with the FP64 run in --precision-sweep "
                         "(float16-nearest / bfloat16-nearest: narrow storage rounded to "
                         "nearest instead of stochastically)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the multi-GPU halo exchange is not implemented; "
              "run with --gpus 1", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import __graft_entry__
    import scenarios as S

    __graft_entry__.ensure_built()  # the in-tree HIP library, built once if missing
    if args.cpu_baseline_full:
        res = cpu_legs("vacuum_s7_r80", warmup=32, steps=32, repeats=3)
        for leg, v in res["legs"].items():
            v["published_steps_per_s"] = PUBLISHED_S7_R80_F32_STEPS_PER_S[leg]
        print(json.dumps({"metric": "CPU steps/s, reference benchmark method", **res}))
        return 0
    torch.cuda.set_device(0)
    if args.workload == "c2":
        print(json.dumps(ensemble_bench(args)))
        return 0
    if args.workload == "c4":
        print(json.dumps(surrogate_bench(args)))
        return 0
    if args.workload == "gemm":
        print(json.dumps(gemm_bench(args)))
        return 0
    if args.precision_sweep:
        sweep = precision_sweep(args)
        print(json.dumps({"metric": "storage precision sweep: error vs FP64 GPU run",
                          "workload": args.workload, "steps": args.steps,
                          "graph_chunk": args.chunk, "results": sweep}))
        return 0
    sc = S.scenario(args.workload)
    chunk = graph_chunk(args.steps, args.chunk)
    t_setup = time.perf_counter()
    sim = S.build_sim(sc, compile_step=True, compile_chunk_size=chunk)
    if sc["seed_fields"]:
        S.apply_seed_fields(sim, lambda k, v: setattr(sim, k, v))
    pts = S.receiver_points(sc)
    v, l, w = S.receiver_tables(sim.mesh, sim.altitudes_m, pts)
    setup_s = time.perf_counter() - t_setup
    print(f"setup {setup_s:.1f} s  V={sim.mesh.n_vertices} R={len(sim.radial_midpoints_m)} "
          f"dt={sim.time_step_s:.4e}",
          file=sys.stderr, flush=True)
    # Receivers, the trace buffer, the per-step source currents and the step
    # graph are prepared before the clock starts: the timed region is exactly
    # one geo_step call of K steps (graph replays of `chunk` steps, receiver
    # sampling inside the H kernel), bracketed by device synchronisation.
    sim.record_er_observations(v, l, w, args.warmup)
    sim.prepare_compiled_step()
    traces_dev = torch.zeros((args.steps + 1, len(pts)), dtype=sim.backend.compute_dtype,
                             device=sim.backend.torch_device)
    currents = sim._source_currents(args.steps)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    from ionosphere_fdtd_amd import _native

    _native.check(sim._lib.geo_step(sim._plan, currents.data_ptr(), args.steps, chunk,
                                    traces_dev.data_ptr(), sim._stream()), sim._lib)
    ev1.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    event_ms = ev0.elapsed_time(ev1)
    sim.steps += args.steps
    sim.time_s = sim.steps * sim.time_step_s
    traces = traces_dev.cpu().numpy()
    assert np.all(np.isfinite(traces))
    V, R = sim.mesh.n_vertices, len(sim.radial_midpoints_m)
    ms_per_step = elapsed * 1e3 / args.steps
    value = args.steps * V * R / elapsed / 1e6
    print(f"timed {args.steps} steps in {elapsed:.4f} s (events {event_ms:.3f} ms, "
          f"graph chunk {chunk})", file=sys.stderr, flush=True)
    # per-kernel durations with HIP events on the launch stream
    h_ms, e_ms = sim.kernel_times_ms(args.kernel_steps)
    bytes_ = algorithmic_bytes(sim)
    dominant = "h_step_kernel" if h_ms >= e_ms else "e_step_kernel"
    dom_ms = max(h_ms, e_ms)
    dom_bytes = bytes_["h"] if dominant == "h_step_kernel" else bytes_["e"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = bytes_["h"] + bytes_["e"]
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "geodesic FDTD s7 x 64 radial, FP32, exponential loss, "
                               "Drude plasma >60 km, SI ground, 20 Hz source, 32 receivers",
                   "cells": V * R, "dof": int(V * (R + 1) + sim.mesh.n_edges * (2 * R + 1)
                                              + sim.mesh.n_faces * R),
                   "time_step_s": sim.time_step_s, "graph_chunk": chunk,
                   "timed_path": (f"{args.steps // chunk} HIP graph replays of {chunk} steps"
                                  if chunk > 1 else f"{args.steps} single-step launches"),
                   "thread_mapping": {"h_step_kernel": MAPPING_NAMES[sim._layout.h_thread_mapping],
                                      "e_step_kernel": MAPPING_NAMES[sim._layout.e_thread_mapping]},
                   "event_ms_per_step": round(event_ms / args.steps, 5)},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(dominant, args.workload),
                     "algorithmic_bytes": dom_bytes,
                     "algorithmic_bytes_per_kernel": {"h_step_kernel": bytes_["h"],
                                                      "e_step_kernel": bytes_["e"]},
                     "kernel_ms": {"h_step_kernel": round(h_ms, 5), "e_step_kernel": round(e_ms, 5)},
                     "step_fraction": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if args.workload != "c1":
        result["config"]["workload"] = args.workload
    if args.stream_ceiling:
        from ionosphere_fdtd_amd import _native

        probe = {}
        for mode, label in ((0, "copy"), (1, "read"), (2, "write"), (4, "copy_nt"),
                            (5, "read_nt"), (6, "write_nt"), (8, "edge_pattern"), (9, "edge_pattern_indexed")):
            probe[label] = round(_native.stream_probe(torch, 1 << 30, mode), 1)
        result["stream_probe_gbs"] = probe
    if args.copy_calibration:
        # four launches of the library's probe copy kernel, 512 MiB each way:
        # their FETCH_SIZE / WRITE_SIZE readings calibrate the counters
        from ionosphere_fdtd_amd import _native

        _native.stream_probe(torch, 512 << 20, 0, repeats=3)
        result["copy_calibration_bytes_each_way"] = 512 << 20
    if not args.no_cpu_baseline and args.workload == "c1":
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
