"""Benchmark: solves/s + p50/p99 latency at <= 1e-10 relative residual,
FP32 factor + FP64 refinement vs plain FP64, on the wing-box shell
(box 4 x 1 x 0.2 m, 256 x 64 CQUAD4 on each of its four faces, per-element
thickness U[1,4] mm, root clamped, 8 right-hand sides of random nodal forces).

One step = one pass of the hot path over one request: element generation +
deterministic assembly of K, BC partition, numeric factorization, and the 8
solves refined to ||K_ff u - F_ff|| / ||F_ff|| <= 1e-10.  The symbolic analysis
(ordering, supernodes; topology only) is computed once before timing and reused,
as the serving loop does; its time is reported as ``symbolic_s``.

The default run prints, besides the headline, the rooflines of the element and
assembly kernels, the FP32/FP64 ratio of every other configuration with the
ceiling its measured terms allow (``configs``), the serving stream at both
precisions (``serving``) and the CPU baseline.  Prints ONE JSON line.
``--gpus N`` with N > 1 is rejected: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBS = 8000.0
MI355X_FP32_TFLOPS = 157.3
MI355X_FP64_TFLOPS = 78.6
METRIC = "solves/s + p50/p99 latency @1e-10 residual: FP32-factor+refine vs FP64"
FAMILIES = ["element", "assembly", "factor", "solve", "spmv", "factor_update", "factor_panel", "unused7",
            "fwd_flow", "bwd_flow", "factor_schur", "factor_inverse"]
# algorithmic bytes of one element's generation: read its grid points' coordinates,
# its connectivity and its section record, write its matrix once
ELEM_READ = {"CQUAD4": 4 * 24 + 4 * 4 + 8 * 8, "CTRIA3": 3 * 24 + 3 * 4 + 8 * 8, "CTETRA10": 10 * 24 + 10 * 4 + 2 * 8}
ELEM_DOF = {"CQUAD4": 24, "CTRIA3": 18, "CTETRA10": 30}
ELEM_NODES = {"CQUAD4": 4, "CTRIA3": 3, "CTETRA10": 10}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def omp_threads() -> int:
    try:
        return int(os.environ.get("OMP_NUM_THREADS", "")) or os.cpu_count()
    except ValueError:
        return os.cpu_count()


def read_traffic(path: str, precision: int | None):
    """HBM bytes per launch of a kernel measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by a copy kernel of known size), from the committed
    summary; None when absent."""
    key = f"hbm_bytes_per_launch_fp{precision}:" if precision else "hbm_bytes_per_launch:"
    try:
        for line in open(path):
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_cpu_baseline(nx_sample: int):
    """The FP64 CPU oracle on a bounded sample of the workload: the same wing box
    cut to nx_sample spanwise cells (same cross-section, same 8 right-hand sides).
    Element generation and assembly run on all OpenMP threads; the up-looking
    scalar Cholesky factor and the solves on one.  Returns solves/s on that sample."""
    from jf_2026_05_25_amd import meshes
    from oracle import oracle as O
    m = meshes.wing_box(nx_sample, 64, nrhs=8)
    t0 = time.perf_counter()
    K = O.assemble(m)
    t1 = time.perf_counter()
    fixed = O.bc_partition(K, m.spc)
    u, rr, info = O.solve(K, fixed, m.loads, order="nd", xyz=m.xyz)
    dt = time.perf_counter() - t0
    th = omp_threads()
    return {"value": 8.0 / dt, "unit": "solves/s", "cores": th, "kind": "port",
            "sample": f"oracle FP64 assemble+BC+Cholesky+8 solves on the wing box cut to {nx_sample} spanwise "
                      f"cells ({m.ndof} DOFs, {int((fixed == 0).sum())} free); {dt:.2f} s, of which element "
                      f"generation + assembly {t1 - t0:.2f} s on {th} OpenMP threads and the factor + solves on "
                      f"1 thread (scalar up-looking Cholesky); max relative residual {rr.max():.2e}; "
                      f"CPU {cpu_model()}",
            "seconds": dt}


# ------------------------------------------------------------------ one request
class RequestBench:
    """One model on one context: topology set up once, then timed requests
    (re-generate + re-assemble K, BC partition, factor, refined solves) with the
    right-hand sides and solutions in HBM."""

    def __init__(self, model, ctx, tol, max_refine):
        from jf_2026_05_25_amd import solver
        self.solver = solver
        self.model, self.ctx, self.tol, self.max_refine = model, ctx, tol, max_refine
        t0 = time.perf_counter()
        self.mesh = solver.DeviceMesh(model, ctx)
        self.pattern = solver.Pattern(self.mesh)
        self.K = solver.DeviceMatrix(self.mesh, self.pattern, ctx)
        self.t_setup = time.perf_counter() - t0
        self.fixed = solver.bc_partition(self.K, model.spc)
        t0 = time.perf_counter()
        self.S = solver.Symbolic(self.K, self.mesh, self.fixed, ctx)
        self.t_sym = time.perf_counter() - t0
        self.nrhs = model.loads.shape[1]
        Bc = np.ascontiguousarray(model.loads.T)
        self.b = solver.DeviceBuffer(Bc.nbytes, ctx)
        self.b.upload(Bc)
        self.x = solver.DeviceBuffer(Bc.nbytes, ctx)

    def step(self, precision, factor, host_buffers=False):
        self.K.reassemble(self.mesh)
        fx = self.solver.bc_partition(self.K, self.model.spc)
        if not np.array_equal(fx, self.fixed):
            raise RuntimeError("free set changed between steps")
        if factor is None:
            factor = self.solver.Factor(self.S, self.K, precision, self.ctx)
        else:
            factor.refactor(self.K)
        if host_buffers:
            _, rr, it, ok = factor.solve(self.K, self.model.loads, tol=self.tol, max_refine=self.max_refine)
        else:
            rr, it, ok = factor.solve_device(self.K, self.b, self.x, self.nrhs, tol=self.tol,
                                             max_refine=self.max_refine)
        return factor, rr, it, ok

    def run(self, precision, steps, warmup, host_steps=0):
        F = None
        for _ in range(warmup):
            F, rr, it, ok = self.step(precision, F)
        self.ctx.sync()
        lat, conv = [], True
        t_all = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            F, rr, it, ok = self.step(precision, F)
            lat.append(time.perf_counter() - ts)
            conv = conv and ok
        self.ctx.sync()
        total = time.perf_counter() - t_all
        u = self.x.download(np.zeros((self.nrhs, self.model.ndof))).T.copy()
        host_rate = None
        if host_steps:
            t_h = time.perf_counter()
            for _ in range(host_steps):
                F, _, _, _ = self.step(precision, F, host_buffers=True)
            self.ctx.sync()
            host_rate = self.nrhs * host_steps / (time.perf_counter() - t_h)
        # per-family kernel times (HIP events) on one extra instrumented step
        self.ctx.enable_timing(2)
        self.ctx.reset_timing()
        F, _, _, _ = self.step(precision, F)
        fam = {name: self.ctx.kernel_time(i) for i, name in enumerate(FAMILIES)}
        self.ctx.enable_timing(0)
        out = {"total": total, "lat": np.array(lat), "u": u, "relres": rr, "iters": it, "fam": fam,
               "factor_bytes": F.device_bytes(), "converged": conv, "host_rate": host_rate,
               "solves_per_s": self.nrhs * steps / total}
        F.free()
        return out

    def free(self):
        for o in (self.x, self.b, self.S, self.K, self.pattern, self.mesh):
            o.free()


def request_summary(r, nrhs):
    return {"solves_per_s": r["solves_per_s"], "ms_per_request": 1000.0 * r["total"] / len(r["lat"]),
            "p50_latency_ms": float(np.percentile(r["lat"], 50) * 1000),
            "p99_latency_ms": float(np.percentile(r["lat"], 99) * 1000),
            "max_relative_residual": float(np.max(r["relres"])), "converged": bool(r["converged"]),
            "triangular_solves_per_rhs": [int(x) for x in r["iters"]],
            "kernel_ms": {k: round(v[0], 4) for k, v in r["fam"].items() if v[1]}}


def ratio_ceiling(r32, r64):
    """Upper bound on the FP32/FP64 request-rate ratio from the measured terms:
    the same element generation + assembly, an FP32 factor at best twice as fast
    as the FP64 one (twice the MFMA rate, half the bytes), and each FP32
    triangular-solve pair at best half the cost of an FP64 pair, times the
    number of pairs the FP32 factor needs."""
    f32, f64 = r32["fam"], r64["fam"]
    gen = f64["element"][0] + f64["assembly"][0]
    fac64 = f64["factor"][0]
    sol64 = f64["solve"][0] + f64["spmv"][0]
    n64 = max(1.0, float(np.max(r64["iters"])))
    n32 = max(1.0, float(np.max(r32["iters"])))
    t64 = gen + fac64 + sol64
    t32_floor = gen + fac64 / 2 + (sol64 / n64) * n32 / 2
    return t64 / t32_floor


def family_rooflines(fam, model, pattern_nnz):
    """Element generation (q4 prep + wavefront kernel, CTRIA3, CTETRA10 kernels of
    one reassembly) and the assembly kernel against HBM, with the algorithmic
    bytes per launch; and the two together against the fused figure (read the
    mesh once, write K once)."""
    counts = {"CQUAD4": len(model.quads), "CTRIA3": len(model.trias), "CTETRA10": len(model.tets)}
    ntrip = sum(n * ELEM_DOF[k] ** 2 for k, n in counts.items())
    out = {}
    e_ms, e_n = fam["element"]
    if e_n:
        bytes_e = sum(n * (ELEM_READ[k] + 8 * ELEM_DOF[k] ** 2) for k, n in counts.items())
        out["element"] = {"bound": "hbm", "kernel": "q4_prep_global_kernel + q4_elem_kernel (+ t3, tet10)",
                          "achieved": bytes_e / (e_ms * 1e-3) / 1e9, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                          "frac": bytes_e / (e_ms * 1e-3) / 1e9 / MI355X_HBM_GBS, "traffic": None,
                          "bytes_per_launch": bytes_e, "avg_launch_ms": e_ms, "launches": 1,
                          "elements": counts}
    a_ms, a_n = fam["assembly"]
    if a_n:
        bytes_a = 12 * ntrip + 16 * pattern_nnz + 8
        out["assembly"] = {"bound": "hbm", "kernel": "assemble_kernel",
                           "achieved": bytes_a / (a_ms / a_n * 1e-3) / 1e9, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                           "frac": bytes_a / (a_ms / a_n * 1e-3) / 1e9 / MI355X_HBM_GBS, "traffic": None,
                           "bytes_per_launch": bytes_a, "avg_launch_ms": a_ms / a_n, "launches": a_n,
                           "triplets": ntrip, "nnz": pattern_nnz}
    if e_n and a_n:
        fused = 24 * model.n_nodes + sum(n * (4 * ELEM_NODES[k] + 8 + 4 * ELEM_DOF[k] ** 2)
                                         for k, n in counts.items()) + 8 * pattern_nnz
        t = (e_ms + a_ms) * 1e-3
        out["generation_and_assembly"] = {
            "bound": "hbm", "achieved": fused / t / 1e9, "peak": MI355X_HBM_GBS, "unit": "GB/s",
            "frac": fused / t / 1e9 / MI355X_HBM_GBS, "bytes": fused, "ms": e_ms + a_ms,
            "note": "fused figure: coordinates, connectivity, thickness and slot map read once, K values written once"}
    return out


def solve_rooflines(fam, info, nrhs, precision):
    es = 4 if precision == 32 else 8
    tname = "float" if precision == 32 else "double"
    nr_pad = 4 if nrhs <= 4 else (8 if nrhs <= 8 else 16)
    peak = MI355X_FP32_TFLOPS if precision == 32 else MI355X_FP64_TFLOPS
    lines = {}
    n_free = info["n_free"]
    vec = {"fwd_flow": 2 * n_free + 2 * info["upd_rows"], "bwd_flow": 2 * n_free + info["upd_rows"]}
    for fname, kname in (("fwd_flow", f"fwd_flow_kernel<{tname}, {nr_pad}>"),
                         ("bwd_flow", f"bwd_flow_kernel<{tname}, {nr_pad}>")):
        f_ms, f_n = fam[fname]
        if not f_n:
            continue
        per_launch = es * (info["solve_entries"] + nrhs * vec[fname])
        avg_ms = f_ms / f_n
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        lines[fname] = (f_ms, {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": MI355X_HBM_GBS,
                               "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                               "bytes_per_launch": per_launch, "avg_launch_ms": avg_ms, "launches": f_n})
    for fname, kname, fkey in (("factor_update", f"update_kernel<{tname}>", "update_flops"),
                               ("factor_schur", f"schur_kernel<{tname}>", "schur_flops")):
        k_ms, k_n = fam[fname]
        if not k_n:
            continue
        per_launch_flops = info[fkey] / k_n
        avg_ms = k_ms / k_n
        achieved = per_launch_flops / (avg_ms * 1e-3) / 1e12
        lines[fname] = (k_ms, {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": peak,
                               "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                               "flops_per_launch": per_launch_flops, "avg_launch_ms": avg_ms, "launches": k_n})
    return lines


def measure_load_vector(ctx, mesh, model, calls: int = 200):
    """The load vector of one load set on the device (jfem_load_vector_dev): a
    uniform 5 kPa PLOAD4 on every shell element plus the model's first right-hand
    side as FORCE cards, into HBM.  Per call: two launches; the algorithmic bytes
    are the set's card data, the coordinates and the vector written once."""
    from jf_2026_05_25_amd import solver
    ne = len(model.quads) + len(model.trias)
    f = model.loads[:, 0].reshape(-1, 6)[:, :3]
    nodes = np.nonzero(np.any(f != 0.0, axis=1))[0]
    mag = np.linalg.norm(f[nodes], axis=1)
    ls = solver.LoadSet(mesh, pload4=(np.arange(ne), np.full(ne, 5000.0)),
                        force=(nodes, mag, f[nodes] / mag[:, None]))
    F = solver.DeviceBuffer(8 * model.ndof, ctx)
    ls.vector_dev(F.ptr)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(calls):
        ls.vector_dev(F.ptr)
    ctx.sync()
    us = (time.perf_counter() - t0) / calls * 1e6
    n = model.n_nodes
    npl_slots = 4 * len(model.quads) + 3 * len(model.trias)
    algo = (24 * n + 8 * (n + 1) + 48 * n                 # coordinates, per-grid offsets, F
            + 4 * ne + 8 * ne + 4 * npl_slots + 32 * ne    # cards, connectivity, nodal shares written
            + 4 * npl_slots + 32 * npl_slots               # per-grid card lists and shares read
            + 8 * (n + 1) + 32 * len(nodes))               # FORCE offsets and records
    ls.free()
    F.free()
    return {"us_per_call": us, "pload4_cards": int(ne), "force_cards": int(len(nodes)),
            "algorithmic_bytes": int(algo), "achieved_GBps": algo / (us * 1e-6) / 1e9,
            "note": "wall clock per call over 200 back-to-back calls (two launches each); "
                    "launch-latency bound at this size"}


def config_line(name, model, ctx, args, steps, warmup):
    """FP32 and FP64 requests of one configuration: rates, ratio and its ceiling."""
    rb = RequestBench(model, ctx, args.tol, args.max_refine)
    info = rb.S.info
    try:
        r = {p: rb.run(p, steps, warmup) for p in (32, 64)}
    finally:
        rb.free()
    err = float(np.max(np.linalg.norm(r[32]["u"] - r[64]["u"], axis=0) / np.linalg.norm(r[64]["u"], axis=0)))
    return {"workload": name, "dofs": model.ndof, "free_dofs": int(info["n_free"]),
            "rhs_per_request": model.nrhs,
            "fp32": request_summary(r[32], model.nrhs), "fp64": request_summary(r[64], model.nrhs),
            "fp32_vs_fp64": r[32]["solves_per_s"] / r[64]["solves_per_s"],
            "fp32_vs_fp64_ceiling": ratio_ceiling(r[32], r[64]), "fp32_refined_error_vs_fp64": err}


# ------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["wingbox", "plate", "block", "variants", "serving"], default="wingbox",
                    help="wingbox (default, the benchmark line): 4x1x0.2 m wing box, 256x64 CQUAD4 per face, 8 RHS; "
                         "plate: 40x40 CQUAD4 flat plate; block: 24^3 hexes x 5 CTETRA10; variants: 64 design "
                         "variants of a 128x128 stiffened plate, 4 RHS, one symbolic analysis; serving: Poisson "
                         "stream of shell requests (2k-200k DOFs, 10% CTRIA3, 1-16 RHS)")
    ap.add_argument("--requests", type=int, default=512, help="serving: number of requests")
    ap.add_argument("--slots", type=int, default=8, help="serving: concurrent request slots (HIP streams)")
    ap.add_argument("--utilization", type=float, default=0.7, help="serving: offered load / measured service rate")
    ap.add_argument("--variants", type=int, default=64, help="variants: number of design variants")
    ap.add_argument("--nx", type=int, default=256, help="spanwise cells (256 = the benchmark workload)")
    ap.add_argument("--nw", type=int, default=64, help="cells across each face")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-refine", type=int, default=200,
                    help="bound on preconditioner applications (triangular solves) per request, minus one")
    ap.add_argument("--cpu-sample-nx", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the per-configuration FP32/FP64 lines")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream in the default run")
    ap.add_argument("--only", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before timing, run a fill and an elementwise multiply over this many MiB (torch) so "
                         "that hardware counters (FETCH_SIZE / WRITE_SIZE) can be scaled by kernels of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures the single-GPU path; N > 1 is not run "
                                   "(see DESIGN.md: how the path would shard)", "n_gpus": args.gpus}))
        sys.exit(2)

    from jf_2026_05_25_amd import meshes, solver

    if args.calibrate_copy_mib > 0:
        # two elementwise kernels of known traffic: a fill (writes N bytes) and
        # dst = 2 * src (reads N, writes N)
        import torch
        nel = args.calibrate_copy_mib * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        dst = src * 2.0
        torch.cuda.synchronize()
        del src, dst

    if args.workload == "variants":
        print(json.dumps(run_variants(args)))
        return
    if args.workload == "serving":
        print(json.dumps(run_serving_both(args)))
        return
    if args.workload == "plate":
        model = meshes.flat_plate(40, 40)
        workload = "flat plate 0.5x0.5 m, 40x40 CQUAD4, t=2 mm, edge clamped, 1 kPa pressure, 1 RHS"
        data = "synthetic (generated flat plate, uniform pressure)"
    elif args.workload == "block":
        model = meshes.solid_block(24, nrhs=1)
        workload = "unit cube, 24^3 hexes x 5 CTETRA10, base clamped, random top traction, 1 RHS"
        data = "synthetic (generated tetrahedral block, random top-face traction)"
    else:
        model = meshes.wing_box(args.nx, args.nw, nrhs=8)
        workload = f"wing box 4x1x0.2 m, {args.nx}x{args.nw} CQUAD4 per face, 8 RHS"
        data = "synthetic (generated wing-box mesh, random per-element thickness and nodal forces)"
    ctx = solver.Context(0)
    rb = RequestBench(model, ctx, args.tol, args.max_refine)
    info = rb.S.info
    nrhs = rb.nrhs
    results = {}
    for prec in ([32, 64] if args.only == "both" else ([32] if args.only == "fp32" else [64])):
        results[prec] = rb.run(prec, args.steps, args.warmup, host_steps=max(1, min(args.steps, 5)))

    main_prec = 32 if 32 in results else 64
    r = results[main_prec]
    out = {
        "metric": METRIC,
        "value": r["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * r["total"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 factor + fp64 refinement" if main_prec == 32 else "fp64",
        "data": data,
        "config": {"workload": workload,
                   "dofs": model.ndof, "free_dofs": int(info["n_free"]), "cquad4": int(len(model.quads)),
                   "ctria3": int(len(model.trias)), "ctetra10": int(len(model.tets)),
                   "rhs_per_request": nrhs, "tol": args.tol},
        "p50_latency_ms": float(np.percentile(r["lat"], 50) * 1000),
        "p99_latency_ms": float(np.percentile(r["lat"], 99) * 1000),
        "max_relative_residual": float(np.max(r["relres"])),
        "converged": bool(r["converged"]),
        "triangular_solves_per_rhs": [int(x) for x in r["iters"]],
        "dof_per_s": info["n_free"] * nrhs * args.steps / r["total"],
        "host_buffers_solves_per_s": r["host_rate"],
        "symbolic_s": rb.t_sym,
        "setup_s": rb.t_setup,
        "symbolic": {k: (float(v) if isinstance(v, float) else int(v)) for k, v in info.items()},
    }
    if 64 in results and 32 in results:
        r64 = results[64]
        out["fp64_solves_per_s"] = r64["solves_per_s"]
        out["fp32_vs_fp64_speedup"] = r["solves_per_s"] / r64["solves_per_s"]
        out["fp32_vs_fp64_ceiling"] = ratio_ceiling(r, r64)
        out["fp64_p50_latency_ms"] = float(np.percentile(r64["lat"], 50) * 1000)
        out["fp64_p99_latency_ms"] = float(np.percentile(r64["lat"], 99) * 1000)
        out["fp64_max_relative_residual"] = float(np.max(r64["relres"]))
        out["fp64_converged"] = bool(r64["converged"])
        out["fp64_host_buffers_solves_per_s"] = r64["host_rate"]
        out["fp64_triangular_solves_per_rhs"] = [int(x) for x in r64["iters"]]
        u32, u64 = r["u"], r64["u"]
        out["fp32_refined_error_vs_fp64"] = float(np.max(np.linalg.norm(u32 - u64, axis=0) /
                                                         np.linalg.norm(u64, axis=0)))
        out["fp64_kernel_ms"] = {k: round(v[0], 4) for k, v in r64["fam"].items() if v[1]}
    # kernel-family breakdown; roofline of the dominant kernel of the headline path
    # (largest total time in the instrumented step) among the factor's trailing
    # update (matrix cores) and the two dataflow triangular-solve kernels (HBM:
    # per launch they stream the factor entries one solve needs plus the vectors);
    # element generation and assembly beside them
    fam = r["fam"]
    out["kernel_ms"] = {k: round(v[0], 4) for k, v in fam.items() if v[1]}
    lines = solve_rooflines(fam, info, nrhs, main_prec)
    counters = args.workload == "wingbox" and args.nx == 256 and args.nw == 64   # the workload the counters were read on
    if lines:
        dom = max(lines, key=lambda k: lines[k][0])
        roof = lines[dom][1]
        if counters:
            kfile = roof["kernel"].split("<")[0]
            roof["traffic"] = read_traffic(os.path.join(ROOT, "profiles", f"{kfile}_wingbox_traffic.txt"), main_prec)
        out["roofline"] = roof
        out["other_rooflines"] = {k: v[1] for k, v in lines.items() if k != dom}
        gen = family_rooflines(fam, model, rb.pattern.nnz)
        if counters:
            for k, kfile in (("element", "q4_elem_kernel"), ("assembly", "assemble_kernel")):
                if k in gen:
                    gen[k]["traffic"] = read_traffic(os.path.join(ROOT, "profiles", f"{kfile}_wingbox_traffic.txt"),
                                                     None)
        out["other_rooflines"].update(gen)
    if args.workload == "wingbox":
        out["load_vector"] = measure_load_vector(ctx, rb.mesh, model)
    rb.free()
    if args.workload == "wingbox" and not args.no_configs and args.only == "both":
        cfg = {"C1_plate_40x40": config_line("flat plate 40x40 CQUAD4, 1 RHS", meshes.flat_plate(40, 40), ctx, args,
                                             20, 3),
               "C3_block_24": config_line("solid block 24^3 x 5 CTETRA10, 1 RHS", meshes.solid_block(24, nrhs=1),
                                          ctx, args, 3, 1)}
        v = run_variants(args, count=min(args.variants, 64))
        cfg["C5_variants"] = {k: v[k] for k in ("config", "value", "fp64_solves_per_s", "fp32_vs_fp64",
                                                "p50_latency_ms", "p99_latency_ms", "fp64_p50_latency_ms",
                                                "fp64_p99_latency_ms", "max_relative_residual",
                                                "fp64_max_relative_residual", "fp32_refined_error_vs_fp64",
                                                "launch_batch") if k in v}
        cfg["C2_wingbox"] = {"fp32_vs_fp64": out.get("fp32_vs_fp64_speedup"),
                             "fp32_vs_fp64_ceiling": out.get("fp32_vs_fp64_ceiling")}
        cfg["sol105_plate_128"] = buckling_line(ctx, 128)
        out["configs"] = cfg
    if args.workload == "wingbox" and not args.no_serving and args.only == "both":
        out["serving"] = run_serving_both(args, as_section=True)
    if not args.no_cpu_baseline and args.workload == "wingbox":
        cb = run_cpu_baseline(args.cpu_sample_nx)
        cb.pop("seconds", None)
        out["cpu_baseline"] = cb
    ctx.close()
    print(json.dumps(out))


def buckling_line(ctx, n: int = 128, roots: int = 4, reps: int = 5):
    """SOL 105 on the simply supported n x n plate under unit uniaxial compression
    (meshes.ss_plate_compression): the static preload (FP64 factor, refined), the
    differential stiffness on K's pattern (per-element generation and assembly,
    timed over ``reps`` calls), and the first ``roots`` buckling roots by subspace
    inverse iteration on the static factor (EIGB 0 < lambda, NDP = roots).  The
    first root is compared with Timoshenko-Gere's sigma_cr = 4 pi^2 D / (b^2 h)."""
    from jf_2026_05_25_amd import buckling, meshes, solver
    m = meshes.ss_plate_compression(n)
    asm = solver.assemble_stiffness(m, ctx)
    cache = solver.LinearSolveCache()
    t0 = time.perf_counter()
    u, fixed, _, sdiag = solver.apply_bc_and_solve(asm, m, precision=64, linear_cache=cache)
    t_static = time.perf_counter() - t0
    u = u.reshape(m.ndof, -1)[:, 0]
    fx, S, F = next(iter(cache.entries.values()))[:3]
    G = buckling.GeometricStiffness(asm, u, ctx)
    G.free()
    t0 = time.perf_counter()
    for _ in range(reps):
        G = buckling.GeometricStiffness(asm, u, ctx)
        G.free()
    t_kg = (time.perf_counter() - t0) / reps
    G = buckling.GeometricStiffness(asm, u, ctx)
    t0 = time.perf_counter()
    res = buckling.buckling_roots(asm.K, G, fx, F, (0.0, math.inf, roots))
    t_roots = time.perf_counter() - t0
    G.free(); F.free(); S.free()
    D = 210e9 * 0.01 ** 3 / (12 * (1 - 0.3 ** 2))
    scr = 4 * math.pi ** 2 * D / 0.01
    nq = len(m.quads)
    kg_bytes = nq * (4 * 24 + 16 + 64 + 4 * 48 + 4608) + 12 * 576 * nq + 16 * asm.pattern.nnz
    out = {"workload": f"SOL 105: simply supported {n}x{n} CQUAD4 plate under uniaxial compression, "
                       f"first {roots} roots",
           "dofs": m.ndof, "free_dofs": res.diagnostics["free_dofs"], "roots": [float(x) for x in res.eigenvalues],
           "first_root_vs_timoshenko": float(res.eigenvalues[0] / scr),
           "static_preload_ms": 1000 * t_static, "kg_generation_assembly_ms": 1000 * t_kg,
           "kg_algorithmic_bytes": kg_bytes, "kg_achieved_GBps": kg_bytes / t_kg / 1e9,
           "roots_ms": 1000 * t_roots, "subspace_iterations": res.diagnostics["iterations"],
           "block": res.diagnostics["block"], "triangular_solves": res.diagnostics["triangular_solves"],
           "max_mode_residual": float(max(res.diagnostics["mode_residual"])),
           "note": "Kg timed with the preload upload and a stream synchronization per call; roots: host "
                   "projected eigenproblem, device products and refined solves"}
    asm.K.free(); asm.pattern.free(); asm.mesh.free()
    return out


def run_variants(args, count=None, streams: int = 4):
    """Design variants of the 128 x 128 stiffened plate (per-element thickness
    x U[0.8, 1.2]), 4 RHS each, sharing one symbolic analysis.  One step runs
    every variant: element generation + assembly of all variants in one launch
    sequence, the BC partition of each, the numeric factors of all variants in
    one launch sequence, then each variant's 4 refined solves (spread over
    ``streams`` HIP streams).  Sections and right-hand sides are resident in HBM
    before timing; a variant's latency runs from the step's start to its solve's
    end."""
    import threading
    from jf_2026_05_25_amd import meshes, solver
    count = count or args.variants
    base = meshes.stiffened_plate(128, nrhs=4)
    variants = meshes.design_variants(base, count)
    ctxs = [solver.Context(0) for _ in range(streams)]
    ctx = ctxs[0]
    mesh = solver.DeviceMesh(base, ctx)
    pattern = solver.Pattern(mesh)
    t0 = time.perf_counter()
    batch = solver.DeviceBatch(mesh, pattern, variants, ctx)
    t_upload = time.perf_counter() - t0
    batch.assemble()
    fixed = solver.bc_partition(batch.matrix(0), base.spc)
    t0 = time.perf_counter()
    S = solver.Symbolic(batch.matrix(0), mesh, fixed, ctx)
    t_sym = time.perf_counter() - t0
    nrhs = base.nrhs
    bufs = []
    for v in variants:
        Bc = np.ascontiguousarray(v.loads.T)
        b = solver.DeviceBuffer(Bc.nbytes, ctx); b.upload(Bc)
        bufs.append((b, solver.DeviceBuffer(Bc.nbytes, ctx)))

    def step(prec):
        ts = time.perf_counter()
        batch.assemble()
        for v, m in enumerate(variants):
            if not np.array_equal(solver.bc_partition(batch.matrix(v), m.spc), fixed):
                raise RuntimeError("free set changed between variants")
        batch.factor(S, prec)
        t_fac = time.perf_counter() - ts
        done = [0.0] * len(variants)
        rr = [None] * len(variants)
        its = [None] * len(variants)
        errs = []

        def worker(k):
            c = ctxs[k]
            for v in range(k, len(variants), streams):
                try:
                    b, x = bufs[v]
                    r, it, ok = batch.factor_view(v).solve_device(batch.matrix(v), b, x, nrhs, tol=args.tol,
                                                                  max_refine=args.max_refine, ctx=c)
                    rr[v], its[v] = r, it
                    if not ok:
                        errs.append(f"variant {v} not converged")
                except Exception as e:   # reported after the step
                    errs.append(repr(e))
                done[v] = time.perf_counter() - ts
        th = [threading.Thread(target=worker, args=(k,)) for k in range(streams)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errs:
            raise RuntimeError(errs[0])
        return time.perf_counter() - ts, t_fac, done, rr, its

    res = {}
    for prec in (32, 64):
        step(prec)                       # warm-up
        for c in ctxs:
            c.sync()
        ctx.enable_timing(1)
        ctx.reset_timing()
        total, t_fac, done, rr, its = step(prec)
        fam = {name: ctx.kernel_time(i) for i, name in enumerate(FAMILIES[:3])}
        ctx.enable_timing(0)
        us = [bufs[v][1].download(np.zeros((nrhs, base.ndof))).T.copy() for v in range(len(variants))]
        res[prec] = {"total": total, "t_fac": t_fac, "lat": np.array(done), "rr": max(float(np.max(r)) for r in rr),
                     "its": max(int(np.max(i)) for i in its), "u": us, "fam": fam}
    for b, x in bufs:
        b.free(); x.free()
    batch.free(); S.free(); pattern.free(); mesh.free()
    for c in ctxs:
        c.close()
    n_solves = nrhs * len(variants)
    r32, r64 = res[32], res[64]
    err = max(float(np.max(np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)))
              for a, b in zip(r32["u"], r64["u"]))
    return {"metric": METRIC,
            "value": n_solves / r32["total"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": 1000.0 * r32["total"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 factor + fp64 refinement",
            "data": "synthetic (generated stiffened plate, thickness variants, random nodal forces)",
            "config": {"workload": f"{len(variants)} design variants of a 128x128 CQUAD4 stiffened plate, 4 RHS each, "
                                   "one symbolic, batched generation + assembly + factor",
                       "dofs": base.ndof, "free_dofs": int(S.info["n_free"]), "cquad4": int(len(base.quads)),
                       "rhs_per_request": nrhs, "variants": len(variants), "tol": args.tol},
            "launch_batch": {"variants_per_launch": len(variants), "solve_streams": streams,
                             "fp32_generation_assembly_ms": r32["fam"]["element"][0] + r32["fam"]["assembly"][0],
                             "fp32_factor_ms": r32["fam"]["factor"][0],
                             "fp64_factor_ms": r64["fam"]["factor"][0],
                             "fp32_until_factors_ms": 1000 * r32["t_fac"], "fp64_until_factors_ms": 1000 * r64["t_fac"],
                             "section_upload_ms": 1000 * t_upload},
            "p50_latency_ms": float(np.percentile(r32["lat"], 50) * 1000),
            "p99_latency_ms": float(np.percentile(r32["lat"], 99) * 1000),
            "max_relative_residual": r32["rr"], "max_triangular_solves": r32["its"],
            "fp64_solves_per_s": n_solves / r64["total"], "fp32_vs_fp64": r64["total"] / r32["total"],
            "fp64_p50_latency_ms": float(np.percentile(r64["lat"], 50) * 1000),
            "fp64_p99_latency_ms": float(np.percentile(r64["lat"], 99) * 1000),
            "fp64_max_relative_residual": r64["rr"], "fp64_max_triangular_solves": r64["its"],
            "fp32_refined_error_vs_fp64": err, "symbolic_s": t_sym,
            "symbolic": {k: (float(v) if isinstance(v, float) else int(v)) for k, v in S.info.items()}}


def run_serving(args, precision, reqs, slots, realtime=True):
    """One pass of the stream through a fresh server at ``precision``; with
    ``realtime`` False every request is visible at once (capacity)."""
    from jf_2026_05_25_amd import serving
    srv = serving.SolveServer(slots=slots, precision=precision, tol=args.tol)
    try:
        t0 = time.perf_counter()
        res = srv.serve(reqs, realtime=realtime)
        wall = time.perf_counter() - t0
        st = serving.summarize(res)
        peak = srv.peak_active
    finally:
        srv.close()
    ok = [r for r in res if r.ok]
    st.update({"wall_s": wall, "peak_concurrent_requests": peak,
               "fp64_fallbacks": sum(1 for r in ok if r.timings.get("factor_precision", precision) != precision),
               "factor_cache_hits": sum(1 for r in ok if r.cache_hit),
               "mean_stage_s": serving.stage_times(ok),
               "failed": [r.error for r in res if not r.ok][:4]})
    if not realtime:
        st["solves_per_s"] = st["solves"] / wall if wall > 0 else 0.0
    return st


def run_serving_both(args, as_section=False):
    """Serving loop on a Poisson stream of shell requests (continuous batching over
    ``slots`` HIP streams, admission by free HBM, device-resident caches), once
    with FP32 factors and once with FP64 factors, on the same arrivals.  The
    arrival rate is ``utilization`` x the single-stream FP32 service rate measured
    on a calibration sample of the same distribution."""
    from jf_2026_05_25_amd import serving
    srv = serving.SolveServer(slots=1, precision=32, tol=args.tol)
    try:
        calib = serving.poisson_stream(16, 1.0, seed=7, precision=32)
        t0 = time.perf_counter()
        for q in calib:
            srv.process(srv.ctxs[0], q, t0)
            srv.clear_caches()
        service = (time.perf_counter() - t0) / len(calib)
    finally:
        srv.close()
    rate = args.utilization / service
    out, cap = {}, {}
    for prec in (32, 64):
        reqs = serving.poisson_stream(args.requests, rate, precision=prec)
        out[prec] = run_serving(args, prec, reqs, args.slots)
        # capacity: the same requests all arriving at t = 0 (the slots never idle)
        cap[prec] = run_serving(args, prec, [dataclasses.replace(q, arrival=0.0) for q in reqs], args.slots,
                                realtime=False)
    s32, s64 = out[32], out[64]
    cfg = (f"serving: {args.requests} requests, Poisson arrivals at {rate:.1f}/s ({args.utilization:.0%} of the "
           f"calibrated single-stream FP32 service rate), 2k-200k DOFs log-uniform, 10% CTRIA3, 1-16 RHS, "
           f"{args.slots} slots")
    sec = {"workload": cfg, "fp32": s32, "fp64": s64,
           "fp32_vs_fp64": s32["solves_per_s"] / s64["solves_per_s"] if s64["solves_per_s"] else None,
           "capacity": {"note": "the same requests all visible at t = 0: solves/s the slots sustain",
                        "fp32": cap[32], "fp64": cap[64],
                        "fp32_vs_fp64": cap[32]["solves_per_s"] / cap[64]["solves_per_s"]},
           "calibrated_service_ms": 1000.0 * service}
    if as_section:
        return sec
    return {"metric": METRIC, "value": s32["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
            "steps": args.requests, "warmup": len(calib),
            "ms_per_step": 1000.0 * s32["wall_s"] / max(1, args.requests), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 factor + fp64 refinement (fp64 factor on fp32 breakdown)",
            "data": "synthetic (Poisson stream of generated flat-plate / box shells, random loads)",
            "config": {"workload": cfg, "tol": args.tol},
            "p50_latency_ms": s32["p50_latency_ms"], "p99_latency_ms": s32["p99_latency_ms"],
            "max_relative_residual": s32["max_relative_residual"], "serving": sec}


if __name__ == "__main__":
    main()
