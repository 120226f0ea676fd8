"""Benchmark: solves/s + p50/p99 latency at <= 1e-10 relative residual,
FP32 factor + FP64 refinement vs plain FP64, on the wing-box shell
(box 4 x 1 x 0.2 m, 256 x 64 CQUAD4 on each of its four faces, per-element
thickness U[1,4] mm, root clamped, 8 right-hand sides of random nodal forces).

One step = one pass of the hot path over one request: element generation +
deterministic assembly of K, BC partition, numeric factorization, and the 8
solves refined to ||K_ff u - F_ff|| / ||F_ff|| <= 1e-10.  The symbolic analysis
(ordering, supernodes; topology only) is computed once before timing and reused,
as the serving loop does; its time is reported as ``symbolic_s``.

Prints ONE JSON line.  ``--gpus N`` with N > 1 is rejected: this build measures
one GPU.
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

MI355X_HBM_GBS = 8000.0
MI355X_FP32_TFLOPS = 157.3
MI355X_FP64_TFLOPS = 78.6


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def read_traffic(path: str, precision: int):
    """HBM bytes per launch of a kernel measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by a copy kernel of known size), from the committed
    summary; None when absent."""
    try:
        for line in open(path):
            if line.startswith(f"hbm_bytes_per_launch_fp{precision}:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_cpu_baseline(nx_sample: int):
    """The FP64 CPU oracle (single thread) on a bounded sample of the workload:
    the same wing box cut to nx_sample spanwise cells (same cross-section, same
    8 right-hand sides).  Returns solves/s on that sample."""
    from jf_2026_05_25_amd import meshes
    from oracle import oracle as O
    m = meshes.wing_box(nx_sample, 64, nrhs=8)
    t0 = time.perf_counter()
    K = O.assemble(m)
    fixed = O.bc_partition(K, m.spc)
    u, rr, info = O.solve(K, fixed, m.loads, order="nd", xyz=m.xyz)
    dt = time.perf_counter() - t0
    return {"value": 8.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 assemble+BC+Cholesky+8 solves on the wing box cut to {nx_sample} spanwise "
                      f"cells ({m.ndof} DOFs, {int((fixed == 0).sum())} free); {dt:.2f} s; "
                      f"max relative residual {rr.max():.2e}; CPU {cpu_model()}",
            "seconds": dt}


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
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent request slots (HIP streams)")
    ap.add_argument("--utilization", type=float, default=0.7, help="serving: offered load / measured service rate")
    ap.add_argument("--nx", type=int, default=256, help="spanwise cells (256 = the benchmark workload)")
    ap.add_argument("--nw", type=int, default=64, help="cells across each face")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-refine", type=int, default=200,
                    help="bound on preconditioner applications (triangular solves) per request, minus one")
    ap.add_argument("--cpu-sample-nx", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
        print(json.dumps(run_serving(args)))
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
    t0 = time.perf_counter()
    mesh = solver.DeviceMesh(model, ctx)
    pattern = solver.Pattern(mesh)
    K = solver.DeviceMatrix(mesh, pattern, ctx)
    t_setup = time.perf_counter() - t0
    fixed = solver.bc_partition(K, model.spc)
    t0 = time.perf_counter()
    S = solver.Symbolic(K, mesh, fixed, ctx)
    t_sym = time.perf_counter() - t0
    info = S.info
    B = model.loads
    nrhs = B.shape[1]
    # right-hand sides and solutions stay in HBM: the timed region moves nothing over PCIe
    # but the BC masks and per-RHS residuals
    Bc = np.ascontiguousarray(B.T)                 # ndof x nrhs column-major
    b_dev = solver.DeviceBuffer(Bc.nbytes, ctx)
    b_dev.upload(Bc)
    x_dev = solver.DeviceBuffer(Bc.nbytes, ctx)

    def step(precision, factor, host_buffers=False):
        """One request: reassemble, BC, factor, nrhs refined solves. Returns (factor, relres, iters, ok)."""
        K.reassemble(mesh)
        fx = solver.bc_partition(K, model.spc)
        if not np.array_equal(fx, fixed):
            raise RuntimeError("free set changed between steps")
        if factor is None:
            factor = solver.Factor(S, K, precision, ctx)
        else:
            factor.refactor(K)
        if host_buffers:
            _, rr, it, ok = factor.solve(K, B, tol=args.tol, max_refine=args.max_refine)
        else:
            rr, it, ok = factor.solve_device(K, b_dev, x_dev, nrhs, tol=args.tol, max_refine=args.max_refine)
        return factor, rr, it, ok

    results = {}
    for prec in ([32, 64] if args.only == "both" else ([32] if args.only == "fp32" else [64])):
        F = None
        for _ in range(args.warmup):
            F, rr, it, ok = step(prec, F)
        ctx.sync()
        lat = []
        conv = True
        t_all = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            F, rr, it, ok = step(prec, F)
            lat.append(time.perf_counter() - ts)
            conv = conv and ok
        ctx.sync()
        total = time.perf_counter() - t_all
        u = x_dev.download(np.zeros((nrhs, model.ndof))).T.copy()
        # the same steps with host arrays in and out (PCIe-inclusive; not the headline)
        nh = max(1, min(args.steps, 5))
        t_h = time.perf_counter()
        for _ in range(nh):
            F, _, _, _ = step(prec, F, host_buffers=True)
        ctx.sync()
        host_rate = nrhs * nh / (time.perf_counter() - t_h)
        # per-family kernel times (HIP events) on one extra instrumented step
        ctx.enable_timing(2)
        ctx.reset_timing()
        F, rr_t, it_t, ok_t = step(prec, F)
        fam = {name: ctx.kernel_time(i) for i, name in enumerate(
            ["element", "assembly", "factor", "solve", "spmv", "factor_update", "factor_panel", "extend_add",
             "fwd_flow", "bwd_flow"])}
        ctx.enable_timing(0)
        results[prec] = {"total": total, "lat": np.array(lat), "u": u, "relres": rr, "iters": it, "fam": fam,
                         "factor_bytes": F.device_bytes(), "converged": conv, "host_rate": host_rate}
        F.free()

    main_prec = 32 if 32 in results else 64
    r = results[main_prec]
    solves_per_s = nrhs * args.steps / r["total"]
    out = {
        "metric": "solves/s + p50/p99 latency @1e-10 residual: FP32-factor+refine vs FP64",
        "value": solves_per_s,
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
        "symbolic_s": t_sym,
        "setup_s": t_setup,
        "symbolic": {k: (float(v) if isinstance(v, float) else int(v)) for k, v in info.items()},
    }
    if 64 in results and 32 in results:
        r64 = results[64]
        out["fp64_solves_per_s"] = nrhs * args.steps / r64["total"]
        out["fp32_vs_fp64_speedup"] = solves_per_s / out["fp64_solves_per_s"]
        out["fp64_p50_latency_ms"] = float(np.percentile(r64["lat"], 50) * 1000)
        out["fp64_p99_latency_ms"] = float(np.percentile(r64["lat"], 99) * 1000)
        out["fp64_max_relative_residual"] = float(np.max(r64["relres"]))
        out["fp64_converged"] = bool(r64["converged"])
        out["fp64_host_buffers_solves_per_s"] = r64["host_rate"]
        out["fp64_triangular_solves_per_rhs"] = [int(x) for x in r64["iters"]]
        u32, u64 = r["u"], r64["u"]
        out["fp32_refined_error_vs_fp64"] = float(np.max(np.linalg.norm(u32 - u64, axis=0) /
                                                         np.linalg.norm(u64, axis=0)))
        out["fp64_kernel_ms"] = {k: round(v[0], 4) for k, v in r64["fam"].items()}
    # kernel-family breakdown; roofline of the dominant kernel of the headline path
    # (largest total time in the instrumented step) among the factor's trailing
    # update (matrix cores) and the two dataflow triangular-solve kernels (HBM:
    # per launch they stream the factor entries one solve needs plus the vectors)
    fam = r["fam"]
    out["kernel_ms"] = {k: round(v[0], 4) for k, v in fam.items()}
    es = 4 if main_prec == 32 else 8
    tname = "float" if main_prec == 32 else "double"
    nr_pad = 4 if nrhs <= 4 else (8 if nrhs <= 8 else 16)
    peak = MI355X_FP32_TFLOPS if main_prec == 32 else MI355X_FP64_TFLOPS
    lines = {}
    n_free = info["n_free"]
    vec = {"fwd_flow": 2 * n_free + 2 * info["upd_rows"], "bwd_flow": 2 * n_free + info["upd_rows"]}
    for fname, kname in (("fwd_flow", f"fwd_flow_kernel<{tname}, {nr_pad}>"), ("bwd_flow", f"bwd_flow_kernel<{tname}, {nr_pad}>")):
        f_ms, f_n = fam[fname]
        if not f_n:
            continue
        per_launch = es * (info["solve_entries"] + nrhs * vec[fname])
        avg_ms = f_ms / f_n
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        lines[fname] = (f_ms, {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": MI355X_HBM_GBS,
                               "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                               "bytes_per_launch": per_launch, "avg_launch_ms": avg_ms, "launches": f_n})
    upd_ms, upd_n = fam["factor_update"]
    if upd_n:
        per_launch_flops = info["update_flops"] / upd_n
        avg_ms = upd_ms / upd_n
        achieved = per_launch_flops / (avg_ms * 1e-3) / 1e12
        lines["factor_update"] = (upd_ms, {"bound": "mfma", "kernel": f"update_kernel<{tname}>", "achieved": achieved,
                                           "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                                           "flops_per_launch": per_launch_flops, "avg_launch_ms": avg_ms,
                                           "launches": upd_n})
    if lines:
        dom = max(lines, key=lambda k: lines[k][0])
        roof = lines[dom][1]
        if args.workload == "wingbox" and args.nx == 256 and args.nw == 64:   # the workload the counters were read on
            kfile = roof["kernel"].split("<")[0]
            roof["traffic"] = read_traffic(os.path.join(ROOT, "profiles", f"{kfile}_wingbox_traffic.txt"), main_prec)
        out["roofline"] = roof
        out["other_rooflines"] = {k: v[1] for k, v in lines.items() if k != dom}
    if args.workload == "wingbox":
        out["load_vector"] = measure_load_vector(ctx, mesh, model)
    if not args.no_cpu_baseline and args.workload == "wingbox":
        cb = run_cpu_baseline(args.cpu_sample_nx)
        cb.pop("seconds", None)
        out["cpu_baseline"] = cb
    print(json.dumps(out))


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


def run_variants(args):
    """64 design variants of the 128 x 128 stiffened plate (per-element thickness
    x U[0.8, 1.2]), 4 RHS each, one symbolic analysis shared by all; per variant:
    section update, element generation + assembly, BC partition, refactor, solve."""
    from jf_2026_05_25_amd import meshes, solver
    base = meshes.stiffened_plate(128, nrhs=4)
    variants = meshes.design_variants(base, 64)
    ctx = solver.Context(0)
    mesh = solver.DeviceMesh(base, ctx)
    pattern = solver.Pattern(mesh)
    K = solver.DeviceMatrix(mesh, pattern, ctx)
    fixed = solver.bc_partition(K, base.spc)
    t0 = time.perf_counter()
    S = solver.Symbolic(K, mesh, fixed, ctx)
    t_sym = time.perf_counter() - t0
    res = {}
    for prec in (32, 64):
        F = None
        lat, rrs, its, conv, us = [], [], [], True, []
        for v in variants[:1]:          # warm-up
            mesh.update_props(v); K.reassemble(mesh)
            F = solver.Factor(S, K, prec, ctx)
            F.solve(K, v.loads, tol=args.tol, max_refine=args.max_refine)
        ctx.sync()
        t_all = time.perf_counter()
        for v in variants:
            ts = time.perf_counter()
            mesh.update_props(v)
            K.reassemble(mesh)
            fx = solver.bc_partition(K, v.spc)
            if not np.array_equal(fx, fixed):
                raise RuntimeError("free set changed between variants")
            F.refactor(K)
            u, rr, it, ok = F.solve(K, v.loads, tol=args.tol, max_refine=args.max_refine)
            lat.append(time.perf_counter() - ts)
            rrs.append(float(np.max(rr))); its.append(int(np.max(it))); conv = conv and ok
            us.append(u)
        total = time.perf_counter() - t_all
        res[prec] = {"total": total, "lat": np.array(lat), "rr": max(rrs), "its": max(its), "conv": conv, "u": us}
        F.free()
    n_solves = 4 * len(variants)
    r32, r64 = res[32], res[64]
    err = max(float(np.max(np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0))) for a, b in zip(r32["u"], r64["u"]))
    return {"metric": "solves/s + p50/p99 latency @1e-10 residual: FP32-factor+refine vs FP64",
            "value": n_solves / r32["total"], "unit": "solves/s", "n_gpus": 1, "steps": len(variants), "warmup": 1,
            "ms_per_step": 1000.0 * r32["total"] / len(variants), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 factor + fp64 refinement",
            "data": "synthetic (generated stiffened plate, 64 thickness variants, random nodal forces)",
            "config": {"workload": "64 design variants of a 128x128 CQUAD4 stiffened plate, 4 RHS each, one symbolic",
                       "dofs": base.ndof, "free_dofs": int(S.info["n_free"]), "cquad4": int(len(base.quads)),
                       "rhs_per_request": 4, "tol": args.tol},
            "p50_latency_ms": float(np.percentile(r32["lat"], 50) * 1000),
            "p99_latency_ms": float(np.percentile(r32["lat"], 99) * 1000),
            "max_relative_residual": r32["rr"], "converged": r32["conv"], "max_triangular_solves": r32["its"],
            "fp64_solves_per_s": n_solves / r64["total"],
            "fp64_p50_latency_ms": float(np.percentile(r64["lat"], 50) * 1000),
            "fp64_p99_latency_ms": float(np.percentile(r64["lat"], 99) * 1000),
            "fp64_max_relative_residual": r64["rr"], "fp64_converged": r64["conv"],
            "fp32_refined_error_vs_fp64": err, "symbolic_s": t_sym,
            "symbolic": {k: (float(v) if isinstance(v, float) else int(v)) for k, v in S.info.items()}}


def run_serving(args):
    """Serving loop on a Poisson stream of shell requests: continuous batching over
    ``slots`` HIP streams, admission by free HBM, device-resident caches.  The
    arrival rate is set to ``utilization`` x the single-stream service rate measured
    on a calibration sample of the same distribution."""
    from jf_2026_05_25_amd import serving
    srv = serving.SolveServer(slots=args.slots, precision=32, tol=args.tol)
    try:
        calib = serving.poisson_stream(16, 1.0, seed=7, precision=32)
        t0 = time.perf_counter()
        for q in calib:
            srv.process(srv.ctxs[0], q, t0)
        srv.factors.clear()
        service = (time.perf_counter() - t0) / len(calib)
        rate = args.utilization / service
        reqs = serving.poisson_stream(args.requests, rate, precision=32)
        t0 = time.perf_counter()
        res = srv.serve(reqs, realtime=True)
        wall = time.perf_counter() - t0
        st = serving.summarize(res)
    finally:
        srv.close()
    ok = [r for r in res if r.ok]
    return {"metric": "solves/s + p50/p99 latency @1e-10 residual: FP32-factor+refine vs FP64",
            "value": st["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": len(reqs), "warmup": len(calib),
            "ms_per_step": 1000.0 * wall / max(1, len(reqs)), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 factor + fp64 refinement (fp64 factor on fp32 breakdown)",
            "data": "synthetic (Poisson stream of generated flat-plate / box shells, random loads)",
            "config": {"workload": f"serving: {len(reqs)} requests, Poisson arrivals at {rate:.1f}/s "
                                   f"({args.utilization:.0%} of the calibrated single-stream service rate), "
                                   f"2k-200k DOFs log-uniform, 10% CTRIA3, 1-16 RHS, {args.slots} slots",
                       "tol": args.tol},
            "p50_latency_ms": st["p50_latency_ms"], "p99_latency_ms": st["p99_latency_ms"],
            "requests_ok": st["ok"], "requests": st["requests"], "requests_per_s": st["requests_per_s"],
            "dof_solves_per_s": st["dof_solves_per_s"], "max_relative_residual": st["max_relative_residual"],
            "fp64_fallbacks": sum(1 for r in ok if r.timings.get("factor_precision") == 64),
            "factor_cache_hits": sum(1 for r in ok if r.cache_hit), "makespan_s": st["makespan_s"],
            "calibrated_service_ms": 1000.0 * service}


if __name__ == "__main__":
    main()
