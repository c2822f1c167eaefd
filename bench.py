"""Per-step particle pipeline throughput on one MI355X.

Workload (default): dam-break block 100x100x100 = 1,000,000 fluid particles in
the corner of a 200x100x200-cell open tank with 3 wall layers, l0 = 5 mm,
re = 3.1 l0, jitter +-0.05 l0 (SplitMix64 seed 0), dt = 1e-4.  One step =
neighbor CSR + surface labels + moment operators + provision + PPE assembly +
PPE solve + pressure correction.  Every timed step starts from the same initial
state (restored on the device inside the timed loop).  Inputs are resident in
HBM before timing starts.

Prints one JSON line.  `value` is particle-steps/s of the mixed-precision step
(FP32 SpMV values, FP64 vectors/dots/residual, FP64 residual replacements)
solved to the 1e-6 relative true residual at which the reference's own AMGX
solve stops (ppe_matrix.cu:607); its pressure, grad p and velocity errors
against the FP64 oracle step (PPE solved to 1e-10) are measured in the same run
and `error_target.met` says whether pressure and velocity stay within the 1e-5
relative-error bar.  The plain FP64 GPU step (PPE to 1e-10) is timed
beside it (`speedup_vs_fp64_gpu`), and both modes are also timed at one common
tolerance (`same_tolerance`).  Each step is timed on its own (host clock around
a synchronised step); `ms_per_step` is the median of the timed steps.
`roofline` is the PPE SpMV (the dominant kernel), measured with HIP events on
the library stream during the timed steps.  `cpu_baseline` is the FP64
C++/OpenMP oracle on the host: the whole step (every stage, the PPE solved to a
1e-10 relative true residual with the same preconditioner) on the same
workload, the median of --cpu-repeats runs.  The last oracle step is also the
reference for the per-field errors of the GPU step (`errors_vs_cpu_oracle`).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# PPE tolerances (relative true residual): the plain FP64 step solves to 1e-10;
# the mixed step to 1e-6, where the reference's own AMGX solve stops
# (ppe_matrix.cu:607), and its pressure and velocity must stay within
# ERROR_TARGET (relative L2) of the FP64 oracle step.
FP64_TOLERANCE = 1.0e-10
MIXED_TOLERANCE = 1.0e-6
ERROR_TARGET = 1.0e-5


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--case", default="M1",
                   help="M1 (1M dam-break), C2 (250k dam-break), C1 (8k hydrostatic), "
                        "C3 (64 batched tanks), C5 (4.1M splash)")
    p.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--shadow-dimension", type=int, default=0,
                   help="IDR(s) shadow dimension (0: the library default)")
    p.add_argument("--coarse-cells", type=int, default=0,
                   help="cells per side of the coarse-space aggregates (0: the library "
                        "default, -1: no coarse space)")
    p.add_argument("--no-graph", action="store_true",
                   help="skip the extra run that replays the captured IDR cycles "
                        "(LSMPS_FLAG_GRAPH); the timed steps always launch directly, "
                        "which is what the HIP-event SpMV timing needs")
    p.add_argument("--no-adjoint", action="store_true",
                   help="skip timing the adjoint of one step (after the timed steps)")
    p.add_argument("--no-autotune", action="store_true",
                   help="fixed lanes per particle instead of the autotuner's choice "
                        "(LSMPS_FLAG_AUTOTUNE, tuned in the warm-up steps)")
    p.add_argument("--schwarz-fp32", action="store_true",
                   help="mixed mode: apply the Schwarz block inverses from FP32 instead of "
                        "their scaled FP16 copies (LSMPS_FLAG_SCHWARZ_FP32)")
    p.add_argument("--tolerance", type=float, default=0.0,
                   help="PPE relative true residual of the timed step (0: 1e-6 for mixed, "
                        "1e-10 for fp64)")
    p.add_argument("--no-same-tolerance", action="store_true",
                   help="skip the runs that time both modes at one common tolerance")
    p.add_argument("--inner-tolerance", type=float, default=0.0,
                   help="mixed mode: residual drop between reliable updates (0: the default)")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--cpu-repeats", type=int, default=5,
                   help="full oracle steps timed on the host (the median is reported)")
    p.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                   help="only run a plain device-to-device copy of GIB GiB (5 launches) and "
                        "print its rate: the reference kernel for scaling HBM byte counters")
    return p.parse_args()


def make_case(cases, name):
    if name == "M1":
        return cases.metric_case()
    if name == "C2":
        return cases.dam_break()
    if name == "C1":
        return cases.hydrostatic()
    if name == "C3":
        return cases.batched_tanks()
    if name == "C5":
        return cases.splash()
    raise SystemExit(f"unknown case {name}")


def run_gpu(pkg, rt, case, precision, steps, warmup, timing=True, adjoint=False, shadow=0,
            coarse=0, graph=True, inner_tolerance=0.0, autotune=True, schwarz_fp32=False,
            tolerance=0.0):
    import numpy as np
    import torch
    abi = pkg.abi
    flags = ((abi.FLAG_KERNEL_TIMING if timing else 0) | (abi.FLAG_GRAPH if graph else 0) |
             (abi.FLAG_AUTOTUNE if autotune else 0) |
             (abi.FLAG_SCHWARZ_FP32 if schwarz_fp32 else 0))
    ctx = rt.Context(precision, flags=flags)
    state = rt.DeviceState(case)
    opts = abi.default_solver_options()
    if shadow > 0:
        opts.shadow_dimension = shadow
    if coarse > 0:
        opts.coarse_cells = coarse
    elif coarse < 0:
        opts.preconditioner = abi.PRECOND_SCHWARZ
    if inner_tolerance > 0:
        opts.inner_tolerance = inner_tolerance
    if tolerance > 0:
        opts.tolerance = tolerance
    st = rt.Stepper(ctx, state, case.config, 256, opts)
    if "fluid_offsets" in case.extra:  # batched tanks
        st.set_segments(case.extra["fluid_offsets"], case.extra["wall_offsets"],
                        case.extra["grids"])
    snap = state.snapshot()
    reports = []
    for _ in range(warmup):
        state.restore(snap)
        st.step(raise_on_nonconvergence=False)
    rt.check(rt.load().lsmps_context_kernel_stats(ctx.handle, None, None, None, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stage = np.zeros(7)
    step_ms = []
    for _ in range(steps):
        ts = time.perf_counter()
        state.restore(snap)
        reports.append(st.step(raise_on_nonconvergence=False))
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        stage += np.array(st.stage_times())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # outputs of the last timed step (every step starts from the same state)
    grad, _ = st.correction_vectors()
    fields = {"pressure": state.fluid["pressure"].cpu().numpy(), "grad": grad,
              "velocity": [state.fluid[k].cpu().numpy() for k in ("vx", "vy", "vz")]}
    import ctypes as C
    launches, ms, nbytes = C.c_int64(), C.c_double(), C.c_double()
    rt.check(rt.load().lsmps_context_kernel_stats(ctx.handle, C.byref(launches), C.byref(ms),
                                                  C.byref(nbytes), 0))
    pressure = state.fluid["pressure"].cpu().numpy()
    lanes, tuned_keys = ctx.lanes()
    adj = None
    if adjoint:
        # one recorded step, then the reverse mode of it (outside the timed region)
        st.set_mode(abi.STEP_RECORD)
        state.restore(snap)
        st.step(raise_on_nonconvergence=False)
        f = state.fluid
        vbar = 2.0 * torch.stack([f["vx"], f["vy"], f["vz"]])
        st.adjoint(vbar, raise_on_nonconvergence=False)  # allocates the adjoint workspace
        torch.cuda.synchronize()
        ta = time.perf_counter()
        res = st.adjoint(vbar, raise_on_nonconvergence=False)
        torch.cuda.synchronize()
        adj = {"ms": (time.perf_counter() - ta) * 1e3,
               "transpose_iterations": res["report"]["iterations"],
               "transpose_relative_residual": res["report"]["relative_residual"]}
    out = {
        "ms_per_step": float(np.median(step_ms)),
        "ms_per_step_mean": (t1 - t0) * 1e3 / steps,
        "step_ms": step_ms,
        "fields": fields,
        "stage_ms": (stage / steps).tolist(),
        "reports": reports,
        "spmv_launches": launches.value,
        "spmv_ms": ms.value,
        "spmv_bytes": nbytes.value,
        "pressure": pressure,
        "bytes": st.bytes(),
        "adjoint": adj,
        "lanes": lanes,
        "tuned_keys": tuned_keys,
    }
    st.close()
    ctx.close()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pkg, case, threads, repeats, tolerance):
    """FP64 oracle on the host: the whole step with the PPE solved to
    `tolerance`, `repeats` times from the same state (median reported); returns
    the timing and, as the error reference, the last step's outputs with the PPE
    solved to FP64_TOLERANCE (one more untimed solve when the two differ)."""
    import numpy as np
    from oracle.oracle import Oracle
    abi = pkg.abi
    orc = Oracle(abi, threads=threads or None)
    runs = []
    for _ in range(max(1, repeats)):
        t = {}
        t0 = time.perf_counter()
        nb = orc.neighbors(case)
        cls = orc.classify(case, nb, diagnostics=False)
        t["geometry"] = time.perf_counter() - t0
        c2 = case.copy()
        c2.fluid["surface_type"] = cls["surface_type"]
        t0 = time.perf_counter()
        vel = orc.moments(c2, nb, abi.MOMENT_VELOCITY_WALL_DIRICHLET_A)
        pre = orc.moments(c2, nb, abi.MOMENT_PRESSURE_WALL_NEUMANN_A)
        t["moments"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        tmp = orc.provision(c2, nb, vel["inverse"])
        t["provision"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ppe = orc.assemble(c2, nb, tmp, vel["inverse"], pre["inverse"],
                           neighbor_capacity=case.n_fluid * 256)
        t["assembly"] = time.perf_counter() - t0
        opt = abi.default_solver_options()
        opt.clamp_negative = 0
        opt.tolerance = tolerance
        t0 = time.perf_counter()
        coarse = orc.coarse_aggregation(case, nb, cls["surface_type"], opt.coarse_cells)
        p, rep = orc.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], ppe["rhs"],
                           opt, coarse=coarse)
        t["solve"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        cor = orc.correction(c2, nb, tmp, np.maximum(p, 0.0), pre["inverse"])
        t["correction"] = time.perf_counter() - t0
        runs.append((sum(t.values()), t, rep))
    totals = [r[0] for r in runs]
    med = float(np.median(totals))
    _, t, rep = runs[-1]
    ref_rep = rep
    if tolerance != FP64_TOLERANCE:
        opt.tolerance = FP64_TOLERANCE
        p, ref_rep = orc.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], ppe["rhs"],
                               opt, coarse=coarse)
        cor = orc.correction(c2, nb, tmp, np.maximum(p, 0.0), pre["inverse"])
    return {
        "value": case.n_fluid / med,
        "unit": "particle-steps/s",
        "cores": orc.threads,
        "kind": "port",
        "sample": (f"FP64 oracle, the whole step on the same {case.n_fluid}-particle workload "
                   f"(PPE solved to the {tolerance:g} relative true residual with the same IDR(s) and "
                   f"preconditioner: {rep['iterations']} iterations), median of {len(runs)} "
                   f"steps ({', '.join(f'{x:.2f}' for x in totals)} s) on {orc.threads} threads "
                   f"of {cpu_model()}"),
        "cpu_model": cpu_model(),
        "step_seconds": totals,
        "stage_seconds": {k: round(v, 3) for k, v in t.items()},
        "iterations": rep["iterations"],
        "relative_residual": rep["relative_residual"],
        "tolerance": tolerance,
        "reference_tolerance": FP64_TOLERANCE,
        "reference_iterations": ref_rep["iterations"],
        "reference": {"pressure": np.maximum(p, 0.0), "grad": cor["grad"],
                      "velocity": [cor["fluid"][k] for k in ("vx", "vy", "vz")],
                      "u_star": [tmp[k] for k in ("vx", "vy", "vz")]},
    }


def field_errors(got, ref):
    """Max / RMS absolute and relative-L2 errors of pressure, grad p and the
    corrected velocity (velocity relative to the predictor u*, whose magnitude
    the correction nearly cancels)."""
    import numpy as np
    out = {}
    for key, scale_key in (("pressure", None), ("grad", None), ("velocity", "u_star")):
        g = np.concatenate([np.ravel(x) for x in (got[key] if key != "pressure" else [got[key]])])
        r = np.concatenate([np.ravel(x) for x in (ref[key] if key != "pressure" else [ref[key]])])
        scale = np.linalg.norm(np.concatenate([np.ravel(x) for x in ref[scale_key]])) \
            if scale_key else np.linalg.norm(r)
        d = g - r
        out[key] = {"max_abs": float(np.max(np.abs(d))), "rms": float(np.sqrt(np.mean(d * d))),
                    "rel_l2": float(np.linalg.norm(d) / max(scale, 1e-300))}
        if scale_key:
            out[key]["rel_l2_of_" + key] = float(np.linalg.norm(d) /
                                                 max(np.linalg.norm(r), 1e-300))
    return out


def copy_calibration(gib):
    """A plain copy of known size: it reads and writes exactly `nbytes` each
    launch, so FETCH_SIZE / WRITE_SIZE of this kernel give the counters' scale."""
    import torch
    nbytes = int(gib * (1 << 30)) // 8 * 8
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(5):
        dst.copy_(src)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / 5
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "avg_ms": ms,
                      "GBps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)


def measured_traffic(case, precision):
    """HBM bytes per SpMV launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (scaled by a plain-copy calibration), with the
    profile it comes from and the commit it was measured at; (None, None) when
    no profile covers the workload.  The counters cannot be read during the
    timed run itself, so this is a recorded measurement, not a live one."""
    path = os.path.join(ROOT, "profiles", "spmv_dia_sell_m1_mixed_traffic.json")
    if case != "M1" or precision != "mixed" or not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    fetch = d["fetch_kib"] * 1024 * d["copy_bytes"] / (d["copy_fetch_kib"] * 1024)
    write = d["write_kib"] * 1024 * d["copy_bytes"] / (d["copy_write_kib"] * 1024)
    source = {"profile": "profiles/" + os.path.basename(path), "kernel": d["kernel"],
              "measured_at_commit": d.get("measured_at_commit")}
    return fetch + write, source


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(requested --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    import numpy as np
    from __graft_entry__ import load_package
    pkg = load_package()
    from importlib import import_module
    rt = import_module("lsmps3d_amd.runtime")
    abi = pkg.abi
    case = make_case(pkg.cases, args.case)
    prec = abi.PRECISION_MIXED if args.precision == "mixed" else abi.PRECISION_FP64
    print(f"[bench] case {args.case}: {case.n_fluid} fluid, {case.n_wall} wall particles; "
          f"grid {list(case.config.cell_dims)}", file=sys.stderr, flush=True)
    mixed = prec == abi.PRECISION_MIXED
    tol = args.tolerance if args.tolerance > 0 else (MIXED_TOLERANCE if mixed else FP64_TOLERANCE)
    main_run = run_gpu(pkg, rt, case, prec, args.steps, args.warmup,
                       adjoint=not args.no_adjoint, shadow=args.shadow_dimension,
                       coarse=args.coarse_cells, graph=False,
                       inner_tolerance=args.inner_tolerance, autotune=not args.no_autotune,
                       schwarz_fp32=args.schwarz_fp32, tolerance=tol)
    iters = [r["iterations"] for r in main_run["reports"]]
    print(f"[bench] {args.precision}: {main_run['ms_per_step']:.1f} ms/step, iterations {iters}, "
          f"stages {np.round(main_run['stage_ms'], 2).tolist()}", file=sys.stderr, flush=True)
    fp64_run = None
    if not args.no_fp64 and prec != abi.PRECISION_FP64:
        fp64_run = run_gpu(pkg, rt, case, abi.PRECISION_FP64, 5, 1,
                           shadow=args.shadow_dimension, coarse=args.coarse_cells, graph=False,
                           autotune=not args.no_autotune, tolerance=FP64_TOLERANCE)
        print(f"[bench] fp64: {fp64_run['ms_per_step']:.1f} ms/step, iterations "
              f"{[r['iterations'] for r in fp64_run['reports']]}", file=sys.stderr, flush=True)
    traffic, traffic_source = measured_traffic(args.case, args.precision)
    avg_ms = main_run["spmv_ms"] / max(main_run["spmv_launches"], 1)
    avg_bytes = main_run["spmv_bytes"] / max(main_run["spmv_launches"], 1)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    value = case.n_fluid / (main_run["ms_per_step"] * 1e-3)
    line = {
        "metric": "particle-steps/s at 1M particles, FP32+refine vs FP64 at <=1e-5 rel err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "ms_per_step_mean": main_run["ms_per_step_mean"],
        "step_ms": [round(x, 3) for x in main_run["step_ms"]],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == abi.PRECISION_MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": {"M1": "dam-break 100x100x100 (1,000,000 fluid) in 200x100x200 tank",
                                "C2": "dam-break 100x50x50 (250,000 fluid) in 200x50x100 tank",
                                "C1": "hydrostatic 20x20x20 (8,000 fluid)",
                                "C3": "64 batched tanks, 2k-60k fluid each, one launch per stage",
                                "C5": "splash 160^3 (4,096,000 fluid), top 8 layers lifted"}[args.case],
                   "n_fluid": case.n_fluid, "n_wall": case.n_wall, "l0_m": case.config.particle_spacing,
                   "re_over_l0": 3.1, "dt_s": case.config.time_step, "solver": f"IDR({args.shadow_dimension or (16 if args.coarse_cells < 0 else 14)}); Schwarz blocks on wrong-sign rows, Jacobi elsewhere" + ("" if args.coarse_cells < 0 else f", coarse space of {args.coarse_cells or pkg.abi.default_solver_options().coarse_cells}^3-cell aggregates") + ("" if args.precision != "mixed" else ("; Schwarz inverses FP32" + ("" if args.schwarz_fp32 else ", applied from scaled FP16 copies"))),
                   "tolerance": tol, "precision": args.precision},
        "solver_iterations": iters,
        "lanes_per_particle": main_run["lanes"],
        "autotuned_keys": main_run["tuned_keys"] if not args.no_autotune else 0,
        "schwarz_blocks": main_run["reports"][-1]["schwarz_blocks"],
        "schwarz_rows": main_run["reports"][-1]["schwarz_rows"],
        "coarse_unknowns": main_run["reports"][-1]["coarse_unknowns"],
        "relative_residual": [r["relative_residual"] for r in main_run["reports"]],
        "stage_ms": dict(zip(["geometry", "moments", "provision", "assembly", "solve",
                              "correction", "total"], np.round(main_run["stage_ms"], 3).tolist())),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": "PPE SpMV (IDR shadow-dot epilogue)",
                     "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": avg_bytes,
                     "launches": main_run["spmv_launches"]},
    }
    if main_run["adjoint"] is not None:
        line["adjoint"] = main_run["adjoint"]
        print(f"[bench] adjoint of one step: {main_run['adjoint']}", file=sys.stderr, flush=True)
    if fp64_run is not None:
        line["fp64_ms_per_step"] = fp64_run["ms_per_step"]
        line["fp64_iterations"] = [r["iterations"] for r in fp64_run["reports"]]
        line["speedup_vs_fp64_gpu"] = fp64_run["ms_per_step"] / main_run["ms_per_step"]
        line["fp64_tolerance"] = FP64_TOLERANCE
        p64 = fp64_run["pressure"]
        line["pressure_rel_l2_vs_fp64_gpu"] = float(
            np.linalg.norm(main_run["pressure"] - p64) / max(np.linalg.norm(p64), 1e-300))
        if not args.no_same_tolerance:
            # both modes at one tolerance: the FP64 step at the mixed step's, and
            # the mixed step at the FP64 step's
            kw = dict(shadow=args.shadow_dimension, coarse=args.coarse_cells, graph=False,
                      autotune=not args.no_autotune)
            f_at = run_gpu(pkg, rt, case, abi.PRECISION_FP64, 3, 1, tolerance=tol, **kw) \
                if tol != FP64_TOLERANCE else fp64_run
            m_at = run_gpu(pkg, rt, case, prec, 3, 1, inner_tolerance=args.inner_tolerance,
                           schwarz_fp32=args.schwarz_fp32, tolerance=FP64_TOLERANCE, **kw) \
                if tol != FP64_TOLERANCE else main_run
            line["same_tolerance"] = [
                {"tolerance": t, "mixed_ms_per_step": m["ms_per_step"],
                 "fp64_ms_per_step": f["ms_per_step"],
                 "mixed_iterations": m["reports"][-1]["iterations"],
                 "fp64_iterations": f["reports"][-1]["iterations"],
                 "speedup": f["ms_per_step"] / m["ms_per_step"]}
                for t, m, f in ((tol, main_run, f_at), (FP64_TOLERANCE, m_at, fp64_run))]
            print(f"[bench] same tolerance: {line['same_tolerance']}", file=sys.stderr,
                  flush=True)
    if not args.no_graph:
        graph_run = run_gpu(pkg, rt, case, prec, 5, 1, timing=False,
                            shadow=args.shadow_dimension, coarse=args.coarse_cells, graph=True,
                            inner_tolerance=args.inner_tolerance,
                            autotune=not args.no_autotune, schwarz_fp32=args.schwarz_fp32,
                            tolerance=tol)
        line["graph"] = {
            "ms_per_step": graph_run["ms_per_step"],
            "graph_cycles": [r["graph_cycles"] for r in graph_run["reports"]],
            "pressure_bit_identical": bool(np.array_equal(graph_run["pressure"],
                                                          main_run["pressure"])),
            "note": "the same step with the IDR cycles replayed from captured HIP graphs "
                    "(device-side loop control, two cycles in flight), no kernel timing"}
        print(f"[bench] graph replays: {graph_run['ms_per_step']:.1f} ms/step", file=sys.stderr,
              flush=True)
    if not args.no_cpu:
        cb = cpu_baseline(pkg, case, args.cpu_threads, args.cpu_repeats, tol)
        print(f"[bench] cpu oracle: {cb['sample']}", file=sys.stderr, flush=True)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        line["cpu_baseline"]["cpu_model"] = cb["cpu_model"]
        line["cpu_baseline"]["step_seconds"] = [round(x, 3) for x in cb["step_seconds"]]
        line["cpu_baseline"]["iterations"] = cb["iterations"]
        line["cpu_baseline"]["tolerance"] = cb["tolerance"]
        line["speedup_vs_cpu"] = value / cb["value"]
        errs = field_errors(main_run["fields"], cb["reference"])
        line["errors_vs_cpu_oracle"] = {"precision": args.precision, "tolerance": tol,
                                        "oracle_tolerance": FP64_TOLERANCE, **errs}
        print(f"[bench] errors vs oracle: {errs}", file=sys.stderr, flush=True)
    if "errors_vs_cpu_oracle" in line:
        e = line["errors_vs_cpu_oracle"]
        p_err, v_err = e["pressure"]["rel_l2"], e["velocity"]["rel_l2"]
        line["error_target"] = {"pressure_rel_l2": p_err, "velocity_rel_l2_of_u_star": v_err,
                                "velocity_rel_l2_of_velocity": e["velocity"]["rel_l2_of_velocity"],
                                "against": "cpu oracle step (PPE to 1e-10)",
                                "bound": ERROR_TARGET,
                                "met": bool(p_err <= ERROR_TARGET and v_err <= ERROR_TARGET)}
    elif "pressure_rel_l2_vs_fp64_gpu" in line:
        p_err = line["pressure_rel_l2_vs_fp64_gpu"]
        line["error_target"] = {"pressure_rel_l2": p_err, "against": "fp64 gpu step (PPE to 1e-10)",
                                "bound": ERROR_TARGET, "met": bool(p_err <= ERROR_TARGET)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
