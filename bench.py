#!/usr/bin/env python
"""Benchmark of the GSIS loop on one MI355X.

One "step" is one outer iteration of Hybrid_Time_advance_LU_SGS over the
whole mesh and velocity grid: kinetic BCs, Shakhov RHS, LU-SGS sweep, moments,
DETFP, Convert_NS and, in plain GSIS (mode M1, the default), the synthetic NS
solve iterated to its tolerance.  Default workload: the production cavity
(256 x 256 uniform cells, 96 x 96 non-uniform velocity grid dense near zero
on [-6,6]^2, Kn = 0.05, lid 0.2, diffuse walls at T = 1, FP64), started from
the gas at rest; synthetic data (no files).

Prints one JSON line: throughput in cell-velocity updates/s (all inputs
resident in HBM), a roofline entry for the dominant kernel timed with HIP
events on the library's stream (and for each other kernel of the step), and
a CPU baseline (the CPU oracle on a bounded sample of the same workload, on
this host's cores).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "hybrid2d-2--sis---shakov---coette--9-6_amd"
HBM_PEAK_GBS = 8000.0


def _load(name, path, pkg_dir=None):
    if name in sys.modules:
        return sys.modules[name]
    kw = {"submodule_search_locations": [str(pkg_dir)]} if pkg_dir else {}
    spec = importlib.util.spec_from_file_location(name, path, **kw)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def time_to_converge(pkg, solver, args, tol=1e-8, field_tol=1e-4, chunk=50, max_steps=6000):
    """The second half of the metric on the same cavity: time for the plain
    FP64 GPU path to bring the residual below ``tol``, and time for the FP64
    and the FP32-sweep paths to bring rho, u, T, sigma_xy and q within
    ``field_tol`` relative L2 of the converged FP64 fields.  First-order
    upwind (the reference's Iflag_Scheme=0): with the NND limiter the residual
    of these cases stalls at 1e-2..1 (limiter cycling, the reference's scheme
    as written), so no run reaches 1e-8.  Only the GPU stepping is timed (the
    residual and field checks between chunks are not).  Also the FP64 time
    to the same residual of the planar Couette configuration (4 x 128 cells,
    32^2 Newton-Cotes velocities, Kn = 0.1, walls at -+0.1)."""
    case = make_case(pkg, args.n, args.nv, args.kn, args.lid, args.dt)
    case.const["scheme"] = pkg.capi.SCHEME_UP1

    def run(prec, stop, c=case):
        s = solver.GsisSolver(c, precision=prec, kstep_show=chunk)
        s.initialize()
        s.counters()
        k, t = 0, 0.0
        info = None
        while k < max_steps:
            t0 = time.perf_counter()
            s.step(chunk)
            s.counters()
            t += time.perf_counter() - t0
            k += chunk
            done, info = stop(s)
            if done:
                break
        S = s.macro()["s"]
        s.close()
        return k, t, info, S, k < max_steps or bool(done)

    def res_below(s):
        r = max(float(x) for x in s.residual())
        return r < tol, r
    k64, t64, r64, ref, ok64 = run(pkg.capi.PREC_FP64, res_below)

    def fields_close(s):
        e = pkg.fields.field_errors(s.macro()["s"], ref, case.nx, case.ny)
        return max(e.values()) <= field_tol, e
    k64f, t64f, e64f, _, ok64f = run(pkg.capi.PREC_FP64, fields_close)
    k32f, t32f, e32f, _, ok32f = run(pkg.capi.PREC_FP32, fields_close)
    # the planar Couette configuration (4 x 128 cells, 32^2 velocities), same
    # scheme, FP64 to the same residual
    c1 = pkg.cases.couette_case(ny_cells=128, nx_cells=4, nv=32, kn=0.1, uwall=0.1, dt=1.0)
    c1.const["scheme"] = pkg.capi.SCHEME_UP1
    kc1, tc1, rc1, _, okc1 = run(pkg.capi.PREC_FP64, res_below, c1)
    return {
        "scheme": "UP1", "residual_tol": tol, "field_tol_rel_l2": field_tol, "check_every": chunk,
        "fp64_to_residual": {"steps": k64, "s": t64, "residual": r64, "reached": ok64},
        "fp64_to_field_tol": {"steps": k64f, "s": t64f, "errors": e64f, "reached": ok64f},
        "fp32_to_field_tol": {"steps": k32f, "s": t32f, "errors": e32f, "reached": ok32f},
        "fp32_field_tol_speedup_vs_fp64_converged": t64 / t32f if t32f > 0 else None,
        "couette_4x128_v32_fp64_to_residual": {"steps": kc1, "s": tc1, "residual": rc1, "reached": okc1},
    }


def make_case(pkg, n, nv, kn, lid, dt):
    return pkg.cases.cavity_case(n, nv, kn=kn, lid=lid, dt=dt, name=f"cavity{n}_v{nv}_kn{kn}")


def cpu_baseline(case, steps, mode, gpu_its, ns_sample=10):
    """The oracle (CPU restatement, OpenMP over velocities) on a bounded
    sample of the same workload, in cell-velocity updates/s.  M0: `steps`
    whole steps.  M1: one kinetic step (M0) plus `ns_sample` iterations of the
    synthetic NS inner loop, extrapolated to a GSIS step with the GPU's inner
    iteration counts of the timed steps (the same algorithm, so the same
    counts).  Returns (value, seconds spent, description)."""
    orc = _load("gsis_oracle_py", ROOT / "oracle" / "oracle.py")
    o = orc.OracleSolver(case, mode=0)
    o.initialize()
    t0 = time.perf_counter()
    o.step(steps)
    t_kin = (time.perf_counter() - t0) / steps
    spent = t_kin * steps
    if mode == 0:
        o.close()
        return case.ncell * case.nv / t_kin, spent, f"{steps} step(s) of the same case on the C oracle"
    assert o.convert_ns() == 0
    t0 = time.perf_counter()
    o.ns_advance(ns_sample)
    t_it = (time.perf_counter() - t0) / ns_sample
    spent += t_it * ns_sample
    o.close()
    its = sum(gpu_its) / max(len(gpu_its), 1)
    return (case.ncell * case.nv / (t_kin + its * t_it), spent,
            f"C oracle: {steps} kinetic step(s) ({t_kin:.2f} s each) and {ns_sample} synthetic NS iterations "
            f"({1e3 * t_it:.1f} ms each) of the same case, extrapolated to the GPU's {its:.0f} inner iterations "
            f"per step")


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of
    its dispatches give the factor between the counters and HBM bytes."""
    import torch

    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "copies": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="cells per side")
    ap.add_argument("--nv", type=int, default=96, help="velocities per direction")
    ap.add_argument("--kn", type=float, default=0.05)
    ap.add_argument("--lid", type=float, default=0.2)
    ap.add_argument("--dt", type=float, default=1.0)
    ap.add_argument("--mode", choices=("m1", "m0"), default="m1",
                    help="m1: GSIS with the synthetic NS solve in every step; m0: the kinetic step alone")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle steps of the CPU-baseline sample (0: skip)")
    ap.add_argument("--kernel", choices=("auto", "split", "tiles", "strip16"), default="auto",
                    help="kinetic kernels: one-pass column-wavefront kernel (auto: 8 columns per workgroup, "
                         "strip16: 16), separate RHS and sweep kernels, or the one-pass velocity-tile walker")
    ap.add_argument("--scheme", choices=("nnd", "weno3", "up1"), default="nnd",
                    help="kinetic flux scheme (the reference's Iflag_scheme 1, 3, 0)")
    ap.add_argument("--tiles", type=int, default=0, help="velocity tiles per workgroup of the tile walker (0: auto)")
    ap.add_argument("--converge", type=int, default=0,
                    help="1: also measure time-to-converge on a cavity (first-order upwind), 0: skip")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy a device buffer of BYTES bytes 5 times (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding velocity space across GPUs is not built.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    pkg = _load("gsis_amd", PKG_DIR / "__init__.py", PKG_DIR)
    solver = _load("gsis_amd.solver", PKG_DIR / "solver.py")
    case = make_case(pkg, args.n, args.nv, args.kn, args.lid, args.dt)
    case.const["scheme"] = {"nnd": pkg.capi.SCHEME_NND, "weno3": pkg.capi.SCHEME_WENO3,
                            "up1": pkg.capi.SCHEME_UP1}[args.scheme]
    prec = pkg.capi.PREC_FP32 if args.precision == "fp32" else pkg.capi.PREC_FP64
    mode = pkg.capi.MODE_M1 if args.mode == "m1" else pkg.capi.MODE_M0
    ksel = {"auto": pkg.capi.KERNEL_AUTO, "split": pkg.capi.KERNEL_SPLIT, "tiles": pkg.capi.KERNEL_TILES,
            "strip16": pkg.capi.KERNEL_STRIP16}[args.kernel]
    # the kinetic RHS of the last timed step is kept, for the residual field
    s = solver.GsisSolver(case, precision=prec, mode=mode, kstep_show=args.warmup + args.steps,
                          split_kernels=ksel, walk_tiles=args.tiles)
    s.initialize()
    print(f"# case {case.name}: {case.ncell} cells x {case.nv} velocities, precision {args.precision}, "
          f"scheme {args.scheme}, mode {args.mode}", flush=True)
    for k in range(args.warmup):
        s.step(1)
        print(f"# warmup step {k + 1}: inner iterations {s.counters()[2]}", flush=True)
    s.counters()  # synchronise the library's stream
    s.timing(1)
    its = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.step(1)
        its.append(s.counters()[2])  # synchronises: the inner iteration count of this step
    t1 = time.perf_counter()
    ms, nt = s.timing(0)
    res = s.residual()
    fp32_err = None
    if args.precision == "fp32":
        # the same steps on the FP64 path (the one checked against the oracle):
        # relative L2 error of rho, u, T, sigma_xy and q of the FP32 sweeps
        ref = solver.GsisSolver(case, precision=pkg.capi.PREC_FP64, mode=mode)
        ref.initialize()
        ref.step(args.warmup + args.steps)
        fp32_err = pkg.fields.field_errors(s.macro()["s"], ref.macro()["s"], case.nx, case.ny)
        ref.close()
    s.close()
    el = t1 - t0
    units = case.ncell * case.nv
    value = units * args.steps / el
    esz = 4 if args.precision == "fp32" else 8
    # Algorithmic HBM bytes per launch of each kernel.  Kinetic kernels, per
    # interior cell-velocity: the one-pass kernels (k_strip, k_walk) read and
    # write G and H once (4*esz); the split path's RHS kernel reads F and
    # writes the RHS (4*esz), its sweep reads RHS + F and writes F (6*esz);
    # moments read F over the padded ring (2*esz).  The synthetic NS loop, per
    # swept cell and inner iteration: read and write the conserved state U
    # (4 FP64 each way, 64 B); it is bound by its serial LU-SGS recurrence, not
    # by these bytes.
    fused = args.kernel != "split"
    one = "walk" if args.kernel == "tiles" else "strip"
    names = [one, "moments"] if fused else ["rhs", "sweep", "moments"]
    tot = [ms[1], ms[2]] if fused else list(ms[:3])
    if mode == pkg.capi.MODE_M1:
        names.append("ns_loop")
        tot.append(ms[3])
    avg_ms = [m / max(nt, 1) for m in tot]
    ring_units = (case.nx + 1) * (case.ny + 1) * case.nv
    swept = max(case.nx - 5, 0) * max(case.ny - 5, 0)
    per_launch = {"strip": 4 * esz * units, "walk": 4 * esz * units, "rhs": 4 * esz * units,
                  "sweep": 6 * esz * units, "moments": 2 * esz * ring_units,
                  "ns_loop": 64 * swept * (sum(its) / max(len(its), 1))}
    rooflines = {}
    for q, nm in enumerate(names):
        ach = per_launch[nm] / (avg_ms[q] * 1e-3) / 1e9 if avg_ms[q] > 0 else 0.0
        rooflines[nm] = {"bytes_per_launch": per_launch[nm], "avg_launch_ms": avg_ms[q], "achieved": ach,
                         "frac": ach / HBM_PEAK_GBS}
    k = max(range(len(names)), key=lambda q: avg_ms[q])
    kin = 0 if fused else 1
    step_alg = 4 * esz * units / (el / args.steps) / 1e9
    # HBM bytes per launch measured by rocprofv3 (FETCH_SIZE / WRITE_SIZE,
    # scaled by the copy-kernel calibration) for this workload, when committed
    traffic, traffic_src = {}, None
    tfile = ROOT / "profiles" / f"hbm_traffic_cavity{args.n}_v{args.nv}.json"
    if tfile.exists() and fused and args.scheme == "nnd":
        rec = json.loads(tfile.read_text()).get(f"cavity{args.n}_v{args.nv}_kn{args.kn}", {})
        for nm in names:
            kern = rec.get(args.precision, {}).get("k_" + nm)
            if kern:
                traffic[nm] = kern["traffic_bytes"]
                traffic_src = f"{tfile.relative_to(ROOT)} (rocprofv3 FETCH_SIZE+WRITE_SIZE)"
    dom = names[k]
    out = {
        "metric": "cell·velocity updates/s + time-to-converge vs plain GSIS at 1e-4 rel L2",
        "value": value,
        "unit": "cell·velocity updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-sweep/f64-moments",
        "data": "synthetic",
        "config": {"workload": f"{'plain GSIS (synthetic NS solve in every step)' if args.mode == 'm1' else 'kinetic step alone (no synthetic solve)'} "
                               f"lid-driven cavity {args.n}x{args.n} cells, {args.nv}x{args.nv} sinh velocity grid "
                               f"on [-6,6]^2, Kn={args.kn}, lid {args.lid}, dt={args.dt}, {args.precision}, "
                               f"{args.scheme.upper()} fluxes, from rest",
                   "cells": case.ncell, "velocities": case.nv, "mode": args.mode},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": rooflines[dom]["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": rooflines[dom]["frac"], "traffic": traffic.get(dom),
                     "traffic_source": traffic_src if dom in traffic else None,
                     "bytes_per_launch": rooflines[dom]["bytes_per_launch"],
                     "avg_launch_ms": rooflines[dom]["avg_launch_ms"],
                     "kernels": {nm: dict(rooflines[nm], traffic=traffic.get(nm)) for nm in names},
                     "step_algorithmic_GBs": step_alg, "step_frac": step_alg / HBM_PEAK_GBS},
        "residual": list(map(float, res)),
    }
    if mode == pkg.capi.MODE_M1:
        kin_ms = sum(avg_ms[:len(names) - 1])
        out["synthetic"] = {"inner_iterations": its, "ns_ms_per_step": avg_ms[-1],
                            "ns_ms_per_inner_iteration": avg_ms[-1] / max(sum(its) / len(its), 1),
                            "kinetic_ms_per_step": kin_ms,
                            "kinetic_only_updates_per_s": units / (kin_ms * 1e-3) if kin_ms > 0 else None}
        out["roofline"]["note"] = ("the synthetic NS loop dominates the GSIS step; it is bound by the serial "
                                   "anti-diagonal LU-SGS recurrence (one plane after another), not by HBM")
    if fp32_err is not None:
        out["fp32_rel_l2_vs_fp64"] = fp32_err
    if args.converge and args.precision == "fp64" and args.kernel == "auto" and args.scheme == "nnd":
        out["time_to_converge"] = time_to_converge(pkg, solver, args)
    if args.cpu_steps > 0:
        ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v, secs, desc = cpu_baseline(case, args.cpu_steps, 1 if args.mode == "m1" else 0, its)
        out["cpu_baseline"] = {"value": v, "unit": "cell·velocity updates/s", "cores": ncores, "kind": "port",
                               "sample": f"{desc} ({secs:.1f} s, OpenMP {ncores} threads)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
