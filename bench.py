#!/usr/bin/env python
"""Benchmark of the GSIS kinetic loop on one MI355X.

One "step" is one outer iteration of Hybrid_Time_advance_LU_SGS over the
whole mesh and velocity grid (kinetic BCs, Shakhov RHS, LU-SGS sweep, moments,
DETFP, Convert_NS).  Workload: the plain GSIS lid-driven cavity of the
baseline (128 x 128 uniform cells, 64 x 64 non-uniform velocity grid dense
near zero on [-6,6]^2, Kn = 0.1, lid 0.15, diffuse walls at T = 1, FP64),
started from the gas at rest; synthetic data (no files).

Prints one JSON line: throughput in cell-velocity updates/s (all inputs
resident in HBM), a roofline entry for the dominant kernel timed with HIP
events on the library's stream, and a CPU baseline (the CPU oracle on a
bounded sample of the same workload, on this host's cores).
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


def cpu_baseline(case, steps):
    """The oracle (CPU restatement, OpenMP over velocities) on `steps` steps of
    the same case; returns cell-velocity updates/s."""
    orc = _load("gsis_oracle_py", ROOT / "oracle" / "oracle.py")
    o = orc.OracleSolver(case)
    o.initialize()
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    o.close()
    return case.ncell * case.nv * steps / (t1 - t0), t1 - t0


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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="cells per side")
    ap.add_argument("--nv", type=int, default=64, help="velocities per direction")
    ap.add_argument("--kn", type=float, default=0.1)
    ap.add_argument("--lid", type=float, default=0.15)
    ap.add_argument("--dt", type=float, default=1.0)
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps of the CPU-baseline sample (0: skip)")
    ap.add_argument("--kstep-show", type=int, default=50,
                    help="residual cadence (the reference's Kstep_show; the kinetic RHS is kept every N steps)")
    ap.add_argument("--kernel", choices=("auto", "split", "tiles", "strip16"), default="auto",
                    help="kinetic kernels: one-pass column-wavefront kernel (auto: 8 columns per workgroup, "
                         "strip16: 16), separate RHS and sweep kernels, or the one-pass velocity-tile walker")
    ap.add_argument("--scheme", choices=("nnd", "weno3", "up1"), default="nnd",
                    help="kinetic flux scheme (the reference's Iflag_scheme 1, 3, 0)")
    ap.add_argument("--tiles", type=int, default=0, help="velocity tiles per workgroup of the tile walker (0: auto)")
    ap.add_argument("--converge", type=int, default=1,
                    help="1: also measure time-to-converge on the same cavity (first-order upwind), 0: skip")
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
    ksel = {"auto": pkg.capi.KERNEL_AUTO, "split": pkg.capi.KERNEL_SPLIT, "tiles": pkg.capi.KERNEL_TILES,
            "strip16": pkg.capi.KERNEL_STRIP16}[args.kernel]
    s = solver.GsisSolver(case, precision=prec, kstep_show=args.kstep_show, split_kernels=ksel,
                          walk_tiles=args.tiles)
    s.initialize()
    print(f"# case {case.name}: {case.ncell} cells x {case.nv} velocities, precision {args.precision}, "
          f"scheme {args.scheme}", flush=True)
    s.step(args.warmup)
    s.counters()  # synchronise the library's stream
    s.timing(1)
    t0 = time.perf_counter()
    s.step(args.steps)
    s.counters()
    t1 = time.perf_counter()
    ms, nt = s.timing(0)
    res = s.residual()
    fp32_err = None
    if args.precision == "fp32":
        # the same steps on the FP64 path (the one checked against the oracle):
        # relative L2 error of rho, u, T, sigma_xy and q of the FP32 sweeps
        ref = solver.GsisSolver(case, precision=pkg.capi.PREC_FP64, kstep_show=args.kstep_show)
        ref.initialize()
        ref.step(args.warmup + args.steps)
        fp32_err = pkg.fields.field_errors(s.macro()["s"], ref.macro()["s"], case.nx, case.ny)
        ref.close()
    s.close()
    el = t1 - t0
    units = case.ncell * case.nv
    value = units * args.steps / el
    esz = 4 if args.precision == "fp32" else 8
    # algorithmic HBM bytes per launch of each kinetic kernel, per interior
    # cell-velocity: the one-pass kernels (k_strip, k_walk) read and write G
    # and H once (4*esz); the split path's RHS kernel reads F and writes the
    # RHS (4*esz) and its sweep reads RHS + F and writes F (6*esz); moments
    # read F over the padded ring.  The whole step's algorithmic figure is 4*esz.
    fused = args.kernel != "split"
    one = "walk" if args.kernel == "tiles" else "strip"
    names = [one, "moments"] if fused else ["rhs", "sweep", "moments"]
    tot = [ms[1], ms[2]] if fused else list(ms)
    per_unit = {"strip": 4 * esz, "walk": 4 * esz, "rhs": 4 * esz, "sweep": 6 * esz, "moments": 2 * esz}
    avg_ms = [m / max(nt, 1) for m in tot]
    k = max(range(len(names)), key=lambda q: avg_ms[q])
    ring_units = (case.nx + 1) * (case.ny + 1) * case.nv
    bytes_launch = per_unit[names[k]] * (ring_units if names[k] == "moments" else units)
    achieved = bytes_launch / (avg_ms[k] * 1e-3) / 1e9
    step_alg = 4 * esz * units / (el / args.steps) / 1e9
    # HBM bytes per launch of the dominant kernel measured by rocprofv3
    # (FETCH_SIZE / WRITE_SIZE, scaled by the copy-kernel calibration) for this
    # workload, when a profile of it is committed
    traffic, traffic_src = None, None
    tfile = ROOT / "profiles" / f"hbm_traffic_cavity{args.n}_v{args.nv}.json"
    if tfile.exists() and fused and args.scheme == "nnd":
        rec = json.loads(tfile.read_text()).get(f"cavity{args.n}_v{args.nv}_kn{args.kn}", {})
        kern = rec.get(args.precision, {}).get("k_" + names[k])
        if kern:
            traffic, traffic_src = kern["traffic_bytes"], f"{tfile.relative_to(ROOT)} (rocprofv3 FETCH_SIZE+WRITE_SIZE)"
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
        "config": {"workload": f"plain GSIS lid-driven cavity {args.n}x{args.n} cells, {args.nv}x{args.nv} sinh "
                               f"velocity grid on [-6,6]^2, Kn={args.kn}, lid {args.lid}, dt={args.dt}, "
                               f"{args.precision}, {args.scheme.upper()} fluxes, from rest",
                   "cells": case.ncell, "velocities": case.nv},
        "roofline": {"bound": "hbm", "kernel": names[k], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": avg_ms[k],
                     "kernel_ms": dict(zip(names, avg_ms)),
                     "step_algorithmic_GBs": step_alg, "step_frac": step_alg / HBM_PEAK_GBS},
        "residual": list(map(float, res)),
    }
    if fp32_err is not None:
        out["fp32_rel_l2_vs_fp64"] = fp32_err
    if args.converge and args.precision == "fp64" and args.kernel == "auto" and args.scheme == "nnd":
        out["time_to_converge"] = time_to_converge(pkg, solver, args)
    if args.cpu_steps > 0:
        ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v, secs = cpu_baseline(case, args.cpu_steps)
        out["cpu_baseline"] = {"value": v, "unit": "cell·velocity updates/s", "cores": ncores, "kind": "port",
                               "sample": f"{args.cpu_steps} steps of the same case on the C oracle "
                                         f"({secs:.1f} s, OpenMP {ncores} threads)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
