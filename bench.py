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
        # the same criterion on both sides: fields within field_tol of the converged FP64 fields
        "fp32_speedup_same_criterion": t64f / t32f if t32f > 0 else None,
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


def bench_batch(pkg, solver, args):
    """The batched configuration: `instances` Couette and cavity instances
    drawn from `seed` (mesh side 64-256, Kn 1e-3-10, wall or lid speed
    0.05-0.3, velocity grid 24^2-64^2), FP32 sweeps with FP64 moments, the
    kinetic step of all of them per graph replay.  The autotuner's trials
    (4 candidates x 1 step) come before the warm-up steps."""
    batch = _load("gsis_amd.batch", PKG_DIR / "batch.py")
    t0 = time.perf_counter()
    m1 = args.mode == "m1"
    # GSIS instances: first-order upwind (with NND the cavity runs do not converge), MassCorrect, at most 1000 inner
    # iterations per step; the kinetic step alone: NND
    cases = batch.instances(args.instances, seed=args.seed, precision_scheme=pkg.capi.SCHEME_UP1 if m1 else None)
    solvers = []
    for c in cases:
        s = solver.GsisSolver(c, precision=pkg.capi.PREC_FP32, kstep_show=1 << 30,
                              mode=pkg.capi.MODE_M1 if m1 else pkg.capi.MODE_M0, ns_maxit=1000 if m1 else 0)
        s.initialize()
        if m1:
            s.set_mass_correct(True)
        solvers.append(s)
    b = batch.Batch(solvers)
    units = sum(c.ncell * c.nv for c in cases)
    ring = sum((c.nx + 1) * (c.ny + 1) * c.nv for c in cases)
    print(f"# {len(cases)} instances, {units:.4g} cell-velocities, set up in {time.perf_counter() - t0:.1f} s",
          flush=True)
    nw, order, tune_ms = b.tune(1)
    print(f"# autotuner: {nw} columns per workgroup, order {order}, {tune_ms:.2f} ms per step", flush=True)
    b.step(args.warmup)
    solvers[0].counters()
    t1 = time.perf_counter()
    b.step(args.steps)
    el = time.perf_counter() - t1
    its = [sv.counters()[2] for sv in solvers] if m1 else None
    ms = el / args.steps * 1e3
    alg = (16 * units + 8 * ring) / (ms * 1e-3) / 1e9  # FP32: read+write G,H once; moments read G,H
    ncou = sum("couette" in c.name for c in cases)
    out = {
        "metric": "cell·velocity updates/s (batched instances)",
        "value": units * args.steps / el,
        "unit": "cell·velocity updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-sweep/f64-moments", "data": "synthetic",
        "config": {"workload": f"batch of {len(cases)} instances from seed {args.seed}: {ncou} planar Couette "
                               f"(4 x side cells) and {len(cases) - ncou} lid-driven cavities (side^2 cells), side "
                               "64-256, Kn log-uniform 1e-3-10, wall/lid speed 0.05-0.3, velocity grid 24^2-64^2, "
                               + ("GSIS (mode M1: UP1 fluxes, MassCorrect, at most 1000 inner iterations), the inner "
                                  "loops of all instances in one launch inside the step's graph" if m1 else
                                  "NND fluxes, kinetic step (mode M0)") + ", one hipGraph replay per step",
                   "instances": len(cases), "cell_velocities": units, "mode": args.mode},
        "autotuner": {"columns_per_workgroup": nw, "order": "largest first" if order else "as drawn",
                      "ms_per_step": tune_ms, "synthetic_solves_in_graph": b.ns_graph()},
        "inner_iterations_last_step": {"max": max(its), "mean": sum(its) / len(its)} if its else None,
        "roofline": {"bound": "hbm", "kernel": "batched step", "achieved": alg, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / HBM_PEAK_GBS, "traffic": None,
                     "note": "algorithmic bytes of the batched kinetic launch and the moments over the step time"},
    }
    b.close()
    for s in solvers:
        s.close()
    print(json.dumps(out), flush=True)


def bench_production(pkg, solver, args):
    """Time to 1e-4 relative L2 (rho, u, T, sigma_xy, q) of plain FP64 GSIS
    (synthetic solve to ns_tol every step, the reference's control) and of the
    fast form (FP32 sweeps, at most --fast-maxit inner iterations per step,
    the batched graph-replayed kinetic step, optionally the learned closure),
    both against the last step of the plain FP64 run.  First-order upwind
    fluxes: with NND the cavity's residual stalls.  Only the stepping is
    timed; the field checks between steps are not."""
    batch = _load("gsis_amd.batch", PKG_DIR / "batch.py")
    closure = _load("gsis_amd.closure", PKG_DIR / "closure.py")
    case = make_case(pkg, args.n, args.nv, args.kn, args.lid, args.dt)
    case.const["scheme"] = pkg.capi.SCHEME_UP1
    nx, ny = case.nx, case.ny

    def run(prec, maxit, steps, use_batch, use_closure, ref=None, tol=1e-4, cfl=20.0, dtmax=10.0, mode=None):
        case.const["cfl"] = cfl
        case.const["dtmax"] = dtmax
        mode = pkg.capi.MODE_M1 if mode is None else mode
        s = solver.GsisSolver(case, precision=prec, mode=mode, ns_maxit=maxit, kstep_show=1 << 30)
        s.initialize()
        if args.mass_correct and mode == pkg.capi.MODE_M1:
            s.set_mass_correct(True)
        b = batch.Batch([s]) if use_batch else None
        net = None
        if use_closure:
            net = closure.trained_network()
            if args.closure_warmup > 0:
                closure.set_warmup(s, net, args.closure_warmup, every=50)
            if args.closure:
                closure.set_closure(s, net, gate=1.0, every=max(1, min(maxit, 50)) if maxit else 50)
        snaps, t, its, hit = [], 0.0, [], None
        for k in range(1, steps + 1):
            t0 = time.perf_counter()
            if b is not None:
                b.step(1)
            else:
                s.step(1)
            its.append(s.counters()[2])
            t += time.perf_counter() - t0
            S = s.macro()["s"]
            if ref is None:
                snaps.append((k, t, S.copy()))
            else:
                e = max(pkg.fields.field_errors(S, ref, nx, ny).values())
                if hit is None and e < tol:
                    hit = (k, t, e)
                    break
            if k % (10 if mode == pkg.capi.MODE_M1 else 500) == 0:
                print(f"# {'fp32' if prec else 'fp64'} mode {mode} maxit {maxit}: step {k}, {t:.1f} s, "
                      f"inner its {its[-1]}", flush=True)
        if b is not None:
            b.close()
        s.close()
        return snaps, t, its, hit

    units = case.ncell * case.nv
    snaps, t_ref, its_ref, _ = run(pkg.capi.PREC_FP64, 0, args.ref_steps, False, False)
    ref = snaps[-1][2]
    hit_plain = None
    for k, t, S in snaps:
        e = max(pkg.fields.field_errors(S, ref, nx, ny).values())
        if e < 1e-4:
            hit_plain = (k, t, e)
            break
    fprec = pkg.capi.PREC_FP32 if args.fast_precision == "fp32" else pkg.capi.PREC_FP64
    _, t_fast, its_fast, hit_fast = run(fprec, args.fast_maxit, args.ref_steps, True,
                                        bool(args.closure) or args.closure_warmup > 0, ref=ref, cfl=args.fast_cfl,
                                        dtmax=args.fast_dtmax)
    # the kinetic step alone (no synthetic solve) to the same reference
    dvm = {}
    if args.dvm_steps > 0:
        for tag, prec, use_b in (("dvm_fp64", pkg.capi.PREC_FP64, False), ("dvm_fp32_batched", pkg.capi.PREC_FP32, True)):
            _, _, its_d, hit_d = run(prec, 0, args.dvm_steps, use_b, False, ref=ref, mode=pkg.capi.MODE_M0)
            dvm[tag] = hit_d

    def leg(hit, its, esz):
        if hit is None:
            return {"reached": False}
        k, t, e = hit
        # algorithmic bytes of the kinetic kernels per step (read+write G,H;
        # moments read G,H) and of the inner loop (64 B per swept cell and iteration)
        swept = max(nx - 5, 0) * max(ny - 5, 0)
        byt = k * (6 * esz * units) + 64 * swept * sum(its[:k])
        return {"reached": True, "steps": k, "seconds": t, "rel_l2": e, "inner_iterations": int(sum(its[:k])),
                "achieved_GBs": byt / t / 1e9, "frac_of_hbm": byt / t / 1e9 / HBM_PEAK_GBS}

    out = {
        "metric": "time to 1e-4 rel L2 vs plain FP64 GSIS",
        "config": {"workload": f"lid-driven cavity {args.n}x{args.n} cells, {args.nv}x{args.nv} sinh velocity grid, "
                               f"Kn={args.kn}, lid {args.lid}, dt={args.dt}, UP1 fluxes, from rest; reference = step "
                               f"{args.ref_steps} of the plain FP64 run"},
        "plain_fp64": leg(hit_plain, its_ref, 8),
        "fast": dict(leg(hit_fast, its_fast, 4 if args.fast_precision == "fp32" else 8),
                     precision=("fp32 sweeps / fp64 moments and synthetic solve" if args.fast_precision == "fp32"
                                else "fp64"),
                     inner_iterations_cap=args.fast_maxit, cfl=args.fast_cfl, dtmax=args.fast_dtmax,
                     batched_graph=True, closure=bool(args.closure), closure_warmup_steps=args.closure_warmup),
        "plain_run": {"steps": args.ref_steps, "seconds": t_ref},
        "mass_correct": bool(args.mass_correct),
    }
    for tag, hit in dvm.items():
        out[tag] = leg(hit, [0] * (hit[0] if hit else 0), 8 if tag == "dvm_fp64" else 4)
        if hit and hit_plain:
            out[tag]["speedup_vs_plain_gsis"] = hit_plain[1] / hit[1]
    if hit_plain and hit_fast:
        out["speedup"] = hit_plain[1] / hit_fast[1]
    print(json.dumps(out), flush=True)


def bench_decks(pkg, solver, args, nsteps=2150, every=50):
    """The reference's shipped cylinder and channel decks as shipped (NND,
    dt_global = 0.01, mode M0), `nsteps` steps on the GPU: the residual every
    `every` steps, and the step at which a run stops with T < 0.1."""
    golden = ROOT / "tests" / "golden"
    out = {"metric": "residual history of the shipped decks", "steps": nsteps, "decks": {}}
    for name in ("channel", "cylinder"):
        case = pkg.cases.shipped_case(golden / name, name)
        rec = {}
        for leg, make in [("gpu", lambda: solver.GsisSolver(case, kstep_show=every))]:
            s = make()
            s.initialize()
            hist, stop = [], None
            t0 = time.perf_counter()
            for k in range(every, nsteps + 1, every):
                try:
                    s.step(every)
                    r = [float(x) for x in s.residual()]
                except Exception as e:  # T < 0.1: the reference's stop
                    stop = {"before_step": k, "error": str(e)[:120]}
                    break
                hist.append((k, r))
            rec[leg] = {"history": hist, "stopped": stop, "seconds": time.perf_counter() - t0}
            print(f"# {name} {leg}: {len(hist) * every} steps, last {hist[-1] if hist else None}, stop {stop}",
                  flush=True)
            s.close()
        out["decks"][name] = rec
    print(json.dumps(out), flush=True)


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
    ap.add_argument("--workload", choices=("cavity", "batch", "production", "decks"), default="cavity",
                    help="cavity: one lid-driven cavity (the headline); batch: the heterogeneous batch of "
                         "--instances Couette and cavity instances (FP32 sweeps, kinetic step, graph replays); "
                         "production: time to 1e-4 relative L2 of plain FP64 GSIS and of the fast form")
    ap.add_argument("--ref-steps", type=int, default=240, help="production: steps of the plain FP64 GSIS run")
    ap.add_argument("--fast-maxit", type=int, default=0,
                    help="production: cap on inner iterations per step of the fast form (0: none)")
    ap.add_argument("--fast-cfl", type=float, default=20.0, help="production: CFL of the fast form's synthetic solve")
    ap.add_argument("--fast-dtmax", type=float, default=10.0,
                    help="production: largest local time step of the fast form's synthetic solve")
    ap.add_argument("--closure", type=int, default=0, help="production: 1 uses the trained closure in the fast form")
    ap.add_argument("--closure-warmup", type=int, default=0,
                    help="production: the trained network's HoT replace the DVM's in the fast form's first N steps")
    ap.add_argument("--fast-precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--dvm-steps", type=int, default=0,
                    help="production: also run the kinetic step alone (FP64, and FP32 batched) up to this many steps")
    ap.add_argument("--mass-correct", type=int, default=1,
                    help="production: 1 runs both forms with the reference's MassCorrect after the moments "
                         "(without it GSIS loses mass in the closed cavity every step and does not converge)")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261202)
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
    if args.workload == "batch":
        bench_batch(pkg, solver, args)
        return
    if args.workload == "production":
        bench_production(pkg, solver, args)
        return
    if args.workload == "decks":
        bench_decks(pkg, solver, args)
        return
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
