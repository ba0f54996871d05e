"""Train the learned HoT closure on fields the DVM solver generates itself, on
the GPU, and evaluate it on held-out cavities.

    python hybrid2d-2--sis---shakov---coette--9-6_amd/train_closure.py --out closure_hot.npz --report report.json

Training set: states near convergence (the last `keep` of `steps` GSIS
steps, mode M1, first-order upwind fluxes, the scheme whose cavity runs
converge; with the NND limiter their residual stalls) of the planar Couette
flow (4 x 128 cells, 32^2 velocities, walls +-0.1) at Kn 0.01, 0.03, 0.1,
0.3 and 1, and of the lid-driven cavity (lid 0.15) on 128^2 cells with 64^2
velocities at Kn 0.01, 0.1 and 1 and on 64^2 cells with 32^2 velocities at
Kn 0.02, 0.05, 0.2 and 0.5.  Each sample is one closure face: its network
inputs and the Convert_NS HoT.  Held out: cavities with Kn drawn
log-uniformly from [0.005, 2].  Training runs gsis_mlp_train (BF16-MFMA
forward and backward, Adam) on the standardised samples.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import pathlib
import sys
import time

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent


def _pkg():
    if "gsis_amd" in sys.modules:
        return sys.modules["gsis_amd"]
    spec = importlib.util.spec_from_file_location("gsis_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gsis_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def harvest(case, steps: int, first: int = 1, ns_maxit: int = 3000, log=print):
    """(x, y, residuals) of the closure faces after each step >= first of a
    GSIS run of `steps` steps from the case's initial state."""
    pkg = _pkg()
    from gsis_amd import closure, solver
    s = solver.GsisSolver(case, mode=pkg.capi.MODE_M1, ns_maxit=ns_maxit)
    s.initialize()
    xs, ys, res = [], [], []
    t0 = time.time()
    for k in range(1, steps + 1):
        s.step(1)
        if k >= first:
            x, y = closure.samples(s)
            xs.append(x)
            ys.append(y)
        if k % 5 == 0 or k == steps:
            r = s.residual()
            res.append((k, r.tolist(), s.counters()[2]))
            log(f"  {case.name} step {k}: inner its {s.counters()[2]}, residual "
                f"{np.array2string(r, precision=2)}, {time.time() - t0:.1f} s", flush=True)
    s.close()
    return np.concatenate(xs), np.concatenate(ys), res


def upwind(pkg, case):
    case.const["scheme"] = pkg.capi.SCHEME_UP1
    return case


def training_cases(pkg, steps: int, keep: int):
    """(case, steps, first kept step) of every training run."""
    cou = [upwind(pkg, pkg.cases.couette_case(ny_cells=128, nx_cells=4, nv=32, kn=kn, uwall=0.1, dt=1.0))
           for kn in (0.01, 0.03, 0.1, 0.3, 1.0)]
    big = [upwind(pkg, pkg.cases.cavity_case(128, 64, kn=kn, lid=0.15, dt=1.0)) for kn in (0.01, 0.1, 1.0)]
    small = [upwind(pkg, pkg.cases.cavity_case(64, 32, kn=kn, lid=0.15, dt=1.0)) for kn in (0.02, 0.05, 0.2, 0.5)]
    runs = [(c, 100, 61) for c in cou]
    runs += [(c, steps, steps - keep + 1) for c in big]
    runs += [(c, 2 * steps, 2 * steps - 2 * keep + 1) for c in small]
    return runs


def heldout_kn(count: int, seed: int):
    rng = np.random.default_rng(seed)
    return sorted(float(math.exp(v)) for v in rng.uniform(math.log(0.005), math.log(2.0), count))


def evaluate(net, x, y, gate):
    """Relative L2 error per output, and the fraction of faces the gate accepts."""
    p = net.forward(x)
    err = [float(np.linalg.norm(p[:, k] - y[:, k]) / max(np.linalg.norm(y[:, k]), 1e-300)) for k in range(y.shape[1])]
    sd = net.scaling[21:26]
    acc = float(np.mean(np.all(np.abs(p - y) / sd <= gate, axis=1)))
    return err, acc


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(PKG_DIR / "closure_hot.npz"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--gate", type=float, default=1.0)
    ap.add_argument("--heldout", type=int, default=4)
    ap.add_argument("--heldout-n", type=int, default=64)
    ap.add_argument("--heldout-steps", type=int, default=100)
    ap.add_argument("--steps", type=int, default=60, help="GSIS steps of each 128^2 training cavity (twice that at 64^2)")
    ap.add_argument("--keep", type=int, default=20, help="last steps of each 128^2 run kept (twice that at 64^2)")
    ap.add_argument("--ns-maxit", type=int, default=1000)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the training steps (quick runs)")
    args = ap.parse_args(argv)
    pkg = _pkg()
    from gsis_amd import closure
    t0 = time.time()
    xs, ys, runs = [], [], {}
    for case, steps, first in training_cases(pkg, args.steps, args.keep):
        n = max(2, int(round(steps * args.scale)))
        f = max(1, min(n, int(round(first * args.scale))))
        print(f"harvest {case.name}: {n} GSIS steps, kept from step {f}", flush=True)
        x, y, res = harvest(case, n, first=f, ns_maxit=args.ns_maxit)
        xs.append(x)
        ys.append(y)
        runs[case.name] = {"steps": n, "first_kept": f, "samples": int(x.shape[0]), "history": res}
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    ok = np.all(np.isfinite(x), axis=1) & np.all(np.isfinite(y), axis=1)
    x, y = x[ok], y[ok]
    print(f"{x.shape[0]} samples in {time.time() - t0:.1f} s", flush=True)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(x.shape[0])
    nval = x.shape[0] // 20
    val, tr = perm[:nval], perm[nval:]
    scaling = closure.scaling_of(x[tr], y[tr])
    net = closure.Mlp(closure.init_params(args.seed), scaling)
    t1 = time.time()
    hist = []
    for lr, ep in ((args.lr, args.epochs), (args.lr * 0.2, max(1, args.epochs // 3))):
        h = net.train(x[tr], y[tr], epochs=ep, batch=args.batch, lr=lr, seed=args.seed + len(hist))
        hist += [float(v) for v in h]
        print(f"  lr {lr:g}: {ep} epochs, loss {h[0]:.4g} -> {h[-1]:.4g}", flush=True)
    t_train = time.time() - t1
    vloss, _ = net.loss_grad(x[val][:8192], y[val][:8192])
    err_val, acc_val = evaluate(net, x[val], y[val], args.gate)
    print(f"trained in {t_train:.1f} s; validation loss {vloss:.4g}, rel L2 per output "
          f"{np.round(err_val, 4).tolist()}, accepted {acc_val:.3f}", flush=True)
    held = {}
    for kn in heldout_kn(args.heldout, args.seed):
        case = upwind(pkg, pkg.cases.cavity_case(args.heldout_n, 32, kn=kn, lid=0.15, dt=1.0))
        hx, hy, _ = harvest(case, args.heldout_steps, first=args.heldout_steps, ns_maxit=args.ns_maxit)
        e, a = evaluate(net, hx, hy, args.gate)
        held[f"{kn:.4g}"] = {"rel_l2": e, "accepted": a, "faces": int(hx.shape[0])}
        print(f"held-out Kn {kn:.4g}: rel L2 {np.round(e, 4).tolist()}, accepted {a:.3f}", flush=True)
    meta = {"samples": int(x.shape[0]), "train": int(tr.size), "validation": int(nval), "epochs": len(hist),
            "batch": args.batch, "lr": args.lr, "seed": args.seed, "loss_per_epoch": hist,
            "validation_loss": vloss, "validation_rel_l2": err_val, "validation_accepted": acc_val,
            "gate": args.gate, "train_seconds": t_train, "runs": runs, "heldout": held}
    closure.save(args.out, net.params(), scaling, meta)
    if args.report:
        pathlib.Path(args.report).write_text(json.dumps(meta, indent=1))
    net.close()
    print(f"wrote {args.out} ({time.time() - t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
