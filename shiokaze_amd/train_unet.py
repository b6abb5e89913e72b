"""Trains the learned preconditioner on one GPU and evaluates it in flexible CG.

Training data come from the plain solver: for each scene of the irregular
generator (20 seeded boxes/spheres, pool y < 0.4 plus 8 liquid spheres), the
projection is solved with Jacobi-PCG in FP64 to a tight residual; the scaled
system (b^ = D^-1/2 b, mask, couplings) and its solution x^ form one sample.
Residuals r_k = b^ - A^ x^_k of capped solves, with the error x^ - x^_k as label,
add samples whose spectrum looks like what flexible CG hands the preconditioner
in later iterations.  Scenes stream through a ring buffer in HBM; every training
step is one sample through the library's own BF16 MFMA forward and backward
kernels and an Adam update (shk_unet_train_step).

Evaluation projects held-out scenes twice: Preconditioner=unet (flexible CG with
the trained network) and the plain Jacobi-PCG (mixed precision), both to the same
relative residual, and reports iterations, times and fallbacks.

    python -m shiokaze_amd.train_unet --n 128 --scenes 2000 --out weights.bin
"""
from __future__ import annotations

import argparse
import collections
import json
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import scene as scn
from .learned import UNet, residual
from .projection import MacPressureSolver3

FIELDS = ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")


def log(msg: str) -> None:
    print(f"[train] {msg}", flush=True)


def _dev(sc):
    return {k: torch.from_numpy(np.ascontiguousarray(sc[k])).cuda() for k in FIELDS}


def _project(ps: MacPressureSolver3, d: dict, dt: float = 1.0) -> dict:
    vel = [d[k].clone() for k in ("u", "v", "w")]
    return ps.project(dt, vel, d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])


class SampleMaker:
    """Scaled systems and labels of scenes, from the library's own projection."""

    def __init__(self, n: int, kind: int, label_residual: float, capped=(8, 40)):
        self.n, self.kind, self.capped = n, kind, capped
        self.ps = MacPressureSolver3((n, n, n), 1.0 / n, config={
            "Precision": "fp64", "Preconditioner": "jacobi", "Residual": label_residual,
            "MaxIterations": 100000})

    def samples(self, sc: dict):
        d = _dev(sc)
        res = _project(self.ps, d)
        if res["unknowns"] == 0 or res["status"] != 0:
            return []
        s = self.ps.export_scaled()
        base = (s["mask"], s["hx"], s["hy"], s["hz"])
        out = [(s["rhat"], *base, s["xhat"])]
        xstar = s["xhat"]
        for k in self.capped:
            self.ps.configure("MaxIterations", k)
            _project(self.ps, d)
            self.ps.configure("MaxIterations", 100000)
            xk = self.ps.export_scaled()["xhat"]
            rk = residual(s["hx"], s["hy"], s["hz"], xk, s["rhat"]) * s["mask"]
            out.append((rk, *base, (xstar - xk) * s["mask"]))
        return out

    def close(self):
        self.ps.close()


def _scene_stream(kind, n, seeds, workers):
    """Host scene generation on a thread pool (the generator releases the GIL), at most
    2 * workers scenes ahead of the consumer; with a list of sizes, seed s is generated
    at n[s % len(n)]."""
    sizes = n if isinstance(n, (list, tuple)) else [n]
    seeds = list(seeds)
    with ThreadPoolExecutor(workers) as ex:
        pending = collections.deque()
        it = iter(seeds)
        for s in it:
            pending.append(ex.submit(scn.generate, kind, sizes[s % len(sizes)], s))
            if len(pending) >= 2 * workers:
                break
        while pending:
            sc = pending.popleft().result()
            nxt = next(it, None)
            if nxt is not None:
                pending.append(ex.submit(scn.generate, kind, sizes[nxt % len(sizes)], nxt))
            yield sc


def evaluate(weights: str, n: int, kind: int, seeds, residual_tol: float, workers: int = 8) -> dict:
    """Flexible CG with the network against Jacobi-PCG (mixed) on held-out scenes."""
    rows = []
    cfg_plain = {"Precision": "mixed", "Preconditioner": "jacobi", "Residual": residual_tol}
    cfg_unet = {"Preconditioner": "unet", "UnetWeights": weights, "Residual": residual_tol}
    with MacPressureSolver3((n, n, n), 1.0 / n, config=cfg_plain) as plain, \
            MacPressureSolver3((n, n, n), 1.0 / n, config=cfg_unet) as learned:
        for i, sc in enumerate(_scene_stream(kind, n, seeds, workers)):
            d = _dev(sc)
            row = {}
            for name, ps in (("plain", plain), ("unet", learned)):
                _project(ps, d)  # warm (allocations, weights)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = _project(ps, d)
                torch.cuda.synchronize()
                row[name] = {"ms": 1e3 * (time.perf_counter() - t0), "iters": r["count"], "outer": r["outer"],
                             "relres": r["true_relres_inf"], "status": r["status"], "fallback": r["fallback"]}
            rows.append(row)
            if i % 10 == 0:
                log(f"eval {i}: plain {row['plain']['iters']} it {row['plain']['ms']:.1f} ms | unet "
                    f"{row['unet']['iters']} it {row['unet']['ms']:.1f} ms fb={row['unet']['fallback']}")
    def agg(name, key):
        return float(np.mean([r[name][key] for r in rows]))
    return {
        "scenes": len(rows),
        "plain_iters": agg("plain", "iters"), "unet_iters": agg("unet", "iters"),
        "plain_ms": agg("plain", "ms"), "unet_ms": agg("unet", "ms"),
        "unet_fallbacks": int(sum(r["unet"]["fallback"] for r in rows)),
        "unet_max_relres": float(max(r["unet"]["relres"] for r in rows)),
        "plain_max_relres": float(max(r["plain"]["relres"] for r in rows)),
        "all_converged": bool(all(r["unet"]["status"] == 0 and r["plain"]["status"] == 0 for r in rows)),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--sizes", default=None,
                    help="comma-separated grid sizes to train on in turn (default: --n); seed s uses sizes[s %% k]")
    ap.add_argument("--eval-sizes", default=None, help="comma-separated grid sizes to evaluate on (default: --n)")
    ap.add_argument("--kind", type=int, default=3)
    ap.add_argument("--scenes", type=int, default=2000, help="training scenes (seeds 0..scenes-1)")
    ap.add_argument("--eval-seed0", type=int, default=2000)
    ap.add_argument("--eval", type=int, default=200, help="held-out scenes (seeds eval_seed0..)")
    ap.add_argument("--buffer", type=int, default=64, help="scenes kept in the HBM ring buffer")
    ap.add_argument("--steps-per-scene", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4)
    ap.add_argument("--omega", type=float, default=0.8, help="weighted-Jacobi smoothing weight")
    ap.add_argument("--smooth-steps", type=int, default=2,
                    help="smoothing steps of the training loss's preconditioner")
    ap.add_argument("--chebyshev", type=float, default=0.0,
                    help="> 1: the training loss's smoothing steps use Chebyshev weights on [2/ratio, 2]")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--label-residual", type=float, default=1e-10)
    ap.add_argument("--eval-residual", type=float, default=1e-6)
    ap.add_argument("--time-budget", type=float, default=600.0, help="seconds of training before evaluation")
    ap.add_argument("--init", default=None, help="weights file to start from")
    ap.add_argument("--out", default=None, help="weights file to write (required unless --eval-only)")
    ap.add_argument("--eval-only", default=None, metavar="WEIGHTS",
                    help="skip training: evaluate this weights file on the held-out scenes")
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--report", default=None, help="JSON file for the training curve and evaluation")
    args = ap.parse_args(argv)
    if args.eval_only:
        report = {"weights": args.eval_only, "eval": {}}
        for n in ([int(v) for v in args.eval_sizes.split(",")] if args.eval_sizes else [args.n]):
            count = args.eval if n <= 128 else max(1, args.eval // 4)
            ev = evaluate(args.eval_only, n, args.kind, range(args.eval_seed0, args.eval_seed0 + count),
                          args.eval_residual, workers=args.workers)
            log(f"eval {n}: " + json.dumps(ev))
            report["eval"][str(n)] = ev
        if args.report:
            with open(args.report, "w") as f:
                json.dump(report, f, indent=1)
        return report
    if not args.out:
        ap.error("--out is required when training")

    rng = np.random.default_rng(args.seed)
    net = UNet(seed=args.seed)
    net.set_smoothing(args.smooth_steps, args.chebyshev)
    if args.init:
        net.load(args.init)
    sizes = [int(v) for v in args.sizes.split(",")] if args.sizes else [args.n]
    makers = {n: SampleMaker(n, args.kind, args.label_residual) for n in sizes}
    ring: list = []
    curve = []
    t0 = time.perf_counter()
    step = 0
    total_steps = args.scenes * args.steps_per_scene
    log(f"sizes={sizes} scenes={args.scenes} params={net.num_params} steps<={total_steps} budget={args.time_budget}s")
    t_gen = t_train = 0.0
    seen = 0
    ema = None
    for sc in _scene_stream(args.kind, sizes, range(args.scenes), args.workers):
        ta = time.perf_counter()
        samples = makers[sc["shape"][0]].samples(sc)
        torch.cuda.synchronize()
        t_gen += time.perf_counter() - ta
        if not samples:
            continue
        seen += 1
        ring.append(samples)
        if len(ring) > args.buffer:
            ring.pop(0)
        tb = time.perf_counter()
        for _ in range(args.steps_per_scene):
            frac = min(1.0, step / max(1, total_steps))
            lr = args.lr * (args.lr_final / args.lr) ** frac
            smp = ring[rng.integers(len(ring))]
            r, mask, hx, hy, hz, xs = smp[rng.integers(len(smp))]
            loss = net.train_step(r, mask, hx, hy, hz, xs, omega=args.omega, lr=lr)
            ema = loss if ema is None else 0.98 * ema + 0.02 * loss
            step += 1
        t_train += time.perf_counter() - tb
        el = time.perf_counter() - t0
        if seen % 25 == 0:
            curve.append({"scene": seen, "step": step, "loss_ema": ema, "elapsed_s": el})
            log(f"scene {seen} step {step} loss(ema) {ema:.4f} lr {lr:.2e} gen {t_gen:.0f}s train {t_train:.0f}s "
                f"({1e3 * t_train / max(1, step):.1f} ms/step)")
            net.save(args.out)
        if el > args.time_budget:
            log(f"time budget reached after {seen} scenes")
            break
    net.save(args.out)
    for m in makers.values():
        m.close()
    net.close()
    log(f"saved {args.out} after {step} steps on {seen} scenes")
    report = {"sizes": sizes, "scenes_used": seen, "steps": step, "final_loss_ema": ema, "train_s": t_train,
              "gen_s": t_gen, "curve": curve, "smoothing": {"steps": args.smooth_steps, "chebyshev": args.chebyshev}}
    if args.eval > 0:
        report["eval"] = {}
        for n in ([int(v) for v in args.eval_sizes.split(",")] if args.eval_sizes else [args.n]):
            count = args.eval if n <= 128 else max(1, args.eval // 4)
            ev = evaluate(args.out, n, args.kind, range(args.eval_seed0, args.eval_seed0 + count), args.eval_residual,
                          workers=args.workers)
            log(f"eval {n}: " + json.dumps(ev))
            report["eval"][str(n)] = ev
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    return report


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
