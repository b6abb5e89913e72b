"""Benchmark: pressure solves/sec of the MAC liquid projection on one MI355X.

A step is one whole projection (shk_project: fractions, labels, coefficients,
rhs, Krylov solve to the relative max-norm residual, pressure scatter, velocity
update) of the same synthetic scene, with the input velocity restored first; all
fields stay resident in HBM.  Default workload: the pool+sphere scene at 256^3,
FP32 stencil/AXPY/Jacobi inner CG with FP64 dots and FP64 iterative refinement,
to 1e-6.

Prints one JSON line (rank 0).  `roofline` is for the dominant Krylov kernel:
algorithmic bytes per launch / its average launch duration, timed with HIP events
on the library's stream (ProfileKernels=1) over --profile-steps further steps of
the same context right after the timed ones (the events add a little time per
launch, so the timed steps themselves run without them).
`cpu_baseline` is the oracle (the FP64 CPU restatement of the reference's path) on the
host: the whole projection with the CG-equivalent solve to the same stop, timed in full
on the host threads, beside the reference-faithful serial CG (its discarded MIC(0) work
included) extrapolated from 10 and 30 timed iterations.  The same leg solves the system
to 1e-12 with the oracle's FP64 Jacobi-PCG: the truth of `accuracy`.
`accuracy` holds the pressure rel-L2 against that truth and the FP64 post-projection
divergence, with `meets_bar` for the bars (1e-5, 1e-6) and `bar_point`, the loosest
max-norm stop at which the solver meets both bars and its solves/s.
`north_star` times plain FP64 GPU CG (the reference's iteration), multigrid PCG and
the learned preconditioner in flexible CG in the same run, with their ratios to plain
FP64 CG, their accuracy and their rooflines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "pressure solves/sec at 256^3 to 1e-6 rel. residual; PCG iters; % HBM peak"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# HBM bytes per brick cell of one launch measured with rocprofv3 --pmc (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, factors from a 256 MiB copy in the same run), mixed precision, Jacobi,
# pool+sphere scene: at 256^3 profiles/hbm_traffic_pool_sphere_256_mixed_jacobi.md, at 512^3
# profiles/krylov_pool_sphere_512_mixed_fp64.md.  Reported as roofline.traffic only for
# those workloads.
MEASURED_TRAFFIC_PER_CELL = {256: {"spmv_dot": 26.11, "axpy_norm": 24.03},
                             512: {"spmv_dot": 25.42, "axpy_norm": 24.01}}


def kernel_bytes_per_cell(kernel: str, precision: str, precond: str) -> int:
    """Algorithmic HBM bytes per brick cell of one launch of each Krylov kernel:
    every array element the kernel needs is read or written once.  With Jacobi the
    system is the symmetrically scaled one (unit diagonal, no diagonal array)."""
    t = 8 if precision == "fp64" else 4      # vector element
    c = 8 if precision == "fp64" else 4      # coefficient element
    jac = precond in ("jacobi", "mg", "unet")  # all three run CG on the scaled unit-diagonal system
    if kernel == "spmv_dot":                  # r, s in; cx, cy, cz (+ diag) in; s', q out
        return 4 * t + 3 * c + (0 if jac else c)
    if kernel == "axpy_norm":                 # x, r, s, q in; x, r out (+ 1/sqrt(diag) for the FP64 stop test)
        return 6 * t + (8 if (jac and precision == "fp64") else 0)
    raise ValueError(kernel)


# Algorithmic HBM bytes per launch of the liquid-step kernels, in units of G = n^3
# (every face array holds ~G faces): value + mask read and written once each.
LIQUID_BYTES_PER_CELL = {
    "k_extrapolate_round": 3,      # 3 axes x the mask read once (in place: only the activation shell's
                                   # neighbour values are read and its faces written, a few % of the grid)
    "k_semi_lagrangian_fwd": 32,   # u, v, w + masks read (15), fluid (4), v0, lo, hi, band written (13)
    "k_semi_lagrangian_bwd": 32,   # v0 x3 + masks (15), vin, lo, hi, band (13), result written (4)
    "k_constrain": 5,              # per axis: the mask and one solid node per face (the snapshot is read only
                                   # inside solids, and only changed faces are written)
    "k_max_u2": 15,                # u, v, w + masks read once
    "k_project_activity": 30,      # 3 axes x (mask read and written, area and rho read)
    "k_gravity": 30,               # 3 axes x (4+1 read, 4+1 written)
    "k_advect_cell": 1,            # per pass: the band mask; only band cells (a few % of the grid) read
                                   # the level set and six faces and write their results
}


def liquid_leg(args, torch, shk, S, sc, dev, cfg, steps=3):
    """Times macliquid3's step at the bench size on the bench scene, stage by stage (each
    library call is synchronous): the velocity half, the level-set tracker (advection,
    redistancing to width 3, extrapolation into the solid; maclevelsetsurfacetracker3.cpp:
    46-62) and the projection; and the CPU oracle of the same stages on a (up to) 128^3
    sample of the same scene."""
    n = args.n
    shape = (n, n, n)
    ps = shk.MacPressureSolver3(shape, sc["dx"], config=cfg)
    lv = shk.MacLiquidVelocity(ps)
    vel = [dev[k].clone() for k in ("u", "v", "w")]
    svel = [torch.zeros_like(v) for v in vel]
    masks = [torch.ones(v.shape, dtype=torch.uint8, device="cuda") for v in vel]
    ps.project(1.0 / 256, vel, dev["solid"], svel, dev["fluid"])   # a divergence-free start
    lv.project_activity(masks)
    lv.step(vel, masks, dev["solid"], svel, dev["fluid"])            # warm-up step
    # the level set's narrow band (half width 3 dx, array3::set_as_levelset) for the
    # level-set advection stage; the projection keeps using the scene's level set
    bw = 3.0 * ps.dx
    band = (dev["fluid"].abs() < bw).to(torch.uint8)
    phi0 = dev["fluid"].clamp(-bw, bw)
    stages = {"max_u": 0.0, "extrapolate": 0.0, "constrain": 0.0, "advect_levelset": 0.0, "redistance": 0.0,
              "extrapolate_levelset": 0.0, "advect": 0.0, "gravity": 0.0, "project": 0.0}
    # the solid's node band as the reference keeps it (|phi_s| < 3 dx, -/+ 3 dx off it)
    sband = (dev["solid"].abs() < bw).to(torch.uint8)
    solid_b = torch.where(sband > 0, dev["solid"], torch.where(dev["solid"] < 0, -bw, bw)).contiguous()
    widths = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        max_u = lv.compute_max_u(vel, masks)
        stages["max_u"] += time.perf_counter() - t
        dt = lv.timestepper.advance(max_u, ps.dx)
        width = int(lv.extend_width + lv.timestepper.current_cfl)
        widths.append(width)
        t = time.perf_counter()
        lv.extrapolate(vel, masks, width)
        stages["extrapolate"] += time.perf_counter() - t
        t = time.perf_counter()
        lv.constrain(dev["solid"], vel, masks)
        stages["constrain"] += time.perf_counter() - t
        phi = phi0.clone()
        torch.cuda.synchronize()
        t = time.perf_counter()
        b = band.clone()
        lv.advect_levelset(dt, phi, b, vel, masks, bw)
        stages["advect_levelset"] += time.perf_counter() - t
        t = time.perf_counter()
        lv.redistance(phi, b, 3)
        stages["redistance"] += time.perf_counter() - t
        t = time.perf_counter()
        lv.extrapolate_levelset(solid_b, phi, b, bw, solid_band=sband)
        stages["extrapolate_levelset"] += time.perf_counter() - t
        t = time.perf_counter()
        lv.advect(dt, vel, masks, dev["fluid"])
        stages["advect"] += time.perf_counter() - t
        t = time.perf_counter()
        lv.add_gravity(dt, vel, masks)
        stages["gravity"] += time.perf_counter() - t
        t = time.perf_counter()
        res = ps.project(dt, vel, dev["solid"], svel, dev["fluid"])
        lv.project_activity(masks)
        stages["project"] += time.perf_counter() - t
    total = time.perf_counter() - t_all
    ps.close()
    tracker_keys = ("advect_levelset", "redistance", "extrapolate_levelset")
    tracker = sum(stages[k] for k in tracker_keys)
    velocity_half = sum(v for k, v in stages.items() if k != "project" and k not in tracker_keys)
    out = {"steps_per_s": round(steps / total, 3), "velocity_half_ms": round(1e3 * velocity_half / steps, 3),
           "levelset_tracker_ms": round(1e3 * tracker / steps, 3),
           "stage_ms": {k: round(1e3 * v / steps, 3) for k, v in stages.items()}, "extend_widths": widths,
           "last_projection_iters": res["count"], "preconditioner": cfg.get("Preconditioner"),
           "advect_algorithmic_bytes": int((LIQUID_BYTES_PER_CELL["k_semi_lagrangian_fwd"]
                                            + LIQUID_BYTES_PER_CELL["k_semi_lagrangian_bwd"]) * 3 * n ** 3)}
    if not args.no_cpu_baseline:
        # the same stages in the CPU oracle (one core) on a 64^3 sample of the scene
        from oracle import oracle as O
        m = min(n, 128)
        small = S.generate(args.scene, m, args.seed)
        v = [np.ascontiguousarray(small[k]) for k in ("u", "v", "w")]
        areas, rhos = O.area_fraction(small["solid"]), O.fluid_fraction(small["fluid"])
        msk = [((a != 0) & (r != 0)).astype(np.uint8) for a, r in zip(areas, rhos)]   # after a projection
        tc = {}
        t = time.perf_counter()
        O.compute_max_u(v, msk)
        tc["max_u"] = time.perf_counter() - t
        t = time.perf_counter()
        v, msk, _ = O.extrapolate(v, msk, 4)
        tc["extrapolate"] = time.perf_counter() - t
        t = time.perf_counter()
        v = O.constrain_velocity(small["dx"], small["solid"], v, msk)
        tc["constrain"] = time.perf_counter() - t
        bws = 3.0 * small["dx"]
        t = time.perf_counter()
        sb = (np.abs(small["fluid"]) < bws).astype(np.uint8)
        ph = O.advect_levelset(small["dx"], 0.5 * small["dx"], np.clip(small["fluid"], -bws, bws), sb, v, msk, bws)
        tc["advect_levelset"] = time.perf_counter() - t
        t = time.perf_counter()
        ph, sb, _ = O.redistance(small["dx"], ph, sb, 3)
        tc["redistance"] = time.perf_counter() - t
        snb = (np.abs(small["solid"]) < bws).astype(np.uint8)
        sol = np.where(snb > 0, small["solid"], np.where(small["solid"] < 0, -bws, bws)).astype(np.float32)
        t = time.perf_counter()
        O.extrapolate_levelset(small["dx"], sol, ph, sb, bws, solid_band=snb)
        tc["extrapolate_levelset"] = time.perf_counter() - t
        t = time.perf_counter()
        v = O.advect_velocity(small["dx"], 0.5 * small["dx"], v, msk, small["fluid"])
        tc["advect"] = time.perf_counter() - t
        t = time.perf_counter()
        O.add_gravity(0.001, (0.0, -9.8, 0.0), v, msk)
        tc["gravity"] = time.perf_counter() - t
        cpu_all = sum(tc.values())
        scale = (n / m) ** 3
        out["cpu_oracle"] = {"sample": f"{m}^3 scene {args.scene}, every stage once, one host core",
                             "stages_ms_sample": round(1e3 * cpu_all, 2),
                             "stages_ms_extrapolated_to_n": round(1e3 * cpu_all * scale, 1),
                             "stage_ms_sample": {k: round(1e3 * x, 2) for k, x in tc.items()}}
        out["stages_gpu_vs_cpu"] = round(cpu_all * scale / ((velocity_half + tracker) / steps), 1)
    return out


PRESSURE_BAR = 1e-5   # pressure rel-L2 against the FP64 solve to 1e-12
DIVERGENCE_BAR = 1e-6  # FP64 post-projection divergence relative to the initial one
# candidate max-norm stops for the loosest setting that meets both bars
BAR_STOPS = (1e-6, 5e-7, 2e-7, 1e-7, 5e-8, 2e-8, 1e-8, 5e-9, 2e-9, 1e-9, 5e-10, 2e-10, 1e-10)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """Host threads for the oracle's threaded solves: the process's CPU share (at most 16,
    the share of one GPU on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_leg(args, sc, residual):
    """The oracle on the host: the system (fractions, labels, CSR assembly), the FP64
    truth (Jacobi-PCG to 1e-12 on T threads), the CG-equivalent solve to the bench
    tolerance timed in full on T threads, and the reference-faithful CG (MIC(0) factor and
    two discarded triangular solves per iteration, one thread) timed for 10 iterations
    and extrapolated to the CG count."""
    from oracle import oracle as O
    T = cpu_threads()
    O.set_threads(T)
    sy = O.system(sc)
    log(f"cpu: system N={sy['n']} in {sy['seconds']:.2f}s on {T} threads")
    xt, it_t, rr_t, st_t, sec_t = O.pcg_threaded(sy["rowptr"], sy["cols"], sy["vals"], sy["rhs"], 1e-12, 100000,
                                                 O.SHK_OR_JACOBI)
    log(f"cpu: truth Jacobi-PCG to 1e-12: {it_t} iterations in {sec_t:.1f}s (status {st_t})")
    xc, it_c, rr_c, st_c, sec_c = O.pcg_threaded(sy["rowptr"], sy["cols"], sy["vals"], sy["rhs"], residual, 100000,
                                                 O.SHK_OR_CG)
    log(f"cpu: CG to {residual:g}: {it_c} iterations in {sec_c:.1f}s")
    # the scatter and velocity update of the same projection (threaded in the oracle)
    t0 = time.perf_counter()
    p = O.scatter(sy["index_map"], xc)
    areas, rhos = O.area_fraction(sc["solid"]), O.fluid_fraction(sc["fluid"])
    O.update_velocity(1.0, sc["dx"], p, areas, rhos, [sc["us"], sc["vs"], sc["ws"]], [sc["u"], sc["v"], sc["w"]])
    sec_u = time.perf_counter() - t0
    total = sy["seconds"] + sec_c + sec_u
    # reference-faithful single-thread sample
    O.set_threads(1)
    t10 = O.pcg(sy["rowptr"], sy["cols"], sy["vals"], sy["rhs"], residual, 10, O.SHK_OR_REFERENCE)
    ta = time.perf_counter()
    O.pcg(sy["rowptr"], sy["cols"], sy["vals"], sy["rhs"], residual, 10, O.SHK_OR_REFERENCE)
    t10s = time.perf_counter() - ta
    ta = time.perf_counter()
    O.pcg(sy["rowptr"], sy["cols"], sy["vals"], sy["rhs"], residual, 30, O.SHK_OR_REFERENCE)
    t30s = time.perf_counter() - ta
    del t10
    per_it = (t30s - t10s) / 20.0
    fixed = max(t10s - 10 * per_it, 0.0)
    ref_total = sy["seconds"] + fixed + it_c * per_it + sec_u
    O.set_threads(T)
    lab = sy["index_map"] >= 0
    truth = np.zeros(sy["index_map"].shape)
    truth[lab] = xt[sy["index_map"][lab]]
    cg_dense = np.zeros(sy["index_map"].shape)
    cg_dense[lab] = xc[sy["index_map"][lab]]
    cpu = {"value": 1.0 / total, "unit": "solves/s", "cores": T, "kind": "port",
           "sample": (f"{args.n}^3 scene {args.scene}: the whole projection in the oracle (fractions, labels, CSR "
                      f"assembly, CG to {residual:g} with the reference's max-norm stop, scatter, velocity update), "
                      f"timed in full on {T} host threads: {it_c} iterations, solve {sec_c:.2f} s"),
           "ms_per_solve": round(1e3 * total, 1), "iters": it_c, "cpu_model": cpu_model(),
           "reference_faithful_1_core": {
               "solves_per_s_extrapolated": 1.0 / ref_total, "ms_per_iteration": round(1e3 * per_it, 2),
               "note": (f"the reference's serial PCG (MIC(0) factor, two discarded triangular solves per iteration) "
                        f"on one thread, timed for 10 and 30 iterations and extrapolated to {it_c} iterations")},
           "truth": {"mode": "Jacobi-PCG", "residual": 1e-12, "iters": it_t, "seconds": round(sec_t, 2),
                     "status": st_t}}
    return cpu, truth, cg_dense


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scene", type=int, default=1, help="1 pool+sphere, 3 irregular")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--preconditioner", default="jacobi", choices=["none", "jacobi", "unet", "mg"])
    ap.add_argument("--residual", type=float, default=1e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-north-star", action="store_true",
                    help="skip the FP64 plain CG, multigrid and learned-preconditioner lines")
    ap.add_argument("--extra", action="store_true", help="also time FP64 Jacobi, the 2D projection and the liquid step")
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="steps after the timed ones with per-kernel HIP event timing")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device copy of MIB MiB first (a known-size kernel for HBM counter scaling)")
    ap.add_argument("--unet-weights", default=str(ROOT / "shiokaze_amd" / "weights" / "unet_irregular_128.bin"),
                    help="trained network for the learned-preconditioner line")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="extra shk_configure keys for the timed context")
    ap.add_argument("--liquid", action="store_true",
                    help="also time the liquid step around the projection (extrapolation, constraint, "
                         "MacCormack advection, gravity, projection) and its CPU oracle on a 128^3 sample")
    ap.add_argument("--dim", type=int, default=3, choices=[2, 3],
                    help="2: the 2D projection (shk_project2d) on an n^2 scene instead")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the projection runs on a single MI355X (no multi-GPU path is built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import shiokaze_amd
    from shiokaze_amd import scene as S

    if args.dim == 2:
        bench_2d(args, torch, shiokaze_amd, S)
        return

    if args.calibrate_copy > 0:
        src = torch.ones(args.calibrate_copy * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)  # first touch of the destination pages
        torch.cuda.synchronize()
        # three timed copies: the device-to-device copy rate of this box (a measured HBM
        # ceiling for streams that read and write)
        best = float("inf")
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            dst.copy_(src)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        mib = args.calibrate_copy
        log(f"calibration copy: {mib} MiB read + {mib} MiB written; best of 3: {best * 1e6:.1f} us = "
            f"{2 * mib * (1 << 20) / best / 1e9:.1f} GB/s")
        del src, dst

    n = args.n
    t0 = time.time()
    sc = S.generate(args.scene, n, args.seed)
    log(f"scene {args.scene} n={n} generated in {time.time() - t0:.1f}s (dropped {sc['dropped']})")
    dev = {k: torch.from_numpy(np.ascontiguousarray(sc[k])).cuda()
           for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
    vel0 = [dev[k].clone() for k in ("u", "v", "w")]
    vel = (dev["u"], dev["v"], dev["w"])
    svel = (dev["us"], dev["vs"], dev["ws"])

    def make(config):
        return shiokaze_amd.MacPressureSolver3((n, n, n), sc["dx"], config=config)

    def step(ps):
        for a in range(3):
            vel[a].copy_(vel0[a])
        return ps.project(1.0, vel, dev["solid"], svel, dev["fluid"])

    def timed(ps, k):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(k):
            r = step(ps)
        torch.cuda.synchronize()
        return k / (time.perf_counter() - t1), r

    cfg = {"Residual": args.residual, "Precision": args.precision, "Preconditioner": args.preconditioner}
    if args.preconditioner == "unet":
        cfg["UnetWeights"] = args.unet_weights
    for kv in args.set:
        k, v = kv.split("=", 1)
        cfg[k] = v
    ps = make(cfg)
    for _ in range(args.warmup):
        r = step(ps)
        log(f"warmup: iters={r['count']} outer={r['outer']} status={r['status_name']} "
            f"setup={r['ms_setup']:.2f}ms solve={r['ms_solve']:.2f}ms update={r['ms_update']:.2f}ms")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    results = []
    for _ in range(args.steps):
        results.append(step(ps))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    if not args.no_profile and args.profile_steps > 0:
        # per-kernel durations: HIP events recorded on the library's stream around
        # every Krylov launch of these extra steps (same context, same scene)
        ps.configure("ProfileKernels", 1)
        for _ in range(args.profile_steps):
            step(ps)
        torch.cuda.synchronize()
    ms_per_step = 1e3 * elapsed / args.steps
    last = results[-1]
    value = args.steps / elapsed
    log(f"timed: {args.steps} steps in {elapsed:.3f}s -> {value:.3f} solves/s, iters={last['count']} "
        f"outer={last['outer']} true relres={last['true_relres_inf']:.3e}")

    n_cells = last["active_bricks"] * 512
    kernels = {}
    roofline = None
    if not args.no_profile:
        for k in ("spmv_dot", "axpy_norm"):
            cnt, ms = ps.kernel_stats(k)
            if cnt == 0:
                continue
            avg_us = 1e3 * ms / cnt
            bpc = kernel_bytes_per_cell(k, args.precision, args.preconditioner)
            traffic = bpc * n_cells
            gbs = traffic / (avg_us * 1e-6) / 1e9
            kernels[k] = {"launches": cnt, "avg_us": round(avg_us, 3), "bytes_per_cell": bpc,
                          "bytes_per_launch": traffic, "GBs": round(gbs, 1),
                          "frac": round(gbs / HBM_PEAK_GBS, 4), "total_ms": round(ms, 3)}
        if kernels:
            dom = max((k for k in kernels if "GBs" in kernels[k]), key=lambda k: kernels[k]["total_ms"])
            kd = kernels[dom]
            measured = (args.scene == 1 and n in MEASURED_TRAFFIC_PER_CELL and args.precision == "mixed"
                        and args.preconditioner == "jacobi" and not args.set)
            traffic = round(MEASURED_TRAFFIC_PER_CELL[n][dom] * n_cells) if measured else None
            roofline = {"bound": "hbm", "kernel": dom, "achieved": kd["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": kd["frac"], "traffic": traffic,
                        "bytes_per_launch": kd["bytes_per_launch"], "avg_launch_us": kd["avg_us"]}
            solve_bytes = sum(kernels[k].get("bytes_per_launch", 0) * kernels[k]["launches"] for k in kernels)
            solve_ms = sum(kernels[k]["total_ms"] for k in kernels)
            roofline["krylov_loop_GBs"] = round(solve_bytes / (solve_ms * 1e-3) / 1e9, 1)

    # ---- the CPU leg: the oracle's FP64 truth, its CG-equivalent solve timed in full and
    # the reference-faithful serial CG (see cpu_leg)
    cpu = None
    truth_dense = cpu_cg_dense = None
    if not args.no_cpu_baseline:
        cpu, truth_dense, cpu_cg_dense = cpu_leg(args, sc, args.residual)
        if n <= 96:
            # small grids (config 1): the whole reference-faithful projection on one host
            # core, and the GPU pressure compared with the CPU one of the same run
            from oracle import oracle as O
            O.set_threads(1)
            rr = O.project_timed(sc, tol=args.residual, mode=O.SHK_OR_REFERENCE)
            total_ms = float(sum(rr["t_ms"][:3]))
            p_gpu = ps.get_pressure().cpu().numpy().astype(np.float64)
            p_cpu = np.asarray(rr["pressure"], dtype=np.float64)
            cpu["reference_faithful_1_core"] = {
                "solves_per_s": 1e3 / total_ms, "iters": rr["iters"],
                "note": "the whole reference-faithful projection timed on one thread (not extrapolated)",
                "pressure_rel_l2_gpu_vs_cpu": float(np.linalg.norm(p_gpu - p_cpu)
                                                    / max(np.linalg.norm(p_cpu), 1e-300))}
            O.set_threads(cpu_threads())
        log(f"cpu baseline: {cpu}")

    # ---- accuracy against the FP64 truth (the oracle's Jacobi-PCG to 1e-12 on the host;
    # without the CPU leg, the GPU's FP64 Jacobi-PCG to 1e-12)
    accuracy = {}
    truth_x = None
    if not args.no_accuracy:
        sysb = ps.export_system()
        lab = sysb["label"].bool()
        if truth_dense is not None:
            xt = torch.from_numpy(truth_dense).cuda().reshape(lab.shape)[lab]
            truth_src = "cpu oracle FP64 Jacobi-PCG to 1e-12"
        else:
            truth = make({"Residual": 1e-12, "Precision": "fp64", "Preconditioner": "jacobi", "CheckInterval": 128})
            step(truth)
            xt = truth.export_system()["x"][lab]
            truth.close()
            truth_src = "gpu FP64 Jacobi-PCG to 1e-12"
        truth_x = (lab, xt)

        def perr(ctx):
            x = ctx.export_system()["x"][lab]
            return float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt))

        div_before = ps.divergence(1.0, vel0, svel, fp64_update=False)
        div_after = ps.divergence(1.0, vel0, svel, fp64_update=True)
        p_rel = perr(ps)
        accuracy = {
            "truth": truth_src,
            "pressure_rel_l2": p_rel,
            "post_div_rel_fp64": div_after / div_before,
            "true_relres_inf": last["true_relres_inf"],
            "true_relres_l2": last["true_relres_l2"],
            "bars": {"pressure_rel_l2": PRESSURE_BAR, "post_div_rel_fp64": DIVERGENCE_BAR},
            "meets_bar": bool(p_rel <= PRESSURE_BAR and div_after / div_before <= DIVERGENCE_BAR),
        }

        def bar_point(base_cfg, stops=BAR_STOPS, k=3):
            """The loosest max-norm stop of `stops` at which the solver meets both bars,
            with its solves/s (1 warm-up + k timed solves)."""
            for tol in stops:
                ctx = make(dict(base_cfg, Residual=tol))
                r = step(ctx)
                e = perr(ctx)
                if e <= PRESSURE_BAR and r["true_relres_inf"] <= DIVERGENCE_BAR and r["status"] == 0:
                    sps, r = timed(ctx, k)
                    out = {"residual": tol, "solves_per_s": sps, "iters": r["count"], "outer": r["outer"],
                           "pressure_rel_l2": perr(ctx), "post_div_rel_fp64": r["true_relres_inf"],
                           "meets_bar": True}
                    ctx.close()
                    return out
                ctx.close()
            return {"meets_bar": False, "stops_tried": list(stops)}

        accuracy["bar_point"] = bar_point(cfg)
        log(f"accuracy: {accuracy}")

    # ---- the north star's comparison: plain FP64 GPU CG (the reference's iteration),
    # multigrid PCG and the learned preconditioner in flexible CG, each at the bench
    # tolerance and at the loosest stop meeting the accuracy bars
    north = {}
    plain_sps = None
    if not args.no_north_star:
        pcg = make({"Residual": args.residual, "Precision": "fp64", "Preconditioner": "none", "CheckInterval": 64})
        step(pcg)
        plain_sps, rp = timed(pcg, 2)
        ent = {"solves_per_s": plain_sps, "iters": rp["count"], "true_relres_inf": rp["true_relres_inf"]}
        pcg.configure("ProfileKernels", 1)
        step(pcg)
        cells = rp["active_bricks"] * 512
        for k in ("spmv_dot", "axpy_norm"):
            cnt, ms = pcg.kernel_stats(k)
            if cnt:
                bpl = kernel_bytes_per_cell(k, "fp64", "none") * cells
                gbs = bpl / (ms / cnt * 1e-3) / 1e9
                ent[k] = {"avg_us": round(1e3 * ms / cnt, 3), "bytes_per_launch": bpl,
                          "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if truth_x is not None:
            ent["pressure_rel_l2"] = perr(pcg)
        if cpu_cg_dense is not None:
            # the same algorithm and stop on the CPU (the oracle's CG, as the reference runs it)
            lab0 = truth_x[0] if truth_x is not None else pcg.export_system()["label"].bool()
            xg = pcg.export_system()["x"][lab0]
            xc = torch.from_numpy(cpu_cg_dense).cuda().reshape(lab0.shape)[lab0]
            ent["pressure_rel_l2_vs_cpu_cg_same_stop"] = float(torch.linalg.norm(xg - xc) / torch.linalg.norm(xc))
            ent["cpu_cg_iters"] = cpu["iters"]
        pcg.close()
        north["fp64_plain_cg"] = ent
        north["headline_vs_fp64_plain_cg"] = round(value / plain_sps, 2)
        solvers = [("mg_pcg", {"Precision": "mixed", "Preconditioner": "mg"})]
        if Path(args.unet_weights).exists():
            solvers.append(("learned_fcg", {"Precision": "mixed", "Preconditioner": "unet",
                                            "UnetWeights": args.unet_weights}))
        for name, scfg in solvers:
            ctx = make(dict(scfg, Residual=args.residual))
            step(ctx)
            sps, r = timed(ctx, 3)
            e = {"solves_per_s": sps, "iters": r["count"], "outer": r["outer"], "fallback": r["fallback"],
                 "true_relres_inf": r["true_relres_inf"], "vs_fp64_plain_cg": round(sps / plain_sps, 2)}
            if truth_x is not None:
                e["pressure_rel_l2"] = perr(ctx)
                e["meets_bar"] = bool(e["pressure_rel_l2"] <= PRESSURE_BAR and r["true_relres_inf"] <= DIVERGENCE_BAR)
            if name == "learned_fcg":
                from shiokaze_amd.learned import flops_per_apply
                e["weights"] = Path(args.unet_weights).name
                ctx.configure("ProfileKernels", 1)
                step(ctx)
                cnt, ms = ctx.kernel_stats("unet_apply")
                if cnt:
                    fl = flops_per_apply((n, n, n))
                    tf = fl / (ms / cnt * 1e-3) / 1e12
                    e["net_apply_ms"] = round(ms / cnt, 4)
                    e["mfma"] = {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                                 "frac": round(tf / 2500.0, 4), "flops_per_apply": fl}
            ctx.close()
            if truth_x is not None:
                bp = bar_point(scfg)
                if bp.get("meets_bar"):
                    bp["vs_fp64_plain_cg"] = round(bp["solves_per_s"] / plain_sps, 2)
                e["bar_point"] = bp
            north[name] = e
        if accuracy.get("bar_point", {}).get("meets_bar"):
            accuracy["bar_point"]["vs_fp64_plain_cg"] = round(accuracy["bar_point"]["solves_per_s"] / plain_sps, 2)
        log(f"north star: {north}")

    extra = {}
    if args.extra:
        pj = make({"Residual": args.residual, "Precision": "fp64", "Preconditioner": "jacobi"})
        step(pj)
        sps, rj = timed(pj, 2)
        extra["fp64_jacobi_pcg"] = {"solves_per_s": sps, "iters": rj["count"]}
        pj.close()
        extra["projection_2d"] = bench_2d_extra(args, torch, shiokaze_amd)
        extra["csr_solve"] = csr_leg(args, torch, shiokaze_amd, make, step)
        log(f"csr solve: {extra['csr_solve']}")
    if args.liquid or args.extra:
        extra["liquid_step"] = liquid_leg(args, torch, shiokaze_amd, S, sc, dev, cfg)
        log(f"liquid step: {extra['liquid_step']}")

    prec_name = {"mixed": "fp32 stencil/axpy + fp64 dots/refinement", "fp64": "fp64"}[args.precision]
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if args.precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": (f"{'pool+sphere' if args.scene == 1 else 'irregular'} scene {n}^3, {prec_name}, "
                         f"{args.preconditioner} preconditioner, to {args.residual:g} relative max-norm residual"),
            "n": n, "scene": args.scene, "seed": args.seed, "unknowns": last["unknowns"],
            "active_bricks": last["active_bricks"], "pcg_iters": last["count"], "outer": last["outer"],
            "precision": args.precision, "preconditioner": args.preconditioner,
            "ms_setup": round(last["ms_setup"], 3), "ms_solve": round(last["ms_solve"], 3),
            "ms_update": round(last["ms_update"], 3),
            **({"fallback": last.get("fallback", 0), "set": args.set} if args.preconditioner == "unet" else {}),
        },
        "roofline": roofline,
        "kernels": kernels,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
        "north_star": north,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    ps.close()


def system_csr(torch, sysd):
    """The projection's FP64 system as CSR on the device (rows = unknowns in cell order):
    the diagonal and -c for each coupling between two unknowns, columns sorted."""
    lab = sysd["label"].bool().reshape(-1)
    nz, ny, nx = sysd["label"].shape
    row = torch.full((lab.numel(),), -1, dtype=torch.int64, device=lab.device)
    nrow = int(lab.sum())
    row[lab] = torch.arange(nrow, device=lab.device)
    cells = torch.nonzero(lab).reshape(-1)
    rs, cs, vs = [row[cells]], [row[cells]], [sysd["diag"].reshape(-1)[cells]]
    coord = (cells % nx, (cells // nx) % ny, cells // (nx * ny))
    for key, stride, q in (("cx", 1, coord[0]), ("cy", nx, coord[1]), ("cz", nx * ny, coord[2])):
        c = sysd[key].reshape(-1)
        hi = cells[q >= 1]  # the low face's coupling is stored at its high cell
        lo = hi - stride
        ok = (c[hi] != 0) & lab[lo]
        hi, lo = hi[ok], lo[ok]
        v = -c[hi]
        rs += [row[hi], row[lo]]
        cs += [row[lo], row[hi]]
        vs += [v, v]
    r, cidx, v = torch.cat(rs), torch.cat(cs), torch.cat(vs)
    order = torch.argsort(r * nrow + cidx)
    r, cidx, v = r[order], cidx[order], v[order]
    rowptr = torch.zeros(nrow + 1, dtype=torch.int64, device=lab.device)
    rowptr[1:] = torch.cumsum(torch.bincount(r, minlength=nrow), 0)
    b = sysd["rhs"].reshape(-1)[cells].contiguous()
    return rowptr, cidx.contiguous(), v.contiguous(), b


def csr_leg(args, torch, shiokaze_amd, make, step):
    """shk_solve_csr (the solve() entry for callers holding a CSR matrix) on the bench
    scene's FP64 system: Jacobi-PCG and the aggregation-multigrid PCG (its host setup
    included, as it runs on every solve), each to the bench tolerance."""
    ps = make({"Residual": args.residual, "Precision": "fp64", "Preconditioner": "jacobi"})
    step(ps)
    sysd = ps.export_system()
    rowptr, cols, vals, b = system_csr(torch, sysd)
    del sysd
    ps.close()
    out = {"rows": int(b.numel()), "nnz": int(vals.numel())}
    for name in ("jacobi", "mg"):
        s = shiokaze_amd.PCGSolver(config={"Residual": args.residual, "Preconditioner": name})
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.solve(rowptr, cols, vals, b)  # mg: builds the hierarchy
        torch.cuda.synchronize()
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, r = s.solve(rowptr, cols, vals, b)  # mg: the same matrix, the hierarchy reused
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s.close()
        out[name] = {"solves_per_s": round(1.0 / dt, 3), "iters": r["count"], "ms": round(1e3 * dt, 2),
                     "first_solve_ms": round(1e3 * first, 2), "true_relres_inf": r["true_relres_inf"],
                     "status": r["status"]}
    return out


def bench_2d_extra(args, torch, shiokaze_amd):
    """The 2D projection (macpressuresolver2) on a 2048^2 pool+disc scene: mixed Jacobi
    PCG and plain FP64 CG, with the 2D stencil kernel's roofline."""
    from shiokaze_amd import scene as S2
    sc2 = S2.generate2d(S2.POOL_SPHERE, 2048, 0)
    d2 = {k: torch.from_numpy(np.ascontiguousarray(sc2[k])).cuda() for k in ("fluid", "solid", "u", "v", "us", "vs")}
    v20 = [d2["u"].clone(), d2["v"].clone()]
    e2 = {"grid": "2048x2048 pool+disc"}
    for name, cfg2 in (("mixed_jacobi", {"Precision": "mixed", "Preconditioner": "jacobi"}),
                       ("fp64_plain_cg", {"Precision": "fp64", "Preconditioner": "none", "CheckInterval": 64})):
        p2 = shiokaze_amd.MacPressureSolver2((2048, 2048), sc2["dx"], config=dict(cfg2, Residual=args.residual))

        def step2():
            d2["u"].copy_(v20[0])
            d2["v"].copy_(v20[1])
            return p2.project(1.0, (d2["u"], d2["v"]), d2["solid"], (d2["us"], d2["vs"]), d2["fluid"])
        step2()
        t1 = time.perf_counter()
        r2 = step2()
        ent = {"solves_per_s": 1.0 / (time.perf_counter() - t1), "iters": r2["count"],
               "unknowns": r2["unknowns"], "true_relres_inf": r2["true_relres_inf"]}
        p2.configure("ProfileKernels", 1)
        step2()
        cnt, ms = p2.kernel_stats("spmv_dot")
        if cnt:
            bpl = kernel_bytes_per_cell("spmv_dot", cfg2["Precision"], cfg2["Preconditioner"]) * r2["active_bricks"] * 512
            gbs = bpl / (ms / cnt * 1e-3) / 1e9
            ent["spmv_dot"] = {"avg_us": round(1e3 * ms / cnt, 3), "bytes_per_launch": bpl,
                               "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        e2[name] = ent
        p2.close()
    e2["mixed_jacobi_vs_fp64_plain_cg"] = round(e2["mixed_jacobi"]["solves_per_s"]
                                                / e2["fp64_plain_cg"]["solves_per_s"], 2)
    return e2


def bench_2d(args, torch, shiokaze_amd, S):
    """The 2D projection (macpressuresolver2): the same step, roofline, accuracy and
    CPU-baseline legs on an n^2 scene (pool+disc or irregular)."""
    if args.preconditioner not in ("none", "jacobi", "mg"):
        print("the 2D projection runs Preconditioner none, jacobi or mg", file=sys.stderr)
        sys.exit(2)
    n = args.n
    sc = S.generate2d(args.scene, n, args.seed)
    dev = {k: torch.from_numpy(np.ascontiguousarray(sc[k])).cuda() for k in ("fluid", "solid", "u", "v", "us", "vs")}
    vel0 = [dev["u"].clone(), dev["v"].clone()]

    def make(config):
        return shiokaze_amd.MacPressureSolver2((n, n), sc["dx"], config=config)

    def step(ps):
        dev["u"].copy_(vel0[0])
        dev["v"].copy_(vel0[1])
        return ps.project(1.0, (dev["u"], dev["v"]), dev["solid"], (dev["us"], dev["vs"]), dev["fluid"])

    cfg = {"Residual": args.residual, "Precision": args.precision, "Preconditioner": args.preconditioner}
    for kv in args.set:
        k, v = kv.split("=", 1)
        cfg[k] = v
    ps = make(cfg)
    for _ in range(args.warmup):
        step(ps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        last = step(ps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    log(f"2D timed: {args.steps} steps in {elapsed:.3f}s -> {value:.3f} solves/s, iters={last['count']}")
    cells = last["active_bricks"] * 512
    kernels, roofline = {}, None
    if not args.no_profile:
        ps.configure("ProfileKernels", 1)
        for _ in range(args.profile_steps):
            step(ps)
        for k in ("spmv_dot", "axpy_norm"):
            cnt, ms = ps.kernel_stats(k)
            if cnt:
                bpl = kernel_bytes_per_cell(k, args.precision, args.preconditioner) * cells
                gbs = bpl / (ms / cnt * 1e-3) / 1e9
                kernels[k] = {"launches": cnt, "avg_us": round(1e3 * ms / cnt, 3), "bytes_per_launch": bpl,
                              "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "total_ms": round(ms, 3)}
        if kernels:
            dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
            kd = kernels[dom]
            roofline = {"bound": "hbm", "kernel": dom, "achieved": kd["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": kd["frac"], "traffic": None, "bytes_per_launch": kd["bytes_per_launch"],
                        "avg_launch_us": kd["avg_us"]}
    accuracy = {}
    if not args.no_accuracy:
        sysb = ps.export_system()
        lab = sysb["label"].bool()
        svel = (dev["us"], dev["vs"])
        div_before = ps.divergence(1.0, vel0, svel, fp64_update=False)
        div_after = ps.divergence(1.0, vel0, svel, fp64_update=True)
        truth = make({"Residual": 1e-12, "Precision": "fp64", "Preconditioner": "jacobi", "CheckInterval": 128})
        step(truth)
        xt = truth.export_system()["x"][lab]
        xb = sysb["x"][lab]
        accuracy = {"pressure_rel_l2_vs_fp64_1e-12": float(torch.linalg.norm(xb - xt) / torch.linalg.norm(xt)),
                    "post_div_rel_fp64": div_after / div_before, "true_relres_inf": last["true_relres_inf"]}
        truth.close()
    # plain FP64 CG on the GPU: the reference's iteration count (for the CPU extrapolation)
    plain = make({"Residual": args.residual, "Precision": "fp64", "Preconditioner": "none", "CheckInterval": 64})
    step(plain)
    t1 = time.perf_counter()
    rp = step(plain)
    plain_sps = 1.0 / (time.perf_counter() - t1)
    plain.close()
    extra = {"fp64_plain_cg": {"solves_per_s": plain_sps, "iters": rp["count"]},
             "vs_fp64_plain_cg": round(value / plain_sps, 2)}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.set_threads(1)
        tk = {}
        for k in (10, 30):
            ta = time.perf_counter()
            O.project2d(sc, tol=args.residual, max_iterations=k, mode=O.SHK_OR_REFERENCE)
            tk[k] = time.perf_counter() - ta
        per_it = (tk[30] - tk[10]) / 20.0
        base = max(tk[10] - 10 * per_it, 0.0)
        total = base + rp["count"] * per_it
        cpu = {"value": 1.0 / total, "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": (f"{n}^2 2D scene {args.scene}: the oracle's 2D projection with the reference-faithful CG "
                          f"(MIC(0) factor, two discarded triangular solves per iteration) run for 10 and 30 "
                          f"iterations and extrapolated to {rp['count']} (the plain CG count); "
                          f"{1e3 * per_it:.2f} ms/iteration, single thread"),
               "ms_per_solve": round(1e3 * total, 1)}
    prec_name = {"mixed": "fp32 stencil/axpy + fp64 dots/refinement", "fp64": "fp64"}[args.precision]
    out = {
        "metric": "pressure solves/sec of the 2D projection to 1e-6 rel. residual; PCG iters; % HBM peak",
        "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64" if args.precision == "mixed" else "fp64", "data": "synthetic",
        "config": {"workload": (f"2D {'pool+disc' if args.scene == 1 else 'irregular'} scene {n}^2, {prec_name}, "
                                f"{args.preconditioner} preconditioner, to {args.residual:g} relative max-norm residual"),
                   "n": n, "dim": 2, "scene": args.scene, "seed": args.seed, "unknowns": last["unknowns"],
                   "active_bricks": last["active_bricks"], "pcg_iters": last["count"], "outer": last["outer"],
                   "precision": args.precision, "preconditioner": args.preconditioner},
        "roofline": roofline, "kernels": kernels, "accuracy": accuracy, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(out), flush=True)
    ps.close()


if __name__ == "__main__":
    main()
