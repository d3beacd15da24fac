#!/usr/bin/env python3
"""Throughput of the 2D EM-PIC step on one MI355X.

Workload: the two-stream configuration at scale (2048x2048 periodic cells,
dx = 0.5 d_e, dt = 0.2/w_pe, 100 electrons per cell in two counter-streaming
beams at +-0.2c with vth = 0.02c, immobile ions), i.e. 419,430,400 particles,
in both forms: FP32 split precision (the headline value) and FP64.  A step is
the reference time-loop body (pic2d.f90:171-188): zero J, push (gather + Boris
+ move + zigzag deposit + wrap), field solve, and the counting sort every 50
steps.  Particles start resident in HBM (device-side seeded loader), sorted
once as the reference does at start-up (:161).

Prints one JSON line.  --gpus N > 1 is not supported on this single-GPU path.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s FP32-split vs FP64 on 1 MI355X; adjoint/forward time ratio"
COMPAT_MEANING = {
    7: "reference behaviours on: Boris rotation with the pre-kick u (pic2d.f90:985-987), "
       "mirrored rho weights (:1285-1295, rho not deposited here), jy fold column mx+2 dropped "
       "(:1156,1159), which breaks charge conservation at the x seam",
    0: "textbook Boris, node-CIC rho, every jy column folded (charge-conserving)"}
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(precision, cfg, steps, warmup, sort_every, seed, push_order=True, compat=7):
    """Times `steps` steps of one form.  The step counter is set (as a restart
    from a checkpoint at that step would set it, pbpic_set_counters) so that
    the timed window holds the sort of the reference cadence: its first sort
    falls min(steps, sort_every) // 2 steps into the window and sort_every
    steps after the previous sort (the untimed steps before the window are
    the warm-up plus what that takes), so it sorts a full cycle of drift, as
    in a long run.  A window of a whole number of sort cycles holds exactly
    steps / sort_every sorts; any other holds one more or one fewer, and
    ms_per_step then charges the sorts at the cadence: the measured mean sort
    time x steps / sort_every in place of the sorts that ran
    (ms_per_step_wall is the raw wall clock)."""
    import torch
    import pbpic_amd

    sp = cfg["species"][0]
    ctx = pbpic_amd.Context(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"], cfg["cva"],
                            cfg["mime"], cfg["in0"], precision=precision, compat=compat,
                            sort_every=sort_every, deposit_rho=False, push_order=push_order,
                            rho_background=cfg["rho_background"])
    npart = cfg["nx"] * cfg["ny"] * sp["ppc"]
    k = ctx.add_species(sp["q"], npart)
    ctx.init_maxwellian(k, sp["ppc"], sp["vth"], sp["vdrift"], counter_stream=sp["counter"],
                        seed=seed, stream=0)
    ctx.countingsort()                                  # pic2d.f90:161
    lead = min(steps, sort_every) // 2 + 1 if sort_every > 0 else 0
    if sort_every > 0:
        # untimed steps: at least `warmup`, and as many more as put the
        # in-window sort sort_every steps after the last one (the initial
        # sort or one in the warm-up), so that it sorts a full cycle's drift
        pre = warmup + (sort_every - lead - warmup) % sort_every
    else:
        pre = warmup
    ctx.step(pre)
    if sort_every > 0:
        nstep, je = ctx.counters()
        base = nstep - nstep % sort_every + sort_every - lead
        ctx.set_counters(base, je)
    d0 = ctx.diag()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    d1 = ctx.diag()
    ctx.close()
    wall = t1 - t0
    sort_ms, nsorts = prof["sort"]
    out = dict(npart=npart, wall=wall, prof=prof, diag0=d0, diag1=d1, sorts_in_window=nsorts,
               ms_per_step_wall=1e3 * wall / steps, untimed_steps=pre)
    if sort_every > 0:
        if nsorts == 0:
            raise SystemExit(f"timed window of {steps} steps holds no sort (sort_every="
                             f"{sort_every}): the step it reports would not be the stated step")
        per_sort = sort_ms / nsorts
        ms = (1e3 * wall - sort_ms + per_sort * steps / sort_every) / steps
        out.update(sort_ms_per_sort=per_sort, sorts_charged=steps / sort_every)
    else:
        ms = 1e3 * wall / steps
    out.update(ms_per_step=ms, ps=npart / (ms * 1e-3))
    return out


# HBM bytes per push launch at 2048x2048 x 100 ppc, the mean over a 50-step
# sort cycle: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a
# 4 GiB copy), from the commands in
# profiles/push_kernel_traffic_two_stream_2048x2048_ppc100.md.
TRAFFIC_2048_PPC100 = {1: 10.266e9 + 10.576e9, 0: 18.721e9 + 19.391e9}
TRAFFIC_SOURCE = "profiles/push_kernel_traffic_two_stream_2048x2048_ppc100.md"


def roofline(precision, res, cfg):
    """Push kernel: algorithmic bytes per launch = particles x (read + write of
    one particle record) + cells x (6 staggered field planes read, FP64, + 3
    current planes flushed as 64-bit fixed-point sums; the FP64 form's second
    word per node is left out, as overhead of its exact summation)."""
    sp = 44 if precision == 0 else 24          # i32 cell + 5 x f64 | 5 x f32
    wj = 8
    ncell = cfg["nx"] * cfg["ny"]
    bytes_launch = res["npart"] * 2 * sp + ncell * (6 * 8 + 3 * wj)
    ms, cnt = res["prof"]["push"]
    avg_s = ms / cnt / 1e3
    achieved = bytes_launch / avg_s / 1e9
    at_scale = (cfg["nx"], cfg["ny"], cfg["species"][0]["ppc"]) == (2048, 2048, 100)
    traffic = TRAFFIC_2048_PPC100[precision] if at_scale else None
    return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=achieved / HBM_PEAK_GBS, traffic=traffic,
                traffic_source=TRAFFIC_SOURCE if at_scale else None, kernel="push_kernel",
                bytes_per_launch=bytes_launch, avg_ms=avg_s * 1e3)


def step_roofline(precision, res, cfg, sort_every):
    """Whole step against B_alg = 2 S_p + B_cell/ppc + B_sort/K_sort."""
    ppc = cfg["species"][0]["ppc"]
    sp, bcell, bsort = (44, 192, 96) if precision == 0 else (24, 96, 56)
    balg = 2 * sp + bcell / ppc + bsort / sort_every
    gbs = balg * res["ps"] / 1e9
    return dict(bytes_per_particle_step=balg, achieved=gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=gbs / HBM_PEAK_GBS)


def adjoint_leg(precision, seed, nsteps=200, ckpt=20):
    """The differentiable-PIC configuration: 256x256 periodic thermal plasma
    with the CPU-case parameters, 32 particles/cell/species, Ey = A sin(2 pi
    x/Lx) with A = 0.01, electron drift u_x = 0.05c, loss = integrated J.E
    over 200 steps, checkpoints every 20 steps.  Times the plain forward (the
    step loop) against the checkpointed forward and the adjoint sweep."""
    import numpy as np
    import torch
    import pbpic_amd
    from pbpic_amd import configs

    cfg = configs.thermal_c0(256, 256, 32)
    x = np.arange(cfg["nx"])[None, :] * np.ones((cfg["ny"], 1))
    ey = 0.01 * np.sin(2 * np.pi * x / cfg["nx"])

    def make():
        ctx = pbpic_amd.Context(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"], cfg["cva"],
                                cfg["mime"], cfg["in0"], precision=precision, sort_every=50,
                                deposit_rho=False)
        for k, sp in enumerate(cfg["species"]):
            ctx.add_species(sp["q"], cfg["nx"] * cfg["ny"] * sp["ppc"])
            ctx.init_maxwellian(k, sp["ppc"], sp["vth"], 0.05 if sp["q"] < 0 else 0.0,
                                seed=seed, stream=k)
        ctx.set_field("ey", ey)
        ctx.sync()
        return ctx

    ctx = make()
    ctx.step(5)                        # warm the kernels up on a throw-away copy
    ctx.close()
    ctx = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(nsteps)
    ctx.sync()
    t_fwd = time.perf_counter() - t0
    ctx.close()

    ctx = make()
    npart = [ctx.num_particles(k) for k in range(len(cfg["species"]))]
    pdt = torch.float64 if precision == 0 else torch.float32
    fg = torch.zeros(6, cfg["ny"], cfg["nx"], dtype=torch.float64, device="cuda")
    pg = [torch.zeros(n, dtype=pdt, device="cuda") for n in npart for _ in range(5)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = ctx.adjoint_forward(nsteps, ckpt)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.adjoint_backward([fg[k].data_ptr() for k in range(6)], [t.data_ptr() for t in pg])
    ctx.sync()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ctx.close()
    gA = float((fg[1].cpu().numpy() * (ey / 0.01)).sum())     # dL/dA through Ey = A sin
    return dict(forward_s=t_fwd, ckpt_forward_s=t1 - t0, adjoint_s=t2 - t1,
                ratio=(t2 - t1) / t_fwd, ratio_total=(t2 - t0) / t_fwd, loss=loss, dloss_dA=gA,
                particles=sum(npart), steps=nsteps, ckpt_every=ckpt)


def harris_leg(precision, steps, warmup, seed):
    """The Harris current-sheet configuration at its full size: 1024 x 512
    domain cells (plus two guard rows) over 25.6 x 12.8 d_i with conducting y
    walls, 64 particles per cell and species (67,108,864 particles), the psi
    perturbation, sort every 50 steps.  Returns particle-steps/s and the
    total-energy change over the timed steps."""
    import torch
    import pbpic_amd
    from pbpic_amd import configs

    cfg = configs.harris_c1()
    h = dict(cfg["harris"], seed=seed)
    n = h["ppc"] * cfg["nx"] * (cfg["ny"] - 2)
    ctx = pbpic_amd.Context(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"], cfg["cva"],
                            cfg["mime"], cfg["in0"], precision=precision, sort_every=50,
                            deposit_rho=False, ybound="conducting")
    si = ctx.add_species(1, n)
    se = ctx.add_species(-1, n)
    ctx.init_harris(si, se, **h)
    ctx.countingsort()
    ctx.step(warmup)
    d0 = ctx.diag()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    d1 = ctx.diag()
    ctx.close()
    tot = lambda d: d["ke_e"] + d["ke_i"] + d["w_e"] + d["w_b"]
    return dict(particles=2 * n, ps=2 * n * steps / wall, ms_per_step=1e3 * wall / steps,
                energy_change=(tot(d1) - tot(d0)) / tot(d0), tiles_spilled=d1["tiles_spilled"])


def adjoint_c4_leg(seed, nsteps=1000, ckpt=50, precision=1):
    """The large differentiable configuration: the Harris sheet at 1024 x 1024
    domain cells (conducting y walls), 64 particles per cell and species
    (134,217,728 particles), 1000 steps, loss = J.E over the central 4 x 2 d_i
    diffusion region, gradient with respect to the initial Bz (and every
    other initial field and particle), checkpoints every 50 steps, split
    precision.  Times the plain forward against the checkpointed forward
    (which ends by allocating the backward's tape) and the adjoint sweep,
    and reports the device memory free before them."""
    import numpy as np
    import torch
    import pbpic_amd
    from pbpic_amd import configs

    cfg = configs.harris_c4()
    h = dict(cfg["harris"], seed=seed)
    n = h["ppc"] * cfg["nx"] * (cfg["ny"] - 2)

    def make():
        ctx = pbpic_amd.Context(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"],
                                cfg["cva"], cfg["mime"], cfg["in0"], precision=precision,
                                sort_every=50, deposit_rho=False, ybound="conducting")
        si = ctx.add_species(1, n)
        se = ctx.add_species(-1, n)
        ctx.init_harris(si, se, **h)
        ctx.countingsort()
        ctx.sync()
        return ctx

    ctx = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(nsteps)
    ctx.sync()
    t_fwd = time.perf_counter() - t0
    ctx.close()
    torch.cuda.empty_cache()

    ctx = make()
    ctx.set_loss_region(configs.loss_region(cfg))
    pdt = torch.float64 if precision == 0 else torch.float32
    fg = torch.zeros(6, cfg["ny"], cfg["nx"], dtype=torch.float64, device="cuda")
    pg = [torch.zeros(n, dtype=pdt, device="cuda") for _ in range(2) for _ in range(5)]
    free0 = torch.cuda.mem_get_info()[0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = ctx.adjoint_forward(nsteps, ckpt)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.adjoint_backward([fg[k].data_ptr() for k in range(6)], [t.data_ptr() for t in pg])
    ctx.sync()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ctx.close()
    gbz = fg[5].cpu().numpy()
    del fg, pg
    torch.cuda.empty_cache()
    return dict(particles=2 * n, steps=nsteps, ckpt_every=ckpt, forward_s=t_fwd,
                ckpt_forward_s=t1 - t0, adjoint_s=t2 - t1, ratio=(t2 - t1) / t_fwd,
                ratio_total=(t2 - t0) / t_fwd, loss=loss,
                dloss_dbz_l2=float(np.sqrt((gbz ** 2).sum())),
                dloss_dbz_maxabs=float(np.abs(gbz).max()),
                device_free_gb_before=free0 / 1e9)


def cpu_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when the environment sets
    it (the share of host cores given to one GPU), else every core this
    process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(budget_s, seed, nx=2048, ny=2048, ppc=100, sort_every=50):
    """The oracle (CPU restatement of pic2d.f90, AoS `particle` records in the
    reference's default FP32) on the benchmark workload itself: the two-stream
    configuration at nx x ny x ppc.  It steps for about budget_s seconds
    (a bounded sample of the 500-step run), then one counting sort is timed
    and charged at 1/sort_every per step, as the reference sorts every 50
    steps (pic2d.f90:180-182).  The push is threaded over contiguous particle
    chunks with private current arrays, the sort over contiguous chunks with
    per-chunk counts (the serial loop's result)."""
    import numpy as np
    import oracle
    from pbpic_amd import configs

    threads = cpu_threads()
    cfg = configs.two_stream_c2(nx, ny, ppc)
    sp = cfg["species"][0]
    t_init = time.perf_counter()
    p = oracle.init_maxwellian(cfg["nx"], cfg["ny"], sp["ppc"], sp["vth"], sp["vdrift"],
                               counter_stream=True, seed=seed)
    n = len(p["grid"])
    o = oracle.Oracle(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"], cfg["cva"],
                      cfg["mime"], cfg["in0"], prec=4, threads=threads)
    o.set_particles(p["grid"], p["dx"], p["dy"], p["ux"], p["uy"], p["uz"],
                    np.full(n, -1, np.int32))
    del p
    t_init = time.perf_counter() - t_init
    # the loader writes cell order, which is the sorted order (pic2d.f90:161)
    o.step(1, sort_every=0)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 500:
        o.step(1, sort_every=0)
        steps += 1
    t_steps = time.perf_counter() - t0
    t1 = time.perf_counter()
    o.countingsort()
    t_sort = time.perf_counter() - t1
    o.close()
    per_step = t_steps / steps + t_sort / sort_every
    return dict(value=n / per_step, unit="particle-steps/s", cores=threads, kind="port",
                host_cpus=os.cpu_count(), host_cpus_usable=len(os.sched_getaffinity(0)),
                s_per_step=t_steps / steps, s_per_sort=t_sort, init_s=t_init,
                sample=f"two-stream {nx}x{ny} x {ppc} ppc ({n} electrons), FP32 (reference "
                       f"default real): {steps} timed steps of the 500-step run plus one timed "
                       f"counting sort charged at 1/{sort_every} per step, {threads} OpenMP "
                       f"threads")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--ppc", type=int, default=100)
    ap.add_argument("--sort-every", type=int, default=50)
    ap.add_argument("--order", choices=("push", "cell"), default="push",
                    help="particle order the sort writes: the push kernel's lane order inside "
                         "each 4096-particle tile (default) or the reference's cell order")
    ap.add_argument("--forms", default="split,fp64")
    ap.add_argument("--compat", type=int, default=7, choices=(0, 7),
                    help="reference behaviours: 7 = all on (the reference's own arithmetic), "
                         "0 = textbook Boris, unmirrored rho, no dropped jy column")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="stepping budget of the CPU baseline on the full workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-harris", action="store_true")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the 1024x1024 Harris adjoint leg (about 240 GB of HBM)")
    ap.add_argument("--c4-steps", type=int, default=1000,
                    help="steps of the 1024x1024 Harris adjoint leg (its configuration: 1000)")
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--calib-copy-gib", type=float, default=0.0,
                    help="first run one plain device copy of this many GiB (a known-size "
                         "kernel that calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the PIC step "
                                   "is not sharded across GPUs in this build"}))
        sys.exit(2)

    import torch
    from pbpic_amd import configs
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    cfg = configs.two_stream_c2(args.nx, args.ny, args.ppc)
    if args.calib_copy_gib > 0:
        nb = int(args.calib_copy_gib * (1 << 30))
        a = torch.ones(nb, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {nb} bytes done")
        del a, b
        torch.cuda.empty_cache()
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    for f in forms:
        prec = 1 if f == "split" else 0
        log(f"[bench] form={f} nx={args.nx} ny={args.ny} ppc={args.ppc} steps={args.steps} "
            f"warmup={args.warmup}")
        res[f] = run_form(prec, cfg, args.steps, args.warmup, args.sort_every, args.seed,
                          push_order=args.order == "push", compat=args.compat)
        r = res[f]
        log(f"[bench] form={f}: {r['ps']:.4e} particle-steps/s, {r['ms_per_step']:.3f} ms/step, "
            f"kernels(ms,count)={r['prof']}")
    head = "split" if "split" in res else forms[0]
    hp = 1 if head == "split" else 0
    out = {
        "metric": METRIC,
        "value": res[head]["ps"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "split" else "f64",
        "data": "synthetic (device-side seeded loader)",
        "config": {"workload": f"two-stream {args.nx}x{args.ny}, {args.ppc} electrons/cell "
                               f"({res[head]['npart']} particles), K_sort={args.sort_every}, "
                               f"form={head}",
                   "nx": args.nx, "ny": args.ny, "ppc": args.ppc,
                   "sort_every": args.sort_every, "particle_order": args.order,
                   "compat": args.compat, "compat_meaning": COMPAT_MEANING[args.compat],
                   "parallelism": "single GPU"},
        "timing": {"ms_per_step_wall": res[head]["ms_per_step_wall"],
                   "sorts_in_window": res[head]["sorts_in_window"],
                   "sorts_charged": res[head].get("sorts_charged"),
                   "sort_ms_per_sort": res[head].get("sort_ms_per_sort"),
                   "rule": "ms_per_step = (wall - in-window sort time + mean sort time x "
                           "steps/sort_every) / steps; the window's first sort falls "
                           "min(steps, sort_every)//2 + 1 steps in"},
        "roofline": roofline(hp, res[head], cfg),
        "step_roofline": step_roofline(hp, res[head], cfg, args.sort_every),
        "kernel_ms": {k: (v[0] / max(v[1], 1), v[1]) for k, v in res[head]["prof"].items()},
    }
    if "fp64" in res and head != "fp64":
        r = res["fp64"]
        out["fp64"] = {"value": r["ps"], "ms_per_step": r["ms_per_step"],
                       "ms_per_step_wall": r["ms_per_step_wall"],
                       "sorts_in_window": r["sorts_in_window"],
                       "sort_ms_per_sort": r.get("sort_ms_per_sort"),
                       "roofline": roofline(0, r, cfg), "step_roofline": step_roofline(0, r, cfg, args.sort_every),
                       "kernel_ms": {k: (v[0] / max(v[1], 1), v[1]) for k, v in r["prof"].items()}}
        out["ratio_split_vs_fp64"] = res[head]["ps"] / r["ps"]
    out["adjoint_forward_ratio"] = None
    if not args.no_adjoint:
        adj = {}
        for f in forms:
            log(f"[bench] adjoint leg form={f}")
            adj[f] = adjoint_leg(1 if f == "split" else 0, args.seed)
            log(f"[bench] adjoint form={f}: {adj[f]}")
        out["adjoint"] = dict(workload="differentiable PIC: 256x256 thermal plasma, 32 ppc/species, "
                                       "200 steps, checkpoint every 20, loss = integrated J.E",
                              **adj)
        # the larger of the forms' adjoint-sweep / plain-forward time ratios
        out["adjoint_forward_ratio"] = max(a["ratio"] for a in adj.values())
        if not args.no_c4:
            log(f"[bench] adjoint leg: Harris 1024x1024, {args.c4_steps} steps, split")
            c4 = adjoint_c4_leg(args.seed, nsteps=args.c4_steps)
            log(f"[bench] adjoint C4: {c4}")
            out["adjoint_c4"] = dict(workload="differentiable PIC, large: Harris sheet 1024x1024 "
                                              "(+2 guard rows), conducting y walls, 64 ppc/species, "
                                              f"{args.c4_steps} steps, checkpoints every 50, loss = "
                                              "J.E over the central 4x2 d_i, split precision", **c4)
            out["adjoint_forward_ratio"] = max(out["adjoint_forward_ratio"], c4["ratio"])
        out["adjoint_forward_ratio_over"] = ("256x256 thermal (both forms)" +
                                             ("" if args.no_c4 else " and 1024x1024 Harris (split)"))
    if not args.no_harris:
        hr = {}
        for f in forms:
            log(f"[bench] Harris-sheet leg form={f}")
            hr[f] = harris_leg(1 if f == "split" else 0, 100, 10, args.seed)
            log(f"[bench] Harris form={f}: {hr[f]}")
        out["harris_c1"] = dict(workload="Harris sheet 1024x512 (+2 guard rows), 25.6x12.8 d_i, "
                                         "conducting y walls, 64 ppc/species, psi perturbation, "
                                         "100 timed steps, K_sort=50", **hr)
        if "split" in hr and "fp64" in hr:
            out["harris_c1"]["ratio_split_vs_fp64"] = hr["split"]["ps"] / hr["fp64"]["ps"]
    out["diag"] = {f: {k: res[f]["diag1"][k] for k in ("ke_e", "w_e", "w_b", "je_int", "gauss_dlinf", "tiles_spilled")}
                   for f in res}
    if not args.no_cpu:
        log(f"[bench] cpu baseline ({args.cpu_seconds:.0f} s budget)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, args.seed, args.nx, args.ny,
                                           args.ppc, args.sort_every)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
