"""Benchmark: leaf-cell RK3 updates/s, learned selector + FP32 flux vs plain FP64, one MI355X.

A step is one pass of the reference's solute loop body (solution.F90:183-215): λ (running
max), dt from CFL, three SSP-RK3 stages (level-jump slot values, fused flux + update, ghost
fill) and the residual, all on the device.

The workload runs as its spec says: the 4-bubble shock on a 1024x512 base, level 4, with a
regrid every 10 steps (solution.F90:408-421). The plain FP64 path (the reference's
arithmetic, bit-exact with the CPU oracle) and the learned path advance side by side in
cycles of 10 steps; after each cycle the FP64 context regrids on its own sensors and the
learned context regrids on the same sensors, so both stay on identical meshes and L1/L∞(ρ)
between them measures the learned path's error alone. `value` = updates/s of the learned
path over the timed steps (K, in cycles that start right after a regrid, so the steps after
each regrid, whose copied values make many leaves troubled, are included); regrids are timed
apart and folded into `value_regrid_amortised`. The learned path: per step the troubled-cell
network runs on the step's state (BF16 MFMA), cleared leaves take the linear-weight FP32
reconstruction, troubled ones the full characteristic WCNS3 in FP32, the substep and RK blend
are FP64, and a 2% sample of cleared leaves is re-checked by the full method (the judge).

Selector training follows the workload spec: pre-trained on ~20 M labelled leaves dumped from
plain runs of the vortex and 4-quadrant Riemann configurations (with their regrids), then
fine-tuned online on the target workload after a cycle whose judge rejection rate exceeds 1%.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fdm-cell-based-amr-cartesian_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_config(name: str, precision: int, selector: int = 0, refine_goal: float = 1.4,
                    coarse_goal: float = None):
    """Solver configuration of a named workload. refine_goal / coarse_goal: the regrid's
    refinement and coarsening goals (multiples of the sensor's RMS, solution.F90:519-535);
    coarse_goal None keeps the solver's default."""
    import fdmamr as F

    cg = {} if coarse_goal is None else dict(coarse_goal_rot=coarse_goal, coarse_goal_div=coarse_goal)

    if name == "vortex":
        # isentropic vortex, periodic [0,10]^2, base 64x64, lvlAMR 2 (the reference's criterion)
        return F.Config(h=10.0 / 64, localdt=1, cfl=0.5, amr=1, lvl_amr=2, pstep_adapt=30,
                        t_end=1e9, pre_adapt=2, flux_precision=precision, selector=selector)
    if name == "riemann4":
        # seeded 4-quadrant Riemann problem, [0,1]^2, base 256x256, lvl <= 4 (density gradient)
        return F.Config(xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, h=1.0 / 256, gamma=1.4,
                        bc=(F.BC_TRANSMISSIVE,) * 4, ic_kind=F.IC_RIEMANN4, seed=1,
                        ic_param=(0.5, 0.5), localdt=1, cfl=0.5, t_end=0.2, amr=1, lvl_amr=4,
                        pstep_adapt=10, amr_sensor=1, pre_adapt=4, wls_local=1,
                        refine_goal_rot=refine_goal, refine_goal_div=refine_goal,
                        flux_precision=precision, selector=selector, **cg)
    if name == "shock4bubbles":
        # Mach 3 shock through 4 seeded bubbles, [0,4]x[0,2], base 1024x512, lvl <= 4
        g, M = 1.4, 3.0
        p2 = 1 + 2 * g / (g + 1) * (M * M - 1)
        r2 = (g + 1) * M * M / ((g - 1) * M * M + 2)
        u2 = M * np.sqrt(g) * (1 - 1 / r2)
        return F.Config(xmin=0.0, xmax=4.0, ymin=0.0, ymax=2.0, h=1.0 / 256, gamma=g,
                        bc=(F.BC_INFLOW, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE),
                        inflow=(r2, u2, 0.0, p2), ic_kind=F.IC_SHOCK_BUBBLES, seed=2,
                        ic_param=(0.5, M, 4, 0.05, 0.15, 0.1, 5.0), localdt=1, cfl=0.5,
                        t_end=1e9, amr=1, lvl_amr=4, pstep_adapt=10, amr_sensor=1, pre_adapt=4,
                        wls_local=1, refine_goal_rot=refine_goal, refine_goal_div=refine_goal,
                        flux_precision=precision, selector=selector, **cg)
    raise SystemExit(f"unknown workload {name}")


def _oracle_kwargs(cfg):
    """The oracle's parameters for a solver configuration (same mesh, IC and AMR settings)."""
    return dict(xmin=cfg.xmin, xmax=cfg.xmax, ymin=cfg.ymin, ymax=cfg.ymax, h=cfg.h, gamma=cfg.gamma,
                Ma=cfg.Ma, tstep=cfg.tstep, cfl=cfg.cfl, t_end=cfg.t_end, iflag_localdt=cfg.localdt,
                iflag_amr=cfg.amr, lvl_amr=cfg.lvl_amr, pstep_adapt=cfg.pstep_adapt,
                nstep_adapt=cfg.nstep_adapt, refine_goal_rot=cfg.refine_goal_rot,
                refine_goal_div=cfg.refine_goal_div, coarse_goal_rot=cfg.coarse_goal_rot,
                coarse_goal_div=cfg.coarse_goal_div, amr_sensor=cfg.amr_sensor, bc_xm=cfg.bc[0],
                bc_xp=cfg.bc[1], bc_ym=cfg.bc[2], bc_yp=cfg.bc[3], inflow=cfg.inflow,
                ic_kind=cfg.ic_kind, seed=cfg.seed,
                ic_param=tuple(cfg.ic_param) + (0.0,) * (8 - len(cfg.ic_param)),
                pre_adapt=cfg.pre_adapt, wls_local=cfg.wls_local)


def cpu_baseline(cfg5, tree, lam, budget_s: float = 8.0):
    """The reference's OpenMP CPU path as the CPU oracle restates it (FP64; the stage's leaf
    loop is an OpenMP dynamic-schedule loop over cell_NS, as solution.F90:279-284) on all the
    host's cores (OMP_NUM_THREADS, threads bound close):
    * the vortex configuration (base 64x64, level 2, regrid every 30 steps): whole steps;
    * a slice of the measured 4-bubble mesh (imported into the oracle): three RK3 stages over
      the cell_NS leaves [N/3, N/3 + n), n sized for about budget_s of work. `value` is this
      slice's rate, the measured workload's."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import oracle as O
    import fdmamr as F

    cores = O.set_threads(0)
    vcfg = workload_config("vortex", F.FLUX_FP64)
    o = O.Oracle(**_oracle_kwargs(vcfg))
    o.step()
    t0 = time.perf_counter()
    upd = steps = 0
    while time.perf_counter() - t0 < budget_s and steps < 100:
        upd += o.n_leaf
        assert o.step() == 0
        steps += 1
    dt_v = time.perf_counter() - t0
    vort = dict(value=upd / dt_v, unit="leaf-updates/s", leaves=o.n_leaf, steps=steps, seconds=round(dt_v, 2))
    o.close()
    t0 = time.perf_counter()
    o = O.Oracle(**_oracle_kwargs(cfg5))
    o.import_tree(tree)
    t_imp = time.perf_counter() - t0
    N = o.n_leaf
    i0, n = N // 3, min(N // 3, 100_000)
    t0 = time.perf_counter()
    o.rk3_range(i0, i0 + n, 1e-6, float(lam[0]), float(lam[1]))
    per = time.perf_counter() - t0
    n = int(min(N // 3, max(n, n * budget_s / max(per, 1e-6) / 2)))
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < budget_s:
        o.rk3_range(i0, i0 + n, 1e-6, float(lam[0]), float(lam[1]))
        reps += 1
    dt_s = time.perf_counter() - t0
    o.close()
    return dict(value=reps * n / dt_s, unit="leaf-updates/s", cores=cores, kind="port",
                sample=f"oracle (C, FP64, OpenMP dynamic over cell_NS, {cores} threads): 4-bubble mesh "
                       f"({N} leaves, imported in {t_imp:.0f}s), cell_NS leaves [{i0}, {i0 + n}) x {reps} "
                       f"RK3 steps in {dt_s:.1f}s",
                vortex=vort)


def measure_traffic(args, mesh_file):
    """HBM bytes of one stage's flux/update launches of the learned path, from two rocprofv3
    counter runs of a short copy of this benchmark (FETCH_SIZE, then WRITE_SIZE, in KiB per
    dispatch). Each counter is scaled by what it reads for a plain 256 MiB device copy in the
    same run (on MI355X FETCH_SIZE counts a 128-byte read as 64 bytes). Returns None when
    rocprofv3 is unavailable or a run fails. The runs load the measured mesh and state from
    mesh_file instead of building them again."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if prof is None:
        log("traffic: rocprofv3 not found")
        return None
    copy_mib = 256
    child = [sys.executable, os.path.abspath(__file__), "--steps", "2", "--warmup", "1",
             "--no-cpu", "--no-traffic", "--traffic-child",
             "--copy-calibration", str(copy_mib), "--workload", args.workload,
             "--refine-goal", str(args.refine_goal), "--develop-t", str(args.develop_t),
             "--mesh-file", mesh_file,
             "--threshold", str(args.threshold), "--label-bound", str(args.label_bound)]
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="fdmamr_pmc_")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--"] + child
        log(f"traffic: {counter} counter run")
        t0 = time.perf_counter()
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=400)
        except subprocess.TimeoutExpired:
            log(f"traffic: {counter} run timed out")
            return None
        log(f"traffic: {counter} run done in {time.perf_counter() - t0:.1f}s")
        info = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not info or not files:
            log(f"traffic: {counter} run failed (rc {r.returncode})")
            return None
        n_leaf = info[-1]["n_leaf"]
        # one launch per stage: 384 claimer blocks for the troubled leaves, one block per 256
        # cleared-leaf slots, 1536 claimer blocks (STAGE_EARLY, STAGE_LATE in fdmamr_device.hpp)
        grid = (384 + (n_leaf + 255) // 256 + 1536) * 256
        rows = sorted(csv.DictReader(open(files[0])), key=lambda x: int(x["Dispatch_Id"]))
        shutil.rmtree(d, ignore_errors=True)
        # the calibration copies are the run's first three copy dispatches (the solver's own
        # device-to-device copies come later)
        copy_kib = [float(x["Counter_Value"]) for x in rows if "copyBuffer" in x["Kernel_Name"]][:3]
        mark = max(i for i, x in enumerate(rows) if "k_positivity" in x["Kernel_Name"])
        stage = [x for x in rows[mark:] if "k_stage_mixed" in x["Kernel_Name"] and int(x["Grid_Size"]) == grid]
        n_stage = len(stage)
        if not copy_kib or not n_stage:
            log(f"traffic: {counter} dispatches not found")
            return None
        scale = 3 * (copy_mib << 20) / (sum(copy_kib) * 1024.0)
        res[counter] = dict(kib_per_stage=sum(float(x["Counter_Value"]) for x in stage) / n_stage,
                            copy_scale=scale, stages=n_stage)
    fetch = res["FETCH_SIZE"]["kib_per_stage"] * 1024 * res["FETCH_SIZE"]["copy_scale"]
    write = res["WRITE_SIZE"]["kib_per_stage"] * 1024 * res["WRITE_SIZE"]["copy_scale"]
    log(f"traffic per stage: read {fetch / 1e6:.1f} MB, write {write / 1e6:.1f} MB "
        f"(copy scales {res['FETCH_SIZE']['copy_scale']:.3f}, {res['WRITE_SIZE']['copy_scale']:.3f})")
    return {"bytes_per_stage": fetch + write, "read_bytes": fetch, "write_bytes": write,
            "fetch_copy_scale": res["FETCH_SIZE"]["copy_scale"],
            "write_copy_scale": res["WRITE_SIZE"]["copy_scale"], "stages": res["FETCH_SIZE"]["stages"]}


def clone(cfg, tree, prim):
    import fdmamr as F
    s = F.Solver(cfg)
    m, n = int(tree["num"][:, 1].max()), int(tree["num"][:, 0].max())
    s.import_tree(m, n, tree)
    s.build_tables()
    s.set_state(prim)
    return s


def pretrain_selector(seed, rows_target, iters, threshold, label_bound):
    """Pre-training as the selector's spec has it: labelled leaves dumped from plain FP32-flux
    runs of the 4-quadrant Riemann configuration (base 256x256, level 4, regrid every 10 steps)
    and of the vortex (base 64x64, level 2, regrid every 30), about rows_target rows, taken
    right after each regrid (copied values: many troubled leaves) and every second step after.
    Trained by the library's own MFMA forward/backward kernels, Adam, three learning rates."""
    import fdmamr as F
    t0 = time.perf_counter()
    trainer = F.Solver(workload_config("riemann4", F.FLUX_FP32)).setup()
    trainer.selector_config(threshold, 0.02, label_bound)
    trainer.selector_init(seed)
    vort = F.Solver(workload_config("vortex", F.FLUX_FP32)).setup()
    vort.selector_config(threshold, 0.02, label_bound)
    vort.selector_set_weights(trainer.selector_get_weights())
    rows = pos = 0
    vrows = 0
    for step in range(100):
        if step % 30 == 0 and step:
            vort.regrid()
        if step % 2 == 0 or step % 30 == 1:
            vrows, _ = vort.selector_collect()
        vort.run_steps(1)
    step = 0
    while rows < rows_target - vrows and trainer.tt < 0.2:
        if step % 10 == 0 and step:
            trainer.regrid()
        if step % 2 == 0 or step % 10 == 1:
            rows, p = trainer.selector_collect()
            pos += p
        trainer.run_steps(1)
        step += 1
    t_data = time.perf_counter() - t0
    # the two row sets live in their own contexts: each trains the same weights in turn
    losses = []
    for k, lr in enumerate((3e-3, 1e-3, 3e-4)):
        losses.append(trainer.selector_train(iters // 3, batch=8192, lr=lr, pos_weight=4.0, seed=seed + k))
        vort.selector_set_weights(trainer.selector_get_weights())
        vort.selector_train(max(iters // 30, 1), batch=8192, lr=lr, pos_weight=4.0, seed=seed + 10 + k)
        trainer.selector_set_weights(vort.selector_get_weights())
    w = trainer.selector_get_weights()
    log(f"selector pre-trained: {rows} Riemann rows ({pos} troubled) from {step} steps to t = {trainer.tt:.4f} "
        f"+ {vrows} vortex rows, loss {losses[0]:.4f} -> {losses[-1]:.4f}, data {t_data:.1f}s, "
        f"total {time.perf_counter() - t0:.1f}s")
    info = dict(rows=rows + vrows, riemann_rows=rows, riemann_troubled=pos, vortex_rows=vrows,
                riemann_steps=step, iters=iters, batch=8192, loss=losses[-1])
    trainer.close()
    vort.close()
    return w, info


def measured_config(args, precision, selector=0):
    import fdmamr as F
    dev = args.workload == "shock4bubbles" and args.develop_t > 0
    return workload_config(args.workload, precision, selector=selector, refine_goal=args.refine_goal,
                           coarse_goal=args.refine_goal / 3.0 if dev else None)


def build_mesh(args, F):
    """The FP64 context on the measured mesh and state, right after a regrid.

    shock4bubbles with --develop-t: the flow is evolved at a coarse maximum level (the
    standard goals, regrid every 40 steps) to a developed state (the shock through the
    bubbles, reflected waves and vortices), regridded --refine-passes times at level 4 with
    the measured goals (new leaves start from their parent's value, as the reference's
    refinement does), then run for --spinup-cycles cycles of --regrid-every steps, each
    ending in a regrid, until the mesh has settled. Otherwise: the workload's initial
    adaptation, then the same cycles."""
    t0 = time.perf_counter()
    if args.workload == "shock4bubbles" and args.develop_t > 0:
        dev = workload_config(args.workload, F.FLUX_FP64, refine_goal=0.3)
        dev.lvl_amr = dev.pre_adapt = args.develop_lvl
        s = F.Solver(dev).setup()
        steps = 0
        while s.tt < args.develop_t:
            s.run_steps(40)
            s.regrid()
            steps += 40
        log(f"developed at level {args.develop_lvl}: {steps} steps to t = {s.tt:.4f}, "
            f"{s.counts()['n_leaf']} leaves, {time.perf_counter() - t0:.1f}s")
        s.cfg = measured_config(args, F.FLUX_FP64)
        for _ in range(args.refine_passes):
            s.regrid()
    else:
        s = F.Solver(measured_config(args, F.FLUX_FP64)).setup()
    counts = []
    for _ in range(args.spinup_cycles):
        s.run_steps(args.regrid_every)
        s.regrid()
        counts.append(s.n_leaf)
    cnt = s.counts()
    log(f"workload {args.workload}: {cnt['n_leaf']} leaves (after each spin-up cycle: {counts}), "
        f"{cnt['n_slot_wls']} WLS slots, {cnt['n_slot_restrict']} restriction slots, l_max {cnt['l_max']}, "
        f"t = {s.tt:.4f}, setup {time.perf_counter() - t0:.1f}s")
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps, in cycles of --regrid-every steps that start right after a regrid")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps of both paths before the cycles")
    ap.add_argument("--workload", default="shock4bubbles")
    ap.add_argument("--refine-goal", type=float, default=0.03,
                    help="refinement goal of the measured mesh (its coarsening goal is a third); 0.03 "
                         "settles the 4-bubble mesh at about 19 M leaves")
    ap.add_argument("--develop-t", type=float, default=0.4,
                    help="shock4bubbles: evolve the flow to this time at --develop-lvl (regrid "
                         "every 40 steps) before refining to level 4; 0 measures on the initial "
                         "condition's mesh")
    ap.add_argument("--develop-lvl", type=int, default=1)
    ap.add_argument("--refine-passes", type=int, default=3,
                    help="regrids at level 4 after the development")
    ap.add_argument("--spinup-cycles", type=int, default=4,
                    help="cycles of --regrid-every FP64 steps and a regrid before the paired run")
    ap.add_argument("--regrid-every", type=int, default=10)
    ap.add_argument("--tune-cycles", type=int, default=8,
                    help="paired cycles before the timed ones; the selector is fine-tuned after "
                         "each whose judge rejection rate exceeds 1%%")
    ap.add_argument("--accuracy-cycles", type=int, default=10,
                    help="paired cycles after the timed ones, before L1/Linf(rho) is taken")
    ap.add_argument("--pretrain-rows", type=int, default=20_000_000)
    ap.add_argument("--train-iters", type=int, default=4000)
    ap.add_argument("--tune-iters", type=int, default=1000,
                    help="minibatches of each online fine-tune")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--threshold", type=float, default=0.5,
                    help="selector probability above which a leaf takes the full path")
    ap.add_argument("--label-bound", type=float, default=0.05,
                    help="largest |w1 - w1_lin| of a leaf that still counts as cleared")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of this many MiB (scale for HBM counters)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs that measure HBM traffic")
    ap.add_argument("--regrid-phases", action="store_true",
                    help="print the phase times of the learned context's last timed regrid (stderr)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--mesh-file", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus must be 1"}))
        raise SystemExit(2)

    if args.copy_calibration:
        # torch (when used) is loaded before the solver library, as in the tests
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 device copies of {n} bytes")
        del src, dst

    import fdmamr as F

    if args.mesh_file:
        # a counter run: the parent's mesh, state and fine-tuned weights
        z = np.load(args.mesh_file)
        tree = {k: z[k] for k in z.files if k not in ("prim", "weights")}
        sel = clone(measured_config(args, F.FLUX_FP32, selector=1), tree, z["prim"])
        sel.selector_config(args.threshold, 0.02, args.label_bound)
        sel.selector_set_weights(z["weights"])
        sel.run_steps(args.warmup)
        # marker dispatch for the counter parser, then the learned steps it measures
        sel.check_positivity()
        sel.run_steps(args.steps)
        print(json.dumps({"n_leaf": sel.n_leaf, "steps": args.steps}))
        return

    w, pre = pretrain_selector(7, args.pretrain_rows, args.train_iters, args.threshold, args.label_bound)
    s64 = build_mesh(args, F)
    sel = clone(measured_config(args, F.FLUX_FP32, selector=1), s64.tree(), s64.get_state())
    sel.selector_config(args.threshold, 0.02, args.label_bound)
    sel.selector_set_weights(w)
    sel.tt, sel.lam = s64.tt, s64.lam.copy()

    def run(s, n):
        t0 = time.perf_counter()
        s.run_steps(n)  # synchronises at the end
        return time.perf_counter() - t0

    def regrid_both(phases=False):
        # the FP64 context on its own sensors, the learned one on the same sensors (same mesh);
        # phases: the library prints the learned context's regrid phases (FDMAMR_TIMING)
        t0 = time.perf_counter()
        rot, div = s64.sensors()
        s64.regrid_with_sensors(rot, div)
        t64 = time.perf_counter() - t0
        if phases:
            os.environ["FDMAMR_TIMING"] = "1"
        t0 = time.perf_counter()
        try:
            sel.regrid_with_sensors(rot, div)
        finally:
            if phases:
                os.environ.pop("FDMAMR_TIMING", None)
        return t64, time.perf_counter() - t0

    def judge_rate(st):
        return st["rejected"] / max(st["judged"], 1)

    # warm-up steps, then every cycle starts right after a regrid
    run(s64, args.warmup)
    run(sel, args.warmup)
    regrid_both()
    E = args.regrid_every
    tune_log = []
    tuning = True  # the pre-trained network has not seen this workload
    rows = 0
    for cyc in range(args.tune_cycles):
        sel.selector_stats(reset=True)
        if tuning:
            # rows from the cycle's first two states (just after the regrid, and one step on),
            # added to those of earlier fine-tune cycles (up to about 150 M rows)
            if rows + 2 * sel.n_leaf > 150_000_000:
                sel.selector_clear_dataset()
            sel.selector_collect()
            run(sel, 1)
            rows, _ = sel.selector_collect()
            sel.selector_train(args.tune_iters, batch=8192, lr=5e-4, pos_weight=4.0, seed=100 + cyc)
            run(sel, E - 1)
        else:
            run(sel, E)
        run(s64, E)
        st = sel.selector_stats(reset=True)
        rate = judge_rate(st)
        tune_log.append(dict(cycle=cyc, tuned=tuning, rejection=round(rate, 5), leaves=sel.n_leaf))
        tuning = rate > 0.01
        regrid_both()
    sel.selector_clear_dataset()
    log("fine-tune cycles: " + ", ".join(f"{d['rejection']:.4f}{'*' if d['tuned'] else ''}" for d in tune_log)
        + " (* = tuned in that cycle)")
    w_tuned = sel.selector_get_weights()

    # timed cycles (regrids between them timed apart)
    n_timed_cycles = max(1, (args.steps + E - 1) // E)
    K = n_timed_cycles * E
    tree_t, prim_t = sel.tree(), sel.get_state()
    lam_t, tt_t = sel.lam.copy(), sel.tt
    wall_l = wall64 = reg_l = reg64 = 0.0
    upd = 0
    sel.selector_stats(reset=True)
    for cyc in range(n_timed_cycles):
        upd += E * sel.n_leaf
        wall_l += run(sel, E)
        wall64 += run(s64, E)
        a, b = regrid_both(phases=args.regrid_phases and cyc == n_timed_cycles - 1)
        reg64 += a
        reg_l += b
    st = sel.selector_stats(reset=True)
    troubled_frac = st["troubled"] / max(upd, 1)
    ups = upd / wall_l
    ups64 = upd / wall64
    log(f"timed: learned+fp32 {K} steps {wall_l * 1e3:.1f} ms, fp64 {wall64 * 1e3:.1f} ms, "
        f"judge rejection {judge_rate(st):.4f}, regrid {reg_l / n_timed_cycles:.2f}s (fp64 context "
        f"{reg64 / n_timed_cycles:.2f}s)")
    err_timed = l1_linf(s64, sel)
    # accuracy horizon: more paired cycles, then L1/Linf(rho) on the common mesh
    for cyc in range(args.accuracy_cycles):
        run(sel, E)
        run(s64, E)
        regrid_both()
    total_steps = args.warmup + E * (args.tune_cycles + n_timed_cycles + args.accuracy_cycles)
    err = l1_linf(s64, sel)
    log(f"L1/Linf(rho) learned vs fp64 after {total_steps} steps and "
        f"{1 + args.tune_cycles + n_timed_cycles + args.accuracy_cycles} regrids: {err}")

    # the same timed steps on the FP32 path with every leaf troubled (no selector): what the
    # network adds over FP32 alone
    s32 = clone(measured_config(args, F.FLUX_FP32, selector=0), tree_t, prim_t)
    s32.tt, s32.lam = tt_t, lam_t.copy()
    run(s32, 1)
    s32.set_state(prim_t)
    wall32 = run(s32, K)
    ups32 = K * s32.n_leaf / wall32
    s32.close()

    # kernel timings with HIP events, in passes of their own from the timed cycles' start
    # (event records add gaps to the stream, so they stay out of `value`)
    def kernel_ms(s, label):
        s.profile(True)
        if s is sel:
            s.selector_stats(reset=True)
        s.run_steps(E)
        ms, nl = s.profile_read()
        sel_ms = sel_n = tiles = 0
        if s is sel:
            sel_ms, sel_n, tiles = s.profile_read_select()
        s.profile(False)
        log(f"{label}: stage kernels {ms / max(nl, 1):.4f} ms avg over {nl} stages"
            + (f"; selector {sel_ms / max(sel_n, 1):.4f} ms avg over {sel_n} launches, "
               f"{tiles / max(sel_n, 1):.0f} network tiles per launch" if s is sel else ""))
        return ms / max(nl, 1), sel_ms / max(sel_n, 1), tiles / max(sel_n, 1)

    for s in (sel, s64):
        s.import_tree(*_dims(tree_t), tree_t)
        s.build_tables()
        s.set_state(prim_t)
        s.tt, s.lam = tt_t, lam_t.copy()
    N = sel.n_leaf
    l_max = sel.counts()["l_max"]
    k_l, k_sel, tiles = kernel_ms(sel, "learned+fp32")
    k64, _, _ = kernel_ms(s64, "fp64")
    # per-leaf algorithmic HBM bytes of one stage's flux/update launch, averaged over the 3
    # stages: stencil ids 32, own double4 state in 32, out 32, U_n 32 in stages 2 and 3
    bytes_per_leaf = 32 + 32 + 32 + 32 * 2.0 / 3.0
    achieved = N * bytes_per_leaf / (k_l * 1e-3) / 1e9
    # MFMA work of the selector: 28 16x16x32 BF16 MFMAs per 16-leaf tile (3 + 2 + 2 per
    # 16-unit output block of the three hidden layers), 16384 flops each; algorithmic
    # 2 * (75*64 + 64*64 + 64*64 + 64) = 26112 flops per leaf the network runs on
    mfma_flops = tiles * 28 * 16384
    alg_flops = tiles * 16 * 26112
    mfma = dict(kernel="k_select", ms=k_sel, network_tiles=tiles, mfma_tflops=mfma_flops / (k_sel * 1e-3) / 1e12,
                algorithmic_tflops=alg_flops / (k_sel * 1e-3) / 1e12, peak_tflops=2500.0,
                frac=mfma_flops / (k_sel * 1e-3) / 2.5e15)

    traffic = None
    if not args.no_traffic:
        import tempfile
        sel.close()
        s64.close()
        fd, mesh_file = tempfile.mkstemp(prefix="fdmamr_mesh_", suffix=".npz")
        os.close(fd)
        try:
            np.savez(mesh_file, prim=prim_t, weights=w_tuned, **tree_t)
            traffic = measure_traffic(args, mesh_file)
        finally:
            os.remove(mesh_file)
    out = {
        "metric": "leaf-cell RK3 updates/s: learned+FP32 vs plain FP64 at L1(rho) err <=1e-4",
        "value": ups,
        "unit": "leaf-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall_l / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-flux/fp64-update",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_leaf": N, "l_max": l_max,
                   "regrid_every": E, "refine_goal": args.refine_goal,
                   "coarse_goal": (args.refine_goal / 3.0 if args.workload == "shock4bubbles"
                                   and args.develop_t > 0 else None),
                   "developed": ({"t": args.develop_t, "max_level": args.develop_lvl,
                                  "refine_passes": args.refine_passes, "spinup_cycles": args.spinup_cycles}
                                 if args.workload == "shock4bubbles" and args.develop_t > 0 else None),
                   "base": "1024x512" if args.workload == "shock4bubbles" else None,
                   "selector_threshold": args.threshold, "label_bound": args.label_bound,
                   "pretrain": pre},
        "fp64_value": ups64,
        "speedup_vs_fp64": ups / ups64,
        "fp32_all_troubled_value": ups32,
        "speedup_fp32_all_troubled_vs_fp64": ups32 / ups64,
        "l1_rho_vs_fp64": err[0],
        "linf_rho_vs_fp64": err[1],
        "accuracy_steps": total_steps,
        "l1_rho_vs_fp64_after_timed": err_timed[0],
        "troubled_fraction": troubled_frac,
        "judge": dict(st, rejection_rate=judge_rate(st)),
        "fine_tune_cycles": tune_log,
        "regrid_s": reg_l / n_timed_cycles,
        "regrid_fp64_s": reg64 / n_timed_cycles,
        "value_regrid_amortised": upd / (wall_l + reg_l),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_stage"] if traffic else None,
                     "kernel": "k_stage_mixed (cleared + troubled leaves of one RK stage)",
                     "kernel_ms": k_l, "bytes_per_leaf_launch": bytes_per_leaf,
                     "fp64_kernel_ms": k64, "traffic_detail": traffic},
        "mfma": mfma,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(measured_config(args, F.FLUX_FP64), tree_t, lam_t)
    print(json.dumps(out))


def _dims(tree):
    return int(tree["num"][:, 1].max()), int(tree["num"][:, 0].max())


def l1_linf(s64, sel):
    """L1 (area-weighted) and Linf of rho between two contexts on the same mesh."""
    a = s64.get_state()[:, 0]
    b = sel.get_state()[:, 0]
    lvl = s64.geometry()["lvl"]
    area = (0.5 ** lvl) ** 2
    return float(np.sum(np.abs(a - b) * area) / np.sum(area)), float(np.max(np.abs(a - b)))


if __name__ == "__main__":
    main()
