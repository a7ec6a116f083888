"""Benchmark: the cell boundary-integral matvec (MyMatMult of common/ModVelSolver.F90:523-601)
on the 512-cell batch at SH degree 16 (nlat 17 x 34, 20% volume fraction, PME 128^3,
B-spline order 6), plain FP64, one MI355X.

One step = one full matvec b = A u with u and b resident in HBM: SH synthesis u -> g,
surface splines of g detJ, real-space double-layer pair sums with self and
near-singular corrections and the periodic linear term, PME (spread, FFTs, Green's
function, gather), + g, SH analysis.

Prints one JSON line.  `value` is cell-matvecs per second (cells / seconds per
matvec).  `roofline` describes the dominant kernel measured with HIP events on the
library's stream; `cpu_baseline` times the CPU restatement of the same algorithm
(oracle/) on this host on a bounded sample.  Beside it: `chosen` (the split-precision
form of the same matvec and its error), `timestep` (the 200-cell tube timestep, GMRES
seconds per timestep in the plain and chosen forms, with its own CPU baseline and the
learned warm start's series, and the factors each layer of the chosen form contributes)
and the 4,096-cell mixed batch (C5).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

try:  # before the library, so both share torch's HIP runtime
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes per k_pair launch on C2 from two rocprofv3 counter runs (FETCH_SIZE,
# WRITE_SIZE), scaled by the counters' reading for a plain 512 MiB copy (fetch x2.0,
# write x1.0); commands and table in profiles/cell_matvec_C2_fp64_hbm_traffic.md
K_PAIR_C2_TRAFFIC_BYTES = int((5363.8 + 342.8) * 1e6)


def progress(msg):
    """One line on stderr when a block of the benchmark ends (the JSON line stays the
    only output on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    """The host threads this process may use: OMP_NUM_THREADS when set (the GPU box sets
    it to its per-GPU CPU share), else the CPUs in the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(w, sample_every=32, threads=None):
    """Time the oracle's matvec pieces on this host: SH/spline/PME stages in full, the
    real-space stage on every `sample_every`-th target (scaled back up)."""
    from oracle import oracle_py
    threads = threads or host_threads()
    oracle_py.Oracle.threads(threads)
    o = oracle_py.from_workload(w)
    dof = o.cell_dof()
    u = np.random.default_rng(7).standard_normal(dof)
    tl = o.tlist_rbc()
    tx, ti, act, A = tl
    c2 = -1.0 / (4.0 * np.pi)
    t = {}
    t0 = time.perf_counter()
    g = o.glob_sph_trans(c=u, direction=-1)
    t["sh_synth"] = time.perf_counter() - t0
    npt = o.np
    gc = np.zeros(3 * npt)
    off = 0
    for nl in w.cells.nlat:
        n = int(nl) * 2 * int(nl)
        for d in range(3):
            gc[3 * off + n * d: 3 * off + n * (d + 1)] = g[off + npt * d: off + npt * d + n]
        off += n
    t0 = time.perf_counter()
    o.set_density(2, gc)
    t["splines"] = time.perf_counter() - t0
    sub = np.zeros_like(act)
    sub[::sample_every] = 1
    t0 = time.perf_counter()
    o.add_int_on_rbcs(0.0, c2, (tx, ti, sub, A))
    t["real_space_sample"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    v = o.pme(0.0, c2, tl, cells=True, walls=False)
    t["pme"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.glob_sph_trans(v=v, direction=1)
    t["sh_anal"] = time.perf_counter() - t0
    total = (t["sh_synth"] + t["splines"] + sample_every * t["real_space_sample"] + t["pme"]
             + t["sh_anal"])
    return total, t, threads


MLP_C3 = os.path.join(ROOT, "rbc3d_amd", "data", "mlp_C3.npz")
# GMRES restart length of the chosen form: 150 and longer take the same iterations on
# the tube (133 to 1e-8 against 910 at PETSc's default 30)
CHOSEN_RESTART = 150


def timestep_bench(steps, warmup, maxit=4000, forms=None):
    """One timestep of the 200-cell tube at viscosity ratio 1 (ModTimeInt.F90:716-808 +
    ModNoSlip.F90:45-150): Solve_RBC_Vel (right-hand side: cell single layer, walls,
    PME; no cell GMRES at ratio 1; density g := v), NoSlipWall (residual velocity,
    wall-traction GMRES to 1e-8 relative residual, update), and Solve_RBC_Vel again with
    the new wall traction (the velocity the step hands on).  The wall traction is reset
    to zero before every step, as at the first step of a run, so every step does the
    same work.

    Forms, every one GMRES to the same 1e-8 relative residual: "plain" = FP64,
    GMRES(30) from zero (the reference's solver); "split" = the split form, GMRES(30);
    "fp64_restart150" = FP64, GMRES(150); "chosen" = the split form with GMRES(150);
    "chosen_learned" = the chosen form warm-started by the network
    (rbc3d_noslip_wall_learned, trained by python -m rbc3d_amd.learned train).  Each
    form's cell velocity is compared with the plain form's."""
    import ctypes as C
    import torch
    import rbc3d_amd
    from rbc3d_amd import configs
    w = configs.make_workload("C3", seed=1)
    g = rbc3d_amd.from_workload(w)
    nv = w.wall[2]
    f0 = np.zeros(3 * nv)
    g.prepare_sing_int_on_wall()
    nnz = C.c_longlong()
    rbc3d_amd._call("rbc3d_wall_self_nnz", g.h, C.byref(nnz))
    stream = torch.cuda.ExternalStream(g.stream())
    mlp = None
    if os.path.exists(MLP_C3):
        from rbc3d_amd import learned
        mlp = learned.load_model(MLP_C3)

    def run(form):
        prec = g.PREC_FP64 if form in ("plain", "fp64_restart150") else g.PREC_SPLIT
        restart = 30 if form in ("plain", "split") else CHOSEN_RESTART
        rtol = 1e-8
        g.set_precision(prec)
        g.set_gmres_restart(restart)
        total_ms, its, res, v = 0.0, [], None, None
        # the timed steps run without the stage events (a GMRES iteration is ~0.6 ms and
        # records ~16 of them); one more step with the events gives the breakdown
        for k in range(warmup + steps + 1):
            if k == warmup + steps:
                g.timing(-1)
                g.timing(1)
            g.set_wall_density(f0)
            g.sync()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record(stream)
            g.solve_rbc_vel()
            if form == "chosen_learned":
                res = g.noslip_wall_learned(mlp, rtol, maxit)
            else:
                res = g.noslip_wall(rtol, maxit)
            v, _, _ = g.solve_rbc_vel()
            e1.record(stream)
            g.sync()
            wall = time.perf_counter() - t0
            if warmup <= k < warmup + steps:
                total_ms += max(e0.elapsed_time(e1), 1e3 * wall)
                its.append(res[0])
        stages = g.stage_times()
        g.timing(0)
        g.set_precision(g.PREC_FP64)
        g.set_gmres_restart(30)
        out = {"s_per_timestep": round(total_ms / steps * 1e-3, 5), "rtol": rtol,
               "gmres_restart": restart,
               "gmres_iterations": its, "gmres_residual": res[1], "gmres_converged": res[2],
               "wall_velocity_after": res[3],
               "stages_ms_per_step": {k: round(v_[0], 4) for k, v_ in stages.items() if v_[1]}}
        if form == "chosen_learned":
            out["fallback"] = res[4]
        return out, v

    cells = w.cells
    out = {
        "workload": f"C3: {cells.ncell} biconcave cells (radius 1.722, nlat 13x26, 25% "
                    f"haematocrit) in an R=8, Lz=32 periodic tube, {nv} wall vertices "
                    f"({w.wall[3]} Tri3), box {w.Lb}, PME {w.Nb}, viscosity ratio 1",
        "steps": steps, "wall_self_matrix_nnz": int(nnz.value)}
    plain, vp = run("plain")
    progress(f"tube timestep plain: {plain['s_per_timestep']} s")
    out["plain"] = plain
    only = forms
    forms = ["split", "fp64_restart150", "chosen"] + (["chosen_learned"] if mlp is not None else [])
    if only is not None:  # a subset, for profiling one form (the plain form always runs)
        forms = [f for f in forms if f in only]
    for form in forms:
        r, v = run(form)
        progress(f"tube timestep {form}: {r['s_per_timestep']} s")
        r["velocity_rel_err_vs_plain"] = float(np.abs(v - vp).max() / np.abs(vp).max())
        r["speedup_vs_plain"] = round(plain["s_per_timestep"] / r["s_per_timestep"], 3)
        out[form] = r
    if mlp is None:
        out["chosen_learned"] = {"skipped": f"no trained network at {MLP_C3}"}
    elif only is None:
        out["learned_series"] = learned_series_bench(mlp)
    # what each layer of the chosen form contributes, as ratios of seconds per timestep
    def sec(f):
        r = out.get(f, {})
        return r.get("s_per_timestep") if isinstance(r, dict) else None
    fac = {}
    if sec("fp64_restart150"):
        fac["restart_150_vs_30_fp64"] = round(plain["s_per_timestep"] / sec("fp64_restart150"), 3)
    if sec("fp64_restart150") and sec("chosen"):
        fac["precision_split_vs_fp64_at_150"] = round(sec("fp64_restart150") / sec("chosen"), 3)
    if sec("split"):
        fac["precision_split_vs_fp64_at_30"] = round(plain["s_per_timestep"] / sec("split"), 3)
    if sec("chosen") and sec("chosen_learned"):
        fac["learned_warm_start_vs_chosen"] = round(sec("chosen") / sec("chosen_learned"), 3)
    if sec("chosen"):
        fac["chosen_vs_plain"] = round(plain["s_per_timestep"] / sec("chosen"), 3)
    out["factors"] = fac
    # the headline fields (plain FP64, the reference's arithmetic)
    out["s_per_timestep"] = plain["s_per_timestep"]
    out["gmres_iterations"] = plain["gmres_iterations"]
    return out


def timestep_cpu_baseline(niter, threads=None, o=None):
    """The oracle's tube timestep on this host, from a bounded sample: one Solve_RBC_Vel
    and two applications of the no-slip operator timed, the timestep estimated as two
    Solve_RBC_Vel plus `niter` (the plain GMRES(30) count) operator applications; the
    wall self-matrix assembly is reported beside it, not counted (it is done once, here
    on every host thread).  Pass `o` to reuse a prepared oracle."""
    from oracle import oracle_py
    from rbc3d_amd import configs
    threads = threads or host_threads()
    t_prep = None
    if o is None:
        oracle_py.Oracle.threads(host_threads())
        w = configs.make_workload("C3", seed=1)
        o = oracle_py.from_workload(w)
        t0 = time.perf_counter()
        o.prepare_wall()
        t_prep = time.perf_counter() - t0
    oracle_py.Oracle.threads(threads)
    o.set_wall_density(np.zeros(3 * oracle_py.L().orc_ntargets_wall(o.h)))
    t0 = time.perf_counter()
    o.solve_rbc_vel()
    t_rhs = time.perf_counter() - t0
    u = np.random.default_rng(3).standard_normal(o.wall_dof())
    t0 = time.perf_counter()
    for _ in range(2):
        o.wall_matmult(u)
    t_mv = (time.perf_counter() - t0) / 2
    est = 2 * t_rhs + niter * t_mv
    return {"value": round(est, 2), "unit": "s/timestep", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement, OpenMP {threads} threads): one Solve_RBC_Vel and "
                      f"two no-slip operator applications timed; timestep = 2 x Solve_RBC_Vel "
                      f"+ {niter} operator applications (the plain GMRES(30) count)",
            "solve_rbc_vel_s": round(t_rhs, 3), "wall_matvec_s": round(t_mv, 3),
            "wall_self_matrix_assembly_s": None if t_prep is None else round(t_prep, 3)}, o


def learned_series_bench(mlp, seed=201, steps=4):
    """The warm start where it is meant to work: consecutive steps of a tube series
    (held-out placement, cells moving rigidly with their mean velocity, dt = 0.05).
    From the same state at every step: plain FP64 NoSlipWall from zero, the chosen
    form (split, GMRES(150)) from zero and the chosen form with the network's warm start,
    all to the same absolute residual 1e-8 |b_0| (|b_0|: the first step's right-hand
    side)."""
    import rbc3d_amd
    from rbc3d_amd import configs, learned
    w = configs.make_workload("C3", seed=seed)
    g = rbc3d_amd.from_workload(w)
    g.prepare_sing_int_on_wall()
    atol = 1e-8 * learned.rhs_norm(g)
    recs = learned.evaluate(g, w.cells, mlp, steps, 0.05, 1e-8, 4000, split=True, atol=atol,
                            restart=CHOSEN_RESTART, cold=True)
    g.close()
    ip = [r["plain_iters"] for r in recs]
    il = [r["learned_iters"] for r in recs]
    tp = sum(r["plain_s"] for r in recs)
    tl = sum(r["learned_s"] for r in recs)
    tc = sum(r["cold_s"] for r in recs)
    return {"workload": f"C3 series, placement seed {seed}, {steps} steps (step 0 starts from "
                        "zero traction)",
            "plain_iterations": ip, "chosen_iterations": [r["cold_iters"] for r in recs],
            "chosen_learned_iterations": il,
            "plain_s_per_step": round(tp / steps, 5), "chosen_s_per_step": round(tc / steps, 5),
            "chosen_learned_s_per_step": round(tl / steps, 5),
            "speedup_vs_plain": round(tp / tl, 3),
            "learned_speedup_vs_chosen": round(tc / tl, 3),
            "fallbacks": sum(int(r["fallback"]) for r in recs),
            "velocity_rel_err_vs_plain_max": max(r["vel_diff"] for r in recs)}


def stage_rates(w, st, split):
    """Rates of the matrix-core and bandwidth stages of one cell matvec from its stage
    times (ms): the Legendre GEMMs of the two global SH transforms (synthesis of the
    density, analysis of the velocity; 2 nlat nlat0 (nlat0 + 1) flop per cell and
    component each) against the matrix cores' peak -- over the whole SH-transform
    stage, which also holds the longitude DFTs and the packing, so a lower bound -- and
    the PME stages' algorithmic bytes for the double layer (n_f = 9 spread fields:
    8 (13 N_s + 7 N_t) + 8 Nb^3 (3 n_f + n_f + 3 + 6 + 3)) against HBM."""
    cells = w.cells
    nl = np.asarray(cells.nlat, np.float64)
    nl0 = np.asarray(cells.nlat0, np.float64)
    f_leg = 2 * 3 * float(np.sum(2 * nl * nl0 * (nl0 + 1)))
    nb3 = float(np.prod(w.Nb))
    ns = float(cells.npoint)
    b_pme = 8 * (13 * ns + 7 * ns) + 8 * nb3 * (3 * 9 + 9 + 3 + 6 + 3)
    leg_ms = st.get("sh_glob", 0.0)
    pme_ms = sum(st.get(k, 0.0) for k in ("spread_prep", "spread", "fft", "green", "gather"))
    return {"sh_legendre": {"flop": f_leg, "sh_stage_ms": round(leg_ms, 4),
                            "tflops": round(f_leg / (leg_ms * 1e-3) / 1e12, 3) if leg_ms else None,
                            # split: three BF16 products per FP32 product
                            "peak_tflops": round(2500.0 / 3, 1) if split else PEAK_FP64_TFLOPS},
            "pme": {"algorithmic_bytes": b_pme, "ms": round(pme_ms, 4),
                    "gb_per_s": round(b_pme / (pme_ms * 1e-3) / 1e9, 1) if pme_ms else None,
                    "hbm_peak_gb_per_s": PEAK_HBM_GBS}}


def c5_bench(steps, warmup):
    """Cell-BIE matvec on the large mixed batch: 3,584 biconcave discs (nlat 13) and 512
    spheres of radius 1.4 (nlat 17) at 30% volume fraction, PME 256^3, P=6, viscosity
    ratio 5; two grid types in one launch of every kernel, and the matvec at fixed
    geometry replayed from a recorded hipGraph in both forms.  Plain FP64 and the split
    form on the same vector (error against FP64), each with its stage breakdown (one
    extra matvec with the stage events), the Legendre GEMMs' rate against the matrix
    cores' peak and the PME stages' algorithmic HBM rate."""
    import torch
    import rbc3d_amd
    from rbc3d_amd import configs
    w = configs.make_workload("C5", seed=1)
    g = rbc3d_amd.from_workload(w)
    g.set_geometry_reuse(False)
    dof = g.cell_dof()
    dev = torch.device("cuda:0")
    u = torch.randn(dof, dtype=torch.float64, device=dev,
                    generator=torch.Generator(device=dev).manual_seed(5))
    b = torch.empty(dof, dtype=torch.float64, device=dev)
    stream = torch.cuda.ExternalStream(g.stream())
    cells = w.cells

    def run(prec):
        g.set_precision(prec)
        for _ in range(warmup):
            g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        g.sync()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        e1.record(stream)
        g.sync()
        torch.cuda.synchronize()
        ms = max(e0.elapsed_time(e1), 1e3 * (time.perf_counter() - t0)) / steps
        g.timing(-1)
        g.timing(1)
        g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        g.sync()
        st = {k: v[0] for k, v in g.stage_times().items() if v[1]}
        g.timing(0)
        out = b.clone()
        g.set_precision(g.PREC_FP64)
        rec = {"ms_per_step": round(ms, 4), "cell_matvecs_per_s": round(cells.ncell / (ms * 1e-3), 3),
               "stages_ms": {k: round(v, 4) for k, v in st.items()}}
        rec.update(stage_rates(w, st, prec == g.PREC_SPLIT))
        return rec, out

    plain, b64 = run(g.PREC_FP64)
    split, b32 = run(g.PREC_SPLIT)
    split["rel_err_vs_plain"] = float((b32 - b64).abs().max() / b64.abs().max())
    npairs, _ = g.pair_stats()

    def run_graph(prec):
        """The matvec as a GMRES solve applies it: geometry reused, replayed from one
        recorded hipGraph."""
        g.set_precision(prec)
        g.set_geometry_reuse(True)
        g.set_graph(True)
        for _ in range(max(warmup, 2)):  # the first records the graph
            g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        g.sync()
        assert g.graph_ready()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        e1.record(stream)
        g.sync()
        torch.cuda.synchronize()
        ms = max(e0.elapsed_time(e1), 1e3 * (time.perf_counter() - t0)) / steps
        out = b.clone()
        g.set_graph(False)
        g.set_geometry_reuse(False)
        g.set_precision(g.PREC_FP64)
        return {"ms_per_step": round(ms, 4),
                "cell_matvecs_per_s": round(cells.ncell / (ms * 1e-3), 3)}, out

    graph64, bg64 = run_graph(g.PREC_FP64)
    graph32, bg32 = run_graph(g.PREC_SPLIT)
    graph64["equal_to_eager"] = bool(torch.equal(bg64, b64))
    graph32["rel_err_vs_plain"] = float((bg32 - b64).abs().max() / b64.abs().max())
    return {"workload": f"C5: {cells.ncell} cells (3584 discs nlat 13, 512 spheres R 1.4 "
                        f"nlat 17), box {w.Lb[0]:.3f}^3 (30% volume fraction), PME {w.Nb[0]}^3, "
                        f"P={w.P}, lambda=5, one cell-BIE matvec",
            "nodes": cells.npoint, "pairs_per_matvec": int(npairs), "steps": steps,
            "ms_per_step": plain["ms_per_step"], "cell_matvecs_per_s": plain["cell_matvecs_per_s"],
            "plain": plain, "split": split,
            "graph": {"what": "geometry reused (the matvecs of one GMRES solve), each stage one "
                              "launch over both grid types, the whole matvec replayed from "
                              "one recorded hipGraph",
                      "fp64": graph64, "split": graph32}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-only", action="store_true")
    ap.add_argument("--sample-every", type=int, default=32)
    ap.add_argument("--no-timestep", action="store_true",
                    help="skip the tube timestep (GMRES seconds per timestep)")
    ap.add_argument("--timestep-steps", type=int, default=3)
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the 4,096-cell mixed-size matvec (C5)")
    ap.add_argument("--timestep-forms", default=None,
                    help="comma list of timestep forms beside plain (default: all, and the "
                         "learned series)")
    ap.add_argument("--timestep-only", action="store_true",
                    help="only the tube timestep (prints its JSON object)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain 512 MiB device copy first (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.timestep_only:
        forms = args.timestep_forms.split(",") if args.timestep_forms else None
        print(json.dumps(timestep_bench(args.timestep_steps, args.warmup, forms=forms)))
        return

    from rbc3d_amd import configs
    w = configs.make_workload(args.config, seed=1)
    ncell = w.cells.ncell

    if args.cpu_only:
        tot, parts, thr = cpu_baseline(w, args.sample_every)
        print(json.dumps({"cpu_s_per_matvec": tot, "parts": parts, "threads": thr}))
        return

    import torch
    import rbc3d_amd

    sysg = rbc3d_amd.from_workload(w)
    dof = sysg.cell_dof()
    dev = torch.device("cuda:0")
    u = torch.randn(dof, dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(3))
    b = torch.empty(dof, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(sysg.stream())

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    def timed_matvecs(reuse, steps, warmup):
        """K matvecs bracketed by HIP events on the library stream (and the host clock)."""
        sysg.set_geometry_reuse(reuse)
        for _ in range(warmup):
            sysg.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        sysg.sync()
        sysg.timing(-1)
        sysg.timing(1)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        sysg.sync()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            sysg.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        e1.record(stream)
        sysg.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = sysg.stage_times()
        sysg.timing(0)
        return max(e0.elapsed_time(e1), 1e3 * wall) / steps, st

    # the headline: every matvec does all of its work, geometry included (as the
    # reference's MyMatMult does)
    ms_per_step, stages = timed_matvecs(False, args.steps, args.warmup)
    progress(f"{args.config} plain matvec: {ms_per_step:.3f} ms")
    # within one GMRES solve the geometry is fixed: neighbour lists and projections
    # computed by the first matvec are reused by the others
    ms_reuse, stages_reuse = timed_matvecs(True, args.steps, 1)
    npairs, _ = sysg.pair_stats()
    nproj, nnear = sysg.near_stats()
    b_plain = b.clone()
    # the split form on the same vector: time and error against the plain result
    sysg.set_precision(sysg.PREC_SPLIT)
    ms_split, stages_split = timed_matvecs(False, args.steps, 1)
    ms_split_reuse, _ = timed_matvecs(True, args.steps, 1)
    torch.cuda.synchronize()
    split_err = float((b - b_plain).abs().max() / b_plain.abs().max())
    progress(f"{args.config} reused {ms_reuse:.3f} ms, split {ms_split:.3f} ms (err {split_err:.2e})")
    sysg.set_precision(sysg.PREC_FP64)

    # dominant kernel and its roofline: the real-space pair kernel is FP64-compute
    # bound; algorithmic work = 28 flop per in-cutoff (target, source) pair for the
    # double layer alone (c1 = 0 in this matvec)
    dom = max(stages.items(), key=lambda kv: kv[1][0])
    pair_ms, pair_n = stages["pair"]
    pair_avg_ms = pair_ms / max(pair_n, 1)
    pair_tflops = 28.0 * npairs / (pair_avg_ms * 1e-3) / 1e12
    roofline = {"kernel": "k_pair (double-layer real-space pair sums)", "bound": "fp64-valu",
                "compute": "FP64 vector ALU (no matrix shape; MI355X FP64 vector peak 78.6 "
                           "TFLOP/s)",
                "achieved": round(pair_tflops, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(pair_tflops / PEAK_FP64_TFLOPS, 4),
                "traffic": K_PAIR_C2_TRAFFIC_BYTES if args.config == "C2" else None,
                "traffic_unit": "bytes per launch (HBM read + write, rocprofv3 counters)",
                "avg_launch_ms": round(pair_avg_ms, 4), "pairs_per_launch": int(npairs),
                "flop_per_pair": 28}

    out = {
        "metric": "cell-matvecs/s (plain FP64 cell-BIE matvec)",
        "value": round(ncell / (ms_per_step * 1e-3), 3),
        "unit": "cell-matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random biconcave discs, random SH densities)",
        "config": {"workload": f"{args.config}: {ncell} biconcave cells, nlat 17x34, box "
                   f"{w.Lb[0]:.3f}^3 (20% volume fraction), PME {w.Nb[0]}^3, P={w.P}, "
                   f"lambda=5, one cell-BIE matvec", "cells": ncell, "nodes": w.cells.npoint,
                   "dof": dof},
        "roofline": roofline,
        "dominant_stage": {"name": dom[0], "ms_per_step": round(dom[1][0] / args.steps, 4)},
        "stages_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in stages.items() if v[1]},
        "near_singular": {"projections": int(nproj), "corrections": int(nnear)},
        "rates": stage_rates(w, {k: v[0] / args.steps for k, v in stages.items() if v[1]}, False),
        "chosen": {
            "what": "split form (FP32 PME, split-BF16 MFMA Legendre GEMMs; pair sums and "
                    "singular quadrature FP64), same matvec",
            "ms_per_step": round(ms_split, 4),
            "cell_matvecs_per_s": round(ncell / (ms_split * 1e-3), 3),
            "geometry_reused_ms_per_step": round(ms_split_reuse, 4),
            "rel_err_vs_plain": split_err,
            "rates": stage_rates(w, {k: v[0] / args.steps for k, v in stages_split.items() if v[1]},
                                 True),
            "stages_ms_per_step": {k: round(v[0] / args.steps, 4)
                                   for k, v in stages_split.items() if v[1]}},
        "geometry_reused": {
            "what": "matvecs 2..K of a GMRES solve at fixed geometry: neighbouring-cell "
                    "lists, target binning and closest-point projections from the first",
            "ms_per_step": round(ms_reuse, 4),
            "cell_matvecs_per_s": round(ncell / (ms_reuse * 1e-3), 3),
            "stages_ms_per_step": {k: round(v[0] / args.steps, 4)
                                   for k, v in stages_reuse.items() if v[1]}},
    }
    if not args.no_timestep:
        out["timestep"] = timestep_bench(args.timestep_steps, 1)
        progress("tube timestep done")
        # the metric's second half: GMRES seconds per tube timestep, chosen vs plain
        ts = out["timestep"]
        ch, pl = ts.get("chosen", {}), ts.get("plain", {})
        out["gmres_s_per_timestep"] = {
            "plain_fp64_gmres30": pl.get("s_per_timestep"),
            "chosen_split_gmres150": ch.get("s_per_timestep"),
            "speedup": ch.get("speedup_vs_plain"),
            "velocity_rel_err": ch.get("velocity_rel_err_vs_plain")}
    if not args.no_c5:
        out["C5"] = c5_bench(3, 1)
        progress(f"C5 matvec: {out['C5']['ms_per_step']:.3f} ms")
    if not args.no_cpu:
        thr = host_threads()
        tot, parts, thr = cpu_baseline(w, args.sample_every, thr)
        one_every = 8 * args.sample_every
        progress(f"CPU baseline ({thr} threads): {tot:.2f} s per matvec")
        tot1, parts1, _ = cpu_baseline(w, one_every, 1)
        progress(f"CPU baseline (1 thread): {tot1:.2f} s per matvec")
        out["cpu_baseline"] = {"value": round(ncell / tot, 4), "unit": "cell-matvecs/s",
                               "cores": thr, "kind": "port",
                               "sample": f"oracle (C restatement, OpenMP {thr} threads): SH, "
                               f"splines and PME in full; real-space stage on every "
                               f"{args.sample_every}th target, scaled by {args.sample_every}",
                               "s_per_matvec": round(tot, 3),
                               "parts_s": {k: round(v, 3) for k, v in parts.items()},
                               "host": host_info(),
                               "one_thread": {
                                   "value": round(ncell / tot1, 5), "cores": 1,
                                   "s_per_matvec": round(tot1, 3),
                                   "sample": f"the same on 1 thread (one MPI rank's share), "
                                             f"real-space stage on every {one_every}th target, "
                                             f"scaled by {one_every}",
                                   "parts_s": {k: round(v, 3) for k, v in parts1.items()}},
                               "gpu_speedup": round(out["value"] / (ncell / tot), 1),
                               "gpu_speedup_vs_one_thread": round(out["value"] / (ncell / tot1), 1)}
        if "timestep" in out:
            ts = out["timestep"]
            niter = int(np.mean(ts["plain"]["gmres_iterations"]))
            cb, o3 = timestep_cpu_baseline(niter, thr)
            progress(f"tube CPU baseline ({thr} threads): {cb['value']} s per timestep")
            cb1, _ = timestep_cpu_baseline(niter, 1, o3)
            progress(f"tube CPU baseline (1 thread): {cb1['value']} s per timestep")
            cb["one_thread"] = cb1
            cb["host"] = host_info()
            ts["cpu_baseline"] = cb
            for f in ("plain", "chosen"):
                ts[f]["speedup_vs_cpu"] = round(cb["value"] / ts[f]["s_per_timestep"], 1)
                ts[f]["speedup_vs_cpu_one_thread"] = round(cb1["value"] / ts[f]["s_per_timestep"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
