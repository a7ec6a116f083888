"""The learned warm start of NoSlipWall (common/ModNoSlip.F90:45-150): training data
from the plain solver on tube time series, training of the per-wall-vertex network
with the library's own BF16 kernels (rbc3d_mlp_*), and its evaluation against plain
GMRES (which is both the fallback and the measure of success).

One step of a series, at viscosity ratio 1 (the tube configuration): Solve_RBC_Vel
gives the cells' velocity; the network's inputs are read at the wall vertices (the
cells' velocity there, the current traction f0, the vertex normal); NoSlipWall with
the plain solver gives the converged traction f, and f - f0 is the training target;
then every cell moves rigidly with the mean velocity of its nodes for dt.  Membrane
mechanics are not part of this library, so cells keep their shape and force density: the
series exercises how the wall traction follows cells that move along the tube.

The network itself (9 -> 64 -> 64 -> 64 -> 3, GELU, BF16 weights, FP32
accumulation) is defined in rbc3d_amd/csrc/mlp.hip.
"""
from __future__ import annotations

import math
import time
from dataclasses import replace

import numpy as np

try:  # before the library, so both share torch's HIP runtime
    import torch  # noqa: F401
except ImportError:
    pass


def cell_offsets(cells):
    n = (cells.nlat * 2 * cells.nlat).astype(np.int64)
    return np.concatenate([[0], np.cumsum(n)]), n


def move_cells(cells, v, dt):
    """Translate every cell by dt times the mean velocity of its nodes; v is the
    node-ordered (npoint,3) velocity flattened in Fortran order."""
    off, n = cell_offsets(cells)
    npt_all = int(off[-1])
    V = v.reshape(3, npt_all)
    x = cells.x.copy()
    for k in range(cells.ncell):
        vm = V[:, off[k]:off[k + 1]].mean(1)
        b = 3 * off[k]
        for d in range(3):
            x[b + d * n[k]: b + (d + 1) * n[k]] += dt * vm[d]
    return replace(cells, x=x)


def features(g, normals=None):
    """Network inputs at the wall vertices, (nvert, 9) float32: the cells' velocity,
    the current wall traction and the vertex normal."""
    nv = g.nvert
    u = g.wall_cell_velocity().reshape(3, nv).T
    f0 = g.wall_density().reshape(3, nv).T
    nrm = (g.wall_vertex_normals() if normals is None else normals).reshape(3, nv).T
    return np.ascontiguousarray(np.hstack([u, f0, nrm]), dtype=np.float32), f0


def rhs_norm(g):
    """|u_res| over the wall vertices: the norm of NoSlipWall's right-hand side (up to
    the duplicated seam vertices)."""
    return float(np.linalg.norm(g.wall_residual_vel()))


def step_rtol(g, rtol, atol):
    """The relative tolerance that stops GMRES at max(rtol |b|, atol)."""
    if not atol:
        return rtol
    return max(rtol, atol / max(rhs_norm(g), 1e-300))


def series(g, cells, nsteps, dt, rtol, maxit, log=None, atol=None):
    """Run a tube time series with the plain solver; yields (X, T, info) per step with
    X the inputs (nvert, 9) and T = f - f0 (nvert, 3), both float32.  GMRES stops at
    max(rtol |b|, atol): in a series |b| is only the change since the last step, and a
    tolerance relative to it alone asks for ever smaller absolute residuals."""
    nv = g.nvert
    normals = g.wall_vertex_normals()
    for step in range(nsteps):
        v, _, _ = g.solve_rbc_vel()
        X, f0 = features(g, normals)
        rt = step_rtol(g, rtol, atol)
        t0 = time.perf_counter()
        it, res, conv, _ = g.noslip_wall(rt, maxit)
        dtw = time.perf_counter() - t0
        f = g.wall_density().reshape(3, nv).T
        T = np.ascontiguousarray(f - f0, dtype=np.float32)
        info = {"step": step, "niter": it, "converged": conv, "s": dtw}
        if log:
            log(info)
        yield X, T, info
        cells = move_cells(cells, v, dt)
        g.set_cells(cells)
        g.set_cell_density(1, cells.f)
        g.set_cell_density(2, cells.g)


def norm_stats(X, T):
    mean = X.mean(0, dtype=np.float64)
    std = X.std(0, dtype=np.float64)
    std = np.where(std > 1e-12, std, 1.0)
    scale = np.sqrt((T.astype(np.float64) ** 2).mean(0))
    scale = np.where(scale > 1e-30, scale, 1.0)
    return mean, std, scale


def train(mlp, Xd, Td, epochs=30, batch=65536, lr=2e-3, seed=0, log=None):
    """Adam on the device arrays Xd (N,9), Td (N,3) (torch float32 tensors on the GPU);
    cosine learning-rate decay to lr/20.  Returns the loss after each epoch (on a fixed
    sample of up to 2^18 points)."""
    import torch
    mlp.use_torch_stream()  # the batches are gathered on torch's stream
    N = Xd.shape[0]
    gen = torch.Generator(device=Xd.device).manual_seed(seed)
    nb = max(1, N // batch)
    total = epochs * nb
    hist = []
    k = 0
    ev = torch.randperm(N, generator=gen, device=Xd.device)[: min(N, 1 << 18)]
    Xe, Te = Xd[ev].contiguous(), Td[ev].contiguous()
    for ep in range(epochs):
        perm = torch.randperm(N, generator=gen, device=Xd.device)
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            xb, tb = Xd[idx].contiguous(), Td[idx].contiguous()
            a = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * k / total)))
            mlp.adam_step_dev(xb.shape[0], xb.data_ptr(), tb.data_ptr(), a)
            k += 1
        g = torch.empty(mlp.nparams, dtype=torch.float32, device=Xd.device)
        loss = mlp.grad_dev(Xe.shape[0], Xe.data_ptr(), Te.data_ptr(), g.data_ptr())
        hist.append(loss)
        if log:
            log({"epoch": ep, "loss": loss})
    return hist


def evaluate(g, cells, mlp, nsteps, dt, rtol=1e-8, maxit=4000, log=None, split=False,
             atol=None, restart=30, cold=False):
    """From the same state at every step: plain NoSlipWall from zero (FP64, GMRES(30))
    and the warm-started one (in the split form when `split`, the cells' velocity after
    it too; GMRES(`restart`)), both to max(rtol |b|, atol); with `cold`, also the
    warm-started arm's settings from zero, which isolates what the network adds.  The
    series continues from the plain solution.  Returns per-step records (iterations,
    seconds, cell-velocity difference)."""
    out = []
    for step in range(nsteps):
        f0 = g.wall_density()
        v0, _, _ = g.solve_rbc_vel()
        rtol_s = step_rtol(g, rtol, atol)
        g.set_wall_density(f0)
        g.sync()
        t0 = time.perf_counter()
        ip, rp, cp, _ = g.noslip_wall(rtol_s, maxit)
        tp = time.perf_counter() - t0
        fp = g.wall_density()
        vp, _, _ = g.solve_rbc_vel()
        rel = lambda v: float(np.abs(v - vp).max() / np.abs(vp).max())
        rec = {"step": step, "rtol": rtol_s, "plain_iters": ip, "plain_s": tp, "plain_conv": cp}
        g.set_gmres_restart(restart)
        if split:
            g.set_precision(g.PREC_SPLIT)
        try:
            if cold:
                g.set_wall_density(f0)
                g.sync()
                t0 = time.perf_counter()
                ic, _, cc, _ = g.noslip_wall(rtol_s, maxit)
                tc = time.perf_counter() - t0
                vc, _, _ = g.solve_rbc_vel()
                rec.update(cold_iters=ic, cold_s=tc, cold_conv=cc, cold_vel_diff=rel(vc))
            g.set_wall_density(f0)
            g.sync()
            t0 = time.perf_counter()
            il, rl, cl, _, fb = g.noslip_wall_learned(mlp, rtol_s, maxit)
            tl = time.perf_counter() - t0
            vl, _, _ = g.solve_rbc_vel()
        finally:
            g.set_precision(g.PREC_FP64)
            g.set_gmres_restart(30)
        rec.update(learned_iters=il, learned_s=tl, learned_conv=cl, fallback=fb,
                   vel_diff=rel(vl))
        out.append(rec)
        if log:
            log(rec)
        g.set_wall_density(fp)
        cells = move_cells(cells, v0, dt)
        g.set_cells(cells)
        g.set_cell_density(1, cells.f)
        g.set_cell_density(2, cells.g)
    return out


def save_model(path, mlp, mean, std, scale, meta=None):
    np.savez(path, params=mlp.get_params(), mean=np.asarray(mean, np.float64),
             std=np.asarray(std, np.float64), scale=np.asarray(scale, np.float64),
             meta=np.asarray(repr(meta or {})))


def load_model(path):
    import rbc3d_amd
    z = np.load(path, allow_pickle=False)
    m = rbc3d_amd.Mlp()
    m.set_params(z["params"])
    m.set_norm(z["mean"], z["std"], z["scale"])
    return m


def _seeds(spec):
    a, b = spec.split("-")
    return list(range(int(a), int(b) + 1))


def init_params(rng):
    """He-style initialisation of the FP32 parameter vector (W ~ N(0, 2 / fan_in),
    b = 0, last layer scaled by 0.1), in the order of include/rbc3d.h."""
    parts = []
    for shp, last in [((64, 9), False), ((64,), None), ((64, 64), False), ((64,), None),
                      ((64, 64), False), ((64,), None), ((3, 64), True), ((3,), None)]:
        if last is None:
            parts.append(np.zeros(shp))
        else:
            wgt = rng.standard_normal(shp) * np.sqrt(2.0 / shp[1])
            parts.append(wgt * (0.1 if last else 1.0))
    return np.concatenate([q.reshape(-1) for q in parts]).astype(np.float32)


def main(argv=None):
    """Train the network on tube series of the training seeds and save it, or evaluate
    a saved network on held-out seeds (one JSON line per step and a summary)."""
    import argparse
    import json
    import rbc3d_amd
    from rbc3d_amd import configs
    ap = argparse.ArgumentParser(prog="python -m rbc3d_amd.learned")
    ap.add_argument("phase", choices=["train", "eval"])
    ap.add_argument("--config", default="C3")
    ap.add_argument("--seeds", default=None, help="first-last (train: 101-120, eval: 201-205)")
    ap.add_argument("--steps", type=int, default=None, help="steps per seed (train 100, eval 8)")
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--rtol", type=float, default=None, help="train: 1e-6 targets; eval: 1e-8")
    ap.add_argument("--atol-rel", type=float, default=None,
                    help="absolute tolerance as a multiple of the first step's |b| "
                         "(train: 1e-6, eval: 1e-8)")
    ap.add_argument("--time-budget", type=float, default=600.0,
                    help="train: stop generating data after this many seconds")
    ap.add_argument("--maxit", type=int, default=4000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--split", action="store_true", help="eval: the learned solve in the split form")
    ap.add_argument("--restart", type=int, default=30,
                    help="eval: GMRES restart length of the warm-started solve")
    ap.add_argument("--cold", action="store_true",
                    help="eval: also the warm-started solve's settings from zero")
    ap.add_argument("--model", default="mlp_model.npz")
    a = ap.parse_args(argv)
    say = lambda d: print(json.dumps(d), flush=True)
    if a.phase == "train":
        import torch
        seeds = _seeds(a.seeds or "101-120")
        steps = a.steps or 100
        rtol = a.rtol or 1e-6
        dev = torch.device("cuda:0")
        Xs, Ts = [], []
        t0 = time.perf_counter()
        used = []
        for s in seeds:
            if time.perf_counter() - t0 > a.time_budget:
                say({"time_budget_reached": True, "seeds_used": used})
                break
            w = configs.make_workload(a.config, seed=s)
            g = rbc3d_amd.from_workload(w)
            g.prepare_sing_int_on_wall()
            atol = (a.atol_rel or 1e-6) * rhs_norm(g)
            for X, T, info in series(g, w.cells, steps, a.dt, rtol, a.maxit, atol=atol):
                Xs.append(torch.from_numpy(X).to(dev))
                Ts.append(torch.from_numpy(T).to(dev))
                if info["step"] % 10 == 0 or info["step"] == steps - 1:
                    say({"seed": s, **info, "elapsed_s": round(time.perf_counter() - t0, 1)})
            used.append(s)
            g.close()
        Xd, Td = torch.cat(Xs), torch.cat(Ts)
        del Xs, Ts
        sel = torch.randperm(Xd.shape[0], device=dev)[: min(Xd.shape[0], 1 << 21)]
        mean, std, scale = norm_stats(Xd[sel].cpu().numpy(), Td[sel].cpu().numpy())
        m = rbc3d_amd.Mlp()
        m.set_params(init_params(np.random.default_rng(0)))
        m.set_norm(mean, std, scale)
        hist = train(m, Xd, Td, a.epochs, a.batch, a.lr, log=say)
        save_model(a.model, m, mean, std, scale,
                   {"config": a.config, "seeds": used, "steps": steps,
                    "dt": a.dt, "rtol": rtol, "epochs": a.epochs, "samples": int(Xd.shape[0]),
                    "loss": hist[-1]})
        say({"saved": a.model, "samples": int(Xd.shape[0]), "loss_first": hist[0],
             "loss_last": hist[-1], "elapsed_s": round(time.perf_counter() - t0, 1)})
        return 0
    seeds = _seeds(a.seeds or "201-205")
    steps = a.steps or 8
    rtol = a.rtol or 1e-8
    m = load_model(a.model)
    allrec = []
    for s in seeds:
        w = configs.make_workload(a.config, seed=s)
        g = rbc3d_amd.from_workload(w)
        g.prepare_sing_int_on_wall()
        atol = (a.atol_rel or 1e-8) * rhs_norm(g)
        recs = evaluate(g, w.cells, m, steps, a.dt, rtol, a.maxit,
                        log=lambda r, s=s: say({"seed": s, **r}), split=a.split, atol=atol,
                        restart=a.restart, cold=a.cold)
        allrec += recs
        g.close()
    ip = np.array([r["plain_iters"] for r in allrec], float)
    il = np.array([r["learned_iters"] for r in allrec], float)
    tp = np.array([r["plain_s"] for r in allrec])
    tl = np.array([r["learned_s"] for r in allrec])
    summ = {"summary": True, "steps": len(allrec), "plain_iters_mean": ip.mean(),
            "learned_iters_mean": il.mean(), "plain_s_mean": tp.mean(), "learned_s_mean": tl.mean(),
            "fallbacks": int(sum(r["fallback"] for r in allrec)),
            "vel_diff_max": max(r["vel_diff"] for r in allrec)}
    if a.cold:
        summ["cold_iters_mean"] = float(np.mean([r["cold_iters"] for r in allrec]))
        summ["cold_s_mean"] = float(np.mean([r["cold_s"] for r in allrec]))
    say(summ)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
