"""Batched Monte Carlo ensembles of independent Richards problems.

An ensemble is many independent realisations (own mesh, own log-normal K
field, own infiltration rate).  On one GPU they are advanced together as one
block-diagonal problem: the instances' nodes and elements are stored back to
back in one context, so every residual, J v, Gram-Schmidt pass, colouring
sweep and preconditioner launch covers all instances at once (one launch per
kernel per ensemble instead of one per instance), node tiles and
block-Jacobi blocks never couple different instances (there are no entries
between them), and the integrator measures every error and Newton norm per
instance (the maximum over instances of each one's weighted RMS,
fvp_integrator_set_segments) while all instances share t, h and the BDF order.

The reference runs one problem per process; this module has no counterpart in
it.  Instances of one ensemble must share the spatial dimension (the residual
kernel is specialised per dimension); a mixed 2D/3D draw runs as two batches.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from .meshgen import Mesh, box_mesh_2d, gaussian_field
from .problems import SOIL_CLASSES, Problem, _zone_row, box3d_problem


@dataclasses.dataclass
class Draw:
    index: int
    dim: int
    n_nodes: int
    sigma_lnK: float
    rain: float
    seed: int


def draw_instances(n=256, seed=4242, n_min=5000, n_max=60000):
    """Per-instance parameters: node count ~ U[n_min, n_max], sigma_lnK ~
    U[0.5, 2.0], infiltration ~ U[0.01, 0.1] m/day, jitter seed 1000 + i,
    2D for even seeds and 3D for odd ones."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s = 1000 + i
        out.append(Draw(i, 2 if s % 2 == 0 else 3, int(rng.integers(n_min, n_max + 1)),
                        float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.01, 0.1)), s))
    return out


def slice_instance(d: Draw) -> Problem:
    """2D slice 10 m x 5 m of ~n_nodes nodes (jittered lattice, Delaunay),
    homogeneous loam with a log-normal K field (mean ln 0.25 m/day, 2 m / 0.5 m
    correlation), psi0 = 1 - y, infiltration on the strip x in [4, 6] m."""
    nx = max(5, int(round(np.sqrt(2.0 * d.n_nodes))))
    ny = max(3, int(round(d.n_nodes / nx)))
    mesh = box_mesh_2d(nx, ny, 10.0, 5.0, seed=d.seed)
    spacing = (10.0 / (nx - 1), 5.0 / (ny - 1))
    field = gaussian_field((nx, ny), spacing, (2.0, 0.5), d.sigma_lnK, d.seed + 1) + np.log(0.25)
    cen = mesh.xyz[mesh.conn].mean(axis=1)
    idx = [np.clip(np.rint(cen[:, a] / spacing[a]).astype(np.int64), 0, (nx, ny)[a] - 1) for a in range(2)]
    alpha, n, tr, ts = SOIL_CLASSES["loam"]
    zones = np.array([_zone_row(1.0, ts, alpha, n, tr / ts, 1e-5)])
    bc_type = np.array([0, 3, 3], dtype=np.int32)
    bc = np.zeros((3, 6))
    bc[2, 0] = -d.rain
    return Problem(f"slice{d.index}", mesh, zones, np.zeros(mesh.n_elements, dtype=np.int32),
                   np.exp(field[idx[0], idx[1]]), bc_type, bc, 1000.0, 1.0 - mesh.xyz[:, 1], 0.5, "day")


def box_instance(d: Draw) -> Problem:
    """3D box of the 1M-node workload's statistics at ~n_nodes nodes
    (4 nz x 4 nz x nz lattice)."""
    nz = max(4, int(round((d.n_nodes / 16.0) ** (1.0 / 3.0))))
    p = box3d_problem(nx=4 * nz, ny=4 * nz, nz=nz, seed_mesh=d.seed, seed_field=d.seed + 1,
                      seed_layers=d.seed + 2, sigma_lnK=d.sigma_lnK, rain=d.rain)
    p.name = f"box{d.index}"
    p.t_end = 0.5
    return p


def make_instance(d: Draw) -> Problem:
    return slice_instance(d) if d.dim == 2 else box_instance(d)


def union_problem(problems):
    """One problem holding the instances back to back.  Tag 1 (no flow) is
    shared, each instance's tag 2 (its infiltration face) becomes tag 2 + i;
    zone tables are merged by row.  Returns (problem, node offsets)."""
    dims = {p.mesh.dim for p in problems}
    if len(dims) != 1:
        raise ValueError("an ensemble batch needs instances of one dimension")
    dim = dims.pop()
    rho0 = problems[0].rho0
    if any(p.rho0 != rho0 for p in problems):
        raise ValueError("instances must share rho0")
    zone_rows, ezone, xyz, conn, ftags, nptr, ntags, zfield, ekmul, psi0 = [], [], [], [], [], [], [], [], [], []
    offsets = [0]
    bc_type = [0, 3]
    bc = [np.zeros(6), np.zeros(6)]
    for i, p in enumerate(problems):
        if not (np.array_equal(p.bc_type, [0, 3, 3]) and p.bc[1, 0] == 0.0):
            raise ValueError("instances must have tags 0 (interior), 1 (no flow), 2 (prescribed flux)")
        zmap = []
        for row in p.zones:
            for k, r in enumerate(zone_rows):
                if np.array_equal(r, row):
                    zmap.append(k)
                    break
            else:
                zone_rows.append(row.copy())
                zmap.append(len(zone_rows) - 1)
        off = offsets[-1]
        m = p.mesh
        ezone.append(np.asarray(zmap, dtype=np.int32)[p.ezone])
        xyz.append(m.xyz)
        conn.append(m.conn + off)
        ft = m.face_tags.copy()
        ft[ft == 2] = 2 + i
        ftags.append(ft)
        nt = m.node_tags.copy()
        nt[nt == 2] = 2 + i
        ntags.append(nt)
        nptr.append(m.node_tag_ptr[1:] + (nptr[-1][-1] if nptr else 0))
        zfield.append(m.zone)
        ekmul.append(p.ekmul)
        psi0.append(p.psi0)
        bc_type.append(3)
        bc.append(p.bc[2].copy())
        offsets.append(off + m.n_nodes)
    if len(zone_rows) > 16:
        raise ValueError("more than 16 distinct zones")
    mesh = Mesh(dim, np.vstack(xyz), np.vstack(conn).astype(np.int32), np.vstack(ftags).astype(np.int32),
                np.concatenate([[0]] + nptr).astype(np.int32), np.concatenate(ntags).astype(np.int32),
                np.concatenate(zfield).astype(np.int32))
    prob = Problem("ensemble", mesh, np.array(zone_rows), np.concatenate(ezone), np.concatenate(ekmul),
                   np.array(bc_type, dtype=np.int32), np.array(bc), rho0, np.concatenate(psi0),
                   problems[0].t_end, problems[0].time_unit)
    return prob, np.array(offsets, dtype=np.int64)


def run_batch(problems, t_end=0.5, rtol=1e-6, atol=1e-6, precon=None, block_rows=0, stream=None, progress=20.0,
              log=None):
    """Advance one batch (instances of one dimension) to t_end on cuda:0.
    Returns per-instance final heads, the integrator statistics and the wall
    time of the advance (set-up excluded)."""
    import torch

    from .lib import PRECON_BJILU0_F32, Context, Integrator
    union, offsets = union_problem(problems)
    if log is not None:
        log(f"  union of {len(problems)} instances: {union.mesh.n_nodes} nodes, {union.mesh.n_elements} elements")
    t_setup = time.perf_counter()
    ctx = Context(union)
    if stream is not None:
        ctx.set_stream(stream)
    psi0 = torch.tensor(union.psi0, device="cuda:0")
    integ = Integrator(ctx, psi0, rtol=rtol, atol=atol, max_order=3,
                       precon=PRECON_BJILU0_F32 if precon is None else precon, block_rows=block_rows)
    integ.set_segments(offsets)
    if log is not None:
        log(f"  set-up {time.perf_counter() - t_setup:.1f}s")
    out = torch.empty_like(psi0)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    last = w0
    while integ.stats()["t"] < t_end * (1 - 1e-12):
        integ.step(t_end)
        if log is not None and time.perf_counter() - last > progress:  # long batches report progress
            last = time.perf_counter()
            st = integ.stats()
            log(f"  t={st['t']:.4g} h={st['h_last']:.3g}: {st['steps']} steps, {st['gmres_iters']} GMRES "
                f"iterations, {last - w0:.1f}s")
    integ.advance(t_end, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    psi = out.cpu().numpy()
    heads = [psi[offsets[i]:offsets[i + 1]] for i in range(len(problems))]
    return heads, integ.stats(), wall


def run_one(problem, t_end=0.5, rtol=1e-6, atol=1e-6, precon=None, overlap=1, budget_s=None):
    """Advance one instance on its own context (own HIP stream) to t_end, or
    until budget_s wall seconds have passed.  Returns a result dict."""
    import torch

    from .lib import PRECON_BJILU0_F32, Context, Integrator
    w0 = time.perf_counter()
    ctx = Context(problem)
    psi0 = torch.tensor(problem.psi0, device="cuda:0")
    torch.cuda.synchronize()
    integ = Integrator(ctx, psi0, rtol=rtol, atol=atol, max_order=3,
                       precon=PRECON_BJILU0_F32 if precon is None else precon, overlap=overlap)
    w1 = time.perf_counter()
    status = "done"
    while integ.stats()["t"] < t_end * (1 - 1e-12):
        if budget_s is not None and time.perf_counter() - w1 > budget_s:
            status = "budget"
            break
        rc = integ.step(t_end)
        if rc != 0:
            status = f"failed ({rc})"
            break
    out = torch.empty_like(psi0)
    ctx.synchronize()
    st = integ.stats()
    integ.advance(st["t"], out)
    ctx.synchronize()
    heads = out.cpu().numpy()
    integ.close()
    ctx.close()
    return {"t": st["t"], "status": status, "steps": st["steps"], "newton_fails": st["newton_fails"],
            "error_fails": st["error_fails"], "gmres_iters": st["gmres_iters"], "setup_s": w1 - w0,
            "run_s": time.perf_counter() - w1, "finite": bool(np.isfinite(heads).all()), "heads": heads}


def run_concurrent(problems, t_end=0.5, streams=4, budget_s=None, log=None, **kw):
    """Every instance on its own context and HIP stream, `streams` of them in
    flight at once (one host thread each; the library releases the
    interpreter during its calls), largest first.  Each instance keeps its own
    step size, order and Newton / linear-solver history, so a stiff instance
    slows only itself.  Returns the per-instance results and the wall time."""
    from concurrent.futures import ThreadPoolExecutor
    order = sorted(range(len(problems)), key=lambda i: -problems[i].mesh.n_nodes)
    res = [None] * len(problems)
    w0 = time.perf_counter()

    def work(i):
        r = run_one(problems[i], t_end=t_end, budget_s=budget_s, **kw)
        if log is not None:
            log(f"  instance {i}: {problems[i].mesh.n_nodes} nodes, t={r['t']:.4g} ({r['status']}), "
                f"{r['steps']} steps, {r['newton_fails']} Newton failures, {r['run_s']:.1f}s")
        return i, r

    with ThreadPoolExecutor(max_workers=streams) as pool:
        for i, r in pool.map(work, order):
            res[i] = r
    return res, time.perf_counter() - w0


def _build(d):
    return make_instance(d)


def main(argv=None):
    """Run a Monte Carlo ensemble to t_end on cuda:0 and print one JSON line:
    python -m fvmpor_amd.ensemble [--n 256] [--t-end 0.5] [--mode concurrent|batch]
    [--streams 4] [--budget 30]."""
    import argparse
    import json
    import sys
    from concurrent.futures import ProcessPoolExecutor

    import torch

    from .lib import PRECON_BJILU0_F32
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--t-end", type=float, default=0.5)
    ap.add_argument("--n-min", type=int, default=5000)
    ap.add_argument("--n-max", type=int, default=60000)
    ap.add_argument("--workers", type=int, default=16, help="host processes building the meshes")
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--mode", default="concurrent", choices=["concurrent", "batch"])
    ap.add_argument("--streams", type=int, default=4, help="instances in flight (concurrent mode)")
    ap.add_argument("--budget", type=float, default=30.0, help="wall seconds per instance (concurrent mode)")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)
    draws = draw_instances(args.n, n_min=args.n_min, n_max=args.n_max)
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=args.workers) as pool:
        probs = list(pool.map(_build, draws, chunksize=4))
    log(f"built {len(probs)} meshes in {time.perf_counter() - t0:.1f}s")
    out = {"instances": len(probs), "t_end": args.t_end, "mode": args.mode, "rtol": args.rtol,
           "nodes": int(sum(p.mesh.n_nodes for p in probs))}
    if args.mode == "concurrent":
        res, wall = run_concurrent(probs, t_end=args.t_end, streams=args.streams, budget_s=args.budget, log=log,
                                   rtol=args.rtol, atol=args.rtol)
        done = [i for i, r in enumerate(res) if r["status"] == "done"]
        days = sum(r["t"] for r in res)
        out.update({
            "streams": args.streams, "budget_s_per_instance": args.budget, "wall_s": wall,
            "completed": len(done), "instance_days": days,
            "instance_days_per_s": days / wall,
            "node_days_per_s": sum(r["t"] * p.mesh.n_nodes for r, p in zip(res, probs)) / wall,
            "all_finite": bool(all(r["finite"] for r in res)),
            "steps": int(sum(r["steps"] for r in res)), "newton_fails": int(sum(r["newton_fails"] for r in res)),
            "gmres_iters": int(sum(r["gmres_iters"] for r in res)),
            "stragglers": [{"index": draws[i].index, "dim": draws[i].dim, "nodes": probs[i].mesh.n_nodes,
                            "sigma_lnK": round(draws[i].sigma_lnK, 3), "rain": round(draws[i].rain, 4),
                            "t": res[i]["t"], "status": res[i]["status"], "steps": res[i]["steps"],
                            "newton_fails": res[i]["newton_fails"]}
                           for i in range(len(res)) if res[i]["status"] != "done"],
        })
    else:
        out["batches"] = []
        wall_total = 0.0
        for dim in (2, 3):
            batch = [p for p in probs if p.mesh.dim == dim]
            if not batch:
                continue
            heads, st, wall = run_batch(batch, t_end=args.t_end, rtol=args.rtol, atol=args.rtol,
                                        precon=PRECON_BJILU0_F32, log=log)
            torch.cuda.synchronize()
            nodes = sum(p.mesh.n_nodes for p in batch)
            wall_total += wall
            out["batches"].append({"dim": dim, "instances": len(batch), "nodes": nodes, "wall_s": wall,
                                   "steps": st["steps"], "newton_iters": st["newton_iters"],
                                   "gmres_iters": st["gmres_iters"], "precon_setups": st["precon_setups"],
                                   "finite": bool(all(np.isfinite(h).all() for h in heads))})
            log(f"{dim}D batch: {len(batch)} instances, {nodes} nodes, {wall:.2f}s, {st['steps']} steps")
        out["wall_s"] = wall_total
        out["instance_days_per_s"] = len(probs) * args.t_end / wall_total
        out["node_days_per_s"] = out["nodes"] * args.t_end / wall_total
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
