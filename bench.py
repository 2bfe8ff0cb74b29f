"""Benchmark: simulated days per wall-second on the 1M-node 3D heterogeneous
Richards box (100 x 100 x 20 m, jittered 160 x 160 x 40 lattice, Delaunay
tetrahedra, log-normal Ks, four van Genuchten soil classes, water table at
5 m, 0.02 m/day rain), BDF / Newton-GMRES / block-Jacobi ILU(0) on one MI355X:
FP32 preconditioner plus learned Newton initial iterate against the plain FP64
path, at rtol = atol = 1e-6.

A "step" is one accepted time step of the integrator (begin-step upwinding,
predictor, Newton iterations each with a residual, matrix-free GMRES with
block-Jacobi ILU(0) of the FD Jacobian, whose quotients use the coloured
method's increments and are formed element by element).  Headline run: FP32
preconditioner + learned initial iterate (the integrator backs off from the
network when its proposals do not halve the preconditioned first residual;
"learned_start_used" counts its tries); W untimed warm-up steps from t = 0,
then exactly K timed steps; value = simulated days advanced in the timed steps
/ wall seconds.  The plain FP64 run and the FP32-preconditioner-only run
integrate the same simulated window (W warm-up steps from t = 0, then to the
headline run's end time) and are reported beside it with their rates, GMRES
iterations, final-head difference and mass-balance error.  The network is
trained before the runs (untimed) by the library's own kernels on pairs the
plain solver records on a fresh-seed 200k-node box of the same statistics.
Inputs live in HBM before the timed regions start.

Prints one JSON line (see the fields below).  Roofline: the residual's
dominant kernel (the element-flux kernel of the launch shape the autotuner
picks), its algorithmic bytes per launch divided by its average duration
measured with HIP events on the context stream over the headline run's timed
steps; peak 8 TB/s.  The whole residual's B_F = 48 N + 28 T + N_edges (state,
coordinates, connectivity, conductivity, class, one upwind byte per edge) over
its measured time is reported beside it.  CPU baseline: the CPU oracle (a C
restatement of the reference's residual) timed on a bounded sample, scaled to
sim-days/wall-s (see "sample" in the output).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "sim-days/wall-s, 1M-node 3D Richards: FP32-precon+learned vs plain FP64 @1e-6"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=160)
    ap.add_argument("--ny", type=int, default=160)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--atol", type=float, default=1e-6)
    ap.add_argument("--precon", default="bjilu0", choices=["bjilu0", "ilu0", "jacobi"],
                    help="preconditioner of the plain FP64 run")
    ap.add_argument("--block-rows", type=int, default=1024,
                    help="block-Jacobi rows per block (0 = library default; 1024 measured fastest on this box)")
    ap.add_argument("--overlap", type=int, default=1, help="block-Jacobi overlap layers (restricted additive Schwarz)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    ap.add_argument("--launch", default="auto",
                    help="residual launch: 'auto' (autotuner at the initial state) or variant:tile_nodes:xcd_map, "
                         "variant 0 = element owner, 1 = LDS window")
    ap.add_argument("--no-fp32", action="store_true",
                    help="headline run with the FP64 preconditioner (no FP32 variant)")
    ap.add_argument("--cpu-nx", type=int, default=80, help="lattice size of the CPU sample box")
    ap.add_argument("--no-learned", action="store_true", help="headline run without the learned initial iterate")
    ap.add_argument("--train-nz", type=int, default=23, help="training box lattice 4nz x 4nz x nz")
    ap.add_argument("--train-t", type=float, default=0.25, help="simulated days of training-pair recording")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--probe-copy", type=int, default=0, help="MiB of a plain device copy after the runs")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(args, gpu_window_days, gpu_window_residuals, n_nodes_gpu):
    """Oracle residual evaluations on a smaller box of the same generator and
    statistics, time per node-evaluation scaled to the full mesh, times the
    residual evaluations the GPU needed for the timed window."""
    from fvmpor_amd.problems import box3d_problem
    from oracle.oracle import OracleProblem
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nx = args.cpu_nx
    prob = box3d_problem(nx=nx, ny=nx, nz=max(4, nx // 4))
    op = OracleProblem(prob)
    psi = prob.psi0 - 0.5
    psip = np.zeros_like(psi)
    op.begin_step(0.0, psi, psip)
    op.residual(0.0, psi, psip)
    n_eval, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 10.0 or n_eval < 3:
        op.residual(0.0, psi, psip)
        n_eval += 1
    per_eval = (time.perf_counter() - t0) / n_eval
    per_eval_full = per_eval * n_nodes_gpu / prob.mesh.n_nodes
    cpu_seconds = per_eval_full * gpu_window_residuals
    return {
        "value": gpu_window_days / cpu_seconds,
        "unit": "sim-days/wall-s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{n_eval} CPU-oracle residual evaluations on the {prob.mesh.n_nodes}-node box of the same "
                   f"generator ({per_eval*1e3:.1f} ms each), scaled by node count to {per_eval_full:.2f} s per "
                   f"1M-node residual and by the {gpu_window_residuals} residual evaluations the GPU run "
                   f"needed for the timed steps; preconditioner and Krylov vector work not counted"),
    }


def measured_traffic(n_nodes, n_elements, launch):
    """HBM bytes per launch of the element kernel measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE counter runs (scaled by a plain copy of known size)
    for this workload and launch shape, from profiles/; None when no
    measurement matches."""
    path = os.path.join(ROOT, "profiles", "owner_kernel_box3d_1M_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = json.load(f)
    if (m.get("nodes"), m.get("elements")) != (n_nodes, n_elements):
        return None
    key = [launch["variant"], launch["tile_nodes"], launch["xcd_map"]]
    for r in m.get("measurements", []):
        if r.get("launch") == key:
            return r
    return None


def run_window(args, prob, ctx, psi0, precon, t_stop=None, steps=None, learned=None, profile=False):
    """One integrator from t = 0 on the context: W untimed warm-up steps, then
    either exactly `steps` timed steps or steps up to t_stop.  Returns the
    window's stats, its wall time, the final heads and the profile figures."""
    import torch
    from fvmpor_amd.lib import Integrator
    integ = Integrator(ctx, psi0, rtol=args.rtol, atol=args.atol, max_order=3, precon=precon,
                       block_rows=args.block_rows, overlap=args.overlap)
    if learned is not None:
        integ.set_learned(*learned)
    T_END = 1.0
    def step(t_stop):
        rc = integ.step(t_stop)
        if rc != 0:
            raise RuntimeError(f"integrator step failed with status {rc} at t={integ.stats()['t']:.4g}")
    for _ in range(args.warmup):
        step(T_END)
    torch.cuda.synchronize()
    s0 = integ.stats()
    if profile:
        ctx.profile(True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    i = 0
    while (steps is not None and i < steps) or (steps is None and integ.stats()["t"] < t_stop):
        step(T_END if steps is not None else t_stop)
        i += 1
        if i % 10 == 0:
            s = integ.stats()
            log(f"  step {i}: t={s['t']:.4e} d, h={s['h_last']:.2e}, order {s['order_last']}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    s1 = integ.stats()
    prof = None
    if profile:
        prof = {k: ctx.profile_get(k) for k in (0, 1, 2, 3)}
        ctx.profile(False)
    out = torch.empty_like(psi0)
    integ.advance(s1["t"], out)
    torch.cuda.synchronize()
    psi = out.cpu().numpy()
    integ.close()
    d = {k: s1[k] - s0[k] for k in ("steps", "newton_iters", "gmres_iters", "precon_setups", "residuals",
                                     "learned_tried", "learned_used")}
    return {"t0": s0["t"], "t1": s1["t"], "wall": wall, "d": d, "psi": psi, "prof": prof}


def train_network(args):
    """Pairs from the plain solver on a fresh-seed box of the same statistics
    (the library's recording mode), then Adam on the device."""
    from fvmpor_amd.learn import train_on
    from fvmpor_amd.problems import box3d_problem
    nz = args.train_nz
    t0 = time.perf_counter()
    tp = box3d_problem(nx=4 * nz, ny=4 * nz, nz=nz, seed_mesh=501, seed_field=502, seed_layers=503, sigma_lnK=1.0,
                       rain=0.02)
    log(f"training box N={tp.mesh.n_nodes} built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    params, hist, npairs, _ = train_on([tp], args.train_t, epochs=args.epochs, lr=1e-3,
                                       log=lambda m: log("  " + m) if m.startswith("epoch") and
                                       int(m.split()[1][:-1]) % 10 == 9 else None)
    info = {"train_nodes": tp.mesh.n_nodes, "pairs": npairs, "epochs": args.epochs, "loss_first": hist[0],
            "loss_last": hist[-1], "train_s": time.perf_counter() - t0}
    log(f"network trained: {info}")
    return params, info


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from fvmpor_amd.diagnostics import inflow_rate, total_mass
    from fvmpor_amd.learn import node_graph, static_features
    from fvmpor_amd.lib import (BLOCK_ROWS, PRECON_BJILU0, PRECON_BJILU0_F32, PRECON_ILU0, PRECON_JACOBI, Context,
                                Mpnn)
    from fvmpor_amd.problems import box3d_problem

    params, train_info = (None, None) if args.no_learned else train_network(args)
    t0 = time.perf_counter()
    prob = box3d_problem(nx=args.nx, ny=args.ny, nz=args.nz)
    log(f"mesh N={prob.mesh.n_nodes} T={prob.mesh.n_elements} built in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    ctx = Context(prob)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    psi0 = torch.tensor(prob.psi0, device=dev)
    pc64 = {"bjilu0": PRECON_BJILU0, "ilu0": PRECON_ILU0, "jacobi": PRECON_JACOBI}[args.precon]
    ncol, _ = ctx.colours()
    if args.launch == "auto":
        launch = ctx.autotune(0.0, psi0, torch.zeros_like(psi0))
    else:
        v, tn, xm = (int(x) for x in args.launch.split(":"))
        ctx.set_residual_launch(v, tn, xm)
        launch = {"variant": v, "tile_nodes": tn, "xcd_map": xm}
    learned = None
    if params is not None:
        ia, ja = node_graph(prob.mesh)
        model = Mpnn(ia, ja, layers=3)
        model.set_params(params)
        stat = torch.tensor(static_features(prob), device=dev)
        torch.cuda.synchronize()
        learned = (model, stat)
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f}s, {ncol} colours, N_edges={ctx.NE}, residual launch {launch}")

    head_pc = PRECON_BJILU0_F32 if args.precon == "bjilu0" and not args.no_fp32 else pc64
    head_name = ("FP32-precon" if head_pc == PRECON_BJILU0_F32 else "plain FP64") + \
        (" + learned initial iterate" if learned is not None else "")
    log(f"headline run: {head_name}")
    head = run_window(args, prob, ctx, psi0, head_pc, steps=args.steps, learned=learned, profile=True)
    tb = head["t1"]
    log(f"plain FP64 run over the same window")
    plain = run_window(args, prob, ctx, psi0, pc64, t_stop=tb)
    fp32 = None
    if learned is not None and head_pc == PRECON_BJILU0_F32:
        log(f"FP32-precon run (no network) over the same window")
        fp32 = run_window(args, prob, ctx, psi0, PRECON_BJILU0_F32, t_stop=tb)

    m0, q = total_mass(prob, prob.psi0), inflow_rate(prob)

    def mb(psi):
        return float((total_mass(prob, psi) - m0 - q * tb) / (q * tb))

    def summary(w, ref=None):
        r = {"value": (w["t1"] - w["t0"]) / w["wall"], "unit": "sim-days/wall-s", "wall_s": w["wall"],
             "t_window": [w["t0"], w["t1"]], "steps": w["d"]["steps"],
             "ms_per_step": w["wall"] / max(w["d"]["steps"], 1) * 1e3, "newton_iters": w["d"]["newton_iters"],
             "gmres_iters": w["d"]["gmres_iters"], "precon_setups": w["d"]["precon_setups"],
             "residual_evals": w["d"]["residuals"], "mass_balance_err": mb(w["psi"])}
        if ref is not None:
            r["head_rel_err_vs_plain_fp64"] = float(np.abs(w["psi"] - ref["psi"]).max() / np.abs(ref["psi"]).max())
        if w["d"]["learned_tried"]:
            r["learned_start_used"] = [w["d"]["learned_used"], w["d"]["learned_tried"]]
        return r

    days = head["t1"] - head["t0"]
    wall = head["wall"]
    N, T, NE = ctx.N, ctx.T, ctx.NE
    (ek_ms, ek_n), (rs_ms, rs_n), (pc_ms, pc_n), (jk_ms, jk_n) = (head["prof"][k] for k in (0, 1, 2, 3))
    bf = 48 * N + 28 * T + NE  # one residual: psi, psi', F, xyz per node; conn, K, class per tet; upwind per edge
    owner = launch["variant"] == 0
    # dominant kernel's own algorithmic bytes: the element-owner kernel gathers xyz, psi, k_r once per node,
    # accumulates 8 B per node and streams conn, K, class, upwind bits per tet; the window kernel is the
    # whole residual but the van Genuchten pass
    kbytes = 48 * N + 26 * T if owner else bf
    kname = "owner_kernel<3> (residual element fluxes)" if owner else "tile_lds_kernel<3> (residual)"
    ek_avg_s = ek_ms / max(ek_n, 1) * 1e-3
    achieved = kbytes / ek_avg_s / 1e9
    traffic = measured_traffic(N, T, launch)
    rs_avg_s = rs_ms / max(rs_n, 1) * 1e-3
    log(f"timed {args.steps} steps: {days:.4e} days in {wall:.2f}s; element kernel {ek_avg_s*1e6:.1f} us x {ek_n} "
        f"(J v: {jk_ms/max(jk_n,1)*1e3:.1f} us x {jk_n}), "
        f"residual {rs_ms/max(rs_n,1)*1e3:.1f} us x {rs_n}, precon solve {pc_ms/max(pc_n,1)*1e3:.1f} us x {pc_n}")
    plain_s = summary(plain)
    head_s = summary(head, plain)
    out = {
        "metric": METRIC,
        "value": days / wall,
        "unit": "sim-days/wall-s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded jittered-lattice Delaunay mesh, FFT log-normal Ks, random soil layers",
        "config": {
            "workload": f"3D box {args.nx}x{args.ny}x{args.nz} lattice ({N} nodes, {T} tets), BDF(<=3) "
                        f"Newton-GMRES, block-Jacobi ILU(0) ({args.block_rows or BLOCK_ROWS}-row blocks, "
                        f"{args.overlap} overlap layer(s)), rtol=atol={args.rtol:g}; headline: {head_name}, "
                        f"timed steps {args.warmup + 1}..{args.warmup + args.steps} from t=0",
            "nodes": N, "elements": T, "edges": NE, "precon_plain": args.precon,
            "block_rows": args.block_rows or BLOCK_ROWS, "overlap": args.overlap,
            "sim_days_timed": days, "t_start": head["t0"], "t_end": head["t1"],
            "newton_iters": head["d"]["newton_iters"], "gmres_iters": head["d"]["gmres_iters"],
            "precon_setups": head["d"]["precon_setups"], "residual_evals": head["d"]["residuals"],
            "residual_us": rs_ms / max(rs_n, 1) * 1e3,
            "residual_launch": launch,
            "residual_GBps_BF": bf / rs_avg_s / 1e9,
            "precon_solve_us": pc_ms / max(pc_n, 1) * 1e3,
            "jv_element_kernel_us": jk_ms / max(jk_n, 1) * 1e3,
            "jv_element_kernel_launches": jk_n,
            "parallelism": "1 GPU",
        },
        "headline": dict(head_s, name=head_name),
        "plain_fp64": plain_s,
        "speedup_vs_plain_fp64": head_s["value"] / plain_s["value"],
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_source": "profiles/owner_kernel_box3d_1M_traffic.json" if traffic else None,
            "bytes_per_launch": kbytes,
            "avg_launch_us": ek_avg_s * 1e6,
            "launches": ek_n,
        },
    }
    if fp32 is not None:
        out["fp32_precon_only"] = summary(fp32, plain)
    if train_info is not None:
        out["network"] = train_info
    if args.probe_copy:
        # a plain device copy of known size, for scaling the HBM byte counters of a profiler run
        a = torch.empty(args.probe_copy << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args, days, head["d"]["residuals"], N)
        except Exception as exc:  # the GPU measurement stands on its own
            log(f"cpu baseline failed: {exc!r}")
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
