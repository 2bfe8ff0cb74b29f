"""Benchmark: simulated days per wall-second of the plain FP64 path on the 1M-node
3D heterogeneous Richards box (100 x 100 x 20 m, jittered 160 x 160 x 40 lattice,
Delaunay tetrahedra, log-normal Ks, four van Genuchten soil classes, water table
at 5 m, 0.02 m/day rain), BDF / Newton-GMRES / block-Jacobi ILU(0) on one MI355X.

A "step" is one accepted time step of the integrator (begin-step upwinding,
predictor, Newton iterations each with a residual, matrix-free GMRES with
block-Jacobi ILU(0) of the coloured FD Jacobian).  W untimed warm-up steps from t = 0,
then exactly K timed steps; value = simulated days advanced in the timed
steps / wall seconds.  Inputs live in HBM before the timed region starts.

Prints one JSON line (see the fields below).  Roofline: the residual's
dominant kernel (the element-flux kernel of the launch shape the autotuner
picks), its algorithmic bytes per launch divided by its average duration
measured with HIP events on the context stream over the timed steps; peak
8 TB/s.  The whole residual's B_F = 48 N + 28 T + N_edges (state, coordinates,
connectivity, conductivity, class, one upwind byte per edge) over its measured
time is reported beside it.  The FP32-preconditioner variant runs over the
same simulated window.  CPU baseline: the CPU oracle (a C restatement of the
reference's residual) timed on a bounded sample, scaled to sim-days/wall-s
(see "sample" in the output).
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
    ap.add_argument("--precon", default="bjilu0", choices=["bjilu0", "bjilu0_f32", "ilu0", "jacobi"])
    ap.add_argument("--block-rows", type=int, default=0, help="block-Jacobi rows per block (0 = library default)")
    ap.add_argument("--overlap", type=int, default=1, help="block-Jacobi overlap layers (restricted additive Schwarz)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    ap.add_argument("--launch", default="auto",
                    help="residual launch: 'auto' (autotuner at the initial state) or variant:tile_nodes:xcd_map, "
                         "variant 0 = element owner, 1 = LDS window")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the FP32-preconditioner run over the same simulated window")
    ap.add_argument("--cpu-nx", type=int, default=80, help="lattice size of the CPU sample box")
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


def fp32_window(args, prob, ctx, psi0, st0, st1, wall64, psi64):
    """The FP32-preconditioner variant over the same simulated window: a second
    integrator on the same context from t = 0 takes the same W untimed warm-up
    steps, then timed steps with the FP64 run's end time as stop time, so both
    runs end at the same t.  Reports its rate, the GMRES-iteration delta, the
    final-head difference to the FP64 run and both runs' mass-balance errors."""
    import torch
    from fvmpor_amd.diagnostics import inflow_rate, total_mass
    from fvmpor_amd.lib import Integrator, PRECON_BJILU0_F32
    tb = st1["t"]
    integ = Integrator(ctx, psi0, rtol=args.rtol, atol=args.atol, max_order=3, precon=PRECON_BJILU0_F32,
                       block_rows=args.block_rows, overlap=args.overlap)
    for _ in range(args.warmup):
        integ.step(1.0)
    torch.cuda.synchronize()
    s0 = integ.stats()
    w0 = time.perf_counter()
    while integ.stats()["t"] < tb:
        integ.step(tb)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    s1 = integ.stats()
    out = torch.empty_like(psi0)
    integ.advance(tb, out)
    torch.cuda.synchronize()
    psi32 = out.cpu().numpy()
    ta = s0["t"]
    m0, q = total_mass(prob, prob.psi0), inflow_rate(prob)

    def mb(psi):
        return (total_mass(prob, psi) - m0 - q * tb) / (q * tb)

    d = lambda k: s1[k] - s0[k]
    return {
        "value": (tb - ta) / wall,
        "unit": "sim-days/wall-s",
        "steps": d("steps"),
        "ms_per_step": wall / max(d("steps"), 1) * 1e3,
        "speedup_vs_fp64": wall64 / wall,
        "gmres_iters": d("gmres_iters"),
        "gmres_iters_fp64": st1["gmres_iters"] - st0["gmres_iters"],
        "newton_iters": d("newton_iters"),
        "head_rel_err_vs_fp64": float(np.abs(psi32 - psi64).max() / np.abs(psi64).max()),
        "mass_balance_err": float(mb(psi32)),
        "mass_balance_err_fp64": float(mb(psi64)),
        "t_window": [ta, tb],
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from fvmpor_amd.lib import BLOCK_ROWS, Context, Integrator, PRECON_BJILU0, PRECON_BJILU0_F32, PRECON_ILU0, PRECON_JACOBI
    from fvmpor_amd.problems import box3d_problem

    t0 = time.perf_counter()
    prob = box3d_problem(nx=args.nx, ny=args.ny, nz=args.nz)
    log(f"mesh N={prob.mesh.n_nodes} T={prob.mesh.n_elements} built in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    ctx = Context(prob)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    psi0 = torch.tensor(prob.psi0, device=dev)
    pc = {"bjilu0": PRECON_BJILU0, "bjilu0_f32": PRECON_BJILU0_F32, "ilu0": PRECON_ILU0,
          "jacobi": PRECON_JACOBI}[args.precon]
    integ = Integrator(ctx, psi0, rtol=args.rtol, atol=args.atol, max_order=3, precon=pc,
                       block_rows=args.block_rows, overlap=args.overlap)
    ncol, _ = ctx.colours()
    if args.launch == "auto":
        launch = ctx.autotune(0.0, psi0, torch.zeros_like(psi0))
    else:
        v, tn, xm = (int(x) for x in args.launch.split(":"))
        ctx.set_residual_launch(v, tn, xm)
        launch = {"variant": v, "tile_nodes": tn, "xcd_map": xm}
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f}s, {ncol} colours, N_edges={ctx.NE}, residual launch {launch}")
    T_END = 1.0
    for i in range(args.warmup):
        integ.step(T_END)
    torch.cuda.synchronize()
    st0 = integ.stats()
    ctx.profile(True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for i in range(args.steps):
        integ.step(T_END)
        if (i + 1) % 10 == 0:
            s = integ.stats()
            log(f"step {i + 1}/{args.steps}: t={s['t']:.4e} d, h={s['h_last']:.2e}, order {s['order_last']}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    st1 = integ.stats()
    ek_ms, ek_n = ctx.profile_get(0)
    rs_ms, rs_n = ctx.profile_get(1)
    pc_ms, pc_n = ctx.profile_get(2)
    ctx.profile(False)
    out64 = torch.empty_like(psi0)
    integ.advance(st1["t"], out64)
    torch.cuda.synchronize()
    psi_ref = out64.cpu().numpy()
    fp32 = None
    if not args.no_fp32 and args.precon == "bjilu0":
        fp32 = fp32_window(args, prob, ctx, psi0, st0, st1, wall, psi_ref)
    days = st1["t"] - st0["t"]
    N, T, NE = ctx.N, ctx.T, ctx.NE
    bf = 48 * N + 28 * T + NE  # one residual: psi, psi', F, xyz per node; conn, K, class per tet; upwind per edge
    owner = launch["variant"] == 0
    # dominant kernel's own algorithmic bytes: the element-owner kernel gathers xyz, psi, k_r once per node,
    # accumulates 8 B per node and streams conn, K, class, upwind bits per tet; the window kernel is the
    # whole residual but the van Genuchten pass
    kbytes = 48 * N + 26 * T if owner else bf
    kname = "owner_kernel<3> (residual element fluxes)" if owner else "tile_lds_kernel<3> (residual)"
    ek_avg_s = ek_ms / max(ek_n, 1) * 1e-3
    achieved = kbytes / ek_avg_s / 1e9
    rs_avg_s = rs_ms / max(rs_n, 1) * 1e-3
    window_res = st1["residuals"] - st0["residuals"]
    log(f"timed {args.steps} steps: {days:.4e} days in {wall:.2f}s; element kernel {ek_avg_s*1e6:.1f} us x {ek_n}, "
        f"residual {rs_ms/max(rs_n,1)*1e3:.1f} us x {rs_n}, precon solve {pc_ms/max(pc_n,1)*1e3:.1f} us x {pc_n}")
    log(f"stats {st1}")
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
            "workload": f"plain FP64 3D box {args.nx}x{args.ny}x{args.nz} lattice ({N} nodes, {T} tets), "
                        f"BDF(<=3) Newton-GMRES {args.precon}, rtol=atol={args.rtol:g}, timed steps "
                        f"{args.warmup + 1}..{args.warmup + args.steps} from t=0",
            "nodes": N, "elements": T, "edges": NE, "precon": args.precon,
            "block_rows": args.block_rows or BLOCK_ROWS, "overlap": args.overlap,
            "sim_days_timed": days, "t_start": st0["t"], "t_end": st1["t"],
            "newton_iters": st1["newton_iters"] - st0["newton_iters"],
            "gmres_iters": st1["gmres_iters"] - st0["gmres_iters"],
            "precon_setups": st1["precon_setups"] - st0["precon_setups"],
            "residual_evals": window_res,
            "residual_us": rs_ms / max(rs_n, 1) * 1e3,
            "residual_launch": launch,
            "residual_GBps_BF": bf / rs_avg_s / 1e9,
            "precon_solve_us": pc_ms / max(pc_n, 1) * 1e3,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": kbytes,
            "avg_launch_us": ek_avg_s * 1e6,
            "launches": ek_n,
        },
    }
    if fp32 is not None:
        out["fp32_precon"] = fp32
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args, days, window_res, N)
        except Exception as exc:  # the GPU measurement stands on its own
            log(f"cpu baseline failed: {exc!r}")
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
