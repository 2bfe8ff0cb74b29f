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
+ OpenMP restatement of the reference's FP64 path) on the same 1M-node box and
the host's cores: the median residual time, and the whole Newton-GMRES path
over the first --cpu-steps steps of the plain FP64 run's timed window (see
"sample" in the output), with the GPU's rate on the same steps beside it.

"full_run_1day" holds the metric's literal run, t = 0 to 1 day, for each
configuration, and "gate" the metric's conditions (final-head error <= 1e-6
against the plain FP64 run, >= 3x over it, the learned start not slower than
the FP32 preconditioner alone).
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
# the CPU baseline's OpenMP threads stay on neighbouring cores (read when the runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")

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
    ap.add_argument("--precon", default="bjilu0", choices=["bjilu0", "ilu0", "jacobi", "schur"],
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
    ap.add_argument("--no-full", action="store_true", help="skip the full runs to t = 1 day")
    ap.add_argument("--order", default="tile", choices=["tile", "rcm", "lattice"],
                    help="node numbering of the box: compact tiles (default), reverse Cuthill-McKee, generator order")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU oracle's Newton-GMRES sample")
    ap.add_argument("--no-learned", action="store_true", help="headline run without the learned initial iterate")
    ap.add_argument("--train-nz", type=int, default=23, help="training box lattice 4nz x 4nz x nz")
    ap.add_argument("--train-t", type=float, default=0.25, help="simulated days of training-pair recording")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--probe-copy", type=int, default=0, help="MiB of a plain device copy after the runs")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, prob, plain, gpu_res_us):
    """The CPU oracle (a C + OpenMP restatement of the reference's FP64 path)
    on the benchmark's own 1M-node box, in this process on the host's cores:
    (1) the median of >= 20 residual evaluations; (2) the whole Newton-GMRES
    path -- coloured FD Jacobian, ILU(0), GMRES(20) with FD J v, IDA's Newton
    and linear tolerances and Jacobian-refresh policy, its own upwinding per
    step -- over the first steps of the plain FP64 run's timed window (same
    starting state, same step sizes), over
    --cpu-steps steps (default 2, about 10-30 s of CPU work); value =
    simulated days of those steps / wall seconds.  The replay starts from one
    stored solution, so its first step is BDF1 and later ones follow the
    GPU's orders as far as the history allows."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    op = O.OracleProblem(prob)
    build_s = time.perf_counter() - t0
    psi = plain["psi_start"]
    psip = np.zeros_like(psi)
    op.begin_step(plain["t0"], psi, psip)
    op.residual(plain["t0"], psi, psip)
    times = []
    while len(times) < 20 or sum(times) < 3.0:
        a = time.perf_counter()
        op.residual(plain["t0"], psi, psip)
        times.append(time.perf_counter() - a)
    res_s = float(np.median(times))
    # the 2D slice (the CPU-runnable configuration): residual median
    from fvmpor_amd.problems import slice2d_problem
    sp_ = slice2d_problem()
    sop = O.OracleProblem(sp_)
    spsi, spsip = sp_.psi0.copy(), np.zeros(sp_.mesh.n_nodes)
    sop.begin_step(0.0, spsi, spsip)
    stimes = []
    while len(stimes) < 20 or sum(stimes) < 1.0:
        a = time.perf_counter()
        sop.residual(0.0, spsi, spsip)
        stimes.append(time.perf_counter() - a)
    hs = np.asarray(plain["hs"], dtype=np.float64)
    ks = np.asarray(plain["ks"], dtype=np.int32)
    gpu_step_s = np.asarray(plain["step_s"], dtype=np.float64)
    done = min(args.cpu_steps, len(hs))
    a = time.perf_counter()
    _, st_tot = O.integrate_replay_opts(op, psi, hs[:done], ks[:done], None, t0=plain["t0"], rtol=args.rtol,
                                        atol=args.atol, newton_tol=0.33, lin_abs=0.05 * 0.33, jac_policy=2,
                                        ida_test=True, max_newton=12, dim_krylov=20, max_restarts=5)
    wall = time.perf_counter() - a
    days = float(hs[:done].sum())
    gpu_same = days / float(gpu_step_s[:done].sum())
    return {
        "value": days / wall,
        "unit": "sim-days/wall-s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "sample": (f"C/OpenMP oracle on the same {prob.mesh.n_nodes}-node box: {done} step(s) of the plain FP64 "
                   f"run's timed window from t={plain['t0']:.4g} d (h={', '.join(f'{h:.3g}' for h in hs[:done])} d), "
                   f"coloured FD Jacobian + ILU(0) + GMRES(20) with FD J v and IDA tolerances, own upwinding; "
                   f"{st_tot['newton_iters']} Newton, {st_tot['gmres_iters']} GMRES iterations, "
                   f"{st_tot['jac_setups']} Jacobian set-up(s), {st_tot['residuals']} residuals in {wall:.1f} s"),
        "residual_median_s": res_s,
        "residual_evals_timed": len(times),
        "slice_residual_median_s": float(np.median(stimes)),
        "slice_nodes": sp_.mesh.n_nodes,
        "gpu_residual_us": gpu_res_us,
        "oracle_build_s": build_s,
        "gpu_plain_fp64_same_steps": gpu_same,
        "gpu_over_cpu_same_steps": gpu_same / (days / wall),
    }


def slice_residual_gpu():
    """The 2D slice (the CPU-runnable configuration, 4 141 nodes): median of 50
    whole residual evaluations (HIP events) at the hydrostatic start."""
    import torch
    from fvmpor_amd.lib import Context
    from fvmpor_amd.problems import slice2d_problem
    prob = slice2d_problem()
    ctx = Context(prob)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    psi = torch.tensor(prob.psi0, device="cuda:0")
    psip = torch.zeros_like(psi)
    F = torch.empty_like(psi)
    ctx.begin_step(0.0, psi, psip)
    for _ in range(5):
        ctx.residual(0.0, psi, psip, F)
    us = []
    for _ in range(50):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.residual(0.0, psi, psip, F)
        e1.record(stream)
        e1.synchronize()
        us.append(e0.elapsed_time(e1) * 1e3)
    ctx.close()
    return {"nodes": prob.mesh.n_nodes, "elements": prob.mesh.n_elements, "gpu_residual_median_us": float(np.median(us))}


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


def run_window(args, prob, ctx, psi0, precon, t_stop=None, steps=None, learned=None, profile=False, keep_start=False):
    """One integrator from t = 0 on the context: W untimed warm-up steps, then
    either exactly `steps` timed steps or steps up to t_stop.  Returns the
    window's stats, its wall time, the final heads, the step sizes and the
    wall time at each step's end, the profile figures and (keep_start) the
    heads at the window's start."""
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
    psi_start = None
    if keep_start:
        out = torch.empty_like(psi0)
        integ.advance(s0["t"], out)
        psi_start = out.cpu().numpy()
    if profile:
        ctx.profile(True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    i = 0
    hs, ks, ends = [], [], []
    while (steps is not None and i < steps) or (steps is None and integ.stats()["t"] < t_stop):
        step(T_END if steps is not None else t_stop)
        hs.append(integ.stats()["h_last"])
        ks.append(integ.stats()["order_last"])
        ends.append(time.perf_counter() - w0)
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
    return {"t0": s0["t"], "t1": s1["t"], "wall": wall, "d": d, "psi": psi, "prof": prof, "hs": hs, "ks": ks,
            "step_s": np.diff([0.0] + ends).tolist(), "psi_start": psi_start}


def full_run(args, ctx, psi0, precon, learned=None):
    """The whole run from t = 0 to 1 day (untimed set-up excluded)."""
    import torch
    from fvmpor_amd.lib import Integrator
    integ = Integrator(ctx, psi0, rtol=args.rtol, atol=args.atol, max_order=3, precon=precon,
                       block_rows=args.block_rows, overlap=args.overlap)
    if learned is not None:
        integ.set_learned(*learned)
    out = torch.empty_like(psi0)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    rc = integ.advance(1.0, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    st = integ.stats()
    integ.close()
    if rc != 0:
        raise RuntimeError(f"run to 1 day stopped at t={st['t']:.4g} (status {rc})")
    r = {"value": 1.0 / wall, "unit": "sim-days/wall-s", "wall_s": wall, "steps": st["steps"],
         "newton_iters": st["newton_iters"], "gmres_iters": st["gmres_iters"], "precon_setups": st["precon_setups"],
         "error_fails": st["error_fails"], "newton_fails": st["newton_fails"]}
    if st["learned_tried"]:
        r["learned_start_used"] = [st["learned_used"], st["learned_tried"]]
    return r, out.cpu().numpy()


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
    from fvmpor_amd.lib import (BLOCK_ROWS, PRECON_BJILU0, PRECON_BJILU0_F32, PRECON_ILU0, PRECON_JACOBI,
                                PRECON_SCHUR, Context,
                                Mpnn)
    from fvmpor_amd.problems import box3d_problem

    params, train_info = (None, None) if args.no_learned else train_network(args)
    t0 = time.perf_counter()
    prob = box3d_problem(nx=args.nx, ny=args.ny, nz=args.nz, order=args.order)
    log(f"mesh N={prob.mesh.n_nodes} T={prob.mesh.n_elements} built in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    ctx = Context(prob)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    psi0 = torch.tensor(prob.psi0, device=dev)
    pc64 = {"bjilu0": PRECON_BJILU0, "ilu0": PRECON_ILU0, "jacobi": PRECON_JACOBI,
            "schur": PRECON_SCHUR}[args.precon]
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
    log(f"setup {time.perf_counter() - t0:.1f}s, N_edges={ctx.NE}, residual launch {launch}")

    head_pc = PRECON_BJILU0_F32 if args.precon == "bjilu0" and not args.no_fp32 else pc64
    head_name = ("FP32-precon" if head_pc == PRECON_BJILU0_F32 else "plain FP64") + \
        (" + learned initial iterate" if learned is not None else "")
    log(f"headline run: {head_name}")
    head = run_window(args, prob, ctx, psi0, head_pc, steps=args.steps, learned=learned, profile=True)
    tb = head["t1"]
    log(f"plain FP64 run over the same window")
    plain = run_window(args, prob, ctx, psi0, pc64, t_stop=tb, keep_start=True)
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
            "nodes": N, "elements": T, "edges": NE, "precon_plain": args.precon, "node_order": args.order,
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
    # the metric's own run: t = 0 to 1 day, each configuration
    if not args.no_full:
        log("full runs to t = 1 day")
        full = {}
        fp, psi_p = full_run(args, ctx, psi0, pc64)
        full["plain_fp64"] = fp
        runs = [("headline", head_pc, learned)]
        if fp32 is not None:
            runs.append(("fp32_precon_only", PRECON_BJILU0_F32, None))
        for name, pc, lr in runs:
            r, psi_r = full_run(args, ctx, psi0, pc, lr)
            r["head_rel_err_vs_plain_fp64"] = float(np.abs(psi_r - psi_p).max() / np.abs(psi_p).max())
            r["speedup_vs_plain_fp64"] = r["value"] / fp["value"]
            full[name] = r
        out["full_run_1day"] = full
    # gates of the metric: head error <= 1e-6 against FP64, and the learned
    # start must not lose to the FP32 preconditioner alone
    src = out.get("full_run_1day", {})
    h_err = src.get("headline", head_s).get("head_rel_err_vs_plain_fp64")
    sp = src.get("headline", {}).get("speedup_vs_plain_fp64", out["speedup_vs_plain_fp64"])
    gate = {"head_err_le_1e-6": h_err is not None and h_err <= 1e-6, "speedup_vs_plain_fp64": sp,
            "speedup_ge_3x": sp >= 3.0, "source": "full_run_1day" if src else "timed window"}
    if fp32 is not None:
        f_val = src.get("fp32_precon_only", out["fp32_precon_only"])["value"]
        h_val = src.get("headline", head_s)["value"]
        gate["learned_not_slower_than_fp32_only"] = h_val >= f_val
    gate["pass"] = all(v for k, v in gate.items() if isinstance(v, bool))
    out["gate"] = gate
    if train_info is not None:
        out["network"] = train_info
    if args.probe_copy:
        # a plain device copy of known size, for scaling the HBM byte counters of a profiler run
        a = torch.empty(args.probe_copy << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    out["slice_residual"] = slice_residual_gpu()
    if not args.no_cpu:
        try:
            log("CPU oracle baseline")
            out["cpu_baseline"] = cpu_baseline(args, prob, plain, rs_ms / max(rs_n, 1) * 1e3)
        except Exception as exc:  # the GPU measurement stands on its own
            log(f"cpu baseline failed: {exc!r}")
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
