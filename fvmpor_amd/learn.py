"""Learned Newton initial iterate: data generation, training and use.

The network (fvp_mpnn_*, BF16 matrix-core kernels) maps per-node features of
a BDF step to the Newton correction the step needs, in error-weight units:
target = ewt * (psi_corrected - psi_pred).  The integrator builds the
dynamic feature columns on the device every step attempt (psi_n, the scaled
predictor increment, log10 h, order); this module supplies the static
columns and drives data generation (the plain solver in recording mode) and
training (forward, backward and Adam on the device).  The plain predictor
stays the fallback: the integrator keeps whichever starting point has the
smaller preconditioned first Newton residual, and backs off (skips the
network for 4, 8, ... up to 64 step attempts) after three tries in a row
whose proposal did not halve it.  No counterpart in the reference.

Static feature columns (N x 16, columns 0..4 are left to the integrator):
  5  height, normalised to [0, 1] over the mesh
  6  mean ln K_zz of the elements around the node (zone K times multiplier)
  7..10  share of the node's elements in zones 0..3
  11 prescribed-flux boundary node (1 / 0)
  12 its prescribed flux (m per time unit, positive into the domain) x 10
  13 Dirichlet node (1 / 0)
  14 mean ln alpha_vG,  15 mean n_vG of the elements around the node
"""
from __future__ import annotations

import time

import numpy as np

N_FEATURES = 16


def node_graph(mesh):
    """CSR adjacency (ia, ja) of the mesh's nodes without the diagonal:
    nodes sharing an element, sorted columns, symmetric."""
    conn = np.asarray(mesh.conn, dtype=np.int64)
    n = mesh.n_nodes
    npe = conn.shape[1]
    pairs = [conn[:, a] * n + conn[:, b] for a in range(npe) for b in range(npe) if a != b]
    keys = np.unique(np.concatenate(pairs))
    rows, cols = keys // n, keys % n
    ia = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=ia[1:])
    return ia, cols.astype(np.int32)


def static_features(problem) -> np.ndarray:
    """(N, 16) float32 static columns of a problem (the layout above)."""
    mesh = problem.mesh
    n, dim = mesh.n_nodes, mesh.dim
    conn = np.asarray(mesh.conn, dtype=np.int64)
    npe = conn.shape[1]
    x = np.zeros((n, N_FEATURES), dtype=np.float64)
    z = mesh.xyz[:, dim - 1]
    x[:, 5] = (z - z.min()) / max(z.max() - z.min(), 1e-30)
    zones = np.asarray(problem.zones, dtype=np.float64)
    ez = np.asarray(problem.ezone, dtype=np.int64)
    lnk = np.log(zones[ez, 2] * np.asarray(problem.ekmul, dtype=np.float64))
    cnt = np.bincount(conn.ravel(), minlength=n).astype(np.float64)
    cnt = np.maximum(cnt, 1.0)

    def node_mean(v):
        return np.bincount(conn.ravel(), weights=np.repeat(v, npe), minlength=n) / cnt

    x[:, 6] = node_mean(lnk)
    for zc in range(min(4, len(zones))):
        x[:, 7 + zc] = node_mean((ez == zc).astype(np.float64))
    ntp = np.asarray(mesh.node_tag_ptr)
    nt = np.asarray(mesh.node_tags)
    for i_tag, (btype, row) in enumerate(zip(problem.bc_type, problem.bc)):
        if i_tag == 0 or btype == 0:
            continue
        nodes = np.repeat(np.arange(n), np.diff(ntp))[nt == i_tag]
        if btype == 3:
            if row[0] != 0.0:
                x[nodes, 11] = 1.0
                x[nodes, 12] = -row[0] * 10.0
        elif btype in (1, 4):
            x[nodes, 13] = 1.0
    x[:, 14] = node_mean(np.log(zones[ez, 4]))
    x[:, 15] = node_mean(zones[ez, 5])
    return x.astype(np.float32)


def record_pairs(problem, t_end, rtol=1e-6, atol=1e-6, precon=None, max_pairs=10000, offsets=None):
    """Run the plain solver to t_end with recording on; returns the per-step
    (features, targets) pairs as device tensors and the integrator stats."""
    import torch

    from .lib import PRECON_BJILU0_F32, Context, Integrator
    dev = torch.device("cuda:0")
    ctx = Context(problem)
    psi0 = torch.tensor(problem.psi0, device=dev)
    integ = Integrator(ctx, psi0, rtol=rtol, atol=atol, max_order=3,
                       precon=PRECON_BJILU0_F32 if precon is None else precon)
    if offsets is not None:
        integ.set_segments(offsets)
    stat = torch.tensor(static_features(problem), device=dev)
    feat = torch.zeros((ctx.N, N_FEATURES), dtype=torch.float32, device=dev)
    targ = torch.zeros(ctx.N, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    integ.set_recording(feat, targ, stat)
    pairs = []
    while integ.stats()["t"] < t_end * (1 - 1e-12) and len(pairs) < max_pairs:
        rc = integ.step(t_end)
        if rc != 0:
            raise RuntimeError(f"integrator step failed ({rc})")
        ctx.synchronize()
        pairs.append((feat.clone(), targ.clone()))
    torch.cuda.synchronize()
    st = integ.stats()
    integ.set_recording(None, None)
    integ.close()
    ctx.close()
    return pairs, st


def train(model, pairs, epochs=20, lr=1e-3, seed=0, log=None):
    """Adam on the device over the recorded pairs (one pair per update,
    shuffled each epoch).  Returns the mean loss per epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(pairs)):
            x, t = pairs[i]
            tot += model.loss_grad(x, t)
            model.adam(lr)
        hist.append(tot / max(1, len(pairs)))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
    model.synchronize()
    return hist


def train_on(problems, t_end, layers=3, epochs=30, lr=1e-3, seed=0, rtol=1e-6, atol=1e-6, log=None):
    """Record training pairs with the plain solver on the problems (advanced
    together as one block-diagonal batch, per-instance error norms) and train
    a fresh network on them.  Returns (flat parameters, loss history, number
    of pairs, recording stats)."""
    from .ensemble import union_problem
    from .lib import Mpnn
    if len(problems) > 1:
        prob, off = union_problem(problems)
    else:
        prob, off = problems[0], None
    pairs, st = record_pairs(prob, t_end, rtol=rtol, atol=atol, offsets=off)
    ia, ja = node_graph(prob.mesh)
    model = Mpnn(ia, ja, layers=layers, seed=seed)
    hist = train(model, pairs, epochs=epochs, lr=lr, seed=seed, log=log)
    params = model.get_params()
    model.close()
    return params, hist, len(pairs), st


def run(problem, t_end, model_params=None, layers=3, rtol=1e-6, atol=1e-6, precon=None, trust=2.0, band=0.5):
    """Integrate to t_end on cuda:0, with the learned initial iterate when
    model_params (a flat parameter array) is given.  Returns (final heads,
    stats, wall seconds of the advance)."""
    import torch

    from .lib import PRECON_BJILU0_F32, Context, Integrator, Mpnn
    dev = torch.device("cuda:0")
    ctx = Context(problem)
    psi0 = torch.tensor(problem.psi0, device=dev)
    integ = Integrator(ctx, psi0, rtol=rtol, atol=atol, max_order=3,
                       precon=PRECON_BJILU0_F32 if precon is None else precon, overlap=1)
    model = None
    if model_params is not None:
        ia, ja = node_graph(problem.mesh)
        model = Mpnn(ia, ja, layers=layers)
        model.set_params(model_params)
        stat = torch.tensor(static_features(problem), device=dev)
        torch.cuda.synchronize()
        integ.set_learned(model, stat, trust=trust, band=band)
    out = torch.empty_like(psi0)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    rc = integ.advance(t_end, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    st = integ.stats()
    if rc != 0:
        raise RuntimeError(f"integration stopped at t={st['t']:.4g} (status {rc}): {st}")
    heads = out.cpu().numpy()
    integ.close()
    if model is not None:
        model.close()
    ctx.close()
    return heads, st, wall


def main(argv=None):
    """Train on fresh-seed boxes of the 1M-node workload's statistics and
    compare plain and learned runs on the workload's own box; one JSON line:
    python -m fvmpor_amd.learn [--train-nz 23] [--instances 2] [--eval-nz 40]."""
    import argparse
    import json
    import sys

    from .problems import box3d_problem
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-nz", type=int, default=23)
    ap.add_argument("--instances", type=int, default=2)
    ap.add_argument("--eval-nz", type=int, default=40)
    ap.add_argument("--t-train", type=float, default=0.25)
    ap.add_argument("--t-eval", type=float, default=0.25)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args(argv)
    rng = np.random.default_rng(2024)

    def box(nz, seed):
        return box3d_problem(nx=4 * nz, ny=4 * nz, nz=nz, seed_mesh=seed, seed_field=seed + 1, seed_layers=seed + 2,
                             sigma_lnK=float(rng.uniform(0.8, 1.3)), rain=float(rng.uniform(0.015, 0.03)))
    def say(msg):
        print(msg, file=sys.stderr, flush=True)
    train_p = [box(args.train_nz, 500 + 10 * i) for i in range(args.instances)]
    say(f"training meshes: {[p.mesh.n_nodes for p in train_p]} nodes")
    t0 = time.perf_counter()
    params, hist, npairs, st = train_on(train_p, args.t_train, epochs=args.epochs, lr=args.lr,
                                        log=lambda m: print(m, file=sys.stderr, flush=True))
    t_train = time.perf_counter() - t0
    say(f"trained in {t_train:.1f} s on {npairs} pairs")
    # held out: the 1M-node workload's own box (mesh seed 2, field seed 3, layer seed 4,
    # sigma_lnK 1, rain 0.02 m/day), none of whose seeds the training boxes use
    test = box3d_problem(nx=4 * args.eval_nz, ny=4 * args.eval_nz, nz=args.eval_nz)
    say(f"held-out mesh: {test.mesh.n_nodes} nodes")
    hp, sp, wp = run(test, args.t_eval)
    say(f"plain run: {wp:.2f} s, {sp['steps']} steps")
    hl, sl, wl = run(test, args.t_eval, model_params=params)
    say(f"learned run: {wl:.2f} s, {sl['steps']} steps")
    out = {"train_nodes": int(sum(p.mesh.n_nodes for p in train_p)), "pairs": npairs, "train_s": t_train,
           "loss_first": hist[0], "loss_last": hist[-1], "eval_nodes": test.mesh.n_nodes, "t_eval": args.t_eval,
           "plain": {"wall_s": wp, "steps": sp["steps"], "newton": sp["newton_iters"], "gmres": sp["gmres_iters"]},
           "learned": {"wall_s": wl, "steps": sl["steps"], "newton": sl["newton_iters"], "gmres": sl["gmres_iters"],
                       "used": sl["learned_used"], "tried": sl["learned_tried"]},
           "head_rel_diff": float(np.abs(hl - hp).max() / np.abs(hp).max())}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
