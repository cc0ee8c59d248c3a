"""Training and evaluation of the learned FGMRES preconditioner.

Data: seeded optically thick, highly scattering 3D instances (n^3 cells, DG Q1,
level-symmetric S8, sigma_t log-uniform in [1, 1000] and c uniform in
[0.99, 0.9999] on 4^3-cell blocks, block sources U[0, 1], vacuum).  Each
instance contributes (residual, correction) pairs that the operator's own
sweeps generate: along the error propagation e_{k+1} = e_k - A e_k of source
iteration, (r, e) = (A e_k, e_k) is an exact pair, and the network learns the
coarse part of the scattered correction e - r = e_{k+1}.  Two chains per
instance: one from the source-driven right-hand side D L^-1 (q + inflow), one
from a random cell field.

Per coarse cell the network sees r_hat = R r / s (s = RMS of R r), tau_hat and
kappa_hat, and predicts t = R (e - r) / s through the fixed prior g (see
ouranos_amd/csrc/learned.hpp for the feature definitions).

    python -m ouranos_amd.learned train --instances 2000 --out weights.npy
    python -m ouranos_amd.learned evaluate --weights weights.npy --instances 200
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import ouranos_amd as ou
from ouranos_amd import problems as P

TRAIN_SEED0 = 1_000_000   # instance i of the training set uses seed TRAIN_SEED0 + i
EVAL_SEED0 = 2_000_000    # held-out instances
RECORD = (0, 2, 6, 14)    # recorded steps k of each chain
N_STEPS = 15


class Family:
    """A seeded family of training/evaluation instances.
    thick:    32^3 DG Q1, LS-S8, 4^3-cell blocks, sigma_t log-U[1,1000],
              c U[0.99,0.9999] (the learned-preconditioner data set), FP64 sweeps;
    headline: 64^3 DG Q2, product GLC 12x24, L=1, 8^3-cell blocks with the
              headline materials (sigma_t log-U[1,500], c U[0.99,0.9999]); on
              the 32^3 coarse grid its coarse cells and blocks have the optical
              sizes of the 128^3 headline problem's; FP32 sweeps;
    headline128: the headline problem's own generator and size (128^3 DG Q2,
              GLC 12x24, L=1, 16^3-cell blocks) at other seeds; FP32 sweeps."""

    NAMES = ("thick", "headline", "headline128")

    def __init__(self, name):
        if name not in self.NAMES:
            raise ValueError(f"unknown family {name}")
        self.name = name
        self.precision = ou.FP64 if name == "thick" else ou.FP32
        self.quad = (ou.Quadrature(ou.LS, 3, sn=8) if name == "thick" else
                     ou.Quadrature(ou.PRODUCT_GLC, 3, n_polar=12, n_azim=24, L_max=1))

    def spec(self, seed):
        if self.name == "thick":
            return P.thick_3d(n=32, block=4, seed=seed, quad=self.quad)
        if self.name == "headline128":
            return P.headline_3d(n=128, block=16, seed=seed, quad=self.quad)
        return P.headline_3d(n=64, block=8, seed=seed, quad=self.quad)


def instance(seed, n=32, quad=None):
    return P.thick_3d(n=n, block=4, seed=seed, quad=quad)


def random_field(spec, seed):
    """A random moment-0 field: per-cell normal values (higher moments zero)."""
    rng = np.random.default_rng(seed)
    e = np.zeros(spec.size)
    cells = rng.normal(size=spec.n_cells)
    e[:spec.n_dofs] = np.repeat(cells, spec.n_loc)
    return e


def features_of(rk, mat):
    """Network features [G^3][12] of the octant-restricted residual rk [G^3][8]
    and the material features mat [G^3][3]; returns (features, scale s)."""
    s = float(np.sqrt(np.mean(rk * rk)))
    f = np.zeros((rk.shape[0], ou.UNet.N_FEATURES), dtype=np.float32)
    if s > 0:
        f[:, 0:8] = rk / s
    f[:, 8:11] = mat
    return f, s


KRYLOV_STEPS = 90   # plain GMRES(30) applications per instance
KRYLOV_STRIDE = 6   # record every 6th Arnoldi pair
FGMRES_RESTART = 30  # restart length of the on-policy FGMRES pairs


def pairs_of(op, spec, grid, seed, sources=("si", "krylov"), net=None):
    """Features [k][G^3][12] and targets [k][G^3][8] (float32) of one instance.
    Sources: "si" = the error-propagation chains of source iteration from the
    right-hand side and from a random field; "krylov" = the Arnoldi vectors
    of plain GMRES(30) on the instance's source problem; "fgmres" = the
    preconditioned vectors of FGMRES(FGMRES_RESTART) with `net` itself (the
    residuals the network meets when it is used)."""
    _, mat = op.coarse_features(grid)
    chunks = []
    if "si" in sources:
        for e0 in (None, random_field(spec, seed)):
            chunks.append(op.learned_pairs(grid, N_STEPS, RECORD, e0=e0))
    if "krylov" in sources:
        chunks.append(op.learned_pairs_krylov(grid, KRYLOV_STEPS, stride=KRYLOV_STRIDE,
                                              max_records=KRYLOV_STEPS // KRYLOV_STRIDE))
    if "fgmres" in sources:
        if net is None:
            raise ValueError("the fgmres source needs the network")
        chunks.append(op.learned_pairs_krylov(grid, KRYLOV_STEPS, stride=KRYLOV_STRIDE,
                                              max_records=KRYLOV_STEPS // KRYLOV_STRIDE,
                                              restart=FGMRES_RESTART, net=net))
    feats, targets = [], []
    for r, e in chunks:
        for rk, ek in zip(r, e):
            f, s = features_of(rk, mat)
            if not s > 0:
                continue
            feats.append(f)
            targets.append(((ek - rk) / s).astype(np.float32))
    return feats, targets


def generate(n_instances, seed0, grid=32, n=32, log=print, sources=("si", "krylov"),
             family=None, net=None):
    fam = family or Family("thick")
    feats, targets = [], []
    t0 = time.perf_counter()
    for i in range(n_instances):
        spec = fam.spec(seed0 + i) if family else instance(seed0 + i, n=n, quad=fam.quad)
        op = spec.operator(fam.precision)
        f, t = pairs_of(op, spec, grid, seed0 + i, sources, net=net)
        op.close()
        feats += f
        targets += t
        if (i + 1) % max(1, n_instances // 10) == 0:
            log(f"  generated {i + 1}/{n_instances} instances, {len(feats)} pairs, "
                f"{time.perf_counter() - t0:.1f} s")
    return np.stack(feats), np.stack(targets)


def train(net, feats, targets, epochs, batch, lr, seed=0, log=print):
    rng = np.random.default_rng(seed)
    n = feats.shape[0]
    steps = epochs * (n // batch)
    history = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        losses = []
        for b in range(n // batch):
            idx = np.sort(order[b * batch:(b + 1) * batch])
            # cosine decay from lr to lr/20
            lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * step / max(1, steps))))
            losses.append(net.train_step(feats[idx], targets[idx], lr_t))
            step += 1
        history.append(float(np.mean(losses)))
        log(f"  epoch {ep + 1}/{epochs}: mean loss {history[-1]:.4f} "
            f"({time.perf_counter() - t0:.1f} s)")
    return history


def train_streaming(net, n_instances, seed0, grid, chunk, passes, batch, lr, val=None,
                    sources=("si", "krylov"), seed=0, log=print, family=None):
    """Online training over n_instances generated chunk by chunk: each chunk's
    pairs get `passes` shuffled passes, then are dropped (the data set of 2000
    instances does not fit in host memory at once).  Cosine learning-rate
    decay over the whole run.  Returns per-chunk mean losses (and validation
    losses if `val` = (features, targets))."""
    rng = np.random.default_rng(seed)
    n_chunks = (n_instances + chunk - 1) // chunk
    est_steps = None
    step = 0
    hist, vhist = [], []
    t0 = time.perf_counter()
    for c in range(n_chunks):
        lo = c * chunk
        # the fgmres source draws its pairs from the network as trained so far
        F, T = generate(min(chunk, n_instances - lo), seed0 + lo, grid=grid,
                        log=lambda m: None, sources=sources, family=family, net=net)
        if est_steps is None:
            est_steps = n_chunks * passes * (F.shape[0] // batch)
        losses = []
        for _ in range(passes):
            order = rng.permutation(F.shape[0])
            for b in range(F.shape[0] // batch):
                idx = np.sort(order[b * batch:(b + 1) * batch])
                lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * min(1.0, step / est_steps))))
                losses.append(net.train_step(F[idx], T[idx], lr_t))
                step += 1
        hist.append(float(np.mean(losses)))
        msg = (f"  chunk {c + 1}/{n_chunks}: {F.shape[0]} pairs, mean loss {hist[-1]:.4f}, "
               f"{step} steps, {time.perf_counter() - t0:.0f} s")
        if val is not None:
            vhist.append(validation_loss(net, val[0], val[1], batch))
            msg += f", validation loss {vhist[-1]:.4f}"
        log(msg)
        del F, T
    return hist, vhist


def validation_loss(net, feats, targets, batch):
    out = []
    for b in range(0, feats.shape[0] - batch + 1, batch):
        loss, _ = net.loss_grad(feats[b:b + batch], targets[b:b + batch])
        out.append(loss)
    return float(np.mean(out)) if out else float("nan")


def evaluate(net, n_instances, seed0=EVAL_SEED0, n=32, tol=1e-8, restart=30, max_it=4000,
             log=print):
    """Plain GMRES(restart) vs FGMRES(restart) with the network, FP64 operator,
    on held-out instances; returns per-instance records."""
    quad = ou.Quadrature(ou.LS, 3, sn=8)
    rows = []
    for i in range(n_instances):
        spec = instance(seed0 + i, n=n, quad=quad)
        op = spec.operator(ou.FP64)
        phi, a = op.solve(solver=ou.GMRES, tol=tol, max_it=max_it, restart=restart,
                          raise_on_fail=False)
        phl, b = op.solve_learned(net, tol=tol, max_it=max_it, restart=restart,
                                  raise_on_fail=False)
        op.close()
        diff = float(np.linalg.norm(phl[:spec.n_dofs] - phi[:spec.n_dofs]) /
                     np.linalg.norm(phi[:spec.n_dofs]))
        rows.append({"seed": seed0 + i, "plain_its": a["iterations"],
                     "plain_s": a["seconds"], "plain_conv": a["converged"],
                     "learned_its": b["iterations"], "learned_s": b["seconds"],
                     "learned_conv": b["converged"], "scalar_flux_diff": diff})
        if (i + 1) % max(1, n_instances // 10) == 0:
            pi = np.mean([r["plain_its"] for r in rows])
            li = np.mean([r["learned_its"] for r in rows])
            log(f"  evaluated {i + 1}/{n_instances}: mean iterations plain {pi:.1f}, "
                f"learned {li:.1f}")
    return rows


def gcr(apply_a, b, precond, tol, restart, max_it, x_exact=None, progress=None, device=None):
    """Flexible GCR (Eisenstat, Elman and Schultz, SIAM J. Numer. Anal. 20
    (1983)) restarted every `restart` steps: step k applies the preconditioner
    to the current residual, z_k = M(r_k, x_k), makes A z_k orthonormal to the
    previous images (classical Gram-Schmidt, two passes) and minimises the
    residual over the search directions.  Unlike FGMRES the preconditioner
    always sees the true residual, whose exact correction A^-1 r_k = x* - x_k
    is known when the solution is: that is what the ideal preconditioner of
    `ceiling` uses.  Returns (x, steps, relative residual history).
    device: a torch device name ("cuda") keeps the search directions there
    (the headline family's 60 vectors of 28 M values are slow in numpy);
    apply_a and precond still take and return numpy arrays."""
    if device is not None:
        return _gcr_torch(apply_a, b, precond, tol, restart, max_it, progress, device)
    x = np.zeros_like(b)
    r = b.copy()
    bn = float(np.linalg.norm(b))
    hist = [1.0]
    ps, qs = [], []
    k = 0
    while k < max_it and hist[-1] >= tol:
        z = precond(r, x)
        w = apply_a(z)
        k += 1
        for _ in range(2):
            for p, q in zip(ps, qs):
                a = float(q @ w)
                w -= a * q
                z -= a * p
        nw = float(np.linalg.norm(w))
        if not nw > 0.0:
            break
        q, p = w / nw, z / nw
        alpha = float(q @ r)
        x += alpha * p
        r -= alpha * q
        ps.append(p)
        qs.append(q)
        if len(ps) == restart:
            ps, qs = [], []
            r = b - apply_a(x)  # restart from the true residual
        hist.append(float(np.linalg.norm(r)) / bn)
        if progress and k % 20 == 0:
            progress(k, hist[-1])
    return x, k, hist


def _gcr_torch(apply_a, b, precond, tol, restart, max_it, progress, device):
    """gcr() with the vectors in torch tensors on `device` (same arithmetic, FP64)."""
    import torch
    dev = torch.device(device)

    def T(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

    def N(t):
        return t.cpu().numpy()

    bt = T(b)
    x = torch.zeros_like(bt)
    r = bt.clone()
    bn = float(torch.linalg.norm(bt))
    hist = [1.0]
    ps, qs = [], []
    k = 0
    while k < max_it and hist[-1] >= tol:
        z = T(precond(N(r), N(x)))
        w = T(apply_a(N(z)))
        k += 1
        for _ in range(2):
            for p, q in zip(ps, qs):
                a = float(torch.dot(q, w))
                w -= a * q
                z -= a * p
        nw = float(torch.linalg.norm(w))
        if not nw > 0.0:
            break
        q, p = w / nw, z / nw
        alpha = float(torch.dot(q, r))
        x += alpha * p
        r -= alpha * q
        ps.append(p)
        qs.append(q)
        if len(ps) == restart:
            ps, qs = [], []
            r = bt - T(apply_a(N(x)))  # restart from the true residual
        hist.append(float(torch.linalg.norm(r)) / bn)
        if progress and k % 20 == 0:
            progress(k, hist[-1])
    return N(x), k, hist


def ceiling(net, n_instances=3, seed0=EVAL_SEED0, n=32, grid=32, tol=1e-8, restart=30,
            max_it=600, log=print, family="thick", device=None):
    """How far the network's correction space can go.  On held-out instances
    of the thick family (FP64 operator), flexible GCR(restart) to `tol` with
      plain    z = r;
      learned  z = r + P s N(R r / s), the network as the FGMRES preconditioner
               applies it;
      ideal    z = r + P R (A^-1 r - r), with A^-1 r = x* - x exact (x* from
               GMRES to 1e-12): the correction the network is trained to
               predict, at its exact value, in the same coarse corner space;
      ideal_moments (more than one moment) the same corner correction for
               every moment, not only the scalar flux: a richer space;
      exact    z = A^-1 r (a one-step solve; checks the driver).
    family "headline" runs the headline family instead (64^3 Q2, GLC 12x24,
    L=1; see Family), FP64 operator.  Returns one record per instance."""
    fam = Family(family)
    rows = []
    for i in range(n_instances):
        spec = (instance(seed0 + i, n=n, quad=fam.quad) if family == "thick"
                else fam.spec(seed0 + i))
        op = spec.operator(ou.FP64)
        rhs = _rhs_of(op)
        x_star, info = op.solve_rhs(rhs, solver=ou.GMRES, tol=1e-12, max_it=4000)
        _, mat = op.coarse_features(grid)

        def coarse(v):
            return op.coarse_features(grid, v)[0]

        def prec_learned(r, x):
            rk = coarse(r)
            f, s_r = features_of(rk, mat)
            if not s_r > 0:
                return r.copy()
            pred = net.predict(f[None])[0]
            return op.coarse_prolong(grid, pred * s_r, z=r)

        def prec_ideal(r, x):
            e = x_star - x
            return op.coarse_prolong(grid, (coarse(e) - coarse(r)).astype(np.float32), z=r)

        nd, n_mom = spec.n_dofs, op.n // spec.n_dofs

        def moment_field(v, m):
            """Moment m of v, placed where the transfers read moment 0."""
            out = np.zeros(op.n)
            out[:nd] = v[m * nd:(m + 1) * nd]
            return out

        def prec_ideal_moments(r, x):
            # the ideal corner correction of every moment, not only the scalar flux
            e = x_star - x
            z = r.copy()
            for m in range(n_mom):
                corners = coarse(moment_field(e, m)) - coarse(moment_field(r, m))
                z[m * nd:(m + 1) * nd] += op.coarse_prolong(grid, corners.astype(np.float32))[:nd]
            return z

        def prec_exact(r, x):
            return x_star - x

        row = {"seed": seed0 + i, "reference_gmres_iterations": info["iterations"]}
        methods = [("plain", lambda r, x: r.copy())]
        if grid == net.grid:  # the network runs on the grid it was trained for
            methods.append(("learned", prec_learned))
        methods.append(("ideal", prec_ideal))
        if n_mom > 1:
            methods.append(("ideal_moments", prec_ideal_moments))
        methods.append(("exact", prec_exact))
        for name, m in methods:
            t0 = time.perf_counter()
            x, k, hist = gcr(op.Apply, rhs, m, tol, restart, max_it,
                             progress=lambda k, h: log(f"    {name} step {k}: {h:.2e}"),
                             device=device)
            diff = float(np.linalg.norm(x[:spec.n_dofs] - x_star[:spec.n_dofs]) /
                         np.linalg.norm(x_star[:spec.n_dofs]))
            row[name] = {"steps": k, "rel_residual": hist[-1], "scalar_flux_diff": diff,
                         "seconds": time.perf_counter() - t0}
            log(f"  instance {seed0 + i}: GCR({restart}) {name}: {k} steps, residual "
                f"{hist[-1]:.2e}, scalar-flux difference {diff:.1e}")
        op.close()
        rows.append(row)
    return rows


def _rhs_of(op):
    """The instance's right-hand side D L^-1 (q + inflow) (ouranos.cc:149-158):
    the sweep of the source with a zero scattering source."""
    op.compute_scattering_source(np.zeros(op.n))
    return op.sweep(None, with_source=True)


def save(net, path):
    """Parameters (.npy, float32) plus a JSON sidecar with grid and mode."""
    np.save(path, net.get_params())
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump({"grid": net.grid, "mode": net.mode, "n_params": net.n_params}, f)


def load(path, max_batch=1, grid=None):
    """Loads saved weights.  The network is fully convolutional, so `grid`
    may override the coarse grid it was trained on."""
    with open(os.path.splitext(path)[0] + ".json") as f:
        meta = json.load(f)
    net = ou.UNet(grid=grid or meta["grid"], max_batch=max_batch)
    net.mode = meta["mode"]
    net.set_params(np.load(path, allow_pickle=False))
    return net


def summarize(rows):
    pi = np.array([r["plain_its"] for r in rows], dtype=float)
    li = np.array([r["learned_its"] for r in rows], dtype=float)
    ps = np.array([r["plain_s"] for r in rows])
    ls = np.array([r["learned_s"] for r in rows])
    return {"instances": len(rows),
            "plain_iterations_mean": float(pi.mean()), "learned_iterations_mean": float(li.mean()),
            "iteration_ratio_mean": float(np.mean(pi / np.maximum(li, 1))),
            "plain_seconds_total": float(ps.sum()), "learned_seconds_total": float(ls.sum()),
            "speedup_time": float(ps.sum() / max(ls.sum(), 1e-12)),
            "plain_converged": int(sum(r["plain_conv"] for r in rows)),
            "learned_converged": int(sum(r["learned_conv"] for r in rows)),
            "max_scalar_flux_diff": float(max(r["scalar_flux_diff"] for r in rows))}


def main(argv=None):
    global KRYLOV_STEPS, KRYLOV_STRIDE, FGMRES_RESTART
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--instances", type=int, default=2000)
    t.add_argument("--val-instances", type=int, default=20)
    t.add_argument("--chunk", type=int, default=100)
    t.add_argument("--passes", type=int, default=3)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--grid", type=int, default=32)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    t.add_argument("--mode", type=int, default=0)
    t.add_argument("--family", default="thick", choices=list(Family.NAMES))
    t.add_argument("--sources", default="si,krylov")
    t.add_argument("--init", default=None, help="start from these weights")
    t.add_argument("--seed-offset", type=int, default=0,
                   help="first training instance seed = TRAIN_SEED0 + offset")
    t.add_argument("--krylov-steps", type=int, default=KRYLOV_STEPS)
    t.add_argument("--krylov-stride", type=int, default=KRYLOV_STRIDE)
    t.add_argument("--pair-restart", type=int, default=FGMRES_RESTART,
                   help="restart length of the on-policy FGMRES pairs")
    t.add_argument("--eval-instances", type=int, default=0)
    c = sub.add_parser("ceiling")
    c.add_argument("--weights", required=True)
    c.add_argument("--instances", type=int, default=3)
    c.add_argument("--restart", type=int, default=30)
    c.add_argument("--family", default="thick", choices=list(Family.NAMES))
    c.add_argument("--grid", type=int, default=32,
                   help="coarse grid of the ideal corrections (the network runs only on its own)")
    c.add_argument("--device", default=None,
                   help="torch device for the GCR vectors (e.g. cuda); numpy by default")
    c.add_argument("--out", default=None)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", required=True)
    e.add_argument("--instances", type=int, default=200)
    e.add_argument("--grid", type=int, default=32)
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "train":
        KRYLOV_STEPS, KRYLOV_STRIDE = a.krylov_steps, a.krylov_stride
        FGMRES_RESTART = a.pair_restart
        fam = Family(a.family)
        sources = tuple(a.sources.split(","))
        net = ou.UNet(grid=a.grid, max_batch=a.batch, seed=a.seed)
        net.mode = a.mode
        if a.init:
            init = load(a.init)
            net.mode = init.mode
            net.set_params(init.get_params())
            init.close()
        log(f"[learned] generating {a.val_instances} validation instances ({a.family})")
        Fv, Tv = generate(a.val_instances, EVAL_SEED0 + 100_000, grid=a.grid, log=log,
                          sources=sources, family=fam, net=net)
        v0 = validation_loss(net, Fv, Tv, a.batch)
        log(f"[learned] validation loss before training {v0:.4f}; training on {a.instances} "
            f"instances in chunks of {a.chunk}, {a.passes} passes each")
        t0 = time.perf_counter()
        hist, vhist = train_streaming(net, a.instances, TRAIN_SEED0 + a.seed_offset, a.grid,
                                      a.chunk, a.passes,
                                      a.batch, a.lr, val=(Fv, Tv), seed=a.seed, log=log,
                                      sources=sources, family=fam)
        secs = time.perf_counter() - t0
        v1 = vhist[-1]
        log(f"[learned] validation loss after training {v1:.4f} ({secs:.0f} s)")
        save(net, a.out)
        rep = {"instances": a.instances, "chunk": a.chunk, "passes": a.passes,
               "batch": a.batch, "lr": a.lr, "mode": a.mode, "seconds": secs,
               "train_loss": hist, "val_loss": vhist, "val_loss_before": v0,
               "val_loss_after": v1, "family": a.family, "family_doc": Family.__doc__,
               "sources": {"si": "SI error-propagation chains (RHS and random field, steps "
                                 f"{list(RECORD)})",
                           "krylov": f"plain GMRES(30) Arnoldi pairs (every {KRYLOV_STRIDE}th "
                                     f"of {KRYLOV_STEPS})",
                           "fgmres": f"FGMRES({FGMRES_RESTART}) pairs (A z_j, z_j) with the "
                                     f"network being trained (every {KRYLOV_STRIDE}th of "
                                     f"{KRYLOV_STEPS})"},
               "sources_used": list(sources),
               "train_command": "python -m ouranos_amd.learned " + " ".join(argv or sys.argv[1:])}
        if a.eval_instances:
            rows = evaluate(net, a.eval_instances, log=log)
            rep["evaluation"] = summarize(rows)
            log(json.dumps(rep["evaluation"]))
        with open(os.path.splitext(a.out)[0] + "_report.json", "w") as f:
            json.dump(rep, f, indent=1)
        log("[learned] done")
    elif a.cmd == "ceiling":
        if a.device:
            # torch's HIP runtime initialises before the library's (the other
            # order leaves torch without a device)
            import torch
            torch.zeros(1, device=a.device)
        net = load(a.weights)
        rows = ceiling(net, a.instances, grid=a.grid, restart=a.restart, log=log,
                       family=a.family, device=a.device)
        summary = {m: float(np.mean([r[m]["steps"] for r in rows]))
                   for m in ("plain", "learned", "ideal", "ideal_moments", "exact")
                   if all(m in r for r in rows)}
        log(json.dumps({"mean_steps": summary}))
        if a.out:
            with open(a.out, "w") as f:
                json.dump({"mean_steps": summary, "rows": rows,
                           "command": "python -m ouranos_amd.learned " +
                                      " ".join(argv or sys.argv[1:])}, f, indent=1)
    else:
        net = load(a.weights)
        rows = evaluate(net, a.instances, log=log)
        s = summarize(rows)
        log(json.dumps(s))
        if a.out:
            with open(a.out, "w") as f:
                json.dump({"summary": s, "rows": rows}, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
