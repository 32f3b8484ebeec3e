"""Training of the learned pressure preconditioner on the GPU.

The training systems are the GFM pressure systems of random droplet configurations, built by
the library itself: the level set of a union of circles (heavy phase inside, phi = max over
the circles of r_k - |x - c_k|), its curvature and Heaviside, the divergence of a smooth random
velocity, assembled by asd2_gfm_assemble (the path of MACSolver2D::SolvePoisson, reference
src/mac2d.cpp:1249) and exported in FP64.  The default distribution is the seeded-droplet one of
the GFM benchmark case: 1-20 non-overlapping circles of radius uniform in [0.02, 0.1], density
ratio log-uniform in [10, 1000], right-hand side minus the divergence of a smooth random velocity;
FAMILIES adds larger droplets with surface tension and one to three large droplets at high
density ratio.

The network serves as the bottom of the multilevel hierarchy (the level whose coarsening equals
the network's, 32 x 32 blocks: a 64^2 image at 2048^2).  Its training inputs are the coarse
corrections g that the solver itself hands to that level (asd2_poisson_record_coarse, recorded
during multilevel solves and then during learned solves with the network being trained), with
that level's Galerkin operator P^T A P and the energy g.A^-1 g of each input (a sparse LU on the
host).  The loss is the relative A-norm error left by the best step along the network's output
(asd2_unet_train_step_coarse); the network's passes and the Adam steps are the library's kernels.
train_hybrid is the production recipe, and WEIGHTS names the parameters it produced, which
`pretrained` loads.

The fine-level samples (make_samples / train: residuals after k iterations of a solver, with the
network's own block restriction) serve networks used as the whole coarse correction.
"""
from __future__ import annotations

import time

import numpy as np

from . import (GRID_P, GRID_U, GRID_V, PREC_MIXED, PRECOND_LEARNED, PRECOND_MULTILEVEL, Context, DeviceArray,
               UNet, make_fluid)

G = 3


def _centres(n, dx, g=G):
    x = (np.arange(n + 2 * g) + 0.5 - g) * dx
    return np.meshgrid(x, x, indexing="ij")


def droplets(rng, n, dx, count=(1, 20), radius=(0.02, 0.1)):
    """Signed distance of `count` non-overlapping circles in the unit box (padded layout)."""
    k = int(rng.integers(count[0], count[1] + 1))
    circles = []
    for _ in range(10000):
        if len(circles) == k:
            break
        r = rng.uniform(*radius)
        c = rng.uniform([r, r], [1 - r, 1 - r])
        if all(np.hypot(*(c - c2)) > r + r2 + 2 * dx for c2, r2 in circles):
            circles.append((c, r))
    X, Y = _centres(n, dx)
    phi = np.full(X.shape, -np.inf)
    for c, r in circles:
        phi = np.maximum(phi, r - np.hypot(X - c[0], Y - c[1]))
    return np.ascontiguousarray(phi)


def smooth_velocity(rng, n, dx, modes=8):
    """A sum of low Fourier modes on the MAC faces (u on W faces, v on S faces)."""
    xc = (np.arange(n + 2 * G) + 0.5 - G) * dx
    xf = (np.arange(n + 2 * G) - G) * dx
    Xu, Yu = np.meshgrid(xf, xc, indexing="ij")
    Xv, Yv = np.meshgrid(xc, xf, indexing="ij")
    u = np.zeros_like(Xu)
    v = np.zeros_like(Xv)
    for _ in range(modes):
        kx, ky = rng.integers(1, 6, size=2)
        a, b = rng.standard_normal(2)
        px, py = rng.uniform(0, 2 * np.pi, size=2)
        u += a * np.sin(np.pi * kx * Xu + px) * np.cos(np.pi * ky * Yu + py)
        v += b * np.cos(np.pi * kx * Xv + px) * np.sin(np.pi * ky * Yv + py)
    return np.ascontiguousarray(u / modes), np.ascontiguousarray(v / modes)


class Samples:
    """Training systems on the device: (S, n + 6, n + 6) FP32 stacks of residual, Jacobi
    inverse, W-face and S-face coefficients and diagonal."""

    def __init__(self, arrays, n):
        self.n = n
        self.count = arrays[0].shape[0]
        self.dev = [DeviceArray.from_host(np.ascontiguousarray(a, np.float32)) for a in arrays]
        self.stride = (n + 2 * G) ** 2 * 4  # bytes per sample

    def batch(self, start):
        """device views of samples start, start + 1, ... (contiguous)"""
        return [_View(d.ptr + start * self.stride) for d in self.dev]


class _View:
    def __init__(self, ptr):
        self.ptr = ptr


def system_context(n):
    """A context for n x n droplet systems (Neumann pressure walls, no-slip velocity walls)."""
    dx = 1.0 / n
    ctx = Context(n, n, dx, dx)
    ctx.set_bc(GRID_P, ("neumann",) * 4)
    ctx.set_bc(GRID_U, ("dirichlet",) * 4)
    ctx.set_bc(GRID_V, ("dirichlet",) * 4)
    return ctx


def build_system(ctx, rng, rho_h=1000.0, ratio=(10.0, 1000.0), sigma=(0.0, 0.0), count=(1, 20),
                 radius=(0.02, 0.1)):
    """Assembles the GFM pressure system of a random droplet field in ctx (the path of
    MACSolver2D::SolvePoisson): circles, curvature, Heaviside, divergence of a smooth random
    velocity, density ratio rho_h / rho_l log-uniform in `ratio`, surface tension uniform in
    `sigma`."""
    n = ctx.nx
    dx = 1.0 / n
    ls, kap, H, div, u, v = (ctx.array() for _ in range(6))
    ls.upload(droplets(rng, n, dx, count=count, radius=radius))
    ctx.bc_apply(GRID_P, ls)
    ctx.kappa(ls, kap)
    ctx.bc_apply(GRID_P, kap)
    ctx.heaviside(ls, H)
    uu, vv = smooth_velocity(rng, n, dx)
    u.upload(uu)
    v.upload(vv)
    ctx.bc_apply(GRID_U, u)
    ctx.bc_apply(GRID_V, v)
    ctx.divergence(u, v, div)
    rl = rho_h / float(np.exp(rng.uniform(np.log(ratio[0]), np.log(ratio[1]))))
    fl = make_fluid(rho_h, rl, 1.0, 0.01, float(rng.uniform(*sigma)), 0.3 * dx)
    ctx.gfm_assemble(fl, ls, kap, H, div)
    ctx.sync()
    return fl


def residual_after(ctx, k, solver, ps, y, rhs):
    """b - A x_k after k iterations of `solver` ("jacobi": mixed-precision Jacobi-PCG,
    "multilevel": flexible PCG with the multilevel coarse correction) from x_0 = 0"""
    if k == 0:
        return rhs.copy()
    ps.upload(np.zeros_like(rhs))
    precond = PRECOND_MULTILEVEL if solver == "multilevel" else 0
    ctx.poisson_solve(ps, Context.opts(tol=1e-30, max_iter=k, precision=PREC_MIXED, precond=precond),
                      raise_on_fail=False)
    ctx.gfm_apply(ps, y)
    return rhs - y.download()


def make_samples(n, systems=6, snapshots=(("jacobi", 0), ("jacobi", 10), ("jacobi", 40), ("jacobi", 120)),
                 seed=0, log=None, **kw):
    """`systems` random droplet systems on an n x n grid, one sample per snapshot (solver, k): the
    residual after k iterations of that solver (k = 0: the right-hand side).  A bare k means
    ("jacobi", k)."""
    rng = np.random.default_rng(seed)
    cols = [[] for _ in range(5)]
    ctx = system_context(n)
    ps, y = ctx.array(), ctx.array()
    t0 = time.perf_counter()
    snaps = [("jacobi", s) if isinstance(s, (int, np.integer)) else tuple(s) for s in snapshots]
    for s in range(systems):
        build_system(ctx, rng, **kw)
        ax, ay, dg, rhs = ctx.gfm_export()
        dinv = np.zeros_like(dg)
        inner = (slice(G, G + n), slice(G, G + n))
        dinv[inner] = 1.0 / dg[inner]
        for solver, k in snaps:
            r = residual_after(ctx, k, solver, ps, y, rhs)
            rm = np.zeros_like(r)
            rm[inner] = r[inner]
            for col, a in zip(cols, (rm, dinv, ax, ay, dg)):
                col.append(a.astype(np.float32))
        if log:
            log(f"system {s + 1}/{systems} ({time.perf_counter() - t0:.1f} s)")
    order = rng.permutation(len(cols[0]))
    return Samples([np.stack([col[i] for i in order]) for col in cols], n)


def train(net: UNet, samples: Samples, steps=300, batch=4, lr=2e-3, lr_final=5e-4, seed=0, log=None):
    """Adam on the energy loss over random contiguous windows of the (shuffled) samples;
    returns the per-step losses."""
    rng = np.random.default_rng(seed)
    batch = min(batch, samples.count, net.max_batch)
    dx = 1.0 / samples.n
    losses = []
    for step in range(steps):
        start = int(rng.integers(0, samples.count - batch + 1))
        rate = lr if step < 0.75 * steps else lr_final
        losses.append(net.train_step(batch, *samples.batch(start), dx, dx, rate))
        if log and (step % 50 == 0 or step == steps - 1):
            log(f"step {step}: loss {losses[-1]:.3f}")
    return losses


def save_params(net: UNet, path: str, meta: dict | None = None):
    """FP32 parameter vector (asd2_unet_get_params order) as raw little-endian floats, with a
    JSON side file of its training record"""
    import json
    v = net.get_params()
    v.astype("<f4").tofile(path)
    with open(path + ".json", "w") as f:
        json.dump({"params": int(v.size), **(meta or {})}, f, indent=1)


def load_params(net: UNet, path: str):
    v = np.fromfile(path, dtype="<f4")
    if v.size != net.nparams:
        raise ValueError(f"{path} holds {v.size} parameters, the network has {net.nparams}")
    net.set_params(v)


# ---------------------------------------------------------------------------- coarse-level training
def coarse_operator(ax, ay, dg, c, g=G):
    """P^T A P of c x c blocks from padded FP64 fine arrays (ax / ay = W / S face coefficient of each
    cell), in the layout of the solver's level arrays: (H + 1) x (W + 1), acx = W-face coefficient,
    acy = S-face coefficient, dc = diagonal (the block's fine diagonals plus twice its inner faces)"""
    nx, ny = ax.shape[0] - 2 * g, ax.shape[1] - 2 * g
    Hc, Wc = nx // c, ny // c
    AX, AY, D = ax[g:g + nx, g:g + ny], ay[g:g + nx, g:g + ny], dg[g:g + nx, g:g + ny]
    acx = np.zeros((Hc + 1, Wc + 1))
    acy = np.zeros((Hc + 1, Wc + 1))
    dc = np.zeros((Hc + 1, Wc + 1))
    acx[:Hc, :Wc] = AX[::c, :].reshape(Hc, Wc, c).sum(axis=2)
    acy[:Hc, :Wc] = AY[:, ::c].reshape(Hc, c, Wc).sum(axis=1)
    inner_i = (np.arange(nx) % c != 0)[:, None]
    inner_j = (np.arange(ny) % c != 0)[None, :]
    Dn = D + 2.0 * AX * inner_i + 2.0 * AY * inner_j
    dc[:Hc, :Wc] = Dn.reshape(Hc, c, Wc, c).sum(axis=(1, 3))
    return acx, acy, dc


def coarse_matrix(acx, acy, dc):
    """the level operator (layout of coarse_operator) as a scipy CSR matrix over the H x W cells
    (unknown I * W + J)"""
    import scipy.sparse as sp
    Hc, Wc = acx.shape[0] - 1, acx.shape[1] - 1
    idx = np.arange(Hc * Wc).reshape(Hc, Wc)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [dc[:Hc, :Wc].ravel()]
    for w, a, b in ((acx[1:Hc, :Wc], idx[1:, :], idx[:-1, :]), (acy[:Hc, 1:Wc], idx[:, 1:], idx[:, :-1])):
        rows += [a.ravel(), b.ravel()]
        cols += [b.ravel(), a.ravel()]
        vals += [w.ravel(), w.ravel()]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(Hc * Wc, Hc * Wc))


class CoarseSamples:
    """Coarse training systems on the device (FP32): g [S][H][W], acx / acy / dc [S][(H + 1)(W + 1)],
    energy [S] = g.A^-1 g"""

    def __init__(self, g, acx, acy, dc, energy, h2):
        self.count = g.shape[0]
        self.H, self.W = g.shape[1], g.shape[2]
        self.h2 = h2
        self.host = [np.ascontiguousarray(a, np.float32) for a in (g, acx, acy, dc, energy)]
        self.dev = [DeviceArray.from_host(a) for a in self.host]
        self.strides = [self.H * self.W * 4] + [(self.H + 1) * (self.W + 1) * 4] * 3 + [4]

    @classmethod
    def union(cls, pools, seed=0):
        """the samples of several pools (same grid and coarsening) in one shuffled pool"""
        if any(p.h2 != pools[0].h2 or (p.H, p.W) != (pools[0].H, pools[0].W) for p in pools):
            raise ValueError("pools of different levels")
        order = np.random.default_rng(seed).permutation(sum(p.count for p in pools))
        cols = [np.concatenate([p.host[k] for p in pools])[order] for k in range(5)]
        return cls(*cols, pools[0].h2)

    def batch(self, start):
        return [_View(d.ptr + start * st) for d, st in zip(self.dev, self.strides)]


def make_coarse_samples(n, coarsen, systems=16, iters=(0, 2, 5, 10, 20, 40, 70, 100), seed=0, net=None, log=None,
                        **kw):
    """Coarse-correction inputs g the solver itself produces on the level of coarsening `coarsen`
    (recorded by asd2_poisson_record_coarse at the given iterations of a multilevel solve, or of
    the learned solve with `net` attached), with that level's Galerkin operator and the energy g.A^-1 g of each
    input (a sparse LU of the level operator), for `systems` random droplet systems on an n x n grid."""
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(seed)
    ctx = system_context(n)
    H = W = n // coarsen
    rec = DeviceArray((len(iters), H, W))
    gs, ops, es = [], [], []
    t0 = time.perf_counter()
    for s in range(systems):
        build_system(ctx, rng, **kw)
        rec.upload(np.zeros((len(iters), H, W)))
        ctx.record_coarse(H, W, list(iters), rec)
        if net is not None:
            ctx.set_preconditioner(net)
        ctx.poisson_solve(ctx.array(), Context.opts(tol=1e-10, max_iter=max(iters) + 1, precision=PREC_MIXED,
                                                    precond=PRECOND_LEARNED if net is not None else PRECOND_MULTILEVEL),
                          raise_on_fail=False)
        g = rec.download()
        ax, ay, dg, _ = ctx.gfm_export()
        op = coarse_operator(ax, ay, dg, coarsen)
        lu = spla.splu(coarse_matrix(*op).tocsc())
        for k in range(len(iters)):
            if np.any(g[k] != 0.0):
                gv = g[k].ravel()
                gs.append(g[k])
                ops.append(op)
                es.append(float(gv @ lu.solve(gv)))
        if log:
            log(f"system {s + 1}/{systems}: {len(gs)} samples ({time.perf_counter() - t0:.1f} s)")
    order = rng.permutation(len(gs))
    g = np.stack([gs[i] for i in order])
    acx, acy, dc = (np.stack([ops[i][k] for i in order]) for k in range(3))
    return CoarseSamples(g, acx, acy, dc, np.array(es)[order], (1.0 / n) ** 2 / coarsen)


def train_coarse(net: UNet, samples: CoarseSamples, steps=2000, batch=8, lr=2e-3, lr_final=5e-4, seed=0, log=None,
                 relative=True):
    """Adam on the coarse energy loss over random contiguous windows of the samples (relative: the
    A-norm error left relative to the sample's energy; else relative to its Jacobi scale); returns
    the per-step losses"""
    rng = np.random.default_rng(seed)
    batch = min(batch, samples.count, net.max_batch)
    losses = []
    for step in range(steps):
        start = int(rng.integers(0, samples.count - batch + 1))
        rate = lr if step < 0.75 * steps else lr_final
        g, acx, acy, dc, e = samples.batch(start)
        losses.append(net.train_step_coarse(batch, g, acx, acy, dc, samples.h2, samples.h2, rate,
                                            e if relative else None))
        if log and (step % 200 == 0 or step == steps - 1):
            log(f"step {step}: loss {losses[-1]:.4f}")
    return losses


# droplet families of the training systems (keyword arguments of build_system): the seeded-droplet
# distribution, larger droplets with surface tension, and one to three large droplets at high
# density ratio (the stationary-droplet step's kind of system)
FAMILIES = {
    "seeded": {},
    "large": dict(radius=(0.02, 0.3), sigma=(0.0, 1.0)),
    "few_large": dict(count=(1, 3), radius=(0.15, 0.3), ratio=(100.0, 1000.0), sigma=(0.5, 1.5)),
}


def learned_iterations(net, contexts, tol=1e-8):
    """iterations of the learned solve (mixed precision, `net` attached) on each context's system"""
    its = []
    for ctx in contexts:
        ctx.set_preconditioner(net)
        st = ctx.poisson_solve(ctx.array(), Context.opts(tol=tol, precision=PREC_MIXED, precond=PRECOND_LEARNED),
                               raise_on_fail=False)
        ctx.set_preconditioner(None)
        its.append(int(st.iterations) if st.converged == 1 else 10 ** 6)
    return its


def train_hybrid(n=2048, coarsen=32, systems=(32, 16, 16), steps=6000, rounds=3, candidates=2, validation=3, seed=0,
                 log=None, after_round=None):
    """The network for the bottom of the n x n hierarchy (the level of coarsening `coarsen`).

    Round 1 records coarse-correction inputs from multilevel solves of the training systems (the
    FAMILIES, `systems` of each) and trains `candidates` networks from different initial weights on
    them; the one whose learned solve takes the fewest iterations on `validation` systems of each
    family (seeds disjoint from the training ones) goes on.  Every further round records inputs from
    the learned solve with the current network (its own input distribution), adds them to the pool
    and continues training on all of it.  Loss: the relative A-norm error of the best step along the
    network's correction.  Returns the network and a record of the training."""
    vals = []
    for k, fam in enumerate(FAMILIES):
        for v in range(validation):
            ctx = system_context(n)
            build_system(ctx, np.random.default_rng(seed + 700 + 10 * k + v), **FAMILIES[fam])
            vals.append(ctx)
    record = {"grid": n, "coarsen": coarsen, "image": [n // coarsen, n // coarsen], "steps": steps,
              "families": {f: {kk: list(vv) for kk, vv in kw.items()} for f, kw in FAMILIES.items()},
              "systems": list(systems), "rounds": []}
    pools, net = [], None
    for rnd in range(rounds):
        t0 = time.perf_counter()
        for k, (fam, count) in enumerate(zip(FAMILIES, systems)):
            pools.append(make_coarse_samples(n, coarsen, systems=count, seed=seed + 1000 * rnd + 100 * k, net=net,
                                             **FAMILIES[fam]))
            if log:
                log(f"round {rnd + 1} {fam}: {pools[-1].count} samples ({time.perf_counter() - t0:.1f} s)")
        union = CoarseSamples.union(pools, seed=seed + rnd)
        t1 = time.perf_counter()
        rec = {"samples": union.count, "data_s": t1 - t0}
        if rnd == 0:
            best = None
            rec["candidates"] = []
            for c in range(candidates):
                cand = UNet(n, n, max_batch=8, seed=seed + 7 + 4 * c, coarsen=coarsen)
                losses = train_coarse(cand, union, steps=steps, batch=8, seed=seed + 9 + 4 * c, lr=2e-3, lr_final=5e-4)
                its = learned_iterations(cand, vals)
                rec["candidates"].append({"loss_last200": float(np.mean(losses[-200:])), "validation_iters": its})
                if log:
                    log(f"round 1 candidate {c + 1}: loss {np.mean(losses[-200:]):.4f}, validation iterations {its}")
                if best is None or sum(its) < best[0]:
                    best = (sum(its), cand, c)
            net = best[1]
            rec["chosen"] = best[2]
        else:
            losses = train_coarse(net, union, steps=steps, batch=8, seed=seed + 31 * rnd, lr=1e-3, lr_final=2.5e-4)
            rec["loss_last200"] = float(np.mean(losses[-200:]))
            rec["validation_iters"] = learned_iterations(net, vals)
        rec["train_s"] = time.perf_counter() - t1
        record["rounds"].append(rec)
        if log:
            log(f"round {rnd + 1}: {rec}")
        if after_round:
            after_round(rnd, net)
    return net, record


WEIGHTS = {32: "droplet_2048_c32.f32"}  # coarsening -> parameter file


def pretrained(n, coarsen=32, max_batch=1):
    """the committed network for the bottom of the n x n hierarchy (trained by train_hybrid on the
    droplet families at 2048^2: the weights are convolutional and serve any grid whose level of
    coarsening `coarsen` is a whole image), or None when there is none for this coarsening"""
    import os
    name = WEIGHTS.get(coarsen)
    if name is None or n % coarsen:
        return None
    net = UNet(n, n, max_batch=max_batch, coarsen=coarsen)
    load_params(net, os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", name))
    return net


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the learned coarse solve on the GPU and save its parameters")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--coarsen", type=int, default=32)
    ap.add_argument("--systems", default="32,16,16", help="training systems per family and round")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--candidates", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    net, rec = train_hybrid(a.n, a.coarsen, tuple(int(s) for s in a.systems.split(",")), a.steps, a.rounds,
                            a.candidates, seed=a.seed, log=lambda m: print(m, flush=True))
    save_params(net, a.out, rec)
    print(json.dumps(rec))
