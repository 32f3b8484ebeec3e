"""Training of the learned pressure preconditioner on the GPU.

The training systems are the GFM pressure systems of random droplet configurations, built by
the library itself: the level set of a union of circles (heavy phase inside, phi = max over
the circles of r_k - |x - c_k|), its curvature and Heaviside, the divergence of a smooth random
velocity, assembled by asd2_gfm_assemble (the path of MACSolver2D::SolvePoisson, reference
src/mac2d.cpp:1249) and exported in FP64.  The default distribution is the seeded-droplet one of
the GFM benchmark case: 1-20 non-overlapping circles of radius uniform in [0.02, 0.1], density
ratio log-uniform in [10, 1000], right-hand side minus the divergence of a smooth random velocity.
The residual samples are the right-hand side and the residuals b - A x_k after k iterations of a
solver: the mixed-precision Jacobi-PCG, or the flexible PCG with the multilevel coarse correction,
whose residuals are the ones a good coarse correction leaves behind.  numpy only draws the random
fields and stacks host copies into the FP32 padded layout of asd2_unet_train_step; the network's
forward and backward passes and the Adam steps are the library's kernels.
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


class CoarseSamples:
    """Coarse training systems on the device (FP32): g [S][H][W], acx / acy / dc [S][(H + 1)(W + 1)]"""

    def __init__(self, g, acx, acy, dc, h2):
        self.count = g.shape[0]
        self.H, self.W = g.shape[1], g.shape[2]
        self.h2 = h2
        self.dev = [DeviceArray.from_host(np.ascontiguousarray(a, np.float32)) for a in (g, acx, acy, dc)]
        self.strides = [self.H * self.W * 4] + [(self.H + 1) * (self.W + 1) * 4] * 3

    def batch(self, start):
        return [_View(d.ptr + start * st) for d, st in zip(self.dev, self.strides)]


def make_coarse_samples(n, coarsen, systems=16, iters=(0, 2, 5, 10, 20, 40, 70, 100), seed=0, net=None, log=None,
                        **kw):
    """Coarse-correction inputs g the solver itself produces on the level of coarsening `coarsen`
    (recorded by asd2_poisson_record_coarse at the given iterations of a multilevel solve, or of
    the learned solve with `net` attached), with that level's Galerkin operator, for `systems`
    random droplet systems on an n x n grid."""
    rng = np.random.default_rng(seed)
    ctx = system_context(n)
    H = W = n // coarsen
    rec = DeviceArray((len(iters), H, W))
    gs, ops = [], []
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
        for k in range(len(iters)):
            if np.any(g[k] != 0.0):
                gs.append(g[k])
                ops.append(op)
        if log:
            log(f"system {s + 1}/{systems}: {len(gs)} samples ({time.perf_counter() - t0:.1f} s)")
    order = rng.permutation(len(gs))
    g = np.stack([gs[i] for i in order])
    acx, acy, dc = (np.stack([ops[i][k] for i in order]) for k in range(3))
    return CoarseSamples(g, acx, acy, dc, (1.0 / n) ** 2 / coarsen)


def train_coarse(net: UNet, samples: CoarseSamples, steps=2000, batch=8, lr=2e-3, lr_final=5e-4, seed=0, log=None):
    """Adam on the coarse energy loss over random contiguous windows of the samples; returns the
    per-step losses"""
    rng = np.random.default_rng(seed)
    batch = min(batch, samples.count, net.max_batch)
    losses = []
    for step in range(steps):
        start = int(rng.integers(0, samples.count - batch + 1))
        rate = lr if step < 0.75 * steps else lr_final
        losses.append(net.train_step_coarse(batch, *samples.batch(start), samples.h2, samples.h2, rate))
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


def train_hybrid(n=2048, coarsen=32, systems=(32, 16, 16), steps=3000, rounds=2, seed=0, log=None):
    """The network for the bottom of the n x n hierarchy (the level of coarsening `coarsen`): round 1
    trains on coarse-correction inputs recorded from multilevel solves, every further round adds
    inputs recorded from the learned solve with the network of the round before (its own input
    distribution).  Returns the network and a record of the training."""
    net = UNet(n, n, max_batch=8, seed=seed + 7, coarsen=coarsen)
    record = {"grid": n, "coarsen": coarsen, "image": [n // coarsen, n // coarsen], "rounds": []}
    pools = []
    for rnd in range(rounds):
        t0 = time.perf_counter()
        new = []
        for k, (fam, count) in enumerate(zip(FAMILIES, systems)):
            new.append(make_coarse_samples(n, coarsen, systems=count, seed=seed + 1000 * rnd + 100 * k,
                                           net=net if rnd > 0 else None, **FAMILIES[fam]))
        pools.extend(new)
        t1 = time.perf_counter()
        losses = []
        schedule = [(p, steps if i == len(pools) - len(new) else steps // 2) for i, p in enumerate(pools)]
        for i, (p, k) in enumerate(schedule):
            lr = 2e-3 if rnd == 0 else 1e-3
            losses += train_coarse(net, p, steps=k, batch=8, seed=seed + 17 * rnd + i, lr=lr, lr_final=lr / 4)
        t2 = time.perf_counter()
        rec = {"samples": [p.count for p in new], "data_s": t1 - t0, "train_s": t2 - t1, "steps": len(losses),
               "loss_first": float(losses[0]), "loss_last100": float(np.mean(losses[-100:]))}
        record["rounds"].append(rec)
        if log:
            log(f"round {rnd + 1}: {rec}")
    return net, record


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the learned coarse solve on the GPU and save its parameters")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--coarsen", type=int, default=32)
    ap.add_argument("--systems", default="32,16,16", help="training systems per family and round")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    net, rec = train_hybrid(a.n, a.coarsen, tuple(int(s) for s in a.systems.split(",")), a.steps, a.rounds, a.seed,
                            log=lambda m: print(m, flush=True))
    rec["families"] = {k: {kk: list(vv) for kk, vv in v.items()} for k, v in FAMILIES.items()}
    save_params(net, a.out, rec)
    print(json.dumps(rec))
