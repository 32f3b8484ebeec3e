"""FP64 CPU restatement of the learned preconditioner's network (forward and backward).

TEST INFRASTRUCTURE ONLY: this module is the checker of the GPU network kernels
(asd2_amd/csrc/unet.hip) and of their training gradients; the product never imports it.

The network has no reference counterpart (the reference preconditions with Jacobi only,
src/poisson2d.cpp:346-347).  It is a 4-level U-Net on the coarse grid (H = nx / c rows along
x, W = ny / c columns along y, channels last):

    input  (4): coarse Jacobi-scaled residual after a Jacobi step, log10 coarse face
                coefficients (x, y faces) and log10 coarse diagonal, each about [-1, 1]
    enc0: conv3x3 4->16, relu; conv3x3 16->16, relu                 (H x W)
    enc1: avgpool2; conv3x3 16->32, relu; conv3x3 32->32, relu      (H/2)
    enc2: avgpool2; conv3x3 32->64, relu; conv3x3 64->64, relu      (H/4)
    enc3: avgpool2; conv3x3 64->128, relu; conv3x3 128->128, relu   (H/8)
    dec2: [enc2, up(enc3)] conv3x3 192->64, relu; conv3x3 64->64, relu
    dec1: [enc1, up(dec2)] conv3x3 96->32, relu; conv3x3 32->32, relu
    dec0: [enc0, up(dec1)] conv3x3 48->16, relu; conv3x3 16->16, relu
    out:  conv1x1 16->1 (linear)

"up" is nearest-neighbour upsampling, convolutions see zeros outside the grid.  The network
runs on a c x c block coarsening of the grid and supplies the direction of a coarse correction
added to Jacobi (the preconditioner and its loss are stated below).
Weights are stored [tap][c_out][c_in] with tap = 3 (di + 1) + (dj + 1).
"""
from __future__ import annotations

import numpy as np

C = (16, 32, 64, 128)
CIN = 4

# (name, c_in, c_out, kind, level) in execution order
LAYERS = [
    ("e0a", CIN, C[0]), ("e0b", C[0], C[0]),
    ("e1a", C[0], C[1]), ("e1b", C[1], C[1]),
    ("e2a", C[1], C[2]), ("e2b", C[2], C[2]),
    ("e3a", C[2], C[3]), ("e3b", C[3], C[3]),
    ("d2a", C[2] + C[3], C[2]), ("d2b", C[2], C[2]),
    ("d1a", C[1] + C[2], C[1]), ("d1b", C[1], C[1]),
    ("d0a", C[0] + C[1], C[0]), ("d0b", C[0], C[0]),
]
OUT = ("out", C[0], 1)


def flatten(p):
    """Parameters as the GPU library's FP32 vector: conv weights [9][c_out][c_in] in layer
    order, then all conv biases, then the 16 output weights and the output bias."""
    parts = [p[name + ".w"].ravel() for name, _, _ in LAYERS]
    parts += [p[name + ".b"].ravel() for name, _, _ in LAYERS]
    parts += [p["out.w"].ravel(), p["out.b"].ravel()]
    return np.concatenate(parts).astype(np.float32)


def unflatten(v):
    p, off = {}, 0
    for name, ci, co in LAYERS:
        p[name + ".w"] = np.asarray(v[off:off + 9 * co * ci], np.float64).reshape(9, co, ci)
        off += 9 * co * ci
    for name, ci, co in LAYERS:
        p[name + ".b"] = np.asarray(v[off:off + co], np.float64)
        off += co
    p["out.w"] = np.asarray(v[off:off + C[0]], np.float64).reshape(1, 1, C[0])
    off += C[0]
    p["out.b"] = np.asarray(v[off:off + 1], np.float64)
    return p


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed: int = 0):
    """He-uniform initial weights (the GPU trainer draws its own; this is for tests)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in LAYERS:
        lim = np.sqrt(6.0 / (9 * ci))
        p[name + ".w"] = rng.uniform(-lim, lim, (9, co, ci))
        p[name + ".b"] = rng.uniform(-0.1, 0.1, co)
    p["out.w"] = rng.uniform(-0.1, 0.1, (1, 1, C[0]))
    p["out.b"] = np.zeros(1)
    return p


# ------------------------------------------------------------------ primitive ops (HWC arrays)
def conv3x3(x, w, b):
    """x (H, W, ci), w (9, co, ci), b (co) -> (H, W, co); zero padding."""
    H, W_, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    y = np.broadcast_to(b, (H, W_, len(b))).copy()
    for t in range(9):
        di, dj = t // 3 - 1, t % 3 - 1
        y += xp[1 + di:1 + di + H, 1 + dj:1 + dj + W_, :] @ w[t].T
    return y


def conv3x3_backward(x, w, dy):
    """gradients of conv3x3 w.r.t. x, w, b."""
    H, W_, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    dyp = np.pad(dy, ((1, 1), (1, 1), (0, 0)))
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    for t in range(9):
        di, dj = t // 3 - 1, t % 3 - 1
        xs = xp[1 + di:1 + di + H, 1 + dj:1 + dj + W_, :]
        dw[t] = np.einsum("hwo,hwi->oi", dy, xs)
        # y[p] += x[p + d] w  =>  dx[q] += dy[q - d] w
        dx += dyp[1 - di:1 - di + H, 1 - dj:1 - dj + W_, :] @ w[t]
    return dx, dw, dy.sum(axis=(0, 1))


def pool2(x):
    H, W_, c = x.shape
    return x.reshape(H // 2, 2, W_ // 2, 2, c).mean(axis=(1, 3))


def pool2_backward(dy):
    return np.repeat(np.repeat(dy, 2, axis=0), 2, axis=1) * 0.25


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up2_backward(dy):
    H, W_, c = dy.shape
    return dy.reshape(H // 2, 2, W_ // 2, 2, c).sum(axis=(1, 3))


def relu(x):
    return np.maximum(x, 0.0)


# ------------------------------------------------------------------ network
def forward(p, f, bf=False, masks=None):
    """f (H, W, 4) -> y (H, W) and the activations needed by backward.  bf=True rounds the
    conv weights, the features and every stored activation to BF16 as the GPU kernels do.
    masks=None: the ReLU network.  masks = the activations of a ReLU pass: the linear pass of
    the network with that pass's ReLU pattern frozen and no biases (each conv output is
    multiplied by mask > 0), a linear function of f."""
    q = bf16 if bf else (lambda x: x)
    f = q(f)
    a = {"in": f, "bf": bf, "masks": masks}

    def cr(name, x):
        a[name + ".x"] = x
        if masks is None:
            y = q(relu(conv3x3(x, q(p[name + ".w"]), p[name + ".b"])))
        else:
            y = q(conv3x3(x, q(p[name + ".w"]), np.zeros(p[name + ".b"].shape)) * (masks[name] > 0))
        a[name] = y
        return y

    e0 = cr("e0b", cr("e0a", f))
    e1 = cr("e1b", cr("e1a", q(pool2(e0))))
    e2 = cr("e2b", cr("e2a", q(pool2(e1))))
    e3 = cr("e3b", cr("e3a", q(pool2(e2))))
    d2 = cr("d2b", cr("d2a", np.concatenate([e2, up2(e3)], axis=2)))
    d1 = cr("d1b", cr("d1a", np.concatenate([e1, up2(d2)], axis=2)))
    d0 = cr("d0b", cr("d0a", np.concatenate([e0, up2(d1)], axis=2)))
    a["out.x"] = d0
    y = d0 @ p["out.w"][0, 0] + (p["out.b"][0] if masks is None else 0.0)
    return y, a


def backward(p, a, dy):
    """dy (H, W) -> parameter gradients (same keys as p) and d(input).  After a bf=True
    forward, every stored gradient tensor is rounded to BF16 where the GPU stores one.  After
    a masked (linear) forward the ReLU derivatives are the frozen masks and the biases, which
    that pass does not use, get zero gradients."""
    bf = a.get("bf", False)
    masks = a.get("masks")
    q = bf16 if bf else (lambda x: x)
    g = {}
    d0 = a["out.x"]
    g["out.w"] = np.einsum("hw,hwc->c", dy, d0)[None, None, :]
    g["out.b"] = np.array([dy.sum() if masks is None else 0.0])

    def mask(name, d):
        return d * ((a[name] if masks is None else masks[name]) > 0)

    def crb(name, dpre):
        """dpre: gradient at the pre-activation of `name` -> (dw, db), d(input)"""
        dx, dw, db = conv3x3_backward(a[name + ".x"], q(p[name + ".w"]), dpre)
        g[name + ".w"] = dw
        g[name + ".b"] = db if masks is None else np.zeros_like(db)
        return dx

    c0, c1, c2 = C[0], C[1], C[2]
    d = q(mask("d0b", dy[:, :, None] * p["out.w"][0, 0][None, None, :]))
    d = q(mask("d0a", crb("d0b", d)))
    dx = q(crb("d0a", d))
    de0, dd1 = dx[:, :, :c0], q(mask("d1b", up2_backward(dx[:, :, c0:])))
    d = q(mask("d1a", crb("d1b", dd1)))
    dx = q(crb("d1a", d))
    de1, dd2 = dx[:, :, :c1], q(mask("d2b", up2_backward(dx[:, :, c1:])))
    d = q(mask("d2a", crb("d2b", dd2)))
    dx = q(crb("d2a", d))
    de2, de3 = dx[:, :, :c2], q(mask("e3b", up2_backward(dx[:, :, c2:])))
    d = q(mask("e3a", crb("e3b", de3)))
    dp = q(crb("e3a", d))
    d = q(mask("e2b", de2 + pool2_backward(dp)))
    d = q(mask("e2a", crb("e2b", d)))
    dp = q(crb("e2a", d))
    d = q(mask("e1b", de1 + pool2_backward(dp)))
    d = q(mask("e1a", crb("e1b", d)))
    dp = q(crb("e1a", d))
    d = q(mask("e0b", de0 + pool2_backward(dp)))
    d = q(mask("e0a", crb("e0b", d)))
    df = crb("e0a", d)
    return g, df


# ------------------------------------------------------------------ preconditioner and loss
# The network acts on a coarsening of the grid by c x c blocks (c a power of two; c = 1 keeps
# the grid).  With P the piecewise-constant prolongation (a coarse value copied to the c^2
# fine cells of its block) and P^T the block sum, the preconditioner is
#
#     z = D^-1 r + gamma P y,   y = net(features),   gamma = (y.g) / (y.A_c y),
#     g = P^T (r - A D^-1 r),   A_c = P^T A P,
#
# i.e. Jacobi plus a learned coarse correction along y scaled by the step that minimises the
# energy E(v) = v.A v - 2 v.r (E(D^-1 r + gamma P y) = E(D^-1 r) - (y.g)^2 / y.A_c y).  The exact
# coarse correction y = A_c^-1 g minimises E over the coarse space; the network approximates it.
# y is the network's linear pass on channel 0 with the ReLU pattern of a pass on the coefficient
# channels alone: for one system the map g -> y is linear, so the preconditioner varies with r
# only through the scalar gamma and CG keeps its Krylov acceleration (a ReLU network on the
# residual itself makes every iteration's preconditioner different and loses it).
# For a 5-point A, A_c is again 5-point: a coarse face couples the two blocks through the c fine
# faces on the block boundary, and the coarse diagonal collects the block's fine diagonal plus
# twice each face inside the block.

def coarse_op(ax_full, ay_full, dg, c):
    """A_c = P^T A P as (acx (Hc + 1, Wc), acy (Hc, Wc + 1), dc (Hc, Wc)), face layout as
    apply_op's."""
    H, W_ = dg.shape
    Hc, Wc = H // c, W_ // c
    acx = ax_full[::c, :].reshape(Hc + 1, Wc, c).sum(axis=2)
    acy = ay_full[:, ::c].reshape(Hc, c, Wc + 1).sum(axis=1)
    inner_x = ax_full[:H, :].copy()
    inner_x[::c, :] = 0.0          # faces on block boundaries are not inside a block
    inner_y = ay_full[:, :W_].copy()
    inner_y[:, ::c] = 0.0
    dc = block_sum(dg + 2.0 * inner_x + 2.0 * inner_y, c)
    return acx, acy, dc


def block_sum(v, c):
    H, W_ = v.shape
    return v.reshape(H // c, c, W_ // c, c).sum(axis=(1, 3))


def prolong(y, c):
    return np.repeat(np.repeat(y, c, axis=0), c, axis=1)


def _lg(v, h2):
    t = -v * h2
    return np.where(t > 0, (np.log10(np.maximum(t, 1e-300)) + 1.5) / 1.5, -2.0)


def features(r, ax_full, ay_full, dg, h2x, h2y, c):
    """Interior arrays -> (Hc, Wc, 4) network inputs: g / dc / s (s = rms over the coarse cells
    of g / dc), log10 coarse W-face and S-face coefficients and log10 coarse diagonal, each
    scaled by (c h)^2 / c^2 = h^2 / c so that a uniform coefficient gives the fine value; plus
    the pieces the preconditioner needs."""
    zj = r / dg
    w = r - apply_op(ax_full, ay_full, dg, zj)
    g = block_sum(w, c)
    acx, acy, dc = coarse_op(ax_full, ay_full, dg, c)
    u = g / dc
    s = np.sqrt(np.mean(u * u))
    s = s if s > 0 else 1.0
    h2 = min(h2x, h2y)
    fd = (np.log10(dc * h2 / c) - 0.5) / 1.5
    f = np.stack([u / s, _lg(acx[:-1, :], h2x / c), _lg(acy[:, :-1], h2y / c), fd], axis=2)
    return f, dict(zj=zj, w=w, g=g, acx=acx, acy=acy, dc=dc, s=s)


def gamma_of(y, g, acx, acy, dc):
    """Energy-optimal step along P y; 0 when y carries no energy."""
    den = float(np.sum(y * apply_op(acx, acy, dc, y)))
    num = float(np.sum(y * g))
    return num / den if den > 0 and np.isfinite(num / den) else 0.0


def apply_op(ax_full, ay_full, dg, z):
    """5-point operator on the interior: (A z)_c = dg z + ax_c z_W + ax_E z_E + ay_c z_S + ay_N z_N,
    ax_full / ay_full of shape (H + 1, W) / (H, W + 1) hold the coefficients of every face."""
    H, W_ = z.shape
    zp = np.pad(z, 1)
    return (dg * z + ax_full[:H, :] * zp[:-2, 1:-1] + ax_full[1:, :] * zp[2:, 1:-1]
            + ay_full[:, :W_] * zp[1:-1, :-2] + ay_full[:, 1:] * zp[1:-1, 2:])


def static_masks(p, f, bf=False):
    """ReLU activations of the network on the coefficient channels alone (channel 0 = 0): the
    masks of the linear pass, fixed for one system."""
    fs = f.copy()
    fs[:, :, 0] = 0.0
    _, a = forward(p, fs, bf=bf)
    return {name: a[name] for name, _, _ in LAYERS}


def network_output(p, f, bf=False, linear=True):
    """y of the preconditioner: the linear pass on channel 0 (masks from the coefficient
    channels) or, linear=False, the plain ReLU network on all four channels."""
    if not linear:
        return forward(p, f, bf=bf)
    fl = np.zeros_like(f)
    fl[:, :, 0] = f[:, :, 0]
    return forward(p, fl, bf=bf, masks=static_masks(p, f, bf))


def precondition(p, r, ax_full, ay_full, dg, h2x, h2y, c=1, bf=False, linear=True):
    f, aux = features(r, ax_full, ay_full, dg, h2x, h2y, c)
    y, a = network_output(p, f, bf=bf, linear=linear)
    gam = gamma_of(y, aux["g"], aux["acx"], aux["acy"], aux["dc"])
    z = aux["zj"] + gam * prolong(y, c)
    aux.update(y=y, a=a, f=f, gamma=gam)
    return z, aux


def loss_and_grad(p, r, ax_full, ay_full, dg, h2x, h2y, c=1, bf=False, linear=True):
    """L = E(z) / r.D^-1 r with E(v) = v.A v - 2 v.r and dL/dparams.  E(z) + e.A e = |z - e|_A^2
    for e = A^-1 r: the loss is the A-norm error of z as an approximate solve, shifted by a
    constant of the system, so it weights the smooth error that decides the CG iteration count
    (the squared residual |A z - r|^2 weights error by A^2 and ignores it).  With the optimal
    gamma, E(z) = E(D^-1 r) - (y.g)^2 / y.A_c y and dL/dy = -2 gamma (g - gamma A_c y) / r.D^-1 r.
    In the linear form the masks are treated as constants (they change with the weights only
    where a pre-activation crosses zero)."""
    z, aux = precondition(p, r, ax_full, ay_full, dg, h2x, h2y, c, bf, linear)
    cn = float(np.sum(r * aux["zj"]))
    L = float(np.sum(z * apply_op(ax_full, ay_full, dg, z)) - 2.0 * np.sum(z * r)) / cn
    gam = aux["gamma"]
    dy = -2.0 * gam * (aux["g"] - gam * apply_op(aux["acx"], aux["acy"], aux["dc"], aux["y"])) / cn
    g, _ = backward(p, aux["a"], dy)
    return L, g
