"""The CPU oracle of the learned preconditioner (oracle/unet.py) pinned by mathematics: every
analytic gradient against central finite differences in FP64, the adjoint identity of each
primitive, the coarse operator against the dense Galerkin product P^T A P, the energy loss
against a direct solve, the optimality of the coarse step, and the preconditioner's positive
homogeneity.  (The network has no reference counterpart, so nothing else can pin it.)"""
import numpy as np
import pytest

from oracle import unet as U


def _system(rng, H, W):
    """A random SPD 5-point operator with density-like face coefficients."""
    beta_x = 10.0 ** rng.uniform(-3, 0, (H + 1, W))
    beta_y = 10.0 ** rng.uniform(-3, 0, (H, W + 1))
    beta_x[0, :] = 0.0
    beta_x[-1, :] = 0.0
    beta_y[:, 0] = 0.0
    beta_y[:, -1] = 0.0
    beta_x[0, 3] = 0.5  # a Dirichlet-like boundary face keeps the operator definite
    h2 = 1.0 / (H * W)
    ax = -beta_x / h2
    ay = -beta_y / h2
    dg = (beta_x[:-1, :] + beta_x[1:, :] + beta_y[:, :-1] + beta_y[:, 1:]) / h2
    return ax, ay, dg, h2


def test_conv_adjoint_and_pool_up_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 8, 4))
    w = rng.standard_normal((9, 5, 4))
    b = rng.standard_normal(5)
    dy = rng.standard_normal((6, 8, 5))
    y = U.conv3x3(x, w, np.zeros(5))
    dx, dw, db = U.conv3x3_backward(x, w, dy)
    # <dy, conv(x)> = <dx, x> (linear in x), and = <dw, w> (linear in w)
    assert np.isclose(np.sum(dy * y), np.sum(dx * x))
    assert np.isclose(np.sum(dy * y), np.sum(dw * w))
    assert np.allclose(db, dy.sum(axis=(0, 1)))
    assert np.allclose(U.conv3x3(x, w, b) - y, b)
    p = rng.standard_normal((6, 8, 3))
    q = rng.standard_normal((3, 4, 3))
    assert np.isclose(np.sum(U.pool2(p) * q), np.sum(p * U.pool2_backward(q)))
    assert np.isclose(np.sum(U.up2(q) * p), np.sum(q * U.up2_backward(p)))


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((5, 7, 3))
    w = rng.standard_normal((9, 2, 3))
    b = rng.standard_normal(2)
    y = U.conv3x3(x, w, b)
    i, j, o = 0, 6, 1  # a corner output: taps outside the grid see zeros
    ref = b[o]
    for t in range(9):
        ii, jj = i + t // 3 - 1, j + t % 3 - 1
        if 0 <= ii < 5 and 0 <= jj < 7:
            ref += w[t, o] @ x[ii, jj]
    assert np.isclose(y[i, j, o], ref)


def _dense(ax, ay, dg):
    H, W = dg.shape
    n = H * W
    cols = np.eye(n)
    return np.stack([U.apply_op(ax, ay, dg, cols[:, k].reshape(H, W)).ravel() for k in range(n)], axis=1)


@pytest.mark.parametrize("c,linear", [(1, True), (2, True), (2, False)])
def test_loss_gradient_matches_finite_differences(c, linear):
    rng = np.random.default_rng(3 + c)
    H = W = 16 * c
    ax, ay, dg, h2 = _system(rng, H, W)
    r = rng.standard_normal((H, W))
    p = U.init_params(4)
    L, g = U.loss_and_grad(p, r, ax, ay, dg, h2, h2, c, linear=linear)
    assert np.isfinite(L)
    eps = 1e-6
    checked = 0
    for key in ["e0a.w", "e0b.b", "e1b.w", "e2a.w", "e3a.w", "e3b.b", "d2a.w", "d1a.b", "d0a.w", "d0b.w",
                "out.w", "out.b"]:
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in p[key].shape)
            old = p[key][idx]
            p[key][idx] = old + eps
            lp, _ = U.loss_and_grad(p, r, ax, ay, dg, h2, h2, c, linear=linear)
            p[key][idx] = old - eps
            lm, _ = U.loss_and_grad(p, r, ax, ay, dg, h2, h2, c, linear=linear)
            p[key][idx] = old
            fd = (lp - lm) / (2 * eps)
            an = g[key][idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)) + 1e-7, (key, idx, fd, an)
            checked += 1
    assert checked == 36


def test_input_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = U.init_params(6)
    f = rng.standard_normal((8, 8, U.CIN))
    dy = rng.standard_normal((8, 8))
    y, a = U.forward(p, f)
    _, df = U.backward(p, a, dy)
    eps = 1e-6
    for _ in range(8):
        idx = tuple(rng.integers(0, s) for s in f.shape)
        fp, fm = f.copy(), f.copy()
        fp[idx] += eps
        fm[idx] -= eps
        fd = (np.sum(U.forward(p, fp)[0] * dy) - np.sum(U.forward(p, fm)[0] * dy)) / (2 * eps)
        assert abs(fd - df[idx]) <= 1e-6 * max(1.0, abs(fd))


def test_linear_pass_is_linear_and_masked():
    """The masked pass is additive and homogeneous in channel 0 for any sign, ignores the
    coefficient channels of its input, and equals the ReLU pass where that pass's own pattern
    is used with zero biases."""
    rng = np.random.default_rng(13)
    p = U.init_params(14)
    f = rng.standard_normal((16, 16, U.CIN))
    m = U.static_masks(p, f)
    x1, x2 = np.zeros_like(f), np.zeros_like(f)
    x1[:, :, 0] = rng.standard_normal((16, 16))
    x2[:, :, 0] = rng.standard_normal((16, 16))
    y1, _ = U.forward(p, x1, masks=m)
    y2, _ = U.forward(p, x2, masks=m)
    y12, _ = U.forward(p, x1 - 2.5 * x2, masks=m)
    assert np.allclose(y12, y1 - 2.5 * y2, rtol=1e-12, atol=1e-12 * np.abs(y1).max())
    yn, _ = U.network_output(p, np.concatenate([x1[:, :, :1], f[:, :, 1:]], axis=2))
    assert np.allclose(yn, y1, rtol=1e-13, atol=0)
    p0 = {k: (np.zeros_like(v) if k.endswith(".b") else v) for k, v in p.items()}
    yr, a = U.forward(p0, x1)
    yl, _ = U.forward(p0, x1, masks={name: a[name] for name, _, _ in U.LAYERS})
    assert np.allclose(yl, yr, rtol=1e-12, atol=1e-14)


def test_preconditioner_positively_homogeneous():
    rng = np.random.default_rng(7)
    ax, ay, dg, h2 = _system(rng, 32, 32)
    p = U.init_params(8)
    r = rng.standard_normal((32, 32))
    z1, _ = U.precondition(p, r, ax, ay, dg, h2, h2, 2)
    z2, _ = U.precondition(p, 3.5 * r, ax, ay, dg, h2, h2, 2)
    assert np.allclose(z2, 3.5 * z1, rtol=1e-12, atol=0)


@pytest.mark.parametrize("c", [2, 4])
def test_coarse_operator_is_galerkin_product(c):
    rng = np.random.default_rng(20 + c)
    H, W = 4 * c, 6 * c
    ax, ay, dg, h2 = _system(rng, H, W)
    Am = _dense(ax, ay, dg)
    P = np.kron(np.kron(np.eye(H // c), np.ones((c, 1))), np.kron(np.eye(W // c), np.ones((c, 1))))
    acx, acy, dc = U.coarse_op(ax, ay, dg, c)
    assert np.allclose(_dense(acx, acy, dc), P.T @ Am @ P, rtol=1e-12, atol=1e-9 * np.abs(Am).max())
    v = rng.standard_normal((H // c, W // c))
    assert np.allclose(U.prolong(v, c).ravel(), P @ v.ravel())
    assert np.allclose(U.block_sum(U.prolong(v, c), c), c * c * v)


def test_coarse_step_is_energy_optimal():
    """gamma minimises E along P y, E never exceeds Jacobi's, and the exact coarse solve
    y = A_c^-1 g has gamma = 1 and the lowest energy over the coarse space."""
    rng = np.random.default_rng(30)
    c, H = 2, 16
    ax, ay, dg, h2 = _system(rng, H, H)
    r = rng.standard_normal((H, H))
    Am = _dense(ax, ay, dg)

    def E(v):
        return v.ravel() @ Am @ v.ravel() - 2.0 * v.ravel() @ r.ravel()

    f, aux = U.features(r, ax, ay, dg, h2, h2, c)
    acx, acy, dc, g = aux["acx"], aux["acy"], aux["dc"], aux["g"]
    y = rng.standard_normal(g.shape)
    gam = U.gamma_of(y, g, acx, acy, dc)
    e0 = E(aux["zj"] + gam * U.prolong(y, c))
    for t in (0.9, 1.1, -1.0):
        assert e0 <= E(aux["zj"] + t * gam * U.prolong(y, c)) + 1e-12 * abs(e0)
    assert e0 <= E(aux["zj"])
    ystar = np.linalg.solve(_dense(acx, acy, dc), g.ravel()).reshape(g.shape)
    assert np.isclose(U.gamma_of(ystar, g, acx, acy, dc), 1.0, rtol=1e-9)
    estar = E(aux["zj"] + U.prolong(ystar, c))
    for _ in range(5):
        y2 = ystar + 0.1 * rng.standard_normal(g.shape)
        assert estar <= E(aux["zj"] + U.gamma_of(y2, g, acx, acy, dc) * U.prolong(y2, c)) + 1e-12 * abs(estar)


def test_energy_loss_is_shifted_a_norm_error():
    """L + e.A e / c = |z - e|_A^2 / c with e = A^-1 r (direct sparse solve), c = r.D^-1 r."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(11)
    H = W = 16
    ax, ay, dg, h2 = _system(rng, H, W)
    r = rng.standard_normal((H, W))
    p = U.init_params(12)
    p["out.w"][:] = rng.uniform(-0.1, 0.1, p["out.w"].shape)
    L, _ = U.loss_and_grad(p, r, ax, ay, dg, h2, h2, 2)
    z, _ = U.precondition(p, r, ax, ay, dg, h2, h2, 2)
    Am = _dense(ax, ay, dg)
    assert np.allclose(Am, Am.T, rtol=0, atol=1e-9 * np.abs(Am).max())
    e = spla.spsolve(sp.csr_matrix(Am), r.ravel()).reshape(H, W)
    c = np.sum(r * r / dg)
    d = (z - e).ravel()
    assert np.isclose(L + e.ravel() @ Am @ e.ravel() / c, d @ Am @ d / c, rtol=1e-9)
    # the loss of Jacobi alone is R - 2 with R the Rayleigh quotient of D^-1/2 A D^-1/2
    p["out.w"][:] = 0.0
    L0, _ = U.loss_and_grad(p, r, ax, ay, dg, h2, h2, 2)
    zj = (r / dg).ravel()
    assert np.isclose(L0, (zj @ Am @ zj) / c - 2.0, rtol=1e-12)


def test_zero_network_is_jacobi():
    rng = np.random.default_rng(9)
    ax, ay, dg, h2 = _system(rng, 16, 16)
    p = U.init_params(10)
    p["out.w"][:] = 0.0
    p["out.b"][:] = 0.0
    r = rng.standard_normal((16, 16))
    z, _ = U.precondition(p, r, ax, ay, dg, h2, h2, 2)
    assert np.allclose(z, r / dg, rtol=1e-14, atol=0)
