"""The committed learned coarse solve (asd2_amd/weights/droplet_2048_c32.f32, trained by
asd2_amd.learned.train_hybrid on the GPU) on held-out seeded-droplet systems: the distribution of
the GFM benchmark case, droplet layouts drawn from seeds that no training or validation system used.
The network is the bottom of the multilevel hierarchy (the level of 32 x 32 blocks, a 64^2 image at
2048^2, 16^2 at 512^2).  Each case is solved with the trained network, with a zero network (the
same solver with the bottom of its hierarchy removed: the ablation) and with the FP64 Jacobi-PCG
the reference uses (src/poisson2d.cpp:345-396); the learned solution is checked against an FP64
solution (the oracle's FP64 PCG at 512^2, the library's FP64 PCG at larger sizes, whose parity with
the oracle other tests pin) and its residual is recomputed on the host with the oracle's stencil."""
import numpy as np
import pytest

import asd2_amd as A
from asd2_amd import learned as LN
from tests.gfm_common import interior

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,seed", [(512, 901), (512, 903), (1024, 904), (2048, 901), (2048, 902)])
def test_committed_network_held_out(oracle, n, seed):
    from oracle import make_grid
    ctx = LN.system_context(n)
    LN.build_system(ctx, np.random.default_rng(seed))
    net = LN.pretrained(n)
    assert net is not None
    ctx.set_preconditioner(net)
    ps = ctx.array()
    st = ctx.poisson_solve(ps, ctx.opts(tol=1e-10, precond=A.PRECOND_LEARNED, precision=A.PREC_MIXED))
    zero = A.UNet(n, n, max_batch=1, coarsen=32)
    zero.set_params(np.zeros(zero.nparams, np.float32))
    ctx.set_preconditioner(zero)
    cap = 10 * st.iterations
    st_z = ctx.poisson_solve(ctx.array(), ctx.opts(tol=1e-10, max_iter=cap, precond=A.PRECOND_LEARNED,
                                                   precision=A.PREC_MIXED), raise_on_fail=False)
    ctx.set_preconditioner(None)
    ps64 = ctx.array()
    st64 = ctx.poisson_solve(ps64, ctx.opts(tol=1e-10, precision=A.PREC_FP64, mode=A.CG_STANDARD))
    ax, ay, dg, rhs = ctx.gfm_export()
    gr = make_grid(n, n, 1.0 / n, 1.0 / n)
    if n <= 512:
        x_ref, _, _, _ = oracle.pcg(gr, ax, ay, dg, rhs, 1e-12, 200000, 1)
    else:
        x_ref = ps64.download()
    x = ps.download()
    r = rhs - oracle.stencil_apply(gr, ax, ay, dg, x)
    res = float(np.linalg.norm(interior(gr, r)) / np.linalg.norm(interior(gr, rhs)))
    err = _rel(interior(gr, x), interior(gr, x_ref))
    zero_its = st_z.iterations if st_z.converged == 1 else f">{cap}"
    print(f"{n}^2 seed {seed}: learned {st.iterations} iters ({st.reliable_updates} updates), zero network "
          f"{zero_its}, FP64 Jacobi-PCG {st64.iterations}; host residual {res:.2e}, rel L2 {err:.2e}")
    assert st.converged == 1 and st.fallbacks == 0
    assert res <= 1.01e-10 and err <= 1e-6
    # what the network gains: the solver without it takes at least 4x the iterations
    assert st_z.converged != 1 or st_z.iterations >= 4 * st.iterations
    assert st.iterations * 10 <= st64.iterations
