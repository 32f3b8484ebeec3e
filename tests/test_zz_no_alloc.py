"""asd2_ls_advect_noalloc (no device allocation; stage arrays in the library's static per-GPU
buffer) against asd2_ls_advect, through the C ABI: bit-identical level sets on two calls in a row,
the second on a smaller grid (the buffer then holds the first call's values), with Neumann and with
CUSTOM pressure sides; refusals of grids beyond the buffer and of level-set-free sides."""
import numpy as np
import pytest

from oracle import make_grid
from tests import scenarios as S

pytestmark = pytest.mark.gpu


def _advect_both(A, nx, ny, bc, seed):
    dx = 1.0 / max(nx, ny)
    ctx = A.Context(nx, ny, dx, dx)
    ctx.set_bc(A.GRID_P, bc)
    phi = S.circle_sdf(nx, ny, dx, dx, 0.45, 0.55, 0.22)
    phi += S.random_field(seed, phi.shape, 1e-3)
    u, v = S.single_vortex(nx, ny, dx, dx)
    du, dv = ctx.array(u), ctx.array(v)
    a, b = ctx.array(phi), ctx.array(phi)
    dt = 0.5 * dx
    print(f"noalloc: grid {nx}x{ny} bc {bc} seed {seed}", flush=True)
    for _ in range(2):  # two steps: the second starts from the first's ghosts
        ctx.ls_advect(a, du, dv, dt)
        ctx.ls_advect_noalloc(b, du, dv, dt)
    ctx.sync()
    return a.download(), b.download(), phi, u, v, dt, dx


@pytest.mark.parametrize("bc", [("neumann",) * 4, ("neumann", "custom", "dirichlet", "neumann")])
def test_noalloc_matches_advect_on_two_grids(oracle, bc):
    import asd2_amd as A
    for nx, ny, seed in ((150, 132, 1), (41, 37, 2)):  # second call smaller
        ref, got, phi, u, v, dt, dx = _advect_both(A, nx, ny, bc, seed)
        assert np.array_equal(ref, got), f"{nx}x{ny}: max diff {np.abs(ref - got).max()}"
        if bc == ("neumann",) * 4:
            gr = make_grid(nx, ny, dx, dx)
            want = phi.copy()
            for _ in range(2):
                oracle.solve_levelset(gr, want, u, v, dt)
            assert np.abs(got - want).max() <= 1e-12


def test_noalloc_refuses_what_it_cannot_hold():
    import asd2_amd as A
    big = A.Context(2100, 2100, 1.0 / 2100, 1.0 / 2100)  # (2106)^2 padded cells > the buffer
    big.set_bc(A.GRID_P, ("neumann",) * 4)
    tiny = A.DeviceArray((8,))
    with pytest.raises(A.Asd2Error) as e:
        big.ls_advect_noalloc(tiny, tiny, tiny, 1e-3)
    assert e.value.code == A.ERR_ARG and "exceed" in str(e.value)
    small = A.Context(16, 16, 1.0 / 16, 1.0 / 16)
    small.set_bc(A.GRID_P, ("lsfree", "neumann", "neumann", "neumann"))
    arr = small.array()
    with pytest.raises(A.Asd2Error) as e:
        small.ls_advect_noalloc(arr, arr, arr, 1e-3)
    assert e.value.code == A.ERR_UNSUPPORTED
