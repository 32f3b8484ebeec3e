"""Python binding of the gfx950 two-phase hot path (C ABI of ``include/asd2_hip.h``).

The library ``libasd2_hip.so`` is built in-tree by ``asd2_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback.  Device memory is allocated through
the library itself (``DeviceArray``), so no other GPU runtime user is needed in-process.
Arrays use the reference layout (padded by g = 3, j fastest), i.e. numpy arrays of shape
``(nx + 6, ny + 6)`` in C order.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libasd2_hip.so")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_HIP, ERR_NOT_CONVERGED, ERR_NONFINITE, ERR_STATE = range(7)
BC = {"periodic": 0, "neumann": 1, "dirichlet": 2, "axisym": 3, "wall": 4,
      "inlet": 5, "outlet": 6, "pressure": 7, "lsfree": 8, "custom": 9}
GRID_P, GRID_U, GRID_V = 0, 1, 2
CG_REFERENCE, CG_STANDARD = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_MULTILEVEL = 0, 1, 2
MOM_RHS, MOM_CONVECTION, MOM_VISCOSITY = 0, 1, 2
KID = {"weno_stage": 1, "reinit": 2, "pcg64_spmv": 3, "pcg64_update": 4, "pcg32_spmv": 5,
       "pcg32_update": 6, "gfm": 7, "coarse": 8, "fcg_begin": 9, "fcg_spmv": 10, "fcg_update": 11,
       "fast_pass": 12, "fast_gemm": 13, "fast_tridiag": 14, "momentum": 15}


class Asd2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"asd2 error {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("num_bc_grid", C.c_int),
                ("base_x", C.c_double), ("base_y", C.c_double), ("dx", C.c_double), ("dy", C.c_double)]


class Fluid(C.Structure):
    _fields_ = [("rho_h", C.c_double), ("rho_l", C.c_double), ("mu_h", C.c_double),
                ("mu_l", C.c_double), ("sigma", C.c_double), ("dt", C.c_double),
                ("g", C.c_double), ("g_axis", C.c_int), ("gravity_fix", C.c_int),
                ("re", C.c_double), ("we", C.c_double), ("fr", C.c_double), ("cfl", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("mode", C.c_int),
                ("precision", C.c_int), ("precond", C.c_int), ("reliable_delta", C.c_double),
                ("check_every", C.c_int)]


class LsDiag(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("cells", "mass", "centroid_x", "centroid_y", "length", "circularity",
                                          "err_linf", "err_l1", "err_l2", "ref_cells", "mass_loss_pct")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("reliable_updates", C.c_int), ("fallbacks", C.c_int),
                ("converged", C.c_int), ("rel_residual", C.c_double),
                ("true_rel_residual", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def make_fluid(rho_h, rho_l, mu_h, mu_l, sigma, dt, g=0.0, g_axis=1, gravity_fix=1,
               L=1.0, U=1.0, cfl=0.5) -> Fluid:
    """Dimensionless groups as built by the dimensional MACSolver2D constructor
    (reference src/mac2d.cpp:28-49)."""
    fl = Fluid()
    fl.rho_h, fl.rho_l, fl.mu_h, fl.mu_l = rho_h, rho_l, mu_h, mu_l
    fl.sigma, fl.dt, fl.g, fl.g_axis, fl.gravity_fix = sigma, dt, g, g_axis, gravity_fix
    fl.re = rho_h * L * U / mu_h if mu_h != 0 else float("inf")
    fl.we = rho_h * L * U * U / sigma if sigma != 0 else float("inf")
    fl.fr = U * U / (g * L) if g != 0 else float("inf")
    fl.cfl = cfl
    return fl


_lib = None
_VP = C.c_void_p


def load() -> C.CDLL:
    """Load libasd2_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m asd2_amd.build` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    sigs = {
        "asd2_last_error": ([], C.c_char_p), "asd2_version": ([], C.c_int),
        "asd2_device_count": ([P(C.c_int)], C.c_int),
        "asd2_device_name": ([C.c_char_p, C.c_int], C.c_int),
        "asd2_malloc": ([P(_VP), C.c_size_t], C.c_int), "asd2_free": ([_VP], C.c_int),
        "asd2_memcpy_h2d": ([_VP, _VP, C.c_size_t], C.c_int),
        "asd2_memcpy_d2h": ([_VP, _VP, C.c_size_t], C.c_int),
        "asd2_memcpy_d2d": ([_VP, _VP, C.c_size_t], C.c_int),
        "asd2_device_synchronize": ([], C.c_int),
        "asd2_stream_copy": ([_VP, _VP, _VP, C.c_size_t], C.c_int),
        "asd2_fill": ([_VP, _VP, C.c_size_t, C.c_double], C.c_int),
        "asd2_mac_pressure": ([_VP, _VP, _VP, C.c_double], C.c_int),
        "asd2_create": ([P(Grid), P(_VP)], C.c_int), "asd2_destroy": ([_VP], C.c_int),
        "asd2_synchronize": ([_VP], C.c_int),
        "asd2_set_bc": ([_VP, C.c_int, P(C.c_int), P(C.c_double)], C.c_int),
        "asd2_bc_apply": ([_VP, C.c_int, _VP], C.c_int),
        "asd2_ls_advect": ([_VP, _VP, _VP, _VP, C.c_double], C.c_int),
        "asd2_ls_advect_noalloc": ([_VP, _VP, _VP, _VP, C.c_double], C.c_int),
        "asd2_ls_weno_rhs": ([_VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_ls_reinit_minrk2": ([_VP, _VP, C.c_int], C.c_int),
        "asd2_ls_reinit_sussman": ([_VP, _VP, C.c_int], C.c_int),
        "asd2_ls_subcell_absgrad": ([_VP, _VP, _VP, _VP], C.c_int),
        "asd2_ls_eno_absgrad": ([_VP, _VP, _VP, _VP], C.c_int),
        "asd2_ls_reinit_iterations": ([C.c_double, C.c_double], C.c_int),
        "asd2_mac_kappa": ([_VP, _VP, _VP], C.c_int),
        "asd2_mac_heaviside": ([_VP, _VP, _VP, C.c_int], C.c_int),
        "asd2_mac_divergence": ([_VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_update_vel": ([_VP, P(Fluid), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_update_dt": ([_VP, P(Fluid), _VP, _VP, P(C.c_double)], C.c_int),
        "asd2_mac_intermediate_vel": ([_VP, P(Fluid), C.c_int, _VP, _VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_gfm_assemble": ([_VP, P(Fluid), _VP, _VP, _VP, _VP], C.c_int),
        "asd2_gfm_assemble_axisym": ([_VP, P(Fluid), _VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_divergence_axisym": ([_VP, _VP, _VP, _VP, C.c_int], C.c_int),
        "asd2_mac_kappa_axisym": ([_VP, _VP, _VP], C.c_int),
        "asd2_mac_update_dt_axisym": ([_VP, P(Fluid), _VP, _VP, _VP, P(C.c_double)], C.c_int),
        "asd2_tdma": ([_VP, C.c_int, C.c_int, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_rhs_u_axisym": ([_VP, P(Fluid), _VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_rhs_v_axisym": ([_VP, P(Fluid), _VP, _VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_momentum_term_axisym": ([_VP, P(Fluid), C.c_int, C.c_int, _VP, _VP, _VP, _VP, _VP, _VP],
                                          C.c_int),
        "asd2_mac_uhat_axisym": ([_VP, P(Fluid), _VP, _VP, _VP, _VP], C.c_int),
        "asd2_mac_vhat_axisym": ([_VP, P(Fluid), _VP, _VP, _VP, _VP], C.c_int),
        "asd2_gfm_export": ([_VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_gfm_import": ([_VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_gfm_apply": ([_VP, _VP, _VP], C.c_int),
        "asd2_poisson_solve": ([_VP, P(SolveOpts), _VP, P(SolveStats)], C.c_int),
        "asd2_poisson_bicgstab": ([_VP, P(SolveOpts), _VP, P(SolveStats)], C.c_int),
        "asd2_csr_pcg": ([_VP, C.c_int, _VP, _VP, _VP, _VP, _VP, _VP, P(SolveOpts), P(SolveStats)], C.c_int),
        "asd2_csr_gs": ([_VP, C.c_int, _VP, _VP, _VP, _VP, _VP, C.c_int, P(SolveStats)], C.c_int),
        "asd2_csr_bicgstab": ([_VP, C.c_int, _VP, _VP, _VP, _VP, _VP, _VP, P(SolveOpts), P(SolveStats)],
                              C.c_int),
        "asd2_ls_reinit_original": ([_VP, _VP, C.c_int], C.c_int),
        "asd2_ls_reinit_first_time": ([_VP, _VP, C.c_int], C.c_int),
        "asd2_ls_weno_absgrad": ([_VP, _VP, _VP, _VP], C.c_int),
        "asd2_ls_heaviside_deriv": ([_VP, _VP, _VP], C.c_int),
        "asd2_ls_smoothed_sign": ([_VP, _VP, _VP], C.c_int),
        "asd2_ls_sussman_constraint": ([_VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_ls_unit_weno5": ([_VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int], C.c_int),
        "asd2_ls_diagnostics": ([_VP, _VP, _VP, P(LsDiag)], C.c_int),
        "asd2_mac_momentum_term": ([_VP, P(Fluid), C.c_int, C.c_int, _VP, _VP, _VP, _VP, _VP], C.c_int),
        "asd2_unet_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, P(_VP)], C.c_int),
        "asd2_unet_destroy": ([_VP], C.c_int),
        "asd2_unet_param_count": ([_VP, P(C.c_longlong)], C.c_int),
        "asd2_unet_set_params": ([_VP, _VP, C.c_longlong], C.c_int),
        "asd2_unet_get_params": ([_VP, _VP, C.c_longlong], C.c_int),
        "asd2_unet_get_grads": ([_VP, _VP, C.c_longlong], C.c_int),
        "asd2_unet_forward": ([_VP, C.c_int, _VP, _VP], C.c_int),
        "asd2_unet_backward": ([_VP, C.c_int, _VP], C.c_int),
        "asd2_unet_forward_linear": ([_VP, C.c_int, _VP, _VP], C.c_int),
        "asd2_unet_backward_linear": ([_VP, C.c_int, _VP], C.c_int),
        "asd2_unet_train_step": ([_VP, C.c_int, _VP, _VP, _VP, _VP, _VP, C.c_double, C.c_double, C.c_float,
                                  P(C.c_float)], C.c_int),
        "asd2_set_preconditioner": ([_VP, _VP], C.c_int),
        "asd2_unet_train_step_coarse": ([_VP, C.c_int, _VP, _VP, _VP, _VP, _VP, C.c_double, C.c_double,
                                         C.c_float, C.POINTER(C.c_float)], C.c_int),
        "asd2_poisson_record_coarse": ([_VP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _VP], C.c_int),
        "asd2_fast_poisson": ([_VP, _VP, _VP, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
        "asd2_timer_start": ([_VP], C.c_int),
        "asd2_timer_stop": ([_VP, P(C.c_double)], C.c_int),
        "asd2_profile_kernel": ([_VP, C.c_int], C.c_int),
        "asd2_profile_read": ([_VP, P(C.c_int), P(C.c_double)], C.c_int),
        "asd2_profile_read_each": ([_VP, P(C.c_float), C.c_int, P(C.c_int)], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(code: int) -> int:
    if code != OK:
        raise Asd2Error(code, load().asd2_last_error().decode(errors="replace"))
    return code


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(load().asd2_device_name(buf, 256))
    return buf.value.decode()


class DeviceArray:
    """A zero-initialised device allocation owned by Python."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape) if not isinstance(shape, int) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _VP()
        check(load().asd2_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(load().asd2_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(load().asd2_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def copy_from(self, other: "DeviceArray") -> "DeviceArray":
        assert other.nbytes == self.nbytes
        check(load().asd2_memcpy_d2d(self.ptr, other.ptr, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            load().asd2_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _bc_codes(kinds):
    return (C.c_int * 4)(*[BC[k] if isinstance(k, str) else int(k) for k in kinds])


@dataclass
class Stats:
    iterations: int
    reliable_updates: int
    fallbacks: int
    converged: int
    rel_residual: float
    true_rel_residual: float


class UNet:
    """The learned preconditioner's network (asd2_unet) for nx x ny grids coarsened by c x c
    blocks: its images are the (nx / c) x (ny / c) coarse grids."""

    def __init__(self, nx, ny, max_batch=1, seed=0, coarsen=1):
        self.lib = load()
        h = _VP()
        check(self.lib.asd2_unet_create(nx, ny, coarsen, max_batch, seed, C.byref(h)))
        self.h = h.value
        self.nx, self.ny, self.max_batch, self.coarsen = nx, ny, max_batch, coarsen
        self.H, self.W = nx // coarsen, ny // coarsen
        cnt = C.c_longlong(0)
        check(self.lib.asd2_unet_param_count(self.h, C.byref(cnt)))
        self.nparams = cnt.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.asd2_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, v):
        v = np.ascontiguousarray(v, np.float32)
        check(self.lib.asd2_unet_set_params(self.h, v.ctypes.data, v.size))

    def get_params(self):
        v = np.empty(self.nparams, np.float32)
        check(self.lib.asd2_unet_get_params(self.h, v.ctypes.data, v.size))
        return v

    def get_grads(self):
        v = np.empty(self.nparams, np.float32)
        check(self.lib.asd2_unet_get_grads(self.h, v.ctypes.data, v.size))
        return v

    def forward(self, feat):
        """feat (B, H, W, 4) -> y (B, H, W) on the coarse grid"""
        feat = np.ascontiguousarray(feat, np.float32)
        b = feat.shape[0]
        d_f = DeviceArray.from_host(feat)
        d_y = DeviceArray((b, self.H, self.W), np.float32)
        check(self.lib.asd2_unet_forward(self.h, b, d_f.ptr, d_y.ptr))
        return d_y.download()

    def backward(self, dy):
        dy = np.ascontiguousarray(dy, np.float32)
        d_dy = DeviceArray.from_host(dy)
        check(self.lib.asd2_unet_backward(self.h, dy.shape[0], d_dy.ptr))

    def forward_linear(self, feat):
        """the preconditioner's linear form: feat (B, H, W, 4) -> y (B, H, W), linear in
        channel 0 with the ReLU pattern of channels 1..3"""
        feat = np.ascontiguousarray(feat, np.float32)
        b = feat.shape[0]
        d_f = DeviceArray.from_host(feat)
        d_y = DeviceArray((b, self.H, self.W), np.float32)
        check(self.lib.asd2_unet_forward_linear(self.h, b, d_f.ptr, d_y.ptr))
        return d_y.download()

    def backward_linear(self, dy):
        dy = np.ascontiguousarray(dy, np.float32)
        d_dy = DeviceArray.from_host(dy)
        check(self.lib.asd2_unet_backward_linear(self.h, dy.shape[0], d_dy.ptr))

    def train_step_coarse(self, batch, d_g, d_acx, d_acy, d_dc, h2x, h2y, lr, d_energy=None):
        """one Adam step on `batch` coarse systems (device FP32: g [batch][H][W], acx / acy / dc
        [batch][(H + 1) (W + 1)], energy g.A^-1 g [batch] or None); returns the loss before the step"""
        loss = C.c_float(0.0)
        check(self.lib.asd2_unet_train_step_coarse(self.h, batch, d_g.ptr, d_acx.ptr, d_acy.ptr, d_dc.ptr,
                                                   d_energy.ptr if d_energy is not None else None, h2x, h2y, lr,
                                                   C.byref(loss)))
        return loss.value

    def train_step(self, batch, d_r, d_dinv, d_ax, d_ay, d_dg, dx, dy, lr):
        """one Adam step; d_* are device arrays of `batch` FP32 padded systems; returns the loss"""
        loss = C.c_float(0.0)
        check(self.lib.asd2_unet_train_step(self.h, batch, d_r.ptr, d_dinv.ptr, d_ax.ptr, d_ay.ptr, d_dg.ptr,
                                            dx, dy, lr, C.byref(loss)))
        return loss.value


class Context:
    """One grid on the device (the C ABI's asd2_ctx) with its BCs, stream and workspace."""

    def __init__(self, nx, ny, dx, dy, base_x=0.0, base_y=0.0, num_bc_grid=3):
        self.lib = load()
        g = Grid(nx, ny, num_bc_grid, base_x, base_y, dx, dy)
        h = _VP()
        check(self.lib.asd2_create(C.byref(g), C.byref(h)))
        self.h = h.value
        self.nx, self.ny, self.g, self.dx, self.dy = nx, ny, num_bc_grid, dx, dy
        self.shape = (nx + 2 * num_bc_grid, ny + 2 * num_bc_grid)

    def close(self):
        if getattr(self, "h", None):
            self.lib.asd2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, host: np.ndarray | None = None) -> DeviceArray:
        d = DeviceArray(self.shape)
        if host is not None:
            d.upload(host)
        return d

    def sync(self):
        check(self.lib.asd2_synchronize(self.h))

    def stream_copy(self, dst: DeviceArray, src: DeviceArray):
        assert dst.nbytes == src.nbytes and dst.dtype == np.float64
        check(self.lib.asd2_stream_copy(self.h, dst.ptr, src.ptr, dst.nbytes // 8))

    def set_bc(self, grid_kind, kinds, consts=(0.0, 0.0, 0.0, 0.0)):
        check(self.lib.asd2_set_bc(self.h, grid_kind, _bc_codes(kinds), (C.c_double * 4)(*consts)))

    def bc_apply(self, grid_kind, d: DeviceArray):
        check(self.lib.asd2_bc_apply(self.h, grid_kind, d.ptr))

    # level set ---------------------------------------------------------------------
    def ls_advect(self, ls, u, v, dt):
        check(self.lib.asd2_ls_advect(self.h, ls.ptr, u.ptr, v.ptr, dt))

    def ls_advect_noalloc(self, ls, u, v, dt):
        """asd2_ls_advect_noalloc: the same advection with the stage arrays in the library's
        static per-GPU buffer (no device allocation)"""
        check(self.lib.asd2_ls_advect_noalloc(self.h, ls.ptr, u.ptr, v.ptr, dt))

    def ls_weno_rhs(self, ls, uc, vc, L):
        check(self.lib.asd2_ls_weno_rhs(self.h, ls.ptr, uc.ptr, vc.ptr, L.ptr))

    def ls_reinit_minrk2(self, ls, iters):
        check(self.lib.asd2_ls_reinit_minrk2(self.h, ls.ptr, iters))

    def ls_reinit_sussman(self, ls, iters):
        check(self.lib.asd2_ls_reinit_sussman(self.h, ls.ptr, iters))

    def ls_subcell_absgrad(self, ls, ls0, out):
        check(self.lib.asd2_ls_subcell_absgrad(self.h, ls.ptr, ls0.ptr, out.ptr))

    def ls_eno_absgrad(self, ls, ls0, out):
        check(self.lib.asd2_ls_eno_absgrad(self.h, ls.ptr, ls0.ptr, out.ptr))

    def ls_reinit_original(self, ls, iters):
        check(self.lib.asd2_ls_reinit_original(self.h, ls.ptr, iters))

    def ls_reinit_first_time(self, ls, iters):
        check(self.lib.asd2_ls_reinit_first_time(self.h, ls.ptr, iters))

    def ls_weno_absgrad(self, ls, ls0, out):
        check(self.lib.asd2_ls_weno_absgrad(self.h, ls.ptr, ls0.ptr, out.ptr))

    def ls_heaviside_deriv(self, ls, out):
        check(self.lib.asd2_ls_heaviside_deriv(self.h, ls.ptr, out.ptr))

    def ls_smoothed_sign(self, ls, out):
        check(self.lib.asd2_ls_smoothed_sign(self.h, ls.ptr, out.ptr))

    def ls_sussman_constraint(self, ls, ls0, hd, out):
        check(self.lib.asd2_ls_sussman_constraint(self.h, ls.ptr, ls0.ptr, hd.ptr if hd is not None else None,
                                                  out.ptr))

    def ls_unit_weno5(self, F, d, n, g):
        """One line of n + 2g values -> (FP, FM) on the same padded line."""
        dF = DeviceArray.from_host(np.asarray(F, np.float64))
        dP, dM = DeviceArray((n + 2 * g,)), DeviceArray((n + 2 * g,))
        check(self.lib.asd2_ls_unit_weno5(self.h, dF.ptr, dP.ptr, dM.ptr, d, n, g))
        return dP.download(), dM.download()

    def ls_diagnostics(self, ls, ref=None) -> dict:
        """Mass, centroid, interface length, circularity and (with ref) the error norms and mass
        loss of a level-set field (asd2_ls_diagnostics)."""
        d = LsDiag()
        check(self.lib.asd2_ls_diagnostics(self.h, ls.ptr, None if ref is None else ref.ptr, C.byref(d)))
        return d.as_dict()

    # MAC ---------------------------------------------------------------------------
    def kappa(self, ls, kap):
        check(self.lib.asd2_mac_kappa(self.h, ls.ptr, kap.ptr))

    def heaviside(self, ls, H, smooth=False):
        check(self.lib.asd2_mac_heaviside(self.h, ls.ptr, H.ptr, 1 if smooth else 0))

    def divergence(self, u, v, div):
        check(self.lib.asd2_mac_divergence(self.h, u.ptr, v.ptr, div.ptr))

    def fill(self, arr, value):
        check(self.lib.asd2_fill(self.h, arr.ptr, arr.nbytes // 8, C.c_double(value)))

    def pressure(self, ps, p, dt):
        """p = ps / dt on the interior (MACSolver2D::SolvePoisson's last loop)"""
        check(self.lib.asd2_mac_pressure(self.h, ps.ptr, p.ptr, C.c_double(dt)))

    def update_vel(self, fl, u, v, us, vs, ps, ls, kap, H):
        check(self.lib.asd2_mac_update_vel(self.h, C.byref(fl), u.ptr, v.ptr, us.ptr, vs.ptr, ps.ptr,
                                           ls.ptr, kap.ptr, H.ptr))

    def update_dt(self, fl, u, v) -> float:
        out = C.c_double(0.0)
        check(self.lib.asd2_mac_update_dt(self.h, C.byref(fl), u.ptr, v.ptr, C.byref(out)))
        return out.value

    def intermediate_vel(self, fl, order, ls, u, v, uhat, vhat, Hs):
        check(self.lib.asd2_mac_intermediate_vel(self.h, C.byref(fl), order, ls.ptr, u.ptr, v.ptr,
                                                 uhat.ptr, vhat.ptr, Hs.ptr))

    def momentum_term(self, fl, grid_kind, term, ls, u, v, Hs, out):
        check(self.lib.asd2_mac_momentum_term(self.h, C.byref(fl), grid_kind, term, ls.ptr, u.ptr, v.ptr,
                                              Hs.ptr, out.ptr))

    # axisymmetric (x = r from the grid's base x, y = z) ----------------------------
    def kappa_axisym(self, ls, kap):
        check(self.lib.asd2_mac_kappa_axisym(self.h, ls.ptr, kap.ptr))

    def divergence_axisym(self, u, v, div, for_poisson):
        check(self.lib.asd2_mac_divergence_axisym(self.h, u.ptr, v.ptr, div.ptr, 1 if for_poisson else 0))

    def update_dt_axisym(self, fl, u, v, kap) -> float:
        out = C.c_double(0.0)
        check(self.lib.asd2_mac_update_dt_axisym(self.h, C.byref(fl), u.ptr, v.ptr, kap.ptr, C.byref(out)))
        return out.value

    def gfm_assemble_axisym(self, fl, ls, kap, H, div4p):
        check(self.lib.asd2_gfm_assemble_axisym(self.h, C.byref(fl), ls.ptr, kap.ptr,
                                                H.ptr if H is not None else None, div4p.ptr))

    def rhs_u_axisym(self, fl, ls, u, v, H, rhs):
        check(self.lib.asd2_mac_rhs_u_axisym(self.h, C.byref(fl), ls.ptr, u.ptr, v.ptr, H.ptr, rhs.ptr))

    def rhs_v_axisym(self, fl, ls, u, uhat, v, H, rhs):
        check(self.lib.asd2_mac_rhs_v_axisym(self.h, C.byref(fl), ls.ptr, u.ptr, uhat.ptr, v.ptr, H.ptr, rhs.ptr))

    def momentum_term_axisym(self, fl, grid_kind, term, ls, u, uhat, v, H, out):
        check(self.lib.asd2_mac_momentum_term_axisym(self.h, C.byref(fl), grid_kind, term, ls.ptr, u.ptr,
                                                     uhat.ptr if uhat is not None else None, v.ptr, H.ptr,
                                                     out.ptr))

    def uhat_axisym(self, fl, ls, rhs, H, uhat):
        check(self.lib.asd2_mac_uhat_axisym(self.h, C.byref(fl), ls.ptr, rhs.ptr, H.ptr, uhat.ptr))

    def vhat_axisym(self, fl, ls, rhs, H, vhat):
        check(self.lib.asd2_mac_vhat_axisym(self.h, C.byref(fl), ls.ptr, rhs.ptr, H.ptr, vhat.ptr))

    def tdma(self, a, b, c, d):
        """count tridiagonal systems of n unknowns ([n][count] device arrays); the solution into d"""
        n, count = d.shape
        check(self.lib.asd2_tdma(self.h, n, count, a.ptr, b.ptr, c.ptr, d.ptr))

    # GFM / Poisson -----------------------------------------------------------------
    def gfm_assemble(self, fl, ls, kap, H, div):
        check(self.lib.asd2_gfm_assemble(self.h, C.byref(fl), ls.ptr, kap.ptr,
                                         H.ptr if H is not None else None, div.ptr))

    def gfm_export(self):
        outs = [self.array() for _ in range(4)]
        check(self.lib.asd2_gfm_export(self.h, *[o.ptr for o in outs]))
        return [o.download() for o in outs]

    def gfm_import(self, ax, ay, diag, rhs):
        ds = [self.array(a) for a in (ax, ay, diag, rhs)]
        check(self.lib.asd2_gfm_import(self.h, *[d.ptr for d in ds]))
        self.sync()

    def record_coarse(self, level_h, level_w, iters, d_out):
        """arms asd2_poisson_record_coarse for the next coarse-corrected solve"""
        it = (C.c_int * len(iters))(*iters)
        check(self.lib.asd2_poisson_record_coarse(self.h, level_h, level_w, len(iters), it, d_out.ptr))

    def set_preconditioner(self, net):
        check(self.lib.asd2_set_preconditioner(self.h, net.h if net is not None else None))

    def gfm_apply(self, x, y):
        check(self.lib.asd2_gfm_apply(self.h, x.ptr, y.ptr))

    def fast_poisson(self, rhs, ps, len_x, len_y, dx, dy, q=0.0):
        """the constant-coefficient fast solve of PoissonSolver2D::MKL_2FUniform_2D(ps, rhs, lenX,
        lenY, dx, dy, PBC) with the context's pressure sides: ps (device, padded) receives the
        solution at every node of the mesh of cell centres, ghosts included"""
        check(self.lib.asd2_fast_poisson(self.h, rhs.ptr, ps.ptr, len_x, len_y, dx, dy, q))

    @staticmethod
    def opts(tol=1e-8, max_iter=100000, mode=CG_STANDARD, precision=PREC_FP64,
             precond=PRECOND_JACOBI, reliable_delta=0.0, check_every=0) -> SolveOpts:
        return SolveOpts(tol, max_iter, mode, precision, precond, reliable_delta, check_every)

    def poisson_solve(self, ps, opts: SolveOpts, raise_on_fail=True) -> Stats:
        st = SolveStats()
        code = self.lib.asd2_poisson_solve(self.h, C.byref(opts), ps.ptr, C.byref(st))
        if raise_on_fail:
            check(code)
        return Stats(**st.as_dict())

    def poisson_bicgstab(self, ps, opts: SolveOpts, raise_on_fail=True) -> Stats:
        st = SolveStats()
        code = self.lib.asd2_poisson_bicgstab(self.h, C.byref(opts), ps.ptr, C.byref(st))
        if raise_on_fail:
            check(code)
        return Stats(**st.as_dict())

    def csr_pcg(self, rowptr, cols, vals, mdiag, b, opts: SolveOpts, bicgstab=False):
        n = len(b)
        fn = self.lib.asd2_csr_bicgstab if bicgstab else self.lib.asd2_csr_pcg
        d_rp = DeviceArray.from_host(np.asarray(rowptr, np.int32))
        d_c = DeviceArray.from_host(np.asarray(cols, np.int32))
        d_v = DeviceArray.from_host(np.asarray(vals, np.float64))
        d_m = DeviceArray.from_host(np.asarray(mdiag, np.float64))
        d_b = DeviceArray.from_host(np.asarray(b, np.float64))
        d_x = DeviceArray((n,))
        st = SolveStats()
        check(fn(self.h, n, d_rp.ptr, d_c.ptr, d_v.ptr, d_m.ptr, d_b.ptr, d_x.ptr, C.byref(opts), C.byref(st)))
        return d_x.download(), Stats(**st.as_dict())

    def csr_gs(self, rowptr, cols, vals, b, x0, max_iter):
        """PoissonSolver2D::GS_2FUniform_2D (src/poisson2d.cpp:7-88) on a zero-based CSR matrix
        from the start vector x0: returns (x, stats); stats.iterations = sweeps,
        stats.rel_residual = the last relative change."""
        n = len(b)
        d_rp = DeviceArray.from_host(np.asarray(rowptr, np.int32))
        d_c = DeviceArray.from_host(np.asarray(cols, np.int32))
        d_v = DeviceArray.from_host(np.asarray(vals, np.float64))
        d_b = DeviceArray.from_host(np.asarray(b, np.float64))
        d_x = DeviceArray.from_host(np.asarray(x0, np.float64))
        st = SolveStats()
        check(self.lib.asd2_csr_gs(self.h, n, d_rp.ptr, d_c.ptr, d_v.ptr, d_b.ptr, d_x.ptr, int(max_iter),
                                   C.byref(st)))
        return d_x.download(), Stats(**st.as_dict())

    # timing --------------------------------------------------------------------------
    def timer_start(self):
        check(self.lib.asd2_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_double(0.0)
        check(self.lib.asd2_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def profile_kernel(self, kid: int):
        check(self.lib.asd2_profile_kernel(self.h, kid))

    def profile_read(self):
        n = C.c_int(0)
        ms = C.c_double(0.0)
        check(self.lib.asd2_profile_read(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_read_each(self, capacity: int = 1 << 16) -> np.ndarray:
        """Per-launch durations (ms) of the profiled kernel family, in launch order."""
        buf = np.zeros(capacity, dtype=np.float32)
        n = C.c_int(0)
        check(self.lib.asd2_profile_read_each(self.h, buf.ctypes.data_as(C.POINTER(C.c_float)), capacity, C.byref(n)))
        return buf[:min(n.value, capacity)].copy()
