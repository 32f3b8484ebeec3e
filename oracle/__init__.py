"""ctypes binding of the FP64 oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this package; the HIP product path never does.

``Oracle`` wraps ``_build/liboracle.so`` (the C restatement in ``asd2_oracle.c``);
``Reference`` wraps ``_ref/libasd2ref.so`` (the reference's own level-set and BC
sources compiled in place, present only where the reference checkout was available
at build time).  Arrays are numpy float64 in the reference layout
(padded by g, j fastest: ``idx(i, j) = j + (ny + 2g) * i``), i.e. numpy shape
``(nx + 2g, ny + 2g)`` in C order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libasd2ref.so")
REF_SRC = "/root/reference"

BC = {"periodic": 0, "neumann": 1, "dirichlet": 2, "axisym": 3, "wall": 4,
      "inlet": 5, "outlet": 6, "pressure": 7, "lsfree": 8, "custom": 9}


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("g", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double),
                ("bc_p", C.c_int * 4), ("c_p", C.c_double * 4),
                ("bc_u", C.c_int * 4), ("c_u", C.c_double * 4),
                ("bc_v", C.c_int * 4), ("c_v", C.c_double * 4)]


class Fluid(C.Structure):
    _fields_ = [("rho_h", C.c_double), ("rho_l", C.c_double), ("mu_h", C.c_double),
                ("mu_l", C.c_double), ("sigma", C.c_double), ("dt", C.c_double),
                ("g", C.c_double), ("g_axis", C.c_int), ("gravity_fix", C.c_int),
                ("re", C.c_double), ("we", C.c_double), ("fr", C.c_double),
                ("cfl", C.c_double)]


def make_grid(nx, ny, dx, dy, g=3, bc_p=("neumann",) * 4, c_p=(0.0,) * 4,
              bc_u=("dirichlet",) * 4, c_u=(0.0,) * 4, bc_v=("dirichlet",) * 4,
              c_v=(0.0,) * 4) -> Grid:
    gr = Grid()
    gr.nx, gr.ny, gr.g, gr.dx, gr.dy = nx, ny, g, dx, dy
    for k in range(4):
        gr.bc_p[k] = BC[bc_p[k]] if isinstance(bc_p[k], str) else int(bc_p[k])
        gr.bc_u[k] = BC[bc_u[k]] if isinstance(bc_u[k], str) else int(bc_u[k])
        gr.bc_v[k] = BC[bc_v[k]] if isinstance(bc_v[k], str) else int(bc_v[k])
        gr.c_p[k], gr.c_u[k], gr.c_v[k] = c_p[k], c_u[k], c_v[k]
    return gr


def make_fluid(rho_h, rho_l, mu_h, mu_l, sigma, dt, g=0.0, g_axis=1, gravity_fix=1,
               L=1.0, U=1.0, cfl=0.5) -> Fluid:
    """Dimensionless groups follow the dimensional constructor, mac2d.cpp:28-49."""
    fl = Fluid()
    fl.rho_h, fl.rho_l, fl.mu_h, fl.mu_l = rho_h, rho_l, mu_h, mu_l
    fl.sigma, fl.dt, fl.g, fl.g_axis, fl.gravity_fix = sigma, dt, g, g_axis, gravity_fix
    fl.re = rho_h * L * U / mu_h if mu_h != 0 else float("inf")
    fl.we = rho_h * L * U * U / sigma if sigma != 0 else float("inf")
    fl.fr = U * U / (g * L) if g != 0 else float("inf")
    fl.cfl = cfl
    return fl


def shape(gr: Grid):
    return (gr.nx + 2 * gr.g, gr.ny + 2 * gr.g)


_P = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def build(ref: bool = True) -> None:
    """Compile the oracle (and, where the reference checkout exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REF_SRC}"], check=True)


class Oracle:
    """The C restatement.  Each method cites the reference function it restates."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            build(ref=False)
        self.lib = C.CDLL(path)
        L = self.lib
        G = C.POINTER(Grid)
        F = C.POINTER(Fluid)
        sig = {
            "or_bc_p": [G, _P], "or_bc_u": [G, _P], "or_bc_v": [G, _P],
            "or_unit_hjweno5": [_P, _P, _P, C.c_double, C.c_int, C.c_int],
            "or_hjweno5_ls": [G, _P, _P, _P, _P],
            "or_solve_levelset": [G, _P, _P, _P, C.c_double],
            "or_eno_absgrad": [G, _P, _P, _P], "or_subcell_absgrad": [G, _P, _P, _P],
            "or_smoothed_sign": [G, _P, _P], "or_heaviside_deriv": [G, _P, _P],
            "or_sussman_constraint": [G, _P, _P, _P, _P],
            "or_reinit_minrk2": [G, _P, C.c_int], "or_reinit_sussman": [G, _P, C.c_int],
            "or_reinit_original": [G, _P, C.c_int], "or_reinit_first_time": [G, _P, C.c_int],
            "or_weno_absgrad": [G, _P, _P, _P],
            "or_update_kappa": [G, _P, _P], "or_heaviside": [G, _P, _P],
            "or_smooth_heaviside": [G, _P, _P], "or_divergence": [G, _P, _P, _P],
            "or_update_vel": [G, F, _P, _P, _P, _P, _P, _P, _P, _P],
            "or_convection_u": [G, _P, _P, _P], "or_convection_v": [G, _P, _P, _P],
            "or_viscosity_u": [G, F, _P, _P, _P, _P, _P],
            "or_viscosity_v": [G, F, _P, _P, _P, _P, _P],
            "or_intermediate_vel": [G, F, C.c_int, _P, _P, _P, _P, _P, _P],
            "or_gfm_assemble_stencil": [G, F, _P, _P, _P, _P, _P, _P, _P, _P],
            "or_stencil_apply": [G, _P, _P, _P, _P, _P],
            "or_gfm_assemble_axisym": [G, F, C.c_double, _P, _P, _P, _P, _P, _P, _P, _P, _P],
            "or_divergence_axisym": [G, C.c_double, _P, _P, _P, C.c_int],
            "or_tdma": [_P, _P, _P, _P, C.c_int],
            "or_update_kappa_axisym": [G, C.c_double, _P, _P],
            "or_viscosity_u_axisym": [G, F, _P, _P, _P, _P],
            "or_viscosity_v_axisym": [G, F, _P, _P, _P, _P, _P],
            "or_rhs_u_axisym": [G, F, _P, _P, _P, _P, _P],
            "or_rhs_v_axisym": [G, F, _P, _P, _P, _P, _P, _P],
            "or_uhat_axisym": [G, F, _P, _P, _P, _P],
            "or_vhat_axisym": [G, F, _P, _P, _P, _P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None
        L.or_ls_diagnostics.argtypes = [G, _P, C.c_void_p, C.c_double, C.c_double, _P]
        L.or_ls_diagnostics.restype = None
        L.or_reinit_iterations.argtypes = [C.c_double, C.c_double]
        L.or_reinit_iterations.restype = C.c_int
        L.or_update_dt.argtypes = [G, F, _P, _P]
        L.or_update_dt.restype = C.c_double
        L.or_update_dt_axisym.argtypes = [G, F, _P, _P, _P]
        L.or_update_dt_axisym.restype = C.c_double
        IP = C.POINTER(C.c_int)
        L.or_pcg_stencil.argtypes = [G, _P, _P, _P, _P, _P, C.c_double, C.c_int, C.c_int, IP, _P]
        L.or_pcg_stencil.restype = C.c_int
        L.or_bicgstab_stencil.argtypes = [G, _P, _P, _P, _P, _P, C.c_double, C.c_int, IP, _P]
        L.or_bicgstab_stencil.restype = C.c_int
        L.or_pcg_csr.argtypes = [C.c_int, IP, IP, _P, _P, _P, _P, C.c_double, C.c_int, C.c_int, IP, _P]
        L.or_pcg_csr.restype = C.c_int
        L.or_gs_csr.argtypes = [C.c_int, IP, IP, _P, _P, _P, C.c_int, IP, _P]
        L.or_gs_csr.restype = C.c_int

    # --- boundary conditions (bc2d.cpp) ---------------------------------------------
    def bc_p(self, gr, a):
        self.lib.or_bc_p(C.byref(gr), _ptr(a)); return a

    def bc_u(self, gr, a):
        self.lib.or_bc_u(C.byref(gr), _ptr(a)); return a

    def bc_v(self, gr, a):
        self.lib.or_bc_v(C.byref(gr), _ptr(a)); return a

    # --- level set (levelset2d.cpp) ---------------------------------------------------
    def unit_hjweno5(self, F, d, n, g=3):
        FP = np.zeros_like(F); FM = np.zeros_like(F)
        self.lib.or_unit_hjweno5(_ptr(F), _ptr(FP), _ptr(FM), d, n, g)
        return FP, FM

    def hjweno5_ls(self, gr, ls, uc, vc):
        L = np.zeros_like(ls)
        self.lib.or_hjweno5_ls(C.byref(gr), _ptr(ls), _ptr(uc), _ptr(vc), _ptr(L))
        return L

    def solve_levelset(self, gr, ls, u, v, dt):
        self.lib.or_solve_levelset(C.byref(gr), _ptr(ls), _ptr(u), _ptr(v), dt); return ls

    def eno_absgrad(self, gr, ls, ls0):
        o = np.zeros_like(ls); self.lib.or_eno_absgrad(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(o)); return o

    def subcell_absgrad(self, gr, ls, ls0):
        o = np.zeros_like(ls); self.lib.or_subcell_absgrad(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(o)); return o

    def smoothed_sign(self, gr, ls):
        o = np.zeros_like(ls); self.lib.or_smoothed_sign(C.byref(gr), _ptr(ls), _ptr(o)); return o

    def heaviside_deriv(self, gr, ls):
        o = np.zeros_like(ls); self.lib.or_heaviside_deriv(C.byref(gr), _ptr(ls), _ptr(o)); return o

    def sussman_constraint(self, gr, ls, ls0, hd):
        o = np.zeros_like(ls)
        self.lib.or_sussman_constraint(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(hd), _ptr(o))
        return o

    def reinit_minrk2(self, gr, ls, iters):
        self.lib.or_reinit_minrk2(C.byref(gr), _ptr(ls), iters); return ls

    def reinit_sussman(self, gr, ls, iters):
        self.lib.or_reinit_sussman(C.byref(gr), _ptr(ls), iters); return ls

    def reinit_original(self, gr, ls, iters):
        self.lib.or_reinit_original(C.byref(gr), _ptr(ls), iters); return ls

    def reinit_first_time(self, gr, ls, iters):
        self.lib.or_reinit_first_time(C.byref(gr), _ptr(ls), iters); return ls

    def weno_absgrad(self, gr, ls, ls0):
        o = np.zeros_like(ls); self.lib.or_weno_absgrad(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(o)); return o

    def reinit_iterations(self, eps, max_time):
        return self.lib.or_reinit_iterations(eps, max_time)

    # --- MAC helpers (mac2d.cpp) -------------------------------------------------------
    def update_kappa(self, gr, ls, kappa=None):
        k = np.zeros_like(ls) if kappa is None else kappa
        self.lib.or_update_kappa(C.byref(gr), _ptr(ls), _ptr(k)); return k

    DIAG_FIELDS = ("cells", "mass", "centroid_x", "centroid_y", "length", "circularity",
                   "err_linf", "err_l1", "err_l2", "ref_cells", "mass_loss_pct")

    def ls_diagnostics(self, gr, ls, ref=None, base=(0.0, 0.0)):
        """Mass and error norms of tests/LargeTests/test_levelset2d_large.cpp:94-127, plus centroid,
        interface length and circularity (no reference counterpart); a dict."""
        out = np.zeros(11)
        rp = None if ref is None else ref.ctypes.data_as(C.c_void_p)
        self.lib.or_ls_diagnostics(C.byref(gr), _ptr(ls), rp, base[0], base[1], _ptr(out))
        return dict(zip(self.DIAG_FIELDS, out.tolist()))

    def heaviside(self, gr, ls):
        H = np.zeros_like(ls); self.lib.or_heaviside(C.byref(gr), _ptr(ls), _ptr(H)); return H

    def smooth_heaviside(self, gr, ls):
        H = np.zeros_like(ls); self.lib.or_smooth_heaviside(C.byref(gr), _ptr(ls), _ptr(H)); return H

    def divergence(self, gr, u, v):
        d = np.zeros_like(u); self.lib.or_divergence(C.byref(gr), _ptr(u), _ptr(v), _ptr(d)); return d

    def update_vel(self, gr, fl, u, v, us, vs, ps, ls, kappa, H):
        self.lib.or_update_vel(C.byref(gr), C.byref(fl), _ptr(u), _ptr(v), _ptr(us), _ptr(vs),
                               _ptr(ps), _ptr(ls), _ptr(kappa), _ptr(H))
        return u, v

    def update_dt(self, gr, fl, u, v):
        return self.lib.or_update_dt(C.byref(gr), C.byref(fl), _ptr(u), _ptr(v))

    def convection_u(self, gr, u, v):
        o = np.zeros_like(u); self.lib.or_convection_u(C.byref(gr), _ptr(u), _ptr(v), _ptr(o)); return o

    def convection_v(self, gr, u, v):
        o = np.zeros_like(u); self.lib.or_convection_v(C.byref(gr), _ptr(u), _ptr(v), _ptr(o)); return o

    def viscosity_u(self, gr, fl, u, v, ls, H):
        o = np.zeros_like(u)
        self.lib.or_viscosity_u(C.byref(gr), C.byref(fl), _ptr(u), _ptr(v), _ptr(ls), _ptr(H), _ptr(o))
        return o

    def viscosity_v(self, gr, fl, u, v, ls, H):
        o = np.zeros_like(u)
        self.lib.or_viscosity_v(C.byref(gr), C.byref(fl), _ptr(u), _ptr(v), _ptr(ls), _ptr(H), _ptr(o))
        return o

    def intermediate_vel(self, gr, fl, order, ls, u, v, Hs):
        uh = np.zeros_like(u); vh = np.zeros_like(v)
        self.lib.or_intermediate_vel(C.byref(gr), C.byref(fl), order, _ptr(ls), _ptr(u), _ptr(v),
                                     _ptr(uh), _ptr(vh), _ptr(Hs))
        return uh, vh

    # --- GFM + solvers -----------------------------------------------------------------
    def gfm_assemble(self, gr, fl, ls, kappa, H, div):
        ax = np.zeros_like(ls); ay = np.zeros_like(ls); dg = np.zeros_like(ls); rhs = np.zeros_like(ls)
        self.lib.or_gfm_assemble_stencil(C.byref(gr), C.byref(fl), _ptr(ls), _ptr(kappa), _ptr(H),
                                         _ptr(div), _ptr(ax), _ptr(ay), _ptr(dg), _ptr(rhs))
        return ax, ay, dg, rhs

    def gfm_assemble_axisym(self, gr, fl, base_r, ls, kappa, H, div):
        """(ax, ay, ae, diag, rhs) of MACSolver2DAxisym::SolvePoisson's system"""
        out = [np.zeros_like(ls) for _ in range(5)]
        self.lib.or_gfm_assemble_axisym(C.byref(gr), C.byref(fl), base_r, _ptr(ls), _ptr(kappa), _ptr(H),
                                        _ptr(div), *[_ptr(a) for a in out])
        return tuple(out)

    def divergence_axisym(self, gr, base_r, u, v, for_poisson):
        d = np.zeros_like(u)
        self.lib.or_divergence_axisym(C.byref(gr), base_r, _ptr(u), _ptr(v), _ptr(d), int(for_poisson))
        return d

    def update_kappa_axisym(self, gr, base_r, ls):
        k = np.zeros_like(ls)
        self.lib.or_update_kappa_axisym(C.byref(gr), base_r, _ptr(ls), _ptr(k))
        return k

    def update_dt_axisym(self, gr, fl, u, v, kappa):
        return self.lib.or_update_dt_axisym(C.byref(gr), C.byref(fl), _ptr(u), _ptr(v), _ptr(kappa))

    def viscosity_u_axisym(self, gr, fl, v, ls, H):
        o = np.zeros_like(v)
        self.lib.or_viscosity_u_axisym(C.byref(gr), C.byref(fl), _ptr(v), _ptr(ls), _ptr(H), _ptr(o))
        return o

    def viscosity_v_axisym(self, gr, fl, u, uhat, ls, H):
        o = np.zeros_like(u)
        self.lib.or_viscosity_v_axisym(C.byref(gr), C.byref(fl), _ptr(u), _ptr(uhat), _ptr(ls), _ptr(H), _ptr(o))
        return o

    def rhs_u_axisym(self, gr, fl, ls, u, v, H):
        o = np.zeros_like(u)
        self.lib.or_rhs_u_axisym(C.byref(gr), C.byref(fl), _ptr(ls), _ptr(u), _ptr(v), _ptr(H), _ptr(o))
        return o

    def rhs_v_axisym(self, gr, fl, ls, u, uhat, v, H):
        o = np.zeros_like(v)
        self.lib.or_rhs_v_axisym(C.byref(gr), C.byref(fl), _ptr(ls), _ptr(u), _ptr(uhat), _ptr(v), _ptr(H),
                                 _ptr(o))
        return o

    def uhat_axisym(self, gr, fl, ls, rhs, H):
        o = np.zeros_like(rhs)
        self.lib.or_uhat_axisym(C.byref(gr), C.byref(fl), _ptr(ls), _ptr(rhs), _ptr(H), _ptr(o))
        return o

    def vhat_axisym(self, gr, fl, ls, rhs, H):
        o = np.zeros_like(rhs)
        self.lib.or_vhat_axisym(C.byref(gr), C.byref(fl), _ptr(ls), _ptr(rhs), _ptr(H), _ptr(o))
        return o

    def tdma(self, a, b, c, d):
        """x of the tridiagonal system (copies in, the reference's recurrence)"""
        a, b, c, d = (np.array(v, dtype=np.float64) for v in (a, b, c, d))
        self.lib.or_tdma(_ptr(a), _ptr(b), _ptr(c), _ptr(d), len(d))
        return d

    def stencil_apply(self, gr, ax, ay, dg, p):
        q = np.zeros_like(p)
        self.lib.or_stencil_apply(C.byref(gr), _ptr(ax), _ptr(ay), _ptr(dg), _ptr(p), _ptr(q))
        return q

    def pcg(self, gr, ax, ay, dg, b, tol, maxit, mode=1):
        x = np.zeros_like(b); it = C.c_int(0); rr = C.c_double(0.0)
        st = self.lib.or_pcg_stencil(C.byref(gr), _ptr(ax), _ptr(ay), _ptr(dg), _ptr(b), _ptr(x),
                                     tol, maxit, mode, C.byref(it), C.byref(rr))
        return x, it.value, rr.value, st

    def bicgstab(self, gr, ax, ay, dg, b, tol, maxit):
        x = np.zeros_like(b); it = C.c_int(0); rr = C.c_double(0.0)
        st = self.lib.or_bicgstab_stencil(C.byref(gr), _ptr(ax), _ptr(ay), _ptr(dg), _ptr(b), _ptr(x),
                                          tol, maxit, C.byref(it), C.byref(rr))
        return x, it.value, rr.value, st

    def pcg_csr(self, rowptr, cols, vals, mdiag, b, tol, maxit, mode=1):
        IP = C.POINTER(C.c_int)
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        cl = np.ascontiguousarray(cols, dtype=np.int32)
        x = np.zeros_like(b); it = C.c_int(0); rr = C.c_double(0.0)
        st = self.lib.or_pcg_csr(len(b), rp.ctypes.data_as(IP), cl.ctypes.data_as(IP), _ptr(vals),
                                 _ptr(mdiag), _ptr(b), _ptr(x), tol, maxit, mode, C.byref(it), C.byref(rr))
        return x, it.value, rr.value, st

    def gs_csr(self, rowptr, cols, vals, b, x0, maxit):
        """GS_2FUniform_2D (src/poisson2d.cpp:7-88) on a zero-based CSR matrix from x0:
        returns (x, sweeps, last relative change)."""
        IP = C.POINTER(C.c_int)
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        cl = np.ascontiguousarray(cols, dtype=np.int32)
        x = np.array(x0, dtype=np.float64, copy=True)
        it = C.c_int(0); err = C.c_double(0.0)
        st = self.lib.or_gs_csr(len(b), rp.ctypes.data_as(IP), cl.ctypes.data_as(IP), _ptr(vals), _ptr(b),
                                _ptr(x), maxit, C.byref(it), C.byref(err))
        if st != 0:
            raise MemoryError("or_gs_csr")
        return x, it.value, err.value


class Reference:
    """The reference's own src/levelset2d.cpp + src/bc2d.cpp (oracle/_ref)."""

    def __init__(self, path: str = REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        G = C.POINTER(Grid)
        sig = {
            "ref_bc_p": [G, _P], "ref_bc_u": [G, _P], "ref_bc_v": [G, _P],
            "ref_unit_hjweno5": [_P, _P, _P, C.c_double, C.c_int, C.c_int],
            "ref_hjweno5_ls": [G, _P, _P, _P, _P],
            "ref_solve_levelset": [G, _P, _P, _P, C.c_double],
            "ref_eno_absgrad": [G, _P, _P, _P], "ref_subcell_absgrad": [G, _P, _P, _P],
            "ref_smoothed_sign": [G, _P, _P], "ref_heaviside_deriv": [G, _P, _P],
            "ref_sussman_constraint": [G, _P, _P, _P, _P],
            "ref_reinit_minrk2": [G, _P, C.c_double], "ref_reinit_sussman": [G, _P, C.c_double],
            "ref_reinit_original": [G, _P, C.c_double], "ref_reinit_first_time": [G, _P],
            "ref_weno_absgrad": [G, _P, _P, _P],
        }
        for name, args in sig.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = C.c_int

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_PATH)

    def bc_p(self, gr, a):
        self.lib.ref_bc_p(C.byref(gr), _ptr(a)); return a

    def bc_u(self, gr, a):
        self.lib.ref_bc_u(C.byref(gr), _ptr(a)); return a

    def bc_v(self, gr, a):
        self.lib.ref_bc_v(C.byref(gr), _ptr(a)); return a

    def unit_hjweno5(self, F, d, n, g=3):
        FP = np.zeros_like(F); FM = np.zeros_like(F)
        self.lib.ref_unit_hjweno5(_ptr(F), _ptr(FP), _ptr(FM), d, n, g)
        return FP, FM

    def hjweno5_ls(self, gr, ls, uc, vc):
        L = np.zeros_like(ls)
        self.lib.ref_hjweno5_ls(C.byref(gr), _ptr(ls), _ptr(uc), _ptr(vc), _ptr(L))
        return L

    def solve_levelset(self, gr, ls, u, v, dt):
        self.lib.ref_solve_levelset(C.byref(gr), _ptr(ls), _ptr(u), _ptr(v), dt); return ls

    def eno_absgrad(self, gr, ls, ls0):
        o = np.zeros_like(ls); self.lib.ref_eno_absgrad(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(o)); return o

    def subcell_absgrad(self, gr, ls, ls0):
        o = np.zeros_like(ls); self.lib.ref_subcell_absgrad(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(o)); return o

    def smoothed_sign(self, gr, ls):
        o = np.zeros_like(ls); self.lib.ref_smoothed_sign(C.byref(gr), _ptr(ls), _ptr(o)); return o

    def heaviside_deriv(self, gr, ls):
        o = np.zeros_like(ls); self.lib.ref_heaviside_deriv(C.byref(gr), _ptr(ls), _ptr(o)); return o

    def sussman_constraint(self, gr, ls, ls0, hd):
        o = np.zeros_like(ls)
        self.lib.ref_sussman_constraint(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(hd), _ptr(o))
        return o

    def reinit_minrk2(self, gr, ls, max_time):
        self.lib.ref_reinit_minrk2(C.byref(gr), _ptr(ls), max_time); return ls

    def reinit_sussman(self, gr, ls, max_time):
        self.lib.ref_reinit_sussman(C.byref(gr), _ptr(ls), max_time); return ls

    def reinit_original(self, gr, ls, max_time):
        self.lib.ref_reinit_original(C.byref(gr), _ptr(ls), max_time); return ls

    def reinit_first_time(self, gr, ls):
        self.lib.ref_reinit_first_time(C.byref(gr), _ptr(ls)); return ls

    def weno_absgrad(self, gr, ls, ls0):
        o = np.zeros_like(ls); self.lib.ref_weno_absgrad(C.byref(gr), _ptr(ls), _ptr(ls0), _ptr(o)); return o
