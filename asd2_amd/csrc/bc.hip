// bc.hip -- ghost-cell fill of the staggered grids (reference src/bc2d.cpp).
#include <algorithm>

#include "internal.h"

namespace asd2 {

struct BcArgs {
    int kind[4];
    double c[4];
};

// ------------------------------------------------------------------ P grid
// One thread per grid line: a j-line (interior j) gets its W ghosts, then its E ghosts; an
// i-line (interior i) its S ghosts, then its N ghosts, each in the reference's sequential order
// (needed by the LS-free extrapolation, bc2d.cpp:896-934), so the fill follows the order W, E,
// S, N of ApplyBC_P_2D (bc2d.cpp:346-353) on every line even where a side could reach the
// opposite side's ghosts (grids below g cells are refused at context creation); the lines of
// different threads never overlap (W/E lines cover interior j, S/N lines interior i, corner
// ghosts are never written, bc2d.cpp:810-832).  Periodic sides copy the opposite interior rows
// (bc2d.cpp:786-808); the reference's east side writes rows nx .. nx + g - 1, inside the domain,
// from rows g .. 2g - 1, and leaves the east ghosts as they were: reproduced as written.  Those
// interior rows are what the S/N lines read, so a grid with a periodic side fills its W/E lines
// and its S/N lines in two launches.
__device__ __forceinline__ void bc_p_side(const Geom& G, const BcArgs& A, double* __restrict__ a, int side, int line) {
    const int nx = G.nx, ny = G.ny, g = G.g;
    const int k = A.kind[side];
    const double c = A.c[side];
    const bool neu = (k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_WALL);
    const bool dir = (k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET || k == ASD2_BC_OUTLET ||
                      k == ASD2_BC_PRESSURE);
    const bool lsf = (k == ASD2_BC_LSFREE);
    const bool per = (k == ASD2_BC_PERIODIC);
    if (side == 0) {
        const int j = g + line;
        if (per) {
            for (int i = 0; i < g; i++) a[G.idx(i, j)] = a[G.idx(nx + i, j)];
        } else if (lsf) {
            for (int i = g - 1; i >= 0; i--)
                a[G.idx(i, j)] = a[G.idx(i + 1, j)] + G.dx * (a[G.idx(i + 1, j)] - a[G.idx(i + 2, j)]) / G.dx;
        } else if (neu || dir) {
            for (int i = 0; i < g; i++) {
                const double m = a[G.idx(2 * g - i - 1, j)];
                a[G.idx(i, j)] = neu ? m : (-m + 2.0 * c);
            }
        }
    } else if (side == 1) {
        const int j = g + line;
        if (per) {
            for (int i = 0; i < g; i++) a[G.idx(nx + i, j)] = a[G.idx(i + g, j)];
        } else if (lsf) {
            for (int i = 0; i < g; i++)
                a[G.idx(nx + g + i, j)] = a[G.idx(nx + g + i - 1, j)]
                    + G.dx * (a[G.idx(nx + g + i - 1, j)] - a[G.idx(nx + g + i - 2, j)]) / G.dx;
        } else if (neu || dir) {
            for (int i = 0; i < g; i++) {
                const double m = a[G.idx(nx + i, j)];
                a[G.idx(nx + 2 * g - i - 1, j)] = neu ? m : (-m + 2.0 * c);
            }
        }
    } else if (side == 2) {
        const int i = g + line;
        if (per) {
            for (int j = 0; j < g; j++) a[G.idx(i, j)] = a[G.idx(i, ny + j)];
        } else if (lsf) {
            for (int j = g - 1; j >= 0; j--)
                a[G.idx(i, j)] = a[G.idx(i, j + 1)] + G.dy * (a[G.idx(i, j + 1)] - a[G.idx(i, j + 2)]) / G.dy;
        } else if (neu || dir) {
            for (int j = 0; j < g; j++) {
                const double m = a[G.idx(i, 2 * g - j - 1)];
                a[G.idx(i, j)] = neu ? m : (-m + 2.0 * c);
            }
        }
    } else {
        const int i = g + line;
        if (per) {
            for (int j = 0; j < g; j++) a[G.idx(i, ny + j + g)] = a[G.idx(i, j + g)];
        } else if (lsf) {
            for (int j = 0; j < g; j++)
                a[G.idx(i, ny + g + j)] = a[G.idx(i, ny + g + j - 1)]
                    + G.dy * (a[G.idx(i, ny + g + j - 1)] - a[G.idx(i, ny + g + j - 2)]) / G.dy;
        } else if (neu || dir) {
            for (int j = 0; j < g; j++) {
                const double m = a[G.idx(i, ny + j)];
                a[G.idx(i, ny + 2 * g - j - 1)] = neu ? m : (-m + 2.0 * c);
            }
        }
    }
}

// lines first .. last - 1 of the W/E (t < ny) and S/N (t >= ny) line numbering
__global__ __launch_bounds__(256) void k_bc_p(Geom G, BcArgs A, double* __restrict__ a, int first, int last) {
    const int t = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= last) return;
    if (t < G.ny) {
        bc_p_side(G, A, a, 0, t);
        bc_p_side(G, A, a, 1, t);
    } else if (t < G.ny + G.nx) {
        bc_p_side(G, A, a, 2, t - G.ny);
        bc_p_side(G, A, a, 3, t - G.ny);
    }
}

int bc_apply_p(asd2_ctx* ctx, double* d) {
    BcArgs A;
    bool periodic = false;
    for (int s = 0; s < 4; s++) {
        A.kind[s] = ctx->bc[ASD2_GRID_P][s];
        A.c[s] = ctx->bcc[ASD2_GRID_P][s];
        periodic = periodic || A.kind[s] == ASD2_BC_PERIODIC;
    }
    const int nx = ctx->geo.nx, ny = ctx->geo.ny;
    if (periodic) {  // W/E lines, then S/N lines (see bc_p_side)
        hipLaunchKernelGGL(k_bc_p, dim3(cdiv(ny, 256)), dim3(256), 0, ctx->stream, ctx->geo, A, d, 0, ny);
        hipLaunchKernelGGL(k_bc_p, dim3(cdiv(nx, 256)), dim3(256), 0, ctx->stream, ctx->geo, A, d, ny, ny + nx);
    } else {
        hipLaunchKernelGGL(k_bc_p, dim3(cdiv(nx + ny, 256)), dim3(256), 0, ctx->stream, ctx->geo, A, d, 0, nx + ny);
    }
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

// ------------------------------------------------------------------ U / V grids
// One workgroup walks the sides W, E, S, N in the reference order with a barrier
// between sides: the U/V rules overlap at corners (bc2d.cpp:599-627,705-733), so the
// side order is part of the result.  Within a side all writes are independent.
// grid-stride loop over the cells of one ghost-fill statement (one block: block-stride)
#define PFOR(count) \
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < (count); q += gridDim.x * blockDim.x)

__device__ void u_side(const Geom& G, int side, int k, double c, double* a) {
    const int nx = G.nx, ny = G.ny, g = G.g;
    const int H = ny + 2 * g;
    if (side == 0) {
        if (k == ASD2_BC_PERIODIC) {
            PFOR((g + 1) * ny) { const int i = q / ny, j = g + q % ny; a[G.idx(i, j)] = a[G.idx(nx - 1 + i, j)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_OUTLET) {
            PFOR(H) a[G.idx(g, q)] = a[G.idx(g + 1, q)];
            PFOR(g * (ny + 1)) { const int i = q / (ny + 1), j = g + q % (ny + 1); a[G.idx(i, j)] = a[G.idx(2 * g - i, j)]; }
        } else if (k == ASD2_BC_AXISYM || k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(H) a[G.idx(g, q)] = c;
            PFOR(g * ny) { const int i = q / ny, j = g + q % ny; a[G.idx(i, j)] = -a[G.idx(2 * g - i, j)] + 2.0 * c; }
        }
    } else if (side == 1) {
        if (k == ASD2_BC_PERIODIC) {
            PFOR(g * ny) { const int i = q / ny, j = g + q % ny; a[G.idx(nx + i + g, j)] = a[G.idx(i + g, j)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_OUTLET) {
            PFOR(H) a[G.idx(nx + g, q)] = a[G.idx(nx + g - 1, q)];
            PFOR((g - 1) * (ny + 1)) {
                const int i = 1 + q / (ny + 1), j = g + q % (ny + 1);
                a[G.idx(nx + 2 * g - i, j)] = a[G.idx(nx + i, j)];
            }
        } else if (k == ASD2_BC_AXISYM || k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(H) a[G.idx(nx + g, q)] = c;
            PFOR((g - 1) * ny) {
                const int i = 1 + q / ny, j = g + q % ny;
                a[G.idx(nx + 2 * g - i, j)] = -a[G.idx(nx + i, j)] + 2.0 * c;
            }
        }
    } else if (side == 2) {
        if (k == ASD2_BC_PERIODIC) {
            PFOR((nx - 1) * g) { const int i = g + 1 + q / g, j = q % g; a[G.idx(i, j)] = a[G.idx(i, ny + j)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_OUTLET) {
            PFOR((nx - 1) * g) { const int i = g + 1 + q / g, j = q % g; a[G.idx(i, j)] = a[G.idx(i, 2 * g - j - 1)]; }
        } else if (k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(nx * g) { const int i = g + q / g, j = q % g; a[G.idx(i, j)] = -a[G.idx(i, 2 * g - j - 1)] + 2.0 * c; }
        }
    } else {
        if (k == ASD2_BC_PERIODIC) {
            PFOR((nx - 1) * g) { const int i = g + 1 + q / g, j = q % g; a[G.idx(i, ny + j + g)] = a[G.idx(i, j + g)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_OUTLET) {
            PFOR((nx - 1) * g) {
                const int i = g + 1 + q / g, j = q % g;
                a[G.idx(i, ny + 2 * g - j - 1)] = a[G.idx(i, ny + j)];
            }
        } else if (k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(nx * g) {
                const int i = g + q / g, j = q % g;
                a[G.idx(i, ny + 2 * g - j - 1)] = -a[G.idx(i, ny + j)] + 2.0 * c;
            }
        }
    }
}

__device__ void v_side(const Geom& G, int side, int k, double c, double* a) {
    const int nx = G.nx, ny = G.ny, g = G.g;
    const int W = nx + 2 * g;
    if (side == 0) {
        if (k == ASD2_BC_PERIODIC) {
            PFOR(g * (ny - 1)) { const int i = q / (ny - 1), j = g + 1 + q % (ny - 1); a[G.idx(i, j)] = a[G.idx(nx + i, j)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_OUTLET) {
            PFOR(g * (ny - 1)) { const int i = q / (ny - 1), j = g + 1 + q % (ny - 1); a[G.idx(i, j)] = a[G.idx(2 * g - i - 1, j)]; }
        } else if (k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(g * (ny - 1)) {
                const int i = q / (ny - 1), j = g + 1 + q % (ny - 1);
                a[G.idx(i, j)] = -a[G.idx(2 * g - i - 1, j)] + 2.0 * c;
            }
        }
    } else if (side == 1) {
        if (k == ASD2_BC_PERIODIC) {
            PFOR(g * (ny - 1)) { const int i = q / (ny - 1), j = g + 1 + q % (ny - 1); a[G.idx(nx + i + g, j)] = a[G.idx(i + g, j)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_OUTLET) {
            PFOR(g * (ny - 1)) {
                const int i = q / (ny - 1), j = g + 1 + q % (ny - 1);
                a[G.idx(nx + 2 * g - i - 1, j)] = a[G.idx(nx + i, j)];
            }
        } else if (k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(g * (ny - 1)) {
                const int i = q / (ny - 1), j = g + 1 + q % (ny - 1);
                a[G.idx(nx + 2 * g - i - 1, j)] = -a[G.idx(nx + i, j)] + 2.0 * c;
            }
        }
    } else if (side == 2) {
        if (k == ASD2_BC_PERIODIC) {
            PFOR(nx * (g + 1)) { const int i = g + q / (g + 1), j = q % (g + 1); a[G.idx(i, j)] = a[G.idx(i, ny - 1 + j)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_OUTLET) {
            PFOR(W) a[G.idx(q, g)] = a[G.idx(q, g + 1)];
            PFOR(nx * g) { const int i = g + q / g, j = q % g; a[G.idx(i, j)] = a[G.idx(i, 2 * g - j)]; }
        } else if (k == ASD2_BC_AXISYM || k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(W) a[G.idx(q, g)] = c;
            PFOR(nx * g) { const int i = g + q / g, j = q % g; a[G.idx(i, j)] = -a[G.idx(i, 2 * g - j)] + 2.0 * c; }
        }
    } else {
        if (k == ASD2_BC_PERIODIC) {
            PFOR(nx * g) { const int i = g + q / g, j = q % g; a[G.idx(i, ny + j + g)] = a[G.idx(i, j + g)]; }
        } else if (k == ASD2_BC_NEUMANN || k == ASD2_BC_OUTLET) {
            PFOR(W) a[G.idx(q, ny + g)] = a[G.idx(q, ny + g - 1)];
            PFOR(nx * (g - 1)) {
                const int i = g + q / (g - 1), j = 1 + q % (g - 1);
                a[G.idx(i, ny + 2 * g - j)] = a[G.idx(i, ny + j)];
            }
        } else if (k == ASD2_BC_AXISYM || k == ASD2_BC_WALL || k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET) {
            PFOR(W) a[G.idx(q, ny + g)] = c;
            PFOR(nx * (g - 1)) {
                const int i = g + q / (g - 1), j = 1 + q % (g - 1);
                a[G.idx(i, ny + 2 * g - j)] = -a[G.idx(i, ny + j)] + 2.0 * c;
            }
        }
    }
}
#undef PFOR

// Velocity ghost cells in the reference's side order W, E, S, N, on many blocks.  The W and E
// statements write disjoint cells and, unless a side is periodic, read only rows they do not
// write, so they share one launch; likewise S and N, which read what W and E wrote.  A periodic
// side reads the row (column) the opposite side writes first, so a periodic pair runs one side
// per launch.  Inside a side no statement reads a cell another one writes.
__global__ __launch_bounds__(256) void k_bc_uv(Geom G, BcArgs A, int is_v, int side0, int side1,
                                               double* __restrict__ a) {
    for (int side = side0; side < side1; side++) {
        if (is_v) v_side(G, side, A.kind[side], A.c[side], a);
        else u_side(G, side, A.kind[side], A.c[side], a);
    }
}

int bc_apply_uv(asd2_ctx* ctx, int grid_kind, double* d) {
    BcArgs A;
    for (int s = 0; s < 4; s++) { A.kind[s] = ctx->bc[grid_kind][s]; A.c[s] = ctx->bcc[grid_kind][s]; }
    const Geom& G = ctx->geo;
    const int cells = (std::max(G.nx, G.ny) + 2 * G.g) * (G.g + 1);  // largest statement
    const int nb = std::max(1, std::min(128, (cells + 255) / 256));
    const int is_v = grid_kind == ASD2_GRID_V ? 1 : 0;
    const bool tiny = G.nx < 2 * G.g + 2 || G.ny < 2 * G.g + 2;  // mirrored rows could reach the far side
    for (int pair = 0; pair < 4; pair += 2) {
        if (tiny || A.kind[pair] == ASD2_BC_PERIODIC || A.kind[pair + 1] == ASD2_BC_PERIODIC) {
            for (int s = pair; s < pair + 2; s++)
                hipLaunchKernelGGL(k_bc_uv, dim3(nb), dim3(256), 0, ctx->stream, G, A, is_v, s, s + 1, d);
        } else {
            hipLaunchKernelGGL(k_bc_uv, dim3(nb), dim3(256), 0, ctx->stream, G, A, is_v, pair, pair + 2, d);
        }
    }
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

}  // namespace asd2

using namespace asd2;

extern "C" int asd2_bc_apply(asd2_ctx* ctx, int grid_kind, double* d_arr) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_arr) { set_error("null array"); return ASD2_ERR_ARG; }
    if (grid_kind == ASD2_GRID_P) return bc_apply_p(ctx, d_arr);
    if (grid_kind == ASD2_GRID_U || grid_kind == ASD2_GRID_V) return bc_apply_uv(ctx, grid_kind, d_arr);
    set_error("bad grid kind %d", grid_kind);
    return ASD2_ERR_ARG;
}
