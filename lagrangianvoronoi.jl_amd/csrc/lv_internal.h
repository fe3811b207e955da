// lv_internal.h -- shared types and device helpers of liblvhip.so (gfx950).
//
// Every translation unit is compiled with -ffp-contract=off: Voronoi vertices are produced by a
// sequence of cuts whose rounding must equal the reference's plain IEEE FP64 evaluation order
// (no FMA, correctly rounded division and sqrt), or neighbour topology stops being bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lvhip.h"

#define LV_MAXE LV_MAX_EDGES   // edge slots per polygon (ELL layout and clipper working set)
#define LV_MAXPATH 512         // magic-path nodes kept on the device
#define LV_NKTIMERS 4

// Grid constants, passed to kernels by value.
struct lv_grid_params {
    double bxmin0, bxmin1, bxmax0, bxmax1;   // boundary_rect
    double cmin0, cmin1, cmax0, cmax1;       // cropping_rect
    double origin0, origin1;                 // cell-list origin = cropping.xmin - (h, h)
    double h, dr, r_max, rr_max, xperiod, yperiod;
    int n1, n2, xperiodic, yperiodic;
    int npath, has_next;
    double next_rr;
};

// Polygon mesh in ELL layout, slot-major so that lane i of a wavefront reading slot k of
// particles i..i+63 touches one contiguous run: plane p of slot k of particle i is at
// ev[(p*LV_MAXE + k)*ld + i], label at el[k*ld + i].
struct lv_mesh_view {
    const int *ne;
    const double *ev;
    const int *el;
    int ld;
};

#define LV_HIP_CHECK(expr)                                                                          \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) return lv_fail_hip(ctx, _e, #expr, __FILE__, __LINE__);               \
    } while (0)

// ------------------------------------------------------------------------------------------------
// Device geometry helpers: each is the reference expression with its exact operation order.
// ------------------------------------------------------------------------------------------------

// get_arrow(q, x, grid) added to x: y = x + get_arrow(q.x, x), src/voronoigrid.jl:116-125
__device__ __forceinline__ void lv_arrow_target(const lv_grid_params &g, double q1, double q2, double x1, double x2,
                                                double &y1, double &y2)
{
    double v1 = q1 - x1, v2 = q2 - x2;
    if (g.xperiodic && fabs(v1) > 0.5 * g.xperiod) {
        double sp = ((v1 > 0.0) ? 1.0 : -1.0) * g.xperiod;
        v1 = v1 - sp * 1.0;
        v2 = v2 - sp * 0.0;
    }
    if (g.yperiodic && fabs(v2) > 0.5 * g.yperiod) {
        double sp = ((v2 > 0.0) ? 1.0 : -1.0) * g.yperiod;
        v1 = v1 - sp * 0.0;
        v2 = v2 - sp * 1.0;
    }
    y1 = x1 + v1;
    y2 = x2 + v2;
}

// lr_ratio(dx, e) = sqrt(|v1 - v2|^2 / |dx|^2), src/polygon.jl:228-232
__device__ __forceinline__ double lv_lr_ratio(double dx1, double dx2, double v1x, double v1y, double v2x, double v2y)
{
    double a = v1x - v2x, b = v1y - v2y;
    double l2 = a * a + b * b;
    double r2 = dx1 * dx1 + dx2 * dx2;
    return sqrt(l2 / r2);
}

__device__ __forceinline__ double lv_norm2(double a, double b) { return sqrt(a * a + b * b); }

// findkey, src/neighborlist.jl:47-52: 1-based (i1, i2); returns false if the key is off the grid
__device__ __forceinline__ bool lv_findkey(const lv_grid_params &g, double x1, double x2, int &i1, int &i2)
{
    double fa = floor((x1 - g.origin0) / g.h);
    double fb = floor((x2 - g.origin1) / g.h);
    if (!(fa == fa) || !(fb == fb) || fa < -2.0e9 || fa > 2.0e9 || fb < -2.0e9 || fb > 2.0e9) {
        i1 = i2 = -1;
        return false;
    }
    i1 = (int)fa + 1;
    i2 = (int)fb + 1;
    return i1 >= 1 && i1 <= g.n1 && i2 >= 1 && i2 <= g.n2;
}

// Wavefront / block reductions of doubles (64-lane wavefronts, blocks of up to 1024 threads).
__device__ __forceinline__ double lv_wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int BLOCK>
__device__ __forceinline__ double lv_block_sum(double v, double *smem)
{
    v = lv_wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (BLOCK > 64) {
        if (lane == 0) smem[w] = v;
        __syncthreads();
        if (w == 0) {
            v = (lane < BLOCK / 64) ? smem[lane] : 0.0;
            v = lv_wave_sum(v);
        }
        __syncthreads();
    }
    return v;  // valid in thread 0
}

// ------------------------------------------------------------------------------------------------
// Host-side launch wrappers implemented in the kernel translation units.
// ------------------------------------------------------------------------------------------------
struct lv_ctx;
int lv_fail_hip(lv_ctx *ctx, hipError_t e, const char *what, const char *file, int line);
int lv_fail(lv_ctx *ctx, int code, const char *msg);

// lv_scan.hip
size_t lv_scan_temp_bytes(int64_t n);
hipError_t lv_exclusive_scan_i32(const int *in, int *out, int64_t n, int *d_total, void *temp, hipStream_t s);

// lv_celllist.hip
size_t lv_celllist_temp_bytes(int64_t max_items, int64_t ncells);
hipError_t lv_build_cell_list(const lv_grid_params &g, int n, const double *x, int *cell_start, int *cell_labels,
                              double *cell_x, int64_t max_items, int *d_nitems, void *temp, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1, bool timed, int *h_nitems = nullptr);

// lv_clip.hip.  retry: n + 1 ints of work space (the seeds whose polygon needs the full capacity).
hipError_t lv_launch_clip(const lv_grid_params &g, int n, const double *x, const int *cell_start,
                          const int *cell_labels, const double *cell_x, const int *path_i, const double *path_rr,
                          int *ne, double *ev, int *el, int ld, double *area, double *cent, int *err, int *retry,
                          hipStream_t s, int m);

// Device pointers of the particle fields (PolygonNS, src/celldefs.jl:7-39) as SoA planes.
struct lv_fields {
    double *x, *v, *dv, *mom;                          // [2][ld]
    double *rho, *e, *P, *c2, *mass, *energy, *quality, *mu;
    double *D;                                         // [4][ld], column-major 2x2
    double *area, *cent;                               // cached area(p), centroid(p) of the mesh
    int *phase;
};

// Edge k of particle i in the ELL mesh
#define LV_EV(M, p, k, i) ((M).ev[(size_t)((p) * LV_MAXE + (k)) * (M).ld + (i)])
#define LV_EL(M, k, i) ((M).el[(size_t)(k) * (M).ld + (i)])

// lv_stencil.hip
hipError_t lv_launch_move(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int *err,
                          hipStream_t s);
hipError_t lv_launch_gravity(lv_fields F, int n, double g1, double g2, double dt, hipStream_t s);
hipError_t lv_launch_find_rho(lv_fields F, int n, hipStream_t s);
hipError_t lv_launch_eos(lv_fields F, int n, int kind, double gamma, double P0, double Pmin, hipStream_t s);
// with_find_D: find_D! in pressure_step!'s second pass (step! calls it right after)
hipError_t lv_launch_pressure_step(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt,
                                   hipStream_t s, bool with_find_D = false);
hipError_t lv_launch_find_D(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, hipStream_t s);
// with_find_dv: find_dv!(dv_dt, alpha) in viscous_step!'s second pass
hipError_t lv_launch_viscous(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, double avdr,
                             hipStream_t s, bool with_find_dv = false, double dv_dt = 0.0, double alpha = 0.0);
struct lv_vbc {
    double v[8];
};
hipError_t lv_launch_friction(lv_fields F, lv_mesh_view M, int n, lv_vbc vbc, double dt, hipStream_t s);
hipError_t lv_launch_find_dv(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, double alpha,
                             hipStream_t s);
hipError_t lv_launch_relaxation(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int rusanov,
                                hipStream_t s);

// lv_mls.hip: nearest_polygon, movingls (order 1, 2, 3: N = 2, 5, 9 coefficients per polygon, row-major)
// and point_value over the cell list of the last remesh
hipError_t lv_launch_movingls(const lv_grid_params &g, const double *x, int n, const int *cell_start,
                              const int *cell_labels, const int *path_i, const double *path_rr, int order,
                              const double *f, double h, double *out, hipStream_t s);
hipError_t lv_launch_nearest(const lv_grid_params &g, const double *x, int n, const int *cell_start,
                             const int *cell_labels, const int *path_i, const double *path_rr, int m,
                             const double *q, int *idx, hipStream_t s);
hipError_t lv_launch_point_value(const lv_grid_params &g, const double *x, int n, const int *cell_start,
                                 const int *cell_labels, const int *path_i, const double *path_rr, int m,
                                 const double *q, const double *f, double h, double *out, hipStream_t s);
