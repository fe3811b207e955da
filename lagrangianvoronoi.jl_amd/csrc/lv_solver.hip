// lv_solver.hip -- the pressure system of find_pressure! (src/pressure.jl:89-224): CSR assembly of
// PressureOperator (refresh!(A), :104-117), the right-hand side (refresh!(solver), :162-203), and the
// Krylov solvers that replace Krylov.jl's minres! (:219):
//   * FP64 Jacobi-PCG;
//   * FP32 Jacobi-PCG inside FP64 iterative refinement (FP64 residual, FP64 dots, FP32 SpMV/AXPY).
//
// The operator is kept in the reference's own form y_i = d_i x_i + sum_j c_ij (x_i - x_j): the
// difference form keeps the nearly-constant pressure modes accurate in FP32, which the folded
// diagonal form would lose to cancellation.  One CG iteration is three launches: SpMV fused with the
// p.q dot; the x/r/z update fused with the r.z and r.r dots; the p update.  Dots reduce per block
// into a partials array and the last block to finish (atomic ticket) sums the partials in a fixed
// order, so results are deterministic, and then computes alpha/beta/convergence on the device: the
// host reads one flag per batch of iterations instead of one per dot.
#include <string.h>

#include "lv_solver.h"

namespace {

constexpr int B = 256;

// ---- assembly ----------------------------------------------------------------------------------

__global__ __launch_bounds__(B) void k_degree(lv_mesh_view M, int n, int *__restrict__ deg)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    int d = 0;
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) d += (LV_EL(M, k, i) >= 0);
    deg[i] = d;
}

// refresh!(A, grid, dt), src/pressure.jl:104-117
__global__ __launch_bounds__(B) void k_assemble(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                                const int *__restrict__ row_ptr, int *__restrict__ cols,
                                                double *__restrict__ val64, float *__restrict__ val32,
                                                double *__restrict__ diag64, float *__restrict__ diag32,
                                                double *__restrict__ dinv64, float *__restrict__ dinv32,
                                                int *__restrict__ nonsingular)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double rho = F.rho[i];
    const double d = F.mass[i] / (((rho * rho) * F.c2[i]) * (dt * dt));
    double dd = d;
    int o = row_ptr[i];
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0) continue;
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr =
            lv_lr_ratio(x1 - y1, x2 - y2, LV_EV(M, 0, k, i), LV_EV(M, 1, k, i), LV_EV(M, 2, k, i), LV_EV(M, 3, k, i));
        const double c = lrr * (0.5 / rho + 0.5 / F.rho[j]);
        cols[o] = j;
        val64[o] = c;
        val32[o] = (float)c;
        dd += c;
        o++;
    }
    // the system is singular iff every d_i == 0 (c2 = Inf): one flag write per wavefront at most
    if (__any(d != 0.0) && (threadIdx.x & 63) == 0 && *nonsingular == 0) atomicOr(nonsingular, 1);
    diag64[i] = d;
    diag32[i] = (float)d;
    dinv64[i] = 1.0 / dd;
    dinv32[i] = (float)(1.0 / dd);
}

// refresh!(solver) first pass, src/pressure.jl:167-187: b, the initial guess P and GP
__global__ __launch_bounds__(B) void k_rhs1(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                            lv_vbc vbc, double *__restrict__ b, double *__restrict__ P0,
                                            double *__restrict__ GP)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double Pi = F.P[i];
    const double pv1 = F.v[i], pv2 = F.v[n + i];
    double bi = (F.area[i] * Pi) / ((F.rho[i] * F.c2[i]) * (dt * dt));
    double gx = 0.0, gy = 0.0;
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        const double dv1 = pv1 - F.v[j], dv2 = pv2 - F.v[n + j];
        bi -= (lrr / dt) * (dv1 * (m1 - y1) + dv2 * (m2 - y2));
        const double w = lrr * (Pi - F.P[j]);
        gx -= w * (m1 - x1);
        gy -= w * (m2 - x2);
    }
    for (int k = 0; k < ne; k++) {  // boundaries(p)
        const int lab = LV_EL(M, k, i);
        if (lab >= 0) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        const double dS1 = e1y - e2y, dS2 = e2x - e1x;
        const int slot = -lab - 1;
        const double vb1 = vbc.v[2 * slot], vb2 = vbc.v[2 * slot + 1];
        bi -= (dS1 * (vb1 - pv1) + dS2 * (vb2 - pv2)) / dt;
    }
    const double mi = F.mass[i];
    GP[i] = gx / mi;
    GP[n + i] = gy / mi;
    b[i] = bi;
    P0[i] = Pi;
}

// refresh!(solver) gp_step pass, src/pressure.jl:189-202
__global__ __launch_bounds__(B) void k_rhs2(lv_grid_params g, lv_fields F, lv_mesh_view M, int n,
                                            const double *__restrict__ GP, double *__restrict__ b)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double g1 = GP[i], g2 = GP[n + i];
    double bi = b[i];
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        const double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
        bi += lrr * ((g1 - GP[j]) * (m1 - z1) + (g2 - GP[n + j]) * (m2 - z2));
    }
    b[i] = bi;
}

__global__ __launch_bounds__(B) void k_write_P(lv_fields F, int n, const double *__restrict__ P)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) F.P[i] = P[i];
}

// ---- reductions --------------------------------------------------------------------------------

// Publish this block's NV partial sums; returns true in the last block to arrive, after it has
// summed all partials into out[0..NV) in a fixed order.
template <int NV>
__device__ __forceinline__ bool block_reduce_last(double (&v)[NV], double *__restrict__ partials, int nb,
                                                  unsigned *__restrict__ ticket, double (&out)[NV])
{
    __shared__ double sm[NV][B / 64];
    __shared__ bool last;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double s = lv_wave_sum(v[q]);
        if (lane == 0) sm[q][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double s = 0.0;
            for (int k = 0; k < B / 64; k++) s += sm[q][k];
            partials[(size_t)q * nb + blockIdx.x] = s;
        }
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        last = (t == (unsigned)nb - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double s = 0.0;
        for (int k = threadIdx.x; k < nb; k += B) s += partials[(size_t)q * nb + k];
        s = lv_wave_sum(s);
        __syncthreads();
        if (lane == 0) sm[q][w] = s;
        __syncthreads();
        double t = 0.0;
        for (int k = 0; k < B / 64; k++) t += sm[q][k];
        out[q] = t;
    }
    if (threadIdx.x == 0) *ticket = 0u;
    return true;
}

// ---- FP64 Jacobi-PCG ----------------------------------------------------------------------------

// sums of b and x (singular systems: b is projected on 1-perp, x keeps its mean)
__global__ __launch_bounds__(B) void k_means(int n, const double *__restrict__ b, const double *__restrict__ x,
                                             double *__restrict__ partials, lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        v[0] = b[i];
        v[1] = x[i];
    }
    double o[2];
    if (block_reduce_last<2>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->bmean = S->singular ? o[0] / n : 0.0;
        S->xmean0 = o[1] / n;
    }
}

// y = A x in the reference's difference form
template <typename T>
__device__ __forceinline__ T spmv_row(int i, const int *__restrict__ rp, const int *__restrict__ cols,
                                      const T *__restrict__ vals, const T *__restrict__ diag, const T *__restrict__ x)
{
    const T xi = x[i];
    T acc = diag[i] * xi;
    const int e = rp[i + 1];
    for (int o = rp[i]; o < e; o++) acc += vals[o] * (xi - x[cols[o]]);
    return acc;
}

// r = (b - bmean) - A x; z = dinv r; p = z; dots r.z, r.r, b'.b'
__global__ __launch_bounds__(B) void k_cg64_init(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                                 const double *__restrict__ vals, const double *__restrict__ diag,
                                                 const double *__restrict__ dinv, const double *__restrict__ b,
                                                 const double *__restrict__ x, double *__restrict__ r,
                                                 double *__restrict__ z, double *__restrict__ p,
                                                 double *__restrict__ partials, lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double bi = b[i] - S->bmean;
        const double ri = bi - spmv_row<double>(i, rp, cols, vals, diag, x);
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
        v[2] = bi * bi;
    }
    double o[3];
    if (block_reduce_last<3>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->rz = o[0];
        S->rr = o[1];
        S->bb = o[2];
        S->thr2 = S->tol2 * o[2];
        S->iters = 0;
        S->done = (o[1] <= S->thr2) ? 1 : 0;
    }
}

__global__ __launch_bounds__(B) void k_cg64_spmv(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                                 const double *__restrict__ vals, const double *__restrict__ diag,
                                                 const double *__restrict__ p, double *__restrict__ q,
                                                 double *__restrict__ partials, lv_cg_scalars *S)
{
    if (S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const double qi = spmv_row<double>(i, rp, cols, vals, diag, p);
        q[i] = qi;
        v[0] = p[i] * qi;
    }
    double o[1];
    if (block_reduce_last<1>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->pq = o[0];
        S->alpha = S->rz / o[0];
    }
}

__global__ __launch_bounds__(B) void k_cg64_update(int n, double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, const double *__restrict__ p,
                                                   const double *__restrict__ q, const double *__restrict__ dinv,
                                                   double *__restrict__ partials, lv_cg_scalars *S)
{
    if (S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    const double alpha = S->alpha;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
    }
    double o[2];
    if (block_reduce_last<2>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->beta = o[0] / S->rz;
        S->rz = o[0];
        S->rr = o[1];
        S->iters += 1;
        if (o[1] <= S->thr2) S->done = 1;
        else if (S->iters >= S->itmax || !(o[1] == o[1])) S->done = 2;
    }
}

__global__ __launch_bounds__(B) void k_cg64_pupdate(int n, const double *__restrict__ z, double *__restrict__ p,
                                                    const lv_cg_scalars *S)
{
    if (S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) p[i] = z[i] + S->beta * p[i];
}

// true residual r = (b - bmean) - A x and |r|^2 (FP64), for refinement sweeps and final checks
__global__ __launch_bounds__(B) void k_resid64(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                               const double *__restrict__ vals, const double *__restrict__ diag,
                                               const double *__restrict__ b, const double *__restrict__ x,
                                               double *__restrict__ r, double *__restrict__ partials,
                                               lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        const double bi = b[i] - S->bmean;
        const double ri = bi - spmv_row<double>(i, rp, cols, vals, diag, x);
        r[i] = ri;
        v[0] = ri * ri;
        v[1] = bi * bi;
    }
    double o[2];
    if (block_reduce_last<2>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->rr64 = o[0];
        S->bb = o[1];
        S->thr2 = S->tol2 * o[1];
    }
}

// restore the mean of x to that of the initial guess (singular systems only)
__global__ __launch_bounds__(B) void k_sum1(int n, const double *__restrict__ x, double *__restrict__ partials,
                                            lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    double v[1] = {(i < n) ? x[i] : 0.0};
    double o[1];
    if (block_reduce_last<1>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) S->xmean = o[0] / n;
}

__global__ __launch_bounds__(B) void k_shift(int n, double *__restrict__ x, const lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) x[i] -= (S->xmean - S->xmean0);
}

// ---- FP32 inner PCG of the refinement --------------------------------------------------------

// r32 = r64 / |r64| (projected on 1-perp for singular systems), e = 0, z = dinv r, p = z
__global__ __launch_bounds__(B) void k_cg32_init(int n, const double *__restrict__ r64, const float *__restrict__ dinv,
                                                 float *__restrict__ e, float *__restrict__ r, float *__restrict__ z,
                                                 float *__restrict__ p, double *__restrict__ partials,
                                                 lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const double inv = 1.0 / sqrt(S->rr64);
    double v[2] = {0.0, 0.0};
    if (i < n) {
        const float ri = (float)(r64[i] * inv);
        const float zi = dinv[i] * ri;
        e[i] = 0.0f;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        v[0] = (double)ri * (double)zi;
        v[1] = (double)ri * (double)ri;
    }
    double o[2];
    if (block_reduce_last<2>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->rz = o[0];
        S->rr = o[1];
        S->thr2 = S->inner_tol2 * o[1];
        S->iters = 0;
        S->done = (o[1] <= S->thr2) ? 1 : 0;
    }
}

__global__ __launch_bounds__(B) void k_cg32_spmv(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                                 const float *__restrict__ vals, const float *__restrict__ diag,
                                                 const float *__restrict__ p, float *__restrict__ q,
                                                 double *__restrict__ partials, lv_cg_scalars *S)
{
    if (S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const float qi = spmv_row<float>(i, rp, cols, vals, diag, p);
        q[i] = qi;
        v[0] = (double)p[i] * (double)qi;
    }
    double o[1];
    if (block_reduce_last<1>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->pq = o[0];
        S->alpha = S->rz / o[0];
    }
}

__global__ __launch_bounds__(B) void k_cg32_update(int n, float *__restrict__ e, float *__restrict__ r,
                                                   float *__restrict__ z, const float *__restrict__ p,
                                                   const float *__restrict__ q, const float *__restrict__ dinv,
                                                   double *__restrict__ partials, lv_cg_scalars *S)
{
    if (S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    const float alpha = (float)S->alpha;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        e[i] += alpha * p[i];
        const float ri = r[i] - alpha * q[i];
        const float zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        v[0] = (double)ri * (double)zi;
        v[1] = (double)ri * (double)ri;
    }
    double o[2];
    if (block_reduce_last<2>(v, partials, gridDim.x, &S->ticket, o) && threadIdx.x == 0) {
        S->beta = o[0] / S->rz;
        S->rz = o[0];
        S->rr = o[1];
        S->iters += 1;
        if (o[1] <= S->thr2) S->done = 1;
        else if (S->iters >= S->itmax || !(o[1] == o[1])) S->done = 2;
    }
}

__global__ __launch_bounds__(B) void k_cg32_pupdate(int n, const float *__restrict__ z, float *__restrict__ p,
                                                    const lv_cg_scalars *S)
{
    if (S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    const float beta = (float)S->beta;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// x += |r64| * e
__global__ __launch_bounds__(B) void k_correct(int n, double *__restrict__ x, const float *__restrict__ e,
                                               const lv_cg_scalars *S)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const double s = sqrt(S->rr64);
    if (i < n) x[i] += s * (double)e[i];
}

inline dim3 grid_of(int n) { return dim3((n + B - 1) / B); }

}  // namespace

// ---- host wrappers ---------------------------------------------------------------------------------

hipError_t lv_launch_degree(lv_mesh_view M, int n, int *deg, hipStream_t s)
{
    hipLaunchKernelGGL(k_degree, grid_of(n), dim3(B), 0, s, M, n, deg);
    return hipGetLastError();
}

hipError_t lv_launch_assemble(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt,
                              lv_pressure_system &A, int *nonsingular, hipStream_t s)
{
    hipLaunchKernelGGL(k_assemble, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, (const int *)A.row_ptr, A.cols,
                       A.val64, A.val32, A.diag64, A.diag32, A.dinv64, A.dinv32, nonsingular);
    return hipGetLastError();
}

hipError_t lv_launch_rhs(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int gp_step,
                         lv_vbc vbc, double *b, double *P0, double *GP, hipStream_t s)
{
    hipLaunchKernelGGL(k_rhs1, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, vbc, b, P0, GP);
    if (gp_step) hipLaunchKernelGGL(k_rhs2, grid_of(n), dim3(B), 0, s, g, F, M, n, (const double *)GP, b);
    return hipGetLastError();
}

hipError_t lv_launch_write_P(lv_fields F, int n, const double *P, hipStream_t s)
{
    hipLaunchKernelGGL(k_write_P, grid_of(n), dim3(B), 0, s, F, n, P);
    return hipGetLastError();
}

size_t lv_partials_bytes(int n) { return (size_t)4 * ((n + B - 1) / B) * sizeof(double); }

namespace {

struct timer_scope {
    lv_timer *t;
    hipStream_t s;
    timer_scope(lv_timer *t_, hipStream_t s_) : t(t_), s(s_)
    {
        if (t && t->on) lv_timer_begin(t, s);
    }
    ~timer_scope()
    {
        if (t && t->on) lv_timer_end(t, s);
    }
};

hipError_t read_scalars(lv_cg_scalars *dS, lv_cg_scalars *hS, hipStream_t s)
{
    hipError_t e = hipMemcpyAsync(hS, dS, sizeof(lv_cg_scalars), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

}  // namespace

// Runs CG64 iterations until done; batches of `batch` iterations between host checks.
static hipError_t cg64_iterate(const lv_pressure_system &A, lv_solver_work &W, int n, hipStream_t s, lv_timer *t64,
                               int batch, lv_cg_scalars &h)
{
    const dim3 G = grid_of(n);
    for (;;) {
        for (int k = 0; k < batch; k++) {
            {
                timer_scope ts(t64, s);
                hipLaunchKernelGGL(k_cg64_spmv, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                                   (const double *)A.val64, (const double *)A.diag64, (const double *)W.p, W.q,
                                   W.partials, W.S);
            }
            hipLaunchKernelGGL(k_cg64_update, G, dim3(B), 0, s, n, W.x, W.r, W.z, (const double *)W.p,
                               (const double *)W.q, (const double *)A.dinv64, W.partials, W.S);
            hipLaunchKernelGGL(k_cg64_pupdate, G, dim3(B), 0, s, n, (const double *)W.z, W.p,
                               (const lv_cg_scalars *)W.S);
        }
        hipError_t e = read_scalars(W.S, &h, s);
        if (e != hipSuccess) return e;
        if (t64) lv_timer_flush(t64);
        if (h.done) return hipSuccess;
    }
}

static hipError_t cg32_iterate(const lv_pressure_system &A, lv_solver_work &W, int n, hipStream_t s, lv_timer *t32,
                               int batch, lv_cg_scalars &h)
{
    const dim3 G = grid_of(n);
    for (;;) {
        for (int k = 0; k < batch; k++) {
            {
                timer_scope ts(t32, s);
                hipLaunchKernelGGL(k_cg32_spmv, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                                   (const float *)A.val32, (const float *)A.diag32, (const float *)W.p32, W.q32,
                                   W.partials, W.S);
            }
            hipLaunchKernelGGL(k_cg32_update, G, dim3(B), 0, s, n, W.e32, W.r32, W.z32, (const float *)W.p32,
                               (const float *)W.q32, (const float *)A.dinv32, W.partials, W.S);
            hipLaunchKernelGGL(k_cg32_pupdate, G, dim3(B), 0, s, n, (const float *)W.z32, W.p32,
                               (const lv_cg_scalars *)W.S);
        }
        hipError_t e = read_scalars(W.S, &h, s);
        if (e != hipSuccess) return e;
        if (t32) lv_timer_flush(t32);
        if (h.done) return hipSuccess;
    }
}

// Solves A x = b (device vectors; x holds the initial guess).  Writes iterations, sweeps and the
// final FP64 relative residual.  Returns hipSuccess; *converged tells whether tol was met.
hipError_t lv_solve_device(const lv_pressure_system &A, lv_solver_work &W, int n, const double *b, double *x,
                           const lv_solve_opts &opt, hipStream_t s, lv_timer *t32, lv_timer *t64, int *iters,
                           int *sweeps, double *relres, bool *converged)
{
    const dim3 G = grid_of(n);
    lv_cg_scalars h;
    memset(&h, 0, sizeof(h));
    h.tol2 = opt.tol * opt.tol;
    h.inner_tol2 = opt.inner_tol * opt.inner_tol;
    h.itmax = opt.itmax > 0 ? opt.itmax : 1000;
    h.singular = A.singular;
    hipError_t e = hipMemcpyAsync(W.S, &h, sizeof(h), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(W.x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_means, G, dim3(B), 0, s, n, b, (const double *)W.x, W.partials, W.S);
    const int batch = 8;
    int total = 0, nsweep = 0;
    if (opt.method == LV_SOLVER_PCG64) {
        hipLaunchKernelGGL(k_cg64_init, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                           (const double *)A.val64, (const double *)A.diag64, (const double *)A.dinv64, b,
                           (const double *)W.x, W.r, W.z, W.p, W.partials, W.S);
        e = cg64_iterate(A, W, n, s, t64, batch, h);
        if (e != hipSuccess) return e;
        total = h.iters;
    } else {
        const int outer_max = opt.outer_max > 0 ? opt.outer_max : 20;
        for (;;) {
            {
                timer_scope ts(t64, s);
                hipLaunchKernelGGL(k_resid64, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                                   (const double *)A.val64, (const double *)A.diag64, b, (const double *)W.x, W.r,
                                   W.partials, W.S);
            }
            e = read_scalars(W.S, &h, s);
            if (e != hipSuccess) return e;
            if (h.rr64 <= h.thr2 || nsweep >= outer_max || !(h.rr64 == h.rr64)) break;
            hipLaunchKernelGGL(k_cg32_init, G, dim3(B), 0, s, n, (const double *)W.r, (const float *)A.dinv32, W.e32,
                               W.r32, W.z32, W.p32, W.partials, W.S);
            e = cg32_iterate(A, W, n, s, t32, batch, h);
            if (e != hipSuccess) return e;
            total += h.iters;
            nsweep++;
            hipLaunchKernelGGL(k_correct, G, dim3(B), 0, s, n, W.x, (const float *)W.e32, (const lv_cg_scalars *)W.S);
        }
    }
    // final FP64 residual of the returned x
    if (A.singular) {
        hipLaunchKernelGGL(k_sum1, G, dim3(B), 0, s, n, (const double *)W.x, W.partials, W.S);
        hipLaunchKernelGGL(k_shift, G, dim3(B), 0, s, n, W.x, (const lv_cg_scalars *)W.S);
    }
    hipLaunchKernelGGL(k_resid64, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                       (const double *)A.val64, (const double *)A.diag64, b, (const double *)W.x, W.r, W.partials,
                       W.S);
    e = read_scalars(W.S, &h, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(x, W.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    *iters = total;
    *sweeps = nsweep;
    *relres = (h.bb > 0.0) ? sqrt(h.rr64 / h.bb) : sqrt(h.rr64);
    *converged = (h.rr64 <= h.thr2);
    return hipSuccess;
}
