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
//
// A kernel that produces dot products writes one partial sum per block with plain stores.  The
// kernel that consumes them re-reduces all NB partials in every block, in the same fixed order,
// so every block obtains bit-identical totals without atomics or device-scope fences (on a
// multi-XCD GPU a release fence per block costs an L2 write-back).  The kernel boundary makes the
// partials visible.  Block 0 also stores the totals in lv_cg_scalars for the host and later kernels.

template <int NV>
__device__ __forceinline__ void block_partials(const double (&v)[NV], double *__restrict__ partials, int nb)
{
    __shared__ double sm[NV][B / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double s = lv_wave_sum(v[q]);
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
    }
}

template <int NV>
__device__ __forceinline__ void reduce_all(const double *__restrict__ partials, int nb, double (&out)[NV])
{
    __shared__ double sm[NV][B / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double s = 0.0;
        for (int k = threadIdx.x; k < nb; k += B) s += partials[(size_t)q * nb + k];
        s = lv_wave_sum(s);
        if (lane == 0) sm[q][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double t = 0.0;
        for (int k = 0; k < B / 64; k++) t += sm[q][k];
        out[q] = t;
    }
    __syncthreads();
}

// single-block finalisers of the one-off reductions
enum { FIN_MEANS = 0, FIN_INIT64 = 1, FIN_RESID = 2, FIN_SUM1 = 3, FIN_INIT32 = 4 };

__global__ __launch_bounds__(B) void k_finalize(const double *__restrict__ partials, int nb, int mode, int n,
                                                double inner_tol2, lv_cg_scalars *S)
{
    double o[3];
    {
        double t3[3];
        reduce_all<3>(partials, nb, t3);
        o[0] = t3[0];
        o[1] = t3[1];
        o[2] = t3[2];
    }
    if (threadIdx.x != 0) return;
    switch (mode) {
    case FIN_MEANS:
        S->bmean = S->singular ? o[0] / n : 0.0;
        S->xmean0 = o[1] / n;
        break;
    case FIN_INIT64:
        S->rz[0] = o[0];
        S->rr = o[1];
        S->bb = o[2];
        S->thr2 = S->tol2 * o[2];
        S->iters = 0;
        S->done = (o[1] <= S->thr2) ? 1 : 0;
        break;
    case FIN_RESID:
        S->rr64 = o[0];
        S->bb = o[1];
        S->thr2 = S->tol2 * o[1];
        break;
    case FIN_SUM1:
        S->xmean = o[0] / n;
        break;
    case FIN_INIT32:
        S->rz[0] = o[0];
        S->rr = o[1];
        S->thr2 = inner_tol2 * o[1];
        S->iters = 0;
        S->done = (o[1] <= S->thr2) ? 1 : 0;
        break;
    }
}

// ---- FP64 Jacobi-PCG ----------------------------------------------------------------------------

// sums of b and x (singular systems: b is projected on 1-perp, x keeps its mean)
__global__ __launch_bounds__(B) void k_means(int n, const double *__restrict__ b, const double *__restrict__ x,
                                             double *__restrict__ partials)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        v[0] += b[i];
        v[1] += x[i];
    }
    block_partials<3>(v, partials, gridDim.x);
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

// r = (b - bmean) - A x; z = dinv r; p = z; partials r.z, r.r, b'.b'
__global__ __launch_bounds__(B) void k_cg64_init(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                                 const double *__restrict__ vals, const double *__restrict__ diag,
                                                 const double *__restrict__ dinv, const double *__restrict__ b,
                                                 const double *__restrict__ x, double *__restrict__ r,
                                                 double *__restrict__ z, double *__restrict__ p,
                                                 double *__restrict__ partials, const lv_cg_scalars *S)
{
    const double bmean = S->bmean;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const double bi = b[i] - bmean;
        const double ri = bi - spmv_row<double>(i, rp, cols, vals, diag, x);
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        v[0] += ri * zi;
        v[1] += ri * ri;
        v[2] += bi * bi;
    }
    block_partials<3>(v, partials, gridDim.x);
}

template <typename T>
__global__ __launch_bounds__(B) void k_cg_spmv(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                               const T *__restrict__ vals, const T *__restrict__ diag,
                                               const T *__restrict__ p, T *__restrict__ q,
                                               double *__restrict__ Ppq, const lv_cg_scalars *S)
{
    if (S->done) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const T qi = spmv_row<T>(i, rp, cols, vals, diag, p);
        q[i] = qi;
        v[0] += (double)p[i] * (double)qi;
    }
    block_partials<1>(v, Ppq, gridDim.x);
}

// alpha = rz/pq; x += alpha p; r -= alpha q; z = dinv r; partials r.z, r.r
template <typename T>
__global__ __launch_bounds__(B) void k_cg_update(int n, T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                                 const T *__restrict__ p, const T *__restrict__ q,
                                                 const T *__restrict__ dinv, const double *__restrict__ Ppq,
                                                 double *__restrict__ Prr, lv_cg_scalars *S, int par)
{
    if (S->done) return;
    double pq[1];
    reduce_all<1>(Ppq, gridDim.x, pq);
    const double alpha = S->rz[par] / pq[0];
    const T a = (T)alpha;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        v[0] += (double)ri * (double)zi;
        v[1] += (double)ri * (double)ri;
    }
    block_partials<2>(v, Prr, gridDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->pq = pq[0];
        S->alpha = alpha;
    }
}

// beta = rz_new/rz; p = z + beta p; convergence bookkeeping in block 0
template <typename T>
__global__ __launch_bounds__(B) void k_cg_pupdate(int n, const T *__restrict__ z, T *__restrict__ p,
                                                  const double *__restrict__ Prr, lv_cg_scalars *S, int par)
{
    if (S->done) return;
    double o[2];
    reduce_all<2>(Prr, gridDim.x, o);
    const double beta = o[0] / S->rz[par];
    const T bt = (T)beta;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) p[i] = z[i] + bt * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->rz[par ^ 1] = o[0];
        S->beta = beta;
        S->rr = o[1];
        S->iters += 1;
        if (o[1] <= S->thr2) S->done = 1;
        else if (S->iters >= S->itmax || !(o[1] == o[1])) S->done = 2;
    }
}

// true residual r = (b - bmean) - A x; partials |r|^2, |b'|^2 (FP64)
__global__ __launch_bounds__(B) void k_resid64(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                               const double *__restrict__ vals, const double *__restrict__ diag,
                                               const double *__restrict__ b, const double *__restrict__ x,
                                               double *__restrict__ r, double *__restrict__ partials,
                                               const lv_cg_scalars *S)
{
    const double bmean = S->bmean;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const double bi = b[i] - bmean;
        const double ri = bi - spmv_row<double>(i, rp, cols, vals, diag, x);
        r[i] = ri;
        v[0] += ri * ri;
        v[1] += bi * bi;
    }
    block_partials<3>(v, partials, gridDim.x);
}

__global__ __launch_bounds__(B) void k_sum1(int n, const double *__restrict__ x, double *__restrict__ partials)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) v[0] += x[i];
    block_partials<3>(v, partials, gridDim.x);
}

// restore the mean of x to that of the initial guess (singular systems only)
__global__ __launch_bounds__(B) void k_shift(int n, double *__restrict__ x, const lv_cg_scalars *S)
{
    const double d = S->xmean - S->xmean0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) x[i] -= d;
}

// ---- FP32 inner PCG of the refinement --------------------------------------------------------

// r32 = r64 / |r64|, e = 0, z = dinv r, p = z; partials r.z, r.r
__global__ __launch_bounds__(B) void k_cg32_init(int n, const double *__restrict__ r64, const float *__restrict__ dinv,
                                                 float *__restrict__ e, float *__restrict__ r, float *__restrict__ z,
                                                 float *__restrict__ p, double *__restrict__ partials,
                                                 const lv_cg_scalars *S)
{
    const double inv = 1.0 / sqrt(S->rr64);
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const float ri = (float)(r64[i] * inv);
        const float zi = dinv[i] * ri;
        e[i] = 0.0f;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        v[0] += (double)ri * (double)zi;
        v[1] += (double)ri * (double)ri;
    }
    block_partials<3>(v, partials, gridDim.x);
}

// x += |r64| * e
__global__ __launch_bounds__(B) void k_correct(int n, double *__restrict__ x, const float *__restrict__ e,
                                               const lv_cg_scalars *S)
{
    const double s = sqrt(S->rr64);
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) x[i] += s * (double)e[i];
}

inline dim3 grid_of(int n) { return dim3((n + B - 1) / B); }

// grid of the reduction kernels: enough 256-thread blocks for ~8 wavefronts per SIMD on 256 CUs,
// capped so that re-reducing the partials stays cheap
constexpr int NBMAX = 2048;
inline int nblocks_of(int n)
{
    int nb = (n + B - 1) / B;
    return nb < NBMAX ? (nb > 0 ? nb : 1) : NBMAX;
}

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

size_t lv_partials_bytes(int n)
{
    (void)n;
    return (size_t)8 * NBMAX * sizeof(double);
}

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

// partial-sum regions: P3 = one-off reductions (3 values), Ppq (1), Prr (2)
struct parts {
    double *p3, *pq, *rr;
};
inline parts parts_of(double *base) { return parts{base, base + 3 * NBMAX, base + 4 * NBMAX}; }

// CG iterations (three launches each) until the device flags completion; the host checks the flag
// once per batch of iterations.
template <typename T>
hipError_t cg_iterate(int n, const int *rp, const int *cols, const T *vals, const T *diag, const T *dinv, T *x, T *r,
                      T *z, T *p, T *q, const parts &P, lv_cg_scalars *dS, hipStream_t s, lv_timer *tm, int batch,
                      lv_cg_scalars &h)
{
    const int nb = nblocks_of(n);
    int k = 0;
    for (;;) {
        for (int j = 0; j < batch; j++, k++) {
            const int par = k & 1;
            {
                timer_scope ts(tm, s);
                hipLaunchKernelGGL(k_cg_spmv<T>, dim3(nb), dim3(B), 0, s, n, rp, cols, vals, diag, (const T *)p, q,
                                   P.pq, (const lv_cg_scalars *)dS);
            }
            hipLaunchKernelGGL(k_cg_update<T>, dim3(nb), dim3(B), 0, s, n, x, r, z, (const T *)p, (const T *)q, dinv,
                               (const double *)P.pq, P.rr, dS, par);
            hipLaunchKernelGGL(k_cg_pupdate<T>, dim3(nb), dim3(B), 0, s, n, (const T *)z, p, (const double *)P.rr,
                               dS, par);
        }
        hipError_t e = read_scalars(dS, &h, s);
        if (e != hipSuccess) return e;
        if (tm) lv_timer_flush(tm);
        if (h.done) return hipSuccess;
    }
}

}  // namespace

// Solves A x = b (device vectors; x holds the initial guess).  Writes iterations, sweeps and the
// final FP64 relative residual.  *converged tells whether tol was met.
hipError_t lv_solve_device(const lv_pressure_system &A, lv_solver_work &W, int n, const double *b, double *x,
                           const lv_solve_opts &opt, hipStream_t s, lv_timer *t32, lv_timer *t64, int *iters,
                           int *sweeps, double *relres, bool *converged)
{
    const int nb = nblocks_of(n);
    const dim3 G(nb);
    const parts P = parts_of(W.partials);
    lv_cg_scalars h;
    memset(&h, 0, sizeof(h));
    h.tol2 = opt.tol * opt.tol;
    h.itmax = opt.itmax > 0 ? opt.itmax : 1000;
    h.singular = A.singular;
    hipError_t e = hipMemcpyAsync(W.S, &h, sizeof(h), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(W.x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_means, G, dim3(B), 0, s, n, b, (const double *)W.x, P.p3);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_MEANS, n, 0.0, W.S);
    const int batch = 16;
    int total = 0, nsweep = 0;
    if (opt.method == LV_SOLVER_PCG64) {
        hipLaunchKernelGGL(k_cg64_init, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                           (const double *)A.val64, (const double *)A.diag64, (const double *)A.dinv64, b,
                           (const double *)W.x, W.r, W.z, W.p, P.p3, (const lv_cg_scalars *)W.S);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_INIT64, n, 0.0,
                           W.S);
        e = read_scalars(W.S, &h, s);
        if (e != hipSuccess) return e;
        if (!h.done) {
            e = cg_iterate<double>(n, A.row_ptr, A.cols, A.val64, A.diag64, A.dinv64, W.x, W.r, W.z, W.p, W.q, P,
                                   W.S, s, t64, batch, h);
            if (e != hipSuccess) return e;
        }
        total = h.iters;
    } else {
        const int outer_max = opt.outer_max > 0 ? opt.outer_max : 20;
        const double floor_tol = opt.inner_tol > 0.0 ? opt.inner_tol : 1e-4;
        for (;;) {
            {
                timer_scope ts(t64, s);
                hipLaunchKernelGGL(k_resid64, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                                   (const double *)A.val64, (const double *)A.diag64, b, (const double *)W.x, W.r,
                                   P.p3, (const lv_cg_scalars *)W.S);
            }
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_RESID, n, 0.0,
                               W.S);
            e = read_scalars(W.S, &h, s);
            if (e != hipSuccess) return e;
            if (h.rr64 <= h.thr2 || nsweep >= outer_max || !(h.rr64 == h.rr64)) break;
            // ask the FP32 sweep for just the reduction still missing (with a 0.3 margin), but not
            // below what FP32 arithmetic can deliver in one sweep
            double need = 0.3 * sqrt(h.thr2 / h.rr64);
            double itol = need > floor_tol ? need : floor_tol;
            if (itol > 0.5) itol = 0.5;
            hipLaunchKernelGGL(k_cg32_init, G, dim3(B), 0, s, n, (const double *)W.r, (const float *)A.dinv32, W.e32,
                               W.r32, W.z32, W.p32, P.p3, (const lv_cg_scalars *)W.S);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_INIT32, n,
                               itol * itol, W.S);
            e = read_scalars(W.S, &h, s);
            if (e != hipSuccess) return e;
            if (!h.done) {
                e = cg_iterate<float>(n, A.row_ptr, A.cols, A.val32, A.diag32, A.dinv32, W.e32, W.r32, W.z32, W.p32,
                                      W.q32, P, W.S, s, t32, batch, h);
                if (e != hipSuccess) return e;
            }
            total += h.iters;
            nsweep++;
            hipLaunchKernelGGL(k_correct, G, dim3(B), 0, s, n, W.x, (const float *)W.e32, (const lv_cg_scalars *)W.S);
        }
    }
    // final FP64 residual of the returned x
    if (A.singular) {
        hipLaunchKernelGGL(k_sum1, G, dim3(B), 0, s, n, (const double *)W.x, P.p3);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_SUM1, n, 0.0, W.S);
        hipLaunchKernelGGL(k_shift, G, dim3(B), 0, s, n, W.x, (const lv_cg_scalars *)W.S);
    }
    hipLaunchKernelGGL(k_resid64, G, dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                       (const double *)A.val64, (const double *)A.diag64, b, (const double *)W.x, W.r, P.p3,
                       (const lv_cg_scalars *)W.S);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_RESID, n, 0.0, W.S);
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
