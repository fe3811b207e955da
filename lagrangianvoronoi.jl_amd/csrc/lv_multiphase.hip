// lv_multiphase.hip -- multiphase_projection! (src/relaxation.jl:75-206): MultiphaseProjector,
// refresh!(MultiphaseSolver) and the dv update, with the linear solve on the device.
//
// Only edges between particles of different phases contribute, so the operator couples interface
// particles only and its other rows are zero.  The interface set is compacted once per call (flag,
// scan, scatter) and every kernel runs over it: rows are gathered from the full-size mesh and
// fields, and the unknowns are compact vectors.  Krylov.jl's minres! (atol = rtol = 1e-4, itmax =
// 200, zero initial guess) is replaced by unpreconditioned CG from zero with the same stopping rule
// |r| <= atol + rtol |b| (the operator is symmetric positive semi-definite and b lies in its range,
// so CG converges to the same solution; the iterates differ).  Every FP64 expression keeps the
// reference's operand order; dot products reduce through fixed-order block partials.
#include "lv_internal.h"
#include "lv_reduce.h"

namespace {

constexpr int B = 256;

inline dim3 g1(int64_t count) { return dim3((unsigned)std::max<int64_t>(1, (count + B - 1) / B)); }

// rr is double-buffered by iteration parity (iteration k reads rr[k&1], writes rr[(k+1)&1]) so that no
// block reads a value another block of the same launch is replacing; the iteration index comes from
// the host for the same reason
struct mph_scalars {
    double rr[2], rr0, pq, thr2;
    int iters, itmax, done;
};

// 1 if particle i has a neighbour of another phase
__global__ __launch_bounds__(B) void k_mph_flag(lv_fields F, lv_mesh_view M, int n, int *__restrict__ flag)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const int ph = F.phase[i];
    int f = 0;
    for (int k = 0; k < M.ne[i]; k++) {
        const int j = LV_EL(M, k, i);
        if (j >= 0 && F.phase[j] != ph) f = 1;
    }
    flag[i] = f;
}

__global__ __launch_bounds__(B) void k_mph_compact(int n, const int *__restrict__ flag, const int *__restrict__ pos,
                                                   int *__restrict__ list, int *__restrict__ inv)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) {
        list[pos[i]] = i;
        inv[i] = pos[i];
    } else {
        inv[i] = -1;
    }
}

// refresh!(solver::MultiphaseSolver), src/relaxation.jl:165-180, for the interface rows; partial |b|^2
__global__ __launch_bounds__(B) void k_mph_rhs(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, int m,
                                               const int *__restrict__ list, double *__restrict__ b,
                                               double *__restrict__ part)
{
    double acc[1] = {0.0};
    for (int c = blockIdx.x * B + threadIdx.x; c < m; c += gridDim.x * B) {
        const int i = list[c];
        const double x1 = F.x[i], x2 = F.x[n + i];
        const double d1 = F.dv[i], d2 = F.dv[n + i];
        const int ph = F.phase[i];
        double bi = 0.0;
        for (int k = 0; k < M.ne[i]; k++) {
            const int j = LV_EL(M, k, i);
            if (j < 0 || F.phase[j] == ph) continue;
            const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i),
                         e2y = LV_EV(M, 3, k, i);
            double y1, y2;
            lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
            const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
            const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
            const double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
            const double q1 = F.dv[j], q2 = F.dv[n + j];
            const double a = (d1 - q1) * (m1 - z1) + (d2 - q2) * (m2 - z2);
            const double cc = (d1 + q1) * (x1 - y1) + (d2 + q2) * (x2 - y2);
            bi -= lrr * (a - 0.5 * cc);
        }
        b[c] = bi;
        acc[0] += bi * bi;
    }
    block_partials<1>(acc, part, gridDim.x);
}

// mul!(res, A::MultiphaseProjector, x), src/relaxation.jl:92-107 (first pass): T = interface
// gradient of x per unit area, over compact rows (x and T compact)
__global__ __launch_bounds__(B) void k_mph_grad(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, int m,
                                                const int *__restrict__ list, const int *__restrict__ inv,
                                                const double *__restrict__ x, double *__restrict__ T,
                                                const mph_scalars *S)
{
    if (S && S->done) return;
    const int c = blockIdx.x * B + threadIdx.x;
    if (c >= m) return;
    const int i = list[c];
    const double x1 = F.x[i], x2 = F.x[n + i];
    const int ph = F.phase[i];
    double t1 = 0.0, t2 = 0.0;
    for (int k = 0; k < M.ne[i]; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0 || F.phase[j] == ph) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        const int cj = inv[j];
        if (cj < 0) continue;  // a one-sided neighbour relation (degenerate mesh) couples nothing
        const double w = lrr * (x[c] - x[cj]);
        t1 -= w * (m1 - x1);
        t2 -= w * (m2 - x2);
    }
    const double A = F.area[i];
    T[c] = t1 / A;
    T[m + c] = t2 / A;
}

// second pass (src/relaxation.jl:108-124): y = A x from T; partial x.y
__global__ __launch_bounds__(B) void k_mph_div(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, int m,
                                               const int *__restrict__ list, const int *__restrict__ inv,
                                               const double *__restrict__ T, const double *__restrict__ x,
                                               double *__restrict__ y, double *__restrict__ part,
                                               const mph_scalars *S)
{
    if (S && S->done) return;
    double acc[1] = {0.0};
    for (int c = blockIdx.x * B + threadIdx.x; c < m; c += gridDim.x * B) {
        const int i = list[c];
        const double x1 = F.x[i], x2 = F.x[n + i];
        const int ph = F.phase[i];
        const double Ti1 = T[c], Ti2 = T[m + c];
        double ri = 0.0;
        for (int k = 0; k < M.ne[i]; k++) {
            const int j = LV_EL(M, k, i);
            if (j < 0 || F.phase[j] == ph) continue;
            const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i),
                         e2y = LV_EV(M, 3, k, i);
            double y1, y2;
            lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
            const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
            const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
            const double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
            const int cj = inv[j];
            if (cj < 0) continue;
            const double Tj1 = T[cj], Tj2 = T[m + cj];
            const double a = (Ti1 - Tj1) * (m1 - z1) + (Ti2 - Tj2) * (m2 - z2);
            const double cc = (Ti1 + Tj1) * (x1 - y1) + (Ti2 + Tj2) * (x2 - y2);
            ri -= lrr * (a - 0.5 * cc);
        }
        y[c] = ri;
        if (part) acc[0] += x[c] * ri;
    }
    if (part) block_partials<1>(acc, part, gridDim.x);
}

// CG start: x = 0, r = p = b; |b|^2 -> threshold (atol + rtol |b|)^2
__global__ __launch_bounds__(B) void k_mph_init(int m, const double *__restrict__ b, double *__restrict__ x,
                                                double *__restrict__ r, double *__restrict__ p,
                                                const double *__restrict__ part, int nb, double atol, double rtol,
                                                int itmax, mph_scalars *S)
{
    double o[1];
    reduce_all<1>(part, nb, o);
    for (int c = blockIdx.x * B + threadIdx.x; c < m; c += gridDim.x * B) {
        x[c] = 0.0;
        r[c] = b[c];
        p[c] = b[c];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double thr = atol + rtol * sqrt(o[0]);
        S->rr[0] = o[0];
        S->rr0 = o[0];
        S->thr2 = thr * thr;
        S->iters = 0;
        S->itmax = itmax;
        S->done = (o[0] <= thr * thr) ? 1 : 0;
    }
}

// alpha = rr / p.q; x += alpha p; r -= alpha q; partial |r|^2
__global__ __launch_bounds__(B) void k_mph_update(int m, double *__restrict__ x, double *__restrict__ r,
                                                  const double *__restrict__ p, const double *__restrict__ q,
                                                  const double *__restrict__ Ppq, double *__restrict__ Prr, int nb,
                                                  int par, mph_scalars *S)
{
    if (S->done) return;
    double o[1];
    reduce_all<1>(Ppq, nb, o);
    const double alpha = (o[0] > 0.0) ? S->rr[par] / o[0] : 0.0;
    double acc[1] = {0.0};
    for (int c = blockIdx.x * B + threadIdx.x; c < m; c += gridDim.x * B) {
        x[c] += alpha * p[c];
        const double rc = r[c] - alpha * q[c];
        r[c] = rc;
        acc[0] += rc * rc;
    }
    block_partials<1>(acc, Prr, gridDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) S->pq = o[0];
}

// beta = rr_new / rr; convergence; p = r + beta p (iteration k, par = k & 1)
__global__ __launch_bounds__(B) void k_mph_dir(int m, const double *__restrict__ r, double *__restrict__ p,
                                               const double *__restrict__ Prr, int nb, int k, int par,
                                               mph_scalars *S)
{
    if (S->done) return;
    double o[1];
    reduce_all<1>(Prr, nb, o);
    const int it = k + 1;
    const bool conv = o[0] <= S->thr2;
    const bool stop = conv || it >= S->itmax || !(S->pq > 0.0) || !(o[0] == o[0]);
    const double beta = o[0] / S->rr[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->iters = it;
        S->rr[par ^ 1] = o[0];
        if (stop) S->done = conv ? 1 : 2;
    }
    if (stop) return;
    for (int c = blockIdx.x * B + threadIdx.x; c < m; c += gridDim.x * B) p[c] = r[c] + beta * p[c];
}

// the dv update of multiphase_projection!, src/relaxation.jl:193-204
__global__ __launch_bounds__(B) void k_mph_apply_dv(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, int m,
                                                    const int *__restrict__ list, const int *__restrict__ inv,
                                                    const double *__restrict__ res, double quality_threshold)
{
    const int c = blockIdx.x * B + threadIdx.x;
    if (c >= m) return;
    const int i = list[c];
    if (F.quality[i] < quality_threshold) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const int ph = F.phase[i];
    const double A = F.area[i];
    double dv1 = F.dv[i], dv2 = F.dv[n + i];
    for (int k = 0; k < M.ne[i]; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0 || F.phase[j] == ph) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        const int cj = inv[j];
        if (cj < 0) continue;
        const double w = lrr * (res[c] - res[cj]);
        dv1 += (w * (m1 - x1)) / A;
        dv2 += (w * (m2 - x2)) / A;
    }
    F.dv[i] = dv1;
    F.dv[n + i] = dv2;
}

// full-size vectors <-> compact
__global__ __launch_bounds__(B) void k_mph_gather(int m, const int *__restrict__ list, const double *__restrict__ a,
                                                  double *__restrict__ c)
{
    const int k = blockIdx.x * B + threadIdx.x;
    if (k < m) c[k] = a[list[k]];
}
__global__ __launch_bounds__(B) void k_mph_scatter(int n, const int *__restrict__ inv, const double *__restrict__ c,
                                                   double *__restrict__ a)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) a[i] = (inv[i] >= 0) ? c[inv[i]] : 0.0;
}

}  // namespace

#define MPH_CHECK(expr)                  \
    do {                                 \
        hipError_t _e = (expr);          \
        if (_e != hipSuccess) return _e; \
    } while (0)

// Work space of the multiphase solve (device, sized for cap particles) -------------------------------
struct lv_mph_work {
    int *flag, *pos, *list, *inv;
    void *scan_tmp;
    double *b, *x, *r, *p, *q, *T, *part;
    mph_scalars *S;
    int cap;
};

hipError_t lv_mph_alloc(lv_mph_work **out, int cap)
{
    lv_mph_work *w = new lv_mph_work();
    const size_t c = (size_t)(cap > 0 ? cap : 1);
    w->cap = cap;
    MPH_CHECK(hipMalloc(&w->flag, sizeof(int) * (c + 1)));
    MPH_CHECK(hipMalloc(&w->pos, sizeof(int) * (c + 1)));
    MPH_CHECK(hipMalloc(&w->list, sizeof(int) * c));
    MPH_CHECK(hipMalloc(&w->inv, sizeof(int) * c));
    MPH_CHECK(hipMalloc(&w->scan_tmp, lv_scan_temp_bytes((int64_t)c + 1)));
    double **vec[] = {&w->b, &w->x, &w->r, &w->p, &w->q};
    for (double **v : vec) MPH_CHECK(hipMalloc(v, sizeof(double) * c));
    MPH_CHECK(hipMalloc(&w->T, sizeof(double) * 2 * c));
    MPH_CHECK(hipMalloc(&w->part, sizeof(double) * 2 * LV_NBMAX));
    MPH_CHECK(hipMalloc(&w->S, sizeof(mph_scalars)));
    *out = w;
    return hipSuccess;
}

void lv_mph_free(lv_mph_work *w)
{
    if (!w) return;
    void *ptrs[] = {w->flag, w->pos, w->list, w->inv, w->scan_tmp, w->b, w->x, w->r, w->p, w->q, w->T, w->part, w->S};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete w;
}

// the interface set; returns its size m (host)
static hipError_t mph_compact(lv_mph_work *w, lv_fields F, lv_mesh_view M, int n, int *m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mph_flag, g1(n), dim3(B), 0, s, F, M, n, w->flag);
    MPH_CHECK(hipMemsetAsync(w->flag + n, 0, sizeof(int), s));
    MPH_CHECK(lv_exclusive_scan_i32(w->flag, w->pos, (int64_t)n + 1, nullptr, w->scan_tmp, s));
    hipLaunchKernelGGL(k_mph_compact, g1(n), dim3(B), 0, s, n, (const int *)w->flag, (const int *)w->pos, w->list,
                       w->inv);
    MPH_CHECK(hipMemcpyAsync(m, w->pos + n, sizeof(int), hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

// multiphase_projection!: the right-hand side, CG to |r| <= atol + rtol |b|, the dv update.
// *iters = CG iterations (-1 when itmax stopped it, as the reference's unconverged MINRES, whose
// result is applied all the same); *relres = |r| / |b| at the end (0 without an interface).
hipError_t lv_mph_projection(lv_mph_work *w, const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n,
                             double quality_threshold, double atol, double rtol, int itmax, int *iters, double *relres,
                             hipStream_t s)
{
    int m = 0;
    MPH_CHECK(mph_compact(w, F, M, n, &m, s));
    *iters = 0;
    *relres = 0.0;
    if (m == 0) return hipSuccess;
    const int nb = lv_nblocks(m);
    double *Ppq = w->part, *Prr = w->part + LV_NBMAX;
    hipLaunchKernelGGL(k_mph_rhs, dim3(nb), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list, w->b, Prr);
    hipLaunchKernelGGL(k_mph_init, dim3(nb), dim3(B), 0, s, m, (const double *)w->b, w->x, w->r, w->p,
                       (const double *)Prr, nb, atol, rtol, itmax, w->S);
    mph_scalars h;
    for (int k0 = 0;; k0 += 16) {
        for (int k = k0; k < k0 + 16; k++) {
            hipLaunchKernelGGL(k_mph_grad, g1(m), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list,
                               (const int *)w->inv, (const double *)w->p, w->T, (const mph_scalars *)w->S);
            hipLaunchKernelGGL(k_mph_div, dim3(nb), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list,
                               (const int *)w->inv, (const double *)w->T, (const double *)w->p, w->q, Ppq,
                               (const mph_scalars *)w->S);
            hipLaunchKernelGGL(k_mph_update, dim3(nb), dim3(B), 0, s, m, w->x, w->r, (const double *)w->p,
                               (const double *)w->q, (const double *)Ppq, Prr, nb, k & 1, w->S);
            hipLaunchKernelGGL(k_mph_dir, dim3(nb), dim3(B), 0, s, m, (const double *)w->r, w->p,
                               (const double *)Prr, nb, k, k & 1, w->S);
        }
        MPH_CHECK(hipMemcpyAsync(&h, w->S, sizeof(h), hipMemcpyDeviceToHost, s));
        MPH_CHECK(hipStreamSynchronize(s));
        if (h.done) break;
    }
    hipLaunchKernelGGL(k_mph_apply_dv, g1(m), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list, (const int *)w->inv,
                       (const double *)w->x, quality_threshold);
    *iters = (h.done == 1) ? h.iters : -1;
    const double rr = h.rr[h.iters & 1];
    *relres = (h.rr0 > 0.0) ? sqrt(rr / h.rr0) : 0.0;
    return hipGetLastError();
}

// the operator and right-hand side on full-size host-layout vectors (device pointers), for tests
hipError_t lv_mph_apply(lv_mph_work *w, const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, const double *x,
                        double *y, hipStream_t s)
{
    int m = 0;
    MPH_CHECK(mph_compact(w, F, M, n, &m, s));
    if (m > 0) {
        hipLaunchKernelGGL(k_mph_gather, g1(m), dim3(B), 0, s, m, (const int *)w->list, x, w->p);
        hipLaunchKernelGGL(k_mph_grad, g1(m), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list, (const int *)w->inv,
                           (const double *)w->p, w->T, (const mph_scalars *)nullptr);
        hipLaunchKernelGGL(k_mph_div, dim3(lv_nblocks(m)), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list,
                           (const int *)w->inv, (const double *)w->T, (const double *)w->p, w->q, (double *)nullptr,
                           (const mph_scalars *)nullptr);
    }
    hipLaunchKernelGGL(k_mph_scatter, g1(n), dim3(B), 0, s, n, (const int *)w->inv, (const double *)w->q, y);
    return hipGetLastError();
}

hipError_t lv_mph_rhs(lv_mph_work *w, const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double *b,
                      hipStream_t s)
{
    int m = 0;
    MPH_CHECK(mph_compact(w, F, M, n, &m, s));
    if (m > 0)
        hipLaunchKernelGGL(k_mph_rhs, dim3(lv_nblocks(m)), dim3(B), 0, s, g, F, M, n, m, (const int *)w->list, w->b,
                           w->part);
    hipLaunchKernelGGL(k_mph_scatter, g1(n), dim3(B), 0, s, n, (const int *)w->inv, (const double *)w->b, b);
    return hipGetLastError();
}
