// lv_mls.hip -- nearest_polygon (src/voronoigrid.jl:154-178), movingls with the Gaussian kernel
// (src/movingls.jl:41-96) and point_value (:140-145), one thread per query or per polygon.
//
// Both walks follow the reference over the cell list of the last remesh: magic-path nodes in
// order, each cell's labels in their sorted order (periodic images included, as the reference's
// cells hold them), each candidate taken at its own seed position q.x.
//   * nearest_polygon measures |q.x - x|^2 without a minimum image and stops at the first node
//     whose rr exceeds the best distance found;
//   * movingls takes y = p.x + get_arrow(q.x, p.x) and stops at the first node with rr > h (the
//     reference compares the squared node bound with h itself).  Only cells with rr < h^2 can hold
//     a candidate with |y - p.x| < h, so the walk is exact while h^2 <= rr_max, i.e. h <= r_max,
//     which the host checks.
// The sums a = sum w (f(q) - f(p)) dx and R = sum w dx dx^T accumulate in that order in FP64; the
// 2x2 system is solved by Cramer's rule, the 5x5 and 9x9 ones by Gaussian elimination with
// partial pivoting.
#include "lv_internal.h"

namespace {

constexpr int B = 256;

// nearest_polygon: the index of the nearest seed, -1 if the path prefix ends before the search does
__device__ int nearest(const lv_grid_params &g, const double *__restrict__ x, int n, const int *__restrict__ cell_start,
                       const int *__restrict__ cell_labels, const int *__restrict__ path_i,
                       const double *__restrict__ path_rr, double q1, double q2)
{
    int k1, k2;
    lv_findkey(g, q1, q2, k1, k2);
    if (k1 < 0) return -1;
    double best = __builtin_inf();
    int ib = -1;
    int k = 0;
    for (; k < g.npath; k++) {
        if (path_rr[k] > best) break;
        const int c1 = k1 + path_i[2 * k], c2 = k2 + path_i[2 * k + 1];
        if (c1 < 1 || c1 > g.n1 || c2 < 1 || c2 > g.n2) continue;
        const int c = (c1 - 1) + (c2 - 1) * g.n1;
        for (int s = cell_start[c]; s < cell_start[c + 1]; s++) {
            const int j = cell_labels[s];
            const double a = x[j] - q1, b = x[n + j] - q2;
            const double rr = a * a + b * b;
            if (rr < best) {
                best = rr;
                ib = j;
            }
        }
    }
    if (k == g.npath && g.has_next && !(g.next_rr > best)) return -1;  // the walk would go on
    return ib;
}

template <int N>
__device__ __forceinline__ void power_vector(double d1, double d2, double (&p)[N])
{
    p[0] = d1;
    p[1] = d2;
    if (N >= 5) {
        p[2] = d1 * d1;
        p[3] = d1 * d2;
        p[4] = d2 * d2;
    }
    if (N >= 9) {
        p[5] = (d1 * d1) * d1;
        p[6] = (d1 * d1) * d2;
        p[7] = (d1 * d2) * d2;
        p[8] = (d2 * d2) * d2;
    }
}

// gaussian_kernel(h, r) = exp(-0.5 (r/h)^2) / sqrt(2 pi h^2), src/movingls.jl:22-24
__device__ __forceinline__ double gaussian(double h, double r)
{
    const double q = r / h;
    return exp(-0.5 * (q * q)) / sqrt((2.0 * 3.141592653589793) * (h * h));
}

// movingls(T, grid, p, fun; h) for polygon i and the scalar field f: the N Taylor coefficients
template <int N>
__device__ void movingls_at(const lv_grid_params &g, const double *__restrict__ x, int n,
                            const int *__restrict__ cell_start, const int *__restrict__ cell_labels,
                            const int *__restrict__ path_i, const double *__restrict__ path_rr,
                            const double *__restrict__ f, double h, int i, double (&a)[N])
{
    const double x1 = x[i], x2 = x[n + i];
    const double fp = f[i];
    double R[N][N];
#pragma unroll
    for (int r = 0; r < N; r++) {
        a[r] = 0.0;
#pragma unroll
        for (int c = 0; c < N; c++) R[r][c] = 0.0;
    }
    int k1, k2;
    lv_findkey(g, x1, x2, k1, k2);
    for (int k = 0; k < g.npath; k++) {
        if (path_rr[k] > h) break;
        const int c1 = k1 + path_i[2 * k], c2 = k2 + path_i[2 * k + 1];
        if (c1 < 1 || c1 > g.n1 || c2 < 1 || c2 > g.n2) continue;
        const int c = (c1 - 1) + (c2 - 1) * g.n1;
        for (int s = cell_start[c]; s < cell_start[c + 1]; s++) {
            const int j = cell_labels[s];
            double y1, y2;
            lv_arrow_target(g, x[j], x[n + j], x1, x2, y1, y2);
            if (x1 == y1 && x2 == y2) continue;
            const double e1 = x1 - y1, e2 = x2 - y2;
            const double r = sqrt(e1 * e1 + e2 * e2);
            if (!(r < h)) continue;
            const double w = gaussian(h, r);
            const double df = f[j] - fp;
            double dx[N];
            power_vector<N>(y1 - x1, y2 - x2, dx);
#pragma unroll
            for (int q = 0; q < N; q++) a[q] += (w * df) * dx[q];
#pragma unroll
            for (int q = 0; q < N; q++)
#pragma unroll
                for (int t = 0; t < N; t++) R[q][t] += (w * dx[q]) * dx[t];
        }
    }
    if (N == 2) {
        const double det = R[0][0] * R[1][1] - R[0][1] * R[1][0];
        const double b0 = a[0], b1 = a[1];
        a[0] = (R[1][1] * b0 - R[0][1] * b1) / det;
        a[1] = (R[0][0] * b1 - R[1][0] * b0) / det;
        return;
    }
    // Gaussian elimination with partial pivoting
#pragma unroll
    for (int p = 0; p < N; p++) {
        int m = p;
        double best = fabs(R[p][p]);
#pragma unroll
        for (int r = p + 1; r < N; r++)
            if (fabs(R[r][p]) > best) {
                best = fabs(R[r][p]);
                m = r;
            }
#pragma unroll
        for (int r = p + 1; r < N; r++)
            if (r == m) {
#pragma unroll
                for (int c = 0; c < N; c++) {
                    const double t = R[p][c];
                    R[p][c] = R[r][c];
                    R[r][c] = t;
                }
                const double t = a[p];
                a[p] = a[r];
                a[r] = t;
            }
#pragma unroll
        for (int r = p + 1; r < N; r++) {
            const double l = R[r][p] / R[p][p];
#pragma unroll
            for (int c = p; c < N; c++) R[r][c] -= l * R[p][c];
            a[r] -= l * a[p];
        }
    }
#pragma unroll
    for (int p = N - 1; p >= 0; p--) {
        double s = a[p];
#pragma unroll
        for (int c = p + 1; c < N; c++) s -= R[p][c] * a[c];
        a[p] = s / R[p][p];
    }
}

struct mls_args {
    lv_grid_params g;
    const double *x;
    int n;
    const int *cell_start, *cell_labels, *path_i;
    const double *path_rr;
};

template <int N>
__global__ __launch_bounds__(B) void k_movingls(mls_args A, const double *__restrict__ f, double h,
                                                double *__restrict__ out)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= A.n) return;
    double a[N];
    movingls_at<N>(A.g, A.x, A.n, A.cell_start, A.cell_labels, A.path_i, A.path_rr, f, h, i, a);
#pragma unroll
    for (int q = 0; q < N; q++) out[(size_t)i * N + q] = a[q];
}

__global__ __launch_bounds__(B) void k_nearest(mls_args A, int m, const double *__restrict__ q, int *__restrict__ idx)
{
    const int t = blockIdx.x * B + threadIdx.x;
    if (t >= m) return;
    idx[t] = nearest(A.g, A.x, A.n, A.cell_start, A.cell_labels, A.path_i, A.path_rr, q[t], q[m + t]);
}

// point_value(grid, x, fun) = fun(p) + dot(L, x - p.x) with p the nearest polygon and L its linear
// expansion at h = grid.h; NaN where the nearest polygon is not found
__global__ __launch_bounds__(B) void k_point_value(mls_args A, int m, const double *__restrict__ q,
                                                   const double *__restrict__ f, double h, double *__restrict__ out)
{
    const int t = blockIdx.x * B + threadIdx.x;
    if (t >= m) return;
    const double q1 = q[t], q2 = q[m + t];
    const int p = nearest(A.g, A.x, A.n, A.cell_start, A.cell_labels, A.path_i, A.path_rr, q1, q2);
    if (p < 0) {
        out[t] = __builtin_nan("");
        return;
    }
    double L[2];
    movingls_at<2>(A.g, A.x, A.n, A.cell_start, A.cell_labels, A.path_i, A.path_rr, f, h, p, L);
    out[t] = f[p] + (L[0] * (q1 - A.x[p]) + L[1] * (q2 - A.x[A.n + p]));
}

inline dim3 grid_for(int m) { return dim3((unsigned)((m + B - 1) / B > 0 ? (m + B - 1) / B : 1)); }

}  // namespace

hipError_t lv_launch_movingls(const lv_grid_params &g, const double *x, int n, const int *cell_start,
                              const int *cell_labels, const int *path_i, const double *path_rr, int order,
                              const double *f, double h, double *out, hipStream_t s)
{
    const mls_args A{g, x, n, cell_start, cell_labels, path_i, path_rr};
    if (n <= 0) return hipSuccess;
    if (order == 1) hipLaunchKernelGGL(k_movingls<2>, grid_for(n), dim3(B), 0, s, A, f, h, out);
    else if (order == 2) hipLaunchKernelGGL(k_movingls<5>, grid_for(n), dim3(B), 0, s, A, f, h, out);
    else if (order == 3) hipLaunchKernelGGL(k_movingls<9>, grid_for(n), dim3(B), 0, s, A, f, h, out);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t lv_launch_nearest(const lv_grid_params &g, const double *x, int n, const int *cell_start,
                             const int *cell_labels, const int *path_i, const double *path_rr, int m,
                             const double *q, int *idx, hipStream_t s)
{
    const mls_args A{g, x, n, cell_start, cell_labels, path_i, path_rr};
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nearest, grid_for(m), dim3(B), 0, s, A, m, q, idx);
    return hipGetLastError();
}

hipError_t lv_launch_point_value(const lv_grid_params &g, const double *x, int n, const int *cell_start,
                                 const int *cell_labels, const int *path_i, const double *path_rr, int m,
                                 const double *q, const double *f, double h, double *out, hipStream_t s)
{
    const mls_args A{g, x, n, cell_start, cell_labels, path_i, path_rr};
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_point_value, grid_for(m), dim3(B), 0, s, A, m, q, f, h, out);
    return hipGetLastError();
}
