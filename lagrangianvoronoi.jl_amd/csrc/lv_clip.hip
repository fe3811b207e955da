// lv_clip.hip -- Voronoi polygon construction by half-plane clipping: the cutting loop of remesh!
// (src/voronoigrid.jl:53-81, 89-108), voronoicut!(p, y, label) (src/polygon.jl:51-97), influence_rr
// (:101-107), sort_edges! (src/IO.jl:35-48) and the area/centroid read-outs (src/polygon.jl:114-122,
// 201-219), fused into one kernel.
//
// One thread owns one seed.  Its polygon lives in LDS as five slot-major planes (v1x, v1y, v2x, v2y,
// label), so lane t of the block addresses bank-conflict-free 8-byte words.  The first launch gives
// each polygon CAP_SMALL slots (18 KB per 64-seed block, twice the waves per CU of LV_MAXE slots); a
// seed whose polygon outgrows them (0.4% at a 2048^2 jittered lattice, where no polygon passes 10
// edges while it is clipped) is queued and clipped again from the start by a second launch with
// LV_MAXE slots, so the result is the same either way.  The edge
// list keeps the reference's storage semantics exactly: cut edges are rewritten in place, deleted
// edges are replaced by the last edge (FastVector deleteat!, src/fastvector.jl:44-50) and re-tested,
// and the new edge is pushed last.  With the same candidate order (magic path, then each cell's
// labels ascending) and IEEE FP64 without contraction, every vertex, label and edge position equals
// the reference's bit for bit.  Candidates come from the sorted cell list, which carries each
// label's seed position next to it.
#include "lv_internal.h"

namespace {

constexpr int CB = 64;          // seeds per block (one wavefront)
constexpr int CAP_SMALL = 8;    // edge capacity of the first launch's LDS polygon
constexpr int RETRY_BLOCKS = 1024;  // blocks of the second launch (grid-stride over the queue)
constexpr double SIGNUM_EPS = 2.0 * 2.220446049250313e-16;  // 2*eps(Float64), src/polygon.jl:2

template <int CAP>
struct lds_poly {
    double v1x[CAP][CB], v1y[CAP][CB], v2x[CAP][CB], v2y[CAP][CB];
    int lab[CAP][CB];
};

__device__ __forceinline__ int signum(double v)
{
    if (v < -SIGNUM_EPS) return -1;
    if (v > SIGNUM_EPS) return 1;
    return 0;
}

template <int CAP>
__device__ __forceinline__ double influence_rr(const lds_poly<CAP> &P, int t, int len, double x1, double x2)
{
    double rr = 0.0;
    for (int k = 0; k < len; k++) {
        double a = P.v1x[k][t] - x1, b = P.v1y[k][t] - x2;
        double r = 4.0 * (a * a + b * b);
        if (r > rr) rr = r;
    }
    return rr;
}

// voronoicut!(p, y, label): 1 if an edge was added, 0 if not, -1 on overflow
template <int CAP>
__device__ __forceinline__ int cut(lds_poly<CAP> &P, int t, int &len, double x1, double x2, double y1, double y2,
                                   int label)
{
    const double d1 = y1 - x1, d2 = y2 - x2;
    const double m1 = 0.5 * (y1 + x1), m2 = 0.5 * (y2 + x2);
    const double c = d1 * m1 + d2 * m2;
    const double qnan = __builtin_nan("");
    double X1 = qnan, X2 = qnan, Y1 = qnan, Y2 = qnan;
    int i = 0;
    while (i < len) {
        const double e1x = P.v1x[i][t], e1y = P.v1y[i][t], e2x = P.v2x[i][t], e2y = P.v2y[i][t];
        const double f1 = (d1 * e1x + d2 * e1y) - c;
        const double f2 = (d1 * e2x + d2 * e2y) - c;
        const int s1 = signum(f1), s2 = signum(f2);
        const int s12 = s1 + s2;
        if ((0 <= s12 && s12 <= 1) && ((s1 | s2) != 0)) {
            Y1 = X1;
            Y2 = X2;
            const double tt = 1.0 / (f1 - f2);
            X1 = tt * (f1 * e2x - f2 * e1x);
            X2 = tt * (f1 * e2y - f2 * e1y);
            if (s1 == 0) { X1 = e1x; X2 = e1y; }
            if (s2 == 0) { X1 = e2x; X2 = e2y; }
            if (s1 == 1) { P.v1x[i][t] = X1; P.v1y[i][t] = X2; }
            else { P.v2x[i][t] = X1; P.v2y[i][t] = X2; }
        }
        if (1 <= s12) {
            const int last = len - 1;
            if (last > i) {
                P.v1x[i][t] = P.v1x[last][t];
                P.v1y[i][t] = P.v1y[last][t];
                P.v2x[i][t] = P.v2x[last][t];
                P.v2y[i][t] = P.v2y[last][t];
                P.lab[i][t] = P.lab[last][t];
            }
            len--;
        } else {
            i++;
        }
    }
    if (!(isnan(Y1) && isnan(Y2)) && (X1 != Y1 || X2 != Y2)) {
        if (len >= CAP) return -1;
        const double cr = (Y1 - X1) * (x2 - X2) - (Y2 - X2) * (x1 - X1);  // cross2(Y - X, x - X)
        if (cr > 0.0) {  // invert(e), src/IO.jl:5-7
            P.v1x[len][t] = Y1; P.v1y[len][t] = Y2; P.v2x[len][t] = X1; P.v2y[len][t] = X2;
        } else {
            P.v1x[len][t] = X1; P.v1y[len][t] = X2; P.v2x[len][t] = Y1; P.v2y[len][t] = Y2;
        }
        P.lab[len][t] = label;
        len++;
        return 1;
    }
    return 0;
}

template <int CAP>
__device__ __forceinline__ void swap_slots(lds_poly<CAP> &P, int t, int a, int b)
{
    double tx;
    tx = P.v1x[a][t]; P.v1x[a][t] = P.v1x[b][t]; P.v1x[b][t] = tx;
    tx = P.v1y[a][t]; P.v1y[a][t] = P.v1y[b][t]; P.v1y[b][t] = tx;
    tx = P.v2x[a][t]; P.v2x[a][t] = P.v2x[b][t]; P.v2x[b][t] = tx;
    tx = P.v2y[a][t]; P.v2y[a][t] = P.v2y[b][t]; P.v2y[b][t] = tx;
    int tl = P.lab[a][t]; P.lab[a][t] = P.lab[b][t]; P.lab[b][t] = tl;
}

// one seed's polygon, clipped in P's LDS slots; false when it outgrew them and the kernel defers it
template <int CAP, bool DEFER>
__device__ bool clip_seed(const lv_grid_params &g, int n, int i, const double *__restrict__ x,
                          const int *__restrict__ cell_start, const int *__restrict__ cell_labels,
                          const double *__restrict__ cell_x, int m, const int *__restrict__ path_i,
                          const double *__restrict__ path_rr, int *__restrict__ ne, double *__restrict__ ev,
                          int *__restrict__ el, int ld, double *__restrict__ area, double *__restrict__ cent,
                          int *__restrict__ err, lds_poly<CAP> &P, int t)
{
    const double x1 = x[i], x2 = x[n + i];

    // reset!(p, cropping_rect), src/polygon.jl:37-47: DOWN(B->A), LEFT(A->D), UP(D->C), RIGHT(C->B)
    {
        const double Ax = g.cmin0, Ay = g.cmin1, Cx = g.cmax0, Cy = g.cmax1;
        const double Bx = Cx, By = Ay, Dx = Ax, Dy = Cy;
        P.v1x[0][t] = Bx; P.v1y[0][t] = By; P.v2x[0][t] = Ax; P.v2y[0][t] = Ay; P.lab[0][t] = -3;
        P.v1x[1][t] = Ax; P.v1y[1][t] = Ay; P.v2x[1][t] = Dx; P.v2y[1][t] = Dy; P.lab[1][t] = -4;
        P.v1x[2][t] = Dx; P.v1y[2][t] = Dy; P.v2x[2][t] = Cx; P.v2y[2][t] = Cy; P.lab[2][t] = -1;
        P.v1x[3][t] = Cx; P.v1y[3][t] = Cy; P.v2x[3][t] = Bx; P.v2y[3][t] = By; P.lab[3][t] = -2;
    }
    int len = 4;
    double prr = influence_rr(P, t, len, x1, x2);
    int k1, k2;
    lv_findkey(g, x1, x2, k1, k2);
    int status = LV_OK;
    for (int k = 0;; k++) {
        if (k == g.npath) {
            if (g.has_next && !(g.next_rr > prr)) status = LV_E_MESH_DESTROYED;  // src/voronoigrid.jl:63-65
            break;
        }
        const double rr = path_rr[k];
        if (rr > prr) break;
        const int c1 = k1 + path_i[2 * k], c2 = k2 + path_i[2 * k + 1];
        if (c1 < 1 || c1 > g.n1 || c2 < 1 || c2 > g.n2) continue;
        const int c = (c1 - 1) + (c2 - 1) * g.n1;
        const int s0 = cell_start[c], s1 = cell_start[c + 1];
        for (int s = s0; s < s1; s++) {
            const int j = cell_labels[s];
            double y1, y2;
            lv_arrow_target(g, cell_x[s], cell_x[m + s], x1, x2, y1, y2);
            const double a = x1 - y1, b = x2 - y2;
            if ((x1 == y1 && x2 == y2) || (a * a + b * b > prr)) continue;
            const int r = cut(P, t, len, x1, x2, y1, y2, j);
            if (r < 0) {
                status = LV_E_POLY_OVERFLOW;
                break;
            }
            if (r) prr = influence_rr(P, t, len, x1, x2);
        }
        if (status != LV_OK) break;
    }
    if (DEFER && status == LV_E_POLY_OVERFLOW) return false;
    if (status != LV_OK) {
        atomicCAS(err, 0, status);
        ne[i] = 0;
        area[i] = 0.0;
        cent[i] = 0.0;
        cent[n + i] = 0.0;
        return true;
    }
    // sort_edges!, src/IO.jl:35-48
    for (int a = 0; a < len; a++) {
        const double lx = P.v2x[a][t], ly = P.v2y[a][t];
        for (int b = a + 1; b < len; b++) {
            if (lx == P.v1x[b][t] && ly == P.v1y[b][t]) {
                if (b != a + 1) swap_slots(P, t, a + 1, b);
                break;
            }
        }
    }
    // area(p) and centroid(p) in edge order, src/polygon.jl:114-122, 201-219
    double A = 0.0, Ac = 0.0, cx = 0.0, cy = 0.0;
    for (int k = 0; k < len; k++) {
        const double v1x = P.v1x[k][t], v1y = P.v1y[k][t], v2x = P.v2x[k][t], v2y = P.v2y[k][t];
        const double p1 = v1x - x1, p2 = v1y - x2, q1 = v2x - x1, q2 = v2y - x2;
        const double h = 0.5 * fabs(p1 * q2 - p2 * q1);
        A += h;
        Ac += h;
        cx += (h * ((x1 + v1x) + v2x)) / 3.0;
        cy += (h * ((x2 + v1y) + v2y)) / 3.0;
        ev[(size_t)(0 * LV_MAXE + k) * ld + i] = v1x;
        ev[(size_t)(1 * LV_MAXE + k) * ld + i] = v1y;
        ev[(size_t)(2 * LV_MAXE + k) * ld + i] = v2x;
        ev[(size_t)(3 * LV_MAXE + k) * ld + i] = v2y;
        el[(size_t)k * ld + i] = P.lab[k][t];
    }
    ne[i] = len;
    area[i] = A;
    cent[i] = cx / Ac;
    cent[n + i] = cy / Ac;
    return true;
}

// The seed block a workgroup clips.  Workgroups are dispatched round-robin over the 8 XCDs, so
// workgroup b runs on XCD b mod 8; mapping it to block b would give each XCD every 8th block, and on a
// lattice whose rows span a multiple of 8 blocks that is the same few column strips of the domain in
// every row.  The clipping work per seed varies across the domain (it depends on where the seeds sit
// in their cell-list cells), so XCDs with the costly strips ran up to 28% longer while the others
// idled for the last third of the launch.  Rotating each group of 8 workgroups by a hash of its group
// index (the top 3 bits of a multiplicative hash, so that no period of the domain lines up with it)
// gives every XCD an even share of every column; the last, partial group keeps the identity.
__device__ __forceinline__ int clip_block(int b, int nb)
{
    const unsigned q = (unsigned)b >> 3;
    if (8 * ((int)q + 1) > nb) return b;
    const unsigned rot = (q * 2654435761u) >> 29;
    return (int)((q << 3) | ((b + rot) & 7));
}

// every seed with CAP_SMALL slots; the ones that outgrow them go to the queue retry[1..retry[0]]
__global__ __launch_bounds__(CB) void k_clip(lv_grid_params g, int n, const double *__restrict__ x,
                                             const int *__restrict__ cell_start, const int *__restrict__ cell_labels,
                                             const double *__restrict__ cell_x, int m, const int *__restrict__ path_i,
                                             const double *__restrict__ path_rr, int *__restrict__ ne,
                                             double *__restrict__ ev, int *__restrict__ el, int ld,
                                             double *__restrict__ area, double *__restrict__ cent,
                                             int *__restrict__ err, int *__restrict__ retry)
{
    __shared__ lds_poly<CAP_SMALL> P;
    const int t = threadIdx.x;
    const int i = clip_block(blockIdx.x, gridDim.x) * CB + t;
    if (i >= n) return;  // no block-wide barriers below
    if (!clip_seed<CAP_SMALL, true>(g, n, i, x, cell_start, cell_labels, cell_x, m, path_i, path_rr, ne, ev, el, ld,
                                    area, cent, err, P, t))
        retry[1 + atomicAdd(&retry[0], 1)] = i;
}

// the queued seeds with LV_MAXE slots (grid-stride over the queue)
__global__ __launch_bounds__(CB) void k_clip_retry(lv_grid_params g, int n, const double *__restrict__ x,
                                                   const int *__restrict__ cell_start,
                                                   const int *__restrict__ cell_labels,
                                                   const double *__restrict__ cell_x, int m,
                                                   const int *__restrict__ path_i,
                                                   const double *__restrict__ path_rr, int *__restrict__ ne,
                                                   double *__restrict__ ev, int *__restrict__ el, int ld,
                                                   double *__restrict__ area, double *__restrict__ cent,
                                                   int *__restrict__ err, const int *__restrict__ retry)
{
    __shared__ lds_poly<LV_MAXE> P;
    const int t = threadIdx.x;
    const int cnt = retry[0];
    for (int q = blockIdx.x * CB + t; q < cnt; q += gridDim.x * CB)
        clip_seed<LV_MAXE, false>(g, n, retry[1 + q], x, cell_start, cell_labels, cell_x, m, path_i, path_rr, ne, ev,
                                  el, ld, area, cent, err, P, t);
}

}  // namespace

hipError_t lv_launch_clip(const lv_grid_params &g, int n, const double *x, const int *cell_start,
                          const int *cell_labels, const double *cell_x, const int *path_i, const double *path_rr,
                          int *ne, double *ev, int *el, int ld, double *area, double *cent, int *err, int *retry,
                          hipStream_t s, int m)
{
    if (n <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(retry, 0, sizeof(int), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_clip, dim3((n + CB - 1) / CB), dim3(CB), 0, s, g, n, x, cell_start, cell_labels, cell_x, m,
                       path_i, path_rr, ne, ev, el, ld, area, cent, err, retry);
    const int nb = std::min(RETRY_BLOCKS, (n + CB - 1) / CB);
    hipLaunchKernelGGL(k_clip_retry, dim3(nb), dim3(CB), 0, s, g, n, x, cell_start, cell_labels, cell_x, m, path_i,
                       path_rr, ne, ev, el, ld, area, cent, err, (const int *)retry);
    return hipGetLastError();
}
