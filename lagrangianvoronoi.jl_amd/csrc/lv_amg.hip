// lv_amg.hip -- multilevel preconditioner for the pressure system of find_pressure! (src/pressure.jl:89-224).
//
// No counterpart in the reference (it runs unpreconditioned MINRES, src/pressure.jl:219).  A
// one-ring preconditioner can shrink the condition number of the Voronoi Laplacian by a constant
// factor only; the smooth error modes need a coarse correction.
//
// Hierarchy (rebuilt for every assembled operator):
//   level 0: the particles and A itself (ELL, FP32);
//   level 1: the cells of a regular grid of width >= CELL_DR dr (2 dr: about 4 particles a cell)
//            over the boundary rectangle; every particle is aggregated into the cell holding it.
//            Cells of 4 dr (16 particles) coarsen faster but cost 20% more CG iterations on the
//            2048^2 benchmark steps (171 against 142 at 2 dr) for less work per cycle.  The Galerkin operator P^T A P of the
//            piecewise-constant P keeps the reference's difference form:
//              d_I = sum_{i in I} d_i,   c_IJ = sum_{i in I, j in J} c_ij,
//            stored as a 3x3 stencil (a coupling that would reach beyond the neighbouring cells,
//            only possible on strongly distorted meshes, is dropped from both sides, which keeps
//            the coarse operator positive semi-definite);
//   level l+1: 2 x 2 blocks of level-l cells, the same Galerkin rule (4 x 4 blocks coarsen too fast
//            for the V-cycle: on the 2048^2 benchmark system PCG needs ~60 iterations with them,
//            ~40 with 2 x 2);
//   the coarsest level (<= LV_AMG_DENSE_MAX cells) is inverted densely once per assembly.
// The levels of at most TAIL_MAX cells run inside one single-workgroup kernel per cycle (k_tail):
// they are too small to fill the device, and one launch replaces two per level.  Above them, levels
// 2 and up run in pairs, each pair as one down and one up kernel (k_st_down2, k_st_up2) that
// recompute a halo around each workgroup's tile instead of waiting for a second launch.
// Every sum is a gather in a fixed order (per-aggregate row lists are sorted), so the hierarchy
// and the preconditioner are deterministic.
//
// Cycle: symmetric V(1,1).  The fine smoother is the learned SPAI (M of lv_spai.hip, scaled so
// that the largest eigenvalue of M A is ~1.2, estimated by power iteration at set-up) or, without
// it, damped Jacobi (2/3 D^-1); the coarse smoothers are damped Jacobi.  The cycle is an SPD
// linear operator, so CG stays valid; it runs in FP32 inside the FP64 iterative refinement, where
// Jacobi-PCG remains the fallback and the FP64 residual remains the judge.
#include <string.h>

#include <algorithm>

#include "lv_amg.h"
#include "lv_ell.h"
#include "lv_reduce.h"

namespace {

constexpr int B = 256;
constexpr float OMEGA = 2.0f / 3.0f;  // damped Jacobi
constexpr float FINE_TARGET = 1.2f;   // lambda_max of the scaled learned smoother times A
constexpr int CF = 2;                 // coarse levels: CF x CF blocks of the finer level's cells
// Over-correction: the prolongated coarse correction is scaled by ALPHA_FINE into the fine level and
// by ALPHA_COARSE into a coarse level, except for corrections from the levels of at most TAIL_MAX
// cells (solved nearly exactly, where a scaling only overshoots).  The piecewise-constant
// prolongation leaves the coarse correction short in energy; the scaling keeps the cycle SPD for
// any factor >= 0 (I - a P B_c P^T A stays <= I in the A inner product when B_c is SPD) and cuts
// the PCG iterations on the 2048^2 benchmark system from 41 to ~34.
constexpr float ALPHA_FINE = 1.6f;
constexpr float ALPHA_COARSE = 1.4f;
constexpr int CF2 = CF * CF;
constexpr double CELL_DR = 2.0;  // level-1 cell width in units of dr
constexpr int RL = 4;            // lanes that restrict one level-1 cell (about 4 rows each)

// the sum of RL consecutive lanes by a fixed shuffle tree, in every one of them
__device__ __forceinline__ float sum_lanes(float s)
{
#pragma unroll
    for (int m = 1; m < RL; m <<= 1) s += __shfl_xor(s, m);
    return s;
}
__device__ __forceinline__ double sum_lanes(double s)
{
#pragma unroll
    for (int m = 1; m < RL; m <<= 1) s += __shfl_xor(s, m);
    return s;
}

inline dim3 g1(int64_t count) { return dim3((unsigned)std::max<int64_t>(1, (count + B - 1) / B)); }

// 3x3 stencil slots: k = 0..7 for the offsets (dx, dy), dx-major, skipping (0, 0)
__device__ __forceinline__ int slot_of(int dx, int dy)
{
    const int q = (dx + 1) * 3 + (dy + 1);
    return q > 4 ? q - 1 : q;
}
__device__ __forceinline__ void offset_of(int k, int &dx, int &dy)
{
    const int q = k >= 4 ? k + 1 : k;
    dx = q / 3 - 1;
    dy = q % 3 - 1;
}
// a cell offset on a periodic axis of K cells, folded into [-(K-1)/2, K/2]
__device__ __forceinline__ int fold(int d, int K, int per)
{
    if (!per) return d;
    d %= K;
    if (d > K / 2) d -= K;
    if (d < -((K - 1) / 2)) d += K;
    return d;
}
__device__ __forceinline__ int wrap(int i, int K, int per, bool &ok)
{
    if (per) {
        ok = true;
        return (i % K + K) % K;
    }
    ok = (i >= 0 && i < K);
    return i;
}

// ---- level 1: aggregation of the particles -----------------------------------------------------

// ctr[key] += 1 for every active lane, by runs: lanes of a wavefront with equal keys in consecutive
// lanes (the active lanes being a prefix) form a run, whose first lane adds the run's length in one
// atomic; each lane gets the old value plus its rank in the run, its own slot among the key's.
// Consecutive rows mostly share a level-1 cell, so this halves the atomics, which bound both
// kernels below.  Every lane of the wavefront calls it.
__device__ __forceinline__ int run_atomic_add(int *ctr, int key, bool active)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const unsigned long long am = __ballot(active);
    const bool pact = lane > 0 && ((am >> (lane - 1)) & 1ull);
    const bool head = active && (!pact || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (every lane at lane 63)
    const int hb = 63 - __clzll((long long)(heads & upto));
    const int h = active ? hb : lane;
    const unsigned long long above = heads & ~upto;
    const int nexth = above ? __ffsll((long long)above) - 1 : 64;
    const int nact = 64 - __clzll((long long)am);
    const int runlen = (nexth < nact ? nexth : nact) - lane;
    int old = 0;
    if (head) old = atomicAdd(&ctr[key], runlen);
    return __shfl(old, h) + (lane - h);
}

__global__ __launch_bounds__(B) void k_agg_key(int n, const double *__restrict__ x, double x0, double y0, double hx,
                                               double hy, int Kx, int Ky, int *__restrict__ agg,
                                               int *__restrict__ cnt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const bool act = i < n;
    int a = 0;
    if (act) {
        int ix = (int)floor((x[i] - x0) / hx), iy = (int)floor((x[n + i] - y0) / hy);
        ix = ix < 0 ? 0 : (ix >= Kx ? Kx - 1 : ix);
        iy = iy < 0 ? 0 : (iy >= Ky ? Ky - 1 : iy);
        a = ix * Ky + iy;
        agg[i] = a;
    }
    (void)run_atomic_add(cnt, a, act);
}

// (the order of a cell's rows in list is then fixed by k_agg_sort)
__global__ __launch_bounds__(B) void k_agg_fill(int n, const int *__restrict__ agg, const int *__restrict__ start,
                                                int *__restrict__ cursor, int *__restrict__ list)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const bool act = i < n;
    const int a = act ? agg[i] : 0;
    const int p = run_atomic_add(cursor, a, act);
    if (act) list[start[a] + p] = i;
}

// per cell: its rows sorted (a fixed order for every sum over a cell's rows, here and in the
// aggregate order)
__global__ __launch_bounds__(B) void k_agg_sort(int na, const int *__restrict__ start, int *__restrict__ list)
{
    const int I = blockIdx.x * B + threadIdx.x;
    if (I >= na) return;
    const int s0 = start[I], s1 = start[I + 1];
    for (int a = s0 + 1; a < s1; a++) {
        const int v = list[a];
        int b = a;
        while (b > s0 && list[b - 1] > v) {
            list[b] = list[b - 1];
            b--;
        }
        list[b] = v;
    }
}

// per cell, RL lanes: the Galerkin sums of d and of the couplings to the 8 neighbouring cells in
// FP64, lane q over the cell's sorted rows q, q + RL, ..., then the lanes' sums added by sum_lanes'
// fixed tree
__global__ __launch_bounds__(B) void k_coarse1(int na, int Kx, int Ky, int xper, int yper,
                                               const int *__restrict__ start, const int *__restrict__ list,
                                               const int *__restrict__ agg, int n, const int *__restrict__ deg,
                                               const int *__restrict__ ecol, const double *__restrict__ ev,
                                               const double *__restrict__ d64, float *__restrict__ dout,
                                               float *__restrict__ cout, float *__restrict__ dinv)
{
    const int t = blockIdx.x * B + threadIdx.x;
    const int I = t / RL, q = t % RL;
    const bool live = I < na;  // every lane takes part in the shuffles below
    const int Ic = live ? I : 0;
    const int s0 = start[Ic], s1 = live ? start[Ic + 1] : s0;
    const int Ix = Ic / Ky, Iy = Ic % Ky;
    double d = 0.0, c[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int a = s0 + q; a < s1; a += RL) {
        const int i = list[a];
        d += d64[i];
        const int dg = deg[i];
        for (int k = 0; k < dg; k++) {
            const size_t e = (size_t)k * n + i;
            const int J = agg[ecol[e]];
            if (J == Ic) continue;
            const int dx = fold(J / Ky - Ix, Kx, xper), dy = fold(J % Ky - Iy, Ky, yper);
            if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || (dx == 0 && dy == 0)) continue;
            c[slot_of(dx, dy)] += ev[e];
        }
    }
    d = sum_lanes(d);
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = sum_lanes(c[k]);
    if (!live || q != 0) return;
    if (s1 == s0) d = 1.0;  // an empty cell is a decoupled identity row
    double sum = d;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        cout[(size_t)k * na + I] = (float)c[k];
        sum += c[k];
    }
    dout[I] = (float)d;
    dinv[I] = sum > 0.0 ? (float)(1.0 / sum) : 1.0f;
}

// level l+1 from level l: CF x CF blocks of cells
__global__ __launch_bounds__(B) void k_coarsen(int Kxc, int Kyc, int Kxf, int Kyf, int xper, int yper,
                                               const float *__restrict__ df, const float *__restrict__ cf,
                                               float *__restrict__ dc, float *__restrict__ cc,
                                               float *__restrict__ dinv)
{
    const int P = blockIdx.x * B + threadIdx.x;
    const int nc = Kxc * Kyc, nf = Kxf * Kyf;
    if (P >= nc) return;
    const int PX = P / Kyc, PY = P % Kyc;
    double d = 0.0, c[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int cx1 = std::min(CF * PX + CF, Kxf), cy1 = std::min(CF * PY + CF, Kyf);
    for (int cx = CF * PX; cx < cx1; cx++)
        for (int cy = CF * PY; cy < cy1; cy++) {
            const int f = cx * Kyf + cy;
            d += df[f];
            for (int k = 0; k < 8; k++) {
                const float ck = cf[(size_t)k * nf + f];
                if (ck == 0.0f) continue;
                int ox, oy;
                offset_of(k, ox, oy);
                bool okx, oky;
                const int nx = wrap(cx + ox, Kxf, xper, okx), ny = wrap(cy + oy, Kyf, yper, oky);
                if (!okx || !oky) continue;
                const int dx = fold(nx / CF - PX, Kxc, xper), dy = fold(ny / CF - PY, Kyc, yper);
                if (dx == 0 && dy == 0) continue;
                c[slot_of(dx, dy)] += ck;
            }
        }
    double sum = d;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        cc[(size_t)k * nc + P] = (float)c[k];
        sum += c[k];
    }
    dc[P] = (float)d;
    dinv[P] = sum > 0.0 ? (float)(1.0 / sum) : 1.0f;
}

// the coarsest level as a dense FP64 matrix (plus sigma/nc everywhere for a singular system, which
// makes it definite without changing its action on mean-free vectors), inverted in place in LDS by
// Gauss-Jordan in one workgroup (an SPD matrix needs no pivoting); the inverse is stored
// column-major in FP32.  In LDS the rows have the fixed stride DM, so that an entry's row and column
// are a shift and a mask of its index, and the 1024 lanes take DM * DM / 1024 entries each.
constexpr int DM = LV_AMG_DENSE_MAX, DTB = 1024;
static_assert(DM == 64 && (DM * DM) % DTB == 0, "dense level: 64 x 64 entries over the workgroup");
__global__ __launch_bounds__(DTB) void k_dense_inverse(int Kx, int Ky, int xper, int yper, const float *__restrict__ d,
                                                       const float *__restrict__ c, int singular,
                                                       float *__restrict__ inv)
{
    const int nc = Kx * Ky;
    __shared__ double M[DM * DM];
    __shared__ double colp[DM], rowdiag[DM];
    for (int t = threadIdx.x; t < DM * DM; t += DTB) M[t] = 0.0;
    __syncthreads();
    if (threadIdx.x < nc) {  // row I by lane I
        const int I = threadIdx.x, Ix = I / Ky, Iy = I % Ky;
        double diag = d[I];
        for (int k = 0; k < 8; k++) {
            const double ck = c[(size_t)k * nc + I];
            if (ck == 0.0) continue;
            int ox, oy;
            offset_of(k, ox, oy);
            bool okx, oky;
            const int jx = wrap(Ix + ox, Kx, xper, okx), jy = wrap(Iy + oy, Ky, yper, oky);
            if (!okx || !oky) continue;
            diag += ck;
            M[I * DM + jx * Ky + jy] -= ck;
        }
        M[I * DM + I] += diag;
        rowdiag[I] = diag;
    }
    __syncthreads();
    if (singular) {
        __shared__ double sig;
        if (threadIdx.x == 0) {
            double dsum = 0.0;
            for (int I = 0; I < nc; I++) dsum += rowdiag[I];
            sig = dsum / nc / nc;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < DM * DM; t += DTB)
            if (t / DM < nc && t % DM < nc) M[t] += sig;
        __syncthreads();
    }
    // in place: after pivot p, row p holds row p of the partial inverse and column p the
    // eliminated multipliers
    for (int p = 0; p < nc; p++) {
        const double piv = M[p * DM + p];
        if (threadIdx.x < nc) colp[threadIdx.x] = M[threadIdx.x * DM + p];
        __syncthreads();
        if (threadIdx.x < nc) {
            const int j = threadIdx.x;
            M[p * DM + j] = (j == p ? 1.0 : M[p * DM + j]) / piv;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DM * DM / DTB; q++) {
            const int t = threadIdx.x + q * DTB, i = t / DM, j = t % DM;
            if (i >= nc || j >= nc || i == p) continue;
            M[t] = (j == p ? 0.0 : M[t]) - colp[i] * M[p * DM + j];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < DM * DM; t += DTB) {
        const int i = t / DM, j = t % DM;
        if (i < nc && j < nc) inv[j * nc + i] = (float)M[t];
    }
}

// ---- fine level -----------------------------------------------------------------------------------
//
// The fine kernels read the operator through a row view (lv_ell.h): the operator's own ELL (FP32
// values, particle order) or the aggregate-order copy built at set-up (BF16 values, 16-bit column
// offsets), in which case every fine vector of the cycle is in the aggregate order too.

template <class V>
struct fine_ops {
    int n;
    V A;
    const float *diag, *dinv, *mval, *mdg;
    float scale, alpha;
};

// (A z)_i in the difference form with the first 8 slots unrolled and predicated
template <class V>
__device__ __forceinline__ float a_row(int i, const fine_ops<V> &F, const float *__restrict__ z)
{
    constexpr int U = 8;
    const int n = F.n;
    const float zi = z[i];
    float acc = F.diag[i] * zi;
    int c[U];
    float v[U], zj[U];
    const int d = lv_ell_head<U>(F.A, i, n, c, v);
#pragma unroll
    for (int k = 0; k < U; k++) zj[k] = z[c[k]];
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (zi - zj[k]);
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        acc += F.A.value(e) * (zi - z[F.A.column(e, i, n)]);
    }
    return acc;
}

// (M r)_i of the learned preconditioner (on A's pattern, in the same order as the view), or D^-1 r
template <class V>
__device__ __forceinline__ float m_row(int i, const fine_ops<V> &F, const float *__restrict__ r)
{
    if (!F.mval) return F.dinv[i] * r[i];
    constexpr int U = 8;
    const int n = F.n;
    float acc = F.mdg[i] * r[i];
    int c[U];
    float v[U], rj[U];
    const int d = lv_ell_head_with<U>(F.A, F.mval, i, n, c, v);
#pragma unroll
    for (int k = 0; k < U; k++) rj[k] = r[c[k]];
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * rj[k];
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        acc += F.mval[e] * r[F.A.column(e, i, n)];
    }
    return acc;
}

// z = scale * M r
template <class V>
__global__ __launch_bounds__(B) void k_fine_pre(fine_ops<V> F, const float *__restrict__ r, float *__restrict__ z,
                                                const lv_cg_scalars *S)
{
    if (S && S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= F.n) return;
    z[i] = F.scale * m_row(i, F, r);
}

// t = r - A z
template <class V>
__global__ __launch_bounds__(B) void k_fine_resid(fine_ops<V> F, const float *__restrict__ r,
                                                  const float *__restrict__ z, float *__restrict__ t,
                                                  const lv_cg_scalars *S)
{
    if (S && S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= F.n) return;
    t[i] = r[i] - a_row(i, F, z);
}

// rc[I] = sum over the rows of cell I of t: RL lanes per cell read its (sorted) row list
// contiguously and reduce by a fixed shuffle tree; list null: the rows of cell I are
// start[I]..start[I+1]-1 themselves (aggregate order)
__global__ __launch_bounds__(B) void k_restrict1(int na, const int *__restrict__ start, const int *__restrict__ list,
                                                 const float *__restrict__ t, float *__restrict__ rc,
                                                 const lv_cg_scalars *S)
{
    if (S && S->done) return;
    const int tid = blockIdx.x * B + threadIdx.x;
    const int I = tid / RL, q = tid % RL;
    float s = 0.0f;
    if (I < na) {
        const int s1 = start[I + 1];
        if (list)
            for (int a = start[I] + q; a < s1; a += RL) s += t[list[a]];
        else
            for (int a = start[I] + q; a < s1; a += RL) s += t[a];
    }
    s = sum_lanes(s);
    if (q == 0 && I < na) rc[I] = s;
}

// Jacobi fine smoother, down, one row: z0 = w D^-1 r and t = r - A z0 (z0 of a neighbour is
// recomputed from its r); returns t, stores z0
template <class V>
__device__ __forceinline__ float down_row_j(const fine_ops<V> &F, int i, const float *__restrict__ r,
                                            float *__restrict__ z0)
{
    const int n = F.n;
    constexpr int U = 8;
    const float ri = r[i];
    const float zi = F.scale * F.dinv[i] * ri;
    float acc = F.diag[i] * zi;
    int c[U];
    float v[U];
    const int d = lv_ell_head<U>(F.A, i, n, c, v);
    float rj[U], dj[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        rj[k] = r[c[k]];
        dj[k] = F.dinv[c[k]];
    }
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (zi - F.scale * dj[k] * rj[k]);
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        const int j = F.A.column(e, i, n);
        acc += F.A.value(e) * (zi - F.scale * F.dinv[j] * r[j]);
    }
    z0[i] = zi;
    return ri - acc;
}

template <class V>
__global__ __launch_bounds__(B) void k_fine_down_j(fine_ops<V> F, const float *__restrict__ r,
                                                   float *__restrict__ z0, float *__restrict__ t,
                                                   const lv_cg_scalars *S)
{
    if (S && S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= F.n) return;
    t[i] = down_row_j(F, i, r, z0);
}

// The down pass and the restriction in one kernel, for the aggregate order (the rows of level-1
// cell I are start[I]..start[I+1]-1): block b owns the cells whose first row lies in
// [FUSE_ROWS b, FUSE_ROWS (b+1)) (blk[b].x..blk[b+1].x-1), keeps their residuals t in LDS and sums
// each cell's as k_restrict1 does (RL lanes, the same order and shuffle tree, so rc is the same).
// A block holds at most FUSE_ROWS - 1 + (largest cell) rows; set-up checks that this fits FUSE_CAP.
// With z0 given each thread takes FUSE_RPT rows (FUSE_ROWS a little under FUSE_RPT B), whose loads
// are in flight together.
constexpr int FUSE_RPT = 2;
constexpr int FUSE_ROWS = FUSE_RPT * 256 - 32;
constexpr int FUSE_CAP = 2048;

// t = r - A z0 for one row when z0 = w D^-1 r is already stored (the same values down_row_j forms)
template <class V>
__device__ __forceinline__ float down_row_z0(const fine_ops<V> &F, int i, const float *__restrict__ r,
                                             const float *__restrict__ z0)
{
    const int n = F.n;
    constexpr int U = 8;
    const float zi = z0[i];
    float acc = F.diag[i] * zi;
    int c[U];
    float v[U], zj[U];
    const int d = lv_ell_head<U>(F.A, i, n, c, v);
#pragma unroll
    for (int k = 0; k < U; k++) zj[k] = z0[c[k]];
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (zi - zj[k]);
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        acc += F.A.value(e) * (zi - z0[F.A.column(e, i, n)]);
    }
    return r[i] - acc;
}

// With z1 (level 1 above the tail), each cell's pre-smoothed z1 = w D1^-1 rc is stored too, as
// level 1's down pass would form it.
//
// blk[b] = (first cell, first row) of block b, so the block's bounds take one round trip, with the
// done flag (S->done, or without a solve the zero pair blk[nbf + 1]).  The
// restriction's row-list bounds and level-1 diagonals for its first FUSE_RPT trips of B lanes (a
// block's cells have about as many lanes as rows) are loaded before the rows, not after the
// barrier: a block's life is a chain of dependent round trips, and the blocks on a CU take several
// turns.  A lane past the block's cells loads cell na - 1's and drops them; dinv1 (level 1's
// diagonal) is read whether or not z1 is stored.
template <class V, bool Z0>
__global__ __launch_bounds__(B) void k_fine_down_jr(fine_ops<V> F, const float *__restrict__ r,
                                                    float *__restrict__ z0, const int *__restrict__ start,
                                                    const int2 *__restrict__ blk, int na, float *__restrict__ rc,
                                                    const float *__restrict__ dinv1, float *__restrict__ z1,
                                                    const int *__restrict__ done)
{
    __shared__ float tl[FUSE_CAP];
    const int b = lv_xcd_block(blockIdx.x, gridDim.x);
    const int2 p0 = blk[b], p1 = blk[b + 1];
    const int dn = *done;
    // the bounds and the flag in one round trip: without this the bounds' loads sink below the test
    asm volatile("" ::"s"(p0.x), "s"(p0.y), "s"(p1.x), "s"(p1.y), "s"(dn));
    if (dn) return;
    const int c0 = p0.x, c1 = p1.x, R0 = p0.y, R1 = p1.y;
    int sa[FUSE_RPT], sb[FUSE_RPT];
    float dv[FUSE_RPT];
#pragma unroll
    for (int p = 0; p < FUSE_RPT; p++) {
        const int I = c0 + (p * B + (int)threadIdx.x) / RL;
        const int Ic = I < na ? I : na - 1;
        sa[p] = start[Ic];
        sb[p] = start[Ic + 1];
        dv[p] = dinv1[Ic];
    }
    if constexpr (Z0) {
        // rows i0 + q B; all but the first are clamped into the block's rows so that every row's
        // loads are issued without a branch (a clamped duplicate is computed and dropped)
        for (int i0 = R0 + (int)threadIdx.x; i0 < R1; i0 += FUSE_RPT * B) {
            float t[FUSE_RPT];
#pragma unroll
            for (int q = 0; q < FUSE_RPT; q++) {
                const int i = i0 + q * B;
                t[q] = down_row_z0(F, i < R1 ? i : R1 - 1, r, z0);
            }
#pragma unroll
            for (int q = 0; q < FUSE_RPT; q++)
                if (i0 + q * B < R1) tl[i0 + q * B - R0] = t[q];
        }
    } else {
        for (int i = R0 + (int)threadIdx.x; i < R1; i += B) tl[i - R0] = down_row_j(F, i, r, z0);
    }
    __syncthreads();
    const int lanes = (c1 - c0) * RL;
    // one trip of B lanes, RL per cell; the trip count is uniform, so every lane shuffles
    auto restrict_trip = [&](int base, int a0, int s1, float di) {
        const int idx = base + (int)threadIdx.x;
        const int I = c0 + idx / RL, q = idx % RL;
        float s = 0.0f;
        if (idx < lanes)
            for (int a = a0 + q; a < s1; a += RL) s += tl[a - R0];
        s = sum_lanes(s);
        if (q == 0 && idx < lanes) {
            rc[I] = s;
            if (z1) z1[I] = OMEGA * di * s;
        }
    };
#pragma unroll
    for (int p = 0; p < FUSE_RPT; p++)
        if (p * B < lanes) restrict_trip(p * B, sa[p], sb[p], dv[p]);
    for (int base = FUSE_RPT * B; base < lanes; base += B) {
        const int I = c0 + (base + (int)threadIdx.x) / RL;
        const int Ic = I < na ? I : na - 1;
        restrict_trip(base, start[Ic], start[Ic + 1], dinv1[Ic]);
    }
}

// blk[b] = (the first cell I with start[I] >= FUSE_ROWS b, start[I]) for b = 0..nbf-1,
// blk[nbf] = (na, n) (start[na] = n, nbf = n / FUSE_ROWS + 1) and blk[nbf + 1] = (0, 0), the done flag
// of a cycle outside a solve.  Thread I writes the b in
// (start[I-1], start[I]] / FUSE_ROWS; flag[1] is set when a cell has more than FUSE_CAP - FUSE_ROWS
// rows.
__global__ __launch_bounds__(B) void k_fuse_blocks(int na, int nbf, const int *__restrict__ start,
                                                   int2 *__restrict__ blk, int *__restrict__ flag)
{
    const int I = blockIdx.x * B + threadIdx.x;
    if (I > na) return;
    const int sI = start[I];
    const int lo = I == 0 ? 0 : start[I - 1] / FUSE_ROWS + 1;
    const int hi = sI / FUSE_ROWS;
    for (int b = lo; b <= hi && b < nbf; b++) blk[b] = make_int2(I, sI);
    if (I == na) {
        blk[nbf] = make_int2(na, sI);
        blk[nbf + 1] = make_int2(0, 0);
    }
    if (I < na && start[I + 1] - sI > FUSE_CAP - FUSE_ROWS) atomicOr(flag + 1, 1);
}

// up: z1 = z0 + alpha P ec and t = r - A z1 in one pass (z1 of a neighbour from its z0 and its
// cell's ec)
template <class V>
__device__ __forceinline__ float up_row(const fine_ops<V> &F, int i, const float *__restrict__ z0,
                                        const int *__restrict__ agg, const float *__restrict__ ec,
                                        const float *__restrict__ r, float &z1)
{
    constexpr int U = 8;
    const int n = F.n;
    // three dependent stages (columns; their z0 and cells; the cells' corrections), each issued for
    // all of the first U slots before the next
    int c[U];
    float v[U];
    const int d = lv_ell_head<U>(F.A, i, n, c, v);
    float zj[U];
    int aj[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        zj[k] = z0[c[k]];
        aj[k] = agg[c[k]];
    }
    z1 = z0[i] + F.alpha * ec[agg[i]];
    float acc = F.diag[i] * z1;
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (z1 - (zj[k] + F.alpha * ec[aj[k]]));
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        const int j = F.A.column(e, i, n);
        acc += F.A.value(e) * (z1 - (z0[j] + F.alpha * ec[agg[j]]));
    }
    return r[i] - acc;
}

// learned smoother, up: z = z1 and t = r - A z1 (the post-smoothing follows in k_fine_post)
template <class V>
__global__ __launch_bounds__(B) void k_fine_up(fine_ops<V> F, const float *__restrict__ r,
                                               const float *__restrict__ z0, const int *__restrict__ agg,
                                               const float *__restrict__ ec, float *__restrict__ z,
                                               float *__restrict__ t, const lv_cg_scalars *S)
{
    if (S && S->done) return;
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= F.n) return;
    float z1;
    const float ti = up_row(F, i, z0, agg, ec, r, z1);
    z[i] = z1;
    t[i] = ti;
}

// prolongation in place: z0 += alpha ec[cell of the row], four rows per thread with 16-byte loads
// and stores of agg and z0 (both start 256-byte aligned); the last thread takes the n % 4 rows
constexpr int PROLONG_RPT = 4;
__global__ __launch_bounds__(B) void k_prolong1(int n, const int *__restrict__ agg, const float *__restrict__ ec,
                                                float alpha, float *__restrict__ z0, const lv_cg_scalars *S)
{
    typedef int i4 __attribute__((ext_vector_type(4)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    if (S && S->done) return;
    const int i = PROLONG_RPT * (blockIdx.x * B + threadIdx.x);
    if (i + 3 < n) {
        const i4 a = *reinterpret_cast<const i4 *>(agg + i);
        f4 z = *reinterpret_cast<const f4 *>(z0 + i);
        z.x += alpha * ec[a.x];
        z.y += alpha * ec[a.y];
        z.z += alpha * ec[a.z];
        z.w += alpha * ec[a.w];
        *reinterpret_cast<f4 *>(z0 + i) = z;
    } else {
        for (int k = i; k < n; k++) z0[k] += alpha * ec[agg[k]];
    }
}

// Jacobi smoother, up (after k_prolong1): z = z1 + w D^-1 (r - A z1); partials r.z, r.r of the final z
template <class V>
__global__ __launch_bounds__(B) void k_fine_up_j(fine_ops<V> F, const float *__restrict__ r,
                                                 const float *__restrict__ z1, float *__restrict__ z,
                                                 double *__restrict__ Prr, const lv_cg_scalars *S)
{
    if (S && S->done) return;
    double v[2] = {0.0, 0.0};
    // (the compact overflow of lv_ell.h does not pay here: 49.4 against 48.5 us per launch at 2048^2)
    for (int i = lv_xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x; i < F.n; i += gridDim.x * B) {
        const float ti = r[i] - a_row(i, F, z1);
        const float zi = z1[i] + F.scale * F.dinv[i] * ti;
        z[i] = zi;
        const float ri = r[i];
        v[0] += (double)ri * (double)zi;
        v[1] += (double)ri * (double)ri;
    }
    if (Prr) block_partials<2>(v, Prr, gridDim.x);
}

// z += scale * M t; partials r.z, r.r of the final z (grid-stride over lv_nblocks(n) blocks)
template <class V>
__global__ __launch_bounds__(B) void k_fine_post(fine_ops<V> F, const float *__restrict__ r,
                                                 const float *__restrict__ t, float *__restrict__ z,
                                                 double *__restrict__ Prr, const lv_cg_scalars *S)
{
    if (S && S->done) return;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < F.n; i += gridDim.x * B) {
        const float zi = z[i] + F.scale * m_row(i, F, t);
        z[i] = zi;
        const float ri = r[i];
        v[0] += (double)ri * (double)zi;
        v[1] += (double)ri * (double)ri;
    }
    if (Prr) block_partials<2>(v, Prr, gridDim.x);
}

// ---- coarse levels (3x3 stencil) ---------------------------------------------------------------

struct st_level {
    int Kx, Ky, xper, yper;
    const float *d, *c, *dinv;
    float alpha;  // factor on the correction prolongated into this level
};

// wrap(i, K, per) for i in [-1, K] (a cell index plus an offset of at most one): a compare and an
// add instead of two remainders
// (selects, not branches: a branch here splits the stencil kernels' loads into basic blocks that
// the compiler then waits between)
__device__ __forceinline__ int wrap_near(int i, int K, int per, bool &ok)
{
    const int w = i < 0 ? i + K : (i >= K ? i - K : i);
    ok = per || (i >= 0 && i < K);
    return per ? w : i;
}

// the 8 neighbour cells of (cx, cy) on level L and their couplings; a neighbour outside a
// non-periodic level gets coupling 0 and index f, so every load below is in bounds and all of them
// can be issued at once (the coupling slot is loaded whatever the neighbour, and then selected: a
// load under the condition would make the gathers wait for it).  With PJ, also each neighbour's
// parent on the next level (Kyc columns).
__device__ __forceinline__ void st_nbrs(const st_level &L, int cx, int cy, int f, float ck[8], int J[8],
                                        int *PJ = nullptr, int Kyc = 0)
{
    const int n = L.Kx * L.Ky;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int ox, oy;
        offset_of(k, ox, oy);
        bool okx, oky;
        const int jx = wrap_near(cx + ox, L.Kx, L.xper, okx), jy = wrap_near(cy + oy, L.Ky, L.yper, oky);
        const bool ok = okx && oky;
        J[k] = ok ? jx * L.Ky + jy : f;
        const float c = L.c[(size_t)k * n + f];
        ck[k] = ok ? c : 0.0f;
        if (PJ) PJ[k] = ok ? (jx / CF) * Kyc + jy / CF : (cx / CF) * Kyc + cy / CF;
    }
}

// down, one lane: child q of parent P (CF2 lanes per parent): z_f = w D^-1 r_f and the child's
// residual t_f = r_f - (A z)_f, with the neighbours' z recomputed from their r (0 outside the level)
// (ZGIVEN: z = w D^-1 r of the level is already stored, the same values, and only gathered)
template <bool ZGIVEN = false>
__device__ __forceinline__ float st_down_lane(const st_level &L, int Kyc, int nc, int P, int q,
                                              const float *__restrict__ r, float *__restrict__ z)
{
    if (P >= nc) return 0.0f;
    const int Px = P / Kyc;
    const int cx = CF * Px + q / CF, cy = CF * (P - Px * Kyc) + q % CF;
    if (cx >= L.Kx || cy >= L.Ky) return 0.0f;
    const int f = cx * L.Ky + cy;
    float ck[8], zj[8];
    int J[8];
    st_nbrs(L, cx, cy, f, ck, J);
    const float rf = r[f];
    float zf;
    if (ZGIVEN) {
#pragma unroll
        for (int k = 0; k < 8; k++) zj[k] = z[J[k]];
        zf = z[f];
    } else {
        float rj[8], dj[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            rj[k] = r[J[k]];
            dj[k] = L.dinv[J[k]];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) zj[k] = OMEGA * dj[k] * rj[k];
        zf = OMEGA * L.dinv[f] * rf;
    }
    float acc = L.d[f] * zf;
#pragma unroll
    for (int k = 0; k < 8; k++) acc += ck[k] * (zf - zj[k]);
    if (!ZGIVEN) z[f] = zf;
    return rf - acc;
}

// the children's residuals of CF2 consecutive lanes summed by a fixed shuffle tree (every lane of
// the wave takes part)
__device__ __forceinline__ float sum_children(float t)
{
#pragma unroll
    for (int m = 1; m < CF2; m <<= 1) t += __shfl_xor(t, m);
    return t;
}

// up, one cell: z1 = z + ec[parent] (neighbours' z1 recomputed), t = r - A z1, out = z1 + w D^-1 t
__device__ __forceinline__ float st_up_cell(const st_level &L, int Kyc, int f, const float *__restrict__ r,
                                            const float *__restrict__ z, const float *__restrict__ ec,
                                            float *__restrict__ out)
{
    const int cx = f / L.Ky, cy = f - cx * L.Ky;
    float ck[8], zj[8], ej[8];
    int J[8], PJ[8];
    st_nbrs(L, cx, cy, f, ck, J, PJ, Kyc);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        zj[k] = z[J[k]];
        ej[k] = L.alpha * ec[PJ[k]];
    }
    const float z1 = z[f] + L.alpha * ec[(cx / CF) * Kyc + cy / CF];
    float acc = L.d[f] * z1;
#pragma unroll
    for (int k = 0; k < 8; k++) acc += ck[k] * (z1 - (zj[k] + ej[k]));
    const float v = z1 + OMEGA * L.dinv[f] * (r[f] - acc);
    out[f] = v;
    return v;
}

// W-cycle, after the first visit of level L returned t: acc = t and r = r - A t (the right-hand
// side of the second visit)
__device__ __forceinline__ void st_wsave_cell(const st_level &L, int f, float *__restrict__ r,
                                              const float *__restrict__ t, float *__restrict__ acc)
{
    const int cx = f / L.Ky, cy = f - cx * L.Ky;
    float ck[8], tj[8];
    int J[8];
    st_nbrs(L, cx, cy, f, ck, J);
#pragma unroll
    for (int k = 0; k < 8; k++) tj[k] = t[J[k]];
    const float tf = t[f];
    float a = L.d[f] * tf;
#pragma unroll
    for (int k = 0; k < 8; k++) a += ck[k] * (tf - tj[k]);
    acc[f] = tf;
    r[f] -= a;
}

// down on level L: pre-smoothed z of L (given, or stored here) and the right-hand side rc of the
// next level
template <bool ZGIVEN>
__global__ __launch_bounds__(B) void k_st_down(st_level L, int Kyc, int nc, const float *__restrict__ r,
                                               float *__restrict__ z, float *__restrict__ rc, const lv_cg_scalars *S)
{
    if (S && S->done) return;
    // (workgroups in the XCD-aware order: the parent rows above and below a workgroup's, whose
    // children it gathers, are mostly on its own XCD)
    const int tid = lv_xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x;
    const int P = tid / CF2, q = tid % CF2;
    const float t = sum_children(st_down_lane<ZGIVEN>(L, Kyc, nc, P, q, r, z));
    if (q == 0 && P < nc) rc[P] = t;
}

// up on level L: the final correction out of L from the next level's correction ec.  PROLONG (level
// 1 in the aggregate order, whose cell f owns rows start[f]..start[f+1]-1): also the prolongation
// z0 += alpha0 out[f] of those rows, as k_prolong1 forms it
template <bool PROLONG>
__global__ __launch_bounds__(B) void k_st_up(st_level L, int Kyc, const float *__restrict__ r,
                                             const float *__restrict__ z, const float *__restrict__ ec,
                                             float *__restrict__ out, const lv_cg_scalars *S,
                                             const int *__restrict__ start, float alpha0, float *__restrict__ z0)
{
    if (S && S->done) return;
    const int f = lv_xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x;  // (as k_st_down)
    if (f >= L.Kx * L.Ky) return;
    // the prolongation's rows: bounds first, then the first PH rows' z0 loaded together (a row past
    // the cell's reloads its last, or row 0 for an empty cell at the end), so that the cell's rows
    // take one round trip after the stencil's loads instead of one each
    constexpr int PH = 4;
    int b0 = 0, e = 0;
    float zz[PH];
    if (PROLONG) {
        b0 = start[f];
        e = start[f + 1];
        const int last = e > 0 ? e - 1 : 0;
#pragma unroll
        for (int h = 0; h < PH; h++) zz[h] = z0[b0 + h < e ? b0 + h : last];
    }
    const float v = st_up_cell(L, Kyc, f, r, z, ec, out);
    if (PROLONG) {
        // (the rows' values in registers before the stores' branches, which would otherwise take
        // the first row's load with them)
        asm volatile("" : "+v"(zz[0]), "+v"(zz[1]), "+v"(zz[2]), "+v"(zz[3]));
        const float a = alpha0 * v;
#pragma unroll
        for (int h = 0; h < PH; h++)
            if (b0 + h < e) z0[b0 + h] = zz[h] + a;
        for (int i = b0 + PH; i < e; i++) z0[i] += a;
    }
}

// ---- two coarse levels in one kernel ------------------------------------------------------------
//
// Levels l and l+1 above the tail run their down passes as one kernel and their up passes as another.
// A block owns a tile of cells and recomputes in LDS what its stencils read from the cells around the
// tile (a halo), with the expressions and the order of st_down_lane, sum_children and st_up_cell, so
// every value is bit for bit the one the one-level kernels give.  Cells are addressed by unwrapped
// coordinates; on a periodic axis the width of level l is a multiple of 4, so that the parent map
// x -> x >> 1 commutes with the wrap on levels l and l+1.  Level l is at least PAIR_MIN_K cells wide
// on both axes, so the unwrapped coordinates of every region stay within one width of the level and
// one conditional add or subtract wraps them.  The down tile is small (PD = 4: 4 lanes of work per
// cell of level l+2) so that the small levels spread over many workgroups; 8 x 8 tiles left one
// wavefront per SIMD working through six halo cells a lane, no faster than two one-level kernels.

constexpr int PD = 4;   // down: a block owns PD x PD cells of level l+2
constexpr int PU = 16;  // up: a block owns PU x PU cells of level l, one a lane
constexpr int PAIR_MIN_K = 24;
static_assert(PAIR_MIN_K >= 4 * PD + 8 && PAIR_MIN_K >= PU + 8, "every region within one width of its level");

struct st_pair {
    st_level L0, L1;  // levels l and l+1
    int Kx2, Ky2;     // level l+2
};

// cell (x, y) of a Kx x Ky level, unwrapped within one width: its index, or 0 with ok false outside a
// non-periodic axis (so that every load below has a valid address and none waits behind a branch)
__device__ __forceinline__ int cell_of(int x, int y, int Kx, int Ky, int xper, int yper, bool &ok)
{
    bool okx, oky;
    x = wrap_near(x, Kx, xper, okx);
    y = wrap_near(y, Ky, yper, oky);
    ok = okx && oky;
    return ok ? x * Ky + y : 0;
}

// the couplings of cell f = (x, y) (unwrapped) of level L and the mask of its neighbours inside the
// level; a neighbour outside gets coupling 0 (st_nbrs)
__device__ __forceinline__ unsigned st_coefs(const st_level &L, int x, int y, int f, float ck[8])
{
    const int n = L.Kx * L.Ky;
    unsigned ok = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int ox, oy;
        offset_of(k, ox, oy);
        const bool in = (L.xper || (x + ox >= 0 && x + ox < L.Kx)) && (L.yper || (y + oy >= 0 && y + oy < L.Ky));
        ok |= (unsigned)in << k;
        const float c = L.c[(size_t)k * n + f];
        ck[k] = in ? c : 0.0f;
    }
    return ok;
}

// the residual r - A z of one cell from the z of its 3 x 3 neighbourhood in an LDS region of row
// length SR (zc at offset 0), as st_down_lane forms it
template <int SR>
__device__ __forceinline__ float st_resid_lds(float rf, float d, const float ck[8], unsigned ok,
                                              const float *__restrict__ zc)
{
    const float zf = zc[0];
    float acc = d * zf;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int ox, oy;
        offset_of(k, ox, oy);
        const float zj = ((ok >> k) & 1) ? zc[ox * SR + oy] : zf;
        acc += ck[k] * (zf - zj);
    }
    return rf - acc;
}

// down on levels l and l+1: the pre-smoothed z of both, the right-hand side r1 of level l+1 and r2
// of level l+2, as k_st_down<false> on level l and then on level l+1.  Regions of a block (level,
// first cell relative to its tile's first cell, side): A (l, -3, 4PD+6) z and r, B (l, -2, 4PD+4)
// the residual, C (l+1, -1, 2PD+2) r and z, D (l+1, 0, 2PD) the residual.
__global__ __launch_bounds__(B) void k_st_down2(st_pair Q, const float *__restrict__ r0, float *__restrict__ z0,
                                                float *__restrict__ r1, float *__restrict__ z1,
                                                float *__restrict__ r2, const lv_cg_scalars *S)
{
    constexpr int SA = 4 * PD + 6, SB = 4 * PD + 4, SC = 2 * PD + 2, SD = 2 * PD;
    constexpr int NA = SA * SA, NB = SB * SB, NC = SC * SC;
    constexpr int IA = (NA + B - 1) / B, IB = (NB + B - 1) / B, IC = (NC + B - 1) / B;
    static_assert(SD * SD <= B, "region D: at most one cell a lane");
    __shared__ float zA[NA], rA[NA], tB[NB], rC[NC], zC[NC], tD[SD * SD];
    if (S && S->done) return;
    const st_level &L0 = Q.L0, &L1 = Q.L1;
    const int qx = blockIdx.x * PD, qy = blockIdx.y * PD;  // the tile's first cell on level l+2
    const int ax = 4 * qx - 3, ay = 4 * qy - 3, cx = 2 * qx - 1, cy = 2 * qy - 1;
    const int tid = threadIdx.x;
    // every global load first, unconditionally (fA..: -1 outside the level or past the region)
    int fA[IA], fB[IB], fC[IC];
    float ra[IA], da[IA], dB[IB], cB[IB][8], dvC[IC], cD[8];
    unsigned okB[IB];
#pragma unroll
    for (int i = 0; i < IA; i++) {
        const int s = tid + i * B, sa = s < NA ? s : NA - 1;  // lanes past the region load a cell of it
        bool ok;
        const int f = cell_of(ax + sa / SA, ay + sa % SA, L0.Kx, L0.Ky, L0.xper, L0.yper, ok);
        ra[i] = r0[f];
        da[i] = L0.dinv[f];
        fA[i] = (ok && s < NA) ? f : -1;
    }
#pragma unroll
    for (int i = 0; i < IB; i++) {
        const int s = tid + i * B, sb = s < NB ? s : NB - 1, x = ax + 1 + sb / SB, y = ay + 1 + sb % SB;
        bool ok;
        const int f = cell_of(x, y, L0.Kx, L0.Ky, L0.xper, L0.yper, ok);
        dB[i] = L0.d[f];
        okB[i] = st_coefs(L0, x, y, f, cB[i]);
        fB[i] = (ok && s < NB) ? f : -1;
    }
#pragma unroll
    for (int i = 0; i < IC; i++) {
        const int s = tid + i * B, sc = s < NC ? s : NC - 1;
        bool ok;
        const int f = cell_of(cx + sc / SC, cy + sc % SC, L1.Kx, L1.Ky, L1.xper, L1.yper, ok);
        dvC[i] = L1.dinv[f];
        fC[i] = (ok && s < NC) ? f : -1;
    }
    const int td = tid < SD * SD ? tid : 0, dxs = td / SD, dys = td % SD, Xd = cx + 1 + dxs, Yd = cy + 1 + dys;
    bool okd;
    int fD = cell_of(Xd, Yd, L1.Kx, L1.Ky, L1.xper, L1.yper, okd);
    const float dD = L1.d[fD];
    const unsigned okD = st_coefs(L1, Xd, Yd, fD, cD);
    fD = (okd && tid < SD * SD) ? fD : -1;
    // A: z = w D^-1 r of level l
#pragma unroll
    for (int i = 0; i < IA; i++) {
        const int s = tid + i * B;
        if (s < NA) {
            zA[s] = fA[i] >= 0 ? OMEGA * da[i] * ra[i] : 0.0f;
            rA[s] = fA[i] >= 0 ? ra[i] : 0.0f;
        }
    }
    __syncthreads();
    // B: the residual of level l; the tile's own cells store their z
#pragma unroll
    for (int i = 0; i < IB; i++) {
        const int s = tid + i * B;
        if (s >= NB) continue;
        const int sx = s / SB, sy = s % SB;
        if (fB[i] < 0) {
            tB[s] = 0.0f;
            continue;
        }
        const int a = (sx + 1) * SA + sy + 1;
        tB[s] = st_resid_lds<SA>(rA[a], dB[i], cB[i], okB[i], zA + a);
        const int x = ax + 1 + sx, y = ay + 1 + sy;
        if (sx >= 2 && sx < 4 * PD + 2 && sy >= 2 && sy < 4 * PD + 2 && x < L0.Kx && y < L0.Ky) z0[fB[i]] = zA[a];
    }
    __syncthreads();
    // C: r of level l+1, the sum of its children's residuals in sum_children's order, and its z
#pragma unroll
    for (int i = 0; i < IC; i++) {
        const int s = tid + i * B;
        if (s >= NC) continue;
        const int sx = s / SC, sy = s % SC;
        if (fC[i] < 0) {
            rC[s] = 0.0f;
            zC[s] = 0.0f;
            continue;
        }
        const int b = 2 * sx * SB + 2 * sy;
        const float r = (tB[b] + tB[b + 1]) + (tB[b + SB] + tB[b + SB + 1]);
        rC[s] = r;
        zC[s] = OMEGA * dvC[i] * r;
        const int X = cx + sx, Y = cy + sy;
        if (sx >= 1 && sx < 2 * PD + 1 && sy >= 1 && sy < 2 * PD + 1 && X < L1.Kx && Y < L1.Ky) r1[fC[i]] = r;
    }
    __syncthreads();
    // D: the residual of level l+1 on the tile's children; they store their z
    if (tid < SD * SD) {
        const int c = (dxs + 1) * SC + dys + 1;
        tD[tid] = fD >= 0 ? st_resid_lds<SC>(rC[c], dD, cD, okD, zC + c) : 0.0f;
        if (fD >= 0 && Xd < L1.Kx && Yd < L1.Ky) z1[fD] = zC[c];
    }
    __syncthreads();
    // E: r of level l+2
    if (tid < PD * PD) {
        const int sx = tid / PD, sy = tid % PD, X = qx + sx, Y = qy + sy;
        if (X < Q.Kx2 && Y < Q.Ky2) {
            const int b = 2 * sx * SD + 2 * sy;
            r2[X * Q.Ky2 + Y] = (tD[b] + tD[b + 1]) + (tD[b + SD] + tD[b + SD + 1]);
        }
    }
}

// up on levels l+1 and l: the final correction t0 of level l from the final correction t2 of level
// l+2, as k_st_up<false> on level l+1 and then on level l (level l+1's own correction stays in LDS).
// Regions of a block (level, first cell relative to the tile's, side): Z0 (l, -1, PU+2) z, V (l+1,
// -1, PU/2+2) the correction, Z1 (l+1, -2, PU/2+4) z, E (l+2, -1, PU/4+2) t2.
__global__ __launch_bounds__(B) void k_st_up2(st_pair Q, const float *__restrict__ r0, const float *__restrict__ z0,
                                              const float *__restrict__ r1, const float *__restrict__ z1,
                                              const float *__restrict__ t2, float *__restrict__ t0,
                                              const lv_cg_scalars *S)
{
    constexpr int SZ0 = PU + 2, SV = PU / 2 + 2, SZ1 = PU / 2 + 4, SE = PU / 4 + 2;
    constexpr int NZ0 = SZ0 * SZ0, IZ0 = (NZ0 + B - 1) / B;
    static_assert(PU * PU == B && SZ1 * SZ1 <= B && SV * SV <= B, "one cell of each region a lane");
    __shared__ float zl[NZ0], vm[SV * SV], zm[SZ1 * SZ1], e2[SE * SE];
    if (S && S->done) return;
    const st_level &L0 = Q.L0, &L1 = Q.L1;
    const int x0 = blockIdx.x * PU, y0 = blockIdx.y * PU;  // the tile's first cell on level l
    const int vx = x0 / 2 - 1, vy = y0 / 2 - 1, ex = x0 / 4 - 1, ey = y0 / 4 - 1;
    const int tid = threadIdx.x;
    // every global load first, unconditionally (from cell 0 outside a level or a region)
    float zv[IZ0];
#pragma unroll
    for (int i = 0; i < IZ0; i++) {
        const int s = tid + i * B, sz = s < NZ0 ? s : NZ0 - 1;  // lanes past the region load a cell of it
        bool ok;
        const int f = cell_of(x0 - 1 + sz / SZ0, y0 - 1 + sz % SZ0, L0.Kx, L0.Ky, L0.xper, L0.yper, ok);
        const float v = z0[f];
        zv[i] = ok ? v : 0.0f;
    }
    bool okm, oke;
    const int tm = tid < SZ1 * SZ1 ? tid : 0, te = tid < SE * SE ? tid : 0;
    const int fm = cell_of(vx - 1 + tm / SZ1, vy - 1 + tm % SZ1, L1.Kx, L1.Ky, L1.xper, L1.yper, okm);
    const int fe = cell_of(ex + te / SE, ey + te % SE, Q.Kx2, Q.Ky2, L1.xper, L1.yper, oke);
    const float zm_ = z1[fm], ev_ = t2[fe];
    const float zmv = okm ? zm_ : 0.0f, ev = oke ? ev_ : 0.0f;
    const int tv = tid < SV * SV ? tid : 0, vxs = tv / SV, vys = tv % SV, Xv = vx + vxs, Yv = vy + vys;
    bool okv;
    int fV = cell_of(Xv, Yv, L1.Kx, L1.Ky, L1.xper, L1.yper, okv);
    float cV[8];
    const float dV = L1.d[fV], dvV = L1.dinv[fV], rV = r1[fV];
    const unsigned okV = st_coefs(L1, Xv, Yv, fV, cV);
    fV = (okv && tid < SV * SV) ? fV : -1;
    const int txs = tid / PU, tys = tid % PU, x = x0 + txs, y = y0 + tys;
    const bool own = x < L0.Kx && y < L0.Ky;
    const int f = own ? x * L0.Ky + y : 0;
    float cT[8];
    const float dT = L0.d[f], dvT = L0.dinv[f], rT = r0[f];
    const unsigned okT = st_coefs(L0, x, y, f, cT);
#pragma unroll
    for (int i = 0; i < IZ0; i++)
        if (tid + i * B < NZ0) zl[tid + i * B] = zv[i];
    if (tid < SZ1 * SZ1) zm[tid] = zmv;
    if (tid < SE * SE) e2[tid] = ev;
    __syncthreads();
    // V: level l+1's final correction (st_up_cell with ec = t2)
    if (tid < SV * SV) {
        float v = 0.0f;
        if (fV >= 0) {
            const float a = L1.alpha;
            const int zc = (vxs + 1) * SZ1 + vys + 1;
            const float pe = e2[((Xv >> 1) - ex) * SE + ((Yv >> 1) - ey)];
            const float z1c = zm[zc] + a * pe;
            float acc = dV * z1c;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                int ox, oy;
                offset_of(k, ox, oy);
                const bool in = (okV >> k) & 1;
                const float zj = in ? zm[zc + ox * SZ1 + oy] : zm[zc];
                const float ej = a * (in ? e2[(((Xv + ox) >> 1) - ex) * SE + (((Yv + oy) >> 1) - ey)] : pe);
                acc += cV[k] * (z1c - (zj + ej));
            }
            v = z1c + OMEGA * dvV * (rV - acc);
        }
        vm[tid] = v;
    }
    __syncthreads();
    // the tile: level l's final correction (st_up_cell with ec = level l+1's)
    if (own) {
        const float a = L0.alpha;
        const int zc = (txs + 1) * SZ0 + tys + 1;
        const float pv = vm[((x >> 1) - vx) * SV + ((y >> 1) - vy)];
        const float z1c = zl[zc] + a * pv;
        float acc = dT * z1c;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            int ox, oy;
            offset_of(k, ox, oy);
            const bool in = (okT >> k) & 1;
            const float zj = in ? zl[zc + ox * SZ0 + oy] : zl[zc];
            const float ej = a * (in ? vm[(((x + ox) >> 1) - vx) * SV + (((y + oy) >> 1) - vy)] : pv);
            acc += cT[k] * (z1c - (zj + ej));
        }
        t0[f] = z1c + OMEGA * dvT * (rT - acc);
    }
}

// The small levels of the cycle in one workgroup: a program of operations on levels first..last
// (last: the dense one), built at set-up for a V- or W-cycle, each operation followed by a barrier
// (the levels' vectors stay in global memory, which the barrier makes visible across the
// workgroup).  Reads the right-hand side of level first, writes its final correction.
constexpr int TB = 1024;
constexpr int TAIL_MAX = 1024;  // levels of at most this many cells run in the tail
enum tail_op { OP_DOWN = 0, OP_DENSE = 1, OP_UP = 2, OP_WSAVE = 3, OP_WADD = 4 };
constexpr int OP_SHIFT = 5;  // an operation is op << OP_SHIFT | level
static_assert(LV_AMG_MAXLEV < (1 << OP_SHIFT), "levels fit the operation code");
struct tail_levels {
    int first, last, xper, yper;
    int Kx[LV_AMG_MAXLEV + 1], Ky[LV_AMG_MAXLEV + 1];
    float alpha[LV_AMG_MAXLEV + 1];
    const float *d[LV_AMG_MAXLEV + 1], *c[LV_AMG_MAXLEV + 1], *dinv[LV_AMG_MAXLEV + 1];
    float *r[LV_AMG_MAXLEV + 1], *z[LV_AMG_MAXLEV + 1], *t[LV_AMG_MAXLEV + 1], *acc[LV_AMG_MAXLEV + 1];
};

// where the tail keeps its levels: global memory (the levels' own arrays) or LDS (offsets into
// one workgroup array); both give the same values, operation by operation
struct tail_in_global {
    const tail_levels &T;
    const float *inv_;
    __device__ const float *d(int l) const { return T.d[l]; }
    __device__ const float *c(int l) const { return T.c[l]; }
    __device__ const float *dinv(int l) const { return T.dinv[l]; }
    __device__ float *r(int l) const { return T.r[l]; }
    __device__ float *z(int l) const { return T.z[l]; }
    __device__ float *t(int l) const { return T.t[l]; }
    __device__ float *acc(int l) const { return T.acc[l]; }
    __device__ const float *inv() const { return inv_; }
};

// offsets (in floats) of every array of the tail levels in the LDS copy; total <= TAIL_LDS
constexpr int TAIL_LDS = 35 * 1024;
struct tail_layout {
    int d[LV_AMG_MAXLEV + 1], c[LV_AMG_MAXLEV + 1], dinv[LV_AMG_MAXLEV + 1];
    int r[LV_AMG_MAXLEV + 1], z[LV_AMG_MAXLEV + 1], t[LV_AMG_MAXLEV + 1], acc[LV_AMG_MAXLEV + 1];
    int inv, coef, total;  // [0, coef): the constant arrays, then the work vectors
};

struct tail_in_lds {
    float *lds;
    const tail_layout &Y;
    __device__ const float *d(int l) const { return lds + Y.d[l]; }
    __device__ const float *c(int l) const { return lds + Y.c[l]; }
    __device__ const float *dinv(int l) const { return lds + Y.dinv[l]; }
    __device__ float *r(int l) const { return lds + Y.r[l]; }
    __device__ float *z(int l) const { return lds + Y.z[l]; }
    __device__ float *t(int l) const { return lds + Y.t[l]; }
    __device__ float *acc(int l) const { return lds + Y.acc[l]; }
    __device__ const float *inv() const { return lds + Y.inv; }
};

// the program's operations, each followed by a barrier that makes its writes visible to the
// workgroup
template <class M>
__device__ __forceinline__ void run_tail(const tail_levels &T, const M &m, const int *prog, int nprog)
{
    for (int pc = 0; pc < nprog; pc++) {
        const int code = prog[pc];
        const int op = code >> OP_SHIFT, l = code & ((1 << OP_SHIFT) - 1);
        const st_level L{T.Kx[l], T.Ky[l], T.xper, T.yper, m.d(l), m.c(l), m.dinv(l), T.alpha[l]};
        const int n = T.Kx[l] * T.Ky[l];
        if (op == OP_DOWN) {
            const int Kyc = T.Ky[l + 1], nc = T.Kx[l + 1] * T.Ky[l + 1];
            float *rl = m.r(l), *zl = m.z(l), *rc = m.r(l + 1);
            for (int base = 0; base < nc * CF2; base += TB) {  // uniform trip count: every lane shuffles
                const int tid = base + threadIdx.x;
                const int P = tid / CF2, q = tid % CF2;
                const float t = sum_children(st_down_lane(L, Kyc, nc, P, q, rl, zl));
                if (q == 0 && P < nc) rc[P] = t;
            }
        } else if (op == OP_DENSE) {
            // t = inv r: 16 lanes per row (LV_AMG_DENSE_MAX * 16 == TB)
            const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
            const float *rd = m.r(l), *iv = m.inv();
            float acc = 0.0f;
            if (i < n)
                for (int j = q; j < n; j += 16) acc += iv[j * n + i] * rd[j];
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            acc += __shfl_xor(acc, 4);
            acc += __shfl_xor(acc, 8);
            if (q == 0 && i < n) m.t(l)[i] = acc;
        } else if (op == OP_UP) {
            const float *rl = m.r(l), *zl = m.z(l), *ec = m.t(l + 1);
            float *out = m.t(l);
            for (int f = threadIdx.x; f < n; f += TB) st_up_cell(L, T.Ky[l + 1], f, rl, zl, ec, out);
        } else if (op == OP_WSAVE) {
            float *rl = m.r(l), *tl = m.t(l), *al = m.acc(l);
            for (int f = threadIdx.x; f < n; f += TB) st_wsave_cell(L, f, rl, tl, al);
        } else {  // OP_WADD
            float *tl = m.t(l);
            const float *al = m.acc(l);
            for (int f = threadIdx.x; f < n; f += TB) tl[f] += al[f];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(TB) void k_tail(tail_levels T, const int *__restrict__ prog, int nprog,
                                             const float *__restrict__ inv, const lv_cg_scalars *S)
{
    if (S && S->done) return;
    run_tail(T, tail_in_global{T, inv}, prog, nprog);
}

// The same program with the tail levels in LDS: the couplings, the dense inverse, the program
// and the right-hand side of the first level are copied in, every operation then reads and
// writes LDS only, and the final correction of the first level is copied out.
// The constant arrays come from img (k_tail_pack's image of them in this layout) by one flat copy
// whose loads are all issued before the done flag is tested and the first LDS store is made; a copy
// loop per array would wait for a global round trip per iteration.
__global__ __launch_bounds__(TB) void k_tail_lds(tail_levels T, tail_layout Y, const int *__restrict__ prog,
                                                 int nprog, const float *__restrict__ img, const lv_cg_scalars *S)
{
    __shared__ __align__(16) float lds[TAIL_LDS];
    __shared__ int lprog[LV_AMG_MAXPROG];
    const int first = T.first;
    constexpr int PER4 = (TAIL_LDS / 4 + TB - 1) / TB;  // float4s per lane
    const int n4 = Y.coef / 4;
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 *img4 = reinterpret_cast<const f4 *>(img);
    f4 v[PER4];
#pragma unroll
    for (int q = 0; q < PER4; q++)
        if (q * TB + (int)threadIdx.x < n4) v[q] = img4[q * TB + threadIdx.x];
    const int n0 = T.Kx[first] * T.Ky[first];
    const int f0 = threadIdx.x;  // n0 <= TAIL_MAX == TB: one right-hand side entry per lane
    const float r0 = (f0 < n0) ? T.r[first][f0] : 0.0f;
    const int pr = (int)threadIdx.x < nprog ? prog[threadIdx.x] : 0;  // nprog <= LV_AMG_MAXPROG <= TB
    if (S && S->done) return;  // tested after the loads above are issued (they read valid arrays)
    f4 *lds4 = reinterpret_cast<f4 *>(lds);
#pragma unroll
    for (int q = 0; q < PER4; q++)
        if (q * TB + (int)threadIdx.x < n4) lds4[q * TB + threadIdx.x] = v[q];
    for (int f = 4 * n4 + threadIdx.x; f < Y.coef; f += TB) lds[f] = img[f];
    if (f0 < n0) lds[Y.r[first] + f0] = r0;
    if ((int)threadIdx.x < nprog) lprog[threadIdx.x] = pr;
    __syncthreads();
    run_tail(T, tail_in_lds{lds, Y}, lprog, nprog);
    for (int f = threadIdx.x; f < n0; f += TB) T.t[first][f] = lds[Y.t[first] + f];
}
static_assert(B == 256 && FUSE_ROWS + 64 <= FUSE_CAP, "fused blocks fit their LDS rows");
static_assert(LV_AMG_DENSE_MAX * 16 == TB, "the dense solve of the tail maps 16 lanes to a row");
static_assert(TAIL_MAX <= TB && LV_AMG_MAXPROG <= TB, "one right-hand side entry and one operation per lane");
static_assert(TAIL_LDS % 4 == 0, "the LDS image is copied as float4");

// the tail levels' constant arrays (diagonals, couplings, dense inverse) into img in the LDS layout
__global__ __launch_bounds__(B) void k_tail_pack(tail_levels T, tail_layout Y, const float *__restrict__ inv,
                                                 float *__restrict__ img)
{
    const int stride = gridDim.x * B;
    for (int l = T.first; l < T.last; l++) {
        const int n = T.Kx[l] * T.Ky[l];
        for (int f = blockIdx.x * B + threadIdx.x; f < n; f += stride) {
            img[Y.d[l] + f] = T.d[l][f];
            img[Y.dinv[l] + f] = T.dinv[l][f];
        }
        for (int f = blockIdx.x * B + threadIdx.x; f < 8 * n; f += stride) img[Y.c[l] + f] = T.c[l][f];
    }
    const int nd = T.Kx[T.last] * T.Ky[T.last];
    for (int f = blockIdx.x * B + threadIdx.x; f < nd * nd; f += stride) img[Y.inv + f] = inv[f];
}

// ---- power iteration for the scale of the learned fine smoother ----------------------------------

// +-1 probe hashed from the particle index (list: the particle of each row in the aggregate order)
__global__ __launch_bounds__(B) void k_pw_init(int n, const int *__restrict__ list, float *__restrict__ v)
{
    const int p = blockIdx.x * B + threadIdx.x;
    if (p >= n) return;
    const int i = list ? list[p] : p;
    uint64_t z = 0x5EED5EEDull + ((uint64_t)i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v[p] = (z >> 63) ? 1.0f : -1.0f;
}

// u = M (A v) computed row-wise through the work vector t = A v: two launches
template <class V>
__global__ __launch_bounds__(B) void k_pw_av(fine_ops<V> F, const float *__restrict__ v, float *__restrict__ t)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < F.n) t[i] = a_row(i, F, v);
}
// u = M t; partials |u|^2, |v|^2
template <class V>
__global__ __launch_bounds__(B) void k_pw_mt(fine_ops<V> F, const float *__restrict__ t,
                                             const float *__restrict__ v, float *__restrict__ u,
                                             double *__restrict__ part)
{
    double s[2] = {0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < F.n; i += gridDim.x * B) {
        const float ui = m_row(i, F, t);
        u[i] = ui;
        s[0] += (double)ui * (double)ui;
        s[1] += (double)v[i] * (double)v[i];
    }
    block_partials<2>(s, part, gridDim.x);
}
__global__ __launch_bounds__(B) void k_pw_norm(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    double o[2];
    reduce_all<2>(part, nb, o);
    if (threadIdx.x == 0) {
        out[0] = o[0];
        out[1] = o[1];
    }
}
// v = u / |u|
__global__ __launch_bounds__(B) void k_pw_scale(int n, const float *__restrict__ u, const double *__restrict__ nrm,
                                                float *__restrict__ v)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) v[i] = (float)((double)u[i] / sqrt(nrm[0]));
}

// ---- the aggregate order -----------------------------------------------------------------------

// iperm[list[p]] = p and the level-1 cell of row p
__global__ __launch_bounds__(B) void k_perm_index(int n, const int *__restrict__ list, const int *__restrict__ agg,
                                                  int *__restrict__ iperm, int *__restrict__ aggp)
{
    const int p = blockIdx.x * B + threadIdx.x;
    if (p >= n) return;
    const int i = list[p];
    iperm[i] = p;
    aggp[p] = agg[i];
}

// row p of the aggregate order from row list[p] of the operator (same slot order): degree, column
// offsets taken modulo n into [-n/2, n/2], FP32 and BF16 couplings, d, D^-1, and the learned M;
// slots LV_OVF_BASE.. (offsets, FP32 couplings) also into the compact overflow of p's 64-row group
// when the group has at most ocap overflow slots (lv_ell.h); *flag is raised when an offset needs
// more than 16 bits
__global__ __launch_bounds__(B) void k_perm_ell(int n, const int *__restrict__ list, const int *__restrict__ iperm,
                                                const int *__restrict__ deg, const int *__restrict__ ecol,
                                                const float *__restrict__ ev, const float *__restrict__ diag,
                                                const float *__restrict__ dinv, const float *__restrict__ mval,
                                                const float *__restrict__ mdg, uint8_t *__restrict__ pdeg,
                                                int16_t *__restrict__ poff, float *__restrict__ pev32,
                                                uint16_t *__restrict__ pev16, float *__restrict__ pdiag,
                                                float *__restrict__ pdinv, float *__restrict__ pmval,
                                                float *__restrict__ pmdg, int *__restrict__ flag,
                                                int16_t *__restrict__ ooff, float *__restrict__ oev32, int ocap)
{
    const int p = blockIdx.x * B + threadIdx.x;
    // everything of row i loaded as soon as i is known, unconditionally (a lane past n takes row
    // list[n-1] and drops it): the degree, the diagonals and the first U slots' columns and
    // couplings; then their rows' new positions gathered together (a slot past the degree reads
    // the allocated slot, and its column is taken as the row's own so that the gather stays in
    // bounds; neither is used)
    const int i = list[p < n ? p : n - 1];
    constexpr int U = 8;
    const int di = deg[i];
    const float dgi = diag[i], dvi = dinv[i];
    int cj[U];
    float vj[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        const size_t e = (size_t)k * n + i;
        cj[k] = ecol[e];
        vj[k] = ev[e];
    }
    const int d = p < n ? di : 0;
    const lv_ovf_pos q = lv_ovf_prefix(d);  // the block's wavefronts are whole 64-row groups
    if (p >= n) return;
    const bool compact = ocap > 0 && q.tot <= ocap;
    const size_t g0 = (size_t)(p >> 6) * LV_OVF_CAP;
    pdeg[p] = (uint8_t)d;
    pdiag[p] = dgi;
    pdinv[p] = dvi;
    if (mdg) pmdg[p] = mdg[i];
    bool wide = false;
    int pj[U];
#pragma unroll
    for (int k = 0; k < U; k++) pj[k] = k < d ? cj[k] : i;
#pragma unroll
    for (int k = 0; k < U; k++) pj[k] = iperm[pj[k]];
    auto put = [&](int k, int col_new, float v) {
        const size_t e = (size_t)k * n + i, ep = (size_t)k * n + p;
        int off = col_new - p;
        if (off > n / 2) off -= n;
        if (off < -(n / 2)) off += n;
        wide |= (off < -32768 || off > 32767);
        poff[ep] = (int16_t)off;
        pev32[ep] = v;
        pev16[ep] = lv_bf16_bits(v);
        if (mval) pmval[ep] = mval[e];
        if (compact && k >= LV_OVF_BASE) {
            const size_t eo = g0 + (size_t)(q.pos + (k - LV_OVF_BASE));
            ooff[eo] = (int16_t)off;
            oev32[eo] = v;
        }
    };
#pragma unroll
    for (int k = 0; k < U; k++)
        if (k < d) put(k, pj[k], vj[k]);
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        put(k, iperm[ecol[e]], ev[e]);
    }
    if (wide) atomicOr(flag, 1);
}

}  // namespace

#define AMG_CHECK(expr)                  \
    do {                                 \
        hipError_t _e = (expr);          \
        if (_e != hipSuccess) return _e; \
    } while (0)

template <typename T>
static hipError_t grow(T **p, size_t *cap, size_t want)
{
    if (*cap >= want && *p) return hipSuccess;
    if (*p) AMG_CHECK(hipFree(*p));
    *p = nullptr;
    AMG_CHECK(hipMalloc((void **)p, sizeof(T) * (want ? want : 1)));
    *cap = want;
    return hipSuccess;
}

hipError_t lv_amg_init(lv_amg *G)
{
    memset(G, 0, sizeof(*G));
    AMG_CHECK(hipMalloc(&G->dense_inv, sizeof(float) * LV_AMG_DENSE_MAX * LV_AMG_DENSE_MAX));
    AMG_CHECK(hipMalloc(&G->tail_img, sizeof(float) * TAIL_LDS));
    AMG_CHECK(hipMalloc(&G->pw, sizeof(double) * (2 * LV_NBMAX + 2)));
    AMG_CHECK(hipMalloc(&G->prog, sizeof(int) * LV_AMG_MAXPROG));
    AMG_CHECK(hipMalloc(&G->pflag, 2 * sizeof(int)));
    return hipSuccess;
}

void lv_amg_free(lv_amg *G)
{
    void *ptrs[] = {G->agg, G->cnt, G->start, G->cursor, G->list, G->scan_tmp, G->t0, G->z0, G->dense_inv, G->tail_img,
                    G->pw, G->pv, G->pu, G->prog, G->iperm, G->aggp, G->pdeg, G->poff, G->pev32, G->pev16,
                    G->pdiag32, G->pdinv32, G->pmval, G->pmdg, G->pflag, G->bfirst, G->ooff, G->oev32};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int l = 0; l <= LV_AMG_MAXLEV; l++) {
        lv_amg_level &L = G->lev[l];
        void *q[] = {L.d, L.c, L.dinv, L.r, L.z, L.t, L.acc};
        for (void *p : q)
            if (p) (void)hipFree(p);
    }
    memset(G, 0, sizeof(*G));
}

static hipError_t level_alloc(lv_amg_level &L, int Kx, int Ky)
{
    L.Kx = Kx;
    L.Ky = Ky;
    L.n = Kx * Ky;
    if (L.cap < (size_t)L.n || !L.d) {
        void *q[] = {L.d, L.c, L.dinv, L.r, L.z, L.t, L.acc};
        for (void *p : q)
            if (p) AMG_CHECK(hipFree(p));
        const size_t m = (size_t)L.n;
        AMG_CHECK(hipMalloc(&L.d, sizeof(float) * m));
        AMG_CHECK(hipMalloc(&L.c, sizeof(float) * 8 * m));
        AMG_CHECK(hipMalloc(&L.dinv, sizeof(float) * m));
        AMG_CHECK(hipMalloc(&L.r, sizeof(float) * m));
        AMG_CHECK(hipMalloc(&L.z, sizeof(float) * m));
        AMG_CHECK(hipMalloc(&L.t, sizeof(float) * m));
        AMG_CHECK(hipMalloc(&L.acc, sizeof(float) * m));
        L.cap = m;
    }
    return hipSuccess;
}

static fine_ops<lv_ell_own<float>> fine_own(const lv_amg *G, const lv_pressure_system &A)
{
    fine_ops<lv_ell_own<float>> F;
    F.n = G->n0;
    F.A = lv_ell_own<float>{A.deg, A.ecol, A.ev32};
    F.diag = A.diag32;
    F.dinv = A.dinv32;
    F.mval = G->fine_learned ? G->mval : nullptr;
    F.mdg = G->fine_learned ? G->mdg : nullptr;
    F.scale = G->fine_scale;
    F.alpha = G->fine_alpha;
    return F;
}

// the aggregate order with the couplings in V's precision (BF16 for the cycle, FP32 for the power
// iteration of the smoother scale)
template <class V, typename VAL>
static fine_ops<V> fine_agg(const lv_amg *G, const VAL *val)
{
    fine_ops<V> F;
    F.n = G->n0;
    F.A = V{G->pdeg, G->poff, val};
    if constexpr (std::is_same<VAL, float>::value) {  // the compact overflow holds the FP32 couplings
        F.A.ocol = G->ooff;
        F.A.oval = G->oev32;
        F.A.ocap = G->ovf_cap;
    }
    F.diag = G->pdiag32;
    F.dinv = G->pdinv32;
    F.mval = G->fine_learned ? G->pmval : nullptr;
    F.mdg = G->fine_learned ? G->pmdg : nullptr;
    F.scale = G->fine_scale;
    F.alpha = G->fine_alpha;
    return F;
}

// 10 steps of power iteration on M A (fine view F): the largest eigenvalue's estimate, 0 if none
template <class V>
static hipError_t power_lambda(lv_amg *G, const fine_ops<V> &Fo, const int *list, double *lam, hipStream_t s)
{
    const int n = G->n0;
    const int nb = lv_nblocks(n);
    double *part = G->pw, *nrm = G->pw + 2 * LV_NBMAX;
    hipLaunchKernelGGL(k_pw_init, g1(n), dim3(B), 0, s, n, list, G->pv);
    double h[2] = {0.0, 0.0};
    for (int it = 0; it < 10; it++) {
        hipLaunchKernelGGL(k_pw_av<V>, g1(n), dim3(B), 0, s, Fo, (const float *)G->pv, G->t0);
        hipLaunchKernelGGL(k_pw_mt<V>, dim3(nb), dim3(B), 0, s, Fo, (const float *)G->t0, (const float *)G->pv,
                           G->pu, part);
        hipLaunchKernelGGL(k_pw_norm, dim3(1), dim3(B), 0, s, (const double *)part, nb, nrm);
        hipLaunchKernelGGL(k_pw_scale, g1(n), dim3(B), 0, s, n, (const float *)G->pu, (const double *)nrm, G->pv);
    }
    AMG_CHECK(hipMemcpyAsync(h, nrm, sizeof(h), hipMemcpyDeviceToHost, s));
    AMG_CHECK(hipStreamSynchronize(s));
    *lam = (h[1] > 0.0) ? sqrt(h[0] / h[1]) : 0.0;
    return hipSuccess;
}

// the operations of the single-workgroup kernel for the visit of level l (last: the dense level);
// with W, every coarse level below l is visited twice, the second time on the residual the first
// correction leaves (OP_WSAVE), the two corrections summed (OP_WADD).  false if the program would
// not fit.
static bool tail_visit(int l, int last, bool W, int *prog, int &np)
{
    auto emit = [&](int op, int lev) {
        if (np >= LV_AMG_MAXPROG) return false;
        prog[np++] = (op << OP_SHIFT) | lev;
        return true;
    };
    if (l == last) return emit(OP_DENSE, l);
    if (!emit(OP_DOWN, l) || !tail_visit(l + 1, last, W, prog, np)) return false;
    if (W && l + 1 < last)
        if (!emit(OP_WSAVE, l + 1) || !tail_visit(l + 1, last, W, prog, np) || !emit(OP_WADD, l + 1)) return false;
    return emit(OP_UP, l);
}

// the program of the tail: the visit of its first level, itself W-visited when W
static int tail_program(int first, int last, bool W, int *prog)
{
    int np = 0;
    bool ok = tail_visit(first, last, W, prog, np);
    if (ok && W && first < last) {
        ok = np < LV_AMG_MAXPROG;
        if (ok) prog[np++] = (OP_WSAVE << OP_SHIFT) | first;
        ok = ok && tail_visit(first, last, W, prog, np) && np < LV_AMG_MAXPROG;
        if (ok) prog[np++] = (OP_WADD << OP_SHIFT) | first;
    }
    if (ok) return np;
    np = 0;
    tail_visit(first, last, false, prog, np);
    return np;
}

// the levels ltail..nlev of the single-workgroup kernel
static tail_levels tail_levels_of(const lv_amg *G)
{
    tail_levels T;
    memset(&T, 0, sizeof(T));
    T.first = G->ltail;
    T.last = G->nlev;
    T.xper = G->xper;
    T.yper = G->yper;
    for (int l = G->ltail; l <= G->nlev; l++) {
        const lv_amg_level &L = G->lev[l];
        T.Kx[l] = L.Kx;
        T.Ky[l] = L.Ky;
        T.alpha[l] = L.alpha;
        T.acc[l] = L.acc;
        T.d[l] = L.d;
        T.c[l] = L.c;
        T.dinv[l] = L.dinv;
        T.r[l] = L.r;
        T.z[l] = L.z;
        T.t[l] = L.t;
    }
    return T;
}

// their LDS layout: the constant arrays first (diagonals, couplings, dense inverse), then the work
// vectors; about 14 floats per cell of the levels above the dense one, whose total stays below
// twice TAIL_MAX
static tail_layout tail_layout_of(const lv_amg *G)
{
    tail_layout Y;
    memset(&Y, 0, sizeof(Y));
    int off = 0;
    auto put = [&off](int count) {
        const int o = off;
        off += count;
        return o;
    };
    for (int l = G->ltail; l < G->nlev; l++) {
        const int nl = G->lev[l].n;
        Y.d[l] = put(nl);
        Y.dinv[l] = put(nl);
        Y.c[l] = put(8 * nl);
    }
    Y.inv = put(G->lev[G->nlev].n * G->lev[G->nlev].n);
    Y.coef = off;
    for (int l = G->ltail; l <= G->nlev; l++) {
        const int nl = G->lev[l].n;
        if (l < G->nlev) {
            Y.z[l] = put(nl);
            Y.acc[l] = put(nl);
        }
        Y.r[l] = put(nl);
        Y.t[l] = put(nl);
    }
    Y.total = off;
    return Y;
}

// the tail runs in LDS when its levels fit (the usual case), else in global memory
static bool tail_fits_lds(const lv_amg *G, const tail_layout &Y)
{
    return Y.total <= TAIL_LDS && G->nprog <= LV_AMG_MAXPROG;
}

hipError_t lv_amg_setup(lv_amg *G, const lv_grid_params &g, const double *x, const lv_pressure_system &A, int n,
                        const float *mval, const float *mdg, int layout, hipStream_t s)
{
    G->n0 = n;
    G->xper = g.xperiodic;
    G->yper = g.yperiodic;
    const double Lx = g.bxmax0 - g.bxmin0, Ly = g.bxmax1 - g.bxmin1;
    const int Kx = std::max(1, (int)floor(Lx / (CELL_DR * g.dr))), Ky = std::max(1, (int)floor(Ly / (CELL_DR * g.dr)));
    const int na = Kx * Ky;
    G->na = na;
    AMG_CHECK(grow(&G->agg, &G->cap_agg, (size_t)n));
    AMG_CHECK(grow(&G->list, &G->cap_list, (size_t)n));
    AMG_CHECK(grow(&G->t0, &G->cap_t0, (size_t)n));
    AMG_CHECK(grow(&G->z0, &G->cap_z0, (size_t)n));
    AMG_CHECK(grow(&G->cnt, &G->cap_cnt, (size_t)na + 1));
    AMG_CHECK(grow(&G->start, &G->cap_start, (size_t)na + 1));
    AMG_CHECK(grow(&G->cursor, &G->cap_cursor, (size_t)na + 1));
    AMG_CHECK(grow((char **)&G->scan_tmp, &G->cap_scan, lv_scan_temp_bytes((int64_t)na + 1)));
    AMG_CHECK(hipMemsetAsync(G->cnt, 0, sizeof(int) * (na + 1), s));
    AMG_CHECK(hipMemsetAsync(G->cursor, 0, sizeof(int) * (na + 1), s));
    hipLaunchKernelGGL(k_agg_key, g1(n), dim3(B), 0, s, n, x, g.bxmin0, g.bxmin1, Lx / Kx, Ly / Ky, Kx, Ky, G->agg,
                       G->cnt);
    AMG_CHECK(lv_exclusive_scan_i32(G->cnt, G->start, (int64_t)na + 1, nullptr, G->scan_tmp, s));
    hipLaunchKernelGGL(k_agg_fill, g1(n), dim3(B), 0, s, n, (const int *)G->agg, (const int *)G->start, G->cursor,
                       G->list);
    AMG_CHECK(level_alloc(G->lev[1], Kx, Ky));
    hipLaunchKernelGGL(k_agg_sort, g1(na), dim3(B), 0, s, na, (const int *)G->start, G->list);
    hipLaunchKernelGGL(k_coarse1, g1((int64_t)na * RL), dim3(B), 0, s, na, Kx, Ky, G->xper, G->yper,
                       (const int *)G->start, (const int *)G->list,
                       (const int *)G->agg, n, (const int *)A.deg, (const int *)A.ecol, (const double *)A.ev64,
                       (const double *)A.diag64, G->lev[1].d, G->lev[1].c, G->lev[1].dinv);
    int l = 1;
    while (G->lev[l].n > LV_AMG_DENSE_MAX && l < LV_AMG_MAXLEV) {
        const lv_amg_level &F = G->lev[l];
        AMG_CHECK(level_alloc(G->lev[l + 1], (F.Kx + CF - 1) / CF, (F.Ky + CF - 1) / CF));
        lv_amg_level &C = G->lev[l + 1];
        hipLaunchKernelGGL(k_coarsen, g1(C.n), dim3(B), 0, s, C.Kx, C.Ky, F.Kx, F.Ky, G->xper, G->yper,
                           (const float *)F.d, (const float *)F.c, C.d, C.c, C.dinv);
        l++;
    }
    if (G->lev[l].n > LV_AMG_DENSE_MAX) return hipErrorInvalidValue;
    G->nlev = l;
    // over-correction factors, the tail's first level and its program (a W-cycle for a singular
    // system, whose coarse levels carry no compressibility term to damp the smooth error)
    G->fine_alpha = G->lev[1].n > TAIL_MAX ? ALPHA_FINE : 1.0f;
    for (int k = 1; k <= G->nlev; k++)
        G->lev[k].alpha = (k < G->nlev && G->lev[k + 1].n > TAIL_MAX) ? ALPHA_COARSE : 1.0f;
    int lt = G->nlev;
    while (lt > 1 && G->lev[lt - 1].n <= TAIL_MAX) lt--;
    G->ltail = lt;
    // pairs of levels above the tail that run as one kernel each way, from the tail upwards (the
    // smallest levels gain most): level l at least PAIR_MIN_K wide, a multiple of 4 on a periodic axis
    memset(G->pair, 0, sizeof(G->pair));
    for (int k = lt - 2; k >= 2;) {
        const lv_amg_level &L = G->lev[k];
        const bool ok = L.Kx >= PAIR_MIN_K && L.Ky >= PAIR_MIN_K && (!G->xper || L.Kx % 4 == 0) &&
                        (!G->yper || L.Ky % 4 == 0);
        G->pair[k] = ok ? 1 : 0;
        k -= ok ? 2 : 1;
    }
    G->wcycle = A.singular ? 1 : 0;
    G->nprog = tail_program(lt, G->nlev, G->wcycle != 0, G->hprog);
    AMG_CHECK(hipMemcpyAsync(G->prog, G->hprog, sizeof(int) * G->nprog, hipMemcpyHostToDevice, s));
    const lv_amg_level &D = G->lev[l];
    hipLaunchKernelGGL(k_dense_inverse, dim3(1), dim3(DTB), 0, s, D.Kx, D.Ky, G->xper, G->yper, (const float *)D.d,
                       (const float *)D.c, A.singular, G->dense_inv);
    {
        const tail_layout Y = tail_layout_of(G);
        if (tail_fits_lds(G, Y))
            hipLaunchKernelGGL(k_tail_pack, dim3(64), dim3(B), 0, s, tail_levels_of(G), Y,
                               (const float *)G->dense_inv, G->tail_img);
    }
    // fine smoother
    G->fine_learned = (mval && mdg) ? 1 : 0;
    G->mval = mval;
    G->mdg = mdg;
    G->fine_scale = OMEGA;
    // the aggregate order: the rows of each level-1 cell consecutive (list, sorted within a cell)
    G->perm_ok = 0;
    G->nbfuse = 0;
    G->z0_in = 0;
    if (layout == 0 || layout >= 2) {
        const size_t m = (size_t)LV_MAXE * n;
        AMG_CHECK(grow(&G->iperm, &G->cap_iperm, (size_t)n));
        AMG_CHECK(grow(&G->aggp, &G->cap_aggp, (size_t)n));
        AMG_CHECK(grow(&G->pdeg, &G->cap_pdeg, (size_t)n));
        AMG_CHECK(grow(&G->poff, &G->cap_poff, m));
        AMG_CHECK(grow(&G->pev32, &G->cap_pev32, m));
        AMG_CHECK(grow(&G->pev16, &G->cap_pev16, m));
        const size_t mo = ((size_t)n / 64 + 1) * LV_OVF_CAP;
        AMG_CHECK(grow(&G->ooff, &G->cap_ooff, mo));
        AMG_CHECK(grow(&G->oev32, &G->cap_oev32, mo));
        // (below 65536 rows launches, not the padding, set the kernels' time: no compact copy)
        G->ovf_cap = (layout == 0 && n >= 65536) ? LV_OVF_CAP : layout == 3 ? 16 : 0;
        AMG_CHECK(grow(&G->pdiag32, &G->cap_pdiag32, (size_t)n));
        AMG_CHECK(grow(&G->pdinv32, &G->cap_pdinv32, (size_t)n));
        if (G->fine_learned) {
            AMG_CHECK(grow(&G->pmval, &G->cap_pmval, m));
            AMG_CHECK(grow(&G->pmdg, &G->cap_pmdg, (size_t)n));
        }
        AMG_CHECK(hipMemsetAsync(G->pflag, 0, 2 * sizeof(int), s));
        hipLaunchKernelGGL(k_perm_index, g1(n), dim3(B), 0, s, n, (const int *)G->list, (const int *)G->agg, G->iperm,
                           G->aggp);
        hipLaunchKernelGGL(k_perm_ell, g1(n), dim3(B), 0, s, n, (const int *)G->list, (const int *)G->iperm,
                           (const int *)A.deg, (const int *)A.ecol, (const float *)A.ev32, (const float *)A.diag32,
                           (const float *)A.dinv32, G->fine_learned ? mval : nullptr,
                           G->fine_learned ? mdg : nullptr, G->pdeg, G->poff, G->pev32, G->pev16, G->pdiag32,
                           G->pdinv32, G->pmval, G->pmdg, G->pflag, G->ooff, G->oev32, G->ovf_cap);
        // the blocks of the fused down pass (Jacobi smoother)
        const int nbf = n / FUSE_ROWS + 1;
        AMG_CHECK(grow(&G->bfirst, &G->cap_bfirst, 2 * ((size_t)nbf + 2)));
        hipLaunchKernelGGL(k_fuse_blocks, g1((int64_t)na + 1), dim3(B), 0, s, na, nbf, (const int *)G->start,
                           reinterpret_cast<int2 *>(G->bfirst), G->pflag);
        int flags[2] = {1, 1};
        AMG_CHECK(hipMemcpyAsync(flags, G->pflag, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        AMG_CHECK(hipStreamSynchronize(s));
        G->perm_ok = flags[0] ? 0 : 1;
        G->nbfuse = (G->perm_ok && !flags[1]) ? nbf : 0;
        G->z0_in = (G->nbfuse > 0 && !G->fine_learned) ? 1 : 0;
    }
    // the learned smoother's scale: lambda_max(M A) -> FINE_TARGET
    if (G->fine_learned) {
        AMG_CHECK(grow(&G->pv, &G->cap_pv, (size_t)n));
        AMG_CHECK(grow(&G->pu, &G->cap_pu, (size_t)n));
        G->fine_scale = 1.0f;
        double lam = 0.0;
        if (G->perm_ok)
            AMG_CHECK(power_lambda(G, fine_agg<lv_ell_agg32>(G, (const float *)G->pev32), G->list, &lam, s));
        else
            AMG_CHECK(power_lambda(G, fine_own(G, A), nullptr, &lam, s));
        G->fine_scale = (lam > 0.0 && lam == lam) ? (float)(FINE_TARGET / lam) : 0.0f;
        if (!(G->fine_scale > 0.0f)) {  // no usable estimate: damped Jacobi
            G->fine_learned = 0;
            G->fine_scale = OMEGA;
        }
    }
    return hipGetLastError();
}

// levels l and l+1 for the pair kernels
static st_pair st_pair_of(const lv_amg *G, int l)
{
    const lv_amg_level &A = G->lev[l], &C = G->lev[l + 1], &C2 = G->lev[l + 2];
    return st_pair{st_level{A.Kx, A.Ky, G->xper, G->yper, A.d, A.c, A.dinv, A.alpha},
                   st_level{C.Kx, C.Ky, G->xper, G->yper, C.d, C.c, C.dinv, C.alpha}, C2.Kx, C2.Ky};
}

// the cycle on a fine view F; agg: level-1 cell of each fine row; list: the fine rows of each cell
// (null when they are consecutive)
template <class V>
static hipError_t vcycle(const lv_amg *G, const fine_ops<V> &F, const int *agg, const int *list, const float *r,
                         float *z, double *Prr, const lv_cg_scalars *S, hipStream_t s, bool z0_ready)
{
    const int n = G->n0;
    const lv_amg_level &L1 = G->lev[1];
    bool z1_given = false;  // the fused down pass stored level 1's pre-smoothed z
    // level 0, down: z0 (in G->z0) and its residual t (in G->t0), restricted to level 1
    if (G->fine_learned) {
        hipLaunchKernelGGL(k_fine_pre<V>, g1(n), dim3(B), 0, s, F, r, G->z0, S);
        hipLaunchKernelGGL(k_fine_resid<V>, g1(n), dim3(B), 0, s, F, r, (const float *)G->z0, G->t0, S);
        hipLaunchKernelGGL(k_restrict1, g1((int64_t)G->na * RL), dim3(B), 0, s, G->na, (const int *)G->start, list,
                           (const float *)G->t0, L1.r, S);
    } else if (!list && G->nbfuse > 0) {
        // level 1 above the tail: its pre-smoothed z stored here, gathered by its down pass
        z1_given = G->ltail > 1;
        float *z1 = z1_given ? L1.z : nullptr;
        const int *done = S ? &S->done : G->bfirst + 2 * (G->nbfuse + 1);
        if (z0_ready && G->z0_in)
            hipLaunchKernelGGL((k_fine_down_jr<V, true>), dim3(G->nbfuse), dim3(B), 0, s, F, r, G->z0,
                               (const int *)G->start, reinterpret_cast<const int2 *>(G->bfirst), G->na, L1.r,
                               (const float *)L1.dinv, z1, done);
        else
            hipLaunchKernelGGL((k_fine_down_jr<V, false>), dim3(G->nbfuse), dim3(B), 0, s, F, r, G->z0,
                               (const int *)G->start, reinterpret_cast<const int2 *>(G->bfirst), G->na, L1.r,
                               (const float *)L1.dinv, z1, done);
    } else {
        hipLaunchKernelGGL(k_fine_down_j<V>, g1(n), dim3(B), 0, s, F, r, G->z0, G->t0, S);
        hipLaunchKernelGGL(k_restrict1, g1((int64_t)G->na * RL), dim3(B), 0, s, G->na, (const int *)G->start, list,
                           (const float *)G->t0, L1.r, S);
    }
    // coarse levels above the tail, down: each level's pre-smoothed z and the next level's
    // right-hand side
    const int lt = G->ltail;
    const bool pairs = !G->level_kernels;
    for (int l = 1; l < lt; l++) {
        const lv_amg_level &L = G->lev[l], &C = G->lev[l + 1];
        const st_level sl{L.Kx, L.Ky, G->xper, G->yper, L.d, L.c, L.dinv, L.alpha};
        if (pairs && G->pair[l]) {
            const lv_amg_level &C2 = G->lev[l + 2];
            const st_pair q = st_pair_of(G, l);
            hipLaunchKernelGGL(k_st_down2, dim3((C2.Kx + PD - 1) / PD, (C2.Ky + PD - 1) / PD), dim3(B), 0, s, q,
                               (const float *)L.r, L.z, C.r, C.z, C2.r, S);
            l++;
            continue;
        }
        if (l == 1 && z1_given)
            hipLaunchKernelGGL(k_st_down<true>, g1((int64_t)C.n * CF2), dim3(B), 0, s, sl, C.Ky, C.n,
                               (const float *)L.r, L.z, C.r, S);
        else
            hipLaunchKernelGGL(k_st_down<false>, g1((int64_t)C.n * CF2), dim3(B), 0, s, sl, C.Ky, C.n,
                               (const float *)L.r, L.z, C.r, S);
    }
    // the tail: levels lt..nlev down, dense, and up; the final correction of level lt in its t
    const tail_levels T = tail_levels_of(G);
    const tail_layout Y = tail_layout_of(G);
    if (tail_fits_lds(G, Y))
        hipLaunchKernelGGL(k_tail_lds, dim3(1), dim3(TB), 0, s, T, Y, (const int *)G->prog, G->nprog,
                           (const float *)G->tail_img, S);
    else
        hipLaunchKernelGGL(k_tail, dim3(1), dim3(TB), 0, s, T, (const int *)G->prog, G->nprog,
                           (const float *)G->dense_inv, S);
    // coarse levels above the tail, up: each level's final correction in t; level 1 in the aggregate
    // order with the Jacobi smoother also prolongates into z0
    const bool prolong_in_up = lt > 1 && !G->fine_learned && list == nullptr;
    for (int l = lt - 1; l >= 1; l--) {
        if (pairs && l >= 3 && G->pair[l - 1]) {
            const lv_amg_level &L = G->lev[l - 1], &C = G->lev[l], &C2 = G->lev[l + 1];
            hipLaunchKernelGGL(k_st_up2, dim3((L.Kx + PU - 1) / PU, (L.Ky + PU - 1) / PU), dim3(B), 0, s,
                               st_pair_of(G, l - 1), (const float *)L.r, (const float *)L.z, (const float *)C.r,
                               (const float *)C.z, (const float *)C2.t, L.t, S);
            l--;
            continue;
        }
        const lv_amg_level &L = G->lev[l], &C = G->lev[l + 1];
        const st_level sl{L.Kx, L.Ky, G->xper, G->yper, L.d, L.c, L.dinv, L.alpha};
        if (l == 1 && prolong_in_up)
            hipLaunchKernelGGL(k_st_up<true>, g1(L.n), dim3(B), 0, s, sl, C.Ky, (const float *)L.r,
                               (const float *)L.z, (const float *)C.t, L.t, S, (const int *)G->start, G->fine_alpha,
                               G->z0);
        else
            hipLaunchKernelGGL(k_st_up<false>, g1(L.n), dim3(B), 0, s, sl, C.Ky, (const float *)L.r,
                               (const float *)L.z, (const float *)C.t, L.t, S, (const int *)nullptr, 0.0f,
                               (float *)nullptr);
    }
    // level 0, up: prolongation, residual and post-smoothing; partials of r.z, r.r
    if (G->fine_learned) {
        hipLaunchKernelGGL(k_fine_up<V>, g1(n), dim3(B), 0, s, F, r, (const float *)G->z0, agg, (const float *)L1.t,
                           z, G->t0, S);
        hipLaunchKernelGGL(k_fine_post<V>, dim3(lv_nblocks(n)), dim3(B), 0, s, F, r, (const float *)G->t0, z, Prr, S);
    } else {
        if (!prolong_in_up)
            hipLaunchKernelGGL(k_prolong1, g1((n + PROLONG_RPT - 1) / PROLONG_RPT), dim3(B), 0, s, n, agg,
                               (const float *)L1.t, G->fine_alpha, G->z0, S);
        hipLaunchKernelGGL(k_fine_up_j<V>, dim3(lv_nblocks(n)), dim3(B), 0, s, F, r, (const float *)G->z0, z, Prr, S);
    }
    return hipGetLastError();
}

hipError_t lv_amg_vcycle(const lv_amg *G, const lv_pressure_system &A, const float *r, float *z, double *Prr,
                         const lv_cg_scalars *S, hipStream_t s, bool z0_ready)
{
    if (G->perm_ok)
        return vcycle(G, fine_agg<lv_ell_agg16>(G, (const uint16_t *)G->pev16), (const int *)G->aggp,
                      (const int *)nullptr, r, z, Prr, S, s, z0_ready);
    return vcycle(G, fine_own(G, A), (const int *)G->agg, (const int *)G->list, r, z, Prr, S, s, false);
}
