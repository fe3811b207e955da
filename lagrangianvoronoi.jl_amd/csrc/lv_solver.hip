// lv_solver.hip -- the pressure system of find_pressure! (src/pressure.jl:89-224): CSR assembly of
// PressureOperator (refresh!(A), :104-117), the right-hand side (refresh!(solver), :162-203), and the
// Krylov solvers that replace Krylov.jl's minres! (:219):
//   * FP64 Jacobi-PCG;
//   * FP32 PCG inside FP64 iterative refinement (FP64 residual, FP64 dots, FP32 SpMV/AXPY), with
//     Jacobi or the learned preconditioner of lv_spai.hip (Jacobi-PCG redoes any sweep the learned
//     preconditioner fails to converge).
//
// The operator is kept in the reference's own form y_i = d_i x_i + sum_j c_ij (x_i - x_j): the
// difference form keeps the nearly-constant pressure modes accurate in FP32, which the folded
// diagonal form would lose to cancellation.  One Jacobi-PCG iteration is two launches: the SpMV of z
// fused with the p/q recurrences and the p.q dot (A p_new = A z + beta A p_old), and the x/r/z update
// fused with the r.z and r.r dots; the learned preconditioner adds one ELL pass z = M r.  Dots reduce
// per block into a partials array that every block of the consuming kernel re-reduces in the same
// fixed order (no atomics, no fences), so results are deterministic; alpha, beta and convergence are
// evaluated on the device, which also reports its progress to pinned host memory so that the host
// keeps one iteration queued ahead and stops launching at convergence without synchronising.
#include <string.h>

#include "lv_solver.h"
#include "lv_reduce.h"
#include "lv_amg.h"
#include "lv_ell.h"
#include <chrono>
#include <type_traits>

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

// refresh!(A, grid, dt), src/pressure.jl:104-117 (mode 0), or the operator of ViscousSolver's
// assemble!, src/viscous_solver.jl:89-113 (mode 1: d_i = m_i/dt, c_ij = 0.5 lrr (mu_i + mu_j))
// The P-independent parts of refresh!(solver) for the interior edge to y (src/pressure.jl:173-180,
// 189-197): w = lrr (m - x_i) and u = lrr (m - z) into the edge's mesh slot e; returns the edge's
// divergence term of b
__device__ __forceinline__ double rhs_edge_geom(const lv_rhs_geom &R, size_t e, double lrr, double x1, double x2,
                                                double y1, double y2, double m1, double m2, double dv1, double dv2,
                                                double dt)
{
    const double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
    R.w1[e] = lrr * (m1 - x1);
    R.w2[e] = lrr * (m2 - x2);
    R.u1[e] = lrr * (m1 - z1);
    R.u2[e] = lrr * (m2 - z2);
    return (lrr / dt) * (dv1 * (m1 - y1) + dv2 * (m2 - y2));
}

// the boundary-flux term of b for a boundary edge with label lab (src/pressure.jl:181-186)
__device__ __forceinline__ double rhs_bdary(const lv_vbc &vbc, int lab, double e1x, double e1y, double e2x,
                                            double e2y, double pv1, double pv2, double dt)
{
    const double dS1 = e1y - e2y, dS2 = e2x - e1x;
    const int slot = -lab - 1;
    const double vb1 = vbc.v[2 * slot], vb2 = vbc.v[2 * slot + 1];
    return (dS1 * (vb1 - pv1) + dS2 * (vb2 - pv2)) / dt;
}

// the right-hand side's edge slots LV_OVF_BASE.. of row i (just written to the planes by this
// thread; a boundary slot's values are whatever the planes hold, and unused) into the compact copy
// of its 64-row group, when the group has at most R.ocap of them.  Every active lane of the
// wavefront takes part in the ballots.
__device__ __forceinline__ void rhs_geom_compact(const lv_mesh_view &M, const lv_rhs_geom &R, int i, int n, int ne)
{
    const lv_ovf_pos q = lv_ovf_prefix(ne);
    if (q.tot > R.ocap) return;
    const size_t g0 = (size_t)(i >> 6) * LV_OVF_CAP;
    for (int k = LV_OVF_BASE; k < ne; k++) {
        const size_t e = (size_t)k * n + i, eo = g0 + (size_t)(q.pos + (k - LV_OVF_BASE));
        R.oel[eo] = LV_EL(M, k, i);
        R.ow1[eo] = R.w1[e];
        R.ow2[eo] = R.w2[e];
        R.ou1[eo] = R.u1[e];
        R.ou2[eo] = R.u2[e];
    }
}

constexpr int ACSR = 2048;  // CSR entries of a k_assemble block staged in LDS (32 KB)

// one row of k_assemble
__device__ __forceinline__ void assemble_row(const lv_grid_params &g, const lv_fields &F, const lv_mesh_view &M, int n,
                                             double dt, int mode, int i, const int *__restrict__ row_ptr, bool stage,
                                             int cbase, int *s_col, double *s_v64, float *s_v32,
                                             int *__restrict__ cols, double *__restrict__ val64,
                                             float *__restrict__ val32, double *__restrict__ diag64,
                                             float *__restrict__ diag32, double *__restrict__ dinv64,
                                             float *__restrict__ dinv32, int *__restrict__ ecol,
                                             double *__restrict__ ev64, float *__restrict__ ev32,
                                             int *__restrict__ nonsingular, const lv_vbc &vbc, const lv_rhs_geom &R,
                                             int *__restrict__ ocol, double *__restrict__ oev64,
                                             float *__restrict__ oev32, uint8_t *__restrict__ deg8,
                                             int16_t *__restrict__ ecol16, int16_t *__restrict__ ocol16,
                                             int *__restrict__ wide)
{
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double rho = F.rho[i];
    const double d = (mode == 0) ? F.mass[i] / (((rho * rho) * F.c2[i]) * (dt * dt)) : F.mass[i] / dt;
    const double mui = F.mu[i];
    const bool geom = R.w1 != nullptr;
    const double pv1 = F.v[i], pv2 = F.v[n + i];
    double dd = d, bd = 0.0;
    int o = row_ptr[i];
    // the solver's ELL copy keeps each row's neighbours sorted by column, so that slot k of a
    // wavefront gathers from one lattice neighbourhood; the CSR copy keeps the edge order
    int sc[LV_MAXE];
    double sv[LV_MAXE];
    int m = 0;
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        if (j < 0) {
            if (geom) bd -= rhs_bdary(vbc, j, e1x, e1y, e2x, e2y, pv1, pv2, dt);
            continue;
        }
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double c = (mode == 0) ? lrr * (0.5 / rho + 0.5 / F.rho[j]) : (0.5 * lrr) * (mui + F.mu[j]);
        if (stage) {
            s_col[o - cbase] = j;
            s_v64[o - cbase] = c;
            s_v32[o - cbase] = (float)c;
        } else {
            cols[o] = j;
            val64[o] = c;
            val32[o] = (float)c;
        }
        dd += c;
        o++;
        if (geom) {
            const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
            bd -= rhs_edge_geom(R, (size_t)k * n + i, lrr, x1, x2, y1, y2, m1, m2, pv1 - F.v[j], pv2 - F.v[n + j],
                                dt);
        }
        int q = m++;
        while (q > 0 && sc[q - 1] > j) {
            sc[q] = sc[q - 1];
            sv[q] = sv[q - 1];
            q--;
        }
        sc[q] = j;
        sv[q] = c;
    }
    // the 16-bit copy's offsets j - i, taken into [-(n/2), n/2] modulo n (lv_ell_col undoes it)
    auto off16 = [&](int j) {
        int off = j - i;
        if (off > n / 2) off -= n;
        if (off < -(n / 2)) off += n;
        return off;
    };
    bool w16 = false;
    for (int k = 0; k < m; k++) {
        const size_t es = (size_t)k * n + i;
        ecol[es] = sc[k];
        ev64[es] = sv[k];
        ev32[es] = (float)sv[k];
        const int off = off16(sc[k]);
        w16 |= (off < -32768 || off > 32767);
        ecol16[es] = (int16_t)off;
    }
    deg8[i] = (uint8_t)m;
    // slots LV_OVF_BASE.. also into the compact copy of the row's 64-row group (a block's
    // wavefronts are whole groups), when the group has at most LV_OVF_CAP of them
    const lv_ovf_pos q = lv_ovf_prefix(m);
    if (ocol && q.tot <= LV_OVF_CAP) {
        const size_t g0 = (size_t)(i >> 6) * LV_OVF_CAP;
        for (int k = LV_OVF_BASE; k < m; k++) {
            const size_t eo = g0 + (size_t)(q.pos + (k - LV_OVF_BASE));
            ocol[eo] = sc[k];
            oev64[eo] = sv[k];
            oev32[eo] = (float)sv[k];
            ocol16[eo] = (int16_t)off16(sc[k]);
        }
    }
    if (__any(w16) && (threadIdx.x & 63) == 0) atomicOr(wide, 1);
    if (geom) {
        R.bdiv[i] = bd;
        R.den[i] = (rho * F.c2[i]) * (dt * dt);
        if (R.ocap > 0) rhs_geom_compact(M, R, i, n, ne);
    }
    // the system is singular iff every d_i == 0 (c2 = Inf): one flag write per wavefront at most
    if (__any(d != 0.0) && (threadIdx.x & 63) == 0 && *nonsingular == 0) atomicOr(nonsingular, 1);
    diag64[i] = d;
    diag32[i] = (float)d;
    dinv64[i] = 1.0 / dd;
    dinv32[i] = (float)(1.0 / dd);
}

// refresh!(A) (src/pressure.jl:104-117): the CSR copy in edge order, the ELL copy with each row's
// columns sorted, diagonals and Jacobi inverses.  With R.w1 (mode 0, find_pressure!) the same edge
// loop also stores the P-independent parts of refresh!(solver) (vbc: boundary_velocity per label).
__global__ __launch_bounds__(B) void k_assemble(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt, int mode,
                                                const int *__restrict__ row_ptr, int *__restrict__ cols,
                                                double *__restrict__ val64, float *__restrict__ val32,
                                                double *__restrict__ diag64, float *__restrict__ diag32,
                                                double *__restrict__ dinv64, float *__restrict__ dinv32,
                                                int *__restrict__ ecol, double *__restrict__ ev64,
                                                float *__restrict__ ev32, int *__restrict__ nonsingular, lv_vbc vbc,
                                                lv_rhs_geom R, int *__restrict__ ocol, double *__restrict__ oev64,
                                                float *__restrict__ oev32, uint8_t *__restrict__ deg8,
                                                int16_t *__restrict__ ecol16, int16_t *__restrict__ ocol16,
                                                int *__restrict__ wide)
{
    // The block's rows own one contiguous span of the CSR copy; it is formed in LDS and stored with
    // coalesced writes after the barrier (each thread's own entries would be 64 scattered stores an
    // edge), unless it is longer than ACSR entries.
    __shared__ int s_col[ACSR];
    __shared__ double s_v64[ACSR];
    __shared__ float s_v32[ACSR];
    const int i = blockIdx.x * B + threadIdx.x;
    const int rb1 = (int)blockIdx.x * B + B < n ? (int)blockIdx.x * B + B : n;
    const int cbase = row_ptr[blockIdx.x * B], clen = row_ptr[rb1] - cbase;
    const bool stage = clen <= ACSR;
    if (i < n) {  // rows past n hold no slots: their lanes stay out of the group's ballots below
        assemble_row(g, F, M, n, dt, mode, i, row_ptr, stage, cbase, s_col, s_v64, s_v32, cols, val64, val32,
                     diag64, diag32, dinv64, dinv32, ecol, ev64, ev32, nonsingular, vbc, R, ocol, oev64, oev32,
                     deg8, ecol16, ocol16, wide);
    }
    if (stage) {
        __syncthreads();
        for (int t = threadIdx.x; t < clen; t += B) {
            cols[cbase + t] = s_col[t];
            val64[cbase + t] = s_v64[t];
            val32[cbase + t] = s_v32[t];
        }
    }
}

// ---- refresh!(solver) split by what depends on P ---------------------------------------------------
//
// During the Picard iterations of find_pressure! (src/pressure.jl:215-224) the mesh, v and the
// boundary velocity are fixed; only P changes.  So b = A P/(rho c2 dt^2) + bdiv + [it > 1] corr(GP) with
//   bdiv_i = -sum_j (lrr/dt) (v_i - v_j).(m - y) - sum_boundary dS.(vbc - v_i)/dt     (:173-182)
//   GP_i   = -(1/m_i) sum_j (P_i - P_j) w_ij,   w_ij = lrr (m - x_i)                   (:180, 184)
//   corr_i =  sum_j (GP_i - GP_j).u_ij,          u_ij = lrr (m - z)                     (:189-202)
// k_assemble (or k_rhs_geom) evaluates bdiv, the compressibility denominator and w, u once per
// find_pressure!, w and u into the mesh's edge slots.  Each Picard iteration then streams w or u and
// gathers P or GP only.  b's terms are grouped differently from the reference's single running sum
// (rounding-level differences).

// the same P-independent parts on their own, for an operator assembled without them (the order of
// the sums as in k_assemble)
__global__ __launch_bounds__(B) void k_rhs_geom(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                                lv_vbc vbc, lv_rhs_geom R)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double pv1 = F.v[i], pv2 = F.v[n + i];
    double bd = 0.0;
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        if (j < 0) {
            bd -= rhs_bdary(vbc, j, e1x, e1y, e2x, e2y, pv1, pv2, dt);
            continue;
        }
        double y1, y2;
        lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        bd -= rhs_edge_geom(R, (size_t)k * n + i, lrr, x1, x2, y1, y2, m1, m2, pv1 - F.v[j], pv2 - F.v[n + j], dt);
    }
    R.bdiv[i] = bd;
    R.den[i] = (F.rho[i] * F.c2[i]) * (dt * dt);
    if (R.ocap > 0) rhs_geom_compact(M, R, i, n, ne);
}

// The edge slots of row i for the Picard right-hand side: neighbour (the row itself on a boundary
// slot or past ne) and two weight planes (0 there).  Slots 0..LV_OVF_BASE-1 come from the planes,
// the later ones from the compact copy of the row's 64-row group when it holds them (q: the group's
// ballot prefix), else from the planes.  The first U slots are loaded unconditionally (both copies
// hold every slot they may read, the compact index clamped into the group's region), the pointers
// chosen by selects, so that the loads and the gathers that follow are issued together.
struct rhs_slots {
    const int *el;
    const double *a, *b;
    size_t g0;
    int pos, n, i;
    bool compact;
    __device__ size_t index(int k) const
    {
        if (k < LV_OVF_BASE || !compact) return (size_t)k * n + i;
        const int o = pos + (k - LV_OVF_BASE);
        return g0 + (size_t)(o < LV_OVF_CAP ? o : LV_OVF_CAP - 1);
    }
};
__device__ __forceinline__ rhs_slots rhs_slots_of(const lv_mesh_view &M, const double *a, const double *b,
                                                  const int *oel, const double *oa, const double *ob, int ocap,
                                                  lv_ovf_pos q, int i, int n)
{
    const bool compact = ocap > 0 && q.tot <= ocap;
    return rhs_slots{M.el, a, b, (size_t)(i >> 6) * LV_OVF_CAP, q.pos, n, i, compact};
}
template <int U>
__device__ __forceinline__ void rhs_head(const rhs_slots &S, const int *oel, const double *oa, const double *ob,
                                         int ne, int (&c)[U], double (&va)[U], double (&vb)[U])
{
    int j[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        const bool ov = k >= LV_OVF_BASE && S.compact;
        const size_t e = S.index(k);
        j[k] = (ov ? oel : S.el)[e];
        va[k] = (ov ? oa : S.a)[e];
        vb[k] = (ov ? ob : S.b)[e];
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
        const bool on = (k < ne) & (j[k] >= 0);  // (not &&: a branch here would sink the loads above into it)
        c[k] = on ? j[k] : S.i;
        va[k] = on ? va[k] : 0.0;
        vb[k] = on ? vb[k] : 0.0;
    }
}

// GP from P; with Pout also the write-back of P into the particles (find_pressure!'s p.P = P), with
// Pout2 a second copy of P (the Picard window's record of the last iterate)
__global__ __launch_bounds__(B) void k_gp(lv_mesh_view M, int n, lv_rhs_geom R, const double *__restrict__ P,
                                          const double *__restrict__ mass, double *__restrict__ GP,
                                          double *__restrict__ Pout, double *__restrict__ Pout2)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const int ne = i < n ? M.ne[i] : 0;
    const lv_ovf_pos q = R.ocap > 0 ? lv_ovf_prefix(ne) : lv_ovf_pos{0, 0};  // the whole wavefront, before any lane leaves
    if (i >= n) return;
    constexpr int U = 8;
    int c[U];
    double a1[U], a2[U], pj[U];
    const rhs_slots S = rhs_slots_of(M, R.w1, R.w2, R.oel, R.ow1, R.ow2, R.ocap, q, i, n);
    rhs_head<U>(S, R.oel, R.ow1, R.ow2, ne, c, a1, a2);
    const double Pi = P[i];
#pragma unroll
    for (int k = 0; k < U; k++) pj[k] = P[c[k]];
    double gx = 0.0, gy = 0.0;
#pragma unroll
    for (int k = 0; k < U; k++) {
        const double dp = Pi - pj[k];
        gx -= dp * a1[k];
        gy -= dp * a2[k];
    }
    for (int k = U; k < ne; k++) {
        const size_t e = S.index(k);
        const int j = (S.compact ? R.oel : M.el)[e];
        if (j < 0) continue;
        const double dp = Pi - P[j];
        gx -= dp * (S.compact ? R.ow1 : R.w1)[e];
        gy -= dp * (S.compact ? R.ow2 : R.w2)[e];
    }
    const double mi = mass[i];
    GP[i] = gx / mi;
    GP[n + i] = gy / mi;
    if (Pout) Pout[i] = Pi;
    if (Pout2) Pout2[i] = Pi;
}

// b = A P/(rho c2 dt^2) + bdiv (+ corr(GP) when gp); with P0 also the initial guess P0 = P
__global__ __launch_bounds__(B) void k_rhs_b(lv_mesh_view M, int n, int gp, lv_rhs_geom R,
                                             const double *__restrict__ GP, const double *__restrict__ area,
                                             const double *__restrict__ P, double *__restrict__ b,
                                             double *__restrict__ P0)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const int ne = i < n ? M.ne[i] : 0;
    const lv_ovf_pos q = (gp && R.ocap > 0) ? lv_ovf_prefix(ne) : lv_ovf_pos{0, 0};  // before any lane leaves
    if (i >= n) return;
    const double Pi = P[i];
    double bi = (area[i] * Pi) / R.den[i] + R.bdiv[i];
    if (gp) {
        constexpr int U = 8;
        int c[U];
        double a1[U], a2[U], g1[U], g2[U];
        const rhs_slots S = rhs_slots_of(M, R.u1, R.u2, R.oel, R.ou1, R.ou2, R.ocap, q, i, n);
        rhs_head<U>(S, R.oel, R.ou1, R.ou2, ne, c, a1, a2);
#pragma unroll
        for (int k = 0; k < U; k++) {
            g1[k] = GP[c[k]];
            g2[k] = GP[n + c[k]];
        }
        const double gi1 = GP[i], gi2 = GP[n + i];
        double corr = 0.0;
#pragma unroll
        for (int k = 0; k < U; k++) corr += (gi1 - g1[k]) * a1[k] + (gi2 - g2[k]) * a2[k];
        for (int k = U; k < ne; k++) {
            const size_t e = S.index(k);
            const int j = (S.compact ? R.oel : M.el)[e];
            if (j < 0) continue;
            corr += (gi1 - GP[j]) * (S.compact ? R.ou1 : R.u1)[e] + (gi2 - GP[n + j]) * (S.compact ? R.ou2 : R.u2)[e];
        }
        bi += corr;
    }
    b[i] = bi;
    if (P0) P0[i] = Pi;
}

__global__ __launch_bounds__(B) void k_write_P(lv_fields F, int n, const double *__restrict__ P)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) F.P[i] = P[i];
}

// single-block finalisers of the one-off reductions
enum { FIN_MEANS = 0, FIN_INIT64 = 1, FIN_RESID = 2, FIN_SUM1 = 3, FIN_INIT32 = 4, FIN_RESID0 = 5 };

__global__ __launch_bounds__(B) void k_finalize(const double *__restrict__ partials, int nb, int mode, int n,
                                                double inner_tol2, lv_cg_scalars *S, lv_progress *pub = nullptr,
                                                int seq = 0)
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
    case FIN_INIT64:  // P3 holds |b'|^2
        S->bb = o[0];
        S->thr2 = S->tol2 * o[0];
        S->iters = 0;
        S->done = 0;
        break;
    case FIN_RESID:
        S->rr64 = o[0];
        S->bb = o[1];
        S->thr2 = S->tol2 * o[1];
        if (pub) {  // the host's tests after a sweep, without a copy and a synchronisation
            pub->rr64 = S->rr64;
            pub->rr64_0 = S->rr64_0;
            pub->bb = S->bb;
            pub->thr2 = S->thr2;
            pub->iters = S->iters;
            pub->skip = S->skip;
            __threadfence_system();
            __hip_atomic_store(&pub->fseq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        break;
    case FIN_RESID0: {  // the initial residual: the host's tests before the first sweep, made here
        // (inner_tol2 carries the per-sweep floor of the inner tolerance, not squared)
        const double rr = o[0], thr2 = S->tol2 * o[1];
        S->rr64 = rr;
        S->rr64_0 = rr;
        S->bb = o[1];
        S->thr2 = thr2;
        S->skip = (rr <= thr2 || !(rr == rr)) ? 1 : 0;
        const double need = 0.8 * sqrt(thr2 / rr);
        double itol = need > inner_tol2 ? need : inner_tol2;
        if (itol > 0.8) itol = 0.8;
        S->itol2 = itol * itol;
        break;
    }
    case FIN_SUM1:
        S->xmean = o[0] / n;
        S->xshift = S->xmean - S->xmean0;
        break;
    case FIN_INIT32:  // reads the (r.z, r.r) partials of the inner init: |r32|^2 = 1 up to rounding
        S->thr2 = inner_tol2 * o[1];
        S->iters = 0;
        S->done = 0;
        break;
    }
}

// thr2 of an inner FP32 solve from the (r.z, r.r) partials of its init.  dev: the first sweep, whose
// inner tolerance and skip flag k_finalize (FIN_RESID0) left in S; a skipped sweep starts done, and
// its first k_cg_s reports that to the host.
__global__ __launch_bounds__(B) void k_finalize2(const double *__restrict__ Prr, int nb, double inner_tol2,
                                                 lv_cg_scalars *S, int dev)
{
    double o[2];
    reduce_all<2>(Prr, nb, o);
    if (threadIdx.x != 0) return;
    S->thr2 = (dev ? S->itol2 : inner_tol2) * o[1];
    S->iters = 0;
    S->done = (dev && S->skip) ? 1 : 0;
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

// (A x)_i in the reference's difference form, from the slot-major ELL copy.  The first 8 slots are
// unrolled and predicated so that all their column and coefficient loads are in flight before the
// dependent gathers of x; rows with more than 8 neighbours finish in a short loop.
template <typename T, class V>
__device__ __forceinline__ T spmv_row(int i, int n, const V &A, const T *__restrict__ diag, const T *__restrict__ x)
{
    constexpr int U = 8;
    const T xi = x[i];
    T acc = diag[i] * xi;
    int c[U];
    T v[U], xj[U];
    const int d = lv_ell_head<U>(A, i, n, c, v);
#pragma unroll
    for (int k = 0; k < U; k++) xj[k] = x[c[k]];
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (xi - xj[k]);
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        acc += A.value(e) * (xi - x[A.column(e, i, n)]);
    }
    return acc;
}

// r = (b - bmean) - A x; z = dinv r; p = z; partials r.z, r.r, b'.b'
// Grid-stride passes over the rows whose wavefronts are whole 64-row groups: the compact overflow's
// prefix of the next pass is formed during this one (lv_ell.h)
// (COMPACT false: no compact copy, and no ballots)
#define LV_OVF_PASSES(A, n, i0, stride, body)                                                 \
    if constexpr (COMPACT) {                                                                    \
        int dn_ = (i0) < (n) ? (A).degree(i0) : 0;                                              \
        lv_ovf_pos qn_ = lv_ovf_prefix(dn_);                                                    \
        for (int i = (i0); i < (n); i += (stride)) {                                            \
            const int d_ = dn_;                                                                 \
            const lv_ovf_pos q_ = qn_;                                                          \
            dn_ = i + (stride) < (n) ? (A).degree(i + (stride)) : 0;                            \
            body;                                                                               \
            qn_ = lv_ovf_prefix(dn_);                                                           \
        }                                                                                       \
    } else {                                                                                    \
        for (int i = (i0); i < (n); i += (stride)) {                                            \
            const int d_ = (A).degree(i);                                                       \
            const lv_ovf_pos q_{0, 0};                                                          \
            body;                                                                               \
        }                                                                                       \
    }

template <bool COMPACT>
__global__ __launch_bounds__(B) void k_cg64_init(int n, lv_ell_own<double> A, const double *__restrict__ diag,
                                                 const double *__restrict__ dinv, const double *__restrict__ b,
                                                 const double *__restrict__ x, double *__restrict__ r,
                                                 double *__restrict__ z, double *__restrict__ Prr,
                                                 double *__restrict__ P3, const lv_cg_scalars *S)
{
    const double bmean = S->bmean;
    double v[2] = {0.0, 0.0}, w[3] = {0.0, 0.0, 0.0};
    LV_OVF_PASSES(A, n, blockIdx.x * B + threadIdx.x, gridDim.x * B, {
        const double bi = b[i] - bmean;
        const double ri = bi - lv_ell_row_ovf<double>(A, i, n, d_, q_, diag[i], x[i], x);
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        v[0] += ri * zi;
        v[1] += ri * ri;
        w[0] += bi * bi;
    })
    block_partials<2>(v, Prr, gridDim.x);
    block_partials<3>(w, P3, gridDim.x);
}

// One CG iteration is two launches.
//
// S (SpMV): reduces r.z and r.r of the previous U, checks convergence, then for every row
//   w_i = (A z)_i,  q_i = w_i + beta q_i,  p_i = z_i + beta p_i,  partial p.q
// using A p_new = A z + beta A p_old, so only z is gathered and p, q are touched row-locally.
// The block's slice of the CSR arrays is first staged through LDS with coalesced loads.
// U (update): alpha = r.z / p.q;  x += alpha p;  r -= alpha q;  z = dinv r;  partials r.z, r.r.
// (wavefronts per SIMD asked of the register allocator: 8 for FP32, where with the compact
// overflow's prefix the kernel would otherwise take 66 VGPRs and 7 wavefronts and run 25% slower;
// 7 for FP64 with the compact copy (74 VGPRs and 6 otherwise), nothing for FP64 without it (72 VGPRs,
// 7 wavefronts; asking 7 there spills))
template <typename T, bool COMPACT>
struct cg_waves {
    static constexpr int value = 8;
};
template <>
struct cg_waves<double, true> {
    static constexpr int value = 7;
};
template <>
struct cg_waves<double, false> {
    static constexpr int value = 1;
};
template <typename T, class V, bool COMPACT>
__global__ __launch_bounds__(B, (cg_waves<T, COMPACT>::value)) void k_cg_s(int n, V A, const T *__restrict__ diag,
                                            const T *__restrict__ z, T *__restrict__ p, T *__restrict__ q,
                                            const double *__restrict__ Prr, double *__restrict__ Ppq,
                                            lv_cg_scalars *S, lv_progress *prog, int par, int k, int first,
                                            int seen_here)
{
    // the partials and the scalars the test and beta need in one round trip
    double pp[2][LV_RED_PER];
    reduce_all_load<2>(Prr, gridDim.x, pp);
    const int dn = S->done, itmax = S->itmax;
    const double thr2 = S->thr2, rzp = S->rz[par];
    reduce_all_pin<2>(pp);
    asm volatile("" ::"s"(dn), "s"(itmax), "s"(thr2), "s"(rzp));
    if (dn) {
        // a sweep that starts done (nothing left to solve): its first launch tells the host
        if (k == 0 && prog && blockIdx.x == 0 && threadIdx.x == 0)
            __hip_atomic_store(&prog->done, dn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    double o[2];
    reduce_all_sum<2>(Prr, gridDim.x, pp, o);
    const double rz = o[0], rr = o[1];
    const bool conv = rr <= thr2;
    const bool stop = conv || k >= itmax || !(rr == rr);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->rz[par ^ 1] = rz;
        S->rr = rr;
        if (k == 0) S->rr0 = rr;
        S->iters = k;
        if (stop) S->done = conv ? 1 : 2;
        if (prog) {
            if (stop) __hip_atomic_store(&prog->done, conv ? 1 : 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (seen_here) __hip_atomic_store(&prog->seen, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    if (stop) return;
    const T beta = first ? (T)0 : (T)(rz / rzp);
    double pq = 0.0;
    // grid-stride passes over the rows, each pass one contiguous window whose eighth per XCD is
    // contiguous too (lv_xcd_block); two rows per thread per pass keep twice the loads in flight
    const int stride = gridDim.x * B;
    const int istart = lv_xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x;
    // compact overflow: a wavefront's rows i0 (and i1) are one 64-row group, and the group's prefix
    // of the next pass is formed during this one, so that its overflow loads issue with the slots
    int dn0 = 0, dn1 = 0;
    lv_ovf_pos qn0{0, 0}, qn1{0, 0};
    if constexpr (COMPACT) {
        dn0 = istart < n ? A.degree(istart) : 0;
        dn1 = istart + stride < n ? A.degree(istart + stride) : 0;
        qn0 = lv_ovf_prefix(dn0);
        qn1 = lv_ovf_prefix(dn1);
    }
    for (int i0 = istart; i0 < n; i0 += 2 * stride) {
        const int i1 = i0 + stride;
        const T z0 = z[i0];
        T w0, z1 = (T)0, w1 = (T)0;
        if constexpr (COMPACT) {
            const int d0 = dn0, d1 = dn1;
            const lv_ovf_pos q0 = qn0, q1 = qn1;
            const int j0 = i0 + 2 * stride, j1 = j0 + stride;
            dn0 = j0 < n ? A.degree(j0) : 0;
            dn1 = j1 < n ? A.degree(j1) : 0;
            w0 = lv_ell_row_ovf<T>(A, i0, n, d0, q0, diag[i0], z0, z);
            if (i1 < n) {
                z1 = z[i1];
                w1 = lv_ell_row_ovf<T>(A, i1, n, d1, q1, diag[i1], z1, z);
            }
            qn0 = lv_ovf_prefix(dn0);
            qn1 = lv_ovf_prefix(dn1);
        } else {
            w0 = spmv_row<T>(i0, n, A, diag, z);
            if (i1 < n) {
                z1 = z[i1];
                w1 = spmv_row<T>(i1, n, A, diag, z);
            }
        }
        const T q0 = first ? w0 : w0 + beta * q[i0];
        const T p0 = first ? z0 : z0 + beta * p[i0];
        q[i0] = q0;
        p[i0] = p0;
        pq += (double)p0 * (double)q0;
        if (i1 < n) {
            const T q1 = first ? w1 : w1 + beta * q[i1];
            const T p1 = first ? z1 : z1 + beta * p[i1];
            q[i1] = q1;
            p[i1] = p1;
            pq += (double)p1 * (double)q1;
        }
    }
    double v[1] = {pq};
    block_partials<1>(v, Ppq, gridDim.x);
}

// x += alpha p, r -= alpha q.  With Jacobi (JACOBI) also z = D^-1 r and the partials of r.z, r.r
// for the next k_cg_s; otherwise the partials of r.r alone (Pur), which k_conv_check tests before
// the preconditioner is applied, and with z0 the multilevel cycle's pre-smoothed z0 = sc D^-1 r.
template <typename T, bool JACOBI>
__global__ __launch_bounds__(B) void k_cg_u(int n, T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                            const T *__restrict__ p, const T *__restrict__ q,
                                            const T *__restrict__ dinv, const double *__restrict__ Ppq,
                                            double *__restrict__ Prr, double *__restrict__ Pur, lv_cg_scalars *S,
                                            int par, T *__restrict__ z0, T sc)
{
    // the partials, the flag and rz in one round trip (every block starts here at once)
    double pp[1][LV_RED_PER];
    reduce_all_load<1>(Ppq, gridDim.x, pp);
    const int dn = S->done;
    const double rzn = S->rz[par ^ 1];
    reduce_all_pin<1>(pp);
    asm volatile("" ::"s"(dn), "s"(rzn));
    if (dn) return;
    double pq[1];
    reduce_all_sum<1>(Ppq, gridDim.x, pp, pq);
    const double alpha = rzn / pq[0];
    const T a = (T)alpha;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        if (!JACOBI && z0) z0[i] = sc * dinv[i] * ri;
        if (JACOBI) {
            const T zi = dinv[i] * ri;
            z[i] = zi;
            v[0] += (double)ri * (double)zi;
        }
        v[1] += (double)ri * (double)ri;
    }
    if (JACOBI) {
        block_partials<2>(v, Prr, gridDim.x);
    } else {
        const double w[1] = {v[1]};
        block_partials<1>(w, Pur, gridDim.x);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->pq = pq[0];
        S->alpha = alpha;
    }
}

// the convergence test of k_cg_s, made on the residual k_cg_u has just updated (after k+1
// updates), so that a converged iteration skips the preconditioner; it also marks iteration k as
// seen (its test ran), which lets the host launch iteration k+1 while the device applies the
// preconditioner
__global__ __launch_bounds__(B) void k_conv_check(const double *__restrict__ Pur, int nb, lv_cg_scalars *S,
                                                  lv_progress *prog, int k)
{
    // the partials, the flag and the threshold in one round trip
    double p[1][LV_RED_PER];
    reduce_all_load<1>(Pur, nb, p);
    const int dn = S->done;
    const double thr2 = S->thr2;
    reduce_all_pin<1>(p);
    asm volatile("" ::"s"(dn), "s"(thr2));
    if (dn) return;
    double o[1];
    reduce_all_sum<1>(Pur, nb, p, o);
    if (threadIdx.x == 0) {
        if (o[0] <= thr2) {
            S->rr = o[0];
            S->iters = k + 1;
            S->done = 1;
            if (prog) __hip_atomic_store(&prog->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        if (prog) __hip_atomic_store(&prog->seen, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// (M r)_i with the learned preconditioner in ELL slots (same columns as A's ELL copy): the first 8
// slots loaded unconditionally (lv_ell_head) so their loads are in flight before the gathers of r
__device__ __forceinline__ float apply_m_row(int i, int n, const int *__restrict__ deg, const int *__restrict__ ecol,
                                             const float *__restrict__ mval, const float *__restrict__ mdg,
                                             const float *__restrict__ r, float ri)
{
    constexpr int U = 8;
    float acc = mdg[i] * ri;
    int c[U];
    float v[U], rj[U];
    const int d = lv_ell_head_with<U>(lv_ell_own<float>{deg, ecol, mval}, mval, i, n, c, v);
#pragma unroll
    for (int k = 0; k < U; k++) rj[k] = r[c[k]];
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * rj[k];
    for (int k = U; k < d; k++) {
        const size_t e = (size_t)k * n + i;
        acc += mval[e] * r[ecol[e]];
    }
    return acc;
}

// z = M r; partials r.z and r.r for the next S kernel (two rows per thread per pass)
__global__ __launch_bounds__(B) void k_apply_m32(int n, const int *__restrict__ deg, const int *__restrict__ ecol,
                                                 const float *__restrict__ mval, const float *__restrict__ mdg,
                                                 const float *__restrict__ r, float *__restrict__ z,
                                                 double *__restrict__ Prr, const lv_cg_scalars *S)
{
    if (S->done) return;
    double v[2] = {0.0, 0.0};
    const int stride = gridDim.x * B;
    for (int i0 = blockIdx.x * B + threadIdx.x; i0 < n; i0 += 2 * stride) {
        const int i1 = i0 + stride;
        const float r0 = r[i0];
        const float z0 = apply_m_row(i0, n, deg, ecol, mval, mdg, r, r0);
        z[i0] = z0;
        v[0] += (double)r0 * (double)z0;
        v[1] += (double)r0 * (double)r0;
        if (i1 < n) {
            const float r1 = r[i1];
            const float z1 = apply_m_row(i1, n, deg, ecol, mval, mdg, r, r1);
            z[i1] = z1;
            v[0] += (double)r1 * (double)z1;
            v[1] += (double)r1 * (double)r1;
        }
    }
    block_partials<2>(v, Prr, gridDim.x);
}

// b = (m/dt) v_c and x = v_c for one velocity component (ViscousSolver's assemble!, src/viscous_solver.jl:97-102)
__global__ __launch_bounds__(B) void k_visc_rhs(int n, const double *__restrict__ mass, const double *__restrict__ vc,
                                                double dt, double *__restrict__ b, double *__restrict__ x)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    b[i] = (mass[i] / dt) * vc[i];
    x[i] = vc[i];
}

// (A (x - d))_i with every x read shifted by d, i.e. the product with the stored values x - d
__device__ __forceinline__ double spmv_row_shifted(int i, int n, const lv_ell_own<double> &A,
                                                   const double *__restrict__ diag, const double *__restrict__ x,
                                                   double d)
{
    constexpr int U = 8;
    const double xi = x[i] - d;
    double acc = diag[i] * xi;
    int c[U];
    double v[U], xj[U];
    const int deg = lv_ell_head<U>(A, i, n, c, v);
#pragma unroll
    for (int k = 0; k < U; k++) xj[k] = x[c[k]] - d;
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (xi - xj[k]);
    for (int k = U; k < deg; k++) {
        const size_t e = (size_t)k * n + i;
        acc += A.value(e) * (xi - (x[A.column(e, i, n)] - d));
    }
    return acc;
}

// true residual r = (b - bmean) - A x; partials |r|^2, |b'|^2 (FP64).  With Pxs (the partial sums
// of x written by k_correct, singular systems) the residual is that of x - d, d = mean(x) - xmean0,
// the x that k_shift stores at the end of the solve; block 0 records d.
template <bool COMPACT, class V = lv_ell_own<double>>
__global__ __launch_bounds__(B) void k_resid64(int n, V A, const double *__restrict__ diag,
                                               const double *__restrict__ b, const double *__restrict__ x,
                                               double *__restrict__ r, double *__restrict__ partials,
                                               const double *__restrict__ Pxs, lv_cg_scalars *S)
{
    const double bmean = S->bmean;
    double d = 0.0;
    if (Pxs) {
        double t[1];
        reduce_all<1>(Pxs, gridDim.x, t);
        d = t[0] / n - S->xmean0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S->xshift = d;
    double v[3] = {0.0, 0.0, 0.0};
    LV_OVF_PASSES(A, n, blockIdx.x * B + threadIdx.x, gridDim.x * B, {
        const double bi = b[i] - bmean;
        // (with Pxs, the product with the stored x - d: every x read shifted, as k_shift will store it)
        const double ri = bi - (Pxs ? lv_ell_row_ovf<double>(A, i, n, d_, q_, diag[i], x[i] - d, x, d)
                                    : lv_ell_row_ovf<double>(A, i, n, d_, q_, diag[i], x[i], x));
        r[i] = ri;
        v[0] += ri * ri;
        v[1] += bi * bi;
    })
    block_partials<3>(v, partials, gridDim.x);
}

// y = A x (FP64, the operator's own ELL)
template <bool COMPACT, class V = lv_ell_own<double>>
__global__ __launch_bounds__(B) void k_spmv64(int n, V A, const double *__restrict__ diag,
                                              const double *__restrict__ x, double *__restrict__ y)
{
    LV_OVF_PASSES(A, n, blockIdx.x * B + threadIdx.x, gridDim.x * B,
                  { y[i] = lv_ell_row_ovf<double>(A, i, n, d_, q_, diag[i], x[i], x); })
}

__global__ __launch_bounds__(B) void k_sum1(int n, const double *__restrict__ x, double *__restrict__ partials)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) v[0] += x[i];
    block_partials<3>(v, partials, gridDim.x);
}

// restore the mean of x to that of the initial guess (singular systems only): x -= xshift
__global__ __launch_bounds__(B) void k_shift(int n, double *__restrict__ x, const lv_cg_scalars *S)
{
    const double d = S->xshift;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) x[i] -= d;
}

// ---- FP32 inner PCG of the refinement --------------------------------------------------------

// A first sweep that k_finalize (FIN_RESID0) flagged skip (initial residual within tol, zero, or
// NaN) must leave x as it is: its FP32 vectors are zeros and its correction scale is 0, so neither
// 1/sqrt(0) nor a NaN residual reaches x.
__device__ inline bool sweep_skipped(const lv_cg_scalars *S, int first) { return first && S->skip; }

// r32 = r64 / |r64|, e = 0, z = dinv r, p = z; partials r.z, r.r.  With perm, the FP32 vectors
// are in the aggregate order: r32[i] comes from r64[perm[i]].  With z0 also z0 = sc dinv r (the
// multilevel cycle's pre-smoothed correction).  first: this is the solve's first sweep.
__global__ __launch_bounds__(B) void k_cg32_init(int n, const double *__restrict__ r64, const int *__restrict__ perm,
                                                 const float *__restrict__ dinv, float *__restrict__ e,
                                                 float *__restrict__ r, float *__restrict__ z,
                                                 double *__restrict__ Prr, const lv_cg_scalars *S,
                                                 float *__restrict__ z0, float sc, int first)
{
    const bool skip = sweep_skipped(S, first);
    const double inv = skip ? 0.0 : 1.0 / sqrt(S->rr64);
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const float ri = skip ? 0.0f : (float)(r64[perm ? perm[i] : i] * inv);
        const float zi = dinv[i] * ri;
        e[i] = 0.0f;
        r[i] = ri;
        z[i] = zi;
        if (z0) z0[i] = sc * dinv[i] * ri;
        v[0] += (double)ri * (double)zi;
        v[1] += (double)ri * (double)ri;
    }
    block_partials<2>(v, Prr, gridDim.x);
}

// x += |r64| e (e in the aggregate order when iperm is given: x[i] takes e[iperm[i]]); with Pxs
// also the partial sums of the new x (for the shift of a singular system).  A skipped first sweep
// leaves x unchanged (and still writes its sums).
__global__ __launch_bounds__(B) void k_correct(int n, double *__restrict__ x, const float *__restrict__ e,
                                               const int *__restrict__ iperm, double *__restrict__ Pxs,
                                               const lv_cg_scalars *S, int first)
{
    const bool skip = sweep_skipped(S, first);
    const double s = skip ? 0.0 : sqrt(S->rr64);
    double v[1] = {0.0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const double xi = skip ? x[i] : x[i] + s * (double)e[iperm ? iperm[i] : i];
        x[i] = xi;
        v[0] += xi;
    }
    if (Pxs) block_partials<1>(v, Pxs, gridDim.x);
}

// ---- small FP64 vector algebra for the Picard extrapolation -----------------------------------

// operands of the multi-dot and combination kernels; the a's and x's may be FP32 (the Picard
// window's stored directions), everything is accumulated in FP64
template <typename TA>
struct lv_vec8 {
    const TA *a[8];
    const double *b[8];
};
template <typename TX>
struct lv_comb8 {
    const TX *x[8];
    double c[8];
};

// M <= 8 dot products a_q . b_q in one pass (M a template parameter, so that a row's loads are
// issued together rather than each behind a branch on the count)
template <typename TA, int M>
__global__ __launch_bounds__(B) void k_multidot(int n, lv_vec8<TA> V, double *__restrict__ partials)
{
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
#pragma unroll
        for (int q = 0; q < M; q++) v[q] += (double)V.a[q][i] * V.b[q][i];
    }
    block_partials<8>(v, partials, gridDim.x);
}

__global__ __launch_bounds__(B) void k_multidot_fin(const double *__restrict__ partials, int nb,
                                                    double *__restrict__ out)
{
    double o[8];
    reduce_all<8>(partials, nb, o);
    if (threadIdx.x < 8) out[threadIdx.x] = o[threadIdx.x];
}

// y = alpha*y + sum_q c_q x_q  (y may be one of the x_q only if its c is applied via alpha).  The
// coefficients are C.c scaled by device values dc (the output of k_multidot_fin): dmode 0 none,
// 1 c_q dc[q], 2 c_q / sqrt(dc[0]) (0 when dc[0] is not positive and finite).  Accumulated in FP64
// and rounded once to y's type.
// (M, the number of x's, a template parameter as in k_multidot)
template <typename TX, typename TY, int M>
__global__ __launch_bounds__(B) void k_combine(int n, TY *__restrict__ y, double alpha, lv_comb8<TX> C,
                                               const double *__restrict__ dc, int dmode)
{
    double c[8];
#pragma unroll
    for (int q = 0; q < 8; q++) c[q] = C.c[q];
    if (dmode == 1) {
#pragma unroll
        for (int q = 0; q < M; q++) c[q] = C.c[q] * dc[q];
    } else if (dmode == 2) {
        const double h = dc[0];
        const double f = (h > 0.0 && h <= 1.7976931348623157e308) ? 1.0 / sqrt(h) : 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) c[q] = C.c[q] * f;
    }
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        double acc = (alpha != 0.0) ? alpha * (double)y[i] : 0.0;  // alpha = 0 never reads y (may be uninitialised)
#pragma unroll
        for (int q = 0; q < M; q++) acc += c[q] * (double)C.x[q][i];
        y[i] = (TY)acc;
    }
}

// A x = b' - r: recovers A P of a finished solve from its final residual (b' = b - bmean)
__global__ __launch_bounds__(B) void k_apply_from_residual(int n, const double *__restrict__ b,
                                                          const double *__restrict__ r, double *__restrict__ Ax,
                                                          const lv_cg_scalars *S)
{
    const double bmean = S->bmean;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) Ax[i] = (b[i] - bmean) - r[i];
}

inline dim3 grid_of(int n) { return dim3((n + B - 1) / B); }

constexpr int NBMAX = LV_NBMAX;
inline int nblocks_of(int n) { return lv_nblocks(n); }

}  // namespace

// ---- host wrappers ---------------------------------------------------------------------------------

hipError_t lv_launch_degree(lv_mesh_view M, int n, int *deg, hipStream_t s)
{
    hipLaunchKernelGGL(k_degree, grid_of(n), dim3(B), 0, s, M, n, deg);
    return hipGetLastError();
}

hipError_t lv_launch_assemble(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int mode,
                              lv_pressure_system &A, int *nonsingular, int *wide, hipStream_t s, const lv_vbc *vbc,
                              const lv_rhs_geom *R)
{
    const lv_vbc v = vbc ? *vbc : lv_vbc{};
    const lv_rhs_geom r = (R && mode == 0) ? *R : lv_rhs_geom{};
    hipLaunchKernelGGL(k_assemble, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, mode, (const int *)A.row_ptr, A.cols,
                       A.val64, A.val32, A.diag64, A.diag32, A.dinv64, A.dinv32, A.ecol, A.ev64, A.ev32, nonsingular,
                       v, r, A.ocol, A.oev64, A.oev32, A.deg8, A.ecol16, A.ocol16, wide);
    return hipGetLastError();
}

hipError_t lv_launch_rhs_geom(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, lv_vbc vbc,
                              const lv_rhs_geom &R, hipStream_t s)
{
    hipLaunchKernelGGL(k_rhs_geom, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, vbc, R);
    return hipGetLastError();
}

hipError_t lv_launch_gp(lv_mesh_view M, int n, const lv_rhs_geom &R, const double *P, const double *mass, double *GP,
                        double *Pout, double *Pout2, hipStream_t s)
{
    hipLaunchKernelGGL(k_gp, grid_of(n), dim3(B), 0, s, M, n, R, P, mass, GP, Pout, Pout2);
    return hipGetLastError();
}

hipError_t lv_launch_rhs_b(lv_mesh_view M, int n, int gp_step, const lv_rhs_geom &R, const double *GP,
                           const double *area, const double *P, double *b, double *P0, hipStream_t s)
{
    hipLaunchKernelGGL(k_rhs_b, grid_of(n), dim3(B), 0, s, M, n, gp_step, R, GP, area, P, b, P0);
    return hipGetLastError();
}

hipError_t lv_launch_visc_rhs(int n, const double *mass, const double *vc, double dt, double *b, double *x,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_visc_rhs, grid_of(n), dim3(B), 0, s, n, mass, vc, dt, b, x);
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

// S into h through the pinned staging copy pin (a copy into pageable memory would be staged too)
hipError_t read_scalars(lv_cg_scalars *dS, lv_cg_scalars *pin, lv_cg_scalars *h, hipStream_t s)
{
    hipError_t e = hipMemcpyAsync(pin, dS, sizeof(lv_cg_scalars), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    *h = *pin;
    return hipSuccess;
}

// The scalars k_finalize published for sweep decision seq into h (the fields the host tests).  The
// host spins on the pinned record; after 20 ms without it (a kernel failed, or the device is
// busy elsewhere) it falls back to the synchronising read, whose values are then current.
hipError_t await_published(lv_progress *hprog, int seq, lv_cg_scalars *dS, lv_cg_scalars *pin, lv_cg_scalars *h,
                           hipStream_t s)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (int spin = 1;; spin++) {
        if (__atomic_load_n(&hprog->fseq, __ATOMIC_ACQUIRE) == seq) {
            h->rr64 = hprog->rr64;
            h->rr64_0 = hprog->rr64_0;
            h->bb = hprog->bb;
            h->thr2 = hprog->thr2;
            h->iters = hprog->iters;
            h->skip = hprog->skip;
            return hipSuccess;
        }
        if ((spin & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20))
            return read_scalars(dS, pin, h, s);
    }
}

__global__ void k_init_scalars(lv_cg_scalars *S, double tol2, int itmax, int singular)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    lv_cg_scalars z = {};
    z.tol2 = tol2;
    z.itmax = itmax;
    z.singular = singular;
    *S = z;
}

// the operator's own ELL as the kernels read it, with its compact overflow when it has one
// (systems below LV_OVF_MIN_ROWS rows leave it out: there the kernels' launch latency, not the
// padding's bytes, sets their time, and the ballots cost more than they save: Gresho at 100^2 took
// 11.8 against 11.2 ms per step with it)
#define LV_OVF_MIN_ROWS 65536
// (amg_layout 2 leaves it out too, for measurement and tests)
static int own_ocap(const lv_pressure_system &A, int n, int layout)
{
    return (A.ocol && n >= LV_OVF_MIN_ROWS && layout != 2) ? LV_OVF_CAP : 0;
}
static lv_ell_own<double> own64(const lv_pressure_system &A, int n, int layout = 0)
{
    lv_ell_own<double> v{A.deg, A.ecol, A.ev64};
    v.ocol = A.ocol;
    v.oval = A.oev64;
    v.ocap = own_ocap(A, n, layout);
    return v;
}
// the FP64 operator through its 16-bit offsets and uint8 degrees (valid when A.narrow)
using lv_ell_own64s = lv_ell_view<uint8_t, int16_t, double>;
static lv_ell_own64s own64s(const lv_pressure_system &A, int n, int layout = 0)
{
    lv_ell_own64s v{A.deg8, A.ecol16, A.ev64};
    v.ocol = A.ocol16;
    v.oval = A.oev64;
    v.ocap = own_ocap(A, n, layout);
    return v;
}
static lv_ell_own<float> own32(const lv_pressure_system &A, int n, int layout = 0)
{
    lv_ell_own<float> v{A.deg, A.ecol, A.ev32};
    v.ocol = A.ocol;
    v.oval = A.oev32;
    v.ocap = own_ocap(A, n, layout);
    return v;
}

// kern<true> (with the compact copy) or kern<false>, by the view's ocap
#define LV_LAUNCH_BY_OCAP(kern, view, ...)                 \
    do {                                                   \
        if ((view).ocap > 0)                               \
            hipLaunchKernelGGL((kern<true>), __VA_ARGS__);  \
        else                                               \
            hipLaunchKernelGGL((kern<false>), __VA_ARGS__); \
    } while (0)

// r = b - A x in FP64 (k_resid64) and y = A x (k_spmv64), through the 16-bit copy of the operator's
// columns when every offset fits (2 bytes a slot and 1 a row instead of 4 and 4)
template <class V>
static void launch_resid64_v(const V &Av, dim3 G, hipStream_t s, int n, const double *diag, const double *b,
                             const double *x, double *r, double *p3, const double *Pxs, lv_cg_scalars *S)
{
    if (Av.ocap > 0)
        hipLaunchKernelGGL((k_resid64<true, V>), G, dim3(B), 0, s, n, Av, diag, b, x, r, p3, Pxs, S);
    else
        hipLaunchKernelGGL((k_resid64<false, V>), G, dim3(B), 0, s, n, Av, diag, b, x, r, p3, Pxs, S);
}
static void launch_resid64(const lv_pressure_system &A, int layout, dim3 G, hipStream_t s, int n, const double *b,
                           const double *x, double *r, double *p3, const double *Pxs, lv_cg_scalars *S)
{
    if (A.narrow)
        launch_resid64_v(own64s(A, n, layout), G, s, n, (const double *)A.diag64, b, x, r, p3, Pxs, S);
    else
        launch_resid64_v(own64(A, n, layout), G, s, n, (const double *)A.diag64, b, x, r, p3, Pxs, S);
}
template <class V>
static void launch_spmv64_v(const V &Av, dim3 G, hipStream_t s, int n, const double *diag, const double *x, double *y)
{
    if (Av.ocap > 0)
        hipLaunchKernelGGL((k_spmv64<true, V>), G, dim3(B), 0, s, n, Av, diag, x, y);
    else
        hipLaunchKernelGGL((k_spmv64<false, V>), G, dim3(B), 0, s, n, Av, diag, x, y);
}

// partial-sum regions: P3 = one-off reductions (3 values), Ppq (1), Prr (2)
struct parts {
    double *p3, *pq, *rr, *ur;
};
inline parts parts_of(double *base) { return parts{base, base + 3 * NBMAX, base + 4 * NBMAX, base + 6 * NBMAX}; }

// CG iterations (two launches each; with a preconditioner also the convergence test and its
// application) until the device flags completion; *dflag is the done flag (1 converged, 2 stopped).
// It returns without synchronising: S (iterations, residuals) is read by the caller's next read.
// The host launches iteration k once the device has marked iteration k - 1 seen in prog (pinned
// host memory): k_cg_s marks it when it starts, or, with a preconditioner, k_conv_check once its
// test has run, so the host launches the next iteration while the device applies the
// preconditioner.  It stops launching as soon as prog->done is set: at most one iteration (with a
// preconditioner, one application of it) runs as no-ops after convergence.  If the record does not
// advance for 20 ms the host synchronises and reads the scalars instead, so the loop ends whatever
// the record shows.
template <typename T, class V>
hipError_t cg_iterate(int n, V A, const T *diag, const T *dinv, T *x, T *r, T *z, T *p, T *q, const parts &P,
                      lv_cg_scalars *dS, lv_progress *hprog, lv_progress *dprog, lv_cg_scalars *pin, hipStream_t s,
                      lv_timer *tm, lv_cg_scalars &h, const lv_precond *M, int *dflag)
{
    const int nb = nblocks_of(n);
    // no kernel of this solve that stores progress is in flight: since the previous sweep the caller
    // has read the FP64 residual (published by k_finalize, which runs after every earlier kernel of
    // the stream has finished) or synchronised before a fallback sweep
    __atomic_store_n(&hprog->seen, -1, __ATOMIC_RELAXED);
    __atomic_store_n(&hprog->done, 0, __ATOMIC_RELEASE);
    for (int k = 0;; k++) {
        if (k > 0) {
            bool stop = false;
            const auto t0 = std::chrono::steady_clock::now();
            for (int spin = 1;; spin++) {
                const int d = __atomic_load_n(&hprog->done, __ATOMIC_ACQUIRE);
                if (d) {
                    *dflag = d;
                    stop = true;
                    break;
                }
                if (__atomic_load_n(&hprog->seen, __ATOMIC_ACQUIRE) >= k - 1) break;
                if ((spin & 255) == 0 &&
                    std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
                    hipError_t e = read_scalars(dS, pin, &h, s);  // every launched kernel has finished
                    if (e != hipSuccess) return e;
                    *dflag = h.done;
                    stop = h.done != 0;
                    break;
                }
            }
            if (stop) break;
        }
        const int par = k & 1;
        {
            timer_scope ts(tm, s);
            if (A.ocap > 0)
                hipLaunchKernelGGL((k_cg_s<T, V, true>), dim3(nb), dim3(B), 0, s, n, A, diag, (const T *)z, p, q,
                                   (const double *)P.rr, P.pq, dS, dprog, par, k, k == 0 ? 1 : 0, M ? 0 : 1);
            else
                hipLaunchKernelGGL((k_cg_s<T, V, false>), dim3(nb), dim3(B), 0, s, n, A, diag, (const T *)z, p, q,
                                   (const double *)P.rr, P.pq, dS, dprog, par, k, k == 0 ? 1 : 0, M ? 0 : 1);
        }
        if (M) {
            // the cycle's pre-smoothed z0 stored with the residual update (FP32 solves only)
            T *z0 = nullptr;
            T sc = (T)0;
            if constexpr (std::is_same<T, float>::value)
                if (M->amg && M->amg->z0_in) {
                    z0 = M->amg->z0;
                    sc = M->amg->fine_scale;
                }
            hipLaunchKernelGGL((k_cg_u<T, false>), dim3(nb), dim3(B), 0, s, n, x, r, z, (const T *)p,
                               (const T *)q, dinv, (const double *)P.pq, P.rr, P.ur, dS, par, z0, sc);
            hipLaunchKernelGGL(k_conv_check, dim3(1), dim3(B), 0, s, (const double *)P.ur, nb, dS, dprog, k);
            if (M->amg) {
                hipError_t e = lv_amg_vcycle(M->amg, *M->A, (const float *)r, (float *)z, P.rr,
                                             (const lv_cg_scalars *)dS, s, z0 != nullptr);
                if (e != hipSuccess) return e;
            } else if constexpr (std::is_same<V, lv_ell_own<float>>::value) {
                hipLaunchKernelGGL(k_apply_m32, dim3(nb), dim3(B), 0, s, n, A.deg, A.col, M->mval, M->mdg,
                                   (const float *)r, (float *)z, P.rr, (const lv_cg_scalars *)dS);
            } else {
                return hipErrorInvalidValue;  // the learned one-level preconditioner is in the operator's order
            }
        } else {
            hipLaunchKernelGGL((k_cg_u<T, true>), dim3(nb), dim3(B), 0, s, n, x, r, z, (const T *)p,
                               (const T *)q, dinv, (const double *)P.pq, P.rr, P.ur, dS, par, (T *)nullptr,
                               (T)0);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

// Solves A x = b (device vectors; x holds the initial guess and is iterated in place).  Writes
// iterations, sweeps and the final FP64 relative residual.  *converged tells whether tol was met.
hipError_t lv_solve_device(const lv_pressure_system &A, lv_solver_work &W, int n, const double *b, double *x,
                           const lv_solve_opts &opt, hipStream_t s, lv_timer *t32, lv_timer *t64, int *iters,
                           int *sweeps, double *relres, double *relres0, bool *converged, const lv_precond *M,
                           int *fallbacks)
{
    *relres0 = -1.0;
    *fallbacks = 0;
    const int nb = nblocks_of(n);
    const dim3 G(nb);
    const parts P = parts_of(W.partials);
    lv_cg_scalars h;
    memset(&h, 0, sizeof(h));
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, s, W.S, opt.tol * opt.tol,
                       opt.itmax > 0 ? opt.itmax : 1000, (int)A.singular);
    hipError_t e = hipSuccess;
    // the means of b and x serve singular systems only (bmean stays the 0 k_init_scalars set, and xmean0
    // is read only where x's mean is restored)
    if (A.singular) {
        hipLaunchKernelGGL(k_means, G, dim3(B), 0, s, n, b, (const double *)x, P.p3);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_MEANS, n, 0.0,
                           W.S);
    }
    int total = 0, nsweep = 0, skipped = 0;
    if (opt.method == LV_SOLVER_PCG64) {
        LV_LAUNCH_BY_OCAP(k_cg64_init, own64(A, n, opt.amg_layout), G, dim3(B), 0, s, n, own64(A, n, opt.amg_layout), (const double *)A.diag64, (const double *)A.dinv64, b,
                           (const double *)x, W.r, W.z, P.rr, P.p3, (const lv_cg_scalars *)W.S);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_INIT64, n, 0.0,
                           W.S);
        int done = 0;
        e = cg_iterate<double>(n, own64(A, n, opt.amg_layout), (const double *)A.diag64,
                               (const double *)A.dinv64, x, W.r, W.z, W.p, W.q, P, W.S, W.hprog,
                               W.dprog, W.hS, s, t64, h, (const lv_precond *)nullptr, &done);
        if (e != hipSuccess) return e;
        e = read_scalars(W.S, W.hS, &h, s);
        if (e != hipSuccess) return e;
        total = h.iters;
        *relres0 = (h.bb > 0.0) ? sqrt(h.rr0 / h.bb) : sqrt(h.rr0);
    } else {
        const int outer_max = opt.outer_max > 0 ? opt.outer_max : 20;
        const double floor_tol = opt.inner_tol > 0.0 ? opt.inner_tol : 1e-4;
        // singular systems: the residual after a correction is that of x shifted back to the initial
        // guess's mean (k_correct writes the sums of x, k_resid64 shifts what it reads); k_shift
        // stores that x when the sweeps end, so the last residual is the returned x's
        double *Pxs = A.singular ? P.ur : nullptr;
        // the iterations of a sweep are still in S at the next sweep's read of the FP64 residual
        // (k_resid64 and k_finalize leave them), so a sweep ends without a synchronising read.
        // Nor does the first sweep wait for one: k_finalize (FIN_RESID0) makes the tests below on
        // the device, and a first sweep with nothing to solve starts done and changes nothing
        // (its correction is zero).
        bool pending = false;
        double itol = 0.0;
        for (;;) {
            launch_resid64(A, opt.amg_layout, G, s, n, b, (const double *)x, W.r, P.p3,
                           nsweep > 0 ? (const double *)Pxs : (const double *)nullptr, W.S);
            if (nsweep == 0) {
                hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_RESID0,
                                   n, floor_tol, W.S);
            } else {
                const int seq = __atomic_load_n(&W.hprog->fseq, __ATOMIC_ACQUIRE) + 1;
                hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_RESID, n,
                                   0.0, W.S, W.dprog, seq);
                e = await_published(W.hprog, seq, W.S, W.hS, &h, s);
                if (e != hipSuccess) return e;
                if (pending) total += h.iters;
                pending = false;
                if (nsweep == 1) {
                    *relres0 = (h.bb > 0.0) ? sqrt(h.rr64_0 / h.bb) : sqrt(h.rr64_0);
                    skipped = h.skip;  // the first sweep had nothing to solve (not reported as a sweep)
                }
                if (h.rr64 <= h.thr2 || nsweep >= outer_max || !(h.rr64 == h.rr64)) break;
                // ask the FP32 sweep for just the reduction still missing, with a margin of 0.8 for
                // the FP32 residual's drift from the FP64 one (a few per cent at these reductions: at
                // margins up to 0.9 no 2048^2 benchmark solve needed an extra sweep), but not below
                // what FP32 arithmetic can deliver in one sweep (the first sweep's is the same rule,
                // evaluated by k_finalize)
                const double need = 0.8 * sqrt(h.thr2 / h.rr64);
                itol = need > floor_tol ? need : floor_tol;
                if (itol > 0.8) itol = 0.8;
            }
            const lv_precond *Mi = (opt.method == LV_SOLVER_SPAI || opt.method == LV_SOLVER_AMG) ? M : nullptr;
            // with the multilevel preconditioner in the aggregate order, the whole FP32 sweep runs in
            // that order (the FP64 residual and correction stay in the particles' order)
            const lv_amg *Gp = (Mi && Mi->amg && Mi->amg->perm_ok) ? Mi->amg : nullptr;
            const float *dinv32 = Gp ? (const float *)Gp->pdinv32 : (const float *)A.dinv32;
            for (int attempt = 0; attempt < 2; attempt++) {
                float *z0 = (Mi && Mi->amg && Mi->amg->z0_in) ? Mi->amg->z0 : nullptr;
                hipLaunchKernelGGL(k_cg32_init, G, dim3(B), 0, s, n, (const double *)W.r,
                                   Gp ? (const int *)Gp->list : (const int *)nullptr, dinv32, W.e32, W.r32, W.z32,
                                   P.rr, (const lv_cg_scalars *)W.S, z0, z0 ? Mi->amg->fine_scale : 0.0f,
                                   nsweep == 0 ? 1 : 0);
                // k_finalize2 clears the done flag of the previous sweep, so it precedes the first
                // application of the learned preconditioner (which is skipped while done is set)
                hipLaunchKernelGGL(k_finalize2, dim3(1), dim3(B), 0, s, (const double *)P.rr, nb, itol * itol, W.S,
                                   nsweep == 0 ? 1 : 0);
                if (Mi && Mi->amg) {
                    e = lv_amg_vcycle(Mi->amg, A, (const float *)W.r32, W.z32, P.rr, (const lv_cg_scalars *)W.S, s,
                                      z0 != nullptr);
                    if (e != hipSuccess) return e;
                } else if (Mi) {
                    hipLaunchKernelGGL(k_apply_m32, G, dim3(B), 0, s, n, (const int *)A.deg, (const int *)A.ecol,
                                       Mi->mval, Mi->mdg, (const float *)W.r32, W.z32, P.rr,
                                       (const lv_cg_scalars *)W.S);
                }
                int done = 0;
                if (Gp)
                    e = cg_iterate<float>(n, lv_ell_agg32{Gp->pdeg, Gp->poff, Gp->pev32, Gp->ooff, Gp->oev32, Gp->ovf_cap},
                                          (const float *)Gp->pdiag32,
                                          dinv32, W.e32, W.r32, W.z32, W.p32, W.q32, P, W.S, W.hprog, W.dprog,
                                          W.hS, s, t32, h, Mi, &done);
                else
                    e = cg_iterate<float>(n, own32(A, n, opt.amg_layout), (const float *)A.diag32,
                                          dinv32, W.e32, W.r32, W.z32, W.p32, W.q32, P, W.S, W.hprog, W.dprog,
                                          W.hS, s, t32, h, Mi, &done);
                if (e != hipSuccess) return e;
                // Jacobi-PCG is the fallback: a sweep of the learned preconditioner that stalls or breaks
                // down is redone with Jacobi (after a synchronising read, which also counts its iterations)
                if (!Mi || done == 1) break;
                e = read_scalars(W.S, W.hS, &h, s);
                if (e != hipSuccess) return e;
                total += h.iters;
                (*fallbacks)++;
                Mi = nullptr;
            }
            pending = true;
            hipLaunchKernelGGL(k_correct, G, dim3(B), 0, s, n, x, (const float *)W.e32,
                               Gp ? (const int *)Gp->iperm : (const int *)nullptr, Pxs, (const lv_cg_scalars *)W.S,
                               nsweep == 0 ? 1 : 0);
            nsweep++;
        }
        if (Pxs && nsweep > 0) hipLaunchKernelGGL(k_shift, G, dim3(B), 0, s, n, x, (const lv_cg_scalars *)W.S);
    }
    if (opt.method == LV_SOLVER_PCG64) {
        // final FP64 residual of the returned x
        if (A.singular) {
            hipLaunchKernelGGL(k_sum1, G, dim3(B), 0, s, n, (const double *)x, P.p3);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_SUM1, n, 0.0,
                               W.S);
            hipLaunchKernelGGL(k_shift, G, dim3(B), 0, s, n, x, (const lv_cg_scalars *)W.S);
        }
        launch_resid64(A, opt.amg_layout, G, s, n, b, (const double *)x, W.r, P.p3, (const double *)nullptr, W.S);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(B), 0, s, (const double *)P.p3, nb, (int)FIN_RESID, n, 0.0,
                           W.S);
        e = read_scalars(W.S, W.hS, &h, s);
        if (e != hipSuccess) return e;
    }
    *iters = total;
    *sweeps = nsweep - skipped;
    *relres = (h.bb > 0.0) ? sqrt(h.rr64 / h.bb) : sqrt(h.rr64);
    *converged = (h.rr64 <= h.thr2);
    return hipSuccess;
}

// z = M r for the inner solve's preconditioner (multilevel cycle, learned M, or Jacobi), applied in
// FP32 to FP64 device vectors
namespace {

__global__ void k_set_running(lv_cg_scalars *S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) S->done = 0;
}
// b[i] = a[idx[i]] (idx null: a[i]) with the precision change
__global__ __launch_bounds__(B) void k_to32(int n, const double *__restrict__ a, const int *__restrict__ idx,
                                            float *__restrict__ b)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = (float)a[idx ? idx[i] : i];
}
__global__ __launch_bounds__(B) void k_to64(int n, const float *__restrict__ a, const int *__restrict__ idx,
                                            double *__restrict__ b)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = (double)a[idx ? idx[i] : i];
}
__global__ __launch_bounds__(B) void k_jacobi32(int n, const float *__restrict__ dinv, const float *__restrict__ r,
                                                float *__restrict__ z)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) z[i] = dinv[i] * r[i];
}

}  // namespace

hipError_t lv_apply_precond(const lv_pressure_system &A, lv_solver_work &W, int n, const lv_precond *M,
                            const double *r, double *z, hipStream_t s)
{
    const parts P = parts_of(W.partials);
    // the multilevel cycle in the aggregate order takes r and returns z in that order
    const lv_amg *Gp = (M && M->amg && M->amg->perm_ok) ? M->amg : nullptr;
    hipLaunchKernelGGL(k_to32, grid_of(n), dim3(B), 0, s, n, r, Gp ? (const int *)Gp->list : (const int *)nullptr,
                       W.r32);
    hipLaunchKernelGGL(k_set_running, dim3(1), dim3(64), 0, s, W.S);
    if (M && M->amg) {
        hipError_t e = lv_amg_vcycle(M->amg, A, (const float *)W.r32, W.z32, P.rr, (const lv_cg_scalars *)W.S, s);
        if (e != hipSuccess) return e;
    } else if (M && M->mval) {
        hipLaunchKernelGGL(k_apply_m32, dim3(nblocks_of(n)), dim3(B), 0, s, n, (const int *)A.deg,
                           (const int *)A.ecol, M->mval, M->mdg, (const float *)W.r32, W.z32, P.rr,
                           (const lv_cg_scalars *)W.S);
    } else {
        hipLaunchKernelGGL(k_jacobi32, grid_of(n), dim3(B), 0, s, n, (const float *)A.dinv32, (const float *)W.r32,
                           W.z32);
    }
    hipLaunchKernelGGL(k_to64, grid_of(n), dim3(B), 0, s, n, (const float *)W.z32,
                       Gp ? (const int *)Gp->iperm : (const int *)nullptr, z);
    return hipGetLastError();
}

// ---- host helpers for the Picard extrapolation ------------------------------------------------

namespace {
template <typename TA>
hipError_t dots_device(int n, const TA *const *a, const double *const *b, int m, double *partials, double *dout,
                       hipStream_t s)
{
    lv_vec8<TA> V;
    for (int q = 0; q < 8; q++) {
        V.a[q] = (q < m) ? a[q] : a[0];
        V.b[q] = (q < m) ? b[q] : b[0];
    }
    const int nb = nblocks_of(n);
    switch (m) {
#define LV_MD(M)                                                                                         \
    case M:                                                                                              \
        hipLaunchKernelGGL((k_multidot<TA, M>), dim3(nb), dim3(B), 0, s, n, V, partials);             \
        break;
        LV_MD(1) LV_MD(2) LV_MD(3) LV_MD(4) LV_MD(5) LV_MD(6) LV_MD(7) LV_MD(8)
#undef LV_MD
    default:
        hipLaunchKernelGGL((k_multidot<TA, 0>), dim3(nb), dim3(B), 0, s, n, V, partials);
    }
    hipLaunchKernelGGL(k_multidot_fin, dim3(1), dim3(B), 0, s, (const double *)partials, nb, dout);
    return hipGetLastError();
}

template <typename TX, typename TY>
hipError_t combine(int n, TY *y, double alpha, const TX *const *x, const double *c, int m, hipStream_t s,
                   const double *dc, int dmode)
{
    lv_comb8<TX> C;
    for (int q = 0; q < 8; q++) {  // unused slots (q >= m, never read) hold a valid device pointer
        C.x[q] = (q < m) ? x[q] : (m > 0 ? x[0] : reinterpret_cast<const TX *>(y));
        C.c[q] = (q < m) ? c[q] : 0.0;
    }
    const dim3 G(nblocks_of(n));
    switch (m) {
#define LV_CB(M)                                                                                         \
    case M:                                                                                              \
        hipLaunchKernelGGL((k_combine<TX, TY, M>), G, dim3(B), 0, s, n, y, alpha, C, dc, dmode);      \
        break;
        LV_CB(1) LV_CB(2) LV_CB(3) LV_CB(4) LV_CB(5) LV_CB(6) LV_CB(7) LV_CB(8)
#undef LV_CB
    default:
        hipLaunchKernelGGL((k_combine<TX, TY, 0>), G, dim3(B), 0, s, n, y, alpha, C, dc, dmode);
    }
    return hipGetLastError();
}
}  // namespace

hipError_t lv_dots_device(int n, const double *const *a, const double *const *b, int m, double *partials,
                          double *dout, hipStream_t s)
{
    return dots_device(n, a, b, m, partials, dout, s);
}
hipError_t lv_dots_device(int n, const float *const *a, const double *const *b, int m, double *partials,
                          double *dout, hipStream_t s)
{
    return dots_device(n, a, b, m, partials, dout, s);
}

hipError_t lv_combine(int n, double *y, double alpha, const double *const *x, const double *c, int m, hipStream_t s,
                      const double *dc, int dmode)
{
    return combine(n, y, alpha, x, c, m, s, dc, dmode);
}
hipError_t lv_combine(int n, double *y, double alpha, const float *const *x, const double *c, int m, hipStream_t s,
                      const double *dc, int dmode)
{
    return combine(n, y, alpha, x, c, m, s, dc, dmode);
}
hipError_t lv_combine(int n, float *y, double alpha, const double *const *x, const double *c, int m, hipStream_t s,
                      const double *dc, int dmode)
{
    return combine(n, y, alpha, x, c, m, s, dc, dmode);
}

hipError_t lv_spmv64(const lv_pressure_system &A, int n, const double *x, double *y, hipStream_t s)
{
    if (A.narrow)
        launch_spmv64_v(own64s(A, n), dim3(nblocks_of(n)), s, n, (const double *)A.diag64, x, y);
    else
        launch_spmv64_v(own64(A, n), dim3(nblocks_of(n)), s, n, (const double *)A.diag64, x, y);
    return hipGetLastError();
}

hipError_t lv_apply_from_residual(int n, const double *b, const double *r, double *Ax, const lv_solver_work &W,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_from_residual, dim3(nblocks_of(n)), dim3(B), 0, s, n, b, r, Ax,
                       (const lv_cg_scalars *)W.S);
    return hipGetLastError();
}
