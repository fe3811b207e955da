// lv_stencil.hip -- per-particle physics operators over the Voronoi mesh: move!, gravity_step!,
// find_rho!, the equations of state, pressure_step!, find_D!, viscous_step!, bdary_friction!,
// find_dv! and relaxation_step!.
//
// One thread per particle; each walks its own ELL edge slots (coalesced across the wavefront) and
// gathers neighbour fields by label.  Every two-pass operator is two launches, because the second
// pass of the reference reads neighbour values the first pass wrote (e.g. pressure_step! reads the
// updated q.v).  A first pass only ever writes fields of its own particle that no neighbour reads in
// that pass, so it updates in place exactly as the reference's @batch loops do.  Expression order
// follows the reference line cited at each kernel; norms are sqrt(a*a + b*b).
#include "lv_internal.h"

namespace {

constexpr int B = 256;

__device__ __forceinline__ void nbr_y(const lv_grid_params &g, const lv_fields &F, int n, int j, double x1, double x2,
                                      double &y1, double &y2)
{
    lv_arrow_target(g, F.x[j], F.x[n + j], x1, x2, y1, y2);
}

// least_positive_residue, src/voronoigrid.jl:181-183
__device__ __forceinline__ double lpr(double x, double d) { return fmod(fmod(x, d) + d, d); }

// try_move!, src/move.jl:23-33: 1 moved, 0 not inside, -1 NaN velocity
__device__ __forceinline__ int try_move(const lv_grid_params &g, double &x1, double &x2, double v1, double v2,
                                        double dt)
{
    if (isnan(v1) || isnan(v2)) return -1;
    const double a1 = x1 + dt * v1, a2 = x2 + dt * v2;
    // periodic_wrap, src/voronoigrid.jl:186-192 (Bool*Float64 gives a signed zero when false)
    const double wx = lpr(a1 - g.bxmin0, g.xperiod) + g.bxmin0;
    const double wy = lpr(a2 - g.bxmin1, g.yperiod) + g.bxmin1;
    const double tx = g.xperiodic ? (wx - a1) : copysign(0.0, wx - a1);
    const double ty = g.yperiodic ? (wy - a2) : copysign(0.0, wy - a2);
    const double r1 = (a1 + tx * 1.0) + ty * 0.0;
    const double r2 = (a2 + tx * 0.0) + ty * 1.0;
    if ((g.bxmin0 <= r1 && r1 <= g.bxmax0) && (g.bxmin1 <= r2 && r2 <= g.bxmax1)) {
        x1 = r1;
        x2 = r2;
        return 1;
    }
    return 0;
}

// move!, src/move.jl:9-21 (the position update; the caller remeshes)
__global__ __launch_bounds__(B) void k_move(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                            int *err)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    double x1 = F.x[i], x2 = F.x[n + i];
    double v1 = F.v[i], v2 = F.v[n + i];
    int r = try_move(g, x1, x2, v1, v2, dt);
    if (r == 0) {
        const int ne = M.ne[i];
        for (int k = 0; k < ne; k++) {  // boundaries(p): project v to the tangent space
            if (LV_EL(M, k, i) >= 0) continue;
            double n1 = LV_EV(M, 1, k, i) - LV_EV(M, 3, k, i);
            double n2 = LV_EV(M, 2, k, i) - LV_EV(M, 0, k, i);
            const double nn = lv_norm2(n1, n2);  // normal_vector(e), src/polygon.jl:153-156
            n1 = n1 / nn;
            n2 = n2 / nn;
            const double vn = v1 * n1 + v2 * n2;
            v1 = v1 - vn * n1;
            v2 = v2 - vn * n2;
        }
        r = try_move(g, x1, x2, v1, v2, dt);
        if (r == 0) {
            v1 = 0.0;
            v2 = 0.0;
            r = try_move(g, x1, x2, v1, v2, dt);
        }
    }
    if (r < 0) atomicCAS(err, 0, LV_E_NAN_VELOCITY);
    F.x[i] = x1;
    F.x[n + i] = x2;
    F.v[i] = v1;
    F.v[n + i] = v2;
}

// gravity_step!, src/pressure.jl:77-82
__global__ __launch_bounds__(B) void k_gravity(lv_fields F, int n, double g1, double g2, double dt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    F.v[i] += dt * g1;
    F.v[n + i] += dt * g2;
}

// find_rho!, src/pressure.jl:36-40
__global__ __launch_bounds__(B) void k_find_rho(lv_fields F, int n)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    F.rho[i] = F.mass[i] / F.area[i];
}

// stiffened_eos! (kind 1, src/pressure.jl:64-70) and ideal_eos! (kind 2, :49-55)
__global__ __launch_bounds__(B) void k_eos(lv_fields F, int n, int kind, double gamma, double P0, double Pmin)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double rho = F.mass[i] / F.area[i];
    const double v1 = F.v[i], v2 = F.v[n + i];
    const double ein = F.e[i] - 0.5 * (v1 * v1 + v2 * v2);  // eint, :32-34
    const double P = ((gamma - 1.0) * rho) * ein;
    F.rho[i] = rho;
    F.P[i] = P;
    if (kind == 1) {
        F.c2[i] = (gamma * (P + P0)) / rho;
    } else {
        const double pm = (Pmin > P) ? Pmin : P;
        F.c2[i] = (gamma * pm) / rho;
    }
}

// Edge slots of a row in batches of EB: the batch's labels and edge vertices are loaded together
// (a slot past the row's last reads the row's last slot again, a cache hit), then the
// neighbours' fields together (a boundary slot or one past the row gathers the row itself), and
// only then is the batch applied, edge by edge in the reference's order.  A loop that loads each
// edge's label, then its vertices and its neighbour's fields, waits for three or four round trips
// per edge.
constexpr int EB = 4;
struct edge_batch {
    int j[EB];
    bool on[EB];
    double e1x[EB], e1y[EB], e2x[EB], e2y[EB];
};
__device__ __forceinline__ void load_edges(const lv_mesh_view &M, int i, int k0, int ne, edge_batch &E)
{
#pragma unroll
    for (int u = 0; u < EB; u++) {
        const int k = k0 + u < ne ? k0 + u : ne - 1;
        const int j = LV_EL(M, k, i);
        E.e1x[u] = LV_EV(M, 0, k, i);
        E.e1y[u] = LV_EV(M, 1, k, i);
        E.e2x[u] = LV_EV(M, 2, k, i);
        E.e2y[u] = LV_EV(M, 3, k, i);
        E.on[u] = (k0 + u < ne) & (j >= 0);
        E.j[u] = E.on[u] ? j : i;
    }
}

// pressure_step! pass 1, src/pressure.jl:11-17: v += dt/m * lrr * (P_i - P_j) * (m - x)
__global__ __launch_bounds__(B) void k_pstep1(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    double v1 = F.v[i], v2 = F.v[n + i];
    const double Pi = F.P[i], dtm = dt / F.mass[i];
    const int ne = M.ne[i];
    for (int k0 = 0; k0 < ne; k0 += EB) {
        edge_batch E;
        load_edges(M, i, k0, ne, E);
        double xj1[EB], xj2[EB], Pj[EB];
#pragma unroll
        for (int u = 0; u < EB; u++) {
            const int j = E.j[u];
            xj1[u] = F.x[j];
            xj2[u] = F.x[n + j];
            Pj[u] = F.P[j];
        }
#pragma unroll
        for (int u = 0; u < EB; u++) {
            if (!E.on[u]) continue;
            double y1, y2;
            lv_arrow_target(g, xj1[u], xj2[u], x1, x2, y1, y2);
            const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, E.e1x[u], E.e1y[u], E.e2x[u], E.e2y[u]);
            const double m1 = 0.5 * (E.e1x[u] + E.e2x[u]), m2 = 0.5 * (E.e1y[u] + E.e2y[u]);
            const double w = (dtm * lrr) * (Pi - Pj[u]);
            v1 = v1 + w * (m1 - x1);
            v2 = v2 + w * (m2 - x2);
        }
    }
    F.v[i] = v1;
    F.v[n + i] = v2;
}

// pressure_step! pass 2, src/pressure.jl:18-24: e -= dt*lrr/m * ((m-x).(P_i v_i) - (m-y).(P_j v_j))
// WITH_D: find_D! (src/diffusion.jl:8-19) in the same pass, as step! calls it right after
// pressure_step! (both read the velocities pass 1 left and the same edges; the sums are k_find_D's)
template <bool WITH_D>
__global__ __launch_bounds__(B) void k_pstep2(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double Pi = F.P[i], mi = F.mass[i];
    const double vi1 = F.v[i], vi2 = F.v[n + i];
    const double pv1 = Pi * vi1, pv2 = Pi * vi2;
    double ei = F.e[i];
    double D11 = 0.0, D21 = 0.0, D12 = 0.0, D22 = 0.0;
    const int ne = M.ne[i];
    for (int k0 = 0; k0 < ne; k0 += EB) {
        edge_batch E;
        load_edges(M, i, k0, ne, E);
        double xj1[EB], xj2[EB], Pq[EB], vq1[EB], vq2[EB];
#pragma unroll
        for (int u = 0; u < EB; u++) {
            const int j = E.j[u];
            xj1[u] = F.x[j];
            xj2[u] = F.x[n + j];
            Pq[u] = F.P[j];
            vq1[u] = F.v[j];
            vq2[u] = F.v[n + j];
        }
#pragma unroll
        for (int u = 0; u < EB; u++) {
            if (!E.on[u]) continue;
            double y1, y2;
            lv_arrow_target(g, xj1[u], xj2[u], x1, x2, y1, y2);
            const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, E.e1x[u], E.e1y[u], E.e2x[u], E.e2y[u]);
            const double m1 = 0.5 * (E.e1x[u] + E.e2x[u]), m2 = 0.5 * (E.e1y[u] + E.e2y[u]);
            const double Pj = Pq[u];
            const double vj1 = vq1[u], vj2 = vq2[u];
            const double qv1 = Pj * vj1, qv2 = Pj * vj2;
            const double a = (m1 - x1) * pv1 + (m2 - x2) * pv2;
            const double b = (m1 - y1) * qv1 + (m2 - y2) * qv2;
            ei = ei - ((dt * lrr) / mi) * (a - b);
            if (WITH_D) {
                const double a1 = vi1 - vj1, a2 = vi2 - vj2;
                const double b1 = m1 - y1, b2 = m2 - y2;
                D11 = D11 + lrr * (a1 * b1);
                D21 = D21 + lrr * (a2 * b1);
                D12 = D12 + lrr * (a1 * b2);
                D22 = D22 + lrr * (a2 * b2);
            }
        }
    }
    F.e[i] = ei;
    if (WITH_D) {
        const double A = F.area[i];
        D11 = D11 / A;
        D21 = D21 / A;
        D12 = D12 / A;
        D22 = D22 / A;
        F.D[i] = 0.5 * (D11 + D11);
        F.D[n + i] = 0.5 * (D21 + D12);
        F.D[2 * n + i] = 0.5 * (D12 + D21);
        F.D[3 * n + i] = 0.5 * (D22 + D22);
    }
}

// find_D!, src/diffusion.jl:8-19
__global__ __launch_bounds__(B) void k_find_D(lv_grid_params g, lv_fields F, lv_mesh_view M, int n)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double pv1 = F.v[i], pv2 = F.v[n + i];
    double D11 = 0.0, D21 = 0.0, D12 = 0.0, D22 = 0.0;
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        double y1, y2;
        nbr_y(g, F, n, j, x1, x2, y1, y2);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
        const double a1 = pv1 - F.v[j], a2 = pv2 - F.v[n + j];
        const double b1 = m1 - y1, b2 = m2 - y2;
        D11 = D11 + lrr * (a1 * b1);
        D21 = D21 + lrr * (a2 * b1);
        D12 = D12 + lrr * (a1 * b2);
        D22 = D22 + lrr * (a2 * b2);
    }
    const double A = F.area[i];
    D11 = D11 / A;
    D21 = D21 / A;
    D12 = D12 / A;
    D22 = D22 / A;
    F.D[i] = 0.5 * (D11 + D11);
    F.D[n + i] = 0.5 * (D21 + D12);
    F.D[2 * n + i] = 0.5 * (D12 + D21);
    F.D[3 * n + i] = 0.5 * (D22 + D22);
}

// getS, src/diffusion.jl:22-29 (column-major S), from a cell's D, mu and rho
__device__ __forceinline__ void getS(double D11, double D21, double D12, double D22, double mu0, double rho,
                                        double dr, double S[4])
{
    const double divv = ((D11 * 1.0 + D21 * 0.0) + D12 * 0.0) + D22 * 1.0;
    double mu = mu0;
    if (divv < 0.0) mu -= (divv * rho) * (dr * dr);
    const double tm = 2.0 * mu;
    S[0] = tm * (D11 - (divv * 1.0) / 3.0);
    S[1] = tm * (D21 - (divv * 0.0) / 3.0);
    S[2] = tm * (D12 - (divv * 0.0) / 3.0);
    S[3] = tm * (D22 - (divv * 1.0) / 3.0);
}

// viscous_step! pass 1, src/diffusion.jl:41-46: v -= dt*lrr/m * (S_p - S_q) * (m - x)
// (three-argument s*M*w evaluated as M*(w*s))
__global__ __launch_bounds__(B) void k_visc1(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                             double avdr)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    double v1 = F.v[i], v2 = F.v[n + i];
    double Sp[4];
    getS(F.D[i], F.D[n + i], F.D[2 * n + i], F.D[3 * n + i], F.mu[i], F.rho[i], avdr, Sp);
    const double mi = F.mass[i];
    const int ne = M.ne[i];
    for (int k0 = 0; k0 < ne; k0 += EB) {
        edge_batch E;
        load_edges(M, i, k0, ne, E);
        double xj1[EB], xj2[EB], D11[EB], D21[EB], D12[EB], D22[EB], muj[EB], rhoj[EB];
#pragma unroll
        for (int u = 0; u < EB; u++) {
            const int j = E.j[u];
            xj1[u] = F.x[j];
            xj2[u] = F.x[n + j];
            D11[u] = F.D[j];
            D21[u] = F.D[n + j];
            D12[u] = F.D[2 * n + j];
            D22[u] = F.D[3 * n + j];
            muj[u] = F.mu[j];
            rhoj[u] = F.rho[j];
        }
#pragma unroll
        for (int u = 0; u < EB; u++) {
            if (!E.on[u]) continue;
            double y1, y2;
            lv_arrow_target(g, xj1[u], xj2[u], x1, x2, y1, y2);
            const double m1 = 0.5 * (E.e1x[u] + E.e2x[u]), m2 = 0.5 * (E.e1y[u] + E.e2y[u]);
            double Sq[4];
            getS(D11[u], D21[u], D12[u], D22[u], muj[u], rhoj[u], avdr, Sq);
            const double sc = (dt * lv_lr_ratio(x1 - y1, x2 - y2, E.e1x[u], E.e1y[u], E.e2x[u], E.e2y[u])) / mi;
            const double M11 = Sp[0] - Sq[0], M21 = Sp[1] - Sq[1], M12 = Sp[2] - Sq[2], M22 = Sp[3] - Sq[3];
            const double w1 = (m1 - x1) * sc, w2 = (m2 - x2) * sc;
            v1 = v1 - (M11 * w1 + M12 * w2);
            v2 = v2 - (M21 * w1 + M22 * w2);
        }
    }
    F.v[i] = v1;
    F.v[n + i] = v2;
}

// quality and dv of find_dv!, src/relaxation.jl:10-25, from the neighbour distances rmin, rmax (the
// centroid cached by the clipper, same expression order)
__device__ __forceinline__ void find_dv_row(lv_fields &F, int n, int i, double x1, double x2, double rmin, double rmax,
                                            double dt, double alpha)
{
    const double q = rmin / rmax;
    F.quality[i] = q;
    const double D11 = F.D[i], D21 = F.D[n + i], D12 = F.D[2 * n + i], D22 = F.D[3 * n + i];
    const double nD = sqrt(((D11 * D11 + D21 * D21) + D12 * D12) + D22 * D22);
    const double lambda = (alpha * nD) / (q * q);
    const double f = lambda / (1.0 + dt * lambda);
    F.dv[i] = f * (F.cent[i] - x1);
    F.dv[n + i] = f * (F.cent[n + i] - x2);
}

// viscous_step! pass 2, src/diffusion.jl:47-52.  WITH_DV: find_dv!(dt_dv, alpha) in the same pass (it
// reads D, the positions and the mesh, none of which the viscous step or bdary_friction! change)
template <bool WITH_DV>
__global__ __launch_bounds__(B) void k_visc2(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                             double avdr, double dt_dv, double alpha)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    double ei = F.e[i];
    double Sp[4];
    getS(F.D[i], F.D[n + i], F.D[2 * n + i], F.D[3 * n + i], F.mu[i], F.rho[i], avdr, Sp);
    const double pv1 = F.v[i], pv2 = F.v[n + i];
    const double sv1 = Sp[0] * pv1 + Sp[2] * pv2;
    const double sv2 = Sp[1] * pv1 + Sp[3] * pv2;
    const double mi = F.mass[i];
    double rmax = 0.0, rmin = __builtin_inf();
    const int ne = M.ne[i];
    for (int k0 = 0; k0 < ne; k0 += EB) {
        edge_batch E;
        load_edges(M, i, k0, ne, E);
        double xj1[EB], xj2[EB], D11[EB], D21[EB], D12[EB], D22[EB], muj[EB], rhoj[EB], vq1[EB], vq2[EB];
#pragma unroll
        for (int u = 0; u < EB; u++) {
            const int j = E.j[u];
            xj1[u] = F.x[j];
            xj2[u] = F.x[n + j];
            D11[u] = F.D[j];
            D21[u] = F.D[n + j];
            D12[u] = F.D[2 * n + j];
            D22[u] = F.D[3 * n + j];
            muj[u] = F.mu[j];
            rhoj[u] = F.rho[j];
            vq1[u] = F.v[j];
            vq2[u] = F.v[n + j];
        }
#pragma unroll
        for (int u = 0; u < EB; u++) {
            if (!E.on[u]) continue;
            double y1, y2;
            lv_arrow_target(g, xj1[u], xj2[u], x1, x2, y1, y2);
            const double m1 = 0.5 * (E.e1x[u] + E.e2x[u]), m2 = 0.5 * (E.e1y[u] + E.e2y[u]);
            double Sq[4];
            getS(D11[u], D21[u], D12[u], D22[u], muj[u], rhoj[u], avdr, Sq);
            const double qv1 = Sq[0] * vq1[u] + Sq[2] * vq2[u];
            const double qv2 = Sq[1] * vq1[u] + Sq[3] * vq2[u];
            const double a = (m1 - x1) * sv1 + (m2 - x2) * sv2;
            const double b = (m1 - y1) * qv1 + (m2 - y2) * qv2;
            ei = ei + ((dt * lv_lr_ratio(x1 - y1, x2 - y2, E.e1x[u], E.e1y[u], E.e2x[u], E.e2y[u])) / mi) * (a - b);
            if (WITH_DV) {
                const double r = lv_norm2(x1 - y1, x2 - y2);
                if (r > rmax) rmax = r;
                if (r < rmin) rmin = r;
            }
        }
    }
    F.e[i] = ei;
    if (WITH_DV) find_dv_row(F, n, i, x1, x2, rmin, rmax, dt_dv, alpha);
}

// bdary_friction!, src/diffusion.jl:64-78, vDirichlet(m) = vbc[label]; charfun(m) is false on the
// walls whose table entry is NaN (their edges are skipped, as the reference's `continue` does)
__global__ __launch_bounds__(B) void k_friction(lv_fields F, lv_mesh_view M, int n, lv_vbc vbc, double dt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    double v1 = F.v[i], v2 = F.v[n + i];
    double ei = F.e[i];
    const double mu = F.mu[i], mi = F.mass[i];
    double tmp = 1.0;
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int lab = LV_EL(M, k, i);
        if (lab >= 0) continue;
        const int slot = -lab - 1;
        const double vd1 = vbc.v[2 * slot], vd2 = vbc.v[2 * slot + 1];
        if (isnan(vd1) || isnan(vd2)) continue;
        const double e1x = LV_EV(M, 0, k, i), e1y = LV_EV(M, 1, k, i), e2x = LV_EV(M, 2, k, i), e2y = LV_EV(M, 3, k, i);
        const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
        double n1 = e1y - e2y, n2 = e2x - e1x;
        const double nn = lv_norm2(n1, n2);
        n1 = n1 / nn;
        n2 = n2 / nn;
        const double len = lv_norm2(e1x - e2x, e1y - e2y);
        const double lrr = len / fabs((m1 - x1) * n1 + (m2 - x2) * n2);
        const double ml = mu * lrr;
        const double f1 = (ml * vd1) / mi, f2 = (ml * vd2) / mi;
        tmp += ((dt * mu) * lrr) / mi;
        ei += dt * (f1 * v1 + f2 * v2);
        v1 = v1 + dt * f1;
        v2 = v2 + dt * f2;
    }
    F.e[i] = ei;
    F.v[i] = v1 / tmp;
    F.v[n + i] = v2 / tmp;
}

// find_dv!, src/relaxation.jl:10-25 (centroid(p) cached by the clipper, same expression order)
__global__ __launch_bounds__(B) void k_find_dv(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                               double alpha)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    double rmax = 0.0, rmin = __builtin_inf();
    const int ne = M.ne[i];
    for (int k = 0; k < ne; k++) {
        const int j = LV_EL(M, k, i);
        if (j < 0) continue;
        double y1, y2;
        nbr_y(g, F, n, j, x1, x2, y1, y2);
        const double r = lv_norm2(x1 - y1, x2 - y2);
        if (r > rmax) rmax = r;
        if (r < rmin) rmin = r;
    }
    find_dv_row(F, n, i, x1, x2, rmin, rmax, dt, alpha);
}

// relaxation_step! pass 1, src/relaxation.jl:37-59 (writes mass, momentum, energy of p only)
__global__ __launch_bounds__(B) void k_relax1(lv_grid_params g, lv_fields F, lv_mesh_view M, int n, double dt,
                                              int rusanov)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double x1 = F.x[i], x2 = F.x[n + i];
    const double rp = F.rho[i], ep = F.e[i];
    const double pv1 = F.v[i], pv2 = F.v[n + i];
    const double pd1 = F.dv[i], pd2 = F.dv[n + i];
    const int php = F.phase[i];
    double mass = F.mass[i];
    double mo1 = mass * pv1, mo2 = mass * pv2;
    double en = mass * ep;
    const double na = lv_norm2(pd1, pd2);
    const int ne = M.ne[i];
    for (int k0 = 0; k0 < ne; k0 += EB) {
        edge_batch E;
        load_edges(M, i, k0, ne, E);
        double xj1[EB], xj2[EB], rq_[EB], eq_[EB], vq1[EB], vq2[EB], dq1[EB], dq2[EB];
        int phq[EB];
#pragma unroll
        for (int u = 0; u < EB; u++) {
            const int j = E.j[u];
            phq[u] = F.phase[j];
            xj1[u] = F.x[j];
            xj2[u] = F.x[n + j];
            rq_[u] = F.rho[j];
            eq_[u] = F.e[j];
            vq1[u] = F.v[j];
            vq2[u] = F.v[n + j];
            dq1[u] = F.dv[j];
            dq2[u] = F.dv[n + j];
        }
#pragma unroll
        for (int u = 0; u < EB; u++) {
            if (!E.on[u]) continue;
            if (!(php == phq[u])) continue;
            const double e1x = E.e1x[u], e1y = E.e1y[u], e2x = E.e2x[u], e2y = E.e2y[u];
            double y1, y2;
            lv_arrow_target(g, xj1[u], xj2[u], x1, x2, y1, y2);
            const double lrr = lv_lr_ratio(x1 - y1, x2 - y2, e1x, e1y, e2x, e2y);
            const double m1 = 0.5 * (e1x + e2x), m2 = 0.5 * (e1y + e2y);
            const double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
            const double rq = rq_[u], eq = eq_[u];
            const double qv1 = vq1[u], qv2 = vq2[u];
            const double qd1 = dq1[u], qd2 = dq2[u];
            const double pdvpq = pd1 * (x1 - y1) + pd2 * (x2 - y2);
            const double qdvpq = qd1 * (x1 - y1) + qd2 * (x2 - y2);
            const double pdvmz = pd1 * (m1 - z1) + pd2 * (m2 - z2);
            const double qdvmz = qd1 * (m1 - z1) + qd2 * (m2 - z2);
            const double dl = dt * lrr;
            const double pr_mz = pdvmz * rp, qr_mz = qdvmz * rq, pr_pq = pdvpq * rp, qr_pq = qdvpq * rq;
            mass += dl * ((pr_mz - qr_mz) - 0.5 * (pr_pq + qr_pq));
            mo1 += dl * ((pr_mz * pv1 - qr_mz * qv1) - 0.5 * (pr_pq * pv1 + qr_pq * qv1));
            mo2 += dl * ((pr_mz * pv2 - qr_mz * qv2) - 0.5 * (pr_pq * pv2 + qr_pq * qv2));
            en += dl * ((pr_mz * ep - qr_mz * eq) - 0.5 * (pr_pq * ep + qr_pq * eq));
            if (rusanov) {  // src/relaxation.jl:52-58
                const double nb = lv_norm2(qd1, qd2);
                const double a = (nb > na) ? nb : na;
                const double l = lv_norm2(e1x - e2x, e1y - e2y);
                const double w = ((0.5 * dt) * l) * a;
                mass += w * (rq - rp);
                mo1 += w * (rq * qv1 - rp * pv1);
                mo2 += w * (rq * qv2 - rp * pv2);
                en += w * (rq * eq - rp * ep);
            }
        }
    }
    F.mass[i] = mass;
    F.mom[i] = mo1;
    F.mom[n + i] = mo2;
    F.energy[i] = en;
}

// relaxation_step! pass 2, src/relaxation.jl:60-64
__global__ __launch_bounds__(B) void k_relax2(lv_fields F, int n, double dt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double m = F.mass[i];
    F.v[i] = F.mom[i] / m;
    F.v[n + i] = F.mom[n + i] / m;
    F.e[i] = F.energy[i] / m;
    F.x[i] += dt * F.dv[i];
    F.x[n + i] += dt * F.dv[n + i];
}

inline dim3 grid_of(int n) { return dim3((n + B - 1) / B); }

}  // namespace

hipError_t lv_launch_move(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int *err,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_move, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, err);
    return hipGetLastError();
}

hipError_t lv_launch_gravity(lv_fields F, int n, double g1, double g2, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_gravity, grid_of(n), dim3(B), 0, s, F, n, g1, g2, dt);
    return hipGetLastError();
}

hipError_t lv_launch_find_rho(lv_fields F, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_find_rho, grid_of(n), dim3(B), 0, s, F, n);
    return hipGetLastError();
}

hipError_t lv_launch_eos(lv_fields F, int n, int kind, double gamma, double P0, double Pmin, hipStream_t s)
{
    hipLaunchKernelGGL(k_eos, grid_of(n), dim3(B), 0, s, F, n, kind, gamma, P0, Pmin);
    return hipGetLastError();
}

hipError_t lv_launch_pressure_step(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt,
                                   hipStream_t s, bool with_find_D)
{
    hipLaunchKernelGGL(k_pstep1, grid_of(n), dim3(B), 0, s, g, F, M, n, dt);
    if (with_find_D) hipLaunchKernelGGL(k_pstep2<true>, grid_of(n), dim3(B), 0, s, g, F, M, n, dt);
    else hipLaunchKernelGGL(k_pstep2<false>, grid_of(n), dim3(B), 0, s, g, F, M, n, dt);
    return hipGetLastError();
}

hipError_t lv_launch_find_D(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_find_D, grid_of(n), dim3(B), 0, s, g, F, M, n);
    return hipGetLastError();
}

hipError_t lv_launch_viscous(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, double avdr,
                             hipStream_t s, bool with_find_dv, double dv_dt, double alpha)
{
    hipLaunchKernelGGL(k_visc1, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, avdr);
    if (with_find_dv)
        hipLaunchKernelGGL(k_visc2<true>, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, avdr, dv_dt, alpha);
    else
        hipLaunchKernelGGL(k_visc2<false>, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, avdr, 0.0, 0.0);
    return hipGetLastError();
}

hipError_t lv_launch_friction(lv_fields F, lv_mesh_view M, int n, lv_vbc vbc, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_friction, grid_of(n), dim3(B), 0, s, F, M, n, vbc, dt);
    return hipGetLastError();
}

hipError_t lv_launch_find_dv(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, double alpha,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_find_dv, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, alpha);
    return hipGetLastError();
}

hipError_t lv_launch_relaxation(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int rusanov,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_relax1, grid_of(n), dim3(B), 0, s, g, F, M, n, dt, rusanov);
    hipLaunchKernelGGL(k_relax2, grid_of(n), dim3(B), 0, s, F, n, dt);
    return hipGetLastError();
}
