// lv_solver.h -- pressure-system storage, solver workspace and kernel timers of liblvhip.so.
#pragma once

#include "lv_internal.h"
#include "lv_ell.h"

// Device-resident CG scalars: written by the last block of each reduction, read by the next launch.
// rz is double-buffered by iteration parity: the p-update of iteration k reads rz[k&1] in every
// block while block 0 writes rz[(k+1)&1].
struct lv_cg_scalars {
    double rz[2], rr, rr0, pq, alpha, beta, bb, thr2, tol2, rr64, bmean, xmean0, xmean, xshift;
    double rr64_0, itol2;  // the initial FP64 residual^2; the first sweep's inner tolerance^2
    int iters, itmax, done, singular;
    int skip;  // the initial residual already meets the tolerance (or is NaN): no sweep is needed
};

// PressureOperator (src/pressure.jl:89-101) in CSR: off-diagonal coefficients c_ij in the edge order
// of row i, the compressibility diagonal d_i, and the Jacobi inverse 1/(d_i + sum_j c_ij), each in
// FP64 and FP32.
//
// The solver reads a slot-major ELL copy of the same coefficients: slot k of row i at [k*n + i],
// deg[i] slots per row, so slot k of a wavefront's 64 rows is one contiguous run.
struct lv_pressure_system {
    int *row_ptr, *cols;
    double *val64, *diag64, *dinv64;
    float *val32, *diag32, *dinv32;
    int *deg, *ecol;
    double *ev64;
    float *ev32;
    // the compact copy of slots LV_OVF_BASE.. per 64-row group (lv_ell.h): columns, FP64 and FP32
    // values, for the groups with at most LV_OVF_CAP such slots
    int *ocol;
    double *oev64;
    float *oev32;
    // the same columns as 16-bit offsets j - i (taken modulo n, lv_ell.h) with uint8 degrees, for the
    // FP64 passes; narrow: every offset of the last assembly fits, so they hold the operator
    uint8_t *deg8;
    int16_t *ecol16, *ocol16;
    int narrow;
    int64_t nnz, nnz_cap;
    int singular;  // every d_i == 0 (c2 = Inf): a graph Laplacian with the constants as null space
    int valid;
};

// Progress of a running CG solve, stored by the device in pinned host memory: seen = the last
// iteration whose convergence test ran, done = the solve's done flag.  The host reads it without
// synchronising to keep its launches one iteration ahead of the device.  After a refinement
// sweep's FP64 residual, k_finalize also publishes there what the host tests (the scalars of
// lv_cg_scalars of those names) and then fseq, the number the host asked for.
struct lv_progress {
    int seen, done;
    int fseq, iters, skip, pad;
    double rr64, rr64_0, bb, thr2;
};

struct lv_solver_work {
    double *r, *z, *p, *q;
    float *e32, *r32, *z32, *p32, *q32;
    double *partials;
    lv_cg_scalars *S;
    lv_progress *hprog, *dprog;  // the same pinned record: host and device addresses
    lv_cg_scalars *hS;           // pinned host copy of S for the synchronising reads
};

#define LV_TPOOL 1024
struct lv_timer {
    int on, used;
    int every, pending;  // events around every `every`-th launch; pending: this launch is sampled
    int64_t seen;
    double total_ms;
    int64_t launches;
    double bytes_per_launch;
    hipEvent_t ev[2 * LV_TPOOL];
};
void lv_timer_begin(lv_timer *t, hipStream_t s);
void lv_timer_end(lv_timer *t, hipStream_t s);
void lv_timer_flush(lv_timer *t);

// a preconditioner in the solver's ELL slots: z_i = mdg_i r_i + sum_k mval[k*n + i] r[ecol[k*n + i]]
struct lv_amg;
// The inner solve's preconditioner: the learned M in A's ELL slots, or (amg set) the multilevel
// cycle of lv_amg.hip on the operator A.
struct lv_precond {
    const float *mval = nullptr, *mdg = nullptr;
    const lv_amg *amg = nullptr;
    const lv_pressure_system *A = nullptr;
};

size_t lv_partials_bytes(int n);
hipError_t lv_apply_precond(const lv_pressure_system &A, lv_solver_work &W, int n, const lv_precond *M,
                            const double *r, double *z, hipStream_t s);
hipError_t lv_launch_degree(lv_mesh_view M, int n, int *deg, hipStream_t s);
struct lv_rhs_geom;
// refresh!(A) into A (mode 0 the pressure operator, 1 ViscousSolver's); with R (mode 0) the same pass
// also stores the P-independent parts of refresh!(solver) for boundary velocity vbc
// (wide: set when a column offset does not fit the 16-bit copy)
hipError_t lv_launch_assemble(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, int mode,
                              lv_pressure_system &A, int *nonsingular, int *wide, hipStream_t s,
                              const lv_vbc *vbc = nullptr, const lv_rhs_geom *R = nullptr);
// The P-independent parts of refresh!(solver) for one find_pressure!: per mesh edge slot
// w = lrr (m - x_i) and u = lrr (m - z) (each two planes of LV_MAXE x ld doubles, interior edges only),
// per row the divergence and boundary-flux part of b (bdiv) and rho c2 dt^2 (den).
struct lv_rhs_geom {
    double *w1, *w2, *u1, *u2, *bdiv, *den;
    // the mesh slots LV_OVF_BASE.. of each 64-row group in a compact copy (lv_ell.h), labels and the
    // four planes, for the groups with at most ocap of them (ocap 0: no copy; the writer and the
    // readers of one set of planes see the same ocap)
    int *oel;
    double *ow1, *ow2, *ou1, *ou2;
    int ocap;
};
hipError_t lv_launch_rhs_geom(const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double dt, lv_vbc vbc,
                              const lv_rhs_geom &R, hipStream_t s);
// GP of P (Pout, Pout2: also P written there, either may be null); b of P with (gp_step) or without the
// GP term (P0: also P0 = P)
hipError_t lv_launch_gp(lv_mesh_view M, int n, const lv_rhs_geom &R, const double *P, const double *mass, double *GP,
                        double *Pout, double *Pout2, hipStream_t s);
hipError_t lv_launch_rhs_b(lv_mesh_view M, int n, int gp_step, const lv_rhs_geom &R, const double *GP,
                           const double *area, const double *P, double *b, double *P0, hipStream_t s);
hipError_t lv_launch_visc_rhs(int n, const double *mass, const double *vc, double dt, double *b, double *x,
                              hipStream_t s);
hipError_t lv_launch_write_P(lv_fields F, int n, const double *P, hipStream_t s);
hipError_t lv_solve_device(const lv_pressure_system &A, lv_solver_work &W, int n, const double *b, double *x,
                           const lv_solve_opts &opt, hipStream_t s, lv_timer *t32, lv_timer *t64, int *iters,
                           int *sweeps, double *relres, double *relres0, bool *converged, const lv_precond *M,
                           int *fallbacks);

// Picard extrapolation helpers (vectors of length n on stream s; the window's stored directions
// are FP32, everything else FP64; sums in FP64)
// the dots into the device array dout[0..m) without a host copy
hipError_t lv_dots_device(int n, const double *const *a, const double *const *b, int m, double *partials,
                          double *dout, hipStream_t s);
hipError_t lv_dots_device(int n, const float *const *a, const double *const *b, int m, double *partials,
                          double *dout, hipStream_t s);
// y = alpha y + sum_q c_q x_q, the c_q scaled by device values: dmode 0 none, 1 c_q dc[q],
// 2 c_q / sqrt(dc[0]) (0 unless dc[0] is positive and finite)
hipError_t lv_combine(int n, double *y, double alpha, const double *const *x, const double *c, int m, hipStream_t s,
                      const double *dc = nullptr, int dmode = 0);
hipError_t lv_combine(int n, double *y, double alpha, const float *const *x, const double *c, int m, hipStream_t s,
                      const double *dc = nullptr, int dmode = 0);
hipError_t lv_combine(int n, float *y, double alpha, const double *const *x, const double *c, int m, hipStream_t s,
                      const double *dc = nullptr, int dmode = 0);
hipError_t lv_apply_from_residual(int n, const double *b, const double *r, double *Ax, const lv_solver_work &W,
                                  hipStream_t s);
// y = A x in FP64
hipError_t lv_spmv64(const lv_pressure_system &A, int n, const double *x, double *y, hipStream_t s);
