/*
 * lvhip.h -- C ABI of liblvhip.so, the MI355X (gfx950) implementation of LagrangianVoronoi.jl's
 * per-step hot path: remesh (cell list + Voronoi clipping), the pressure Poisson solve and the
 * particle updates.
 *
 * Conventions
 *   - Every entry point returns an int status (LV_OK = 0).  No C++ exception crosses the boundary.
 *   - Particle labels are 0-based.  Boundary edge labels are exactly the reference's
 *     BDARY_UP = -1, BDARY_RIGHT = -2, BDARY_DOWN = -3, BDARY_LEFT = -4 (src/polygon.jl:4-7).
 *   - Host arrays are structure-of-arrays planes: a 2-vector field of n particles is
 *     double[2*n] with component k of particle i at [k*n + i]; the 2x2 tensor D is double[4*n] in
 *     Julia's column-major order (D11, D21, D12, D22).
 *   - Host buffers are copied during the call and never retained.  Device memory belongs to the
 *     context.  A context is not thread-safe; it owns one HIP stream of default priority.
 *   - Citations are file:line of OndrejKincl/LagrangianVoronoi.jl v0.1.0.
 */
#ifndef LVHIP_H
#define LVHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------------- */
#define LV_OK 0
#define LV_E_MESH_DESTROYED 1 /* throw("The Voronoi Mesh has been destroyed."), src/voronoigrid.jl:63-65 */
#define LV_E_NAN_VELOCITY 2   /* throw("Velocity field invalidated."), src/move.jl:24-26 */
#define LV_E_POLY_OVERFLOW 3  /* a polygon needs more than LV_MAX_EDGES edges */
#define LV_E_NOT_CONVERGED 4  /* pressure solve missed its tolerance (the reference ignores this) */
#define LV_E_HIP 5            /* HIP runtime error; lv_last_error() has the text */
#define LV_E_INVALID 6        /* invalid argument */
#define LV_E_NOMEM 7          /* device allocation failed */
#define LV_E_NO_MESH 8        /* an operation needs a mesh and remesh has not run since x changed */

#define LV_MAX_EDGES 16       /* edges a polygon may have (the reference's FastVector grows freely) */

/* ---- field ids for lv_upload / lv_download (PolygonNS fields, src/celldefs.jl:7-39) ------------ */
#define LV_FIELD_X 0         /* double[2n]  generating seed x */
#define LV_FIELD_V 1         /* double[2n]  velocity v */
#define LV_FIELD_DV 2        /* double[2n]  repair velocity dv */
#define LV_FIELD_MOMENTUM 3  /* double[2n] */
#define LV_FIELD_RHO 4       /* double[n] */
#define LV_FIELD_E 5         /* double[n]   specific energy */
#define LV_FIELD_P 6         /* double[n] */
#define LV_FIELD_C2 7        /* double[n] */
#define LV_FIELD_MASS 8      /* double[n] */
#define LV_FIELD_ENERGY 9    /* double[n] */
#define LV_FIELD_QUALITY 10  /* double[n] */
#define LV_FIELD_MU 11       /* double[n] */
#define LV_FIELD_D 12        /* double[4n] */
#define LV_FIELD_PHASE 13    /* int32[n] */
#define LV_FIELD_AREA 14     /* double[n]   area(p) of the current mesh (read-only) */
#define LV_FIELD_CENTROID 15 /* double[2n]  centroid(p) of the current mesh (read-only) */
#define LV_NFIELDS 16

/* ---- pressure solver methods ------------------------------------------------------------------ */
#define LV_SOLVER_PCG64 0  /* FP64 Jacobi-PCG */
#define LV_SOLVER_MIXED 1  /* FP32 Jacobi-PCG inside FP64 iterative refinement */
#define LV_SOLVER_SPAI 2   /* FP32 PCG with the learned SPAI preconditioner inside FP64 refinement;
                              falls back to LV_SOLVER_MIXED when it stalls */
#define LV_SOLVER_AMG 3    /* FP32 PCG with a multilevel (aggregation) V-cycle inside FP64 refinement; the
                              learned SPAI is its fine-level smoother when parameters are loaded (damped
                              Jacobi otherwise); falls back to LV_SOLVER_MIXED when it stalls */

typedef struct lv_ctx lv_ctx;

/* VoronoiGrid{T}(boundary_rect, dr; h = 2dr, r_max = 10dr, xperiodic, yperiodic),
 * src/voronoigrid.jl:26-49.  h <= 0 or r_max <= 0 select the reference defaults. */
typedef struct {
    double xmin[2], xmax[2]; /* boundary_rect corners */
    double dr;
    double h;
    double r_max;
    int32_t xperiodic, yperiodic;
    int32_t capacity;        /* most particles the context will hold */
} lv_grid_desc;

typedef struct {
    int32_t method;       /* LV_SOLVER_* */
    double tol;           /* stop at ||b - A P||_2 <= tol * ||b||_2 (FP64-evaluated) */
    int32_t itmax;        /* iteration cap per linear solve (inner iterations for MIXED/SPAI) */
    double inner_tol;     /* MIXED/SPAI: relative reduction asked of each FP32 inner solve */
    int32_t outer_max;    /* MIXED/SPAI: cap on refinement sweeps */
    double vbc[8];        /* boundary_velocity per label: vbc[2*(-label-1) + comp], label -1..-4 */
    int32_t picard_window; /* find_pressure!: 0 warm-starts each Picard solve from the previous P, as the
                              reference does; k > 0 starts it from the previous P plus the minimal-residual
                              combination of the last k Picard corrections (every solve still meets tol) */
    int32_t amg_layout;   /* AMG: 0 runs the inner solve and the cycle with the rows renumbered cell by cell
                              (16-bit column offsets, BF16 couplings in the cycle) when the offsets fit,
                              the slots past the sixth of each 64-row group read from a compact copy;
                              1 keeps the operator's own row order (int32 columns, FP32 couplings);
                              2 is 0 without the compact copies (also the one of the operator's own
                              ELL that the FP64 kernels read, and find_pressure!'s of the right-hand
                              side's edge geometry), 3 uses the aggregate order's only for
                              groups of at most 16 such slots (the same results bit for bit; for
                              measurement and tests).  Systems below 65536 rows use no compact copy. */
    int32_t picard_carry; /* with picard_window > 0: 1 also carries the first two Picard corrections
                             into the next find_pressure! on the same particles, where (multiplied by
                             the new operator) they improve the guesses of Picard solves 2 to 4;
                             every solve still meets tol */
    int32_t amg_level_kernels; /* AMG: 0 runs pairs of the coarse levels above the single-workgroup
                                  tail as one down and one up kernel each, 1 every such level as its
                                  own two kernels (the same results bit for bit; for measurement and
                                  tests) */
} lv_solve_opts;

typedef struct {
    int32_t iters[32];    /* Krylov iterations of each Picard solve */
    int32_t outer[32];    /* refinement sweeps of each Picard solve (MIXED/SPAI) */
    double relres[32];    /* final FP64 relative residual of each Picard solve */
    double relres0[32];   /* FP64 relative residual of each solve's initial guess */
    double solve_ms[32];  /* device time of each linear solve (HIP events on the context stream) */
    int32_t total_iters;
    int32_t fallbacks;    /* SPAI solves that fell back to Jacobi */
    int32_t nsolves;
    int32_t multiphase_iters;   /* lv_step_fused with multiphase: CG iterations (-1: itmax reached) */
    double multiphase_relres;   /* ... and its final |r| / |b| */
    int32_t agg_order;          /* 1: the FP32 inner solves ran in the multilevel preconditioner's aggregate order */
} lv_solve_stats;

/* One step!(sim, t) of tests/taylorgreen.jl:61-71 and its relatives (examples/cavity.jl:70-80,
 * examples/rayleightaylor.jl:90-101 without the multiphase projection):
 *   move! -> [gravity_step!] -> [stiffened_eos! | ideal_eos!] -> find_pressure! -> pressure_step!
 *   -> find_D! -> [viscous_step!] -> [bdary_friction!] -> find_dv! -> relaxation_step! */
typedef struct {
    double dt;
    int32_t eos;          /* 0 none, 1 stiffened_eos!(gamma, P0), 2 ideal_eos!(gamma; Pmin), 3 find_rho! only
                             (rho = m/area, src/pressure.jl:36-40: an incompressible flow's density kept
                             equal to its cells' mass per area, which examples/cavity.jl's step! omits) */
    double gamma, P0, Pmin;
    int32_t gravity;      /* 1: gravity_step!(g, dt) after move! */
    double g[2];
    int32_t niter;        /* Picard iterations of find_pressure! (reference default 10) */
    lv_solve_opts solve;
    int32_t viscous;      /* 1: viscous_step!(grid, dt; artificial_viscosity); 2: viscous_step!(ViscousSolver(grid), dt)
                             (implicit, src/viscous_solver.jl:120-128) solved with `solve`'s method to viscous_tol */
    int32_t artificial_viscosity;
    int32_t friction;     /* 1: bdary_friction!(vDirichlet, dt) with vDirichlet per label */
    double friction_vbc[8];  /* as lv_bdary_friction's table: NaN skips a wall (charfun false there) */
    double alpha;         /* find_dv!(dt, alpha) */
    int32_t rusanov;      /* relaxation_step!(dt; rusanov) */
    int32_t multiphase;   /* 1: multiphase_projection! after find_dv! (examples/rayleightaylor.jl:99) */
    double multiphase_quality; /* MultiphaseSolver(grid; quality_threshold), reference default 0.25 */
    double multiphase_tol;     /* atol = rtol of its solve (the reference's 1e-4) */
    int32_t multiphase_itmax;  /* its iteration cap (the reference's 200) */
    double viscous_tol;        /* viscous == 2: tolerance of the implicit viscous solve (the reference's 1e-6) */
} lv_step_desc;

/* ---- context ---------------------------------------------------------------------------------- */
int lv_create(const lv_grid_desc *desc, lv_ctx **out);  /* VoronoiGrid ctor, src/voronoigrid.jl:26-49 */
int lv_destroy(lv_ctx *ctx);
const char *lv_last_error(lv_ctx *ctx);
int lv_sync(lv_ctx *ctx);
/* grid geometry: n1, n2 (cell counts), npath (magic-path nodes within r_max) and
 * out6 = cropping xmin(2), cropping xmax(2), rr of the first node beyond r_max (-1 if none), rr_max */
int lv_grid_info(lv_ctx *ctx, int32_t *n1, int32_t *n2, int32_t *npath, double *out6);
/* magic_path prefix, src/neighborlist.jl:29-41: offsets and rr of the nodes with rr <= rr_max */
int lv_grid_path(lv_ctx *ctx, int32_t *i1, int32_t *i2, double *rr);

/* ---- particle state --------------------------------------------------------------------------- */
/* push!(grid.polygons, T(x = x)) for n seeds at once (src/populate.jl pattern): sets the particle
 * count, uploads x, zeroes every other field (PolygonNS defaults, src/celldefs.jl:7-27) and marks
 * the mesh stale. */
int lv_set_particles(lv_ctx *ctx, int32_t n, const double *x);
int lv_num_particles(lv_ctx *ctx, int32_t *n);
int lv_upload(lv_ctx *ctx, int32_t field, const void *host);
int lv_download(lv_ctx *ctx, int32_t field, void *host);
/* device pointer of a field (for zero-copy interop with other device code on the same stream) */
int lv_field_device_ptr(lv_ctx *ctx, int32_t field, void **dptr);

/* ---- mesh ------------------------------------------------------------------------------------- */
int lv_remesh(lv_ctx *ctx);                                  /* remesh!, src/voronoigrid.jl:89-108 */
int lv_mesh_size(lv_ctx *ctx, int64_t *nedges);
/* edges of every polygon after sort_edges! (src/IO.jl:35-48), CSR by particle:
 * edge_ptr[n+1], labels[nedges], verts[4*nedges] = (v1x, v1y, v2x, v2y) per edge */
int lv_get_mesh(lv_ctx *ctx, int32_t *edge_ptr, int32_t *labels, double *verts);
/* the cell list of the last remesh (CellList cells, src/neighborlist.jl:11-67): cell c (0-based,
 * c = (i1-1) + (i2-1)*n1) holds cell_labels[cell_start[c] .. cell_start[c+1]) in ascending order */
int lv_cell_list_size(lv_ctx *ctx, int64_t *nitems);
int lv_get_cell_list(lv_ctx *ctx, int32_t *cell_start, int32_t *cell_labels);

/* ---- physics operators (each is the reference function of the same name) ---------------------- */
int lv_move(lv_ctx *ctx, double dt);                          /* move!, src/move.jl:9-21 (remeshes) */
int lv_gravity_step(lv_ctx *ctx, const double g[2], double dt);  /* src/pressure.jl:77-82 */
int lv_find_rho(lv_ctx *ctx);                                 /* src/pressure.jl:36-40 */
int lv_eos_stiffened(lv_ctx *ctx, double gamma, double P0);   /* src/pressure.jl:64-70 */
int lv_eos_ideal(lv_ctx *ctx, double gamma, double Pmin);     /* src/pressure.jl:49-55 */
/* find_pressure!(solver, dt, niter; boundary_velocity), src/pressure.jl:215-224 */
int lv_find_pressure(lv_ctx *ctx, double dt, int32_t niter, const lv_solve_opts *opts, lv_solve_stats *stats);
int lv_pressure_step(lv_ctx *ctx, double dt);                 /* src/pressure.jl:10-25 */
int lv_find_D(lv_ctx *ctx);                                   /* src/diffusion.jl:8-19 */
int lv_viscous_step(lv_ctx *ctx, double dt, int32_t artificial_viscosity); /* src/diffusion.jl:39-53 */
/* bdary_friction!(grid, vDirichlet, dt; charfun), src/diffusion.jl:64-78: vDirichlet as a per-label table
 * vbc[2*(-label-1) + comp], evaluated once per wall, not per edge midpoint; a NaN entry makes charfun
 * false on that wall (its edges are skipped) */
int lv_bdary_friction(lv_ctx *ctx, const double vbc[8], double dt);
int lv_find_dv(lv_ctx *ctx, double dt, double alpha);         /* src/relaxation.jl:10-25 */
int lv_relaxation_step(lv_ctx *ctx, double dt, int32_t rusanov); /* src/relaxation.jl:36-73 (remeshes) */
/* the whole step!, device resident; stats may be NULL */
int lv_step_fused(lv_ctx *ctx, const lv_step_desc *desc, lv_solve_stats *stats);

/* viscous_step!(solver::ViscousSolver, dt), src/viscous_solver.jl:120-128: assembles the implicit
 * operator (d_i = m_i/dt, c_ij = 0.5 lrr (mu_i + mu_j), :89-113) into the solver's system and solves
 * it once per velocity component from the initial guess v; v is overwritten.  opts == NULL: mixed
 * precision PCG to 1e-6 with itmax 2000 (the reference's minres! tolerances).  Afterwards the solver's
 * system holds this operator: lv_get_pressure_operator and lv_solve_pressure refuse it until the
 * pressure operator is assembled again. */
int lv_viscous_implicit(lv_ctx *ctx, double dt, const lv_solve_opts *opts, lv_solve_stats *stats);

/* multiphase_projection!(MultiphaseSolver(grid; quality_threshold)), src/relaxation.jl:190-206:
 * projects dv so that no area crosses a phase boundary.  The reference's minres!(atol = rtol = 1e-4,
 * itmax = 200) is CG with the same stopping rule here; iters = -1 when itmax stopped it (the update
 * is applied all the same, as the reference does after its warning). */
int lv_multiphase_projection(lv_ctx *ctx, double quality_threshold, double atol, double rtol, int32_t itmax,
                             int32_t *iters, double *relres);
/* refresh!(::MultiphaseSolver) (src/relaxation.jl:165-180) and mul!(res, ::MultiphaseProjector, x)
 * (:92-125) on host vectors of n doubles (zero rows off the interface) */
int lv_multiphase_rhs(lv_ctx *ctx, double *b);
int lv_multiphase_apply(lv_ctx *ctx, const double *x, double *y);

/* ---- pressure system access (PressureOperator/PressureSolver, src/pressure.jl:89-203) ---------- */
/* refresh!(A, grid, dt), src/pressure.jl:104-117, into device CSR; returns nnz of the off-diagonal */
int lv_assemble_pressure(lv_ctx *ctx, double dt, int64_t *nnz);
/* the assembled operator: diag[n], row_ptr[n+1], cols[nnz], coef[nnz] (y_i = d_i x_i + sum c_ij (x_i - x_j));
 * cols and coef may be NULL (nnz = row_ptr[n]) */
int lv_get_pressure_operator(lv_ctx *ctx, double *diag, int32_t *row_ptr, int32_t *cols, double *coef);
/* refresh!(solver, dt, gp_step, boundary_velocity), src/pressure.jl:162-203: b[n] and GP[2n] */
int lv_pressure_rhs(lv_ctx *ctx, double dt, int32_t gp_step, const double vbc[8], double *b, double *GP);
/* solve the assembled operator for b with initial guess x (in/out) */
int lv_solve_pressure(lv_ctx *ctx, const double *b, double *x, const lv_solve_opts *opts, lv_solve_stats *stats);
/* z = B r with the FP32 preconditioner of the inner solve of `method` (LV_SOLVER_MIXED: Jacobi,
 * LV_SOLVER_SPAI: the learned M, LV_SOLVER_AMG: the multilevel V-cycle) on the assembled operator;
 * host vectors of n doubles.  No counterpart in the reference; it is the hook a Julia caller uses
 * to hand this preconditioner to Krylov.jl (src/pressure.jl:219 runs minres! unpreconditioned).
 * The multilevel cycle keeps the amg_layout and amg_level_kernels of the last solve on this context
 * (0 before any). */
int lv_apply_preconditioner(lv_ctx *ctx, int32_t method, const double *r, double *z);

/* ---- initialisation and interpolation ------------------------------------------------------- */
/* the iterations of populate_lloyd! (src/populate.jl:132-145) on the current seeds: niterations
 * times remesh! then x = centroid(p) for every polygon, then remesh! */
int lv_lloyd(lv_ctx *ctx, int32_t niterations);
/* nearest_polygon(grid, x), src/voronoigrid.jl:154-178, for m query points q (planes q1..., q2...)
 * over the cell list of the last remesh: idx[t] = 0-based polygon index.  LV_E_INVALID when the
 * search would leave the magic-path prefix (a query farther than r_max from every seed). */
int lv_nearest_polygon(lv_ctx *ctx, int32_t m, const double *q, int32_t *idx);
/* movingls(T, grid, p, p -> field[comp]; h, gaussian_kernel), src/movingls.jl:41-96, for every
 * polygon: order 1, 2, 3 = LinearExpansion, QuadraticExpansion, CubicExpansion (2, 5, 9
 * coefficients per polygon, row-major in coeffs).  h <= 0 selects grid.h; h must not exceed r_max. */
int lv_movingls(lv_ctx *ctx, int32_t order, int32_t field, int32_t comp, double h, double *coeffs);
/* point_value(grid, x, p -> field[comp]), src/movingls.jl:140-145, at m query points */
int lv_point_values(lv_ctx *ctx, int32_t m, const double *q, int32_t field, int32_t comp, double *out);
/* lv_point_values without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops of probe queries, allocation-free harnesses).
 * The queries and values of one call are staged in a static per-device array of 1.5 MiB, so m is at
 * most LV_NOALLOC_MAX_QUERIES (LV_E_INVALID beyond).  Calls that overlap on one GPU (contexts on
 * several threads) race on that array; results are identical to lv_point_values. */
#define LV_NOALLOC_MAX_QUERIES 65536
int lv_point_values_noalloc(lv_ctx *ctx, int32_t m, const double *q, int32_t field, int32_t comp, double *out);

/* ---- output (src/IO.jl:9-71) ------------------------------------------------------------------ */
/* export_grid(grid, filename, vars...), src/IO.jl:9-34: the polygons as VTK XML PolyData (.vtp is
 * appended to filename unless it ends so), each polygon's vertices the v1 of its edges in their
 * chained order, with one cell-data column per field id in `fields` (LV_FIELD_X..LV_FIELD_PHASE):
 * Float64 for numbers, (x, y, 0) for vectors and for D its first two column-major entries, as
 * append_datasets! (src/IO.jl:61-71) does.  compress != 0 zlib-compresses every array (WriteVTK's
 * default); an unknown field is LV_E_INVALID ("Cannot export variable ... because it does not
 * exist.", src/IO.jl:65). */
int lv_export_grid(lv_ctx *ctx, const char *filename, const int32_t *fields, int32_t nfields, int32_t compress);
/* export_points(grid, filename, vars...), src/IO.jl:50-59: the seeds as a cloud of vertices */
int lv_export_points(lv_ctx *ctx, const char *filename, const int32_t *fields, int32_t nfields, int32_t compress);

/* ---- learned SPAI preconditioner --------------------------------------------------------------
 * No counterpart in the reference (it runs unpreconditioned MINRES, src/pressure.jl:219).  A
 * message-passing network over the Voronoi edges emits the coefficients of a symmetric, strictly
 * diagonally dominant (hence SPD) approximate inverse on the pressure operator's own pattern; it is
 * used by LV_SOLVER_SPAI, with Jacobi-PCG as fallback.  Parameters are a flat float32 vector. */
int lv_spai_num_params(int64_t *count);
int lv_spai_set_params(lv_ctx *ctx, const float *params);  /* params == NULL unloads the network */
int lv_spai_get_params(lv_ctx *ctx, float *params);
/* one Adam step on the assembled operator; probes: the current P (scaled) and nprobe-1 Rademacher
 * vectors drawn from seed; loss is the value before the update */
int lv_spai_train_step(lv_ctx *ctx, double lr, int32_t nprobe, uint64_t seed, double *loss);
/* loss and gradient of every parameter at the current parameters (grad may be NULL) */
int lv_spai_loss_grad(lv_ctx *ctx, int32_t nprobe, uint64_t seed, int32_t use_data, double *loss, float *grad);
/* network outputs on the assembled operator: o_edge[nnz] (CSR order), o_node[n], and the scaled
 * preconditioner's off-diagonal w[nnz] and diagonal mdiag[n] (any pointer may be NULL) */
int lv_spai_outputs(lv_ctx *ctx, float *o_edge, float *o_node, float *w, float *mdiag);

/* ---- timing ----------------------------------------------------------------------------------- */
/* When enabled (on != 0), the library brackets launches of the named kernel with HIP events on
 * the context stream and accumulates their device time: every launch for on == 1, every on-th
 * launch for on > 1 (each event pair adds a few microseconds between the launches it brackets);
 * launches counts the bracketed ones. kernel ids: 0 = FP32 SpMV, 1 = FP64 SpMV, 2 = clipping,
 * 3 = cell-list scatter.  The algorithmic bytes of kernel 0 follow the layout of the
 * last solve (the operator's own ELL, or the aggregate order's 8-bit degrees and 16-bit offsets). */
int lv_timing_enable(lv_ctx *ctx, int32_t kernel_id, int32_t on);
int lv_timing_read(lv_ctx *ctx, int32_t kernel_id, double *total_ms, int64_t *launches, double *bytes_per_launch);
/* launches of the kernel since lv_timing_enable, bracketed or not */
int lv_timing_seen(lv_ctx *ctx, int32_t kernel_id, int64_t *seen);

#ifdef __cplusplus
}
#endif
#endif
