// lv_amg.h -- multilevel (aggregation) preconditioner for the pressure system.
#pragma once

#include "lv_solver.h"

#define LV_AMG_MAXLEV 16
#define LV_AMG_DENSE_MAX 64  // the coarsest level is inverted densely once it has at most this many cells

// Level l >= 1: a Kx x Ky grid of cells; the operator in the reference's difference form,
// (A y)_I = d_I y_I + sum_k c_Ik (y_I - y_{I + off_k}), over the 8 neighbour directions.
struct lv_amg_level {
    int Kx, Ky, n;
    float *d, *c, *dinv;  // d[n], c[8n] slot-major, dinv[n] = 1 / (d + sum c)
    float *r, *z, *t;     // right-hand side, pre-smoothed correction, final correction of the cycle
    float *acc;           // the first correction of a W-cycle visit
    float alpha;          // factor on the correction prolongated into this level from the next
    size_t cap;
};

#define LV_AMG_MAXPROG 1024  // operations of the single-workgroup program for the smallest levels

struct lv_amg {
    int nlev;       // coarse levels; level nlev is the dense one
    int n0, na;     // fine rows, level-1 cells
    int xper, yper;
    int *agg, *cnt, *start, *cursor, *list;
    void *scan_tmp;
    size_t cap_agg, cap_list, cap_t0, cap_z0, cap_cnt, cap_start, cap_cursor, cap_scan;
    float *t0, *z0; // fine residual and pre-smoothed correction
    lv_amg_level lev[LV_AMG_MAXLEV + 1];
    float *dense_inv;    // nc * nc, column-major
    float *tail_img;     // the tail levels' constant arrays in their LDS layout (packed per set-up)
    double *pw;          // power-iteration partials and norms
    float *pv, *pu;      // power-iteration vectors (n0)
    size_t cap_pv, cap_pu;
    int ltail;           // first level of the single-workgroup kernel (levels ltail..nlev)
    int pair[LV_AMG_MAXLEV + 1];  // pair[l]: levels l and l+1 (both above ltail) run as one down and
                                  // one up kernel
    int level_kernels;   // 1: every level above ltail runs as its own kernels (pairs not used)
    int wcycle;          // 1: the levels of that kernel are visited by a W-cycle
    int *prog;           // its operations (device), nprog of them; hprog: the host copy
    int nprog;
    int hprog[LV_AMG_MAXPROG];
    // the fine level in the aggregate order (perm_ok): row p is particle list[p]; the inner FP32
    // solve and the cycle then run on this copy of the operator
    int perm_ok;
    int *iperm, *aggp;   // particle -> row of the aggregate order; level-1 cell of each row
    uint8_t *pdeg;       // degrees
    int16_t *poff;       // column offsets j - i (mod n), slot-major
    float *pev32;        // couplings, FP32 (Krylov operator, power iteration)
    uint16_t *pev16;     // couplings, BF16 (the cycle's smoothing and residuals)
    float *pdiag32, *pdinv32, *pmval, *pmdg;
    int16_t *ooff;       // the compact overflow of slots LV_OVF_BASE.. per 64-row group (lv_ell.h),
    float *oev32;        // read by the Krylov operator: offsets and FP32 couplings, LV_OVF_CAP per group
    int ovf_cap;         // groups with at most this many overflow slots use it (0: none)
    size_t cap_ooff, cap_oev32;
    int *pflag;          // [0] set when a column offset does not fit 16 bits, [1] when a level-1 cell
                         // is too large for the fused down pass
    int *bfirst;         // fused down pass: (first level-1 cell, first row) of each block (nbfuse + 1 pairs), then (0, 0)
    size_t cap_bfirst;
    int nbfuse;          // its blocks; 0: the down pass and the restriction run as two kernels
    int z0_in;           // 1: the fused down pass can take z0 = fine_scale D^-1 r from the caller
                         // (stored row-locally by the Krylov kernel that updates r)
    size_t cap_iperm, cap_aggp, cap_pdeg, cap_poff, cap_pev32, cap_pev16, cap_pdiag32, cap_pdinv32, cap_pmval,
        cap_pmdg;
    float fine_alpha;    // factor on the correction prolongated into the fine level
    float fine_scale;    // the fine smoother applies fine_scale * M (learned) or fine_scale * D^-1 (Jacobi)
    int fine_learned;    // 1: the learned preconditioner's M is the fine smoother
    const float *mval, *mdg;
};

hipError_t lv_amg_init(lv_amg *G);
void lv_amg_free(lv_amg *G);
// Builds the hierarchy for the assembled operator A (device positions x of the particles, grid of
// the boundary rectangle).  mval/mdg: the learned preconditioner in A's ELL slots, or null for a
// damped-Jacobi fine smoother.  layout 0: the aggregate order when its column offsets fit 16 bits
// (perm_ok), else the operator's own; 1: the operator's own order.
hipError_t lv_amg_setup(lv_amg *G, const lv_grid_params &g, const double *x, const lv_pressure_system &A, int n,
                        const float *mval, const float *mdg, int layout, hipStream_t s);
// z = B r for the symmetric V(1,1) cycle B; writes the (r.z, r.r) partials of lv_nblocks(n) blocks
// that the next CG kernel reduces.  Every kernel is a no-op once S->done is set.  With perm_ok, r
// and z are in the aggregate order.  z0_ready (only when G->z0_in): G->z0 already holds
// fine_scale D^-1 r, bit for bit as the cycle would compute it.
hipError_t lv_amg_vcycle(const lv_amg *G, const lv_pressure_system &A, const float *r, float *z, double *Prr,
                         const lv_cg_scalars *S, hipStream_t s, bool z0_ready = false);
