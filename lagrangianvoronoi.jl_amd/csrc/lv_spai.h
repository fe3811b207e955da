// lv_spai.h -- learned SPAI preconditioner: parameter layout, workspace, host entry points.
#pragma once

#include <algorithm>

#include "lv_solver.h"

#define LV_SPAI_WIDTH 64
// flat FP32 parameter vector: node encoder Wn[4][64] bn[64], edge encoder We[4][64] be[64], three
// message-passing layers W_l[192][64] b_l[64], edge decoder ve[64] ce, node decoder un[64] cn
#define LV_SPAI_OFF_WN 0
#define LV_SPAI_OFF_BN 256
#define LV_SPAI_OFF_WE 320
#define LV_SPAI_OFF_BE 576
#define LV_SPAI_OFF_W(l) (640 + (l) * (192 * 64 + 64))
#define LV_SPAI_OFF_B(l) (LV_SPAI_OFF_W(l) + 192 * 64)
#define LV_SPAI_OFF_VE (640 + 3 * (192 * 64 + 64))
#define LV_SPAI_OFF_CE (LV_SPAI_OFF_VE + 64)
#define LV_SPAI_OFF_UN (LV_SPAI_OFF_CE + 1)
#define LV_SPAI_OFF_CN (LV_SPAI_OFF_UN + 64)
#define LV_SPAI_NPARAM (LV_SPAI_OFF_CN + 1)

struct lv_spai {
    // parameters and optimiser state (FP32), BF16 copies of the layer weights for the MFMA
    float *theta, *grad, *adam_m, *adam_v;
    __bf16 *Wt, *Wk;
    double *scal;
    int adam_t, has_params;
    int n, E;
    // graph, features, activations of the last forward pass
    int *src, *rev;
    float *a, *xf, *ff, *rfl, *sqD;
    float *Hf;
    __bf16 *Hb, *Eb;
    float *oe, *on, *w, *mdiag, *delta;
    // training workspace
    float *v, *u, *y, *gy, *G, *doe, *don, *dE, *dH, *dX;
    __bf16 *dZb;
    float *part;
    double *lpart;
    // the preconditioner in the solver's ELL slots
    float *mval, *mdg;
    size_t cap_src, cap_rev, cap_a, cap_xf, cap_ff, cap_rfl, cap_sqD, cap_Hf, cap_Hb, cap_Eb, cap_oe, cap_on, cap_w,
        cap_mdiag, cap_delta, cap_v, cap_u, cap_y, cap_gy, cap_G, cap_doe, cap_don, cap_dE, cap_dH, cap_dX, cap_dZb,
        cap_part, cap_lpart, cap_mval, cap_mdg;
};

int64_t lv_spai_param_count();
hipError_t lv_spai_init(lv_spai *S);
void lv_spai_free(lv_spai *S);
hipError_t lv_spai_set_params(lv_spai *S, const float *host, hipStream_t s);
hipError_t lv_spai_get_params(lv_spai *S, float *host, hipStream_t s);
hipError_t lv_spai_forward(lv_spai *S, const lv_pressure_system &A, int n, hipStream_t s);
hipError_t lv_spai_infer(lv_spai *S, const lv_pressure_system &A, int n, hipStream_t s);
hipError_t lv_spai_emit(lv_spai *S, const lv_pressure_system &A, int n, hipStream_t s);
hipError_t lv_spai_loss_grad(lv_spai *S, const lv_pressure_system &A, int n, const double *x, int P, uint64_t seed,
                             bool want_grad, double *loss_host, hipStream_t s);
hipError_t lv_spai_adam(lv_spai *S, double lr, hipStream_t s);
