// lv_api.hip -- the C ABI of liblvhip.so: context, particle state, remesh and the step operators.
//
// A context is one VoronoiGrid (src/voronoigrid.jl:14-50) whose particles, mesh, cell list and
// pressure system live in device memory as SoA planes.  Every call enqueues on the context's own
// stream; calls that can raise one of the reference's errors synchronise once to read the device
// error word and map it to its status code.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "lv_spai.h"
#include "lv_amg.h"
#include "lv_vtk.h"

struct lv_mph_work;
hipError_t lv_mph_alloc(lv_mph_work **out, int cap);
void lv_mph_free(lv_mph_work *w);
hipError_t lv_mph_projection(lv_mph_work *w, const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n,
                             double quality_threshold, double atol, double rtol, int itmax, int *iters, double *relres,
                             hipStream_t s);
hipError_t lv_mph_apply(lv_mph_work *w, const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, const double *x,
                        double *y, hipStream_t s);
hipError_t lv_mph_rhs(lv_mph_work *w, const lv_grid_params &g, lv_fields F, lv_mesh_view M, int n, double *b,
                      hipStream_t s);

struct lv_ctx {
    lv_grid_params g;
    int cap = 0, n = 0;
    hipStream_t s = nullptr;
    lv_fields F{};
    // mesh (ELL, slot-major, leading dimension n)
    int *ne = nullptr;
    double *ev = nullptr;
    int *el = nullptr;
    int *clip_retry = nullptr;  // [0] count, [1..] the seeds whose polygon outgrew the small LDS capacity
    bool mesh_valid = false;   // the mesh belongs to the current positions
    // cell list
    int *cell_start = nullptr, *cell_labels = nullptr;
    double *cell_x = nullptr;
    int64_t max_items = 0;
    int nitems = 0;
    void *cl_temp = nullptr;
    int *d_nitems = nullptr;
    // magic path prefix
    int *path_i = nullptr;
    double *path_rr = nullptr;
    std::vector<int> h_path_i;
    std::vector<double> h_path_rr;
    // pressure system
    lv_pressure_system A{};
    lv_solver_work W{};
    double *b = nullptr, *P0 = nullptr, *GP = nullptr;
    lv_rhs_geom rg{};         // the P-independent parts of refresh!(solver) (allocated on first use)
    bool rg_alloc = false;
    int *deg = nullptr;
    void *scan_temp = nullptr;
    int *d_flags = nullptr;   // [0] error word, [1] nonsingular flag, [2] a column offset wider than 16 bits
    lv_timer *timers[LV_NKTIMERS] = {nullptr, nullptr, nullptr, nullptr};
    // Picard extrapolation: orthonormal A-images q_k of past Picard corrections and their
    // preimages z_k (A z_k = q_k), the last two Picard iterates and their A-images
    float *pw_q[8] = {nullptr}, *pw_z[8] = {nullptr};  // orthonormal directions, FP32
    double *pw_P[2] = {nullptr, nullptr}, *pw_AP[2] = {nullptr, nullptr};
    double *pw_t1 = nullptr, *pw_t2 = nullptr, *pw_dots = nullptr;
    int pw_cap = 0;
    // carried Picard directions (lv_solve_opts.picard_carry): the previous call's first two
    // corrections (pc_z, pc_cnt of them, valid for pc_n particles), the ones this call is recording
    // (pc_new), and this call's orthonormal set made of the carried ones (pc_q = A pc_o)
    double *pc_z[2] = {nullptr}, *pc_new[2] = {nullptr};
    float *pc_q[2] = {nullptr}, *pc_o[2] = {nullptr};
    int pc_cnt = 0, pc_n = -1;
    // learned preconditioner: network state and the preconditioner of the last assembled operator
    lv_spai spai{};
    int spai_ok = 0;          // lv_spai_init succeeded
    int64_t assembly_id = 0;  // bumped by every assembly
    int64_t spai_emitted_for = -1;
    // multilevel preconditioner of the last assembled operator
    lv_amg amg{};
    int amg_ok = 0;
    int64_t amg_built_for = -1;
    int amg_learned = -1;  // the fine smoother the hierarchy was built with (1 learned, 0 Jacobi)
    int amg_layout = -1;   // the fine-level layout it was asked for (lv_solve_opts.amg_layout)
    int spmv32_agg = 0;    // the last FP32 inner solve ran in the aggregate order
    int A_kind = 0;        // what A holds: 0 the pressure operator, 1 the implicit viscous operator
    // multiphase projection work space (allocated on first use)
    lv_mph_work *mph = nullptr;
    // solve timing: event pairs recorded around each solve and read when the API call returns its
    // statistics (one wait per call instead of one per solve); out: where each pair's time goes
    static constexpr int NSEV = 32;
    hipEvent_t sev[2 * NSEV] = {};
    double *sev_out[NSEV] = {};
    int sev_ready = 0, sev_n = 0;
    // the device the context was created on, and that device's address of the allocation-free
    // entry point's staging array
    int device = 0;
    double *scratch = nullptr;
    char msg[512] = {0};
};

int lv_fail(lv_ctx *ctx, int code, const char *msg)
{
    if (ctx) snprintf(ctx->msg, sizeof(ctx->msg), "%s", msg);
    return code;
}

int lv_fail_hip(lv_ctx *ctx, hipError_t e, const char *what, const char *file, int line)
{
    if (ctx) snprintf(ctx->msg, sizeof(ctx->msg), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
                      what, file, line);
    return (e == hipErrorOutOfMemory) ? LV_E_NOMEM : LV_E_HIP;
}

// ---- kernel timers ---------------------------------------------------------------------------------

// An event pair costs a few microseconds of GPU time between the launches it brackets, so a timer
// may sample: it brackets every `every`-th launch and averages over those.
void lv_timer_begin(lv_timer *t, hipStream_t s)
{
    t->pending = (t->seen++ % (t->every > 0 ? t->every : 1)) == 0;
    if (!t->pending) return;
    if (t->used == LV_TPOOL) lv_timer_flush(t);
    hipEventRecord(t->ev[2 * t->used], s);
}

void lv_timer_end(lv_timer *t, hipStream_t s)
{
    if (!t->pending) return;
    hipEventRecord(t->ev[2 * t->used + 1], s);
    t->used++;
    t->pending = 0;
}

void lv_timer_flush(lv_timer *t)
{
    for (int k = 0; k < t->used; k++) {
        float ms = 0.0f;
        hipEventSynchronize(t->ev[2 * k + 1]);
        if (hipEventElapsedTime(&ms, t->ev[2 * k], t->ev[2 * k + 1]) == hipSuccess) t->total_ms += ms;
        t->launches++;
    }
    t->used = 0;
}

static lv_timer *timer_if_on(lv_ctx *ctx, int id) { return (ctx->timers[id] && ctx->timers[id]->on) ? ctx->timers[id] : nullptr; }

// ---- set-up ---------------------------------------------------------------------------------------

namespace {

struct pnode {
    int i1, i2;
    long long s, gen;
    double rr;
};

// magic_path, src/neighborlist.jl:29-41: nodes sorted by rr = h^2*(max(0,|i1|-1)^2 + max(0,|i2|-1)^2),
// ties by i1^2 + i2^2, then by generation order (i2 outer over (1-n1):(n1-1), i1 inner over
// (1-n2):(n2-1)), both sorts stable.  Only the nodes with rr <= rr_max and the first node past it
// decide anything, so only small offsets are enumerated.
void build_path(lv_grid_params &g, std::vector<int> &pi, std::vector<double> &prr)
{
    const int lim1 = g.n2 - 1, lim2 = g.n1 - 1;
    const int kmax = (int)ceil(g.r_max / g.h) + 3;
    const int m1 = std::min(kmax, lim1), m2 = std::min(kmax, lim2);
    std::vector<pnode> nodes;
    nodes.reserve((size_t)(2 * m1 + 1) * (2 * m2 + 1));
    for (int i2 = -m2; i2 <= m2; i2++)
        for (int i1 = -m1; i1 <= m1; i1++) {
            long long a = std::max(0, std::abs(i1) - 1), b = std::max(0, std::abs(i2) - 1);
            pnode p;
            p.i1 = i1;
            p.i2 = i2;
            p.s = (long long)i1 * i1 + (long long)i2 * i2;
            p.gen = (long long)(i2 - (1 - g.n1)) * (2LL * g.n2 - 1) + (i1 - (1 - g.n2));
            p.rr = (g.h * g.h) * (double)(a * a + b * b);
            nodes.push_back(p);
        }
    std::stable_sort(nodes.begin(), nodes.end(), [](const pnode &p, const pnode &q) {
        if (p.rr != q.rr) return p.rr < q.rr;
        if (p.s != q.s) return p.s < q.s;
        return p.gen < q.gen;
    });
    size_t np = 0;
    while (np < nodes.size() && !(nodes[np].rr > g.rr_max)) np++;
    pi.clear();
    prr.clear();
    for (size_t k = 0; k < np; k++) {
        pi.push_back(nodes[k].i1);
        pi.push_back(nodes[k].i2);
        prr.push_back(nodes[k].rr);
    }
    g.npath = (int)np;
    g.has_next = (np < nodes.size()) ? 1 : 0;
    g.next_rr = g.has_next ? nodes[np].rr : INFINITY;
}

template <typename T>
hipError_t dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void **)p, sizeof(T) * count);
}

void dfree(void *p)
{
    if (p) hipFree(p);
}

double *field_ptr(lv_ctx *c, int f, size_t *count, size_t *elem, bool *readonly)
{
    const size_t n = (size_t)c->n;
    *elem = sizeof(double);
    *readonly = false;
    switch (f) {
    case LV_FIELD_X: *count = 2 * n; return c->F.x;
    case LV_FIELD_V: *count = 2 * n; return c->F.v;
    case LV_FIELD_DV: *count = 2 * n; return c->F.dv;
    case LV_FIELD_MOMENTUM: *count = 2 * n; return c->F.mom;
    case LV_FIELD_RHO: *count = n; return c->F.rho;
    case LV_FIELD_E: *count = n; return c->F.e;
    case LV_FIELD_P: *count = n; return c->F.P;
    case LV_FIELD_C2: *count = n; return c->F.c2;
    case LV_FIELD_MASS: *count = n; return c->F.mass;
    case LV_FIELD_ENERGY: *count = n; return c->F.energy;
    case LV_FIELD_QUALITY: *count = n; return c->F.quality;
    case LV_FIELD_MU: *count = n; return c->F.mu;
    case LV_FIELD_D: *count = 4 * n; return c->F.D;
    case LV_FIELD_PHASE: *count = n; *elem = sizeof(int); return (double *)c->F.phase;
    case LV_FIELD_AREA: *count = n; *readonly = true; return c->F.area;
    case LV_FIELD_CENTROID: *count = 2 * n; *readonly = true; return c->F.cent;
    default: return nullptr;
    }
}

lv_mesh_view mesh_view(lv_ctx *c) { return lv_mesh_view{c->ne, c->ev, c->el, c->n}; }

int read_flags(lv_ctx *ctx, int *flags2)
{
    LV_HIP_CHECK(hipMemcpyAsync(flags2, ctx->d_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int clear_flags(lv_ctx *ctx)
{
    LV_HIP_CHECK(hipMemsetAsync(ctx->d_flags, 0, 3 * sizeof(int), ctx->s));
    return LV_OK;
}

int check_ctx(lv_ctx *ctx, bool need_mesh)
{
    if (!ctx) return LV_E_INVALID;
    if (ctx->n <= 0) return lv_fail(ctx, LV_E_INVALID, "no particles");
    if (need_mesh && !ctx->mesh_valid) return lv_fail(ctx, LV_E_NO_MESH, "positions changed since the last remesh");
    return LV_OK;
}

const char *error_text(int code)
{
    switch (code) {
    case LV_E_MESH_DESTROYED: return "The Voronoi Mesh has been destroyed.";
    case LV_E_NAN_VELOCITY: return "Velocity field invalidated.";
    case LV_E_POLY_OVERFLOW: return "A Voronoi polygon exceeded LV_MAX_EDGES edges.";
    default: return "error";
    }
}

}  // namespace

// ---- staging of the allocation-free entry point ------------------------------------------------------

// lv_point_values_noalloc stages one call's queries (2 m doubles) and values (m doubles) here, for at
// most LV_NOALLOC_MAX_QUERIES queries (1.5 MiB).  Each device the library runs on holds its own copy
// until the process ends, and every call on a device shares it.  In host code the name is only a
// stand-in for the symbol: kernels and copies take the device address that hipGetSymbolAddress
// returns for the current device (scratch_address).
static __device__ __align__(256) double scratch[3 * (size_t)LV_NOALLOC_MAX_QUERIES];

namespace {

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_addr;  // per device id; null until resolved

// The device address of `scratch` on `device`, which must be the current device; resolved once per
// device (under a lock, since contexts may be created from several threads).
hipError_t scratch_address(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_addr[device] = (double *)p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

}  // namespace

extern "C" {

int lv_create(const lv_grid_desc *d, lv_ctx **out)
{
    if (!d || !out) return LV_E_INVALID;
    *out = nullptr;
    const double h = (d->h > 0.0) ? d->h : 2.0 * d->dr;            // h = 2dr, src/voronoigrid.jl:28
    const double r_max = (d->r_max > 0.0) ? d->r_max : 10.0 * d->dr;  // r_max = 10dr
    if (!(h > 0.0) || !(d->dr > 0.0) || d->capacity <= 0) return LV_E_INVALID;
    if (!(d->xmax[0] > d->xmin[0]) || !(d->xmax[1] > d->xmin[1])) return LV_E_INVALID;
    lv_ctx *ctx = new lv_ctx();
    lv_grid_params &g = ctx->g;
    memset(&g, 0, sizeof(g));
    g.bxmin0 = d->xmin[0];
    g.bxmin1 = d->xmin[1];
    g.bxmax0 = d->xmax[0];
    g.bxmax1 = d->xmax[1];
    g.xperiodic = d->xperiodic ? 1 : 0;
    g.yperiodic = d->yperiodic ? 1 : 0;
    // cropping rect, src/voronoigrid.jl:29-30 (r_max*periodic is r_max or +0.0)
    const double rx = g.xperiodic ? r_max : 0.0, ry = g.yperiodic ? r_max : 0.0;
    g.cmin0 = (g.bxmin0 - rx * 1.0) - ry * 0.0;
    g.cmin1 = (g.bxmin1 - rx * 0.0) - ry * 1.0;
    g.cmax0 = (g.bxmax0 + rx * 1.0) + ry * 0.0;
    g.cmax1 = (g.bxmax1 + rx * 0.0) + ry * 1.0;
    g.xperiod = g.bxmax0 - g.bxmin0;
    g.yperiod = g.bxmax1 - g.bxmin1;
    g.h = h;
    g.dr = d->dr;
    g.r_max = r_max;
    g.rr_max = r_max * r_max;
    // CellList, src/neighborlist.jl:23-25
    g.origin0 = g.cmin0 - h;
    g.origin1 = g.cmin1 - h;
    g.n1 = (int)floor((g.cmax0 - g.cmin0) / h) + 3;
    g.n2 = (int)floor((g.cmax1 - g.cmin1) / h) + 3;
    build_path(g, ctx->h_path_i, ctx->h_path_rr);
    if (g.npath > LV_MAXPATH) {
        delete ctx;
        return LV_E_INVALID;
    }
    ctx->cap = d->capacity;
    const size_t cap = (size_t)d->capacity;
    const int64_t ncells = (int64_t)g.n1 * g.n2;
    hipError_t e = hipGetDevice(&ctx->device);
    if (e == hipSuccess) e = scratch_address(ctx->device, &ctx->scratch);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->s, hipStreamNonBlocking);
#define ALLOC(p, cnt)                                   \
    if (e == hipSuccess) e = dalloc(&(p), (size_t)(cnt));
    ALLOC(ctx->F.x, 2 * cap);
    ALLOC(ctx->F.v, 2 * cap);
    ALLOC(ctx->F.dv, 2 * cap);
    ALLOC(ctx->F.mom, 2 * cap);
    ALLOC(ctx->F.rho, cap);
    ALLOC(ctx->F.e, cap);
    ALLOC(ctx->F.P, cap);
    ALLOC(ctx->F.c2, cap);
    ALLOC(ctx->F.mass, cap);
    ALLOC(ctx->F.energy, cap);
    ALLOC(ctx->F.quality, cap);
    ALLOC(ctx->F.mu, cap);
    ALLOC(ctx->F.D, 4 * cap);
    ALLOC(ctx->F.area, cap);
    ALLOC(ctx->F.cent, 2 * cap);
    ALLOC(ctx->F.phase, cap);
    ALLOC(ctx->ne, cap);
    ALLOC(ctx->ev, 4 * LV_MAXE * cap);
    ALLOC(ctx->el, LV_MAXE * cap);
    ALLOC(ctx->clip_retry, cap + 1);
    ALLOC(ctx->cell_start, ncells + 1);
    ALLOC(ctx->d_nitems, 1);
    ALLOC(ctx->d_flags, 3);
    ALLOC(ctx->path_i, 2 * (size_t)std::max(1, g.npath));
    ALLOC(ctx->path_rr, (size_t)std::max(1, g.npath));
    // pressure system and solver work
    ALLOC(ctx->A.row_ptr, cap + 1);
    ALLOC(ctx->A.diag64, cap);
    ALLOC(ctx->A.dinv64, cap);
    ALLOC(ctx->A.diag32, cap);
    ALLOC(ctx->A.dinv32, cap);
    ALLOC(ctx->deg, cap + 1);
    ALLOC(ctx->A.ecol, (size_t)LV_MAXE * cap);
    ALLOC(ctx->A.ev64, (size_t)LV_MAXE * cap);
    ALLOC(ctx->A.ev32, (size_t)LV_MAXE * cap);
    ALLOC(ctx->A.ocol, ((size_t)cap / 64 + 1) * LV_OVF_CAP);
    ALLOC(ctx->A.oev64, ((size_t)cap / 64 + 1) * LV_OVF_CAP);
    ALLOC(ctx->A.oev32, ((size_t)cap / 64 + 1) * LV_OVF_CAP);
    ALLOC(ctx->A.deg8, cap);
    ALLOC(ctx->A.ecol16, (size_t)LV_MAXE * cap);
    ALLOC(ctx->A.ocol16, ((size_t)cap / 64 + 1) * LV_OVF_CAP);
    ALLOC(ctx->b, cap);
    ALLOC(ctx->P0, cap);
    ALLOC(ctx->GP, 2 * cap);
    ALLOC(ctx->W.r, cap);
    ALLOC(ctx->W.z, cap);
    ALLOC(ctx->W.p, cap);
    ALLOC(ctx->W.q, cap);
    ALLOC(ctx->W.e32, cap);
    ALLOC(ctx->W.r32, cap);
    ALLOC(ctx->W.z32, cap);
    ALLOC(ctx->W.p32, cap);
    ALLOC(ctx->W.q32, cap);
    ALLOC(ctx->W.S, 1);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&ctx->W.hprog, sizeof(lv_progress), hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) {
        memset(ctx->W.hprog, 0, sizeof(lv_progress));
        e = hipHostGetDevicePointer((void **)&ctx->W.dprog, ctx->W.hprog, 0);
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->W.hS, sizeof(lv_cg_scalars), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&ctx->W.partials, lv_partials_bytes((int)cap));
    if (e == hipSuccess) e = hipMalloc(&ctx->scan_temp, lv_scan_temp_bytes((int64_t)cap + 1) + 256);
#undef ALLOC
    if (e == hipSuccess && g.npath > 0) {
        e = hipMemcpy(ctx->path_i, ctx->h_path_i.data(), sizeof(int) * 2 * g.npath, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(ctx->path_rr, ctx->h_path_rr.data(), sizeof(double) * g.npath, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(ctx->d_flags, 0, 3 * sizeof(int));
    if (e == hipSuccess) {
        e = lv_spai_init(&ctx->spai);
        ctx->spai_ok = (e == hipSuccess);
    }
    if (e == hipSuccess) {
        e = lv_amg_init(&ctx->amg);
        ctx->amg_ok = (e == hipSuccess);
    }
    for (int k = 0; k < LV_NKTIMERS && e == hipSuccess; k++) {
        ctx->timers[k] = new lv_timer();
        memset(ctx->timers[k], 0, sizeof(lv_timer));
        for (int q = 0; q < 2 * LV_TPOOL && e == hipSuccess; q++) e = hipEventCreate(&ctx->timers[k]->ev[q]);
    }
    if (e != hipSuccess) {
        int code = lv_fail_hip(ctx, e, "lv_create", __FILE__, __LINE__);
        lv_destroy(ctx);
        return code;
    }
    *out = ctx;
    return LV_OK;
}

int lv_destroy(lv_ctx *ctx)
{
    if (!ctx) return LV_OK;
    if (ctx->s) hipStreamSynchronize(ctx->s);
    void *ptrs[] = {ctx->F.x, ctx->F.v, ctx->F.dv, ctx->F.mom, ctx->F.rho, ctx->F.e, ctx->F.P, ctx->F.c2,
                    ctx->F.mass, ctx->F.energy, ctx->F.quality, ctx->F.mu, ctx->F.D, ctx->F.area, ctx->F.cent,
                    ctx->F.phase, ctx->ne, ctx->ev, ctx->el, ctx->clip_retry, ctx->cell_start, ctx->cell_labels,
                    ctx->cell_x,
                    ctx->cl_temp, ctx->d_nitems, ctx->d_flags, ctx->path_i, ctx->path_rr, ctx->A.row_ptr,
                    ctx->A.cols, ctx->A.val64, ctx->A.val32, ctx->A.diag64, ctx->A.dinv64, ctx->A.diag32,
                    ctx->A.dinv32, ctx->deg, ctx->A.ecol, ctx->A.ev64, ctx->A.ev32, ctx->A.ocol, ctx->A.oev64,
                    ctx->A.oev32, ctx->A.deg8, ctx->A.ecol16, ctx->A.ocol16, ctx->b, ctx->P0, ctx->GP, ctx->W.r, ctx->W.z, ctx->W.p,
                    ctx->W.q, ctx->W.e32, ctx->W.r32, ctx->W.z32, ctx->W.p32, ctx->W.q32, ctx->W.S,
                    ctx->W.partials, ctx->scan_temp};
    for (void *p : ptrs) dfree(p);
    if (ctx->W.hprog) (void)hipHostFree(ctx->W.hprog);
    if (ctx->W.hS) (void)hipHostFree(ctx->W.hS);
    if (ctx->spai_ok) lv_spai_free(&ctx->spai);
    if (ctx->amg_ok) lv_amg_free(&ctx->amg);
    lv_mph_free(ctx->mph);
    for (int k = 0; k < 8; k++) {
        dfree(ctx->pw_q[k]);
        dfree(ctx->pw_z[k]);
    }
    for (int k = 0; k < 2; k++) {
        dfree(ctx->pw_P[k]);
        dfree(ctx->pw_AP[k]);
    }
    dfree(ctx->pw_t1);
    dfree(ctx->pw_t2);
    dfree(ctx->pw_dots);
    for (int k = 0; k < 2; k++) {
        dfree(ctx->pc_z[k]);
        dfree(ctx->pc_new[k]);
        dfree(ctx->pc_q[k]);
        dfree(ctx->pc_o[k]);
    }
    for (double *p : {ctx->rg.w1, ctx->rg.w2, ctx->rg.u1, ctx->rg.u2, ctx->rg.bdiv, ctx->rg.den, ctx->rg.ow1,
                      ctx->rg.ow2, ctx->rg.ou1, ctx->rg.ou2})
        dfree(p);
    dfree(ctx->rg.oel);
    for (int k = 0; k < LV_NKTIMERS; k++) {
        if (!ctx->timers[k]) continue;
        for (int q = 0; q < 2 * LV_TPOOL; q++)
            if (ctx->timers[k]->ev[q]) hipEventDestroy(ctx->timers[k]->ev[q]);
        delete ctx->timers[k];
    }
    for (int q = 0; q < 2 * lv_ctx::NSEV; q++)
        if (ctx->sev[q]) hipEventDestroy(ctx->sev[q]);
    if (ctx->s) hipStreamDestroy(ctx->s);
    delete ctx;
    return LV_OK;
}

const char *lv_last_error(lv_ctx *ctx) { return ctx ? ctx->msg : "null context"; }

int lv_sync(lv_ctx *ctx)
{
    if (!ctx) return LV_E_INVALID;
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_grid_info(lv_ctx *ctx, int32_t *n1, int32_t *n2, int32_t *npath, double *out6)
{
    if (!ctx) return LV_E_INVALID;
    const lv_grid_params &g = ctx->g;
    if (n1) *n1 = g.n1;
    if (n2) *n2 = g.n2;
    if (npath) *npath = g.npath;
    if (out6) {
        out6[0] = g.cmin0;
        out6[1] = g.cmin1;
        out6[2] = g.cmax0;
        out6[3] = g.cmax1;
        out6[4] = g.has_next ? g.next_rr : -1.0;
        out6[5] = g.rr_max;
    }
    return LV_OK;
}

int lv_grid_path(lv_ctx *ctx, int32_t *i1, int32_t *i2, double *rr)
{
    if (!ctx) return LV_E_INVALID;
    for (int k = 0; k < ctx->g.npath; k++) {
        i1[k] = ctx->h_path_i[2 * k];
        i2[k] = ctx->h_path_i[2 * k + 1];
        rr[k] = ctx->h_path_rr[k];
    }
    return LV_OK;
}

int lv_set_particles(lv_ctx *ctx, int32_t n, const double *x)
{
    if (!ctx || n < 0 || n > ctx->cap || (n > 0 && !x)) return LV_E_INVALID;
    ctx->n = n;
    ctx->mesh_valid = false;
    ctx->A.valid = 0;
    ctx->pc_cnt = 0;  // a new particle set: no carried Picard directions
    const size_t cap = (size_t)ctx->cap;
    double *zero2[] = {ctx->F.v, ctx->F.dv, ctx->F.mom, ctx->F.cent};
    for (double *p : zero2) LV_HIP_CHECK(hipMemsetAsync(p, 0, sizeof(double) * 2 * cap, ctx->s));
    double *zero1[] = {ctx->F.rho, ctx->F.e, ctx->F.P, ctx->F.c2, ctx->F.mass,
                       ctx->F.energy, ctx->F.quality, ctx->F.mu, ctx->F.area};
    for (double *p : zero1) LV_HIP_CHECK(hipMemsetAsync(p, 0, sizeof(double) * cap, ctx->s));
    LV_HIP_CHECK(hipMemsetAsync(ctx->F.D, 0, sizeof(double) * 4 * cap, ctx->s));
    LV_HIP_CHECK(hipMemsetAsync(ctx->F.phase, 0, sizeof(int) * cap, ctx->s));
    LV_HIP_CHECK(hipMemsetAsync(ctx->ne, 0, sizeof(int) * cap, ctx->s));
    if (n > 0) LV_HIP_CHECK(hipMemcpyAsync(ctx->F.x, x, sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_num_particles(lv_ctx *ctx, int32_t *n)
{
    if (!ctx || !n) return LV_E_INVALID;
    *n = ctx->n;
    return LV_OK;
}

int lv_upload(lv_ctx *ctx, int32_t field, const void *host)
{
    if (!ctx || !host) return LV_E_INVALID;
    size_t count, elem;
    bool ro;
    void *p = field_ptr(ctx, field, &count, &elem, &ro);
    if (!p || ro) return lv_fail(ctx, LV_E_INVALID, "invalid or read-only field");
    if (count == 0) return LV_OK;
    LV_HIP_CHECK(hipMemcpyAsync(p, host, count * elem, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    if (field == LV_FIELD_X) {
        ctx->mesh_valid = false;
        ctx->pc_cnt = 0;  // positions from elsewhere: no carried Picard corrections
    }
    return LV_OK;
}

int lv_download(lv_ctx *ctx, int32_t field, void *host)
{
    if (!ctx || !host) return LV_E_INVALID;
    size_t count, elem;
    bool ro;
    void *p = field_ptr(ctx, field, &count, &elem, &ro);
    if (!p) return lv_fail(ctx, LV_E_INVALID, "invalid field");
    if (count == 0) return LV_OK;
    LV_HIP_CHECK(hipMemcpyAsync(host, p, count * elem, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_field_device_ptr(lv_ctx *ctx, int32_t field, void **dptr)
{
    if (!ctx || !dptr) return LV_E_INVALID;
    size_t count, elem;
    bool ro;
    void *p = field_ptr(ctx, field, &count, &elem, &ro);
    if (!p) return LV_E_INVALID;
    *dptr = p;
    return LV_OK;
}

// ---- remesh ----------------------------------------------------------------------------------------

int lv_remesh(lv_ctx *ctx)
{
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    const int n = ctx->n;
    const int64_t ncells = (int64_t)ctx->g.n1 * ctx->g.n2;
    if (ctx->max_items == 0) {
        int64_t want = (int64_t)n + n / 4 + 4096;
        if (!ctx->g.xperiodic && !ctx->g.yperiodic) want = n + 1;
        want = std::min<int64_t>(want, 9LL * n);
        ctx->max_items = want;
        LV_HIP_CHECK(dalloc(&ctx->cell_labels, want));
        LV_HIP_CHECK(dalloc(&ctx->cell_x, 2 * want));
        LV_HIP_CHECK(hipMalloc(&ctx->cl_temp, lv_celllist_temp_bytes(std::max<int64_t>(want, n + 1), ncells)));
    }
    rc = clear_flags(ctx);
    if (rc) return rc;
    lv_timer *tc = timer_if_on(ctx, 3);
    if (tc && tc->used == LV_TPOOL) lv_timer_flush(tc);
    for (int attempt = 0; attempt < 2; attempt++) {
        hipError_t e = lv_build_cell_list(ctx->g, n, ctx->F.x, ctx->cell_start, ctx->cell_labels, ctx->cell_x,
                                          ctx->max_items, ctx->d_nitems, ctx->cl_temp, ctx->s,
                                          tc ? tc->ev[2 * tc->used] : nullptr, tc ? tc->ev[2 * tc->used + 1] : nullptr,
                                          tc != nullptr, &ctx->nitems);
        if (e == hipErrorInvalidValue && attempt == 0) {
            // more periodic images than buffered: grow to the count the scan produced and rebuild
            int m = 0;
            LV_HIP_CHECK(hipMemcpy(&m, ctx->d_nitems, sizeof(int), hipMemcpyDeviceToHost));
            if (m <= 0) return lv_fail(ctx, LV_E_INVALID, "cell list: no items");
            dfree(ctx->cell_labels);
            dfree(ctx->cell_x);
            dfree(ctx->cl_temp);
            ctx->cell_labels = nullptr;
            ctx->cell_x = nullptr;
            ctx->cl_temp = nullptr;
            int64_t want = (int64_t)m + m / 8 + 4096;
            ctx->max_items = want;
            LV_HIP_CHECK(dalloc(&ctx->cell_labels, want));
            LV_HIP_CHECK(dalloc(&ctx->cell_x, 2 * want));
            LV_HIP_CHECK(hipMalloc(&ctx->cl_temp, lv_celllist_temp_bytes(std::max<int64_t>(want, n + 1), ncells)));
            continue;
        }
        LV_HIP_CHECK(e);
        if (tc) tc->used++;
        break;
    }
    lv_timer *tk = timer_if_on(ctx, 2);
    if (tk) lv_timer_begin(tk, ctx->s);
    LV_HIP_CHECK(lv_launch_clip(ctx->g, n, ctx->F.x, ctx->cell_start, ctx->cell_labels, ctx->cell_x, ctx->path_i,
                                ctx->path_rr, ctx->ne, ctx->ev, ctx->el, n, ctx->F.area, ctx->F.cent, ctx->d_flags,
                                ctx->clip_retry, ctx->s, (int)ctx->max_items));
    if (tk) lv_timer_end(tk, ctx->s);
    int flags[2];
    rc = read_flags(ctx, flags);
    if (rc) return rc;
    if (tk) lv_timer_flush(tk);
    if (tc) lv_timer_flush(tc);
    ctx->A.valid = 0;
    if (flags[0] != 0) {
        ctx->mesh_valid = false;
        return lv_fail(ctx, flags[0], error_text(flags[0]));
    }
    ctx->mesh_valid = true;
    return LV_OK;
}

int lv_mesh_size(lv_ctx *ctx, int64_t *nedges)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    std::vector<int> ne(ctx->n);
    LV_HIP_CHECK(hipMemcpy(ne.data(), ctx->ne, sizeof(int) * ctx->n, hipMemcpyDeviceToHost));
    int64_t t = 0;
    for (int v : ne) t += v;
    *nedges = t;
    return LV_OK;
}

int lv_get_mesh(lv_ctx *ctx, int32_t *edge_ptr, int32_t *labels, double *verts)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    const int n = ctx->n;
    std::vector<int> ne(n), el((size_t)LV_MAXE * n);
    std::vector<double> ev((size_t)4 * LV_MAXE * n);
    LV_HIP_CHECK(hipMemcpy(ne.data(), ctx->ne, sizeof(int) * n, hipMemcpyDeviceToHost));
    LV_HIP_CHECK(hipMemcpy(el.data(), ctx->el, sizeof(int) * LV_MAXE * n, hipMemcpyDeviceToHost));
    LV_HIP_CHECK(hipMemcpy(ev.data(), ctx->ev, sizeof(double) * 4 * LV_MAXE * n, hipMemcpyDeviceToHost));
    int64_t o = 0;
    for (int i = 0; i < n; i++) {
        edge_ptr[i] = (int32_t)o;
        for (int k = 0; k < ne[i]; k++) {
            labels[o] = el[(size_t)k * n + i];
            for (int p = 0; p < 4; p++) verts[4 * o + p] = ev[(size_t)(p * LV_MAXE + k) * n + i];
            o++;
        }
    }
    edge_ptr[n] = (int32_t)o;
    return LV_OK;
}

// ---- VTK export (src/IO.jl:9-71) ------------------------------------------------------------------

namespace {

// the exportable PolygonNS fields (src/celldefs.jl:7-39) by name: field id and whether it is a
// Number (Float64 column) or converts to Vec3 (append_datasets!, src/IO.jl:61-71)
struct export_field {
    const char *name;
    int id;
};
const export_field kExport[] = {{"x", LV_FIELD_X},         {"v", LV_FIELD_V},         {"dv", LV_FIELD_DV},
                                {"momentum", LV_FIELD_MOMENTUM}, {"rho", LV_FIELD_RHO}, {"e", LV_FIELD_E},
                                {"P", LV_FIELD_P},         {"c2", LV_FIELD_C2},       {"mass", LV_FIELD_MASS},
                                {"energy", LV_FIELD_ENERGY}, {"quality", LV_FIELD_QUALITY}, {"mu", LV_FIELD_MU},
                                {"D", LV_FIELD_D},         {"phase", LV_FIELD_PHASE}};

static std::string vtp_path(const char *filename)
{
    std::string f(filename);
    if (f.size() < 4 || f.compare(f.size() - 4, 4, ".vtp") != 0) f += ".vtp";
    return f;
}

// the cell-data columns of the requested fields: Float64 for numbers (the phase too), Vec3 (x, y,
// 0) for vectors, and for the matrix D its first two entries in column-major order, as
// Vec3(D) does
int export_columns(lv_ctx *ctx, const int32_t *fields, int32_t nfields, std::vector<std::vector<double>> &cols,
                   std::vector<lv_vtk_array> &arr)
{
    const int n = ctx->n;
    cols.assign(nfields, {});
    arr.assign(nfields, {});
    for (int a = 0; a < nfields; a++) {
        const export_field *ef = nullptr;
        for (const export_field &e : kExport)
            if (e.id == fields[a]) ef = &e;
        if (!ef) {
            char msg[160];
            snprintf(msg, sizeof(msg), "Cannot export variable %d because it does not exist.", (int)fields[a]);
            return lv_fail(ctx, LV_E_INVALID, msg);
        }
        size_t count, elem;
        bool ro;
        void *p = field_ptr(ctx, ef->id, &count, &elem, &ro);
        std::vector<double> &c = cols[a];
        if (ef->id == LV_FIELD_PHASE) {
            std::vector<int> h(n);
            if (n) LV_HIP_CHECK(hipMemcpy(h.data(), p, sizeof(int) * n, hipMemcpyDeviceToHost));
            c.assign(h.begin(), h.end());
            arr[a] = lv_vtk_array{ef->name, 1, c.data()};
            continue;
        }
        std::vector<double> h(count);
        if (count) LV_HIP_CHECK(hipMemcpy(h.data(), p, sizeof(double) * count, hipMemcpyDeviceToHost));
        if (count == (size_t)n) {
            c = std::move(h);
            arr[a] = lv_vtk_array{ef->name, 1, c.data()};
        } else {
            c.resize((size_t)3 * n);
            for (int i = 0; i < n; i++) {
                c[3 * (size_t)i] = h[i];
                c[3 * (size_t)i + 1] = h[(size_t)n + i];
                c[3 * (size_t)i + 2] = 0.0;
            }
            arr[a] = lv_vtk_array{ef->name, 3, c.data()};
        }
    }
    return LV_OK;
}

}  // namespace

int lv_export_grid(lv_ctx *ctx, const char *filename, const int32_t *fields, int32_t nfields, int32_t compress)
{
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (!filename || nfields < 0 || (nfields > 0 && !fields)) return lv_fail(ctx, LV_E_INVALID, "lv_export_grid");
    const int n = ctx->n;
    std::vector<int> ne(n), dummy;
    std::vector<double> ev((size_t)2 * LV_MAXE * n);
    if (n) {
        LV_HIP_CHECK(hipMemcpy(ne.data(), ctx->ne, sizeof(int) * n, hipMemcpyDeviceToHost));
        // the v1 planes of the edge slots
        LV_HIP_CHECK(hipMemcpy(ev.data(), ctx->ev, sizeof(double) * 2 * LV_MAXE * n, hipMemcpyDeviceToHost));
    }
    // every polygon's edges in their (chained) order: its vertices are the edges' v1 (src/IO.jl:17-27)
    std::vector<double> pts;
    std::vector<int64_t> conn, offs(n);
    int64_t o = 0;
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < ne[i]; k++) {
            pts.push_back(ev[(size_t)k * n + i]);
            pts.push_back(ev[(size_t)(LV_MAXE + k) * n + i]);
            pts.push_back(0.0);
            conn.push_back(o++);
        }
        offs[i] = o;
    }
    std::vector<std::vector<double>> cols;
    std::vector<lv_vtk_array> arr;
    rc = export_columns(ctx, fields, nfields, cols, arr);
    if (rc) return rc;
    const std::string path = vtp_path(filename);
    if (lv_write_vtp(path.c_str(), o, pts.data(), LV_VTK_POLYS, n, conn.data(), offs.data(), arr.data(), nfields,
                     compress) != 0)
        return lv_fail(ctx, LV_E_INVALID, "cannot write the VTK file");
    return LV_OK;
}

int lv_export_points(lv_ctx *ctx, const char *filename, const int32_t *fields, int32_t nfields, int32_t compress)
{
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (!filename || nfields < 0 || (nfields > 0 && !fields)) return lv_fail(ctx, LV_E_INVALID, "lv_export_points");
    const int n = ctx->n;
    std::vector<double> x((size_t)2 * n), pts((size_t)3 * n);
    if (n) LV_HIP_CHECK(hipMemcpy(x.data(), ctx->F.x, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
    std::vector<int64_t> conn(n), offs(n);
    for (int i = 0; i < n; i++) {
        pts[3 * (size_t)i] = x[i];
        pts[3 * (size_t)i + 1] = x[(size_t)n + i];
        pts[3 * (size_t)i + 2] = 0.0;
        conn[i] = i;
        offs[i] = i + 1;
    }
    std::vector<std::vector<double>> cols;
    std::vector<lv_vtk_array> arr;
    rc = export_columns(ctx, fields, nfields, cols, arr);
    if (rc) return rc;
    const std::string path = vtp_path(filename);
    if (lv_write_vtp(path.c_str(), n, pts.data(), LV_VTK_VERTS, n, conn.data(), offs.data(), arr.data(), nfields,
                     compress) != 0)
        return lv_fail(ctx, LV_E_INVALID, "cannot write the VTK file");
    return LV_OK;
}

// ---- interpolation (src/movingls.jl, src/voronoigrid.jl:154-178) ------------------------------------

namespace {

// the device plane of component comp of a Float64 field, or null
const double *field_plane(lv_ctx *ctx, int field, int comp)
{
    size_t count, elem;
    bool ro;
    double *p = field_ptr(ctx, field, &count, &elem, &ro);
    if (!p || elem != sizeof(double) || ctx->n <= 0) return nullptr;
    const size_t comps = count / (size_t)ctx->n;
    if (comp < 0 || (size_t)comp >= comps) return nullptr;
    return p + (size_t)comp * ctx->n;
}

int need_cell_list(lv_ctx *ctx)
{
    if (!ctx->mesh_valid) return lv_fail(ctx, LV_E_INVALID, "the cell list is older than the positions: remesh first");
    return LV_OK;
}

// device scratch of one call, freed on every return path
struct dev_buf {
    void *p = nullptr;
    ~dev_buf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int lv_nearest_polygon(lv_ctx *ctx, int32_t m, const double *q, int32_t *idx)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (m < 0 || (m > 0 && (!q || !idx))) return lv_fail(ctx, LV_E_INVALID, "lv_nearest_polygon");
    if ((rc = need_cell_list(ctx))) return rc;
    if (m == 0) return LV_OK;
    dev_buf dq, di;
    LV_HIP_CHECK(hipMalloc(&dq.p, sizeof(double) * 2 * m));
    LV_HIP_CHECK(hipMalloc(&di.p, sizeof(int) * m));
    LV_HIP_CHECK(hipMemcpyAsync(dq.p, q, sizeof(double) * 2 * m, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(lv_launch_nearest(ctx->g, ctx->F.x, ctx->n, ctx->cell_start, ctx->cell_labels, ctx->path_i,
                                   ctx->path_rr, m, (const double *)dq.p, (int *)di.p, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(idx, di.p, sizeof(int) * m, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    for (int t = 0; t < m; t++)
        if (idx[t] < 0) return lv_fail(ctx, LV_E_INVALID, "nearest polygon beyond r_max of a query point");
    return LV_OK;
}

int lv_movingls(lv_ctx *ctx, int32_t order, int32_t field, int32_t comp, double h, double *coeffs)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (order < 1 || order > 3 || !coeffs) return lv_fail(ctx, LV_E_INVALID, "lv_movingls: order 1, 2 or 3");
    if ((rc = need_cell_list(ctx))) return rc;
    const double *f = field_plane(ctx, field, comp);
    if (!f) return lv_fail(ctx, LV_E_INVALID, "lv_movingls: not a Float64 field component");
    if (!(h > 0.0)) h = ctx->g.h;
    if (h > ctx->g.r_max) return lv_fail(ctx, LV_E_INVALID, "lv_movingls: h beyond r_max");
    const int N = order == 1 ? 2 : (order == 2 ? 5 : 9);
    const size_t bytes = sizeof(double) * (size_t)N * ctx->n;
    dev_buf d;
    LV_HIP_CHECK(hipMalloc(&d.p, bytes));
    LV_HIP_CHECK(lv_launch_movingls(ctx->g, ctx->F.x, ctx->n, ctx->cell_start, ctx->cell_labels, ctx->path_i,
                                    ctx->path_rr, order, f, h, (double *)d.p, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(coeffs, d.p, bytes, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_point_values(lv_ctx *ctx, int32_t m, const double *q, int32_t field, int32_t comp, double *out)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (m < 0 || (m > 0 && (!q || !out))) return lv_fail(ctx, LV_E_INVALID, "lv_point_values");
    if ((rc = need_cell_list(ctx))) return rc;
    const double *f = field_plane(ctx, field, comp);
    if (!f) return lv_fail(ctx, LV_E_INVALID, "lv_point_values: not a Float64 field component");
    if (m == 0) return LV_OK;
    dev_buf dq, dout;
    LV_HIP_CHECK(hipMalloc(&dq.p, sizeof(double) * 2 * m));
    LV_HIP_CHECK(hipMalloc(&dout.p, sizeof(double) * m));
    LV_HIP_CHECK(hipMemcpyAsync(dq.p, q, sizeof(double) * 2 * m, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(lv_launch_point_value(ctx->g, ctx->F.x, ctx->n, ctx->cell_start, ctx->cell_labels, ctx->path_i,
                                       ctx->path_rr, m, (const double *)dq.p, f, ctx->g.h, (double *)dout.p, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(out, dout.p, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    for (int t = 0; t < m; t++)
        if (out[t] != out[t]) return lv_fail(ctx, LV_E_INVALID, "nearest polygon beyond r_max of a query point");
    return LV_OK;
}

// lv_point_values with the queries and values staged in the static `scratch` array instead of
// device buffers allocated per call: no allocation of any kind, for callers that may not allocate.
// Concurrent calls on one device race on that array (calls on one context are serialised anyway).
int lv_point_values_noalloc(lv_ctx *ctx, int32_t m, const double *q, int32_t field, int32_t comp, double *out)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (m < 0 || (m > 0 && (!q || !out))) return lv_fail(ctx, LV_E_INVALID, "lv_point_values_noalloc");
    if (m > LV_NOALLOC_MAX_QUERIES)
        return lv_fail(ctx, LV_E_INVALID, "lv_point_values_noalloc: more than LV_NOALLOC_MAX_QUERIES (65536) "
                                          "query points in one call");
    if ((rc = need_cell_list(ctx))) return rc;
    const double *f = field_plane(ctx, field, comp);
    if (!f) return lv_fail(ctx, LV_E_INVALID, "lv_point_values_noalloc: not a Float64 field component");
    if (m == 0) return LV_OK;
    if (!ctx->scratch) return lv_fail(ctx, LV_E_HIP, "lv_point_values_noalloc: staging array not resolved");
    double *dq = ctx->scratch;                       // q1[m], q2[m]
    double *dout = ctx->scratch + 2 * (size_t)m;     // the m values
    LV_HIP_CHECK(hipMemcpyAsync(dq, q, sizeof(double) * 2 * m, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(lv_launch_point_value(ctx->g, ctx->F.x, ctx->n, ctx->cell_start, ctx->cell_labels, ctx->path_i,
                                       ctx->path_rr, m, (const double *)dq, f, ctx->g.h, dout, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(out, dout, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    for (int t = 0; t < m; t++)
        if (out[t] != out[t]) return lv_fail(ctx, LV_E_INVALID, "nearest polygon beyond r_max of a query point");
    return LV_OK;
}

// populate_lloyd!'s iterations (src/populate.jl:132-145): niterations times remesh! and x = centroid(p)
int lv_lloyd(lv_ctx *ctx, int32_t niterations)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (niterations < 0) return lv_fail(ctx, LV_E_INVALID, "lv_lloyd");
    for (int it = 0; it < niterations; it++) {
        if ((rc = lv_remesh(ctx))) return rc;
        LV_HIP_CHECK(hipMemcpyAsync(ctx->F.x, ctx->F.cent, sizeof(double) * 2 * ctx->n, hipMemcpyDeviceToDevice,
                                    ctx->s));
        ctx->mesh_valid = false;
    }
    return lv_remesh(ctx);
}

int lv_cell_list_size(lv_ctx *ctx, int64_t *nitems)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    *nitems = ctx->nitems;
    return LV_OK;
}

int lv_get_cell_list(lv_ctx *ctx, int32_t *cell_start, int32_t *cell_labels)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    const int64_t ncells = (int64_t)ctx->g.n1 * ctx->g.n2;
    LV_HIP_CHECK(hipMemcpy(cell_start, ctx->cell_start, sizeof(int) * (ncells + 1), hipMemcpyDeviceToHost));
    LV_HIP_CHECK(hipMemcpy(cell_labels, ctx->cell_labels, sizeof(int) * ctx->nitems, hipMemcpyDeviceToHost));
    return LV_OK;
}

// ---- operators -------------------------------------------------------------------------------------

int lv_move(lv_ctx *ctx, double dt)
{
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    rc = clear_flags(ctx);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_move(ctx->g, ctx->F, mesh_view(ctx), ctx->n, dt, ctx->d_flags, ctx->s));
    int flags[2];
    rc = read_flags(ctx, flags);
    if (rc) return rc;
    ctx->mesh_valid = false;
    if (flags[0]) return lv_fail(ctx, flags[0], error_text(flags[0]));
    return lv_remesh(ctx);
}

int lv_gravity_step(lv_ctx *ctx, const double g[2], double dt)
{
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_gravity(ctx->F, ctx->n, g[0], g[1], dt, ctx->s));
    return LV_OK;
}

int lv_find_rho(lv_ctx *ctx)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_find_rho(ctx->F, ctx->n, ctx->s));
    return LV_OK;
}

int lv_eos_stiffened(lv_ctx *ctx, double gamma, double P0)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_eos(ctx->F, ctx->n, 1, gamma, P0, 0.0, ctx->s));
    return LV_OK;
}

int lv_eos_ideal(lv_ctx *ctx, double gamma, double Pmin)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_eos(ctx->F, ctx->n, 2, gamma, 0.0, Pmin, ctx->s));
    return LV_OK;
}

// the system of the inner solver: mode 0 the pressure operator, mode 1 ViscousSolver's operator; with
// rhs_vbc (mode 0) the assembly also stores the P-independent parts of refresh!(solver) in ctx->rg
static int assemble_system(lv_ctx *ctx, double dt, int mode, int64_t *nnz, const lv_vbc *rhs_vbc = nullptr)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    const int n = ctx->n;
    lv_mesh_view M = mesh_view(ctx);
    LV_HIP_CHECK(lv_launch_degree(M, n, ctx->deg, ctx->s));
    // row_ptr[0..n] = exclusive scan of the n degrees followed by one extra slot: deg[n] is 0
    LV_HIP_CHECK(hipMemsetAsync(ctx->deg + n, 0, sizeof(int), ctx->s));
    LV_HIP_CHECK(lv_exclusive_scan_i32(ctx->deg, ctx->A.row_ptr, n + 1, nullptr, ctx->scan_temp, ctx->s));
    int total = 0;
    LV_HIP_CHECK(hipMemcpyAsync(&total, ctx->A.row_ptr + n, sizeof(int), hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    if (total > ctx->A.nnz_cap) {
        dfree(ctx->A.cols);
        dfree(ctx->A.val64);
        dfree(ctx->A.val32);
        int64_t want = (int64_t)total + total / 8 + 1024;
        LV_HIP_CHECK(dalloc(&ctx->A.cols, want));
        LV_HIP_CHECK(dalloc(&ctx->A.val64, want));
        LV_HIP_CHECK(dalloc(&ctx->A.val32, want));
        ctx->A.nnz_cap = want;
    }
    ctx->A.nnz = total;
    ctx->A.deg = ctx->deg;
    rc = clear_flags(ctx);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_assemble(ctx->g, ctx->F, M, n, dt, mode, ctx->A, ctx->d_flags + 1, ctx->d_flags + 2,
                                    ctx->s, rhs_vbc, rhs_vbc ? &ctx->rg : nullptr));
    int flags[3];
    LV_HIP_CHECK(hipMemcpyAsync(flags, ctx->d_flags, 3 * sizeof(int), hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    ctx->A.singular = (flags[1] == 0);
    ctx->A.narrow = (flags[2] == 0);
    ctx->assembly_id++;
    ctx->A.valid = 1;
    ctx->A_kind = mode;
    if (nnz) *nnz = total;
    return LV_OK;
}

int lv_assemble_pressure(lv_ctx *ctx, double dt, int64_t *nnz) { return assemble_system(ctx, dt, 0, nnz); }

static int need_pressure_operator(lv_ctx *ctx)
{
    if (!ctx->A.valid) return lv_fail(ctx, LV_E_INVALID, "pressure operator not assembled");
    if (ctx->A_kind != 0)
        return lv_fail(ctx, LV_E_INVALID, "the last assembled operator is the implicit viscous one; assemble the "
                                          "pressure operator first");
    return LV_OK;
}

int lv_get_pressure_operator(lv_ctx *ctx, double *diag, int32_t *row_ptr, int32_t *cols, double *coef)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    rc = need_pressure_operator(ctx);
    if (rc) return rc;
    const int n = ctx->n;
    LV_HIP_CHECK(hipMemcpy(diag, ctx->A.diag64, sizeof(double) * n, hipMemcpyDeviceToHost));
    LV_HIP_CHECK(hipMemcpy(row_ptr, ctx->A.row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (cols) LV_HIP_CHECK(hipMemcpy(cols, ctx->A.cols, sizeof(int) * ctx->A.nnz, hipMemcpyDeviceToHost));
    if (coef) LV_HIP_CHECK(hipMemcpy(coef, ctx->A.val64, sizeof(double) * ctx->A.nnz, hipMemcpyDeviceToHost));
    return LV_OK;
}

static int ensure_rhs_geom(lv_ctx *ctx)
{
    if (ctx->rg_alloc) return LV_OK;
    const size_t cap = (size_t)ctx->cap;
    LV_HIP_CHECK(dalloc(&ctx->rg.w1, (size_t)LV_MAXE * cap));
    LV_HIP_CHECK(dalloc(&ctx->rg.w2, (size_t)LV_MAXE * cap));
    LV_HIP_CHECK(dalloc(&ctx->rg.u1, (size_t)LV_MAXE * cap));
    LV_HIP_CHECK(dalloc(&ctx->rg.u2, (size_t)LV_MAXE * cap));
    LV_HIP_CHECK(dalloc(&ctx->rg.bdiv, cap));
    LV_HIP_CHECK(dalloc(&ctx->rg.den, cap));
    const size_t mo = (cap / 64 + 1) * LV_OVF_CAP;
    LV_HIP_CHECK(dalloc(&ctx->rg.oel, mo));
    for (double **p : {&ctx->rg.ow1, &ctx->rg.ow2, &ctx->rg.ou1, &ctx->rg.ou2}) LV_HIP_CHECK(dalloc(p, mo));
    ctx->rg.ocap = 0;
    ctx->rg_alloc = true;
    return LV_OK;
}

int lv_pressure_rhs(lv_ctx *ctx, double dt, int32_t gp_step, const double vbc[8], double *b, double *GP)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    lv_vbc v;
    for (int k = 0; k < 8; k++) v.v[k] = vbc ? vbc[k] : 0.0;
    // the same split evaluation as find_pressure!: the P-independent parts, GP, then b
    rc = ensure_rhs_geom(ctx);
    if (rc) return rc;
    ctx->rg.ocap = ctx->n >= 65536 ? LV_OVF_CAP : 0;
    const int n = ctx->n;
    const lv_mesh_view M = mesh_view(ctx);
    LV_HIP_CHECK(lv_launch_rhs_geom(ctx->g, ctx->F, M, n, dt, v, ctx->rg, ctx->s));
    LV_HIP_CHECK(lv_launch_gp(M, n, ctx->rg, ctx->F.P, ctx->F.mass, ctx->GP, nullptr, nullptr, ctx->s));
    LV_HIP_CHECK(lv_launch_rhs_b(M, n, gp_step, ctx->rg, ctx->GP, ctx->F.area, ctx->F.P, ctx->b, ctx->P0, ctx->s));
    if (b) LV_HIP_CHECK(hipMemcpyAsync(b, ctx->b, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->s));
    if (GP) LV_HIP_CHECK(hipMemcpyAsync(GP, ctx->GP, sizeof(double) * 2 * ctx->n, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

static void default_opts(lv_solve_opts &o)
{
    memset(&o, 0, sizeof(o));
    o.method = LV_SOLVER_PCG64;
    o.tol = 1e-10;
    o.itmax = 1000;
    o.inner_tol = 1e-4;
    o.outer_max = 20;
    o.picard_window = 0;
}

// The preconditioner of LV_SOLVER_SPAI (the learned M) or LV_SOLVER_AMG (the multilevel cycle,
// with the learned M as fine smoother when parameters are loaded), built once per assembly.
static int spai_prepare(lv_ctx *ctx, const lv_solve_opts &opt, lv_precond *M, bool *have)
{
    *have = false;
    const bool amg = (opt.method == LV_SOLVER_AMG);
    if (opt.method != LV_SOLVER_SPAI && !amg) return LV_OK;
    const bool learned = ctx->spai_ok && ctx->spai.has_params;
    if (learned && ctx->spai_emitted_for != ctx->assembly_id) {
        LV_HIP_CHECK(lv_spai_infer(&ctx->spai, ctx->A, ctx->n, ctx->s));
        LV_HIP_CHECK(lv_spai_emit(&ctx->spai, ctx->A, ctx->n, ctx->s));
        ctx->spai_emitted_for = ctx->assembly_id;
    }
    if (!amg) {
        if (!learned) return LV_OK;
        M->mval = ctx->spai.mval;
        M->mdg = ctx->spai.mdg;
        *have = true;
        return LV_OK;
    }
    if (!ctx->amg_ok) return lv_fail(ctx, LV_E_INVALID, "multilevel preconditioner unavailable");
    if (ctx->amg_built_for != ctx->assembly_id || ctx->amg_learned != (int)learned ||
        ctx->amg_layout != opt.amg_layout) {
        LV_HIP_CHECK(lv_amg_setup(&ctx->amg, ctx->g, ctx->F.x, ctx->A, ctx->n, learned ? ctx->spai.mval : nullptr,
                                  learned ? ctx->spai.mdg : nullptr, opt.amg_layout, ctx->s));
        ctx->amg_built_for = ctx->assembly_id;
        ctx->amg_learned = (int)learned;
        ctx->amg_layout = opt.amg_layout;
    }
    ctx->amg.level_kernels = opt.amg_level_kernels ? 1 : 0;
    M->amg = &ctx->amg;
    M->A = &ctx->A;
    *have = true;
    return LV_OK;
}

// the solve times still to be read: each pair's elapsed time into its slot
static void flush_solve_times(lv_ctx *ctx)
{
    for (int k = 0; k < ctx->sev_n; k++) {
        float ms = 0.0f;
        if (hipEventSynchronize(ctx->sev[2 * k + 1]) == hipSuccess &&
            hipEventElapsedTime(&ms, ctx->sev[2 * k], ctx->sev[2 * k + 1]) == hipSuccess)
            *ctx->sev_out[k] = ms;
    }
    ctx->sev_n = 0;
}

// an event pair for one solve's time, whose value goes to *out when the times are flushed
static int solve_time_pair(lv_ctx *ctx, double *out, hipEvent_t **pair)
{
    if (!ctx->sev_ready) {
        for (int q = 0; q < 2 * lv_ctx::NSEV; q++) LV_HIP_CHECK(hipEventCreate(&ctx->sev[q]));
        ctx->sev_ready = 1;
    }
    if (ctx->sev_n == lv_ctx::NSEV) flush_solve_times(ctx);
    const int k = ctx->sev_n++;
    ctx->sev_out[k] = out;
    *pair = &ctx->sev[2 * k];
    return LV_OK;
}

// flushes the solve times when an API call that fills statistics returns, on every path (the
// statistics belong to the caller and are only valid during the call)
struct solve_time_scope {
    lv_ctx *ctx;
    ~solve_time_scope() { flush_solve_times(ctx); }
};

// one linear solve on the assembled operator, device vectors b and x (x in/out)
static int solve_on_device(lv_ctx *ctx, const double *db, double *dx, const lv_solve_opts &opt, lv_solve_stats *st)
{
    lv_precond M{};
    bool haveM = false;
    // the solve time includes the network inference when this solve is the first on a new operator
    hipEvent_t *ev = nullptr;
    if (st && st->nsolves < 32) {
        int rc = solve_time_pair(ctx, &st->solve_ms[st->nsolves], &ev);
        if (rc) return rc;
        LV_HIP_CHECK(hipEventRecord(ev[0], ctx->s));
    }
    int rc = spai_prepare(ctx, opt, &M, &haveM);
    if (rc) {
        if (ev) ctx->sev_n--;  // this solve's pair (the last one) is dropped
        return rc;
    }
    int iters = 0, sweeps = 0, fb = 0;
    double relres = 0.0, relres0 = 0.0;
    bool conv = false;
    ctx->spmv32_agg = (haveM && M.amg && M.amg->perm_ok) ? 1 : 0;
    if (st) st->agg_order = ctx->spmv32_agg;
    hipError_t e = lv_solve_device(ctx->A, ctx->W, ctx->n, db, dx, opt, ctx->s, timer_if_on(ctx, 0),
                                   timer_if_on(ctx, 1), &iters, &sweeps, &relres, &relres0, &conv,
                                   haveM ? &M : nullptr, &fb);
    if (opt.method == LV_SOLVER_SPAI && !haveM) fb++;  // no trained network: Jacobi
    if (e != hipSuccess) {
        if (ev) ctx->sev_n--;
        return lv_fail_hip(ctx, e, "lv_solve_device", __FILE__, __LINE__);
    }
    if (ev) LV_HIP_CHECK(hipEventRecord(ev[1], ctx->s));
    if (st && st->nsolves < 32) {
        int k = st->nsolves;
        st->iters[k] = iters;
        st->outer[k] = sweeps;
        st->relres[k] = relres;
        st->relres0[k] = relres0;
    }
    if (st) {
        st->nsolves++;
        st->total_iters += iters;
        st->fallbacks += fb;
    }
    return conv ? LV_OK : LV_E_NOT_CONVERGED;
}

int lv_apply_preconditioner(lv_ctx *ctx, int32_t method, const double *r, double *z)
{
    if (!ctx || !r || !z) return LV_E_INVALID;
    if (!ctx->A.valid) return lv_fail(ctx, LV_E_INVALID, "no assembled pressure operator");
    if (method != LV_SOLVER_MIXED && method != LV_SOLVER_SPAI && method != LV_SOLVER_AMG)
        return lv_fail(ctx, LV_E_INVALID, "lv_apply_preconditioner: method has no FP32 preconditioner");
    lv_solve_opts opt{};
    opt.method = method;
    opt.amg_layout = ctx->amg_layout > 0 ? ctx->amg_layout : 0;  // the layout of the last multilevel solve
    opt.amg_level_kernels = ctx->amg.level_kernels;               // and its coarse-level kernels
    lv_precond M{};
    bool have = false;
    int rc = spai_prepare(ctx, opt, &M, &have);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * ctx->n;
    LV_HIP_CHECK(hipMemcpyAsync(ctx->W.r, r, bytes, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(lv_apply_precond(ctx->A, ctx->W, ctx->n, have ? &M : nullptr, ctx->W.r, ctx->W.z, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(z, ctx->W.z, bytes, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_solve_pressure(lv_ctx *ctx, const double *b, double *x, const lv_solve_opts *opts, lv_solve_stats *stats)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    solve_time_scope times{ctx};
    rc = need_pressure_operator(ctx);
    if (rc) return rc;
    lv_solve_opts o;
    if (opts) o = *opts;
    else default_opts(o);
    const int n = ctx->n;
    LV_HIP_CHECK(hipMemcpyAsync(ctx->b, b, sizeof(double) * n, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(ctx->P0, x, sizeof(double) * n, hipMemcpyHostToDevice, ctx->s));
    if (stats) memset(stats, 0, sizeof(*stats));
    rc = solve_on_device(ctx, ctx->b, ctx->P0, o, stats);
    LV_HIP_CHECK(hipMemcpyAsync(x, ctx->P0, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return rc;
}

static int ensure_picard_window(lv_ctx *ctx, int w)
{
    const size_t cap = (size_t)ctx->cap;
    if (ctx->pw_cap == 0) {
        for (int k = 0; k < 2; k++) {
            LV_HIP_CHECK(dalloc(&ctx->pw_P[k], cap));
            LV_HIP_CHECK(dalloc(&ctx->pw_AP[k], cap));
        }
        LV_HIP_CHECK(dalloc(&ctx->pw_t1, cap));
        LV_HIP_CHECK(dalloc(&ctx->pw_t2, cap));
        LV_HIP_CHECK(dalloc(&ctx->pw_dots, 8));
    }
    for (int k = ctx->pw_cap; k < w; k++) {
        LV_HIP_CHECK(dalloc(&ctx->pw_q[k], cap));
        LV_HIP_CHECK(dalloc(&ctx->pw_z[k], cap));
    }
    if (w > ctx->pw_cap) ctx->pw_cap = w;
    return LV_OK;
}

// Adds the Picard correction (P_cur - P_prev, AP_cur - AP_prev) to the window: one round of
// classical Gram-Schmidt against the stored q's (applied identically to the z's), then
// normalisation; the oldest direction makes room when the window is full.  (A second round left
// the iterations of the 2048^2 benchmark steps unchanged and cost 0.8 ms per step.)  The dots stay on the
// device (the combinations read them there), so the host never waits; a correction with no
// component left enters as a zero direction, which adds nothing to later projections.
// the direction in pw_t1 with its image A pw_t1 in pw_t2 (both overwritten) into the set (Q, Z) of
// at most w
static int picard_add_pair(lv_ctx *ctx, float **Q, float **Z, int w, int *cnt)
{
    const int n = ctx->n;
    hipStream_t s = ctx->s;
    const double *xp[8];
    const float *xf[8];
    double c[8];
    if (*cnt > 0) {
        const double *b[8];
        for (int k = 0; k < *cnt; k++) {
            xf[k] = Q[k];
            b[k] = ctx->pw_t2;
        }
        LV_HIP_CHECK(lv_dots_device(n, xf, b, *cnt, ctx->W.partials, ctx->pw_dots, s));
        for (int k = 0; k < *cnt; k++) c[k] = -1.0;  // -<q_k, t2>
        LV_HIP_CHECK(lv_combine(n, ctx->pw_t2, 1.0, xf, c, *cnt, s, ctx->pw_dots, 1));
        for (int k = 0; k < *cnt; k++) xf[k] = Z[k];
        LV_HIP_CHECK(lv_combine(n, ctx->pw_t1, 1.0, xf, c, *cnt, s, ctx->pw_dots, 1));
    }
    {
        const double *a[1] = {ctx->pw_t2}, *b[1] = {ctx->pw_t2};
        LV_HIP_CHECK(lv_dots_device(n, a, b, 1, ctx->W.partials, ctx->pw_dots, s));
    }
    if (*cnt == w) {  // drop the oldest direction
        float *q0 = Q[0], *z0 = Z[0];
        for (int k = 1; k < w; k++) {
            Q[k - 1] = Q[k];
            Z[k - 1] = Z[k];
        }
        Q[w - 1] = q0;
        Z[w - 1] = z0;
        (*cnt)--;
    }
    xp[0] = ctx->pw_t2;
    c[0] = 1.0;  // 1 / |t2|
    LV_HIP_CHECK(lv_combine(n, Q[*cnt], 0.0, xp, c, 1, s, ctx->pw_dots, 2));
    xp[0] = ctx->pw_t1;
    LV_HIP_CHECK(lv_combine(n, Z[*cnt], 0.0, xp, c, 1, s, ctx->pw_dots, 2));
    (*cnt)++;
    return LV_OK;
}

static int picard_add_direction(lv_ctx *ctx, int w, int *cnt, const double *Pcur, const double *Pprev,
                                const double *APcur, const double *APprev)
{
    const double *xp[2] = {Pcur, Pprev};
    const double c[2] = {1.0, -1.0};
    LV_HIP_CHECK(lv_combine(ctx->n, ctx->pw_t1, 0.0, xp, c, 2, ctx->s));
    xp[0] = APcur;
    xp[1] = APprev;
    LV_HIP_CHECK(lv_combine(ctx->n, ctx->pw_t2, 0.0, xp, c, 2, ctx->s));
    return picard_add_pair(ctx, ctx->pw_q, ctx->pw_z, w, cnt);
}

// P_cur - P_prev into y
static int picard_record(lv_ctx *ctx, double *y, const double *Pcur, const double *Pprev)
{
    const double *xp[2] = {Pcur, Pprev};
    const double c[2] = {1.0, -1.0};
    LV_HIP_CHECK(lv_combine(ctx->n, y, 0.0, xp, c, 2, ctx->s));
    return LV_OK;
}

// P0 += sum_k <q_k, r> z_k over the set (Q, Z): the guess whose residual is r minus its projection
// on the q's.  r is b - AP_cur (APcur given), or pw_t2 as a previous projection left it (APcur null);
// with keep, pw_t2 is left holding the projected residual.
static int picard_project(lv_ctx *ctx, float *const *Q, float *const *Z, int cnt, const double *APcur, bool keep)
{
    const int n = ctx->n;
    hipStream_t s = ctx->s;
    const double *xp[8] = {ctx->b, APcur};
    double c[8] = {1.0, -1.0};
    if (APcur) LV_HIP_CHECK(lv_combine(n, ctx->pw_t2, 0.0, xp, c, 2, s));
    const float *xf[8];
    const double *b[8];
    for (int k = 0; k < cnt; k++) {
        xf[k] = Q[k];
        b[k] = ctx->pw_t2;
        c[k] = 1.0;  // <q_k, r>
    }
    if (cnt == 0) return LV_OK;
    LV_HIP_CHECK(lv_dots_device(n, xf, b, cnt, ctx->W.partials, ctx->pw_dots, s));
    for (int k = 0; k < cnt; k++) xf[k] = Z[k];
    LV_HIP_CHECK(lv_combine(n, ctx->P0, 1.0, xf, c, cnt, s, ctx->pw_dots, 1));
    if (keep) {
        for (int k = 0; k < cnt; k++) {
            xf[k] = Q[k];
            c[k] = -1.0;
        }
        LV_HIP_CHECK(lv_combine(n, ctx->pw_t2, 1.0, xf, c, cnt, s, ctx->pw_dots, 1));
    }
    return LV_OK;
}

// find_pressure!(solver, dt, niter; boundary_velocity), src/pressure.jl:215-224
int lv_find_pressure(lv_ctx *ctx, double dt, int32_t niter, const lv_solve_opts *opts, lv_solve_stats *stats)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    solve_time_scope times{ctx};
    lv_solve_opts o;
    if (opts) o = *opts;
    else default_opts(o);
    if (stats) memset(stats, 0, sizeof(*stats));
    lv_vbc v;
    for (int k = 0; k < 8; k++) v.v[k] = o.vbc[k];
    rc = ensure_rhs_geom(ctx);
    if (rc) return rc;
    // the compact copy of the right-hand side's edge slots past the sixth, as the operator's (not below
    // 65536 rows; amg_layout 2 turns it off with the operator's)
    ctx->rg.ocap = (ctx->n >= 65536 && o.amg_layout != 2) ? LV_OVF_CAP : 0;
    rc = assemble_system(ctx, dt, 0, nullptr, &v);  // refresh!(A), and the fixed parts of refresh!(solver)
    if (rc) return rc;
    const int w = std::min(std::max(o.picard_window, 0), 8);
    if (w > 0) {
        rc = ensure_picard_window(ctx, w);
        if (rc) return rc;
    }
    int cnt = 0, have = 0;  // directions in the window; finished Picard iterates remembered (<= 2)
    // carried directions: the previous call's first two corrections, multiplied by this call's
    // operator and made orthonormal, improve the guesses of Picard solves 2 to 4 by a projection of
    // their own after the window's (the window's directions are not made orthogonal to them, so
    // the later solves' window is the same as without them); this call records its own
    const bool carry = w > 0 && o.picard_carry != 0 && niter >= 4;
    int ccnt = 0;
    if (carry) {
        for (int k = 0; k < 2; k++) {
            double **bufs[2] = {&ctx->pc_z[k], &ctx->pc_new[k]};
            for (double **b : bufs)
                if (!*b) LV_HIP_CHECK(dalloc(b, (size_t)ctx->cap));
            float **fbufs[2] = {&ctx->pc_q[k], &ctx->pc_o[k]};
            for (float **b : fbufs)
                if (!*b) LV_HIP_CHECK(dalloc(b, (size_t)ctx->cap));
        }
        if (ctx->pc_n == ctx->n) {
            for (int k = 0; k < ctx->pc_cnt; k++) {
                const double *xp[1] = {ctx->pc_z[k]};
                const double c1[1] = {1.0};
                LV_HIP_CHECK(lv_combine(ctx->n, ctx->pw_t1, 0.0, xp, c1, 1, ctx->s));
                LV_HIP_CHECK(lv_spmv64(ctx->A, ctx->n, ctx->pc_z[k], ctx->pw_t2, ctx->s));
                rc = picard_add_pair(ctx, ctx->pc_q, ctx->pc_o, 2, &ccnt);
                if (rc) return rc;
            }
        }
    } else {
        ctx->pc_cnt = 0;
    }
    int status = LV_OK;
    for (int it = 1; it <= niter; it++) {
        // refresh!(solver, dt, it > 1, boundary_velocity): the first iteration reads p.P (and takes it
        // as the initial guess); later ones read the previous solution in P0, write it back into the
        // particles (p.P = P, src/pressure.jl:221-223) and add the GP term
        if (it == 1) {
            LV_HIP_CHECK(lv_launch_rhs_b(mesh_view(ctx), ctx->n, 0, ctx->rg, ctx->GP, ctx->F.area, ctx->F.P, ctx->b,
                                         ctx->P0, ctx->s));
        } else {
            // (with the Picard window, also its record of the last iterate, pw_P[0])
            LV_HIP_CHECK(lv_launch_gp(mesh_view(ctx), ctx->n, ctx->rg, ctx->P0, ctx->F.mass, ctx->GP, ctx->F.P,
                                      (w > 0 && have >= 1) ? ctx->pw_P[0] : nullptr, ctx->s));
            LV_HIP_CHECK(lv_launch_rhs_b(mesh_view(ctx), ctx->n, 1, ctx->rg, ctx->GP, ctx->F.area, ctx->P0, ctx->b,
                                         nullptr, ctx->s));
        }
        if (w > 0 && have >= 1) {
            if (have >= 2) {
                // the corrections P_2 - P_1 and P_3 - P_2, recorded for the next call
                if (carry && (it == 3 || it == 4)) {
                    rc = picard_record(ctx, ctx->pc_new[it - 3], ctx->pw_P[0], ctx->pw_P[1]);
                    if (rc) return rc;
                }
                rc = picard_add_direction(ctx, w, &cnt, ctx->pw_P[0], ctx->pw_P[1], ctx->pw_AP[0], ctx->pw_AP[1]);
                if (rc) return rc;
            }
            const bool use_carried = ccnt > 0 && it <= 4;
            if (cnt > 0) {
                rc = picard_project(ctx, ctx->pw_q, ctx->pw_z, cnt, ctx->pw_AP[0], use_carried);
                if (rc) return rc;
            }
            if (use_carried) {
                rc = picard_project(ctx, ctx->pc_q, ctx->pc_o, ccnt, cnt > 0 ? nullptr : ctx->pw_AP[0], false);
                if (rc) return rc;
            }
        }
        rc = solve_on_device(ctx, ctx->b, ctx->P0, o, stats);
        if (rc == LV_E_NOT_CONVERGED) status = rc;
        else if (rc) return rc;
        if (w > 0 && it < niter) {  // remember P_it (copied by the next iteration's k_gp) and A P_it = b' - r
            std::swap(ctx->pw_P[0], ctx->pw_P[1]);
            std::swap(ctx->pw_AP[0], ctx->pw_AP[1]);
            LV_HIP_CHECK(lv_apply_from_residual(ctx->n, ctx->b, ctx->W.r, ctx->pw_AP[0], ctx->W, ctx->s));
            have = std::min(have + 1, 2);
        }
    }
    LV_HIP_CHECK(lv_launch_write_P(ctx->F, ctx->n, ctx->P0, ctx->s));
    if (carry) {  // this call's corrections become the carried ones
        for (int k = 0; k < 2; k++) std::swap(ctx->pc_z[k], ctx->pc_new[k]);
        ctx->pc_cnt = 2;
        ctx->pc_n = ctx->n;
    }
    if (status) lv_fail(ctx, status, "pressure solve did not reach its tolerance");
    return status;
}

int lv_pressure_step(lv_ctx *ctx, double dt)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_pressure_step(ctx->g, ctx->F, mesh_view(ctx), ctx->n, dt, ctx->s));
    return LV_OK;
}

int lv_find_D(lv_ctx *ctx)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_find_D(ctx->g, ctx->F, mesh_view(ctx), ctx->n, ctx->s));
    return LV_OK;
}

int lv_viscous_step(lv_ctx *ctx, double dt, int32_t artificial_viscosity)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    const double avdr = artificial_viscosity ? ctx->g.dr : 0.0;  // src/diffusion.jl:40
    LV_HIP_CHECK(lv_launch_viscous(ctx->g, ctx->F, mesh_view(ctx), ctx->n, dt, avdr, ctx->s));
    return LV_OK;
}

int lv_bdary_friction(lv_ctx *ctx, const double vbc[8], double dt)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    lv_vbc v;
    for (int k = 0; k < 8; k++) v.v[k] = vbc ? vbc[k] : 0.0;
    LV_HIP_CHECK(lv_launch_friction(ctx->F, mesh_view(ctx), ctx->n, v, dt, ctx->s));
    return LV_OK;
}

int lv_find_dv(lv_ctx *ctx, double dt, double alpha)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_find_dv(ctx->g, ctx->F, mesh_view(ctx), ctx->n, dt, alpha, ctx->s));
    return LV_OK;
}

int lv_relaxation_step(lv_ctx *ctx, double dt, int32_t rusanov)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    LV_HIP_CHECK(lv_launch_relaxation(ctx->g, ctx->F, mesh_view(ctx), ctx->n, dt, rusanov, ctx->s));
    ctx->mesh_valid = false;
    return lv_remesh(ctx);  // src/relaxation.jl:72
}

// ---- implicit viscosity (src/viscous_solver.jl:43-128) ------------------------------------------

// viscous_step!(solver::ViscousSolver, dt): the operator once, then one solve per velocity component
// (the reference's 2n x 2n system is two copies of the same n x n block), each from the initial
// guess v and right-hand side (m/dt) v; v is overwritten by the solution
int lv_viscous_implicit(lv_ctx *ctx, double dt, const lv_solve_opts *opts, lv_solve_stats *stats)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    solve_time_scope times{ctx};
    lv_solve_opts o;
    if (opts) {
        o = *opts;
    } else {
        default_opts(o);
        o.method = LV_SOLVER_MIXED;
        o.tol = 1e-6;   // minres!(...; atol = 1e-6, rtol = 1e-6, itmax = 2000), src/viscous_solver.jl:122-126
        o.itmax = 2000;
    }
    rc = assemble_system(ctx, dt, 1, nullptr);
    if (rc) return rc;
    const int n = ctx->n;
    int status = LV_OK;
    for (int c = 0; c < 2; c++) {
        double *vc = ctx->F.v + (size_t)c * n;
        LV_HIP_CHECK(lv_launch_visc_rhs(n, ctx->F.mass, vc, dt, ctx->b, ctx->P0, ctx->s));
        rc = solve_on_device(ctx, ctx->b, ctx->P0, o, stats);
        if (rc == LV_E_NOT_CONVERGED) status = rc;
        else if (rc) return rc;
        LV_HIP_CHECK(hipMemcpyAsync(vc, ctx->P0, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->s));
    }
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    if (status) lv_fail(ctx, status, "implicit viscous solve did not reach its tolerance");
    return status;
}

// ---- multiphase projection (src/relaxation.jl:75-206) -------------------------------------------

static int mph_ready(lv_ctx *ctx)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (!ctx->mph) LV_HIP_CHECK(lv_mph_alloc(&ctx->mph, ctx->cap));
    return LV_OK;
}

int lv_multiphase_projection(lv_ctx *ctx, double quality_threshold, double atol, double rtol, int32_t itmax,
                             int32_t *iters, double *relres)
{
    int rc = mph_ready(ctx);
    if (rc) return rc;
    if (itmax < 1 || !(atol >= 0.0) || !(rtol >= 0.0)) return lv_fail(ctx, LV_E_INVALID, "multiphase options");
    int it = 0;
    double rel = 0.0;
    LV_HIP_CHECK(lv_mph_projection(ctx->mph, ctx->g, ctx->F, mesh_view(ctx), ctx->n, quality_threshold, atol, rtol,
                                   itmax, &it, &rel, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    if (iters) *iters = it;
    if (relres) *relres = rel;
    return LV_OK;
}

int lv_multiphase_rhs(lv_ctx *ctx, double *b)
{
    int rc = mph_ready(ctx);
    if (rc) return rc;
    if (!b) return LV_E_INVALID;
    LV_HIP_CHECK(lv_mph_rhs(ctx->mph, ctx->g, ctx->F, mesh_view(ctx), ctx->n, ctx->W.z, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(b, ctx->W.z, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_multiphase_apply(lv_ctx *ctx, const double *x, double *y)
{
    int rc = mph_ready(ctx);
    if (rc) return rc;
    if (!x || !y) return LV_E_INVALID;
    const size_t bytes = sizeof(double) * ctx->n;
    LV_HIP_CHECK(hipMemcpyAsync(ctx->W.r, x, bytes, hipMemcpyHostToDevice, ctx->s));
    LV_HIP_CHECK(lv_mph_apply(ctx->mph, ctx->g, ctx->F, mesh_view(ctx), ctx->n, ctx->W.r, ctx->W.z, ctx->s));
    LV_HIP_CHECK(hipMemcpyAsync(y, ctx->W.z, bytes, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

int lv_step_fused(lv_ctx *ctx, const lv_step_desc *d, lv_solve_stats *stats)
{
    if (!d) return LV_E_INVALID;
    int rc = lv_move(ctx, d->dt);
    if (rc) return rc;
    if (d->gravity) {
        rc = lv_gravity_step(ctx, d->g, d->dt);
        if (rc) return rc;
    }
    if (d->eos == 1) rc = lv_eos_stiffened(ctx, d->gamma, d->P0);
    else if (d->eos == 2) rc = lv_eos_ideal(ctx, d->gamma, d->Pmin);
    else if (d->eos == 3) rc = lv_find_rho(ctx);
    if (rc) return rc;
    int prc = lv_find_pressure(ctx, d->dt, d->niter, &d->solve, stats);
    if (prc && prc != LV_E_NOT_CONVERGED) return prc;
    // pressure_step! with find_D! in its second pass
    LV_HIP_CHECK(lv_launch_pressure_step(ctx->g, ctx->F, mesh_view(ctx), ctx->n, d->dt, ctx->s, true));
    bool dv_done = false;
    if (d->viscous == 2) {  // viscous_step!(ViscousSolver(grid), dt) with the step's solver method
        lv_solve_opts o = d->solve;
        o.tol = d->viscous_tol > 0.0 ? d->viscous_tol : 1e-6;
        o.itmax = 2000;
        rc = lv_viscous_implicit(ctx, d->dt, &o, stats);
        if (rc && rc != LV_E_NOT_CONVERGED) return rc;
        if (rc) prc = rc;
    } else if (d->viscous) {  // viscous_step! with find_dv! in its second pass
        const double avdr = d->artificial_viscosity ? ctx->g.dr : 0.0;  // src/diffusion.jl:40
        LV_HIP_CHECK(lv_launch_viscous(ctx->g, ctx->F, mesh_view(ctx), ctx->n, d->dt, avdr, ctx->s, true, d->dt,
                                       d->alpha));
        dv_done = true;
    }
    if (d->friction) {
        rc = lv_bdary_friction(ctx, d->friction_vbc, d->dt);
        if (rc) return rc;
    }
    if (!dv_done) {
        rc = lv_find_dv(ctx, d->dt, d->alpha);
        if (rc) return rc;
    }
    if (d->multiphase) {  // multiphase_projection!(MultiphaseSolver(grid; quality_threshold)), defaults of :197
        int it = 0;
        double rel = 0.0;
        rc = lv_multiphase_projection(ctx, d->multiphase_quality, d->multiphase_tol, d->multiphase_tol,
                                      d->multiphase_itmax > 0 ? d->multiphase_itmax : 200, &it, &rel);
        if (rc) return rc;
        if (stats) {
            stats->multiphase_iters = it;
            stats->multiphase_relres = rel;
        }
    }
    rc = lv_relaxation_step(ctx, d->dt, d->rusanov);
    if (rc) return rc;
    return prc;
}

// ---- learned SPAI preconditioner ---------------------------------------------------------------

int lv_spai_num_params(int64_t *count)
{
    if (!count) return LV_E_INVALID;
    *count = lv_spai_param_count();
    return LV_OK;
}

int lv_spai_set_params(lv_ctx *ctx, const float *params)
{
    if (!ctx || !ctx->spai_ok) return LV_E_INVALID;
    if (!params) {  // unload: LV_SOLVER_SPAI falls back to Jacobi, LV_SOLVER_AMG smooths with Jacobi
        ctx->spai.has_params = 0;
        ctx->spai_emitted_for = -1;
        return LV_OK;
    }
    LV_HIP_CHECK(lv_spai_set_params(&ctx->spai, params, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    ctx->spai_emitted_for = -1;
    return LV_OK;
}

int lv_spai_get_params(lv_ctx *ctx, float *params)
{
    if (!ctx || !params || !ctx->spai_ok) return LV_E_INVALID;
    LV_HIP_CHECK(lv_spai_get_params(&ctx->spai, params, ctx->s));
    return LV_OK;
}

static int spai_check(lv_ctx *ctx)
{
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (!ctx->spai_ok || !ctx->spai.has_params) return lv_fail(ctx, LV_E_INVALID, "SPAI parameters not set");
    if (!ctx->A.valid) return lv_fail(ctx, LV_E_INVALID, "pressure operator not assembled");
    return LV_OK;
}

int lv_spai_loss_grad(lv_ctx *ctx, int32_t nprobe, uint64_t seed, int32_t use_data, double *loss, float *grad)
{
    int rc = spai_check(ctx);
    if (rc) return rc;
    if (nprobe < 1 || nprobe > 64) return LV_E_INVALID;
    LV_HIP_CHECK(lv_spai_forward(&ctx->spai, ctx->A, ctx->n, ctx->s));
    ctx->spai_emitted_for = -1;
    double L = 0.0;
    LV_HIP_CHECK(lv_spai_loss_grad(&ctx->spai, ctx->A, ctx->n, use_data ? ctx->F.P : nullptr, nprobe, seed,
                                   grad != nullptr, &L, ctx->s));
    if (grad)
        LV_HIP_CHECK(hipMemcpyAsync(grad, ctx->spai.grad, sizeof(float) * lv_spai_param_count(),
                                    hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    if (loss) *loss = L;
    return LV_OK;
}

int lv_spai_train_step(lv_ctx *ctx, double lr, int32_t nprobe, uint64_t seed, double *loss)
{
    int rc = spai_check(ctx);
    if (rc) return rc;
    if (nprobe < 1 || nprobe > 64) return LV_E_INVALID;
    LV_HIP_CHECK(lv_spai_forward(&ctx->spai, ctx->A, ctx->n, ctx->s));
    double L = 0.0;
    LV_HIP_CHECK(lv_spai_loss_grad(&ctx->spai, ctx->A, ctx->n, ctx->F.P, nprobe, seed, true, &L, ctx->s));
    LV_HIP_CHECK(lv_spai_adam(&ctx->spai, lr, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    ctx->spai_emitted_for = -1;
    if (loss) *loss = L;
    return LV_OK;
}

int lv_spai_outputs(lv_ctx *ctx, float *o_edge, float *o_node, float *w, float *mdiag)
{
    int rc = spai_check(ctx);
    if (rc) return rc;
    LV_HIP_CHECK(lv_spai_infer(&ctx->spai, ctx->A, ctx->n, ctx->s));
    ctx->spai_emitted_for = -1;
    const size_t E = (size_t)ctx->A.nnz, n = (size_t)ctx->n;
    if (o_edge) LV_HIP_CHECK(hipMemcpyAsync(o_edge, ctx->spai.oe, sizeof(float) * E, hipMemcpyDeviceToHost, ctx->s));
    if (o_node) LV_HIP_CHECK(hipMemcpyAsync(o_node, ctx->spai.on, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->s));
    if (w) LV_HIP_CHECK(hipMemcpyAsync(w, ctx->spai.w, sizeof(float) * E, hipMemcpyDeviceToHost, ctx->s));
    if (mdiag) LV_HIP_CHECK(hipMemcpyAsync(mdiag, ctx->spai.mdiag, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->s));
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    return LV_OK;
}

// ---- timing ------------------------------------------------------------------------------------

int lv_timing_enable(lv_ctx *ctx, int32_t kernel_id, int32_t on)
{
    if (!ctx || kernel_id < 0 || kernel_id >= LV_NKTIMERS) return LV_E_INVALID;
    lv_timer *t = ctx->timers[kernel_id];
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    lv_timer_flush(t);
    t->on = on ? 1 : 0;
    t->every = on > 1 ? on : 1;
    t->seen = 0;
    t->pending = 0;
    t->total_ms = 0.0;
    t->launches = 0;
    return LV_OK;
}

int lv_timing_seen(lv_ctx *ctx, int32_t kernel_id, int64_t *seen)
{
    if (!ctx || kernel_id < 0 || kernel_id >= LV_NKTIMERS || !seen) return LV_E_INVALID;
    *seen = ctx->timers[kernel_id]->seen;
    return LV_OK;
}

int lv_timing_read(lv_ctx *ctx, int32_t kernel_id, double *total_ms, int64_t *launches, double *bytes_per_launch)
{
    if (!ctx || kernel_id < 0 || kernel_id >= LV_NKTIMERS) return LV_E_INVALID;
    lv_timer *t = ctx->timers[kernel_id];
    LV_HIP_CHECK(hipStreamSynchronize(ctx->s));
    lv_timer_flush(t);
    if (total_ms) *total_ms = t->total_ms;
    if (launches) *launches = t->launches;
    if (bytes_per_launch) {
        // algorithmic (compulsory) bytes of one launch at the current sizes
        const double n = ctx->n, nnz = (double)ctx->A.nnz, m = ctx->nitems;
        const double nc = (double)ctx->g.n1 * ctx->g.n2;
        double b = 0.0;
        // CG SpMV kernel: deg + diag + z + p (read, write) + q (read, write) per row, col + value per
        // nonzero (aggregate order: 1-byte degree, 2-byte column offset)
        if (kernel_id == 0) b = ctx->spmv32_agg ? 1.0 * n + 4.0 * n * 6 + 6.0 * nnz : 4.0 * n + 4.0 * n * 6 + 8.0 * nnz;
        else if (kernel_id == 1) b = 4.0 * n + 8.0 * n * 6 + 12.0 * nnz;
        else if (kernel_id == 2) b = 16.0 * n + 20.0 * m + 4.0 * (nc + 1) + 4.0 * n + 36.0 * 6.0 * n + 24.0 * n;
        else if (kernel_id == 3) b = 3.0 * (8.0 + 8.0 + 4.0) * m;              // 3 radix passes over (key, label)
        *bytes_per_launch = b;
    }
    return LV_OK;
}

}  // extern "C"
