/*
 * lvoracle.c -- CPU FP64 restatement of LagrangianVoronoi.jl's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity tests, the smoke check and the CPU
 * baseline leg of the benchmark call it to check or time the product.  The product library never
 * links it, and nothing here is ever the thing that ships.
 *
 * Every function restates one reference function, cited as file:line of OndrejKincl/LagrangianVoronoi.jl
 * v0.1.0, with Julia's left-to-right evaluation order and no FMA contraction (build with
 * -ffp-contract=off).  Labels are 0-based; boundary labels stay -1..-4 (src/polygon.jl:4-7), so a
 * label < 0 marks a boundary edge.  Vector fields are two planes: comp k of particle i is a[k*n + i].
 * The 2x2 tensor D is four planes in Julia's column-major order (D11, D21, D12, D22).
 *
 * Where the reference leans on a library whose rounding cannot be pinned offline, the choice is
 * stated next to the code: StaticArrays `norm` is taken as sqrt(a1*a1 + a2*a2) and LinearAlgebra's
 * three-argument `s*M*v` as M*(v*s) (ULP-level only).  Krylov.jl MINRES is replaced by a tight
 * Jacobi-PCG solve of the same linear system (src/pressure.jl:219).
 *
 * Parity unpinned (no offline check can fix these; each is named where it is used):
 *   - Krylov.jl 0.9.8 minres! iterates at the rounding level: lvo_minres restates the same
 *     Paige-Saunders recurrences and matches SciPy's minres iterate by iterate
 *     (tests/test_oracle_pins.py), but Krylov.jl itself is not here to compare with; the parity
 *     target of find_pressure! is the solution of the same system to the solve tolerance;
 *   - StaticArrays 1.9.8 norm/dot rounding (ULP level; absorbed by the 1e-12 bounds);
 *   - the reference's Polyester-threaded insertion order (nondeterministic); the serial order
 *     (ascending label per cell) is restated instead;
 *   - Julia's sort! stability for magic-path ties: stable is assumed (Julia 1.10 default), and
 *     the resulting path is pinned against a brute-force stable sort of every offset.
 * Everything else is pinned by the tests under tests/ (brute force, Qhull, closed forms,
 * conservation, the reference's own Taylor-Green thresholds).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LV_OK 0
#define LV_E_MESH_DESTROYED 1
#define LV_E_NAN_VELOCITY 2
#define LV_E_POLY_OVERFLOW 3
#define LV_E_NOT_CONVERGED 4
#define LV_E_INVALID 6

#define BDARY_UP (-1)
#define BDARY_RIGHT (-2)
#define BDARY_DOWN (-3)
#define BDARY_LEFT (-4)

/* SIGNUM_EPS = 2*eps(Float64), src/polygon.jl:2 */
static const double SIGNUM_EPS = 2.0 * 2.220446049250313e-16;

typedef struct {
    double bxmin[2], bxmax[2]; /* boundary_rect */
    double dr, h, r_max;
    int xperiodic, yperiodic;
} lvo_desc;

typedef struct {
    lvo_desc d;
    double rr_max, xperiod, yperiod;
    double cmin[2], cmax[2]; /* cropping_rect */
    double origin[2];
    int n1, n2;
    int npath;     /* magic-path nodes with rr <= rr_max, in path order */
    int has_next;  /* does a node with rr > rr_max exist on the path? */
    double next_rr;
    int *pi1, *pi2;
    double *prr;
} lvo_grid;

typedef struct {
    int n, maxe;
    double *x, *v, *dv, *mom;                      /* [2][n] */
    double *rho, *e, *P, *c2, *mass, *energy, *quality, *mu; /* [n] */
    double *D;                                     /* [4][n] */
    int *phase;                                    /* [n] */
    int *ne;                                       /* [n] edge counts */
    double *ev;                                    /* [n][maxe][4] = v1x v1y v2x v2y */
    int *el;                                       /* [n][maxe] labels */
} lvo_state;

typedef struct {
    double v1x, v1y, v2x, v2y;
    int label;
} edge_t;

/* ------------------------------------------------------------------------------------------------
 * Grid set-up: VoronoiGrid ctor (src/voronoigrid.jl:26-49) and CellList ctor (src/neighborlist.jl:18-43)
 * ---------------------------------------------------------------------------------------------- */

typedef struct {
    int i1, i2;
    long long s;   /* i1^2 + i2^2, first sort key (src/neighborlist.jl:40) */
    long long gen; /* generation order; Julia's default sort! is stable */
    double rr;
} pnode;

static int pnode_cmp(const void *a, const void *b)
{
    const pnode *p = (const pnode *)a, *q = (const pnode *)b;
    if (p->rr < q->rr) return -1;
    if (p->rr > q->rr) return 1;
    if (p->s < q->s) return -1;
    if (p->s > q->s) return 1;
    return (p->gen < q->gen) ? -1 : (p->gen > q->gen);
}

/* magic_path rr = h^2*(max(0,|i1|-1)^2 + max(0,|i2|-1)^2), src/neighborlist.jl:34 */
static double path_rr(double h, int i1, int i2)
{
    long long a = abs(i1) - 1, b = abs(i2) - 1;
    if (a < 0) a = 0;
    if (b < 0) b = 0;
    return (h * h) * (double)(a * a + b * b);
}

void lvo_grid_free(lvo_grid *g)
{
    if (!g) return;
    free(g->pi1);
    free(g->pi2);
    free(g->prr);
    free(g);
}

lvo_grid *lvo_grid_new(const lvo_desc *d)
{
    if (!(d->h > 0.0)) return NULL; /* "h must be positive", src/neighborlist.jl:19-21 */
    lvo_grid *g = (lvo_grid *)calloc(1, sizeof(lvo_grid));
    g->d = *d;
    /* src/voronoigrid.jl:29-33: r_max*periodic is Float64*Bool (r_max or +0.0) */
    double rx = d->xperiodic ? d->r_max : 0.0;
    double ry = d->yperiodic ? d->r_max : 0.0;
    g->cmin[0] = (d->bxmin[0] - rx * 1.0) - ry * 0.0;
    g->cmin[1] = (d->bxmin[1] - rx * 0.0) - ry * 1.0;
    g->cmax[0] = (d->bxmax[0] + rx * 1.0) + ry * 0.0;
    g->cmax[1] = (d->bxmax[1] + rx * 0.0) + ry * 1.0;
    g->xperiod = d->bxmax[0] - d->bxmin[0];
    g->yperiod = d->bxmax[1] - d->bxmin[1];
    g->rr_max = d->r_max * d->r_max;
    /* src/neighborlist.jl:23-25 */
    g->origin[0] = g->cmin[0] - d->h;
    g->origin[1] = g->cmin[1] - d->h;
    g->n1 = (int)floor((g->cmax[0] - g->cmin[0]) / d->h) + 3;
    g->n2 = (int)floor((g->cmax[1] - g->cmin[1]) / d->h) + 3;
    /* Path generation bounds swap n1/n2 exactly as src/neighborlist.jl:30-31 does:
     * i2 in (1-n1):(n1-1) (outer), i1 in (1-n2):(n2-1) (inner). Only the prefix with rr <= rr_max
     * and the first node after it are ever read, so only nodes with small |i| are enumerated. */
    int lim1 = g->n2 - 1, lim2 = g->n1 - 1; /* |i1| <= lim1, |i2| <= lim2 */
    int kmax = (int)ceil(d->r_max / d->h) + 3;
    int m1 = kmax < lim1 ? kmax : lim1;
    int m2 = kmax < lim2 ? kmax : lim2;
    int cap = (2 * m1 + 1) * (2 * m2 + 1);
    pnode *nodes = (pnode *)malloc(sizeof(pnode) * (size_t)cap);
    int cnt = 0;
    for (int i2 = -m2; i2 <= m2; i2++)
        for (int i1 = -m1; i1 <= m1; i1++) {
            pnode *p = &nodes[cnt++];
            p->i1 = i1;
            p->i2 = i2;
            p->s = (long long)i1 * i1 + (long long)i2 * i2;
            p->gen = (long long)(i2 - (1 - g->n1)) * (2LL * g->n2 - 1) + (i1 - (1 - g->n2));
            p->rr = path_rr(d->h, i1, i2);
        }
    qsort(nodes, (size_t)cnt, sizeof(pnode), pnode_cmp);
    int np = 0;
    while (np < cnt && !(nodes[np].rr > g->rr_max)) np++;
    g->npath = np;
    g->pi1 = (int *)malloc(sizeof(int) * (size_t)(np + 1));
    g->pi2 = (int *)malloc(sizeof(int) * (size_t)(np + 1));
    g->prr = (double *)malloc(sizeof(double) * (size_t)(np + 1));
    for (int k = 0; k < np; k++) {
        g->pi1[k] = nodes[k].i1;
        g->pi2[k] = nodes[k].i2;
        g->prr[k] = nodes[k].rr;
    }
    /* The first node beyond rr_max: the smallest rr among in-bound nodes with rr > rr_max.  If
     * enumeration was clipped by the grid bounds in every direction, no such node exists. */
    g->has_next = 0;
    g->next_rr = INFINITY;
    if (np < cnt) {
        g->has_next = 1;
        g->next_rr = nodes[np].rr;
    } else if (m1 < lim1 || m2 < lim2) {
        /* unreachable with kmax margin of 3 cells, kept for completeness */
        g->has_next = 1;
        g->next_rr = path_rr(d->h, m1 + 1, 0);
    }
    free(nodes);
    return g;
}

int lvo_grid_info(const lvo_grid *g, int *n1, int *n2, int *npath, double *out6)
{
    *n1 = g->n1;
    *n2 = g->n2;
    *npath = g->npath;
    out6[0] = g->cmin[0];
    out6[1] = g->cmin[1];
    out6[2] = g->cmax[0];
    out6[3] = g->cmax[1];
    out6[4] = g->has_next ? g->next_rr : -1.0;
    out6[5] = g->rr_max;
    return 0;
}

int lvo_grid_path(const lvo_grid *g, int *i1, int *i2, double *rr)
{
    for (int k = 0; k < g->npath; k++) {
        i1[k] = g->pi1[k];
        i2[k] = g->pi2[k];
        rr[k] = g->prr[k];
    }
    return g->npath;
}

/* ------------------------------------------------------------------------------------------------
 * Cell list: findkey/insert!/insert_periodic! (src/neighborlist.jl:47-67, src/voronoigrid.jl:130-147)
 * ---------------------------------------------------------------------------------------------- */

/* findkey, src/neighborlist.jl:47-52: returns 0-based linear cell id or -1 if out of the grid */
static long long findcell(const lvo_grid *g, double x1, double x2, int *k1, int *k2)
{
    double a = x1 - g->origin[0];
    double b = x2 - g->origin[1];
    double fa = floor(a / g->d.h), fb = floor(b / g->d.h);
    long long i1 = (long long)fa + 1, i2 = (long long)fb + 1;
    if (k1) *k1 = (int)i1;
    if (k2) *k2 = (int)i2;
    if (!(fa == fa) || !(fb == fb)) return -1;
    if (i1 < 1 || i1 > g->n1 || i2 < 1 || i2 > g->n2) return -1;
    return (i1 - 1) + (i2 - 1) * (long long)g->n1;
}

/* The (up to 9) image positions inserted for one seed, in insert order (src/voronoigrid.jl:130-147). */
static int image_positions(const lvo_grid *g, double x1, double x2, double *px, double *py)
{
    double P = g->xperiod, Q = g->yperiod;
    int k = 0;
    px[k] = x1; py[k] = x2; k++;
    if (g->d.xperiodic) {
        px[k] = x1 + P * 1.0; py[k] = x2 + P * 0.0; k++;
        px[k] = x1 - P * 1.0; py[k] = x2 - P * 0.0; k++;
    }
    if (g->d.yperiodic) {
        px[k] = x1 + Q * 0.0; py[k] = x2 + Q * 1.0; k++;
        px[k] = x1 - Q * 0.0; py[k] = x2 - Q * 1.0; k++;
    }
    if (g->d.xperiodic && g->d.yperiodic) {
        px[k] = (x1 + P * 1.0) + Q * 0.0; py[k] = (x2 + P * 0.0) + Q * 1.0; k++;
        px[k] = (x1 + P * 1.0) - Q * 0.0; py[k] = (x2 + P * 0.0) - Q * 1.0; k++;
        px[k] = (x1 - P * 1.0) + Q * 0.0; py[k] = (x2 - P * 0.0) + Q * 1.0; k++;
        px[k] = (x1 - P * 1.0) - Q * 0.0; py[k] = (x2 - P * 0.0) - Q * 1.0; k++;
    }
    return k;
}

/* Build the cell list in serial insertion order (julia -t 1): cell c holds labels
 * cell_labels[cell_start[c] .. cell_start[c+1]).  Returns the number of inserts, or -1 if cap is short. */
long long lvo_cell_list(const lvo_grid *g, int n, const double *x, int *cell_start, int *cell_labels, long long cap)
{
    long long nc = (long long)g->n1 * g->n2;
    memset(cell_start, 0, sizeof(int) * (size_t)(nc + 1));
    double px[9], py[9];
    long long total = 0;
    for (int i = 0; i < n; i++) {
        int k = image_positions(g, x[i], x[n + i], px, py);
        for (int j = 0; j < k; j++) {
            long long c = findcell(g, px[j], py[j], NULL, NULL);
            if (c >= 0) {
                cell_start[c + 1]++;
                total++;
            }
        }
    }
    if (total > cap) return -1;
    for (long long c = 0; c < nc; c++) cell_start[c + 1] += cell_start[c];
    int *fill = (int *)calloc((size_t)nc, sizeof(int));
    for (int i = 0; i < n; i++) {
        int k = image_positions(g, x[i], x[n + i], px, py);
        for (int j = 0; j < k; j++) {
            long long c = findcell(g, px[j], py[j], NULL, NULL);
            if (c >= 0) cell_labels[cell_start[c] + fill[c]++] = i;
        }
    }
    free(fill);
    return total;
}

/* ------------------------------------------------------------------------------------------------
 * Polygon primitives (src/polygon.jl)
 * ---------------------------------------------------------------------------------------------- */

static inline int signum(double x) /* src/polygon.jl:9-16 */
{
    if (x < -SIGNUM_EPS) return -1;
    if (x > SIGNUM_EPS) return 1;
    return 0;
}

/* reset!, src/polygon.jl:37-47 */
static int poly_reset(edge_t *E, const double *cmin, const double *cmax)
{
    double Ax = cmin[0], Ay = cmin[1], Cx = cmax[0], Cy = cmax[1];
    double Bx = Cx, By = Ay, Dx = Ax, Dy = Cy;
    E[0] = (edge_t){Bx, By, Ax, Ay, BDARY_DOWN};
    E[1] = (edge_t){Ax, Ay, Dx, Dy, BDARY_LEFT};
    E[2] = (edge_t){Dx, Dy, Cx, Cy, BDARY_UP};
    E[3] = (edge_t){Cx, Cy, Bx, By, BDARY_RIGHT};
    return 4;
}

/* voronoicut!(p, y, label), src/polygon.jl:51-97.  Returns 1 if a new edge was added, 0 if not,
 * -1 on capacity overflow. */
static int poly_cut(edge_t *E, int *len, int cap, double x1, double x2, double y1, double y2, int label)
{
    double d1 = y1 - x1, d2 = y2 - x2;
    double m1 = 0.5 * (y1 + x1), m2 = 0.5 * (y2 + x2);
    double c = d1 * m1 + d2 * m2;
    int i = 0;
    double X1 = NAN, X2 = NAN, Y1 = NAN, Y2 = NAN;
    while (i < *len) {
        edge_t e = E[i];
        double f1 = (d1 * e.v1x + d2 * e.v1y) - c;
        double f2 = (d1 * e.v2x + d2 * e.v2y) - c;
        int s1 = signum(f1), s2 = signum(f2);
        int s12 = s1 + s2;
        if ((0 <= s12 && s12 <= 1) && ((s1 | s2) != 0)) {
            Y1 = X1;
            Y2 = X2;
            double t = 1.0 / (f1 - f2);
            X1 = t * (f1 * e.v2x - f2 * e.v1x);
            X2 = t * (f1 * e.v2y - f2 * e.v1y);
            if (s1 == 0) { X1 = e.v1x; X2 = e.v1y; }
            if (s2 == 0) { X1 = e.v2x; X2 = e.v2y; }
            if (s1 == 1) E[i] = (edge_t){X1, X2, e.v2x, e.v2y, e.label};
            else E[i] = (edge_t){e.v1x, e.v1y, X1, X2, e.label};
        }
        if (1 <= s12) { /* deleteat!, src/fastvector.jl:44-50 */
            if (*len - 1 > i) E[i] = E[*len - 1];
            (*len)--;
        } else {
            i++;
        }
    }
    if (!(isnan(Y1) && isnan(Y2)) && (X1 != Y1 || X2 != Y2)) {
        edge_t ne = {X1, X2, Y1, Y2, label};
        double cr = (Y1 - X1) * (x2 - X2) - (Y2 - X2) * (x1 - X1); /* cross2(Y - X, x - X) */
        if (cr > 0.0) ne = (edge_t){Y1, Y2, X1, X2, label};
        if (*len >= cap) return -1;
        E[(*len)++] = ne;
        return 1;
    }
    return 0;
}

/* influence_rr, src/polygon.jl:101-107 */
static double poly_influence_rr(const edge_t *E, int len, double x1, double x2)
{
    double rr = 0.0;
    for (int k = 0; k < len; k++) {
        double a = E[k].v1x - x1, b = E[k].v1y - x2;
        double t = 4.0 * (a * a + b * b);
        if (t > rr) rr = t;
    }
    return rr;
}

/* sort_edges!, src/IO.jl:35-48 */
static void poly_sort_edges(edge_t *E, int len)
{
    for (int i = 0; i < len; i++) {
        double lx = E[i].v2x, ly = E[i].v2y;
        for (int j = i + 1; j < len; j++) {
            if (lx == E[j].v1x && ly == E[j].v1y) {
                edge_t t = E[i + 1];
                E[i + 1] = E[j];
                E[j] = t;
                break;
            }
        }
    }
}

/* get_arrow, src/voronoigrid.jl:116-125: y = x + get_arrow(q, x) */
static inline void arrow_target(const lvo_grid *g, double q1, double q2, double x1, double x2, double *y1, double *y2)
{
    double v1 = q1 - x1, v2 = q2 - x2;
    if (g->d.xperiodic && fabs(v1) > 0.5 * g->xperiod) {
        double s = (v1 > 0) ? 1.0 : -1.0;
        double sp = s * g->xperiod;
        v1 = v1 - sp * 1.0;
        v2 = v2 - sp * 0.0;
    }
    if (g->d.yperiodic && fabs(v2) > 0.5 * g->yperiod) {
        double s = (v2 > 0) ? 1.0 : -1.0;
        double sp = s * g->yperiod;
        v1 = v1 - sp * 0.0;
        v2 = v2 - sp * 1.0;
    }
    *y1 = x1 + v1;
    *y2 = x2 + v2;
}

#define ORACLE_EDGE_CAP 256

/* voronoicut!(grid, poly), src/voronoigrid.jl:53-81, plus sort_edges!.  E must hold the reset polygon. */
static int clip_polygon(const lvo_grid *g, int n, const double *x, const int *cell_start, const int *cell_labels,
                        int i, edge_t *E, int *len)
{
    double x1 = x[i], x2 = x[n + i];
    double prr = poly_influence_rr(E, *len, x1, x2);
    int k1, k2;
    findcell(g, x1, x2, &k1, &k2);
    for (int k = 0; k <= g->npath; k++) {
        double rr;
        int o1, o2;
        if (k < g->npath) {
            rr = g->prr[k];
            o1 = g->pi1[k];
            o2 = g->pi2[k];
        } else {
            if (!g->has_next) break;
            if (g->next_rr > prr) break;
            return LV_E_MESH_DESTROYED; /* src/voronoigrid.jl:63-65 */
        }
        if (rr > prr) break;
        long long c1 = (long long)k1 + o1, c2 = (long long)k2 + o2;
        if (c1 < 1 || c1 > g->n1 || c2 < 1 || c2 > g->n2) continue;
        long long c = (c1 - 1) + (c2 - 1) * (long long)g->n1;
        for (int s = cell_start[c]; s < cell_start[c + 1]; s++) {
            int j = cell_labels[s];
            double y1, y2;
            arrow_target(g, x[j], x[n + j], x1, x2, &y1, &y2);
            double a = x1 - y1, b = x2 - y2;
            if ((x1 == y1 && x2 == y2) || (a * a + b * b > prr)) continue;
            int r = poly_cut(E, len, ORACLE_EDGE_CAP, x1, x2, y1, y2, j);
            if (r < 0) return LV_E_POLY_OVERFLOW;
            if (r) prr = poly_influence_rr(E, *len, x1, x2);
        }
    }
    poly_sort_edges(E, *len);
    return LV_OK;
}

/* remesh!, src/voronoigrid.jl:89-108.  Output: edges of polygon i in ev/el slots [i*maxe, i*maxe+ne[i]). */
int lvo_remesh(const lvo_grid *g, lvo_state *s)
{
    int n = s->n;
    long long cap = 9LL * n + 16;
    long long nc = (long long)g->n1 * g->n2;
    int *cell_start = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
    int *cell_labels = (int *)malloc(sizeof(int) * (size_t)cap);
    long long tot = lvo_cell_list(g, n, s->x, cell_start, cell_labels, cap);
    int err = (tot < 0) ? LV_E_INVALID : LV_OK;
    if (err == LV_OK) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            edge_t E[ORACLE_EDGE_CAP];
            int len = poly_reset(E, g->cmin, g->cmax);
            int r = clip_polygon(g, n, s->x, cell_start, cell_labels, i, E, &len);
            if (r == LV_OK && len > s->maxe) r = LV_E_POLY_OVERFLOW;
            if (r != LV_OK) {
#pragma omp critical
                { if (err == LV_OK || r < err) err = r; }
                s->ne[i] = 0;
                continue;
            }
            s->ne[i] = len;
            for (int k = 0; k < len; k++) {
                double *o = &s->ev[((size_t)i * s->maxe + k) * 4];
                o[0] = E[k].v1x; o[1] = E[k].v1y; o[2] = E[k].v2x; o[3] = E[k].v2y;
                s->el[(size_t)i * s->maxe + k] = E[k].label;
            }
        }
    }
    free(cell_start);
    free(cell_labels);
    return err;
}

/* ------------------------------------------------------------------------------------------------
 * Geometry read-outs (src/polygon.jl, src/geometry.jl)
 * ---------------------------------------------------------------------------------------------- */

#define EV(s, i, k) (&(s)->ev[((size_t)(i) * (s)->maxe + (k)) * 4])
#define EL(s, i, k) ((s)->el[(size_t)(i) * (s)->maxe + (k)])

/* area(p), src/polygon.jl:114-122 */
static double poly_area(const lvo_state *s, int i)
{
    int n = s->n;
    double x1 = s->x[i], x2 = s->x[n + i];
    double A = 0.0;
    for (int k = 0; k < s->ne[i]; k++) {
        const double *e = EV(s, i, k);
        double a1 = e[0] - x1, a2 = e[1] - x2, b1 = e[2] - x1, b2 = e[3] - x2;
        A += 0.5 * fabs(a1 * b2 - a2 * b1);
    }
    return A;
}

/* centroid(p), src/polygon.jl:210-219 with tri_area :201-203 */
static void poly_centroid(const lvo_state *s, int i, double *c1, double *c2)
{
    int n = s->n;
    double x1 = s->x[i], x2 = s->x[n + i];
    double A = 0.0, cx = 0.0, cy = 0.0;
    for (int k = 0; k < s->ne[i]; k++) {
        const double *e = EV(s, i, k);
        double a1 = e[0] - x1, a2 = e[1] - x2, b1 = e[2] - x1, b2 = e[3] - x2;
        double dA = 0.5 * fabs(a1 * b2 - a2 * b1);
        A += dA;
        cx += (dA * ((x1 + e[0]) + e[2])) / 3.0;
        cy += (dA * ((x2 + e[1]) + e[3])) / 3.0;
    }
    *c1 = cx / A;
    *c2 = cy / A;
}

/* lr_ratio(dx, e), src/polygon.jl:228-232: |e| / |dx| */
static inline double lr_ratio(double dx1, double dx2, const double *e)
{
    double a = e[0] - e[2], b = e[1] - e[3];
    double l2 = a * a + b * b;
    double r2 = dx1 * dx1 + dx2 * dx2;
    return sqrt(l2 / r2);
}

/* StaticArrays norm of a 2-vector (taken as the unscaled sqrt of the sum of squares) */
static inline double norm2v(double a, double b) { return sqrt(a * a + b * b); }

int lvo_area(const lvo_state *s, double *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) out[i] = poly_area(s, i);
    return 0;
}

int lvo_centroid(const lvo_state *s, double *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) poly_centroid(s, i, &out[i], &out[s->n + i]);
    return 0;
}

/* Neighbour position y = p.x + get_arrow(q.x, p.x) (src/iterators.jl:29) */
static inline void nbr_y(const lvo_grid *g, const lvo_state *s, int i, int j, double *y1, double *y2)
{
    arrow_target(g, s->x[j], s->x[s->n + j], s->x[i], s->x[s->n + i], y1, y2);
}

/* ------------------------------------------------------------------------------------------------
 * EOS and simple updates (src/pressure.jl:32-82)
 * ---------------------------------------------------------------------------------------------- */

int lvo_find_rho(lvo_state *s) /* src/pressure.jl:36-40 */
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) s->rho[i] = s->mass[i] / poly_area(s, i);
    return 0;
}

static inline double eint(const lvo_state *s, int i) /* src/pressure.jl:32-34 */
{
    double v1 = s->v[i], v2 = s->v[s->n + i];
    return s->e[i] - 0.5 * (v1 * v1 + v2 * v2);
}

int lvo_eos_stiffened(lvo_state *s, double gamma, double P0) /* src/pressure.jl:64-70 */
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        s->rho[i] = s->mass[i] / poly_area(s, i);
        s->P[i] = ((gamma - 1.0) * s->rho[i]) * eint(s, i);
        s->c2[i] = (gamma * (s->P[i] + P0)) / s->rho[i];
    }
    return 0;
}

int lvo_eos_ideal(lvo_state *s, double gamma, double Pmin) /* src/pressure.jl:49-55 */
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        s->rho[i] = s->mass[i] / poly_area(s, i);
        s->P[i] = ((gamma - 1.0) * s->rho[i]) * eint(s, i);
        double pm = (Pmin > s->P[i]) ? Pmin : s->P[i];
        s->c2[i] = (gamma * pm) / s->rho[i];
    }
    return 0;
}

int lvo_gravity(lvo_state *s, double g1, double g2, double dt) /* src/pressure.jl:77-82 */
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        s->v[i] += dt * g1;
        s->v[s->n + i] += dt * g2;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * move!, src/move.jl:9-33 (positions only; the caller remeshes afterwards as :20 does)
 * ---------------------------------------------------------------------------------------------- */

static inline double lpr(double x, double d) { return fmod(fmod(x, d) + d, d); } /* src/voronoigrid.jl:181-183 */

static int try_move(const lvo_grid *g, lvo_state *s, int i, double dt)
{
    int n = s->n;
    double v1 = s->v[i], v2 = s->v[n + i];
    if (isnan(v1) || isnan(v2)) return -1;
    double a1 = s->x[i] + dt * v1, a2 = s->x[n + i] + dt * v2;
    /* periodic_wrap, src/voronoigrid.jl:186-192 */
    double wx = lpr(a1 - g->d.bxmin[0], g->xperiod) + g->d.bxmin[0];
    double wy = lpr(a2 - g->d.bxmin[1], g->yperiod) + g->d.bxmin[1];
    double tx = g->d.xperiodic ? (wx - a1) : copysign(0.0, wx - a1);
    double ty = g->d.yperiodic ? (wy - a2) : copysign(0.0, wy - a2);
    double r1 = (a1 + tx * 1.0) + ty * 0.0;
    double r2 = (a2 + tx * 0.0) + ty * 1.0;
    if ((g->d.bxmin[0] <= r1 && r1 <= g->d.bxmax[0]) && (g->d.bxmin[1] <= r2 && r2 <= g->d.bxmax[1])) {
        s->x[i] = r1;
        s->x[n + i] = r2;
        return 1;
    }
    return 0;
}

int lvo_move_positions(const lvo_grid *g, lvo_state *s, double dt)
{
    int err = LV_OK;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        int n = s->n;
        int r = try_move(g, s, i, dt);
        if (r == 0) {
            for (int k = 0; k < s->ne[i]; k++) { /* boundaries(p): project v to the tangent */
                if (EL(s, i, k) >= 0) continue;
                const double *e = EV(s, i, k);
                double n1 = e[1] - e[3], n2 = e[2] - e[0];
                double nn = norm2v(n1, n2);
                n1 = n1 / nn;
                n2 = n2 / nn;
                double vn = s->v[i] * n1 + s->v[n + i] * n2;
                s->v[i] = s->v[i] - vn * n1;
                s->v[n + i] = s->v[n + i] - vn * n2;
            }
            r = try_move(g, s, i, dt);
            if (r == 0) {
                s->v[i] = 0.0;
                s->v[n + i] = 0.0;
                r = try_move(g, s, i, dt);
            }
        }
        if (r < 0) {
#pragma omp critical
            err = LV_E_NAN_VELOCITY;
        }
    }
    return err;
}

/* ------------------------------------------------------------------------------------------------
 * Pressure system (src/pressure.jl:89-224)
 * ---------------------------------------------------------------------------------------------- */

/* refresh!(A), src/pressure.jl:104-117: diag[i] and coef[i*maxe+k] for every edge slot (0 on boundary
 * edges, which the operator skips). */
int lvo_pressure_operator(const lvo_grid *g, const lvo_state *s, double dt, double *diag, double *coef)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        int n = s->n;
        diag[i] = s->mass[i] / (((s->rho[i] * s->rho[i]) * s->c2[i]) * (dt * dt));
        for (int k = 0; k < s->maxe; k++) coef[(size_t)i * s->maxe + k] = 0.0;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(s->x[i] - y1, s->x[n + i] - y2, EV(s, i, k));
            coef[(size_t)i * s->maxe + k] = lrr * (0.5 / s->rho[i] + 0.5 / s->rho[j]);
        }
    }
    return 0;
}

/* mul!(y, A, x), src/pressure.jl:119-130 */
int lvo_apply_A(const lvo_state *s, const double *diag, const double *coef, const double *xv, double *yv)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        double acc = diag[i] * xv[i];
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            acc += coef[(size_t)i * s->maxe + k] * (xv[i] - xv[j]);
        }
        yv[i] = acc;
    }
    return 0;
}

/* Per-label boundary velocity table: vbc[2*(-label-1) + comp] for labels -1..-4 (zero_vbc by default,
 * src/pressure.jl:205-207). */

/* refresh!(solver), src/pressure.jl:162-203.  b, P0 (the initial guess) and GP ([2][n]) are outputs. */
int lvo_rhs(const lvo_grid *g, const lvo_state *s, double dt, int gp_step, const double *vbc,
            double *b, double *P0, double *GP)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double A = poly_area(s, i);
        double bi = (A * s->P[i]) / ((s->rho[i] * s->c2[i]) * (dt * dt));
        P0[i] = s->P[i];
        double gx = 0.0, gy = 0.0;
        double x1 = s->x[i], x2 = s->x[n + i];
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double dv1 = s->v[i] - s->v[j], dv2 = s->v[n + i] - s->v[n + j];
            bi -= (lrr / dt) * (dv1 * (m1 - y1) + dv2 * (m2 - y2));
            double w = lrr * (s->P[i] - s->P[j]);
            gx -= w * (m1 - x1);
            gy -= w * (m2 - x2);
        }
        for (int k = 0; k < s->ne[i]; k++) {
            int lab = EL(s, i, k);
            if (lab >= 0) continue;
            const double *e = EV(s, i, k);
            double dS1 = e[1] - e[3], dS2 = e[2] - e[0];
            double vb1 = vbc ? vbc[2 * (-lab - 1)] : 0.0, vb2 = vbc ? vbc[2 * (-lab - 1) + 1] : 0.0;
            bi -= (dS1 * (vb1 - s->v[i]) + dS2 * (vb2 - s->v[n + i])) / dt;
        }
        GP[i] = gx / s->mass[i];
        GP[n + i] = gy / s->mass[i];
        b[i] = bi;
    }
    if (gp_step) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            double x1 = s->x[i], x2 = s->x[n + i];
            double bi = b[i];
            for (int k = 0; k < s->ne[i]; k++) {
                int j = EL(s, i, k);
                if (j < 0) continue;
                const double *e = EV(s, i, k);
                double y1, y2;
                nbr_y(g, s, i, j, &y1, &y2);
                double lrr = lr_ratio(x1 - y1, x2 - y2, e);
                double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
                double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
                bi += lrr * ((GP[i] - GP[j]) * (m1 - z1) + (GP[n + i] - GP[n + j]) * (m2 - z2));
            }
            b[i] = bi;
        }
    }
    return 0;
}

/* The operator of lvo_apply_A compacted to CSR rows (interior edges in edge order), so that a solve
 * streams 12 bytes per coupling instead of the padded per-polygon slots.  Row sums are evaluated
 * exactly as lvo_apply_A evaluates them. */
typedef struct {
    int n;
    int *rp, *ci;
    double *cv;
    const double *diag;
} lvo_csr;

static void csr_build(const lvo_state *s, const double *diag, const double *coef, lvo_csr *A)
{
    int n = s->n;
    A->n = n;
    A->diag = diag;
    A->rp = (int *)malloc(sizeof(int) * (n + 1));
    A->rp[0] = 0;
    for (int i = 0; i < n; i++) {
        int c = 0;
        for (int k = 0; k < s->ne[i]; k++) c += (EL(s, i, k) >= 0);
        A->rp[i + 1] = A->rp[i] + c;
    }
    A->ci = (int *)malloc(sizeof(int) * ((size_t)A->rp[n] + 1));
    A->cv = (double *)malloc(sizeof(double) * ((size_t)A->rp[n] + 1));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        int o = A->rp[i];
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            A->ci[o] = j;
            A->cv[o] = coef[(size_t)i * s->maxe + k];
            o++;
        }
    }
}

static void csr_free(lvo_csr *A)
{
    free(A->rp);
    free(A->ci);
    free(A->cv);
}

static inline double csr_row(const lvo_csr *A, const double *xv, int i)
{
    double acc = A->diag[i] * xv[i];
    for (int o = A->rp[i]; o < A->rp[i + 1]; o++) acc += A->cv[o] * (xv[i] - xv[A->ci[o]]);
    return acc;
}

/* Reductions over a fixed partition of the rows into NCHUNK chunks, summed in chunk order, so that
 * a solve is bitwise reproducible for any thread count. */
#define NCHUNK 256
#define CHUNK_LO(c, n) ((int)(((int64_t)(n) * (c)) / NCHUNK))

/* Jacobi-PCG on the system of refresh!(A) (the build's replacement for Krylov.minres!, which the
 * reference calls at src/pressure.jl:219).  Stops at ||b - A x||_2 <= tol*||b||_2.  If every diagonal
 * entry is 0 (c2 = Inf), the operator is a graph Laplacian: b is projected on the complement of the
 * constants and x keeps the mean of its initial guess.  Returns iterations, or -1 if not converged. */
int lvo_pcg(const lvo_state *s, const double *diag, const double *coef, const double *b_in, double *x,
            double tol, int itmax, double *relres_out)
{
    int n = s->n;
    lvo_csr A;
    csr_build(s, diag, coef, &A);
    double *b = (double *)malloc(sizeof(double) * n);
    double *r = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n);
    double *dinv = (double *)malloc(sizeof(double) * n);
    double part[3][NCHUNK];
    int singular = 1;
    for (int i = 0; i < n; i++) {
        double dd = diag[i];
        if (dd != 0.0) singular = 0;
        for (int o = A.rp[i]; o < A.rp[i + 1]; o++) dd += A.cv[o];
        dinv[i] = 1.0 / dd;
        b[i] = b_in[i];
    }
    double xmean0 = 0.0;
    if (singular) {
        double bm = 0.0;
        for (int i = 0; i < n; i++) { bm += b[i]; xmean0 += x[i]; }
        bm /= n;
        xmean0 /= n;
        for (int i = 0; i < n; i++) b[i] -= bm;
    }
    double bn = 0.0;
    for (int i = 0; i < n; i++) bn += b[i] * b[i];
    bn = sqrt(bn);
    double rz = 0.0, rn = 0.0;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; c++) {
        double a0 = 0.0, a1 = 0.0;
        for (int i = CHUNK_LO(c, n); i < CHUNK_LO(c + 1, n); i++) {
            r[i] = b[i] - csr_row(&A, x, i);
            z[i] = dinv[i] * r[i];
            p[i] = z[i];
            a0 += r[i] * z[i];
            a1 += r[i] * r[i];
        }
        part[0][c] = a0;
        part[1][c] = a1;
    }
    for (int c = 0; c < NCHUNK; c++) { rz += part[0][c]; rn += part[1][c]; }
    int it = 0;
    double thr = tol * bn;
    while (sqrt(rn) > thr && it < itmax) {
        double pq = 0.0;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < NCHUNK; c++) {
            double a0 = 0.0;
            for (int i = CHUNK_LO(c, n); i < CHUNK_LO(c + 1, n); i++) {
                q[i] = csr_row(&A, p, i);
                a0 += p[i] * q[i];
            }
            part[0][c] = a0;
        }
        for (int c = 0; c < NCHUNK; c++) pq += part[0][c];
        if (!(pq > 0.0) || !(rz > 0.0)) break; /* breakdown at the attainable accuracy */
        double alpha = rz / pq;
        double rz2 = 0.0;
        rn = 0.0;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < NCHUNK; c++) {
            double a0 = 0.0, a1 = 0.0;
            for (int i = CHUNK_LO(c, n); i < CHUNK_LO(c + 1, n); i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
                z[i] = dinv[i] * r[i];
                a0 += r[i] * z[i];
                a1 += r[i] * r[i];
            }
            part[0][c] = a0;
            part[1][c] = a1;
        }
        for (int c = 0; c < NCHUNK; c++) { rz2 += part[0][c]; rn += part[1][c]; }
        double beta = rz2 / rz;
        rz = rz2;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        it++;
        if (it % 50 == 0 || !(sqrt(rn) > thr)) { /* true residual check against recurrence drift */
#pragma omp parallel for schedule(static)
            for (int c = 0; c < NCHUNK; c++) {
                double a1 = 0.0;
                for (int i = CHUNK_LO(c, n); i < CHUNK_LO(c + 1, n); i++) {
                    double t = b[i] - csr_row(&A, x, i);
                    a1 += t * t;
                }
                part[1][c] = a1;
            }
            rn = 0.0;
            for (int c = 0; c < NCHUNK; c++) rn += part[1][c];
        }
    }
    if (singular) {
        double xm = 0.0;
        for (int i = 0; i < n; i++) xm += x[i];
        xm /= n;
        for (int i = 0; i < n; i++) x[i] -= (xm - xmean0);
    }
    if (relres_out) *relres_out = (bn > 0.0) ? sqrt(rn) / bn : sqrt(rn);
    int conv = !(sqrt(rn) > thr);
    free(b); free(r); free(z); free(p); free(q); free(dinv);
    csr_free(&A);
    return conv ? it : -1;
}

/* MINRES (Paige & Saunders 1975, unpreconditioned, in the recurrences of their SYMMLQ/MINRES paper as
 * SciPy's minres also writes them), warm-started from x: the solver the reference calls at
 * src/pressure.jl:219 (Krylov.jl minres!, atol = rtol = 1e-6, itmax = 1000).  It stops when the
 * residual estimate |phibar| <= atol + rtol*beta1 (beta1 = |b - A x0|, the rule Krylov.jl documents
 * for minres!) or when |A r| / (|A| |r|) <= rtol (a least-squares solution of an inconsistent singular
 * system).  Pinned against SciPy's minres, iterate by iterate (tests/test_oracle_pins.py); Krylov.jl
 * 0.9.8 itself is not available, so its rounding stays unpinned.  The function is for studying the
 * reference's loose solve (the cavity failure), not a parity target of the GPU solve.  Returns iterations, or
 * -1 when itmax stopped it; *relres_out = |b - A x| / |b| of the returned x. */
int lvo_minres(const lvo_state *s, const double *diag, const double *coef, const double *b, double *x,
               double atol, double rtol, int itmax, double *relres_out)
{
    int n = s->n;
    lvo_csr A;
    csr_build(s, diag, coef, &A);
    double *r1 = (double *)malloc(sizeof(double) * n), *r2 = (double *)malloc(sizeof(double) * n);
    double *y = (double *)malloc(sizeof(double) * n), *v = (double *)malloc(sizeof(double) * n);
    double *w = (double *)calloc(n, sizeof(double)), *w1 = (double *)calloc(n, sizeof(double));
    double *w2 = (double *)calloc(n, sizeof(double));
    double bn = 0.0, beta1 = 0.0;
    for (int i = 0; i < n; i++) {
        r1[i] = b[i] - csr_row(&A, x, i);
        bn += b[i] * b[i];
        beta1 += r1[i] * r1[i];
    }
    bn = sqrt(bn);
    beta1 = sqrt(beta1);
    for (int i = 0; i < n; i++) { y[i] = r1[i]; r2[i] = r1[i]; }
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, tnorm2 = 0.0;
    double cs = -1.0, sn = 0.0;
    const double eps_stop = atol + rtol * beta1;
    int it = 0, done = !(beta1 > eps_stop);
    while (!done && it < itmax) {
        it++;
        double sc = 1.0 / beta;
        for (int i = 0; i < n; i++) v[i] = sc * y[i];
        double alfa = 0.0;
        for (int i = 0; i < n; i++) y[i] = csr_row(&A, v, i);
        if (it >= 2)
            for (int i = 0; i < n; i++) y[i] -= (beta / oldb) * r1[i];
        for (int i = 0; i < n; i++) alfa += v[i] * y[i];
        for (int i = 0; i < n; i++) {
            y[i] -= (alfa / beta) * r2[i];
            r1[i] = r2[i];
            r2[i] = y[i];
        }
        oldb = beta;
        beta = 0.0;
        for (int i = 0; i < n; i++) beta += r2[i] * r2[i];
        beta = sqrt(beta);
        tnorm2 += alfa * alfa + oldb * oldb + beta * beta;
        double oldeps = epsln;
        double delta = cs * dbar + sn * alfa;
        double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        double root = sqrt(gbar * gbar + dbar * dbar);
        double gamma = sqrt(gbar * gbar + beta * beta);
        if (gamma < 2.220446049250313e-16) gamma = 2.220446049250313e-16;
        cs = gbar / gamma;
        sn = beta / gamma;
        double phi = cs * phibar;
        phibar = sn * phibar;
        double denom = 1.0 / gamma;
        for (int i = 0; i < n; i++) {
            w1[i] = w2[i];
            w2[i] = w[i];
            w[i] = (v[i] - oldeps * w1[i] - delta * w2[i]) * denom;
            x[i] += phi * w[i];
        }
        double Anorm = sqrt(tnorm2);
        double test2 = (Anorm > 0.0) ? root / Anorm : INFINITY;
        if (fabs(phibar) <= eps_stop || test2 <= rtol || !(beta > 0.0)) done = 1;
    }
    double rn = 0.0;
    for (int i = 0; i < n; i++) { double t = b[i] - csr_row(&A, x, i); rn += t * t; }
    if (relres_out) *relres_out = (bn > 0.0) ? sqrt(rn) / bn : sqrt(rn);
    free(r1); free(r2); free(y); free(v); free(w); free(w1); free(w2);
    csr_free(&A);
    return done ? it : -1;
}

/* find_pressure!, src/pressure.jl:215-224: refresh!(A), then niter x [RHS, solve warm-started from P,
 * write P back].  method 0: Jacobi-PCG to tol (the build's solver); 1: lvo_minres with
 * atol = rtol = tol (the reference's solver and tolerances). */
int lvo_find_pressure_m(const lvo_grid *g, lvo_state *s, double dt, int niter, const double *vbc, double tol,
                        int itmax, int *iters_out, int method);
int lvo_find_pressure(const lvo_grid *g, lvo_state *s, double dt, int niter, const double *vbc, double tol,
                      int itmax, int *iters_out)
{
    return lvo_find_pressure_m(g, s, dt, niter, vbc, tol, itmax, iters_out, 0);
}

int lvo_find_pressure_m(const lvo_grid *g, lvo_state *s, double dt, int niter, const double *vbc, double tol,
                        int itmax, int *iters_out, int method)
{
    int n = s->n;
    double *diag = (double *)malloc(sizeof(double) * n);
    double *coef = (double *)malloc(sizeof(double) * (size_t)n * s->maxe);
    double *b = (double *)malloc(sizeof(double) * n);
    double *P = (double *)malloc(sizeof(double) * n);
    double *GP = (double *)malloc(sizeof(double) * 2 * n);
    lvo_pressure_operator(g, s, dt, diag, coef);
    int err = LV_OK;
    for (int it = 1; it <= niter; it++) {
        lvo_rhs(g, s, dt, it > 1, vbc, b, P, GP);
        double rr;
        int k = method == 1 ? lvo_minres(s, diag, coef, b, P, tol, tol, itmax, &rr)
                            : lvo_pcg(s, diag, coef, b, P, tol, itmax, &rr);
        if (iters_out) iters_out[it - 1] = k;
        if (k < 0) err = LV_E_NOT_CONVERGED;
        for (int i = 0; i < n; i++) s->P[i] = P[i];
    }
    free(diag); free(coef); free(b); free(P); free(GP);
    return err;
}

/* pressure_step!, src/pressure.jl:10-25 */
int lvo_pressure_step(const lvo_grid *g, lvo_state *s, double dt)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double v1 = s->v[i], v2 = s->v[n + i];
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double w = ((dt / s->mass[i]) * lrr) * (s->P[i] - s->P[j]);
            v1 = v1 + w * (m1 - x1);
            v2 = v2 + w * (m2 - x2);
        }
        s->v[i] = v1; /* pass 1 reads no neighbour v, so in place is race-free */
        s->v[n + i] = v2;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double ei = s->e[i];
        double Pi = s->P[i];
        double pv1 = Pi * s->v[i], pv2 = Pi * s->v[n + i];
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double Pj = s->P[j];
            double qv1 = Pj * s->v[j], qv2 = Pj * s->v[n + j];
            double a = (m1 - x1) * pv1 + (m2 - x2) * pv2;
            double bq = (m1 - y1) * qv1 + (m2 - y2) * qv2;
            ei = ei - ((dt * lrr) / s->mass[i]) * (a - bq);
        }
        s->e[i] = ei; /* pass 2 reads no neighbour e */
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Viscosity (src/diffusion.jl)
 * ---------------------------------------------------------------------------------------------- */

/* find_D!, src/diffusion.jl:8-19 */
int lvo_find_D(const lvo_grid *g, lvo_state *s)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double D11 = 0.0, D21 = 0.0, D12 = 0.0, D22 = 0.0;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double a1 = s->v[i] - s->v[j], a2 = s->v[n + i] - s->v[n + j];
            double b1 = m1 - y1, b2 = m2 - y2;
            D11 = D11 + lrr * (a1 * b1);
            D21 = D21 + lrr * (a2 * b1);
            D12 = D12 + lrr * (a1 * b2);
            D22 = D22 + lrr * (a2 * b2);
        }
        double A = poly_area(s, i);
        D11 = D11 / A; D21 = D21 / A; D12 = D12 / A; D22 = D22 / A;
        s->D[i] = 0.5 * (D11 + D11);
        s->D[n + i] = 0.5 * (D21 + D12);
        s->D[2 * n + i] = 0.5 * (D12 + D21);
        s->D[3 * n + i] = 0.5 * (D22 + D22);
    }
    return 0;
}

/* getS, src/diffusion.jl:22-29: S as four column-major entries */
static inline void getS(const lvo_state *s, int i, double dr, double *S)
{
    int n = s->n;
    double D11 = s->D[i], D21 = s->D[n + i], D12 = s->D[2 * n + i], D22 = s->D[3 * n + i];
    double divv = ((D11 * 1.0 + D21 * 0.0) + D12 * 0.0) + D22 * 1.0;
    double mu = s->mu[i];
    if (divv < 0.0) mu -= (divv * s->rho[i]) * (dr * dr);
    double tm = 2.0 * mu;
    S[0] = tm * (D11 - (divv * 1.0) / 3.0);
    S[1] = tm * (D21 - (divv * 0.0) / 3.0);
    S[2] = tm * (D12 - (divv * 0.0) / 3.0);
    S[3] = tm * (D22 - (divv * 1.0) / 3.0);
}

/* viscous_step!, src/diffusion.jl:39-53 */
int lvo_viscous_step(const lvo_grid *g, lvo_state *s, double dt, int artificial_viscosity)
{
    int n = s->n;
    double avdr = artificial_viscosity ? g->d.dr : 0.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double v1 = s->v[i], v2 = s->v[n + i];
        double Sp[4];
        getS(s, i, avdr, Sp);
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double Sq[4];
            getS(s, j, avdr, Sq);
            double sc = (dt * lr_ratio(x1 - y1, x2 - y2, e)) / s->mass[i];
            double M11 = Sp[0] - Sq[0], M21 = Sp[1] - Sq[1], M12 = Sp[2] - Sq[2], M22 = Sp[3] - Sq[3];
            double w1 = (m1 - x1) * sc, w2 = (m2 - x2) * sc; /* s*M*w == M*(w*s) */
            v1 = v1 - (M11 * w1 + M12 * w2);
            v2 = v2 - (M21 * w1 + M22 * w2);
        }
        s->v[i] = v1; /* pass 1 reads neighbours' D, mu, rho only */
        s->v[n + i] = v2;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double ei = s->e[i];
        double Sp[4];
        getS(s, i, avdr, Sp);
        double sv1 = Sp[0] * s->v[i] + Sp[2] * s->v[n + i];
        double sv2 = Sp[1] * s->v[i] + Sp[3] * s->v[n + i];
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double Sq[4];
            getS(s, j, avdr, Sq);
            double qv1 = Sq[0] * s->v[j] + Sq[2] * s->v[n + j];
            double qv2 = Sq[1] * s->v[j] + Sq[3] * s->v[n + j];
            double a = (m1 - x1) * sv1 + (m2 - x2) * sv2;
            double bq = (m1 - y1) * qv1 + (m2 - y2) * qv2;
            ei = ei + ((dt * lr_ratio(x1 - y1, x2 - y2, e)) / s->mass[i]) * (a - bq);
        }
        s->e[i] = ei; /* pass 2 reads no neighbour e */
    }
    return 0;
}

/* bdary_friction!, src/diffusion.jl:64-78, with vDirichlet given per boundary label (vbc table) and
 * charfun = everywhere */
int lvo_bdary_friction(lvo_state *s, const double *vbc, double dt)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double tmp = 1.0;
        double v1 = s->v[i], v2 = s->v[n + i];
        double ei = s->e[i];
        for (int k = 0; k < s->ne[i]; k++) {
            int lab = EL(s, i, k);
            if (lab >= 0) continue;
            double vd1 = vbc[2 * (-lab - 1)], vd2 = vbc[2 * (-lab - 1) + 1];
            if (isnan(vd1) || isnan(vd2)) continue;  /* charfun(m) false on this wall: `continue`, :70 */
            const double *e = EV(s, i, k);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double n1 = e[1] - e[3], n2 = e[2] - e[0];
            double nn = norm2v(n1, n2);
            n1 = n1 / nn;
            n2 = n2 / nn;
            double len = norm2v(e[0] - e[2], e[1] - e[3]);
            double lrr = len / fabs((m1 - x1) * n1 + (m2 - x2) * n2);
            double ml = s->mu[i] * lrr;
            double f1 = (ml * vd1) / s->mass[i], f2 = (ml * vd2) / s->mass[i];
            tmp += ((dt * s->mu[i]) * lrr) / s->mass[i];
            ei += dt * (f1 * v1 + f2 * v2);
            v1 = v1 + dt * f1;
            v2 = v2 + dt * f2;
        }
        s->e[i] = ei;
        s->v[i] = v1 / tmp;
        s->v[n + i] = v2 / tmp;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Relaxation (src/relaxation.jl:10-73)
 * ---------------------------------------------------------------------------------------------- */

/* find_dv!, src/relaxation.jl:10-25 */
int lvo_find_dv(const lvo_grid *g, lvo_state *s, double dt, double alpha)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double c1, c2;
        poly_centroid(s, i, &c1, &c2);
        double rmax = 0.0, rmin = INFINITY;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double r = norm2v(x1 - y1, x2 - y2);
            if (r > rmax) rmax = r;
            if (r < rmin) rmin = r;
        }
        double q = rmin / rmax;
        s->quality[i] = q;
        double D11 = s->D[i], D21 = s->D[n + i], D12 = s->D[2 * n + i], D22 = s->D[3 * n + i];
        double nD = sqrt(((D11 * D11 + D21 * D21) + D12 * D12) + D22 * D22);
        double lambda = (alpha * nD) / (q * q);
        double f = lambda / (1.0 + dt * lambda);
        s->dv[i] = f * (c1 - x1);
        s->dv[n + i] = f * (c2 - x2);
    }
    return 0;
}

/* relaxation_step!, src/relaxation.jl:36-73 without the trailing remesh! (the caller remeshes).
 * Pass 1 reads only q.rho, q.v, q.e, q.dv, q.phase of neighbours and writes p.mass, p.momentum,
 * p.energy, so updating in place is race-free, exactly as the @batch loop is. */
int lvo_relaxation_step(const lvo_grid *g, lvo_state *s, double dt, int rusanov)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double rp = s->rho[i], ep = s->e[i];
        double pv1 = s->v[i], pv2 = s->v[n + i];
        double pd1 = s->dv[i], pd2 = s->dv[n + i];
        double mass = s->mass[i];
        double mo1 = mass * pv1, mo2 = mass * pv2;
        double en = mass * ep;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            if (!(s->phase[i] == s->phase[j])) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
            double rq = s->rho[j], eq = s->e[j];
            double qv1 = s->v[j], qv2 = s->v[n + j];
            double qd1 = s->dv[j], qd2 = s->dv[n + j];
            double pdvpq = pd1 * (x1 - y1) + pd2 * (x2 - y2);
            double qdvpq = qd1 * (x1 - y1) + qd2 * (x2 - y2);
            double pdvmz = pd1 * (m1 - z1) + pd2 * (m2 - z2);
            double qdvmz = qd1 * (m1 - z1) + qd2 * (m2 - z2);
            double dl = dt * lrr;
            mass += dl * (((pdvmz * rp) - (qdvmz * rq)) - 0.5 * ((pdvpq * rp) + (qdvpq * rq)));
            mo1 += dl * ((((pdvmz * rp) * pv1) - ((qdvmz * rq) * qv1)) - 0.5 * (((pdvpq * rp) * pv1) + ((qdvpq * rq) * qv1)));
            mo2 += dl * ((((pdvmz * rp) * pv2) - ((qdvmz * rq) * qv2)) - 0.5 * (((pdvpq * rp) * pv2) + ((qdvpq * rq) * qv2)));
            en += dl * ((((pdvmz * rp) * ep) - ((qdvmz * rq) * eq)) - 0.5 * (((pdvpq * rp) * ep) + ((qdvpq * rq) * eq)));
            if (rusanov) {
                double na = norm2v(pd1, pd2), nb = norm2v(qd1, qd2);
                double a = (nb > na) ? nb : na;
                double l = norm2v(e[0] - e[2], e[1] - e[3]);
                double w = ((0.5 * dt) * l) * a;
                mass += w * (rq - rp);
                mo1 += w * (rq * qv1 - rp * pv1);
                mo2 += w * (rq * qv2 - rp * pv2);
                en += w * (rq * eq - rp * ep);
            }
        }
        s->mass[i] = mass;
        s->mom[i] = mo1;
        s->mom[n + i] = mo2;
        s->energy[i] = en;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        s->v[i] = s->mom[i] / s->mass[i];
        s->v[n + i] = s->mom[n + i] / s->mass[i];
        s->e[i] = s->energy[i] / s->mass[i];
        s->x[i] += dt * s->dv[i];
        s->x[n + i] += dt * s->dv[n + i];
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Multiphase projection (src/relaxation.jl:75-206): MultiphaseProjector, refresh!(MultiphaseSolver)
 * and the dv update of multiphase_projection!.  Only edges between particles of different phases
 * contribute.  Krylov.jl's minres! (atol = rtol = 1e-4, itmax = 200) is replaced by the same
 * unpreconditioned CG as the product, run to the caller's tolerance.
 * ---------------------------------------------------------------------------------------------- */

/* refresh!(solver::MultiphaseSolver), src/relaxation.jl:165-180 */
int lvo_multiphase_rhs(const lvo_grid *g, const lvo_state *s, double *b)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double d1 = s->dv[i], d2 = s->dv[n + i];
        double bi = 0.0;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0 || s->phase[j] == s->phase[i]) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
            double q1 = s->dv[j], q2 = s->dv[n + j];
            double a = (d1 - q1) * (m1 - z1) + (d2 - q2) * (m2 - z2);
            double c = (d1 + q1) * (x1 - y1) + (d2 + q2) * (x2 - y2);
            bi -= lrr * (a - 0.5 * c);
        }
        b[i] = bi;
    }
    return 0;
}

/* mul!(res, A::MultiphaseProjector, x), src/relaxation.jl:92-125: T = the interface gradient of x
 * per unit area (2n scratch), then res */
int lvo_multiphase_apply(const lvo_grid *g, const lvo_state *s, const double *x, double *T, double *res)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double t1 = 0.0, t2 = 0.0;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0 || s->phase[j] == s->phase[i]) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double w = lrr * (x[i] - x[j]);
            t1 -= w * (m1 - x1);
            t2 -= w * (m2 - x2);
        }
        double A = poly_area(s, i);
        T[i] = t1 / A;
        T[n + i] = t2 / A;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double x1 = s->x[i], x2 = s->x[n + i];
        double ri = 0.0;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0 || s->phase[j] == s->phase[i]) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double z1 = 0.5 * (x1 + y1), z2 = 0.5 * (x2 + y2);
            double a = (T[i] - T[j]) * (m1 - z1) + (T[n + i] - T[n + j]) * (m2 - z2);
            double c = (T[i] + T[j]) * (x1 - y1) + (T[n + i] + T[n + j]) * (x2 - y2);
            ri -= lrr * (a - 0.5 * c);
        }
        res[i] = ri;
    }
    return 0;
}

/* the dv update of multiphase_projection!, src/relaxation.jl:190-205 */
int lvo_multiphase_update(const lvo_grid *g, lvo_state *s, const double *res, double quality_threshold)
{
    int n = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        if (s->quality[i] < quality_threshold) continue;
        double x1 = s->x[i], x2 = s->x[n + i];
        double A = poly_area(s, i);
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0 || s->phase[j] == s->phase[i]) continue;
            const double *e = EV(s, i, k);
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(x1 - y1, x2 - y2, e);
            double m1 = 0.5 * (e[0] + e[2]), m2 = 0.5 * (e[1] + e[3]);
            double w = lrr * (res[i] - res[j]);
            s->dv[i] += (w * (m1 - x1)) / A;
            s->dv[n + i] += (w * (m2 - x2)) / A;
        }
    }
    return 0;
}

/* multiphase_projection!(solver), src/relaxation.jl:190-206, with CG from x = 0 stopping at
 * |r| <= atol + rtol |b| (Krylov.jl's criterion); returns the iterations, -1 if itmax was hit */
int lvo_multiphase_projection(const lvo_grid *g, lvo_state *s, double quality_threshold, double atol, double rtol,
                              int itmax, double *res_out)
{
    int n = s->n;
    double *b = calloc((size_t)n, sizeof(double)), *x = calloc((size_t)n, sizeof(double));
    double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
    double *T = malloc(sizeof(double) * 2 * n);
    lvo_multiphase_rhs(g, s, b);
    double bb = 0.0;
    for (int i = 0; i < n; i++) {
        r[i] = b[i];
        p[i] = b[i];
        bb += b[i] * b[i];
    }
    double thr = atol + rtol * sqrt(bb), rr = bb;
    int it = 0, ok = sqrt(rr) <= thr;
    while (!ok && it < itmax) {
        lvo_multiphase_apply(g, s, p, T, q);
        double pq = 0.0;
        for (int i = 0; i < n; i++) pq += p[i] * q[i];
        if (!(pq > 0.0)) break;
        double alpha = rr / pq, rr1 = 0.0;
        for (int i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            rr1 += r[i] * r[i];
        }
        it++;
        if (sqrt(rr1) <= thr) {
            ok = 1;
            break;
        }
        double beta = rr1 / rr;
        rr = rr1;
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    }
    lvo_multiphase_update(g, s, x, quality_threshold);
    if (res_out) memcpy(res_out, x, sizeof(double) * n);
    free(b); free(x); free(r); free(p); free(q); free(T);
    return ok ? it : -1;
}

/* ------------------------------------------------------------------------------------------------
 * Implicit viscosity (src/viscous_solver.jl:43-128): ViscousSolver's operator in the difference form
 * of the pressure operator, d_i = m_i/dt and c_ij = 0.5 lrr (mu_i + mu_j) (assemble!, :89-113; the
 * 2n x 2n system is two copies of this n x n block, one per velocity component), and viscous_step!
 * with Jacobi-PCG in place of minres! (atol = rtol = 1e-6, itmax = 2000) from the initial guess v.
 * ---------------------------------------------------------------------------------------------- */

int lvo_viscous_operator(const lvo_grid *g, const lvo_state *s, double dt, double *diag, double *coef)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        int n = s->n;
        diag[i] = s->mass[i] / dt;
        for (int k = 0; k < s->maxe; k++) coef[(size_t)i * s->maxe + k] = 0.0;
        for (int k = 0; k < s->ne[i]; k++) {
            int j = EL(s, i, k);
            if (j < 0) continue;
            double y1, y2;
            nbr_y(g, s, i, j, &y1, &y2);
            double lrr = lr_ratio(s->x[i] - y1, s->x[n + i] - y2, EV(s, i, k));
            coef[(size_t)i * s->maxe + k] = (0.5 * lrr) * (s->mu[i] + s->mu[j]);
        }
    }
    return 0;
}

/* viscous_step!(solver::ViscousSolver, dt), src/viscous_solver.jl:120-128; returns the larger of the
 * two components' iteration counts (-1 if either missed tol) */
int lvo_viscous_implicit(const lvo_grid *g, lvo_state *s, double dt, double tol, int itmax)
{
    int n = s->n;
    double *diag = malloc(sizeof(double) * n), *coef = malloc(sizeof(double) * (size_t)n * s->maxe);
    double *b = malloc(sizeof(double) * n), *x = malloc(sizeof(double) * n);
    lvo_viscous_operator(g, s, dt, diag, coef);
    int worst = 0;
    for (int c = 0; c < 2; c++) {
        for (int i = 0; i < n; i++) {
            b[i] = (s->mass[i] / dt) * s->v[(size_t)c * n + i];
            x[i] = s->v[(size_t)c * n + i];
        }
        double rr = 0.0;
        int it = lvo_pcg(s, diag, coef, b, x, tol, itmax, &rr);
        if (it < 0 || worst < 0) worst = -1;
        else if (it > worst) worst = it;
        for (int i = 0; i < n; i++) s->v[(size_t)c * n + i] = x[i];
    }
    free(diag); free(coef); free(b); free(x);
    return worst;
}
