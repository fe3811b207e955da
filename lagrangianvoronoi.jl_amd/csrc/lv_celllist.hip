// lv_celllist.hip -- cell-list rebuild: CellList insert!/insert_periodic! (src/neighborlist.jl:47-67,
// src/voronoigrid.jl:130-147) as a stable LSD radix sort of cell keys.
//
// The reference pushes label i (and up to 8 periodic images) into FastVector cells in label order,
// so each cell holds its labels in ascending order.  Here every in-grid image becomes an item
// (key = linear cell id, value = label) emitted in (label, image) order; a stable radix sort on
// the key alone then leaves each cell's items in exactly the reference's order.
//
// Radix sort: 8-bit digits, ceil(bits(ncells)/8) passes.  Per pass: per-tile digit histograms
// (LDS atomics), one exclusive scan over the (digit, tile) table, then a scatter in which each
// 256-thread block ranks its 2048-item tile stably: within a wavefront by ballot match on the
// digit bits (peer masks + popcount), across wavefronts and rounds by running LDS counters.
#include "lv_internal.h"

namespace {

constexpr int RB = 256;                 // threads per block
constexpr int RR = 8;                   // rounds (items per thread)
constexpr int RT = RB * RR;             // items per tile
constexpr int NW = RB / 64;

// In-grid images of seed x in insert_periodic! order; returns how many, writes their cell ids.
// the keys of a seed and of its periodic images, in the order insert_periodic! visits them.  An
// image's coordinate is x1, x1 + P or x1 - P (and x2, x2 + Q or x2 - Q; the reference's
// zero-weighted terms add 0.0, which changes no floor), so the exact divisions of findkey are made
// once per distinct coordinate: 6 instead of 18
__device__ __forceinline__ bool lv_key_axis(double x, double origin, double h, int nmax, int &i)
{
    const double f = floor((x - origin) / h);
    if (!(f == f) || f < -2.0e9 || f > 2.0e9) return false;
    i = (int)f + 1;
    return i >= 1 && i <= nmax;
}

__device__ __forceinline__ int seed_images(const lv_grid_params &g, double x1, double x2, int *cells)
{
    const double P = g.xperiod, Q = g.yperiod;
    int ix[3], iy[3];
    bool ox[3], oy[3];
    ox[0] = lv_key_axis(x1, g.origin0, g.h, g.n1, ix[0]);
    oy[0] = lv_key_axis(x2, g.origin1, g.h, g.n2, iy[0]);
    ox[1] = ox[2] = oy[1] = oy[2] = false;
    if (g.xperiodic) {
        ox[1] = lv_key_axis(x1 + P * 1.0, g.origin0, g.h, g.n1, ix[1]);
        ox[2] = lv_key_axis(x1 - P * 1.0, g.origin0, g.h, g.n1, ix[2]);
    }
    if (g.yperiodic) {
        oy[1] = lv_key_axis(x2 + Q * 1.0, g.origin1, g.h, g.n2, iy[1]);
        oy[2] = lv_key_axis(x2 - Q * 1.0, g.origin1, g.h, g.n2, iy[2]);
    }
    // (a, b): the image's x and y variants, in the order of the reference's image list
    int pa[9], pb[9], k = 0;
    pa[k] = 0; pb[k] = 0; k++;
    if (g.xperiodic) {
        pa[k] = 1; pb[k] = 0; k++;
        pa[k] = 2; pb[k] = 0; k++;
    }
    if (g.yperiodic) {
        pa[k] = 0; pb[k] = 1; k++;
        pa[k] = 0; pb[k] = 2; k++;
    }
    if (g.xperiodic && g.yperiodic) {
        pa[k] = 1; pb[k] = 1; k++;
        pa[k] = 1; pb[k] = 2; k++;
        pa[k] = 2; pb[k] = 1; k++;
        pa[k] = 2; pb[k] = 2; k++;
    }
    int m = 0;
    for (int j = 0; j < k; j++)
        if (ox[pa[j]] && oy[pb[j]]) cells[m++] = (ix[pa[j]] - 1) + (iy[pb[j]] - 1) * g.n1;
    return m;
}

__global__ __launch_bounds__(256) void k_img_count(lv_grid_params g, int n, const double *__restrict__ x,
                                                   int *__restrict__ cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cells[9];
    cnt[i] = seed_images(g, x[i], x[n + i], cells);
}

__global__ __launch_bounds__(256) void k_img_emit(lv_grid_params g, int n, const double *__restrict__ x,
                                                  const int *__restrict__ off, unsigned *__restrict__ keys,
                                                  int *__restrict__ vals, int64_t cap)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cells[9];
    int m = seed_images(g, x[i], x[n + i], cells);
    int64_t o = off[i];
    for (int j = 0; j < m; j++) {
        if (o + j < cap) {
            keys[o + j] = (unsigned)cells[j];
            vals[o + j] = i;
        }
    }
}

__global__ __launch_bounds__(RB) void k_radix_hist(const unsigned *__restrict__ keys, int m, int shift,
                                                   int ntiles, int *__restrict__ tile_cnt)
{
    __shared__ int hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RT;
#pragma unroll
    for (int r = 0; r < RR; r++) {
        int j = base + r * RB + threadIdx.x;
        if (j < m) atomicAdd(&hist[(keys[j] >> shift) & 255u], 1);
    }
    __syncthreads();
    tile_cnt[threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(RB) void k_radix_scatter(const unsigned *__restrict__ kin, const int *__restrict__ vin,
                                                      unsigned *__restrict__ kout, int *__restrict__ vout, int m,
                                                      int shift, int ntiles, const int *__restrict__ tile_off)
{
    __shared__ int run[256];        // items of each digit already placed by earlier rounds
    __shared__ int wcnt[NW][256];   // items of each digit in each wavefront of this round
    __shared__ int toff[256];       // global offset of (digit, this tile)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    run[tid] = 0;
    toff[tid] = tile_off[tid * ntiles + blockIdx.x];
#pragma unroll
    for (int q = 0; q < NW; q++) wcnt[q][tid] = 0;
    __syncthreads();
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int base = blockIdx.x * RT;
    for (int r = 0; r < RR; r++) {
        const int j = base + r * RB + tid;
        const bool valid = j < m;
        unsigned key = valid ? kin[j] : 0u;
        int val = valid ? vin[j] : 0;
        const unsigned d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            unsigned long long bb = __ballot(valid && bit);
            peers &= bit ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = toff[d] + run[d] + rank;
            for (int q = 0; q < w; q++) pos += wcnt[q][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            add += wcnt[q][tid];
            wcnt[q][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

// cell_start[c] = first sorted item with key >= c (every cell written exactly once) and the seed
// position of every item, so the clipper reads candidate label and position from one run
// (cell_x holds two planes of stride cx_ld).
__global__ __launch_bounds__(256) void k_cell_bounds(const unsigned *__restrict__ keys, const int *__restrict__ vals,
                                                     int m, int ncells, int *__restrict__ cell_start,
                                                     int *__restrict__ cell_labels, double *__restrict__ cell_x,
                                                     const double *__restrict__ x, int n, int64_t cx_ld)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    int k = (int)keys[s];
    int prev = (s > 0) ? (int)keys[s - 1] : -1;
    for (int c = prev + 1; c <= k; c++) cell_start[c] = s;
    if (s == m - 1)
        for (int c = k + 1; c <= ncells; c++) cell_start[c] = m;
    int lab = vals[s];
    cell_labels[s] = lab;
    cell_x[s] = x[lab];
    cell_x[cx_ld + s] = x[n + lab];
}

__global__ void k_fill_i32(int *p, int64_t n, int v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline int64_t ntiles_of(int64_t m) { return (m + RT - 1) / RT; }

}  // namespace

// temp layout: [cnt/off: n+1 ints][keys A][vals A][keys B][vals B][tile table: 256*T + 1][scan temp]
size_t lv_celllist_temp_bytes(int64_t max_items, int64_t ncells)
{
    (void)ncells;
    int64_t T = ntiles_of(max_items);
    size_t b = 0;
    b += (size_t)(max_items + 1) * 4;             // per-seed counts / offsets (n <= max_items)
    b += (size_t)max_items * 4 * 4;               // two key/value ping-pong buffers
    b += (size_t)(256 * T + 1) * 4;               // tile table
    size_t sc1 = lv_scan_temp_bytes(256 * T), sc2 = lv_scan_temp_bytes(max_items);
    b += (sc1 > sc2 ? sc1 : sc2) + 256;
    return b;
}

// Builds the cell list of n seeds.  d_nitems receives the item count (host reads it to size the
// sort, and returns it in *h_nitems when given).  Returns hipErrorInvalidValue if the items exceed
// max_items.
hipError_t lv_build_cell_list(const lv_grid_params &g, int n, const double *x, int *cell_start, int *cell_labels,
                              double *cell_x, int64_t max_items, int *d_nitems, void *temp, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1, bool timed, int *h_nitems)
{
    const int64_t ncells = (int64_t)g.n1 * g.n2;
    char *t = (char *)temp;
    int *cnt = (int *)t;
    t += (size_t)(max_items + 1) * 4;
    unsigned *kA = (unsigned *)t;
    t += (size_t)max_items * 4;
    int *vA = (int *)t;
    t += (size_t)max_items * 4;
    unsigned *kB = (unsigned *)t;
    t += (size_t)max_items * 4;
    int *vB = (int *)t;
    t += (size_t)max_items * 4;
    int64_t Tmax = ntiles_of(max_items);
    int *tile = (int *)t;
    t += (size_t)(256 * Tmax + 1) * 4;
    t = (char *)(((uintptr_t)t + 255) & ~(uintptr_t)255);
    void *scan_tmp = t;

    if (n <= 0) {
        if (h_nitems) *h_nitems = 0;
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((ncells + 1 + 255) / 256)), dim3(256), 0, s, cell_start,
                           ncells + 1, 0);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_img_count, dim3((n + 255) / 256), dim3(256), 0, s, g, n, x, cnt);
    hipError_t e = lv_exclusive_scan_i32(cnt, cnt, n, d_nitems, scan_tmp, s);
    if (e != hipSuccess) return e;
    int m = 0;
    e = hipMemcpyAsync(&m, d_nitems, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (h_nitems) *h_nitems = m;
    if (m > max_items || m <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_img_emit, dim3((n + 255) / 256), dim3(256), 0, s, g, n, x, (const int *)cnt, kA, vA,
                       (int64_t)max_items);

    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < ncells) bits++;
    const int passes = (bits + 7) / 8;
    const int T = (int)ntiles_of(m);
    unsigned *kin = kA, *kout = kB;
    int *vin = vA, *vout = vB;
    if (timed) hipEventRecord(ev0, s);
    for (int p = 0; p < passes; p++) {
        hipLaunchKernelGGL(k_radix_hist, dim3(T), dim3(RB), 0, s, (const unsigned *)kin, m, 8 * p, T, tile);
        e = lv_exclusive_scan_i32(tile, tile, (int64_t)256 * T, nullptr, scan_tmp, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(T), dim3(RB), 0, s, (const unsigned *)kin, (const int *)vin, kout,
                           vout, m, 8 * p, T, (const int *)tile);
        unsigned *tk = kin; kin = kout; kout = tk;
        int *tv = vin; vin = vout; vout = tv;
    }
    if (timed) hipEventRecord(ev1, s);
    hipLaunchKernelGGL(k_cell_bounds, dim3((m + 255) / 256), dim3(256), 0, s, (const unsigned *)kin,
                       (const int *)vin, m, (int)ncells, cell_start, cell_labels, cell_x, x, n, max_items);
    return hipGetLastError();
}
