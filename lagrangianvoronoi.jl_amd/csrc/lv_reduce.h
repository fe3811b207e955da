// lv_reduce.h -- deterministic block reductions shared by the solver kernels (256-thread blocks).
#pragma once

#include "lv_internal.h"

// grid of the reduction kernels: enough 256-thread blocks for ~8 wavefronts per SIMD on 256 CUs,
// capped so that re-reducing the partials stays cheap.  Every kernel that writes partials for a
// consumer uses exactly this many blocks.
constexpr int LV_NBMAX = 2048;
inline int lv_nblocks(int n)
{
    int nb = (n + 255) / 256;
    return nb < LV_NBMAX ? (nb > 0 ? nb : 1) : LV_NBMAX;
}

namespace {

// ---- XCD-aware work assignment --------------------------------------------------------------------
//
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  A kernel that
// gathers neighbours maps workgroup b to the logical block lv_xcd_block(b, nb): XCD x then owns
// the contiguous logical range [x q + min(x, r), ...) (q = nb / 8, r = nb % 8), so the rows one
// XCD works on at a time, and the neighbours they gather, are contiguous and share its L2.  The
// map is a bijection of [0, nb).
__device__ __forceinline__ int lv_xcd_block(int b, int nb)
{
    const int x = b & 7, k = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + k;
}

// ---- reductions --------------------------------------------------------------------------------
//
// A kernel that produces dot products writes one partial sum per block with plain stores.  The
// kernel that consumes them re-reduces all NB partials in every block, in the same fixed order,
// so every block obtains bit-identical totals without atomics or device-scope fences (on a
// multi-XCD GPU a release fence per block costs an L2 write-back).  The kernel boundary makes the
// partials visible.  Block 0 also stores the totals in lv_cg_scalars for the host and later kernels.

template <int NV>
__device__ __forceinline__ void block_partials(const double (&v)[NV], double *__restrict__ partials, int nb)
{
    __shared__ double sm[NV][256 / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double s = lv_wave_sum(v[q]);
        if (lane == 0) sm[q][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double s = 0.0;
            for (int k = 0; k < 256 / 64; k++) s += sm[q][k];
            partials[(size_t)q * nb + blockIdx.x] = s;
        }
    }
}

// reduce_all in two parts: the loads of a thread's first LV_RED_PER partials per value, then the
// sums, so that a kernel can issue its other loads before it needs the sums
constexpr int LV_RED_PER = (LV_NBMAX + 255) / 256;

template <int NV>
__device__ __forceinline__ void reduce_all_load(const double *__restrict__ partials, int nb,
                                                double (&p)[NV][LV_RED_PER])
{
    // each thread sums partials threadIdx.x, +256, ... in that order; the first LV_NBMAX of them in
    // an unrolled loop without branches (a slot past nb loads a valid partial and is not added), so
    // that all of a thread's loads are in flight together: a rolled loop waits for one L2 round trip
    // per partial
    const int last = nb > 0 ? nb - 1 : 0;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double *pq = partials + (size_t)q * nb;
#pragma unroll
        for (int j = 0; j < LV_RED_PER; j++) {
            const int k = (int)threadIdx.x + 256 * j;
            p[q][j] = pq[k < nb ? k : last];
        }
    }
}

template <int NV>
__device__ __forceinline__ void reduce_all_sum(const double *__restrict__ partials, int nb,
                                               const double (&p)[NV][LV_RED_PER], double (&out)[NV])
{
    __shared__ double sm[NV][256 / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double *pq = partials + (size_t)q * nb;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < LV_RED_PER; j++) {
            const int k = (int)threadIdx.x + 256 * j;
            s = (k < nb) ? s + p[q][j] : s;
        }
        for (int k = (int)threadIdx.x + 256 * LV_RED_PER; k < nb; k += 256) s += pq[k];
        s = lv_wave_sum(s);
        if (lane == 0) sm[q][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double t = 0.0;
        for (int k = 0; k < 256 / 64; k++) t += sm[q][k];
        out[q] = t;
    }
    __syncthreads();
}

template <int NV>
__device__ __forceinline__ void reduce_all(const double *__restrict__ partials, int nb, double (&out)[NV])
{
    double p[NV][LV_RED_PER];
    reduce_all_load<NV>(partials, nb, p);
    reduce_all_sum<NV>(partials, nb, p, out);
}

// the loaded partials in registers here: the loads are issued before a following branch (which
// the compiler would otherwise take them behind)
template <int NV>
__device__ __forceinline__ void reduce_all_pin(double (&p)[NV][LV_RED_PER])
{
#pragma unroll
    for (int q = 0; q < NV; q++)
#pragma unroll
        for (int j = 0; j < LV_RED_PER; j++) asm volatile("" : "+v"(p[q][j]));
}

}  // namespace
