// lv_scan.hip -- device-wide exclusive prefix sum of int32 (reduce-then-scan, 2048-item tiles).
//
// Used for the cell-list item offsets, the radix-sort digit offsets and the CSR row pointer.
// Three launches: per-tile totals, one-block scan of the tile totals (looping over chunks with a
// carry), per-tile scan with the tile's offset.  Deterministic: integer sums in a fixed tree.
#include "lv_internal.h"

namespace {

constexpr int SB = 256;          // threads per block
constexpr int SI = 8;            // items per thread
constexpr int ST = SB * SI;      // items per tile

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive scan over the block; *total receives the block sum
template <int BLOCK>
__device__ __forceinline__ int block_excl_scan(int v, int *smem, int &total)
{
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) smem[w] = inc;
    __syncthreads();
    if (w == 0) {
        int s = (lane < NW) ? smem[lane] : 0;
        s = wave_incl_scan(s);
        if (lane < NW) smem[NW + lane] = s;
    }
    __syncthreads();
    int wpre = (w > 0) ? smem[NW + w - 1] : 0;
    total = smem[2 * NW - 1];
    __syncthreads();
    return wpre + inc - v;
}

__global__ __launch_bounds__(SB) void k_scan_reduce(const int *__restrict__ in, int64_t n, int *__restrict__ bsum)
{
    __shared__ int smem[2 * (SB / 64)];
    const int64_t base = (int64_t)blockIdx.x * ST;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SI; k++) {
        int64_t j = base + (int64_t)k * SB + threadIdx.x;
        if (j < n) s += in[j];
    }
    int tot;
    block_excl_scan<SB>(s, smem, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_bsums(int *__restrict__ bsum, int nb, int *__restrict__ total)
{
    __shared__ int smem[2 * 16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c = 0; c < nb; c += 1024) {
        int j = c + threadIdx.x;
        int v = (j < nb) ? bsum[j] : 0;
        int tot;
        int ex = block_excl_scan<1024>(v, smem, tot);
        int cr = carry;
        if (j < nb) bsum[j] = cr + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = cr + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SB) void k_scan_apply(const int *__restrict__ in, int *__restrict__ out, int64_t n,
                                                   const int *__restrict__ bsum)
{
    __shared__ int smem[2 * (SB / 64)];
    const int64_t base = (int64_t)blockIdx.x * ST + (int64_t)threadIdx.x * SI;  // blocked arrangement
    int v[SI];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SI; k++) {
        int64_t j = base + k;
        v[k] = (j < n) ? in[j] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_excl_scan<SB>(s, smem, tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SI; k++) {
        int64_t j = base + k;
        if (j < n) out[j] = ex;
        ex += v[k];
    }
}

}  // namespace

size_t lv_scan_temp_bytes(int64_t n)
{
    int64_t nb = (n + ST - 1) / ST;
    return (size_t)(nb + 1) * sizeof(int);
}

// out[j] = sum_{k<j} in[j]; *d_total (device, may be null) = sum of all.  in and out may alias.
hipError_t lv_exclusive_scan_i32(const int *in, int *out, int64_t n, int *d_total, void *temp, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int nb = (int)((n + ST - 1) / ST);
    int *bsum = (int *)temp;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SB), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(1024), 0, s, bsum, nb, d_total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SB), 0, s, in, out, n, (const int *)bsum);
    return hipGetLastError();
}
