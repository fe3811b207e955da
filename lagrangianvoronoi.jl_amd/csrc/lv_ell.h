// lv_ell.h -- row access to the slot-major ELL copies of the pressure operator.
//
// Slot k of row i sits at [k*n + i], deg[i] slots per row, so slot k of a wavefront's 64 rows is
// one contiguous run.  Two encodings:
//   * the operator's own (lv_pressure_system): int32 degrees, int32 column indices, FP64 or FP32
//     values, rows in particle order;
//   * the aggregate order of the multilevel preconditioner (lv_amg): rows renumbered so that the
//     particles of one level-1 cell are consecutive, uint8 degrees, int16 column offsets
//     (j - i, taken modulo n), FP32 values for the Krylov operator and BF16 values for the cycle.
// An ell_view couples the three arrays; its methods hide the encoding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ int lv_ell_deg(const int *d, int i) { return d[i]; }
__device__ __forceinline__ int lv_ell_deg(const uint8_t *d, int i) { return (int)d[i]; }

__device__ __forceinline__ int lv_ell_col(const int *c, size_t e, int, int) { return c[e]; }
__device__ __forceinline__ int lv_ell_col(const int16_t *c, size_t e, int i, int n)
{
    int j = i + (int)c[e];
    j += (j < 0) ? n : 0;
    j -= (j >= n) ? n : 0;
    return j;
}

__device__ __forceinline__ double lv_ell_val(const double *v, size_t e) { return v[e]; }
__device__ __forceinline__ float lv_ell_val(const float *v, size_t e) { return v[e]; }
// BF16 held as its 16 high bits
__device__ __forceinline__ float lv_ell_val(const uint16_t *v, size_t e)
{
    return __uint_as_float((uint32_t)v[e] << 16);
}

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t lv_bf16_bits(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename DEG, typename COL, typename VAL>
struct lv_ell_view {
    const DEG *deg;
    const COL *col;
    const VAL *val;
    __device__ __forceinline__ int degree(int i) const { return lv_ell_deg(deg, i); }
    __device__ __forceinline__ int column(size_t e, int i, int n) const { return lv_ell_col(col, e, i, n); }
    __device__ __forceinline__ auto value(size_t e) const { return lv_ell_val(val, e); }
};

// The first U slots of row i: columns and values, loaded unconditionally (the arrays hold LV_MAXE
// slots per row, so every load is in bounds) and replaced by (i, 0) beyond the row's degree, so
// the compiler issues all of them, and the gathers that follow, without waiting slot by slot.
// Returns the degree; slots U.. are the caller's.
template <int U, class V, typename T>
__device__ __forceinline__ int lv_ell_head(const V &A, int i, int n, int (&c)[U], T (&v)[U])
{
    const int d = A.degree(i);
#pragma unroll
    for (int k = 0; k < U; k++) {
        const size_t e = (size_t)k * n + i;
        const int ck = A.column(e, i, n);
        const T vk = (T)A.value(e);
        const bool on = k < d;
        c[k] = on ? ck : i;
        v[k] = on ? vk : (T)0;
    }
    return d;
}

// the same with the values taken from a second array on the same slots (w)
template <int U, class V, typename T>
__device__ __forceinline__ int lv_ell_head_with(const V &A, const T *__restrict__ w, int i, int n, int (&c)[U],
                                                T (&v)[U])
{
    const int d = A.degree(i);
#pragma unroll
    for (int k = 0; k < U; k++) {
        const size_t e = (size_t)k * n + i;
        const int ck = A.column(e, i, n);
        const T vk = w[e];
        const bool on = k < d;
        c[k] = on ? ck : i;
        v[k] = on ? vk : (T)0;
    }
    return d;
}

template <typename T>
using lv_ell_own = lv_ell_view<int, int, T>;  // the operator's own ELL
using lv_ell_agg32 = lv_ell_view<uint8_t, int16_t, float>;     // aggregate order, FP32 values
using lv_ell_agg16 = lv_ell_view<uint8_t, int16_t, uint16_t>;  // aggregate order, BF16 values
