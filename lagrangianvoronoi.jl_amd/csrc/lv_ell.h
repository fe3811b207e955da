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

// The aggregate order also keeps a compact copy of every slot past LV_OVF_BASE.  Rows are grouped
// by 64 (group g = rows 64g .. 64g+63, the rows of one wavefront in a grid-stride pass); the slots
// LV_OVF_BASE.. of a group's rows sit in row order at [g*LV_OVF_CAP + pos], pos counting the group's
// earlier overflow slots, when the group has at most ocap of them.  A wavefront then reads
// LV_OVF_BASE slots per row plus the overflow slots that exist, where reading slots up to 8 of every
// row read the padding of every row below degree 8 (12 bytes per row at the lattice's mean degree
// 6).  The slot-major arrays still hold every slot: groups with more overflow, and kernels whose
// wavefronts do not cover whole groups, read those.  The Krylov operator's pass reads the compact
// copy (59.7 against 62.7 us per launch at 2048^2); the cycle's BF16 passes do not, where it did
// not pay.
#define LV_OVF_BASE 6
#define LV_OVF_CAP 64

template <typename DEG, typename COL, typename VAL>
struct lv_ell_view {
    const DEG *deg;
    const COL *col;
    const VAL *val;
    const COL *ocol = nullptr;  // the compact overflow (aggregate order only)
    const VAL *oval = nullptr;
    int ocap = 0;               // groups with at most this many overflow slots use it (0: none)
    __device__ __forceinline__ int degree(int i) const { return lv_ell_deg(deg, i); }
    __device__ __forceinline__ int column(size_t e, int i, int n) const { return lv_ell_col(col, e, i, n); }
    __device__ __forceinline__ auto value(size_t e) const { return lv_ell_val(val, e); }
};

// Where row i's overflow slots start in its group, and the group's total: the exclusive prefix and
// the sum of max(d - LV_OVF_BASE, 0) over the wavefront's lanes, by one ballot per overflow depth.
// Every lane of the wavefront that holds a row of the group must call it (d = 0 for rows past n).
struct lv_ovf_pos {
    int pos, tot;
};
__device__ __forceinline__ lv_ovf_pos lv_ovf_prefix(int d)
{
    const int k = d > LV_OVF_BASE ? d - LV_OVF_BASE : 0;
    lv_ovf_pos q{0, 0};
    for (int j = 1; j <= 16; j++) {
        const unsigned long long m = __ballot(k >= j);
        if (m == 0ull) break;
        q.pos += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        q.tot += __popcll(m);
    }
    return q;
}

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

// di x_i + sum_k v_k (x_i - x_{c_k}) over row i's d slots (the reference's difference form, slots in
// order), from the LV_OVF_BASE slot-major slots and the compact overflow when the group's q.tot
// fits (else slots LV_OVF_BASE.. slot-major).  The first 8 slots are loaded unconditionally, as in
// lv_ell_head; the overflow entries of rows below degree 8 are clamped into the group's region and
// dropped.
// (with shift, every gathered x_j is read as x_j - shift, for a product with the stored x - shift;
// shift 0 leaves the values exactly as they are)
template <typename T, class V>
__device__ __forceinline__ T lv_ell_row_ovf(const V &A, int i, int n, int d, lv_ovf_pos q, T di, T xi,
                                            const T *__restrict__ x, T shift = (T)0)
{
    constexpr int U = 8;
    int c[U];
    T v[U], xj[U];
#pragma unroll
    for (int k = 0; k < LV_OVF_BASE; k++) {
        const size_t e = (size_t)k * n + i;
        const int ck = A.column(e, i, n);
        const T vk = (T)A.value(e);
        const bool on = k < d;
        c[k] = on ? ck : i;
        v[k] = on ? vk : (T)0;
    }
    const bool compact = A.ocap > 0 && q.tot <= A.ocap;
    const size_t g0 = (size_t)(i >> 6) * LV_OVF_CAP;
    if (compact) {
#pragma unroll
        for (int k = LV_OVF_BASE; k < U; k++) {
            const int o = q.pos + (k - LV_OVF_BASE);
            const size_t e = g0 + (size_t)(o < LV_OVF_CAP ? o : LV_OVF_CAP - 1);
            const int ck = lv_ell_col(A.ocol, e, i, n);
            const T vk = (T)lv_ell_val(A.oval, e);
            const bool on = k < d;
            c[k] = on ? ck : i;
            v[k] = on ? vk : (T)0;
        }
    } else {
#pragma unroll
        for (int k = LV_OVF_BASE; k < U; k++) {
            const size_t e = (size_t)k * n + i;
            const int ck = A.column(e, i, n);
            const T vk = (T)A.value(e);
            const bool on = k < d;
            c[k] = on ? ck : i;
            v[k] = on ? vk : (T)0;
        }
    }
    T acc = di * xi;
#pragma unroll
    for (int k = 0; k < U; k++) xj[k] = x[c[k]] - shift;
#pragma unroll
    for (int k = 0; k < U; k++) acc += v[k] * (xi - xj[k]);
    for (int k = U; k < d; k++) {
        const size_t e = compact ? g0 + (size_t)(q.pos + (k - LV_OVF_BASE)) : (size_t)k * n + i;
        const int ck = compact ? lv_ell_col(A.ocol, e, i, n) : A.column(e, i, n);
        const T vk = compact ? (T)lv_ell_val(A.oval, e) : (T)A.value(e);
        acc += vk * (xi - (x[ck] - shift));
    }
    return acc;
}

// views that can carry the compact overflow (every ELL view; ocap == 0 when it is absent)
template <class V>
struct lv_ell_compact {
    static constexpr bool value = false;
};
template <typename DEG, typename COL, typename VAL>
struct lv_ell_compact<lv_ell_view<DEG, COL, VAL>> {
    static constexpr bool value = true;
};

template <typename T>
using lv_ell_own = lv_ell_view<int, int, T>;  // the operator's own ELL
using lv_ell_agg32 = lv_ell_view<uint8_t, int16_t, float>;     // aggregate order, FP32 values
using lv_ell_agg16 = lv_ell_view<uint8_t, int16_t, uint16_t>;  // aggregate order, BF16 values
