// lv_spai.hip -- learned sparse-approximate-inverse preconditioner for the pressure system.
//
// No counterpart in the reference (which runs unpreconditioned MINRES, src/pressure.jl:219).
//
// Preconditioner.  With D_i the diagonal of A (d_i + sum_j c_ij) and a_ij = c_ij / sqrt(D_i D_j) the
// Jacobi-scaled couplings, the preconditioner of the scaled system lives on A's own pattern:
//     Mt_ij = a_ij exp((o_ij + o_ji)/2),   Mt_ii = sum_j |Mt_ij| + delta_i,   delta_i = r_i exp(o_i)
// with r_i = max(d_i/D_i, R_FLOOR).  It is symmetric and strictly diagonally dominant with a positive
// diagonal, hence SPD for any network output, so CG stays valid.  A zero output gives the first-order
// Neumann approximate inverse I + N of the scaled matrix.  M = D^-1/2 Mt D^-1/2 is applied as one
// ELL pass per CG iteration.
//
// Network (message passing over the directed Voronoi edges, width 64, 3 layers):
//     h_i  = ReLU(x_i Wn + bn)                      x_i = (r_i, log10(r_i)/12 + 1, deg_i/8, 1)
//     e_p  = ReLU(f_p We + be)                      f_p = (a_p, a_p deg_i/6, c_p/D_i, 1)
//     layer l:  m_p = ReLU([h_src ; h_dst ; e_p] W_l + b_l)   (BF16 MFMA 32x32x16, FP32 accumulate)
//               h_i += mean_{p in row i} m_p ;  e_p = m_p
//     o_p = e_p . ve + ce ,   o_i = h_i . un + cn
// Training minimises the relative spread  L = S / c^2,  S = (1/P) sum_q |Mt At v_q - c v_q|^2 / |v_q|^2,
// c = (1/P) sum_q v_q.Mt At v_q / |v_q|^2 (CG is invariant to the scale of M), over probe vectors
// v_q (the current pressure solution in scaled form plus Rademacher vectors), with Adam; the backward pass
// is hand-written (dX = dZ W^T and dW = X^T dZ on BF16 MFMA, FP32 accumulation).
#include <string.h>

#include "lv_spai.h"

namespace {

constexpr int B = 256;
constexpr int H = LV_SPAI_WIDTH;   // 64
constexpr int KX = 3 * H;          // 192: [h_src ; h_dst ; e]
constexpr float R_FLOOR = 1e-4f;
constexpr float OCLAMP = 8.0f;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 v8bf;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float v16f;

__device__ __forceinline__ __bf16 f2bf(float f) { return (__bf16)f; }
__device__ __forceinline__ float bf2f(__bf16 b) { return (float)b; }
__device__ __forceinline__ float clampo(float o) { return fminf(fmaxf(o, -OCLAMP), OCLAMP); }

// ---- graph and features ------------------------------------------------------------------------

// node features, r_i = d_i / D_i (floored), sqrt(D_i), and src of every directed edge of row i
__global__ __launch_bounds__(B) void k_graph_nodes(int n, const int *__restrict__ rp, const double *__restrict__ diag,
                                                   const double *__restrict__ dinv, int *__restrict__ src,
                                                   float *__restrict__ xf, float *__restrict__ r_out,
                                                   float *__restrict__ sqD)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double Di = 1.0 / dinv[i];
    const int p0 = rp[i], p1 = rp[i + 1];
    const float deg = (float)(p1 - p0);
    const double r = diag[i] / Di;
    r_out[i] = fmaxf((float)r, R_FLOOR);
    sqD[i] = (float)sqrt(Di);
    *(float4 *)(xf + 4 * i) = make_float4((float)r, log10f(fmaxf((float)r, 1e-12f)) / 12.0f + 1.0f, deg / 8.0f, 1.0f);
    for (int p = p0; p < p1; p++) src[p] = i;
}

// per directed edge p = (i, j): scaled coupling a_p = c_p / sqrt(D_i D_j), edge features, and the
// reverse edge (j, i) (-1 if row j does not list i)
__global__ __launch_bounds__(B) void k_graph_edges(int E, const int *__restrict__ rp, const int *__restrict__ cols,
                                                   const double *__restrict__ val, const double *__restrict__ dinv,
                                                   const int *__restrict__ src, int *__restrict__ rev,
                                                   float *__restrict__ a, float *__restrict__ ff)
{
    const int p = blockIdx.x * B + threadIdx.x;
    if (p >= E) return;
    const int i = src[p], j = cols[p];
    const double Di = 1.0 / dinv[i], Dj = 1.0 / dinv[j];
    const float deg = (float)(rp[i + 1] - rp[i]);
    const float ap = (float)(val[p] / sqrt(Di * Dj));
    a[p] = ap;
    *(float4 *)(ff + 4 * (size_t)p) = make_float4(ap, ap * deg / 6.0f, (float)(val[p] / Di), 1.0f);
    int q = -1;
    for (int s = rp[j]; s < rp[j + 1]; s++)
        if (cols[s] == i) {
            q = s;
            break;
        }
    rev[p] = q;
}

// ---- encoders (4 -> 64, FP32) ------------------------------------------------------------------

// out[t][c] = ReLU(sum_k in[t][k] W[k][c] + b[c]) for t < count; FP32 (optional) and BF16 copies.
// One thread per (row, 8 channels): 16-byte loads and stores.
__global__ __launch_bounds__(B) void k_encode(int count, const float *__restrict__ in, const float *__restrict__ W,
                                              const float *__restrict__ b, float *__restrict__ outf,
                                              __bf16 *__restrict__ outb)
{
    __shared__ __attribute__((aligned(16))) float sw[5 * H];  // W[4][64], then b[64]
    for (int k = threadIdx.x; k < 5 * H; k += B) sw[k] = (k < 4 * H) ? W[k] : b[k - 4 * H];
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)count * (H / 8)) return;
    const int64_t r = t / (H / 8);
    const int c0 = (int)(t % (H / 8)) * 8;
    const float4 x = *(const float4 *)(in + 4 * r);
    float y[8];
    v8bf o;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int c = c0 + j;
        float z = sw[4 * H + c];
        z += x.x * sw[c];
        z += x.y * sw[H + c];
        z += x.z * sw[2 * H + c];
        z += x.w * sw[3 * H + c];
        y[j] = fmaxf(z, 0.0f);
        o[j] = f2bf(y[j]);
    }
    if (outf) {
        *(float4 *)(outf + r * H + c0) = make_float4(y[0], y[1], y[2], y[3]);
        *(float4 *)(outf + r * H + c0 + 4) = make_float4(y[4], y[5], y[6], y[7]);
    }
    *(v8bf *)(outb + r * H + c0) = o;
}

// ---- message-passing layer, forward (BF16 MFMA) --------------------------------------------------

// One wavefront computes a 32-edge x 64-channel tile of m = ReLU([h_src ; h_dst ; e] W + b).
// Wt is W transposed, channel-major: Wt[c*192 + k] (BF16), so a lane's 8 k-values are contiguous.
__global__ __launch_bounds__(256) void k_mp_forward(int E, const int *__restrict__ src, const int *__restrict__ dst,
                                                    const __bf16 *__restrict__ Hb, const __bf16 *__restrict__ Ein,
                                                    const __bf16 *__restrict__ Wt, const float *__restrict__ bias,
                                                    __bf16 *__restrict__ Mout)
{
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t e0 = tile * 32;
    if (e0 >= E) return;
    const int r = lane & 31, half = lane >> 5;
    const int64_t p = e0 + r;
    const bool valid = p < E;
    const int s_node = valid ? src[p] : 0;
    const int d_node = valid ? dst[p] : 0;
    v16f acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int s = 0; s < KX / 16; s++) {
        const int k0 = 16 * s + 8 * half;
        v8bf av;
        if (valid) {
            const __bf16 *ap;
            if (k0 < H) ap = Hb + (size_t)s_node * H + k0;
            else if (k0 < 2 * H) ap = Hb + (size_t)d_node * H + (k0 - H);
            else ap = Ein + (size_t)p * H + (k0 - 2 * H);
            av = *(const v8bf *)ap;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) av[j] = f2bf(0.0f);
        }
        const v8bf b0 = *(const v8bf *)(Wt + (size_t)r * KX + k0);
        const v8bf b1 = *(const v8bf *)(Wt + (size_t)(r + 32) * KX + k0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int row = 8 * (i / 4) + 4 * half + (i % 4);
        const int64_t q = e0 + row;
        if (q < E) {
            Mout[q * H + r] = f2bf(fmaxf(acc0[i] + bias[r], 0.0f));
            Mout[q * H + r + 32] = f2bf(fmaxf(acc1[i] + bias[r + 32], 0.0f));
        }
    }
}

// h_new[i] = h[i] + mean_{p in row i} m[p]  (FP32 master, BF16 copy for the next layer).
// One thread per (node, 8 channels): 16-byte loads of the message rows.
__global__ __launch_bounds__(B) void k_aggregate(int n, const int *__restrict__ rp, const float *__restrict__ Hf,
                                                 const __bf16 *__restrict__ Mb, float *__restrict__ Hf_out,
                                                 __bf16 *__restrict__ Hb_out)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)n * (H / 8)) return;
    const int i = (int)(t / (H / 8));
    const int c0 = (int)(t % (H / 8)) * 8;
    const int p0 = rp[i], p1 = rp[i + 1];
    float sum[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int p = p0; p < p1; p++) {
        const v8bf m = *(const v8bf *)(Mb + (size_t)p * H + c0);
#pragma unroll
        for (int j = 0; j < 8; j++) sum[j] += bf2f(m[j]);
    }
    const size_t o = (size_t)i * H + c0;
    const float4 h0 = *(const float4 *)(Hf + o), h1 = *(const float4 *)(Hf + o + 4);
    const float hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
    float y[8];
    v8bf yb;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        y[j] = hv[j] + ((p1 > p0) ? sum[j] / (float)(p1 - p0) : 0.0f);
        yb[j] = f2bf(y[j]);
    }
    *(float4 *)(Hf_out + o) = make_float4(y[0], y[1], y[2], y[3]);
    *(float4 *)(Hf_out + o + 4) = make_float4(y[4], y[5], y[6], y[7]);
    *(v8bf *)(Hb_out + o) = yb;
}

// ---- message-passing layer, inference (BF16 MFMA, persistent waves) ----------------------------

// The transposed product m^T = W^T X^T on v_mfma_f32_32x32x16_bf16: the A operand is a W^T fragment
// (lane l: channel l%32 (+32), k = 16s + 8(l/32) .. +7) and the B operand the gathered feature row of
// edge l%32 at the same k, so lane l ends up holding edge l%32 and channels 8g + 4(l/32) + j (g, j < 4)
// of each 32-channel half.  W^T (padded rows) lives in LDS for the whole kernel; every wave loops
// over 32-edge tiles.  Non-final layers stage the BF16 tile through LDS and write it with 16-byte
// stores; the final layer writes only what the decoders need: o_p = m_p . ve + ce and
// s_p = m_p . un (the node decoder adds the mean of s_p over the row to h . un + cn).
constexpr int WPAD = KX + 8;  // LDS row of W^T (bf16)
constexpr int TPAD = H + 8;   // LDS row of a staged edge (bf16)

typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 v4bf;

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_mp_infer(int E, const int *__restrict__ src, const int *__restrict__ dst,
                                                  const __bf16 *__restrict__ Hb, const __bf16 *__restrict__ Ein,
                                                  const float *__restrict__ ff, const float *__restrict__ We,
                                                  const float *__restrict__ be,
                                                  const __bf16 *__restrict__ Wt, const float *__restrict__ bias,
                                                  __bf16 *__restrict__ Mout, const float *__restrict__ ve,
                                                  const float *__restrict__ ce, const float *__restrict__ un,
                                                  float *__restrict__ oe, float *__restrict__ sn)
{
    __shared__ __attribute__((aligned(16))) __bf16 sW[H * WPAD];
    __shared__ __attribute__((aligned(16))) __bf16 sT[4 * 32 * TPAD];
    __shared__ __attribute__((aligned(16))) float sB[H];
    __shared__ __attribute__((aligned(16))) float sV[H];
    __shared__ __attribute__((aligned(16))) float sU[H];
    __shared__ __attribute__((aligned(16))) float sE[5 * H];  // edge encoder We[4][64], be[64] (FIRST)
    if (FIRST)
        for (int k = threadIdx.x; k < 5 * H; k += 256) sE[k] = (k < 4 * H) ? We[k] : be[k - 4 * H];
    for (int t = threadIdx.x; t < H * (KX / 8); t += 256) {
        const int c = t / (KX / 8), k = (t % (KX / 8)) * 8;
        *(v8bf *)(sW + c * WPAD + k) = *(const v8bf *)(Wt + c * KX + k);
    }
    if (threadIdx.x < H) {
        sB[threadIdx.x] = bias[threadIdx.x];
        if (LAST) {
            sV[threadIdx.x] = ve[threadIdx.x];
            sU[threadIdx.x] = un[threadIdx.x];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 31, half = lane >> 5;
    __bf16 *T = sT + wave * 32 * TPAD;
    const int64_t ntiles = ((int64_t)E + 31) / 32;
    const int64_t tstride = (int64_t)gridDim.x * 4;
    // software pipeline: the next tile's node indices are loaded before the MFMAs of the current
    // tile and its feature fragments right after them, so both latencies overlap computation
    int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    v8bf X[KX / 16];
    int sn_next = 0, dn_next = 0;
    auto load_idx = [&](int64_t t) {
        const int64_t q = t * 32 + r;
        const bool ok = t < ntiles && q < E;
        sn_next = ok ? src[q] : 0;
        dn_next = ok ? dst[q] : 0;
    };
    auto load_feats = [&](int64_t t) {
        const int64_t q = t * 32 + r;
        const bool ok = t < ntiles && q < E;
        float4 fq = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (FIRST && ok) fq = *(const float4 *)(ff + 4 * q);
#pragma unroll
        for (int s = 0; s < KX / 16; s++) {
            const int k0 = 16 * s + 8 * half;
            if (FIRST && ok && k0 >= 2 * H) {
                // e_0 = ReLU(f We + be), the same operation order as k_encode
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int c = k0 - 2 * H + j;
                    float z = sE[4 * H + c];
                    z += fq.x * sE[c];
                    z += fq.y * sE[H + c];
                    z += fq.z * sE[2 * H + c];
                    z += fq.w * sE[3 * H + c];
                    X[s][j] = f2bf(fmaxf(z, 0.0f));
                }
            } else if (ok) {
                const __bf16 *ap;
                if (k0 < H) ap = Hb + (size_t)sn_next * H + k0;
                else if (k0 < 2 * H) ap = Hb + (size_t)dn_next * H + (k0 - H);
                else ap = Ein + (size_t)q * H + (k0 - 2 * H);
                X[s] = *(const v8bf *)ap;
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) X[s][j] = f2bf(0.0f);
            }
        }
    };
    load_idx(tile);
    load_feats(tile);
    for (; tile < ntiles; tile += tstride) {
        const int64_t e0 = tile * 32;
        const int64_t p = e0 + r;
        const bool valid = p < E;
        const int64_t next = tile + tstride;
        load_idx(next);
        v16f acc0 = {0}, acc1 = {0};
#pragma unroll
        for (int s = 0; s < KX / 16; s++) {
            const int k0 = 16 * s + 8 * half;
            const v8bf w0 = *(const v8bf *)(sW + r * WPAD + k0);
            const v8bf w1 = *(const v8bf *)(sW + (r + 32) * WPAD + k0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, X[s], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, X[s], acc1, 0, 0, 0);
        }
        load_feats(next);
        if (!LAST) {
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int c = 8 * g + 4 * half;
                const float4 b0 = *(const float4 *)(sB + c), b1 = *(const float4 *)(sB + c + 32);
                const float bb0[4] = {b0.x, b0.y, b0.z, b0.w}, bb1[4] = {b1.x, b1.y, b1.z, b1.w};
                v4bf o0, o1;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    o0[j] = f2bf(fmaxf(acc0[4 * g + j] + bb0[j], 0.0f));
                    o1[j] = f2bf(fmaxf(acc1[4 * g + j] + bb1[j], 0.0f));
                }
                *(v4bf *)(T + r * TPAD + c) = o0;
                *(v4bf *)(T + r * TPAD + c + 32) = o1;
            }
            // the staged tile is read back by other lanes of this wave: LDS operations of one wave
            // complete in issue order, so only the compiler must not move accesses across here
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int row = 8 * k + (lane >> 3), c = (lane & 7) * 8;
                const int64_t q = e0 + row;
                const v8bf v = *(const v8bf *)(T + row * TPAD + c);
                if (q < E) *(v8bf *)(Mout + q * H + c) = v;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        } else {
            float po = 0.0f, ps = 0.0f;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int c = 8 * g + 4 * half;
                const float4 b0 = *(const float4 *)(sB + c), b1 = *(const float4 *)(sB + c + 32);
                const float4 v0 = *(const float4 *)(sV + c), v1 = *(const float4 *)(sV + c + 32);
                const float4 u0 = *(const float4 *)(sU + c), u1 = *(const float4 *)(sU + c + 32);
                const float bb0[4] = {b0.x, b0.y, b0.z, b0.w}, bb1[4] = {b1.x, b1.y, b1.z, b1.w};
                const float vv0[4] = {v0.x, v0.y, v0.z, v0.w}, vv1[4] = {v1.x, v1.y, v1.z, v1.w};
                const float uu0[4] = {u0.x, u0.y, u0.z, u0.w}, uu1[4] = {u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float m0 = bf2f(f2bf(fmaxf(acc0[4 * g + j] + bb0[j], 0.0f)));
                    const float m1 = bf2f(f2bf(fmaxf(acc1[4 * g + j] + bb1[j], 0.0f)));
                    po += m0 * vv0[j];
                    ps += m0 * uu0[j];
                    po += m1 * vv1[j];
                    ps += m1 * uu1[j];
                }
            }
            po += __shfl_xor(po, 32);
            ps += __shfl_xor(ps, 32);
            if (half == 0 && valid) {
                oe[p] = po + ce[0];
                sn[p] = ps;
            }
        }
    }
}

// o_i = h_i . un + cn + mean_{p in row i} s_p (inference node decoder; eight lanes per node)
__global__ __launch_bounds__(B) void k_node_decode(int n, const int *__restrict__ rp, const float *__restrict__ Hf,
                                                   const float *__restrict__ sn, const float *__restrict__ un,
                                                   const float *__restrict__ cn, float *__restrict__ on)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    const int i = (int)(t >> 3), c0 = (int)(t & 7) * 8;
    float d = 0.0f;
    if (i < n) {
        const float4 h0 = *(const float4 *)(Hf + (size_t)i * H + c0);
        const float4 h1 = *(const float4 *)(Hf + (size_t)i * H + c0 + 4);
        d = h0.x * un[c0] + h0.y * un[c0 + 1] + h0.z * un[c0 + 2] + h0.w * un[c0 + 3] + h1.x * un[c0 + 4] +
            h1.y * un[c0 + 5] + h1.z * un[c0 + 6] + h1.w * un[c0 + 7];
    }
    d += __shfl_xor(d, 1);
    d += __shfl_xor(d, 2);
    d += __shfl_xor(d, 4);
    if (i < n && (t & 7) == 0) {
        const int p0 = rp[i], p1 = rp[i + 1];
        float s = 0.0f;
        for (int p = p0; p < p1; p++) s += sn[p];
        on[i] = d + cn[0] + ((p1 > p0) ? s / (float)(p1 - p0) : 0.0f);
    }
}

// ---- decoders and the preconditioner ----------------------------------------------------------

// o_p = e_p . ve + ce (edges), o_i = h_i . un + cn (nodes)
__global__ __launch_bounds__(B) void k_decode(int count, const __bf16 *__restrict__ Xb, const float *__restrict__ Xf,
                                              const float *__restrict__ v, const float *__restrict__ c0,
                                              float *__restrict__ o)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= count) return;
    float s = c0[0];
    if (Xf) {
        for (int c = 0; c < H; c++) s += Xf[t * H + c] * v[c];
    } else {
        for (int c = 0; c < H; c++) s += bf2f(Xb[t * H + c]) * v[c];
    }
    o[t] = s;
}

// w_p = a_p exp((o_p + o_rev)/2); mdiag_i = sum |w| + r_i exp(o_i)
__global__ __launch_bounds__(B) void k_build_mt(int n, const int *__restrict__ rp, const int *__restrict__ rev,
                                                const float *__restrict__ a, const float *__restrict__ oe,
                                                const float *__restrict__ on, const float *__restrict__ rfl,
                                                float *__restrict__ w, float *__restrict__ mdiag,
                                                float *__restrict__ delta)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    float sw = 0.0f;
    for (int p = rp[i]; p < rp[i + 1]; p++) {
        const int q = rev[p];
        const float os = 0.5f * (oe[p] + ((q >= 0) ? oe[q] : oe[p]));
        const float wp = a[p] * expf(clampo(os));
        w[p] = wp;
        sw += fabsf(wp);
    }
    const float dl = rfl[i] * expf(clampo(on[i]));
    delta[i] = dl;
    mdiag[i] = sw + dl;
}

// M = D^-1/2 Mt D^-1/2 written into the solver's ELL slots (A's ELL copy is sorted by column)
__global__ __launch_bounds__(B) void k_write_m(int n, const int *__restrict__ rp, const int *__restrict__ cols,
                                               const int *__restrict__ deg, const int *__restrict__ ecol,
                                               const float *__restrict__ w, const float *__restrict__ mdiag,
                                               const float *__restrict__ sqD, float *__restrict__ mval,
                                               float *__restrict__ mdg)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const float si = sqD[i];
    const int d = deg[i];
    for (int p = rp[i]; p < rp[i + 1]; p++) {
        const int j = cols[p];
        for (int k = 0; k < d; k++) {
            const size_t e = (size_t)k * n + i;
            if (ecol[e] == j) {
                mval[e] = w[p] / (si * sqD[j]);
                break;
            }
        }
    }
    mdg[i] = mdiag[i] / (si * si);
}

// ---- loss and its gradient w.r.t. the preconditioner entries ------------------------------------

// hash of (seed, index) -> +-1
__device__ __forceinline__ float rademacher(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (z >> 63) ? 1.0f : -1.0f;
}

// probe q, node i at v[q*n + i]: q = 0 is the scaled data vector (if given), the rest Rademacher
__global__ __launch_bounds__(B) void k_probes(int n, int P, const double *__restrict__ x, const float *__restrict__ sqD,
                                              uint64_t seed, float *__restrict__ v)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)n * P) return;
    const int q = (int)(t / n), i = (int)(t % n);
    if (q == 0 && x) v[t] = (float)(x[i]) * sqD[i];
    else v[t] = rademacher(seed, (uint64_t)t);
}

// y = Mt (At v) for all probes; also u = At v (kept for the gradient)
__global__ __launch_bounds__(B) void k_apply_at(int n, int P, const int *__restrict__ rp, const int *__restrict__ cols,
                                                const float *__restrict__ a, const float *__restrict__ v,
                                                float *__restrict__ u)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)n * P) return;
    const int q = (int)(t / n), i = (int)(t % n);
    const float *vq = v + (size_t)q * n;
    float s = vq[i];
    for (int p = rp[i]; p < rp[i + 1]; p++) s -= a[p] * vq[cols[p]];
    u[t] = s;
}

__global__ __launch_bounds__(B) void k_apply_mt(int n, int P, const int *__restrict__ rp, const int *__restrict__ cols,
                                                const float *__restrict__ w, const float *__restrict__ mdiag,
                                                const float *__restrict__ u, float *__restrict__ y)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)n * P) return;
    const int q = (int)(t / n), i = (int)(t % n);
    const float *uq = u + (size_t)q * n;
    float s = mdiag[i] * uq[i];
    for (int p = rp[i]; p < rp[i + 1]; p++) s += w[p] * uq[cols[p]];
    y[t] = s;
}

// per-probe partials (FP64) of |y|^2, v.y and |v|^2, one block row per probe
__global__ __launch_bounds__(B) void k_loss_partials(int n, int P, const float *__restrict__ y,
                                                     const float *__restrict__ v, double *__restrict__ part, int nb)
{
    __shared__ double s0[B / 64], s1[B / 64], s2[B / 64];
    const int q = blockIdx.y;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += nb * B) {
        const double yy = (double)y[(size_t)q * n + i];
        const double vv = (double)v[(size_t)q * n + i];
        a += yy * yy;
        b += vv * yy;
        c += vv * vv;
    }
    a = lv_wave_sum(a);
    b = lv_wave_sum(b);
    c = lv_wave_sum(c);
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a;
        s1[threadIdx.x >> 6] = b;
        s2[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0, tc = 0.0;
        for (int k = 0; k < B / 64; k++) {
            ta += s0[k];
            tb += s1[k];
            tc += s2[k];
        }
        part[(size_t)(3 * q) * nb + blockIdx.x] = ta;
        part[(size_t)(3 * q + 1) * nb + blockIdx.x] = tb;
        part[(size_t)(3 * q + 2) * nb + blockIdx.x] = tc;
    }
}

// Relative spread of Mt At over the probes (invariant to the scale of Mt, as CG is):
//   c = (1/P) sum_q v_q.y_q / |v_q|^2,   S = (1/P) sum_q |y_q - c v_q|^2 / |v_q|^2,   L = S / c^2.
// out[0] = L, out[1] = c, out[2] = S, out[3+q] = 1/|v_q|^2
__global__ __launch_bounds__(64) void k_loss_final(int P, int nb, const double *__restrict__ part,
                                                   double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double yy[64], vy[64], vv[64];
    double c = 0.0;
    for (int q = 0; q < P; q++) {
        double a = 0.0, b = 0.0, d = 0.0;
        for (int k = 0; k < nb; k++) {
            a += part[(size_t)(3 * q) * nb + k];
            b += part[(size_t)(3 * q + 1) * nb + k];
            d += part[(size_t)(3 * q + 2) * nb + k];
        }
        yy[q] = a;
        vy[q] = b;
        vv[q] = d;
        c += (d > 0.0) ? b / d : 0.0;
    }
    c /= P;
    double S = 0.0;
    for (int q = 0; q < P; q++) {
        const double inv = (vv[q] > 0.0) ? 1.0 / vv[q] : 0.0;
        S += (yy[q] - 2.0 * c * vy[q] + c * c * vv[q]) * inv;
        out[3 + q] = inv;
    }
    S /= P;
    out[0] = S / (c * c);
    out[1] = c;
    out[2] = S;
}

// gy = dL/dy = 2 / (P c^2 |v|^2) * (y - (c + S/c) v)   (dS/dc = 0 at the optimal c)
__global__ __launch_bounds__(B) void k_grad_y(int n, int P, const float *__restrict__ y, const float *__restrict__ v,
                                              const double *__restrict__ lossv, float *__restrict__ gy)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)n * P) return;
    const int q = (int)(t / n);
    const double c = lossv[1], S = lossv[2];
    const float scale = (float)(2.0 / (P * c * c) * lossv[3 + q]);
    gy[t] = scale * (y[t] - (float)(c + S / c) * v[t]);
}

// G_p = sum_q gy_i u_dst + gdiag_i sign(w_p);  gdiag_i = sum_q gy_i u_i;  then
// do_p = (G_p + G_rev) w_p / 2 and do_i = gdiag_i delta_i  (G written first, then symmetrised)
__global__ __launch_bounds__(B) void k_grad_w(int n, int P, const int *__restrict__ rp, const int *__restrict__ cols,
                                              const float *__restrict__ w, const float *__restrict__ gy,
                                              const float *__restrict__ u, const float *__restrict__ delta,
                                              float *__restrict__ G, float *__restrict__ don)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    float gd = 0.0f;
    for (int q = 0; q < P; q++) gd += gy[(size_t)q * n + i] * u[(size_t)q * n + i];
    for (int p = rp[i]; p < rp[i + 1]; p++) {
        const int j = cols[p];
        float g = 0.0f;
        for (int q = 0; q < P; q++) g += gy[(size_t)q * n + i] * u[(size_t)q * n + j];
        const float sg = (w[p] > 0.0f) ? 1.0f : ((w[p] < 0.0f) ? -1.0f : 0.0f);
        G[p] = g + gd * sg;
    }
    don[i] = gd * delta[i];
}

__global__ __launch_bounds__(B) void k_grad_o(int E, const int *__restrict__ rev, const float *__restrict__ G,
                                              const float *__restrict__ w, const float *__restrict__ oe,
                                              float *__restrict__ doe)
{
    const int p = blockIdx.x * B + threadIdx.x;
    if (p >= E) return;
    const int q = rev[p];
    const float os = 0.5f * (oe[p] + ((q >= 0) ? oe[q] : oe[p]));
    const float live = (fabsf(os) < OCLAMP) ? 1.0f : 0.0f;
    doe[p] = (q >= 0) ? 0.5f * (G[p] + G[q]) * w[p] * live : G[p] * w[p] * live;
}

__global__ __launch_bounds__(B) void k_clamp_mask(int n, const float *__restrict__ on, float *__restrict__ don)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n && !(fabsf(on[i]) < OCLAMP)) don[i] = 0.0f;
}

// ---- backward of the decoders -----------------------------------------------------------------

// dX[t][c] = do[t] * v[c] (+= if accumulate); partial sums of do[t]*X[t][c] and do[t] per block
__global__ __launch_bounds__(B) void k_decode_back(int count, const __bf16 *__restrict__ Xb,
                                                   const float *__restrict__ Xf, const float *__restrict__ v,
                                                   const float *__restrict__ dout, float *__restrict__ dX, int accum,
                                                   float *__restrict__ part /* [nb][H+1] */)
{
    __shared__ float sp[H + 1];
    if (threadIdx.x <= H) sp[threadIdx.x] = 0.0f;
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t < count) {
        const float g = dout[t];
        for (int c = 0; c < H; c++) {
            const float x = Xf ? Xf[t * H + c] : bf2f(Xb[t * H + c]);
            const float val = g * v[c];
            if (accum) dX[t * H + c] += val;
            else dX[t * H + c] = val;
            atomicAdd(&sp[c], g * x);
        }
        atomicAdd(&sp[H], g);
    }
    __syncthreads();
    if (threadIdx.x <= H) part[(size_t)blockIdx.x * (H + 1) + threadIdx.x] = sp[threadIdx.x];
}

// ---- backward of a message-passing layer ------------------------------------------------------

// dM[p] += dH[src p] / deg(src);  dZ = dM * (M > 0)  (BF16 copy for the MFMA)
__global__ __launch_bounds__(B) void k_mp_dz(int E, const int *__restrict__ src, const int *__restrict__ rp,
                                             const float *__restrict__ dH, const __bf16 *__restrict__ Mb,
                                             float *__restrict__ dM, __bf16 *__restrict__ dZb)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)E * H) return;
    const int64_t p = t / H;
    const int c = (int)(t % H);
    const int i = src[p];
    const float deg = (float)(rp[i + 1] - rp[i]);
    const float g = dM[t] + dH[(size_t)i * H + c] / deg;
    dZb[t] = f2bf((bf2f(Mb[t]) > 0.0f) ? g : 0.0f);
}

// dX = dZ W^T: a wavefront computes 32 edges x 192 features; Wk is W row-major (Wk[k*64 + c])
__global__ __launch_bounds__(256) void k_mp_dx(int E, const __bf16 *__restrict__ dZb, const __bf16 *__restrict__ Wk,
                                               float *__restrict__ dX)
{
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t e0 = tile * 32;
    if (e0 >= E) return;
    const int r = lane & 31, half = lane >> 5;
    const int64_t p = e0 + r;
    const bool valid = p < E;
    v16f acc[KX / 32];
#pragma unroll
    for (int t = 0; t < KX / 32; t++) acc[t] = (v16f){0};
#pragma unroll
    for (int s = 0; s < H / 16; s++) {
        const int k0 = 16 * s + 8 * half;
        v8bf av;
        if (valid) av = *(const v8bf *)(dZb + (size_t)p * H + k0);
        else {
#pragma unroll
            for (int j = 0; j < 8; j++) av[j] = f2bf(0.0f);
        }
#pragma unroll
        for (int t = 0; t < KX / 32; t++) {
            // B[k = channel][col = feature f] = W[f][k] = Wk[f*64 + k]
            const v8bf bv = *(const v8bf *)(Wk + (size_t)(32 * t + r) * H + k0);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < KX / 32; t++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int row = 8 * (i / 4) + 4 * half + (i % 4);
            const int64_t q = e0 + row;
            if (q < E) dX[q * KX + 32 * t + r] = acc[t][i];
        }
}

// dW partial over a chunk of 64 edges per wavefront: dW[f][c] = sum_p X[p][f] dZ[p][c].
// Each wavefront owns one 32x32 (f, c) tile for its chunk; 12 tiles cover 192 x 64.
// Partials land in part[chunk][192*64].
__global__ __launch_bounds__(256) void k_mp_dw(int E, const int *__restrict__ src, const int *__restrict__ dst,
                                               const __bf16 *__restrict__ Hb, const __bf16 *__restrict__ Ein,
                                               const __bf16 *__restrict__ dZb, int chunk, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.y * 4 + (threadIdx.x >> 6);   // 0..11 (3 blocks of 4 waves)
    const int ft = wave / 2, ct = wave % 2;                  // f tile 0..5, c tile 0..1
    const int64_t c0 = (int64_t)blockIdx.x * chunk;
    const int r = lane & 31, half = lane >> 5;
    v16f acc = {0};
    const int f = 32 * ft + r;  // this lane's A row (feature)
    for (int64_t s = 0; s < chunk; s += 16) {
        const int64_t e = c0 + s + 8 * half;  // this lane's 8 edges (k)
        v8bf av, bv;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int64_t p = e + j;
            float x = 0.0f, g = 0.0f;
            if (p < E) {
                if (f < H) x = bf2f(Hb[(size_t)src[p] * H + f]);
                else if (f < 2 * H) x = bf2f(Hb[(size_t)dst[p] * H + (f - H)]);
                else x = bf2f(Ein[(size_t)p * H + (f - 2 * H)]);
                g = bf2f(dZb[(size_t)p * H + 32 * ct + r]);
            }
            av[j] = f2bf(x);
            bv[j] = f2bf(g);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    float *o = part + (size_t)blockIdx.x * (KX * H);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int row = 8 * (i / 4) + 4 * half + (i % 4);
        o[(size_t)(32 * ft + row) * H + 32 * ct + r] = acc[i];
    }
}

// db partial: column sums of dZ per block of edges
__global__ __launch_bounds__(H) void k_mp_db(int E, const __bf16 *__restrict__ dZb, int chunk, float *__restrict__ part)
{
    const int c = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * chunk;
    float s = 0.0f;
    for (int64_t p = c0; p < c0 + chunk && p < E; p++) s += bf2f(dZb[(size_t)p * H + c]);
    part[(size_t)blockIdx.x * H + c] = s;
}

// dH_l[i] = dH_{l+1}[i] + sum_{p in row i} dX[p][0:64] + sum_{p in row i} dX[rev p][64:128];
// dE_l[p] = dX[p][128:192]
__global__ __launch_bounds__(B) void k_mp_dh(int n, const int *__restrict__ rp, const int *__restrict__ rev,
                                             const float *__restrict__ dX, float *__restrict__ dH)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)n * H) return;
    const int i = (int)(t / H), c = (int)(t % H);
    float s = dH[t];
    for (int p = rp[i]; p < rp[i + 1]; p++) {
        s += dX[(size_t)p * KX + c];
        const int q = rev[p];
        if (q >= 0) s += dX[(size_t)q * KX + H + c];
    }
    dH[t] = s;
}

__global__ __launch_bounds__(B) void k_mp_de(int E, const float *__restrict__ dX, float *__restrict__ dE)
{
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t >= (int64_t)E * H) return;
    const int64_t p = t / H;
    const int c = (int)(t % H);
    dE[t] = dX[p * KX + 2 * H + c];
}

// encoder backward partials: dW[k][c] = sum_t in[t][k] dZ[t][c], db[c] = sum_t dZ[t][c], dZ = dOut*(Out>0)
__global__ __launch_bounds__(B) void k_encode_back(int count, const float *__restrict__ in,
                                                   const __bf16 *__restrict__ outb, const float *__restrict__ dout,
                                                   float *__restrict__ part /* [nb][5*H] */)
{
    __shared__ float sp[5 * H];
    for (int k = threadIdx.x; k < 5 * H; k += B) sp[k] = 0.0f;
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * B + threadIdx.x;
    if (t < (int64_t)count * H) {
        const int64_t r = t / H;
        const int c = (int)(t % H);
        const float g = (bf2f(outb[t]) > 0.0f) ? dout[t] : 0.0f;
        if (g != 0.0f) {
#pragma unroll
            for (int k = 0; k < 4; k++) atomicAdd(&sp[k * H + c], in[4 * r + k] * g);
            atomicAdd(&sp[4 * H + c], g);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 5 * H; k += B) part[(size_t)blockIdx.x * 5 * H + k] = sp[k];
}

// out[k] (+)= sum_b part[b][k]  (fixed order: deterministic)
__global__ __launch_bounds__(B) void k_sum_parts(int nb, int width, const float *__restrict__ part,
                                                 float *__restrict__ out, int accum)
{
    const int k = blockIdx.x * B + threadIdx.x;
    if (k >= width) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[(size_t)b * width + k];
    out[k] = accum ? out[k] + (float)s : (float)s;
}

// ---- parameters ------------------------------------------------------------------------------

// Adam (Kingma & Ba 2015) with bias correction; FP32 master parameters
__global__ __launch_bounds__(B) void k_adam(int count, float *__restrict__ theta, const float *__restrict__ g,
                                            float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                                            float c1, float c2, float eps)
{
    const int k = blockIdx.x * B + threadIdx.x;
    if (k >= count) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// BF16 copies of the layer weights: Wt[c][k] (forward B operand) and Wk[k][c] (backward)
__global__ __launch_bounds__(B) void k_pack_w(const float *__restrict__ W, __bf16 *__restrict__ Wt,
                                              __bf16 *__restrict__ Wk)
{
    const int t = blockIdx.x * B + threadIdx.x;
    if (t >= KX * H) return;
    const int k = t / H, c = t % H;
    const __bf16 b = f2bf(W[t]);
    Wk[t] = b;
    Wt[(size_t)c * KX + k] = b;
}

inline dim3 g1(int64_t count) { return dim3((unsigned)std::max<int64_t>(1, (count + B - 1) / B)); }

}  // namespace

// ================================================================================================
// host side
// ================================================================================================

#define SP_CHECK(expr)                      \
    do {                                    \
        hipError_t _e = (expr);             \
        if (_e != hipSuccess) return _e;    \
    } while (0)

template <typename T>
static hipError_t grow(T **p, size_t *cap, size_t want)
{
    if (*cap >= want && *p) return hipSuccess;
    if (*p) SP_CHECK(hipFree(*p));
    *p = nullptr;
    SP_CHECK(hipMalloc((void **)p, sizeof(T) * (want ? want : 1)));
    *cap = want;
    return hipSuccess;
}

static void launch_graph(lv_spai *S, const lv_pressure_system &A, int n, int E, hipStream_t s)
{
    hipLaunchKernelGGL(k_graph_nodes, g1(n), dim3(B), 0, s, n, (const int *)A.row_ptr, (const double *)A.diag64,
                       (const double *)A.dinv64, S->src, S->xf, S->rfl, S->sqD);
    hipLaunchKernelGGL(k_graph_edges, g1(E), dim3(B), 0, s, E, (const int *)A.row_ptr, (const int *)A.cols,
                       (const double *)A.val64, (const double *)A.dinv64, (const int *)S->src, S->rev, S->a, S->ff);
}

int64_t lv_spai_param_count() { return LV_SPAI_NPARAM; }

hipError_t lv_spai_init(lv_spai *S)
{
    memset(S, 0, sizeof(*S));
    SP_CHECK(hipMalloc(&S->theta, sizeof(float) * LV_SPAI_NPARAM));
    SP_CHECK(hipMalloc(&S->grad, sizeof(float) * LV_SPAI_NPARAM));
    SP_CHECK(hipMalloc(&S->adam_m, sizeof(float) * LV_SPAI_NPARAM));
    SP_CHECK(hipMalloc(&S->adam_v, sizeof(float) * LV_SPAI_NPARAM));
    SP_CHECK(hipMalloc(&S->Wt, sizeof(__bf16) * 3 * KX * H));
    SP_CHECK(hipMalloc(&S->Wk, sizeof(__bf16) * 3 * KX * H));
    SP_CHECK(hipMalloc(&S->scal, sizeof(double) * 64));
    SP_CHECK(hipMemset(S->adam_m, 0, sizeof(float) * LV_SPAI_NPARAM));
    SP_CHECK(hipMemset(S->adam_v, 0, sizeof(float) * LV_SPAI_NPARAM));
    S->has_params = 0;
    return hipSuccess;
}

void lv_spai_free(lv_spai *S)
{
    void *ptrs[] = {S->theta, S->grad, S->adam_m, S->adam_v, S->Wt, S->Wk, S->scal, S->src, S->rev, S->a, S->xf,
                    S->ff, S->rfl, S->sqD, S->Hf, S->Hb, S->Eb, S->oe, S->on, S->w, S->mdiag, S->delta, S->v, S->u,
                    S->y, S->gy, S->G, S->doe, S->don, S->dE, S->dH, S->dX, S->dZb, S->part, S->mval, S->mdg, S->lpart};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    memset(S, 0, sizeof(*S));
}

hipError_t lv_spai_set_params(lv_spai *S, const float *host, hipStream_t s)
{
    SP_CHECK(hipMemcpyAsync(S->theta, host, sizeof(float) * LV_SPAI_NPARAM, hipMemcpyHostToDevice, s));
    for (int l = 0; l < 3; l++)
        hipLaunchKernelGGL(k_pack_w, g1(KX * H), dim3(B), 0, s, S->theta + LV_SPAI_OFF_W(l), S->Wt + l * KX * H,
                           S->Wk + l * KX * H);
    SP_CHECK(hipMemsetAsync(S->adam_m, 0, sizeof(float) * LV_SPAI_NPARAM, s));
    SP_CHECK(hipMemsetAsync(S->adam_v, 0, sizeof(float) * LV_SPAI_NPARAM, s));
    S->adam_t = 0;
    S->has_params = 1;
    return hipGetLastError();
}

hipError_t lv_spai_get_params(lv_spai *S, float *host, hipStream_t s)
{
    SP_CHECK(hipMemcpyAsync(host, S->theta, sizeof(float) * LV_SPAI_NPARAM, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

// Graph, features and the forward pass on the assembled operator; leaves w, mdiag, delta, and the
// activations of every layer (kept for the backward pass).
hipError_t lv_spai_forward(lv_spai *S, const lv_pressure_system &A, int n, hipStream_t s)
{
    const int E = (int)A.nnz;
    S->n = n;
    S->E = E;
    SP_CHECK(grow(&S->src, &S->cap_src, (size_t)E));
    SP_CHECK(grow(&S->rev, &S->cap_rev, (size_t)E));
    SP_CHECK(grow(&S->a, &S->cap_a, (size_t)E));
    SP_CHECK(grow(&S->ff, &S->cap_ff, (size_t)4 * E));
    SP_CHECK(grow(&S->oe, &S->cap_oe, (size_t)E));
    SP_CHECK(grow(&S->w, &S->cap_w, (size_t)E));
    SP_CHECK(grow(&S->xf, &S->cap_xf, (size_t)4 * n));
    SP_CHECK(grow(&S->rfl, &S->cap_rfl, (size_t)n));
    SP_CHECK(grow(&S->sqD, &S->cap_sqD, (size_t)n));
    SP_CHECK(grow(&S->on, &S->cap_on, (size_t)n));
    SP_CHECK(grow(&S->mdiag, &S->cap_mdiag, (size_t)n));
    SP_CHECK(grow(&S->delta, &S->cap_delta, (size_t)n));
    SP_CHECK(grow(&S->Hf, &S->cap_Hf, (size_t)4 * n * H));   // levels 0..3
    SP_CHECK(grow(&S->Hb, &S->cap_Hb, (size_t)4 * n * H));
    SP_CHECK(grow(&S->Eb, &S->cap_Eb, (size_t)4 * E * H));   // e_0 .. e_3
    launch_graph(S, A, n, E, s);
    const float *th = S->theta;
    const size_t nh = (size_t)n * H, eh = (size_t)E * H;
    hipLaunchKernelGGL(k_encode, g1((int64_t)n * (H / 8)), dim3(B), 0, s, n, (const float *)S->xf,
                       th + LV_SPAI_OFF_WN, th + LV_SPAI_OFF_BN, S->Hf, S->Hb);
    hipLaunchKernelGGL(k_encode, g1((int64_t)E * (H / 8)), dim3(B), 0, s, E, (const float *)S->ff,
                       th + LV_SPAI_OFF_WE, th + LV_SPAI_OFF_BE, (float *)nullptr, S->Eb);
    const unsigned tiles = (unsigned)((E + 31) / 32);
    for (int l = 0; l < 3; l++) {
        hipLaunchKernelGGL(k_mp_forward, dim3((tiles + 3) / 4), dim3(256), 0, s, E, (const int *)S->src,
                           (const int *)A.cols, (const __bf16 *)(S->Hb + l * nh), (const __bf16 *)(S->Eb + l * eh),
                           (const __bf16 *)(S->Wt + l * KX * H), th + LV_SPAI_OFF_B(l), S->Eb + (l + 1) * eh);
        hipLaunchKernelGGL(k_aggregate, g1((int64_t)n * (H / 8)), dim3(B), 0, s, n, (const int *)A.row_ptr,
                           (const float *)(S->Hf + l * nh), (const __bf16 *)(S->Eb + (l + 1) * eh),
                           S->Hf + (l + 1) * nh, S->Hb + (l + 1) * nh);
    }
    hipLaunchKernelGGL(k_decode, g1(E), dim3(B), 0, s, E, (const __bf16 *)(S->Eb + 3 * eh), (const float *)nullptr,
                       th + LV_SPAI_OFF_VE, th + LV_SPAI_OFF_CE, S->oe);
    hipLaunchKernelGGL(k_decode, g1(n), dim3(B), 0, s, n, (const __bf16 *)nullptr, (const float *)(S->Hf + 3 * nh),
                       th + LV_SPAI_OFF_UN, th + LV_SPAI_OFF_CN, S->on);
    hipLaunchKernelGGL(k_build_mt, g1(n), dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)S->rev,
                       (const float *)S->a, (const float *)S->oe, (const float *)S->on, (const float *)S->rfl, S->w,
                       S->mdiag, S->delta);
    return hipGetLastError();
}

// The network on the assembled operator for the solver: the same outputs as lv_spai_forward (w, mdiag,
// delta, o_edge, o_node) without keeping activations; two ping-pong levels of node and edge features.
hipError_t lv_spai_infer(lv_spai *S, const lv_pressure_system &A, int n, hipStream_t s)
{
    const int E = (int)A.nnz;
    S->n = n;
    S->E = E;
    SP_CHECK(grow(&S->src, &S->cap_src, (size_t)E));
    SP_CHECK(grow(&S->rev, &S->cap_rev, (size_t)E));
    SP_CHECK(grow(&S->a, &S->cap_a, (size_t)E));
    SP_CHECK(grow(&S->ff, &S->cap_ff, (size_t)4 * E));
    SP_CHECK(grow(&S->oe, &S->cap_oe, (size_t)E));
    SP_CHECK(grow(&S->w, &S->cap_w, (size_t)E));
    SP_CHECK(grow(&S->G, &S->cap_G, (size_t)E));  // s_p of the final layer
    SP_CHECK(grow(&S->xf, &S->cap_xf, (size_t)4 * n));
    SP_CHECK(grow(&S->rfl, &S->cap_rfl, (size_t)n));
    SP_CHECK(grow(&S->sqD, &S->cap_sqD, (size_t)n));
    SP_CHECK(grow(&S->on, &S->cap_on, (size_t)n));
    SP_CHECK(grow(&S->mdiag, &S->cap_mdiag, (size_t)n));
    SP_CHECK(grow(&S->delta, &S->cap_delta, (size_t)n));
    SP_CHECK(grow(&S->Hf, &S->cap_Hf, (size_t)2 * n * H));
    SP_CHECK(grow(&S->Hb, &S->cap_Hb, (size_t)2 * n * H));
    SP_CHECK(grow(&S->Eb, &S->cap_Eb, (size_t)2 * E * H));
    launch_graph(S, A, n, E, s);
    const float *th = S->theta;
    const size_t nh = (size_t)n * H, eh = (size_t)E * H;
    hipLaunchKernelGGL(k_encode, g1((int64_t)n * (H / 8)), dim3(B), 0, s, n, (const float *)S->xf,
                       th + LV_SPAI_OFF_WN, th + LV_SPAI_OFF_BN, S->Hf, S->Hb);
    const int64_t tiles = ((int64_t)E + 31) / 32;
    const unsigned blocks = (unsigned)std::min<int64_t>((tiles + 3) / 4, 4096);
    for (int l = 0; l < 3; l++) {
        const int cur = l & 1, nxt = cur ^ 1;
        const __bf16 *Wl = S->Wt + l * KX * H;
        if (l < 2) {
            auto kern = (l == 0) ? k_mp_infer<true, false> : k_mp_infer<false, false>;
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, E, (const int *)S->src, (const int *)A.cols,
                               (const __bf16 *)(S->Hb + cur * nh), (const __bf16 *)(S->Eb + cur * eh),
                               (const float *)S->ff, th + LV_SPAI_OFF_WE, th + LV_SPAI_OFF_BE, Wl,
                               th + LV_SPAI_OFF_B(l), S->Eb + nxt * eh, (const float *)nullptr,
                               (const float *)nullptr, (const float *)nullptr, (float *)nullptr, (float *)nullptr);
            hipLaunchKernelGGL(k_aggregate, g1((int64_t)n * (H / 8)), dim3(B), 0, s, n, (const int *)A.row_ptr,
                               (const float *)(S->Hf + cur * nh), (const __bf16 *)(S->Eb + nxt * eh),
                               S->Hf + nxt * nh, S->Hb + nxt * nh);
        } else {
            hipLaunchKernelGGL((k_mp_infer<false, true>), dim3(blocks), dim3(256), 0, s, E, (const int *)S->src,
                               (const int *)A.cols, (const __bf16 *)(S->Hb + cur * nh),
                               (const __bf16 *)(S->Eb + cur * eh), (const float *)nullptr, (const float *)nullptr,
                               (const float *)nullptr, Wl, th + LV_SPAI_OFF_B(l), (__bf16 *)nullptr,
                               th + LV_SPAI_OFF_VE, th + LV_SPAI_OFF_CE, th + LV_SPAI_OFF_UN, S->oe, S->G);
            hipLaunchKernelGGL(k_node_decode, g1((int64_t)n * 8), dim3(B), 0, s, n, (const int *)A.row_ptr,
                               (const float *)(S->Hf + cur * nh), (const float *)S->G, th + LV_SPAI_OFF_UN,
                               th + LV_SPAI_OFF_CN, S->on);
        }
    }
    hipLaunchKernelGGL(k_build_mt, g1(n), dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)S->rev,
                       (const float *)S->a, (const float *)S->oe, (const float *)S->on, (const float *)S->rfl, S->w,
                       S->mdiag, S->delta);
    return hipGetLastError();
}

// M = D^-1/2 Mt D^-1/2 in the solver's ELL slots (after lv_spai_forward or lv_spai_infer on the same
// operator)
hipError_t lv_spai_emit(lv_spai *S, const lv_pressure_system &A, int n, hipStream_t s)
{
    SP_CHECK(grow(&S->mval, &S->cap_mval, (size_t)LV_MAXE * n));
    SP_CHECK(grow(&S->mdg, &S->cap_mdg, (size_t)n));
    hipLaunchKernelGGL(k_write_m, g1(n), dim3(B), 0, s, n, (const int *)A.row_ptr, (const int *)A.cols,
                       (const int *)A.deg, (const int *)A.ecol, (const float *)S->w, (const float *)S->mdiag,
                       (const float *)S->sqD, S->mval, S->mdg);
    return hipGetLastError();
}

// Loss on P probes (probe 0 = the scaled data vector x if given) and, if want_grad, the gradient of
// every parameter into S->grad.  Requires lv_spai_forward on the same operator first.
hipError_t lv_spai_loss_grad(lv_spai *S, const lv_pressure_system &A, int n, const double *x, int P, uint64_t seed,
                             bool want_grad, double *loss_host, hipStream_t s)
{
    const int E = S->E;
    const size_t nh = (size_t)n * H, eh = (size_t)E * H;
    const size_t np = (size_t)n * P;
    SP_CHECK(grow(&S->v, &S->cap_v, np));
    SP_CHECK(grow(&S->u, &S->cap_u, np));
    SP_CHECK(grow(&S->y, &S->cap_y, np));
    SP_CHECK(grow(&S->gy, &S->cap_gy, np));
    const int nbl = 64;
    SP_CHECK(grow(&S->lpart, &S->cap_lpart, (size_t)3 * P * nbl));
    hipLaunchKernelGGL(k_probes, g1((int64_t)np), dim3(B), 0, s, n, P, x, (const float *)S->sqD, seed, S->v);
    hipLaunchKernelGGL(k_apply_at, g1((int64_t)np), dim3(B), 0, s, n, P, (const int *)A.row_ptr, (const int *)A.cols,
                       (const float *)S->a, (const float *)S->v, S->u);
    hipLaunchKernelGGL(k_apply_mt, g1((int64_t)np), dim3(B), 0, s, n, P, (const int *)A.row_ptr,
                       (const int *)A.cols, (const float *)S->w, (const float *)S->mdiag, (const float *)S->u, S->y);
    hipLaunchKernelGGL(k_loss_partials, dim3(nbl, P), dim3(B), 0, s, n, P, (const float *)S->y, (const float *)S->v,
                       S->lpart, nbl);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, s, P, nbl, (const double *)S->lpart, S->scal);
    if (loss_host) {
        SP_CHECK(hipMemcpyAsync(loss_host, S->scal, sizeof(double), hipMemcpyDeviceToHost, s));
        SP_CHECK(hipStreamSynchronize(s));
    }
    if (!want_grad) return hipGetLastError();

    // ---- backward ----
    SP_CHECK(grow(&S->G, &S->cap_G, (size_t)E));
    SP_CHECK(grow(&S->doe, &S->cap_doe, (size_t)E));
    SP_CHECK(grow(&S->don, &S->cap_don, (size_t)n));
    SP_CHECK(grow(&S->dE, &S->cap_dE, eh));
    SP_CHECK(grow(&S->dH, &S->cap_dH, nh));
    SP_CHECK(grow(&S->dX, &S->cap_dX, (size_t)E * KX));
    SP_CHECK(grow(&S->dZb, &S->cap_dZb, eh));
    const int chunk = 256;
    const int nchunks = (E + chunk - 1) / chunk;
    const int nbE = (int)((E + B - 1) / B), nbN = (int)((n + B - 1) / B);
    const size_t part_want = std::max(std::max((size_t)nchunks * KX * H, (size_t)nbE * (H + 1)),
                                      std::max((size_t)g1((int64_t)E * H).x * 5 * H, (size_t)nbN * (H + 1)));
    SP_CHECK(grow(&S->part, &S->cap_part, part_want));
    float *gr = S->grad;
    const float *th = S->theta;
    SP_CHECK(hipMemsetAsync(gr, 0, sizeof(float) * LV_SPAI_NPARAM, s));
    hipLaunchKernelGGL(k_grad_y, g1((int64_t)np), dim3(B), 0, s, n, P, (const float *)S->y, (const float *)S->v,
                       (const double *)S->scal, S->gy);
    hipLaunchKernelGGL(k_grad_w, g1(n), dim3(B), 0, s, n, P, (const int *)A.row_ptr, (const int *)A.cols,
                       (const float *)S->w, (const float *)S->gy, (const float *)S->u, (const float *)S->delta, S->G,
                       S->don);
    hipLaunchKernelGGL(k_grad_o, g1(E), dim3(B), 0, s, E, (const int *)S->rev, (const float *)S->G,
                       (const float *)S->w, (const float *)S->oe, S->doe);
    hipLaunchKernelGGL(k_clamp_mask, g1(n), dim3(B), 0, s, n, (const float *)S->on, S->don);
    // decoders: dE3 = doe ve, dH3 = don un; parameter partials
    hipLaunchKernelGGL(k_decode_back, dim3(nbE), dim3(B), 0, s, E, (const __bf16 *)(S->Eb + 3 * eh),
                       (const float *)nullptr, th + LV_SPAI_OFF_VE, (const float *)S->doe, S->dE, 0, S->part);
    hipLaunchKernelGGL(k_sum_parts, g1(H + 1), dim3(B), 0, s, nbE, H + 1, (const float *)S->part,
                       gr + LV_SPAI_OFF_VE, 0);
    hipLaunchKernelGGL(k_decode_back, dim3(nbN), dim3(B), 0, s, n, (const __bf16 *)nullptr,
                       (const float *)(S->Hf + 3 * nh), th + LV_SPAI_OFF_UN, (const float *)S->don, S->dH, 0, S->part);
    hipLaunchKernelGGL(k_sum_parts, g1(H + 1), dim3(B), 0, s, nbN, H + 1, (const float *)S->part,
                       gr + LV_SPAI_OFF_UN, 0);
    // message-passing layers, last to first.  Incoming: dH (of h_{l+1}) and dE (of m_l = e_{l+1}).
    const unsigned tiles = (unsigned)((E + 31) / 32);
    for (int l = 2; l >= 0; l--) {
        hipLaunchKernelGGL(k_mp_dz, g1((int64_t)E * H), dim3(B), 0, s, E, (const int *)S->src,
                           (const int *)A.row_ptr, (const float *)S->dH, (const __bf16 *)(S->Eb + (l + 1) * eh), S->dE,
                           S->dZb);
        hipLaunchKernelGGL(k_mp_dx, dim3((tiles + 3) / 4), dim3(256), 0, s, E, (const __bf16 *)S->dZb,
                           (const __bf16 *)(S->Wk + l * KX * H), S->dX);
        hipLaunchKernelGGL(k_mp_dw, dim3(nchunks, 3), dim3(256), 0, s, E, (const int *)S->src, (const int *)A.cols,
                           (const __bf16 *)(S->Hb + l * nh), (const __bf16 *)(S->Eb + l * eh), (const __bf16 *)S->dZb,
                           chunk, S->part);
        hipLaunchKernelGGL(k_sum_parts, g1(KX * H), dim3(B), 0, s, nchunks, KX * H, (const float *)S->part,
                           gr + LV_SPAI_OFF_W(l), 0);
        hipLaunchKernelGGL(k_mp_db, dim3(nchunks), dim3(H), 0, s, E, (const __bf16 *)S->dZb, chunk, S->part);
        hipLaunchKernelGGL(k_sum_parts, g1(H), dim3(B), 0, s, nchunks, H, (const float *)S->part,
                           gr + LV_SPAI_OFF_B(l), 0);
        hipLaunchKernelGGL(k_mp_dh, g1((int64_t)n * H), dim3(B), 0, s, n, (const int *)A.row_ptr,
                           (const int *)S->rev, (const float *)S->dX, S->dH);
        hipLaunchKernelGGL(k_mp_de, g1((int64_t)E * H), dim3(B), 0, s, E, (const float *)S->dX, S->dE);
    }
    // encoders
    const unsigned nbEH = g1((int64_t)E * H).x, nbNH = g1((int64_t)n * H).x;
    hipLaunchKernelGGL(k_encode_back, dim3(nbEH), dim3(B), 0, s, E, (const float *)S->ff, (const __bf16 *)S->Eb,
                       (const float *)S->dE, S->part);
    hipLaunchKernelGGL(k_sum_parts, g1(5 * H), dim3(B), 0, s, (int)nbEH, 5 * H, (const float *)S->part,
                       gr + LV_SPAI_OFF_WE, 0);
    hipLaunchKernelGGL(k_encode_back, dim3(nbNH), dim3(B), 0, s, n, (const float *)S->xf, (const __bf16 *)S->Hb,
                       (const float *)S->dH, S->part);
    hipLaunchKernelGGL(k_sum_parts, g1(5 * H), dim3(B), 0, s, (int)nbNH, 5 * H, (const float *)S->part,
                       gr + LV_SPAI_OFF_WN, 0);
    return hipGetLastError();
}

hipError_t lv_spai_adam(lv_spai *S, double lr, hipStream_t s)
{
    S->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)S->adam_t), c2 = 1.0f - powf(b2, (float)S->adam_t);
    hipLaunchKernelGGL(k_adam, g1(LV_SPAI_NPARAM), dim3(B), 0, s, (int)LV_SPAI_NPARAM, S->theta,
                       (const float *)S->grad, S->adam_m, S->adam_v, (float)lr, b1, b2, c1, c2, 1e-8f);
    for (int l = 0; l < 3; l++)
        hipLaunchKernelGGL(k_pack_w, g1(KX * H), dim3(B), 0, s, S->theta + LV_SPAI_OFF_W(l), S->Wt + l * KX * H,
                           S->Wk + l * KX * H);
    return hipGetLastError();
}
