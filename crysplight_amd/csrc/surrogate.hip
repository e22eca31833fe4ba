// surrogate.hip -- the learned light-response surrogate on gfx950: an MLP
// (4 -> 256 -> 256 -> 256 -> 128, GELU) trained and evaluated with BF16 MFMA GEMMs
// accumulating in FP32, FP32 master weights and Adam, and the Poisson sampling of
// event maps from its predicted mean detection maps.  No reference counterpart: the
// reference only has the Monte Carlo (src/transport.jl, src/kernel.jl).
//
// GEMM: C[M x N] = opA[M x K] * opB[K x N] with bf16 operands in global memory,
// staged through LDS in k-contiguous tiles (transposed on the way in when an operand is
// stored the other way round), multiplied by v_mfma_f32_32x32x16_bf16 (4 waves, a
// 64 x 64 tile per workgroup, 32 x 32 per wave), FP32 accumulation, fused epilogues
// (bias, GELU, bf16/fp32 stores).  M, N multiples of 64 and K a multiple of 32 (the
// host pads).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "crysplight.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int BM = 64, BN = 64, BK = 32;            // BK: granularity of K (and K slices)
constexpr int KP = 128;                             // K staged in LDS per phase
constexpr int LDS_PAD = 8;                          // bf16 elements of row padding

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }   // round to nearest even
__device__ __forceinline__ float from_bf16(__bf16 x) { return (float)x; }

// exact GELU, 0.5 x (1 + erf(x / sqrt 2)), and its derivative
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.3989422804014327f * __expf(-0.5f * x * x);
    return cdf + x * pdf;
}

enum Epi : int {
    EPI_F32 = 0,          // C fp32 = acc
    EPI_BIAS_GELU = 1,    // Z fp32 = acc + bias, H bf16 = gelu(Z)
    EPI_BIAS_F32 = 2,     // C fp32 = acc + bias
    EPI_BIAS_BF16 = 3,    // H bf16 = acc + bias (no activation)
};

struct GemmArgs {
    int M, N, K;
    const __bf16 *A; int lda; int a_t;   // a_t: A stored [K][M] (element (m,k) at k*lda+m)
    const __bf16 *B; int ldb; int b_t;   // b_t: B stored [N][K] (element (k,n) at n*ldb+k)
    float *C; int ldc;                   // fp32 output (EPI_F32, EPI_BIAS_F32, EPI_BIAS_GELU: Z)
    int64_t c_slice;                     // EPI_F32 split over K (gridDim.z > 1): the partial
                                         // product of K slice z goes to C + z * c_slice
    __bf16 *H; int ldh;                  // bf16 output
    const float *bias;                   // [N]
};

// One 64 x 64 output tile (bx, by) of K slice bz of nz.
template <int EPI>
__device__ __forceinline__ void gemm_tile(const GemmArgs &g, int bx, int by, int bz, int nz)
{
    __shared__ __bf16 sA[BM][KP + LDS_PAD];        // [m][k]
    __shared__ __bf16 sB[BN][KP + LDS_PAD];        // [n][k]
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = by * BM, n0 = bx * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    // K slice of this workgroup (nz slices of K / nz, a multiple of BK), staged KP at a
    // time: every load of a phase is in flight at once, then each wave runs KP/16 MFMA
    // steps on it
    const int kc = g.K / nz;
    const int kbeg = bz * kc, kend = kbeg + kc;
    for (int k0 = kbeg; k0 < kend; k0 += KP) {
        const int kp = min(KP, kend - k0);         // a multiple of BK
        const int chunks = kp * (BM / 8);          // 8-element chunks per tile (BM == BN)
        constexpr int CPT = KP * (BM / 8) / 256;   // chunks per thread in a full phase
        bf16x8 va[CPT], vb[CPT];
#pragma unroll
        for (int u = 0; u < CPT; ++u) {            // all loads first ...
            const int e = (tid + u * 256) * 8;
            if (e >= chunks * 8) break;
            va[u] = !g.a_t ? *(const bf16x8 *)(g.A + (size_t)(m0 + e / kp) * g.lda + k0 + e % kp)
                           : *(const bf16x8 *)(g.A + (size_t)(k0 + e / BM) * g.lda + m0 + e % BM);
            vb[u] = g.b_t ? *(const bf16x8 *)(g.B + (size_t)(n0 + e / kp) * g.ldb + k0 + e % kp)
                          : *(const bf16x8 *)(g.B + (size_t)(k0 + e / BN) * g.ldb + n0 + e % BN);
        }
#pragma unroll
        for (int u = 0; u < CPT; ++u) {            // ... then the LDS stores
            const int e = (tid + u * 256) * 8;
            if (e >= chunks * 8) break;
            if (!g.a_t) {
                *(bf16x8 *)&sA[e / kp][e % kp] = va[u];           // row m, 8 consecutive k
            } else {
                const int kk = e / BM, mm = e % BM;               // row k, 8 consecutive m
#pragma unroll
                for (int j = 0; j < 8; ++j) sA[mm + j][kk] = va[u][j];
            }
            if (g.b_t) {
                *(bf16x8 *)&sB[e / kp][e % kp] = vb[u];           // row n, 8 consecutive k
            } else {
                const int kk = e / BN, nn = e % BN;               // row k, 8 consecutive n
#pragma unroll
                for (int j = 0; j < 8; ++j) sB[nn + j][kk] = vb[u][j];
            }
        }
        __syncthreads();
        for (int kk = 0; kk < kp; kk += 16) {
            // lane l: row/col l % 32, k = kk + 8 * (l / 32) .. +7
            const int r = lane & 31, kq = kk + 8 * (lane >> 5);
            const bf16x8 a = *(const bf16x8 *)&sA[wm + r][kq];
            const bf16x8 b = *(const bf16x8 *)&sB[wn + r][kq];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }

    // D layout: lane l holds column l % 32, rows 8*(i/4) + 4*(l/32) + i%4, i = 0..15
    const int col = n0 + wn + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        float v = acc[i];
        if (EPI == EPI_F32) {
            g.C[(size_t)bz * g.c_slice + (size_t)row * g.ldc + col] = v;
        } else if (EPI == EPI_BIAS_F32) {
            g.C[(size_t)row * g.ldc + col] = v + g.bias[col];
        } else if (EPI == EPI_BIAS_GELU) {
            v += g.bias[col];
            if (g.C) g.C[(size_t)row * g.ldc + col] = v;
            g.H[(size_t)row * g.ldh + col] = to_bf16(gelu(v));
        } else {  // EPI_BIAS_BF16
            g.H[(size_t)row * g.ldh + col] = to_bf16(v + g.bias[col]);
        }
    }
}

template <int EPI>
__global__ void __launch_bounds__(256) gemm_bf16(const GemmArgs g)
{
    gemm_tile<EPI>(g, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z);
}

// Several independent split-K products in one launch: product j owns the z range
// [z0[j], z0[j+1]) (its K slices); blocks beyond its tile grid return at once.
constexpr int MAXL = 9;                             // layers of the MLP, at most
struct GroupArgs {
    int n;
    GemmArgs a[MAXL];
    int z0[MAXL + 1];
};

__global__ void __launch_bounds__(256) gemm_grouped(const GroupArgs G)
{
    int j = 0;
    while (j + 1 < G.n && (int)blockIdx.z >= G.z0[j + 1]) ++j;
    const GemmArgs &g = G.a[j];
    if ((int)blockIdx.x >= g.N / BN || (int)blockIdx.y >= g.M / BM) return;
    gemm_tile<EPI_F32>(g, blockIdx.x, blockIdx.y, blockIdx.z - G.z0[j], G.z0[j + 1] - G.z0[j]);
}

// ---------------------------------------------------------------- small kernels ----
// inputs fp32 [M x n_in] -> bf16 [Mpad x kin_pad], normalized: (x - shift) * scale
// bidx (optional, device): the batch index, X advanced by *bidx * bstride (graph replays)
__global__ void pack_inputs(const float *X, int M, int Mpad, int n_in, int kin_pad,
                            const float *shift, const float *scale, __bf16 *out,
                            const int64_t *bidx, int64_t bstride)
{
    if (bidx) X += *bidx * bstride;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)Mpad * kin_pad) return;
    const int r = (int)(i / kin_pad), c = (int)(i % kin_pad);
    float v = 0.0f;
    if (r < M && c < n_in) v = (X[(size_t)r * n_in + c] - shift[c]) * scale[c];
    out[i] = to_bf16(v);
}

// ------------------------------------------------------------ the training step ----
// A step is four launches: mlp_rows (inputs, forward, loss, the backward data path and the
// bias-gradient partials, one workgroup per 32 rows), gemm_grouped (the weight gradients of
// every layer as split-K products), reduce_grads (the partials summed in a fixed order) and
// adam_refresh (Adam, the bf16 weight copies).  Every sum runs in a fixed order, so a
// training run is reproducible.

constexpr int RB = 32;                              // rows per mlp_rows workgroup
constexpr int RT = 512;                             // its threads: 8 waves, one 32-column block each
constexpr int AMAX = 256;                           // widest layer (hidden, n_out, kin_pad)
constexpr int KB = 128;                             // K staged in LDS per phase (mlp_rows)

struct RowArgs {
    int L, M, Mpad, n_in, kin_pad, n_out, n_pois;
    int din[MAXL], dpad[MAXL], dout[MAXL];
    const __bf16 *Wt[MAXL];                         // [out][in_pad]
    const __bf16 *Wn[MAXL];                         // [in][out]
    const float *bias[MAXL];
    __bf16 *HT[MAXL];                               // hidden layers: activation, transposed [out][Mpad]
    __bf16 *dZT[MAXL];                              // gradient of layer l's pre-activation, [out][Mpad]
    __bf16 *XT;                                     // packed inputs, transposed [kin_pad][Mpad]
    float *bpart[MAXL];                             // [n_wg][out] bias-gradient partials
    float *lpart;                                   // [n_wg] loss partials
    const float *Xraw, *T, *W, *shift, *scale;
    float inv_wsum;
    const int64_t *bidx;                            // batch index (graph replays) or null
};

// The weights of a workgroup's GEMMs, in the order mlp_rows runs them (forward layers
// 0 .. L-1 with B = Wt[l], then the backward data products of layers L-1 .. 1 with
// B = Wn[l]), streamed KB columns of K at a time through a register buffer: phase q + 1 is
// loaded as soon as phase q sits in LDS, so its global loads are in flight during phase q's
// MFMAs and, at a GEMM boundary, during the epilogue as well.
struct BStream {
    constexpr static int CPT = AMAX * KB / 8 / RT;  // 8-element chunks per thread at N = AMAX
    const RowArgs &a;
    int Q;                                          // phases in all
    bf16x8 buf[CPT];

    __device__ __forceinline__ void gemm(int g, const __bf16 *&B, int &ldb, int &N, int &K) const
    {
        if (g < a.L) { B = a.Wt[g]; ldb = a.dpad[g]; N = a.dout[g]; K = a.dpad[g]; }
        else { const int l = 2 * a.L - 1 - g; B = a.Wn[l]; ldb = a.dout[l]; N = a.din[l]; K = a.dout[l]; }
    }
    __device__ __forceinline__ BStream(const RowArgs &ra) : a(ra), Q(0)
    {
        for (int g = 0; g < 2 * a.L - 1; ++g) {
            const __bf16 *B; int ldb, N, K;
            gemm(g, B, ldb, N, K);
            Q += (K + KB - 1) / KB;
        }
    }
    // phase q: GEMM columns N of B rows, K range k0 .. k0 + kp - 1
    __device__ __forceinline__ void phase(int q, const __bf16 *&B, int &ldb, int &N, int &k0, int &kp) const
    {
        for (int g = 0;; ++g) {
            int K;
            gemm(g, B, ldb, N, K);
            const int np = (K + KB - 1) / KB;
            if (q < np) { k0 = q * KB; kp = min(KB, K - k0); return; }
            q -= np;
        }
    }
    // A phase is N rows of kp / 8 chunks of 8 elements (kp = 64 or 128): thread t takes the
    // chunks t + u * RT, i.e. rows t / cpr + u * (RT / cpr), elements 8 * (t % cpr) onwards.
    __device__ __forceinline__ static int cpr_shift(int kp) { return kp == KB ? 4 : 3; }
    __device__ __forceinline__ void load(int q)
    {
        if (q >= Q) return;
        const __bf16 *B; int ldb, N, k0, kp;
        phase(q, B, ldb, N, k0, kp);
        const int sh = cpr_shift(kp), tid = threadIdx.x;
        const int row = tid >> sh, rstep = RT >> sh;
        const __bf16 *p = B + (size_t)row * ldb + k0 + ((tid & ((1 << sh) - 1)) << 3);
#pragma unroll
        for (int u = 0; u < CPT; ++u)
            if (row + u * rstep < N) buf[u] = *(const bf16x8 *)(p + (size_t)u * rstep * ldb);
    }
    // the loaded phase into sB[n][k - k0]
    __device__ __forceinline__ void stash(int N, int kp, __bf16 (*sB)[KB + LDS_PAD]) const
    {
        const int sh = cpr_shift(kp), tid = threadIdx.x;
        const int row = tid >> sh, rstep = RT >> sh, kk = (tid & ((1 << sh) - 1)) << 3;
#pragma unroll
        for (int u = 0; u < CPT; ++u)
            if (row + u * rstep < N) *(bf16x8 *)&sB[row + u * rstep][kk] = buf[u];
    }
};

// out[RB x N] = A[RB x K] * B^T for the next GEMM of the stream (phases q .. q + np - 1),
// A resident in LDS.  Wave w owns the 32-column block w, lane element i (0..15) of it at
// (row, col).  pre(i, row, col, u, v) loads what the epilogue of an element needs into u, v
// before the K loop, so those loads complete under the MFMAs; epi(i, row, col, acc, u, v)
// stores the element.
template <class Pre, class Epi>
__device__ __forceinline__ void rows_gemm(const __bf16 (*A)[AMAX + LDS_PAD], int K, int N,
                                          BStream &bs, int &q, __bf16 (*sB)[KB + LDS_PAD],
                                          Pre pre, Epi epi)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool mine = wave < N / 32;
    const int col = wave * 32 + (lane & 31);
    float pu[16], pv[16];
    if (mine) {
#pragma unroll
        for (int i = 0; i < 16; ++i) pre(i, 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), col, pu[i], pv[i]);
    }
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += KB, ++q) {
        const int kp = min(KB, K - k0);             // a multiple of 64
        bs.stash(N, kp, sB);
        __syncthreads();
        bs.load(q + 1);                             // the next phase, maybe of the next GEMM
        if (mine) {
            for (int kk = 0; kk < kp; kk += 16) {
                const int r = lane & 31, kq = kk + 8 * (lane >> 5);
                const bf16x8 a = *(const bf16x8 *)&A[r][k0 + kq];
                const bf16x8 b = *(const bf16x8 *)&sB[wave * 32 + r][kq];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (mine) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            epi(i, 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), col, acc[i], pu[i], pv[i]);
    }
}

// The RB rows of act (bf16, N columns) into dst transposed: dst[c * ld + r0 + r].  A
// thread gathers 8 rows of a column from LDS and stores them as one 16-byte chunk.
__device__ __forceinline__ void rows_store_T(const __bf16 (*act)[AMAX + LDS_PAD], int N, __bf16 *dst,
                                             int64_t ld, int r0)
{
    for (int c = threadIdx.x; c < N * (RB / 8); c += RT) {
        const int col = c / (RB / 8), r = (c % (RB / 8)) * 8;
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = act[r + j][col];
        *(bf16x8 *)(dst + (size_t)col * ld + r0 + r) = v;
    }
}

// column sums of the RB rows of act (bf16) into out[c], c < N, rows in order
__device__ __forceinline__ void rows_colsum(const __bf16 (*act)[AMAX + LDS_PAD], int N, float *out)
{
    for (int c = threadIdx.x; c < N; c += RT) {
        float s = 0.0f;
#pragma unroll 8
        for (int r = 0; r < RB; ++r) s += from_bf16(act[r][c]);
        out[c] = s;
    }
}

// One workgroup per RB rows of the batch:
//   pack (x - shift) * scale to bf16;
//   forward: hidden layers Z = A W^T + b, H = gelu(Z) (bf16) and gelu'(Z) (fp32, for the
//   backward);
//   loss of the last layer's output y against t with weight w, normalized by 1 / sum w:
//     columns c <  n_pois: Poisson NLL of a log-rate, w (e^y - t y), gradient w (e^y - t);
//     columns c >= n_pois: weighted squared error w (y - t)^2, gradient 2 w (y - t);
//   backward data path: dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1});
//   bias-gradient partials (column sums of every dZ_l over the RB rows) and the loss partial.
// Padded rows (>= M) carry zero inputs and a zero gradient.
template <int L>
__global__ void __launch_bounds__(RT) mlp_rows(const RowArgs a)
{
    __shared__ __bf16 act[2][RB][AMAX + LDS_PAD];
    __shared__ __bf16 sB[AMAX][KB + LDS_PAD];
    __shared__ float red[RT / 64];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * RB;
    const int64_t b = a.bidx ? *a.bidx : 0;
    const float *Xraw = a.Xraw + b * (int64_t)a.M * a.n_in;
    const float *T = a.T + b * (int64_t)a.M * a.n_out;
    const float *W = a.W ? a.W + b * (int64_t)a.M * a.n_out : nullptr;
    BStream bs(a);
    bs.load(0);
    int q = 0;

    for (int i = tid; i < RB * a.kin_pad; i += RT) {
        const int r = i / a.kin_pad, c = i % a.kin_pad;
        float v = 0.0f;
        if (r0 + r < a.M && c < a.n_in) v = (Xraw[(size_t)(r0 + r) * a.n_in + c] - a.shift[c]) * a.scale[c];
        act[0][r][c] = to_bf16(v);
    }
    __syncthreads();
    rows_store_T(act[0], a.kin_pad, a.XT, a.Mpad, r0);

    int cur = 0;
    float lsum = 0.0f;
    // gelu'(Z) of the hidden layers, kept in registers: the backward product of layer l + 1
    // has the forward layer l's shape, and every element stays with the lane that made it
    float gpr[L - 1][16];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const int N = a.dout[l];
        __bf16 (*out)[AMAX + LDS_PAD] = act[cur ^ 1];
        const float *bias = a.bias[l];
        if (l + 1 < L) {
            rows_gemm(act[cur], a.dpad[l], N, bs, q, sB,
                      [&](int, int row, int col, float &u, float &) { u = bias[col]; },
                      [&](int i, int row, int col, float v, float u, float) {
                          v += u;
                          gpr[l < L - 1 ? l : 0][i] = gelu_grad(v);
                          out[row][col] = to_bf16(gelu(v));
                      });
        } else {
            const int M = a.M, n_pois = a.n_pois;
            const float inv_wsum = a.inv_wsum;
            rows_gemm(act[cur], a.dpad[l], N, bs, q, sB,
                      [&](int, int row, int col, float &t, float &w) {
                          const int rg = r0 + row;
                          const size_t o = (size_t)rg * N + col;
                          t = rg < M ? T[o] : 0.0f;
                          w = rg < M ? (W ? W[o] : 1.0f) : 0.0f;
                      },
                      [&](int, int row, int col, float y, float t, float w) {
                          y += bias[col];
                          const int rg = r0 + row;
                          float g = 0.0f;
                          if (rg < M) {
                              if (col < n_pois) {
                                  const float e = __expf(y);
                                  lsum += w * (e - t * y) * inv_wsum;
                                  g = w * (e - t) * inv_wsum;
                              } else {
                                  const float d = y - t;
                                  lsum += w * d * d * inv_wsum;
                                  g = 2.0f * w * d * inv_wsum;
                              }
                          }
                          out[row][col] = to_bf16(g);
                      });
        }
        __syncthreads();
        cur ^= 1;
        rows_store_T(act[cur], N, l + 1 < L ? a.HT[l] : a.dZT[l], a.Mpad, r0);
    }
    rows_colsum(act[cur], a.dout[L - 1], a.bpart[L - 1] + (size_t)blockIdx.x * a.dout[L - 1]);

#pragma unroll
    for (int l = L - 1; l >= 1; --l) {
        const int N = a.din[l];                     // = dout[l - 1]
        __bf16 (*out)[AMAX + LDS_PAD] = act[cur ^ 1];
        rows_gemm(act[cur], a.dout[l], N, bs, q, sB,
                  [&](int, int, int, float &, float &) {},
                  [&](int i, int row, int col, float v, float, float) {
                      out[row][col] = to_bf16(v * gpr[l - 1][i]);
                  });
        __syncthreads();
        cur ^= 1;
        rows_store_T(act[cur], N, a.dZT[l - 1], a.Mpad, r0);
        rows_colsum(act[cur], N, a.bpart[l - 1] + (size_t)blockIdx.x * N);
    }

    // loss partial of the workgroup, in a fixed order
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) lsum += __shfl_xor(lsum, m);
    if ((tid & 63) == 0) red[tid >> 6] = lsum;
    __syncthreads();
    if (tid == 0) {
        float t = 0.0f;
#pragma unroll
        for (int w = 0; w < RT / 64; ++w) t += red[w];
        a.lpart[blockIdx.x] = t;
    }
}

struct RedArgs {
    int L, n_wg;
    const float *part[MAXL];                        // split-K partials of dW_l, slice stride dout * dpad
    int ks[MAXL], din[MAXL], dpad[MAXL], dout[MAXL];
    const float *bpart[MAXL];                       // [n_wg][dout]
    int64_t woff[MAXL], boff[MAXL], cum[MAXL + 1];  // cum: outputs (dout * din + dout) before layer l
    float *g;
    const float *lpart;
    float *loss;                                    // loss[ctr[0]] with loss_by_batch, else loss[0]
    int loss_by_batch;
    int64_t *ctr;                                   // batch index and Adam step, advanced here
};

// q[0] + q[stride] + ... (n terms), added in order with 16 loads in flight
__device__ __forceinline__ float sum_strided(const float *q, int n, int64_t stride)
{
    float s = 0.0f;
    int u = 0;
    for (; u + 16 <= n; u += 16) {
        float v[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) v[t] = q[(size_t)(u + t) * stride];
#pragma unroll
        for (int t = 0; t < 16; ++t) s += v[t];
    }
    for (; u < n; ++u) s += q[(size_t)u * stride];
    return s;
}

// The gradient: every weight is the in-order sum of its K-slice partials, every bias the
// in-order sum of the per-workgroup column sums; the loss is the in-order sum of the
// workgroup partials.  With ctr (graph replays) the batch index and the Adam step count
// advance here, after the step's last reader of the batch index and before Adam.
__global__ void __launch_bounds__(256) reduce_grads(const RedArgs a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.cum[a.L];
         i += (int64_t)gridDim.x * blockDim.x) {
        int l = 0;
        while (i >= a.cum[l + 1]) ++l;
        const int64_t j = i - a.cum[l];
        const int64_t nw = (int64_t)a.dout[l] * a.din[l];
        float s = 0.0f;
        if (j < nw) {
            const int o = (int)(j / a.din[l]), k = (int)(j % a.din[l]);
            const float *q = a.part[l] + (size_t)o * a.dpad[l] + k;
            const int64_t slice = (int64_t)a.dout[l] * a.dpad[l];
            int u = 0;
            for (; u + 8 <= a.ks[l]; u += 8) {      // eight loads in flight, added in order
                float v[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) v[t] = q[(size_t)(u + t) * slice];
#pragma unroll
                for (int t = 0; t < 8; ++t) s += v[t];
            }
            for (; u < a.ks[l]; ++u) s += q[(size_t)u * slice];
            a.g[a.woff[l] + j] = s;
        } else {
            a.g[a.boff[l] + (j - nw)] = sum_strided(a.bpart[l] + (j - nw), a.n_wg, a.dout[l]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const float s = sum_strided(a.lpart, a.n_wg, 1);
        a.loss[a.ctr && a.loss_by_batch ? a.ctr[0] : 0] = s;
        if (a.ctr) {
            a.ctr[0] += 1;
            a.ctr[1] += 1;
        }
    }
}

struct AdamArgs {
    int L;
    int64_t woff[MAXL], boff[MAXL];
    int din[MAXL], dpad[MAXL], dout[MAXL];
    __bf16 *Wt[MAXL], *Wn[MAXL];
};

// Adam on the fp32 master weights, then the bf16 copies of every weight in both layouts,
// Wt [out][in_pad] (forward) and Wn [in][out] (backward data path, layers >= 1); the zero
// padding of Wt stays as it is.  dstep (graph replays): the Adam step, already advanced;
// the bias corrections c1, c2 then follow from it on the device, and the learning rate is read
// from dlr.
__global__ void adam_refresh(float *w, const float *g, float *m, float *v, int n, float lr, float b1,
                             float b2, float eps, float c1, float c2, float wd, const int64_t *dstep,
                             const float *dlr, const AdamArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dlr) lr = *dlr;
    if (dstep) {
        const float t = (float)*dstep;
        c1 = 1.0f / (1.0f - powf(b1, t));
        c2 = 1.0f / (1.0f - powf(b2, t));
    }
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    int l = 0;
    while (l + 1 < a.L && (int64_t)i >= a.woff[l + 1]) ++l;
    if ((int64_t)i >= a.boff[l]) return;            // a bias
    const int j = i - (int)a.woff[l], o = j / a.din[l], k = j % a.din[l];
    const __bf16 h = to_bf16(wi);
    a.Wt[l][(size_t)o * a.dpad[l] + k] = h;
    if (l > 0) a.Wn[l][(size_t)k * a.dout[l] + o] = h;
}

__global__ void refresh_bf16(const float *w, int out_dim, int in_dim, int in_pad, __bf16 *Wt,
                             __bf16 *Wn)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= out_dim * in_pad) return;
    const int o = i / in_pad, k = i % in_pad;
    const float x = k < in_dim ? w[(size_t)o * in_dim + k] : 0.0f;
    Wt[i] = to_bf16(x);
    if (Wn) Wn[(size_t)k * out_dim + o] = to_bf16(x);
}

// Poisson sampler of src/events.jl:11-22 (Knuth up to lambda 50, rounded Gaussian above)
// on a per-(event, cell) Philox stream, domain 4.
__device__ __forceinline__ void philox_d(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1, uint32_t o[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

struct PoisRng {
    uint32_t k0, k1, a, b, blk = 0, buf[4], left = 0;
    __device__ double u()
    {
        if (left < 2) { philox_d(a, b, blk++, 4u, k0, k1, buf); left = 4; }
        const uint64_t hi = buf[4 - left] >> 5, lo = buf[5 - left] >> 6;
        left -= 2;
        return ((double)(hi * 67108864ull + lo) + 0.5) * (1.0 / 9007199254740992.0);  // (0,1)
    }
};

__device__ int32_t poisson_draw(PoisRng &r, double lam)
{
    if (!(lam > 0.0)) return 0;
    if (lam > 50.0) {
        const double z = sqrt(-2.0 * log(r.u())) * cos(6.283185307179586 * r.u());
        const double v = rint(lam + sqrt(lam) * z);
        return v > 0.0 ? (int32_t)v : 0;
    }
    const double L = exp(-lam);
    int32_t k = 0;
    double p = 1.0;
    for (;;) {
        p *= r.u();
        if (p < L) return k;
        ++k;
    }
}

// Trilinear lookup of a baked grid [nz][ny][nx][ncell] at mm position, clamped inside.
__device__ __forceinline__ void grid_coords(float x, float L, int n, int &i0, float &f)
{
    float g = x / L * (float)(n - 1);
    g = fminf(fmaxf(g, 0.0f), (float)(n - 1) - 1e-4f);
    i0 = (int)g;
    f = g - (float)i0;
}

// Events from deposits: mean counts lambda_c = sum_d yield * E_d / 1000 * p_c(pos_d) for
// surrogate deposits (use[d] != 0), sampled Poisson per (event, cell); mean arrival
// times combined count-weighted.  One thread per (event, cell); deposits of an event are
// [ev_first[e], ev_first[e+1]).
__global__ void sample_events_grid(const float *grid, int gx, int gy, int gz, int ncell,
                                   float Lx, float Ly, float Lz, const float4 *dep,
                                   const float *dep_e_kev, const uint8_t *use,
                                   const int64_t *ev_first, int64_t nev, double yield_per_mev,
                                   uint64_t seed, int32_t *maps, float *tmean, float *lam_out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nev * ncell) return;
    const int64_t e = i / ncell;
    const int c = (int)(i % ncell);
    double lam = 0.0, tw = 0.0;
    for (int64_t d = ev_first[e]; d < ev_first[e + 1]; ++d) {
        if (!use[d]) continue;
        const float4 q = dep[d];
        int ix, iy, iz;
        float fx, fy, fz;
        grid_coords(q.x, Lx, gx, ix, fx);
        grid_coords(q.y, Ly, gy, iy, fy);
        grid_coords(q.z, Lz, gz, iz, fz);
        float p = 0.0f, t = 0.0f;
#pragma unroll
        for (int corner = 0; corner < 8; ++corner) {
            const int dx = corner & 1, dy = (corner >> 1) & 1, dz = corner >> 2;
            const float w = (dx ? fx : 1.0f - fx) * (dy ? fy : 1.0f - fy) * (dz ? fz : 1.0f - fz);
            const size_t base = ((((size_t)(iz + dz) * gy + iy + dy) * gx + ix + dx) * 2) * ncell;
            p += w * grid[base + c];
            t += w * grid[base + ncell + c];
        }
        const double mu = yield_per_mev * (double)dep_e_kev[d] / 1000.0 * (double)fmaxf(p, 0.0f);
        lam += mu;
        tw += mu * ((double)t + (double)q.w);
    }
    PoisRng r;
    r.k0 = (uint32_t)seed; r.k1 = (uint32_t)(seed >> 32);
    r.a = (uint32_t)i; r.b = (uint32_t)((uint64_t)i >> 32);
    maps[i] = poisson_draw(r, lam);
    if (tmean) tmean[i] = lam > 0.0 ? (float)(tw / lam) : INFINITY;
    if (lam_out) lam_out[i] = (float)lam;
}

__global__ void fill_f32(float *p, float v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

// ================================================================= host side ====

static thread_local std::string s_err;
static int sfail(int code, const std::string &m) { s_err = m; return code; }
#define SCHK(expr)                                                                    \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return sfail(CRYSP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct crysp_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_in = 4, hidden = 256, n_out = 128, n_layers = 4, kin_pad = 64;
    std::vector<int> din, dout, dpad;               // per layer: in, out, padded in
    std::vector<size_t> woff, boff;                 // offsets into the flat fp32 parameter vector
    size_t n_params = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 master, grad, Adam
    std::vector<__bf16 *> Wt, Wn;                   // bf16 [out][in_pad], [in][out]
    float *shift = nullptr, *scale = nullptr;       // input normalization
    int step = 0;
    int n_pois = 0;                                 // leading outputs with the Poisson loss
    // activations for a batch capacity
    int cap = 0;
    __bf16 *X = nullptr;                            // [cap][kin_pad]
    std::vector<float *> Z;                         // fp32 [cap][out]: pre-activations (forward),
                                                    // their gelu derivative (training step)
    std::vector<__bf16 *> H;                        // post-activations bf16 [cap][out]
    float *Y = nullptr;                             // output fp32 [cap][n_out]
    __bf16 *dY = nullptr;
    std::vector<__bf16 *> dZ;                       // [cap][out] per layer
    float *part = nullptr;                          // split-K partial weight gradients, all layers
    std::vector<size_t> part_off;                   // per layer, floats
    float *bpart = nullptr;                         // bias-gradient partials [n_wg][out], all layers
    std::vector<size_t> bpart_off;                  // per layer, floats
    float *lpart = nullptr;                         // loss partials [n_wg]
    float *loss = nullptr;
    int64_t *ctr = nullptr;                         // device: batch index, Adam steps (graph replays)
    float *dlr = nullptr;                           // device: learning rate (graph replays)
    float *dloss = nullptr;                         // device: per-step losses of a call
    int64_t dloss_cap = 0;
    // the captured training step, reused while its arguments stay the same
    hipGraphExec_t gexec = nullptr;
    const void *gX = nullptr, *gT = nullptr, *gW = nullptr;
    int gbatch = 0, gsteps = 0;                     // batch size, steps captured per graph
    bool glosses = false;
    float gb1 = 0, gb2 = 0, geps = 0, gwd = 0, gmrw = 0;
    int64_t hctr[2] = {0, 0};
    double flops_fwd_per_row = 0.0;
};

// K slices for a weight-gradient product (output M x N, reduction over the batch K): the
// output alone is only a few 64x64 tiles, so K is split until about SPLIT_TARGET_WG
// workgroups run per layer (all layers share one launch), keeping each slice a multiple of
// BK and at least two BK steps.
constexpr int SPLIT_TARGET_WG = 256;
static int weight_grad_slices(int M, int N, int K)
{
    const int tiles = (M / BM) * (N / BN);
    int ks = 1;
    while (tiles * ks * 2 <= SPLIT_TARGET_WG && K % (ks * 2 * BK) == 0 && K / (ks * 2) >= 2 * BK)
        ks *= 2;
    return ks;
}

static void mlp_drop_graph(crysp_mlp *n)
{
    if (n->gexec) (void)hipGraphExecDestroy(n->gexec);
    n->gexec = nullptr;
}

static int mlp_alloc_batch(crysp_mlp *n, int cap)
{
    if (cap <= n->cap) return CRYSP_OK;
    mlp_drop_graph(n);                              // it holds the buffers about to be replaced
    n->cap = 0;                                     // a failed allocation below leaves none
    auto fre = [](void *q) { if (q) (void)hipFree(q); };
    fre(n->X); fre(n->Y); fre(n->dY);
    for (auto q : n->Z) fre(q);
    for (auto q : n->H) fre(q);
    for (auto q : n->dZ) fre(q);
    n->Z.assign(n->n_layers, nullptr);
    n->H.assign(n->n_layers, nullptr);
    n->dZ.assign(n->n_layers, nullptr);
    SCHK(hipMalloc(&n->X, (size_t)cap * n->kin_pad * 2));
    SCHK(hipMalloc(&n->Y, (size_t)cap * n->n_out * 4));
    SCHK(hipMalloc(&n->dY, (size_t)cap * n->n_out * 2));
    size_t part = 0, bpart = 0;
    const int n_wg = (cap + RB - 1) / RB;
    n->part_off.assign(n->n_layers, 0);
    n->bpart_off.assign(n->n_layers, 0);
    for (int l = 0; l < n->n_layers; ++l) {
        SCHK(hipMalloc(&n->Z[l], (size_t)cap * n->dout[l] * 4));
        SCHK(hipMalloc(&n->H[l], (size_t)cap * n->dout[l] * 2));
        SCHK(hipMalloc(&n->dZ[l], (size_t)cap * n->dout[l] * 2));
        // the most slices any batch can get: tiles * slices stays within SPLIT_TARGET_WG
        const int tiles = (n->dout[l] / BM) * (n->dpad[l] / BN);
        int ks_max = 1;
        while (tiles * ks_max * 2 <= SPLIT_TARGET_WG) ks_max *= 2;
        n->part_off[l] = part;
        part += (size_t)ks_max * n->dout[l] * n->dpad[l];
        n->bpart_off[l] = bpart;
        bpart += (size_t)n_wg * n->dout[l];
    }
    fre(n->part); fre(n->bpart); fre(n->lpart);
    n->part = nullptr; n->bpart = nullptr; n->lpart = nullptr;
    SCHK(hipMalloc(&n->part, part * 4));
    SCHK(hipMalloc(&n->bpart, bpart * 4));
    SCHK(hipMalloc(&n->lpart, (size_t)n_wg * 4));
    n->cap = cap;
    return CRYSP_OK;
}

template <int EPI>
static void gemm(hipStream_t s, const GemmArgs &a, int k_slices = 1)
{
    dim3 grid(a.N / BN, a.M / BM, k_slices);
    hipLaunchKernelGGL(gemm_bf16<EPI>, grid, dim3(256), 0, s, a);
}


static int mlp_forward_dev(crysp_mlp *n, const float *Xdev, int M, int Mpad,
                           const int64_t *bidx = nullptr)
{
    const int64_t tot = (int64_t)Mpad * n->kin_pad;
    hipLaunchKernelGGL(pack_inputs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, n->stream,
                       Xdev, M, Mpad, n->n_in, n->kin_pad, n->shift, n->scale, n->X, bidx,
                       (int64_t)M * n->n_in);
    const __bf16 *in = n->X;
    int ld_in = n->kin_pad;
    for (int l = 0; l < n->n_layers; ++l) {
        GemmArgs a{};
        a.M = Mpad; a.N = n->dout[l]; a.K = n->dpad[l];
        a.A = in; a.lda = ld_in; a.a_t = 0;
        a.B = n->Wt[l]; a.ldb = n->dpad[l]; a.b_t = 1;
        a.bias = n->p + n->boff[l];
        if (l + 1 < n->n_layers) {
            a.C = n->Z[l]; a.ldc = a.N; a.H = n->H[l]; a.ldh = a.N;
            gemm<EPI_BIAS_GELU>(n->stream, a);
            in = n->H[l];
            ld_in = a.N;
        } else {
            a.C = n->Y; a.ldc = a.N;
            gemm<EPI_BIAS_F32>(n->stream, a);
        }
    }
    return CRYSP_OK;
}

// The gradient of the weighted loss on Mpad rows (M real) into n->g (fp32 flat) and the loss
// into loss_out[0], or loss_out[ctr[0]] with loss_by_batch.  With ctr (graph replays),
// ctr[0] is the batch index whose rows are read, and ctr advances.  Three launches:
// mlp_rows, gemm_grouped (dW of every layer), reduce_grads.
static int mlp_grad_dev(crysp_mlp *n, const float *Xraw, const float *T, const float *W, int M,
                        int Mpad, double wsum, float *loss_out, bool loss_by_batch, int64_t *ctr)
{
    const int L = n->n_layers;
    if (L > MAXL || Mpad % RB || Mpad > n->cap) return sfail(CRYSP_EINVAL, "bad training shape");
    const int n_wg = Mpad / RB;
    RowArgs ra{};
    ra.L = L; ra.M = M; ra.Mpad = Mpad; ra.n_in = n->n_in; ra.kin_pad = n->kin_pad; ra.n_out = n->n_out;
    ra.n_pois = n->n_pois;
    for (int l = 0; l < L; ++l) {
        ra.din[l] = n->din[l]; ra.dpad[l] = n->dpad[l]; ra.dout[l] = n->dout[l];
        ra.Wt[l] = n->Wt[l]; ra.Wn[l] = n->Wn[l]; ra.bias[l] = n->p + n->boff[l];
        ra.HT[l] = n->H[l];
        ra.dZT[l] = l + 1 == L ? n->dY : n->dZ[l];
        ra.bpart[l] = n->bpart + n->bpart_off[l];
    }
    ra.XT = n->X; ra.lpart = n->lpart;
    ra.Xraw = Xraw; ra.T = T; ra.W = W; ra.shift = n->shift; ra.scale = n->scale;
    ra.inv_wsum = (float)(1.0 / wsum);
    ra.bidx = ctr;
    switch (L) {                                    // the layer count is a template argument, so
    case 2: hipLaunchKernelGGL(mlp_rows<2>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    case 3: hipLaunchKernelGGL(mlp_rows<3>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    case 4: hipLaunchKernelGGL(mlp_rows<4>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    case 5: hipLaunchKernelGGL(mlp_rows<5>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    case 6: hipLaunchKernelGGL(mlp_rows<6>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    case 7: hipLaunchKernelGGL(mlp_rows<7>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    case 8: hipLaunchKernelGGL(mlp_rows<8>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    default: hipLaunchKernelGGL(mlp_rows<9>, dim3(n_wg), dim3(RT), 0, n->stream, ra); break;
    }                                               // the layer loops unroll (gelu' in registers)

    // dW_l[out][in] = dZ_l^T [out x Mpad] * Hin [Mpad x in_pad], both operands stored
    // transposed (k = the batch row contiguous), every layer in one launch, as split-K
    // partial products
    GroupArgs ga{};
    ga.n = L;
    RedArgs rd{};
    rd.L = L; rd.n_wg = n_wg;
    int z = 0;
    int64_t cum = 0;
    for (int l = 0; l < L; ++l) {
        const int kin = n->dpad[l], nout = n->dout[l];
        const int ks = weight_grad_slices(nout, kin, Mpad);
        GemmArgs &a = ga.a[l];
        a.M = nout; a.N = kin; a.K = Mpad;
        a.A = ra.dZT[l]; a.lda = Mpad; a.a_t = 0;
        a.B = l == 0 ? ra.XT : ra.HT[l - 1]; a.ldb = Mpad; a.b_t = 1;
        a.C = n->part + n->part_off[l]; a.ldc = kin; a.c_slice = (int64_t)nout * kin;
        ga.z0[l] = z;
        z += ks;
        rd.part[l] = a.C; rd.ks[l] = ks;
        rd.din[l] = n->din[l]; rd.dpad[l] = kin; rd.dout[l] = nout;
        rd.bpart[l] = ra.bpart[l];
        rd.woff[l] = (int64_t)n->woff[l]; rd.boff[l] = (int64_t)n->boff[l];
        rd.cum[l] = cum;
        cum += (int64_t)nout * n->din[l] + nout;
    }
    ga.z0[L] = z;
    rd.cum[L] = cum;
    int gx = 1, gy = 1;
    for (int l = 0; l < L; ++l) {
        gx = std::max(gx, ga.a[l].N / BN);
        gy = std::max(gy, ga.a[l].M / BM);
    }
    hipLaunchKernelGGL(gemm_grouped, dim3(gx, gy, z), dim3(256), 0, n->stream, ga);

    rd.g = n->g; rd.lpart = n->lpart; rd.loss = loss_out; rd.loss_by_batch = loss_by_batch; rd.ctr = ctr;
    hipLaunchKernelGGL(reduce_grads, dim3((unsigned)std::min<int64_t>(1024, (cum + 255) / 256)), dim3(256),
                       0, n->stream, rd);
    return CRYSP_OK;
}

static void mlp_adam(crysp_mlp *n, float lr, float b1, float b2, float eps, float c1, float c2,
                     float wd, const int64_t *dstep, const float *dlr)
{
    AdamArgs aa{};
    aa.L = n->n_layers;
    for (int l = 0; l < n->n_layers; ++l) {
        aa.woff[l] = (int64_t)n->woff[l]; aa.boff[l] = (int64_t)n->boff[l];
        aa.din[l] = n->din[l]; aa.dpad[l] = n->dpad[l]; aa.dout[l] = n->dout[l];
        aa.Wt[l] = n->Wt[l]; aa.Wn[l] = n->Wn[l];
    }
    hipLaunchKernelGGL(adam_refresh, dim3((unsigned)((n->n_params + 255) / 256)), dim3(256), 0, n->stream,
                       n->p, n->g, n->m, n->v, (int)n->n_params, lr, b1, b2, eps, c1, c2, wd, dstep, dlr, aa);
}

static void mlp_refresh(crysp_mlp *n)
{
    for (int l = 0; l < n->n_layers; ++l) {
        const int tot = n->dout[l] * n->dpad[l];
        hipLaunchKernelGGL(refresh_bf16, dim3((tot + 255) / 256), dim3(256), 0, n->stream,
                           n->p + n->woff[l], n->dout[l], n->din[l], n->dpad[l], n->Wt[l],
                           l > 0 ? n->Wn[l] : (__bf16 *)nullptr);
    }
}

static bool is_dev(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// copy host/device input to a temporary device buffer if needed
struct Staged {
    void *d = nullptr;
    bool own = false;
    ~Staged() { if (own && d) (void)hipFree(d); }
};
static int stage(hipStream_t s, const void *src, size_t bytes, Staged &st)
{
    if (!src) { st.d = nullptr; return CRYSP_OK; }
    if (is_dev(src)) { st.d = (void *)src; return CRYSP_OK; }
    SCHK(hipMalloc(&st.d, std::max<size_t>(bytes, 4)));
    st.own = true;
    SCHK(hipMemcpyAsync(st.d, src, bytes, hipMemcpyHostToDevice, s));
    return CRYSP_OK;
}

extern "C" {

const char *crysp_mlp_last_error(void) { return s_err.c_str(); }

crysp_mlp *crysp_mlp_create(int device, int32_t n_in, int32_t hidden, int32_t n_hidden_layers,
                            int32_t n_out, uint64_t seed)
{
    if (n_in < 1 || n_in > 32 || hidden < 64 || hidden > AMAX || hidden % 64 || n_out < 64 ||
        n_out > AMAX || n_out % 64 || n_hidden_layers < 1 || n_hidden_layers + 1 > MAXL) {
        sfail(CRYSP_EINVAL, "need 1 <= n_in <= 32, hidden and n_out multiples of 64 up to 256, "
                            "1 to 8 hidden layers");
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) { sfail(CRYSP_EHIP, "hipSetDevice"); return nullptr; }
    auto *n = new crysp_mlp();
    n->device = device;
    n->n_in = n_in; n->hidden = hidden; n->n_out = n_out;
    n->n_layers = n_hidden_layers + 1;
    n->kin_pad = 64;
    for (int l = 0; l < n->n_layers; ++l) {
        const int di = l == 0 ? n_in : hidden, dp = l == 0 ? n->kin_pad : hidden;
        const int dn = l + 1 == n->n_layers ? n_out : hidden;
        n->din.push_back(di); n->dpad.push_back(dp); n->dout.push_back(dn);
        n->woff.push_back(n->n_params); n->n_params += (size_t)dn * di;
        n->boff.push_back(n->n_params); n->n_params += dn;
        n->flops_fwd_per_row += 2.0 * dn * di;
    }
    bool ok = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&n->p, n->n_params * 4) == hipSuccess;
    ok = ok && hipMalloc(&n->g, n->n_params * 4) == hipSuccess;
    ok = ok && hipMalloc(&n->m, n->n_params * 4) == hipSuccess;
    ok = ok && hipMalloc(&n->v, n->n_params * 4) == hipSuccess;
    ok = ok && hipMalloc(&n->shift, 32 * 4) == hipSuccess && hipMalloc(&n->scale, 32 * 4) == hipSuccess;
    ok = ok && hipMalloc(&n->loss, 64) == hipSuccess;
    ok = ok && hipMalloc(&n->ctr, 2 * sizeof(int64_t)) == hipSuccess;
    ok = ok && hipMalloc(&n->dlr, sizeof(float)) == hipSuccess;
    for (int l = 0; l < n->n_layers && ok; ++l) {
        __bf16 *a = nullptr, *b = nullptr;
        ok = hipMalloc(&a, (size_t)n->dout[l] * n->dpad[l] * 2) == hipSuccess;
        ok = ok && hipMalloc(&b, (size_t)n->dpad[l] * n->dout[l] * 2) == hipSuccess;
        n->Wt.push_back(a); n->Wn.push_back(b);
    }
    if (!ok) { sfail(CRYSP_ENOMEM, "MLP allocation failed"); delete n; return nullptr; }
    // He-uniform initialization on the host from a Philox stream (domain 5), zero biases
    std::vector<float> h(n->n_params, 0.0f);
    uint64_t ctr = 0;
    auto uni = [&]() {
        uint32_t o[4], c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32);
        ++ctr;
        // host Philox (same rounds as the device)
        uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32), a0 = c0, a1 = c1, a2 = 0, a3 = 5u;
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * a0, p1 = (uint64_t)0xCD9E8D57u * a2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ a1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ a3 ^ k1;
            a0 = n0; a1 = (uint32_t)p1; a2 = n2; a3 = (uint32_t)p0;
            k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
        }
        o[0] = a0;
        return ((double)(o[0] >> 8) + 0.5) / 16777216.0;
    };
    for (int l = 0; l < n->n_layers; ++l) {
        const double bound = std::sqrt(6.0 / n->din[l]) * (l + 1 == n->n_layers ? 0.5 : 1.0);
        for (size_t i = 0; i < (size_t)n->dout[l] * n->din[l]; ++i)
            h[n->woff[l] + i] = (float)((2.0 * uni() - 1.0) * bound);
    }
    std::vector<float> sh(32, 0.0f), sc(32, 1.0f);
    if (hipMemcpy(n->p, h.data(), n->n_params * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, n->n_params * 4) != hipSuccess ||
        hipMemset(n->v, 0, n->n_params * 4) != hipSuccess ||
        hipMemset(n->g, 0, n->n_params * 4) != hipSuccess ||
        hipMemcpy(n->shift, sh.data(), 128, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->scale, sc.data(), 128, hipMemcpyHostToDevice) != hipSuccess) {
        sfail(CRYSP_EHIP, "MLP initialization copies failed");
        delete n;
        return nullptr;
    }
    for (int l = 0; l < n->n_layers; ++l)
        (void)hipMemset(n->Wn[l], 0, (size_t)n->dpad[l] * n->dout[l] * 2);
    mlp_refresh(n);
    (void)hipStreamSynchronize(n->stream);
    return n;
}

void crysp_mlp_destroy(crysp_mlp *n)
{
    if (!n) return;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    auto fre = [](void *q) { if (q) (void)hipFree(q); };
    mlp_drop_graph(n);
    fre(n->p); fre(n->g); fre(n->m); fre(n->v); fre(n->shift); fre(n->scale); fre(n->loss); fre(n->ctr);
    fre(n->dlr); fre(n->dloss);
    fre(n->part); fre(n->bpart); fre(n->lpart);
    for (auto q : n->Wt) fre(q);
    for (auto q : n->Wn) fre(q);
    fre(n->X); fre(n->Y); fre(n->dY);
    for (auto q : n->Z) fre(q);
    for (auto q : n->H) fre(q);
    for (auto q : n->dZ) fre(q);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

int64_t crysp_mlp_num_params(const crysp_mlp *n) { return n ? (int64_t)n->n_params : -1; }

// flat fp32 parameters, per layer W [out][in] row-major then b [out]
int crysp_mlp_get_params(crysp_mlp *n, float *out)
{
    if (!n || !out) return sfail(CRYSP_EINVAL, "null argument");
    SCHK(hipMemcpy(out, n->p, n->n_params * 4, hipMemcpyDeviceToHost));
    return CRYSP_OK;
}

int crysp_mlp_set_params(crysp_mlp *n, const float *in)
{
    if (!n || !in) return sfail(CRYSP_EINVAL, "null argument");
    SCHK(hipMemcpy(n->p, in, n->n_params * 4, hipMemcpyHostToDevice));
    mlp_refresh(n);
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
}

// input normalization x' = (x - shift) * scale applied inside the network
int crysp_mlp_set_input_norm(crysp_mlp *n, const float *shift, const float *scale)
{
    if (!n || !shift || !scale) return sfail(CRYSP_EINVAL, "null argument");
    SCHK(hipMemcpy(n->shift, shift, n->n_in * 4, hipMemcpyHostToDevice));
    SCHK(hipMemcpy(n->scale, scale, n->n_in * 4, hipMemcpyHostToDevice));
    return CRYSP_OK;
}

int crysp_mlp_set_poisson_outputs(crysp_mlp *n, int32_t n_pois)
{
    if (!n || n_pois < 0 || n_pois > n->n_out) return sfail(CRYSP_EINVAL, "bad n_pois");
    if (n_pois != n->n_pois) mlp_drop_graph(n);     // the captured step holds the old value
    n->n_pois = n_pois;
    return CRYSP_OK;
}

int crysp_mlp_forward(crysp_mlp *n, const float *X, int64_t M, float *Y)
{
    if (!n || !X || !Y || M < 0) return sfail(CRYSP_EINVAL, "bad arguments");
    if (M == 0) return CRYSP_OK;
    SCHK(hipSetDevice(n->device));
    const int64_t chunk = 65536;
    Staged sx;
    int rc = stage(n->stream, X, (size_t)M * n->n_in * 4, sx);
    if (rc) return rc;
    const bool ydev = is_dev(Y);
    if ((rc = mlp_alloc_batch(n, (int)std::min<int64_t>(chunk, ((M + 63) / 64) * 64)))) return rc;
    for (int64_t r0 = 0; r0 < M; r0 += chunk) {
        const int m = (int)std::min<int64_t>(chunk, M - r0);
        const int mpad = ((m + 63) / 64) * 64;
        if ((rc = mlp_forward_dev(n, (const float *)sx.d + r0 * n->n_in, m, mpad))) return rc;
        SCHK(hipMemcpyAsync(Y + r0 * n->n_out, n->Y, (size_t)m * n->n_out * 4,
                            ydev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, n->stream));
    }
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
}

// gradient of the weighted MSE on one batch (no update): grads flat fp32, loss scalar
int crysp_mlp_gradient(crysp_mlp *n, const float *X, const float *T, const float *W, int32_t M,
                       float *grads, float *loss)
{
    if (!n || !X || !T || M < 1) return sfail(CRYSP_EINVAL, "bad arguments");
    SCHK(hipSetDevice(n->device));
    const int mpad = ((M + 63) / 64) * 64;
    int rc = mlp_alloc_batch(n, mpad);
    if (rc) return rc;
    Staged sx, st, sw;
    if ((rc = stage(n->stream, X, (size_t)M * n->n_in * 4, sx)) ||
        (rc = stage(n->stream, T, (size_t)M * n->n_out * 4, st)) ||
        (rc = stage(n->stream, W, (size_t)M * n->n_out * 4, sw)))
        return rc;
    double wsum = (double)M * n->n_out;
    if (W) {
        std::vector<float> hw((size_t)M * n->n_out);
        SCHK(hipMemcpy(hw.data(), sw.d, hw.size() * 4, hipMemcpyDeviceToHost));
        wsum = 0.0;
        for (float x : hw) wsum += x;
        if (!(wsum > 0.0)) return sfail(CRYSP_EINVAL, "loss weights sum to zero");
    }
    if ((rc = mlp_grad_dev(n, (const float *)sx.d, (const float *)st.d, (const float *)sw.d, M, mpad,
                           wsum, n->loss, false, nullptr)))
        return rc;
    if (grads) SCHK(hipMemcpyAsync(grads, n->g, n->n_params * 4, hipMemcpyDeviceToHost, n->stream));
    if (loss) SCHK(hipMemcpyAsync(loss, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
}

// One Adam step per mini-batch on consecutive device-resident rows: batch b uses rows
// [b*batch, (b+1)*batch) of X, T, W (the caller shuffles between epochs).  The loss of
// every step is written to losses[b] (optional).
int crysp_mlp_train_batches(crysp_mlp *n, const float *X, const float *T, const float *W,
                            int64_t n_batches, int32_t batch, float lr, float beta1, float beta2,
                            float eps, float weight_decay, float mean_row_weight, float *losses)
{
    if (!n || !X || !T || n_batches < 0 || batch < 64 || batch % 64)
        return sfail(CRYSP_EINVAL, "bad arguments (batch must be a multiple of 64)");
    if (!is_dev(X) || !is_dev(T) || (W && !is_dev(W)))
        return sfail(CRYSP_EINVAL, "training arrays must be device-resident");
    SCHK(hipSetDevice(n->device));
    int rc = mlp_alloc_batch(n, batch);
    if (rc) return rc;
    const double wsum = (double)batch * n->n_out * (W ? mean_row_weight : 1.0f);
    if (losses && n_batches > n->dloss_cap) {
        mlp_drop_graph(n);
        if (n->dloss) (void)hipFree(n->dloss);
        n->dloss = nullptr;
        n->dloss_cap = 0;
        SCHK(hipMalloc(&n->dloss, (size_t)n_batches * 4));
        n->dloss_cap = n_batches;
    }
    float *dl = losses && n_batches > 0 ? n->dloss : nullptr;
    // A step (mlp_rows, the grouped weight gradients, their reduction, Adam with the bf16
    // refresh) reads the batch index, the Adam step count and the learning rate from the
    // device, the first two advancing inside the step, so S identical steps captured in one
    // graph run S batches in a row (S the largest divisor of n_batches up to 64): one graph
    // launch per S steps.  The graph is kept for later calls with the same arrays, batch,
    // step count per graph and optimizer settings (an epoch loop over reshuffled copies in
    // the same buffers).
    auto enqueue_step = [&](const float *x, const float *t, const float *w, float *lout,
                            int64_t *ctr, const float *dlr, float c1, float c2) -> int {
        int r;
        if ((r = mlp_grad_dev(n, x, t, w, batch, batch, wsum, lout ? lout : n->loss, lout != nullptr, ctr)))
            return r;
        mlp_adam(n, lr, beta1, beta2, eps, c1, c2, weight_decay, ctr ? ctr + 1 : nullptr, dlr);
        return CRYSP_OK;
    };
    bool graph = false;
    int steps = 1;
    for (int d = (int)std::min<int64_t>(64, n_batches); d >= 1; --d)
        if (n_batches % d == 0) { steps = d; break; }
    if (n_batches > 1) {
        n->hctr[0] = 0;
        n->hctr[1] = n->step;
        SCHK(hipMemcpyAsync(n->ctr, n->hctr, sizeof n->hctr, hipMemcpyHostToDevice, n->stream));
        SCHK(hipMemcpyAsync(n->dlr, &lr, sizeof lr, hipMemcpyHostToDevice, n->stream));
        const bool same = n->gexec && n->gX == X && n->gT == T && n->gW == W && n->gbatch == batch &&
                          n->gsteps == steps &&
                          n->glosses == (dl != nullptr) && n->gb1 == beta1 && n->gb2 == beta2 &&
                          n->geps == eps && n->gwd == weight_decay && n->gmrw == mean_row_weight;
        if (!same) {
            mlp_drop_graph(n);
            SCHK(hipStreamSynchronize(n->stream));
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                for (int k = 0; k < steps && rc == CRYSP_OK; ++k)
                    rc = enqueue_step(X, T, W, dl, n->ctr, n->dlr, 0.0f, 0.0f);
                const hipError_t e = hipStreamEndCapture(n->stream, &g);
                if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
                if (e != hipSuccess || !g || hipGraphInstantiate(&n->gexec, g, nullptr, nullptr, 0) != hipSuccess)
                    n->gexec = nullptr;
                if (g) (void)hipGraphDestroy(g);
            }
            if (!n->gexec) (void)hipGetLastError();     // capture unavailable: launch eagerly
            n->gX = X; n->gT = T; n->gW = W; n->gbatch = batch; n->gsteps = steps; n->glosses = dl != nullptr;
            n->gb1 = beta1; n->gb2 = beta2; n->geps = eps; n->gwd = weight_decay; n->gmrw = mean_row_weight;
        }
        graph = n->gexec != nullptr;
    }
    if (graph) {
        for (int64_t b = 0; b < n_batches; b += steps) SCHK(hipGraphLaunch(n->gexec, n->stream));
        n->step += (int)n_batches;
    } else {
        for (int64_t b = 0; b < n_batches; ++b) {
            const float *x = X + (size_t)b * batch * n->n_in;
            const float *t = T + (size_t)b * batch * n->n_out;
            const float *w = W ? W + (size_t)b * batch * n->n_out : nullptr;
            n->step += 1;
            const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)n->step));
            const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)n->step));
            if ((rc = enqueue_step(x, t, w, dl ? dl + b : nullptr, nullptr, nullptr, c1, c2))) return rc;
        }
    }
    if (dl) SCHK(hipMemcpyAsync(losses, dl, (size_t)n_batches * 4, hipMemcpyDeviceToHost, n->stream));
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
}

double crysp_mlp_flops_per_row(const crysp_mlp *n) { return n ? n->flops_fwd_per_row : 0.0; }

// Poisson-sampled event maps from a baked surrogate grid (see sample_events_grid).
int crysp_surrogate_sample_events(int device, const float *grid, int32_t gx, int32_t gy,
                                  int32_t gz, int32_t ncell, const crysp_box *box,
                                  const float *dep_xyzt, const float *dep_e_kev,
                                  const uint8_t *use, const int64_t *ev_first, int64_t ndep,
                                  int64_t nev, double yield_per_mev, uint64_t seed,
                                  int32_t *maps, float *tmean, float *lam)
{
    if (!grid || !box || !ev_first || !maps || nev < 0 || ndep < 0 || gx < 2 || gy < 2 || gz < 2)
        return sfail(CRYSP_EINVAL, "bad arguments");
    if (nev == 0) return CRYSP_OK;
    SCHK(hipSetDevice(device));
    hipStream_t s = nullptr;
    SCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t gbytes = (size_t)gx * gy * gz * 2 * ncell * 4;
    Staged sg, sd, se, su, sf;
    int rc;
    if ((rc = stage(s, grid, gbytes, sg)) || (rc = stage(s, dep_xyzt, (size_t)ndep * 16, sd)) ||
        (rc = stage(s, dep_e_kev, (size_t)ndep * 4, se)) || (rc = stage(s, use, (size_t)ndep, su)) ||
        (rc = stage(s, ev_first, (size_t)(nev + 1) * 8, sf))) {
        (void)hipStreamDestroy(s);
        return rc;
    }
    const size_t cells = (size_t)nev * ncell;
    const bool mdev = is_dev(maps), tdev = is_dev(tmean), ldev = is_dev(lam);
    int32_t *dm = maps;
    float *dt = tmean, *dlam = lam;
    if (!mdev) SCHK(hipMalloc(&dm, cells * 4));
    if (tmean && !tdev) SCHK(hipMalloc(&dt, cells * 4));
    if (lam && !ldev) SCHK(hipMalloc(&dlam, cells * 4));
    hipLaunchKernelGGL(sample_events_grid, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s,
                       (const float *)sg.d, gx, gy, gz, ncell, box->L[0], box->L[1], box->L[2],
                       (const float4 *)sd.d, (const float *)se.d, (const uint8_t *)su.d,
                       (const int64_t *)sf.d, nev, yield_per_mev, seed, dm, dt, dlam);
    SCHK(hipGetLastError());
    if (!mdev) SCHK(hipMemcpyAsync(maps, dm, cells * 4, hipMemcpyDeviceToHost, s));
    if (tmean && !tdev) SCHK(hipMemcpyAsync(tmean, dt, cells * 4, hipMemcpyDeviceToHost, s));
    if (lam && !ldev) SCHK(hipMemcpyAsync(lam, dlam, cells * 4, hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    if (!mdev) (void)hipFree(dm);
    if (tmean && !tdev) (void)hipFree(dt);
    if (lam && !ldev) (void)hipFree(dlam);
    (void)hipStreamDestroy(s);
    return CRYSP_OK;
}

// Debug / layout check: C = A [M x K] * B^T (B stored [N][K]) in bf16 -> fp32.
int crysp_debug_gemm(int device, int32_t M, int32_t N, int32_t K, const uint16_t *A_bf16,
                     int32_t a_t, const uint16_t *B_bf16, int32_t b_t, float *C)
{
    if (M % BM || N % BN || K % BK || !A_bf16 || !B_bf16 || !C) return sfail(CRYSP_EINVAL, "bad shapes");
    SCHK(hipSetDevice(device));
    __bf16 *a, *b;
    float *c;
    SCHK(hipMalloc(&a, (size_t)M * K * 2));
    SCHK(hipMalloc(&b, (size_t)K * N * 2));
    SCHK(hipMalloc(&c, (size_t)M * N * 4));
    SCHK(hipMemcpy(a, A_bf16, (size_t)M * K * 2, hipMemcpyHostToDevice));
    SCHK(hipMemcpy(b, B_bf16, (size_t)K * N * 2, hipMemcpyHostToDevice));
    GemmArgs g{};
    g.M = M; g.N = N; g.K = K;
    g.A = a; g.lda = a_t ? M : K; g.a_t = a_t;
    g.B = b; g.ldb = b_t ? K : N; g.b_t = b_t;
    g.C = c; g.ldc = N;
    gemm<EPI_F32>(nullptr, g);
    SCHK(hipGetLastError());
    SCHK(hipMemcpy(C, c, (size_t)M * N * 4, hipMemcpyDeviceToHost));
    (void)hipFree(a); (void)hipFree(b); (void)hipFree(c);
    return CRYSP_OK;
}

}  // extern "C"
