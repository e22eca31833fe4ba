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
// the bits of to_bf16(x), in the low half of a word
__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    const __bf16 h = to_bf16(x);
    return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float from_bf16(__bf16 x) { return (float)x; }

// GELU in its tanh form, 0.5 x (1 + tanh(u)) = x s with s = sigma(2u),
// u = sqrt(2/pi) (x + 0.044715 x^3), and its derivative s + x s (1 - s) (2u)'.  The network
// is trained and evaluated with this activation; it differs from the erf form by less than
// 5e-4 everywhere.  s comes from the hardware exp and reciprocal (about 16 instructions for
// both values, against about 60 for erf and its derivative): an exp that overflows gives
// s = 0, one that underflows s = 1, so the activation is finite everywhere.
__device__ __forceinline__ float gelu_s(float x)
{
    const float v = x * __builtin_fmaf(x * x, 0.07135481627159416f, 1.5957691216057308f);   // 2u
    return __builtin_amdgcn_rcpf(1.0f + __expf(-v));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float gelu_grad(float x)
{
    const float s = gelu_s(x);
    const float dv = __builtin_fmaf(x * x, 0.21406444881478248f, 1.5957691216057308f);      // (2u)'
    return __builtin_fmaf(x * s * (1.0f - s), dv, s);
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

constexpr int MAXL = 9;                             // layers of the MLP, at most

// ------------------------------------------------------------- weight gradients ----
// dW_l [out x in_pad] = dZ_l^T H_{l-1} over the batch: A = dZT (out wide) and B = HT (or
// the packed inputs XT, in_pad wide), both in the chunked transposed layout mlp_rows stores:
// element (batch row k, column m) of a width-D operand at ((k / 8) D + m) 8 + k % 8, so the
// 8 batch rows an MFMA operand lane needs form one 16-byte chunk.  The batch is
// split into ks slices of kslice rows (ks at most DW_KS whatever the batch, so the
// reduction stays short); slice s's partial product goes to C + s * out * in_pad and
// reduce_grads adds the slices in order.  A workgroup computes one DW_M x BN tile of one
// slice (BM, BN = 128, or 64 where the layer is not a multiple of 128 wide), four waves of
// BM/2 x BN/2, with K staged
// DW_KP at a time through two LDS buffers: the next stage's global loads are in flight
// during the current stage's MFMAs, one barrier per stage.  The workgroups of the first
// column tile also add up their staged dZ rows: the slice's partial of the bias gradient,
// in a fixed order (each thread half a row per stage, in k order, stage after stage; the
// two halves added last).
constexpr int DW_M = 128, DW_KP = 64, DW_MAXN = 128, DW_KS = 32;
// elements between consecutive 8-row groups of a chunked transposed operand of the given
// width (padding these groups apart measured no faster)
__host__ __device__ constexpr int64_t chunk_stride(int64_t width) { return width * 8; }

struct DwArgs {
    const __bf16 *A[MAXL], *B[MAXL];               // dZT, HT (XT for layer 0)
    int M[MAXL], N[MAXL];                          // out, in_pad (multiples of 64)
    float *C[MAXL];                                // partial products [ks][M][N]
    float *db[MAXL];                               // bias partials [ks][M]
    int t0[MAXL + 1];                              // flat tile ranges of the layers
    int n;                                         // layers
    int kslice;                                    // batch rows per slice, a multiple of DW_KP
};

template <int BM_, int BN>
__device__ __forceinline__ void dw_tile(const DwArgs &a, int l, int tile, int slice,
                                        __bf16 (*sA)[DW_M][DW_KP + LDS_PAD],
                                        __bf16 (*sB)[DW_MAXN][DW_KP + LDS_PAD])
{
    constexpr int MT = BM_ / 64, NT = BN / 64;     // 32 x 32 MFMA tiles per wave: MT x NT
    constexpr int CA = BM_ * (DW_KP / 8) / 256;    // 16-byte chunks per thread per stage
    constexpr int CB = BN * (DW_KP / 8) / 256;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ntn = a.N[l] / BN;
    const int m0 = (tile / ntn) * BM_, n0 = (tile % ntn) * BN;
    const int wm = (wave >> 1) * (BM_ / 2), wn = (wave & 1) * (BN / 2);
    const int64_t Ma = a.M[l], Nb = a.N[l];        // operand widths
    const int64_t kb = (int64_t)slice * a.kslice;
    const int64_t sa = chunk_stride(Ma), sb = chunk_stride(Nb);
    const __bf16 *Ag = a.A[l] + kb / 8 * sa + m0 * 8;
    const __bf16 *Bg = a.B[l] + kb / 8 * sb + n0 * 8;
    const bool rows = n0 == 0 && a.db[l];
    // chunk c = tid + 256 u of a stage of A: column m0 + c % BM_, batch rows k0 + 8 (c / BM_)
    // onwards (B likewise with BN), so the 64 lanes of a wave read 64 consecutive chunks of
    // one 8-row group, one contiguous 1 KB run, and write them to LDS rows 36 dwords apart:
    // each 8 lanes of a 16-byte write cover the 32 banks once
    bf16x8 ra[CA], rb[CB];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < CA; ++u) {
            const int c = tid + 256 * u;
            ra[u] = *(const bf16x8 *)(Ag + (k0 / 8 + c / BM_) * sa + (c % BM_) * 8);
        }
#pragma unroll
        for (int u = 0; u < CB; ++u) {
            const int c = tid + 256 * u;
            rb[u] = *(const bf16x8 *)(Bg + (k0 / 8 + c / BN) * sb + (c % BN) * 8);
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < CA; ++u) {
            const int c = tid + 256 * u;
            *(bf16x8 *)&sA[b][c % BM_][(c / BM_) * 8] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < CB; ++u) {
            const int c = tid + 256 * u;
            *(bf16x8 *)&sB[b][c % BN][(c / BN) * 8] = rb[u];
        }
    };
    f32x16 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
    float rsum = 0.0f;                             // bias rows: row tid / 2, half tid % 2
    const int nst = a.kslice / DW_KP;
    load(0);
    store(0);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int cur = st & 1;
        if (st + 1 < nst) load((st + 1) * DW_KP);
#pragma unroll
        for (int kk = 0; kk < DW_KP; kk += 16) {
            const int r = lane & 31, kq = kk + 8 * (lane >> 5);
            bf16x8 fa[MT], fb[NT];
#pragma unroll
            for (int i = 0; i < MT; ++i) fa[i] = *(const bf16x8 *)&sA[cur][wm + 32 * i + r][kq];
#pragma unroll
            for (int j = 0; j < NT; ++j) fb[j] = *(const bf16x8 *)&sB[cur][wn + 32 * j + r][kq];
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (rows && tid < 2 * BM_) {
            const __bf16 *q = &sA[cur][tid >> 1][(tid & 1) * (DW_KP / 2)];
#pragma unroll
            for (int c = 0; c < DW_KP / 2; c += 8) {
                const bf16x8 v = *(const bf16x8 *)(q + c);
#pragma unroll
                for (int e = 0; e < 8; ++e) rsum += from_bf16(v[e]);
            }
        }
        if (st + 1 < nst) store(cur ^ 1);
        __syncthreads();
    }
    float *C = a.C[l] + (size_t)slice * a.M[l] * a.N[l];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int col = n0 + wn + 32 * j + (lane & 31);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int row = m0 + wm + 32 * i + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                C[(size_t)row * a.N[l] + col] = acc[i][j][e];
            }
        }
    if (rows) {
        const float other = __shfl_xor(rsum, 1);
        if ((tid & 1) == 0 && tid < 2 * BM_) a.db[l][(size_t)slice * a.M[l] + m0 + (tid >> 1)] = rsum + other;
    }
}

__global__ void __launch_bounds__(256) dw_grouped(const DwArgs a)
{
    __shared__ __bf16 sA[2][DW_M][DW_KP + LDS_PAD];
    __shared__ __bf16 sB[2][DW_MAXN][DW_KP + LDS_PAD];
    // Workgroups go to the 8 XCDs in turn (linear id mod 8), and each XCD has its own L2.
    // The tiles of one batch slice read the same rows of the stored activations and
    // gradients, so they are placed on one XCD: XCD x takes slices x S/8 .. (x+1) S/8 - 1,
    // tile by tile, and a tile's re-read of its operand rows can hit that XCD's L2.
    int tid_all = blockIdx.x, slice = blockIdx.y;
    const int T = gridDim.x, S = gridDim.y;
    if (S % 8 == 0) {
        const int id = blockIdx.x + blockIdx.y * T;
        const int j = id / 8;
        slice = (id % 8) * (S / 8) + j / T;
        tid_all = j % T;
    }
    int l = 0;
    while (l + 1 < a.n && tid_all >= a.t0[l + 1]) ++l;
    const int tile = tid_all - a.t0[l];
    const bool wm = a.M[l] % 128 == 0, wn = a.N[l] % 128 == 0;
    if (wm && wn) dw_tile<128, 128>(a, l, tile, slice, sA, sB);
    else if (wm) dw_tile<128, 64>(a, l, tile, slice, sA, sB);
    else if (wn) dw_tile<64, 128>(a, l, tile, slice, sA, sB);
    else dw_tile<64, 64>(a, l, tile, slice, sA, sB);
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
// loss partials, one workgroup per 64 or 128 rows), dw_grouped (the weight gradients of every layer
// as split-K products, with the bias row sums), reduce_grads (the partials summed in a fixed
// order) and adam_refresh (Adam, the bf16 weight copies).  Every sum runs in a fixed order, so a
// training run is reproducible.

constexpr int AMAX = 256;                           // widest layer (hidden, n_out, kin_pad)
constexpr int CBLK = AMAX / 32;                     // 32-column blocks of the widest layer
// mlp_rows: NB 32-row blocks per wave (1 or 2), so 64 NB rows per workgroup in two row
// groups; 16 waves, wave w on column block w % CBLK of row group w / CBLK.  Two blocks per
// wave halve the weight staging and barriers per row; one block gives twice the workgroups,
// for batches too small to fill the chip with the larger ones.
constexpr int RB_MIN = 64, RB_MAX = 128;            // rows per workgroup: NB = 1, 2
constexpr int RT = 64 * CBLK * 2;                   // threads per workgroup
constexpr int KB = 128;                             // K staged in LDS per phase (mlp_rows)

struct RowArgs {
    int L, M, Mpad, n_in, kin_pad, n_out, n_pois;
    int din[MAXL], dpad[MAXL], dout[MAXL];
    const __bf16 *Wt[MAXL];                         // [out][in_pad]
    const __bf16 *Wn[MAXL];                         // [in][out]
    const float *bias[MAXL];
    __bf16 *HT[MAXL];                               // hidden layers: activation, chunked transposed
    __bf16 *dZT[MAXL];                              // gradient of layer l's pre-activation, the same
    __bf16 *XT;                                     // packed inputs, the same (kin_pad wide)
    float *lpart;                                   // [n_wg] loss partials
    const float *Xraw, *T, *W, *shift, *scale;
    const float *R0;                                // rate offsets [M][n_pois] or null
    float inv_wsum;
    const int64_t *bidx;                            // batch index (graph replays) or null
    float *Y;                                       // inference: the outputs [M][n_out] fp32
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
    // the forward GEMMs only (inference), or forward and backward (training step)
    __device__ __forceinline__ BStream(const RowArgs &ra, bool train) : a(ra), Q(0)
    {
        for (int g = 0; g < (train ? 2 * a.L - 1 : a.L); ++g) {
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

// out[64 NB x N] = A[64 NB x K] * B^T (+ bias) for the next GEMM of the stream (phases
// q .. q + np - 1), A resident in LDS, all 16 waves sharing every staged slice of B.  Wave w
// owns column block w % CBLK of the rows of row group w / CBLK: NB 32 x 32 blocks that share
// every B fragment it reads.  A forward layer and its backward product have the same width,
// so each element stays with the lane that made it (gelu' is kept in registers between the
// two).  SPREAD (NB = 2, at most 4 column blocks; for the last layer, whose gelu' is not
// kept): the 16 or fewer blocks go one to a wave instead of two to each of 8 waves, so all
// 16 waves share the epilogue.  Lane element i (0 .. 16 NB - 1) of the wave's blocks sits at
// (row, col); epi(i, row, col, acc + bias[col]) stores it.  The bias (null: none) is read
// before the K loop, so the load completes under the MFMAs.  The K loop ends with a barrier,
// so the epilogue may write over A.
template <int NB, bool SPREAD, class Epi>
__device__ __forceinline__ void rows_gemm(const __bf16 (*A)[AMAX + LDS_PAD], int K, int N,
                                          BStream &bs, int &q, __bf16 (*sB)[KB + LDS_PAD],
                                          const float *bias, Epi epi)
{
    static_assert(!SPREAD || NB == 2, "one block per wave already");
    constexpr bool two = NB == 2 && !SPREAD;        // a second block in acc1
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ncb = N / 32;
    const int cb = SPREAD ? wave % ncb : wave % CBLK;
    const int rbase = SPREAD ? (wave / ncb) * 32 : (wave / CBLK) * 32 * NB;
    const bool mine = SPREAD ? wave < 4 * ncb : cb < ncb;
    const int col = cb * 32 + (lane & 31);
    const float u = mine && bias ? bias[col] : 0.0f;
    f32x16 acc0, acc1;                              // acc1: the second block (two)
#pragma unroll
    for (int i = 0; i < 16; ++i) { acc0[i] = 0.0f; acc1[i] = 0.0f; }
    for (int k0 = 0; k0 < K; k0 += KB, ++q) {
        const int kp = min(KB, K - k0);             // a multiple of 64
        bs.stash(N, kp, sB);
        __syncthreads();
        bs.load(q + 1);                             // the next phase, maybe of the next GEMM
        if (mine) {
            for (int kk = 0; kk < kp; kk += 16) {
                const int r = lane & 31, kq = kk + 8 * (lane >> 5);
                const bf16x8 bv = *(const bf16x8 *)&sB[cb * 32 + r][kq];
                const bf16x8 a0 = *(const bf16x8 *)&A[rbase + r][k0 + kq];
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bv, acc0, 0, 0, 0);
                if constexpr (two) {
                    const bf16x8 a1 = *(const bf16x8 *)&A[rbase + 32 + r][k0 + kq];
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bv, acc1, 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    if (mine) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            epi(i, rbase + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), col, acc0[i] + u);
        if constexpr (two) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                epi(16 + i, rbase + 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), col, acc1[i] + u);
        }
    }
}

// The RB rows of act (bf16, N columns) into dst in the chunked transposed layout of the
// batch: element (row r, column c) at ((r / 8) N + c) 8 + r % 8, i.e. for each 8-row chunk
// of the batch, the columns in order, each as the chunk's 8 rows (16 bytes).  A thread
// gathers 8 rows of a column from LDS and stores them as one chunk; the lanes of a wave take
// 64 consecutive columns of one 8-row chunk (N is a multiple of 64), so their 2-byte LDS
// reads of a row fall in 32 distinct banks and their stores are one contiguous 1 KB run.
// (With dst as [column][batch row] instead, a wave's 64 chunks went to 64 separate 16-byte
// pieces of 64 lines, and lanes along the rows for contiguous runs put every LDS read of a
// wave in one bank: rows 8 apart are 1056 dwords apart, a multiple of the 32 banks.)
template <int RB>
__device__ __forceinline__ void rows_store_T(const __bf16 (*act)[AMAX + LDS_PAD], int N, __bf16 *dst, int r0)
{
    const int64_t sd = chunk_stride(N);
    __bf16 *d = dst + (r0 >> 3) * sd;
    // chunk c = threadIdx.x + u RT is (8-row group c / N, column c % N): one division per
    // thread, then steps of RT (RT / N groups and RT % N columns, with the carry)
    const int dg = RT / N, dc = RT % N;
    int g = threadIdx.x / N, col = threadIdx.x % N;
    while (g < RB / 8) {
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = act[8 * g + j][col];
        *(bf16x8 *)(d + g * sd + col * 8) = v;
        g += dg;
        col += dc;
        if (col >= N) { col -= N; ++g; }
    }
}

// One workgroup per RB = 64 NB rows of the batch:
//   pack (x - shift) * scale to bf16;
//   forward: hidden layers Z = A W^T + b, H = gelu(Z) (bf16) and gelu'(Z) (bf16, for the
//   backward);
//   loss of the last layer's output y against t with weight w, normalized by 1 / sum w:
//     columns c <  n_pois: Poisson NLL of a log-rate, w (e^y - t y), gradient w (e^y - t);
//     columns c >= n_pois: weighted squared error w (y - t)^2, gradient 2 w (y - t);
//   backward data path: dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1});
//   the loss partial (the bias gradients come from dw_grouped, which reads every dZ_l).
// Each layer's rows overwrite the previous layer's in one LDS buffer (a GEMM's K loop has
// ended, with a barrier, before its epilogue writes).
// Padded rows (>= M) carry zero inputs and a zero gradient.
// TRAIN = false is the inference pass: the forward layers only, the last one's output (with
// its bias) stored to a.Y, nothing else written.
template <int L, bool TRAIN, int NB, bool SPREAD_LAST>
__global__ void __launch_bounds__(RT) mlp_rows(const RowArgs a)
{
    constexpr int RB = 64 * NB;
    __shared__ __bf16 act[RB][AMAX + LDS_PAD];
    __shared__ __bf16 sB[AMAX][KB + LDS_PAD];
    __shared__ float red[RT / 64];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * RB;
    const int64_t b = a.bidx ? *a.bidx : 0;
    const float *Xraw = a.Xraw + b * (int64_t)a.M * a.n_in;
    const float *T = a.T + b * (int64_t)a.M * a.n_out;
    const float *W = a.W ? a.W + b * (int64_t)a.M * a.n_out : nullptr;
    const float *R0 = a.R0 ? a.R0 + b * (int64_t)a.M * a.n_pois : nullptr;
    BStream bs(a, TRAIN);
    bs.load(0);
    int q = 0;
    // the product act * B^T of the stream's next GEMM (the block mappings: rows_gemm)
    auto gemm = [&](int K, int N, const float *bias, auto epi) {
        rows_gemm<NB, false>(act, K, N, bs, q, sB, bias, epi);
    };
    auto gemm_last = [&](int K, int N, const float *bias, auto epi) {
        rows_gemm<NB, SPREAD_LAST>(act, K, N, bs, q, sB, bias, epi);
    };

    {   // element i = tid + u RT is (row i / kin_pad, column i % kin_pad), stepped as in rows_store_T
        const int kp = a.kin_pad, dr = RT / kp, dc = RT % kp;
        int r = tid / kp, c = tid % kp;
        while (r < RB) {
            float v = 0.0f;
            if (r0 + r < a.M && c < a.n_in) v = (Xraw[(size_t)(r0 + r) * a.n_in + c] - a.shift[c]) * a.scale[c];
            act[r][c] = to_bf16(v);
            r += dr;
            c += dc;
            if (c >= kp) { c -= kp; ++r; }
        }
    }
    __syncthreads();
    if (TRAIN) rows_store_T<RB>(act, a.kin_pad, a.XT, r0);

    float lsum = 0.0f;
    // gelu'(Z) of the hidden layers, kept in registers: the backward product of layer l + 1
    // has the forward layer l's shape, and every element stays with the lane that made it
    // (as bf16, two to a register: 8 NB registers per hidden layer for the wave's 16 NB
    // elements)
    uint32_t gpr[L - 1][8 * NB];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const int N = a.dout[l];
        if (l + 1 < L) {
            gemm(a.dpad[l], N, a.bias[l], [&](int i, int row, int col, float v) {
                if (TRAIN) {
                    const uint32_t hb = bf16_bits(gelu_grad(v));
                    uint32_t &w = gpr[l < L - 1 ? l : 0][i >> 1];
                    w = (i & 1) ? (w | (hb << 16)) : hb;
                }
                act[row][col] = to_bf16(gelu(v));
            });
        } else if (!TRAIN) {
            const int M = a.M;
            float *Y = a.Y;
            gemm_last(a.dpad[l], N, a.bias[l], [&](int, int row, int col, float y) {
                const int rg = r0 + row;
                if (rg < M) Y[(size_t)rg * N + col] = y;
            });
        } else {
            const int M = a.M, n_pois = a.n_pois;
            const float inv_wsum = a.inv_wsum;
            gemm_last(a.dpad[l], N, a.bias[l], [&](int, int row, int col, float y) {
                // targets, weights and rate offsets are read here (four waves per SIMD hide
                // the latency; the targets alone loaded before the K loop measured no faster,
                // and all three would not fit in registers)
                const int rg = r0 + row;
                float g = 0.0f;
                if (rg < M) {
                    const size_t o = (size_t)rg * N + col;
                    const float t = T[o];
                    const float w = W ? W[o] : 1.0f;
                    const float r = (R0 && col < n_pois) ? R0[(size_t)rg * n_pois + col] : 0.0f;
                    if (col < n_pois) {
                        // rate = e^y + r (r: a known part of the rate, 0 without offsets):
                        // w (rate - t log rate), gradient w (1 - t/rate) e^y
                        const float e = __expf(y);
                        if (R0) {
                            const float rate = e + r;
                            lsum += w * (rate - t * __logf(rate)) * inv_wsum;
                            g = w * (e - t * (e / rate)) * inv_wsum;
                        } else {
                            lsum += w * (e - t * y) * inv_wsum;
                            g = w * (e - t) * inv_wsum;
                        }
                    } else {
                        const float d = y - t;
                        lsum += w * d * d * inv_wsum;
                        g = 2.0f * w * d * inv_wsum;
                    }
                }
                act[row][col] = to_bf16(g);
            });
        }
        __syncthreads();
        if (TRAIN) rows_store_T<RB>(act, N, l + 1 < L ? a.HT[l] : a.dZT[l], r0);
    }
    if (!TRAIN) return;

#pragma unroll
    for (int l = L - 1; l >= 1; --l) {
        const int N = a.din[l];                     // = dout[l - 1]
        gemm(a.dout[l], N, nullptr, [&](int i, int row, int col, float v) {
            const uint32_t w = gpr[l - 1][i >> 1];
            act[row][col] = to_bf16(v * __uint_as_float((i & 1) ? (w & 0xFFFF0000u) : (w << 16)));
        });
        __syncthreads();
        rows_store_T<RB>(act, N, a.dZT[l - 1], r0);
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
    const float *db[MAXL];                          // bias partials of the slices [ks][dout]
    int64_t woff[MAXL], boff[MAXL], cum[MAXL + 1];  // cum: outputs (dout * din + dout) before layer l
    int64_t bcum[MAXL + 1];                         // biases (dout) before layer l
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
// in-order sum of the slices' row sums; the loss is the in-order sum of the workgroup
// partials.  With ctr (graph replays) the batch index and the Adam step count
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
        }
    }
    // biases: the slices' partials in order
    for (int64_t bi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; bi < a.bcum[a.L];
         bi += (int64_t)gridDim.x * blockDim.x) {
        int l = 0;
        while (bi >= a.bcum[l + 1]) ++l;
        const int o = (int)(bi - a.bcum[l]);
        a.g[a.boff[l] + o] = sum_strided(a.db[l] + o, a.ks[l], a.dout[l]);
    }
    // the loss: lane t of the first wave adds partials t, t + 64, ... in order, then a fixed
    // butterfly over the lanes (one thread alone spent a load round trip per 16 partials)
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        const int lane = threadIdx.x;
        float s = 0.0f;
#pragma unroll 8
        for (int u = lane; u < a.n_wg; u += 64) s += a.lpart[u];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
        if (lane == 0) {
            a.loss[a.ctr && a.loss_by_batch ? a.ctr[0] : 0] = s;
            if (a.ctr) {
                a.ctr[0] += 1;
                a.ctr[1] += 1;
            }
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


// ------------------------------------------------------------------ direct light ----
// The part of a point source's light map that reaches the readout plane z = Lz straight
// from the emission point: a photon emitted at depth d = Lz - z with polar angle theta
// (u = cos theta) towards the plane lands at distance rho = d tan theta from the foot
// point (x, y), crosses the crystal -> coupling interface with the Fresnel transmission
// T(u) (zero beyond total internal reflection), survives the bulk over the path d / u, and
// is detected with the PDE in the SiPM cell under the landing point.  A landing point
// outside the face means the photon meets a side wall first: not direct.  For cell c,
//   D_c = PDE / (4 pi) * int dphi [ H(u(rho_in(phi)), k) - H(u(rho_out(phi)), k) ],
//   H(u, k) = int_u^1 T(v) e^{-k / v} dv,   k = d / Lambda  (Lambda: absorption and
//   Rayleigh lengths combined), u(rho) = d / sqrt(rho^2 + d^2),
// where [rho_in, rho_out] is where the ray at azimuth phi lies in the cell.  H is
// tabulated on the host; the azimuth integral runs piecewise between the directions of
// the cell's corners (where rho_in, rho_out change edge) with 8-point Gauss-Legendre.
struct DirectTable {
    const float *H;          // [nk][nu]: H at u = uc + (1 - uc) s^2, s = j / (nu - 1)
    int nu, nk;
    float uc;                // cosine of the critical angle (0 without TIR)
    float kmax;              // k of the last row (0 when nk == 1)
    float inv_lambda;        // 1 / absorption length + 1 / Rayleigh length [1/mm]
    float tan_c;             // tan of the critical angle (+Inf without TIR)
    float pde;
    float Lx, Ly, pitch;
    int nx, ny;
    int images;              // add the side walls' mirror images (air-gap wrap: TIR)
};

__device__ __forceinline__ float direct_H(const DirectTable &t, float u, float k)
{
    float s = u > t.uc ? sqrtf((u - t.uc) / (1.0f - t.uc)) : 0.0f;
    s = fminf(s, 1.0f);
    const float fs = s * (float)(t.nu - 1);
    const int j = min((int)fs, t.nu - 2);
    const float w = fs - (float)j;
    if (t.nk == 1) return t.H[j] + w * (t.H[j + 1] - t.H[j]);
    const float fk = fminf(k / t.kmax, 1.0f) * (float)(t.nk - 1);
    const int m = min((int)fk, t.nk - 2);
    const float wk = fk - (float)m;
    const float *a = t.H + (size_t)m * t.nu, *b = a + t.nu;
    const float ha = a[j] + w * (a[j + 1] - a[j]);
    const float hb = b[j] + w * (b[j + 1] - b[j]);
    return ha + wk * (hb - ha);
}

// int over phi in [a0, a1] of [H(u_in) - H(u_out)] for the ray from (x0, y0) through the
// rectangle [xa, xb] x [ya, yb], 8-point Gauss-Legendre
__device__ float direct_piece(const DirectTable &t, float x0, float y0, float d, float k,
                              float xa, float xb, float ya, float yb, float a0, float a1)
{
    constexpr float X[4] = {0.18343464249564980f, 0.52553240991632899f, 0.79666647741362674f,
                            0.96028985649753623f};
    constexpr float W[4] = {0.36268378337836198f, 0.31370664587788729f, 0.22238103445337447f,
                            0.10122853629037626f};
    const float mid = 0.5f * (a0 + a1), half = 0.5f * (a1 - a0);
    float sum = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float phi = mid + (q < 4 ? -X[q] : X[q - 4]) * half;
        const float w = W[q < 4 ? q : q - 4];
        float sp, cp;
        __sincosf(phi, &sp, &cp);
        float tlo = 0.0f, thi = __builtin_inff();
        if (fabsf(cp) > 1e-30f) {
            const float t1 = (xa - x0) / cp, t2 = (xb - x0) / cp;
            tlo = fmaxf(tlo, fminf(t1, t2));
            thi = fminf(thi, fmaxf(t1, t2));
        } else if (x0 < xa || x0 > xb) {
            continue;
        }
        if (fabsf(sp) > 1e-30f) {
            const float t1 = (ya - y0) / sp, t2 = (yb - y0) / sp;
            tlo = fmaxf(tlo, fminf(t1, t2));
            thi = fminf(thi, fmaxf(t1, t2));
        } else if (y0 < ya || y0 > yb) {
            continue;
        }
        if (!(thi > tlo)) continue;
        const float u_in = d * rsqrtf(tlo * tlo + d * d);
        const float u_out = thi < 1e30f ? d * rsqrtf(thi * thi + d * d) : 0.0f;
        sum += w * (direct_H(t, u_out, k) - direct_H(t, u_in, k));
    }
    return sum * half;
}

// Angles (relative to phic, in (-pi, pi]) where the lines x = xa, xb and y = ya, yb cross
// the circle of radius R about (x0, y0): there the ray's entry or exit distance passes the
// critical-angle radius and the integrand loses smoothness.  Appends to ang[], returns the
// new count.
__device__ __forceinline__ int circle_crossings(float x0, float y0, float R, float xa, float xb,
                                                float ya, float yb, float phic, float *ang, int na)
{
    if (!(R < 1e30f)) return na;
    const float lines[4] = {xa - x0, xb - x0, ya - y0, yb - y0};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const float c = lines[m];
        const float h2 = R * R - c * c;
        if (!(h2 > 0.0f)) continue;
        const float h = sqrtf(h2);
#pragma unroll
        for (int sg = -1; sg <= 1; sg += 2) {
            float v = m < 2 ? atan2f(sg * h, c) : atan2f(c, sg * h);
            v -= phic;
            v = v > 3.14159265f ? v - 6.2831853f : (v < -3.14159265f ? v + 6.2831853f : v);
            ang[na++] = v;
        }
    }
    return na;
}

__device__ __forceinline__ void sort_small(float *a, int n)
{
    for (int i = 1; i < n; ++i)
        for (int b = i; b > 0 && a[b - 1] > a[b]; --b) { const float v = a[b]; a[b] = a[b - 1]; a[b - 1] = v; }
}

// D_c of one cell (column-major cell index c = i + nx j) for a source at (x0, y0, z).  The
// azimuth range splits at the cell's corner directions (the ray's entry and exit edges
// change there) and where the cell's edges cross the critical-angle circle.
__device__ float direct_cell(const DirectTable &t, float x0, float y0, float z, float Lz, int c)
{
    const float d = fmaxf(Lz - z, 1e-5f);
    const int i = c % t.nx, j = c / t.nx;
    const float xa = fminf(i * t.pitch, t.Lx), xb = i == t.nx - 1 ? t.Lx : fminf((i + 1) * t.pitch, t.Lx);
    const float ya = fminf(j * t.pitch, t.Ly), yb = j == t.ny - 1 ? t.Ly : fminf((j + 1) * t.pitch, t.Ly);
    if (!(xb > xa) || !(yb > ya)) return 0.0f;
    // nothing beyond the critical-angle radius d tan(theta_c) lands
    const float ex = fmaxf(fmaxf(xa - x0, x0 - xb), 0.0f), ey = fmaxf(fmaxf(ya - y0, y0 - yb), 0.0f);
    const float R = d * t.tan_c;
    if (ex * ex + ey * ey >= R * R) return 0.0f;
    const float k = d * t.inv_lambda;
    const float cx[4] = {xa, xb, xb, xa}, cy[4] = {ya, ya, yb, yb};
    float ang[12];
    float sum = 0.0f;
    const bool inside = x0 >= xa && x0 <= xb && y0 >= ya && y0 <= yb;
    // directions relative to the cell centre's (the whole circle from inside)
    const float phic = inside ? 0.0f : atan2f(0.5f * (ya + yb) - y0, 0.5f * (xa + xb) - x0);
    float lo = -3.14159265f, hi = 3.14159265f;
    int na = 0;
    for (int m = 0; m < 4; ++m) {
        float v = atan2f(cy[m] - y0, cx[m] - x0) - phic;
        v = v > 3.14159265f ? v - 6.2831853f : (v < -3.14159265f ? v + 6.2831853f : v);
        ang[na++] = v;
    }
    if (!inside) {
        // outside: the cell spans less than pi seen from the foot point
        lo = fminf(fminf(ang[0], ang[1]), fminf(ang[2], ang[3]));
        hi = fmaxf(fmaxf(ang[0], ang[1]), fmaxf(ang[2], ang[3]));
    }
    na = circle_crossings(x0, y0, R, xa, xb, ya, yb, phic, ang, na);
    sort_small(ang, na);
    if (inside) {
        // cyclic: consecutive break angles, the last one round to the first
        for (int m = 0; m < na; ++m) {
            const float a0 = ang[m], a1 = m + 1 < na ? ang[m + 1] : ang[0] + 6.2831853f;
            if (a1 > a0) sum += direct_piece(t, x0, y0, d, k, xa, xb, ya, yb, a0, a1);
        }
    } else {
        for (int m = 0; m + 1 < na; ++m) {
            const float a0 = fmaxf(ang[m], lo), a1 = fminf(ang[m + 1], hi);
            if (a1 > a0) sum += direct_piece(t, x0, y0, d, k, xa, xb, ya, yb, phic + a0, phic + a1);
        }
    }
    return t.pde * 0.0795774715f * sum;             // PDE / (4 pi)
}

// Weight of the computed terms: 1 within IMG_FULL_MM of the readout face (and, for a mirror
// image, of its walls), 0 beyond IMG_ZERO_MM, a cubic smoothstep between.  The terms are
// sharp only there; farther out they are smooth and left to the network, and the taper keeps
// what the network learns smooth across the change.
constexpr float IMG_FULL_MM = 2.0f, IMG_ZERO_MM = 3.0f;
__device__ __forceinline__ float img_taper(float a)
{
    const float u = fminf(fmaxf((IMG_ZERO_MM - a) / (IMG_ZERO_MM - IMG_FULL_MM), 0.0f), 1.0f);
    return u * u * (3.0f - 2.0f * u);
}

// The computed part of cell c's rate: the direct light plus, with t.images, the light that
// reaches the readout face after a specular reflection off one side wall or two (total
// internal reflection at the air gap: every ray in the readout cone meets a side wall beyond
// the crystal-air critical angle, up to a sliver near it for n below about 1.8), i.e. the
// direct light of the mirror sources (-x0, y0), (2Lx - x0, y0), (x0, -y0), (x0, 2Ly - y0) and
// of the four double mirrors, each image weighted by the taper of its walls' distances.
// Everything is weighted by the taper of the depth: the terms are sharp only near the
// readout face (the direct light lands within d tan(theta_c) of the foot point); deeper they
// are smooth and the network learns them with the rest.
__device__ float direct_total(const DirectTable &t, float x0, float y0, float z, float Lz, int c)
{
    const float wd = img_taper(Lz - z);
    if (wd <= 0.0f) return 0.0f;
    float D = direct_cell(t, x0, y0, z, Lz, c);
    if (t.images) {
        const float wx0 = img_taper(x0), wx1 = img_taper(t.Lx - x0);
        const float wy0 = img_taper(y0), wy1 = img_taper(t.Ly - y0);
        const float xs[3] = {x0, -x0, 2.0f * t.Lx - x0}, wxs[3] = {1.0f, wx0, wx1};
        const float ys[3] = {y0, -y0, 2.0f * t.Ly - y0}, wys[3] = {1.0f, wy0, wy1};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                if (a == 0 && b == 0) continue;
                const float w = wxs[a] * wys[b];
                if (w > 0.0f) D += w * direct_cell(t, xs[a], ys[b], z, Lz, c);
            }
    }
    return wd * D;
}

// D[i][c] for n points (x, y, z) = xyz[3i..3i+2]; one thread per (point, cell)
__global__ void direct_light_kernel(const DirectTable t, const float *xyz, int64_t n, float Lz, float *D)
{
    const int ncell = t.nx * t.ny;
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n * ncell) return;
    const int64_t i = g / ncell;
    const int c = (int)(g % ncell);
    D[g] = direct_total(t, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], Lz, c);
}

// ------------------------------------------------------ events from the network ----
// Network inputs of deposits (x, y, depth feature, E): positions folded into the
// fundamental triangle 0 <= x <= y <= L/2 of a square face when `fold` (octant code: bit 0
// x mirrored, bit 1 y mirrored, bit 2 swapped), depth -log(Lz - z + log_depth_mm) when
// log_depth_mm > 0; also the folded positions (for the direct light).
__global__ void surrogate_features(const float4 *dep, const float *e_kev, int64_t n, float Lx,
                                   float Ly, float Lz, int fold, float log_depth_mm, float *X,
                                   float *xyz_f, uint8_t *code)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = dep[i];
    float x = q.x, y = q.y;
    uint8_t o = 0;
    if (fold) {
        const float h = 0.5f * Lx;
        const bool mx = x > h, my = y > h;
        if (mx) x = Lx - x;
        if (my) y = Ly - y;
        const bool sw = x > y;
        if (sw) { const float v = x; x = y; y = v; }
        o = (uint8_t)((mx ? 1 : 0) | (my ? 2 : 0) | (sw ? 4 : 0));
    }
    X[4 * i] = x;
    X[4 * i + 1] = y;
    X[4 * i + 2] = log_depth_mm > 0.0f ? -logf(Lz - q.z + log_depth_mm) : q.z;
    X[4 * i + 3] = e_kev[i];
    xyz_f[3 * i] = x;
    xyz_f[3 * i + 1] = y;
    xyz_f[3 * i + 2] = q.z;
    code[i] = o;
}

// Event maps from per-deposit network outputs: deposit d's per-photon rate of actual cell c
// is p = e^{Y[d][c']} + D[d][c'] (c' = inv[code_d][c], the folded index of c), its mean
// arrival time Y[d][ncell + c'] * t_scale + t_shift after the deposit time;
// lambda_c = sum_d yield * E_d / 1000 * p over the event's deposits with use[d] != 0, and
// maps[e][c] ~ Poisson(lambda_c) (src/events.jl:11-22 algorithm).  One thread per
// (event, cell).
__global__ void sample_events_net(const float *Y, int ldy, const float *D, const uint8_t *code,
                                  const int32_t *inv, int ncell, const float4 *dep,
                                  const float *dep_e_kev, const uint8_t *use,
                                  const int64_t *ev_first, int64_t nev, double yield_per_mev,
                                  float t_shift, float t_scale, uint64_t seed, int32_t *maps,
                                  float *tmean, float *lam_out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nev * ncell) return;
    const int64_t e = i / ncell;
    const int c = (int)(i % ncell);
    double lam = 0.0, tw = 0.0;
    for (int64_t d = ev_first[e]; d < ev_first[e + 1]; ++d) {
        if (!use[d]) continue;
        const int cf = code ? inv[code[d] * ncell + c] : c;
        const float *y = Y + (size_t)d * ldy;
        const float p = __expf(y[cf]) + (D ? D[(size_t)d * ncell + cf] : 0.0f);
        const double mu = yield_per_mev * (double)dep_e_kev[d] / 1000.0 * (double)p;
        lam += mu;
        tw += mu * ((double)(y[ncell + cf] * t_scale + t_shift) + (double)dep[d].w);
    }
    PoisRng r;
    r.k0 = (uint32_t)seed; r.k1 = (uint32_t)(seed >> 32);
    r.a = (uint32_t)i; r.b = (uint32_t)((uint64_t)i >> 32);
    maps[i] = poisson_draw(r, lam);
    if (tmean) tmean[i] = lam > 0.0 ? (float)(tw / lam) : INFINITY;
    if (lam_out) lam_out[i] = (float)lam;
}

// Per-deposit maps of the network in actual cell order: p[d][c] and t[d][c] (ns after the
// deposit).
__global__ void surrogate_maps(const float *Y, int ldy, const float *D, const uint8_t *code,
                               const int32_t *inv, int ncell, int64_t n, float t_shift,
                               float t_scale, float *p, float *t)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n * ncell) return;
    const int64_t d = i / ncell;
    const int c = (int)(i % ncell);
    const int cf = code ? inv[code[d] * ncell + c] : c;
    const float *y = Y + (size_t)d * ldy;
    p[i] = __expf(y[cf]) + (D ? D[(size_t)d * ncell + cf] : 0.0f);
    if (t) t[i] = y[ncell + cf] * t_scale + t_shift;
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
    int n_cu = 256;                                 // compute units of the device
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
    float *bpart = nullptr;                         // the slices' bias row sums [ks][out], all layers
    std::vector<size_t> bpart_off;                  // per layer, floats
    float *lpart = nullptr;                         // loss partials [n_wg]
    float *loss = nullptr;
    int64_t *ctr = nullptr;                         // device: batch index, Adam steps (graph replays)
    float *dlr = nullptr;                           // device: learning rate (graph replays)
    float *dloss = nullptr;                         // device: per-step losses of a call
    int64_t dloss_cap = 0;
    // the captured training step, reused while its arguments stay the same
    hipGraphExec_t gexec = nullptr;
    const void *gX = nullptr, *gT = nullptr, *gW = nullptr, *gR0 = nullptr;
    int gbatch = 0, gsteps = 0;                     // batch size, steps captured per graph
    bool glosses = false;
    float gb1 = 0, gb2 = 0, geps = 0, gwd = 0, gmrw = 0;
    int64_t hctr[2] = {0, 0};
    double flops_fwd_per_row = 0.0;
};

// Batch slices of the weight-gradient products: DW_KS, or fewer where Mpad / DW_KP has
// fewer power-of-two factors, so that every slice is a whole number of DW_KP stages.
static int dw_slices(int Mpad)
{
    int ks = 1;
    while (ks * 2 <= DW_KS && (Mpad / DW_KP) % (ks * 2) == 0) ks *= 2;
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
    const size_t groups = (size_t)(cap + 7) / 8;     // chunked transposed operands
    SCHK(hipMalloc(&n->X, groups * chunk_stride(n->kin_pad) * 2));
    SCHK(hipMalloc(&n->Y, (size_t)cap * n->n_out * 4));
    SCHK(hipMalloc(&n->dY, groups * chunk_stride(n->n_out) * 2));
    size_t part = 0, bpart = 0;
    const int n_wg = (cap + RB_MIN - 1) / RB_MIN;
    n->part_off.assign(n->n_layers, 0);
    n->bpart_off.assign(n->n_layers, 0);
    for (int l = 0; l < n->n_layers; ++l) {
        SCHK(hipMalloc(&n->Z[l], (size_t)cap * n->dout[l] * 4));
        SCHK(hipMalloc(&n->H[l], groups * chunk_stride(n->dout[l]) * 2));
        SCHK(hipMalloc(&n->dZ[l], groups * chunk_stride(n->dout[l]) * 2));
        // the slices' partial weight gradients and bias row sums (at most DW_KS slices)
        n->part_off[l] = part;
        part += (size_t)DW_KS * n->dout[l] * n->dpad[l];
        n->bpart_off[l] = bpart;
        bpart += (size_t)DW_KS * n->dout[l];
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


// mlp_rows for L layers with NB 32-row blocks per wave (64 NB rows per workgroup)
template <bool TRAIN, int NB, bool SL>
static void launch_rows_nb(int L, int n_wg, hipStream_t s, const RowArgs &ra)
{
    const dim3 grid((unsigned)n_wg), block(RT);
    switch (L) {                                    // the layer count is a template argument, so
    case 2: hipLaunchKernelGGL((mlp_rows<2, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    case 3: hipLaunchKernelGGL((mlp_rows<3, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    case 4: hipLaunchKernelGGL((mlp_rows<4, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    case 5: hipLaunchKernelGGL((mlp_rows<5, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    case 6: hipLaunchKernelGGL((mlp_rows<6, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    case 7: hipLaunchKernelGGL((mlp_rows<7, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    case 8: hipLaunchKernelGGL((mlp_rows<8, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    default: hipLaunchKernelGGL((mlp_rows<9, TRAIN, NB, SL>), grid, block, 0, s, ra); break;
    }                                               // the layer loops unroll (gelu' in registers)
}

// Row blocks per wave for a batch of `rows`: two (128-row workgroups) once they give every
// compute unit at least two workgroups, else one (64-row workgroups, twice as many).
static int rows_nb(const crysp_mlp *n, int64_t rows)
{
    return rows >= (int64_t)2 * n->n_cu * RB_MAX ? 2 : 1;
}

// mlp_rows over Mpad rows (a multiple of 64 NB)
// (128-row workgroups spread the last layer one block per wave when it has at most 4
// column blocks)
template <bool TRAIN>
static void launch_rows(int L, int nb, int Mpad, hipStream_t s, const RowArgs &ra)
{
    if (nb == 2 && ra.n_out <= 128) launch_rows_nb<TRAIN, 2, true>(L, Mpad / RB_MAX, s, ra);
    else if (nb == 2) launch_rows_nb<TRAIN, 2, false>(L, Mpad / RB_MAX, s, ra);
    else launch_rows_nb<TRAIN, 1, false>(L, Mpad / RB_MIN, s, ra);
}

// Inference on M rows of device inputs into device outputs Y [M][n_out]: one launch of the
// row-block kernel's forward pass (activations stay in LDS, weights stream through it).
static int mlp_infer_dev(crysp_mlp *n, const float *Xdev, int64_t M, float *Ydev)
{
    const int L = n->n_layers;
    if (L > MAXL) return sfail(CRYSP_EINVAL, "too many layers");
    if (M > ((int64_t)1 << 31) - RB_MAX) return sfail(CRYSP_ERANGE, "more than 2^31 rows");
    const int nb = rows_nb(n, M), rb = 64 * nb;
    RowArgs ra{};
    ra.L = L; ra.M = (int)M; ra.Mpad = (int)(((M + rb - 1) / rb) * rb); ra.n_in = n->n_in;
    ra.kin_pad = n->kin_pad; ra.n_out = n->n_out; ra.n_pois = n->n_pois;
    for (int l = 0; l < L; ++l) {
        ra.din[l] = n->din[l]; ra.dpad[l] = n->dpad[l]; ra.dout[l] = n->dout[l];
        ra.Wt[l] = n->Wt[l]; ra.Wn[l] = n->Wn[l]; ra.bias[l] = n->p + n->boff[l];
    }
    ra.Xraw = Xdev; ra.shift = n->shift; ra.scale = n->scale;
    ra.Y = Ydev;
    launch_rows<false>(L, nb, ra.Mpad, n->stream, ra);
    SCHK(hipGetLastError());
    return CRYSP_OK;
}

// The gradient of the weighted loss on Mpad rows (M real) into n->g (fp32 flat) and the loss
// into loss_out[0], or loss_out[ctr[0]] with loss_by_batch.  With ctr (graph replays),
// ctr[0] is the batch index whose rows are read, and ctr advances.  Three launches:
// mlp_rows, dw_grouped (dW and the bias row sums of every layer), reduce_grads.
static int mlp_grad_dev(crysp_mlp *n, const float *Xraw, const float *T, const float *W,
                        const float *R0, int M, int Mpad, double wsum, float *loss_out,
                        bool loss_by_batch, int64_t *ctr)
{
    const int L = n->n_layers;
    if (L > MAXL || Mpad % RB_MAX || Mpad > n->cap) return sfail(CRYSP_EINVAL, "bad training shape");
    const int nb = rows_nb(n, Mpad);
    const int n_wg = Mpad / (64 * nb);
    RowArgs ra{};
    ra.L = L; ra.M = M; ra.Mpad = Mpad; ra.n_in = n->n_in; ra.kin_pad = n->kin_pad; ra.n_out = n->n_out;
    ra.n_pois = n->n_pois;
    for (int l = 0; l < L; ++l) {
        ra.din[l] = n->din[l]; ra.dpad[l] = n->dpad[l]; ra.dout[l] = n->dout[l];
        ra.Wt[l] = n->Wt[l]; ra.Wn[l] = n->Wn[l]; ra.bias[l] = n->p + n->boff[l];
        ra.HT[l] = n->H[l];
        ra.dZT[l] = l + 1 == L ? n->dY : n->dZ[l];
    }
    ra.XT = n->X; ra.lpart = n->lpart;
    ra.Xraw = Xraw; ra.T = T; ra.W = W; ra.R0 = R0; ra.shift = n->shift; ra.scale = n->scale;
    ra.inv_wsum = (float)(1.0 / wsum);
    ra.bidx = ctr;
    launch_rows<true>(L, nb, Mpad, n->stream, ra);

    // dW_l[out][in] = dZ_l^T [out x Mpad] * Hin [Mpad x in_pad], both operands stored
    // chunked transposed (8 batch rows contiguous), every layer in one launch, as the partial
    // products of ks batch slices (with the bias row sums)
    const int ks = dw_slices(Mpad);
    DwArgs da{};
    da.n = L; da.kslice = Mpad / ks;
    RedArgs rd{};
    rd.L = L; rd.n_wg = n_wg;
    int tiles = 0;
    int64_t cum = 0;
    for (int l = 0; l < L; ++l) {
        const int kin = n->dpad[l], nout = n->dout[l];
        da.A[l] = ra.dZT[l];
        da.B[l] = l == 0 ? ra.XT : ra.HT[l - 1];
        da.M[l] = nout; da.N[l] = kin;
        da.C[l] = n->part + n->part_off[l];
        da.db[l] = n->bpart + n->bpart_off[l];
        da.t0[l] = tiles;
        tiles += (nout / (nout % 128 == 0 ? 128 : 64)) * (kin / (kin % 128 == 0 ? 128 : 64));
        rd.part[l] = da.C[l]; rd.ks[l] = ks; rd.db[l] = da.db[l];
        rd.din[l] = n->din[l]; rd.dpad[l] = kin; rd.dout[l] = nout;
        rd.woff[l] = (int64_t)n->woff[l]; rd.boff[l] = (int64_t)n->boff[l];
        rd.cum[l] = cum;
        cum += (int64_t)nout * n->din[l] + nout;
        rd.bcum[l + 1] = rd.bcum[l] + nout;
    }
    da.t0[L] = tiles;
    rd.cum[L] = cum;
    hipLaunchKernelGGL(dw_grouped, dim3(tiles, ks), dim3(256), 0, n->stream, da);

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

// A copy to or from the network's buffers on its own stream, finished on return.  The
// stream is non-blocking, so a plain hipMemcpy (the null stream) is not ordered after the
// kernels and copies queued on it; every host access to the network's memory goes here.
static int mlp_copy(crysp_mlp *n, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    SCHK(hipMemcpyAsync(dst, src, bytes, kind, n->stream));
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
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
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        n->n_cu = cus;
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
    if (!ok) { sfail(CRYSP_ENOMEM, "MLP allocation failed"); crysp_mlp_destroy(n); return nullptr; }
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
    // all on the network's stream (mlp_refresh then writes every element of Wt and Wn)
    const hipStream_t s = n->stream;
    if (hipMemcpyAsync(n->p, h.data(), n->n_params * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(n->m, 0, n->n_params * 4, s) != hipSuccess ||
        hipMemsetAsync(n->v, 0, n->n_params * 4, s) != hipSuccess ||
        hipMemsetAsync(n->g, 0, n->n_params * 4, s) != hipSuccess ||
        hipMemcpyAsync(n->shift, sh.data(), 128, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(n->scale, sc.data(), 128, hipMemcpyHostToDevice, s) != hipSuccess) {
        (void)hipStreamSynchronize(s);
        sfail(CRYSP_EHIP, "MLP initialization copies failed");
        crysp_mlp_destroy(n);
        return nullptr;
    }
    mlp_refresh(n);
    if (hipStreamSynchronize(s) != hipSuccess) {
        sfail(CRYSP_EHIP, "MLP initialization failed");
        crysp_mlp_destroy(n);
        return nullptr;
    }
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
    SCHK(hipSetDevice(n->device));
    return mlp_copy(n, out, n->p, n->n_params * 4, hipMemcpyDeviceToHost);
}

int crysp_mlp_set_params(crysp_mlp *n, const float *in)
{
    if (!n || !in) return sfail(CRYSP_EINVAL, "null argument");
    SCHK(hipSetDevice(n->device));
    SCHK(hipMemcpyAsync(n->p, in, n->n_params * 4, hipMemcpyHostToDevice, n->stream));
    mlp_refresh(n);
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
}

// input normalization x' = (x - shift) * scale applied inside the network
int crysp_mlp_set_input_norm(crysp_mlp *n, const float *shift, const float *scale)
{
    if (!n || !shift || !scale) return sfail(CRYSP_EINVAL, "null argument");
    SCHK(hipSetDevice(n->device));
    int rc = mlp_copy(n, n->shift, shift, n->n_in * 4, hipMemcpyHostToDevice);
    return rc ? rc : mlp_copy(n, n->scale, scale, n->n_in * 4, hipMemcpyHostToDevice);
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
    Staged sx;
    int rc = stage(n->stream, X, (size_t)M * n->n_in * 4, sx);
    if (rc) return rc;
    if (is_dev(Y)) {
        // device outputs: the fused forward kernel writes them in place
        if ((rc = mlp_infer_dev(n, (const float *)sx.d, M, Y))) return rc;
        SCHK(hipStreamSynchronize(n->stream));
        return CRYSP_OK;
    }
    const int64_t chunk = 65536;
    if ((rc = mlp_alloc_batch(n, (int)std::min<int64_t>(chunk, ((M + 63) / 64) * 64)))) return rc;
    for (int64_t r0 = 0; r0 < M; r0 += chunk) {
        const int m = (int)std::min<int64_t>(chunk, M - r0);
        if ((rc = mlp_infer_dev(n, (const float *)sx.d + r0 * n->n_in, m, n->Y))) return rc;
        SCHK(hipMemcpyAsync(Y + r0 * n->n_out, n->Y, (size_t)m * n->n_out * 4,
                            hipMemcpyDeviceToHost, n->stream));
    }
    SCHK(hipStreamSynchronize(n->stream));
    return CRYSP_OK;
}

// gradient of the weighted MSE on one batch (no update): grads flat fp32, loss scalar
int crysp_mlp_gradient(crysp_mlp *n, const float *X, const float *T, const float *W,
                       const float *R0, int32_t M, float *grads, float *loss)
{
    if (!n || !X || !T || M < 1) return sfail(CRYSP_EINVAL, "bad arguments");
    if (R0 && n->n_pois < 1) return sfail(CRYSP_EINVAL, "rate offsets need Poisson outputs");
    SCHK(hipSetDevice(n->device));
    const int mpad = ((M + RB_MAX - 1) / RB_MAX) * RB_MAX;     // whole row-kernel workgroups
    int rc = mlp_alloc_batch(n, mpad);
    if (rc) return rc;
    Staged sx, st, sw, sr;
    if ((rc = stage(n->stream, X, (size_t)M * n->n_in * 4, sx)) ||
        (rc = stage(n->stream, T, (size_t)M * n->n_out * 4, st)) ||
        (rc = stage(n->stream, W, (size_t)M * n->n_out * 4, sw)) ||
        (rc = stage(n->stream, R0, (size_t)M * n->n_pois * 4, sr)))
        return rc;
    double wsum = (double)M * n->n_out;
    if (W) {
        // host weights are summed where they are; device ones are read back on the
        // network's stream, after the staging copies queued there
        std::vector<float> hw;
        const float *w = W;
        if (!sw.own) {
            hw.resize((size_t)M * n->n_out);
            if ((rc = mlp_copy(n, hw.data(), sw.d, hw.size() * 4, hipMemcpyDeviceToHost))) return rc;
            w = hw.data();
        }
        wsum = 0.0;
        for (size_t i = 0; i < (size_t)M * n->n_out; ++i) wsum += w[i];
        if (!(wsum > 0.0)) return sfail(CRYSP_EINVAL, "loss weights sum to zero");
    }
    if ((rc = mlp_grad_dev(n, (const float *)sx.d, (const float *)st.d, (const float *)sw.d,
                           (const float *)sr.d, M, mpad, wsum, n->loss, false, nullptr)))
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
                            const float *R0, int64_t n_batches, int32_t batch, float lr,
                            float beta1, float beta2, float eps, float weight_decay,
                            float mean_row_weight, float *losses)
{
    if (!n || !X || !T || n_batches < 0 || batch < 64 || batch % 64)
        return sfail(CRYSP_EINVAL, "bad arguments (batch must be a multiple of 64)");
    if (!is_dev(X) || !is_dev(T) || (W && !is_dev(W)) || (R0 && !is_dev(R0)))
        return sfail(CRYSP_EINVAL, "training arrays must be device-resident");
    if (R0 && n->n_pois < 1) return sfail(CRYSP_EINVAL, "rate offsets need Poisson outputs");
    SCHK(hipSetDevice(n->device));
    const int mpad = ((batch + RB_MAX - 1) / RB_MAX) * RB_MAX; // whole row-kernel workgroups
    int rc = mlp_alloc_batch(n, mpad);
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
    auto enqueue_step = [&](const float *x, const float *t, const float *w, const float *r0,
                            float *lout, int64_t *ctr, const float *dlr, float c1, float c2) -> int {
        int r;
        if ((r = mlp_grad_dev(n, x, t, w, r0, batch, mpad, wsum, lout ? lout : n->loss,
                              lout != nullptr, ctr)))
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
        const bool same = n->gexec && n->gX == X && n->gT == T && n->gW == W && n->gR0 == R0 &&
                          n->gbatch == batch &&
                          n->gsteps == steps &&
                          n->glosses == (dl != nullptr) && n->gb1 == beta1 && n->gb2 == beta2 &&
                          n->geps == eps && n->gwd == weight_decay && n->gmrw == mean_row_weight;
        if (!same) {
            mlp_drop_graph(n);
            SCHK(hipStreamSynchronize(n->stream));
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                for (int k = 0; k < steps && rc == CRYSP_OK; ++k)
                    rc = enqueue_step(X, T, W, R0, dl, n->ctr, n->dlr, 0.0f, 0.0f);
                const hipError_t e = hipStreamEndCapture(n->stream, &g);
                if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
                if (e != hipSuccess || !g || hipGraphInstantiate(&n->gexec, g, nullptr, nullptr, 0) != hipSuccess)
                    n->gexec = nullptr;
                if (g) (void)hipGraphDestroy(g);
            }
            if (!n->gexec) (void)hipGetLastError();     // capture unavailable: launch eagerly
            n->gX = X; n->gT = T; n->gW = W; n->gR0 = R0; n->gbatch = batch; n->gsteps = steps; n->glosses = dl != nullptr;
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
            const float *r0 = R0 ? R0 + (size_t)b * batch * n->n_pois : nullptr;
            n->step += 1;
            const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)n->step));
            const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)n->step));
            if ((rc = enqueue_step(x, t, w, r0, dl ? dl + b : nullptr, nullptr, nullptr, c1, c2))) return rc;
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

// =========================================================== surrogate object ====

// Fresnel transmission crystal -> coupling at incidence cosine v (src/optics.jl:10-19 in
// binary64): zero at and beyond total internal reflection
static double fresnel_T64(double n1, double n2, double v)
{
    const double q = n1 / n2;
    const double s2 = q * q * (1.0 - v * v);
    if (s2 >= 1.0) return 0.0;
    const double ct = std::sqrt(1.0 - s2);
    const double rs = (n1 * v - n2 * ct) / (n1 * v + n2 * ct);
    const double rp = (n1 * ct - n2 * v) / (n1 * ct + n2 * v);
    return 1.0 - 0.5 * (rs * rs + rp * rp);
}

struct crysp_surrogate {
    crysp_mlp *net = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    crysp_box box{};
    crysp_surrogate_config cfg{};
    int ncell = 0;
    DirectTable tab{};
    float *dH = nullptr;                            // the H table on the device
    int32_t *dinv = nullptr;                        // [8][ncell] folded index of actual cell c
    // per-deposit buffers, grown on demand
    int64_t cap = 0;
    float *X = nullptr, *xyzf = nullptr, *D = nullptr, *Y = nullptr;
    uint8_t *code = nullptr;
};

static void sur_free_bufs(crysp_surrogate *s)
{
    auto fre = [](void *q) { if (q) (void)hipFree(q); };
    fre(s->X); fre(s->xyzf); fre(s->D); fre(s->Y); fre(s->code);
    s->X = s->xyzf = s->D = s->Y = nullptr;
    s->code = nullptr;
    s->cap = 0;
}

static int sur_reserve(crysp_surrogate *s, int64_t n)
{
    if (n <= s->cap) return CRYSP_OK;
    sur_free_bufs(s);
    const int64_t want = std::max<int64_t>(n, 1024);
    if (hipMalloc(&s->X, (size_t)want * 16) != hipSuccess ||
        hipMalloc(&s->xyzf, (size_t)want * 12) != hipSuccess ||
        hipMalloc(&s->D, (size_t)want * s->ncell * 4) != hipSuccess ||
        hipMalloc(&s->Y, (size_t)want * s->net->n_out * 4) != hipSuccess ||
        hipMalloc(&s->code, (size_t)want) != hipSuccess) {
        (void)hipGetLastError();
        sur_free_bufs(s);
        return sfail(CRYSP_ENOMEM, "surrogate buffers: allocation failed");
    }
    s->cap = want;
    return CRYSP_OK;
}

// features, direct light and the network for n deposits (device), into s->X, s->D, s->Y
static int sur_evaluate(crysp_surrogate *s, const float4 *dep, const float *e_kev, int64_t n)
{
    int rc = sur_reserve(s, n);
    if (rc) return rc;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(surrogate_features, dim3(nb), dim3(256), 0, s->stream, dep, e_kev, n,
                       s->box.L[0], s->box.L[1], s->box.L[2], (int)s->cfg.fold, s->cfg.log_depth_mm,
                       s->X, s->xyzf, s->code);
    if (s->cfg.direct) {
        const int64_t tot = n * s->ncell;
        hipLaunchKernelGGL(direct_light_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                           s->stream, s->tab, (const float *)s->xyzf, n, s->box.L[2], s->D);
    }
    SCHK(hipGetLastError());
    // the network on the library's own MLP stream: order the two streams through a sync
    SCHK(hipStreamSynchronize(s->stream));
    return crysp_mlp_forward(s->net, s->X, n, s->Y);
}

extern "C" {

crysp_surrogate *crysp_surrogate_create(crysp_mlp *net, const crysp_box *box,
                                        const crysp_optical_params *op, const crysp_readout *ro,
                                        const crysp_surrogate_config *cfg)
{
    if (!net || !box || !op || !ro || !cfg) { sfail(CRYSP_EINVAL, "null argument"); return nullptr; }
    const int ncell = ro->grid.nx * ro->grid.ny;
    if (net->n_in != 4 || net->n_out < 2 * ncell) {
        sfail(CRYSP_EINVAL, "the network needs 4 inputs and 2 * nx * ny outputs");
        return nullptr;
    }
    if (cfg->fold && (ro->grid.nx != ro->grid.ny || box->L[0] != box->L[1])) {
        sfail(CRYSP_EINVAL, "folding needs a square face and a square grid");
        return nullptr;
    }
    if (cfg->direct && (ro->disc || ro->two_sided)) {
        sfail(CRYSP_EINVAL, "the direct-light term covers a full single readout face");
        return nullptr;
    }
    if (hipSetDevice(net->device) != hipSuccess) { sfail(CRYSP_EHIP, "hipSetDevice"); return nullptr; }
    auto *s = new crysp_surrogate();
    s->net = net;
    s->device = net->device;
    s->box = *box;
    s->cfg = *cfg;
    s->ncell = ncell;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        sfail(CRYSP_EHIP, "stream creation failed");
        return nullptr;
    }
    // folded index of every actual cell for the 8 octant codes (column-major cells
    // i + n j; bit 0 mirrors i, bit 1 mirrors j, bit 2 swaps them)
    std::vector<int32_t> inv((size_t)8 * ncell, 0);
    const int nn = ro->grid.nx;
    for (int o = 0; o < 8; ++o)
        for (int c = 0; c < ncell; ++c) {
            if (!cfg->fold) { inv[o * ncell + c] = c; continue; }
            const int a1 = c % nn, b1 = c / nn;     // folded cell (a1, b1) -> actual (i, j)
            int a = a1, b = b1;
            if (o & 4) { a = b1; b = a1; }
            const int i = (o & 1) ? nn - 1 - a : a, j = (o & 2) ? nn - 1 - b : b;
            inv[o * ncell + (i + nn * j)] = c;
        }
    // H(u, k) = int_u^1 T(v) e^{-k/v} dv on u = uc + (1 - uc) s^2 (which takes the square-root
    // onset of T at the critical angle), cumulative Simpson from s = 1 down
    const double n1 = op->n_crystal, n2 = op->n_coupling;
    const double uc = n1 > n2 ? std::sqrt(1.0 - (n2 / n1) * (n2 / n1)) : 0.0;
    double inv_lambda = 0.0;
    if (std::isfinite(op->abs_len_mm)) inv_lambda += 1.0 / op->abs_len_mm;
    if (std::isfinite(op->rayleigh_mm)) inv_lambda += 1.0 / op->rayleigh_mm;
    const double kmax = box->L[2] * inv_lambda;
    const int nu = 1025, nk = kmax > 0.0 ? 33 : 1, sub = 16;
    std::vector<float> H((size_t)nu * nk);
    for (int m = 0; m < nk; ++m) {
        const double k = nk == 1 ? 0.0 : kmax * m / (nk - 1);
        auto f = [&](double sv) {                   // integrand in s: T(v) e^{-k/v} dv/ds
            const double v = uc + (1.0 - uc) * sv * sv;
            if (v <= 0.0) return 0.0;
            return fresnel_T64(n1, n2, v) * std::exp(-k / v) * 2.0 * (1.0 - uc) * sv;
        };
        double acc = 0.0;
        H[(size_t)m * nu + nu - 1] = 0.0f;
        for (int j = nu - 1; j > 0; --j) {
            const double s1 = (double)j / (nu - 1), s0 = (double)(j - 1) / (nu - 1);
            const double h = (s1 - s0) / sub;
            double part = 0.0;
            for (int q = 0; q < sub; ++q) {
                const double a = s0 + q * h;
                part += h / 6.0 * (f(a) + 4.0 * f(a + 0.5 * h) + f(a + h));
            }
            acc += part;
            H[(size_t)m * nu + j - 1] = (float)acc;
        }
    }
    s->tab.nu = nu; s->tab.nk = nk; s->tab.uc = (float)uc; s->tab.kmax = (float)kmax;
    s->tab.inv_lambda = (float)inv_lambda;
    s->tab.tan_c = uc > 0.0 ? (float)(std::sqrt(1.0 - uc * uc) / uc) : INFINITY;
    if (n1 <= n2) s->tab.tan_c = INFINITY;
    s->tab.pde = op->pde;
    s->tab.Lx = box->L[0]; s->tab.Ly = box->L[1]; s->tab.pitch = ro->grid.pitch_mm;
    s->tab.nx = ro->grid.nx; s->tab.ny = ro->grid.ny;
    s->tab.images = cfg->images && op->air_gap;
    if (hipMalloc(&s->dH, H.size() * 4) != hipSuccess ||
        hipMalloc(&s->dinv, inv.size() * 4) != hipSuccess ||
        hipMemcpyAsync(s->dH, H.data(), H.size() * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->dinv, inv.data(), inv.size() * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
        (void)hipGetLastError();
        crysp_surrogate_destroy(s);
        sfail(CRYSP_ENOMEM, "surrogate tables: allocation failed");
        return nullptr;
    }
    s->tab.H = s->dH;
    return s;
}

void crysp_surrogate_destroy(crysp_surrogate *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    sur_free_bufs(s);
    if (s->dH) (void)hipFree(s->dH);
    if (s->dinv) (void)hipFree(s->dinv);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int crysp_surrogate_direct_light(crysp_surrogate *s, const float *xyz, int64_t n, float *D)
{
    if (!s || !xyz || !D || n < 0) return sfail(CRYSP_EINVAL, "bad arguments");
    if (n == 0) return CRYSP_OK;
    SCHK(hipSetDevice(s->device));
    Staged sx;
    int rc = stage(s->stream, xyz, (size_t)n * 12, sx);
    if (rc) return rc;
    const bool ddev = is_dev(D);
    float *dD = D;
    if (!ddev) SCHK(hipMalloc(&dD, (size_t)n * s->ncell * 4));
    const int64_t tot = n * s->ncell;
    hipLaunchKernelGGL(direct_light_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s->stream,
                       s->tab, (const float *)sx.d, n, s->box.L[2], dD);
    SCHK(hipGetLastError());
    if (!ddev) SCHK(hipMemcpyAsync(D, dD, (size_t)tot * 4, hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    if (!ddev) (void)hipFree(dD);
    return CRYSP_OK;
}

int crysp_surrogate_predict(crysp_surrogate *s, const float *dep_xyzt, const float *e_kev, int64_t n,
                            float *p, float *t)
{
    if (!s || !dep_xyzt || !e_kev || !p || n < 0) return sfail(CRYSP_EINVAL, "bad arguments");
    if (n == 0) return CRYSP_OK;
    SCHK(hipSetDevice(s->device));
    Staged sd, se;
    int rc;
    if ((rc = stage(s->stream, dep_xyzt, (size_t)n * 16, sd)) ||
        (rc = stage(s->stream, e_kev, (size_t)n * 4, se)))
        return rc;
    if ((rc = sur_evaluate(s, (const float4 *)sd.d, (const float *)se.d, n))) return rc;
    const size_t cells = (size_t)n * s->ncell;
    const bool pdev = is_dev(p), tdev = is_dev(t);
    float *dp = p, *dt = t;
    if (!pdev) SCHK(hipMalloc(&dp, cells * 4));
    if (t && !tdev) SCHK(hipMalloc(&dt, cells * 4));
    hipLaunchKernelGGL(surrogate_maps, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->stream,
                       (const float *)s->Y, s->net->n_out, s->cfg.direct ? (const float *)s->D : nullptr,
                       s->cfg.fold ? (const uint8_t *)s->code : nullptr, (const int32_t *)s->dinv,
                       s->ncell, n, s->cfg.t_shift, s->cfg.t_scale, dp, dt);
    SCHK(hipGetLastError());
    if (!pdev) SCHK(hipMemcpyAsync(p, dp, cells * 4, hipMemcpyDeviceToHost, s->stream));
    if (t && !tdev) SCHK(hipMemcpyAsync(t, dt, cells * 4, hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    if (!pdev) (void)hipFree(dp);
    if (t && !tdev) (void)hipFree(dt);
    return CRYSP_OK;
}

int crysp_surrogate_sample(crysp_surrogate *s, const float *dep_xyzt, const float *dep_e_kev,
                           const uint8_t *use, const int64_t *ev_first, int64_t ndep, int64_t nev,
                           double yield_per_mev, uint64_t seed, int32_t *maps, float *tmean,
                           float *lam)
{
    if (!s || !ev_first || !maps || nev < 0 || ndep < 0 || (ndep > 0 && (!dep_xyzt || !dep_e_kev || !use)))
        return sfail(CRYSP_EINVAL, "bad arguments");
    if (nev == 0) return CRYSP_OK;
    SCHK(hipSetDevice(s->device));
    Staged sd, se, su, sf;
    int rc;
    if ((rc = stage(s->stream, dep_xyzt, (size_t)ndep * 16, sd)) ||
        (rc = stage(s->stream, dep_e_kev, (size_t)ndep * 4, se)) ||
        (rc = stage(s->stream, use, (size_t)ndep, su)) ||
        (rc = stage(s->stream, ev_first, (size_t)(nev + 1) * 8, sf)))
        return rc;
    if (ndep > 0 && (rc = sur_evaluate(s, (const float4 *)sd.d, (const float *)se.d, ndep))) return rc;
    const size_t cells = (size_t)nev * s->ncell;
    const bool mdev = is_dev(maps), tdev = is_dev(tmean), ldev = is_dev(lam);
    int32_t *dm = maps;
    float *dt = tmean, *dl = lam;
    if (!mdev) SCHK(hipMalloc(&dm, cells * 4));
    if (tmean && !tdev) SCHK(hipMalloc(&dt, cells * 4));
    if (lam && !ldev) SCHK(hipMalloc(&dl, cells * 4));
    hipLaunchKernelGGL(sample_events_net, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->stream,
                       (const float *)s->Y, s->net->n_out, s->cfg.direct ? (const float *)s->D : nullptr,
                       s->cfg.fold ? (const uint8_t *)s->code : nullptr, (const int32_t *)s->dinv,
                       s->ncell, (const float4 *)sd.d, (const float *)se.d, (const uint8_t *)su.d,
                       (const int64_t *)sf.d, nev, yield_per_mev, s->cfg.t_shift, s->cfg.t_scale, seed,
                       dm, dt, dl);
    SCHK(hipGetLastError());
    if (!mdev) SCHK(hipMemcpyAsync(maps, dm, cells * 4, hipMemcpyDeviceToHost, s->stream));
    if (tmean && !tdev) SCHK(hipMemcpyAsync(tmean, dt, cells * 4, hipMemcpyDeviceToHost, s->stream));
    if (lam && !ldev) SCHK(hipMemcpyAsync(lam, dl, cells * 4, hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    if (!mdev) (void)hipFree(dm);
    if (tmean && !tdev) (void)hipFree(dt);
    if (lam && !ldev) (void)hipFree(dl);
    return CRYSP_OK;
}

}  // extern "C"
