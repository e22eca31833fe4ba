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
    EPI_GELU_BWD = 4,     // dZ bf16 = acc * gelu'(Zin)  (Zin fp32, same layout as C)
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
    const float *Zin;                    // EPI_GELU_BWD: pre-activation, ldc layout
};

template <int EPI>
__global__ void __launch_bounds__(256) gemm_bf16(const GemmArgs g)
{
    __shared__ __bf16 sA[BM][KP + LDS_PAD];        // [m][k]
    __shared__ __bf16 sB[BN][KP + LDS_PAD];        // [n][k]
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    // K slice of this workgroup (gridDim.z slices of K / gridDim.z, a multiple of BK),
    // staged KP at a time: every load of a phase is in flight at once, then each wave
    // runs KP/16 MFMA steps on it
    const int kc = g.K / (int)gridDim.z;
    const int kbeg = (int)blockIdx.z * kc, kend = kbeg + kc;
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
            g.C[(size_t)blockIdx.z * g.c_slice + (size_t)row * g.ldc + col] = v;
        } else if (EPI == EPI_BIAS_F32) {
            g.C[(size_t)row * g.ldc + col] = v + g.bias[col];
        } else if (EPI == EPI_BIAS_GELU) {
            v += g.bias[col];
            if (g.C) g.C[(size_t)row * g.ldc + col] = v;
            g.H[(size_t)row * g.ldh + col] = to_bf16(gelu(v));
        } else if (EPI == EPI_BIAS_BF16) {
            g.H[(size_t)row * g.ldh + col] = to_bf16(v + g.bias[col]);
        } else {  // EPI_GELU_BWD
            g.H[(size_t)row * g.ldh + col] = to_bf16(v * gelu_grad(g.Zin[(size_t)row * g.ldc + col]));
        }
    }
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

// loss and its gradient over the real rows, normalized by 1 / sum w:
//   columns c <  n_pois: Poisson negative log-likelihood of a log-rate y for an observed
//                        rate t: w (e^y - t y), gradient w (e^y - t);
//   columns c >= n_pois: weighted squared error w (y - t)^2, gradient 2 w (y - t).
__global__ void __launch_bounds__(256) loss_grad(const float *Y, const float *T, const float *W,
                                                 int M, int Mpad, int N, int n_pois,
                                                 float inv_wsum, __bf16 *dY, float *loss_part,
                                                 const int64_t *bidx, int64_t bstride)
{
    __shared__ float red[4];
    if (bidx) {                                     // targets and weights of batch *bidx
        T += *bidx * bstride;
        if (W) W += *bidx * bstride;
    }
    float l = 0.0f;
    const int64_t tot = (int64_t)Mpad * N;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / N), c = (int)(i % N);
        float g = 0.0f;
        if (r < M) {
            const float w = W ? W[i] : 1.0f;
            if (c < n_pois) {
                const float e = __expf(Y[i]);
                l += w * (e - T[i] * Y[i]) * inv_wsum;
                g = w * (e - T[i]) * inv_wsum;
            } else {
                const float d = Y[i] - T[i];
                l += w * d * d * inv_wsum;
                g = 2.0f * w * d * inv_wsum;
            }
        }
        dY[i] = to_bf16(g);
    }
    // fixed-order reduction: the loss is the same on every run
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) l += __shfl_xor(l, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Partial column sums of a bf16 matrix [M x N] (bias gradient): part[s][c] = sum of rows
// s*rows_per_slice .. of column c.  Block x: 256 columns, block y: slice s.
__global__ void col_sum_part(const __bf16 *G, int M, int N, int rows_per_slice, float *part)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int r0 = blockIdx.y * rows_per_slice, r1 = min(M, r0 + rows_per_slice);
    float s = 0.0f;
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s += from_bf16(G[(size_t)r * N + c]);
    part[(size_t)blockIdx.y * N + c] = s;
}

// out[r][c] = sum over s = 0..S-1, in order, of P[s * slice + r * ldp + c], for c < cols
// (split-K partial products, bias partials); out row stride ldo.
__global__ void sum_slices(const float *P, int S, int64_t slice, int rows, int cols, int ldp,
                           float *out, int ldo)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)rows * cols;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / cols), c = (int)(i % cols);
        const float *q = P + (size_t)r * ldp + c;
        float s = 0.0f;
        int k = 0;
        for (; k + 8 <= S; k += 8) {              // eight loads in flight, added in order
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = q[(size_t)(k + j) * slice];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += v[j];
        }
        for (; k < S; ++k) s += q[(size_t)k * slice];
        out[(size_t)r * ldo + c] = s;
    }
}

// Sum of n values by one workgroup in a fixed order, into out[*idx] (out[0] without idx).
__global__ void __launch_bounds__(256) sum_small(const float *x, int n, float *out,
                                                 const int64_t *idx)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[idx ? *idx : 0] = red[0];
}

// Adam on fp32 master weights; refreshes the bf16 copies in both layouts
// Wt [out][in_pad] (forward, k-contiguous) and Wn [in][out] (backward data gradient).
// dstep (optional, device): the steps taken so far; the bias corrections c1, c2 then
// follow from it on the device (graph replays)
__global__ void adam_update(float *w, const float *g, float *m, float *v, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2, float wd, const int64_t *dstep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dstep) {
        const float t = (float)(*dstep + 1);
        c1 = 1.0f / (1.0f - powf(b1, t));
        c2 = 1.0f / (1.0f - powf(b2, t));
    }
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// end of a replayed training step: next batch, one more Adam step
__global__ void train_step_end(int64_t *ctr)
{
    if (threadIdx.x == 0) { ctr[0] += 1; ctr[1] += 1; }
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
    std::vector<float *> Z;                         // pre-activations fp32 [cap][out]
    std::vector<__bf16 *> H;                        // post-activations bf16 [cap][out]
    float *Y = nullptr;                             // output fp32 [cap][n_out]
    __bf16 *dY = nullptr;
    std::vector<__bf16 *> dZ;                       // [cap][out] per layer
    __bf16 *dH = nullptr;                           // scratch [cap][hidden]
    float *part = nullptr;                          // split-K partial weight gradients
    size_t part_cap = 0;                            // floats
    float *bpart = nullptr;                         // bias-gradient slice partials
    float *lpart = nullptr;                         // loss partials [LOSS_BLOCKS]
    float *loss = nullptr;
    int64_t *ctr = nullptr;                         // device: batch index, Adam steps (graph replays)
    int64_t hctr[2] = {0, 0};
    double flops_fwd_per_row = 0.0;
};

// K slices for a weight-gradient product (output M x N, reduction over the batch K): the
// output alone is only a few 64x64 tiles, so K is split until about 512 workgroups run,
// keeping each slice a multiple of BK and at least two BK steps.
constexpr int SPLIT_TARGET_WG = 512;
static int weight_grad_slices(int M, int N, int K)
{
    const int tiles = (M / BM) * (N / BN);
    int ks = 1;
    while (tiles * ks * 2 <= SPLIT_TARGET_WG && K % (ks * 2 * BK) == 0 && K / (ks * 2) >= 2 * BK)
        ks *= 2;
    return ks;
}
constexpr int LOSS_BLOCKS = 256;                    // loss partials per step
constexpr int BIAS_ROWS = 32;                       // rows per bias-gradient slice

static int mlp_alloc_batch(crysp_mlp *n, int cap)
{
    if (cap <= n->cap) return CRYSP_OK;
    auto fre = [](void *q) { if (q) (void)hipFree(q); };
    fre(n->X); fre(n->Y); fre(n->dY); fre(n->dH);
    for (auto q : n->Z) fre(q);
    for (auto q : n->H) fre(q);
    for (auto q : n->dZ) fre(q);
    n->Z.assign(n->n_layers, nullptr);
    n->H.assign(n->n_layers, nullptr);
    n->dZ.assign(n->n_layers, nullptr);
    SCHK(hipMalloc(&n->X, (size_t)cap * n->kin_pad * 2));
    SCHK(hipMalloc(&n->Y, (size_t)cap * n->n_out * 4));
    SCHK(hipMalloc(&n->dY, (size_t)cap * n->n_out * 2));
    SCHK(hipMalloc(&n->dH, (size_t)cap * std::max(n->hidden, n->n_out) * 2));
    size_t part = 0;
    int bias_cols = 0;
    for (int l = 0; l < n->n_layers; ++l) {
        SCHK(hipMalloc(&n->Z[l], (size_t)cap * n->dout[l] * 4));
        SCHK(hipMalloc(&n->H[l], (size_t)cap * n->dout[l] * 2));
        SCHK(hipMalloc(&n->dZ[l], (size_t)cap * n->dout[l] * 2));
        // the most slices any batch can get: tiles * slices stays within SPLIT_TARGET_WG
        const int tiles = (n->dout[l] / BM) * (n->dpad[l] / BN);
        int ks_max = 1;
        while (tiles * ks_max * 2 <= SPLIT_TARGET_WG) ks_max *= 2;
        part = std::max(part, (size_t)ks_max * n->dout[l] * n->dpad[l]);
        bias_cols = std::max(bias_cols, n->dout[l]);
    }
    fre(n->part); fre(n->bpart); fre(n->lpart);
    n->part = nullptr; n->bpart = nullptr; n->lpart = nullptr;
    SCHK(hipMalloc(&n->part, part * 4));
    n->part_cap = part;
    SCHK(hipMalloc(&n->bpart, (size_t)((cap + BIAS_ROWS - 1) / BIAS_ROWS) * bias_cols * 4));
    SCHK(hipMalloc(&n->lpart, LOSS_BLOCKS * 4));
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

// gradients of the weighted loss into n->g (fp32 flat), the loss into *loss_out (device;
// n->loss when null)
static int mlp_backward_dev(crysp_mlp *n, const float *T, const float *W, int M, int Mpad, double wsum,
                            float *loss_out = nullptr, const int64_t *bidx = nullptr)
{
    const int64_t tot = (int64_t)Mpad * n->n_out;
    const int lblocks = (int)std::min<int64_t>(LOSS_BLOCKS, (tot + 255) / 256);
    hipLaunchKernelGGL(loss_grad, dim3(lblocks), dim3(256), 0, n->stream, n->Y, T, W, M, Mpad,
                       n->n_out, n->n_pois, (float)(1.0 / wsum), n->dY, n->lpart, bidx,
                       (int64_t)M * n->n_out);
    hipLaunchKernelGGL(sum_small, dim3(1), dim3(256), 0, n->stream, n->lpart, lblocks,
                       loss_out ? loss_out : n->loss, loss_out ? bidx : nullptr);
    const __bf16 *dcur = n->dY;                    // gradient w.r.t. layer l's output (pre-act)
    for (int l = n->n_layers - 1; l >= 0; --l) {
        const __bf16 *hin = l == 0 ? n->X : n->H[l - 1];
        const int kin = n->dpad[l], nout = n->dout[l];
        // dW[out][in] = dZ^T[out x M] * Hin[M x in]   (A = dZ stored [M][out] -> a_t), as
        // split-K partial products summed in a fixed order; at layer 0 only the real
        // input columns of the padded width are kept
        {
            const int ks = weight_grad_slices(nout, kin, Mpad);
            if ((size_t)ks * nout * kin > n->part_cap) return sfail(CRYSP_EINVAL, "split-K buffer too small");
            GemmArgs a{};
            a.M = nout; a.N = kin; a.K = Mpad;
            a.A = dcur; a.lda = nout; a.a_t = 1;
            a.B = hin; a.ldb = kin; a.b_t = 0;
            a.C = n->part; a.ldc = kin; a.c_slice = (int64_t)nout * kin;
            gemm<EPI_F32>(n->stream, a, ks);
            const int64_t cells = (int64_t)nout * n->din[l];
            hipLaunchKernelGGL(sum_slices, dim3((unsigned)std::min<int64_t>(1024, (cells + 255) / 256)),
                               dim3(256), 0, n->stream, n->part, ks, a.c_slice, nout, n->din[l], kin,
                               n->g + n->woff[l], n->din[l]);
        }
        {
            const int slices = (Mpad + BIAS_ROWS - 1) / BIAS_ROWS;
            hipLaunchKernelGGL(col_sum_part, dim3((nout + 255) / 256, slices), dim3(256), 0,
                               n->stream, dcur, Mpad, nout, BIAS_ROWS, n->bpart);
            hipLaunchKernelGGL(sum_slices, dim3((nout + 255) / 256), dim3(256), 0, n->stream,
                               n->bpart, slices, (int64_t)nout, 1, nout, nout, n->g + n->boff[l], nout);
        }
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l [M x out] * W_l [out x in]) * gelu'(Z_{l-1});  B = Wn [in][out] -> b_t
        {
            GemmArgs a{};
            a.M = Mpad; a.N = n->din[l]; a.K = nout;
            a.A = dcur; a.lda = nout; a.a_t = 0;
            a.B = n->Wn[l]; a.ldb = nout; a.b_t = 1;
            a.H = n->dZ[l - 1]; a.ldh = a.N;
            a.Zin = n->Z[l - 1]; a.ldc = a.N;
            gemm<EPI_GELU_BWD>(n->stream, a);
        }
        dcur = n->dZ[l - 1];
    }
    return CRYSP_OK;
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
    if (n_in < 1 || n_in > 32 || hidden < 64 || hidden % 64 || n_out < 64 || n_out % 64 ||
        n_hidden_layers < 1 || n_hidden_layers > 8) {
        sfail(CRYSP_EINVAL, "need 1 <= n_in <= 32, hidden and n_out multiples of 64");
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
    fre(n->p); fre(n->g); fre(n->m); fre(n->v); fre(n->shift); fre(n->scale); fre(n->loss); fre(n->ctr);
    fre(n->part); fre(n->bpart); fre(n->lpart);
    for (auto q : n->Wt) fre(q);
    for (auto q : n->Wn) fre(q);
    fre(n->X); fre(n->Y); fre(n->dY); fre(n->dH);
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
    if ((rc = mlp_forward_dev(n, (const float *)sx.d, M, mpad))) return rc;
    if ((rc = mlp_backward_dev(n, (const float *)st.d, (const float *)sw.d, M, mpad, wsum))) return rc;
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
    float *dl = nullptr;
    if (losses && n_batches > 0) SCHK(hipMalloc(&dl, (size_t)n_batches * 4));
    // One step (forward, backward, Adam, bf16 refresh) is captured once and replayed per
    // batch: the batch index and the Adam step count live on the device and advance at
    // the end of each replay, so the ~30 small launches of a step cost one graph launch.
    const unsigned adam_blocks = (unsigned)((n->n_params + 255) / 256);
    auto enqueue_step = [&](const float *x, const float *t, const float *w, float *lout,
                            const int64_t *bidx, const int64_t *dstep, float c1, float c2) -> int {
        int r;
        if ((r = mlp_forward_dev(n, x, batch, batch, bidx))) return r;
        if ((r = mlp_backward_dev(n, t, w, batch, batch, wsum, lout, bidx))) return r;
        hipLaunchKernelGGL(adam_update, dim3(adam_blocks), dim3(256), 0, n->stream, n->p, n->g,
                           n->m, n->v, (int)n->n_params, lr, beta1, beta2, eps, c1, c2,
                           weight_decay, dstep);
        mlp_refresh(n);
        return CRYSP_OK;
    };
    hipGraphExec_t exec = nullptr;
    if (n_batches > 1) {
        n->hctr[0] = 0;
        n->hctr[1] = n->step;
        SCHK(hipMemcpyAsync(n->ctr, n->hctr, sizeof n->hctr, hipMemcpyHostToDevice, n->stream));
        SCHK(hipStreamSynchronize(n->stream));
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            rc = enqueue_step(X, T, W, dl, n->ctr, n->ctr + 1, 0.0f, 0.0f);
            if (rc == CRYSP_OK)
                hipLaunchKernelGGL(train_step_end, dim3(1), dim3(64), 0, n->stream, n->ctr);
            const hipError_t e = hipStreamEndCapture(n->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); if (dl) (void)hipFree(dl); return rc; }
            if (e != hipSuccess || !g || hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess)
                exec = nullptr;
            if (g) (void)hipGraphDestroy(g);
        }
        if (!exec) (void)hipGetLastError();         // capture unavailable: launch eagerly
    }
    if (exec) {
        for (int64_t b = 0; b < n_batches; ++b) SCHK(hipGraphLaunch(exec, n->stream));
        n->step += (int)n_batches;
        SCHK(hipStreamSynchronize(n->stream));
        (void)hipGraphExecDestroy(exec);
    } else {
        for (int64_t b = 0; b < n_batches; ++b) {
            const float *x = X + (size_t)b * batch * n->n_in;
            const float *t = T + (size_t)b * batch * n->n_out;
            const float *w = W ? W + (size_t)b * batch * n->n_out : nullptr;
            n->step += 1;
            const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)n->step));
            const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)n->step));
            if ((rc = enqueue_step(x, t, w, dl ? dl + b : nullptr, nullptr, nullptr, c1, c2))) {
                if (dl) (void)hipFree(dl);
                return rc;
            }
        }
    }
    if (dl) {
        SCHK(hipMemcpyAsync(losses, dl, (size_t)n_batches * 4, hipMemcpyDeviceToHost, n->stream));
    }
    SCHK(hipStreamSynchronize(n->stream));
    if (dl) (void)hipFree(dl);
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
