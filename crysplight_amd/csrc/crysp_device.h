// crysp_device.h -- per-lane physics of the optical-photon transport on gfx950.
//
// One lane owns one photon at a time.  Everything here is IEEE binary32 arithmetic
// compiled with -ffp-contract=off and correctly rounded division / sqrt, so each
// operation gives the same bits as the Julia Float32 code it follows.  The
// transcendentals (log, sin, cos) are evaluated in binary64 and rounded once to
// binary32; Julia's Float32 versions are computed the same way and are correctly
// rounded in practice (explicit fma() calls below are part of those binary64 kernels,
// not contractions of the Float32 physics).  Two comparisons are first tried on hardware
// estimates (v_log_f32, v_sqrt_f32, v_rcp_f32) whose error bounds are pinned by tests; a
// lane whose estimate lies within a band of the threshold computes the exact value, so the
// outcome of every comparison is that of the exact arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "math_tables.h"           // sin, cos of k pi/128 (f_sincos); 1/c, log c (f_log)

namespace crysp {

enum : int { ST_FLYING = 0, ST_DETECTED = 1, ST_ABS_BULK = 2, ST_ABS_WALL = 3,
             ST_ABS_SIPM = 4, ST_CAP = 5, ST_ABS_SUR = 6 };
// face codes, src/geometry.jl:4-9
enum : int { F_XP = 1, F_XM = 2, F_YP = 3, F_YM = 4, F_BACK = 5, F_FRONT = 6 };

// flag bits of KParams::flags
enum : uint32_t { K_ABS = 1u, K_RAY = 2u, K_WRAP_SPEC = 4u, K_AIR_GAP = 8u, K_DISC = 16u,
                  K_SUR_SPEC = 32u, K_TWO_SIDED = 64u, K_ROUGH = 128u };

// The optics feature set of a kernel variant.  A specialized variant bakes one flag set
// into its code: its branches fold at compile time, and no flag masks stay live in
// registers across the step.  OPT_GENERIC reads the flags from KParams at run time.  The
// host launches a specialization only when the launch's flags equal it exactly.
constexpr uint32_t OPT_GENERIC = 0xFFFFFFFFu;
constexpr uint32_t OPT_ESR = K_ABS | K_WRAP_SPEC | K_AIR_GAP | K_ROUGH;   // specular wrap
constexpr uint32_t OPT_TEFLON = K_ABS | K_AIR_GAP | K_ROUGH;              // Lambertian wrap
constexpr uint32_t OPT_TEFLON_RAY = K_ABS | K_RAY | K_AIR_GAP | K_ROUGH;  // + Rayleigh

constexpr float TWO_PI_F = 6.2831855f;        // 2f0 * Float32(pi): exact product
constexpr float C_MM_NS = 299.792458f;        // src/transport.jl:4

// Everything a launch needs that is the same for all lanes (kernarg -> SGPRs).
struct KParams {
    float L0, L1, L2;
    float n_crystal, n_coupling, abs_len, ray_len, wrap_R, sigma_alpha, pde;
    float inv_v;                              // n_crystal / c  [ns/mm]
    float neg_abs_ln2;                        // -abs_len * ln2 (the absorption estimate)
    float q_coup, q_air;                      // n_crystal / n_coupling, n_crystal / 1
    float cx, cy, r2, sur_R;
    float pitch, bin_ns;
    float tau_d, tau_r;
    int32_t nx, ny, nbins, maxb;
    uint32_t flags;
    uint32_t key0, key1;                      // Philox key = (seed lo, seed hi)
    int32_t npos;
    uint64_t pid0;
    int64_t n_photons;
    int64_t ndep;
    int64_t nev;
    int32_t chunk;                            // photons per work-queue grab
    int32_t refill_min;                       // refill lanes once this many are idle
    int32_t hist_words;                       // nx*ny*(nbins+1)
    int32_t nbox;                             // crystals of a several-crystal event queue
    // event mode, events owned by workgroups
    int32_t gate;                             // 0 run; 1 run if the deposits are ordered by
                                              // event, 2 if not (device flag bit 8 decides)
    int32_t piece;                            // photons per workgroup claim of an event
    int32_t pool_cap;                         // partial maps the helpers may write
    int64_t help_window;                      // helpers look among the last this many events
};

template <uint32_t OPT>
__device__ __forceinline__ bool has(const KParams &P, uint32_t bit)
{
    return OPT == OPT_GENERIC ? (P.flags & bit) != 0u : (OPT & bit) != 0u;
}

// A field of the running transport kernel's arguments, read from the kernel argument
// segment at each use (KParams first, as the kernel's first argument).  Code that runs
// rarely (a detection, a refill, an emission, the session bookkeeping) reads its parameters
// this way, so they hold no scalar registers across the photon loop: there the compiler
// would keep each one live and, short of scalar registers, spill it to a vector lane, which
// costs a v_readlane (a vector instruction) at every use.  Only for code that
// transport_kernel runs.
template <typename T>
__device__ __forceinline__ T karg_at(size_t off)
{
    const __attribute__((address_space(4))) char *kp =
        (const __attribute__((address_space(4))) char *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(kp));
    return *(const __attribute__((address_space(4))) T *)(kp + off);
}
#define KP(f) karg_at<decltype(KParams::f)>(offsetof(KParams, f))

// ------------------------------------------------------------------ Philox ----
// Philox4x32-10 (src/philox.jl:10-28): ten rounds, the key bumped between rounds.
// The key is uniform: the empty asm makes it opaque here, so the ten bumped keys are
// scalar adds at each use instead of twenty precomputed values held (and spilled)
// across the whole kernel.
__device__ __forceinline__ uint4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1)
{
    asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product each (v_mad_u64_u32) gives both halves
        // and one three-input XOR (v_bitop3_b32, truth table 0x96) for each new word
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        c0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        c1 = (uint32_t)p1;
        c2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// Per-photon stream: counter (pid lo, pid hi, block, 0); four words per block, used in
// order (src/philox.jl:34-57).  Generated words wait in a ring of 16 words per lane in
// LDS: 16 slots x 64 lanes per wavefront, slot-major, so a wavefront's reads and writes
// are free of bank conflicts whatever slot each lane is at.  A new photon starts with
// block 0 (rng_start); after that blocks are generated at the top of each bounce step,
// for all lanes at once, while the ring has room (rng_prefetch), so a step starts with
// 9..16 words in hand and a draw is one LDS read.  Only a step that outruns the
// ring generates a block at the draw itself.
constexpr uint32_t RING_SLOT = 256u;              // bytes between a lane's ring slots
constexpr uint32_t RING_POS_MASK = 15u * RING_SLOT;
constexpr uint32_t RING_WAVE_BYTES = 16u * RING_SLOT;

struct Rng {
    uint32_t lo, hi, blk;                     // counter words of the next block to generate
    uint32_t rd, wr;                          // words drawn / generated, x RING_SLOT
    uint32_t base;                            // this lane's byte offset in ring slot 0
    char *ring;                               // the workgroup's rings (LDS)
};

// ring: LDS of blockDim.x * 64 bytes; blockDim.x is a multiple of 64
__device__ __forceinline__ void rng_bind(Rng &r, void *ring)
{
    r.ring = (char *)ring;
    r.base = (threadIdx.x >> 6) * RING_WAVE_BYTES + (threadIdx.x & 63u) * 4u;
}

__device__ __forceinline__ void rng_init(Rng &r, uint64_t pid)
{
    r.lo = (uint32_t)pid; r.hi = (uint32_t)(pid >> 32); r.blk = 0u;
    r.rd = 0u; r.wr = 0u;
}

// Operation counters of the counting kernel variant, kept in the workgroup's LDS so they
// add no registers: n_photon, n_draw, n_block, n_step, n_fresnel, n_surface_try,
// n_lambert, n_rayleigh_try, n_detect, n_nscat.
template <bool COUNT>
struct OpCount {
    unsigned long long *lds;
    __device__ __forceinline__ void add(int k, uint32_t d = 1u)
    {
        if (COUNT) atomicAdd(&lds[k], (unsigned long long)d);
    }
};

// Byte address of ring position pos: its slot bits over the lane's base (disjoint bits).
__device__ __forceinline__ uint32_t ring_addr(const Rng &r, uint32_t pos)
{
    return (pos & RING_POS_MASK) | r.base;
}

// Generate the next block into the ring (the caller guarantees four free slots).  The
// writer's slot is a multiple of four, so the block's slots are consecutive.
__device__ __forceinline__ void rng_fill(Rng &r, uint32_t k0, uint32_t k1)
{
    const uint4 b = philox10(r.lo, r.hi, r.blk, 0u, k0, k1);
    char *q = r.ring + ring_addr(r, r.wr);
    *(uint32_t *)(q) = b.x;
    *(uint32_t *)(q + RING_SLOT) = b.y;
    *(uint32_t *)(q + 2 * RING_SLOT) = b.z;
    *(uint32_t *)(q + 3 * RING_SLOT) = b.w;
    r.blk += 1u;
    r.wr += 4u * RING_SLOT;
}

// A new photon: block 0, enough for the emission (at most four draws); the step-top
// sites fill the ring further together with the other lanes.
__device__ __forceinline__ void rng_start(Rng &r, uint64_t pid, uint32_t k0, uint32_t k1)
{
    rng_init(r, pid);
    rng_fill(r, k0, k1);
}

// Top of a step: one block while the ring has four free slots, and a second one for a
// lane down to eight words.  A step draws four words on average; photons whose steps
// draw five or six (the escape cone through the air gap keeps its angle at every face of
// the same axis) would otherwise drain the ring and generate inside the step.
__device__ __forceinline__ void rng_prefetch(Rng &r, uint32_t k0, uint32_t k1)
{
    if (r.wr - r.rd <= 12u * RING_SLOT) rng_fill(r, k0, k1);
    if (r.wr - r.rd <= 8u * RING_SLOT) rng_fill(r, k0, k1);
}

// CHECK = false: a draw that cannot outrun the ring.  The top of every step leaves at
// least eight words (rng_prefetch), so the first eight draws of a step need no check;
// such a draw has no branch before its LDS read, which the compiler can then move ahead.
template <bool COUNT, bool CHECK = true>
__device__ __forceinline__ uint32_t rng_next(Rng &r, const KParams &P, OpCount<COUNT> &oc)
{
    if (CHECK && r.rd == r.wr) rng_fill(r, P.key0, P.key1);
    if (COUNT && ((r.rd / RING_SLOT) & 3u) == 0u) oc.add(2);   // a block's first word
    oc.add(1);
    const uint32_t x = *(const uint32_t *)(r.ring + ring_addr(r, r.rd));
    r.rd += RING_SLOT;
    return x;
}

// ((x >>> 8) + 1) * 2^-24 in (0, 1] (src/philox.jl:62): every step exact.
__device__ __forceinline__ float unit_of(uint32_t x)
{
    return ((float)(x >> 8) + 1.0f) * 5.9604645e-08f;
}

template <bool COUNT, bool CHECK = true>
__device__ __forceinline__ float randu(Rng &r, const KParams &P, OpCount<COUNT> &oc)
{
    return unit_of(rng_next<COUNT, CHECK>(r, P, oc));
}

// ----------------------------------------------------------- transcendentals ----
// The transport evaluates log only at randu outputs (0, 1] and sin/cos only at
// phi = 2f0*Float32(pi)*u in (0, 2pi] and at tilts alpha in [0, 1.2).  These versions
// are written for that domain (x in [0, 2pi] for sin/cos, any positive normal float
// for log): the float argument is widened to binary64, reduced exactly, expanded in a
// short series with ~1 ulp binary64 error, and rounded once to binary32 -- the
// correctly rounded result except when the exact value lies within ~2^-51 relative of
// a binary32 rounding boundary.  The tests compare every one of the 2^24 log and
// phase arguments against a correctly rounded reference.

// a * b + c for a binary64 constant c, which the VOP3 fma reads from a scalar register
// pair.  Written as __builtin_fma, the compiler selects the two-address fmac form, whose
// addend must be the destination VGPR pair: every Horner step then copied its constant into
// VGPRs (one v_mov_b64 per term) and the resident constants held 20-30 VGPRs.
__device__ __forceinline__ double fma_k(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// z * b + c for binary64 constants b (a scalar register pair) and c, the first step of a
// Horner chain: c is moved into VGPRs at the use instead of the leading coefficient being
// held in a VGPR pair across the whole kernel.  The product is commutative, so this is the
// same fma as fma_k(b, z, c).
__device__ __forceinline__ double fma_first(double z, double b, double c)
{
    double cv;
    asm volatile("v_mov_b64 %0, %1" : "=v"(cv) : "s"(c));
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(z), "s"(b), "v"(cv));
    return r;
}

// log(x) = e ln2 + log(c) + log1p(r),  x = 2^e m,  m in [sqrt(2)/2, sqrt(2)),  r = m (1/c) - 1
// (one fma, no division): the low exponent bit and the 7 leading mantissa bits of m pick
// one of 129 intervals, whose 1/c (rounded) and log(c) = -log(1/c) come from a table; c is
// 1 for the two intervals next to 1, so log(x) near 1 is log1p(r) alone, accurate relative
// to itself.  |r| <= 2^-7: log1p to r^8 leaves < 2^-56 relative.  Over every randu value
// the binary64 result is within 2^-51.5 relative of the exact value, which comes no closer
// than 2^-47.5 to a binary32 rounding boundary.  Every float argument maps to an index
// inside the table.
__device__ __forceinline__ float f_log(float x)
{
    const uint32_t b = __float_as_uint(x);
    int e = (int)(b >> 23) - 127;
    uint32_t mb = (b & 0x007FFFFFu) | 0x3F800000u;
    if (mb > 0x3FB504F3u) { mb -= 0x00800000u; ++e; }      // m >= sqrt(2): halve it
    const uint32_t i = ((mb >> 16) & 0xFFu) - 0x35u;         // 0..128
    const double inv = LOG_TAB[i][0], T = LOG_TAB[i][1];
    const double m = (double)__uint_as_float(mb);
    const double r = __builtin_fma(m, inv, -1.0);
    const double z = r * r;
    double p = fma_first(r, -1.0 / 8.0, 1.0 / 7.0);
    p = fma_k(p, r, -1.0 / 6.0);
    p = fma_k(p, r, 0.2);
    p = fma_k(p, r, -0.25);
    p = fma_k(p, r, 1.0 / 3.0);
    p = fma_k(p, r, -0.5);
    const double ed = (double)e;
    // ln2 in two parts: ed * LN2_HI is exact for |e| < 2^21
    return (float)__builtin_fma(ed, 0.6931471803691238,
                                __builtin_fma(ed, 1.9082149292705877e-10, T + __builtin_fma(z, p, r)));
}

// sin and cos of x in [0, 2pi (1 + 2^-20)] (the phases 2f0*Float32(pi)*u and the tilts):
// x = k pi/128 + r with k = rint(128 x / pi) in 0..256 and pi/128 in three parts (33 + 33
// + 53 bits, so k * P1 and x - k * P1 are exact), |r| <= pi/256.  sin r to r^7 and
// cos r - 1 to r^6 leave less than 2^-66 relative; with S, C = sin, cos(k pi/128) from a
// table of correctly rounded binary64 values,
//   sin x = S + (S (cos r - 1) + C sin r),   cos x = C + (C (cos r - 1) - S sin r),
// the large term added last.  Over every phase argument the binary64 result is within
// 2^-51.5 relative of the exact value, which comes no closer than 2^-49.2 to a binary32
// rounding boundary, so the rounding to binary32 is the correct one; the tests check all
// 2^24 phases and the tilt domain on the device.  About two thirds of the instructions of
// a quadrant reduction with r^17 / r^16 series and no quadrant selects.  An index outside
// the table (an argument outside the domain) is clamped into it.
__device__ __forceinline__ void f_sincos(float xf, float &sf, float &cf)
{
    const double x = (double)xf;
    const double kd = __builtin_rint(x * 40.74366543152521);       // 128 / pi
    int k = (int)kd;
    k = k < 0 ? 0 : (k > 256 ? 256 : k);
    const double S = SINCOS_TAB[k][0], C = SINCOS_TAB[k][1];
    double r = __builtin_fma(-kd, 1.5707963267341256 / 64.0, x);
    r = __builtin_fma(-kd, 6.077100506303966e-11 / 64.0, r);
    r = __builtin_fma(-kd, 2.0222662487959506e-21 / 64.0, r);
    const double z = r * r;
    double q = fma_first(z, -1.0 / 5040.0, 1.0 / 120.0);
    q = fma_k(q, z, -1.0 / 6.0);
    const double sn = __builtin_fma(r * z, q, r);                  // sin r
    double p = fma_first(z, -1.0 / 720.0, 1.0 / 24.0);
    p = fma_k(p, z, -0.5);
    const double cm1 = p * z;                                      // cos r - 1
    sf = (float)(S + __builtin_fma(S, cm1, C * sn));
    cf = (float)(C + __builtin_fma(C, cm1, -(S * sn)));
}

// sin and cos of a micro-facet tilt alpha >= 0.  Below 1/8 rad, where alpha = sigma |g|
// almost always lies, the Taylor series go straight from alpha with no reduction: the
// first omitted terms, a^13/13! and a^14/14!, are below 2^-66 relative.  Larger tilts
// take f_sincos.
__device__ __forceinline__ void f_sincos_tilt(float af, float &sf, float &cf)
{
    if (!(af < 0.125f)) { f_sincos(af, sf, cf); return; }
    const double a = (double)af;
    const double z = a * a;
    double ps = fma_first(z, -2.505210838544172e-08, 2.7557319223985893e-06);  // -1/11!, 1/9!
    ps = fma_k(ps, z, -0.0001984126984126984);       // -1/7!
    ps = fma_k(ps, z, 0.008333333333333333);         // 1/5!
    ps = fma_k(ps, z, -0.16666666666666666);         // -1/3!
    const double sn = __builtin_fma(a * z, ps, a);
    double pc = fma_first(z, 2.08767569878681e-09, -2.755731922398589e-07);  // 1/12!, -1/10!
    pc = fma_k(pc, z, 2.48015873015873e-05);         // 1/8!
    pc = fma_k(pc, z, -0.001388888888888889);        // -1/6!
    pc = fma_k(pc, z, 0.041666666666666664);         // 1/4!
    pc = __builtin_fma(pc, z, -0.5);                         // -1/2!
    const double cs = __builtin_fma(z, pc, 1.0);
    sf = (float)sn;
    cf = (float)cs;
}

__device__ __forceinline__ float f_cos(float x)
{
    float s, c;
    f_sincos(x, s, c);
    return c;
}

// Correctly rounded binary32 sqrt for 0 and normal arguments -- every argument in the
// transport is 0 or at least 2^-25, so the input scaling of the general sequence is left
// out: the hardware estimate (within 1 ulp) moves to a neighbour when the fma residual
// of that neighbour shows it is the better one.
__device__ __forceinline__ float f_sqrt(float x)
{
    const float s0 = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s0) - 1u);
    const float su = __uint_as_float(__float_as_uint(s0) + 1u);
    const float rd = __builtin_fmaf(-sd, s0, x);
    const float ru = __builtin_fmaf(-su, s0, x);
    float s = rd <= 0.0f ? sd : s0;
    s = ru > 0.0f ? su : s;
    return s;
}

// Hardware log2 estimate (v_log_f32), not correctly rounded: only for decisions whose
// margin covers its error, with the exact f_log behind them.
__device__ __forceinline__ float f_log2_est(float x) { return __builtin_amdgcn_logf(x); }

// Julia Base semantics used by the reference
__device__ __forceinline__ float jmax0(float v) { return (v != v) ? v : (v > 0.0f ? v : 0.0f); }
__device__ __forceinline__ float jclamp(float x, float lo, float hi)
{
    return x > hi ? hi : (x < lo ? lo : x);
}

// ----------------------------------------------------------------- optics ----
// Unpolarized Fresnel with TIR (src/optics.jl:10-19).  q = n1 / n2 is the same binary32
// quotient for every call at an interface, divided once on the host.  sin^2 of the
// refraction angle: TIR when it reaches 1.
__device__ __forceinline__ float fresnel_sin2t(float q, float cosi)
{
    return (q * q) * (1.0f - cosi * cosi);
}

// The reflectance R for sin2t < 1, every operation rounded as the reference rounds it.
__device__ __forceinline__ float fresnel_R(float n1, float n2, float cosi, float sin2t)
{
    const float cost = f_sqrt(1.0f - sin2t);
    const float a = (n1 * cosi - n2 * cost) / (n1 * cosi + n2 * cost);
    const float b = (n1 * cost - n2 * cosi) / (n1 * cost + n2 * cosi);
    return 0.5f * (a * a + b * b);
}

// The same R from the hardware sqrt and reciprocal estimates (about a third of the
// instructions).  Each error term of a numerator is a few 2^-24 times a term of its own
// denominator, so |a|, |b| <= 1 move by a few 2^-23 and R by less than 2^-19, far inside
// FRESNEL_BAND (the tests pin the gap on a dense cosine grid over several interfaces).
__device__ __forceinline__ float fresnel_R_est(float n1, float n2, float cosi, float sin2t)
{
    const float cost = __builtin_amdgcn_sqrtf(1.0f - sin2t);
    const float x1 = n1 * cosi, x2 = n2 * cost, y1 = n1 * cost, y2 = n2 * cosi;
    const float a = (x1 - x2) * __builtin_amdgcn_rcpf(x1 + x2);
    const float b = (y1 - y2) * __builtin_amdgcn_rcpf(y1 + y2);
    return 0.5f * (a * a + b * b);
}

constexpr float FRESNEL_BAND = 6.1035156e-05f;       // 2^-14
constexpr float ABS_BAND = 2.4414062e-04f;           // 2^-12

// The Fresnel reflection decision at an interface: TIR reflects without a draw;
// otherwise one draw u reflects when u < R (src/kernel.jl:220-226, 269-274; the UNIFIED
// tries, src/optics.jl:67-68).  The estimate decides unless u lies within FRESNEL_BAND of
// it; there the exact R decides.  The outcome is always the exact R's.
template <bool COUNT, bool CHECK = true>
__device__ __forceinline__ bool fresnel_reflects(Rng &r, const KParams &P, OpCount<COUNT> &oc,
                                                 float q, float n1, float n2, float cosi)
{
    const float sin2t = fresnel_sin2t(q, cosi);
    if (sin2t >= 1.0f) return true;
    const float u = randu<COUNT, CHECK>(r, P, oc);
    const float re = fresnel_R_est(n1, n2, cosi, sin2t);
    if (u < re - FRESNEL_BAND) return true;
    if (u >= re + FRESNEL_BAND) return false;
    return u < fresnel_R(n1, n2, cosi, sin2t);
}

__device__ __forceinline__ int face_axis(int f) { return f <= 2 ? 1 : (f <= 4 ? 2 : 3); }

// Specular flip of the face-normal component (src/optics.jl:25-28), written as three
// selects so the direction stays in registers (an axis-indexed write would put it in
// scratch memory).
__device__ __forceinline__ void specular(float &ux, float &uy, float &uz, int face)
{
    const int a = face_axis(face);
    ux = a == 1 ? -ux : ux;
    uy = a == 2 ? -uy : uy;
    uz = a == 3 ? -uz : uz;
}

// Lambertian re-injection about the inward normal (src/optics.jl:118-138).
template <bool COUNT>
__device__ __forceinline__ void lambertian(Rng &r, const KParams &P, OpCount<COUNT> &oc,
                                           int face, float &ux, float &uy, float &uz)
{
    oc.add(6);
    const float ct = f_sqrt(randu<COUNT>(r, P, oc));
    const float st = f_sqrt(jmax0(1.0f - ct * ct));
    const float phi = TWO_PI_F * randu<COUNT>(r, P, oc);
    float sp, cp;
    f_sincos(phi, sp, cp);
    const float a = st * cp, b = st * sp;
    switch (face) {
    case F_XP: ux = -ct; uy = a; uz = b; break;
    case F_XM: ux = ct; uy = a; uz = b; break;
    case F_YP: ux = a; uy = -ct; uz = b; break;
    case F_YM: ux = a; uy = ct; uz = b; break;
    case F_BACK: ux = a; uy = b; uz = -ct; break;
    default: ux = a; uy = b; uz = ct; break;
    }
}

// Rayleigh (1 + cos^2) scatter about the current direction (src/optics.jl:87-112).
template <bool COUNT>
__device__ __forceinline__ void rayleigh(Rng &r, const KParams &P, OpCount<COUNT> &oc,
                                         float &ux, float &uy, float &uz)
{
    float mu;
    for (;;) {
        oc.add(7);
        mu = 1.0f - 2.0f * randu<COUNT>(r, P, oc);
        const float u2 = randu<COUNT>(r, P, oc);
        if (2.0f * u2 <= 1.0f + mu * mu) break;
    }
    const float st = f_sqrt(jmax0(1.0f - mu * mu));
    const float phi = TWO_PI_F * randu<COUNT>(r, P, oc);
    float sp, cp;
    f_sincos(phi, sp, cp);
    const float cb = st * cp, sb = st * sp;
    // transverse basis: t1 = (-uy, ux, 0) / |.| if |uz| < 0.9, else (0, -uz, uy) / |.|; the
    // operands are selected first so that one sqrt and one division serve both cases
    const bool near_z = fabsf(uz) < 0.9f;
    const float pa = near_z ? ux : uy, pb = near_z ? uy : uz;
    float inv = 1.0f / f_sqrt(pa * pa + pb * pb);
    const float t1x = near_z ? -uy * inv : 0.0f;
    const float t1y = near_z ? ux * inv : -uz * inv;
    const float t1z = near_z ? 0.0f : uy * inv;
    const float t2x = uy * t1z - uz * t1y;
    const float t2y = uz * t1x - ux * t1z;
    const float t2z = ux * t1y - uy * t1x;
    const float vx = mu * ux + cb * t1x + sb * t2x;
    const float vy = mu * uy + cb * t1y + sb * t2y;
    const float vz = mu * uz + cb * t1z + sb * t2z;
    inv = 1.0f / f_sqrt(vx * vx + vy * vy + vz * vz);
    ux = vx * inv; uy = vy * inv; uz = vz * inv;
}

// UNIFIED specular-lobe interaction at the rough crystal-air interface
// (src/optics.jl:49-81): up to 20 micro-facet tries, fallback ideal specular.
//
// facet_try: the first half of a try -- the Box-Muller tilt (two draws), the azimuth
// (one draw) and the facet normal.  False when the try is void (tilt >= 1.2 rad, or the
// facet faces away from the photon); the caller then moves on to the next try.
template <bool COUNT, bool CHECK = true>
__device__ __forceinline__ bool facet_try(Rng &r, const KParams &P, OpCount<COUNT> &oc,
                                          float ux, float uy, float uz, int a, float sgn,
                                          float &nfx, float &nfy, float &nfz, float &cosi)
{
    oc.add(5);
    const float u1 = randu<COUNT, CHECK>(r, P, oc);
    const float u2 = randu<COUNT, CHECK>(r, P, oc);
    const float g = f_sqrt(-2.0f * f_log(u1)) * f_cos(TWO_PI_F * u2);
    const float alpha = P.sigma_alpha * fabsf(g);
    if (alpha >= 1.2f) return false;
    float sa, ca;
    f_sincos_tilt(alpha, sa, ca);
    const float phi = TWO_PI_F * randu<COUNT, CHECK>(r, P, oc);
    float sp, cp;
    f_sincos(phi, sp, cp);
    const float b = sa * cp, c = sa * sp;
    nfx = a == 1 ? ca * sgn : b;
    nfy = a == 2 ? ca * sgn : (a == 1 ? b : c);
    nfz = a == 3 ? ca * sgn : c;
    cosi = ux * nfx + uy * nfy + uz * nfz;
    return !(cosi <= 0.0f);
}

// The end of a try that Fresnel-reflected: mirror about the facet, accepted when the
// result points into the crystal (against the outward normal, whose only nonzero
// component is on the face axis).
__device__ __forceinline__ bool facet_reflect(float &ux, float &uy, float &uz, float cosi,
                                              float nfx, float nfy, float nfz, int a, float sgn)
{
    const float rx = ux - 2.0f * cosi * nfx;
    const float ry = uy - 2.0f * cosi * nfy;
    const float rz = uz - 2.0f * cosi * nfz;
    const float rn = a == 1 ? rx * sgn : (a == 2 ? ry * sgn : rz * sgn);
    if (rn < 0.0f) { ux = rx; uy = ry; uz = rz; return true; }
    return false;
}

__device__ __forceinline__ float face_sign(int face)
{
    return (face == F_XP || face == F_YP || face == F_BACK) ? 1.0f : -1.0f;
}

// Tries k0..19 of the interaction.  Returns true when the photon is transmitted into
// the gap.
template <bool COUNT>
__device__ __forceinline__ bool unified_surface(Rng &r, const KParams &P, OpCount<COUNT> &oc,
                                                float &ux, float &uy, float &uz, int face,
                                                int k0 = 0)
{
    const int a = face_axis(face);
    const float sgn = face_sign(face);
    for (int k = k0; k < 20; ++k) {
        float nfx, nfy, nfz, cosi;
        if (!facet_try<COUNT>(r, P, oc, ux, uy, uz, a, sgn, nfx, nfy, nfz, cosi)) continue;
        oc.add(4);
        if (!fresnel_reflects<COUNT>(r, P, oc, P.q_air, P.n_crystal, 1.0f, cosi)) return true;
        if (facet_reflect(ux, uy, uz, cosi, nfx, nfy, nfz, a, sgn)) return false;
    }
    specular(ux, uy, uz, face);
    return false;
}

// Emission (src/generation.jl:4-13, src/kernel.jl:175-180) + rise time.
template <bool COUNT>
__device__ __forceinline__ void emit(Rng &r, const KParams &P, OpCount<COUNT> &oc, float t0,
                                     float &ux, float &uy, float &uz, float &t)
{
    const float u1 = randu<COUNT>(r, P, oc);
    const float z = 1.0f - 2.0f * u1;
    const float rr = f_sqrt(jmax0(1.0f - z * z));
    const float phi = TWO_PI_F * randu<COUNT>(r, P, oc);
    float sp, cp;
    f_sincos(phi, sp, cp);
    ux = rr * cp; uy = rr * sp; uz = z;
    t = t0;
    const float tau_d = KP(tau_d), tau_r = KP(tau_r);
    if (tau_d > 0.0f) t += -tau_d * f_log(randu<COUNT>(r, P, oc));
    if (tau_r > 0.0f) t += -tau_r * f_log(randu<COUNT>(r, P, oc));
}

// Emission in batches.  A wave computes the emissions of the next (up to) 64 photons it
// will hand out with all its lanes at once, instead of a few lanes at each refill: lane i
// generates block 0 of its photon's stream and makes emit()'s draws from it.  The results
// wait in the wave's slots in LDS (word-major: word j of slot s at buf[64 j + s]; the
// caller keeps a sixth word per slot): the direction, then two words A, B that hold the
// emission delays or the block's words the emission did not draw (when a delay is off,
// its draw is not made):
//   decay and rise: A = decay delay, B = rise delay   (4 draws)
//   decay only:     A = decay delay, B = word 3       (3 draws)
//   rise only:      A = rise delay,  B = word 3       (3 draws)
//   neither:        A = word 2,      B = word 3       (2 draws)
// The lane that takes the photon restores its stream where emit() would have left it:
// block 1 next, the undrawn words in their ring slots.  Directions and times are the same
// operations on the same words, so the photon is the one emit() makes.
constexpr uint32_t EMB_SLOTS = 64;
constexpr uint32_t EMB_HDR = 6 * EMB_SLOTS;               // header words after the slots
constexpr uint32_t EMB_WAVE_BYTES = (EMB_HDR + 16) * 4;   // a wave's slots and header

__device__ __forceinline__ void emit_batch_slot(const KParams &P, uint64_t pid, uint32_t *buf,
                                                uint32_t s)
{
    const uint4 w = philox10((uint32_t)pid, (uint32_t)(pid >> 32), 0u, 0u, P.key0, P.key1);
    const float z = 1.0f - 2.0f * unit_of(w.x);
    const float rr = f_sqrt(jmax0(1.0f - z * z));
    const float phi = TWO_PI_F * unit_of(w.y);
    float sp, cp;
    f_sincos(phi, sp, cp);
    uint32_t A = w.z, Bw = w.w;
    const float tau_d = KP(tau_d), tau_r = KP(tau_r);
    if (tau_d > 0.0f) {
        A = __float_as_uint(-tau_d * f_log(unit_of(w.z)));
        if (tau_r > 0.0f) Bw = __float_as_uint(-tau_r * f_log(unit_of(w.w)));
    } else if (tau_r > 0.0f) {
        A = __float_as_uint(-tau_r * f_log(unit_of(w.z)));
    }
    buf[s] = __float_as_uint(rr * cp);
    buf[EMB_SLOTS + s] = __float_as_uint(rr * sp);
    buf[2 * EMB_SLOTS + s] = __float_as_uint(z);
    buf[3 * EMB_SLOTS + s] = A;
    buf[4 * EMB_SLOTS + s] = Bw;
}

__device__ __forceinline__ void emit_take(Rng &r, const KParams &P, uint64_t pid,
                                          const uint32_t *buf, uint32_t s, float t0,
                                          float &ux, float &uy, float &uz, float &t)
{
    ux = __uint_as_float(buf[s]);
    uy = __uint_as_float(buf[EMB_SLOTS + s]);
    uz = __uint_as_float(buf[2 * EMB_SLOTS + s]);
    const uint32_t A = buf[3 * EMB_SLOTS + s], Bw = buf[4 * EMB_SLOTS + s];
    const bool dec = KP(tau_d) > 0.0f, ris = KP(tau_r) > 0.0f;
    const uint32_t m = 2u + (dec ? 1u : 0u) + (ris ? 1u : 0u);
    rng_init(r, pid);
    r.blk = 1u;
    r.rd = m * RING_SLOT;
    r.wr = 4u * RING_SLOT;
    if (m <= 3u) *(uint32_t *)(r.ring + ring_addr(r, 3u * RING_SLOT)) = Bw;
    if (m == 2u) *(uint32_t *)(r.ring + ring_addr(r, 2u * RING_SLOT)) = A;
    t = t0;
    if (dec) t += __uint_as_float(A);
    if (ris) t += __uint_as_float(dec ? Bw : A);
}

struct Photon {
    float x, y, z, ux, uy, uz, t;
    int32_t bounces;
};

// One iteration of the bounce loop (src/kernel.jl:185-298).  Returns the photon's
// status after the step (ST_FLYING to continue); `lin` gets the detected SiPM index.
// L0, L1, L2: the crystal's sides (the launch's box, or the photon's own box when
// deposits of several crystals share one queue).
template <bool COUNT, uint32_t OPT>
__device__ __forceinline__ int step(Photon &p, Rng &r, const KParams &P, OpCount<COUNT> &oc,
                                    int &lin, int &nscat, const float L0, const float L1,
                                    const float L2)
{
    oc.add(3);
    // wall_hit (src/geometry.jl:19-32), ties x -> y -> z: (L - x) / u for u > 0, -x / u for
    // u < 0, Inf for u == 0.  The numerator is selected first so that each axis costs one
    // division for all lanes, not one per sign branch (a division by zero is discarded).
    const float inf = __builtin_inff();
    const float qx = (p.ux > 0.0f ? L0 - p.x : -p.x) / p.ux;
    const float qy = (p.uy > 0.0f ? L1 - p.y : -p.y) / p.uy;
    const float qz = (p.uz > 0.0f ? L2 - p.z : -p.z) / p.uz;
    const float dx = p.ux != 0.0f ? qx : inf;
    const float dy = p.uy != 0.0f ? qy : inf;
    const float dz = p.uz != 0.0f ? qz : inf;
    float d;
    int face;
    if (dx <= dy && dx <= dz) { d = dx; face = p.ux > 0.0f ? F_XP : F_XM; }
    else if (dy <= dz)        { d = dy; face = p.uy > 0.0f ? F_YP : F_YM; }
    else                      { d = dz; face = p.uz > 0.0f ? F_BACK : F_FRONT; }

    // bulk absorption (src/kernel.jl:187-197): absorbed when d_abs = -abs_len log(u)
    // satisfies d_abs < d and d_abs <= d_ray.  d_abs enters nothing but this decision, so
    // it is taken on the estimate -abs_len ln2 log2_est(u) (relative error below 2^-21 on
    // every draw value, pinned by the tests) unless the estimate lies within ABS_BAND of
    // m = min(d, d_ray); there the exact d_abs decides.  The outcome is the exact one's.
    float u_abs = 1.0f, d_ray = inf;
    // the step's first draws (at most six: absorption, Rayleigh distance, the first
    // micro-facet try and its Fresnel draw) are within the eight the top of the step holds
    if (has<OPT>(P, K_ABS)) u_abs = randu<COUNT, false>(r, P, oc);
    if (has<OPT>(P, K_RAY)) d_ray = -P.ray_len * f_log(randu<COUNT, false>(r, P, oc));
    if (has<OPT>(P, K_ABS)) {
        const float m = d_ray < d ? d_ray : d;
        const float est = P.neg_abs_ln2 * f_log2_est(u_abs);
        bool absorbed;
        if (est * (1.0f + ABS_BAND) < m) absorbed = true;
        else if (est * (1.0f - ABS_BAND) > m) absorbed = false;
        else {
            const float d_abs = -P.abs_len * f_log(u_abs);
            absorbed = d_abs < d && d_abs <= d_ray;
        }
        if (absorbed) return ST_ABS_BULK;
    }
    if (d_ray < d) {
        p.x += d_ray * p.ux; p.y += d_ray * p.uy; p.z += d_ray * p.uz;
        p.t += d_ray * P.inv_v;
        rayleigh<COUNT>(r, P, oc, p.ux, p.uy, p.uz);
        ++nscat;
        oc.add(9);
        ++p.bounces;
        return p.bounces >= P.maxb ? ST_CAP : ST_FLYING;
    }
    // advance to the wall, snap onto it, clamp the rest (src/kernel.jl:206-218)
    p.x = jclamp(p.x + d * p.ux, 0.0f, L0);
    p.y = jclamp(p.y + d * p.uy, 0.0f, L1);
    p.z = jclamp(p.z + d * p.uz, 0.0f, L2);
    const int a = face_axis(face);
    if (a == 1)      p.x = face == F_XP ? L0 : 0.0f;
    else if (a == 2) p.y = face == F_YP ? L1 : 0.0f;
    else             p.z = face == F_BACK ? L2 : 0.0f;
    p.t += d * P.inv_v;
    const float cosi = a == 1 ? fabsf(p.ux) : (a == 2 ? fabsf(p.uy) : fabsf(p.uz));

    int result = ST_FLYING;
    const bool readout_face = face == F_BACK || (has<OPT>(P, K_TWO_SIDED) && face == F_FRONT);
    // After the optics of the face, at most one more draw decides the photon: the PDE
    // (readout face, inside the active area), the surround reflectivity (readout face,
    // outside it) or the wrap reflectivity (a wrapped face the photon reached).  The three
    // are one draw site with a per-lane threshold.
    int kind = 0;                                   // 1 PDE, 2 surround, 3 wrap
    float thr = 0.0f;
    if (!readout_face && !has<OPT>(P, K_AIR_GAP)) {
        // front-painted wrap in optical contact (src/kernel.jl:252-262)
        kind = 3;
        thr = P.wrap_R;
    } else {
        // Readout face: crystal -> coupling Fresnel against the face normal, then PDE
        // (src/kernel.jl:220-251).  Back-painted wrap: crystal -> air Fresnel, against a
        // sampled micro-facet normal when the surface is rough, then the reflector
        // (src/kernel.jl:263-293).  Both Fresnel evaluations (and the first UNIFIED try's)
        // run as one, with per-lane indices and cosine; each lane's draws keep their order.
        const float sgn = face_sign(face);
        const bool rough = !readout_face && has<OPT>(P, K_ROUGH);
        float cf = cosi, nfx = 0.0f, nfy = 0.0f, nfz = 0.0f;
        bool ok = true;                              // the (first) try reaches Fresnel
        if (rough) ok = facet_try<COUNT, false>(r, P, oc, p.ux, p.uy, p.uz, a, sgn, nfx, nfy, nfz, cf);
        bool refl = false;
        if (ok) {
            oc.add(4);
            refl = fresnel_reflects<COUNT, false>(r, P, oc, readout_face ? P.q_coup : P.q_air,
                                           P.n_crystal, readout_face ? P.n_coupling : 1.0f, cf);
        }
        if (readout_face) {
            if (refl) {
                specular(p.ux, p.uy, p.uz, face);
            } else {
                const float ex = p.x - P.cx, ey = p.y - P.cy;
                const bool inside = !has<OPT>(P, K_DISC) || ex * ex + ey * ey <= P.r2;
                kind = inside ? 1 : 2;
                thr = inside ? P.pde : P.sur_R;
            }
        } else {
            bool transmitted;
            if (!rough) {
                transmitted = !refl;
                if (refl) specular(p.ux, p.uy, p.uz, face);
            } else if (ok && !refl) {
                transmitted = true;
            } else if (ok && facet_reflect(p.ux, p.uy, p.uz, cf, nfx, nfy, nfz, a, sgn)) {
                transmitted = false;
            } else {
                transmitted = unified_surface<COUNT>(r, P, oc, p.ux, p.uy, p.uz, face, 1);
            }
            if (transmitted) {
                kind = 3;
                thr = P.wrap_R;
            }
        }
    }
    if (kind != 0) {
        const bool pass = randu<COUNT>(r, P, oc) <= thr;
        if (kind == 1) {
            if (!pass) return ST_ABS_SIPM;
            const float pitch = KP(pitch);
            const int nx = KP(nx), ny = KP(ny);
            int ii = 1 + (int)(p.x / pitch);
            int jj = 1 + (int)(p.y / pitch);
            ii = ii < nx ? ii : nx;
            jj = jj < ny ? jj : ny;
            const int l = (jj - 1) * nx + ii;
            lin = face == F_FRONT ? -l : l;
            oc.add(8);
            return ST_DETECTED;
        }
        if (!pass) return kind == 2 ? ST_ABS_SUR : ST_ABS_WALL;
        const bool spec = kind == 2 ? has<OPT>(P, K_SUR_SPEC) : has<OPT>(P, K_WRAP_SPEC);
        if (spec) specular(p.ux, p.uy, p.uz, face);
        else lambertian<COUNT>(r, P, oc, face, p.ux, p.uy, p.uz);
    }
    ++p.bounces;
    return p.bounces >= P.maxb ? ST_CAP : result;
}

// Order-preserving u32 key of a float, so atomicMin on keys is min on floats.
__device__ __forceinline__ uint32_t time_key(float t)
{
    const uint32_t b = __float_as_uint(t);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ __forceinline__ float key_time(uint32_t k)
{
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    union { uint32_t u; float f; } v;
    v.u = b;
    return v.f;
}
constexpr uint32_t KEY_INF = 0xFF800000u;     // time_key(+Inf)

}  // namespace crysp
