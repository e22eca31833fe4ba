// transport.hip -- gfx950 kernels of the optical-photon transport and the C-ABI
// (include/crysplight.h) of libcrysplight_hip.so.
//
// Kernel design (one persistent kernel, three output modes):
//  * Photons are numbered k = 0..n-1 inside a call.  A wave grabs CHUNK consecutive
//    photon numbers at a time from one device cursor (one atomic per chunk) and hands
//    them to its lanes as they free up: a lane whose photon ended takes a new photon
//    at the top of the next step, so no lane idles behind the wave's longest photon.
//  * Each loop iteration is ONE bounce step (crysp::step) for every live lane.
//  * Records mode writes the reference's four per-photon records
//    (src/kernel.jl:300-303).  Accumulate mode histograms detections in LDS per
//    workgroup (counts + first-time keys) and flushes once with global atomics;
//    counters are reduced per wave.  Event mode adds each detection to its event's
//    Int32 map with a global atomic and first-hit time with atomicMin on an
//    order-preserving key.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "crysp_device.h"
#include "crysplight.h"

using namespace crysp;

namespace {

enum Mode : int { M_RECORDS = 0, M_ACCUM = 1, M_EVENTS = 2 };

struct KBuffers {
    const float *px, *py, *pz, *pt;                 // records / accumulate inputs
    const float4 *dep;                              // events: (x, y, z, t0) per deposit
    const int64_t *dep_off;                         // events: ndep+1 photon offsets
    const int32_t *dep_event;                       // events: event of each deposit
    uint8_t *status; int16_t *idx; float *tdet; int32_t *bnc;   // records
    uint32_t *counts; uint32_t *tkeys; unsigned long long *counters;  // accumulate
    int32_t *maps; uint32_t *ev_tkeys;              // events
    unsigned long long *ev_tsum;                    // events, optional: sum of (t - t0) in 2^-10 ns
    const float4 *boxes;                            // events, several crystals: (Lx, Ly, Lz, -)
    const int32_t *dep_box;                         // events, several crystals: box of a deposit
    unsigned long long *cursor;                     // work-queue cursor (zeroed); the event
                                                    // cursor when workgroups own events
    unsigned long long *opcounts;                   // 10 op counters (COUNT only)
    unsigned int *flag;                             // events: input error flags
    // events owned by workgroups
    unsigned long long *ev_next;                    // per event: its next unclaimed photon
    const unsigned long long *ev_end;               // per event: one past its last photon
    uint32_t *ev_head;                              // per event: first partial map + 1, 0 none
    uint32_t *pool_link;                            // per partial: the next one + 1, 0 none
    uint32_t *pool_data;                            // per partial: nxy counts, nxy time keys
    uint32_t *sess_ctr;                             // [0] partials used, [1] events written
    const uint32_t *ev_order;                       // events in the order owners claim them
    uint32_t *open_ev;                              // per workgroup: the event it owns and has
                                                    // not yet claimed to the end, or ~0
    float *tmin_out;                                // tmin (f32), written by the event's owner
    int32_t *npe_out;
};

// v, which the compiler cannot see through (a value derived from it is recomputed where
// it is used instead of being kept in a register)
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t first_lane_u64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// deposit holding photon g: largest d with dep_off[d] <= g
__device__ __forceinline__ int64_t find_deposit(const int64_t *off, int64_t ndep, int64_t g)
{
    int64_t lo = 0, hi = ndep;                      // invariant off[lo] <= g < off[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// ---- events owned by workgroups ---------------------------------------------------
// With the deposits ordered by event, each event's photons form one contiguous range.  A
// workgroup opens a session on an event (the next one from the event cursor), claims the
// event's photons `piece` at a time from that event's own cursor and hands them to its
// waves, which add detections to the session's map in LDS.  Once the event's cursor has
// run out (the session is closed) and the last of its photons has ended, one wave writes
// the event's map row, tmin and npe with plain stores: the event has one writer, so no
// global atomic touches the maps.  When every event has an owner, workgroups left without
// work help: they open sessions on events with photons still unclaimed, and their sums go
// to partial maps that finish_sessions adds in afterwards.  Counts are sums and first
// times minima, so the maps equal those of any other order of the same detections.
constexpr int NSESS = 4;                            // session slots per workgroup
constexpr int SESS_MAX_CELLS = 256;                 // SiPMs a session map holds
constexpr uint32_t SS_CLOSED = 0x80000000u;         // the event's photons are all claimed
constexpr uint32_t SS_FREE = 0x40000000u;           // the slot holds no session
constexpr uint32_t SS_COUNT = 0x3FFFFFFFu;          // photons claimed and not yet ended
constexpr uint32_t SS_FLUSHING = SS_CLOSED | SS_FREE;  // claimed by the wave writing it out
constexpr uint32_t CUR_NONE = 0xFFu, CUR_DONE = 0xFEu;
constexpr uint32_t FLAG_UNSORTED = 8u, FLAG_INTERNAL = 16u;
// words of the session area in LDS: lock, current slot, the workgroup's claimed photons
// [wnext, wend) of slot wslot, a spare word, then per slot its state, event and
// partial (-1 for the event's owner), then per slot nxy counts and nxy time keys
enum : int { SH_LOCK = 0, SH_CUR = 1, SH_WNEXT = 2, SH_WEND = 4, SH_WSLOT = 6, SH_SPARE = 7,
             SH_STATE = 8, SH_EV = 8 + NSESS, SH_PART = 8 + 2 * NSESS, SH_MAPS = 32 };
static_assert(SH_PART + NSESS <= SH_MAPS, "session header overlaps the maps");

__host__ __device__ constexpr size_t sess_lds_words(int nxy)
{
    return (size_t)SH_MAPS + (size_t)NSESS * 2 * (size_t)nxy;
}

__device__ __forceinline__ uint32_t lds_acquire(uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_release(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS words the whole wave reads at one address, made wave-uniform (scalar registers)
__device__ __forceinline__ uint32_t lds_uni(const uint32_t *p)
{
    return __builtin_amdgcn_readfirstlane(*p);
}

__device__ __forceinline__ uint64_t lds_u64(const uint32_t *p)
{
    return (uint64_t)lds_uni(p) | ((uint64_t)lds_uni(p + 1) << 32);
}

__device__ __forceinline__ void lds_set_u64(uint32_t *p, uint64_t v)
{
    p[0] = (uint32_t)v;
    p[1] = (uint32_t)(v >> 32);
}

// A field of the transport kernel's KBuffers (its second argument, right after KParams in
// the kernel argument segment) read again at each use, as KP reads KParams: for code that
// runs rarely inside transport_kernel.
static_assert(sizeof(KParams) % alignof(KBuffers) == 0, "KBuffers follows KParams directly");
#define KB(f) karg_at<decltype(KBuffers::f)>(sizeof(KParams) + offsetof(KBuffers, f))

// Session q's sums written out by one wave (every lane; q wave-uniform), the slot freed.
__device__ void sess_flush(uint32_t *S, int q, int nxy, const KBuffers &B)
{
    const int64_t e = (int32_t)lds_uni(&S[SH_EV + q]);
    const int32_t part = (int32_t)lds_uni(&S[SH_PART + q]);
    uint32_t *m = S + SH_MAPS + q * 2 * nxy;
    // destinations: the event's rows (owner) or the partial map (helper)
    uint32_t *dc = part < 0 ? (uint32_t *)KB(maps) + e * nxy : KB(pool_data) + (int64_t)part * 2 * nxy;
    uint32_t *dk = part < 0 ? (uint32_t *)KB(tmin_out) + e * nxy : dc + nxy;
    uint32_t sum = 0u;
#pragma unroll 1
    for (int c = (int)__lane_id(); c < nxy; c += 64) {
        const uint32_t cnt = m[c], key = m[nxy + c];
        m[c] = 0u;
        m[nxy + c] = KEY_INF;
        dc[c] = cnt;
        dk[c] = part < 0 ? __float_as_uint(key_time(key)) : key;
        sum += cnt;
    }
#pragma unroll
    for (int m2 = 32; m2 >= 1; m2 >>= 1) sum += __shfl_xor(sum, m2);
    if (__lane_id() == 0) {
        if (part < 0) {
            KB(npe_out)[e] = (int32_t)sum;
            atomicAdd(&KB(sess_ctr)[1], 1u);
        } else if (sum) {
            // the partial is read by finish_sessions, after this kernel: no fence needed
            KB(pool_link)[part] = atomicExch(&KB(ev_head)[e], (uint32_t)part + 1u);
        }
        lds_release(&S[SH_STATE + q], SS_FREE);
    }
}

// An event a helper can join: one that its owner has not yet claimed to the end (the
// owners publish their current event) with at least two pieces of photons unclaimed, so
// that a helper's partial map pays for itself.  The
// scan starts past this workgroup's own entry, four entries per lane at a time; -1 when no
// owner has such an event.  Wave-uniform.
__device__ int64_t find_help(const KParams &P, const KBuffers &B)
{
    const uint32_t nw = gridDim.x;
    const uint32_t *open = KB(open_ev);
    for (uint32_t o = 0; o < nw; o += 256) {
        uint32_t evs[4];
        uint64_t m[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t k = o + (uint32_t)__lane_id() + 64u * i;
            evs[i] = k < nw ? __hip_atomic_load(&open[(blockIdx.x + 1u + k) % nw], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : ~0u;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) m[i] = __ballot(evs[i] < (uint32_t)KP(nev));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint64_t mm = m[i];
            while (mm) {
                const int l = (int)__builtin_ctzll(mm);
                const int64_t e = (int64_t)__builtin_amdgcn_readlane(evs[i], l);
                const unsigned long long nx =
                    __hip_atomic_load(&KB(ev_next)[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (nx + 2ull * (unsigned long long)KP(piece) <= KB(ev_end)[e]) return e;
                mm &= mm - 1;
            }
        }
    }
    return -1;
}

// Sessions that are closed with no photon in flight, each claimed (compare-and-swap) by
// this wave to write out; a bit mask of slots.  Every wave that ends photons looks here
// after reporting them (or its next take does), so the report that takes a closed session
// to zero is followed by a scan that finds it.  Wave-uniform.
__device__ uint32_t sess_scan(uint32_t *S)
{
    uint32_t mine = 0u;
#pragma unroll
    for (int q = 0; q < NSESS; ++q) {
        if (lds_uni(&S[SH_STATE + q]) != SS_CLOSED) continue;
        uint32_t won = 0u;
        if (__lane_id() == 0) {
            uint32_t expect = SS_CLOSED;
            won = __hip_atomic_compare_exchange_strong(&S[SH_STATE + q], &expect, SS_FLUSHING,
                                                       __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP) ? 1u : 0u;
        }
        if (__builtin_amdgcn_readfirstlane(won)) mine |= 1u << q;
    }
    return mine;
}

// A wave's next photons from its workgroup's sessions, one action per call: take up to a
// chunk of the photons the workgroup has claimed ([b, e) of slot q); or claim the next
// piece of the current session's event (closing the session when the event runs out); or
// open a session on a new event (an owner) or on one with photons left (a helper); or
// find there is no more work (done).  b == e when this call took no photons (also when
// another wave holds the lock).  Sessions found closed with no photon in flight are
// claimed into flush for the caller to write out.  Wave-uniform.
__device__ void sess_take(uint32_t *S, const KParams &P, const KBuffers &B, uint64_t &b,
                          uint64_t &e, int &q, bool &done, uint32_t &flush)
{
    const bool leader = __lane_id() == 0;
    uint32_t locked = 0u;
    if (leader) {
        uint32_t expect = 0u;
        locked = __hip_atomic_compare_exchange_strong(&S[SH_LOCK], &expect, 1u, __ATOMIC_ACQUIRE,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                     ? 1u : 0u;
    }
    if (!__builtin_amdgcn_readfirstlane(locked)) return;
    const uint64_t wn = lds_u64(S + SH_WNEXT), we = lds_u64(S + SH_WEND);
    const uint32_t cur = lds_uni(&S[SH_CUR]);
    if (wn < we) {                                  // photons the workgroup has claimed
        const uint64_t n = we - wn < (uint64_t)KP(chunk) ? we - wn : (uint64_t)KP(chunk);
        b = wn;
        e = wn + n;
        q = (int)lds_uni(&S[SH_WSLOT]);
        if (leader) lds_set_u64(S + SH_WNEXT, wn + n);
    } else if (cur == CUR_DONE) {
        done = true;
    } else if (cur < (uint32_t)NSESS) {             // the next piece of the session's event
        const int64_t ev = (int32_t)lds_uni(&S[SH_EV + cur]);
        const bool helper = (int32_t)lds_uni(&S[SH_PART + cur]) >= 0;
        // owners claim pieces; helpers, who share the event with its owner, half pieces
        const unsigned long long piece =
            (unsigned long long)(helper ? (KP(piece) / 2 > KP(chunk) ? KP(piece) / 2 : KP(chunk)) : KP(piece));
        unsigned long long g = 0;
        if (leader) g = atomicAdd(&KB(ev_next)[ev], piece);
        g = first_lane_u64(g);
        const unsigned long long gend = KB(ev_end)[ev];
        const unsigned long long n = g < gend ? (gend - g < piece ? gend - g : piece) : 0ull;
        const bool close = g + piece >= gend;
        // the claimed photons join the session's count (a claim that reached the event's end
        // closes it; the scan at the end of this call finds it if nothing is in flight)
        if (leader)
            __hip_atomic_fetch_add(&S[SH_STATE + cur], (uint32_t)n + (close ? SS_CLOSED : 0u),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (leader) {
            if (n) {
                const uint64_t n2 = n < (uint64_t)KP(chunk) ? n : (uint64_t)KP(chunk);
                lds_set_u64(S + SH_WNEXT, g + n2);
                lds_set_u64(S + SH_WEND, g + n);
                S[SH_WSLOT] = cur;
            }
            if (close) {
                S[SH_CUR] = CUR_NONE;
                if (!helper) KB(open_ev)[blockIdx.x] = ~0u;     // nothing left to help with
            }
        }
        if (n) {                                    // the wave takes the piece's first chunk
            b = g;
            e = g + (n < (uint64_t)KP(chunk) ? n : (uint64_t)KP(chunk));
            q = (int)cur;
        }
    } else {                                        // open a session in a free slot
        int fq = -1;
        for (int s = 0; s < NSESS; ++s)
            if (fq < 0 && __builtin_amdgcn_readfirstlane(lds_acquire(&S[SH_STATE + s])) == SS_FREE) fq = s;
        if (fq >= 0) {
            unsigned long long ec = 0;
            if (leader) ec = atomicAdd(KB(cursor), 1ull);
            ec = first_lane_u64(ec);
            int64_t ev = -1;
            if (ec < (unsigned long long)KP(nev))   // the event's owner
                ev = (int64_t)__builtin_amdgcn_readfirstlane(KB(ev_order)[ec]);
            int32_t part = -1;
            if (ec >= (unsigned long long)KP(nev)) {
                ev = find_help(P, B);
                if (ev >= 0) {
                    uint32_t pi = 0u;
                    if (leader) pi = atomicAdd(&KB(sess_ctr)[0], 1u);
                    pi = __builtin_amdgcn_readfirstlane(pi);
                    if (pi < (uint32_t)KP(pool_cap)) part = (int32_t)pi;
                    else ev = -1;
                }
            }
            if (leader) {
                if (ev < 0) {
                    S[SH_CUR] = CUR_DONE;
                } else {
                    S[SH_EV + fq] = (uint32_t)ev;
                    S[SH_PART + fq] = (uint32_t)part;
                    S[SH_STATE + fq] = 0u;
                    S[SH_CUR] = (uint32_t)fq;
                    if (part < 0) KB(open_ev)[blockIdx.x] = (uint32_t)ev;     // helpers may join
                }
            }
            if (ev < 0) done = true;
        }
    }
    flush |= sess_scan(S);
    if (leader) lds_release(&S[SH_LOCK], 0u);
}

// Wave timeline (diagnostics): per wave of the last launch, the tick of the device wall
// clock at which it ended, when g_timeline_on is set.  Every context on a device shares them.
constexpr uint32_t TL_MAX_WAVES = 16384;
__device__ unsigned long long g_timeline[TL_MAX_WAVES];
__device__ int g_timeline_on;

// Seven waves per SIMD (at most 72 VGPRs): the waves cover the LDS reads of the Philox
// rings and each other's dependent binary64 chains.
template <int MODE, bool COUNT, bool MULTIBOX = false, uint32_t OPT = OPT_GENERIC,
          bool SESS = false>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7, 7)))
transport_kernel(const KParams P, const KBuffers B)
{
    // dynamic LDS: the lanes' Philox rings (blockDim.x * 64 B), then in accumulate mode
    // the histogram and time keys
    extern __shared__ uint32_t lds_dyn[];
    constexpr bool sess = MODE == M_EVENTS && SESS;
    // emission in batches (not in the counting variant or with sessions); the kernels
    // without sessions keep each wave's batch after the rings
    constexpr bool emb = !COUNT && !sess;
    uint32_t *lds = (uint32_t *)((char *)lds_dyn + 64u * blockDim.x +
                                 (sess ? 0u : (blockDim.x >> 6) * EMB_WAVE_BYTES));
    __shared__ unsigned long long lds_ctr[9];       // accumulate: the Accumulator counters
    const int nxy = P.nx * P.ny;
    if (MODE == M_EVENTS && P.gate) {
        // the device decides between the two event paths: ordered deposits or not
        const bool unsorted =
            (__hip_atomic_load(B.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_UNSORTED) != 0u;
        if (unsorted != (P.gate == 2)) return;
    }

    if (sess) {
        for (int w = threadIdx.x; w < (int)sess_lds_words(nxy); w += blockDim.x) {
            uint32_t v = 0u;
            if (w == SH_CUR) v = CUR_NONE;
            else if (w >= SH_STATE && w < SH_STATE + NSESS) v = SS_FREE;
            else if (w >= SH_MAPS && ((w - SH_MAPS) / nxy) % 2 == 1) v = KEY_INF;
            lds[w] = v;
        }
        __syncthreads();
    }
    if (MODE == M_ACCUM) {
        for (int w = threadIdx.x; w < P.hist_words + nxy; w += blockDim.x)
            lds[w] = w < P.hist_words ? 0u : KEY_INF;
        if (threadIdx.x < 9) lds_ctr[threadIdx.x] = 0ull;
        __syncthreads();
    }

    __shared__ unsigned long long lds_ops[10];
    OpCount<COUNT> oc{lds_ops};
    if (COUNT) {
        if (threadIdx.x < 10) lds_ops[threadIdx.x] = 0ull;
        __syncthreads();
    }

    Photon p{};
    Rng rng;
    rng_bind(rng, lds_dyn);
    rng_init(rng, 0);
    int64_t k = 0;                                  // photon number of this lane
    int32_t ev = 0;                                 // its event (event mode)
    float t_start = 0.0f;                           // its deposit time (event mode)
    float bL0 = P.L0, bL1 = P.L1, bL2 = P.L2;       // its crystal
    int64_t my_dep = 0;                             // deposit of this lane's last photon
    int nscat = 0;
    bool alive = false;

    uint64_t pool_next = 0, pool_end = 0;           // wave-uniform work pool
    int64_t pool_dep = 0;                           // deposit holding the pool chunk's start
    int pool_slot = 0;                              // sessions: the pool's session slot
    uint32_t flush_pending = 0u;                    // sessions: slots to write out (bits)
    uint64_t unreported = 0;                        // sessions: lanes whose photon ended and
                                                    // is not yet taken off its session's count
    uint32_t idle_iters = 0;                        // sessions: iterations of an idle wait
    bool exhausted = false;
    // event mode reads the photon total the offsets scan produced
    const uint64_t N = (MODE == M_EVENTS && P.n_photons < 0) ? (uint64_t)B.dep_off[P.ndep]
                                                             : (uint64_t)P.n_photons;
    const uint32_t wave_lanes = (uint32_t)__popcll(__ballot(true));
    // The wave's batch: EMB_SLOTS slots of six words (word-major: the emission's five
    // words, then the photon's deposit relative to the batch's first), then a header
    // (in 64-bit words): the photons of the last chunk not yet emitted start at
    // res_next (in deposit res_dep; res_left of them); the emitted batch starts at
    // photon eb_lo (in deposit eb_dep), holds cnt photons, of which pos are handed out.
    // Addressed at use from the ring's lane offset (wave w's ring starts at byte
    // 4096 w), through an empty asm so that no register holds the address across the
    // loop; the header lives in LDS for the same reason.
#define EMB_BUF ((uint32_t *)((char *)lds_dyn + 64u * blockDim.x + (opaque_u32(rng.base) >> 12) * EMB_WAVE_BYTES))
    enum { EH_RES_NEXT = 0, EH_RES_DEP = 2, EH_EB_LO = 4, EH_EB_DEP = 6, EH_RES_LEFT = 8,
           EH_POS = 9, EH_CNT = 10 };
    if (emb) {
        uint32_t *h = EMB_BUF + EMB_HDR;
        // every lane writes the same words: a lane reads back what it wrote itself
        lds_set_u64(h + EH_RES_NEXT, 0ull);
        lds_set_u64(h + EH_RES_DEP, 0ull);
        lds_set_u64(h + EH_EB_LO, 0ull);
        lds_set_u64(h + EH_EB_DEP, 0ull);
        h[EH_RES_LEFT] = 0u;
        h[EH_POS] = 0u;
        h[EH_CNT] = 0u;
    }
    // a lane takes batch slot `slot`
    auto emb_take = [&](uint32_t slot) {
        const uint32_t *b = EMB_BUF;
        k = (int64_t)(lds_u64(b + EMB_HDR + EH_EB_LO) + slot);
        float x0, y0, z0, t0;
        if (MODE == M_EVENTS) {
            const int64_t d = (int64_t)lds_u64(b + EMB_HDR + EH_EB_DEP) + b[5 * EMB_SLOTS + slot];
            const float4 q = KB(dep)[d];
            x0 = q.x; y0 = q.y; z0 = q.z; t0 = q.w;
            ev = KB(dep_event)[d];
            t_start = t0;
            if (MULTIBOX) {
                const int32_t bi = B.dep_box[d];
                const float4 bx = B.boxes[(uint32_t)bi < (uint32_t)P.nbox ? bi : 0];
                bL0 = bx.x; bL1 = bx.y; bL2 = bx.z;
            }
        } else {
            const int64_t j = P.npos == 1 ? 0 : k % (int64_t)P.npos;
            x0 = B.px[j]; y0 = B.py[j]; z0 = B.pz[j];
            t0 = B.pt ? B.pt[j] : 0.0f;
        }
        p.x = x0; p.y = y0; p.z = z0; p.bounces = 0;
        emit_take(rng, P, KP(pid0) + (uint64_t)k + 1ull, b, slot, t0, p.ux, p.uy, p.uz, p.t);
        nscat = 0;
        alive = true;
    };

    for (;;) {
        // ---- refill idle lanes, in batches of at least refill_min ------------------
        const uint64_t need = __ballot(!alive);
        const uint32_t nneed = (uint32_t)__popcll(need);
        const bool refill = !exhausted && nneed != 0u &&
                            (nneed >= (uint32_t)P.refill_min || nneed == wave_lanes);
        if (sess) {
            // ended photons are taken off their sessions' counts before their lanes take
            // new photons (or, once no more come, every step); a closed session whose count
            // reaches zero is written out by this wave
            // (a wave's LDS operations take effect in order, so its detections are in the
            // session maps before its report; the reports need no reply)
            if (unreported && (refill || exhausted)) {
                while (unreported) {
                    const int q = __builtin_amdgcn_readlane(ev, (int)__builtin_ctzll(unreported));
                    const uint64_t mq = __ballot(((unreported >> __lane_id()) & 1ull) && ev == q);
                    if (__lane_id() == 0)
                        __hip_atomic_fetch_sub(&lds[SH_STATE + q], (uint32_t)__popcll(mq),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    unreported &= ~mq;
                }
                if (exhausted) flush_pending |= sess_scan(lds);
            }
            while (flush_pending) {
                sess_flush(lds, (int)__builtin_ctz(flush_pending), nxy, B);
                flush_pending &= flush_pending - 1u;
            }
        }
        if (emb && refill) {
            // photons come from the wave's emitted batch; lanes past its end take from
            // the next batch, which the whole wave emits
            const uint32_t rank = lanes_below(need);
            uint32_t pos = lds_uni(EMB_BUF + EMB_HDR + EH_POS);
            const uint32_t avail = lds_uni(EMB_BUF + EMB_HDR + EH_CNT) - pos;
            if (!alive && rank < avail) emb_take(pos + rank);
            if (nneed > avail) {
                uint32_t *h = EMB_BUF + EMB_HDR;
                uint64_t rn = lds_u64(h + EH_RES_NEXT);
                int64_t rd = (int64_t)lds_u64(h + EH_RES_DEP);
                uint32_t left = lds_uni(h + EH_RES_LEFT);
                if (left == 0u) {                           // the next chunk of the queue
                    uint64_t b = 0;
                    const uint32_t chunk = (uint32_t)KP(chunk);
                    if (__lane_id() == __builtin_amdgcn_readfirstlane(__lane_id()))
                        b = atomicAdd(KB(cursor), (unsigned long long)chunk);
                    b = first_lane_u64(b);
                    if (b >= N) exhausted = true;
                    else {
                        rn = b;
                        left = N - b < (uint64_t)chunk ? (uint32_t)(N - b) : chunk;
                        if (MODE == M_EVENTS) rd = find_deposit(KB(dep_off), KP(ndep), (int64_t)b);
                    }
                }
                const uint32_t nb = left < EMB_SLOTS ? left : EMB_SLOTS;
                // the takes above read the old batch before the wave writes the new one,
                // and the takes below read slots other lanes wrote (a wave's LDS
                // operations complete in order; the fences keep the compiler's order)
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                uint32_t last = 0u;
                if (__lane_id() < nb) {
                    const uint64_t kk = rn + __lane_id();
                    emit_batch_slot(P, KP(pid0) + kk + 1ull, EMB_BUF, __lane_id());
                    if (MODE == M_EVENTS) {
                        // the deposit of photon kk: at or after the chunk's, a few at most
                        const int64_t *off = KB(dep_off);
                        int64_t d = rd;
                        while (off[d + 1] <= (int64_t)kk) ++d;
                        last = (uint32_t)(d - rd);
                        EMB_BUF[5 * EMB_SLOTS + __lane_id()] = last;
                    }
                }
                h = EMB_BUF + EMB_HDR;
                lds_set_u64(h + EH_EB_LO, rn);
                lds_set_u64(h + EH_EB_DEP, (uint64_t)rd);
                if (nb) {
                    if (MODE == M_EVENTS) rd += __builtin_amdgcn_readlane(last, (int)nb - 1);
                    rn += nb;
                }
                lds_set_u64(h + EH_RES_NEXT, rn);
                lds_set_u64(h + EH_RES_DEP, (uint64_t)rd);
                h[EH_RES_LEFT] = left - nb;
                h[EH_CNT] = nb;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const uint32_t r2 = rank - avail;
                if (!alive && rank >= avail && r2 < nb) emb_take(r2);
                const uint32_t want = nneed - avail;
                pos = want < nb ? want : nb;
            } else {
                pos += nneed;
            }
            EMB_BUF[EMB_HDR + EH_POS] = pos;
        } else if (refill) {
            const uint32_t rank = lanes_below(need);
            uint64_t p_next = pool_next, p_end = pool_end;
            int64_t p_dep = pool_dep;
            int p_slot = pool_slot;
            const uint64_t avail = p_end - p_next;
            uint64_t base2 = 0, end2 = 0;
            int64_t new_dep = 0;
            int new_slot = 0;
            if (sess && (uint64_t)nneed > avail) {
                bool done = false;
                sess_take(lds, P, B, base2, end2, new_slot, done, flush_pending);
                if (done) exhausted = true;
                if (end2 > base2) new_dep = find_deposit(B.dep_off, P.ndep, (int64_t)base2);
            } else if ((uint64_t)nneed > avail) {
                uint64_t b = 0;
                if (__lane_id() == __builtin_amdgcn_readfirstlane(__lane_id()))
                    b = atomicAdd(B.cursor, (unsigned long long)P.chunk);
                b = first_lane_u64(b);
                base2 = b;
                end2 = b + (uint64_t)P.chunk < N ? b + (uint64_t)P.chunk : N;
                if (b >= N) { base2 = end2 = N; exhausted = true; }
                else if (MODE == M_EVENTS) new_dep = find_deposit(B.dep_off, P.ndep, (int64_t)b);
            }
            bool take = false;
            int64_t start_dep = p_dep;
            int take_slot = p_slot;
            if (!alive) {
                if ((uint64_t)rank < avail) { k = (int64_t)(p_next + rank); take = true; }
                else {
                    const uint64_t r2 = base2 + ((uint64_t)rank - avail);
                    if (r2 < end2) { k = (int64_t)r2; take = true; start_dep = new_dep; take_slot = new_slot; }
                }
            }
            if ((uint64_t)nneed <= avail) p_next += nneed;
            else {
                const uint64_t nx2 = base2 + ((uint64_t)nneed - avail);
                p_next = nx2 < end2 ? nx2 : end2;
                p_end = end2;
                p_dep = new_dep;
                p_slot = new_slot;
            }
            pool_next = p_next;
            pool_end = p_end;
            pool_dep = p_dep;
            pool_slot = p_slot;
            if (take) {
                float x0, y0, z0, t0;
                if (MODE == M_EVENTS) {
                    // photons of a lane only move forward, and the chunk start bounds below
                    // (sessions: a chunk lies in one event, so its start is close enough)
                    int64_t d = sess ? start_dep : (start_dep > my_dep ? start_dep : my_dep);
                    while (B.dep_off[d + 1] <= k) ++d;
                    if (!sess) my_dep = d;
                    const float4 q = B.dep[d];
                    x0 = q.x; y0 = q.y; z0 = q.z; t0 = q.w;
                    ev = sess ? take_slot : B.dep_event[d];     // sessions: the map slot
                    t_start = t0;
                    if (MULTIBOX) {
                        // a crystal index outside the table reads crystal 0 (the check
                        // kernel has flagged the call as failed)
                        const int32_t bi = B.dep_box[d];
                        const float4 bx = B.boxes[(uint32_t)bi < (uint32_t)P.nbox ? bi : 0];
                        bL0 = bx.x; bL1 = bx.y; bL2 = bx.z;
                    }
                } else {
                    const int64_t j = P.npos == 1 ? 0 : k % (int64_t)P.npos;
                    x0 = B.px[j]; y0 = B.py[j]; z0 = B.pz[j];
                    t0 = B.pt ? B.pt[j] : 0.0f;
                }
                rng_start(rng, P.pid0 + (uint64_t)k + 1ull, P.key0, P.key1);
                p.x = x0; p.y = y0; p.z = z0; p.bounces = 0;
                emit<COUNT>(rng, P, oc, t0, p.ux, p.uy, p.uz, p.t);
                nscat = 0;
                alive = true;
                oc.add(0);
            }
        }
        if (!sess && __ballot(alive) == 0ull) break;    // queue drained, all lanes idle
        const bool idle = sess && __ballot(alive) == 0ull;

        // ---- one bounce step ------------------------------------------------------
        bool died = false;
        if (alive) {
            rng_prefetch(rng, P.key0, P.key1);      // the step's Philox block, one site
            int lin = 0;
            const int st = step<COUNT, OPT>(p, rng, P, oc, lin, nscat, bL0, bL1, bL2);
            if (st != ST_FLYING) {
                alive = false;
                died = true;
                if (MODE == M_RECORDS) {
                    if (B.status) B.status[k] = (uint8_t)st;
                    if (B.idx) B.idx[k] = (int16_t)(st == ST_DETECTED ? lin : 0);
                    if (B.tdet) B.tdet[k] = p.t;
                    if (B.bnc) B.bnc[k] = p.bounces;
                } else if (MODE == M_ACCUM) {
                    // counter slots: ndet 0, ndet_front 1, bulk 2, wall 3, sipm 4, cap 5,
                    // surround 6, nscat 7, sum_bounces_det 8 (the statuses map onto 0, 2..6)
                    if (nscat) atomicAdd(&lds_ctr[7], (unsigned long long)nscat);
                    if (st == ST_DETECTED) {
                        int l = lin;
                        if (l < 0) { atomicAdd(&lds_ctr[1], 1ull); l = -l; }
                        const int cell = l - 1;     // (ii-1) + nx*(jj-1)
                        const float q = p.t / P.bin_ns;
                        int bin = (q >= (float)P.nbins) ? P.nbins : (int)q;  // 0-based
                        bin = bin < 0 ? 0 : bin;
                        atomicAdd(&lds[cell + nxy * bin], 1u);
                        atomicMin(&lds[P.hist_words + cell], time_key(p.t));
                        atomicAdd(&lds_ctr[8], (unsigned long long)p.bounces);
                    }
                    const int slot = st == ST_DETECTED ? 0 : (st == ST_CAP ? 5 : (st == ST_ABS_SUR ? 6 : st));
                    atomicAdd(&lds_ctr[slot], 1ull);
                } else if (sess) {
                    if (st == ST_DETECTED) {
                        const int l = lin < 0 ? -lin : lin;
                        uint32_t *m = lds + SH_MAPS + ev * 2 * nxy;
                        atomicAdd(&m[l - 1], 1u);
                        atomicMin(&m[nxy + l - 1], time_key(p.t));
                    }
                } else {
                    if (st == ST_DETECTED && (uint32_t)ev < (uint32_t)KP(nev)) {
                        const int l = lin < 0 ? -lin : lin;
                        const int64_t cell = (int64_t)(l - 1) + (int64_t)(KP(nx) * KP(ny)) * ev;
                        atomicAdd(&KB(maps)[cell], 1);
                        atomicMin(&KB(ev_tkeys)[cell], time_key(p.t));
                        unsigned long long *tsum = KB(ev_tsum);
                        if (tsum) {
                            const float dt = p.t - t_start;
                            atomicAdd(&tsum[cell], (unsigned long long)(dt > 0.0f ? dt * 1024.0f : 0.0f));
                        }
                    }
                }
            }
        }
        if (sess) {
            unreported |= __ballot(died);
            if (idle) {
                // no more work and all lanes idle: done once every ended photon is reported
                // and every session this wave closed is written out
                if (exhausted && !unreported && !flush_pending) {
                    // a last look: a session this wave's reports took to zero
                    const uint32_t last = sess_scan(lds);
                    if (!last) break;
                    flush_pending |= last;
                    continue;
                }
                if (exhausted) continue;
                // another wave holds the lock, or every slot is draining; a wait that
                // outlasts the bound (2^27 naps of about 500 cycles, a minute) ends the
                // wave with the internal-error flag set
                if (++idle_iters > (1u << 27)) {
                    if (__lane_id() == 0) atomicOr(B.flag, FLAG_INTERNAL);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            } else {
                idle_iters = 0;
            }
        }
    }

    // wave timeline (diagnostics): this wave's end on the device wall clock (a file-scope
    // array: its address is a constant, so nothing stays live across the photon loop)
    if (g_timeline_on && __lane_id() == 0) {
        const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (opaque_u32(rng.base) >> 12);
        if (w < TL_MAX_WAVES) g_timeline[w] = wall_clock64();
    }
    if (COUNT) {
        __syncthreads();
        if (threadIdx.x < 10 && lds_ops[threadIdx.x]) atomicAdd(&B.opcounts[threadIdx.x], lds_ops[threadIdx.x]);
    }
    if (MODE == M_ACCUM) {
        __syncthreads();
        if (threadIdx.x < 9 && lds_ctr[threadIdx.x])
            atomicAdd(&B.counters[threadIdx.x], lds_ctr[threadIdx.x]);
        for (int w = threadIdx.x; w < P.hist_words; w += blockDim.x)
            if (lds[w]) atomicAdd(&B.counts[w], lds[w]);
        for (int w = threadIdx.x; w < nxy; w += blockDim.x)
            if (lds[P.hist_words + w] != KEY_INF) atomicMin(&B.tkeys[w], lds[P.hist_words + w]);
    }
}
#undef EMB_BUF

// -------------------------------------------------------------- small kernels ----
__global__ void fill_u32(uint32_t *p, uint32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void keys_to_times(const uint32_t *keys, float *t, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        t[i] = key_time(keys[i]);
}

// A fill that runs only when the device flag's ordering bit says the path it serves is the
// one taken: want_unsorted 1 for the per-detection path, 0 for owned events.
__global__ void fill_u32_gated(uint32_t *p, uint32_t v, int64_t n, const unsigned int *flag,
                               int want_unsorted)
{
    if (((*flag & FLAG_UNSORTED) != 0u) != (want_unsorted != 0)) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// first deposit of event >= e in an event-ordered table
__device__ __forceinline__ int64_t lower_bound_event(const int32_t *ev, int64_t ndep, int64_t e)
{
    int64_t lo = 0, hi = ndep;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)ev[mid] < e) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Photon ranges of the events of an event-ordered deposit table, and the partial-map list
// heads cleared.  Skipped when the device found the table out of order (the other path).
__global__ void event_ranges(const int32_t *dep_event, const int64_t *off, int64_t ndep,
                             int64_t nev, unsigned long long *ev_next, unsigned long long *ev_end,
                             uint32_t *head, const unsigned int *flag, int gated)
{
    if (gated && (*flag & FLAG_UNSORTED)) return;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nev;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lo = lower_bound_event(dep_event, ndep, e);
        const int64_t hi = lower_bound_event(dep_event, ndep, e + 1);
        ev_next[e] = (unsigned long long)off[lo];
        ev_end[e] = (unsigned long long)off[hi];
        head[e] = 0u;
    }
}

// The order in which owners claim events: largest first (by photon count, in powers of
// two), so that the events still running when the last one has an owner are small ones
// and the launch ends with little left to split among helpers.  One workgroup; a counting
// sort over 65 size classes.  Skipped with the other path.
__global__ void __launch_bounds__(1024) order_events(const unsigned long long *ev_next,
                                                     const unsigned long long *ev_end, int64_t nev,
                                                     uint32_t *order, const unsigned int *flag,
                                                     int gated)
{
    if (gated && (*flag & FLAG_UNSORTED)) return;
    __shared__ uint32_t cnt[65];
    auto size_class = [&](int64_t e) {
        const unsigned long long n = ev_end[e] - ev_next[e];
        return n ? 64 - __clzll((long long)n) : 0;
    };
    for (int i = threadIdx.x; i < 65; i += blockDim.x) cnt[i] = 0u;
    __syncthreads();
    for (int64_t e = threadIdx.x; e < nev; e += blockDim.x) atomicAdd(&cnt[size_class(e)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0u;
        for (int b = 64; b >= 0; --b) { const uint32_t c = cnt[b]; cnt[b] = acc; acc += c; }
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < nev; e += blockDim.x)
        order[atomicAdd(&cnt[size_class(e)], 1u)] = (uint32_t)e;
}

// Owned events: the helpers' partial maps added to the owners' rows (counts summed, first
// times the minimum, npe recomputed), one wave per event, and a check that every event
// was written.
__global__ void finish_sessions(const uint32_t *head, const uint32_t *link, const uint32_t *pdata,
                                int32_t *maps, float *tmin, int32_t *npe, int64_t nev, int nxy,
                                const uint32_t *sess_ctr, unsigned int *flag, int gated)
{
    if (gated && (*flag & FLAG_UNSORTED)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && (int64_t)sess_ctr[1] != nev)
        atomicOr(flag, FLAG_INTERNAL);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t e = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; e < nev; e += waves) {
        const uint32_t h = head[e];
        if (!h) continue;
        int32_t s = 0;
        for (int c = (int)__lane_id(); c < nxy; c += 64) {
            int32_t cnt = maps[e * nxy + c];
            uint32_t key = time_key(tmin[e * nxy + c]);
            for (uint32_t p = h; p; p = link[p - 1]) {
                const uint32_t *pd = pdata + (int64_t)(p - 1) * 2 * nxy;
                cnt += (int32_t)pd[c];
                key = pd[nxy + c] < key ? pd[nxy + c] : key;
            }
            maps[e * nxy + c] = cnt;
            tmin[e * nxy + c] = key_time(key);
            s += cnt;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
        if (__lane_id() == 0) npe[e] = s;
    }
}

// per event: npe = sum of its map (src/events.jl:86), tmin from keys
__global__ void finish_events(const int32_t *maps, const uint32_t *keys, float *tmin,
                              int32_t *npe, int64_t nev, int nxy, const unsigned int *flag,
                              int gated)
{
    // one wave per event: lane c takes cells c, c + 64, ... (coalesced), and npe is the sum
    // of the row's counts (integers: the order of the additions does not matter)
    if (gated && !(*flag & FLAG_UNSORTED)) return;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t e = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; e < nev; e += nw) {
        int32_t s = 0;
        for (int c = lane; c < nxy; c += 64) {
            const int64_t i = e * nxy + c;
            s += maps[i];
            tmin[i] = key_time(keys[i]);
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
        if (lane == 0) npe[e] = s;
    }
}

// The set-up of an event launch, in one kernel.  Workgroup 0 owns the flag word and stores
// it once, at its end: it clears the queue cursor (and the op counters), scans the
// device-resident photon counts into offsets (off[ndep] = the total; bit 1: a negative
// count, taken as 0) and checks device-resident index tables (bit 2: an event outside
// [0, nev); bit 4: a crystal outside [0, nbox); bit 8, not an error: events out of order,
// so they are not contiguous photon ranges).  The transport kernel guards both table reads
// itself; the bits turn a bad table into an error for the caller.  The other workgroups
// clear the output cells (counts 0, first-time keys +inf) when the launch fills them here,
// and one further range (the time sums).
struct PrepArgs {
    unsigned long long *cursor;
    unsigned int *flag;                             // 4 words, all rewritten
    const int32_t *nph;                             // device counts, or null (offsets given)
    int64_t *off;                                   // [ndep + 1]
    const int32_t *ev, *box;                        // tables to check, or null
    int64_t ndep, nev;
    int32_t nbox;
    unsigned long long *opc;                        // op counters [10], or null
    uint32_t *maps, *keys;                          // [cells] each, or null
    int64_t cells;
    uint32_t *zero;                                 // [nzero] to clear, or null
    int64_t nzero;
};

// deposits per thread and pass of workgroup 0: its loads are issued together, and each
// thread's run stays in registers between the sum and the offsets
constexpr int PREP_RUN = 16;

__global__ void __launch_bounds__(1024) prepare_events(const PrepArgs a)
{
    const int tid = threadIdx.x;
    if (blockIdx.x > 0) {
        const int64_t nt = (int64_t)(gridDim.x - 1) * blockDim.x;
        const int64_t i0 = (int64_t)(blockIdx.x - 1) * blockDim.x + tid;
        if (a.maps)
            for (int64_t i = i0; i < a.cells; i += nt) { a.maps[i] = 0u; a.keys[i] = KEY_INF; }
        if (a.zero)
            for (int64_t i = i0; i < a.nzero; i += nt) a.zero[i] = 0u;
        return;
    }
    __shared__ int64_t wsum[16];
    __shared__ unsigned int sbad;
    if (tid == 0) sbad = 0u;
    __syncthreads();                                // before any thread's bits reach it
    const int lane = tid & 63, wave = tid >> 6, nwv = blockDim.x >> 6;
    unsigned bad = 0u;
    int64_t carry = 0;
    for (int64_t base = 0; base < a.ndep; base += (int64_t)blockDim.x * PREP_RUN) {
        const int64_t lo = base + (int64_t)tid * PREP_RUN;
        // every load of a run is predicated (no early exit), so they are issued together;
        // past the table's end an event reads as the largest index and a crystal as 0
        if (a.ev) {
            int32_t e[PREP_RUN];
#pragma unroll
            for (int j = 0; j < PREP_RUN; ++j) e[j] = lo + j < a.ndep ? a.ev[lo + j] : INT32_MAX;
            int32_t prev = lo > 0 && lo <= a.ndep ? a.ev[lo - 1] : INT32_MIN;
#pragma unroll
            for (int j = 0; j < PREP_RUN; ++j) {
                if (lo + j < a.ndep && (e[j] < 0 || (int64_t)e[j] >= a.nev)) bad |= 2u;
                if (prev > e[j]) bad |= FLAG_UNSORTED;
                prev = e[j];
            }
        }
        if (a.box) {
            int32_t bx[PREP_RUN];
#pragma unroll
            for (int j = 0; j < PREP_RUN; ++j) bx[j] = lo + j < a.ndep ? a.box[lo + j] : 0;
#pragma unroll
            for (int j = 0; j < PREP_RUN; ++j)
                if (bx[j] < 0 || bx[j] >= a.nbox) bad |= 4u;
        }
        if (!a.nph) continue;
        int32_t v[PREP_RUN];
        int64_t s = 0;
#pragma unroll
        for (int j = 0; j < PREP_RUN; ++j) {
            int32_t x = lo + j < a.ndep ? a.nph[lo + j] : 0;
            if (x < 0) { bad |= 1u; x = 0; }
            v[j] = x;
            s += x;
        }
        // the runs' exclusive prefix: an inclusive scan within each wave, then one over the
        // wave totals
        int64_t x = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        if (wave == 0) {
            int64_t t = lane < nwv ? wsum[lane] : 0;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const int64_t y = __shfl_up(t, d);
                if (lane >= d) t += y;
            }
            if (lane < nwv) wsum[lane] = t;         // inclusive prefix over the waves
        }
        __syncthreads();
        int64_t acc = carry + (wave ? wsum[wave - 1] : 0) + x - s;
        carry += wsum[nwv - 1];
#pragma unroll
        for (int j = 0; j < PREP_RUN; ++j)
            if (lo + j < a.ndep) { a.off[lo + j] = acc; acc += v[j]; }
        __syncthreads();                            // the next pass rewrites wsum
    }
    if (a.nph && tid == 0) a.off[a.ndep] = carry;   // the photon total
    if (a.opc && tid < 10) a.opc[tid] = 0ull;
    if (tid == 0) *a.cursor = 0ull;
    if (bad) atomicOr(&sbad, bad);
    __syncthreads();
    if (tid < 4) a.flag[tid] = tid == 0 ? sbad : 0u;
}

__global__ void debug_draws_kernel(uint64_t seed, uint64_t pid0, int64_t n_pids, int n_draws,
                                   uint32_t *out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_pids) return;
    extern __shared__ uint32_t ring[];
    KParams P{};
    P.key0 = (uint32_t)seed; P.key1 = (uint32_t)(seed >> 32);
    Rng r;
    rng_bind(r, ring);
    rng_init(r, pid0 + (uint64_t)i + 1ull);
    OpCount<false> oc{nullptr};
    for (int d = 0; d < n_draws; ++d) out[i * n_draws + d] = rng_next<false>(r, P, oc);
}

// Largest |fresnel_R_est - fresnel_R| at cosine c over a fixed set of interfaces: the
// crystal/coupling and crystal/air pairs of the workloads, equal indices, and n1 < n2.
__device__ float fresnel_gap(float c)
{
    constexpr float n1s[9] = {2.15f, 2.15f, 1.79f, 1.79f, 1.95f, 1.95f, 1.5f, 1.3f, 1.0f};
    constexpr float n2s[9] = {1.5f, 1.0f, 1.5f, 1.0f, 1.45f, 1.0f, 1.5f, 1.5f, 2.15f};
    float g = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float s2 = fresnel_sin2t(n1s[k] / n2s[k], c);
        if (s2 < 1.0f)
            g = fmaxf(g, fabsf(fresnel_R_est(n1s[k], n2s[k], c, s2) - fresnel_R(n1s[k], n2s[k], c, s2)));
    }
    return g;
}

__global__ void debug_math_kernel(int which, const float *x, int64_t n, float *y)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        float s, c;
        if (which == 0) y[i] = f_log(v);
        else if (which == 3) y[i] = f_sqrt(v);
        else if (which == 4) y[i] = f_cos(v);
        else if (which == 5 || which == 6) { f_sincos_tilt(v, s, c); y[i] = which == 5 ? s : c; }
        else if (which == 7) y[i] = f_log2_est(v);
        else if (which == 8) y[i] = fresnel_gap(v);
        else { f_sincos(v, s, c); y[i] = which == 1 ? s : c; }
    }
}

}  // namespace

// =============================================================== host side ====

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(CRYSP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int B_NBUF_MAX = 32;

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct crysp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    int max_threads_per_cu = 2048;
    size_t lds_per_block = 65536;
    int wg = 256;
    int ppt = 4;
    int chunk = 256;
    int refill_min = 4;
    bool own_events = false;                        // event mode: workgroups own events
    bool timeline = false;                          // record each wave's start and end
    int64_t timeline_waves = 0;                     // waves of the last recorded launch
    void *timeline_dev = nullptr;                   // the device's g_timeline
    int64_t wall_hz = 100000000;                    // device wall clock rate
    int piece = 1024;                               // photons per workgroup claim of an event
    bool timing = false;
    double kernel_ms = 0.0;
    int64_t n_launches = 0;
    bool count_ops = false;
    unsigned long long ops_host[10] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf bufs[B_NBUF_MAX];
    int64_t last_n_hint = 0;
    // captured event-mode graph
    hipGraphExec_t graph = nullptr;
    std::vector<unsigned char> graph_key;
    // the device's copy of the file-scope scratch array (allocation-free entry point)
    void *scratch = nullptr;
};

// Scratch memory of crysp_run_events_noalloc: a file-scope device array, so that the
// entry point needs no allocation.  Each device holds its own copy for the life of the
// process; every call on a device uses that device's copy, so calls that overlap on one
// device race.  Layout: the work-queue cursor, the error flags, the deposits' photon
// offsets (NA_MAX_DEPOSITS + 1 int64).
constexpr int64_t NA_MAX_DEPOSITS = (int64_t)1 << 20;
constexpr size_t NA_CURSOR = 0, NA_FLAG = 64, NA_DEPOFF = 256;
constexpr size_t NA_BYTES = NA_DEPOFF + 8 * (size_t)(NA_MAX_DEPOSITS + 1);
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// The scratch array's address on `device` (current), looked up once per device.  The
// array's host-side name is a stand-in, not a device address: only this pointer is
// passed to kernels and copies.
static void *scratch_address(int device)
{
    static std::mutex mu;
    static void *addr[64] = {};
    static bool done[64] = {};
    if (device < 0 || device >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!done[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
        }
        addr[device] = p;
        done[device] = p != nullptr;
    }
    return addr[device];
}

enum BufId { B_PX, B_PY, B_PZ, B_PT, B_STATUS, B_IDX, B_TDET, B_BNC, B_COUNTS, B_TKEYS,
             B_COUNTERS, B_CURSOR, B_DEP, B_NPH, B_DEPEV, B_DEPOFF, B_MAPS, B_TMIN, B_EVKEYS,
             B_NPE, B_OPS, B_MISC, B_MISC2, B_FLAG, B_BOXTAB, B_DEPBOX, B_EVNEXT, B_EVEND,
             B_EVHEAD, B_POOL, B_SESSCTR, B_EVORDER, B_NBUF };
static_assert(B_NBUF <= B_NBUF_MAX, "buffer table");

static int ensure(crysp_ctx *c, int id, size_t bytes, void **out)
{
    DevBuf &b = c->bufs[id];
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr; b.cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        if (hipMalloc(&b.p, want) != hipSuccess) {
            (void)hipGetLastError();
            return fail(CRYSP_ENOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
        }
        b.cap = want;
    }
    *out = b.p;
    return CRYSP_OK;
}

// true when p is device memory usable in place by kernels on the context's device
static bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) { (void)hipGetLastError(); return false; }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// input: device pointer in place, or a staged copy in buffer `id`
static int stage_in(crysp_ctx *c, int id, const void *src, size_t bytes, const void **dev)
{
    if (!src) { *dev = nullptr; return CRYSP_OK; }
    if (is_device_ptr(src)) { *dev = src; return CRYSP_OK; }
    void *d;
    int rc = ensure(c, id, bytes, &d);
    if (rc) return rc;
    if (bytes) HIPCHK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, c->stream));
    *dev = d;
    return CRYSP_OK;
}

// output: device pointer in place, or buffer `id` copied back by stage_out
static int out_buf(crysp_ctx *c, int id, void *dst, size_t bytes, void **dev)
{
    if (!dst) { *dev = nullptr; return CRYSP_OK; }
    if (is_device_ptr(dst)) { *dev = dst; return CRYSP_OK; }
    return ensure(c, id, bytes, dev);
}

static int stage_out(crysp_ctx *c, void *dst, const void *dev, size_t bytes)
{
    if (!dst || dst == dev || !bytes) return CRYSP_OK;
    HIPCHK(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    return CRYSP_OK;
}

static int fill_params(KParams &P, const crysp_box *box, const crysp_optical_params *op,
                       const crysp_readout *ro, const crysp_emission *em, uint64_t seed,
                       uint64_t pid0, int32_t maxb)
{
    if (!box || !op || !ro) return fail(CRYSP_EINVAL, "null parameter struct");
    if (ro->grid.nx < 1 || ro->grid.ny < 1)
        return fail(CRYSP_EINVAL, "SiPM grid must have nx, ny >= 1");
    if ((int64_t)ro->grid.nx * ro->grid.ny > 32767)
        return fail(CRYSP_ERANGE, "nx*ny > 32767 overflows the Int16 SiPM index");
    if (!(ro->grid.pitch_mm > 0.0f)) return fail(CRYSP_EINVAL, "pitch_mm must be > 0");
    for (int a = 0; a < 3; ++a)
        if (!(box->L[a] > 0.0f) || !std::isfinite(box->L[a]))
            return fail(CRYSP_EINVAL, "box sides must be finite and > 0");
    if (maxb < 1) return fail(CRYSP_EINVAL, "max_bounces must be >= 1");
    std::memset(&P, 0, sizeof P);
    P.L0 = box->L[0]; P.L1 = box->L[1]; P.L2 = box->L[2];
    P.n_crystal = op->n_crystal; P.n_coupling = op->n_coupling;
    P.abs_len = op->abs_len_mm; P.ray_len = op->rayleigh_mm;
    P.wrap_R = op->wrap_R; P.sigma_alpha = op->sigma_alpha; P.pde = op->pde;
    P.inv_v = op->n_crystal / C_MM_NS;
    P.neg_abs_ln2 = -op->abs_len_mm * 0.6931472f;
    P.q_coup = op->n_crystal / op->n_coupling;
    P.q_air = op->n_crystal / 1.0f;
    P.cx = ro->cx; P.cy = ro->cy; P.r2 = ro->r2; P.sur_R = ro->sur_R;
    P.pitch = ro->grid.pitch_mm; P.nx = ro->grid.nx; P.ny = ro->grid.ny;
    P.tau_d = em ? em->tau_decay : 0.0f;
    P.tau_r = em ? em->tau_rise : 0.0f;
    P.maxb = maxb;
    uint32_t f = 0;
    if (std::isfinite(op->abs_len_mm)) f |= K_ABS;
    if (std::isfinite(op->rayleigh_mm)) f |= K_RAY;
    if (op->wrap_specular) f |= K_WRAP_SPEC;
    if (op->air_gap) f |= K_AIR_GAP;
    if (ro->disc) f |= K_DISC;
    if (ro->sur_specular) f |= K_SUR_SPEC;
    if (ro->two_sided) f |= K_TWO_SIDED;
    if (op->sigma_alpha > 0.0f) f |= K_ROUGH;
    P.flags = f;
    P.key0 = (uint32_t)seed; P.key1 = (uint32_t)(seed >> 32);
    P.pid0 = pid0;
    return CRYSP_OK;
}

// LDS of the lanes' Philox rings: 16 words per lane
static size_t ring_bytes(int wg) { return (size_t)wg * 64; }
// the batched-emission slots and header of the kernels without sessions, per wave
static size_t emb_bytes(int wg) { return (size_t)(wg / 64) * EMB_WAVE_BYTES; }
// the transport kernel's static LDS (counters) stays below this
constexpr size_t STATIC_LDS_MARGIN = 256;

static int grid_for(crysp_ctx *c, int64_t n, int wg, int ppt, size_t lds_bytes)
{
    // persistent grid: ceil(n / (wg*ppt)) workgroups, at most what stays resident
    int64_t want = (n + (int64_t)wg * ppt - 1) / ((int64_t)wg * ppt);
    // the transport kernels run at most seven waves per SIMD (four SIMDs per CU)
    int per_cu = std::max(1, std::min(c->max_threads_per_cu, 7 * 4 * 64) / wg);
    if (lds_bytes) per_cu = std::max(1, std::min<int>(per_cu, (int)(160 * 1024 / lds_bytes)));
    int64_t cap = (int64_t)c->n_cu * per_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, cap));
}

// The transport kernel for a launch: the specialization whose optics feature set equals the
// launch's flags (ESR or Teflon over an air gap, Teflon with Rayleigh scattering), else the
// generic kernel.  Records
// mode and the op-counting variant always run the generic kernel.
template <int MODE, bool MULTIBOX = false, bool SESS = false>
static void launch_optics_variant(const KParams &P, const KBuffers &B, dim3 grid, dim3 block,
                                  size_t lds, hipStream_t stream)
{
    if constexpr (MODE != M_RECORDS) {
        if (P.flags == OPT_ESR) {
            hipLaunchKernelGGL((transport_kernel<MODE, false, MULTIBOX, OPT_ESR, SESS>), grid, block,
                               lds, stream, P, B);
            return;
        }
        if (P.flags == OPT_TEFLON) {
            hipLaunchKernelGGL((transport_kernel<MODE, false, MULTIBOX, OPT_TEFLON, SESS>), grid,
                               block, lds, stream, P, B);
            return;
        }
        if (P.flags == OPT_TEFLON_RAY) {
            hipLaunchKernelGGL((transport_kernel<MODE, false, MULTIBOX, OPT_TEFLON_RAY, SESS>), grid,
                               block, lds, stream, P, B);
            return;
        }
    }
    hipLaunchKernelGGL((transport_kernel<MODE, false, MULTIBOX, OPT_GENERIC, SESS>), grid, block,
                       lds, stream, P, B);
}

// The wave timeline buffer for a launch of `waves` waves (cleared by a fill kernel), or
// nothing when the timeline is off.
static int timeline_setup(crysp_ctx *c, int64_t waves)
{
    if (!c->timeline) return CRYSP_OK;
    c->timeline_waves = std::min<int64_t>(waves, TL_MAX_WAVES);
    hipLaunchKernelGGL(fill_u32, dim3(64), dim3(256), 0, c->stream, (uint32_t *)c->timeline_dev, 0u,
                       (int64_t)TL_MAX_WAVES * 2);
    return CRYSP_OK;
}

template <int MODE>
static int launch_transport(crysp_ctx *c, KParams P, KBuffers B, size_t lds_bytes)
{
    const int wg = c->wg;
    lds_bytes += ring_bytes(wg) + emb_bytes(wg);
    const int grid = grid_for(c, P.n_photons, wg, c->ppt, lds_bytes);
    P.chunk = c->chunk;
    P.refill_min = c->refill_min;
    void *cur;
    int rc = ensure(c, B_CURSOR, sizeof(unsigned long long), &cur);
    if (rc) return rc;
    B.cursor = (unsigned long long *)cur;
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, c->stream, (uint32_t *)cur, 0u, 2);
    if ((rc = timeline_setup(c, (int64_t)grid * (wg / 64)))) return rc;
    if (c->count_ops) {
        void *ops;
        rc = ensure(c, B_OPS, 10 * sizeof(unsigned long long), &ops);
        if (rc) return rc;
        B.opcounts = (unsigned long long *)ops;
        hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, c->stream, (uint32_t *)ops, 0u, 20);
    }
    if (c->timing) HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (c->count_ops)
        hipLaunchKernelGGL((transport_kernel<MODE, true>), dim3(grid), dim3(wg), lds_bytes,
                           c->stream, P, B);
    else
        launch_optics_variant<MODE>(P, B, dim3(grid), dim3(wg), lds_bytes, c->stream);
    HIPCHK(hipGetLastError());
    if (c->timing) HIPCHK(hipEventRecord(c->ev1, c->stream));
    return CRYSP_OK;
}

static int finish_timing(crysp_ctx *c)
{
    if (c->count_ops) {
        unsigned long long h[10];
        HIPCHK(hipMemcpyAsync(h, c->bufs[B_OPS].p, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < 10; ++q) c->ops_host[q] += h[q];
    }
    if (!c->timing) return CRYSP_OK;
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->kernel_ms += ms;
    c->n_launches += 1;
    return CRYSP_OK;
}

extern "C" {

const char *crysp_last_error(void) { return g_err.c_str(); }
int crysp_abi_version(void) { return CRYSPLIGHT_ABI_VERSION; }

crysp_ctx *crysp_create(int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        fail(CRYSP_EHIP, "no HIP device available");
        return nullptr;
    }
    if (device < 0 || device >= ndev) {
        fail(CRYSP_EINVAL, "device index out of range");
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        fail(CRYSP_EHIP, "hipSetDevice failed");
        return nullptr;
    }
    crysp_ctx *c = new crysp_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        fail(CRYSP_EHIP, "hipGetDeviceProperties failed");
        return nullptr;
    }
    c->n_cu = prop.multiProcessorCount;
    c->max_threads_per_cu = prop.maxThreadsPerMultiProcessor;
    c->lds_per_block = prop.sharedMemPerBlock;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
        wall_khz > 0)
        c->wall_hz = (int64_t)wall_khz * 1000;
    c->scratch = scratch_address(device);           // device current (hipSetDevice above)
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        fail(CRYSP_EHIP, "stream/event creation failed");
        return nullptr;
    }
    return c;
}

void crysp_destroy(crysp_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (auto &b : c->bufs)
        if (b.p) (void)hipFree(b.p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int crysp_set_schedule(crysp_ctx *c, int32_t wg, int32_t ppt)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (wg < 64 || wg > 1024 || wg % 64) return fail(CRYSP_EINVAL, "wg_size must be 64..1024, multiple of 64");
    if (ring_bytes(wg) + emb_bytes(wg) + STATIC_LDS_MARGIN > c->lds_per_block)
        return fail(CRYSP_ERANGE, "wg_size " + std::to_string(wg) + " needs " +
                                  std::to_string(ring_bytes(wg) + emb_bytes(wg)) +
                                  " B of LDS for the Philox rings and emission slots; "
                                  "the device allows " + std::to_string(c->lds_per_block) + " B per workgroup");
    if (ppt < 1 || ppt > 64) return fail(CRYSP_EINVAL, "photons_per_thread must be 1..64");
    c->wg = wg;
    c->ppt = ppt;
    return CRYSP_OK;
}

int crysp_set_queue(crysp_ctx *c, int32_t chunk, int32_t refill_min)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (chunk < 1 || chunk > (1 << 20)) return fail(CRYSP_EINVAL, "chunk must be 1..2^20");
    if (refill_min < 1 || refill_min > 64) return fail(CRYSP_EINVAL, "refill_min must be 1..64");
    c->chunk = chunk;
    c->refill_min = refill_min;
    return CRYSP_OK;
}

int crysp_set_wave_timeline(crysp_ctx *c, int32_t enable)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    HIPCHK(hipSetDevice(c->device));
    if (!c->timeline_dev)
        HIPCHK(hipGetSymbolAddress(&c->timeline_dev, HIP_SYMBOL(g_timeline)));
    const int on = enable != 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_timeline_on), &on, sizeof on));
    c->timeline = on != 0;
    return CRYSP_OK;
}

int crysp_get_wave_timeline(crysp_ctx *c, uint64_t *out, int64_t cap, int64_t *n_waves,
                            double *ticks_per_s)
{
    if (!c || (cap > 0 && !out)) return fail(CRYSP_EINVAL, "bad timeline arguments");
    const int64_t n = c->timeline_waves;
    if (n_waves) *n_waves = n;
    if (ticks_per_s) *ticks_per_s = (double)c->wall_hz;
    const int64_t m = std::min<int64_t>(n, std::max<int64_t>(cap, 0));
    if (m > 0 && c->timeline_dev) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(out, c->timeline_dev, (size_t)m * 8, hipMemcpyDeviceToHost));
    }
    return CRYSP_OK;
}

int crysp_set_event_sessions(crysp_ctx *c, int32_t enable, int32_t piece)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (piece < 1 || piece > (1 << 24)) return fail(CRYSP_EINVAL, "piece must be 1..2^24");
    c->own_events = enable != 0;
    c->piece = piece;
    return CRYSP_OK;
}

int crysp_get_schedule(crysp_ctx *c, int32_t *wg, int32_t *ppt)
{
    if (!c || !wg || !ppt) return fail(CRYSP_EINVAL, "null argument");
    *wg = c->wg;
    *ppt = c->ppt;
    return CRYSP_OK;
}

int crysp_set_timing(crysp_ctx *c, int32_t enable)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    c->timing = enable != 0;
    return CRYSP_OK;
}

int crysp_kernel_time(crysp_ctx *c, double *total_ms, int64_t *n, int32_t reset)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (total_ms) *total_ms = c->kernel_ms;
    if (n) *n = c->n_launches;
    if (reset) { c->kernel_ms = 0.0; c->n_launches = 0; }
    return CRYSP_OK;
}

int crysp_set_op_counting(crysp_ctx *c, int32_t enable)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    c->count_ops = enable != 0;
    return CRYSP_OK;
}

int crysp_get_op_counts(crysp_ctx *c, int64_t *out, int32_t reset)
{
    if (!c || !out) return fail(CRYSP_EINVAL, "null argument");
    for (int q = 0; q < 10; ++q) out[q] = (int64_t)c->ops_host[q];
    if (reset) std::memset(c->ops_host, 0, sizeof c->ops_host);
    return CRYSP_OK;
}

int crysp_transport_records(crysp_ctx *c, const crysp_box *box,
                            const crysp_optical_params *op, const crysp_readout *ro,
                            const float *px, const float *py, const float *pz,
                            const float *pt, int32_t npos, int64_t n, uint64_t seed,
                            uint64_t pid0, int32_t maxb, const crysp_emission *em,
                            uint8_t *status, int16_t *idx, float *tdet, int32_t *bnc)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (n < 0) return fail(CRYSP_EINVAL, "n_photons < 0");
    if (npos < 1 || !px || !py || !pz) return fail(CRYSP_EINVAL, "need npos >= 1 and px, py, pz");
    KParams P;
    int rc = fill_params(P, box, op, ro, em, seed, pid0, maxb);
    if (rc) return rc;
    if (n == 0) return CRYSP_OK;
    HIPCHK(hipSetDevice(c->device));
    P.n_photons = n;
    P.npos = npos;
    KBuffers B{};
    const size_t np = (size_t)npos * sizeof(float);
    if ((rc = stage_in(c, B_PX, px, np, (const void **)&B.px)) ||
        (rc = stage_in(c, B_PY, py, np, (const void **)&B.py)) ||
        (rc = stage_in(c, B_PZ, pz, np, (const void **)&B.pz)) ||
        (rc = stage_in(c, B_PT, pt, np, (const void **)&B.pt)))
        return rc;
    if ((rc = out_buf(c, B_STATUS, status, (size_t)n, (void **)&B.status)) ||
        (rc = out_buf(c, B_IDX, idx, (size_t)n * 2, (void **)&B.idx)) ||
        (rc = out_buf(c, B_TDET, tdet, (size_t)n * 4, (void **)&B.tdet)) ||
        (rc = out_buf(c, B_BNC, bnc, (size_t)n * 4, (void **)&B.bnc)))
        return rc;
    if ((rc = launch_transport<M_RECORDS>(c, P, B, 0))) return rc;
    if ((rc = stage_out(c, status, B.status, (size_t)n)) ||
        (rc = stage_out(c, idx, B.idx, (size_t)n * 2)) ||
        (rc = stage_out(c, tdet, B.tdet, (size_t)n * 4)) ||
        (rc = stage_out(c, bnc, B.bnc, (size_t)n * 4)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return finish_timing(c);
}

int crysp_transport_accumulate(crysp_ctx *c, const crysp_box *box,
                               const crysp_optical_params *op, const crysp_readout *ro,
                               const crysp_time_binning *tb, const float *px, const float *py,
                               const float *pz, const float *pt, int32_t npos, int64_t n,
                               uint64_t seed, uint64_t pid0, int32_t maxb,
                               const crysp_emission *em, uint32_t *counts, float *first_ns,
                               int64_t *counters)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (n < 0) return fail(CRYSP_EINVAL, "n_photons < 0");
    if (!tb || tb->nbins < 0 || !(tb->bin_ns > 0.0f)) return fail(CRYSP_EINVAL, "bad time binning");
    if (npos < 1 || !px || !py || !pz) return fail(CRYSP_EINVAL, "need npos >= 1 and px, py, pz");
    if (!counts || !first_ns || !counters) return fail(CRYSP_EINVAL, "null output");
    KParams P;
    int rc = fill_params(P, box, op, ro, em, seed, pid0, maxb);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const int64_t nxy = (int64_t)ro->grid.nx * ro->grid.ny;
    const int64_t hist = nxy * ((int64_t)tb->nbins + 1);
    const size_t lds_bytes = (size_t)(hist + nxy) * 4;
    if (lds_bytes + ring_bytes(c->wg) + emb_bytes(c->wg) + STATIC_LDS_MARGIN > c->lds_per_block)
        return fail(CRYSP_ERANGE, "nx*ny*(nbins+2) histogram words plus the rings and emission batches (89 B per lane) exceed "
                                  "the workgroup LDS (" + std::to_string(c->lds_per_block) + " B)");
    P.n_photons = n;
    P.npos = npos;
    P.nbins = tb->nbins;
    P.bin_ns = tb->bin_ns;
    P.hist_words = (int32_t)hist;
    KBuffers B{};
    const size_t np = (size_t)npos * sizeof(float);
    if ((rc = stage_in(c, B_PX, px, np, (const void **)&B.px)) ||
        (rc = stage_in(c, B_PY, py, np, (const void **)&B.py)) ||
        (rc = stage_in(c, B_PZ, pz, np, (const void **)&B.pz)) ||
        (rc = stage_in(c, B_PT, pt, np, (const void **)&B.pt)))
        return rc;
    void *dcounts, *dkeys, *dctr, *dfirst;
    if ((rc = out_buf(c, B_COUNTS, counts, (size_t)hist * 4, &dcounts)) ||
        (rc = ensure(c, B_TKEYS, (size_t)nxy * 4, &dkeys)) ||
        (rc = ensure(c, B_COUNTERS, 16 * 8, &dctr)) ||
        (rc = out_buf(c, B_TMIN, first_ns, (size_t)nxy * 4, &dfirst)))
        return rc;
    B.counts = (uint32_t *)dcounts;
    B.tkeys = (uint32_t *)dkeys;
    B.counters = (unsigned long long *)dctr;
    hipLaunchKernelGGL(fill_u32, dim3(64), dim3(256), 0, c->stream, B.counts, 0u, hist);
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, c->stream, B.tkeys, KEY_INF, nxy);
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, c->stream, (uint32_t *)dctr, 0u, 32);
    if (n > 0 && (rc = launch_transport<M_ACCUM>(c, P, B, lds_bytes))) return rc;
    hipLaunchKernelGGL(keys_to_times, dim3(1), dim3(256), 0, c->stream, B.tkeys,
                       (float *)dfirst, nxy);
    unsigned long long hc[16];
    HIPCHK(hipMemcpyAsync(hc, dctr, sizeof hc, hipMemcpyDeviceToHost, c->stream));
    if ((rc = stage_out(c, counts, dcounts, (size_t)hist * 4)) ||
        (rc = stage_out(c, first_ns, dfirst, (size_t)nxy * 4)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < CRYSP_NCOUNTERS; ++q) counters[q] = (int64_t)hc[q];
    if (n > 0) return finish_timing(c);
    return CRYSP_OK;
}

// Event mode.  Deposit photon offsets come from the host (H2D) when the counts live in
// host memory, or from a device scan when they live in device memory; the transport
// kernel reads the photon total from dep_off[ndep], so nothing forces a host round trip
// and the whole launch sequence is captured once in a hipGraph and replayed while the
// arguments (pointers, sizes, parameters, schedule) stay the same.
static int events_impl(crysp_ctx *c, const crysp_box *box, const crysp_optical_params *op,
                       const crysp_readout *ro, const float *dep_xyzt, const int32_t *dep_nph,
                       const int32_t *dep_event, int64_t ndep, int64_t nev, uint64_t seed,
                       uint64_t pid0, int32_t maxb, const crysp_emission *em, int32_t *maps,
                       float *tmin, int32_t *npe, int64_t *n_out, bool use_graph,
                       unsigned long long *tsum_dev = nullptr, const crysp_box *boxes = nullptr,
                       int32_t nbox = 0, const int32_t *dep_box = nullptr,
                       void *noalloc = nullptr)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (ndep < 0 || nev < 0) return fail(CRYSP_EINVAL, "negative sizes");
    if (ndep > 0 && (!dep_xyzt || !dep_nph || !dep_event))
        return fail(CRYSP_EINVAL, "null deposit arrays");
    if (nev > 0 && (!maps || !tmin || !npe)) return fail(CRYSP_EINVAL, "null outputs");
    if (nev > INT32_MAX) return fail(CRYSP_ERANGE, "more than 2^31-1 events");
    if (noalloc) {
        // everything in device memory already, nothing staged, no graph, no counters
        if (ndep > NA_MAX_DEPOSITS)
            return fail(CRYSP_ERANGE, "more than " + std::to_string(NA_MAX_DEPOSITS) +
                                      " deposits for the allocation-free entry point");
        if ((ndep > 0 && (!is_device_ptr(dep_xyzt) || !is_device_ptr(dep_nph) ||
                          !is_device_ptr(dep_event))) ||
            (nev > 0 && (!is_device_ptr(maps) || !is_device_ptr(tmin) || !is_device_ptr(npe))))
            return fail(CRYSP_EINVAL, "the allocation-free entry point takes device arrays only");
        if (c->count_ops)
            return fail(CRYSP_EINVAL, "operation counting needs the allocating entry point");
        if (boxes || tsum_dev) return fail(CRYSP_EINVAL, "single crystal, no time sums");
    }
    KParams P;
    int rc = fill_params(P, box, op, ro, em, seed, pid0, maxb);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const int nxy = ro->grid.nx * ro->grid.ny;
    const bool nph_dev = ndep > 0 && is_device_ptr(dep_nph);
    P.ndep = ndep;
    P.nev = nev;
    KBuffers B{};
    void *doff, *dflag;
    if (noalloc) {
        doff = (char *)noalloc + NA_DEPOFF;
        dflag = (char *)noalloc + NA_FLAG;
    } else if ((rc = ensure(c, B_DEPOFF, (size_t)(ndep + 1) * 8, &doff)) ||
               (rc = ensure(c, B_FLAG, 16, &dflag))) {
        return rc;
    }
    B.dep_off = (const int64_t *)doff;
    B.flag = (unsigned int *)dflag;
    int64_t ntot_host = -1;
    std::vector<int64_t> off;                       // alive until the final sync
    if (!nph_dev) {
        off.assign(ndep + 1, 0);
        for (int64_t d = 0; d < ndep; ++d) {
            if (dep_nph[d] < 0) return fail(CRYSP_EINVAL, "negative photon count in dep_nph");
            off[d + 1] = off[d] + dep_nph[d];
        }
        ntot_host = off[ndep];
        HIPCHK(hipMemcpyAsync(doff, off.data(), (size_t)(ndep + 1) * 8, hipMemcpyHostToDevice,
                              c->stream));
    }
    const bool ev_dev = ndep > 0 && is_device_ptr(dep_event);
    bool ev_sorted = true;                          // host tables: events in order
    if (!ev_dev)
        for (int64_t d = 0; d < ndep; ++d) {
            if (dep_event[d] < 0 || dep_event[d] >= nev)
                return fail(CRYSP_EINVAL, "dep_event out of range");
            if (d > 0 && dep_event[d] < dep_event[d - 1]) ev_sorted = false;
        }
    if ((rc = stage_in(c, B_DEP, dep_xyzt, (size_t)ndep * 16, (const void **)&B.dep)) ||
        (rc = stage_in(c, B_DEPEV, dep_event, (size_t)ndep * 4, (const void **)&B.dep_event)))
        return rc;
    const int32_t *nph_d = nph_dev ? dep_nph : nullptr;
    void *dmaps, *dtmin, *dnpe;
    const size_t cells = (size_t)nev * nxy;
    if ((rc = out_buf(c, B_MAPS, maps, cells * 4, &dmaps)) ||
        (rc = out_buf(c, B_TMIN, tmin, cells * 4, &dtmin)) ||
        (rc = out_buf(c, B_NPE, npe, (size_t)nev * 4, &dnpe)))
        return rc;
    B.maps = (int32_t *)dmaps;
    // first-hit times are gathered as order-preserving keys in the tmin buffer itself and
    // turned into floats in place by finish_events
    void *dkeys = dtmin;
    B.ev_tkeys = (uint32_t *)dkeys;
    B.ev_tsum = tsum_dev;
    const bool multibox = boxes != nullptr;
    std::vector<float> boxtab;
    if (multibox) {
        if (nbox < 1 || nbox > 64 || !dep_box) return fail(CRYSP_EINVAL, "need 1..64 boxes and dep_box");
        boxtab.resize((size_t)nbox * 4);
        for (int b = 0; b < nbox; ++b) {
            for (int a = 0; a < 3; ++a) {
                if (!(boxes[b].L[a] > 0.0f) || !std::isfinite(boxes[b].L[a]))
                    return fail(CRYSP_EINVAL, "box sides must be finite and > 0");
                boxtab[4 * b + a] = boxes[b].L[a];
            }
            boxtab[4 * b + 3] = 0.0f;
        }
        if (!is_device_ptr(dep_box))
            for (int64_t d = 0; d < ndep; ++d)
                if (dep_box[d] < 0 || dep_box[d] >= nbox) return fail(CRYSP_EINVAL, "dep_box out of range");
        P.nbox = nbox;
        void *dbt;
        if ((rc = ensure(c, B_BOXTAB, (size_t)nbox * 16, &dbt))) return rc;
        HIPCHK(hipMemcpyAsync(dbt, boxtab.data(), (size_t)nbox * 16, hipMemcpyHostToDevice, c->stream));
        B.boxes = (const float4 *)dbt;
        if ((rc = stage_in(c, B_DEPBOX, dep_box, (size_t)ndep * 4, (const void **)&B.dep_box))) return rc;
    }
    void *cur;
    if (noalloc) cur = (char *)noalloc + NA_CURSOR;
    else if ((rc = ensure(c, B_CURSOR, 8, &cur))) return rc;
    B.cursor = (unsigned long long *)cur;
    if (c->count_ops) {
        void *ops;
        if ((rc = ensure(c, B_OPS, 80, &ops))) return rc;
        B.opcounts = (unsigned long long *)ops;
    }
    P.chunk = c->chunk;
    P.refill_min = c->refill_min;
    // the grid only sizes the persistent launch; the kernel drains whatever total it reads
    const int64_t n_hint = ntot_host >= 0 ? ntot_host : (c->last_n_hint > 0 ? c->last_n_hint : (int64_t)1 << 40);
    P.n_photons = -1;                               // read dep_off[ndep] on the device
    const size_t lds_ev = ring_bytes(c->wg) + emb_bytes(c->wg);
    const int grid = grid_for(c, n_hint, c->wg, c->ppt, lds_ev);

    // Events owned by workgroups (one writer per map row, no global atomics on the maps)
    // when the table is ordered by event; a device-resident table is checked on the device,
    // and then both paths are enqueued, each skipping itself unless the flag picks it.
    const size_t lds_sess = ring_bytes(c->wg) + sess_lds_words(nxy) * 4;
    const bool sess = c->own_events && !multibox && !tsum_dev && !c->count_ops && !noalloc &&
                      ndep > 0 && nev > 0 && nxy <= SESS_MAX_CELLS && (ev_dev || ev_sorted) &&
                      lds_sess + STATIC_LDS_MARGIN <= c->lds_per_block;
    const bool run_old = !sess || ev_dev;
    const int grid_sess = sess ? grid_for(c, n_hint, c->wg, c->ppt, lds_sess) : 0;
    KParams Ps = P;
    if (sess) {
        Ps.gate = ev_dev ? 1 : 0;
        Ps.piece = c->piece;
        Ps.pool_cap = 8 * grid_sess;
        Ps.help_window = 2 * (int64_t)grid_sess;
        P.gate = ev_dev ? 2 : 0;
        void *p1, *p2, *p3, *p4, *p5;
        const size_t pool_words = (size_t)Ps.pool_cap * (2 * (size_t)nxy + 1);
        // the counters, then one open-event word per workgroup
        if ((rc = ensure(c, B_EVNEXT, (size_t)nev * 8, &p1)) ||
            (rc = ensure(c, B_EVEND, (size_t)nev * 8, &p2)) ||
            (rc = ensure(c, B_EVHEAD, (size_t)nev * 4, &p3)) ||
            (rc = ensure(c, B_POOL, pool_words * 4, &p4)) ||
            (rc = ensure(c, B_SESSCTR, 16 + (size_t)grid_sess * 4, &p5)))
            return rc;
        void *p6;
        if ((rc = ensure(c, B_EVORDER, (size_t)nev * 4, &p6))) return rc;
        B.ev_order = (const uint32_t *)p6;
        B.open_ev = (uint32_t *)p5 + 4;
        B.ev_next = (unsigned long long *)p1;
        B.ev_end = (const unsigned long long *)p2;
        B.ev_head = (uint32_t *)p3;
        B.pool_link = (uint32_t *)p4;
        B.pool_data = (uint32_t *)p4 + Ps.pool_cap;
        B.sess_ctr = (uint32_t *)p5;
        B.tmin_out = (float *)dtmin;
        B.npe_out = (int32_t *)dnpe;
    }

    // everything the captured nodes read: parameters, every buffer address (the kernel's
    // and the outputs finish_events writes), the launch shape and the variant
    const bool box_dev = multibox && is_device_ptr(dep_box);
    const int64_t tl_waves = (int64_t)std::max(grid, grid_sess) * (c->wg / 64);
    std::vector<unsigned char> key(sizeof P + sizeof B + 3 * sizeof(void *) + 10 * sizeof(int64_t));
    {
        int64_t extra[10] = { (int64_t)grid, (int64_t)c->wg, nev, (int64_t)c->count_ops,
                              (int64_t)c->timing, (int64_t)nph_dev, (int64_t)ev_dev,
                              (int64_t)box_dev, (int64_t)sess, (int64_t)c->piece };
        const void *ptrs[3] = { nph_d, dtmin, dnpe };
        unsigned char *q = key.data();
        std::memcpy(q, &P, sizeof P); q += sizeof P;
        std::memcpy(q, &B, sizeof B); q += sizeof B;
        std::memcpy(q, ptrs, sizeof ptrs); q += sizeof ptrs;
        std::memcpy(q, extra, sizeof extra);
    }
    // timing events bracket the transport kernel when launched eagerly, and the whole
    // graph (a few-microsecond fill/scan/finish around the kernel) when replayed
    auto enqueue = [&](bool capturing) -> int {
        // cursor, flags, offsets, index checks and the output fills in one launch;
        // device-resident index tables are checked on the device (host ones were checked
        // before any launch)
        PrepArgs pa{};
        pa.cursor = (unsigned long long *)cur;
        pa.flag = (unsigned int *)dflag;
        pa.nph = nph_d;
        pa.off = (int64_t *)doff;
        const int32_t *box_chk = box_dev ? B.dep_box : nullptr;
        pa.ev = ndep > 0 && (ev_dev || box_chk) ? B.dep_event : nullptr;
        pa.box = ndep > 0 ? box_chk : nullptr;
        pa.ndep = ndep; pa.nev = nev; pa.nbox = nbox;
        pa.opc = c->count_ops ? B.opcounts : nullptr;
        if (run_old && !sess) {
            pa.maps = (uint32_t *)dmaps; pa.keys = (uint32_t *)dkeys; pa.cells = (int64_t)cells;
        }
        if (tsum_dev) { pa.zero = (uint32_t *)tsum_dev; pa.nzero = (int64_t)cells * 2; }
        const int64_t nfill = std::max<int64_t>(pa.maps ? pa.cells : 0, pa.zero ? pa.nzero : 0);
        const unsigned prep_grid = 1u + (unsigned)std::min<int64_t>(512, (nfill + 4095) / 4096);
        hipLaunchKernelGGL(prepare_events, dim3(prep_grid), dim3(1024), 0, c->stream, pa);
        if (run_old && sess) {                      // only if the device picks this path
            hipLaunchKernelGGL(fill_u32_gated, dim3(1024), dim3(256), 0, c->stream, (uint32_t *)dmaps,
                               0u, (int64_t)cells, (const unsigned int *)dflag, 1);
            hipLaunchKernelGGL(fill_u32_gated, dim3(1024), dim3(256), 0, c->stream, (uint32_t *)dkeys,
                               KEY_INF, (int64_t)cells, (const unsigned int *)dflag, 1);
        }
        if (sess) {
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, c->stream, B.sess_ctr, 0u, 4);
            hipLaunchKernelGGL(fill_u32, dim3((grid_sess + 255) / 256), dim3(256), 0, c->stream,
                               B.open_ev, ~0u, (int64_t)grid_sess);
            hipLaunchKernelGGL(event_ranges, dim3((unsigned)std::min<int64_t>(1024, (nev + 255) / 256)),
                               dim3(256), 0, c->stream, B.dep_event, B.dep_off, ndep, nev, B.ev_next,
                               (unsigned long long *)B.ev_end, B.ev_head, (const unsigned int *)dflag,
                               (int)ev_dev);
            hipLaunchKernelGGL(order_events, dim3(1), dim3(1024), 0, c->stream,
                               (const unsigned long long *)B.ev_next, B.ev_end, nev,
                               (uint32_t *)B.ev_order, (const unsigned int *)dflag, (int)ev_dev);
        }
        if ((rc = timeline_setup(c, tl_waves))) return rc;
        if (ndep > 0) {
            if (c->timing && !capturing) HIPCHK(hipEventRecord(c->ev0, c->stream));
            if (sess)
                launch_optics_variant<M_EVENTS, false, true>(Ps, B, dim3(grid_sess), dim3(c->wg),
                                                             lds_sess, c->stream);
            if (!run_old) {
            } else if (multibox)
                launch_optics_variant<M_EVENTS, true>(P, B, dim3(grid), dim3(c->wg), lds_ev, c->stream);
            else if (c->count_ops)
                hipLaunchKernelGGL((transport_kernel<M_EVENTS, true>), dim3(grid), dim3(c->wg), lds_ev,
                                   c->stream, P, B);
            else
                launch_optics_variant<M_EVENTS>(P, B, dim3(grid), dim3(c->wg), lds_ev, c->stream);
            if (c->timing && !capturing) HIPCHK(hipEventRecord(c->ev1, c->stream));
        }
        const dim3 ev_grid((unsigned)std::min<int64_t>(4096, (nev + 3) / 4));     // a wave per event
        if (nev > 0 && run_old)
            hipLaunchKernelGGL(finish_events, ev_grid, dim3(256), 0, c->stream, (const int32_t *)dmaps,
                               (const uint32_t *)dkeys, (float *)dtmin, (int32_t *)dnpe, nev, nxy,
                               (const unsigned int *)dflag, (int)sess);
        if (sess)
            hipLaunchKernelGGL(finish_sessions, dim3((unsigned)std::min<int64_t>(4096, (nev + 3) / 4)),
                               dim3(256), 0, c->stream,
                               (const uint32_t *)B.ev_head, (const uint32_t *)B.pool_link,
                               (const uint32_t *)B.pool_data, (int32_t *)dmaps, (float *)dtmin,
                               (int32_t *)dnpe, nev, nxy, (const uint32_t *)B.sess_ctr,
                               (unsigned int *)dflag, (int)ev_dev);
        return CRYSP_OK;
    };
    bool launched = false;
    if (use_graph && !noalloc) {
        if (c->graph && key == c->graph_key) {
            if (c->timing) HIPCHK(hipEventRecord(c->ev0, c->stream));
            HIPCHK(hipGraphLaunch(c->graph, c->stream));
            if (c->timing) HIPCHK(hipEventRecord(c->ev1, c->stream));
            launched = true;
        } else {
            if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
            c->graph_key.clear();
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                rc = enqueue(true);
                hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
                if (e == hipSuccess && g &&
                    hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0) == hipSuccess) {
                    c->graph_key = key;
                    if (c->timing) HIPCHK(hipEventRecord(c->ev0, c->stream));
                    HIPCHK(hipGraphLaunch(c->graph, c->stream));
                    if (c->timing) HIPCHK(hipEventRecord(c->ev1, c->stream));
                    launched = true;
                }
                if (g) (void)hipGraphDestroy(g);
            }
            if (!launched) (void)hipGetLastError();   // capture unsupported here: run eagerly
        }
    }
    if (!launched) {
        if ((rc = enqueue(false))) return rc;
        HIPCHK(hipGetLastError());
    }
    int64_t tail[2] = { 0, 0 };                      // photon total, error flags
    HIPCHK(hipMemcpyAsync(&tail[0], (int64_t *)doff + ndep, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&tail[1], dflag, 4, hipMemcpyDeviceToHost, c->stream));
    if ((rc = stage_out(c, maps, dmaps, cells * 4)) ||
        (rc = stage_out(c, tmin, dtmin, cells * 4)) ||
        (rc = stage_out(c, npe, dnpe, (size_t)nev * 4)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last_n_hint = tail[0];
    if (n_out) *n_out = tail[0];
    if (tail[1] & 1) return fail(CRYSP_EINVAL, "negative photon count in dep_nph");
    if (tail[1] & 2) return fail(CRYSP_EINVAL, "dep_event out of range (those deposits' detections were dropped)");
    if (tail[1] & 4) return fail(CRYSP_EINVAL, "dep_box out of range (those deposits were traced in crystal 0)");
    if (tail[1] & FLAG_INTERNAL) return fail(CRYSP_EHIP, "internal error: the event queue did not drain");
    if (ndep > 0) return finish_timing(c);
    return CRYSP_OK;
}

int crysp_run_events(crysp_ctx *c, const crysp_box *box, const crysp_optical_params *op,
                     const crysp_readout *ro, const float *dep_xyzt, const int32_t *dep_nph,
                     const int32_t *dep_event, int64_t ndep, int64_t nev, uint64_t seed,
                     uint64_t pid0, int32_t maxb, const crysp_emission *em, int32_t *maps,
                     float *tmin, int32_t *npe, int64_t *n_out)
{
    return events_impl(c, box, op, ro, dep_xyzt, dep_nph, dep_event, ndep, nev, seed, pid0,
                       maxb, em, maps, tmin, npe, n_out, true);
}

int crysp_run_events_noalloc(crysp_ctx *c, const crysp_box *box, const crysp_optical_params *op,
                             const crysp_readout *ro, const float *dep_xyzt, const int32_t *dep_nph,
                             const int32_t *dep_event, int64_t ndep, int64_t nev, uint64_t seed,
                             uint64_t pid0, int32_t maxb, const crysp_emission *em, int32_t *maps,
                             float *tmin, int32_t *npe, int64_t *n_out)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (!c->scratch) return fail(CRYSP_EHIP, "no address for the device scratch array");
    return events_impl(c, box, op, ro, dep_xyzt, dep_nph, dep_event, ndep, nev, seed, pid0,
                       maxb, em, maps, tmin, npe, n_out, false, nullptr, nullptr, 0, nullptr,
                       c->scratch);
}

int crysp_run_events_multibox(crysp_ctx *c, const crysp_box *boxes, int32_t nbox,
                              const int32_t *dep_box, const crysp_optical_params *op,
                              const crysp_readout *ro, const float *dep_xyzt,
                              const int32_t *dep_nph, const int32_t *dep_event, int64_t ndep,
                              int64_t nev, uint64_t seed, uint64_t pid0, int32_t maxb,
                              const crysp_emission *em, int32_t *maps, float *tmin, int32_t *npe,
                              int64_t *n_out)
{
    if (!boxes) return fail(CRYSP_EINVAL, "null boxes");
    return events_impl(c, &boxes[0], op, ro, dep_xyzt, dep_nph, dep_event, ndep, nev, seed, pid0,
                       maxb, em, maps, tmin, npe, n_out, true, nullptr, boxes, nbox, dep_box);
}

int crysp_light_maps(crysp_ctx *c, const crysp_box *box, const crysp_optical_params *op,
                     const crysp_readout *ro, const float *xyz, const int32_t *nph, int64_t npts,
                     uint64_t seed, uint64_t pid0, int32_t maxb, const crysp_emission *em,
                     int32_t *counts, uint64_t *tsum_fx)
{
    if (!c) return fail(CRYSP_EINVAL, "null context");
    if (npts < 0 || (npts > 0 && (!xyz || !nph || !counts))) return fail(CRYSP_EINVAL, "bad arguments");
    if (!ro) return fail(CRYSP_EINVAL, "null readout");
    if (npts == 0) return CRYSP_OK;
    const int nxy = ro->grid.nx * ro->grid.ny;
    std::vector<float> dep((size_t)npts * 4);
    std::vector<int32_t> ev((size_t)npts);
    for (int64_t i = 0; i < npts; ++i) {
        dep[4 * i] = xyz[3 * i]; dep[4 * i + 1] = xyz[3 * i + 1]; dep[4 * i + 2] = xyz[3 * i + 2];
        dep[4 * i + 3] = 0.0f;
        ev[i] = (int32_t)i;
    }
    std::vector<float> tmin((size_t)npts * nxy);
    std::vector<int32_t> npe((size_t)npts);
    void *dts = nullptr;
    int rc;
    if (tsum_fx && (rc = ensure(c, B_MISC2, (size_t)npts * nxy * 8, &dts))) return rc;
    rc = events_impl(c, box, op, ro, dep.data(), nph, ev.data(), npts, npts, seed, pid0, maxb, em,
                     counts, tmin.data(), npe.data(), nullptr, false, (unsigned long long *)dts);
    if (rc) return rc;
    if (tsum_fx) {
        HIPCHK(hipMemcpy(tsum_fx, dts, (size_t)npts * nxy * 8, hipMemcpyDeviceToHost));
    }
    return CRYSP_OK;
}

int crysp_autotune_events(crysp_ctx *c, const crysp_box *box,
                          const crysp_optical_params *op, const crysp_readout *ro,
                          const float *dep_xyzt, const int32_t *dep_nph,
                          const int32_t *dep_event, int64_t ndep, int64_t nev, uint64_t seed,
                          int32_t maxb, const crysp_emission *em, const int32_t *wgs,
                          int32_t n_wg, const int32_t *ppts, int32_t n_ppt, int32_t reps,
                          float *times_ms, int32_t *best_wg, int32_t *best_ppt)
{
    if (!c || !wgs || !ppts || n_wg < 1 || n_ppt < 1 || reps < 1)
        return fail(CRYSP_EINVAL, "bad autotune arguments");
    const int old_wg = c->wg, old_ppt = c->ppt;
    const bool old_timing = c->timing;
    const double old_ms = c->kernel_ms;
    const int64_t old_n = c->n_launches;
    const int nxy = ro ? ro->grid.nx * ro->grid.ny : 0;
    std::vector<int32_t> hm((size_t)std::max<int64_t>(1, nev) * std::max(1, nxy)),
        hn((size_t)std::max<int64_t>(1, nev));
    std::vector<float> ht(hm.size());
    // outputs go to device scratch so the timing is the kernel's own
    void *dm, *dt;
    int rc;
    if ((rc = ensure(c, B_MISC, hm.size() * 4, &dm)) || (rc = ensure(c, B_MISC2, ht.size() * 4, &dt)))
        return rc;
    void *dnpe;
    if ((rc = ensure(c, B_NPE, hn.size() * 4, &dnpe))) return rc;
    float best = 1e30f;
    int bw = old_wg, bp = old_ppt;
    c->timing = true;
    for (int a = 0; a < n_wg; ++a) {
        for (int b = 0; b < n_ppt; ++b) {
            if (wgs[a] >= 64 && wgs[a] <= 1024 &&
                ring_bytes(wgs[a]) + emb_bytes(wgs[a]) + STATIC_LDS_MARGIN > c->lds_per_block) {
                if (times_ms) times_ms[a * n_ppt + b] = INFINITY;   // does not fit: skipped
                continue;
            }
            if ((rc = crysp_set_schedule(c, wgs[a], ppts[b]))) goto done;
            float tbest = 1e30f;
            for (int r = 0; r < reps; ++r) {
                c->kernel_ms = 0.0;
                c->n_launches = 0;
                rc = events_impl(c, box, op, ro, dep_xyzt, dep_nph, dep_event, ndep, nev, seed,
                                 0, maxb, em, (int32_t *)dm, (float *)dt, (int32_t *)dnpe,
                                 nullptr, false);
                if (rc) goto done;
                tbest = std::min(tbest, (float)c->kernel_ms);
            }
            if (times_ms) times_ms[a * n_ppt + b] = tbest;
            if (tbest < best) { best = tbest; bw = wgs[a]; bp = ppts[b]; }
        }
    }
done:
    c->timing = old_timing;
    c->kernel_ms = old_ms;
    c->n_launches = old_n;
    if (rc) { c->wg = old_wg; c->ppt = old_ppt; return rc; }
    c->wg = bw;
    c->ppt = bp;
    if (best_wg) *best_wg = bw;
    if (best_ppt) *best_ppt = bp;
    return CRYSP_OK;
}

}  // extern "C"

// The launch-knob sweep shared by the tuners: every (wg, ppt, chunk, refill_min) of the
// lists, `reps` timed calls each (best of reps kept), the fastest combination left in the
// context.  run() makes one call on the context's own device buffers.
template <class Run>
static int autotune_sweep(crysp_ctx *c, const int32_t *wgs, int32_t n_wg, const int32_t *ppts,
                          int32_t n_ppt, const int32_t *chunks, int32_t n_chunk,
                          const int32_t *refills, int32_t n_refill, int32_t reps, float *times_ms,
                          int32_t *best4, Run run)
{
    if (!wgs || !ppts || !chunks || !refills || n_wg < 1 || n_ppt < 1 || n_chunk < 1 ||
        n_refill < 1 || reps < 1)
        return fail(CRYSP_EINVAL, "bad autotune arguments");
    const int old_wg = c->wg, old_ppt = c->ppt, old_chunk = c->chunk, old_refill = c->refill_min;
    const bool old_timing = c->timing;
    const double old_ms = c->kernel_ms;
    const int64_t old_n = c->n_launches;
    float best = 1e30f;
    int bw = old_wg, bp = old_ppt, bc = old_chunk, br = old_refill;
    int rc = CRYSP_OK;
    c->timing = true;
    int64_t idx = 0;
    for (int a = 0; a < n_wg && rc == CRYSP_OK; ++a)
        for (int b = 0; b < n_ppt && rc == CRYSP_OK; ++b)
            for (int q = 0; q < n_chunk && rc == CRYSP_OK; ++q)
                for (int r = 0; r < n_refill && rc == CRYSP_OK; ++r, ++idx) {
                    if (wgs[a] >= 64 && wgs[a] <= 1024 &&
                        ring_bytes(wgs[a]) + emb_bytes(wgs[a]) + STATIC_LDS_MARGIN > c->lds_per_block) {
                        if (times_ms) times_ms[idx] = INFINITY;     // does not fit: skipped
                        continue;
                    }
                    if ((rc = crysp_set_schedule(c, wgs[a], ppts[b])) ||
                        (rc = crysp_set_queue(c, chunks[q], refills[r])))
                        break;
                    float tbest = 1e30f;
                    for (int k = 0; k < reps && rc == CRYSP_OK; ++k) {
                        c->kernel_ms = 0.0;
                        c->n_launches = 0;
                        if ((rc = run())) break;
                        tbest = std::min(tbest, (float)c->kernel_ms);
                    }
                    if (rc) break;
                    if (times_ms) times_ms[idx] = tbest;
                    if (tbest < best) { best = tbest; bw = wgs[a]; bp = ppts[b]; bc = chunks[q]; br = refills[r]; }
                }
    c->timing = old_timing;
    c->kernel_ms = old_ms;
    c->n_launches = old_n;
    if (rc) {
        c->wg = old_wg; c->ppt = old_ppt; c->chunk = old_chunk; c->refill_min = old_refill;
        return rc;
    }
    c->wg = bw; c->ppt = bp; c->chunk = bc; c->refill_min = br;
    if (best4) { best4[0] = bw; best4[1] = bp; best4[2] = bc; best4[3] = br; }
    return CRYSP_OK;
}

extern "C" {

int crysp_autotune_events_queue(crysp_ctx *c, const crysp_box *box, const crysp_optical_params *op,
                                const crysp_readout *ro, const float *dep_xyzt,
                                const int32_t *dep_nph, const int32_t *dep_event, int64_t ndep,
                                int64_t nev, uint64_t seed, int32_t maxb, const crysp_emission *em,
                                const int32_t *wgs, int32_t n_wg, const int32_t *ppts, int32_t n_ppt,
                                const int32_t *chunks, int32_t n_chunk, const int32_t *refills,
                                int32_t n_refill, int32_t reps, float *times_ms, int32_t *best4)
{
    if (!c || !ro) return fail(CRYSP_EINVAL, "bad autotune arguments");
    const int nxy = ro->grid.nx * ro->grid.ny;
    const size_t cells = (size_t)std::max<int64_t>(1, nev) * std::max(1, nxy);
    void *dm, *dt, *dn;
    int rc;
    // outputs go to device scratch so the timing is the kernel's own
    if ((rc = ensure(c, B_MISC, cells * 4, &dm)) || (rc = ensure(c, B_MISC2, cells * 4, &dt)) ||
        (rc = ensure(c, B_NPE, (size_t)std::max<int64_t>(1, nev) * 4, &dn)))
        return rc;
    return autotune_sweep(c, wgs, n_wg, ppts, n_ppt, chunks, n_chunk, refills, n_refill, reps,
                          times_ms, best4, [&]() {
        return events_impl(c, box, op, ro, dep_xyzt, dep_nph, dep_event, ndep, nev, seed, 0, maxb,
                           em, (int32_t *)dm, (float *)dt, (int32_t *)dn, nullptr, false);
    });
}

int crysp_autotune_accumulate(crysp_ctx *c, const crysp_box *box, const crysp_optical_params *op,
                              const crysp_readout *ro, const crysp_time_binning *tb, const float *px,
                              const float *py, const float *pz, const float *pt, int32_t npos,
                              int64_t n, uint64_t seed, int32_t maxb, const crysp_emission *em,
                              const int32_t *wgs, int32_t n_wg, const int32_t *ppts, int32_t n_ppt,
                              const int32_t *chunks, int32_t n_chunk, const int32_t *refills,
                              int32_t n_refill, int32_t reps, float *times_ms, int32_t *best4)
{
    if (!c || !ro || !tb || tb->nbins < 0) return fail(CRYSP_EINVAL, "bad autotune arguments");
    const int64_t nxy = (int64_t)ro->grid.nx * ro->grid.ny;
    void *dc, *df;
    int rc;
    if ((rc = ensure(c, B_MISC, (size_t)nxy * (tb->nbins + 1) * 4, &dc)) ||
        (rc = ensure(c, B_MISC2, (size_t)nxy * 4, &df)))
        return rc;
    int64_t ctr[CRYSP_NCOUNTERS];
    return autotune_sweep(c, wgs, n_wg, ppts, n_ppt, chunks, n_chunk, refills, n_refill, reps,
                          times_ms, best4, [&]() {
        return crysp_transport_accumulate(c, box, op, ro, tb, px, py, pz, pt, npos, n, seed, 0, maxb,
                                          em, (uint32_t *)dc, (float *)df, ctr);
    });
}

int crysp_debug_draws(crysp_ctx *c, uint64_t seed, uint64_t pid0, int64_t n_pids,
                      int32_t n_draws, uint32_t *out)
{
    if (!c || !out || n_pids < 0 || n_draws < 0) return fail(CRYSP_EINVAL, "bad arguments");
    if (n_pids == 0 || n_draws == 0) return CRYSP_OK;
    HIPCHK(hipSetDevice(c->device));
    void *d;
    int rc = out_buf(c, B_MISC, out, (size_t)n_pids * n_draws * 4, &d);
    if (rc) return rc;
    hipLaunchKernelGGL(debug_draws_kernel, dim3((unsigned)((n_pids + 255) / 256)), dim3(256), 256 * 64,
                       c->stream, seed, pid0, n_pids, n_draws, (uint32_t *)d);
    HIPCHK(hipGetLastError());
    if ((rc = stage_out(c, out, d, (size_t)n_pids * n_draws * 4))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return CRYSP_OK;
}

int crysp_debug_math(crysp_ctx *c, int32_t which, const float *x, int64_t n, float *y)
{
    if (!c || !x || !y || n < 0 || which < 0 || which > 8) return fail(CRYSP_EINVAL, "bad arguments");
    if (n == 0) return CRYSP_OK;
    HIPCHK(hipSetDevice(c->device));
    const float *dx;
    void *dy;
    int rc;
    if ((rc = stage_in(c, B_MISC, x, (size_t)n * 4, (const void **)&dx)) ||
        (rc = out_buf(c, B_MISC2, y, (size_t)n * 4, &dy)))
        return rc;
    hipLaunchKernelGGL(debug_math_kernel, dim3(2048), dim3(256), 0, c->stream, which, dx, n,
                       (float *)dy);
    HIPCHK(hipGetLastError());
    if ((rc = stage_out(c, y, dy, (size_t)n * 4))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return CRYSP_OK;
}

}  // extern "C"
