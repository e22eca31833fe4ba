/*
 * crysp_oracle.c -- CPU restatement of CryspLight's optical-photon transport.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (crysplight_amd + libcrysplight_hip.so) never links it.
 *
 * It restates, draw for draw, the reference's Julia code:
 *   Philox4x32-10 ............ src/philox.jl:5-62, src/kernel.jl:15-46
 *   box geometry ............. src/geometry.jl:4-32
 *   surface / bulk optics .... src/optics.jl:10-138, src/kernel.jl:50-151
 *   emission ................. src/generation.jl:4-13, src/kernel.jl:175-180
 *   bounce loop + records .... src/kernel.jl:159-305 (== src/transport.jl:118-234)
 *   host reduction ........... src/kernel.jl:338-365, src/transport.jl:64-108
 *   per-event reduction ...... src/events.jl:73-88
 * plus one extension the reference does not have: a rise time (tau_rise > 0 adds a
 * second exponential draw AFTER the decay draw, so tau_rise == 0 consumes exactly the
 * reference's draws).
 *
 * Float32 semantics: every +,-,*,/ and sqrt is an IEEE binary32 operation (build with
 * -ffp-contract=off, no fast-math; x86-64 SSE evaluates float in float).  Julia's
 * Float32 log/sin/cos are evaluated through Float64 kernels and are correctly rounded
 * in practice; this oracle uses (float)f((double)x) with glibc's double routines, the
 * closest available stand-in.  Per-photon parity against Julia's own transcendentals
 * is therefore "parity unpinned" (Julia cannot run here); parity of the HIP kernels
 * against THIS oracle is what the tests enforce.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- layouts ---- */
/* Same byte layout as the Julia isbits structs (src/transport.jl:14-62,
 * src/geometry.jl:11-13) and as include/crysplight.h. */
typedef struct { float L[3]; } or_box;
typedef struct {
    float n_crystal, n_coupling, abs_len_mm, rayleigh_mm, wrap_R;
    uint8_t wrap_specular, air_gap, pad0_, pad1_;
    float sigma_alpha, pde;
} or_optical;
typedef struct { float pitch_mm; int32_t nx, ny; } or_grid;
typedef struct {
    or_grid grid;
    uint8_t disc, pad0_[3];
    float cx, cy, r2, sur_R;
    uint8_t sur_specular, two_sided, pad1_[2];
} or_readout;
typedef struct { float bin_ns; int32_t nbins; } or_tbin;

/* operation counters feeding the roofline's algorithmic-op model */
typedef struct {
    int64_t n_photon, n_draw, n_block, n_step, n_fresnel, n_surface_try,
            n_lambert, n_rayleigh_try, n_detect, n_nscat;
} or_opcount;

enum { ST_DETECTED = 1, ST_ABS_BULK = 2, ST_ABS_WALL = 3, ST_ABS_SIPM = 4,
       ST_CAP = 5, ST_ABS_SUR = 6 };                     /* src/transport.jl:6-12 */
enum { F_XP = 1, F_XM = 2, F_YP = 3, F_YM = 4, F_BACK = 5, F_FRONT = 6 }; /* geometry.jl:4-9 */

static const float C_MM_NS = 299.792458f;               /* src/transport.jl:4 */
static const float TWO_PI_F = 6.2831855f;               /* 2f0 * Float32(pi), exact */

/* ------------------------------------------------------------- Philox ------ */
/* src/philox.jl:10-28: ten rounds, key bumped after rounds 1..9. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        if (r < 9) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* src/philox.jl:34-57: key = (seed lo, seed hi), counter = (pid lo, pid hi, block, 0);
 * one block gives four words consumed in order. */
typedef struct {
    uint32_t key[2], id_lo, id_hi, block, buf[4];
    int i;
    or_opcount *oc;
} or_stream;

static void st_init(or_stream *s, uint64_t seed, uint64_t pid, or_opcount *oc)
{
    s->key[0] = (uint32_t)seed; s->key[1] = (uint32_t)(seed >> 32);
    s->id_lo = (uint32_t)pid;   s->id_hi = (uint32_t)(pid >> 32);
    s->block = 0; s->i = 4; s->oc = oc;
}

static inline uint32_t st_next(or_stream *s)
{
    if (s->i > 3) {
        uint32_t ctr[4] = { s->id_lo, s->id_hi, s->block, 0u };
        or_philox4x32_10(ctr, s->key, s->buf);
        s->block += 1u; s->i = 0;
        if (s->oc) s->oc->n_block++;
    }
    if (s->oc) s->oc->n_draw++;
    return s->buf[s->i++];
}

/* src/philox.jl:62: ((x >>> 8) + 1) * 2^-24, exact, in (0, 1]. */
static inline float u32_to_unit(uint32_t x)
{
    return ((float)(x >> 8) + 1.0f) * 5.9604645e-08f;
}
static inline float randu(or_stream *s) { return u32_to_unit(st_next(s)); }

/* ---------------------------------------------------- Julia-float helpers ---- */
static inline float jlog(float x) { return (float)log((double)x); }
static inline float jsin(float x) { return (float)sin((double)x); }
static inline float jcos(float x) { return (float)cos((double)x); }
/* Base.max(0f0, v): NaN propagates, -0.0 -> +0.0 */
static inline float jmax0(float v) { return isnan(v) ? v : (v > 0.0f ? v : 0.0f); }
/* Base.clamp: ifelse(x > hi, hi, ifelse(x < lo, lo, x)) */
static inline float jclamp(float x, float lo, float hi) { return x > hi ? hi : (x < lo ? lo : x); }

/* ------------------------------------------------------------ geometry ---- */
/* src/geometry.jl:19-32, tie order x -> y -> z */
static inline float wall_hit(float x, float y, float z, float ux, float uy, float uz,
                             const float L[3], int *face)
{
    const float inf = INFINITY;
    float dx = ux > 0.0f ? (L[0] - x) / ux : (ux < 0.0f ? -x / ux : inf);
    float dy = uy > 0.0f ? (L[1] - y) / uy : (uy < 0.0f ? -y / uy : inf);
    float dz = uz > 0.0f ? (L[2] - z) / uz : (uz < 0.0f ? -z / uz : inf);
    if (dx <= dy && dx <= dz) { *face = ux > 0.0f ? F_XP : F_XM; return dx; }
    if (dy <= dz)             { *face = uy > 0.0f ? F_YP : F_YM; return dy; }
    *face = uz > 0.0f ? F_BACK : F_FRONT; return dz;
}

static inline int face_axis(int f) { return f <= 2 ? 1 : (f <= 4 ? 2 : 3); }

/* src/optics.jl:10-19 (x^2 is x*x: Base.literal_pow) */
static inline int fresnel(float n1, float n2, float cosi, float *R)
{
    float q = n1 / n2;
    float sin2t = (q * q) * (1.0f - cosi * cosi);
    if (sin2t >= 1.0f) { *R = 1.0f; return 1; }
    float cost = sqrtf(1.0f - sin2t);
    float a = (n1 * cosi - n2 * cost) / (n1 * cosi + n2 * cost);
    float b = (n1 * cost - n2 * cosi) / (n1 * cost + n2 * cosi);
    *R = 0.5f * (a * a + b * b);
    return 0;
}

/* src/optics.jl:25-28 */
static inline void specular(float *ux, float *uy, float *uz, int face)
{
    int a = face_axis(face);
    if (a == 1) *ux = -*ux; else if (a == 2) *uy = -*uy; else *uz = -*uz;
}

/* src/optics.jl:118-138 / src/kernel.jl:65-85 */
static void lambertian(or_stream *s, int face, float *ux, float *uy, float *uz)
{
    if (s->oc) s->oc->n_lambert++;
    float ct = sqrtf(randu(s));
    float st = sqrtf(jmax0(1.0f - ct * ct));
    float phi = TWO_PI_F * randu(s);
    float a = st * jcos(phi), b = st * jsin(phi);
    switch (face) {
    case F_XP: *ux = -ct; *uy = a;   *uz = b;   break;
    case F_XM: *ux = ct;  *uy = a;   *uz = b;   break;
    case F_YP: *ux = a;   *uy = -ct; *uz = b;   break;
    case F_YM: *ux = a;   *uy = ct;  *uz = b;   break;
    case F_BACK: *ux = a; *uy = b;   *uz = -ct; break;
    default:   *ux = a;   *uy = b;   *uz = ct;  break;
    }
}

/* src/optics.jl:87-112 / src/kernel.jl:87-114 */
static void rayleigh(or_stream *s, float *pux, float *puy, float *puz)
{
    float ux = *pux, uy = *puy, uz = *puz, mu;
    for (;;) {
        if (s->oc) s->oc->n_rayleigh_try++;
        mu = 1.0f - 2.0f * randu(s);
        float u2 = randu(s);
        if (2.0f * u2 <= 1.0f + mu * mu) break;
    }
    float st = sqrtf(jmax0(1.0f - mu * mu));
    float phi = TWO_PI_F * randu(s);
    float cb = st * jcos(phi), sb = st * jsin(phi);
    float t1x, t1y, t1z, inv;
    if (fabsf(uz) < 0.9f) {
        inv = 1.0f / sqrtf(ux * ux + uy * uy);
        t1x = -uy * inv; t1y = ux * inv; t1z = 0.0f;
    } else {
        inv = 1.0f / sqrtf(uy * uy + uz * uz);
        t1x = 0.0f; t1y = -uz * inv; t1z = uy * inv;
    }
    float t2x = uy * t1z - uz * t1y;
    float t2y = uz * t1x - ux * t1z;
    float t2z = ux * t1y - uy * t1x;
    float vx = mu * ux + cb * t1x + sb * t2x;
    float vy = mu * uy + cb * t1y + sb * t2y;
    float vz = mu * uz + cb * t1z + sb * t2z;
    inv = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
    *pux = vx * inv; *puy = vy * inv; *puz = vz * inv;
}

/* src/optics.jl:49-81 / src/kernel.jl:116-151.  Returns 1 if transmitted. */
static int surface(or_stream *s, float *pux, float *puy, float *puz, int face,
                   float n_crystal, float sigma_alpha)
{
    float ux = *pux, uy = *puy, uz = *puz;
    int a = face_axis(face);
    float sgn = (face == F_XP || face == F_YP || face == F_BACK) ? 1.0f : -1.0f;
    float Nx = a == 1 ? sgn : 0.0f, Ny = a == 2 ? sgn : 0.0f, Nz = a == 3 ? sgn : 0.0f;
    for (int k = 0; k < 20; ++k) {
        if (s->oc) s->oc->n_surface_try++;
        /* gaussian (src/optics.jl:38-39): log draw first, then the cos draw */
        float u1 = randu(s);
        float u2 = randu(s);
        float g = sqrtf(-2.0f * jlog(u1)) * jcos(TWO_PI_F * u2);
        float alpha = sigma_alpha * fabsf(g);
        if (alpha >= 1.2f) continue;
        float ca = jcos(alpha), sa = jsin(alpha);
        float phi = TWO_PI_F * randu(s);
        float b = sa * jcos(phi), c = sa * jsin(phi);
        float Nfx = a == 1 ? ca * sgn : b;
        float Nfy = a == 2 ? ca * sgn : (a == 1 ? b : c);
        float Nfz = a == 3 ? ca * sgn : c;
        float cosi = ux * Nfx + uy * Nfy + uz * Nfz;
        if (cosi <= 0.0f) continue;
        float R;
        if (s->oc) s->oc->n_fresnel++;
        int tir = fresnel(n_crystal, 1.0f, cosi, &R);
        int refl = tir;
        if (!tir) refl = randu(s) < R;
        if (refl) {
            float rx = ux - 2.0f * cosi * Nfx;
            float ry = uy - 2.0f * cosi * Nfy;
            float rz = uz - 2.0f * cosi * Nfz;
            if (rx * Nx + ry * Ny + rz * Nz < 0.0f) {
                *pux = rx; *puy = ry; *puz = rz; return 0;
            }
        } else {
            return 1;
        }
    }
    specular(pux, puy, puz, face);
    return 0;
}

/* ------------------------------------------------------------ emission ---- */
/* src/generation.jl:4-9 / src/kernel.jl:50-57 */
static void isotropic(or_stream *s, float *ux, float *uy, float *uz)
{
    float u1 = randu(s);
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(jmax0(1.0f - z * z));
    float phi = TWO_PI_F * randu(s);
    *ux = r * jcos(phi); *uy = r * jsin(phi); *uz = z;
}

/* src/kernel.jl:176-180 (src/generation.jl:12-13) + the rise-time extension */
static float emission(or_stream *s, float t0, float tau_d, float tau_r)
{
    float t = t0;
    if (tau_d > 0.0f) t += -tau_d * jlog(randu(s));
    if (tau_r > 0.0f) t += -tau_r * jlog(randu(s));
    return t;
}

/* ------------------------------------------------------------ transport ---- */
typedef struct { uint8_t status; int16_t idx; float t; int32_t bounces; int32_t nscat; } or_rec;

/* The bounce loop: src/kernel.jl:185-299 (identical draw order to
 * src/transport.jl:126-233).  (x,y,z,ux,uy,uz,t) is the emitted photon. */
static or_rec propagate(or_stream *s, const or_box *box, const or_optical *op,
                        const or_readout *ro, float x, float y, float z,
                        float ux, float uy, float uz, float t, int32_t maxb)
{
    const float *L = box->L;
    const float inv_v = op->n_crystal / C_MM_NS;
    const int abs_on = isfinite(op->abs_len_mm), ray_on = isfinite(op->rayleigh_mm);
    int32_t bounces = 0, nscat = 0;
    int result = 0;
    int16_t out_idx = 0;
    while (result == 0) {
        if (s->oc) s->oc->n_step++;
        int face;
        float d = wall_hit(x, y, z, ux, uy, uz, L, &face);
        float d_abs = INFINITY, d_ray = INFINITY;
        if (abs_on) d_abs = -op->abs_len_mm * jlog(randu(s));
        if (ray_on) d_ray = -op->rayleigh_mm * jlog(randu(s));
        if (d_abs < d && d_abs <= d_ray) {
            result = ST_ABS_BULK;
        } else if (d_ray < d) {
            x += d_ray * ux; y += d_ray * uy; z += d_ray * uz;
            t += d_ray * inv_v;
            rayleigh(s, &ux, &uy, &uz);
            nscat++;
            bounces++;
            if (bounces >= maxb) result = ST_CAP;
        } else {
            x = jclamp(x + d * ux, 0.0f, L[0]);
            y = jclamp(y + d * uy, 0.0f, L[1]);
            z = jclamp(z + d * uz, 0.0f, L[2]);
            int a = face_axis(face);
            if (a == 1)      x = face == F_XP ? L[0] : 0.0f;
            else if (a == 2) y = face == F_YP ? L[1] : 0.0f;
            else             z = face == F_BACK ? L[2] : 0.0f;
            t += d * inv_v;
            float cosi = a == 1 ? fabsf(ux) : (a == 2 ? fabsf(uy) : fabsf(uz));

            if (face == F_BACK || (ro->two_sided && face == F_FRONT)) {
                float R;
                if (s->oc) s->oc->n_fresnel++;
                int tir = fresnel(op->n_crystal, op->n_coupling, cosi, &R);
                int refl = tir;
                if (!tir) refl = randu(s) < R;
                if (refl) {
                    specular(&ux, &uy, &uz, face);
                } else {
                    float ex = x - ro->cx, ey = y - ro->cy;
                    if (!ro->disc || ex * ex + ey * ey <= ro->r2) {
                        if (randu(s) <= op->pde) {
                            const float pitch = ro->grid.pitch_mm;
                            int32_t ii = 1 + (int32_t)(x / pitch);
                            int32_t jj = 1 + (int32_t)(y / pitch);
                            if (ii > ro->grid.nx) ii = ro->grid.nx;
                            if (jj > ro->grid.ny) jj = ro->grid.ny;
                            int32_t lin = (jj - 1) * ro->grid.nx + ii;
                            out_idx = (int16_t)(face == F_FRONT ? -lin : lin);
                            result = ST_DETECTED;
                            if (s->oc) s->oc->n_detect++;
                        } else {
                            result = ST_ABS_SIPM;
                        }
                    } else if (randu(s) <= ro->sur_R) {
                        if (ro->sur_specular) specular(&ux, &uy, &uz, face);
                        else lambertian(s, face, &ux, &uy, &uz);
                    } else {
                        result = ST_ABS_SUR;
                    }
                }
            } else if (!op->air_gap) {
                if (randu(s) <= op->wrap_R) {
                    if (op->wrap_specular) specular(&ux, &uy, &uz, face);
                    else lambertian(s, face, &ux, &uy, &uz);
                } else {
                    result = ST_ABS_WALL;
                }
            } else {
                int transmitted = 0;
                if (op->sigma_alpha > 0.0f) {
                    transmitted = surface(s, &ux, &uy, &uz, face, op->n_crystal, op->sigma_alpha);
                } else {
                    float R;
                    if (s->oc) s->oc->n_fresnel++;
                    int tir = fresnel(op->n_crystal, 1.0f, cosi, &R);
                    int refl = tir;
                    if (!tir) refl = randu(s) < R;
                    if (refl) specular(&ux, &uy, &uz, face);
                    else transmitted = 1;
                }
                if (transmitted) {
                    if (randu(s) <= op->wrap_R) {
                        if (op->wrap_specular) specular(&ux, &uy, &uz, face);
                        else lambertian(s, face, &ux, &uy, &uz);
                    } else {
                        result = ST_ABS_WALL;
                    }
                }
            }
            if (result == 0) {
                bounces++;
                if (bounces >= maxb) result = ST_CAP;
            }
        }
    }
    or_rec r = { (uint8_t)result, out_idx, t, bounces, nscat };
    return r;
}

static void add_counts(or_opcount *dst, const or_opcount *src)
{
    dst->n_photon += src->n_photon; dst->n_draw += src->n_draw; dst->n_block += src->n_block;
    dst->n_step += src->n_step; dst->n_fresnel += src->n_fresnel;
    dst->n_surface_try += src->n_surface_try; dst->n_lambert += src->n_lambert;
    dst->n_rayleigh_try += src->n_rayleigh_try; dst->n_detect += src->n_detect;
    dst->n_nscat += src->n_nscat;
}

/* ------------------------------------------------------------ exports ---- */

/* First n_draws u32 of the streams pid0+1 .. pid0+n_pids (philox.jl:43-57). */
void or_debug_draws(uint64_t seed, uint64_t pid0, int64_t n_pids, int32_t n_draws,
                    uint32_t *out)
{
    for (int64_t k = 0; k < n_pids; ++k) {
        or_stream s;
        st_init(&s, seed, pid0 + (uint64_t)k + 1u, NULL);
        for (int32_t d = 0; d < n_draws; ++d) out[k * n_draws + d] = st_next(&s);
    }
}

/* randu over an array of raw u32 (closed-form check of philox.jl:62) */
void or_u32_to_unit(const uint32_t *x, int64_t n, float *u)
{
    for (int64_t k = 0; k < n; ++k) u[k] = u32_to_unit(x[k]);
}

void or_wall_hit(const float *p, const float *L, float *d, int32_t *face)
{
    int f;
    *d = wall_hit(p[0], p[1], p[2], p[3], p[4], p[5], L, &f);
    *face = f;
}

int32_t or_fresnel(float n1, float n2, float cosi, float *R)
{
    return fresnel(n1, n2, cosi, R);
}

/* Samplers driven by ONE stream (seed, pid), n successive calls, for the moment
 * tests of test/runtests.jl:65-100.  kind: 0 isotropic, 1 lambertian(face),
 * 2 rayleigh about (a0,a1,a2), 3 surface(face, n=a0, sigma=a1, dir=a2..a4), 4 emission
 * (tau_d=a0, tau_r=a1).  out gets 4 floats per call. */
void or_sample(int32_t kind, uint64_t seed, uint64_t pid, int64_t n, int32_t face,
               const float *a, float *out)
{
    or_stream s;
    st_init(&s, seed, pid, NULL);
    for (int64_t k = 0; k < n; ++k) {
        float v[4] = { 0, 0, 0, 0 };
        switch (kind) {
        case 0: isotropic(&s, &v[0], &v[1], &v[2]); break;
        case 1: lambertian(&s, face, &v[0], &v[1], &v[2]); break;
        case 2: v[0] = a[0]; v[1] = a[1]; v[2] = a[2]; rayleigh(&s, &v[0], &v[1], &v[2]); break;
        case 3: v[0] = a[2]; v[1] = a[3]; v[2] = a[4];
                v[3] = (float)surface(&s, &v[0], &v[1], &v[2], face, a[0], a[1]); break;
        default: v[0] = emission(&s, 0.0f, a[0], a[1]); break;
        }
        memcpy(out + 4 * k, v, sizeof v);
    }
}

/* One photon with a GIVEN start (src/transport.jl:118 usage in the tests):
 * stream (seed, pid), no emission draws. */
void or_propagate_one(const or_box *box, const or_optical *op, const or_readout *ro,
                      uint64_t seed, uint64_t pid, const float *xyz_u_t, int32_t maxb,
                      uint8_t *status, int16_t *idx, float *tdet, int32_t *bnc)
{
    or_stream s;
    st_init(&s, seed, pid, NULL);
    const float *p = xyz_u_t;
    or_rec r = propagate(&s, box, op, ro, p[0], p[1], p[2], p[3], p[4], p[5], p[6], maxb);
    *status = r.status; *idx = r.idx; *tdet = r.t; *bnc = r.bounces;
}

/* Records mode: src/kernel.jl:159-305 over ndrange = n.  Photon i (1-based) uses
 * stream pid = i + pid0 and position j = 1 + (i-1) % npos.  nthreads <= 0: all.
 * counts (optional) receives the summed operation counters. */
void or_transport_records(const or_box *box, const or_optical *op, const or_readout *ro,
                          const float *px, const float *py, const float *pz,
                          const float *pt, int32_t npos, int64_t n, uint64_t seed,
                          uint64_t pid0, int32_t maxb, float tau_d, float tau_r,
                          uint8_t *status, int16_t *idx, float *tdet, int32_t *bnc,
                          int64_t *nscat_total, or_opcount *counts, int32_t nthreads)
{
    int64_t nscat = 0;
    or_opcount tot;
    memset(&tot, 0, sizeof tot);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+:nscat)
#endif
    {
        or_opcount loc;
        memset(&loc, 0, sizeof loc);
        or_opcount *ocp = counts ? &loc : NULL;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
        for (int64_t k = 0; k < n; ++k) {
            or_stream s;
            st_init(&s, seed, (uint64_t)k + 1u + pid0, ocp);
            int64_t j = k % npos;
            float ux, uy, uz;
            isotropic(&s, &ux, &uy, &uz);
            float t = emission(&s, pt ? pt[j] : 0.0f, tau_d, tau_r);
            or_rec r = propagate(&s, box, op, ro, px[j], py[j], pz[j], ux, uy, uz, t, maxb);
            if (status) status[k] = r.status;
            if (idx) idx[k] = r.idx;
            if (tdet) tdet[k] = r.t;
            if (bnc) bnc[k] = r.bounces;
            nscat += r.nscat;
            if (ocp) { loc.n_photon++; loc.n_nscat += r.nscat; }
        }
        if (counts) {
#ifdef _OPENMP
#pragma omp critical
#endif
            add_counts(&tot, &loc);
        }
    }
    if (nscat_total) *nscat_total = nscat;
    if (counts) *counts = tot;
}

/* Host reduction of records into an Accumulator: src/kernel.jl:338-365.
 * counts: u32 [nx*ny*(nbins+1)], element (ii, jj, bin) at (ii-1) + nx*((jj-1) + ny*(bin-1));
 * first_ns: f32 [nx*ny] (+Inf where nothing);  ctr[9] = ndet, ndet_front, nabs_bulk,
 * nabs_wall, nabs_sipm, ncap, nabs_sur, nscat(unset here), sum_bounces_det. */
void or_reduce_records(const or_grid *grid, const or_tbin *tb, int64_t n,
                       const uint8_t *status, const int16_t *idx, const float *tdet,
                       const int32_t *bnc, uint32_t *counts, float *first_ns, int64_t *ctr)
{
    const int64_t nx = grid->nx, ny = grid->ny;
    memset(counts, 0, sizeof(uint32_t) * (size_t)(nx * ny * (tb->nbins + 1)));
    for (int64_t k = 0; k < nx * ny; ++k) first_ns[k] = INFINITY;
    for (int k = 0; k < 9; ++k) ctr[k] = 0;
    for (int64_t k = 0; k < n; ++k) {
        int s = status[k];
        if (s == ST_DETECTED) {
            int64_t lin = idx[k];
            if (lin < 0) { ctr[1]++; lin = -lin; }
            int64_t ii = (lin - 1) % nx, jj = (lin - 1) / nx;
            float t = tdet[k];
            double q = (double)(t / tb->bin_ns);     /* unsafe_trunc(Int, t / bin_ns) */
            int64_t bin = q >= (double)tb->nbins ? (int64_t)tb->nbins + 1 : 1 + (int64_t)q;
            if (bin > tb->nbins) bin = tb->nbins + 1;
            if (bin < 1) bin = 1;                    /* negative times: first bin */
            counts[ii + nx * (jj + ny * (bin - 1))] += 1u;
            if (t < first_ns[ii + nx * jj]) first_ns[ii + nx * jj] = t;
            ctr[0]++;
            ctr[8] += bnc[k];
        } else if (s == ST_ABS_BULK) ctr[2]++;
        else if (s == ST_ABS_WALL) ctr[3]++;
        else if (s == ST_ABS_SIPM) ctr[4]++;
        else if (s == ST_ABS_SUR) ctr[6]++;
        else ctr[5]++;
    }
}

/* Event mode.  The photons of deposit k are nph[k] consecutive photons in run order
 * (src/events.jl:109-116: deposit-major within an event, events in order); photon g
 * (0-based over the whole call) uses stream pid = pid0 + g + 1 and starts at its
 * deposit's (x, y, z) with t0 = deposit time.  dep_event[k] is the deposit's event.
 * Per-event reduction: src/events.jl:73-87 (maps summed over both faces via |idx|).
 * maps: i32 [nev*nx*ny] ((ii-1) + nx*(jj-1) + nx*ny*ev); tmin: f32 (+Inf = none);
 * npe: i32 [nev].  Records of every photon are also returned when the pointers are set;
 * counts (optional) receives the summed operation counters. */
void or_run_events(const or_box *box, const or_optical *op, const or_readout *ro,
                   const float *dep_xyzt, const int32_t *dep_nph, const int32_t *dep_event,
                   int64_t ndep, int64_t nev, uint64_t seed, uint64_t pid0, int32_t maxb,
                   float tau_d, float tau_r, int32_t *maps, float *tmin, int32_t *npe,
                   uint8_t *rec_status, int16_t *rec_idx, float *rec_t, or_opcount *counts,
                   int32_t nthreads)
{
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ndep + 1));
    off[0] = 0;
    for (int64_t k = 0; k < ndep; ++k) off[k + 1] = off[k] + dep_nph[k];
    const int64_t ntot = off[ndep];
    uint8_t *st = rec_status ? rec_status : (uint8_t *)malloc((size_t)ntot + 1);
    int16_t *ix = rec_idx ? rec_idx : (int16_t *)malloc(sizeof(int16_t) * (size_t)(ntot + 1));
    float *tt = rec_t ? rec_t : (float *)malloc(sizeof(float) * (size_t)(ntot + 1));
    or_opcount tot;
    memset(&tot, 0, sizeof tot);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        or_opcount loc;
        memset(&loc, 0, sizeof loc);
        or_opcount *ocp = counts ? &loc : NULL;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t k = 0; k < ndep; ++k) {
            const float *d = dep_xyzt + 4 * k;
            for (int64_t g = off[k]; g < off[k + 1]; ++g) {
                or_stream s;
                st_init(&s, seed, pid0 + (uint64_t)g + 1u, ocp);
                float ux, uy, uz;
                isotropic(&s, &ux, &uy, &uz);
                float t = emission(&s, d[3], tau_d, tau_r);
                or_rec r = propagate(&s, box, op, ro, d[0], d[1], d[2], ux, uy, uz, t, maxb);
                st[g] = r.status; ix[g] = r.idx; tt[g] = r.t;
                if (ocp) { loc.n_photon++; loc.n_nscat += r.nscat; }
            }
        }
        if (counts) {
#ifdef _OPENMP
#pragma omp critical
#endif
            add_counts(&tot, &loc);
        }
    }
    if (counts) *counts = tot;
    const int64_t nx = ro->grid.nx, nxy = (int64_t)ro->grid.nx * ro->grid.ny;
    memset(maps, 0, sizeof(int32_t) * (size_t)(nev * nxy));
    memset(npe, 0, sizeof(int32_t) * (size_t)nev);
    for (int64_t k = 0; k < nev * nxy; ++k) tmin[k] = INFINITY;
    for (int64_t k = 0; k < ndep; ++k) {
        const int64_t ev = dep_event[k];
        for (int64_t g = off[k]; g < off[k + 1]; ++g) {
            if (st[g] != ST_DETECTED) continue;
            int64_t lin = ix[g] < 0 ? -(int64_t)ix[g] : ix[g];
            int64_t cell = (lin - 1) % nx + nx * ((lin - 1) / nx) + nxy * ev;
            maps[cell] += 1;
            if (tt[g] < tmin[cell]) tmin[cell] = tt[g];
            npe[ev] += 1;
        }
    }
    if (!rec_status) free(st);
    if (!rec_idx) free(ix);
    if (!rec_t) free(tt);
    free(off);
}

/* The oracle's Float32 transcendentals over an array (which: 0 log, 1 sin, 2 cos). */
void or_math(int32_t which, const float *x, int64_t n, float *y)
{
    for (int64_t k = 0; k < n; ++k)
        y[k] = which == 0 ? jlog(x[k]) : (which == 1 ? jsin(x[k]) : jcos(x[k]));
}

/* Struct sizes, for the layout check against the Julia isbits layout. */
void or_layout(int32_t *sizes)
{
    sizes[0] = (int32_t)sizeof(or_box);
    sizes[1] = (int32_t)sizeof(or_optical);
    sizes[2] = (int32_t)sizeof(or_grid);
    sizes[3] = (int32_t)sizeof(or_readout);
    sizes[4] = (int32_t)sizeof(or_tbin);
    sizes[5] = (int32_t)sizeof(or_opcount);
}

int32_t or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
