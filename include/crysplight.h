/*
 * crysplight.h -- C-ABI of libcrysplight_hip.so, the MI355X (gfx950) optical-photon
 * transport of CryspLight.
 *
 * Conventions
 *  - Every entry point returns 0 (CRYSP_OK) or a negative crysp_error; nothing throws
 *    across the ABI.  crysp_last_error() gives the message of the calling thread's
 *    last failure.
 *  - Array arguments are plain pointers.  Each one may be host memory (pageable or
 *    pinned; the library stages it through its own device buffers) or device memory
 *    from hipMalloc / a torch tensor on the context's device (used in place).  All
 *    calls are synchronous, as the reference's KA.synchronize (src/kernel.jl:335).
 *  - Ordering: the library runs its kernels on a stream of its own and does not wait for
 *    work the caller queued on other streams.  Device arrays must hold their final
 *    contents when a call starts: wait for their producers first (for torch tensors,
 *    synchronize torch's current stream).  When a call returns, its outputs are complete.
 *  - Device-resident index tables (event and crystal indices) are checked on the device;
 *    an index out of range makes the call return CRYSP_EINVAL after its launch.
 *  - Photon ids follow the reference: photon k (1-based within a call) draws from the
 *    Philox stream (seed, pid = k + pid0) (src/kernel.jl:162-163).
 *  - Multi-dimensional outputs are column-major like the Julia arrays they replace:
 *    counts (nx, ny, nbins+1): element (i, j, b) (1-based) at (i-1) + nx*((j-1) + ny*(b-1));
 *    per-event maps (nx, ny, nev): (i-1) + nx*(j-1) + nx*ny*(ev-1).
 *  - One host thread per context at a time (the reference launches from one task).
 */
#ifndef CRYSPLIGHT_H
#define CRYSPLIGHT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRYSPLIGHT_ABI_VERSION 1

/* ---- parameter structs: byte-identical to the reference's isbits structs ---- */

/* Box (src/geometry.jl:11-13): crystal [0,Lx]x[0,Ly]x[0,Lz] in mm, SiPMs at z = Lz. */
typedef struct { float L[3]; } crysp_box;

/* OpticalParams (src/transport.jl:14-24), 32 bytes: bools at offsets 20, 21. */
typedef struct {
    float   n_crystal;
    float   n_coupling;      /* grease (or gap) index at the SiPM face               */
    float   abs_len_mm;      /* bulk absorption length; +Inf disables (no draw)      */
    float   rayleigh_mm;     /* Rayleigh length; +Inf disables (no draw)             */
    float   wrap_R;          /* wrap reflectivity                                    */
    uint8_t wrap_specular;   /* 1 = ESR (specular), 0 = PTFE (Lambertian)            */
    uint8_t air_gap;         /* 1 = back-painted (air gap), 0 = front-painted        */
    uint8_t pad_[2];
    float   sigma_alpha;     /* UNIFIED roughness [rad] of the crystal-air interface */
    float   pde;             /* photon detection efficiency at the readout face      */
} crysp_optical_params;

/* SipmGrid (src/transport.jl:26-30). */
typedef struct { float pitch_mm; int32_t nx, ny; } crysp_sipm_grid;

/* Readout (src/transport.jl:41-50), 36 bytes. */
typedef struct {
    crysp_sipm_grid grid;
    uint8_t disc;            /* 1 = only a disc of radius^2 r2 about (cx, cy) is active */
    uint8_t pad0_[3];
    float   cx, cy, r2;
    float   sur_R;           /* surround reflectivity outside the disc               */
    uint8_t sur_specular;
    uint8_t two_sided;       /* readout on both z faces                              */
    uint8_t pad1_[2];
} crysp_readout;

/* TimeBinning (src/transport.jl:59-62): nbins + 1 bins, the last is overflow. */
typedef struct { float bin_ns; int32_t nbins; } crysp_time_binning;

/* Emission model.  tau_decay > 0: t += Exp(tau_decay) (one draw, src/kernel.jl:177-180).
 * tau_rise > 0 (no reference counterpart): a second draw adds Exp(tau_rise), so the
 * emission pdf is (e^{-t/td} - e^{-t/tr}) / (td - tr); tau_rise = 0 draws as the reference. */
typedef struct { float tau_decay; float tau_rise; } crysp_emission;

/* Photon fates (src/transport.jl:6-12). */
enum crysp_status {
    CRYSP_DETECTED = 1, CRYSP_ABS_BULK = 2, CRYSP_ABS_WALL = 3,
    CRYSP_ABS_SIPM = 4, CRYSP_CAP = 5, CRYSP_ABS_SUR = 6
};

enum crysp_error {
    CRYSP_OK = 0,
    CRYSP_EINVAL = -1,   /* bad argument (null pointer, npos < 1, ...)               */
    CRYSP_EHIP = -2,     /* a HIP runtime call failed                                */
    CRYSP_ENOMEM = -3,   /* device allocation failed                                 */
    CRYSP_ERANGE = -4    /* size outside what the ABI supports (nx*ny > 32767, ...)  */
};

/* Indices of the int64 counters[] array of crysp_transport_accumulate, the fields of
 * the reference Accumulator (src/transport.jl:64-76). */
enum crysp_counter {
    CRYSP_C_NDET = 0, CRYSP_C_NDET_FRONT, CRYSP_C_NABS_BULK, CRYSP_C_NABS_WALL,
    CRYSP_C_NABS_SIPM, CRYSP_C_NCAP, CRYSP_C_NABS_SUR, CRYSP_C_NSCAT,
    CRYSP_C_SUM_BOUNCES_DET, CRYSP_NCOUNTERS
};

typedef struct crysp_ctx crysp_ctx;

/* Context: owns a HIP stream on `device`, cached device buffers, captured hipGraphs
 * and the autotuned launch schedule.  Replaces the ArrayT backend selector
 * (src/kernel.jl:319,330; scripts/shoot_gammas.jl:47). NULL on failure. */
crysp_ctx *crysp_create(int device);
void crysp_destroy(crysp_ctx *ctx);
const char *crysp_last_error(void);
int crysp_abi_version(void);

/* Records mode: bit-for-bit replacement of transport_kernel!
 * (src/kernel.jl:159-305) over ndrange = n_photons.  Photon k (1-based) starts at
 * position j = 1 + (k-1) % npos with t0 = pt[j] (pt may be NULL: zeros), draws an
 * isotropic direction and the emission delay, and is transported until it ends.
 * Outputs per photon: status (crysp_status), idx (SiPM linear index (j-1)*nx + i,
 * negated on the z = 0 face, 0 unless detected), tdet (time at the end [ns]) and
 * bnc (bounces).  Any of the four output pointers may be NULL. */
int crysp_transport_records(crysp_ctx *ctx, const crysp_box *box,
                            const crysp_optical_params *op, const crysp_readout *ro,
                            const float *px, const float *py, const float *pz,
                            const float *pt, int32_t npos, int64_t n_photons,
                            uint64_t seed, uint64_t pid0, int32_t max_bounces,
                            const crysp_emission *em,
                            uint8_t *status, int16_t *idx, float *tdet, int32_t *bnc);

/* Accumulate mode: the same photons reduced on the device into the reference's
 * Accumulator (run_photons_ka! + its host loop, src/kernel.jl:312-367, and
 * run_photons!, src/transport.jl:245-273).  counts: u32 (nx, ny, nbins+1);
 * first_ns: f32 (nx, ny), +Inf where nothing was detected; counters: int64
 * [CRYSP_NCOUNTERS] (crysp_counter).  Outputs are overwritten, not added to. */
int crysp_transport_accumulate(crysp_ctx *ctx, const crysp_box *box,
                               const crysp_optical_params *op, const crysp_readout *ro,
                               const crysp_time_binning *tb,
                               const float *px, const float *py, const float *pz,
                               const float *pt, int32_t npos, int64_t n_photons,
                               uint64_t seed, uint64_t pid0, int32_t max_bounces,
                               const crysp_emission *em,
                               uint32_t *counts, float *first_ns, int64_t *counters);

/* Event mode (replaces the batch loop of run_events!, src/events.jl:59-91,109-121):
 * ndep deposits, deposit d holding dep_xyzt[4d..4d+3] = (x, y, z [mm], t0 [ns]),
 * dep_nph[d] photons and belonging to event dep_event[d] (0-based, < nev).  The photons
 * of all deposits are numbered g = 0.. in deposit order (deposit-major, as the
 * reference pushes them) and photon g uses stream pid = pid0 + g + 1; one persistent
 * launch drains them through a device work queue.  Outputs (overwritten): maps i32
 * (nx, ny, nev), tmin f32 (nx, ny, nev) (+Inf = no hit), npe i32 [nev]; both readout
 * faces count (src/events.jl:78).  n_photons_out (optional) = total photons. */
int crysp_run_events(crysp_ctx *ctx, const crysp_box *box,
                     const crysp_optical_params *op, const crysp_readout *ro,
                     const float *dep_xyzt, const int32_t *dep_nph,
                     const int32_t *dep_event, int64_t ndep, int64_t nev,
                     uint64_t seed, uint64_t pid0, int32_t max_bounces,
                     const crysp_emission *em,
                     int32_t *maps, float *tmin, int32_t *npe, int64_t *n_photons_out);

/* crysp_run_events without any allocation: no reference counterpart.  It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every array must already be device memory on the
 * context's device; no graph is captured and operation counting must be off.  Scratch
 * memory (queue cursor, error flags, deposit photon offsets) is a file-scope device array
 * of about 8 MiB, held by each device that runs the library until the process ends:
 * ndep <= 2^20 (CRYSP_ERANGE beyond).  All calls on one device share that array, so calls
 * that overlap on one device race.  Results equal crysp_run_events'. */
int crysp_run_events_noalloc(crysp_ctx *ctx, const crysp_box *box,
                             const crysp_optical_params *op, const crysp_readout *ro,
                             const float *dep_xyzt, const int32_t *dep_nph,
                             const int32_t *dep_event, int64_t ndep, int64_t nev,
                             uint64_t seed, uint64_t pid0, int32_t max_bounces,
                             const crysp_emission *em, int32_t *maps, float *tmin,
                             int32_t *npe, int64_t *n_photons_out);

/* Event mode with deposits in several crystals sharing one queue (no reference
 * counterpart): deposit d lies in crystal boxes[dep_box[d]] (nbox <= 64); everything
 * else as crysp_run_events. */
int crysp_run_events_multibox(crysp_ctx *ctx, const crysp_box *boxes, int32_t nbox,
                              const int32_t *dep_box, const crysp_optical_params *op,
                              const crysp_readout *ro, const float *dep_xyzt,
                              const int32_t *dep_nph, const int32_t *dep_event, int64_t ndep,
                              int64_t nev, uint64_t seed, uint64_t pid0, int32_t max_bounces,
                              const crysp_emission *em, int32_t *maps, float *tmin, int32_t *npe,
                              int64_t *n_photons_out);

/* Light-response maps (no reference counterpart; the surrogate's Monte Carlo targets):
 * npts point sources at xyz[3i..3i+2] with nph[i] photons each (t0 = 0), one event per
 * point through the event mode.  counts: i32 (nx, ny, npts); tsum_fx (optional): u64
 * (nx, ny, npts) sums of the detected photons' arrival times in units of 2^-10 ns, so
 * mean times are exact integer sums divided on the host in binary64. */
int crysp_light_maps(crysp_ctx *ctx, const crysp_box *box, const crysp_optical_params *op,
                     const crysp_readout *ro, const float *xyz, const int32_t *nph, int64_t npts,
                     uint64_t seed, uint64_t pid0, int32_t max_bounces, const crysp_emission *em,
                     int32_t *counts, uint64_t *tsum_fx);

/* Launch schedule of the transport kernels: workgroup size (64..1024, multiple of 64)
 * and photons per thread (1..8; sets the persistent grid to
 * ceil(n_photons / (wg * ppt)) workgroups, capped at what the device keeps resident).
 * No reference counterpart (KA picks the workgroup, src/kernel.jl:8). */
int crysp_set_schedule(crysp_ctx *ctx, int32_t wg_size, int32_t photons_per_thread);
int crysp_get_schedule(crysp_ctx *ctx, int32_t *wg_size, int32_t *photons_per_thread);

/* Work-queue knobs (no reference counterpart): a wave takes `chunk` consecutive photons
 * per atomic grab of the queue cursor, and hands new photons to its idle lanes once at
 * least refill_min of them are idle (or all are).  Defaults 256 and 4. */
int crysp_set_queue(crysp_ctx *ctx, int32_t chunk, int32_t refill_min);

/* Event-mode work split (no reference counterpart).  enable != 0: when the deposits are
 * ordered by event (non-decreasing dep_event, as the reference pushes them), workgroups
 * own whole events, sum their detections in LDS and write each map row once; they claim
 * an event's photons `piece` at a time (default 1024), take events largest first, and
 * workgroups without work help with events that still have photons.  Tables out of order, several crystals, time sums,
 * operation counting, grids beyond 256 SiPMs and the allocation-free entry point use
 * per-detection atomics instead.  enable = 0 (the default) always uses per-detection
 * atomics.  Outputs are identical either way. */
int crysp_set_event_sessions(crysp_ctx *ctx, int32_t enable, int32_t piece);

/* Wave timeline (diagnostics, no reference counterpart): with enable != 0 every transport
 * launch records, per wave, the device wall-clock tick at which the wave ended.
 * crysp_get_wave_timeline copies the last launch's ticks of up to `cap` waves into out
 * (cap u64; 0 for a wave that never ran), and sets *n_waves (waves of that launch) and
 * *ticks_per_s (the clock's rate).  With the launch's duration it shows how long the
 * launch ran with few waves left: the tail that the longest photons set.  All contexts on
 * a device share one record. */
int crysp_set_wave_timeline(crysp_ctx *ctx, int32_t enable);
int crysp_get_wave_timeline(crysp_ctx *ctx, uint64_t *out, int64_t cap, int64_t *n_waves,
                            double *ticks_per_s);

/* Autotuner (no reference counterpart): times crysp_run_events on the given deposits
 * for every (wg_sizes[a], ppts[b]) with `reps` repetitions each, keeps the fastest as
 * the context's schedule and writes the best-of-reps kernel time [ms] of every pair
 * into times_ms[a * n_ppt + b].  Results do not depend on the schedule. */
int crysp_autotune_events(crysp_ctx *ctx, const crysp_box *box,
                          const crysp_optical_params *op, const crysp_readout *ro,
                          const float *dep_xyzt, const int32_t *dep_nph,
                          const int32_t *dep_event, int64_t ndep, int64_t nev,
                          uint64_t seed, int32_t max_bounces, const crysp_emission *em,
                          const int32_t *wg_sizes, int32_t n_wg,
                          const int32_t *ppts, int32_t n_ppt, int32_t reps,
                          float *times_ms, int32_t *best_wg, int32_t *best_ppt);

/* Autotuners over every launch knob (no reference counterpart): each combination of
 * wg_sizes x ppts x chunks x refill_mins (crysp_set_schedule, crysp_set_queue) is timed
 * `reps` times on the given work (event mode, or accumulate mode for a point source), the
 * best-of-reps kernel time [ms] of combination (a, b, q, r) is written to
 * times_ms[((a * n_ppt + b) * n_chunk + q) * n_refill + r], and the fastest becomes the
 * context's schedule and queue setting (best4 = wg, ppt, chunk, refill_min).  Results do not
 * depend on the knobs. */
int crysp_autotune_events_queue(crysp_ctx *ctx, const crysp_box *box,
                                const crysp_optical_params *op, const crysp_readout *ro,
                                const float *dep_xyzt, const int32_t *dep_nph,
                                const int32_t *dep_event, int64_t ndep, int64_t nev,
                                uint64_t seed, int32_t max_bounces, const crysp_emission *em,
                                const int32_t *wg_sizes, int32_t n_wg, const int32_t *ppts,
                                int32_t n_ppt, const int32_t *chunks, int32_t n_chunk,
                                const int32_t *refill_mins, int32_t n_refill, int32_t reps,
                                float *times_ms, int32_t *best4);
int crysp_autotune_accumulate(crysp_ctx *ctx, const crysp_box *box,
                              const crysp_optical_params *op, const crysp_readout *ro,
                              const crysp_time_binning *tb, const float *px, const float *py,
                              const float *pz, const float *pt, int32_t npos, int64_t n_photons,
                              uint64_t seed, int32_t max_bounces, const crysp_emission *em,
                              const int32_t *wg_sizes, int32_t n_wg, const int32_t *ppts,
                              int32_t n_ppt, const int32_t *chunks, int32_t n_chunk,
                              const int32_t *refill_mins, int32_t n_refill, int32_t reps,
                              float *times_ms, int32_t *best4);

/* Debug: the first n_draws raw u32 words of the streams pid0+1 .. pid0+n_pids,
 * produced on the device by the same Philox code the kernels use
 * (src/philox.jl:43-57).  out: n_pids * n_draws words, pid-major. */
int crysp_debug_draws(crysp_ctx *ctx, uint64_t seed, uint64_t pid0, int64_t n_pids,
                      int32_t n_draws, uint32_t *out);

/* Debug: the device Float32 functions the transport uses, applied to n inputs, for
 * the exhaustive checks of their domains.  which: 0 log, 1 sin, 2 cos (both from the
 * sincos path), 3 sqrt, 4 cos (the stand-alone cos of the Box-Muller gaussian), 5 sin,
 * 6 cos (the micro-facet tilt path, for arguments >= 0), 7 the hardware log2 estimate
 * behind the bulk-absorption decision (not a correctly rounded function), 8 the largest
 * gap between the estimated and the exact Fresnel reflectance at cosine x over a fixed set
 * of interfaces. */
int crysp_debug_math(crysp_ctx *ctx, int32_t which, const float *x, int64_t n, float *y);

/* Kernel timing: when enabled, every transport launch is bracketed by hipEvents on
 * the context's stream; crysp_kernel_time returns the summed kernel time [ms] and
 * the number of launches since the last reset. */
int crysp_set_timing(crysp_ctx *ctx, int32_t enable);
int crysp_kernel_time(crysp_ctx *ctx, double *total_ms, int64_t *n_launches, int32_t reset);

/* Operation counting: when enabled, transport launches use a kernel variant that also
 * counts, per call, [n_photon, n_draw, n_philox_block, n_step, n_fresnel,
 * n_surface_try, n_lambert, n_rayleigh_try, n_detect, n_scatter]; crysp_get_op_counts
 * returns their sums since the last reset (the inputs of the roofline op model). */
int crysp_set_op_counting(crysp_ctx *ctx, int32_t enable);
int crysp_get_op_counts(crysp_ctx *ctx, int64_t *counts10, int32_t reset);

/* ---- host-side event feed (CPU code in the same library) ------------------------
 * The gamma stage that produces deposits for crysp_run_events: the reference's vendored
 * gamma core (the .jl files of src/gamma/) and its bridge (src/gamma_interface.jl), Float64 cm/MeV.
 * Random numbers come from per-event Philox streams (the reference's Xoshiro cannot be
 * reproduced), so events agree with the reference statistically, not bit for bit. */
typedef struct crysp_gamma_material crysp_gamma_material;

/* Gamma.make_material (src/gamma/gmaterials.jl:22-31) from an XCOM table file (rows of
 * 8 numbers: E [MeV], coherent, incoherent, photoelectric, pair nuclear, pair electron,
 * totals; mu/rho [cm^2/g]) or from such rows in memory.  NULL on failure. */
crysp_gamma_material *crysp_gamma_material_load(const char *xcom_path, double density);
crysp_gamma_material *crysp_gamma_material_create(const double *rows8, int32_t nrows,
                                                  double density);
void crysp_gamma_material_free(crysp_gamma_material *m);
const char *crysp_gamma_last_error(void);

/* Gamma.sigma_macro (src/gamma/gmaterials.jl:45-55): (Compton, photoelectric, pair)
 * macroscopic cross sections [1/cm] at e_mev. */
int crysp_gamma_sigma(const crysp_gamma_material *m, double e_mev, double *out3);

/* Gamma.propagate_photon (src/gamma/gtransport.jl:44-81) in the box-centred frame (cm):
 * up to cap records (x, y, z, e_in, e_dep, process) with process 0 photoelectric,
 * 1 Compton, 2 below cut, 3 pair, 4 escape; n_out = number of records. */
int crysp_gamma_history(const crysp_gamma_material *m, const crysp_box *box, double e0_mev,
                        const double *p0_cm, const double *dir0, uint64_t seed,
                        uint64_t stream, double *recs6, int32_t cap, int32_t *n_out);

/* The event loop of run_events! (src/events.jl:93-117): per event a gamma of e0_mev
 * enters the z = 0 face at a uniform point along +z; its deposits (crystal frame:
 * x, y, z [mm], t [ns]) get Poisson(yield_per_mev * E / 1000) photons.  Deposits are
 * written densely (event e owns [ev_first[e], ev_first[e+1])); dep_* arrays need
 * n_events * max_dep_per_event entries.  Truth (optional pointers): edep, xyz1, e1,
 * xyz2, e2, er [keV, mm], int_type (-1 none, 0 photoelectric, X Compton scatters),
 * n_int, as in src/events.jl:25-37. */
int crysp_gamma_events(const crysp_gamma_material *m, const crysp_box *box, int64_t n_events,
                       uint64_t seed, double e0_mev, double yield_per_mev,
                       int32_t max_dep_per_event, float *dep_xyzt, int32_t *dep_nph,
                       int32_t *dep_event, float *dep_e_kev, int8_t *dep_proc,
                       int64_t *ev_first, float *edep, float *xyz1, float *e1, float *xyz2,
                       float *e2, float *er, int8_t *int_type, int16_t *n_int,
                       int64_t *n_dep_out);

/* Synthetic single-deposit events (no reference counterpart; the batched and surrogate
 * workloads): position uniform in the axis-aligned box [lo3, hi3] mm, energy uniform
 * in [e_lo, e_hi] keV, t0 = 0, photons Poisson(yield * E / 1000), or n_fixed each when
 * yield_per_mev <= 0. */
int crysp_synthetic_deposits(int64_t n_events, uint64_t seed, const float *lo3,
                             const float *hi3, float e_lo_kev, float e_hi_kev,
                             double yield_per_mev, int32_t n_fixed, float *dep_xyzt,
                             float *dep_e_kev, int32_t *dep_nph);

/* rand_poisson (src/events.jl:11-22: Knuth for lambda <= 50, rounded Gaussian above)
 * on n independent Philox streams, for the sampler's tests. */
int crysp_poisson_samples(uint64_t seed, double lambda, int64_t n, int64_t *out);

/* ---- learned light-response surrogate (no reference counterpart) ---------------
 * An MLP n_in -> hidden x n_hidden_layers -> n_out with GELU, trained and evaluated by
 * BF16 MFMA GEMMs (v_mfma_f32_32x32x16_bf16) accumulating in FP32, FP32 master weights
 * and Adam.  Parameters are a flat FP32 vector: per layer W [out][in] then b [out].
 * Limits: n_in <= 32; hidden and n_out multiples of 64 up to 256; 1 to 8 hidden layers
 * (a training step keeps 32 rows of every layer in LDS). */
typedef struct crysp_mlp crysp_mlp;

crysp_mlp *crysp_mlp_create(int device, int32_t n_in, int32_t hidden, int32_t n_hidden_layers,
                            int32_t n_out, uint64_t seed);
void crysp_mlp_destroy(crysp_mlp *mlp);
const char *crysp_mlp_last_error(void);
int64_t crysp_mlp_num_params(const crysp_mlp *mlp);
double crysp_mlp_flops_per_row(const crysp_mlp *mlp);
int crysp_mlp_get_params(crysp_mlp *mlp, float *params);
int crysp_mlp_set_params(crysp_mlp *mlp, const float *params);
/* x' = (x - shift) * scale, applied to the n_in inputs inside the network */
int crysp_mlp_set_input_norm(crysp_mlp *mlp, const float *shift, const float *scale);
/* The first n_pois outputs are log-rates trained with a Poisson negative log-likelihood,
 * w (e^y - t y) for an observed rate t; the rest with a weighted squared error. */
int crysp_mlp_set_poisson_outputs(crysp_mlp *mlp, int32_t n_pois);
/* Y [M x n_out] = MLP(X [M x n_in]); host or device pointers */
int crysp_mlp_forward(crysp_mlp *mlp, const float *X, int64_t M, float *Y);
/* Gradient of the loss (normalized by sum W) over one batch (W may be NULL: all ones),
 * without an update: grads [num_params] and loss (host pointers, optional).  R0 (may be
 * NULL) [M x n_pois]: a known part of each Poisson rate, so that the rate is e^y + R0 and
 * the loss w (rate - t log rate). */
int crysp_mlp_gradient(crysp_mlp *mlp, const float *X, const float *T, const float *W,
                       const float *R0, int32_t M, float *grads, float *loss);
/* n_batches Adam steps on consecutive device-resident rows (batch a multiple of 64) of
 * X, T, W and R0 (rate offsets as in crysp_mlp_gradient, rows of n_pois; may be NULL);
 * mean_row_weight = mean over rows of sum(W) / n_out, the loss normalization.
 * losses[b] (host, optional) receives each step's loss.  A step is four kernels; up to 64
 * steps are captured in one hipGraph, which is kept for later calls with the same X, T, W
 * pointers, batch, optimizer settings and loss layout (the learning rate may change: it is
 * read from device memory).  Refill the same buffers between epochs to reuse it. */
int crysp_mlp_train_batches(crysp_mlp *mlp, const float *X, const float *T, const float *W,
                            const float *R0, int64_t n_batches, int32_t batch, float lr,
                            float beta1, float beta2, float eps, float weight_decay,
                            float mean_row_weight, float *losses);

/* Events from a surrogate baked on a regular grid: grid [gz][gy][gx][2][ncell] holds per
 * grid point the per-photon detection probability of every cell and the mean arrival
 * time (ns after the deposit); for event e with deposits [ev_first[e], ev_first[e+1])
 * and use[d] != 0, lambda_c = sum_d yield * E_d / 1000 * p_c(x_d) (trilinear) and
 * maps[e][c] ~ Poisson(lambda_c) (src/events.jl:11-22 algorithm, Philox domain 4);
 * tmean/lam (optional) get the count-weighted mean time and lambda. */
int crysp_surrogate_sample_events(int device, const float *grid, int32_t gx, int32_t gy,
                                  int32_t gz, int32_t ncell, const crysp_box *box,
                                  const float *dep_xyzt, const float *dep_e_kev,
                                  const uint8_t *use, const int64_t *ev_first, int64_t ndep,
                                  int64_t nev, double yield_per_mev, uint64_t seed,
                                  int32_t *maps, float *tmean, float *lam);

/* A light-response surrogate of one crystal/optics set: a trained network (4 inputs,
 * 2 * nx * ny outputs: per-photon log-rates of the cells, then standardized mean arrival
 * times) plus, when cfg.direct, the light that reaches the readout face straight from the
 * deposit, computed, not learned: D_c(x) = PDE / (4 pi) * integral over the directions
 * whose straight path lands in cell c of T(cos theta) exp(-path / Lambda) (Fresnel
 * transmission into the coupling, bulk attenuation over the path), weighted by a smoothstep
 * of the depth Lz - z (1 within 2 mm of the readout face, 0 beyond 3 mm: the term is sharp
 * only there).  The per-photon rate of cell c is e^{y_c} + D_c.  Inputs of the network: (x, y, depth feature, E) with x, y
 * folded into 0 <= x <= y <= L/2 when cfg.fold (square face and grid; outputs in folded
 * cell order), depth feature -log(Lz - z + log_depth_mm) when log_depth_mm > 0, else z.
 * Mean times: y * t_scale + t_shift [ns after the deposit].  The surrogate keeps a
 * reference to `mlp`, which must outlive it.  No reference counterpart. */
typedef struct {
    uint8_t fold;
    uint8_t direct;
    uint8_t images;          /* with direct and an air-gap wrap: add the direct light's
                                mirror images in the side walls (total internal reflection)
                                within 3 mm of a side wall and of the readout face */
    uint8_t pad_;
    float log_depth_mm;
    float t_shift, t_scale;
} crysp_surrogate_config;
typedef struct crysp_surrogate crysp_surrogate;

crysp_surrogate *crysp_surrogate_create(crysp_mlp *mlp, const crysp_box *box,
                                        const crysp_optical_params *op, const crysp_readout *ro,
                                        const crysp_surrogate_config *cfg);
void crysp_surrogate_destroy(crysp_surrogate *s);
/* D [n x nx*ny] (standard cell order i + nx j) at points xyz [n x 3] as given (no folding). */
int crysp_surrogate_direct_light(crysp_surrogate *s, const float *xyz, int64_t n, float *D);
/* Per-photon maps p [n x nx*ny] and mean arrival times t (optional) in actual cell order for
 * deposits dep_xyzt [n x 4] with energies e_kev [n]. */
int crysp_surrogate_predict(crysp_surrogate *s, const float *dep_xyzt, const float *e_kev,
                            int64_t n, float *p, float *t);
/* Events from the network, all on the device: for event e with deposits
 * [ev_first[e], ev_first[e+1]) and use[d] != 0, lambda_c = sum_d yield * E_d / 1000 * p_c(x_d)
 * and maps[e][c] ~ Poisson(lambda_c) (src/events.jl:11-22 algorithm, Philox domain 4);
 * tmean/lam (optional) as crysp_surrogate_sample_events. */
int crysp_surrogate_sample(crysp_surrogate *s, const float *dep_xyzt, const float *dep_e_kev,
                           const uint8_t *use, const int64_t *ev_first, int64_t ndep, int64_t nev,
                           double yield_per_mev, uint64_t seed, int32_t *maps, float *tmean,
                           float *lam);

/* Debug: C [M x N] = A * B in bf16 on the GEMM kernel (A [M][K], or [K][M] when a_t;
 * B [K][N], or [N][K] when b_t), for the MFMA lane-layout check. */
int crysp_debug_gemm(int device, int32_t M, int32_t N, int32_t K, const uint16_t *A_bf16,
                     int32_t a_t, const uint16_t *B_bf16, int32_t b_t, float *C);

#ifdef __cplusplus
}
#endif

#endif /* CRYSPLIGHT_H */
