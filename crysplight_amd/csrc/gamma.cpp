// gamma.cpp -- host-side feed of the event mode: 511 keV gamma transport in the crystal
// (Compton / photoelectric from NIST XCOM tables), per-deposit Poisson photon yields,
// and the synthetic deposit feeds.  Float64, cm / MeV inside, like the reference's
// vendored gamma core; deposits leave in the crystal frame (mm, keV, ns, Float32).
//
// Restates: src/gamma/nist_data.jl:25-71 (XCOM loader, log-log interpolation),
// src/gamma/gmaterials.jl:22-61 (material, sigma_macro, mfp), src/gamma/ggeometry.jl
// (slab exit distance), src/gamma/sampling.jl:6-140 (free path, process choice,
// Klein-Nishina by Butcher-Messel, rotation), src/gamma/gtransport.jl:44-81 (history),
// src/gamma_interface.jl:30-49 (frame change, path-length times, process codes) and
// src/events.jl:10-22,93-117 (Poisson yield, truth record).
//
// Random numbers: the reference draws from Julia's Xoshiro(hash((seed, ev))), which
// cannot be reproduced outside Julia.  Here every event owns a Philox4x32-10 stream
// with key (seed lo, seed hi) and counter (event lo, event hi, block, DOMAIN), DOMAIN
// = 1 for gamma events and 2 for the synthetic feeds (the optical photons use 0), and
// uniforms in [0, 1) take 53 bits from two words.  Event-level results are therefore
// statistically, not bitwise, equivalent to the reference's.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "crysplight.h"

namespace {

void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
            uint32_t out[4])
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// per-event stream of doubles in [0, 1)
struct EventRng {
    uint32_t k0, k1, e0, e1, dom, blk = 0, buf[4] = {0, 0, 0, 0};
    int left = 0;
    EventRng(uint64_t seed, uint64_t ev, uint32_t domain)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), e0((uint32_t)ev),
          e1((uint32_t)(ev >> 32)), dom(domain) {}
    uint32_t u32()
    {
        if (left == 0) { philox(e0, e1, blk++, dom, k0, k1, buf); left = 4; }
        return buf[4 - left--];
    }
    double rand()
    {
        const uint64_t a = u32() >> 5, b = u32() >> 6;
        return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
    }
    double randn()       // Box-Muller (one value per call)
    {
        double u1 = rand();
        const double u2 = rand();
        if (u1 <= 0.0) u1 = 1.0 / 9007199254740992.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

}  // namespace

struct crysp_gamma_material {
    double density = 0.0;
    std::vector<double> E, logE, inc, ph, pair, log_inc, log_ph, log_pair;
};

namespace {

void prelog(const std::vector<double> &f, std::vector<double> &lf)
{
    lf.resize(f.size());
    for (size_t i = 0; i < f.size(); ++i) lf[i] = f[i] > 0.0 ? std::log(f[i]) : -INFINITY;
}

// src/gamma/nist_data.jl:59-71: log-log with a known bracket lo (0-based)
double interp_ll(double lx, const std::vector<double> &lxp, const std::vector<double> &lfp,
                 const std::vector<double> &fp, size_t lo)
{
    const size_t hi = lo + 1;
    if (fp[lo] <= 0.0) return 0.0;
    if (fp[hi] <= 0.0) return fp[lo];
    const double t = (lx - lxp[lo]) / (lxp[hi] - lxp[lo]);
    return std::exp(lfp[lo] + t * (lfp[hi] - lfp[lo]));
}

// src/gamma/gmaterials.jl:45-55
void sigma_macro(const crysp_gamma_material &m, double e, double &sc, double &sph, double &sp)
{
    sc = sph = sp = 0.0;
    const size_t n = m.E.size();
    if (n < 2) return;
    const double lx = std::log(e);
    // searchsortedlast (1-based, 0 if none) clamped to 1..n-1, here 0-based 0..n-2
    size_t last = (size_t)(std::upper_bound(m.E.begin(), m.E.end(), e) - m.E.begin());
    size_t lo = last == 0 ? 0 : last - 1;
    lo = std::min(lo, n - 2);
    sc = interp_ll(lx, m.logE, m.log_inc, m.inc, lo) * m.density;
    sph = interp_ll(lx, m.logE, m.log_ph, m.ph, lo) * m.density;
    sp = interp_ll(lx, m.logE, m.log_pair, m.pair, lo) * m.density;
}

// src/gamma/ggeometry.jl:19-38: distance to exit an origin-centred box
double distance_to_exit(const double p[3], const double d[3], const double h[3])
{
    double tn = -INFINITY, tf = INFINITY;
    for (int i = 0; i < 3; ++i) {
        if (std::fabs(d[i]) > 1e-20) {
            const double t1 = (-h[i] - p[i]) / d[i], t2 = (h[i] - p[i]) / d[i];
            tn = std::max(tn, std::min(t1, t2));
            tf = std::min(tf, std::max(t1, t2));
        } else if (p[i] < -h[i] || p[i] > h[i]) {
            return INFINITY;
        }
    }
    if (tf < tn) return INFINITY;
    return tf > 1e-10 ? tf : INFINITY;
}

constexpr double ME = 0.51099895;          // electron rest mass [MeV]
constexpr double C_CM_NS = 29.9792458;

// src/gamma/sampling.jl:28-40: Klein-Nishina, Butcher-Messel
void sample_compton(double E, EventRng &rng, double &eprime, double &cost)
{
    const double eps0 = ME / (ME + 2.0 * E);
    const double a1 = -std::log(eps0);
    const double a2 = (1.0 - eps0 * eps0) / 2.0;
    for (;;) {
        const double r1 = rng.rand(), r2 = rng.rand(), r3 = rng.rand();
        const double eps = r1 < a1 / (a1 + a2) ? std::pow(eps0, r2)
                                               : std::sqrt(eps0 * eps0 + (1.0 - eps0 * eps0) * r2);
        const double omc = (1.0 - eps) * ME / (E * eps);
        const double sin2 = std::min(1.0, std::max(0.0, omc * (2.0 - omc)));
        const double g = 1.0 - eps * sin2 / (1.0 + eps * eps);
        if (g >= r3) { eprime = eps * E; cost = 1.0 - omc; return; }
    }
}

// src/gamma/sampling.jl:74-87: local (z = ref) to global
void rotate_to_global(double lx, double ly, double lz, const double ref[3], double out[3])
{
    const double n = std::sqrt(ref[0] * ref[0] + ref[1] * ref[1] + ref[2] * ref[2]);
    const double rx = ref[0] / n, ry = ref[1] / n, rz = ref[2] / n;
    if (std::fabs(rz) > 0.99999) {
        const double s = rz > 0 ? 1.0 : (rz < 0 ? -1.0 : 0.0);
        out[0] = lx * s; out[1] = ly * s; out[2] = lz * s;
        return;
    }
    const double rp = std::sqrt(rx * rx + ry * ry);
    const double e1x = -rz * rx / rp, e1y = -rz * ry / rp, e1z = -(rz * rz - 1.0) / rp;
    const double e2x = -ry / rp, e2y = rx / rp;
    out[0] = lx * e1x + ly * e2x + lz * rx;
    out[1] = lx * e1y + ly * e2y + lz * ry;
    out[2] = lx * e1z + lz * rz;
}

enum Proc { P_PHOTO = 0, P_COMPTON = 1, P_BELOWCUT = 2, P_PAIR = 3, P_ESCAPE = 4 };

struct Interaction { double x, y, z, e_in, e_dep; int proc; };

// src/gamma/gtransport.jl:44-81
void propagate_gamma(const crysp_gamma_material &m, const double half[3], double e0,
                     const double p0[3], const double d0[3], EventRng &rng,
                     std::vector<Interaction> &recs, double ecut = 0.010)
{
    double E = e0, pos[3] = { p0[0], p0[1], p0[2] };
    const double dn = std::sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    double dir[3] = { d0[0] / dn, d0[1] / dn, d0[2] / dn };
    recs.clear();
    for (;;) {
        double sc, sph, sp;
        sigma_macro(m, E, sc, sph, sp);
        const double S = sc + sph + sp;
        const double s = S > 0.0 ? -std::log(rng.rand()) / S : INFINITY;
        const double dex = distance_to_exit(pos, dir, half);
        if (s >= dex) {
            for (int i = 0; i < 3; ++i) pos[i] += dex * dir[i];
            recs.push_back({ pos[0], pos[1], pos[2], E, 0.0, P_ESCAPE });
            return;
        }
        for (int i = 0; i < 3; ++i) pos[i] += s * dir[i];
        // sample_process (src/gamma/sampling.jl:15-20): photoelectric is the catch-all
        const double r = rng.rand();
        const double pc = sc / S, pp = sp / S;
        const int proc = r < pc ? P_COMPTON : (r < pc + pp ? P_PAIR : P_PHOTO);
        if (proc != P_COMPTON) {
            recs.push_back({ pos[0], pos[1], pos[2], E, E, proc });
            return;
        }
        double ep, cost;
        sample_compton(E, rng, ep, cost);
        const double phi = 2.0 * M_PI * rng.rand();
        const double sint = std::sqrt(std::max(0.0, 1.0 - cost * cost));
        double nd[3];
        rotate_to_global(sint * std::cos(phi), sint * std::sin(phi), cost, dir, nd);
        recs.push_back({ pos[0], pos[1], pos[2], E, E - ep, P_COMPTON });
        dir[0] = nd[0]; dir[1] = nd[1]; dir[2] = nd[2];
        E = ep;
        if (E < ecut) {
            recs.push_back({ pos[0], pos[1], pos[2], E, E, P_BELOWCUT });
            return;
        }
    }
}

// src/events.jl:11-22
int64_t rand_poisson(EventRng &rng, double lam)
{
    if (lam <= 0.0) return 0;
    if (lam > 50.0) {
        const double v = std::nearbyint(lam + std::sqrt(lam) * rng.randn());
        return v > 0.0 ? (int64_t)v : 0;
    }
    const double L = std::exp(-lam);
    int64_t k = 0;
    double p = 1.0;
    for (;;) {
        p *= rng.rand();
        if (p < L) return k;
        ++k;
    }
}

template <class F>
void parallel_for(int64_t n, F f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 2048) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=, &f]() {
            for (int64_t i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

thread_local std::string g_gerr;

}  // namespace

extern "C" {

const char *crysp_gamma_last_error(void) { return g_gerr.c_str(); }

crysp_gamma_material *crysp_gamma_material_create(const double *table, int32_t nrows,
                                                  double density)
{
    if (!table || nrows < 2 || !(density >= 0.0)) {
        g_gerr = "need >= 2 rows and density >= 0";
        return nullptr;
    }
    auto *m = new crysp_gamma_material();
    m->density = density;
    for (int32_t i = 0; i < nrows; ++i) {
        const double *r = table + 8 * (size_t)i;
        double e = r[0];
        // src/gamma/nist_data.jl:42-48: nudge duplicate (K-edge) energies apart
        if (!m->E.empty() && e <= m->E.back()) e = m->E.back() * (1.0 + 1e-9);
        m->E.push_back(e);
        m->inc.push_back(r[2]);
        m->ph.push_back(r[3]);
        m->pair.push_back(r[4] + r[5]);
    }
    m->logE.resize(m->E.size());
    for (size_t i = 0; i < m->E.size(); ++i) m->logE[i] = std::log(m->E[i]);
    prelog(m->inc, m->log_inc);
    prelog(m->ph, m->log_ph);
    prelog(m->pair, m->log_pair);
    return m;
}

// Reads an XCOM table: rows starting with a digit and holding >= 8 numbers
// (src/gamma/nist_data.jl:25-39).
crysp_gamma_material *crysp_gamma_material_load(const char *xcom_path, double density)
{
    std::ifstream in(xcom_path ? xcom_path : "");
    if (!in) {
        g_gerr = std::string("cannot open XCOM table ") + (xcom_path ? xcom_path : "(null)");
        return nullptr;
    }
    std::vector<double> rows;
    std::string line;
    while (std::getline(in, line)) {
        size_t a = line.find_first_not_of(" \t\r");
        if (a == std::string::npos || line[a] < '0' || line[a] > '9') continue;
        std::istringstream ss(line.substr(a));
        double v[8];
        int k = 0;
        while (k < 8 && (ss >> v[k])) ++k;
        if (k < 8) continue;
        rows.insert(rows.end(), v, v + 8);
    }
    if (rows.size() < 16) {
        g_gerr = "no XCOM data rows found";
        return nullptr;
    }
    return crysp_gamma_material_create(rows.data(), (int32_t)(rows.size() / 8), density);
}

void crysp_gamma_material_free(crysp_gamma_material *m) { delete m; }

int crysp_gamma_sigma(const crysp_gamma_material *m, double e_mev, double *out3)
{
    if (!m || !out3) return CRYSP_EINVAL;
    sigma_macro(*m, e_mev, out3[0], out3[1], out3[2]);
    return CRYSP_OK;
}

// One gamma history from p0 (cm, box-centred) along d0: its interaction records as
// rows (x, y, z [cm], e_in, e_dep [MeV], process), for the vendored-core tests.
int crysp_gamma_history(const crysp_gamma_material *m, const crysp_box *box, double e0_mev,
                        const double *p0, const double *d0, uint64_t seed, uint64_t stream,
                        double *recs6, int32_t cap, int32_t *n_out)
{
    if (!m || !box || !p0 || !d0 || !recs6 || !n_out) return CRYSP_EINVAL;
    const double half[3] = { (double)(box->L[0] / 20.0f), (double)(box->L[1] / 20.0f),
                             (double)(box->L[2] / 20.0f) };
    EventRng rng(seed, stream, 1u);
    std::vector<Interaction> recs;
    propagate_gamma(*m, half, e0_mev, p0, d0, rng, recs);
    *n_out = (int32_t)recs.size();
    for (int32_t i = 0; i < std::min<int32_t>(cap, (int32_t)recs.size()); ++i) {
        const Interaction &r = recs[i];
        double *o = recs6 + 6 * (size_t)i;
        o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.e_in; o[4] = r.e_dep; o[5] = r.proc;
    }
    return (int32_t)recs.size() > cap ? CRYSP_ERANGE : CRYSP_OK;
}

// Gamma events of src/events.jl:93-117: n_events gammas of e0 entering the z = 0 face at
// a uniform point, along +z.  Per event the deposits (crystal frame) and their Poisson
// photon counts (yield_per_mev * E / 1000), and the truth record.  Deposits of event e
// occupy [ev_first[e], ev_first[e+1]); at most max_dep_per_event deposits are expanded into
// photons (later ones still count in edep and n_int).  Outputs sized: dep_* [n_events * max_dep_per_event],
// ev_first [n_events + 1], truth [n_events] (xyz: 3 per event).
int crysp_gamma_events(const crysp_gamma_material *m, const crysp_box *box, int64_t n_events,
                       uint64_t seed, double e0_mev, double yield_per_mev,
                       int32_t max_dep_per_event, float *dep_xyzt, int32_t *dep_nph,
                       int32_t *dep_event, float *dep_e_kev, int8_t *dep_proc,
                       int64_t *ev_first, float *edep, float *xyz1, float *e1, float *xyz2,
                       float *e2, float *er, int8_t *int_type, int16_t *n_int,
                       int64_t *n_dep_out)
{
    if (!m || !box || n_events < 0 || max_dep_per_event < 1 || !dep_xyzt || !dep_nph ||
        !dep_event || !ev_first || !n_dep_out) {
        g_gerr = "bad arguments";
        return CRYSP_EINVAL;
    }
    const float *L = box->L;
    const double half[3] = { (double)(L[0] / 20.0f), (double)(L[1] / 20.0f), (double)(L[2] / 20.0f) };
    const float e0_kev = (float)(e0_mev * 1000.0);
    const int32_t K = max_dep_per_event;
    std::vector<int32_t> nd(n_events, 0);
    parallel_for(n_events, [&](int64_t ev) {
        EventRng rng(seed, (uint64_t)ev + 1u, 1u);
        std::vector<Interaction> recs;
        recs.reserve(16);
        // entry point (src/events.jl:95): uniform on the z = 0 face
        const float ex = (float)(rng.rand() * (double)L[0]);
        const float ey = (float)(rng.rand() * (double)L[1]);
        const float ez = 0.0f;
        // src/gamma_interface.jl:32-33: to the box-centred frame, cm (Float32 arithmetic)
        const double p0[3] = { (double)((ex - L[0] / 2.0f) / 10.0f), (double)((ey - L[1] / 2.0f) / 10.0f),
                               (double)((ez - L[2] / 2.0f) / 10.0f) };
        const double d0[3] = { 0.0, 0.0, 1.0 };
        propagate_gamma(*m, half, e0_mev, p0, d0, rng, recs);
        // src/gamma_interface.jl:36-47: deposits with path-length times
        double xp = p0[0], yp = p0[1], zp = p0[2], path = 0.0;
        int32_t k = 0, ncompt = 0;
        float edep_ev = 0.0f, e1v = 0.0f, e2v = 0.0f;
        const size_t base = (size_t)ev * K;
        for (const Interaction &r : recs) {
            path += std::sqrt((r.x - xp) * (r.x - xp) + (r.y - yp) * (r.y - yp) + (r.z - zp) * (r.z - zp));
            xp = r.x; yp = r.y; zp = r.z;
            if (!(r.e_dep > 0.0)) continue;
            const int proc = r.proc == P_COMPTON ? 1 : (r.proc == P_BELOWCUT ? 2 : 0);
            const float ekev = (float)(r.e_dep * 1000.0);
            if (proc == 1) ++ncompt;
            const float x = (float)(r.x * 10.0 + (double)(L[0] / 2.0f));
            const float y = (float)(r.y * 10.0 + (double)(L[1] / 2.0f));
            const float z = (float)(r.z * 10.0 + (double)(L[2] / 2.0f));
            if (k == 0) {
                e1v = ekev;
                if (xyz1) { xyz1[3 * ev] = x; xyz1[3 * ev + 1] = y; xyz1[3 * ev + 2] = z; }
            } else if (k == 1) {
                e2v = ekev;
                if (xyz2) { xyz2[3 * ev] = x; xyz2[3 * ev + 1] = y; xyz2[3 * ev + 2] = z; }
            }
            edep_ev += ekev;
            if (k < K) {
                float *d = dep_xyzt + 4 * (base + k);
                d[0] = x; d[1] = y; d[2] = z; d[3] = (float)(path / C_CM_NS);
                if (dep_e_kev) dep_e_kev[base + k] = ekev;
                if (dep_proc) dep_proc[base + k] = (int8_t)proc;
                dep_event[base + k] = (int32_t)ev;
                // src/events.jl:111: photons from this deposit
                const int64_t nph = rand_poisson(rng, yield_per_mev * (double)ekev / 1000.0);
                dep_nph[base + k] = (int32_t)std::min<int64_t>(nph, INT32_MAX);
            }
            ++k;
        }
        if (k == 0) {
            if (xyz1) { xyz1[3 * ev] = xyz1[3 * ev + 1] = xyz1[3 * ev + 2] = 0.0f; }
            if (xyz2) { xyz2[3 * ev] = xyz2[3 * ev + 1] = xyz2[3 * ev + 2] = 0.0f; }
        } else if (k == 1 && xyz2) {
            xyz2[3 * ev] = xyz2[3 * ev + 1] = xyz2[3 * ev + 2] = 0.0f;
        }
        if (edep) edep[ev] = edep_ev;
        if (e1) e1[ev] = e1v;
        if (e2) e2[ev] = e2v;
        if (er) er[ev] = k == 0 ? e0_kev : e0_kev - e1v - e2v;
        if (int_type) int_type[ev] = (int8_t)(k == 0 ? -1 : std::min<int32_t>(ncompt, 127));
        if (n_int) n_int[ev] = (int16_t)k;
        nd[ev] = std::min(k, K);
    });
    // compact the per-event slots into a dense deposit table
    int64_t w = 0;
    for (int64_t ev = 0; ev < n_events; ++ev) {
        ev_first[ev] = w;
        const size_t base = (size_t)ev * K;
        for (int32_t k = 0; k < nd[ev]; ++k, ++w) {
            if ((size_t)w == base + k) continue;
            std::memcpy(dep_xyzt + 4 * w, dep_xyzt + 4 * (base + k), 16);
            dep_nph[w] = dep_nph[base + k];
            dep_event[w] = dep_event[base + k];
            if (dep_e_kev) dep_e_kev[w] = dep_e_kev[base + k];
            if (dep_proc) dep_proc[w] = dep_proc[base + k];
        }
    }
    ev_first[n_events] = w;
    *n_dep_out = w;
    return CRYSP_OK;
}

// Synthetic single-deposit events: position uniform in [lo, hi] per axis (mm), energy
// uniform in [e_lo, e_hi] keV, t0 = 0, photons ~ Poisson(yield * E / 1000)
// (yield <= 0: n_fixed photons each).  Philox domain 2.
int crysp_synthetic_deposits(int64_t n_events, uint64_t seed, const float *lo3,
                             const float *hi3, float e_lo_kev, float e_hi_kev,
                             double yield_per_mev, int32_t n_fixed, float *dep_xyzt,
                             float *dep_e_kev, int32_t *dep_nph)
{
    if (n_events < 0 || !lo3 || !hi3 || !dep_xyzt || !dep_nph) {
        g_gerr = "bad arguments";
        return CRYSP_EINVAL;
    }
    parallel_for(n_events, [&](int64_t ev) {
        EventRng rng(seed, (uint64_t)ev + 1u, 2u);
        float *d = dep_xyzt + 4 * ev;
        for (int a = 0; a < 3; ++a)
            d[a] = (float)((double)lo3[a] + rng.rand() * (double)(hi3[a] - lo3[a]));
        d[3] = 0.0f;
        const float e = (float)((double)e_lo_kev + rng.rand() * (double)(e_hi_kev - e_lo_kev));
        if (dep_e_kev) dep_e_kev[ev] = e;
        const int64_t n = yield_per_mev > 0.0 ? rand_poisson(rng, yield_per_mev * (double)e / 1000.0)
                                              : (int64_t)n_fixed;
        dep_nph[ev] = (int32_t)std::min<int64_t>(n, INT32_MAX);
    });
    return CRYSP_OK;
}

// Poisson draws from per-index Philox streams (domain 3), for tests of the sampler.
int crysp_poisson_samples(uint64_t seed, double lam, int64_t n, int64_t *out)
{
    if (n < 0 || !out) return CRYSP_EINVAL;
    parallel_for(n, [&](int64_t i) {
        EventRng rng(seed, (uint64_t)i + 1u, 3u);
        out[i] = rand_poisson(rng, lam);
    });
    return CRYSP_OK;
}

}  // extern "C"
