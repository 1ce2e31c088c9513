// Photoionization source: the reference's X-ray pulse entering the gas
// (INCIDENCE1.m:27-158 with GUASS.m:3-15, findEShell.m:1-12,
// findSigmaPE.m:8-22, findPEangle.m:10-33, findV.m:5-8), on the r-z gas
// grid of the collision step.
//
//   k_photo_attenuate  one thread per ring of the beam: Beer-Lambert down the
//                      z slabs, dN = n N sigma dz, N -= dN, and the number of
//                      macro events round(dN / FN) per (ring, slab) cell
//                      (INCIDENCE1.m:51-60); FP64 with contraction off, so the
//                      integer counts are reproducible on the CPU
//   k_photo_scan       exclusive prefix of the cell counts (slab-major order,
//                      as the reference concatenates its slabs)
//   k_photo_emit       one thread per event: position in its cell, the
//                      photoelectron (Sauter angle by rejection), the Auger
//                      electron of a K-shell event, and the ion with its
//                      thermal velocity plus the recoil (INCIDENCE1.m:96-156)
//   k_photo_deplete    one thread per cell with events: the gas loses the
//                      ionized molecules and their thermal energy
//                      (INCIDENCE1.m:87-92)
//   k_photo_counts     the species counts, capacity clamping and the result
//
// Random numbers: Philox4x32-10, key = seed, counter = (event, pulse, run,
// kSourceBlock + b): b = 0 position and photoelectron azimuth, 1 thermal
// normals, 2 Auger angles, 3.. Sauter proposals (two per block).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "context.h"
#include "device_common.h"

using namespace gasx;
using namespace gasx_host;

namespace {

constexpr uint32_t kSourceBlock = 0x40000000u;   // apart from the collision streams' blocks
constexpr int kSauterBlocks = 64;                // 128 proposals; the last is taken if all fail
constexpr int kMaxSrcRings = 4096;
constexpr long long kMaxSrcEvents = 1ll << 26;   // events per call (more: GASX_E_ARG, nothing emitted)

struct SrcArgs {
    // grid and gas
    int32_t nrings, gnr, gnz;
    double gdr, gdz, phip;
    const float* dens;                // gnz x gnr, iz * gnr + ir
    const float* temp;                // or null: t_scalar
    double t_scalar;
    // physics (host-computed FP64)
    double sigma_pe, fn;
    double v_pe;                      // findV((Ex - E_shell) * 1.6e-16)
    double v_ae;                      // findV(E_aug * 1.6e-16), K shell only
    int32_t auger;
    double sA, sbg, sg0;              // Sauter: A = 1/beta - 1, beta gamma (gamma-1)(gamma-2)/2, g0
    double p_photon;                  // Ex * 1.6e-16 / c
    double m_e, m_mol, k_b;           // 9.11e-31, 2 m_a 1.6726e-27, kB (INCIDENCE1.m:21-24)
    uint32_t key0, key1, pulse, run;
    // work arrays
    const double* ring_photons;       // nrings: photons entering slab 0 per ring
    double* photons_left;             // nrings
    int32_t* counts;                  // ncells (slab-major: k * nrings + i)
    int32_t* offs;                    // ncells + 1
    // outputs
    float *ex, *ey, *ez, *evx, *evy, *evz;   // electrons
    int32_t* e_count;
    int32_t e_cap;
    float *ix, *iy, *iz, *ivx, *ivy, *ivz;   // ions
    int32_t* i_count;
    int32_t i_cap;
    float* dens_out;                  // depletion (or null)
    float* temp_out;
    unsigned long long* result;       // [events, electrons, ions, dropped e, dropped i]
};

__device__ __forceinline__ double u01d(uint32_t x) { return (double)u01(x); }

__global__ void k_photo_attenuate(SrcArgs a)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nrings) return;
    double N = a.ring_photons[i];
    for (int k = 0; k < a.gnz; ++k) {
        const double n = (double)a.dens[k * a.gnr + i];
        double dN = n * N;
        dN = dN * a.sigma_pe;
        dN = dN * a.gdz;                                  // INCIDENCE1.m:57 (A = 1)
        N = N - dN;                                       // :58
        const double m = round(dN / a.fn);               // :59-60, half away from zero
        // a count beyond the per-call limit only has to push the total over it
        a.counts[k * a.nrings + i] = m <= (double)kMaxSrcEvents ? (int32_t)m : (int32_t)(kMaxSrcEvents + 1);
    }
    a.photons_left[i] = N;
}

// Exclusive prefix of n counts by one 1024-thread workgroup (tiles of 8192),
// accumulated in 64 bits; result[0] <- the total.  The int32 offsets are
// used only when the total is within kMaxSrcEvents (then they all fit).
__global__ __launch_bounds__(1024) void k_photo_scan(const int32_t* cnt, int32_t* offs, int n,
                                                     unsigned long long* result)
{
    __shared__ long long s_w[16];
    constexpr int kT = 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    long long carry = 0;
    for (int b = 0; b < n; b += 1024 * kT) {
        const int b0 = b + (int)threadIdx.x * kT;
        long long v[kT], sum = 0;
#pragma unroll
        for (int j = 0; j < kT; ++j) {
            v[j] = b0 + j < n ? (long long)cnt[b0 + j] : 0;
            sum += v[j];
        }
        long long incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const long long t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) s_w[wave] = incl;
        __syncthreads();
        long long woff = 0, tot = 0;
        for (int w = 0; w < 16; ++w) {
            const long long c = s_w[w];
            if (w < wave) woff += c;
            tot += c;
        }
        __syncthreads();
        long long acc = carry + woff + incl - sum;
#pragma unroll
        for (int j = 0; j < kT; ++j) {
            if (b0 + j < n) offs[b0 + j] = (int32_t)acc;
            acc += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        offs[n] = (int32_t)(carry <= kMaxSrcEvents ? carry : 0);
        result[0] = (unsigned long long)carry;
    }
}

// The events of this call: the total, or none beyond the limit.
__device__ __forceinline__ int src_total(const SrcArgs& a)
{
    const unsigned long long t = a.result[0];
    return t <= (unsigned long long)kMaxSrcEvents ? (int)t : 0;
}

// The thermal normals of event e (Box-Muller on block 1).
__device__ __forceinline__ void thermal_normals(const SrcArgs& a, uint32_t e, double& n1, double& n2, double& n3)
{
#pragma clang fp contract(off)
    const U4 R = philox4x32_10(e, a.pulse, a.run, kSourceBlock + 1u, a.key0, a.key1);
    const double a1 = u01d(R.a), a2 = u01d(R.b), a3 = u01d(R.c), a4 = u01d(R.d);
    const double twopi = 6.283185307179586;
    const double r1 = sqrt(-2.0 * log(a1)), r2 = sqrt(-2.0 * log(a3));
    n1 = r1 * cos(twopi * a2);
    n2 = r1 * sin(twopi * a2);
    n3 = r2 * cos(twopi * a4);
}

__device__ __forceinline__ double cell_temperature(const SrcArgs& a, int i, int k)
{
    return a.temp ? (double)a.temp[k * a.gnr + i] : a.t_scalar;
}

// Sauter angle (findPEangle.m:16-33): v = 1 - cos(theta) from the proposal
// density v / (A + v)^3 by its inverse, accepted with probability g(v) / g0.
// Only IEEE basic operations and sqrt: the decisions are reproducible.
__device__ __forceinline__ double sauter_v(const SrcArgs& a, uint32_t e)
{
#pragma clang fp contract(off)
    const double A = a.sA;
    double v = 0.0;
    for (int b = 0; b < kSauterBlocks; ++b) {
        const U4 R = philox4x32_10(e, a.pulse, a.run, kSourceBlock + 3u + (uint32_t)b, a.key0, a.key1);
        const double pr[4] = {u01d(R.a), u01d(R.b), u01d(R.c), u01d(R.d)};
        for (int h = 0; h < 2; ++h) {
            const double r1 = pr[2 * h], r2 = pr[2 * h + 1];
            const double d = (A + 2.0) * (A + 2.0) - 4.0 * r1;
            v = 2.0 * A / d * (2.0 * r1 + (A + 2.0) * sqrt(r1));
            const double gv = (2.0 - v) * (1.0 / (A + v) + a.sbg);
            if (r2 * a.sg0 < gv) return v;
        }
    }
    return v;
}

__global__ void k_photo_emit(SrcArgs a)
{
#pragma clang fp contract(off)
    const int ncells = a.nrings * a.gnz;
    const int total = src_total(a);
    // counts before this source (k_photo_counts updates them afterwards)
    const int e0 = *a.e_count, i0 = a.i_count ? *a.i_count : 0;
    // events past every free electron and ion slot store nothing
    const long long room = a.ix ? std::max((long long)a.e_cap - e0, (long long)a.i_cap - i0)
                                : (long long)a.e_cap - e0;
    const int n_emit = (int)std::max(0ll, std::min((long long)total, room));
    const double twopi = 6.283185307179586;
    const double pi = 3.141592653589793;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n_emit; e += gridDim.x * blockDim.x) {
        int lo = 0, hi = ncells;                   // offs[lo] <= e < offs[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.offs[mid] <= e) lo = mid; else hi = mid;
        }
        // empty cells share the offset of the next: the last cell with offs <= e holds e
        const int c = lo;
        const int ring = c % a.nrings, slab = c / a.nrings;
        const U4 B0 = philox4x32_10((uint32_t)e, a.pulse, a.run, kSourceBlock, a.key0, a.key1);
        // INCIDENCE1.m:62,67-68: z, r uniform in the cell; phi uniform in the wedge
        const double r = ((double)ring + u01d(B0.a)) * a.gdr;
        const double phi = u01d(B0.b) * a.phip;
        const double z = ((double)slab + u01d(B0.c)) * a.gdz;
        const double chi = twopi * u01d(B0.d);          // :102
        const double x = r * cos(phi), y = r * sin(phi);
        // photoelectron, photon along +z (:101-108)
        const double v = sauter_v(a, (uint32_t)e);
        const double ct = 1.0 - v, st = sqrt(v * (2.0 - v));
        const double pvx = a.v_pe * st * cos(chi), pvy = a.v_pe * st * sin(chi), pvz = a.v_pe * ct;
        // ion: thermal velocity (:69-71) and the recoil (:113-155)
        double n1, n2, n3;
        thermal_normals(a, (uint32_t)e, n1, n2, n3);
        const double sig = sqrt(a.k_b * cell_temperature(a, ring, slab) / a.m_mol);
        double mx = a.m_mol * (sig * n1) - a.m_e * pvx;
        double my = a.m_mol * (sig * n2) - a.m_e * pvy;
        double mz = a.p_photon + a.m_mol * (sig * n3) - a.m_e * pvz;
        if (a.auger) {                                   // :130-154
            const U4 B2 = philox4x32_10((uint32_t)e, a.pulse, a.run, kSourceBlock + 2u, a.key0, a.key1);
            const double th = pi * u01d(B2.a), ch = twopi * u01d(B2.b);
            const double ax = a.v_ae * sin(th) * cos(ch), ay = a.v_ae * sin(th) * sin(ch), az = a.v_ae * cos(th);
            mx = mx - a.m_e * ax;
            my = my - a.m_e * ay;
            mz = mz - a.m_e * az;
            const int64_t d = (int64_t)e0 + total + e;   // Auger electrons after every photoelectron
            if (d < a.e_cap) {
                a.ex[d] = (float)x; a.ey[d] = (float)y; a.ez[d] = (float)z;
                a.evx[d] = (float)ax; a.evy[d] = (float)ay; a.evz[d] = (float)az;
            }
        }
        const int64_t d = (int64_t)e0 + e;
        if (d < a.e_cap) {
            a.ex[d] = (float)x; a.ey[d] = (float)y; a.ez[d] = (float)z;
            a.evx[d] = (float)pvx; a.evy[d] = (float)pvy; a.evz[d] = (float)pvz;
        }
        const int64_t di = (int64_t)i0 + e;
        if (a.ix && di < a.i_cap) {
            a.ix[di] = (float)x; a.iy[di] = (float)y; a.iz[di] = (float)z;
            a.ivx[di] = (float)(mx / a.m_mol); a.ivy[di] = (float)(my / a.m_mol); a.ivz[di] = (float)(mz / a.m_mol);
        }
    }
}

// INCIDENCE1.m:87-92 per (ring, slab) cell with events: N' = N - FN count,
// E' = 3/2 kB N T - sum 1/2 FN m v_th^2 (in event order), n' = N'/V,
// T' = E' / (3/2 kB N').  A cell emptied completely keeps T and gets n = 0.
__global__ void k_photo_deplete(SrcArgs a)
{
#pragma clang fp contract(off)
    const int ncells = a.nrings * a.gnz;
    if (src_total(a) == 0) return;                 // nothing emitted (or beyond the limit)
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
        const int cnt = a.offs[c + 1] - a.offs[c];
        if (cnt == 0) continue;
        const int ring = c % a.nrings, slab = c / a.nrings;
        const double T = cell_temperature(a, ring, slab);
        const double sig = sqrt(a.k_b * T / a.m_mol);
        double s2 = 0.0;
        for (int e = a.offs[c]; e < a.offs[c + 1]; ++e) {
            double n1, n2, n3;
            thermal_normals(a, (uint32_t)e, n1, n2, n3);
            const double vx = sig * n1, vy = sig * n2, vz = sig * n3;
            s2 = s2 + ((vx * vx + vy * vy) + vz * vz);
        }
        const double V = a.phip / 2.0 * ((double)(ring + 1) * (ring + 1) - (double)ring * ring) * a.gdr * a.gdr * a.gdz;
        const double n = (double)a.dens[slab * a.gnr + ring];
        const double N = n * V;
        const double Nn = N - a.fn * (double)cnt;
        const double E = 1.5 * a.k_b * N * T - 0.5 * a.fn * a.m_mol * s2;
        const int gi = slab * a.gnr + ring;
        if (Nn > 0.0) {
            a.dens_out[gi] = (float)(Nn / V);
            a.temp_out[gi] = (float)(E / (1.5 * a.k_b * Nn));
        } else {
            a.dens_out[gi] = 0.0f;
            a.temp_out[gi] = (float)T;
        }
    }
}

// Cells without events keep their gas values in the output arrays.
__global__ void k_photo_copy_gas(SrcArgs a)
{
    const int n = a.gnr * a.gnz;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int ring = c % a.gnr, slab = c / a.gnr;
        if (a.dens_out != a.dens) a.dens_out[c] = a.dens[c];
        if (a.temp_out != a.temp) a.temp_out[c] = (float)cell_temperature(a, ring, slab);
    }
}

__global__ void k_photo_counts(SrcArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t total = src_total(a);
    const int64_t ne = total * (a.auger ? 2 : 1);
    const int64_t e0 = *a.e_count;
    const int64_t eadd = std::min<int64_t>(ne, std::max<int64_t>(0, (int64_t)a.e_cap - e0));
    *a.e_count = (int32_t)(e0 + eadd);
    int64_t iadd = 0;
    if (a.ix) {
        const int64_t i0 = *a.i_count;
        iadd = std::min<int64_t>(total, std::max<int64_t>(0, (int64_t)a.i_cap - i0));
        *a.i_count = (int32_t)(i0 + iadd);
    }
    a.result[1] = (unsigned long long)eadd;
    a.result[2] = (unsigned long long)iadd;
    a.result[3] = (unsigned long long)(ne - eadd);
    a.result[4] = (unsigned long long)(a.ix ? total - iadd : 0);
}

// findV.m:5-8 (m = 9.109e-31), speed of kinetic energy ke (J)
double find_v(double ke)
{
#pragma clang fp contract(off)
    const double c = 3e8, m = 9.109e-31;
    const double g = 1.0 + ke / (m * c * c);
    return c * std::sqrt(1.0 - 1.0 / (g * g));
}

double interp1_or_nan(const std::vector<double>& x, const std::vector<double>& y, double q)
{
#pragma clang fp contract(off)
    if (!(q >= x.front()) || !(q <= x.back())) return NAN;
    size_t lo = 0, hi = x.size() - 1;
    while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        if (x[mid] <= q) lo = mid; else hi = mid;
    }
    const double s = (q - x[lo]) / (x[hi] - x[lo]);
    return y[lo] + s * (y[hi] - y[lo]);
}

}  // namespace

extern "C" gasx_status gasx_photoionize(gasx_ctx* ctx, gasx_species* e, gasx_species* ions, const gasx_gas* g,
                                        const gasx_photo_source* src, gasx_photo_result* out)
{
#pragma clang fp contract(off)
    if (!ctx || !e || !g || !src || !out) return GASX_E_ARG;
    if (!e->x || !e->y || !e->z || !e->vx || !e->vy || !e->vz || !e->count || e->capacity < 1 ||
        e->capacity > (int64_t)INT32_MAX - kChunk)
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: invalid electron species");
    if (ions && (!ions->x || !ions->y || !ions->z || !ions->vx || !ions->vy || !ions->vz || !ions->count ||
                 ions->capacity < 1 || ions->capacity > (int64_t)INT32_MAX - kChunk))
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: invalid ion species");
    if (!g->density || g->gnr < 1 || g->gnz < 1 || !(g->gdr > 0) || !(g->gdz > 0))
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: invalid gas description");
    if (!(src->photon_kev > 0) || !(src->n_photons >= 0) || !(src->beam_radius > 0) || !src->binding_ev ||
        src->n_shells < 1 || !src->mac || src->n_mac < 2 || !(src->atomic_mass > 0) || !(src->auger_kev >= 0))
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: invalid source description");
    if ((src->temperature_out != nullptr) != (src->density_out != nullptr))
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: depletion needs both density_out and temperature_out");
    const gasx_params& p = ctx->p;
    // findEShell.m:4-12: the first shell (file order) the photon can ionize
    int shell = -1;
    for (int i = 0; i < src->n_shells; ++i)
        if (src->photon_kev >= src->binding_ev[i] / 1000.0) { shell = i; break; }
    if (shell < 0)
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: photon energy below every binding energy");
    const double e_shell = src->binding_ev[shell] / 1000.0;
    const bool auger = shell != src->n_shells - 1;          // INCIDENCE1.m:124
    // findSigmaPE.m:8-22: MAC interpolated at Ex in MeV, sigma = MAC m_a / N_A 1e-4
    std::vector<double> mx(src->n_mac), my(src->n_mac);
    for (int i = 0; i < src->n_mac; ++i) {
        mx[i] = src->mac[2 * i];
        my[i] = src->mac[2 * i + 1];
        if (i > 0 && !(mx[i] > mx[i - 1]))
            return fail(ctx, GASX_E_ARG, "gasx_photoionize: MAC energies must increase");
    }
    const double mac = interp1_or_nan(mx, my, src->photon_kev / 1000.0);
    if (!(mac >= 0)) return fail(ctx, GASX_E_ARG, "gasx_photoionize: photon energy outside the MAC table");
    const double sigma_pe = mac * src->atomic_mass / 6.02e23 * 1e-4;
    // GUASS.m:3-15: ring i of width gdr carries the integral of x exp(-x^2 / 2 s^2)
    // over it (closed form), s = laser_r / 3, rings up to round(laser_r / gdr)
    const double s = src->beam_radius / 3.0;
    int nrings = (int)std::lround(src->beam_radius / g->gdr);
    nrings = std::max(1, std::min(nrings, std::min(g->gnr, kMaxSrcRings)));
    std::vector<double> ring(nrings);
    double tot = 0.0;
    for (int i = 0; i < nrings; ++i) {
        const double a0 = i * g->gdr, a1 = (i + 1) * g->gdr;
        ring[i] = s * s * (std::exp(-a0 * a0 / (2 * s * s)) - std::exp(-a1 * a1 / (2 * s * s)));
        tot += ring[i];
    }
    for (int i = 0; i < nrings; ++i) ring[i] = src->n_photons * (ring[i] / tot);
    // photoelectron energy and the Sauter constants (findPEangle.m:6-14, m_e = 9.11e-31)
    const double dE = (src->photon_kev - e_shell) * 1.6e-16;   // INCIDENCE1.m:28
    const double erest = 9.11e-31 * 3e8 * 3e8;
    const double gam = 1.0 + dE / erest;
    const double beta = std::sqrt(dE * (dE + 2 * erest)) / (dE + erest);
    SrcArgs a{};
    a.sA = 1.0 / beta - 1.0;
    a.sbg = 0.5 * beta * gam * (gam - 1.0) * (gam - 2.0);
    a.sg0 = 2.0 * (1.0 / a.sA + a.sbg);
    if (!(dE > 0)) return fail(ctx, GASX_E_ARG, "gasx_photoionize: no photoelectron energy");
    a.nrings = nrings; a.gnr = g->gnr; a.gnz = g->gnz;
    a.gdr = g->gdr; a.gdz = g->gdz; a.phip = p.phip;
    a.dens = g->density; a.temp = g->temperature; a.t_scalar = g->t_scalar;
    a.sigma_pe = sigma_pe; a.fn = p.fn;
    a.v_pe = find_v(dE);
    a.v_ae = find_v(src->auger_kev * 1.6e-16);
    a.auger = auger ? 1 : 0;
    a.p_photon = src->photon_kev * 1.6e-16 / 3e8;               // INCIDENCE1.m:113
    a.m_e = 9.11e-31;                                            // :21
    a.m_mol = 2.0 * src->atomic_mass * 1.6726e-27;                // :24
    a.k_b = 1.38064853e-23;                                      // :23
    a.key0 = (uint32_t)src->seed; a.key1 = (uint32_t)(src->seed >> 32);
    a.pulse = src->pulse; a.run = src->run;
    a.ex = e->x; a.ey = e->y; a.ez = e->z; a.evx = e->vx; a.evy = e->vy; a.evz = e->vz;
    a.e_count = e->count; a.e_cap = (int32_t)e->capacity;
    if (ions) {
        a.ix = ions->x; a.iy = ions->y; a.iz = ions->z; a.ivx = ions->vx; a.ivy = ions->vy; a.ivz = ions->vz;
        a.i_count = ions->count; a.i_cap = (int32_t)ions->capacity;
    }
    a.dens_out = src->density_out; a.temp_out = src->temperature_out;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    // work buffers: ring photons and leftovers, counts, offsets, result words
    const size_t ncells = (size_t)nrings * g->gnz;
    if (ncells >= (size_t)INT32_MAX) return fail(ctx, GASX_E_ARG, "gasx_photoionize: source grid too large");
    const size_t bytes = 2 * nrings * sizeof(double) + (2 * ncells + 1) * sizeof(int32_t) + 8 * sizeof(unsigned long long);
    char* w = nullptr;
    HIPCHK(hipMalloc((void**)&w, bytes + 256), "alloc source buffers");
    char* q = w;
    auto take = [&](size_t b) { char* r = q; q += (b + 7) & ~(size_t)7; return r; };
    double* d_ring = (double*)take(nrings * sizeof(double));
    a.ring_photons = d_ring;
    a.photons_left = (double*)take(nrings * sizeof(double));
    a.counts = (int32_t*)take(ncells * sizeof(int32_t));
    a.offs = (int32_t*)take((ncells + 1) * sizeof(int32_t));
    a.result = (unsigned long long*)take(8 * sizeof(unsigned long long));
    std::vector<double> left(nrings);
    unsigned long long res[8] = {};
    hipError_t err = hipMemcpyAsync(d_ring, ring.data(), nrings * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_photo_attenuate, dim3((nrings + 255) / 256), dim3(256), 0, ctx->stream, a);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_photo_scan, dim3(1), dim3(1024), 0, ctx->stream, a.counts, a.offs, (int)ncells, a.result);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_photo_emit, dim3(2048), dim3(256), 0, ctx->stream, a);
        err = hipGetLastError();
    }
    if (err == hipSuccess && a.dens_out) {
        hipLaunchKernelGGL(k_photo_copy_gas, dim3(256), dim3(256), 0, ctx->stream, a);
        err = hipGetLastError();
        if (err == hipSuccess) {
            hipLaunchKernelGGL(k_photo_deplete, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, ctx->stream, a);
            err = hipGetLastError();
        }
    }
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_photo_counts, dim3(1), dim3(64), 0, ctx->stream, a);
        err = hipGetLastError();
    }
    if (err == hipSuccess)
        err = hipMemcpyAsync(left.data(), a.photons_left, nrings * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(res, a.result, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    hipFree(w);
    if (err != hipSuccess) return hip_fail(ctx, err, "photoionization source");
    if (res[0] > (unsigned long long)kMaxSrcEvents)
        return fail(ctx, GASX_E_ARG, "gasx_photoionize: " + std::to_string(res[0]) + " events exceed the limit of " +
                                         std::to_string(kMaxSrcEvents) + " per call (raise FN)");
    double nl = 0.0;
    for (int i = 0; i < nrings; ++i) nl += left[i];
    out->photons_left = nl;
    out->n_events = (int64_t)res[0];
    out->n_electrons = (int64_t)res[1];
    out->n_ions = (int64_t)res[2];
    out->n_dropped_electrons = (int64_t)res[3];
    out->n_dropped_ions = (int64_t)res[4];
    out->sigma_pe = sigma_pe;
    out->electron_kev = src->photon_kev - e_shell;
    out->auger = auger ? 1 : 0;
    out->n_rings = nrings;
    return (res[3] || res[4]) ? fail(ctx, GASX_E_CAPACITY, "gasx_photoionize: capacity exceeded") : GASX_OK;
}
