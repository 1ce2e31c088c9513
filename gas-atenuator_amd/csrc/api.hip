// Host side of libgasx.so: context, table preparation, the C ABI entry points
// and the fused transport loop.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "context.h"

using namespace gasx;
using namespace gasx_host;

// ---------------------------------------------------------------------------
// Fixed staging for gasx_collide_noalloc: one array per GPU in the code
// object, held until the process ends.  Layout in 32-bit words: six FP32
// secondary staging arrays and six ion staging arrays of kNaSlots each, the
// per-segment secondary and ion counts, the append scan (segment offsets,
// group totals) and the append's words.  Every word a call reads is written
// by that call first (each segment below the slot count gets its count, the
// scan writes its offsets, totals and bases), except the one-launch append's
// four sync words, which are zero when the code object loads and which every
// append leaves zero, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kNaSlots = GASX_NOALLOC_MAX_SLOTS;
constexpr int64_t kNaSegs = kNaSlots / kSeg;
static_assert(kNaSlots % kSeg == 0, "whole staging segments");
constexpr int64_t kNaScanWords = ((kNaSegs + (kNaSegs + 1023) / 1024 + 1 + 63) / 64) * 64;
constexpr int64_t kNaStageCnt = 12 * kNaSlots;            // word offsets
constexpr int64_t kNaIonCnt = kNaStageCnt + kNaSegs;
constexpr int64_t kNaScan = kNaIonCnt + kNaSegs;
constexpr int64_t kNaBase = kNaScan + kNaScanWords;
constexpr int64_t kNaWords = kNaBase + 4 + 2 * kAppendSyncWords;
constexpr int kNaMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint32_t scratch[kNaWords];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNaMaxDevices] = {};   // this device's address of `scratch`

// With the context's device current: look up (once per device) where this
// device holds `scratch`.  Host code uses only the returned address.
gasx_status bind_scratch(gasx_ctx* ctx)
{
    if (ctx->device >= kNaMaxDevices) return GASX_OK;      // gasx_collide_noalloc then refuses
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(ctx, GASX_E_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    return GASX_OK;
}
}  // namespace

size_t gasx_host::next_event(gasx_ctx* c)
{
    if (c->ev_used == c->events.size()) {
        hipEvent_t e;
        // timestamps only (read after a stream synchronize): no system-scope fence
        // on record, which would otherwise add ~3 us per recorded launch
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return (size_t)-1;
        c->events.push_back(e);
    }
    return c->ev_used++;
}

namespace {

double interp1_nan(const double* x, const double* y, int n, double xq)
{
    if (!(xq >= x[0]) || !(xq <= x[n - 1])) return NAN;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (x[mid] <= xq) lo = mid; else hi = mid;
    }
    double s = (xq - x[lo]) / (x[hi] - x[lo]);
    return y[lo] + s * (y[hi] - y[lo]);
}

bool valid_species(const gasx_species* s)
{
    return s && s->x && s->y && s->z && s->vx && s->vy && s->vz && s->count && s->capacity > 0 &&
           s->capacity < (int64_t)INT32_MAX - kChunk && s->mass > 0.0;
}

bool valid_fields(const gasx_fields* f) { return f && f->er && f->ez && f->hphi; }

}  // namespace

// ---------------------------------------------------------------------------
// Tables: electron_collision.m:12-41 merges and the find_CS0.m units.
// ---------------------------------------------------------------------------
int gasx::prepare_tables(const gasx_raw_tables* raw, HostTables* h, DevTables* d)
{
    const double sqr_a0 = 2.8002852e-21;   // electron_collision.m:12, find_CS0.m:17
    if (!raw || !raw->t_el || !raw->t_sel || !raw->t_sel1 || !raw->t_del || !raw->t_sdel ||
        !raw->t_ion || !raw->t_exc)
        return 1;
    const int n_small = raw->n_sel1 - 1;                     // electron_collision.m:17
    const int n_dcs = (raw->w_sdel - 1) + (raw->w_del - 1);  // electron_collision.m:23-25
    if (raw->n_el < 2 || raw->n_el > kMaxTab || raw->n_sel < 2 || raw->n_sel > kMaxTab ||
        raw->n_ion < 2 || raw->n_ion > kMaxTab || raw->n_exc < 1 || raw->n_exc > kMaxExc ||
        n_small < 0 || n_dcs < 2 || n_dcs > kMaxTab || n_small + raw->n_el != n_dcs)
        return 2;
    *h = HostTables{};
    h->n_el = raw->n_el;
    for (int i = 0; i < raw->n_el; ++i) {
        h->el_e[i] = raw->t_el[i];
        h->el_s[i] = raw->t_el[raw->n_el + i];
    }
    h->n_sel = raw->n_sel;
    for (int i = 0; i < raw->n_sel; ++i) {
        h->sel_e[i] = raw->t_sel[2 * i];
        h->sel_s[i] = raw->t_sel[2 * i + 1];
    }
    h->n_ion = raw->n_ion;
    for (int i = 0; i < raw->n_ion; ++i) {
        h->ion_e[i] = raw->t_ion[2 * i];
        h->ion_s[i] = raw->t_ion[2 * i + 1];
    }
    h->n_exc = raw->n_exc;
    for (int i = 0; i < raw->n_exc; ++i) {   // columns j, A, omega, v, miu
        h->exc_w[i] = raw->t_exc[5 * i + 0];
        h->exc_a[i] = raw->t_exc[5 * i + 1];
        h->exc_om[i] = raw->t_exc[5 * i + 2];
        h->exc_nu[i] = raw->t_exc[5 * i + 3];
        h->exc_ga[i] = raw->t_exc[5 * i + 4];
        if (!(h->exc_w[i] > 0)) return 3;
    }
    h->n_dcs = n_dcs;
    for (int i = 0; i < n_small; ++i) h->dcs_tcs[i] = raw->t_sel1[2 * i + 1] * 1e-20 / sqr_a0 / 2.0;
    for (int i = 0; i < raw->n_el; ++i) h->dcs_tcs[n_small + i] = h->el_s[i];
    for (int j = 0; j < n_dcs; ++j) {
        const double* tab;
        int w, col;
        if (j < raw->w_sdel - 1) { tab = raw->t_sdel; w = raw->w_sdel; col = j + 1; }
        else { tab = raw->t_del; w = raw->w_del; col = j - (raw->w_sdel - 1) + 1; }
        h->dcs_e[j] = tab[col];
        if (!(tab[col] > 0)) return 4;
        h->dcs_loge[j] = std::log(tab[col]);
        for (int k = 0; k < kNAng; ++k) {
            double v = tab[(k + 1) * w + col];
            if (!(v > 0)) return 5;
            h->dcs_logv[j * kNAng + k] = std::log(v);
        }
    }
    for (int j = 1; j < n_dcs; ++j)
        if (!(h->dcs_loge[j] > h->dcs_loge[j - 1])) return 6;

    // device image: energies in eV, sigma in m^2
    std::memset(d, 0, sizeof(*d));
    SmallTabs& s = d->s;
    s.n_el = h->n_el; s.n_sel = h->n_sel; s.n_ion = h->n_ion; s.n_exc = h->n_exc; s.n_dcs = n_dcs;
    for (int i = 0; i < h->n_el; ++i) {
        s.el_e[i] = (float)(h->el_e[i] * 1000.0);
        s.el_s[i] = (float)(h->el_s[i] * sqr_a0 * 2.0);     // find_CS0.m:21
    }
    for (int i = 0; i < h->n_sel; ++i) {
        s.sel_e[i] = (float)h->sel_e[i];
        s.sel_s[i] = (float)(h->sel_s[i] * 1e-20);           // find_CS0.m:25
    }
    for (int i = 0; i < h->n_ion; ++i) {
        s.ion_e[i] = (float)h->ion_e[i];
        s.ion_s[i] = (float)(h->ion_s[i] * 1e-20);           // find_CS0.m:33
    }
    const double q0 = 6.542e-14;                             // find_CS0.m:42
    for (int i = 0; i < h->n_exc; ++i) {
        s.exc_w[i] = (float)h->exc_w[i];
        s.exc_q[i] = (float)(q0 * h->exc_a[i] / (h->exc_w[i] * h->exc_w[i]) * 1e-4);
        s.exc_a[i] = (float)(h->exc_a[i] / (h->exc_w[i] * h->exc_w[i]));
        s.exc_om[i] = (float)h->exc_om[i];
        s.exc_nu[i] = (float)h->exc_nu[i];
        s.exc_ga[i] = (float)h->exc_ga[i];
        s.exc_l2w[i] = (float)std::log2(h->exc_w[i]);
    }
    double sind[kNAng];
    for (int k = 0; k < kNAng; ++k) {
        sind[k] = (k == 0 || k == 180) ? 0.0 : (k == 90 ? 1.0 : std::sin(k * 3.14159265358979323846 / 180.0));
        s.sind[k] = (float)sind[k];
    }
    for (int j = 0; j < n_dcs; ++j) {
        s.dcs_loge[j] = (float)h->dcs_loge[j];
        double p[kNAng], z = 0.0;
        for (int k = 0; k < kNAng; ++k) {
            d->dcs_logv[j * kNAng + k] = (float)h->dcs_logv[j * kNAng + k];
            p[k] = sind[k] * std::exp(h->dcs_logv[j * kNAng + k]);
            d->dcs_p[j * kNAng + k] = (float)p[k];
        }
        for (int k = 0; k + 1 < kNAng; ++k) z += p[k] + p[k + 1];
        double acc = 0.0;
        float* C = &d->dcs_cdf[j * kNAng];
        C[0] = 0.0f;
        for (int k = 0; k + 1 < kNAng; ++k) {
            acc += p[k] + p[k + 1];
            C[k + 1] = (k + 2 == kNAng) ? 1.0f : (float)(acc / z);
        }
        s.dcs_z[j] = (float)z;
        // guide: largest k in [0, 179] with C[k] <= g / kGuide (on the FP32 values)
        for (int g = 0; g <= kGuide; ++g) {
            const float b = (float)g / (float)kGuide;
            int k = 0;
            while (k < kNAng - 2 && C[k + 1] <= b) ++k;
            s.guide[j][g] = (uint8_t)k;
        }
    }
    return 0;
}

// sigma_T(E) * v(E) maximised over the collision window (find_CS0.m forms,
// FP64) on a dense log grid plus every table node.  Used for the exact
// rejection bound p_T <= dt * n_max * max(v sigma_T).
double gasx::max_rate_coefficient(const HostTables& h, const gasx_params& p)
{
    auto sig_t = [&](double T) {
        double se = T >= 50.0 ? interp1_nan(h.el_e, h.el_s, h.n_el, T / 1000.0) * 2.8002852e-21 * 2.0
                              : interp1_nan(h.sel_e, h.sel_s, h.n_sel, T) * 1e-20;
        double si = interp1_nan(h.ion_e, h.ion_s, h.n_ion, T) * 1e-20;
        double sx = 0.0;
        for (int i = 0; i < h.n_exc; ++i) {
            double a = 6.542e-14 * h.exc_a[i] / (h.exc_w[i] * h.exc_w[i]);
            double b = std::pow(h.exc_w[i] / T, h.exc_om[i]);
            double c = std::pow(1.0 - std::pow(h.exc_w[i] / T, h.exc_ga[i]), h.exc_nu[i]);
            sx += a * b * c;
        }
        return se + si + sx * 1e-4;
    };
    auto speed = [&](double T) {   // push-side electron (EM4.cu KE with field6 constants)
        double g1 = T * p.e_em / (p.me_em * p.c * p.c);
        return p.c * std::sqrt(g1 * (g1 + 2.0)) / (1.0 + g1);
    };
    double best = 0.0;
    auto visit = [&](double T) {
        if (T < 15.6 || T > 20000.0) return;
        double v = speed(T) * sig_t(T);
        if (v > best) best = v;
    };
    const int N = 200000;
    for (int i = 0; i <= N; ++i) visit(15.6 * std::pow(20000.0 / 15.6, (double)i / N));
    for (int i = 0; i < h.n_el; ++i) visit(h.el_e[i] * 1000.0);
    for (int i = 0; i < h.n_sel; ++i) visit(h.sel_e[i]);
    for (int i = 0; i < h.n_ion; ++i) visit(h.ion_e[i]);
    return best;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" int gasx_abi_version(void) { return GASX_ABI_VERSION; }

extern "C" const char* gasx_status_string(gasx_status s)
{
    switch (s) {
    case GASX_OK: return "ok";
    case GASX_E_ARG: return "invalid argument";
    case GASX_E_CAPACITY: return "capacity exceeded";
    case GASX_E_HIP: return "HIP runtime error";
    case GASX_E_NOMEM: return "out of device memory";
    case GASX_E_STATE: return "invalid call sequence";
    case GASX_E_NUMERIC: return "numeric range error";
    }
    return "unknown status";
}

extern "C" const char* gasx_last_error(const gasx_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

static void fill_geometry(gasx_ctx* c)
{
    const gasx_params& p = c->p;
    Geom& g = c->geom;
    g.nr = p.nr; g.nz = p.nz;
    g.dr = (float)p.dr; g.dz = (float)p.dz;
    g.inv_dr = (float)(1.0 / p.dr); g.inv_dz = (float)(1.0 / p.dz);
    g.rp = (float)(p.nr * p.dr); g.lp = (float)(p.nz * p.dz);
    g.phip = (float)p.phip; g.cphip = (float)std::cos(p.phip); g.sphip = (float)std::sin(p.phip);
    g.dt = (float)p.ddt;
    g.c = (float)p.c; g.inv_c2 = (float)(1.0 / (p.c * p.c)); g.mu0 = (float)p.mu0;

    FdtdArgs& a = c->fa;
    a.nr = p.nr; a.nz = p.nz;
    a.n_sub = p.n_sub;
    a.halo = p.n_sub + 3;
    // cycles of 6+ substeps run as two launches of n_sub/2 and the rest:
    // halos of n_sub/2 + 2 instead of n_sub + 3 (measured 15% faster at
    // 256x512 with n_sub = 10)
    a.split = p.n_sub >= 6 ? p.n_sub / 2 : 0;
    a.part = 0;
    a.dtf = p.ddt / p.n_sub;
    a.ch_r = a.dtf / (p.mu0 * p.dr);
    a.ch_z = a.dtf / (p.mu0 * p.dz);
    a.ce = a.dtf / p.eps0;
    a.cez = a.dtf / (p.eps0 * p.dz);
    a.c4 = 4 * a.dtf / (p.eps0 * p.dr);
    a.dr = p.dr;
    a.ez_metric = p.ez_metric;
    a.axis_sign = p.axis_j_sign >= 0 ? 1 : -1;
    a.junit = p.fn * p.e_em * p.c / 4294967296.0;
    a.vcoef = p.phip / 2 * p.dr * p.dr * p.dz;
    // default tiles: 16 x 32 (r x z), shrunk for tiny grids
    a.ti = std::min(GASX_FDTD_TI, p.nr);
    a.tk = std::min(GASX_FDTD_TK, p.nz);
    a.ntr = (p.nr + a.ti - 1) / a.ti;
    a.ntk = (p.nz + a.tk - 1) / a.tk;
}

extern "C" gasx_status gasx_ctx_create(int device, const gasx_params* p, const gasx_raw_tables* t,
                                       gasx_ctx** out)
{
    if (!out) return GASX_E_ARG;
    *out = nullptr;
    if (!p || !t) return GASX_E_ARG;
    if (device < 0) return GASX_E_ARG;   // the library has no CPU backend
    if (p->nr < 2 || p->nz < 2 || !(p->dr > 0) || !(p->dz > 0) || !(p->phip > 0) ||
        !(p->phip < 3.14159265358979 / 2) || !(p->ddt > 0) || p->n_sub < 1 || p->n_sub > 64 ||
        !(p->c > 0) || !(p->eps0 > 0) || !(p->mu0 > 0) || !(p->e_em > 0) || !(p->me_em > 0) ||
        !(p->e_mc > 0) || !(p->me_mc > 0) || !(p->m_ion > 0) || !(p->k_b > 0) ||
        (int64_t)p->nr * p->nz > (int64_t)1 << 28)
        return GASX_E_ARG;
    gasx_ctx* ctx = new (std::nothrow) gasx_ctx();
    if (!ctx) return GASX_E_NOMEM;
    ctx->device = device;
    ctx->p = *p;
    DevTables dt;
    if (prepare_tables(t, &ctx->ht, &dt) != 0) {
        delete ctx;
        return GASX_E_ARG;
    }
    fill_geometry(ctx);
    ctx->max_rate = max_rate_coefficient(ctx->ht, ctx->p);
    gasx_status st = GASX_OK;
    auto chk = [&](hipError_t e, const char* w) {
        if (e != hipSuccess && st == GASX_OK) st = hip_fail(ctx, e, w);
    };
    chk(hipSetDevice(device), "hipSetDevice");
    if (st == GASX_OK) st = bind_scratch(ctx);
    if (st == GASX_OK) chk(fdtd_prepare(), "field kernel LDS limit");
    if (st == GASX_OK) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
        chk(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        ctx->own_stream = true;
        const size_t nh = (size_t)p->nr * p->nz, ner = (size_t)p->nr * (p->nz + 1),
                     nez = (size_t)(p->nr + 1) * p->nz;
        chk(dalloc(&ctx->d_tab, 1), "alloc tables");
        chk(dalloc(&ctx->er32, ner), "alloc er32");
        chk(dalloc(&ctx->ez32, nez), "alloc ez32");
        chk(dalloc(&ctx->h32, nh), "alloc h32");
        chk(dalloc(&ctx->er_b, ner), "alloc er_b");
        chk(dalloc(&ctx->ez_b, nez), "alloc ez_b");
        chk(dalloc(&ctx->h_b, nh), "alloc h_b");
        chk(dalloc(&ctx->er_m, ner), "alloc er_m");
        chk(dalloc(&ctx->ez_m, nez), "alloc ez_m");
        chk(dalloc(&ctx->h_m, nh), "alloc h_m");
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 2; ++k) chk(dalloc(&ctx->q[b][k], nh * kJCopies), "alloc q");
        chk(dalloc(&ctx->jr_tmp, nh), "alloc jr");
        chk(dalloc(&ctx->jz_tmp, nh), "alloc jz");
        chk(dalloc(&ctx->d_counters, (size_t)kNCounters * kCounterSlots), "alloc counters");
        chk(dalloc(&ctx->scan_tmp, 1), "alloc scan");
        chk(dalloc(&ctx->base_tmp, 4 + 2 * kAppendSyncWords), "alloc base");
        chk(dalloc(&ctx->d_tally, 4 * 256 + 8), "alloc tally");
        chk(hipHostMalloc((void**)&ctx->h_count, kSortInfoWords * sizeof(int32_t), hipHostMallocDefault), "alloc pinned count");
        chk(hipEventCreateWithFlags(&ctx->sort_ev, hipEventDisableTiming), "sort event");
        if (st == GASX_OK) {
            chk(hipMemcpy(ctx->d_tab, &dt, sizeof(dt), hipMemcpyHostToDevice), "upload tables");
            chk(hipMemset(ctx->d_counters, 0, (size_t)kNCounters * kCounterSlots * sizeof(unsigned long long)),
                "zero counters");
            chk(hipMemset(ctx->er32, 0, ner * sizeof(float)), "zero er32");
            chk(hipMemset(ctx->ez32, 0, nez * sizeof(float)), "zero ez32");
            chk(hipMemset(ctx->h32, 0, nh * sizeof(float)), "zero h32");
            chk(hipMemset(ctx->base_tmp, 0, (4 + 2 * kAppendSyncWords) * sizeof(int32_t)), "zero append words");
        }
    }
    if (st != GASX_OK) {
        std::string m = ctx->err;
        gasx_ctx_destroy(ctx);
        return st;
    }
    *out = ctx;
    return GASX_OK;
}

extern "C" gasx_status gasx_ctx_destroy(gasx_ctx* ctx)
{
    if (!ctx) return GASX_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    void* bufs[] = {ctx->d_tab, ctx->er32, ctx->ez32, ctx->h32, ctx->er_b, ctx->ez_b, ctx->h_b,
                    ctx->er_m, ctx->ez_m, ctx->h_m,
                    ctx->q[0][0], ctx->q[0][1], ctx->q[1][0], ctx->q[1][1], ctx->jr_tmp, ctx->jz_tmp,
                    ctx->d_counters, ctx->stage_cnt, ctx->istage_cnt, ctx->scan_tmp, ctx->base_tmp,
                    ctx->sort_tmp, ctx->d_tally};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (int c = 0; c < 6; ++c) {
        if (ctx->stage[c]) hipFree(ctx->stage[c]);
        if (ctx->istage[c]) hipFree(ctx->istage[c]);
    }
    if (ctx->h_count) hipHostFree(ctx->h_count);
    if (ctx->sort_ev) hipEventDestroy(ctx->sort_ev);
    for (hipEvent_t e : ctx->events) hipEventDestroy(e);
    batch_release(ctx);
    diff_release(ctx);
    dsmc_release(ctx);
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return GASX_OK;
}

extern "C" gasx_status gasx_ctx_set_stream(gasx_ctx* ctx, void* s)
{
    if (!ctx) return GASX_E_ARG;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
    ctx->stream = (hipStream_t)s;
    return GASX_OK;
}

extern "C" gasx_status gasx_synchronize(gasx_ctx* ctx)
{
    if (!ctx) return GASX_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return GASX_OK;
}

// staging for appends, sized for `cap` slots
static gasx_status ensure_staging(gasx_ctx* ctx, int64_t cap, bool ions)
{
    const int64_t slots = ((cap + kSeg - 1) / kSeg) * kSeg;
    if (slots / kSeg > (int64_t)1024 * 1024)      // the append scan's group limit
        return fail(ctx, GASX_E_ARG, "capacity above 2^28 slots is not supported by the append");
    if (slots > ctx->stage_slots || (ions && !ctx->istage[0])) {
        const int64_t ns = std::max(slots, ctx->stage_slots);
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        for (int c = 0; c < 6; ++c) {
            if (ctx->stage[c]) hipFree(ctx->stage[c]);
            if (ctx->istage[c]) hipFree(ctx->istage[c]);
            ctx->stage[c] = nullptr;
            ctx->istage[c] = nullptr;
        }
        if (ctx->stage_cnt) hipFree(ctx->stage_cnt);
        if (ctx->istage_cnt) hipFree(ctx->istage_cnt);
        if (ctx->scan_tmp) hipFree(ctx->scan_tmp);
        ctx->stage_cnt = ctx->istage_cnt = ctx->scan_tmp = nullptr;
        for (int c = 0; c < 6; ++c) {
            HIPCHK(dalloc(&ctx->stage[c], ns), "alloc stage");
            HIPCHK(dalloc(&ctx->istage[c], ns), "alloc ion stage");
        }
        HIPCHK(dalloc(&ctx->stage_cnt, ns / kSeg), "alloc stage cnt");
        HIPCHK(dalloc(&ctx->istage_cnt, ns / kSeg), "alloc ion stage cnt");
        // segment offsets, then one total per group of 1024 segments
        HIPCHK(dalloc(&ctx->scan_tmp, ns / kSeg + (ns / kSeg + 1023) / 1024 + 1), "alloc scan");
        ctx->stage_slots = ns;
    }
    return GASX_OK;
}

// Where a collision launch stages its secondaries and ions and where their
// append keeps its scan: the context's buffers, or the fixed array.
struct Staging {
    float* sec[6];
    int32_t* sec_cnt;
    float* ion[6];
    int32_t* ion_cnt;
    int32_t* scan_tmp;    // max_segs segment offsets, then the group totals
    int32_t* base_tmp;    // [0..1] secondary append, [2..3] ion append (two-launch path);
                          // then their one-launch sync words, kAppendSyncWords each (zero
                          // between calls)
    int32_t max_segs;
};

static Staging ctx_staging(const gasx_ctx* ctx)
{
    Staging t{};
    for (int c = 0; c < 6; ++c) { t.sec[c] = ctx->stage[c]; t.ion[c] = ctx->istage[c]; }
    t.sec_cnt = ctx->stage_cnt;
    t.ion_cnt = ctx->istage_cnt;
    t.scan_tmp = ctx->scan_tmp;
    t.base_tmp = ctx->base_tmp;
    t.max_segs = (int32_t)(ctx->stage_slots / kSeg);
    return t;
}

static PartArgs make_part(gasx_ctx* ctx, const gasx_species* s)
{
    PartArgs pa{};
    pa.x = s->x; pa.y = s->y; pa.z = s->z; pa.vx = s->vx; pa.vy = s->vy; pa.vz = s->vz;
    pa.xi = s->x; pa.yi = s->y; pa.zi = s->z; pa.vxi = s->vx; pa.vyi = s->vy; pa.vzi = s->vz;
    pa.count = s->count;
    pa.cap = (int32_t)s->capacity;
    pa.q_over_m = (float)(s->qnum * ctx->p.e_em / s->mass);
    pa.kq = (float)(4294967296.0 / ctx->p.c) * s->qnum;
    pa.er = ctx->er32; pa.ez = ctx->ez32; pa.h = ctx->h32;
    for (int c = 0; c < 6; ++c) { pa.sec[c] = ctx->stage[c]; pa.ion[c] = ctx->istage[c]; }
    pa.sec_cnt = ctx->stage_cnt;
    pa.ion_cnt = ctx->istage_cnt;
    pa.counters = ctx->d_counters;
    return pa;
}

gasx_status gasx_host::make_coll(gasx_ctx* ctx, const gasx_gas* g, double dt_coll, uint64_t seed,
                                 uint32_t step, uint32_t run, CollArgs* ca)
{
    *ca = CollArgs{};
    if (!g) return GASX_OK;
    if (!g->density || g->gnr < 1 || g->gnz < 1 || !(g->gdr > 0) || !(g->gdz > 0) ||
        !(g->n_max >= 0) || !(dt_coll > 0))
        return fail(ctx, GASX_E_ARG, "invalid gas description");
    const gasx_params& p = ctx->p;
    ca->tab = ctx->d_tab;
    ca->dens = g->density;
    ca->temp = g->temperature;
    ca->gnr = g->gnr; ca->gnz = g->gnz;
    ca->inv_gdr = (float)(1.0 / g->gdr); ca->inv_gdz = (float)(1.0 / g->gdz);
    ca->t_scalar = (float)g->t_scalar;
    ca->dt_coll = (float)dt_coll;
    double pm = dt_coll * g->n_max * ctx->max_rate * 1.02 + 1e-7;
    ca->pmax = (float)std::min(1.0, pm);
    ca->key0 = (uint32_t)seed; ca->key1 = (uint32_t)(seed >> 32);
    ca->step = step; ca->run = run;
    ca->ev_per_mc2_em = (float)(p.me_em * p.c * p.c / p.e_em);
    {
        const double mc2_ev = p.me_em * p.c * p.c / p.e_em;
        auto v2_of = [&](double T) {   // push-side electron: v^2 = c^2 (1 - 1/gamma^2)
            const double gm = 1.0 + T / mc2_ev;
            return p.c * p.c * (1.0 - 1.0 / (gm * gm));
        };
        ca->v2_lo = (float)v2_of(15.6 * (1.0 - 1e-4));
        ca->v2_hi = (float)v2_of(20000.0 * (1.0 + 1e-4));
    }
    ca->mc2_mc_ev = (float)(p.me_mc * p.c * p.c / p.e_mc);
    ca->ion_vth_k = (float)std::sqrt(p.k_b / p.m_ion);
    ca->exc_form = p.exc_form;
    ca->rot_eps = (float)p.rot_eps;
    return GASX_OK;
}

int gasx_host::particle_grid(const gasx_ctx* ctx, int64_t segments, int per_cu)
{
    const int64_t persistent = std::min<int64_t>((int64_t)ctx->cu_count * per_cu, kCounterSlots);
    const int64_t need = (segments + kBlock / kWaveSize - 1) / (kBlock / kWaveSize);
    return (int)std::max<int64_t>(1, std::min(need, persistent));
}

// With collisions, one generation of resident workgroups: each wave then
// walks the most segments, so its candidate ring fills the most collision
// rounds (a round costs the same however few lanes it uses).  A push alone
// streams better from an oversubscribed grid.
static int particle_blocks(gasx_ctx* ctx, int64_t cap, uint32_t flags)
{
    return particle_grid(ctx, (cap + kSeg - 1) / kSeg, (flags & kCollide) ? particle_occupancy(flags) : 16);
}

// one field cycle: reads (in) writes (out) and refreshes the FP32 copies
static gasx_status fdtd_cycle(gasx_ctx* ctx, const double* er_in, const double* ez_in,
                              const double* h_in, double* er_out, double* ez_out, double* h_out,
                              const unsigned long long* qr, const unsigned long long* qz,
                              const double* jr, const double* jz, unsigned long long* qr_clear,
                              unsigned long long* qz_clear)
{
    FdtdRun r{};
    r.er_in = er_in; r.ez_in = ez_in; r.h_in = h_in;
    r.er_out = er_out; r.ez_out = ez_out; r.h_out = h_out;
    r.er32 = ctx->er32; r.ez32 = ctx->ez32; r.h32 = ctx->h32;
    r.er_mid = ctx->er_m; r.ez_mid = ctx->ez_m; r.h_mid = ctx->h_m;
    r.qr = qr; r.qz = qz; r.jr = jr; r.jz = jz;
    r.qr_clear = qr_clear; r.qz_clear = qz_clear;
    TimedScope ts(ctx, kTFdtd);
    HIPCHK(launch_fdtd(ctx->fa, r, nullptr, 1, qr != nullptr, ctx->stream), "fdtd");
    return GASX_OK;
}

// ---------------------------------------------------------------------------
// FDTD entry points
// ---------------------------------------------------------------------------
extern "C" gasx_status gasx_fdtd(gasx_ctx* ctx, gasx_fields* f, const double* jr, const double* jz,
                                 int n_push)
{
    if (!ctx || !valid_fields(f) || !jr || !jz || n_push < 0) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const gasx_params& p = ctx->p;
    const size_t nh = (size_t)p.nr * p.nz, ner = (size_t)p.nr * (p.nz + 1), nez = (size_t)(p.nr + 1) * p.nz;
    for (int s = 0; s < n_push; ++s) {
        const bool even = (s % 2) == 0;
        gasx_status st = even ? fdtd_cycle(ctx, f->er, f->ez, f->hphi, ctx->er_b, ctx->ez_b, ctx->h_b,
                                           nullptr, nullptr, jr, jz, nullptr, nullptr)
                              : fdtd_cycle(ctx, ctx->er_b, ctx->ez_b, ctx->h_b, f->er, f->ez, f->hphi,
                                           nullptr, nullptr, jr, jz, nullptr, nullptr);
        if (st != GASX_OK) return st;
    }
    if (n_push % 2 == 1) {
        HIPCHK(hipMemcpyAsync(f->er, ctx->er_b, ner * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy");
        HIPCHK(hipMemcpyAsync(f->ez, ctx->ez_b, nez * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy");
        HIPCHK(hipMemcpyAsync(f->hphi, ctx->h_b, nh * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy");
    }
    return GASX_OK;
}

extern "C" gasx_status gasx_fields_sync(gasx_ctx* ctx, const gasx_fields* f)
{
    if (!ctx || !valid_fields(f)) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const gasx_params& p = ctx->p;
    HIPCHK(launch_to_f32(f->er, ctx->er32, (int64_t)p.nr * (p.nz + 1), ctx->stream), "to_f32");
    HIPCHK(launch_to_f32(f->ez, ctx->ez32, (int64_t)(p.nr + 1) * p.nz, ctx->stream), "to_f32");
    HIPCHK(launch_to_f32(f->hphi, ctx->h32, (int64_t)p.nr * p.nz, ctx->stream), "to_f32");
    return GASX_OK;
}

// ---------------------------------------------------------------------------
// Particle entry points
// ---------------------------------------------------------------------------
// jstride > 0: qr, qz are the context's kJCopies-fold accumulators
static gasx_status run_particles(gasx_ctx* ctx, uint32_t flags, const gasx_species* s,
                                 unsigned long long* qr, unsigned long long* qz, const CollArgs& ca,
                                 size_t jstride = 0, int timed_kind = kTParticle, float* const* in = nullptr,
                                 const Staging* stg = nullptr)
{
    PartArgs pa = make_part(ctx, s);
    if (stg) {
        for (int c = 0; c < 6; ++c) { pa.sec[c] = stg->sec[c]; pa.ion[c] = stg->ion[c]; }
        pa.sec_cnt = stg->sec_cnt;
        pa.ion_cnt = stg->ion_cnt;
    }
    if (in) { pa.xi = in[0]; pa.yi = in[1]; pa.zi = in[2]; pa.vxi = in[3]; pa.vyi = in[4]; pa.vzi = in[5]; }
    pa.qr = qr; pa.qz = qz;
    pa.jstride = qr ? jstride : 0;
    const KernelEvents ke = kernel_events(ctx, timed_kind);
    HIPCHK(launch_particles(flags, ctx->geom, pa, ca, particle_blocks(ctx, s->capacity, flags), ctx->stream,
                            ke.start, ke.stop),
           "particles");
    return GASX_OK;
}

// EM2 half drift and / or EM7 deposit (k_drift_deposit); jstride > 0: qr, qz
// are the context's kJCopies-fold accumulators
static gasx_status run_drift_deposit(gasx_ctx* ctx, const gasx_species* s, bool drift, bool deposit,
                                     unsigned long long* qr, unsigned long long* qz, size_t jstride = 0,
                                     int timed_kind = kTParticle, float* const* in = nullptr)
{
    PartArgs pa = make_part(ctx, s);
    // in: read the particles from a sort's permuted copies (the drift writes all six values back)
    if (in) { pa.xi = in[0]; pa.yi = in[1]; pa.zi = in[2]; pa.vxi = in[3]; pa.vyi = in[4]; pa.vzi = in[5]; }
    pa.qr = qr; pa.qz = qz;
    pa.jstride = qr ? jstride : 0;
    const int64_t chunks = (s->capacity + kChunkD - 1) / kChunkD;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(chunks, std::min<int64_t>(
                                                                             (int64_t)ctx->cu_count * 16,
                                                                             kCounterSlots)));
    const KernelEvents ke = kernel_events(ctx, timed_kind);
    HIPCHK(launch_drift_deposit(drift, deposit, false, ctx->geom, pa, blocks, ctx->stream, ke.start, ke.stop),
           "drift/deposit");
    return GASX_OK;
}

extern "C" gasx_status gasx_half_drift(gasx_ctx* ctx, gasx_species* s)
{
    if (!ctx || !valid_species(s)) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    return run_drift_deposit(ctx, s, true, false, nullptr, nullptr);
}

extern "C" gasx_status gasx_deposit(gasx_ctx* ctx, const gasx_species* s, int64_t* qr, int64_t* qz)
{
    if (!ctx || !valid_species(s) || !qr || !qz) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    return run_drift_deposit(ctx, s, false, true, (unsigned long long*)qr, (unsigned long long*)qz);
}

extern "C" gasx_status gasx_current_from_fixed(gasx_ctx* ctx, const int64_t* qr, const int64_t* qz,
                                               double* jr, double* jz)
{
    if (!ctx || !qr || !qz || !jr || !jz) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(launch_j_from_fixed(ctx->fa, (const unsigned long long*)qr, (const unsigned long long*)qz,
                               jr, jz, ctx->stream),
           "j_from_fixed");
    return GASX_OK;
}

extern "C" gasx_status gasx_push(gasx_ctx* ctx, gasx_species* s, const gasx_fields* f)
{
    if (!ctx || !valid_species(s)) return GASX_E_ARG;
    (void)f;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    CollArgs ca{};
    return run_particles(ctx, kPush, s, nullptr, nullptr, ca);
}

extern "C" gasx_status gasx_cross_sections(gasx_ctx* ctx, const float* e_kev, int64_t n, float* se,
                                           float* si, float* sx)
{
    if (!ctx || n < 0 || (n > 0 && (!e_kev || !se || !si || !sx))) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(launch_cross_sections(ctx->d_tab, e_kev, n, se, si, sx, ctx->stream), "cross_sections");
    return GASX_OK;
}

// Appends what the last particle launch over `src` staged into dst (parent
// slot order).  Ions must be appended before the electrons' own secondaries:
// the electron count fixes the number of staged segments.
static gasx_status append_from_stage(gasx_ctx* ctx, const Staging& stg, const gasx_species* src,
                                     gasx_species* dst, bool ion)
{
    float* d[6] = {dst->x, dst->y, dst->z, dst->vx, dst->vy, dst->vz};
    TimedScope ts(ctx, kTAppend);
    HIPCHK(launch_append(d, dst->count, (int32_t)dst->capacity, ion ? stg.ion : stg.sec,
                         ion ? stg.ion_cnt : stg.sec_cnt, src->count, stg.scan_tmp,
                         stg.base_tmp + (ion ? 2 : 0),
                         reinterpret_cast<uint32_t*>(stg.base_tmp + 4 + (ion ? kAppendSyncWords : 0)), stg.max_segs,
                         &ctx->d_counters[ion ? kCDropIon : kCDropSec], ctx->stream),
           "append");
    return GASX_OK;
}

// One collision check of every electron of e, staging into stg, then the
// appends (ions first: they read the same per-segment ranks as the electrons).
static gasx_status collide_staged(gasx_ctx* ctx, const Staging& stg, gasx_species* e, gasx_species* ions,
                                  const CollArgs& ca)
{
    const uint32_t flags = kCollide | kStageSec | (ions ? kStageIon : 0u);
    gasx_status st = run_particles(ctx, flags, e, nullptr, nullptr, ca, 0, kTParticle, nullptr, &stg);
    if (st != GASX_OK) return st;
    if (ions) {
        st = append_from_stage(ctx, stg, e, ions, true);
        if (st != GASX_OK) return st;
    }
    return append_from_stage(ctx, stg, e, e, false);
}

extern "C" gasx_status gasx_collide(gasx_ctx* ctx, gasx_species* e, gasx_species* ions,
                                    const gasx_gas* g, double dt_coll, uint64_t seed, uint32_t step,
                                    uint32_t run)
{
    if (!ctx || !valid_species(e) || !g || (ions && !valid_species(ions))) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    gasx_status st = ensure_staging(ctx, e->capacity, ions != nullptr);
    if (st != GASX_OK) return st;
    CollArgs ca;
    st = make_coll(ctx, g, dt_coll, seed, step, run, &ca);
    if (st != GASX_OK) return st;
    const Staging stg = ctx_staging(ctx);
    return collide_staged(ctx, stg, e, ions, ca);
}

extern "C" gasx_status gasx_collide_noalloc(gasx_ctx* ctx, gasx_species* e, gasx_species* ions,
                                            const gasx_gas* g, double dt_coll, uint64_t seed,
                                            uint32_t step, uint32_t run)
{
    if (!ctx || !valid_species(e) || !g || (ions && !valid_species(ions))) return GASX_E_ARG;
    if (e->capacity > kNaSlots)
        return fail(ctx, GASX_E_ARG, "gasx_collide_noalloc: electron capacity " + std::to_string(e->capacity) +
                                         " exceeds GASX_NOALLOC_MAX_SLOTS (" + std::to_string(kNaSlots) + ")");
    if (!ctx->scratch) return fail(ctx, GASX_E_STATE, "gasx_collide_noalloc: no fixed staging on this device");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    CollArgs ca;
    gasx_status st = make_coll(ctx, g, dt_coll, seed, step, run, &ca);
    if (st != GASX_OK) return st;
    uint32_t* w = static_cast<uint32_t*>(ctx->scratch);   // this device's address of the array
    Staging stg{};
    for (int c = 0; c < 6; ++c) {
        stg.sec[c] = reinterpret_cast<float*>(w + (size_t)c * kNaSlots);
        stg.ion[c] = reinterpret_cast<float*>(w + (size_t)(6 + c) * kNaSlots);
    }
    stg.sec_cnt = reinterpret_cast<int32_t*>(w + kNaStageCnt);
    stg.ion_cnt = reinterpret_cast<int32_t*>(w + kNaIonCnt);
    stg.scan_tmp = reinterpret_cast<int32_t*>(w + kNaScan);
    stg.base_tmp = reinterpret_cast<int32_t*>(w + kNaBase);
    stg.max_segs = (int32_t)((e->capacity + kSeg - 1) / kSeg);   // <= kNaSegs
    return collide_staged(ctx, stg, e, ions, ca);
}

extern "C" gasx_status gasx_collision_bound(gasx_ctx* ctx, const gasx_gas* g, double dt_coll, double* pmax)
{
    if (!ctx || !g || !pmax) return GASX_E_ARG;
    CollArgs ca;
    gasx_status st = make_coll(ctx, g, dt_coll, 0, 0, 0, &ca);
    if (st != GASX_OK) return st;
    *pmax = ca.pmax;
    return GASX_OK;
}

// ---------------------------------------------------------------------------
// Fused loop shared by gasx_em_step (several pushed species, no collisions)
// and gasx_run (electrons with collisions, optional appended ions).
// ---------------------------------------------------------------------------
// Cell sort of one species; sorted_out non-null leaves the permuted particles
// in the sort scratch (their pointers returned) instead of copying them back.
static gasx_status sort_species(gasx_ctx* ctx, gasx_species* s, float** sorted_out);
static gasx_status sort_species_begin(gasx_ctx* ctx, gasx_species* s);
static gasx_status sort_species_finish(gasx_ctx* ctx, gasx_species* s, float** sorted_out);

struct LoopOpts {
    int n_steps = 0;
    bool collide = false, secondaries = false, ions = false;
    int sort_every = 0;
    uint64_t seed = 0;
    uint32_t run_id = 0, step0 = 0;
};

static gasx_status fused_loop(gasx_ctx* ctx, gasx_species* sp, int nsp, gasx_species* ions,
                              gasx_fields* f, const gasx_gas* g, const LoopOpts& o)
{
    const gasx_params& p = ctx->p;
    const size_t nh = (size_t)p.nr * p.nz, ner = (size_t)p.nr * (p.nz + 1), nez = (size_t)(p.nr + 1) * p.nz;
    if (o.n_steps <= 0) return GASX_OK;
    CollArgs ca{};
    if (o.collide) {
        gasx_status st = make_coll(ctx, g, p.ddt, o.seed, o.step0, o.run_id, &ca);
        if (st != GASX_OK) return st;
    }
    if (o.collide) {
        gasx_status st = ensure_staging(ctx, sp[0].capacity, o.ions);
        if (st != GASX_OK) return st;
    }
    // cell-sorted slots keep each chunk's current within one LDS tile; the
    // last species' permuted copies feed the prologue's drift, which writes
    // them back in place of the sort's own copy back
    float* sorted[6] = {};
    int sorted_k = -1;                 // species whose next particle launch reads the sorted copies
    const bool initial_sort = o.sort_every > 0;
    if (initial_sort) {
        for (int k = 0; k < nsp; ++k) {
            gasx_status st = sort_species(ctx, &sp[k], k == nsp - 1 ? sorted : nullptr);
            if (st != GASX_OK) return st;
        }
    }
    // both current buffers, every copy (the field kernel clears copy 0 of the
    // next buffer, the copy reduction zeroes the others after reading them)
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 2; ++k)
            HIPCHK(launch_fill_u64(ctx->q[b][k], (int64_t)(nh * kJCopies), ctx->stream), "zero q");
    // prologue: EM2 half drift + EM7 deposit of every species (electricmagneticfield6.m:141-144)
    for (int k = 0; k < nsp; ++k) {
        gasx_status st = run_drift_deposit(ctx, &sp[k], true, true, ctx->q[0][0], ctx->q[0][1], nh, kTDrift,
                                           initial_sort && k == nsp - 1 ? sorted : nullptr);
        if (st != GASX_OK) return st;
    }
    // the field cycle of step s: E, H from t^s to t^(s+1) with the current deposited before it
    auto field_cycle = [&](int s) -> gasx_status {
        const int cur = s & 1, nxt = cur ^ 1;
        const bool last = s == o.n_steps - 1;
        if (kJCopies > 1) {   // fold the current copies into copy 0
            TimedScope ts(ctx, kTFdtd);
            HIPCHK(launch_j_reduce(ctx->q[cur][0], ctx->q[cur][1], (int64_t)nh, nh, ctx->stream), "j reduce");
        }
        if ((s % 2) == 0)
            return fdtd_cycle(ctx, f->er, f->ez, f->hphi, ctx->er_b, ctx->ez_b, ctx->h_b,
                              ctx->q[cur][0], ctx->q[cur][1], nullptr, nullptr,
                              last ? nullptr : ctx->q[nxt][0], last ? nullptr : ctx->q[nxt][1]);
        return fdtd_cycle(ctx, ctx->er_b, ctx->ez_b, ctx->h_b, f->er, f->ez, f->hphi,
                          ctx->q[cur][0], ctx->q[cur][1], nullptr, nullptr,
                          last ? nullptr : ctx->q[nxt][0], last ? nullptr : ctx->q[nxt][1]);
    };
    bool cycle_ahead = false;          // this step's field cycle was enqueued before a sort
    for (int s = 0; s < o.n_steps; ++s) {
        const int cur = s & 1, nxt = cur ^ 1;
        const bool last = s == o.n_steps - 1;
        gasx_status st;
        if (!cycle_ahead) {
            st = field_cycle(s);
            if (st != GASX_OK) return st;
        }
        cycle_ahead = false;
        ca.step = o.step0 + (uint32_t)s;
        const bool sort_now = o.sort_every > 0 && !last && ((s + 1) % o.sort_every) == 0;
        for (int k = 0; k < nsp; ++k) {
            // two launches: push (+ collide, staged secondaries already
            // drifted), then drift + deposit of the existing slots; the
            // collision phase's registers would otherwise limit the occupancy
            // of the streaming half drift and deposit
            // the particle launch also makes the step's second half drift
            // (except after the last push), so the deposit that follows only
            // reads the particles
            uint32_t flags = kPush | (last ? 0u : kDrift2);
            if (k == 0 && o.collide) {
                flags |= kCollide;
                if (o.secondaries) flags |= kStageSec | (last ? 0u : kDriftSec);
                if (o.ions) flags |= kStageIon;
            }
            st = run_particles(ctx, flags, &sp[k], nullptr, nullptr, ca, 0, kTParticle,
                               k == sorted_k ? sorted : nullptr);
            if (st != GASX_OK) return st;
            if (k == sorted_k) sorted_k = -1;
            // appends first (ions, then the electrons' secondaries, already
            // drifted), so the deposit that follows covers the appended slots too
            if (k == 0 && o.collide) {
                const Staging stg = ctx_staging(ctx);
                if (o.ions) {
                    st = append_from_stage(ctx, stg, &sp[0], ions, true);
                    if (st != GASX_OK) return st;
                }
                if (o.secondaries) {
                    st = append_from_stage(ctx, stg, &sp[0], &sp[0], false);
                    if (st != GASX_OK) return st;
                }
            }
            // one species: the sort's key pass as soon as the positions are
            // final, before the deposit, which does not move them
            if (sort_now && nsp == 1) {
                st = sort_species_begin(ctx, &sp[0]);
                if (st != GASX_OK) return st;
            }
            if (!last) {
                st = run_drift_deposit(ctx, &sp[k], false, true, ctx->q[nxt][0], ctx->q[nxt][1], nh, kTDrift);
                if (st != GASX_OK) return st;
            }
        }
        if (sort_now) {
            if (nsp == 1) {
                // the next step's field cycle does not depend on the particles'
                // order: enqueued ahead, it keeps the GPU busy while the host
                // waits for the slot count
                st = field_cycle(s + 1);
                if (st != GASX_OK) return st;
                cycle_ahead = true;
                st = sort_species_finish(ctx, &sp[0], sorted);
                if (st != GASX_OK) return st;
            } else {
                for (int k = 0; k < nsp; ++k) {
                    st = sort_species(ctx, &sp[k], k == nsp - 1 ? sorted : nullptr);
                    if (st != GASX_OK) return st;
                }
            }
            sorted_k = nsp - 1;
        }
    }
    if (o.n_steps % 2 == 1) {
        HIPCHK(hipMemcpyAsync(f->er, ctx->er_b, ner * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy");
        HIPCHK(hipMemcpyAsync(f->ez, ctx->ez_b, nez * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy");
        HIPCHK(hipMemcpyAsync(f->hphi, ctx->h_b, nh * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy");
    }
    return GASX_OK;
}

extern "C" gasx_status gasx_em_step(gasx_ctx* ctx, gasx_species* sp, int n_species, gasx_fields* f,
                                    int n_push)
{
    if (!ctx || !sp || n_species < 1 || n_species > 8 || !valid_fields(f) || n_push < 0)
        return GASX_E_ARG;
    for (int k = 0; k < n_species; ++k)
        if (!valid_species(&sp[k])) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    LoopOpts o;
    o.n_steps = n_push;
    return fused_loop(ctx, sp, n_species, nullptr, f, nullptr, o);
}

extern "C" gasx_status gasx_run(gasx_ctx* ctx, gasx_species* e, gasx_species* ions, gasx_fields* f,
                                const gasx_gas* g, const gasx_run_opts* ro)
{
    if (!ctx || !valid_species(e) || !valid_fields(f) || !ro || ro->n_steps < 0 ||
        (ions && !valid_species(ions)) || (ro->collide && !g) || ro->sort_every < 0)
        return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    LoopOpts o;
    o.n_steps = ro->n_steps;
    o.collide = ro->collide != 0;
    o.secondaries = o.collide && ro->secondaries != 0;
    o.ions = o.collide && ions != nullptr;
    o.sort_every = ro->sort_every;
    o.seed = ro->seed;
    o.run_id = ro->run_id;
    o.step0 = ro->step0;
    if (ro->time_kernels) start_timing(ctx, ro->time_kernels, ro->time_every);
    gasx_status st = fused_loop(ctx, e, 1, ions, f, g, o);
    ctx->timing = false;
    return st;
}

extern "C" gasx_status gasx_sort(gasx_ctx* ctx, gasx_species* s)
{
    if (!ctx || !valid_species(s)) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    return sort_species(ctx, s, nullptr);
}

// The key pass of a cell sort (sort_begin); the scratch is sized first.
static gasx_status sort_species_begin(gasx_ctx* ctx, gasx_species* s)
{
    size_t need = 0;
    hipError_t e = sort_begin(ctx->geom, s, nullptr, 0, &need, ctx->stream, ctx->h_count, ctx->sort_ev);
    if (e != hipSuccess && e != hipErrorInvalidValue) return hip_fail(ctx, e, "sort size");
    if (need > ctx->sort_bytes) {
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        if (ctx->sort_tmp) hipFree(ctx->sort_tmp);
        ctx->sort_tmp = nullptr;
        HIPCHK(hipMalloc(&ctx->sort_tmp, need), "alloc sort");
        ctx->sort_bytes = need;
    }
    HIPCHK(sort_begin(ctx->geom, s, ctx->sort_tmp, ctx->sort_bytes, nullptr, ctx->stream, ctx->h_count,
                      ctx->sort_ev),
           "sort keys");
    return GASX_OK;
}

// The rest of the sort (sort_finish).  sorted_out non-null: the permuted
// particles stay in the sort scratch (for the next particle launch to read)
// instead of being copied back.
static gasx_status sort_species_finish(gasx_ctx* ctx, gasx_species* s, float** sorted_out)
{
    TimedScope ts(ctx, kTSort);
    HIPCHK(sort_finish(ctx->geom, s, ctx->sort_tmp, ctx->stream, sorted_out, ctx->h_count, ctx->sort_ev), "sort");
    return GASX_OK;
}

static gasx_status sort_species(gasx_ctx* ctx, gasx_species* s, float** sorted_out)
{
    gasx_status st = sort_species_begin(ctx, s);
    if (st != GASX_OK) return st;
    return sort_species_finish(ctx, s, sorted_out);
}

extern "C" gasx_status gasx_get_timing(gasx_ctx* ctx, gasx_timing* out)
{
    if (!ctx || !out) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    double ms[kNTimed] = {};
    for (int k = 0; k < kNTimed; ++k)
        for (auto& pr : ctx->tpairs[k]) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, ctx->events[pr.first], ctx->events[pr.second]), "event time");
            ms[k] += t;
        }
    out->particle_ms = ms[kTParticle];
    out->fdtd_ms = ms[kTFdtd];
    out->append_ms = ms[kTAppend];
    out->sort_ms = ms[kTSort];
    out->n_particle_launches = (int64_t)ctx->tpairs[kTParticle].size();
    out->n_fdtd_launches = (int64_t)ctx->tpairs[kTFdtd].size();
    out->n_append_launches = (int64_t)ctx->tpairs[kTAppend].size();
    out->n_sorts = (int64_t)ctx->tpairs[kTSort].size();
    out->graph_ms = ms[kTGraph];
    out->n_graph_launches = (int64_t)ctx->tpairs[kTGraph].size();
    out->drift_ms = ms[kTDrift];
    out->n_drift_launches = (int64_t)ctx->tpairs[kTDrift].size();
    return GASX_OK;
}

extern "C" gasx_status gasx_tallies(gasx_ctx* ctx, const gasx_fields* f, const gasx_species* sp,
                                    int n_species, double out[4])
{
    if (!ctx || !valid_fields(f) || !out || n_species < 0 || (n_species > 0 && !sp)) return GASX_E_ARG;
    for (int k = 0; k < n_species; ++k)
        if (!valid_species(&sp[k])) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(launch_tallies(ctx->p, f->er, f->ez, f->hphi, sp, n_species, ctx->d_tally,
                          ctx->d_tally + 4 * 256, ctx->stream),
           "tallies");
    HIPCHK(hipMemcpyAsync(out, ctx->d_tally + 4 * 256, 4 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream),
           "copy tallies");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return GASX_OK;
}

extern "C" gasx_status gasx_get_counters(gasx_ctx* ctx, gasx_counters* out, int reset)
{
    if (!ctx || !out) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    std::vector<unsigned long long> slots((size_t)kNCounters * kCounterSlots);
    HIPCHK(hipMemcpyAsync(slots.data(), ctx->d_counters, slots.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, ctx->stream),
           "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    unsigned long long h[kNCounters] = {};
    for (size_t b = 0; b < (size_t)kCounterSlots; ++b)
        for (int k = 0; k < kNCounters; ++k) h[k] += slots[b * kNCounters + k];
    out->n_elastic = (int64_t)h[kCElastic];
    out->n_ionize = (int64_t)h[kCIonize];
    out->n_excite = (int64_t)h[kCExcite];
    out->n_dropped_sec = (int64_t)h[kCDropSec];
    out->n_dropped_ion = (int64_t)h[kCDropIon];
    out->n_bound_violations = (int64_t)h[kCViol];
    out->n_absorbed = (int64_t)h[kCAbsorbed];
    out->n_saturated = (int64_t)h[kCSat];
    out->n_pushed = (int64_t)h[kCPushed];
    if (reset) {
        HIPCHK(hipMemsetAsync(ctx->d_counters, 0, slots.size() * sizeof(unsigned long long), ctx->stream),
               "zero");
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    }
    if (out->n_dropped_sec > 0 || out->n_dropped_ion > 0)
        return fail(ctx, GASX_E_CAPACITY, "appends exceeded capacity");
    return GASX_OK;
}

extern "C" gasx_status gasx_get_count(gasx_ctx* ctx, const gasx_species* s, int64_t* n)
{
    if (!ctx || !s || !s->count || !n) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    int32_t v = 0;
    HIPCHK(hipMemcpyAsync(&v, s->count, sizeof(v), hipMemcpyDeviceToHost, ctx->stream), "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    *n = v;
    return GASX_OK;
}
