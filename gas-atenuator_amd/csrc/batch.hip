// Batched ensemble: many independent runs of the production loop advanced
// together.  Per step one field launch covers every run's tiles and one
// persistent particle launch covers every run's chunks, so small runs share
// the GPU instead of each under-filling it; the whole step sequence can be
// captured in one hipGraph and replayed.  Each run computes exactly what
// gasx_run would for it alone (same Philox counters, same fixed-point
// current, same field arithmetic, same append order).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "context.h"

using namespace gasx;
using namespace gasx_host;

extern "C" int gasx_batch_align(void) { return kChunk; }

namespace {

template <typename T>
void dfree(T*& p)
{
    if (p) hipFree((void*)p);
    p = nullptr;
}

void free_run_buffers(BatchState& b)
{
    dfree(b.er_b); dfree(b.ez_b); dfree(b.h_b);
    dfree(b.er_m); dfree(b.ez_m); dfree(b.h_m);
    dfree(b.er32); dfree(b.ez32); dfree(b.h32);
    for (auto& qq : b.q) { dfree(qq[0]); dfree(qq[1]); }
    dfree(b.runs); dfree(b.ion_runs); dfree(b.fdtd_runs); dfree(b.chunk_start); dfree(b.tmp_run);
    dfree(b.run_counters);
    b.nruns_cap = 0;
    b.uploaded_key.clear();
}

void free_stage_buffers(BatchState& b)
{
    for (auto& s : b.stage) dfree(s);
    dfree(b.offs); dfree(b.stage_cnt);
    b.stage_segs = 0;
    for (auto& s : b.istage) dfree(s);
    dfree(b.istage_cnt);
    b.istage_segs = 0;
}

void drop_graph(BatchState& b)
{
    if (b.exec) hipGraphExecDestroy(b.exec);
    b.exec = nullptr;
    b.graph_key.clear();
}

struct Sizes {
    size_t nh, ner, nez;
};

Sizes sizes(const gasx_ctx* ctx)
{
    const size_t nr = ctx->p.nr, nz = ctx->p.nz;
    return Sizes{nr * nz, nr * (nz + 1), (nr + 1) * nz};
}

gasx_status ensure_batch(gasx_ctx* ctx, int nruns, int64_t stage_segs, bool ions)
{
    BatchState& b = ctx->batch;
    const Sizes z = sizes(ctx);
    if (nruns > b.nruns_cap) {
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        drop_graph(b);
        free_run_buffers(b);
        const size_t n = (size_t)nruns;
        HIPCHK(dalloc(&b.er_b, n * z.ner), "alloc batch er");
        HIPCHK(dalloc(&b.ez_b, n * z.nez), "alloc batch ez");
        HIPCHK(dalloc(&b.h_b, n * z.nh), "alloc batch h");
        HIPCHK(dalloc(&b.er_m, n * z.ner), "alloc batch er mid");
        HIPCHK(dalloc(&b.ez_m, n * z.nez), "alloc batch ez mid");
        HIPCHK(dalloc(&b.h_m, n * z.nh), "alloc batch h mid");
        HIPCHK(dalloc(&b.er32, n * z.ner), "alloc batch er32");
        HIPCHK(dalloc(&b.ez32, n * z.nez), "alloc batch ez32");
        HIPCHK(dalloc(&b.h32, n * z.nh), "alloc batch h32");
        for (auto& qq : b.q)
            for (auto& q : qq) HIPCHK(dalloc(&q, n * z.nh * kJCopies), "alloc batch current");
        HIPCHK(dalloc(&b.runs, n), "alloc batch runs");
        HIPCHK(dalloc(&b.ion_runs, n), "alloc batch ion runs");
        HIPCHK(dalloc(&b.fdtd_runs, 4 * n), "alloc batch field runs");
        HIPCHK(dalloc(&b.chunk_start, n + 1), "alloc batch chunks");
        HIPCHK(dalloc(&b.tmp_run, 2 * n), "alloc batch tmp");
        HIPCHK(dalloc(&b.run_counters, n * kNCounters), "alloc batch counters");
        HIPCHK(hipMemset(b.run_counters, 0, n * kNCounters * sizeof(unsigned long long)), "zero counters");
        b.nruns_cap = nruns;
    }
    if (stage_segs > b.stage_segs) {
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        drop_graph(b);
        free_stage_buffers(b);
        for (auto& s : b.stage) HIPCHK(dalloc(&s, (size_t)stage_segs * kSeg), "alloc batch stage");
        HIPCHK(dalloc(&b.offs, (size_t)stage_segs), "alloc batch offsets");
        HIPCHK(dalloc(&b.stage_cnt, (size_t)stage_segs), "alloc batch stage counts");
        b.stage_segs = stage_segs;
    }
    if (ions && stage_segs > b.istage_segs) {
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        drop_graph(b);
        for (auto& s : b.istage) dfree(s);
        dfree(b.istage_cnt);
        for (auto& s : b.istage) HIPCHK(dalloc(&s, (size_t)stage_segs * kSeg), "alloc batch ion stage");
        HIPCHK(dalloc(&b.istage_cnt, (size_t)stage_segs), "alloc batch ion stage counts");
        b.istage_segs = stage_segs;
    }
    return GASX_OK;
}

bool has_ions(const gasx_batch* b, const gasx_run_opts* o) { return o->collide && b->ix != nullptr; }

// Ranges [base[r], base[r] + capacity[r]) of nruns runs: each valid and none
// overlapping; *slots = the end of the last range.
bool ranges_ok(const int64_t* base, const int64_t* cap, int nruns, int64_t* slots)
{
    std::vector<int> idx(nruns);
    std::iota(idx.begin(), idx.end(), 0);
    for (int r = 0; r < nruns; ++r)
        if (base[r] < 0 || base[r] % kChunk != 0 || cap[r] < 1 || cap[r] > (int64_t)INT32_MAX - kChunk)
            return false;
    std::sort(idx.begin(), idx.end(), [&](int a, int c) { return base[a] < base[c]; });
    for (int k = 1; k < nruns; ++k)
        if (base[idx[k - 1]] + cap[idx[k - 1]] > base[idx[k]]) return false;
    *slots = base[idx.back()] + cap[idx.back()];
    return true;
}

gasx_status validate(gasx_ctx* ctx, const gasx_batch* b, const gasx_run_opts* o, int64_t* stage_segs,
                     int64_t* pool_slots, int64_t* ion_slots)
{
    if (!b || !o || b->nruns < 1 || b->nruns > kMaxRuns || !b->x || !b->y || !b->z || !b->vx || !b->vy ||
        !b->vz || !b->base || !b->capacity || !b->counts || !b->er || !b->ez || !b->hphi)
        return fail(ctx, GASX_E_ARG, "invalid batch description");
    if (o->n_steps < 0 || o->sort_every != 0)
        return fail(ctx, GASX_E_ARG, "batched runs need n_steps >= 0 and sort_every = 0");
    if (o->collide && (!b->density || !b->n_max || b->gnr < 1 || b->gnz < 1 || !(b->gdr > 0) || !(b->gdz > 0)))
        return fail(ctx, GASX_E_ARG, "batched collisions need a gas description");
    int64_t chunks = 0;
    for (int r = 0; r < b->nruns; ++r) {
        if (o->collide && !(b->n_max[r] >= 0))
            return fail(ctx, GASX_E_ARG, "run " + std::to_string(r) + ": bad n_max");
        if (b->capacity[r] >= 1) chunks += (b->capacity[r] + kSeg - 1) / kSeg;
    }
    if (!ranges_ok(b->base, b->capacity, b->nruns, pool_slots))
        return fail(ctx, GASX_E_ARG, "batch runs: a bad base or capacity, or ranges that overlap in the pool");
    *ion_slots = 0;
    if (b->ix) {
        if (!b->iy || !b->iz || !b->ivx || !b->ivy || !b->ivz || !b->ion_base || !b->ion_capacity ||
            !b->ion_counts)
            return fail(ctx, GASX_E_ARG, "invalid batch ion pool");
        if (!ranges_ok(b->ion_base, b->ion_capacity, b->nruns, ion_slots))
            return fail(ctx, GASX_E_ARG, "batch ions: a bad base or capacity, or ranges that overlap in the pool");
    }
    *stage_segs = chunks;
    return GASX_OK;
}

FdtdArgs tuned_fdtd(const gasx_ctx* ctx, const gasx_batch_tuning& t)
{
    FdtdArgs a = ctx->fa;
    a.ti = std::max(1, std::min(t.fdtd_ti, ctx->p.nr));
    a.tk = std::max(1, std::min(t.fdtd_tk, ctx->p.nz));
    a.ntr = (ctx->p.nr + a.ti - 1) / a.ti;
    a.ntk = (ctx->p.nz + a.tk - 1) / a.tk;
    return a;
}

bool tile_fits(const gasx_ctx* ctx, int ti, int tk)
{
    FdtdArgs a = ctx->fa;
    a.ti = std::min(ti, ctx->p.nr);
    a.tk = std::min(tk, ctx->p.nz);
    int k = 0, bs = 0;
    return fdtd_fwd_lds_bytes(a) <= 160 * 1024 && fdtd_fwd_shape(a, &k, &bs);
}

// Everything upload() writes into the device descriptors depends on.
std::vector<unsigned char> upload_key(const gasx_batch* b, const gasx_run_opts* o)
{
    std::vector<unsigned char> k;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        k.insert(k.end(), c, c + n);
    };
    put(&b->nruns, sizeof(b->nruns));
    put(&b->er, sizeof(b->er));
    put(&b->ez, sizeof(b->ez));
    put(&b->hphi, sizeof(b->hphi));
    put(b->base, b->nruns * sizeof(int64_t));
    put(b->capacity, b->nruns * sizeof(int64_t));
    put(&o->collide, sizeof(o->collide));
    if (o->collide) put(b->n_max, b->nruns * sizeof(double));
    put(&o->run_id, sizeof(o->run_id));
    const bool ions = b->ix != nullptr;
    put(&ions, sizeof(ions));
    if (ions) {
        put(b->ion_base, b->nruns * sizeof(int64_t));
        put(b->ion_capacity, b->nruns * sizeof(int64_t));
    }
    return k;
}

// Uploads the run descriptors and the field-kernel run tables (outside any
// capture: they come from pageable host memory).
gasx_status upload(gasx_ctx* ctx, const gasx_batch* b, const gasx_run_opts* o)
{
    BatchState& st = ctx->batch;
    const Sizes z = sizes(ctx);
    const int n = b->nruns;
    std::vector<BatchRun> runs(n), iruns;
    for (int r = 0; r < n; ++r) {
        BatchRun& br = runs[r];
        br.base = b->base[r];
        br.cap = (int32_t)b->capacity[r];
        br.pmax = 1.0f;
        if (o->collide) {
            const double pm = ctx->p.ddt * b->n_max[r] * ctx->max_rate * 1.02 + 1e-7;
            br.pmax = (float)std::min(1.0, pm);
        }
        br.run = o->run_id + (uint32_t)r;
        br.pad_ = 0;
    }
    if (b->ix) {   // the same runs over their ranges of the ion pool
        iruns = runs;
        for (int r = 0; r < n; ++r) {
            iruns[r].base = b->ion_base[r];
            iruns[r].cap = (int32_t)b->ion_capacity[r];
        }
    }
    std::vector<FdtdRun> fr(4 * (size_t)n);
    for (int parity = 0; parity < 2; ++parity)
        for (int last = 0; last < 2; ++last)
            for (int r = 0; r < n; ++r) {
                FdtdRun& R = fr[(size_t)(parity * 2 + last) * n + r];
                R = FdtdRun{};
                double* uer = b->er + (size_t)r * z.ner;
                double* uez = b->ez + (size_t)r * z.nez;
                double* uh = b->hphi + (size_t)r * z.nh;
                double* per = st.er_b + (size_t)r * z.ner;
                double* pez = st.ez_b + (size_t)r * z.nez;
                double* ph = st.h_b + (size_t)r * z.nh;
                if (parity == 0) {
                    R.er_in = uer; R.ez_in = uez; R.h_in = uh;
                    R.er_out = per; R.ez_out = pez; R.h_out = ph;
                } else {
                    R.er_in = per; R.ez_in = pez; R.h_in = ph;
                    R.er_out = uer; R.ez_out = uez; R.h_out = uh;
                }
                R.er32 = st.er32 + (size_t)r * z.ner;
                R.ez32 = st.ez32 + (size_t)r * z.nez;
                R.h32 = st.h32 + (size_t)r * z.nh;
                R.er_mid = st.er_m + (size_t)r * z.ner;
                R.ez_mid = st.ez_m + (size_t)r * z.nez;
                R.h_mid = st.h_m + (size_t)r * z.nh;
                const int cur = parity, nxt = parity ^ 1;   // step s uses q[s & 1]
                R.qr = st.q[cur][0] + (size_t)r * z.nh;
                R.qz = st.q[cur][1] + (size_t)r * z.nh;
                if (!last) {
                    R.qr_clear = st.q[nxt][0] + (size_t)r * z.nh;
                    R.qz_clear = st.q[nxt][1] + (size_t)r * z.nh;
                }
            }
    HIPCHK(hipMemcpyAsync(st.runs, runs.data(), n * sizeof(BatchRun), hipMemcpyHostToDevice, ctx->stream),
           "upload runs");
    if (b->ix)
        HIPCHK(hipMemcpyAsync(st.ion_runs, iruns.data(), n * sizeof(BatchRun), hipMemcpyHostToDevice, ctx->stream),
               "upload ion runs");
    HIPCHK(hipMemcpyAsync(st.fdtd_runs, fr.data(), fr.size() * sizeof(FdtdRun), hipMemcpyHostToDevice,
                          ctx->stream),
           "upload field runs");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    st.uploaded_key = upload_key(b, o);
    return GASX_OK;
}

// Enqueues the whole step sequence (capturable: kernels and device copies only).
gasx_status enqueue_steps(gasx_ctx* ctx, const gasx_batch* b, const gasx_run_opts* o,
                          const gasx_batch_tuning& tune, int64_t stage_segs)
{
    BatchState& st = ctx->batch;
    const Sizes z = sizes(ctx);
    const int n = b->nruns;
    const FdtdArgs fa = tuned_fdtd(ctx, tune);
    CollArgs ca{};
    if (o->collide) {
        gasx_gas g{};
        g.density = b->density;
        g.temperature = nullptr;
        g.gnr = b->gnr; g.gnz = b->gnz; g.gdr = b->gdr; g.gdz = b->gdz;
        g.t_scalar = b->t_scalar;
        g.n_max = 0.0;                                     // per-run bounds come from the descriptors
        gasx_status s = make_coll(ctx, &g, ctx->p.ddt, o->seed, o->step0, o->run_id, &ca);
        if (s != GASX_OK) return s;
    }
    PartArgs pa{};
    pa.x = b->x; pa.y = b->y; pa.z = b->z; pa.vx = b->vx; pa.vy = b->vy; pa.vz = b->vz;
    pa.xi = b->x; pa.yi = b->y; pa.zi = b->z; pa.vxi = b->vx; pa.vyi = b->vy; pa.vzi = b->vz;
    pa.count = b->counts;
    pa.q_over_m = (float)(-1.0 * ctx->p.e_em / ctx->p.me_em);
    pa.kq = -(float)(4294967296.0 / ctx->p.c);
    pa.er = st.er32; pa.ez = st.ez32; pa.h = st.h32;
    for (int c = 0; c < 6; ++c) pa.sec[c] = st.stage[c];
    pa.sec_cnt = st.stage_cnt;
    pa.counters = ctx->d_counters;
    pa.runs = st.runs;
    pa.nruns = n;
    pa.chunk_start = st.chunk_start;
    pa.run_counters = st.run_counters;
    // current layout: [copy][run][cell], copy stride = the allocated runs
    const size_t jstride = (size_t)st.nruns_cap * z.nh;
    pa.jstride = jstride;
    const int grid = particle_grid(ctx, stage_segs, tune.particle_blocks_per_cu);
    float* pool[6] = {b->x, b->y, b->z, b->vx, b->vy, b->vz};
    float* ipool[6] = {b->ix, b->iy, b->iz, b->ivx, b->ivy, b->ivz};
    const bool sec = o->collide && o->secondaries;
    const bool ions = has_ions(b, o);
    if (ions) {
        for (int c = 0; c < 6; ++c) pa.ion[c] = st.istage[c];
        pa.ion_cnt = st.istage_cnt;
    }

    for (int bb = 0; bb < 2; ++bb)
        for (int k = 0; k < 2; ++k)
            HIPCHK(launch_fill_u64(st.q[bb][k], (int64_t)(jstride * kJCopies), ctx->stream), "zero q");
    {   // prologue: EM2 half drift + EM7 deposit of every run
        PartArgs p0 = pa;
        p0.qr = st.q[0][0]; p0.qz = st.q[0][1];
        TimedScope ts(ctx, kTDrift);
        HIPCHK(launch_drift_deposit(true, true, true, ctx->geom, p0, grid, ctx->stream), "drift + deposit");
    }
    for (int s = 0; s < o->n_steps; ++s) {
        const int cur = s & 1, nxt = cur ^ 1;
        const bool last = s == o->n_steps - 1;
        {
            TimedScope ts(ctx, kTFdtd);
            HIPCHK(launch_j_reduce(st.q[cur][0], st.q[cur][1], (int64_t)(n * z.nh), jstride, ctx->stream),
                   "j reduce");
            HIPCHK(launch_fdtd(fa, FdtdRun{}, st.fdtd_runs + (size_t)(cur * 2 + (last ? 1 : 0)) * n, n, 1,
                               ctx->stream),
                   "fdtd");
        }
        // push + half drift (+ collide, secondaries staged already drifted) +
        // the second half drift, then the deposit of the existing slots, as
        // the single-run loop does
        uint32_t flags = kBatch | kPush | (last ? 0u : kDrift2);
        if (o->collide) flags |= kCollide | (sec ? (kStageSec | (last ? 0u : kDriftSec)) : 0u);
        if (ions) flags |= kStageIon;
        PartArgs ps = pa;
        ps.qr = last ? nullptr : st.q[nxt][0];
        ps.qz = last ? nullptr : st.q[nxt][1];
        CollArgs cs = ca;
        cs.step = o->step0 + (uint32_t)s;
        {
            TimedScope ts(ctx, kTParticle);
            PartArgs pp = ps;
            pp.qr = pp.qz = nullptr;
            HIPCHK(launch_particles(flags, ctx->geom, pp, cs, grid, ctx->stream), "particles");
        }
        if (sec || ions) {   // appended first (ions, then secondaries): the deposit then covers the new slots
            TimedScope ts(ctx, kTAppend);
            if (ions)
                HIPCHK(launch_append_batch(ipool, b->ion_counts, st.ion_runs, n, st.istage, st.istage_cnt,
                                           st.chunk_start, st.offs, st.tmp_run, (int32_t)stage_segs,
                                           st.run_counters, &ctx->d_counters[kCDropIon], kCDropIon,
                                           ctx->stream),
                       "append ions");
            if (sec)
                HIPCHK(launch_append_batch(pool, b->counts, st.runs, n, st.stage, st.stage_cnt, st.chunk_start,
                                           st.offs, st.tmp_run, (int32_t)stage_segs, st.run_counters,
                                           &ctx->d_counters[kCDropSec], kCDropSec, ctx->stream),
                       "append");
        }
        if (!last) {
            TimedScope ts(ctx, kTDrift);
            HIPCHK(launch_drift_deposit(false, true, true, ctx->geom, ps, grid, ctx->stream), "deposit");
        }
    }
    if (o->n_steps % 2 == 1) {
        HIPCHK(hipMemcpyAsync(b->er, st.er_b, n * z.ner * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
               "copy");
        HIPCHK(hipMemcpyAsync(b->ez, st.ez_b, n * z.nez * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
               "copy");
        HIPCHK(hipMemcpyAsync(b->hphi, st.h_b, n * z.nh * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
               "copy");
    }
    return GASX_OK;
}

std::vector<unsigned char> graph_key(const gasx_batch* b, const gasx_run_opts* o, const gasx_batch_tuning& t)
{
    std::vector<unsigned char> k;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        k.insert(k.end(), c, c + n);
    };
    put(b, sizeof(*b));
    put(b->base, b->nruns * sizeof(int64_t));
    put(b->capacity, b->nruns * sizeof(int64_t));
    if (o->collide) put(b->n_max, b->nruns * sizeof(double));
    if (b->ix) {
        put(b->ion_base, b->nruns * sizeof(int64_t));
        put(b->ion_capacity, b->nruns * sizeof(int64_t));
    }
    gasx_run_opts oo = *o;
    oo.time_kernels = 0;
    oo.time_every = 0;
    put(&oo, sizeof(oo));
    put(&t, sizeof(t));
    return k;
}

gasx_status run_batch(gasx_ctx* ctx, gasx_batch* b, const gasx_run_opts* o, const gasx_batch_tuning& tune)
{
    int64_t stage_segs = 0, pool = 0, ipool = 0;
    gasx_status s = validate(ctx, b, o, &stage_segs, &pool, &ipool);
    if (s != GASX_OK) return s;
    if (o->n_steps == 0) return GASX_OK;
    s = ensure_batch(ctx, b->nruns, stage_segs, has_ions(b, o));
    if (s != GASX_OK) return s;
    BatchState& st = ctx->batch;
    if (!o->use_graph) {
        s = upload(ctx, b, o);
        if (s != GASX_OK) return s;
        return enqueue_steps(ctx, b, o, tune, stage_segs);
    }
    std::vector<unsigned char> key = graph_key(b, o, tune);
    if (!st.exec || key != st.graph_key) {
        drop_graph(st);
        s = upload(ctx, b, o);
        if (s != GASX_OK) return s;
        // Capture on a private stream (the caller's may be the legacy default
        // stream, which cannot be captured); the graph is launched on the
        // caller's stream, so ordering with its other work is kept.
        if (!st.capture_stream)
            HIPCHK(hipStreamCreateWithFlags(&st.capture_stream, hipStreamNonBlocking), "capture stream");
        const bool was_timing = ctx->timing;
        hipStream_t user = ctx->stream;
        ctx->timing = false;                         // no event records inside the graph
        ctx->stream = st.capture_stream;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        hipGraph_t graph = nullptr;
        if (e == hipSuccess) {
            s = enqueue_steps(ctx, b, o, tune, stage_segs);
            e = hipStreamEndCapture(ctx->stream, &graph);
        }
        ctx->stream = user;
        ctx->timing = was_timing;
        if (s != GASX_OK) {
            if (graph) hipGraphDestroy(graph);
            return s;
        }
        if (e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return hip_fail(ctx, e, "stream capture");
        }
        e = hipGraphInstantiate(&st.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) {
            st.exec = nullptr;
            return hip_fail(ctx, e, "graph instantiate");
        }
        st.graph_key = std::move(key);
    } else if (st.uploaded_key != upload_key(b, o)) {
        // an eager run or an autotune trial since the capture left other
        // descriptors in the buffers the graph reads
        s = upload(ctx, b, o);
        if (s != GASX_OK) return s;
    }
    TimedScope ts(ctx, kTGraph);
    HIPCHK(hipGraphLaunch(st.exec, ctx->stream), "graph launch");
    return GASX_OK;
}

}  // namespace

void gasx_host::batch_release(gasx_ctx* ctx)
{
    BatchState& b = ctx->batch;
    drop_graph(b);
    free_run_buffers(b);
    free_stage_buffers(b);
    if (b.capture_stream) hipStreamDestroy(b.capture_stream);
    b.capture_stream = nullptr;
}

extern "C" gasx_status gasx_batch_run(gasx_ctx* ctx, gasx_batch* b, const gasx_run_opts* o)
{
    if (!ctx || !b || !o) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    if (o->time_kernels) start_timing(ctx, o->time_kernels, o->time_every);
    gasx_status s = run_batch(ctx, b, o, ctx->tuning);
    ctx->timing = false;
    return s;
}

extern "C" gasx_status gasx_batch_set_tuning(gasx_ctx* ctx, const gasx_batch_tuning* t)
{
    if (!ctx || !t || t->fdtd_ti < 1 || t->fdtd_tk < 1 || t->particle_blocks_per_cu < 1 ||
        t->particle_blocks_per_cu > 64 || !tile_fits(ctx, t->fdtd_ti, t->fdtd_tk))
        return GASX_E_ARG;
    ctx->tuning = *t;
    return GASX_OK;
}

extern "C" gasx_status gasx_batch_get_tuning(gasx_ctx* ctx, gasx_batch_tuning* t)
{
    if (!ctx || !t) return GASX_E_ARG;
    *t = ctx->tuning;
    return GASX_OK;
}

extern "C" gasx_status gasx_batch_autotune(gasx_ctx* ctx, gasx_batch* b, const gasx_run_opts* o, int trial_steps,
                                           gasx_batch_tuning* out)
{
    if (!ctx || !b || !o || trial_steps < 1) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    int64_t stage_segs = 0, pool = 0, ipool = 0;
    gasx_status s = validate(ctx, b, o, &stage_segs, &pool, &ipool);
    if (s != GASX_OK) return s;
    const bool ions = has_ions(b, o);
    s = ensure_batch(ctx, b->nruns, stage_segs, ions);
    if (s != GASX_OK) return s;
    const Sizes z = sizes(ctx);
    const size_t n = (size_t)b->nruns;
    // snapshot of everything a trial changes
    const size_t fbytes = n * (z.ner + z.nez + z.nh) * sizeof(double);
    const size_t pbytes = (size_t)pool * sizeof(float);
    const size_t ibytes = ions ? (size_t)ipool * sizeof(float) : 0;
    char* snap = nullptr;
    HIPCHK(hipMalloc((void**)&snap, 6 * pbytes + 6 * ibytes + fbytes + 2 * n * sizeof(int32_t) +
                                        (size_t)kNCounters * kCounterSlots * 8 + n * kNCounters * 8),
           "alloc snapshot");
    float* pools[6] = {b->x, b->y, b->z, b->vx, b->vy, b->vz};
    float* ipools[6] = {b->ix, b->iy, b->iz, b->ivx, b->ivy, b->ivz};
    auto copy_state = [&](bool save) -> hipError_t {
        char* p = snap;
        auto cp = [&](void* live, size_t bytes) -> hipError_t {
            hipError_t e = save ? hipMemcpyAsync(p, live, bytes, hipMemcpyDeviceToDevice, ctx->stream)
                                : hipMemcpyAsync(live, p, bytes, hipMemcpyDeviceToDevice, ctx->stream);
            p += bytes;
            return e;
        };
        hipError_t e = hipSuccess;
        for (int c = 0; c < 6 && e == hipSuccess; ++c) e = cp(pools[c], pbytes);
        if (e == hipSuccess) e = cp(b->er, n * z.ner * sizeof(double));
        if (e == hipSuccess) e = cp(b->ez, n * z.nez * sizeof(double));
        if (e == hipSuccess) e = cp(b->hphi, n * z.nh * sizeof(double));
        if (e == hipSuccess) e = cp(b->counts, n * sizeof(int32_t));
        for (int c = 0; c < 6 && ions && e == hipSuccess; ++c) e = cp(ipools[c], ibytes);
        if (ions && e == hipSuccess) e = cp(b->ion_counts, n * sizeof(int32_t));
        if (e == hipSuccess) e = cp(ctx->d_counters, (size_t)kNCounters * kCounterSlots * 8);
        if (e == hipSuccess) e = cp(ctx->batch.run_counters, n * kNCounters * 8);
        return e;
    };
    hipError_t e = copy_state(true);
    gasx_run_opts to = *o;
    to.n_steps = trial_steps;
    to.use_graph = 0;
    to.time_kernels = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto trial = [&](const gasx_batch_tuning& t, float* ms) -> gasx_status {
        hipError_t he = copy_state(false);
        if (he != hipSuccess) return hip_fail(ctx, he, "restore snapshot");
        he = hipEventRecord(e0, ctx->stream);
        if (he != hipSuccess) return hip_fail(ctx, he, "event");
        gasx_status ts = run_batch(ctx, b, &to, t);
        if (ts != GASX_OK) return ts;
        he = hipEventRecord(e1, ctx->stream);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipEventElapsedTime(ms, e0, e1);
        return he == hipSuccess ? GASX_OK : hip_fail(ctx, he, "trial timing");
    };
    gasx_batch_tuning best = ctx->tuning;
    if (!tile_fits(ctx, best.fdtd_ti, best.fdtd_tk)) best = gasx_batch_tuning{16, 32, 16};
    float best_ms = 0.0f;
    if (e != hipSuccess) s = hip_fail(ctx, e, "snapshot");
    if (s == GASX_OK) s = trial(best, &best_ms);
    // coordinate search: field tile, then particle grid
    const int tiles[][2] = {{16, 32}, {32, 32}, {16, 64}, {32, 16}, {8, 64}, {64, 16}, {8, 32}};
    for (const auto& tl : tiles) {
        if (s != GASX_OK) break;
        if (!tile_fits(ctx, tl[0], tl[1]) || (tl[0] == best.fdtd_ti && tl[1] == best.fdtd_tk)) continue;
        gasx_batch_tuning c = best;
        c.fdtd_ti = tl[0];
        c.fdtd_tk = tl[1];
        float ms = 0.0f;
        s = trial(c, &ms);
        if (s == GASX_OK && ms < best_ms) { best = c; best_ms = ms; }
    }
    // one or two generations of resident workgroups, or a fixed oversubscription
    const uint32_t step_flags = kBatch | kPush | kDrift2 |
                                (o->collide ? kCollide | (o->secondaries ? kStageSec | kDriftSec : 0u) : 0u) |
                                (ions ? kStageIon : 0u);
    const int occ = particle_occupancy(step_flags);
    const int per_cu[] = {occ, 2 * occ, 4, 8, 16, 32};
    for (int k = 0; k < (int)(sizeof(per_cu) / sizeof(per_cu[0])); ++k) {
        const int pc = per_cu[k];
        if (s != GASX_OK) break;
        bool seen = pc == best.particle_blocks_per_cu || pc > 64;
        for (int j = 0; j < k; ++j) seen = seen || per_cu[j] == pc;
        if (seen) continue;
        gasx_batch_tuning c = best;
        c.particle_blocks_per_cu = pc;
        float ms = 0.0f;
        s = trial(c, &ms);
        if (s == GASX_OK && ms < best_ms) { best = c; best_ms = ms; }
    }
    if (s == GASX_OK) {
        e = copy_state(false);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = hip_fail(ctx, e, "restore snapshot");
    } else {
        (void)copy_state(false);
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(snap);
    if (s != GASX_OK) return s;
    ctx->tuning = best;
    if (out) *out = best;
    return GASX_OK;
}

extern "C" gasx_status gasx_batch_counters(gasx_ctx* ctx, int nruns, gasx_counters* out, int reset)
{
    if (!ctx || !out || nruns < 1 || nruns > ctx->batch.nruns_cap) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    std::vector<unsigned long long> h((size_t)nruns * kNCounters);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->batch.run_counters, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream),
           "copy counters");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    for (int r = 0; r < nruns; ++r) {
        const unsigned long long* c = &h[(size_t)r * kNCounters];
        gasx_counters& o = out[r];
        o.n_elastic = (int64_t)c[kCElastic];
        o.n_ionize = (int64_t)c[kCIonize];
        o.n_excite = (int64_t)c[kCExcite];
        o.n_dropped_sec = (int64_t)c[kCDropSec];
        o.n_dropped_ion = (int64_t)c[kCDropIon];
        o.n_bound_violations = (int64_t)c[kCViol];
        o.n_absorbed = (int64_t)c[kCAbsorbed];
        o.n_saturated = (int64_t)c[kCSat];
        o.n_pushed = (int64_t)c[kCPushed];
    }
    if (reset) {
        HIPCHK(hipMemsetAsync(ctx->batch.run_counters, 0, (size_t)ctx->batch.nruns_cap * kNCounters * 8,
                              ctx->stream),
               "zero counters");
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    }
    return GASX_OK;
}
