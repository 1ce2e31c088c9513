// The library context (struct gasx_ctx) and the host helpers shared by the
// translation units that implement the C ABI (api.hip, batch.hip).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "gasx_internal.h"

// Device state of a batched ensemble, sized by the largest batch seen.
struct BatchState {
    int nruns_cap = 0;            // runs the buffers below hold
    int64_t stage_segs = 0;       // staging segments allocated
    double *er_b = nullptr, *ez_b = nullptr, *h_b = nullptr;      // FP64 ping-pong partners, run-major
    double *er_m = nullptr, *ez_m = nullptr, *h_m = nullptr;      // mid buffers of a split field cycle
    float *er32 = nullptr, *ez32 = nullptr, *h32 = nullptr;      // FP32 gather copies, run-major
    unsigned long long* q[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // int64 current, run-major
    gasx::BatchRun* runs = nullptr;             // device descriptors
    gasx::BatchRun* ion_runs = nullptr;         // the runs' ranges of the ion pool
    gasx::FdtdRun* fdtd_runs = nullptr;         // 4 x nruns: [even|odd] x [clear|last]
    int32_t* chunk_start = nullptr;             // nruns + 1 segment prefix
    int32_t* tmp_run = nullptr;                 // 2 x nruns
    int32_t* offs = nullptr;                    // per staging segment
    int32_t* stage_cnt = nullptr;               // per staging segment
    float* stage[6] = {};
    float* istage[6] = {};                      // ion staging (allocated with the first ion batch)
    int32_t* istage_cnt = nullptr;
    int64_t istage_segs = 0;
    unsigned long long* run_counters = nullptr; // nruns x kNCounters
    // captured step sequence (use_graph), replayed while the key matches
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> graph_key;
    // what the device descriptors (runs, fdtd_runs) were last uploaded for:
    // an eager run or an autotune trial in between re-uploads other contents,
    // and a replay then uploads its own again first
    std::vector<unsigned char> uploaded_key;
    hipStream_t capture_stream = nullptr;       // private stream the sequence is captured on
};

// Recorded events, checkpoints and buffers of the differentiable run (adjoint.hip).
struct DiffState;
// Sort buffers, cell map and counters of the DSMC neutral-gas step (dsmc.hip).
struct DsmcState;

struct gasx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    gasx_params p{};
    gasx::Geom geom{};
    gasx::FdtdArgs fa{};
    gasx::HostTables ht{};
    gasx::DevTables* d_tab = nullptr;
    float *er32 = nullptr, *ez32 = nullptr, *h32 = nullptr;      // FP32 gather copies
    double *er_b = nullptr, *ez_b = nullptr, *h_b = nullptr;      // FP64 ping-pong partner
    double *er_m = nullptr, *ez_m = nullptr, *h_m = nullptr;      // mid buffers of a split field cycle
    unsigned long long* q[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [buf][r/z]
    double *jr_tmp = nullptr, *jz_tmp = nullptr;
    unsigned long long* d_counters = nullptr;
    // append staging
    int64_t stage_slots = 0;
    float* stage[6] = {};
    int32_t* stage_cnt = nullptr;
    float* istage[6] = {};
    int32_t* istage_cnt = nullptr;
    int32_t* scan_tmp = nullptr;
    int32_t* base_tmp = nullptr;
    // this device's copy of the library's fixed staging array (gasx_collide_noalloc),
    // from hipGetSymbolAddress at context creation; null if unavailable
    void* scratch = nullptr;
    // sort scratch
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    int32_t* h_count = nullptr;   // kSortInfoWords pinned host words: the sort reads the slot count and the key descents through them
    hipEvent_t sort_ev = nullptr; // recorded after the sort's key pass and count copy
    double* d_tally = nullptr;
    int cu_count = 256;
    double max_rate = 0.0;   // max over the collision window of v(E) sigma_T(E) (m^3/s)
    std::string err;
    // kernel timing (gasx_run_opts.time_kernels): event pairs per launch kind
    bool timing = false;
    unsigned timing_kinds = ~0u;   // bit per TimedKind: the kinds that get event pairs
    int timing_every = 1;          // time launches 0, k, 2k, ... of each kind
    int64_t timing_calls[6] = {};  // launches of each kind since timing was switched on
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    std::vector<std::pair<size_t, size_t>> tpairs[6];   // particle, fdtd, append, sort, graph, drift
    // batched ensemble
    BatchState batch;
    gasx_batch_tuning tuning{16, 32, 16};
    // differentiable run (at most one per context)
    DiffState* diff = nullptr;
    // DSMC neutral gas
    DsmcState* dsmc = nullptr;
};

namespace gasx_host {

enum TimedKind { kTParticle = 0, kTFdtd, kTAppend, kTSort, kTGraph, kTDrift, kNTimed };

size_t next_event(gasx_ctx* c);

// Records an event pair around the enclosed launches when timing is on.
struct TimedScope {
    gasx_ctx* c;
    int kind;
    size_t a = (size_t)-1;
    TimedScope(gasx_ctx* c_, int k) : c(c_), kind(k)
    {
        if (c->timing && (c->timing_kinds >> k & 1u) && c->timing_calls[k]++ % c->timing_every == 0) {
            a = next_event(c);
            if (a != (size_t)-1) (void)hipEventRecord(c->events[a], c->stream);
        }
    }
    ~TimedScope()
    {
        if (c->timing && a != (size_t)-1) {
            size_t b = next_event(c);
            if (b != (size_t)-1) {
                (void)hipEventRecord(c->events[b], c->stream);
                c->tpairs[kind].push_back({a, b});
            }
        }
    }
};

// An event pair that a single launch fills with its own start and end times
// (hipExtLaunchKernel), registered under `kind`; both null when that kind is
// not timed.
struct KernelEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};

inline KernelEvents kernel_events(gasx_ctx* c, int kind)
{
    KernelEvents k;
    if (!c->timing || !(c->timing_kinds >> kind & 1u)) return k;
    if (c->timing_calls[kind]++ % c->timing_every != 0) return k;
    const size_t a = next_event(c);
    if (a == (size_t)-1) return k;
    const size_t b = next_event(c);
    if (b == (size_t)-1) return k;
    k.start = c->events[a];
    k.stop = c->events[b];
    c->tpairs[kind].push_back({a, b});
    return k;
}

// Switches timing on for one run (gasx_run_opts.time_kernels / time_every):
// 2 times the particle-pass launches only (the roofline kernels), any other
// non-zero value every launch; every `every`-th launch of a kind is timed.
inline void start_timing(gasx_ctx* c, int32_t kinds, int32_t every)
{
    c->timing = true;
    c->timing_kinds = kinds == 2 ? (1u << kTParticle) : ~0u;
    c->timing_every = every > 1 ? every : 1;
    for (auto& n : c->timing_calls) n = 0;
    c->ev_used = 0;
    for (auto& v : c->tpairs) v.clear();
}

inline gasx_status fail(gasx_ctx* c, gasx_status s, const std::string& msg)
{
    if (c) c->err = msg;
    return s;
}

inline gasx_status hip_fail(gasx_ctx* c, hipError_t e, const char* where)
{
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    if (c) c->err = m;
    return e == hipErrorOutOfMemory ? GASX_E_NOMEM : GASX_E_HIP;
}

template <typename T>
hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, (n > 0 ? n : 1) * sizeof(T));
}

// Collision arguments of a gas description (rejection bound from the
// context's maximum rate coefficient).
gasx_status make_coll(gasx_ctx* ctx, const gasx_gas* g, double dt_coll, uint64_t seed, uint32_t step,
                      uint32_t run, gasx::CollArgs* ca);
// Persistent particle-kernel grid for `segments` wave segments at most `per_cu` workgroups per CU.
int particle_grid(const gasx_ctx* ctx, int64_t segments, int per_cu);
// Frees the batched-ensemble buffers and graph (batch.hip).
void batch_release(gasx_ctx* ctx);
// Frees the differentiable run (adjoint.hip).
void diff_release(gasx_ctx* ctx);
// Frees the DSMC buffers (dsmc.hip).
void dsmc_release(gasx_ctx* ctx);

}  // namespace gasx_host

#define HIPCHK(expr, where)                                              \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) return gasx_host::hip_fail(ctx, e_, where); \
    } while (0)
