// Internal declarations shared by the HIP translation units of libgasx.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "gasx.h"

namespace gasx {

constexpr int kNAng = 181;       // DCS angles 0..180 deg (ElasticCS_Nitrogen.csv rows 2..182)
constexpr int kMaxTab = 64;
constexpr int kMaxExc = 16;
#ifndef GASX_KPER
#define GASX_KPER 4
#endif
constexpr int kBlock = 256;      // particle kernel workgroup size
constexpr int kWaveSize = 64;
constexpr int kPer = GASX_KPER;  // consecutive slots per lane
constexpr int kSeg = kWaveSize * kPer;  // slots per wave segment: the unit of work and of staging
constexpr int kChunk = kBlock * kPer;   // alignment unit of pooled batch runs (a multiple of kSeg)
#ifndef GASX_FDTD_TI
#define GASX_FDTD_TI 16        // field tile of one workgroup: r cells
#endif
#ifndef GASX_FDTD_TK
#define GASX_FDTD_TK 32        // field tile of one workgroup: z cells
#endif
#ifndef GASX_KPER_DRIFT
#define GASX_KPER_DRIFT 4
#endif
// consecutive slots per lane of the half drift + deposit pass (a multiple of 4:
// float4 loads); more slots per lane merge more same-cell contributions in
// registers before they reach the shared current tile, fewer leave room for
// more resident waves (4 measured best for the deposit alone)
constexpr int kPerD = GASX_KPER_DRIFT;
constexpr int kChunkD = kBlock * kPerD;  // slots per drift/deposit workgroup chunk
constexpr int kCounterSlots = 8192;  // per-block event counter slots (>= particle kernel grid)
// Copies of the fixed-point current accumulators: workgroup b adds into copy
// b % kJCopies, which would spread the atomics on the hot cells of a dense
// spot over as many addresses; the copies are summed (exact integer adds)
// before the field cycle reads the current.  The deposit's LDS tiles leave
// one global atomic per touched cell and chunk, so one copy (no fold pass
// at all) is faster than 2, 4 or 8 on the 256x512 production grid.
constexpr int kJCopies = 1;

constexpr int kGuide = 32;       // guide entries per energy node of the elastic-angle CDF

// Device copy of the prepared N2 tables (electron_collision.m:12-41,
// find_CS0.m:14-56).  All energies in eV, cross sections in m^2.  The small
// part is copied into LDS by the collision kernels.
struct SmallTabs {
    int32_t n_el, n_sel, n_ion, n_exc, n_dcs, pad_;
    float el_e[kMaxTab], el_s[kMaxTab];    // TotalElasticCS_Nitrogen (E>=50 eV), sigma*a0^2*2
    float sel_e[kMaxTab], sel_s[kMaxTab];  // TotalElasticCS_smalleV (E<50 eV), sigma*1e-20
    float ion_e[kMaxTab], ion_s[kMaxTab];  // TotalIonCS, sigma*1e-20
    float exc_w[kMaxExc];                  // W_i (eV)
    float exc_q[kMaxExc];                  // q0*A_i/W_i^2*1e-4  (sigma_exc prefactor, find_CS0.m:50)
    float exc_a[kMaxExc];                  // A_i/W_i^2          (E_dep weights, get_ImpExc2.m:24)
    float exc_om[kMaxExc], exc_nu[kMaxExc], exc_ga[kMaxExc];
    float exc_l2w[kMaxExc];                // log2 W_i
    float dcs_loge[kMaxTab];               // log(E keV) of the merged DCS energies
    float dcs_z[kMaxTab];                  // total interval mass Z of each node (see below)
    float sind[184];                       // sin of integer degrees, exact 0 at 0 and 180
    // guide[j][g]: largest k with cdf_j[k] <= g/kGuide (start of the CDF scan)
    uint8_t guide[kMaxTab][kGuide + 1];
};

struct DevTables {
    SmallTabs s;
    float dcs_logv[kMaxTab * kNAng];       // log DCS, [e*181 + angle]
    // elastic-angle sampler: p_k = sind(k) DCS_k per energy node and the
    // normalized cumulative mass of the 1-degree intervals (m_k = p_k + p_{k+1};
    // cdf[0] = 0 ... cdf[180] = 1)
    float dcs_p[kMaxTab * kNAng];
    float dcs_cdf[kMaxTab * kNAng];
};

// Host-side prepared tables (FP64) + the device image.
struct HostTables {
    int n_el = 0, n_sel = 0, n_ion = 0, n_exc = 0, n_dcs = 0;
    double el_e[kMaxTab], el_s[kMaxTab], sel_e[kMaxTab], sel_s[kMaxTab], ion_e[kMaxTab], ion_s[kMaxTab];
    double exc_w[kMaxExc], exc_a[kMaxExc], exc_om[kMaxExc], exc_nu[kMaxExc], exc_ga[kMaxExc];
    double dcs_e[kMaxTab], dcs_loge[kMaxTab], dcs_tcs[kMaxTab], dcs_logv[kMaxTab * kNAng];
};

int prepare_tables(const gasx_raw_tables* raw, HostTables* h, DevTables* d);
// max over 15.6 eV..20 keV of v(E)*sigma_T(E) (m^3/s), FP64, for the rejection bound
double max_rate_coefficient(const HostTables& h, const gasx_params& p);

// ---------------------------------------------------------------------------
// Kernel argument blocks
// ---------------------------------------------------------------------------
struct Geom {
    int32_t nr, nz;
    float dr, dz, inv_dr, inv_dz, rp, lp, phip, cphip, sphip;
    float dt;          // push step
    float c, inv_c2, mu0;
};

enum : uint32_t {
    kPush = 1u,        // EM4.cu: gather + push + half drift + BC
    kCollide = 2u,     // electron_transport.cu
    kDrift = 4u,       // EM2.cu: half drift + BC
    kDeposit = 8u,     // EM7.cu
    kStageSec = 16u,   // stage ionization secondaries for append
    kStageIon = 32u,   // stage N2+ ions for append
    kBatch = 64u,      // chunks belong to the runs of a batched ensemble
    kDriftSec = 128u,  // half drift + BC of staged secondaries (they join at the step's end, x^{n+3/2})
    kDrift2 = 256u,    // the step's second half drift + BC of every electron, after its collision
                       // (the launch that follows then deposits without writing the particles)
};

constexpr int kMaxRuns = 1024;   // runs per batched ensemble

// One run of a batched ensemble (device array).
struct BatchRun {
    int64_t base;      // first pooled slot of the run
    int32_t cap;       // slots owned
    float pmax;        // collision rejection bound of the run's gas
    uint32_t run;      // Philox counter word 2
    int32_t pad_;
};

struct CollArgs {
    const DevTables* tab;
    const float* dens;
    const float* temp;
    int32_t gnr, gnz;
    float inv_gdr, inv_gdz, t_scalar;
    float dt_coll;
    float pmax;            // rejection bound on p_T (>= every candidate's p_T), <= 1
    uint32_t key0, key1, step, run;
    float ev_per_mc2_em;   // me_em c^2 / e_em (eV): energy of the push's electron
    float mc2_mc_ev;       // me_mc c^2 / e_mc (eV): E_e of electron_transport.cu:52
    float ion_vth_k;       // sqrt(k_b / m_ion)
    int32_t exc_form;
    float rot_eps;
    // squared speeds bracketing the tables' energy window [15.6 eV, 20 keV],
    // widened by 1e-4 in energy: a candidate prefilter that is a superset of
    // collide_candidate's own window test on its FP32 energy
    float v2_lo, v2_hi;
};

struct PartArgs {
    float *x, *y, *z, *vx, *vy, *vz;
    // where k_particles reads the particles (normally x..vz; after a cell sort
    // inside the loop, the sort's permuted copies, saving its copy back)
    const float *xi, *yi, *zi, *vxi, *vyi, *vzi;
    int32_t* count;
    int32_t cap;
    float q_over_m;        // qnum*e_em/mass
    float kq;              // qnum * 2^32 / c (deposit fixed-point scale)
    const float *er, *ez, *h;   // FP32 gather copies
    unsigned long long *qr, *qz;  // int64 fixed-point current (two's complement)
    size_t jstride;               // distance between current copies (0: one copy)
    // staging for appends (SoA, indexed segment*kSeg + rank)
    float* sec[6];
    int32_t* sec_cnt;
    float* ion[6];
    int32_t* ion_cnt;
    unsigned long long* counters;  // gasx_counters layout
    // batched ensemble (kBatch): run r's particles start at runs[r].base,
    // its count is count[r], its fields, current and gas follow run-major
    const BatchRun* runs;
    int32_t nruns;
    int32_t* chunk_start;          // nruns+1 segment prefix of the launch (written by block 0)
    unsigned long long* run_counters;  // nruns x kNCounters
};

enum CounterIdx {
    kCElastic = 0, kCIonize, kCExcite, kCDropSec, kCDropIon, kCViol, kCAbsorbed, kCSat, kCPushed,
    kNCounters
};

struct FdtdArgs {
    int32_t nr, nz;
    int32_t ti, tk, halo;        // tile interior size and halo
    int32_t ntr, ntk;            // tiles per run in r and z
    int32_t n_sub;
    int32_t split;               // 0: one launch per cycle; else substeps of the first of two
    int32_t part;                // set per launch: 0 whole cycle, 1 first part, 2 second part
    double dtf;                  // field substep
    double ch_r, ch_z;           // dtf/(mu dr), dtf/(mu dz)
    double ce, cez, c4;          // dtf/eps, dtf/(eps dz), 4 dtf/(eps dr)
    double dr;
    int32_t ez_metric, axis_sign;
    double junit;                // FN*e_em*c*2^-32 (fixed-point -> A/m^2 times V(i))
    double vcoef;                // PHIp/2*dr^2*dz  (V(i) = vcoef*(2i+1))
};

struct FdtdRun {
    const double *er_in, *ez_in, *h_in;
    double *er_out, *ez_out, *h_out;
    double *er_mid, *ez_mid, *h_mid;       // between the two launches of a split cycle
    float *er32, *ez32, *h32;
    const unsigned long long *qr, *qz;     // fixed-point current (or null)
    const double *jr, *jz;                 // FP64 current (or null)
    unsigned long long *qr_clear, *qz_clear;  // accumulators to zero for the next deposit (or null)
};

// Transposed field cycle: field sensitivities in -> out, current sensitivities out.
struct AdjFdtdRun {
    const double *ler_in, *lez_in, *lh_in;
    double *ler_out, *lez_out, *lh_out;
    double *ljr, *ljz;
    double *ler_mid, *lez_mid, *lh_mid;    // between the two launches of a split cycle
};

// ---------------------------------------------------------------------------
// Launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
// runs_dev == nullptr: single run described by r0
hipError_t launch_fdtd(const FdtdArgs& a, const FdtdRun& r0, const FdtdRun* runs_dev, int nruns,
                       int fixed_j, hipStream_t s);
size_t fdtd_lds_bytes(const FdtdArgs& a);        // the transposed cycle's tile
size_t fdtd_fwd_lds_bytes(const FdtdArgs& a);    // the forward cycle's tile (fields and current)
bool fdtd_fwd_shape(const FdtdArgs& a, int* K, int* BS);   // strip length and workgroup size, if any fits
// transposed cycle (adjoint), single run, tiles of a
hipError_t launch_fdtd_adj(const FdtdArgs& a, const AdjFdtdRun& r, hipStream_t s);
// push / collide / stage (kPush, kCollide, kStageSec, kStageIon, kDriftSec, kBatch)
// Workgroups of the particle kernel variant `flags` resident per CU at once.
int particle_occupancy(uint32_t flags);
// ev0 / ev1 (optional) receive the kernel's own start and end times.
hipError_t launch_particles(uint32_t flags, const Geom& g, const PartArgs& pa, const CollArgs& ca,
                            int max_blocks, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// EM2 half drift and / or EM7 deposit (batch: both); the grid is in kChunk
// chunks.
hipError_t launch_drift_deposit(bool drift, bool deposit, bool batch, const Geom& g, const PartArgs& pa,
                                int blocks, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// Appends what a particle launch staged (copy into dst in parent slot order:
// one launch up to 32768 staging segments when sync is given, else a scan
// launch and a copy launch); the appended slots are not deposited here.
// tmp2: two words of the two-launch path; sync: kAppendSyncWords words of the
// one-launch path, zero before the first call (every one-launch call leaves
// them zero).
constexpr int kAppendSyncWords = 2 + 32768 / 256;
hipError_t launch_append(float* const dst[6], int32_t* dst_count, int32_t cap,
                         float* const stage[6], const int32_t* stage_cnt,
                         const int32_t* src_count, int32_t* scan_tmp, int32_t* tmp2, uint32_t* sync,
                         int32_t max_segs, unsigned long long* dropped, hipStream_t s);
// batched append: per-run scan over the chunks recorded in chunk_start, then copy
hipError_t launch_append_batch(float* const pool[6], int32_t* counts, const BatchRun* runs, int nruns,
                               float* const stage[6], const int32_t* stage_cnt,
                               const int32_t* chunk_start, int32_t* offs, int32_t* tmp_run,
                               int32_t total_chunks, unsigned long long* run_counters,
                               unsigned long long* dropped, int drop_idx, hipStream_t s);
hipError_t fdtd_prepare();
hipError_t launch_cross_sections(const DevTables* t, const float* e_kev, int64_t n, float* se,
                                 float* si, float* sx, hipStream_t s);
hipError_t launch_j_from_fixed(const FdtdArgs& a, const unsigned long long* qr,
                               const unsigned long long* qz, double* jr, double* jz, hipStream_t s);
hipError_t launch_to_f32(const double* src, float* dst, int64_t n, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, hipStream_t s);
// q[c] (c = 0..ncells-1) += sum over copies k = 1..kJCopies-1 of q[k*stride + c],
// for both components; the added copies are zeroed.
hipError_t launch_j_reduce(unsigned long long* qr, unsigned long long* qz, int64_t ncells, size_t stride,
                           hipStream_t s);
hipError_t launch_tallies(const gasx_params& p, const double* er, const double* ez, const double* h,
                          const gasx_species* sp, int nsp, double* partial, double* out,
                          hipStream_t s);
// sorted_out non-null: the permuted arrays are left in the scratch (their six
// pointers returned) instead of being copied back into s
// pinned_info: kSortInfoWords pinned host words the slot count and the key
// descents are read through
constexpr int kSortInfoWords = 2 + 4096;
// The sort in two parts.  sort_begin (scratch null or too small: only the
// size in need_bytes, hipErrorInvalidValue) enqueues the key pass, the copy
// of the slot count and the key descents to pinned_info and records `ready`;
// sort_finish waits for `ready` on the host (not for the whole stream),
// then enqueues the radix sort and the permutation.  Work enqueued between
// the two keeps the GPU busy while the host waits.
hipError_t sort_begin(const Geom& g, gasx_species* s, void* scratch, size_t scratch_bytes, size_t* need_bytes,
                      hipStream_t st, int32_t* pinned_info, hipEvent_t ready);
hipError_t sort_finish(const Geom& g, gasx_species* s, void* scratch, hipStream_t st, float** sorted_out,
                       const int32_t* pinned_info, hipEvent_t ready);

}  // namespace gasx
