// Particle pass of a step: two kernels over the FP32 SoA.
//
// k_particles<F> (the flags select the phases):
//   kPush     EM4.cu:36-177   gather E,H at x^{n+1/2}, relativistic push, half drift, BC
//   kCollide  electron_transport.cu:64-227 (+ find_CS0.m cross sections)
//   kStageSec / kStageIon     electron_collision.m:63-75 append, staged per wave
//                             segment in slot order so the final order is deterministic
//   kDriftSec                 half drift + BC of the staged secondaries
//   kDrift2   EM2.cu:23-57    the step's second half drift + BC of every electron,
//                             after its collision (the next launch only deposits)
//   kBatch                    segments belong to the runs of a batched ensemble
// Each wave owns segments of kSeg consecutive slots, kPer consecutive slots
// per lane (vector loads and stores), with no workgroup barrier, so waves hide
// each other's latency.  Per segment:
//   A  every lane: load, push, boundary conditions, the decision draw; slots
//      with u1 < pmax (the exact rejection bound on p_T) inside the tables'
//      energy window are candidates; the others make their second half drift
//   B  the segment is stored; its candidates go into the wave's LDS ring; a
//      full ring is one round: a decision pass (cross sections, the test
//      against p_T, the branch) over kQEnt entries per lane, then one pass per
//      collision type (elastic / ionization / excitation, second half drift)
//      over that type's entries compacted onto the lanes, so the collision
//      work is dense instead of divergent; secondaries and ions are staged
//
// k_drift_deposit<kDr, kDe, kBat> (k_deposit<kBat>: the deposit alone):
//   kDrift    EM2.cu:23-57    half drift + BC (the prologue of a call)
//   kDeposit  EM7.cu:51-202   fixed-point current at x^{n+3/2} with v^{n+1},
//                             summed in a workgroup LDS tile and flushed into
//                             the workgroup's copy of the current
//
// The slot count is read once at kernel start; appends land in a later
// kernel (launch_append), so the count is stable inside these kernels.
#include <hip/hip_ext.h>

#include "particle_physics.h"

namespace gasx {

// ---- wave-level helpers (no workgroup barriers) ------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWaveSize - 1)); }

// Orders this wave's LDS accesses around a cross-lane exchange.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix (lane order) and wave total of a per-lane count in
// [0, 2^kBits): one ballot per bit.
template <int kBits>
__device__ __forceinline__ int wave_prefix_small(int v, int& total)
{
    const int lane = lane_id();
    const unsigned long long lt = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    int pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < kBits; ++b) {
        const unsigned long long m = __ballot((v >> b) & 1);
        pre += __popcll(m & lt) << b;
        tot += __popcll(m) << b;
    }
    total = tot;
    return pre;
}

constexpr int kPerBits = kPer <= 1 ? 1 : (kPer <= 3 ? 2 : (kPer <= 7 ? 3 : 4));

// Inclusive prefix sum over the wave (lane order); every lane active.  DPP
// within rows of 16 lanes (shifts by 1, 2, 4 and 8; lanes without a source
// add 0), then the row totals: lane 15 of each row into the next row (rows 1
// and 3), lane 31 into rows 2 and 3 -- no LDS round trips, as the cross-lane
// shuffles would take.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);   // row_bcast:15, rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);   // row_bcast:31, rows 2 and 3
    return v;
}

// Exclusive prefix over the block of per-thread counts (thread order) and the
// block total.  Every thread must call it (used once, before the main loop).
__device__ __forceinline__ int block_excl_scan(int v, int* s_wave, int& total)
{
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        const int c = s_wave[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + incl - v;
}

// ---- block-local current tile ---------------------------------------------------
// A chunk's deposit stencils are summed in a 2-D LDS box of at most kTile
// cells and flushed with one global atomic per non-zero cell and component;
// a stencil cell outside the box (or a box too large) goes to global memory
// directly.  Integer sums: the result does not depend on the grouping.
#ifndef GASX_DEP_TILE
#define GASX_DEP_TILE 384
#endif
constexpr int kTile = GASX_DEP_TILE;   // cells (16 bytes of LDS each)
#ifndef GASX_DEP_COPIES
#define GASX_DEP_COPIES 4
#endif
// Copies of the tile, lane l adding into copy l % kTCopies: neighbouring lanes
// of a cell-sorted chunk mostly add into the same cells, and the LDS
// serialises atomics on one address; the copies (offset by two elements so
// that they sit in different banks) are summed when the tile is flushed.
constexpr int kTCopies = GASX_DEP_COPIES;
constexpr int kTStride = kTile + (kTCopies > 1 ? 2 : 0);

struct TileBox {
    int ilo, klo, wi, nk;     // nk == 0: no tile (nothing to deposit, or the box exceeds kTile)
};

// A box that contains the deposit stencil of a particle at radius r, height z:
// the stencil cells computed with fast arithmetic, widened by one cell (the
// exact arithmetic of deposit_cells can land one cell over at a boundary).
__device__ __forceinline__ void deposit_box(const Geom& g, float r, float z, int& i0, int& i1, int& k0,
                                            int& k1)
{
    const int a = (int)floorf(r * g.inv_dr - 0.5f), b = (int)floorf(z * g.inv_dz - 0.5f);
    i0 = max(min(a - 1, g.nr - 1), 0); i1 = max(min(a + 2, g.nr - 1), 0);
    k0 = max(min(b - 1, g.nz - 1), 0); k1 = max(min(b + 2, g.nz - 1), 0);
}

// Minimum (kMin) or maximum of v over the wave, wave-uniform; every lane
// active.  Within each row of 16 lanes by DPP (lane ^ 1, lane ^ 2, the
// mirror in 8, the mirror in 16: each step pairs every lane with one of the
// other half of its group), then across the four rows by reading lanes 0, 16,
// 32 and 48 (no LDS round trips, as the cross-lane shuffles would take).
template <bool kMin>
__device__ __forceinline__ int wave_reduce(int v)
{
    auto op = [](int a, int b) { return kMin ? min(a, b) : max(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));    // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));   // row_mirror
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

// Wave-wide bounding box of [ilo,ihi] x [klo,khi], left in every lane.
__device__ __forceinline__ void wave_box(int& ilo, int& ihi, int& klo, int& khi)
{
    ilo = wave_reduce<true>(ilo);
    ihi = wave_reduce<false>(ihi);
    klo = wave_reduce<true>(klo);
    khi = wave_reduce<false>(khi);
}

// A box of at most `cells` cells, or none (nk == 0).
__device__ __forceinline__ TileBox fit_box(int ilo, int ihi, int klo, int khi, int cells)
{
    TileBox b{0, 0, 0, 0};
    if (ihi >= ilo && khi >= klo && (ihi - ilo + 1) * (khi - klo + 1) <= cells) {
        b.ilo = ilo; b.klo = klo; b.wi = ihi - ilo + 1; b.nk = khi - klo + 1;
    }
    return b;
}

// Block-wide bounding box of [ilo,ihi] x [klo,khi] (ilo > ihi: none) if it
// fits `cells`; the arguments are replaced by the wave's own box and the
// block's box is left in blk[4].  Every thread must call it.
__device__ __forceinline__ TileBox tile_box(int& ilo, int& ihi, int& klo, int& khi, int* s_red, int cells,
                                            int (&blk)[4])
{
    wave_box(ilo, ihi, klo, khi);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
        s_red[wave] = ilo; s_red[nw + wave] = ihi;
        s_red[2 * nw + wave] = klo; s_red[3 * nw + wave] = khi;
    }
    __syncthreads();
    int bil = 0x7fffffff, bih = -1, bkl = 0x7fffffff, bkh = -1;
    for (int w = 0; w < nw; ++w) {
        bil = min(bil, s_red[w]); bih = max(bih, s_red[nw + w]);
        bkl = min(bkl, s_red[2 * nw + w]); bkh = max(bkh, s_red[3 * nw + w]);
    }
    __syncthreads();
    blk[0] = bil; blk[1] = bih; blk[2] = bkl; blk[3] = bkh;
    return fit_box(bil, bih, bkl, bkh, cells);
}

// The summed contributions of a run of particles with one stencil (first cell
// (i0, k0), di / dk = 1 when the stencil spans two cells in r / z).
struct DepRun {
    int i0, k0, di, dk;       // i0 < 0: empty
    long long q[8];           // qr of the 4 cells, then qz
};

__device__ __forceinline__ void run_add(const TileBox& b, const DepRun& d, int nr, unsigned long long* s_tr,
                                        unsigned long long* s_tz, unsigned long long* qr,
                                        unsigned long long* qz)
{
    // the usual case: the whole stencil inside the box, eight tile adds without branches
    const int ti0 = d.i0 - b.ilo, tk0 = d.k0 - b.klo;
    if ((unsigned)ti0 < (unsigned)b.wi && (unsigned)(ti0 + d.di) < (unsigned)b.wi && (unsigned)tk0 < (unsigned)b.nk &&
        (unsigned)(tk0 + d.dk) < (unsigned)b.nk) {
        const int t0 = ti0 + tk0 * b.wi, t1 = t0 + d.di, t2 = t0 + d.dk * b.wi, t3 = t2 + d.di;
        atomicAdd(&s_tr[t0], (unsigned long long)d.q[0]); atomicAdd(&s_tz[t0], (unsigned long long)d.q[4]);
        atomicAdd(&s_tr[t1], (unsigned long long)d.q[1]); atomicAdd(&s_tz[t1], (unsigned long long)d.q[5]);
        atomicAdd(&s_tr[t2], (unsigned long long)d.q[2]); atomicAdd(&s_tz[t2], (unsigned long long)d.q[6]);
        atomicAdd(&s_tr[t3], (unsigned long long)d.q[3]); atomicAdd(&s_tz[t3], (unsigned long long)d.q[7]);
        return;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = d.i0 + ((a & 1) ? d.di : 0), k = d.k0 + ((a & 2) ? d.dk : 0);
        const unsigned long long vr = (unsigned long long)d.q[a], vz = (unsigned long long)d.q[4 + a];
        const int ti = i - b.ilo, tk = k - b.klo;
        if ((unsigned)ti < (unsigned)b.wi && (unsigned)tk < (unsigned)b.nk) {
            const int t = ti + tk * b.wi;     // (adding a zero costs less than testing for it)
            atomicAdd(&s_tr[t], vr);
            atomicAdd(&s_tz[t], vz);
        } else {                                 // outside the box: straight to global
            const int c = i + k * nr;
            if (vr) atomicAdd(&qr[c], vr);
            if (vz) atomicAdd(&qz[c], vz);
        }
    }
}

#ifndef GASX_GATHER_BATCH
#define GASX_GATHER_BATCH 2
#endif
// particles of a lane whose field reads are in flight together (divides kPer)
constexpr int kGatherBatch = GASX_GATHER_BATCH;
static_assert(kPer % kGatherBatch == 0, "gather batch divides the slots per lane");

static_assert(kPer == 4, "one decision draw covers a lane's four slots");

// LDS ring words per candidate: x y z vx vy vz, u1 (u4 after the decision),
// with 9 words u3 (after the decision; with 8 the ionization pass draws it
// again), slot
#ifndef GASX_QW
#define GASX_QW 8
#endif
#ifndef GASX_QCAP
#define GASX_QCAP 256
#endif
constexpr int kQW = GASX_QW;
constexpr int kQSlotRow = kQW - 1;
constexpr int kQCap = GASX_QCAP;  // candidates a wave's ring holds (a round takes them all; a power of two)
constexpr int kQEnt = kQCap / kWaveSize;   // entries a lane decides per round
static_assert(kQW == 8 || kQW == 9, "ring rows");
static_assert((kQCap & (kQCap - 1)) == 0 && kQCap % kWaveSize == 0 && kQCap <= 256,
              "a power of two, whole entries per lane, byte list indices");

// Event counts of a wave added to the block's slot and, for an ensemble, to
// the run's slot (lane 0 adds; several waves share the block slot).
__device__ __forceinline__ void wave_counts_add(unsigned (&v)[kNCounters], unsigned long long* dst_block,
                                                unsigned long long* dst_run)
{
#pragma unroll
    for (int k = 0; k < kNCounters; ++k) {
        unsigned t = v[k];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
        if (lane_id() == 0 && t) {
            atomicAdd(&dst_block[k], (unsigned long long)t);
            if (dst_run) atomicAdd(&dst_run[k], (unsigned long long)t);
        }
        v[k] = 0;
    }
}

// A lane's event counts in one particle launch, two 16-bit counts per
// register (seven counts live across the whole kernel otherwise take seven
// registers).  A lane counts at most a few events per slot of a segment it
// handles (pushes, absorptions, collision types), and the counts are flushed
// at least every kFlushSegs segments a wave handles, so none reaches 2^16.
constexpr int kFlushSegs = 1024;
struct LaneCounts {
    unsigned ei = 0;       // elastic | ionization << 16
    unsigned xv = 0;       // excitation | bound violation << 16
    unsigned sa = 0;       // saturated | absorbed << 16
    unsigned pushed = 0;
    // every lane of the wave calls it (wave_counts_add)
    __device__ __forceinline__ void flush(unsigned long long* dst_block, unsigned long long* dst_run)
    {
        unsigned v[kNCounters] = {};
        v[kCElastic] = ei & 0xffffu; v[kCIonize] = ei >> 16;
        v[kCExcite] = xv & 0xffffu; v[kCViol] = xv >> 16;
        v[kCSat] = sa & 0xffffu; v[kCAbsorbed] = sa >> 16;
        v[kCPushed] = pushed;
        wave_counts_add(v, dst_block, dst_run);
        ei = xv = sa = pushed = 0;
    }
};

template <int N> struct VecF;
template <> struct VecF<1> { using T = float; };
template <> struct VecF<2> { using T = float2; };
template <> struct VecF<4> { using T = float4; };

__device__ __forceinline__ void vload(const float* p, float (&v)[kPer])
{
    using T = typename VecF<kPer>::T;
    const T t = *reinterpret_cast<const T*>(p);
    static_assert(sizeof(T) == kPer * sizeof(float), "vector width");
    __builtin_memcpy(v, &t, sizeof(T));
}

__device__ __forceinline__ void vstore(float* p, const float (&v)[kPer])
{
    using T = typename VecF<kPer>::T;
    T t;
    __builtin_memcpy(&t, v, sizeof(T));
    *reinterpret_cast<T*>(p) = t;
}

// N consecutive floats (N a multiple of 4) as float4 loads / stores.
template <int N>
__device__ __forceinline__ void vload4n(const float* p, float (&v)[N])
{
    static_assert(N % 4 == 0, "float4 pieces");
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        const float4 t = reinterpret_cast<const float4*>(p)[q];
        v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
    }
}

template <int N>
__device__ __forceinline__ void vstore4n(float* p, const float (&v)[N])
{
    static_assert(N % 4 == 0, "float4 pieces");
#pragma unroll
    for (int q = 0; q < N / 4; ++q)
        reinterpret_cast<float4*>(p)[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

// Push and collision pass.  Waves work independently: wave segments of kSeg
// consecutive slots (lane l owns slots seg*kSeg + l*kPer .. + kPer-1),
// grid-stride over segments; no workgroup barrier inside the loop.
//
// Collision candidates (u1 < pmax, energy inside the tables' window) are
// queued in the wave's LDS ring across segments and processed kQCap at a
// time, so a round is full however few candidates a segment has.
// A segment's slots are stored when it is pushed; a round then rewrites the
// particles that were candidates (after a fence on the wave's earlier stores:
// their velocity after the collision and, with kDrift2, their second half
// drift) and stages their secondaries and ions.  The queue is in (segment,
// slot) order, so each staged entry's rank inside its segment -- the append
// order -- is the parent slot order.  In a batched ensemble each wave takes a
// contiguous range of segments, so consecutive segments mostly share a run and
// the ring is drained only where the run changes.
template <uint32_t F>
__device__ __forceinline__ void particles_pass(const Geom& g, const PartArgs& pa0, const CollArgs& ca0)
{
    static_assert((F & (kDrift | kDeposit)) == 0, "drift and deposit run in k_drift_deposit");
    static_assert((F & (kStageSec | kStageIon)) == 0 || (F & kCollide), "staging needs collisions");
    static_assert((F & kDrift2) == 0 || (F & kPush), "the second half drift follows a push");
    constexpr bool kCol = (F & kCollide) != 0;
    constexpr bool kBat = (F & kBatch) != 0;
    constexpr int kWaves = kBlock / kWaveSize;
    __shared__ int s_wave[kWaves];
    __shared__ float s_q[kCol ? kWaves * kQW * kQCap : 1];   // per wave: kQW rows of kQCap
    __shared__ uint8_t s_list[kCol ? kWaves * kQCap : 1];    // per wave: a round's entries by type
    __shared__ int s_cs[kBat ? kMaxRuns + 1 : 1];
    static_assert(sizeof(SmallTabs) % 4 == 0, "SmallTabs is copied in 32-bit words");
    __shared__ uint32_t s_tab_raw[kCol ? sizeof(SmallTabs) / 4 : 1];
    SmallTabs* s_tab = reinterpret_cast<SmallTabs*>(s_tab_raw);
    if (kCol) {   // small collision tables into LDS (find_CS0 tables, excitation fit, guides)
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&ca0.tab->s);
        for (int i = threadIdx.x; i < (int)(sizeof(SmallTabs) / 4); i += blockDim.x) s_tab_raw[i] = src[i];
        __syncthreads();
    }

    int total_segs;
    if constexpr (kBat) {
        // segment prefix over the runs (every block computes it; block 0 records
        // it for the append kernels, which must see the counts of this launch)
        constexpr int kRPer = kMaxRuns / kBlock;
        int loc[kRPer], sum = 0;
#pragma unroll
        for (int j = 0; j < kRPer; ++j) {
            const int r = (int)threadIdx.x * kRPer + j;
            loc[j] = r < pa0.nruns ? (pa0.count[r] + kSeg - 1) / kSeg : 0;
            sum += loc[j];
        }
        int tot;
        int off = block_excl_scan(sum, s_wave, tot);
#pragma unroll
        for (int j = 0; j < kRPer; ++j) {
            const int r = (int)threadIdx.x * kRPer + j;
            if (r < pa0.nruns) s_cs[r] = off;
            off += loc[j];
        }
        if (threadIdx.x == 0) s_cs[pa0.nruns] = tot;
        __syncthreads();
        if (blockIdx.x == 0)
            for (int r = threadIdx.x; r <= pa0.nruns; r += blockDim.x) pa0.chunk_start[r] = s_cs[r];
        total_segs = tot;
    } else {
        total_segs = (*pa0.count + kSeg - 1) / kSeg;
    }
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int nr_ = g.nr, nz_ = g.nz;
    constexpr uintptr_t kAlign = sizeof(typename VecF<kPer>::T) - 1;
    const bool aligned = ((reinterpret_cast<uintptr_t>(pa0.x) | reinterpret_cast<uintptr_t>(pa0.y) |
                           reinterpret_cast<uintptr_t>(pa0.z) | reinterpret_cast<uintptr_t>(pa0.vx) |
                           reinterpret_cast<uintptr_t>(pa0.vy) | reinterpret_cast<uintptr_t>(pa0.vz) |
                           reinterpret_cast<uintptr_t>(pa0.xi) | reinterpret_cast<uintptr_t>(pa0.yi) |
                           reinterpret_cast<uintptr_t>(pa0.zi) | reinterpret_cast<uintptr_t>(pa0.vxi) |
                           reinterpret_cast<uintptr_t>(pa0.vyi) | reinterpret_cast<uintptr_t>(pa0.vzi)) &
                          kAlign) == 0;
    // reading from other arrays than those written (the permuted copies a cell
    // sort left): every slot below the count is written, dead ones included
    const bool moved = pa0.xi != pa0.x;
    LaneCounts lc;
    int cnt_run = -1;
    int segs_done = 0;                        // segments since the last flush of lc (wave-uniform)

    // ---- the wave's candidate ring (wave-uniform state) ---------------------
    float (*Q)[kQCap] = reinterpret_cast<float (*)[kQCap]>(s_q + (kCol ? wave * kQW * kQCap : 0));
    uint8_t* L = s_list + (kCol ? wave * kQCap : 0);
    int q_head = 0, q_n = 0;
    int c_seg = -1, c_sec = 0, c_ion = 0;   // staging counts carried for the last staged segment
    PartArgs qpa = pa0;                       // the queued entries' run (one run at a time)
    CollArgs qca = ca0;
    int q_cs = 0;                             // first staging segment of that run
    const unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    // A candidate's last part: with kDrift2 its second half drift from the
    // mid-step position with its velocity after the collision (null or not),
    // then its six values; otherwise its new velocity, if it changed.
    auto finish = [&](uint32_t slot, float px, float py, float pz, float vx, float vy, float vz, bool changed) {
        if constexpr ((F & kDrift2) != 0) {
            float x2 = px, y2 = py, z2 = pz;
            half_drift(g, x2, y2, z2, vx, vy, vz);
            if (!wedge_bc(g, x2, y2, z2, vx, vy, vz)) {
                z2 = __int_as_float(0x7fc00000);   // absorbed: z = NaN
                lc.sa += 1u << 16;
            }
            qpa.x[slot] = x2; qpa.y[slot] = y2; qpa.z[slot] = z2;
            qpa.vx[slot] = vx; qpa.vy[slot] = vy; qpa.vz[slot] = vz;
        } else {
            if (changed) { qpa.vx[slot] = vx; qpa.vy[slot] = vy; qpa.vz[slot] = vz; }
        }
    };
    // Processes the m (1..kQCap) oldest queued candidates: a decision pass
    // over kQEnt entries per lane (null ones finish there), then one pass per
    // collision type over that type's entries in queue order, compacted onto
    // the lanes, so each type's kinematics run once for up to 64 collisions
    // instead of once per 64 candidates of all types.
    auto round = [&](int m) {
        // the segments' stores (by other lanes) complete before this round rewrites them
        __threadfence_block();
        auto decide = [&](int e) -> int {
            if (e >= m) return 0;
            const int qi = (q_head + e) & (kQCap - 1);
            const float px = Q[0][qi], py = Q[1][qi], pz = Q[2][qi];
            const float vx = Q[3][qi], vy = Q[4][qi], vz = Q[5][qi];
            const uint32_t slot = __float_as_uint(Q[kQSlotRow][qi]);
            const U4 R = philox4x32_10(slot, qca.step, qca.run, 0u, qca.key0, qca.key1);
            CollDecision d;
            collide_decide(g, qca, s_tab, px, py, pz, vx, vy, vz, Q[6][qi], u01(R.b), d);
            lc.xv += (unsigned)((d.ev & kEvViol) != 0) << 16;
            lc.sa += (unsigned)((d.ev & kEvSat) != 0);
            if (d.typ == 0) {
                finish(slot, px, py, pz, vx, vy, vz, false);
            } else {   // the branch uniforms the kinematics pass reads
                Q[6][qi] = u01(R.d);
                if constexpr (kQW == 9) Q[7][qi] = u01(R.c);
            }
            return d.typ;
        };
        int typ[kQEnt];
#pragma unroll
        for (int j = 0; j < kQEnt; ++j) typ[j] = decide(j * kWaveSize + lane);
        // per-type lists of entry indices in queue order: [0, n1) elastic,
        // [n1, n2) ionization, [n2, n3) excitation
        int lim[4];
        lim[0] = 0;
#pragma unroll
        for (int t = 1; t <= 3; ++t) {
            int acc = lim[t - 1];
#pragma unroll
            for (int j = 0; j < kQEnt; ++j) {
                const unsigned long long mj = __ballot(typ[j] == t);
                if (typ[j] == t) L[acc + __popcll(mj & below)] = (uint8_t)(j * kWaveSize + lane);
                acc += __popcll(mj);
            }
            lim[t] = acc;
        }
        wave_sync();
        // elastic
        for (int k0 = lim[0]; k0 < lim[1]; k0 += kWaveSize) {
            const int k = k0 + lane;
            if (k < lim[1]) {
                const int qi = (q_head + L[k]) & (kQCap - 1);
                const float px = Q[0][qi], py = Q[1][qi], pz = Q[2][qi];
                float vx = Q[3][qi], vy = Q[4][qi], vz = Q[5][qi];
                const uint32_t slot = __float_as_uint(Q[kQSlotRow][qi]);
                float T, ve;
                candidate_energy(g, qca, vx, vy, vz, T, ve);
                collide_elastic(qca, s_tab, slot, T, Q[6][qi], vx, vy, vz);
                finish(slot, px, py, pz, vx, vy, vz, true);
                lc.ei += 1u;
            }
        }
        // excitation
        for (int k0 = lim[2]; k0 < lim[3]; k0 += kWaveSize) {
            const int k = k0 + lane;
            if (k < lim[3]) {
                const int qi = (q_head + L[k]) & (kQCap - 1);
                const float px = Q[0][qi], py = Q[1][qi], pz = Q[2][qi];
                float vx = Q[3][qi], vy = Q[4][qi], vz = Q[5][qi];
                const uint32_t slot = __float_as_uint(Q[kQSlotRow][qi]);
                float T, ve;
                candidate_energy(g, qca, vx, vy, vz, T, ve);
                collide_excite(g, qca, s_tab, T, ve, vx, vy, vz);
                finish(slot, px, py, pz, vx, vy, vz, true);
                lc.xv += 1u;
            }
        }
        // ionization: secondaries and ions staged at their rank in the
        // segment (the passes take the ionizations in slot order); the last
        // lane of each segment records the segment's counts so far (a later
        // pass of the same segment overwrites them with the larger totals)
        for (int k0 = lim[1]; k0 < lim[2]; k0 += kWaveSize) {
            const int k = k0 + lane;
            const int cnt = min(lim[2] - k0, kWaveSize);
            const bool act = lane < cnt;
            int gs = 0;
            float px = 0.f, py = 0.f, pz = 0.f;
            CollOut co{};
            if (act) {
                const int qi = (q_head + L[k]) & (kQCap - 1);
                px = Q[0][qi]; py = Q[1][qi]; pz = Q[2][qi];
                float vx = Q[3][qi], vy = Q[4][qi], vz = Q[5][qi];
                const uint32_t slot = __float_as_uint(Q[kQSlotRow][qi]);
                float T, ve, u3;
                if constexpr (kQW == 9) {
                    u3 = Q[7][qi];
                } else {   // the decision's draw again
                    const U4 R = philox4x32_10(slot, qca.step, qca.run, 0u, qca.key0, qca.key1);
                    u3 = u01(R.c);
                }
                candidate_energy(g, qca, vx, vy, vz, T, ve);
                collide_ionize(g, qca, slot, px, py, pz, T, ve, u3, Q[6][qi], (F & kStageIon) != 0, vx, vy,
                               vz, co);
                finish(slot, px, py, pz, vx, vy, vz, true);
                lc.ei += 1u << 16;
                gs = q_cs + (int)(slot / kSeg);
            }
            bool has_sec = (F & kStageSec) && act;
            const bool has_ion = (F & kStageIon) && act;
            float sx = px, sy = py, sz = pz, svx = co.svx, svy = co.svy, svz = co.svz;
            if ((F & kDriftSec) && has_sec) {
                half_drift(g, sx, sy, sz, svx, svy, svz);
                if (!wedge_bc(g, sx, sy, sz, svx, svy, svz)) { has_sec = false; lc.sa += 1u << 16; }
            }
            if constexpr ((F & (kStageSec | kStageIon)) != 0) {
                // rank of each staged entry inside its segment: entries of earlier
                // passes (carried) plus the earlier lanes of this pass in the segment
                const int gprev = __shfl_up(gs, 1, 64);
                const int gnext = __shfl_down(gs, 1, 64);
                const unsigned long long heads = __ballot(act && (lane == 0 || gs != gprev));
                const int head = 63 - __clzll(heads & (below | (1ull << lane)));
                const unsigned long long span = below & ~((head == 0) ? 0ull : ((~0ull) >> (64 - head)));
                int rs = __popcll(__ballot(has_sec) & span), ri = __popcll(__ballot(has_ion) & span);
                if (gs == c_seg) { rs += c_sec; ri += c_ion; }
                if (has_sec) {
                    const int d = gs * kSeg + rs;
                    qpa.sec[0][d] = sx; qpa.sec[1][d] = sy; qpa.sec[2][d] = sz;
                    qpa.sec[3][d] = svx; qpa.sec[4][d] = svy; qpa.sec[5][d] = svz;
                }
                if (has_ion) {   // ion at the ionization site (electron_collision.m:63)
                    const int d = gs * kSeg + ri;
                    qpa.ion[0][d] = px; qpa.ion[1][d] = py; qpa.ion[2][d] = pz;
                    qpa.ion[3][d] = co.ivx; qpa.ion[4][d] = co.ivy; qpa.ion[5][d] = co.ivz;
                }
                const int rs1 = rs + has_sec, ri1 = ri + has_ion;
                if (act && (lane == cnt - 1 || gnext != gs)) {
                    if (F & kStageSec) qpa.sec_cnt[gs] = rs1;
                    if (F & kStageIon) qpa.ion_cnt[gs] = ri1;
                }
                c_seg = __shfl(gs, cnt - 1, 64);
                c_sec = __shfl(rs1, cnt - 1, 64);
                c_ion = __shfl(ri1, cnt - 1, 64);
            }
        }
        wave_sync();   // the ring entries and the lists are free
        q_head = (q_head + m) & (kQCap - 1);
        q_n -= m;
    };
    auto drain = [&]() {
        while (q_n > 0) round(q_n < kQCap ? q_n : kQCap);
    };

    const int n_waves = gridDim.x * kWaves;
    // ensembles on the ring: each wave takes a contiguous range of segments, so
    // consecutive segments mostly share a run and the ring drains only where
    // the run changes; otherwise waves grid-stride over the segments
    int g_begin = blockIdx.x * kWaves + wave, g_end = total_segs, g_step = n_waves;
    if constexpr (kBat) {
        const int per = (total_segs + n_waves - 1) / n_waves;
        g_begin = (blockIdx.x * kWaves + wave) * per;
        g_end = min(total_segs, g_begin + per);
        g_step = 1;
    }
    for (int gseg = g_begin; gseg < g_end; gseg += g_step) {
        if (++segs_done > kFlushSegs) {           // the 16-bit counts stay below 2^16
            if (pa0.counters && (!kBat || cnt_run >= 0))
                lc.flush(pa0.counters + (size_t)blockIdx.x * kNCounters,
                         kBat ? pa0.run_counters + (size_t)cnt_run * kNCounters : nullptr);
            segs_done = 1;
        }
        // seg: index within the population (slots, Philox counters);
        // gseg: index of the staging area
        PartArgs pa = pa0;
        CollArgs ca = ca0;
        int seg = gseg, n, run = 0;
        if constexpr (kBat) {
            int lo = 0, hi = pa0.nruns;            // s_cs[lo] <= gseg < s_cs[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_cs[mid] <= gseg) lo = mid; else hi = mid;
            }
            run = lo;
            seg = gseg - s_cs[lo];
            const BatchRun br = pa0.runs[run];
            pa.x += br.base; pa.y += br.base; pa.z += br.base;
            pa.vx += br.base; pa.vy += br.base; pa.vz += br.base;
            pa.xi += br.base; pa.yi += br.base; pa.zi += br.base;
            pa.vxi += br.base; pa.vyi += br.base; pa.vzi += br.base;
            n = pa0.count[run];
            const size_t nh = (size_t)nr_ * nz_;
            pa.er += (size_t)run * (nh + nr_);
            pa.ez += (size_t)run * (nh + nz_);
            pa.h += (size_t)run * nh;
            const size_t gc = (size_t)ca0.gnr * ca0.gnz;
            if (ca.dens) ca.dens += (size_t)run * gc;
            if (ca.temp) ca.temp += (size_t)run * gc;
            ca.pmax = br.pmax;
            ca.run = br.run;
            if (run != cnt_run) {
                if constexpr (kCol) {
                    // the queued candidates belong to the previous run: process
                    // them (counted for it), then queue this run's
                    drain();
                    qpa = pa;
                    qca = ca;
                    q_cs = s_cs[run];
                }
                if (cnt_run >= 0 && pa0.counters)
                    lc.flush(pa0.counters + (size_t)blockIdx.x * kNCounters,
                             pa0.run_counters + (size_t)cnt_run * kNCounters);
                cnt_run = run;
                segs_done = 0;
            }
        } else {
            n = *pa0.count;
        }
        const int slot0 = seg * kSeg + lane * kPer;
        const bool vec = aligned && slot0 + kPer <= n;
        float x[kPer], y[kPer], z[kPer], vx[kPer], vy[kPer], vz[kPer];
        bool alive[kPer], was_alive[kPer], cand[kPer];
        float u1[kPer];
        if (vec) {
            vload(pa.xi + slot0, x); vload(pa.yi + slot0, y); vload(pa.zi + slot0, z);
            vload(pa.vxi + slot0, vx); vload(pa.vyi + slot0, vy); vload(pa.vzi + slot0, vz);
        } else {
#pragma unroll
            for (int p = 0; p < kPer; ++p) {
                const int slot = slot0 + p;
                x[p] = y[p] = vx[p] = vy[p] = vz[p] = 0.f;
                z[p] = __int_as_float(0x7fc00000);
                if (slot < n) {
                    x[p] = pa.xi[slot]; y[p] = pa.yi[slot]; z[p] = pa.zi[slot];
                    vx[p] = pa.vxi[slot]; vy[p] = pa.vyi[slot]; vz[p] = pa.vzi[slot];
                }
            }
        }
        // ---- A: push + first Philox block; candidates ---------------------------
        // the field reads of kGatherBatch particles are issued together (a
        // slot that is not pushed reads at the origin and drops the values)
        FieldGather G[kGatherBatch];
#pragma unroll
        for (int p = 0; p < kPer; ++p) {
            const int slot = slot0 + p;
            alive[p] = slot < n && !is_dead(z[p]);
            was_alive[p] = alive[p];
            cand[p] = false;
            if ((F & kPush) && p % kGatherBatch == 0) {
#pragma unroll
                for (int q = p; q < p + kGatherBatch; ++q) {
                    const bool on = slot0 + q < n && !is_dead(z[q]);
                    gather_issue(g, pa, on ? x[q] : 0.f, on ? y[q] : 0.f, on ? z[q] : 0.f, G[q - p]);
                }
            }
            if (F & kPush) {
                // without a branch: a dead slot is pushed too (its z stays NaN, and
                // nothing reads its other values)
                push_apply(g, pa, G[p % kGatherBatch], vx[p], vy[p], vz[p]);
                half_drift(g, x[p], y[p], z[p], vx[p], vy[p], vz[p]);
                if (alive[p] && !wedge_bc(g, x[p], y[p], z[p], vx[p], vy[p], vz[p])) {
                    alive[p] = false;
                    lc.sa += 1u << 16;
                }
            }
            if (kCol && p == 0) {   // the decision uniforms of the lane's four slots: one draw
                const U4 D = philox4x32_10((uint32_t)slot0 >> 2, ca.step, ca.run, kDecisionBlock, ca.key0, ca.key1);
                u1[0] = u01(D.a); u1[1] = u01(D.b); u1[2] = u01(D.c); u1[3] = u01(D.d);
            }
            if (kCol && alive[p]) {
                // exact: p_T <= pmax for every particle, and collide_candidate does
                // nothing outside the tables' energy window; the window is tested on
                // v^2 here, widened by 1e-4 in energy (far above the few-ulp error
                // of collide_candidate's FP32 energy), so no particle that
                // collide_candidate would test is dropped
                const float v2 = vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p];
                cand[p] = u1[p] < ca.pmax && v2 >= ca.v2_lo && v2 <= ca.v2_hi;
            }
            if (F & kDrift2) {   // a candidate's is done by its round: selected, not branched
                const bool d2 = alive[p] && !cand[p];
                float x2 = x[p], y2 = y[p], z2 = z[p];
                half_drift(g, x2, y2, z2, vx[p], vy[p], vz[p]);
                x[p] = d2 ? x2 : x[p]; y[p] = d2 ? y2 : y[p]; z[p] = d2 ? z2 : z[p];
                if (d2 && !wedge_bc(g, x[p], y[p], z[p], vx[p], vy[p], vz[p])) {
                    alive[p] = false;
                    lc.sa += 1u << 16;
                }
            }
            if (F & kPush) lc.pushed += was_alive[p];
            if (was_alive[p] && !alive[p]) z[p] = __int_as_float(0x7fc00000);   // absorbed: z = NaN
        }
        {
            // ---- store (a later round rewrites the candidates) ----------------------
            if (vec) {   // a dead slot keeps z = NaN; its other values are the unbranched push's
                vstore(pa.x + slot0, x); vstore(pa.y + slot0, y); vstore(pa.z + slot0, z);
                vstore(pa.vx + slot0, vx); vstore(pa.vy + slot0, vy); vstore(pa.vz + slot0, vz);
            } else {
    #pragma unroll
                for (int p = 0; p < kPer; ++p) {
                    const int slot = slot0 + p;
                    if (slot < n && (was_alive[p] || moved)) {
                        if (alive[p] || moved) {
                            pa.x[slot] = x[p]; pa.y[slot] = y[p]; pa.z[slot] = z[p];
                            pa.vx[slot] = vx[p]; pa.vy[slot] = vy[p]; pa.vz[slot] = vz[p];
                        } else {
                            pa.z[slot] = z[p];
                        }
                    }
                }
            }
            // ---- B: queue the candidates; full rounds as the ring fills -------------
            if constexpr (kCol) {
                int mine = 0;
    #pragma unroll
                for (int p = 0; p < kPer; ++p) mine += cand[p];
                int total;
                const int off = wave_prefix_small<kPerBits>(mine, total);
                if (lane == 0) {   // a round's ionization pass overwrites them with the segment's counts
                    if (F & kStageSec) pa.sec_cnt[gseg] = 0;
                    if (F & kStageIon) pa.ion_cnt[gseg] = 0;
                }
                int done = 0;
                while (done < total) {
                    const int take = min(total - done, kQCap - q_n);
                    int pos = off;
    #pragma unroll
                    for (int p = 0; p < kPer; ++p) {
                        if (cand[p]) {
                            const int r = pos - done;
                            if (r >= 0 && r < take) {
                                int qi = q_head + q_n + r;
                                if (qi >= kQCap) qi -= kQCap;
                                Q[0][qi] = x[p]; Q[1][qi] = y[p]; Q[2][qi] = z[p];
                                Q[3][qi] = vx[p]; Q[4][qi] = vy[p]; Q[5][qi] = vz[p];
                                Q[6][qi] = u1[p];
                                Q[kQSlotRow][qi] = __uint_as_float((uint32_t)(slot0 + p));
                            }
                            ++pos;
                        }
                    }
                    q_n += take;
                    done += take;
                    wave_sync();
                    if (q_n == kQCap) round(kQCap);
                }
            }
        }
    }
    if constexpr (kCol) drain();
    if (pa0.counters && (!kBat || cnt_run >= 0))
        lc.flush(pa0.counters + (size_t)blockIdx.x * kNCounters,
                 kBat ? pa0.run_counters + (size_t)cnt_run * kNCounters : nullptr);
}

// Single run: at least four waves per SIMD.  The 128-register budget that
// allows it costs a few bytes of spill; the extra resident wave hides more of
// the gather and table latency (8% faster at 8e6 particles than at 138
// registers and three waves).  Ensembles keep the unconstrained budget.
template <uint32_t F>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_particles(Geom g, PartArgs pa0,
                                                                                             CollArgs ca0)
{
    particles_pass<F>(g, pa0, ca0);
}

template <uint32_t F>
__global__ __launch_bounds__(kBlock) void k_particles_batch(Geom g, PartArgs pa0, CollArgs ca0)
{
    particles_pass<F>(g, pa0, ca0);
}

// ---- half drift + deposit (EM2.cu:23-57, EM7.cu:51-202) ----------------------
// Workgroups take chunks of kChunkD consecutive slots (thread t owns slots
// chunk*kChunkD + t*kPerD .. + kPerD-1, float4 loads and stores).  Each thread
// sums the contributions of its consecutive particles while they share a
// stencil (cell-sorted slots mostly do), adds each run into the chunk's LDS
// tile, and the tile is flushed into the workgroup's copy of the current.
template <bool kDr, bool kDe, bool kBat>
__device__ __forceinline__ void drift_deposit_pass(const Geom& g, const PartArgs& pa0, int grid)
{
    constexpr int kWaves = kBlock / kWaveSize;
    __shared__ int s_wave[kWaves];
    __shared__ int s_red[4 * kWaves];
    __shared__ unsigned long long s_tr[kDe ? kTCopies * kTStride : 1], s_tz[kDe ? kTCopies * kTStride : 1];
    __shared__ int s_cs[kBat ? kMaxRuns + 1 : 1];
    int total_chunks;
    if constexpr (kBat) {   // chunk prefix over the runs (run bases are multiples of kChunk)
        constexpr int kRPer = kMaxRuns / kBlock;
        int loc[kRPer], sum = 0;
#pragma unroll
        for (int j = 0; j < kRPer; ++j) {
            const int r = (int)threadIdx.x * kRPer + j;
            loc[j] = r < pa0.nruns ? (pa0.count[r] + kChunkD - 1) / kChunkD : 0;
            sum += loc[j];
        }
        int tot;
        int off = block_excl_scan(sum, s_wave, tot);
#pragma unroll
        for (int j = 0; j < kRPer; ++j) {
            const int r = (int)threadIdx.x * kRPer + j;
            if (r < pa0.nruns) s_cs[r] = off;
            off += loc[j];
        }
        if (threadIdx.x == 0) s_cs[pa0.nruns] = tot;
        __syncthreads();
        total_chunks = tot;
    } else {
        total_chunks = (*pa0.count + kChunkD - 1) / kChunkD;
    }
    const int nr_ = g.nr, nz_ = g.nz;
    constexpr uintptr_t kAlign = sizeof(float4) - 1;
    // a single run may read its particles from a sort's permuted copies (xi..vzi)
    // and write them back into x..vz; a batch reads and writes x..vz
    const bool copy_in = !kBat && pa0.xi != pa0.x;
    const bool aligned = ((reinterpret_cast<uintptr_t>(pa0.x) | reinterpret_cast<uintptr_t>(pa0.y) |
                           reinterpret_cast<uintptr_t>(pa0.z) | reinterpret_cast<uintptr_t>(pa0.vx) |
                           reinterpret_cast<uintptr_t>(pa0.vy) | reinterpret_cast<uintptr_t>(pa0.vz) |
                           (kBat ? 0 : reinterpret_cast<uintptr_t>(pa0.xi) | reinterpret_cast<uintptr_t>(pa0.yi) |
                                       reinterpret_cast<uintptr_t>(pa0.zi) | reinterpret_cast<uintptr_t>(pa0.vxi) |
                                       reinterpret_cast<uintptr_t>(pa0.vyi) | reinterpret_cast<uintptr_t>(pa0.vzi))) &
                          kAlign) == 0;
    unsigned cnt[kNCounters] = {};
    int cnt_run = -1;
    for (int chunk = blockIdx.x; chunk < total_chunks; chunk += grid) {   // workgroup-uniform
        PartArgs pa = pa0;
        int lchunk = chunk, n, run = 0;
        if constexpr (kBat) {
            int lo = 0, hi = pa0.nruns;            // s_cs[lo] <= chunk < s_cs[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_cs[mid] <= chunk) lo = mid; else hi = mid;
            }
            run = lo;
            lchunk = chunk - s_cs[lo];
            const int64_t base = pa0.runs[run].base;
            pa.x += base; pa.y += base; pa.z += base;
            pa.vx += base; pa.vy += base; pa.vz += base;
            n = pa0.count[run];
            if (pa.qr) {
                const size_t nh = (size_t)nr_ * nz_;
                pa.qr += (size_t)run * nh;
                pa.qz += (size_t)run * nh;
            }
            if (run != cnt_run) {
                if (cnt_run >= 0 && pa0.counters)
                    wave_counts_add(cnt, pa0.counters + (size_t)blockIdx.x * kNCounters,
                                    pa0.run_counters + (size_t)cnt_run * kNCounters);
                cnt_run = run;
            }
        } else {
            n = *pa0.count;
        }
        if (kDe && pa0.jstride) {                  // this workgroup's copy of the current
            const size_t off = (size_t)(blockIdx.x % kJCopies) * pa0.jstride;
            pa.qr += off;
            pa.qz += off;
        }
        const int slot0 = lchunk * kChunkD + (int)threadIdx.x * kPerD;
        const bool vec = aligned && slot0 + kPerD <= n;
        float x[kPerD], y[kPerD], z[kPerD], vx[kPerD], vy[kPerD], vz[kPerD];
        bool alive[kPerD], was_alive[kPerD];
        const float *sx = pa.x, *sy = pa.y, *sz = pa.z, *svx = pa.vx, *svy = pa.vy, *svz = pa.vz;
        if (copy_in) { sx = pa.xi; sy = pa.yi; sz = pa.zi; svx = pa.vxi; svy = pa.vyi; svz = pa.vzi; }
        if (vec) {
            vload4n(sx + slot0, x); vload4n(sy + slot0, y); vload4n(sz + slot0, z);
            vload4n(svx + slot0, vx); vload4n(svy + slot0, vy); vload4n(svz + slot0, vz);
        } else {
#pragma unroll
            for (int p = 0; p < kPerD; ++p) {
                const int slot = slot0 + p;
                x[p] = y[p] = vx[p] = vy[p] = vz[p] = 0.f;
                z[p] = __int_as_float(0x7fc00000);
                if (slot < n) {
                    x[p] = sx[slot]; y[p] = sy[slot]; z[p] = sz[slot];
                    vx[p] = svx[slot]; vy[p] = svy[slot]; vz[p] = svz[slot];
                }
            }
        }
        // the thread's extent in r^2 and z; its stencil box once, from the extremes
        // (the box is monotone in both)
        float r2lo = INFINITY, r2hi = -INFINITY, zlo = INFINITY, zhi = -INFINITY;
#pragma unroll
        for (int p = 0; p < kPerD; ++p) {
            alive[p] = slot0 + p < n && !is_dead(z[p]);
            was_alive[p] = alive[p];
            if (kDr && alive[p]) {
                half_drift(g, x[p], y[p], z[p], vx[p], vy[p], vz[p]);
                if (!wedge_bc(g, x[p], y[p], z[p], vx[p], vy[p], vz[p])) {
                    alive[p] = false;
                    z[p] = __int_as_float(0x7fc00000);   // absorbed: z = NaN
                    ++cnt[kCAbsorbed];
                }
            }
            if (kDe && alive[p]) {
                const float r2 = x[p] * x[p] + y[p] * y[p];
                r2lo = fminf(r2lo, r2); r2hi = fmaxf(r2hi, r2);
                zlo = fminf(zlo, z[p]); zhi = fmaxf(zhi, z[p]);
            }
        }
        int ilo = 0x7fffffff, ihi = -1, klo = 0x7fffffff, khi = -1;
        if (kDe && r2lo <= r2hi) {
            int i0, i1, k0, k1;
            deposit_box(g, fsqrt(r2lo), zlo, i0, i1, k0, k1);
            ilo = i0; klo = k0;
            deposit_box(g, fsqrt(r2hi), zhi, i0, i1, k0, k1);
            ihi = i1; khi = k1;
        }
        if (kDr) {
            if (vec) {   // dead slots are written back unchanged
                vstore4n(pa.x + slot0, x); vstore4n(pa.y + slot0, y); vstore4n(pa.z + slot0, z);
                vstore4n(pa.vx + slot0, vx); vstore4n(pa.vy + slot0, vy); vstore4n(pa.vz + slot0, vz);
            } else {
#pragma unroll
                for (int p = 0; p < kPerD; ++p) {
                    const int slot = slot0 + p;
                    if (slot < n && (was_alive[p] || copy_in)) {
                        if (alive[p] || copy_in) {
                            pa.x[slot] = x[p]; pa.y[slot] = y[p]; pa.z[slot] = z[p];
                            pa.vx[slot] = vx[p]; pa.vy[slot] = vy[p]; pa.vz[slot] = vz[p];
                        } else {
                            pa.z[slot] = z[p];
                        }
                    }
                }
            }
        }
        if constexpr (kDe) {
            // the workgroup's box in the tile's kTCopies copies (dense chunks:
            // neighbouring lanes add into the same cells); a larger box in the
            // whole tile as one copy (sparse chunks, as the appended
            // secondaries of slowly growing runs are between cell sorts:
            // few particles per cell, so little contention); else each wave
            // its own box in one copy's worth; what lies outside goes to global
            int blk[4];
            TileBox bx = tile_box(ilo, ihi, klo, khi, s_red, kTile, blk);
            int t0 = (int)threadIdx.x, tstep = (int)blockDim.x, tbase = 0, ncop = kTCopies;
            unsigned long long *my_tr, *my_tz;
            if (bx.nk != 0) {                      // workgroup-uniform
                my_tr = s_tr + (threadIdx.x % kTCopies) * kTStride;
                my_tz = s_tz + (threadIdx.x % kTCopies) * kTStride;
            } else {
                ncop = 1;
                bx = fit_box(blk[0], blk[1], blk[2], blk[3], kTCopies * kTStride);
                if (bx.nk == 0) {                  // workgroup-uniform
                    bx = fit_box(ilo, ihi, klo, khi, kTStride);
                    t0 = (int)(threadIdx.x & (kWaveSize - 1));
                    tstep = kWaveSize;
                    tbase = (int)(threadIdx.x / kWaveSize) * kTStride;
                }
                my_tr = s_tr + tbase;
                my_tz = s_tz + tbase;
            }
            for (int t = t0; t < bx.wi * bx.nk; t += tstep) {
                for (int c = 0; c < ncop; ++c) {
                    s_tr[c * kTStride + tbase + t] = 0ull;
                    s_tz[c * kTStride + tbase + t] = 0ull;
                }
            }
            __syncthreads();
            DepRun run_ = {-1, 0, 0, 0, {0, 0, 0, 0, 0, 0, 0, 0}};
#pragma unroll
            for (int p = 0; p < kPerD; ++p) {
                if (alive[p]) {
                    DepCells dc;
                    deposit_cells(g, x[p], y[p], z[p], vx[p], vy[p], vz[p], pa.kq, dc);
                    const int di = dc.i1 - dc.i0, dk = dc.k1 - dc.k0;
                    if (dc.i0 != run_.i0 || dc.k0 != run_.k0 || di != run_.di || dk != run_.dk) {
                        if (run_.i0 >= 0) run_add(bx, run_, nr_, my_tr, my_tz, pa.qr, pa.qz);
                        run_.i0 = dc.i0; run_.k0 = dc.k0; run_.di = di; run_.dk = dk;
#pragma unroll
                        for (int a = 0; a < 4; ++a) { run_.q[a] = dc.qr[a]; run_.q[4 + a] = dc.qz[a]; }
                    } else {
#pragma unroll
                        for (int a = 0; a < 4; ++a) { run_.q[a] += dc.qr[a]; run_.q[4 + a] += dc.qz[a]; }
                    }
                }
            }
            if (run_.i0 >= 0) run_add(bx, run_, nr_, my_tr, my_tz, pa.qr, pa.qz);
            __syncthreads();
            for (int t = t0; t < bx.wi * bx.nk; t += tstep) {
                unsigned long long a = s_tr[tbase + t], c = s_tz[tbase + t];
                for (int k = 1; k < ncop; ++k) {
                    a += s_tr[k * kTStride + tbase + t];
                    c += s_tz[k * kTStride + tbase + t];
                }
                const int cell = (bx.ilo + t % bx.wi) + (bx.klo + t / bx.wi) * nr_;
                if (a) atomicAdd(&pa.qr[cell], a);
                if (c) atomicAdd(&pa.qz[cell], c);
            }
            __syncthreads();                       // the tile is reused by the next chunk
        }
    }
    if (pa0.counters && (!kBat || cnt_run >= 0))
        wave_counts_add(cnt, pa0.counters + (size_t)blockIdx.x * kNCounters,
                        kBat ? pa0.run_counters + (size_t)cnt_run * kNCounters : nullptr);
}

template <bool kDr, bool kDe, bool kBat>
__global__ __launch_bounds__(kBlock) void k_drift_deposit(Geom g, PartArgs pa0)
{
    drift_deposit_pass<kDr, kDe, kBat>(g, pa0, (int)gridDim.x);
}

constexpr int kScanGroup = 1024;         // staging segments per group of the append's scan
constexpr int kMaxScanGroups = 1024;     // staging of up to 2^20 segments

// The deposit alone (every push of the production loop): six waves per SIMD
// (80 registers, a 384-cell tile in four copies) hide more of its LDS and
// memory latency than four.  It runs after the step's appends, so the
// appended secondaries are deposited with the rest through the LDS tiles.
template <bool kBat>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_deposit(Geom g, PartArgs pa0)
{
    drift_deposit_pass<false, true, kBat>(g, pa0, (int)gridDim.x);
}

hipError_t launch_drift_deposit(bool drift, bool deposit, bool batch, const Geom& g, const PartArgs& pa,
                                int blocks, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    blocks = blocks > 0 ? blocks : 1;
    // the events (if any) take the kernel's own start and end times: no marker
    // packets between the launches of the step
    if (batch) {
        if (!deposit) return hipErrorInvalidValue;
        if (drift)
            hipExtLaunchKernelGGL((k_drift_deposit<true, true, true>), dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa);
        else
            hipExtLaunchKernelGGL(k_deposit<true>, dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa);
    } else if (drift && deposit) {
        hipExtLaunchKernelGGL((k_drift_deposit<true, true, false>), dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa);
    } else if (drift) {
        hipExtLaunchKernelGGL((k_drift_deposit<true, false, false>), dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa);
    } else if (deposit) {
        hipExtLaunchKernelGGL(k_deposit<false>, dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- append: exclusive scan of per-chunk counts, then copy ------------------
// src_count: slot count of the population whose kernel staged the entries
// (it fixes the number of chunks); dst_count: the population appended to.
// tmp2[0] <- old dst count, tmp2[1] <- number of staged chunks.
// Exclusive prefix of cnt[0..n) into offs (index order), returned total;
// every thread of the 1024-thread workgroup must call it.  Tiles of 8192
// counts: each thread scans 8 contiguous counts, a wave-shuffle scan and the
// 16 wave totals combine the threads.
__device__ int block_scan_counts(const int32_t* cnt, int32_t* offs, int n, int* s_wtot)
{
    constexpr int kT = 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int carry = 0;
    for (int base = 0; base < n; base += (int)blockDim.x * kT) {
        const int b0 = base + (int)threadIdx.x * kT;
        int v[kT], sum = 0;
#pragma unroll
        for (int j = 0; j < kT; ++j) {
            v[j] = b0 + j < n ? cnt[b0 + j] : 0;
            sum += v[j];
        }
        const int incl = wave_incl_scan(sum);
        if (lane == 63) s_wtot[wave] = incl;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w = 0; w < nw; ++w) {
            const int c = s_wtot[w];
            if (w < wave) woff += c;
            tot += c;
        }
        __syncthreads();                      // s_wtot is rewritten by the next tile
        int acc = carry + woff + incl - sum;
#pragma unroll
        for (int j = 0; j < kT; ++j) {
            if (b0 + j < n) offs[b0 + j] = acc;
            acc += v[j];
        }
        carry += tot;
    }
    return carry;
}

// ---- append -------------------------------------------------------------------
// Two launches, both parallel over the staging segments: k_append_scan gives
// every segment its offset within a group of kScanGroup segments and every
// group its total; k_append_copy prefixes the group totals in LDS and places
// the staged entries at count + offset (parent slot order), then updates the
// count.  src_count fixes the number of staged segments.  The appended slots
// are deposited afterwards, with the rest, by the step's deposit launch.

__global__ __launch_bounds__(kScanGroup) void k_append_scan(const int32_t* src_count,
                                                            const int32_t* dst_count,
                                                            const int32_t* stage_cnt, int32_t* offs,
                                                            int32_t* totals, int32_t* tmp2,
                                                            int32_t max_segs)
{
    __shared__ int s_wtot[kScanGroup / 64];
    int nseg = (*src_count + kSeg - 1) / kSeg;
    if (nseg > max_segs) nseg = max_segs;
    const int c = blockIdx.x * kScanGroup + (int)threadIdx.x;
    const int v = c < nseg ? stage_cnt[c] : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    if (lane == 63) s_wtot[wave] = incl;
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < kScanGroup / 64; ++w) {
        const int t = s_wtot[w];
        if (w < wave) woff += t;
        tot += t;
    }
    if (c < nseg) offs[c] = woff + incl - v;
    if (threadIdx.x == 0) totals[blockIdx.x] = tot;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        tmp2[0] = *dst_count;
        tmp2[1] = nseg;
    }
}

// The segment of entry e of a workgroup's batch: the largest j in [0, kBlock)
// with pre[j] <= e (pre is the batch's exclusive entry prefix, ascending).
__device__ __forceinline__ int batch_segment(const int* pre, int e)
{
    int lo = 0, hi = kBlock;                   // pre[lo] <= e < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// kAppendSplit workgroups share a batch of kBlock staged segments: workgroup
// b copies every kAppendSplit-th run of kBlock entries of batch b / kAppendSplit,
// from run b % kAppendSplit on.  A batch lies inside one scan group
// (kScanGroup is a multiple of kBlock), so its destinations are contiguous,
// from the first segment's offset on; every thread copies whole entries (six
// floats), consecutive threads consecutive destinations.  One workgroup per
// batch left a few hundred workgroups, each waiting on its own loads, for a
// copy of a few hundred thousand entries.
constexpr int kAppendSplit = 8;

__global__ __launch_bounds__(kBlock) void k_append_copy(float* d0, float* d1, float* d2, float* d3,
                                                        float* d4, float* d5, int32_t cap,
                                                        int32_t* dst_count,
                                                        const float* s0, const float* s1,
                                                        const float* s2, const float* s3,
                                                        const float* s4, const float* s5,
                                                        const int32_t* stage_cnt,
                                                        const int32_t* offs, const int32_t* totals,
                                                        int32_t ngroups, const int32_t* tmp2,
                                                        unsigned long long* dropped)
{
    static_assert(kScanGroup % kBlock == 0, "a batch of segments lies inside one scan group");
    __shared__ int s_pre[kMaxScanGroups + 1];
    __shared__ int s_seg[kBlock + 1];
    __shared__ int s_wave[kBlock / 64];
    const int base = tmp2[0];
    const int nseg = tmp2[1];
    // full: everything staged is dropped; only workgroup 0 (the count and the
    // dropped tally) has work
    const int part = (int)blockIdx.x % kAppendSplit;
    if (base >= cap && blockIdx.x != 0) return;
    if ((int)blockIdx.x / kAppendSplit * kBlock >= nseg && blockIdx.x != 0) return;
    // exclusive prefix of the group totals (four per thread)
    {
        constexpr int kG = kMaxScanGroups / kBlock;
        int v[kG], sum = 0;
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            const int i = (int)threadIdx.x * kG + j;
            v[j] = i < ngroups ? totals[i] : 0;
            sum += v[j];
        }
        int tot;
        int acc = block_excl_scan(sum, s_wave, tot);
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            s_pre[threadIdx.x * kG + j] = acc;
            acc += v[j];
        }
        if (threadIdx.x == 0) {
            s_pre[kMaxScanGroups] = tot;
            if (blockIdx.x == 0) {
                const int room = cap - base;
                const int app = tot < room ? tot : room;
                *dst_count = base + app;
                if (tot > app && dropped) atomicAdd(dropped, (unsigned long long)(tot - app));
            }
        }
        __syncthreads();
    }
    if (base >= cap) return;                       // full: everything staged was dropped
    for (int bt = (int)blockIdx.x / kAppendSplit; bt * kBlock < nseg; bt += (int)gridDim.x / kAppendSplit) {
        const int c0 = bt * kBlock;
        const int c = c0 + (int)threadIdx.x;
        const int cnt = c < nseg ? stage_cnt[c] : 0;
        int tot;
        const int pre = block_excl_scan(cnt, s_wave, tot);
        s_seg[threadIdx.x] = pre;
        if (threadIdx.x == 0) s_seg[kBlock] = tot;
        const int first = base + s_pre[c0 / kScanGroup] + offs[c0];
        __syncthreads();
        if (first < cap) {
            for (int e = part * kBlock + (int)threadIdx.x; e < tot; e += kBlock * kAppendSplit) {
                const int dst = first + e;
                if (dst >= cap) break;
                const int j = batch_segment(s_seg, e);
                const int src = (c0 + j) * kSeg + (e - s_seg[j]);
                d0[dst] = s0[src]; d1[dst] = s1[src]; d2[dst] = s2[src];
                d3[dst] = s3[src]; d4[dst] = s4[src]; d5[dst] = s5[src];
            }
        }
        __syncthreads();                           // s_seg is rewritten by the next batch
    }
}

// The append in one launch, for at most kOneMaxSegs staging segments: the
// kOneSplit workgroups of a batch of kBlock segments read the counts of all
// earlier segments themselves (at most kOneMaxSegs words, from L2) for the
// batch's first destination, so no scan launch precedes the copy.  The
// destination count is updated by the workgroup that finishes last, found by
// tickets in two levels: every workgroup takes its batch's ticket (sync[2 +
// batch]) after all its threads have used the old count, the batch's last
// takes the launch's ticket (sync[0]), and the batch of the last staged
// segment publishes the total (sync[1]) before its ticket.  The sync words are
// shared only through device-scope atomics (no fence: an agent-scope fence
// writes the L2 back); the copies and the count reach the next launch at the
// kernel boundary.  Each word is reset by the last to take it, so the launch
// leaves them as it found them (zero before the first launch).
constexpr int kOneMaxSegs = 32768;
constexpr int kOneSyncWords = 2 + kOneMaxSegs / kBlock;
static_assert(kOneSyncWords == kAppendSyncWords, "the callers' sync words");
#ifndef GASX_ONE_SPLIT
#define GASX_ONE_SPLIT 4
#endif
constexpr int kOneSplit = GASX_ONE_SPLIT;

__global__ __launch_bounds__(kBlock) void k_append_one(float* d0, float* d1, float* d2, float* d3, float* d4,
                                                       float* d5, int32_t cap, int32_t* dst_count,
                                                       const float* s0, const float* s1, const float* s2,
                                                       const float* s3, const float* s4, const float* s5,
                                                       const int32_t* stage_cnt, const int32_t* src_count,
                                                       int32_t max_segs, uint32_t* sync,
                                                       unsigned long long* dropped)
{
    __shared__ int s_seg[kBlock + 1];
    __shared__ int s_wave[kBlock / 64];
    const int tid = (int)threadIdx.x;
    const int base = *dst_count;
    const int nseg = (int)min((int64_t)max_segs, ((int64_t)*src_count + kSeg - 1) / kSeg);
    const int bt = (int)blockIdx.x / kOneSplit, part = (int)blockIdx.x % kOneSplit;
    const int c0 = bt * kBlock;
    const bool live = c0 < nseg;                   // workgroup-uniform
    const bool last_batch = live && c0 + kBlock >= nseg;
    // staged entries of the segments before the batch (all below nseg),
    // sixteen loads in flight per thread (two rounds at kOneMaxSegs)
    int before = 0;
    if (live) {
        if ((reinterpret_cast<uintptr_t>(stage_cnt) & 15) == 0) {
            constexpr int kU = 16;
            const int4* c4 = reinterpret_cast<const int4*>(stage_cnt);
            const int nw = c0 / 4;
            for (int i0 = 0; i0 < nw; i0 += kU * kBlock) {
                int4 v[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int i = i0 + u * kBlock + tid;
                    v[u] = i < nw ? c4[i] : make_int4(0, 0, 0, 0);
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) before += v[u].x + v[u].y + v[u].z + v[u].w;
            }
        } else {
            for (int i = tid; i < c0; i += kBlock) before += stage_cnt[i];
        }
    }
    const int cnt = live && c0 + tid < nseg ? stage_cnt[c0 + tid] : 0;
    int first, tot;
    block_excl_scan(before, s_wave, first);
    const int pre = block_excl_scan(cnt, s_wave, tot);
    s_seg[tid] = pre;
    if (tid == 0) s_seg[kBlock] = tot;
    __syncthreads();
    first += base;
    if (live && first < cap) {
        for (int e = part * kBlock + tid; e < tot; e += kBlock * kOneSplit) {
            const int dst = first + e;
            if (dst >= cap) break;
            const int j = batch_segment(s_seg, e);
            const int src = (c0 + j) * kSeg + (e - s_seg[j]);
            d0[dst] = s0[src]; d1[dst] = s1[src]; d2[dst] = s2[src];
            d3[dst] = s3[src]; d4[dst] = s4[src]; d5[dst] = s5[src];
        }
    }
    __syncthreads();                               // every thread has used the old count
    if (tid == 0) {
        if (part == 0 && last_batch) {
            const uint32_t prev = atomicExch(&sync[1], (uint32_t)(first - base + tot));
            asm volatile("" ::"v"(prev));          // performed before the ticket is taken
        }
        if (atomicAdd(&sync[2 + bt], 1u) != kOneSplit - 1) return;   // not the batch's last
        atomicExch(&sync[2 + bt], 0u);
        const uint32_t ticket = atomicAdd(&sync[0], 1u);
        if (ticket == gridDim.x / kOneSplit - 1) {  // every workgroup has read the old count
            const int total = nseg > 0 ? (int)atomicOr(&sync[1], 0u) : 0;
            const int room = cap > base ? cap - base : 0;
            const int app = total < room ? total : room;
            *dst_count = base + app;
            if (total > app && dropped) atomicAdd(dropped, (unsigned long long)(total - app));
            atomicExch(&sync[1], 0u);
            atomicExch(&sync[0], 0u);
        }
    }
}

// scan_tmp: max_segs segment offsets followed by the group totals; tmp2: two
// words of the two-launch path; sync: kOneSyncWords words of the one-launch
// path, zero before the first append and left zero by every one-launch append
hipError_t launch_append(float* const dst[6], int32_t* dst_count, int32_t cap,
                         float* const stage[6], const int32_t* stage_cnt,
                         const int32_t* src_count, int32_t* scan_tmp, int32_t* tmp2, uint32_t* sync,
                         int32_t max_segs, unsigned long long* dropped, hipStream_t s)
{
    if (sync && max_segs >= 1 && max_segs <= kOneMaxSegs) {
        const int blocks = (max_segs + kBlock - 1) / kBlock * kOneSplit;
        hipLaunchKernelGGL(k_append_one, dim3(blocks), dim3(kBlock), 0, s, dst[0], dst[1], dst[2], dst[3], dst[4],
                           dst[5], cap, dst_count, stage[0], stage[1], stage[2], stage[3], stage[4], stage[5],
                           stage_cnt, src_count, max_segs, sync, dropped);
        return hipGetLastError();
    }
    const int ngroups = (max_segs + kScanGroup - 1) / kScanGroup;
    if (ngroups < 1 || ngroups > kMaxScanGroups) return hipErrorInvalidValue;
    int32_t* totals = scan_tmp + max_segs;
    hipLaunchKernelGGL(k_append_scan, dim3(ngroups), dim3(kScanGroup), 0, s, src_count, dst_count,
                       stage_cnt, scan_tmp, totals, tmp2, max_segs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // kAppendSplit workgroups per kBlock staged segments (those past the
    // staged count return at once)
    const int blocks = (max_segs + kBlock - 1) / kBlock * kAppendSplit;
    hipLaunchKernelGGL(k_append_copy, dim3(blocks > 0 ? blocks : kAppendSplit), dim3(kBlock), 0, s, dst[0], dst[1], dst[2],
                       dst[3], dst[4], dst[5], cap, dst_count, stage[0], stage[1], stage[2], stage[3],
                       stage[4], stage[5], stage_cnt, scan_tmp, totals, ngroups, tmp2, dropped);
    return hipGetLastError();
}

// ---- batched append: workgroup r scans run r's staged chunks -----------------
// chunk_start (written by the particle launch) delimits each run's chunks in
// the staging area; tmp_run[2r] <- run r's old count.
__global__ __launch_bounds__(1024) void k_append_scan_batch(int32_t* counts, const BatchRun* runs,
                                                            const int32_t* stage_cnt,
                                                            const int32_t* chunk_start, int32_t* offs,
                                                            int32_t* tmp_run,
                                                            unsigned long long* run_counters,
                                                            unsigned long long* dropped, int drop_idx)
{
    __shared__ int s_wtot[16];
    const int r = blockIdx.x;
    const int c0 = chunk_start[r], nch = chunk_start[r + 1] - c0;
    const int n = counts[r];
    const int total = block_scan_counts(stage_cnt + c0, offs + c0, nch, s_wtot);
    if (threadIdx.x == 0) {
        const int room = runs[r].cap - n;
        const int app = total < room ? total : room;
        tmp_run[2 * r] = n;
        counts[r] = n + app;
        if (total > app) {
            atomicAdd(&run_counters[(size_t)r * kNCounters + drop_idx], (unsigned long long)(total - app));
            atomicAdd(dropped, (unsigned long long)(total - app));
        }
    }
}

// Workgroup b copies the entries of staged chunks [b*kBlock, (b+1)*kBlock),
// which may belong to several runs: each entry finds its chunk (batch prefix)
// and the chunk its run (chunk_start), and lands at the run's old count +
// its offset in the run.
__global__ __launch_bounds__(kBlock) void k_append_copy_batch(float* d0, float* d1, float* d2, float* d3,
                                                              float* d4, float* d5, const BatchRun* runs,
                                                              int nruns, const float* s0, const float* s1,
                                                              const float* s2, const float* s3,
                                                              const float* s4, const float* s5,
                                                              const int32_t* stage_cnt,
                                                              const int32_t* chunk_start,
                                                              const int32_t* offs, const int32_t* tmp_run)
{
    __shared__ int s_cs[kMaxRuns + 1];
    __shared__ int s_seg[kBlock + 1];
    __shared__ int s_wave[kBlock / 64];
    for (int r = threadIdx.x; r <= nruns; r += blockDim.x) s_cs[r] = chunk_start[r];
    __syncthreads();
    const int total = s_cs[nruns];
    for (int bt = blockIdx.x; bt * kBlock < total; bt += gridDim.x) {   // workgroup-uniform
        const int c0 = bt * kBlock;
        const int c = c0 + (int)threadIdx.x;
        const int cnt = c < total ? stage_cnt[c] : 0;
        int tot;
        const int pre = block_excl_scan(cnt, s_wave, tot);
        s_seg[threadIdx.x] = pre;
        if (threadIdx.x == 0) s_seg[kBlock] = tot;
        __syncthreads();
        for (int e = (int)threadIdx.x; e < tot; e += kBlock) {
            const int j = batch_segment(s_seg, e);
            const int ch = c0 + j;
            int lo = 0, hi = nruns;                // s_cs[lo] <= ch < s_cs[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_cs[mid] <= ch) lo = mid; else hi = mid;
            }
            const BatchRun br = runs[lo];
            const int local = tmp_run[2 * lo] + offs[ch] + (e - s_seg[j]);
            if (local < br.cap) {
                const int64_t dst = br.base + local;
                const int src = ch * kSeg + (e - s_seg[j]);
                d0[dst] = s0[src]; d1[dst] = s1[src]; d2[dst] = s2[src];
                d3[dst] = s3[src]; d4[dst] = s4[src]; d5[dst] = s5[src];
            }
        }
        __syncthreads();                           // s_seg is rewritten by the next batch
    }
}

hipError_t launch_append_batch(float* const pool[6], int32_t* counts, const BatchRun* runs, int nruns,
                               float* const stage[6], const int32_t* stage_cnt,
                               const int32_t* chunk_start, int32_t* offs, int32_t* tmp_run,
                               int32_t total_chunks, unsigned long long* run_counters,
                               unsigned long long* dropped, int drop_idx, hipStream_t s)
{
    hipLaunchKernelGGL(k_append_scan_batch, dim3(nruns), dim3(1024), 0, s, counts, runs, stage_cnt,
                       chunk_start, offs, tmp_run, run_counters, dropped, drop_idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int batches = (total_chunks + kBlock - 1) / kBlock;
    const int blocks = batches < 4096 ? (batches > 0 ? batches : 1) : 4096;
    hipLaunchKernelGGL(k_append_copy_batch, dim3(blocks), dim3(kBlock), 0, s, pool[0], pool[1], pool[2],
                       pool[3], pool[4], pool[5], runs, nruns, stage[0], stage[1], stage[2], stage[3],
                       stage[4], stage[5], stage_cnt, chunk_start, offs, tmp_run);
    return hipGetLastError();
}

// ---- cross sections (find_CS0.m) --------------------------------------------
__global__ void k_cross_sections(const DevTables* dt, const float* e_kev, int64_t n, float* se,
                                 float* si, float* sx)
{
    const SmallTabs* t = &dt->s;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float T = e_kev[i] * 1000.0f;
        const float nanv = __int_as_float(0x7fc00000);
        float a, b;
        // elastic: find_CS0.m:16-25 (NaN outside 1 eV .. 20 keV)
        if (e_kev[i] >= 0.05f)
            a = (T >= t->el_e[0] && T <= t->el_e[t->n_el - 1]) ? interp_in(t->el_e, t->el_s, t->n_el, T) : nanv;
        else
            a = (T >= t->sel_e[0] && T <= t->sel_e[t->n_sel - 1]) ? interp_in(t->sel_e, t->sel_s, t->n_sel, T) : nanv;
        b = (T >= t->ion_e[0] && T <= t->ion_e[t->n_ion - 1]) ? interp_in(t->ion_e, t->ion_s, t->n_ion, T) : nanv;
        float s = 0.0f;
        for (int j = 0; j < t->n_exc; ++j) {
            const float wr = t->exc_w[j] / T;
            const float bb = powf(wr, t->exc_om[j]);
            const float cc = powf(1.0f - powf(wr, t->exc_ga[j]), t->exc_nu[j]);
            s += t->exc_q[j] * bb * cc;
        }
        se[i] = a; si[i] = b; sx[i] = s;
    }
}

hipError_t launch_cross_sections(const DevTables* t, const float* e_kev, int64_t n, float* se,
                                 float* si, float* sx, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_cross_sections, dim3((unsigned)blocks), dim3(256), 0, s, t, e_kev, n, se,
                       si, sx);
    return hipGetLastError();
}

// ---- dispatch -----------------------------------------------------------------
template <uint32_t F>
static hipError_t launch_t(const Geom& g, const PartArgs& pa, const CollArgs& ca, int blocks,
                           hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if constexpr ((F & kBatch) != 0)
        hipExtLaunchKernelGGL(k_particles_batch<F>, dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa, ca);
    else
        hipExtLaunchKernelGGL(k_particles<F>, dim3(blocks), dim3(kBlock), 0, s, ev0, ev1, 0, g, pa, ca);
    return hipGetLastError();
}

template <uint32_t F>
static int occupancy_t()
{
    static int cached = 0;          // per instantiation; the device set is homogeneous
    if (cached == 0) {
        int nb = 0;
        hipError_t e;
        if constexpr ((F & kBatch) != 0)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_particles_batch<F>, kBlock, 0);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_particles<F>, kBlock, 0);
        if (e != hipSuccess || nb < 1) nb = 1;
        cached = nb;
    }
    return cached;
}

int particle_occupancy(uint32_t flags)
{
    constexpr uint32_t P = kPush, Co = kCollide, S = kStageSec, I = kStageIon, B = kBatch, DS = kDriftSec,
                       D2 = kDrift2;
    switch (flags) {
    case P: return occupancy_t<P>();
    case Co | S: return occupancy_t<Co | S>();
    case Co | S | I: return occupancy_t<Co | S | I>();
    case P | Co: return occupancy_t<P | Co>();
    case P | Co | S: return occupancy_t<P | Co | S>();
    case P | Co | I: return occupancy_t<P | Co | I>();
    case P | Co | S | I: return occupancy_t<P | Co | S | I>();
    case P | D2: return occupancy_t<P | D2>();
    case P | Co | D2: return occupancy_t<P | Co | D2>();
    case P | Co | I | D2: return occupancy_t<P | Co | I | D2>();
    case P | Co | S | DS | D2: return occupancy_t<P | Co | S | DS | D2>();
    case P | Co | S | I | DS | D2: return occupancy_t<P | Co | S | I | DS | D2>();
    case B | P: return occupancy_t<B | P>();
    case B | P | Co: return occupancy_t<B | P | Co>();
    case B | P | Co | S: return occupancy_t<B | P | Co | S>();
    case B | P | D2: return occupancy_t<B | P | D2>();
    case B | P | Co | D2: return occupancy_t<B | P | Co | D2>();
    case B | P | Co | S | DS | D2: return occupancy_t<B | P | Co | S | DS | D2>();
    case B | P | Co | I: return occupancy_t<B | P | Co | I>();
    case B | P | Co | S | I: return occupancy_t<B | P | Co | S | I>();
    case B | P | Co | I | D2: return occupancy_t<B | P | Co | I | D2>();
    case B | P | Co | S | I | DS | D2: return occupancy_t<B | P | Co | S | I | DS | D2>();
    default: return 1;
    }
}

hipError_t launch_particles(uint32_t flags, const Geom& g, const PartArgs& pa, const CollArgs& ca,
                            int max_blocks, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    const int blocks = max_blocks > 0 ? max_blocks : 1;
    constexpr uint32_t P = kPush, Co = kCollide, S = kStageSec, I = kStageIon, B = kBatch, DS = kDriftSec,
                       D2 = kDrift2;
    switch (flags) {
    case P: return launch_t<P>(g, pa, ca, blocks, s, ev0, ev1);
    case Co | S: return launch_t<Co | S>(g, pa, ca, blocks, s, ev0, ev1);
    case Co | S | I: return launch_t<Co | S | I>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co: return launch_t<P | Co>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | S: return launch_t<P | Co | S>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | I: return launch_t<P | Co | I>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | S | I: return launch_t<P | Co | S | I>(g, pa, ca, blocks, s, ev0, ev1);
    // with the step's second half drift
    case P | D2: return launch_t<P | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | D2: return launch_t<P | Co | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | I | D2: return launch_t<P | Co | I | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | S | DS | D2: return launch_t<P | Co | S | DS | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case P | Co | S | I | DS | D2: return launch_t<P | Co | S | I | DS | D2>(g, pa, ca, blocks, s, ev0, ev1);
    // batched ensemble
    case B | P: return launch_t<B | P>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co: return launch_t<B | P | Co>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | S: return launch_t<B | P | Co | S>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | D2: return launch_t<B | P | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | D2: return launch_t<B | P | Co | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | S | DS | D2: return launch_t<B | P | Co | S | DS | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | I: return launch_t<B | P | Co | I>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | S | I: return launch_t<B | P | Co | S | I>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | I | D2: return launch_t<B | P | Co | I | D2>(g, pa, ca, blocks, s, ev0, ev1);
    case B | P | Co | S | I | DS | D2: return launch_t<B | P | Co | S | I | DS | D2>(g, pa, ca, blocks, s, ev0, ev1);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace gasx
