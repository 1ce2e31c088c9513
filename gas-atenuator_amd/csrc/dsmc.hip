// DSMC neutral-gas step: the reference's molecule transport and collisions
// (main2.m:247-271 with free_motion.m, boundary_conditions_specular.m,
// put_into_cells_symmetry.m, calculate_VC.m and collision_HS_internal.m) on
// the reference's (r, phi, z) cell map (cell_structure.m:6-22), and the gas
// density and temperature that the electron step reads (particle_density0.m,
// count_T_numbensity.m).
//
//   k_dsmc_move<kMove>   per molecule: free flight, the r = Rp wall, the
//                        periodic phi faces, the z walls (FP64, stored FP32),
//                        then the cell of the stored position as a sort key
//   rocPRIM radix sort   (cell, slot) pairs: each cell's molecules in
//                        ascending slot order, the order of the reference's
//                        `a = find(INDEX == indexnum)`
//   k_dsmc_bounds        first and last sorted position of every cell
//   k_dsmc_rare          the molecules near a wall or a face, queued by the
//                        molecule pass, through the whole reference path
//   k_dsmc_unpack        (reorder) the pass's 32-byte records gathered in
//                        sorted order into the molecule arrays
//   k_dsmc_collide       one workgroup per cell: the cell's no-time-counter
//                        attempts in batches of 256 threads.  An attempt's
//                        random numbers come from its own Philox counter
//                        (attempt, cell, step, block), so a batch can draw in
//                        parallel; a thread runs its attempt once no lower
//                        pending thread of the batch holds one of its two
//                        molecules (marks in an LDS hash table).  The result
//                        equals running the attempts one after another.
//   k_dsmc_cellsum /     per cell count and sum of v^2 (fixed reduction
//   k_dsmc_sample_out    order), then density and temperature per cell and
//                        per (ring, layer) of the map
//
// Arithmetic is FP64 with contraction off; positions, velocities and internal
// energies are stored FP32.  The oracle restates the same operations with
// FP32 storage, so the two agree bit for bit except where atan2, sin, cos or
// pow round their last FP64 bit differently.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "context.h"
#include "device_common.h"

using namespace gasx;
using namespace gasx_host;

namespace {

constexpr uint32_t kBlk0 = 0x50000000u;      // sub1, sub2, acceptance, scattering cosine
constexpr uint32_t kBlk1 = 0x50000001u;      // scattering azimuth
constexpr uint32_t kRedraw = 0x50000100u;    // + r: the r-th redraw of an equal pair
constexpr uint32_t kEt = 0x51000000u;        // + j: translational-energy rounds 2j, 2j+1
constexpr uint32_t kEtry = 0x52000000u;      // + j: internal-energy split rounds 2j, 2j+1
constexpr int kMaxRedraw = 4096;
constexpr int kMaxRounds = 65536;
constexpr int kCollThreads = 256;           // attempts per batch: one workgroup per cell
constexpr int kTabW = 4096;                  // LDS marks per cell (molecule index mod kTabW)
constexpr int kMoveBlocksPerCU = 5;          // persistent molecule-pass workgroups per CU (5 waves per SIMD)
constexpr double kPi = 3.141592653589793;    // MATLAB pi
constexpr double kKb = 1.38064853e-23;       // main2.m:5
constexpr int kNDsmcCounters = 7;            // attempts, collisions, SC > SCmax, boundary, r wall, z walls, wraps
constexpr int kCtrStride = 8;                // u64 words per counter slot
constexpr int kCtrSlots = 512;               // workgroup b adds into slot b % kCtrSlots (summed on read):
                                             // thousands of waves adding into one word serialise in L2
constexpr int64_t kMaxCells = 1ll << 24;

struct DMap {
    int32_t nr, nl, nxy;
    int64_t ncell;
    double rp, lp, phip, cp, sp;
    double rp2_lo;            // Rp^2 (1 - 1e-12): x^2 + y^2 below it is surely inside the r wall
    int32_t fast;             // 1: PHIp < pi, the clear-of-walls test below applies
    float inv_dr_f, inv_dz_f;
    const int32_t* pd;        // phi cells per ring
    const int32_t* sum0;      // exclusive prefix, nr + 1
    const double* dphi;       // per ring
    const float* inv_dphi_f;  // 1 / dphi per ring (estimates only)
    const double* vc;         // one cell's volume per ring
    const int32_t* ring_of;   // ring of each cell of a layer (nxy)
};

struct Mol {
    float *x, *y, *z, *vx, *vy, *vz, *ei;
    const int32_t* count;
};

__device__ __forceinline__ double du(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

__device__ __forceinline__ U4 draw(uint32_t a, uint32_t cell, uint32_t step, uint32_t blk, uint32_t k0,
                                   uint32_t k1)
{
    return philox4x32_10(a, cell, step, blk, k0, k1);
}

__device__ __forceinline__ uint32_t word(const U4& q, int i)
{
    return i == 0 ? q.a : (i == 1 ? q.b : (i == 2 ? q.c : q.d));
}

// put_into_cells_symmetry.m:4-16 of the stored position (1-based ceil
// indices clamped into the map), 0-based; ncell for an absorbed slot.
__device__ int64_t dsmc_cell(const DMap& m, double x, double y, double z)
{
#pragma clang fp contract(off)
    if (!(z == z)) return m.ncell;
    const double dz = m.lp / m.nl, dr = m.rp / m.nr;
    const double r = sqrt(x * x + y * y), phi = atan2(y, x);
    const double fz = ceil(z / dz), fr = ceil(r / dr);
    const int iz = fz < 1.0 ? 1 : (fz > m.nl ? m.nl : (int)fz);
    const int ir = fr < 1.0 ? 1 : (fr > m.nr ? m.nr : (int)fr);
    const double fp = ceil(phi / m.dphi[ir - 1]);
    const int pd = m.pd[ir - 1];
    const int ip = fp < 1.0 ? 1 : (fp > pd ? pd : (int)fp);
    return (int64_t)(iz - 1) * m.nxy + m.sum0[ir - 1] + ip - 1;
}

// free_motion.m:3-7 and boundary_conditions_specular.m:6-64 for one molecule
// (c: boundary events, r wall, z walls, phi wraps)
__device__ void dsmc_move_one(const DMap& m, double dt, double& x, double& y, double& z, double& vx, double& vy,
                              double& vz, int c[4])
{
#pragma clang fp contract(off)
    const double Rp = m.rp, Lp = m.lp, pm = m.phip;
    x = x + vx * dt; y = y + vy * dt; z = z + vz * dt;
    double r = sqrt(x * x + y * y), phi = atan2(y, x);
    if (r > Rp) {                                                  // specular_scattering_xy.m:5-40
        c[0]++; c[1]++;
        const double x0 = x - vx * dt, y0 = y - vy * dt;
        const double r0 = sqrt(x0 * x0 + y0 * y0);
        const double t1 = ((r - Rp) / (r - r0)) * dt;
        const double xi = x - vx * t1, yi = y - vy * t1, zi = z - vz * t1;
        const double th = atan2(yi, xi);
        const double ct = cos(th), s = sin(th);
        double vr = vx * ct + vy * s;
        const double vt = -vx * s + vy * ct;
        vr = -vr;
        vx = vr * ct - vt * s;
        vy = vr * s + vt * ct;
        x = xi + vx * t1; y = yi + vy * t1; z = zi + vz * t1;
        r = sqrt(x * x + y * y);
        phi = atan2(y, x);
    }
    while (phi < 0) {                                              // boundary_phi_0.m (as intended)
        c[0]++; c[3]++;
        const double rxy = sqrt(x * x + y * y);
        phi = phi + pm;
        x = rxy * cos(phi); y = rxy * sin(phi);
        const double nvx = vx * m.cp - vy * m.sp, nvy = vx * m.sp + vy * m.cp;
        vx = nvx; vy = nvy;
    }
    while (phi > pm) {                                             // boundary_phi_phimax.m
        c[0]++; c[3]++;
        const double rxy = sqrt(x * x + y * y);
        phi = phi - pm;
        x = rxy * cos(phi); y = rxy * sin(phi);
        const double nvx = vx * m.cp + vy * m.sp, nvy = -vx * m.sp + vy * m.cp;
        vx = nvx; vy = nvy;
    }
    if (z > Lp) {                                                  // specular_scattering_z.m:15-32
        c[0]++; c[2]++;
        const double z0 = z - vz * dt;
        const double t1 = (z - Lp) / (z - z0) * dt;
        const double xi = x - vx * t1, yi = y - vy * t1, zi = z - vz * t1;
        vz = -vz;
        x = xi + vx * t1; y = yi + vy * t1; z = zi + vz * t1;
    }
    if (z < 0) {                                                   // specular_scattering_z.m:17-32
        c[0]++; c[2]++;
        const double z0 = z - vz * dt;
        const double t1 = (0 - z) / (z0 - z) * dt;
        const double xi = x - vx * t1, yi = y - vy * t1, zi = z - vz * t1;
        vz = -vz;
        x = xi + vx * t1; y = yi + vy * t1; z = zi + vz * t1;
    }
}

__device__ __forceinline__ void wave_add_counters(unsigned long long* ctr, int k0, const int* v, int n)
{
    unsigned long long* g = ctr + (size_t)(blockIdx.x % kCtrSlots) * kCtrStride;
    for (int k = 0; k < n; ++k) {
        int s = v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if ((threadIdx.x & 63) == 0 && s) atomicAdd(&g[k0 + k], (unsigned long long)s);
    }
}

// An FP32 estimate e of an exact FP64 quotient whose ceil is wanted: when e
// lies clear of every integer by more than its error bound (a few 1e-7
// relative; the margin allows ten times that), ceil(e) is the exact ceil.
__device__ __forceinline__ bool clear_of_integers(float e)
{
    const float f = e - floorf(e);
    const float mg = 4e-6f * fabsf(e) + 1e-6f;
    return f > mg && f < 1.0f - mg;
}

__device__ __forceinline__ int clamp_ceil(float e, int hi)
{
    const float c = ceilf(e);
    return c < 1.0f ? 1 : (c > (float)hi ? hi : (int)c);
}

// dsmc_cell from FP32 estimates of z/dz, r/dr and phi/dphi when all three are
// clear of integers (the same cell as the FP64 arithmetic, which runs
// otherwise).
__device__ int64_t dsmc_cell_fast(const DMap& m, float x, float y, float z)
{
    if (!(z == z)) return m.ncell;
    const float ez = z * m.inv_dz_f;
    const float er = __builtin_amdgcn_sqrtf(x * x + y * y) * m.inv_dr_f;
    if (clear_of_integers(ez) && clear_of_integers(er)) {
        const int iz = clamp_ceil(ez, m.nl), ir = clamp_ceil(er, m.nr);
        const float ep = atan2f(y, x) * m.inv_dphi_f[ir - 1];
        if (clear_of_integers(ep)) {
            const int ip = clamp_ceil(ep, m.pd[ir - 1]);
            return (int64_t)(iz - 1) * m.nxy + m.sum0[ir - 1] + ip - 1;
        }
    }
    return dsmc_cell(m, x, y, z);
}

// Where the molecule pass puts its results: in place (SoA), or as 32-byte
// records (x, y, z, vx | vy, vz, ei, 0) that the reorder gathers by the
// sorted slot list (one 32-byte read per molecule instead of seven 4-byte
// reads from seven arrays).
struct MoveOut {
    float4* rec;              // null: in place
    uint32_t *keys, *vals;
    int32_t* rq;              // queued slots of molecules near a wall or a face
    int32_t* rq_n;
};

__device__ __forceinline__ void put_molecule(const Mol& mo, const MoveOut& o, int i, float x, float y, float z,
                                             float vx, float vy, float vz, bool vchanged)
{
    if (o.rec) {
        o.rec[2 * (size_t)i] = make_float4(x, y, z, vx);
        o.rec[2 * (size_t)i + 1] = make_float4(vy, vz, mo.ei[i], 0.0f);
    } else {
        mo.x[i] = x; mo.y[i] = y; mo.z[i] = z;
        if (vchanged) { mo.vx[i] = vx; mo.vy[i] = vy; mo.vz[i] = vz; }
    }
}

// Molecule pass: (free flight + walls) and the sort key of every slot below
// the count.  keys = cell (ncell for an absorbed slot), vals = slot.  A
// molecule whose free flight ends clear of the r wall, both phi faces and
// both z walls (tests with margins far above the rounding of the
// reference's atan2 / sqrt comparisons) is exactly what the reference's
// boundary code would leave; the others (about 3% at main2's step) are
// queued (per wave in LDS, flushed to a global list 64 at a time) for
// k_dsmc_rare, so the wall and face code never runs with one lane busy.
template <bool kMove>
__global__ __launch_bounds__(256) void k_dsmc_move(Mol mo, DMap m, double dt, MoveOut o)
{
#pragma clang fp contract(off)
    constexpr int kU = 4;                  // molecules per lane per pass: their loads are issued together
    __shared__ int s_q[4][128];
    const int n = *mo.count;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int* q = s_q[wv];
    int qn = 0;
    const long long stride = (long long)gridDim.x * blockDim.x * kU;
    for (long long base = ((long long)blockIdx.x * blockDim.x + wv * 64) * kU; base < n; base += stride) {
        float fx[kU], fy[kU], fz[kU], fvx[kU], fvy[kU], fvz[kU], fe[kU];
#pragma unroll
        for (int j = 0; j < kU; ++j) {
            const long long i = base + j * 64 + lane;
            fx[j] = fy[j] = fz[j] = fvx[j] = fvy[j] = fvz[j] = fe[j] = 0.0f;
            if (i < n) {
                fx[j] = mo.x[i]; fy[j] = mo.y[i]; fz[j] = mo.z[i];
                fvx[j] = mo.vx[i]; fvy[j] = mo.vy[i]; fvz[j] = mo.vz[i];
                if (o.rec) fe[j] = mo.ei[i];
            }
        }
#pragma unroll
        for (int j = 0; j < kU; ++j) {
            const int i = (int)(base + j * 64 + lane);
            bool rare = false;
            if (base + j * 64 + lane < n) {
                float x1 = fx[j], y1 = fy[j], z1 = fz[j];
                if (kMove && z1 == z1) {
                    const double x = (double)x1 + (double)fvx[j] * dt;
                    const double y = (double)y1 + (double)fvy[j] * dt;
                    const double z = (double)z1 + (double)fvz[j] * dt;
                    const bool clear = m.fast && x * x + y * y < m.rp2_lo && y > 0.0 &&
                                       y * m.cp - x * m.sp < -1e-9 * (fabs(x) + fabs(y)) && z > 0.0 && z < m.lp;
                    if (clear) {
                        x1 = (float)x; y1 = (float)y; z1 = (float)z;
                    } else {
                        rare = true;
                    }
                }
                if (!rare) {
                    if (o.rec) {
                        o.rec[2 * (size_t)i] = make_float4(x1, y1, z1, fvx[j]);
                        o.rec[2 * (size_t)i + 1] = make_float4(fvy[j], fvz[j], fe[j], 0.0f);
                    } else if (kMove) {
                        mo.x[i] = x1; mo.y[i] = y1; mo.z[i] = z1;
                    }
                    o.keys[i] = (uint32_t)dsmc_cell_fast(m, x1, y1, z1);
                    o.vals[i] = (uint32_t)i;
                }
            }
            if (kMove) {
                const unsigned long long mask = __ballot(rare);
                if (mask) {
                    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
                    if (rare) q[qn + rank] = i;
                    qn += __popcll(mask);
                    if (qn >= 64) {
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                        int b = 0;
                        if (lane == 0) b = atomicAdd(o.rq_n, 64);
                        b = __shfl(b, 0, 64);
                        o.rq[b + lane] = q[lane];
                        const int rest = qn - 64;
                        const int carry = lane < rest ? q[lane + 64] : 0;
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                        if (lane < rest) q[lane] = carry;
                        qn = rest;
                    }
                }
            }
        }
    }
    if (kMove && qn > 0) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        int b = 0;
        if (lane == 0) b = atomicAdd(o.rq_n, qn);
        b = __shfl(b, 0, 64);
        if (lane < qn) o.rq[b + lane] = q[lane];
    }
}

// The whole reference path for the queued molecules (dense: one per thread).
__global__ __launch_bounds__(256) void k_dsmc_rare(Mol mo, DMap m, double dt, MoveOut o, unsigned long long* ctr)
{
#pragma clang fp contract(off)
    const int nq = *o.rq_n;
    int c[4] = {0, 0, 0, 0};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nq; k += gridDim.x * blockDim.x) {
        const int j = o.rq[k];
        double x = mo.x[j], y = mo.y[j], z = mo.z[j], vx = mo.vx[j], vy = mo.vy[j], vz = mo.vz[j];
        dsmc_move_one(m, dt, x, y, z, vx, vy, vz, c);
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        put_molecule(mo, o, j, fx, fy, fz, (float)vx, (float)vy, (float)vz, true);
        o.keys[j] = (uint32_t)dsmc_cell_fast(m, fx, fy, fz);
        o.vals[j] = (uint32_t)j;
    }
    wave_add_counters(ctr, 3, c, 4);
}

__global__ void k_dsmc_zero(int32_t* a) { if (threadIdx.x == 0 && blockIdx.x == 0) *a = 0; }

// First and last sorted position of each cell; live = molecules with a cell.
__global__ void k_dsmc_bounds(const uint32_t* keys, const int32_t* count, uint32_t ncell, int32_t* start,
                              int32_t* end, int32_t* live)
{
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        if (k >= ncell) {
            if (i == 0) *live = 0;
            else if (keys[i - 1] < ncell) *live = i;
            continue;
        }
        if (i == 0 || keys[i - 1] != k) start[k] = i;
        if (i == n - 1 || keys[i + 1] != k) {
            end[k] = i + 1;
            if (i == n - 1) *live = n;
        }
    }
}

__global__ void k_dsmc_fill(int32_t* a, int32_t* b, int64_t n, int32_t* live)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = 0;
        b[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *live = 0;
}

// (reorder) the records in sorted order: positions into the molecule arrays,
// (vx, vy, vz, ei) into 16-byte records that the collisions read and write
// (one cache line per molecule touched instead of four; k_dsmc_collide copies
// them into the molecule arrays cell by cell); count = live
__global__ void k_dsmc_unpack(Mol mo, int32_t* count, const float4* rec, const uint32_t* vals, const int32_t* live,
                              float4* vrec)
{
    const int n = *live;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const size_t s = vals[i];
        const float4 a = rec[2 * s], b = rec[2 * s + 1];
        mo.x[i] = a.x; mo.y[i] = a.y; mo.z[i] = a.z;
        vrec[i] = make_float4(a.w, b.x, b.y, b.z);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

struct CollArgsD {
    DMap m;
    float *vx, *vy, *vz, *ei;
    float4* vrec;             // (vx, vy, vz, ei) per slot when the molecules are in cell order
    const int32_t *start, *end;
    const uint32_t* vals;     // sorted slots (null: the molecules are in cell order)
    double sc_max, fn, dt, d2, mass;
    double k1, a1, k2, b1, c2, h;
    float k1f, a1f, k2f, b1f, c2f, hf;   // the same in FP32, for the acceptance estimates
    double alpha;
    int32_t vhs, min_att;
    uint32_t step, key0, key1;
    unsigned long long* ctr;
};

// x^e for x > 0 by exp(e ln x) (argument reduction by powers of 2 and ln 2,
// atanh series for ln on [1/sqrt2, sqrt2), Taylor series for exp on
// |r| <= ln2/2): within 2e-15 relative of the C library's pow (checked on
// 2e6 points, x in [1e-12, 2.3], e in {0.33, 0.76, 1.5}), in far fewer
// registers than the library pow.  Used only to decide the energy
// acceptances that the FP32 estimate leaves open; it decides them as pow
// does unless the uniform lies within 2e-15 of the bound.
__device__ double pow_pos(double x, double e)
{
#pragma clang fp contract(off)
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : __builtin_nan("");
    int k = 0;
    double m = frexp(x, &k);                    // x = m 2^k, m in [0.5, 1)
    if (m < 0.7071067811865476) { m = m * 2.0; k -= 1; }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double t = 1.0 / 25.0;                      // ln m = 2 s (1 + s^2/3 + s^4/5 + ...)
#pragma unroll
    for (int j = 23; j >= 1; j -= 2) t = 1.0 / j + s2 * t;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double lnm = 2.0 * s * t;
    const double y_hi = e * ((double)k * ln2_hi), y_lo = e * ((double)k * ln2_lo + lnm);
    const double y = y_hi + y_lo;
    const double n = rint(y * 1.4426950408889634);
    const double r = (y_hi - n * ln2_hi) + (y_lo - n * ln2_lo);
    double q = 1.0;                             // e^r = 1 + r (1 + r/2 (1 + r/3 (...)))
#pragma unroll
    for (int j = 20; j >= 1; --j) q = 1.0 + r * q * (1.0 / j);
    return ldexp(q, (int)n);
}

__device__ __forceinline__ double dpow(double x, double e) { return e == 1.0 ? x : (e == 0.0 ? 1.0 : pow_pos(x, e)); }

// u < c x^a (1 - x)^b with x = num / den, decided from an FP32 estimate of
// the right side when u is clear of it by far more than the estimate's error
// (a few 1e-7 relative), else by the FP64 expression (the same decision).
__device__ __forceinline__ float pow_est(float x, float e)
{
    return e == 1.0f ? x : (e == 0.0f ? 1.0f : __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x)));
}

__device__ __forceinline__ int decide(float u, float est)
{
    if (!(est == est)) return -1;
    if (u < est * (1.0f - 1e-4f) - 1e-7f) return 1;
    if (u > est * (1.0f + 1e-4f) + 1e-7f) return 0;
    return -1;
}

// collision_HS_internal.m:38-54: does the pair with stored values v[0..7] =
// (vx, vy, vz, ei) of i1, then of i2, collide?
__device__ bool dsmc_accept(const CollArgsD& A, const U4& o0, const float* v, int& n_sc)
{
#pragma clang fp contract(off)
    const double ur = (double)v[0] - (double)v[4], vr = (double)v[1] - (double)v[5], wr = (double)v[2] - (double)v[6];
    const double cr = sqrt(ur * ur + vr * vr + wr * wr);
    const double SC = cr * kPi * A.d2;
    if (SC > A.sc_max) ++n_sc;
    return du(o0.c) < SC / A.sc_max;
}

// collision_HS_internal.m:56-126 for the accepted pair (i1, i2) of attempt a
// (b: the attempt's scattering-cosine word).  kVhs: alpha = 1 (isotropic
// scattering, :91-97), else VSS (:98-113).
template <bool kVhs>
__device__ void dsmc_exchange(const CollArgsD& A, uint32_t a, uint32_t cell, uint32_t b, int i1, int i2,
                              const float* v, int& n_coll)
{
#pragma clang fp contract(off)
    const double vx1 = v[0], vy1 = v[1], vz1 = v[2], e1 = v[3];
    const double vx2 = v[4], vy2 = v[5], vz2 = v[6], e2 = v[7];
    double ur = vx1 - vx2, vr = vy1 - vy2, wr = vz1 - vz2;
    double cr = sqrt(ur * ur + vr * vr + wr * wr);
    const double um = (vx1 + vx2) / 2, vm = (vy1 + vy2) / 2, wm = (vz1 + vz2) / 2;
    const double Ec = 0.25 * A.mass * (cr * cr) + e1 + e2;
    double Et = 0.0;
    U4 q{};
    for (int j = 0; j < kMaxRounds; ++j) {                          // :63-70
        if ((j & 1) == 0) q = draw(a, cell, A.step, kEt + (uint32_t)(j >> 1), A.key0, A.key1);
        const double ue = du(word(q, 2 * (j & 1))), ua = du(word(q, 2 * (j & 1) + 1));
        Et = Ec * ue;
        const float est = pow_est(A.k1f * (float)ue, A.a1f) * pow_est(A.k2f * (float)(1.0 - ue), A.b1f);
        int acc = decide((float)ua, est);
        if (acc < 0) {
            const double pr = dpow(A.k1 * (Et / Ec), A.a1) * dpow(A.k2 * (1 - Et / Ec), A.b1);
            acc = ua < pr ? 1 : 0;
        }
        if (acc) break;
    }
    const double Ei = Ec - Et;
    double etry = 0.0;
    for (int j = 0; j < kMaxRounds; ++j) {                          // :74-83
        if ((j & 1) == 0) q = draw(a, cell, A.step, kEtry + (uint32_t)(j >> 1), A.key0, A.key1);
        const double ue = du(word(q, 2 * (j & 1))), ua = du(word(q, 2 * (j & 1) + 1));
        etry = Ei * ue;
        int acc;
        if (A.h == 0.0) {
            acc = ua < A.c2 ? 1 : 0;                                // (x)^0 (1-x)^0 = 1 exactly
        } else {
            const float est = A.c2f * pow_est((float)ue, A.hf) * pow_est((float)(1.0 - ue), A.hf);
            acc = decide((float)ua, est);
            if (acc < 0) {
                const double pr = A.c2 * dpow(etry / Ei, A.h) * dpow(1 - etry / Ei, A.h);
                acc = ua < pr ? 1 : 0;
            }
        }
        if (acc) break;
    }
    const double cri = 2 * sqrt(Et / A.mass);                       // :86-90
    ur = ur * cri / cr; vr = vr * cri / cr; wr = wr * cri / cr;
    cr = cri;
    const U4 o1 = draw(a, cell, A.step, kBlk1, A.key0, A.key1);
    double urf, vrf, wrf;
    if (kVhs) {                                                     // :91-97
        const double B = 2 * du(b) - 1;
        const double Aa = sqrt(1 - B * B);
        const double C = 2 * kPi * du(o1.a);
        urf = cr * B;
        vrf = cr * Aa * cos(C);
        wrf = cr * Aa * sin(C);
    } else {                                                        // :98-113
        const double B = 2 * pow(du(b), 1 / A.alpha) - 1;
        const double Aa = sqrt(1 - B * B);
        const double C = 2 * kPi * du(o1.a);
        const double S = sin(C), CC = cos(C);
        const double D = sqrt(vr * vr + wr * wr);
        if (D > 1e-6) {
            urf = B * ur + Aa * S * D;
            vrf = B * vr + Aa * (cr * wr * CC - ur * vr * S) / D;
            wrf = B * wr - Aa * (cr * vr * CC + ur * wr * S) / D;   // Bird's sign (the reference's - loses energy)
        } else {
            urf = B * ur;
            vrf = Aa * CC * ur;
            wrf = Aa * S * ur;
        }
    }
    // :84-85 and :115-120
    const float4 n1 = make_float4((float)(um + 0.5 * urf), (float)(vm + 0.5 * vrf), (float)(wm + 0.5 * wrf),
                                  (float)etry);
    const float4 n2 = make_float4((float)(um - 0.5 * urf), (float)(vm - 0.5 * vrf), (float)(wm - 0.5 * wrf),
                                  (float)(Ei - etry));
    if (A.vrec) {
        A.vrec[i1] = n1;
        A.vrec[i2] = n2;
    } else {
        A.vx[i1] = n1.x; A.vy[i1] = n1.y; A.vz[i1] = n1.z; A.ei[i1] = n1.w;
        A.vx[i2] = n2.x; A.vy[i2] = n2.y; A.vz[i2] = n2.z; A.ei[i2] = n2.w;
    }
    ++n_coll;
}

__device__ __forceinline__ void load_pair(const CollArgsD& A, int i1, int i2, float* v)
{
    if (A.vrec) {
        const float4 a = A.vrec[i1], b = A.vrec[i2];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
        v[0] = A.vx[i1]; v[1] = A.vy[i1]; v[2] = A.vz[i1]; v[3] = A.ei[i1];
        v[4] = A.vx[i2]; v[5] = A.vy[i2]; v[6] = A.vz[i2]; v[7] = A.ei[i2];
    }
}

// (cell order) the cell's (vx, vy, vz, ei) records into the molecule arrays
__device__ __forceinline__ void copy_out_cell(const CollArgsD& A, int st, int n)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const float4 a = A.vrec[st + k];
        A.vx[st + k] = a.x; A.vy[st + k] = a.y; A.vz[st + k] = a.z; A.ei[st + k] = a.w;
    }
}

// One workgroup per cell (collision_HS_internal.m:11-136): the cell's
// attempts in batches of kCollThreads, one per thread.  A thread runs its
// attempt in the first round in which no pending thread of a lower index
// holds one of its two molecules (marks: molecule index mod kTabW, the lowest
// pending thread's index by LDS atomicMin; a shared mark only delays), so
// every attempt sees exactly the values the sequential loop would give it.
template <bool kVhs>
__global__ __launch_bounds__(kCollThreads) void k_dsmc_collide(CollArgsD A)
{
#pragma clang fp contract(off)
    __shared__ int tab[kTabW];
    // the round's accepted attempts, exchanged by threads 0 .. n-1 (dense)
    __shared__ int s_nacc;
    __shared__ uint32_t s_att[kCollThreads], s_b[kCollThreads];
    __shared__ int s_i1[kCollThreads], s_i2[kCollThreads];
    __shared__ float s_v[8][kCollThreads];
    const int t = threadIdx.x;
    const int64_t cell = blockIdx.x;
    const int st = A.start[cell], N = A.end[cell] - st;
    if (N <= 2) {                                                   // :26 (uniform over the workgroup)
        if (A.vrec) copy_out_cell(A, st, N);
        return;
    }
    for (int k = t; k < kTabW; k += kCollThreads) tab[k] = kCollThreads;
    const double vc = A.m.vc[A.m.ring_of[cell % A.m.nxy]];
    const double npair = floor(A.sc_max * A.fn * (double)N * (double)(N - 1) / vc * 0.5 * A.dt);   // :27
    long long n_att = (long long)npair;
    if (n_att < A.min_att) n_att = A.min_att;
    int n_a = 0, n_coll = 0, n_sc = 0;
    __syncthreads();
    for (long long b0 = 0; b0 < n_att; b0 += kCollThreads) {
        const long long att = b0 + t;
        bool pending = att < n_att;
        int s1 = 0, s2 = 0;
        U4 o0{};
        if (pending) {
            o0 = draw((uint32_t)att, (uint32_t)cell, A.step, kBlk0, A.key0, A.key1);
            s1 = (int)(du(o0.a) * (double)N);
            s2 = (int)(du(o0.b) * (double)N);
            for (int r = 1; s1 == s2 && r <= kMaxRedraw; ++r) {    // :32-36
                const U4 q = draw((uint32_t)att, (uint32_t)cell, A.step, kRedraw + (uint32_t)r, A.key0, A.key1);
                s1 = (int)(du(q.a) * (double)N);
                s2 = (int)(du(q.b) * (double)N);
            }
            ++n_a;
            if (s1 == s2) pending = false;
        }
        const int h1 = s1 & (kTabW - 1), h2 = s2 & (kTabW - 1);
        // the pair's values, loaded by every thread at once; a thread that
        // waits for a lower one loads them again when it runs
        int i1 = 0, i2 = 0;
        float v[8];
        if (pending) {
            i1 = A.vals ? (int)A.vals[st + s1] : st + s1;
            i2 = A.vals ? (int)A.vals[st + s2] : st + s2;
            load_pair(A, i1, i2, v);
        }
        bool first = true;
        for (;;) {
            if (t == 0) s_nacc = 0;
            if (pending) {
                atomicMin(&tab[h1], t);
                atomicMin(&tab[h2], t);
            }
            __syncthreads();
            const bool ready = pending && tab[h1] == t && tab[h2] == t;
            __syncthreads();
            if (pending) {
                tab[h1] = kCollThreads;
                tab[h2] = kCollThreads;
            }
            if (ready) {
                if (!first) load_pair(A, i1, i2, v);
                if (dsmc_accept(A, o0, v, n_sc)) {
                    const int k = atomicAdd(&s_nacc, 1);
                    s_att[k] = (uint32_t)att;
                    s_b[k] = o0.d;
                    s_i1[k] = i1;
                    s_i2[k] = i2;
#pragma unroll
                    for (int c = 0; c < 8; ++c) s_v[c][k] = v[c];
                }
                pending = false;
            }
            first = false;
            __syncthreads();
            // the ready attempts touch disjoint molecules: any order of their exchanges is the sequential result
            if (t < s_nacc) {
                float w[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) w[c] = s_v[c][t];
                dsmc_exchange<kVhs>(A, s_att[t], (uint32_t)cell, s_b[t], s_i1[t], s_i2[t], w, n_coll);
            }
            // this round's stores before the next round's loads (other threads of the workgroup)
            __threadfence_block();
            if (!__syncthreads_or(pending ? 1 : 0)) break;
        }
    }
    if (A.vrec) copy_out_cell(A, st, N);                            // after the last round's fence and barrier
    const int cv[3] = {n_a, n_coll, n_sc};
    wave_add_counters(A.ctr, 0, cv, 3);
}

// Per cell: count and sum of v^2 (v the speed) over the cell's molecules,
// lane-strided partial sums reduced in a fixed butterfly order.
__global__ __launch_bounds__(256) void k_dsmc_cellsum(int64_t ncell, const int32_t* start, const int32_t* end,
                                                      const uint32_t* vals, const float* vx, const float* vy,
                                                      const float* vz, int32_t* cnt, double* s2)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int64_t cell = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (cell >= ncell) return;
    const int st = start[cell], n = end[cell] - st;
    double s = 0.0;
    for (int k = lane; k < n; k += 64) {
        const int i = vals ? (int)vals[st + k] : st + k;
        const double a = vx[i], b = vy[i], c = vz[i];
        const double v = sqrt(a * a + b * b + c * c);
        s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        cnt[cell] = n;
        s2[cell] = s;
    }
}

// particle_density0.m:35 and count_T_numbensity.m:13 per cell, and the same
// per (ring, layer) of the map: one thread per (layer, ring).
__global__ void k_dsmc_sample_out(DMap m, const int32_t* cnt, const double* s2, double fn, double mass,
                                  float* cdens, float* ctemp, float* rdens, float* rtemp)
{
#pragma clang fp contract(off)
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)m.nl * m.nr) return;
    const int iz = (int)(g / m.nr), ir = (int)(g % m.nr);
    const double vc = m.vc[ir];
    long long rc = 0;
    double rs = 0.0;
    for (int k = m.sum0[ir]; k < m.sum0[ir + 1]; ++k) {
        const int64_t c = (int64_t)iz * m.nxy + k;
        const int n = cnt[c];
        if (cdens) cdens[c] = (float)(fn * (double)n / vc);
        if (ctemp) ctemp[c] = (float)(n ? mass / 3 / kKb * (s2[c] / (double)n) : 0.0);
        rc += n;
        rs += s2[c];
    }
    if (rdens) rdens[g] = (float)(fn * (double)rc / (vc * m.pd[ir]));
    if (rtemp) rtemp[g] = (float)(rc ? mass / 3 / kKb * (rs / (double)rc) : 0.0);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct DsmcState {
    int64_t cap = 0;                       // molecule slots of the sort buffers
    uint32_t *k_in = nullptr, *k_out = nullptr, *v_in = nullptr, *v_out = nullptr;
    void* rtmp = nullptr;
    size_t rtmp_bytes = 0;
    int64_t rec_cap = 0;
    float4* rec = nullptr;                 // reorder records, 2 x float4 per slot
    float4* vrec = nullptr;                // (vx, vy, vz, ei) per slot in cell order
    int64_t rq_cap = 0;
    int32_t* rq = nullptr;                 // wall / face queue (rq_cap slots) and its length
    int32_t* rq_n = nullptr;
    int64_t ncell_cap = 0;
    int32_t *start = nullptr, *end = nullptr, *cnt = nullptr;
    double* s2 = nullptr;
    int32_t* live = nullptr;
    unsigned long long* ctr = nullptr;
    int32_t* h_n = nullptr;                // pinned host word
    // the map on the device, for the parameters it was built from
    gasx_cell_map map{};
    bool have_map = false;
    int32_t* d_pd = nullptr;
    int32_t* d_sum0 = nullptr;
    double* d_dphi = nullptr;
    double* d_vc = nullptr;
    int32_t* d_ring = nullptr;
    float* d_inv_dphi = nullptr;
    int32_t nxy = 0;
};

namespace {

// cell_structure.m:6-22 and calculate_VC.m:5-17 (the same operations as the
// oracle's, in FP64 without contraction)
int64_t build_map_host(const gasx_cell_map& m, std::vector<int32_t>& pd, std::vector<int32_t>& sum0,
                       std::vector<double>& dphi, std::vector<double>& vc)
{
#pragma clang fp contract(off)
    const int rmax = m.n_rings;
    pd.assign(rmax, 0);
    sum0.assign(rmax + 1, 0);
    dphi.assign(rmax, 0.0);
    vc.assign(rmax, 0.0);
    const double dr = 1.0 * m.rp / rmax, dz = 1.0 * m.lp / m.n_layers;
    int64_t acc = 0;
    for (int i = 1; i <= rmax; ++i) {
        const double l = (double)i / rmax * m.rp * m.phip;
        const double c = std::ceil(l / m.delta);
        if (!(c >= 1.0) || c > (double)(1 << 24)) return -1;
        const int32_t p = (int32_t)c;
        const double dp = m.phip / p;
        const double a = (double)i * dr, b = (double)(i - 1) * dr;
        pd[i - 1] = p;
        sum0[i - 1] = (int32_t)acc;
        dphi[i - 1] = dp;
        vc[i - 1] = 0.5 * (a * a - b * b) * dp * dz;
        acc += p;
        if (acc > kMaxCells) return -1;
    }
    sum0[rmax] = (int32_t)acc;
    return acc;
}

bool map_valid(const gasx_cell_map* m)
{
    return m && m->n_rings >= 1 && m->n_layers >= 1 && m->n_rings <= (1 << 20) && m->n_layers <= (1 << 20) &&
           std::isfinite(m->rp) && m->rp > 0 && std::isfinite(m->lp) && m->lp > 0 && std::isfinite(m->phip) &&
           m->phip > 0 && m->phip < 6.283185307179586 && std::isfinite(m->delta) && m->delta > 0;
}

bool mol_valid(const gasx_molecules* mo)
{
    return mo && mo->x && mo->y && mo->z && mo->vx && mo->vy && mo->vz && mo->ei && mo->count && mo->capacity >= 1 &&
           mo->capacity <= (int64_t)INT32_MAX;
}

gasx_status dsmc_state(gasx_ctx* ctx, DsmcState** out)
{
    if (!ctx->dsmc) {
        ctx->dsmc = new DsmcState();
        DsmcState* s = ctx->dsmc;
        HIPCHK(dalloc(&s->ctr, (size_t)kCtrSlots * kCtrStride), "alloc dsmc counters");
        HIPCHK(dalloc(&s->live, 1), "alloc dsmc live");
        HIPCHK(hipHostMalloc((void**)&s->h_n, sizeof(int32_t)), "alloc dsmc pinned word");
        HIPCHK(hipMemsetAsync(s->ctr, 0, (size_t)kCtrSlots * kCtrStride * sizeof(unsigned long long), ctx->stream),
               "clear dsmc counters");
    }
    *out = ctx->dsmc;
    return GASX_OK;
}

// The map on the device (rebuilt when its parameters change).
gasx_status upload_map(gasx_ctx* ctx, DsmcState* s, const gasx_cell_map* m, DMap* dm)
{
    std::vector<int32_t> pd, sum0;
    std::vector<double> dphi, vc;
    const int64_t nxy = build_map_host(*m, pd, sum0, dphi, vc);
    if (nxy < 1 || nxy * (int64_t)m->n_layers > kMaxCells)
        return fail(ctx, GASX_E_ARG, "DSMC cell map: more than 2^24 cells or an empty ring");
    const bool same = s->have_map && s->map.n_rings == m->n_rings && s->map.n_layers == m->n_layers &&
                      s->map.rp == m->rp && s->map.lp == m->lp && s->map.phip == m->phip &&
                      s->map.delta == m->delta;
    if (!same) {
        (void)hipStreamSynchronize(ctx->stream);
        void* old[] = {s->d_pd, s->d_sum0, s->d_dphi, s->d_vc, s->d_ring, s->d_inv_dphi};
        for (void* b : old)
            if (b) hipFree(b);
        s->d_pd = s->d_sum0 = s->d_ring = nullptr;
        s->d_dphi = s->d_vc = nullptr;
        s->d_inv_dphi = nullptr;
        s->have_map = false;
        std::vector<int32_t> ring((size_t)nxy);
        for (int i = 0; i < m->n_rings; ++i)
            for (int k = sum0[i]; k < sum0[i + 1]; ++k) ring[k] = i;
        HIPCHK(dalloc(&s->d_pd, pd.size()), "alloc map");
        HIPCHK(dalloc(&s->d_sum0, sum0.size()), "alloc map");
        HIPCHK(dalloc(&s->d_dphi, dphi.size()), "alloc map");
        HIPCHK(dalloc(&s->d_vc, vc.size()), "alloc map");
        HIPCHK(dalloc(&s->d_ring, ring.size()), "alloc map");
        std::vector<float> inv(dphi.size());
        for (size_t k = 0; k < dphi.size(); ++k) inv[k] = (float)(1.0 / dphi[k]);
        HIPCHK(dalloc(&s->d_inv_dphi, inv.size()), "alloc map");
        HIPCHK(hipMemcpy(s->d_inv_dphi, inv.data(), inv.size() * 4, hipMemcpyHostToDevice), "upload map");
        HIPCHK(hipMemcpy(s->d_pd, pd.data(), pd.size() * 4, hipMemcpyHostToDevice), "upload map");
        HIPCHK(hipMemcpy(s->d_sum0, sum0.data(), sum0.size() * 4, hipMemcpyHostToDevice), "upload map");
        HIPCHK(hipMemcpy(s->d_dphi, dphi.data(), dphi.size() * 8, hipMemcpyHostToDevice), "upload map");
        HIPCHK(hipMemcpy(s->d_vc, vc.data(), vc.size() * 8, hipMemcpyHostToDevice), "upload map");
        HIPCHK(hipMemcpy(s->d_ring, ring.data(), ring.size() * 4, hipMemcpyHostToDevice), "upload map");
        s->map = *m;
        s->have_map = true;
        s->nxy = (int32_t)nxy;
    }
    dm->nr = m->n_rings;
    dm->nl = m->n_layers;
    dm->nxy = s->nxy;
    dm->ncell = (int64_t)s->nxy * m->n_layers;
    dm->rp = m->rp;
    dm->lp = m->lp;
    dm->phip = m->phip;
    dm->cp = std::cos(m->phip);
    dm->sp = std::sin(m->phip);
    dm->rp2_lo = m->rp * m->rp * (1.0 - 1e-12);
    dm->fast = m->phip < 3.141592653589793 ? 1 : 0;
    dm->inv_dr_f = (float)(m->n_rings / m->rp);
    dm->inv_dz_f = (float)(m->n_layers / m->lp);
    dm->inv_dphi_f = s->d_inv_dphi;
    dm->pd = s->d_pd;
    dm->sum0 = s->d_sum0;
    dm->dphi = s->d_dphi;
    dm->vc = s->d_vc;
    dm->ring_of = s->d_ring;
    return GASX_OK;
}

int key_bits(int64_t ncell)
{
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)ncell) ++bits;
    return bits;
}

// Sort buffers for `cap` slots and `ncell` cells.
gasx_status ensure_buffers(gasx_ctx* ctx, DsmcState* s, int64_t cap, int64_t ncell)
{
    if (s->cap < cap) {
        (void)hipStreamSynchronize(ctx->stream);
        void* old[] = {s->k_in, s->k_out, s->v_in, s->v_out, s->rtmp};
        for (void* b : old)
            if (b) hipFree(b);
        s->k_in = s->k_out = s->v_in = s->v_out = nullptr;
        s->rtmp = nullptr;
        s->cap = 0;
        size_t tb = 0;
        HIPCHK(rocprim::radix_sort_pairs((void*)nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                         (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)cap, 0, 32, ctx->stream),
               "radix sort size");
        HIPCHK(dalloc(&s->k_in, cap), "alloc dsmc keys");
        HIPCHK(dalloc(&s->k_out, cap), "alloc dsmc keys");
        HIPCHK(dalloc(&s->v_in, cap), "alloc dsmc slots");
        HIPCHK(dalloc(&s->v_out, cap), "alloc dsmc slots");
        HIPCHK(hipMalloc(&s->rtmp, tb > 0 ? tb : 1), "alloc radix sort scratch");
        s->rtmp_bytes = tb;
        s->cap = cap;
    }
    if (s->ncell_cap < ncell) {
        (void)hipStreamSynchronize(ctx->stream);
        void* old[] = {s->start, s->end, s->cnt, s->s2};
        for (void* b : old)
            if (b) hipFree(b);
        s->start = s->end = s->cnt = nullptr;
        s->s2 = nullptr;
        s->ncell_cap = 0;
        HIPCHK(dalloc(&s->start, ncell), "alloc dsmc cells");
        HIPCHK(dalloc(&s->end, ncell), "alloc dsmc cells");
        HIPCHK(dalloc(&s->cnt, ncell), "alloc dsmc cells");
        HIPCHK(dalloc(&s->s2, ncell), "alloc dsmc cells");
        s->ncell_cap = ncell;
    }
    return GASX_OK;
}

int grid_for(int64_t n, int per_block = 256, int max_blocks = 16384)
{
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (int)b;
}

// The molecule pass runs persistent workgroups, so each wave walks enough
// molecules to fill its queue of wall and face cases many times.
int move_grid(const gasx_ctx* ctx, int64_t n) { return grid_for(n, 256, ctx->cu_count * kMoveBlocksPerCU); }

// The wall / face queue and (with reorder) the record buffer.
gasx_status ensure_move_buffers(gasx_ctx* ctx, DsmcState* s, int64_t cap, bool rec)
{
    if (!s->rq_n) HIPCHK(dalloc(&s->rq_n, 1), "alloc dsmc queue length");
    if (s->rq_cap < cap) {
        (void)hipStreamSynchronize(ctx->stream);
        if (s->rq) hipFree(s->rq);
        s->rq = nullptr;
        s->rq_cap = 0;
        HIPCHK(dalloc(&s->rq, cap), "alloc dsmc wall queue");
        s->rq_cap = cap;
    }
    if (rec && s->rec_cap < cap) {
        (void)hipStreamSynchronize(ctx->stream);
        if (s->rec) hipFree(s->rec);
        s->rec = nullptr;
        s->rec_cap = 0;
        if (s->vrec) hipFree(s->vrec);
        s->vrec = nullptr;
        HIPCHK(dalloc(&s->rec, 2 * (size_t)cap), "alloc dsmc reorder records");
        HIPCHK(dalloc(&s->vrec, (size_t)cap), "alloc dsmc velocity records");
        s->rec_cap = cap;
    }
    return GASX_OK;
}

// The molecule pass (move or not, in place or into records) and its queue of
// wall and face cases.
gasx_status launch_move(gasx_ctx* ctx, DsmcState* s, gasx_molecules* mo, const DMap& dm, bool move, double dt,
                        bool rec)
{
    const Mol m{mo->x, mo->y, mo->z, mo->vx, mo->vy, mo->vz, mo->ei, mo->count};
    MoveOut o{rec ? s->rec : nullptr, s->k_in, s->v_in, s->rq, s->rq_n};
    hipLaunchKernelGGL(k_dsmc_zero, dim3(1), dim3(64), 0, ctx->stream, s->rq_n);
    HIPCHK(hipGetLastError(), "k_dsmc_zero");
    const int blocks = move_grid(ctx, mo->capacity);
    if (move)
        hipLaunchKernelGGL(k_dsmc_move<true>, dim3(blocks), dim3(256), 0, ctx->stream, m, dm, dt, o);
    else
        hipLaunchKernelGGL(k_dsmc_move<false>, dim3(blocks), dim3(256), 0, ctx->stream, m, dm, dt, o);
    HIPCHK(hipGetLastError(), "k_dsmc_move");
    if (move) {
        hipLaunchKernelGGL(k_dsmc_rare, dim3(ctx->cu_count * 4), dim3(256), 0, ctx->stream, m, dm, dt, o, s->ctr);
        HIPCHK(hipGetLastError(), "k_dsmc_rare");
    }
    return GASX_OK;
}

// Keys (after the optional move), the host's slot count, the radix sort and
// the cell bounds; with reorder the molecules are then permuted into cell
// order (absorbed slots dropped).  *sorted_vals = the slot list, or null when
// the molecules are in cell order.
gasx_status dsmc_prepare(gasx_ctx* ctx, gasx_molecules* mo, const gasx_cell_map* map, bool move, double dt,
                         bool reorder, DsmcState** sp, DMap* dm, const uint32_t** sorted_vals, int32_t* n_out)
{
    DsmcState* s = nullptr;
    gasx_status r = dsmc_state(ctx, &s);
    if (r != GASX_OK) return r;
    r = upload_map(ctx, s, map, dm);
    if (r != GASX_OK) return r;
    r = ensure_buffers(ctx, s, mo->capacity, dm->ncell);
    if (r != GASX_OK) return r;
    r = ensure_move_buffers(ctx, s, mo->capacity, reorder);
    if (r != GASX_OK) return r;
    r = launch_move(ctx, s, mo, *dm, move, dt, reorder);
    if (r != GASX_OK) return r;
    HIPCHK(hipMemcpyAsync(s->h_n, mo->count, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream), "read count");
    HIPCHK(hipStreamSynchronize(ctx->stream), "read count");
    const int32_t n = *s->h_n;
    if (n < 0 || n > mo->capacity) return fail(ctx, GASX_E_ARG, "DSMC: molecule count outside [0, capacity]");
    hipLaunchKernelGGL(k_dsmc_fill, dim3(grid_for(dm->ncell)), dim3(256), 0, ctx->stream, s->start, s->end,
                       dm->ncell, s->live);
    HIPCHK(hipGetLastError(), "k_dsmc_fill");
    if (n > 0) {
        size_t tb = s->rtmp_bytes;
        HIPCHK(rocprim::radix_sort_pairs(s->rtmp, tb, s->k_in, s->k_out, s->v_in, s->v_out, (size_t)n, 0,
                                         key_bits(dm->ncell), ctx->stream),
               "radix sort");
        hipLaunchKernelGGL(k_dsmc_bounds, dim3(grid_for(n)), dim3(256), 0, ctx->stream, (const uint32_t*)s->k_out,
                           (const int32_t*)mo->count, (uint32_t)dm->ncell, s->start, s->end, s->live);
        HIPCHK(hipGetLastError(), "k_dsmc_bounds");
    }
    *sorted_vals = s->v_out;
    if (reorder) {
        const Mol m{mo->x, mo->y, mo->z, mo->vx, mo->vy, mo->vz, mo->ei, mo->count};
        hipLaunchKernelGGL(k_dsmc_unpack, dim3(grid_for(n > 0 ? n : 1)), dim3(256), 0, ctx->stream, m, mo->count,
                           (const float4*)s->rec, (const uint32_t*)s->v_out, (const int32_t*)s->live, s->vrec);
        HIPCHK(hipGetLastError(), "k_dsmc_unpack");
        *sorted_vals = nullptr;
    }
    *sp = s;
    *n_out = n;
    return GASX_OK;
}

bool params_valid(const gasx_dsmc_params* p)
{
    return p && std::isfinite(p->dt) && p->dt >= 0 && std::isfinite(p->fn) && p->fn > 0 && std::isfinite(p->mass) &&
           p->mass > 0 && std::isfinite(p->diameter) && p->diameter > 0 && std::isfinite(p->sc_max) &&
           p->sc_max > 0 && std::isfinite(p->omega) && p->omega < 1.5 && std::isfinite(p->zeta) &&
           p->zeta >= 2 && std::isfinite(p->alpha) && p->alpha > 0 && p->min_attempts >= 0 &&
           p->min_attempts <= 1;
}

gasx_status dsmc_collide_launch(gasx_ctx* ctx, gasx_molecules* mo, DsmcState* s, const DMap& dm,
                                const uint32_t* vals, const gasx_dsmc_params* p, uint64_t seed, uint32_t step)
{
#pragma clang fp contract(off)
    CollArgsD A{};
    A.m = dm;
    A.vx = mo->vx; A.vy = mo->vy; A.vz = mo->vz; A.ei = mo->ei;
    A.start = s->start; A.end = s->end;
    A.vals = vals;
    A.vrec = vals ? nullptr : s->vrec;      // in cell order the unpack wrote the velocity records
    A.sc_max = p->sc_max; A.fn = p->fn; A.dt = p->dt;
    A.d2 = p->diameter * p->diameter;
    A.mass = p->mass;
    const double w = p->omega, zeta = p->zeta;
    A.k1 = (zeta + 0.5 - w) / (1.5 - w);                           // collision_HS_internal.m:65
    A.a1 = 1.5 - w;
    A.k2 = (zeta + 0.5 - w) / (zeta - 1);
    A.b1 = zeta - 1;
    A.c2 = std::pow(2.0, zeta - 2);                                 // :77
    A.h = zeta / 2 - 1;
    A.k1f = (float)A.k1; A.a1f = (float)A.a1; A.k2f = (float)A.k2; A.b1f = (float)A.b1;
    A.c2f = (float)A.c2; A.hf = (float)A.h;
    A.alpha = p->alpha;
    A.vhs = std::fabs(p->alpha - 1) < 1e-6 ? 1 : 0;                 // :91
    A.min_att = p->min_attempts;
    A.step = step;
    A.key0 = (uint32_t)seed;
    A.key1 = (uint32_t)(seed >> 32);
    A.ctr = s->ctr;
    if (A.vhs)
        hipLaunchKernelGGL(k_dsmc_collide<true>, dim3((unsigned)dm.ncell), dim3(kCollThreads), 0, ctx->stream, A);
    else
        hipLaunchKernelGGL(k_dsmc_collide<false>, dim3((unsigned)dm.ncell), dim3(kCollThreads), 0, ctx->stream, A);
    HIPCHK(hipGetLastError(), "k_dsmc_collide");
    return GASX_OK;
}

}  // namespace

void gasx_host::dsmc_release(gasx_ctx* ctx)
{
    DsmcState* s = ctx->dsmc;
    if (!s) return;
    void* bufs[] = {s->k_in, s->k_out, s->v_in, s->v_out, s->rtmp, s->start, s->end, s->cnt, s->s2, s->live, s->ctr,
                    s->d_pd, s->d_sum0, s->d_dphi, s->d_vc, s->d_ring, s->d_inv_dphi};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (s->rec) hipFree(s->rec);
    if (s->vrec) hipFree(s->vrec);
    if (s->rq) hipFree(s->rq);
    if (s->rq_n) hipFree(s->rq_n);
    if (s->h_n) hipHostFree(s->h_n);
    delete s;
    ctx->dsmc = nullptr;
}

extern "C" gasx_status gasx_cell_map_info(const gasx_cell_map* m, int32_t* phi_divide, int64_t* cells_per_layer)
{
    if (!map_valid(m) || !cells_per_layer) return GASX_E_ARG;
    std::vector<int32_t> pd, sum0;
    std::vector<double> dphi, vc;
    const int64_t nxy = build_map_host(*m, pd, sum0, dphi, vc);
    if (nxy < 1 || nxy * (int64_t)m->n_layers > kMaxCells) return GASX_E_ARG;
    if (phi_divide) std::copy(pd.begin(), pd.end(), phi_divide);
    *cells_per_layer = nxy;
    return GASX_OK;
}

extern "C" gasx_status gasx_dsmc_cells(gasx_ctx* ctx, const gasx_molecules* mo, const gasx_cell_map* map,
                                       int32_t* cell)
{
    if (!ctx) return GASX_E_ARG;
    if (!mol_valid(mo) || !map_valid(map) || !cell) return fail(ctx, GASX_E_ARG, "gasx_dsmc_cells: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    DsmcState* s = nullptr;
    gasx_status r = dsmc_state(ctx, &s);
    if (r != GASX_OK) return r;
    DMap dm{};
    r = upload_map(ctx, s, map, &dm);
    if (r != GASX_OK) return r;
    r = ensure_buffers(ctx, s, mo->capacity, dm.ncell);
    if (r != GASX_OK) return r;
    const Mol m{mo->x, mo->y, mo->z, mo->vx, mo->vy, mo->vz, mo->ei, mo->count};
    // the sort keys are the cells (the slot column goes to the sort buffer)
    const MoveOut o{nullptr, (uint32_t*)cell, s->v_in, nullptr, nullptr};
    hipLaunchKernelGGL(k_dsmc_move<false>, dim3(move_grid(ctx, mo->capacity)), dim3(256), 0, ctx->stream, m, dm,
                       0.0, o);
    HIPCHK(hipGetLastError(), "k_dsmc_move");
    return GASX_OK;
}

extern "C" gasx_status gasx_dsmc_move(gasx_ctx* ctx, gasx_molecules* mo, const gasx_cell_map* map, double dt)
{
    if (!ctx) return GASX_E_ARG;
    if (!mol_valid(mo) || !map_valid(map) || !std::isfinite(dt) || dt < 0)
        return fail(ctx, GASX_E_ARG, "gasx_dsmc_move: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    DsmcState* s = nullptr;
    gasx_status r = dsmc_state(ctx, &s);
    if (r != GASX_OK) return r;
    DMap dm{};
    r = upload_map(ctx, s, map, &dm);
    if (r != GASX_OK) return r;
    r = ensure_buffers(ctx, s, mo->capacity, dm.ncell);
    if (r != GASX_OK) return r;
    r = ensure_move_buffers(ctx, s, mo->capacity, false);
    if (r != GASX_OK) return r;
    return launch_move(ctx, s, mo, dm, true, dt, false);
}

extern "C" gasx_status gasx_dsmc_collide(gasx_ctx* ctx, gasx_molecules* mo, const gasx_cell_map* map,
                                         const gasx_dsmc_params* p, uint64_t seed, uint32_t step)
{
    if (!ctx) return GASX_E_ARG;
    if (!mol_valid(mo) || !map_valid(map) || !params_valid(p))
        return fail(ctx, GASX_E_ARG, "gasx_dsmc_collide: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    DsmcState* s = nullptr;
    DMap dm{};
    const uint32_t* vals = nullptr;
    int32_t n = 0;
    gasx_status r = dsmc_prepare(ctx, mo, map, false, 0.0, p->reorder != 0, &s, &dm, &vals, &n);
    if (r != GASX_OK || n == 0) return r;
    return dsmc_collide_launch(ctx, mo, s, dm, vals, p, seed, step);
}

extern "C" gasx_status gasx_dsmc_step(gasx_ctx* ctx, gasx_molecules* mo, const gasx_cell_map* map,
                                      const gasx_dsmc_params* p, uint64_t seed, uint32_t step)
{
    if (!ctx) return GASX_E_ARG;
    if (!mol_valid(mo) || !map_valid(map) || !params_valid(p))
        return fail(ctx, GASX_E_ARG, "gasx_dsmc_step: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    DsmcState* s = nullptr;
    DMap dm{};
    const uint32_t* vals = nullptr;
    int32_t n = 0;
    gasx_status r = dsmc_prepare(ctx, mo, map, true, p->dt, p->reorder != 0, &s, &dm, &vals, &n);
    if (r != GASX_OK || n == 0) return r;
    return dsmc_collide_launch(ctx, mo, s, dm, vals, p, seed, step);
}

extern "C" gasx_status gasx_dsmc_sample(gasx_ctx* ctx, const gasx_molecules* mo, const gasx_cell_map* map,
                                        double fn, double mass, float* cell_density, float* cell_temperature,
                                        float* rz_density, float* rz_temperature)
{
    if (!ctx) return GASX_E_ARG;
    if (!mol_valid(mo) || !map_valid(map) || !std::isfinite(fn) || !(fn > 0) || !std::isfinite(mass) || !(mass > 0))
        return fail(ctx, GASX_E_ARG, "gasx_dsmc_sample: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    DsmcState* s = nullptr;
    DMap dm{};
    const uint32_t* vals = nullptr;
    int32_t n = 0;
    gasx_molecules m = *mo;   // read only: no move, no reorder
    gasx_status r = dsmc_prepare(ctx, &m, map, false, 0.0, false, &s, &dm, &vals, &n);
    if (r != GASX_OK) return r;
    hipLaunchKernelGGL(k_dsmc_cellsum, dim3(grid_for(dm.ncell * 64)), dim3(256), 0, ctx->stream, dm.ncell,
                       (const int32_t*)s->start, (const int32_t*)s->end, vals, (const float*)mo->vx,
                       (const float*)mo->vy, (const float*)mo->vz, s->cnt, s->s2);
    HIPCHK(hipGetLastError(), "k_dsmc_cellsum");
    hipLaunchKernelGGL(k_dsmc_sample_out, dim3(grid_for((int64_t)dm.nl * dm.nr)), dim3(256), 0, ctx->stream, dm,
                       (const int32_t*)s->cnt, (const double*)s->s2, fn, mass, cell_density, cell_temperature,
                       rz_density, rz_temperature);
    HIPCHK(hipGetLastError(), "k_dsmc_sample_out");
    return GASX_OK;
}

extern "C" gasx_status gasx_dsmc_get_counters(gasx_ctx* ctx, gasx_dsmc_counters* out, int reset)
{
    if (!ctx || !out) return GASX_E_ARG;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    DsmcState* s = nullptr;
    gasx_status r = dsmc_state(ctx, &s);
    if (r != GASX_OK) return r;
    std::vector<unsigned long long> slots((size_t)kCtrSlots * kCtrStride);
    HIPCHK(hipMemcpyAsync(slots.data(), s->ctr, slots.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          ctx->stream),
           "read dsmc counters");
    HIPCHK(hipStreamSynchronize(ctx->stream), "read dsmc counters");
    unsigned long long h[kNDsmcCounters] = {};
    for (int b = 0; b < kCtrSlots; ++b)
        for (int k = 0; k < kNDsmcCounters; ++k) h[k] += slots[(size_t)b * kCtrStride + k];
    out->n_attempts = (int64_t)h[0];
    out->n_collisions = (int64_t)h[1];
    out->n_sc_exceed = (int64_t)h[2];
    out->n_boundary = (int64_t)h[3];
    out->n_wall_r = (int64_t)h[4];
    out->n_wall_z = (int64_t)h[5];
    out->n_wrap = (int64_t)h[6];
    if (reset)
        HIPCHK(hipMemsetAsync(s->ctr, 0, slots.size() * sizeof(unsigned long long), ctx->stream),
               "clear dsmc counters");
    return GASX_OK;
}
