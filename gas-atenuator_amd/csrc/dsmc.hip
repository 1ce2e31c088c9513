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
//   k_dsmc_gather        (reorder) the molecules permuted into cell order
//   k_dsmc_collide       one wave per cell: the cell's no-time-counter
//                        attempts in batches of 64 lanes.  An attempt's
//                        random numbers come from its own Philox counter
//                        (attempt, cell, step, block), so a batch can draw in
//                        parallel; a lane runs its attempt once no earlier
//                        pending lane of the batch holds one of its two
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
constexpr int kTab = 2048;                   // LDS marks per wave (molecule index mod kTab)
constexpr int kWaves = 4;                    // waves (cells) per collide workgroup
constexpr double kPi = 3.141592653589793;    // MATLAB pi
constexpr double kKb = 1.38064853e-23;       // main2.m:5
constexpr int kNDsmcCounters = 7;            // attempts, collisions, SC > SCmax, boundary, r wall, z walls, wraps
constexpr int64_t kMaxCells = 1ll << 24;

struct DMap {
    int32_t nr, nl, nxy;
    int64_t ncell;
    double rp, lp, phip, cp, sp;
    const int32_t* pd;        // phi cells per ring
    const int32_t* sum0;      // exclusive prefix, nr + 1
    const double* dphi;       // per ring
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

__device__ __forceinline__ void wave_add_counters(unsigned long long* g, int k0, const int* v, int n)
{
    for (int k = 0; k < n; ++k) {
        int s = v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if ((threadIdx.x & 63) == 0 && s) atomicAdd(&g[k0 + k], (unsigned long long)s);
    }
}

// Molecule pass: (free flight + walls) and the sort key of every slot below
// the count.  keys = cell (ncell for an absorbed slot), vals = slot.
template <bool kMove>
__global__ __launch_bounds__(256) void k_dsmc_move(Mol mo, DMap m, double dt, uint32_t* keys, uint32_t* vals,
                                                   unsigned long long* ctr)
{
#pragma clang fp contract(off)
    const int n = *mo.count;
    int c[4] = {0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double x = mo.x[i], y = mo.y[i], z = mo.z[i];
        if (kMove && z == z) {
            double vx = mo.vx[i], vy = mo.vy[i], vz = mo.vz[i];
            const float ovx = mo.vx[i], ovy = mo.vy[i], ovz = mo.vz[i];
            dsmc_move_one(m, dt, x, y, z, vx, vy, vz, c);
            const float fx = (float)x, fy = (float)y, fz = (float)z;
            mo.x[i] = fx; mo.y[i] = fy; mo.z[i] = fz;
            const float fvx = (float)vx, fvy = (float)vy, fvz = (float)vz;
            if (fvx != ovx || fvy != ovy) { mo.vx[i] = fvx; mo.vy[i] = fvy; }   // walls and wraps only
            if (fvz != ovz) mo.vz[i] = fvz;
            x = fx; y = fy; z = fz;
        }
        keys[i] = (uint32_t)dsmc_cell(m, x, y, z);
        vals[i] = (uint32_t)i;
    }
    if (kMove) wave_add_counters(ctr, 3, c, 4);
}

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

// (reorder) molecules of the sorted order into the scratch copies
__global__ void k_dsmc_gather(Mol mo, const uint32_t* vals, const int32_t* live, float* t0, float* t1, float* t2,
                              float* t3, float* t4, float* t5, float* t6)
{
    const int n = *live;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = vals[i];
        t0[i] = mo.x[s]; t1[i] = mo.y[s]; t2[i] = mo.z[s];
        t3[i] = mo.vx[s]; t4[i] = mo.vy[s]; t5[i] = mo.vz[s]; t6[i] = mo.ei[s];
    }
}

// (reorder) the permuted copies back, and count = live
__global__ void k_dsmc_scatter_back(Mol mo, int32_t* count, const int32_t* live, const float* t0, const float* t1,
                                    const float* t2, const float* t3, const float* t4, const float* t5,
                                    const float* t6)
{
    const int n = *live;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        mo.x[i] = t0[i]; mo.y[i] = t1[i]; mo.z[i] = t2[i];
        mo.vx[i] = t3[i]; mo.vy[i] = t4[i]; mo.vz[i] = t5[i]; mo.ei[i] = t6[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

struct CollArgsD {
    DMap m;
    float *vx, *vy, *vz, *ei;
    const int32_t *start, *end;
    const uint32_t* vals;     // sorted slots (null: the molecules are in cell order)
    double sc_max, fn, dt, d2, mass;
    double k1, a1, k2, b1, c2, h;
    double alpha;
    int32_t vhs, min_att;
    uint32_t step, key0, key1;
    unsigned long long* ctr;
};

__device__ __forceinline__ double dpow(double x, double e) { return e == 1.0 ? x : (e == 0.0 ? 1.0 : pow(x, e)); }

// collision_HS_internal.m:38-126 for the pair (i1, i2) of attempt a
__device__ void dsmc_attempt(const CollArgsD& A, uint32_t a, uint32_t cell, const U4& o0, int i1, int i2,
                             int& n_coll, int& n_sc)
{
#pragma clang fp contract(off)
    const double vx1 = A.vx[i1], vy1 = A.vy[i1], vz1 = A.vz[i1], e1 = A.ei[i1];
    const double vx2 = A.vx[i2], vy2 = A.vy[i2], vz2 = A.vz[i2], e2 = A.ei[i2];
    double ur = vx1 - vx2, vr = vy1 - vy2, wr = vz1 - vz2;
    double cr = sqrt(ur * ur + vr * vr + wr * wr);
    const double SC = cr * kPi * A.d2;
    if (SC > A.sc_max) ++n_sc;
    if (!(du(o0.c) < SC / A.sc_max)) return;
    const double um = (vx1 + vx2) / 2, vm = (vy1 + vy2) / 2, wm = (vz1 + vz2) / 2;
    const double Ec = 0.25 * A.mass * (cr * cr) + e1 + e2;
    double Et = 0.0;
    U4 q{};
    for (int j = 0; j < kMaxRounds; ++j) {                          // :63-70
        if ((j & 1) == 0) q = draw(a, cell, A.step, kEt + (uint32_t)(j >> 1), A.key0, A.key1);
        Et = Ec * du(word(q, 2 * (j & 1)));
        const double pr = dpow(A.k1 * (Et / Ec), A.a1) * dpow(A.k2 * (1 - Et / Ec), A.b1);
        if (du(word(q, 2 * (j & 1) + 1)) < pr) break;
    }
    const double Ei = Ec - Et;
    double etry = 0.0;
    for (int j = 0; j < kMaxRounds; ++j) {                          // :74-83
        if ((j & 1) == 0) q = draw(a, cell, A.step, kEtry + (uint32_t)(j >> 1), A.key0, A.key1);
        etry = Ei * du(word(q, 2 * (j & 1)));
        const double pr = A.c2 * dpow(etry / Ei, A.h) * dpow(1 - etry / Ei, A.h);
        if (du(word(q, 2 * (j & 1) + 1)) < pr) break;
    }
    const double cri = 2 * sqrt(Et / A.mass);                       // :86-90
    ur = ur * cri / cr; vr = vr * cri / cr; wr = wr * cri / cr;
    cr = cri;
    const U4 o1 = draw(a, cell, A.step, kBlk1, A.key0, A.key1);
    double urf, vrf, wrf;
    if (A.vhs) {                                                    // :91-97
        const double B = 2 * du(o0.d) - 1;
        const double Aa = sqrt(1 - B * B);
        const double C = 2 * kPi * du(o1.a);
        urf = cr * B;
        vrf = cr * Aa * cos(C);
        wrf = cr * Aa * sin(C);
    } else {                                                        // :98-113
        const double B = 2 * pow(du(o0.d), 1 / A.alpha) - 1;
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
    A.ei[i1] = (float)etry;                                         // :84-85
    A.ei[i2] = (float)(Ei - etry);
    A.vx[i1] = (float)(um + 0.5 * urf); A.vx[i2] = (float)(um - 0.5 * urf);   // :115-120
    A.vy[i1] = (float)(vm + 0.5 * vrf); A.vy[i2] = (float)(vm - 0.5 * vrf);
    A.vz[i1] = (float)(wm + 0.5 * wrf); A.vz[i2] = (float)(wm - 0.5 * wrf);
    ++n_coll;
}

// One wave per cell (collision_HS_internal.m:11-136).  No workgroup barrier:
// the waves of a workgroup handle different cells with different attempt
// counts; each has its own LDS marks.
__global__ __launch_bounds__(64 * kWaves) void k_dsmc_collide(CollArgsD A)
{
#pragma clang fp contract(off)
    __shared__ int s_tab[kWaves][kTab];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t cell = (int64_t)blockIdx.x * kWaves + wv;
    if (cell >= A.m.ncell) return;
    int* tab = s_tab[wv];
    const int st = A.start[cell], N = A.end[cell] - st;
    if (N <= 2) return;                                             // :26
    for (int k = lane; k < kTab; k += 64) tab[k] = 64;
    const double vc = A.m.vc[A.m.ring_of[cell % A.m.nxy]];
    const double npair = floor(A.sc_max * A.fn * (double)N * (double)(N - 1) / vc * 0.5 * A.dt);   // :27
    long long n_att = (long long)npair;
    if (n_att < A.min_att) n_att = A.min_att;
    int n_a = 0, n_coll = 0, n_sc = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (long long b0 = 0; b0 < n_att; b0 += 64) {
        const long long att = b0 + lane;
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
        const int h1 = s1 & (kTab - 1), h2 = s2 & (kTab - 1);
        while (__ballot(pending)) {
            if (pending) {
                atomicMin(&tab[h1], lane);
                atomicMin(&tab[h2], lane);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const bool ready = pending && tab[h1] == lane && tab[h2] == lane;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (pending) {
                tab[h1] = 64;
                tab[h2] = 64;
            }
            if (ready) {
                const int i1 = A.vals ? (int)A.vals[st + s1] : st + s1;
                const int i2 = A.vals ? (int)A.vals[st + s2] : st + s2;
                dsmc_attempt(A, (uint32_t)att, (uint32_t)cell, o0, i1, i2, n_coll, n_sc);
                pending = false;
            }
            // this round's stores before the next round's loads (other lanes of the wave)
            __threadfence_block();
        }
    }
    const int v[3] = {n_a, n_coll, n_sc};
    wave_add_counters(A.ctr, 0, v, 3);
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
    int64_t tmp_cap = 0;
    float* tmp[7] = {};
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
        HIPCHK(dalloc(&s->ctr, kNDsmcCounters), "alloc dsmc counters");
        HIPCHK(dalloc(&s->live, 1), "alloc dsmc live");
        HIPCHK(hipHostMalloc((void**)&s->h_n, sizeof(int32_t)), "alloc dsmc pinned word");
        HIPCHK(hipMemsetAsync(s->ctr, 0, kNDsmcCounters * sizeof(unsigned long long), ctx->stream),
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
        void* old[] = {s->d_pd, s->d_sum0, s->d_dphi, s->d_vc, s->d_ring};
        for (void* b : old)
            if (b) hipFree(b);
        s->d_pd = s->d_sum0 = s->d_ring = nullptr;
        s->d_dphi = s->d_vc = nullptr;
        s->have_map = false;
        std::vector<int32_t> ring((size_t)nxy);
        for (int i = 0; i < m->n_rings; ++i)
            for (int k = sum0[i]; k < sum0[i + 1]; ++k) ring[k] = i;
        HIPCHK(dalloc(&s->d_pd, pd.size()), "alloc map");
        HIPCHK(dalloc(&s->d_sum0, sum0.size()), "alloc map");
        HIPCHK(dalloc(&s->d_dphi, dphi.size()), "alloc map");
        HIPCHK(dalloc(&s->d_vc, vc.size()), "alloc map");
        HIPCHK(dalloc(&s->d_ring, ring.size()), "alloc map");
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
    const Mol m{mo->x, mo->y, mo->z, mo->vx, mo->vy, mo->vz, mo->ei, mo->count};
    const int blocks = grid_for(mo->capacity);
    if (move)
        hipLaunchKernelGGL(k_dsmc_move<true>, dim3(blocks), dim3(256), 0, ctx->stream, m, *dm, dt, s->k_in,
                           s->v_in, s->ctr);
    else
        hipLaunchKernelGGL(k_dsmc_move<false>, dim3(blocks), dim3(256), 0, ctx->stream, m, *dm, dt, s->k_in,
                           s->v_in, s->ctr);
    HIPCHK(hipGetLastError(), "k_dsmc_move");
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
    if (reorder && n > 0) {
        if (s->tmp_cap < mo->capacity) {
            (void)hipStreamSynchronize(ctx->stream);
            for (auto& t : s->tmp) {
                if (t) hipFree(t);
                t = nullptr;
            }
            s->tmp_cap = 0;
            for (auto& t : s->tmp) HIPCHK(dalloc(&t, mo->capacity), "alloc dsmc reorder buffers");
            s->tmp_cap = mo->capacity;
        }
        float** t = s->tmp;
        hipLaunchKernelGGL(k_dsmc_gather, dim3(grid_for(n)), dim3(256), 0, ctx->stream, m,
                           (const uint32_t*)s->v_out, (const int32_t*)s->live, t[0], t[1], t[2], t[3], t[4], t[5],
                           t[6]);
        HIPCHK(hipGetLastError(), "k_dsmc_gather");
        hipLaunchKernelGGL(k_dsmc_scatter_back, dim3(grid_for(n)), dim3(256), 0, ctx->stream, m, mo->count,
                           (const int32_t*)s->live, (const float*)t[0], (const float*)t[1], (const float*)t[2],
                           (const float*)t[3], (const float*)t[4], (const float*)t[5], (const float*)t[6]);
        HIPCHK(hipGetLastError(), "k_dsmc_scatter_back");
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
    A.alpha = p->alpha;
    A.vhs = std::fabs(p->alpha - 1) < 1e-6 ? 1 : 0;                 // :91
    A.min_att = p->min_attempts;
    A.step = step;
    A.key0 = (uint32_t)seed;
    A.key1 = (uint32_t)(seed >> 32);
    A.ctr = s->ctr;
    const int64_t blocks = (dm.ncell + kWaves - 1) / kWaves;
    hipLaunchKernelGGL(k_dsmc_collide, dim3((unsigned)blocks), dim3(64 * kWaves), 0, ctx->stream, A);
    HIPCHK(hipGetLastError(), "k_dsmc_collide");
    return GASX_OK;
}

}  // namespace

void gasx_host::dsmc_release(gasx_ctx* ctx)
{
    DsmcState* s = ctx->dsmc;
    if (!s) return;
    void* bufs[] = {s->k_in, s->k_out, s->v_in, s->v_out, s->rtmp, s->start, s->end, s->cnt, s->s2, s->live, s->ctr,
                    s->d_pd, s->d_sum0, s->d_dphi, s->d_vc, s->d_ring};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (float* t : s->tmp)
        if (t) hipFree(t);
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
    hipLaunchKernelGGL(k_dsmc_move<false>, dim3(grid_for(mo->capacity)), dim3(256), 0, ctx->stream, m, dm, 0.0,
                       (uint32_t*)cell, s->v_in, s->ctr);
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
    const Mol m{mo->x, mo->y, mo->z, mo->vx, mo->vy, mo->vz, mo->ei, mo->count};
    hipLaunchKernelGGL(k_dsmc_move<true>, dim3(grid_for(mo->capacity)), dim3(256), 0, ctx->stream, m, dm, dt,
                       s->k_in, s->v_in, s->ctr);
    HIPCHK(hipGetLastError(), "k_dsmc_move");
    return GASX_OK;
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
    unsigned long long h[kNDsmcCounters] = {};
    HIPCHK(hipMemcpyAsync(h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), "read dsmc counters");
    HIPCHK(hipStreamSynchronize(ctx->stream), "read dsmc counters");
    out->n_attempts = (int64_t)h[0];
    out->n_collisions = (int64_t)h[1];
    out->n_sc_exceed = (int64_t)h[2];
    out->n_boundary = (int64_t)h[3];
    out->n_wall_r = (int64_t)h[4];
    out->n_wall_z = (int64_t)h[5];
    out->n_wrap = (int64_t)h[6];
    if (reset) HIPCHK(hipMemsetAsync(s->ctr, 0, sizeof(h), ctx->stream), "clear dsmc counters");
    return GASX_OK;
}
