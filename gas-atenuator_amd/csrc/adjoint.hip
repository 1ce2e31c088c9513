// Differentiable electron step with collision events frozen pathwise, and its
// hand-written adjoint.
//
// The model (a reference run at theta0 records the events; every later pass
// replays them):
//   prologue   x^{1/2} = BC(x0 + v0 dt/2)                      (EM2.cu)
//              J_0 = deposit(x^{1/2}, v0, w0)                    (EM7.cu, FP64)
//   step s     F_{s+1} = fieldcycle(F_s, J_s)                   (EM3_1.cu / EM3.cu, linear)
//              v^p = push(v_s; E, H of F_{s+1} at x^{1/2}_s)     (EM4.cu)
//              (x_{s+1}, v^-) = BC(x^{1/2}_s + v^p dt/2)
//              v_{s+1} = v^- + dv_s                              frozen collision kick
//              w_{s+1} = w_s L_s                                 likelihood ratio of the
//                  frozen decision: L = n(r)/n0(r) after a collision, else
//                  (1 - p0 n(r)/n0(r)) / (1 - p0); n(r) is the density profile
//                  (piecewise linear in r over the nodes), p0 the recorded p_T
//              (x^{1/2}_{s+1}, v_{s+1}) = BC(x_{s+1} + v_{s+1} dt/2)
//              J_{s+1} = deposit(x^{1/2}_{s+1}, v_{s+1}, w_{s+1})  (not after the last step)
//   objective  W = PHIp eps/2 sum over Er cells with r >= frac*Rp of Er^2 r dr dz
// with the deposit scaled by the source amplitude A (FN -> A FN).  The
// recording pass draws the collision decision and kinematics from the
// production physics (Philox counters as gasx_run, find_CS0.m cross sections,
// electron_transport.cu kinematics) with p0 = |v| dt n0(r) sigma_T(E).
//
// The adjoint runs backwards over checkpoint segments: each segment is
// recomputed from its checkpoint keeping every step's particle state and
// FP32 gather fields, then one particle kernel per step applies the
// transposed Jacobians of deposit, drift, boundary map, kick, weight and push
// (scattering the field sensitivities), and the transposed field cycle
// carries the field sensitivities one push back and yields the current's.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"
#include "particle_physics.h"

using namespace gasx;
using namespace gasx_host;

namespace gasx {

constexpr int kDiffBlock = 256;
constexpr int kMaxNodes = 64;

// Linear map of EM2/EM4's boundary conditions as applied to one particle:
// (x, y) -> sgn R (x, y), (vx, vy) -> R (vx, vy), z and vz negated when
// reflected an odd number of times; dead when absorbed (r > Rp).
struct BcMap {
    float c, s, sgn, zs;
    int dead;
};

// wedge_bc (device_common.h) with its map recorded; same operations.
__device__ BcMap bc_apply(const Geom& g, float& x, float& y, float& z, float& vx, float& vy, float& vz)
{
    BcMap m{1.0f, 0.0f, 1.0f, 1.0f, 0};
    const float r2 = x * x + y * y;
    if (r2 > g.rp * g.rp) { m.dead = 1; return m; }
    if (x > 0.0f) {
        int k = 0, guard = 0;
#pragma unroll 1
        while (y < 0.0f && guard++ < 32) {
            const float x1 = x * g.cphip - y * g.sphip, y1 = x * g.sphip + y * g.cphip;
            const float u1 = vx * g.cphip - vy * g.sphip, w1 = vx * g.sphip + vy * g.cphip;
            x = x1; y = y1; vx = u1; vy = w1; ++k;
        }
#pragma unroll 1
        while (y * g.cphip > x * g.sphip && guard++ < 32) {
            const float x1 = x * g.cphip + y * g.sphip, y1 = -x * g.sphip + y * g.cphip;
            const float u1 = vx * g.cphip + vy * g.sphip, w1 = -vx * g.sphip + vy * g.cphip;
            x = x1; y = y1; vx = u1; vy = w1; --k;
        }
        if (k != 0) sincosf((float)k * g.phip, &m.s, &m.c);
    } else {
        const float sgn = x < 0.0f ? -1.0f : 1.0f;
        const float r = fsqrt(r2);
        float phi = atanf(y / x);
        float rot;
        if (phi < 0.0f) {
            const float k = ceilf(-phi / g.phip);
            phi = phi + k * g.phip;
            rot = k * g.phip;
        } else if (phi > g.phip) {
            const float k = ceilf((phi - g.phip) / g.phip);
            phi = phi - k * g.phip;
            rot = -k * g.phip;
        } else {
            rot = 0.0f;
        }
        if (!(phi == phi)) { phi = 0.0f; rot = 0.0f; }
        float sr, cr, sp, cp;
        __sincosf(rot, &sr, &cr);
        __sincosf(phi, &sp, &cp);
        const float vx1 = vx * cr - vy * sr;
        const float vy1 = vx * sr + vy * cr;
        vx = vx1; vy = vy1;
        x = r * cp; y = r * sp;
        m.c = cr; m.s = sr; m.sgn = sgn;
    }
    if (z > g.lp) { z = 2.0f * g.lp - z; vz = -vz; m.zs = -m.zs; }
    if (z < 0.0f) { z = -z; vz = -vz; m.zs = -m.zs; }
    return m;
}

// Adjoint of the map: lambdas of the outputs in, of the inputs out.
__device__ __forceinline__ void bc_adjoint(const BcMap& m, float& lx, float& ly, float& lz, float& lvx,
                                           float& lvy, float& lvz)
{
    const float ax = m.sgn * (m.c * lx + m.s * ly), ay = m.sgn * (-m.s * lx + m.c * ly);
    const float bx = m.c * lvx + m.s * lvy, by = -m.s * lvx + m.c * lvy;
    lx = ax; ly = ay; lvx = bx; lvy = by;
    lz *= m.zs; lvz *= m.zs;
}

// Gather stencils of EM4.cu with the derivative of each upper weight (1 on the
// stencil's interior, 0 where the edge rules clamp it).
struct Gath {
    int i0, i1, k0, k1, h0, h1, j0, j1;   // Er/H in r (cells), Er in z (nodes), H/Ez in z (cells), Ez in r (nodes)
    float a0, a1, b0, b1, c0, c1, d0, d1;
    float da, db, dc, dd;
};

__device__ __forceinline__ void half_w(float s, int n, int& i0, int& i1, float& w0, float& w1, float& dw)
{
    half_axis(s, n, i0, i1, w0, w1);
    const int a = (int)floorf(s);
    dw = (a < 0 || a >= n - 1) ? 0.0f : 1.0f;
}

__device__ __forceinline__ void node_w(float s, int n, int& i0, int& i1, float& w0, float& w1, float& dw)
{
    node_axis(s, n, i0, i1, w0, w1);
    const int a = (int)floorf(s);
    dw = (a < 0 || a >= n) ? 0.0f : 1.0f;
}

__device__ __forceinline__ void gath(const Geom& g, float r, float z, Gath& w)
{
    const float srn = r * g.inv_dr, szn = z * g.inv_dz;
    half_w(srn - 0.5f, g.nr, w.i0, w.i1, w.a0, w.a1, w.da);
    node_w(szn, g.nz, w.k0, w.k1, w.b0, w.b1, w.db);
    half_w(szn - 0.5f, g.nz, w.h0, w.h1, w.c0, w.c1, w.dc);
    node_w(srn, g.nr, w.j0, w.j1, w.d0, w.d1, w.dd);
}

struct Fld {
    const float *er, *ez, *h;
};

// EM4 push (push_one's arithmetic on explicit gathers), intermediates kept.
struct PushT {
    float r, cph, sph, vr, Er, Ez, H, gm, ux, uy, uz, igp, fr, fz;
    Gath w;
};

__device__ __forceinline__ void push_fwd(const Geom& g, float qm, const Fld& f, float x, float y, float z,
                                         float& vx, float& vy, float& vz, PushT& t)
{
    const float r2 = x * x + y * y;
    t.cph = 1.0f; t.sph = 0.0f; t.r = 0.0f;
    if (r2 > 0.0f) {
        const float ir = frsq(r2);
        t.r = r2 * ir;
        t.cph = x * ir; t.sph = y * ir;
    }
    t.vr = vx * t.cph + vy * t.sph;
    gath(g, t.r, z, t.w);
    const Gath& w = t.w;
    const int nr = g.nr, ld = g.nr + 1;
    t.Er = w.b0 * (w.a0 * f.er[w.i0 + w.k0 * nr] + w.a1 * f.er[w.i1 + w.k0 * nr]) +
           w.b1 * (w.a0 * f.er[w.i0 + w.k1 * nr] + w.a1 * f.er[w.i1 + w.k1 * nr]);
    t.H = w.c0 * (w.a0 * f.h[w.i0 + w.h0 * nr] + w.a1 * f.h[w.i1 + w.h0 * nr]) +
          w.c1 * (w.a0 * f.h[w.i0 + w.h1 * nr] + w.a1 * f.h[w.i1 + w.h1 * nr]);
    t.Ez = w.c0 * (w.d0 * f.ez[w.j0 + w.h0 * ld] + w.d1 * f.ez[w.j1 + w.h0 * ld]) +
           w.c1 * (w.d0 * f.ez[w.j0 + w.h1 * ld] + w.d1 * f.ez[w.j1 + w.h1 * ld]);
    t.fr = qm * (t.Er - vz * g.mu0 * t.H);
    t.fz = qm * (t.Ez + t.vr * g.mu0 * t.H);
    const float v2 = vx * vx + vy * vy + vz * vz;
    t.gm = rsqrtf(1.0f - v2 * g.inv_c2);
    t.ux = t.gm * vx + (t.fr * t.cph) * g.dt;
    t.uy = t.gm * vy + (t.fr * t.sph) * g.dt;
    t.uz = t.gm * vz + t.fz * g.dt;
    t.igp = rsqrtf(1.0f + (t.ux * t.ux + t.uy * t.uy + t.uz * t.uz) * g.inv_c2);
    vx = t.ux * t.igp; vy = t.uy * t.igp; vz = t.uz * t.igp;
}

// Adjoint of push_fwd at input (x, y, z, v): lv holds the lambda of the pushed
// velocity on entry and of the input velocity on return; the position lambdas
// are added to lx, ly, lz; the lambdas of the gathered Er, Ez, H (to be
// spread over t.w's stencils) are returned in gEr, gEz, gH.
__device__ __forceinline__ void push_adj(const Geom& g, float qm, const Fld& f, float x, float y, float z,
                                         float vx, float vy, float vz, const PushT& t, float& lvx,
                                         float& lvy, float& lvz, float& lx, float& ly, float& lz,
                                         float& gEr, float& gEz, float& gH)
{
    const float ic2 = g.inv_c2, dt = g.dt;
    // v' = igp u
    const float udl = t.ux * lvx + t.uy * lvy + t.uz * lvz;
    const float k3 = t.igp * t.igp * t.igp * udl * ic2;
    const float lux = t.igp * lvx - k3 * t.ux, luy = t.igp * lvy - k3 * t.uy, luz = t.igp * lvz - k3 * t.uz;
    // u = gm v + dt (fr cph, fr sph, fz)
    const float lfr = dt * (lux * t.cph + luy * t.sph), lfz = dt * luz;
    float lcph = dt * lux * t.fr, lsph = dt * luy * t.fr;
    const float vdl = vx * lux + vy * luy + vz * luz;
    const float kg = t.gm * t.gm * t.gm * vdl * ic2;
    float ovx = t.gm * lux + kg * vx, ovy = t.gm * luy + kg * vy, ovz = t.gm * luz + kg * vz;
    // fr = qm (Er - vz mu0 H), fz = qm (Ez + vr mu0 H)
    const float lEr_g = qm * lfr, lEz_g = qm * lfz;
    const float lH_g = qm * g.mu0 * (-vz * lfr + t.vr * lfz);
    ovz += -qm * g.mu0 * t.H * lfr;
    const float lvr = qm * g.mu0 * t.H * lfz;
    ovx += t.cph * lvr; ovy += t.sph * lvr;
    lcph += vx * lvr; lsph += vy * lvr;
    lvx = ovx; lvy = ovy; lvz = ovz;
    // gathers
    const Gath& w = t.w;
    const int nr = g.nr, ld = g.nr + 1;
    const float er00 = f.er[w.i0 + w.k0 * nr], er10 = f.er[w.i1 + w.k0 * nr];
    const float er01 = f.er[w.i0 + w.k1 * nr], er11 = f.er[w.i1 + w.k1 * nr];
    const float h00 = f.h[w.i0 + w.h0 * nr], h10 = f.h[w.i1 + w.h0 * nr];
    const float h01 = f.h[w.i0 + w.h1 * nr], h11 = f.h[w.i1 + w.h1 * nr];
    const float ez00 = f.ez[w.j0 + w.h0 * ld], ez10 = f.ez[w.j1 + w.h0 * ld];
    const float ez01 = f.ez[w.j0 + w.h1 * ld], ez11 = f.ez[w.j1 + w.h1 * ld];
    const float dEr_dr = w.da * (w.b0 * (er10 - er00) + w.b1 * (er11 - er01));
    const float dEr_dz = w.db * ((w.a0 * er01 + w.a1 * er11) - (w.a0 * er00 + w.a1 * er10));
    const float dH_dr = w.da * (w.c0 * (h10 - h00) + w.c1 * (h11 - h01));
    const float dH_dz = w.dc * ((w.a0 * h01 + w.a1 * h11) - (w.a0 * h00 + w.a1 * h10));
    const float dEz_dr = w.dd * (w.c0 * (ez10 - ez00) + w.c1 * (ez11 - ez01));
    const float dEz_dz = w.dc * ((w.d0 * ez01 + w.d1 * ez11) - (w.d0 * ez00 + w.d1 * ez10));
    const float lr = (lEr_g * dEr_dr + lH_g * dH_dr + lEz_g * dEz_dr) * g.inv_dr;
    lz += (lEr_g * dEr_dz + lH_g * dH_dz + lEz_g * dEz_dz) * g.inv_dz;
    if (t.r > 0.0f) {
        const float ir = 1.0f / t.r, ir3 = ir * ir * ir;
        lx += lr * x * ir + (lcph * y * y - lsph * x * y) * ir3;
        ly += lr * y * ir + (-lcph * x * y + lsph * x * x) * ir3;
    }
    gEr = lEr_g; gEz = lEz_g; gH = lH_g;
}

// EM7 deposit of weight w in FP64 A/m^2 (K = A FN e qnum), and its adjoint.
struct DepT {
    float r, vr, sr, sz;
    int i0, i1, k0, k1;
    float wr0, wr1, wz0, wz1, dwr, dwz;
};

__device__ __forceinline__ void dep_geom(const Geom& g, float x, float y, float z, float vx, float vy,
                                         DepT& d)
{
    const float r2 = x * x + y * y;
    d.r = sqrtf(r2);
    d.vr = d.r > 0.0f ? (vx * x + vy * y) / d.r : vx;
    d.sr = d.r * g.inv_dr - 0.5f;
    d.sz = z * g.inv_dz - 0.5f;
    half_w(d.sr, g.nr, d.i0, d.i1, d.wr0, d.wr1, d.dwr);
    half_w(d.sz, g.nz, d.k0, d.k1, d.wz0, d.wz1, d.dwz);
}

__device__ __forceinline__ double cell_vol(double vcoef, int i) { return vcoef * (double)(2 * i + 1); }

// EM7 deposit in int64 fixed point (dep_values / block_deposit below): each
// cell gets rint(W w v kq) with kq = A qnum 2^32 / c, and the field cycle
// reads J = FN e c 2^-32 Q / V(i) (= A FN e qnum w W v / V).  Integer sums
// make every replay bitwise reproducible; the adjoint differentiates the
// unrounded map (the rounding is below 2^-32 c of each contribution).

// Adds the deposit's lambdas (of position, velocity, weight) given the current
// lambdas ljr, ljz; returns dW/dA's share (<lambda_J, J> / A) of this particle.
__device__ __forceinline__ double dep_adj(const Geom& g, double K, double A, double vcoef, float x, float y,
                                          float z, float vx, float vy, float vz, float w, const double* ljr,
                                          const double* ljz, float& lx, float& ly, float& lz, float& lvx,
                                          float& lvy, float& lvz, float& lw)
{
    DepT d;
    dep_geom(g, x, y, z, vx, vy, d);
    const int ci[4] = {d.i0, d.i1, d.i0, d.i1}, ck[4] = {d.k0, d.k0, d.k1, d.k1};
    const float wr[4] = {d.wr0, d.wr1, d.wr0, d.wr1}, wz[4] = {d.wz0, d.wz0, d.wz1, d.wz1};
    const float dr_[4] = {-d.dwr, d.dwr, -d.dwr, d.dwr}, dz_[4] = {-d.dwz, -d.dwz, d.dwz, d.dwz};
    double gr = 0.0, gz = 0.0, gsr = 0.0, gsz = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int c = ci[a] + ck[a] * g.nr;
        const double iv = 1.0 / cell_vol(vcoef, ci[a]);
        const double lr_ = ljr[c] * iv, lz_ = ljz[c] * iv;
        gr += (double)(wr[a] * wz[a]) * lr_;
        gz += (double)(wr[a] * wz[a]) * lz_;
        const double mix = (double)d.vr * lr_ + (double)vz * lz_;
        gsr += (double)(dr_[a] * wz[a]) * mix;
        gsz += (double)(wr[a] * dz_[a]) * mix;
    }
    const double Kw = K * (double)w;
    lw += (float)(K * ((double)d.vr * gr + (double)vz * gz));
    lvz += (float)(Kw * gz);
    const float lvr = (float)(Kw * gr);
    const float lsr = (float)(Kw * gsr), lsz = (float)(Kw * gsz);
    lz += lsz * g.inv_dz;
    if (d.r > 0.0f) {
        const float ir = 1.0f / d.r;
        lvx += lvr * x * ir;
        lvy += lvr * y * ir;
        const float lrr = lsr * g.inv_dr;
        lx += lvr * (vx * ir - d.vr * x * ir * ir) + lrr * x * ir;
        ly += lvr * (vy * ir - d.vr * y * ir * ir) + lrr * y * ir;
    } else {
        lvx += lvr;
    }
    return Kw * ((double)d.vr * gr + (double)vz * gz) / A;
}

// Density profile: nodes at (j + 1/2) Rp / nn, linear between, constant outside.
__device__ __forceinline__ float prof(const float* nodes, int nn, float r, float rp, int& j0, float& f,
                                      float& slope)
{
    const float s = r * (float)nn / rp - 0.5f;
    if (s <= 0.0f) { j0 = 0; f = 0.0f; slope = 0.0f; return nodes[0]; }
    if (s >= (float)(nn - 1)) { j0 = nn - 2; f = 1.0f; slope = 0.0f; return nodes[nn - 1]; }
    j0 = (int)floorf(s);
    if (j0 > nn - 2) j0 = nn - 2;
    f = s - (float)j0;
    slope = (nodes[j0 + 1] - nodes[j0]) * (float)nn / rp;
    return nodes[j0] + f * (nodes[j0 + 1] - nodes[j0]);
}

// Per-particle state arrays of the differentiable model (SoA, capacity n).
struct DState {
    float *x, *y, *z, *vx, *vy, *vz, *w;
};

// Frozen events of one step.
struct DEvents {
    float *dvx, *dvy, *dvz, *p0;
    uint8_t* col;
};

struct DiffArgs {
    Geom g;
    int32_t n;             // slots
    float qm;
    double K, A, vcoef;    // adjoint: J = K w W v / (vcoef (2i+1)), K = A FN e qnum
    double kq;             // forward fixed point: A qnum 2^32 / c
    const float* nodes;    // current profile (device, nn)
    const float* nodes0;   // reference profile
    int32_t nn;
    int32_t deposit;       // deposit J_{s+1} (not after the last step)
};

// ---- workgroup-tiled accumulation ---------------------------------------------
// A workgroup's particles are mostly neighbours (the initial order is by
// cell), so their stencils cover a small box of cells: sums go to an LDS tile
// over that box first and then one global add per touched cell, instead of
// one global atomic per particle and cell.  A box larger than the tile falls
// back to global atomics.
constexpr int kTileCells = 1024;

struct Box {
    int ilo, klo, wi, nk, ok;
};

// Bounding box [ilo, ihi] x [klo, khi] over the active lanes; every thread of
// the workgroup calls it.
__device__ Box block_box(bool act, int ilo, int ihi, int klo, int khi, int* s4)
{
    int a = act ? ilo : INT_MAX, b = act ? ihi : INT_MIN, c = act ? klo : INT_MAX, d = act ? khi : INT_MIN;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a = min(a, __shfl_xor(a, o));
        b = max(b, __shfl_xor(b, o));
        c = min(c, __shfl_xor(c, o));
        d = max(d, __shfl_xor(d, o));
    }
    if (threadIdx.x == 0) { s4[0] = INT_MAX; s4[1] = INT_MIN; s4[2] = INT_MAX; s4[3] = INT_MIN; }
    __syncthreads();
    if ((threadIdx.x & (kWaveSize - 1)) == 0) {
        atomicMin(&s4[0], a); atomicMax(&s4[1], b);
        atomicMin(&s4[2], c); atomicMax(&s4[3], d);
    }
    __syncthreads();
    Box bx;
    const bool any = s4[1] >= s4[0];
    bx.ilo = s4[0]; bx.klo = s4[2];
    bx.wi = any ? s4[1] - s4[0] + 1 : 0;
    bx.nk = any ? s4[3] - s4[2] + 1 : 0;
    bx.ok = any && (long long)bx.wi * bx.nk <= kTileCells;
    __syncthreads();   // s4 is reused by the next call
    return bx;
}

// Fixed-point deposit contributions of one particle.
struct DepQ {
    int i0, i1, k0, k1;
    long long qa[4], qb[4];   // cells (i0,k0) (i1,k0) (i0,k1) (i1,k1)
};

__device__ __forceinline__ void dep_values(const Geom& g, double kq, float x, float y, float z, float vx,
                                           float vy, float vz, float w, DepQ& q)
{
    DepT d;
    dep_geom(g, x, y, z, vx, vy, d);
    q.i0 = d.i0; q.i1 = d.i1; q.k0 = d.k0; q.k1 = d.k1;
    const float ww[4] = {d.wr0 * d.wz0, d.wr1 * d.wz0, d.wr0 * d.wz1, d.wr1 * d.wz1};
    const double sr = kq * (double)w * (double)d.vr, sz = kq * (double)w * (double)vz;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        q.qa[a] = llrint((double)ww[a] * sr);
        q.qb[a] = llrint((double)ww[a] * sz);
    }
}

// Adds every active lane's contributions to qr, qz (one copy, nr x nz).
__device__ void block_deposit(const Geom& g, bool act, const DepQ& q, unsigned long long* qr,
                              unsigned long long* qz, unsigned long long* s_qr, unsigned long long* s_qz,
                              int* s4)
{
    const Box bx = block_box(act, q.i0, q.i1, q.k0, q.k1, s4);
    if (bx.ok) {
        const int n = bx.wi * bx.nk;
        for (int c = threadIdx.x; c < n; c += blockDim.x) { s_qr[c] = 0ull; s_qz[c] = 0ull; }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int c = ((a & 1) ? q.i1 : q.i0) - bx.ilo + (((a & 2) ? q.k1 : q.k0) - bx.klo) * bx.wi;
                if (q.qa[a]) atomicAdd(&s_qr[c], (unsigned long long)q.qa[a]);
                if (q.qb[a]) atomicAdd(&s_qz[c], (unsigned long long)q.qb[a]);
            }
        }
        __syncthreads();
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const int gc = bx.ilo + c % bx.wi + (bx.klo + c / bx.wi) * g.nr;
            if (s_qr[c]) atomicAdd(&qr[gc], s_qr[c]);
            if (s_qz[c]) atomicAdd(&qz[gc], s_qz[c]);
        }
    } else if (act) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = ((a & 1) ? q.i1 : q.i0) + ((a & 2) ? q.k1 : q.k0) * g.nr;
            if (q.qa[a]) atomicAdd(&qr[c], (unsigned long long)q.qa[a]);
            if (q.qb[a]) atomicAdd(&qz[c], (unsigned long long)q.qb[a]);
        }
    }
}

// ---- prologue: half drift + deposit -----------------------------------------
__global__ __launch_bounds__(kDiffBlock) void k_diff_prologue(DiffArgs a, DState in, DState out,
                                                              unsigned long long* qr, unsigned long long* qz)
{
    __shared__ unsigned long long s_qr[kTileCells], s_qz[kTileCells];
    __shared__ int s4[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t jo = (size_t)(blockIdx.x % kJCopies) * a.g.nr * a.g.nz;   // this workgroup's current copy
    bool dep = false;
    DepQ q{};
    if (i < a.n) {
        float x = in.x[i], y = in.y[i], z = in.z[i], vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
        const float w = in.w[i];
        if (!is_dead(z)) {
            half_drift(a.g, x, y, z, vx, vy, vz);
            const BcMap m = bc_apply(a.g, x, y, z, vx, vy, vz);
            if (m.dead) z = __int_as_float(0x7fc00000);
            else { dep = true; dep_values(a.g, a.kq, x, y, z, vx, vy, vz, w, q); }
        }
        out.x[i] = x; out.y[i] = y; out.z[i] = z; out.vx[i] = vx; out.vy[i] = vy; out.vz[i] = vz; out.w[i] = w;
    }
    block_deposit(a.g, dep, q, qr + jo, qz + jo, s_qr, s_qz, s4);
}

// ---- one step (record: live collisions at theta0; replay: frozen events) ----
template <bool kRecord>
__global__ __launch_bounds__(kDiffBlock) void k_diff_step(DiffArgs a, CollArgs ca, DState in, DState out,
                                                          Fld f, DEvents ev, unsigned long long* qr,
                                                          unsigned long long* qz)
{
    __shared__ uint32_t s_tab_raw[kRecord ? sizeof(SmallTabs) / 4 : 1];
    __shared__ unsigned long long s_qr[kTileCells], s_qz[kTileCells];
    __shared__ int s4[4];
    if (kRecord) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&ca.tab->s);
        for (int i = threadIdx.x; i < (int)(sizeof(SmallTabs) / 4); i += blockDim.x) s_tab_raw[i] = src[i];
        __syncthreads();
    }
    const SmallTabs* s_tab = reinterpret_cast<const SmallTabs*>(s_tab_raw);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t jo = (size_t)(blockIdx.x % kJCopies) * a.g.nr * a.g.nz;   // this workgroup's current copy
    bool dep = false;
    DepQ q{};
    if (i < a.n) {
        float x = in.x[i], y = in.y[i], z = in.z[i], vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
        float w = in.w[i];
        if (!is_dead(z)) {
            PushT t;
            push_fwd(a.g, a.qm, f, x, y, z, vx, vy, vz, t);
            half_drift(a.g, x, y, z, vx, vy, vz);
            BcMap m = bc_apply(a.g, x, y, z, vx, vy, vz);
            if (m.dead) {
                z = __int_as_float(0x7fc00000);
                if (kRecord) { ev.dvx[i] = ev.dvy[i] = ev.dvz[i] = 0.0f; ev.p0[i] = 0.0f; ev.col[i] = 0; }
            } else {
                float dvx = 0.0f, dvy = 0.0f, dvz = 0.0f, p0 = 0.0f;
                int col = 0;
                int j0;
                float fr_, sl;
                const float r = fsqrt(x * x + y * y);
                const float n0r = prof(a.nodes0, a.nn, r, a.g.rp, j0, fr_, sl);
                if (kRecord) {
                    // p_T of electron_transport.cu:74-77 with the profile density
                    const float v2 = vx * vx + vy * vy + vz * vz;
                    const float T = ev_from_speed2(v2, a.g.inv_c2, ca.ev_per_mc2_em);
                    if (T >= 15.6f && T <= 20000.0f) {
                        float se, si, sx;
                        cross_sections(s_tab, T, se, si, sx);
                        p0 = fsqrt(v2) * ca.dt_coll * n0r * (se + si + sx);
                    }
                    // the production step's counters: u1 from the decision block of
                    // slot i's group of four, u2..u4 from block 0 of slot i
                    const U4 D = philox4x32_10((uint32_t)i >> 2, ca.step, ca.run, kDecisionBlock, ca.key0, ca.key1);
                    const uint32_t dw[4] = {D.a, D.b, D.c, D.d};
                    if (u01(dw[i & 3]) < p0) {
                        const U4 R = philox4x32_10((uint32_t)i, ca.step, ca.run, 0u, ca.key0, ca.key1);
                        float cvx = vx, cvy = vy, cvz = vz;
                        CollOut co;
                        collide_candidate(a.g, ca, s_tab, (uint32_t)i, x, y, z, cvx, cvy, cvz, 0.0f, u01(R.b),
                                          u01(R.c), u01(R.d), false, co);
                        col = (co.ev & kEvMask) != 0;
                        dvx = cvx - vx; dvy = cvy - vy; dvz = cvz - vz;
                    }
                    ev.dvx[i] = dvx; ev.dvy[i] = dvy; ev.dvz[i] = dvz; ev.p0[i] = p0; ev.col[i] = (uint8_t)col;
                } else {
                    dvx = ev.dvx[i]; dvy = ev.dvy[i]; dvz = ev.dvz[i]; p0 = ev.p0[i]; col = ev.col[i];
                }
                vx += dvx; vy += dvy; vz += dvz;
                // likelihood ratio of the frozen decision
                if (p0 < 1.0f && n0r > 0.0f) {
                    const float nr_ = prof(a.nodes, a.nn, r, a.g.rp, j0, fr_, sl);
                    const float qq = nr_ / n0r;
                    w *= col ? qq : (1.0f - p0 * qq) / (1.0f - p0);
                }
                half_drift(a.g, x, y, z, vx, vy, vz);
                m = bc_apply(a.g, x, y, z, vx, vy, vz);
                if (m.dead) z = __int_as_float(0x7fc00000);
                else if (a.deposit) { dep = true; dep_values(a.g, a.kq, x, y, z, vx, vy, vz, w, q); }
            }
        }
        out.x[i] = x; out.y[i] = y; out.z[i] = z; out.vx[i] = vx; out.vy[i] = vy; out.vz[i] = vz; out.w[i] = w;
    }
    if (a.deposit) block_deposit(a.g, dep, q, qr + jo, qz + jo, s_qr, s_qz, s4);
}

// ---- adjoint of one step -------------------------------------------------------
// Recomputes step s from its input state and applies the transposed maps in
// reverse.  lam_in: lambdas of the step's output state (x^{1/2}_{s+1}, v, w)
// (without J_{s+1}'s share, added here from ljr, ljz); lam_out: lambdas of the
// input state.  Field lambdas of F_{s+1} are scattered into copy
// blockIdx % kJCopies of lEr, lEz, lH (copies of the full arrays, back to
// back), which spreads the atomics of the dense spot over as many addresses.
struct Acc {
    double* gnodes;   // nn
    double* gA;       // 1
};

__global__ __launch_bounds__(kDiffBlock) void k_diff_adj_step(DiffArgs a, DState in, Fld f, DEvents ev,
                                                              DState lam_in, DState lam_out,
                                                              const double* ljr, const double* ljz,
                                                              double* lEr, double* lEz, double* lH, Acc acc)
{
    __shared__ double s_g[kMaxNodes + 1];
    __shared__ double s_er[kTileCells], s_ez[kTileCells], s_h[kTileCells];
    __shared__ int s4[4];
    for (int j = threadIdx.x; j <= kMaxNodes; j += blockDim.x) s_g[j] = 0.0;
    __syncthreads();
    {
        const size_t cp = blockIdx.x % kJCopies, nr = a.g.nr, nz = a.g.nz;
        lEr += cp * nr * (nz + 1);
        lEz += cp * (nr + 1) * nz;
        lH += cp * nr * nz;
    }
    // the gathered fields' lambdas of this lane and their stencils
    bool scat = false;
    Gath gw{};
    float gEr = 0.0f, gEz = 0.0f, gH = 0.0f;
    double gA = 0.0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) {
        float lx = 0.f, ly = 0.f, lz = 0.f, lvx = 0.f, lvy = 0.f, lvz = 0.f, lw = 0.f;
        const float x0 = in.x[i], y0 = in.y[i], z0 = in.z[i], vx0 = in.vx[i], vy0 = in.vy[i], vz0 = in.vz[i];
        const float w0 = in.w[i];
        if (!is_dead(z0)) {
            // forward recompute
            float x = x0, y = y0, z = z0, vx = vx0, vy = vy0, vz = vz0;
            PushT t;
            push_fwd(a.g, a.qm, f, x, y, z, vx, vy, vz, t);
            half_drift(a.g, x, y, z, vx, vy, vz);
            const BcMap m1 = bc_apply(a.g, x, y, z, vx, vy, vz);
            if (!m1.dead) {
                const float x1 = x, y1 = y, z1 = z;
                const float dvx = ev.dvx[i], dvy = ev.dvy[i], dvz = ev.dvz[i], p0 = ev.p0[i];
                const int col = ev.col[i];
                vx += dvx; vy += dvy; vz += dvz;
                const float r1 = fsqrt(x1 * x1 + y1 * y1);
                int j0; float fr_, sl0, sl;
                const float n0r = prof(a.nodes0, a.nn, r1, a.g.rp, j0, fr_, sl0);
                const float nr_ = prof(a.nodes, a.nn, r1, a.g.rp, j0, fr_, sl);
                float L = 1.0f, dLdq = 0.0f;
                const bool lr = p0 < 1.0f && n0r > 0.0f;
                if (lr) {
                    const float q = nr_ / n0r;
                    L = col ? q : (1.0f - p0 * q) / (1.0f - p0);
                    dLdq = col ? 1.0f : -p0 / (1.0f - p0);
                }
                const float w1 = w0 * L;
                half_drift(a.g, x, y, z, vx, vy, vz);
                const BcMap m2 = bc_apply(a.g, x, y, z, vx, vy, vz);
                // lambdas of the output state
                float mx = 0.f, my = 0.f, mz = 0.f, mvx = 0.f, mvy = 0.f, mvz = 0.f, mw = 0.f;
                if (!m2.dead) {
                    mx = lam_in.x[i]; my = lam_in.y[i]; mz = lam_in.z[i];
                    mvx = lam_in.vx[i]; mvy = lam_in.vy[i]; mvz = lam_in.vz[i]; mw = lam_in.w[i];
                    if (a.deposit)
                        gA = dep_adj(a.g, a.K, a.A, a.vcoef, x, y, z, vx, vy, vz, w1, ljr, ljz, mx, my, mz, mvx,
                                     mvy, mvz, mw);
                    // second half drift + BC
                    bc_adjoint(m2, mx, my, mz, mvx, mvy, mvz);
                    const float h = 0.5f * a.g.dt;
                    mvx += h * mx; mvy += h * my; mvz += h * mz;
                }
                // weight: w1 = w0 L(q(r1)), q = n(r1)/n0(r1)
                if (lr && mw != 0.0f) {
                    // d W / d n(r1), in FP64: of order W / n, below FP32's range
                    const double cw = (double)mw * (double)w0 * (double)dLdq / (double)n0r;
                    atomicAdd(&s_g[j0], cw * (double)(1.0f - fr_));
                    atomicAdd(&s_g[j0 + 1], cw * (double)fr_);
                    // position: dq/dr = (sl - q sl0) / n0
                    const float dqdr = (sl - (nr_ / n0r) * sl0) / n0r;
                    if (r1 > 0.0f) {
                        const float lr1 = mw * w0 * dLdq * dqdr / r1;
                        mx += lr1 * x1; my += lr1 * y1;
                    }
                }
                lw = mw * L;
                // kick: identity on v; first half drift + BC
                bc_adjoint(m1, mx, my, mz, mvx, mvy, mvz);
                const float h = 0.5f * a.g.dt;
                mvx += h * mx; mvy += h * my; mvz += h * mz;
                // push
                lx = mx; ly = my; lz = mz;
                lvx = mvx; lvy = mvy; lvz = mvz;
                push_adj(a.g, a.qm, f, x0, y0, z0, vx0, vy0, vz0, t, lvx, lvy, lvz, lx, ly, lz, gEr, gEz, gH);
                gw = t.w;
                scat = true;
            }
        }
        lam_out.x[i] = lx; lam_out.y[i] = ly; lam_out.z[i] = lz;
        lam_out.vx[i] = lvx; lam_out.vy[i] = lvy; lam_out.vz[i] = lvz; lam_out.w[i] = lw;
    }
    // dW/dA: wave sums, then one LDS add per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) gA += __shfl_xor(gA, o);
    if ((threadIdx.x & (kWaveSize - 1)) == 0 && gA != 0.0) atomicAdd(&s_g[kMaxNodes], gA);
    // field lambdas: Er on (cells i, nodes k), H on (cells i, cells h), Ez on (nodes j, cells h)
    const int nr = a.g.nr;
    const Box bx = block_box(scat, min(gw.i0, gw.j0), max(gw.i1, gw.j1), min(gw.k0, gw.h0), max(gw.k1, gw.h1), s4);
    const int eri[4] = {gw.i0, gw.i1, gw.i0, gw.i1}, erk[4] = {gw.k0, gw.k0, gw.k1, gw.k1};
    const int hk[4] = {gw.h0, gw.h0, gw.h1, gw.h1}, ezj[4] = {gw.j0, gw.j1, gw.j0, gw.j1};
    const float wer[4] = {gw.a0 * gw.b0, gw.a1 * gw.b0, gw.a0 * gw.b1, gw.a1 * gw.b1};
    const float wh[4] = {gw.a0 * gw.c0, gw.a1 * gw.c0, gw.a0 * gw.c1, gw.a1 * gw.c1};
    const float wez[4] = {gw.d0 * gw.c0, gw.d1 * gw.c0, gw.d0 * gw.c1, gw.d1 * gw.c1};
    if (bx.ok) {
        const int n = bx.wi * bx.nk;
        for (int c = threadIdx.x; c < n; c += blockDim.x) { s_er[c] = 0.0; s_ez[c] = 0.0; s_h[c] = 0.0; }
        __syncthreads();
        if (scat) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                atomicAdd(&s_er[eri[q] - bx.ilo + (erk[q] - bx.klo) * bx.wi], (double)(wer[q] * gEr));
                atomicAdd(&s_h[eri[q] - bx.ilo + (hk[q] - bx.klo) * bx.wi], (double)(wh[q] * gH));
                atomicAdd(&s_ez[ezj[q] - bx.ilo + (hk[q] - bx.klo) * bx.wi], (double)(wez[q] * gEz));
            }
        }
        __syncthreads();
        // a non-zero tile entry was added at an index that exists in that array
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const int ii = bx.ilo + c % bx.wi, kk = bx.klo + c / bx.wi;
            if (s_er[c] != 0.0) atomicAdd(&lEr[ii + kk * nr], s_er[c]);
            if (s_h[c] != 0.0) atomicAdd(&lH[ii + kk * nr], s_h[c]);
            if (s_ez[c] != 0.0) atomicAdd(&lEz[ii + kk * (nr + 1)], s_ez[c]);
        }
    } else if (scat) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            atomicAdd(&lEr[eri[q] + erk[q] * nr], (double)(wer[q] * gEr));
            atomicAdd(&lH[eri[q] + hk[q] * nr], (double)(wh[q] * gH));
            atomicAdd(&lEz[ezj[q] + hk[q] * (nr + 1)], (double)(wez[q] * gEz));
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < a.nn; j += blockDim.x)
        if (s_g[j] != 0.0) atomicAdd(&acc.gnodes[j], s_g[j]);
    if (threadIdx.x == 0 && s_g[kMaxNodes] != 0.0) atomicAdd(acc.gA, s_g[kMaxNodes]);
}

// Adjoint of the prologue's deposit (dW/dA only: the initial state is fixed).
__global__ __launch_bounds__(kDiffBlock) void k_diff_adj_prologue(DiffArgs a, DState st, const double* ljr,
                                                                  const double* ljz, Acc acc)
{
    __shared__ double s_a;
    if (threadIdx.x == 0) s_a = 0.0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n && !is_dead(st.z[i])) {
        float lx = 0.f, ly = 0.f, lz = 0.f, lvx = 0.f, lvy = 0.f, lvz = 0.f, lw = 0.f;
        const double gA = dep_adj(a.g, a.K, a.A, a.vcoef, st.x[i], st.y[i], st.z[i], st.vx[i], st.vy[i],
                                  st.vz[i], st.w[i], ljr, ljz, lx, ly, lz, lvx, lvy, lvz, lw);
        atomicAdd(&s_a, gA);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_a != 0.0) atomicAdd(acc.gA, s_a);
}

// W (added to w_out) and, when lEr is given, dW/dEr into lEr and max |dW/dEr|
// into amax (as the bits of a non-negative double, which order like integers).
__global__ void k_adj_objective(int nr, int nz, double coef, double dr, int i_min, const double* er, double* lEr,
                                double* w_out, unsigned long long* amax)
{
    __shared__ double s_w[256], s_m[256];
    double acc = 0.0, mx = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nr * (nz + 1); c += gridDim.x * blockDim.x) {
        const int i = c % nr;
        double l = 0.0;
        if (i >= i_min) {
            const double r = (i + 0.5) * dr;
            acc += 0.5 * coef * er[c] * er[c] * r;
            l = coef * er[c] * r;
        }
        if (lEr) lEr[c] = l;
        mx = fmax(mx, fabs(l));
    }
    s_w[threadIdx.x] = acc;
    s_m[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_w[threadIdx.x] += s_w[threadIdx.x + o];
            s_m[threadIdx.x] = fmax(s_m[threadIdx.x], s_m[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(w_out, s_w[0]);
        if (amax) atomicMax(amax, (unsigned long long)__double_as_longlong(s_m[0]));
    }
}

__global__ void k_scale_f64(double* p, int64_t n, double s)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] *= s;
}

// dst += sum of the kJCopies copies (stride n); the copies are zeroed.
__global__ void k_fold_copies(double* dst, double* copies, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = dst[i];
#pragma unroll
        for (int k = 0; k < kJCopies; ++k) {
            s += copies[k * n + i];
            copies[k * n + i] = 0.0;
        }
        dst[i] = s;
    }
}

__global__ void k_fill_f64(double* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

}  // namespace gasx

// =====================================================================================
// Host side
// =====================================================================================

// State of one differentiable run: the recorded events, checkpoints and the
// work buffers of the forward and adjoint passes (device memory, owned).
struct DiffState {
    int n = 0, steps = 0, every = 0, nn = 0, nseg = 0;
    double r_frac = 0.9, amp = 1.0;
    float qnum = -1.0f;
    double mass = 0.0;
    uint64_t seed = 0;
    uint32_t run = 0;
    std::vector<double> nodes0, nodes;
    std::vector<void*> owned;
    float *d_nodes = nullptr, *d_nodes0 = nullptr, *d_one = nullptr;
    float* init[7] = {};                 // initial particles x..vz and weight 1
    float *dvx = nullptr, *dvy = nullptr, *dvz = nullptr, *p0 = nullptr;   // [step][slot]
    uint8_t* col = nullptr;
    std::vector<float*> ck_p;            // nseg x 7: state x^{1/2}_{s0} at each checkpoint
    std::vector<double*> ck_f;           // (nseg + 1) x 3: F_{s0} (er, ez, h)
    std::vector<unsigned long long*> ck_q;   // nseg x 2: fixed-point current J_{s0}
    std::vector<float*> seg_p;           // (every + 1) x 7: the states of one segment
    std::vector<float*> seg_f;           // every x 3: FP32 gather copies of F_{s+1}
    double *F[3] = {}, *Fb[3] = {};      // FP64 fields and their ping-pong partner
    unsigned long long *qr = nullptr, *qz = nullptr;   // fixed-point current, kJCopies copies of nr*nz
    double *lF[3] = {}, *lFo[3] = {}, *ljr = nullptr, *ljz = nullptr;   // field lambdas (ping-pong)
    double* lFm[3] = {};                 // mid buffers of a split transposed cycle
    double* lFc[3] = {};                 // kJCopies copies of each field lambda (scatter targets)
    float* lam[2][7] = {};               // particle lambdas (ping-pong)
    double* acc = nullptr;               // nn node gradients, dW/dA, W
    bool have_forward = false;
};

namespace {

int blocks_for(int64_t n, int bs = kDiffBlock)
{
    const int64_t b = (n + bs - 1) / bs;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20));
}

template <typename T>
hipError_t own(DiffState* d, T** p, size_t n)
{
    hipError_t e = dalloc(p, n);
    if (e == hipSuccess) d->owned.push_back((void*)*p);
    else *p = nullptr;
    return e;
}

DState state_of(float* const* p) { return DState{p[0], p[1], p[2], p[3], p[4], p[5], p[6]}; }

DEvents events_of(const DiffState* d, int s)
{
    const size_t o = (size_t)s * d->n;
    return DEvents{d->dvx + o, d->dvy + o, d->dvz + o, d->p0 + o, d->col + o};
}

size_t nh_of(const gasx_ctx* ctx) { return (size_t)ctx->p.nr * ctx->p.nz; }
size_t ner_of(const gasx_ctx* ctx) { return (size_t)ctx->p.nr * (ctx->p.nz + 1); }
size_t nez_of(const gasx_ctx* ctx) { return (size_t)(ctx->p.nr + 1) * ctx->p.nz; }

DiffArgs diff_args(const gasx_ctx* ctx, const DiffState* d, bool deposit)
{
    DiffArgs a{};
    a.g = ctx->geom;
    a.n = d->n;
    a.qm = (float)((double)d->qnum * ctx->p.e_em / d->mass);
    a.K = d->amp * ctx->p.fn * ctx->p.e_em * (double)d->qnum;
    a.kq = d->amp * (double)d->qnum * 4294967296.0 / ctx->p.c;
    a.A = d->amp;
    a.vcoef = ctx->p.phip / 2 * ctx->p.dr * ctx->p.dr * ctx->p.dz;
    a.nodes = d->d_nodes;
    a.nodes0 = d->d_nodes0;
    a.nn = d->nn;
    a.deposit = deposit ? 1 : 0;
    return a;
}

gasx_status zero64(gasx_ctx* ctx, double* p, int64_t n)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(n)), dim3(kDiffBlock), 0, ctx->stream, p, n);
    HIPCHK(hipGetLastError(), "fill");
    return GASX_OK;
}

gasx_status copy_d(gasx_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "copy");
    return GASX_OK;
}

gasx_status copy_state(gasx_ctx* ctx, const DiffState* d, float* const* dst, float* const* src)
{
    for (int c = 0; c < 7; ++c) {
        gasx_status st = copy_d(ctx, dst[c], src[c], (size_t)d->n * sizeof(float));
        if (st != GASX_OK) return st;
    }
    return GASX_OK;
}

// fields (er, ez, h) in the order of a checkpoint
gasx_status copy_fields(gasx_ctx* ctx, double* const* dst, double* const* src)
{
    const size_t n[3] = {ner_of(ctx), nez_of(ctx), nh_of(ctx)};
    for (int c = 0; c < 3; ++c) {
        gasx_status st = copy_d(ctx, dst[c], src[c], n[c] * sizeof(double));
        if (st != GASX_OK) return st;
    }
    return GASX_OK;
}

gasx_status copy_current(gasx_ctx* ctx, unsigned long long* dr_, unsigned long long* dz_,
                         const unsigned long long* sr, const unsigned long long* sz)
{
    const size_t b = nh_of(ctx) * sizeof(unsigned long long);
    gasx_status st = copy_d(ctx, dr_, sr, b);
    return st == GASX_OK ? copy_d(ctx, dz_, sz, b) : st;
}

// copies = kJCopies: every copy; 1: copy 0 only
gasx_status zero_current(gasx_ctx* ctx, DiffState* d, int copies = kJCopies)
{
    const int64_t n = (int64_t)nh_of(ctx) * copies;
    HIPCHK(launch_fill_u64(d->qr, n, ctx->stream), "fill");
    HIPCHK(launch_fill_u64(d->qz, n, ctx->stream), "fill");
    return GASX_OK;
}

// sums the current copies into copy 0 (and zeroes the others)
gasx_status reduce_current(gasx_ctx* ctx, DiffState* d)
{
    HIPCHK(launch_j_reduce(d->qr, d->qz, (int64_t)nh_of(ctx), nh_of(ctx), ctx->stream), "current reduce");
    return GASX_OK;
}

// Collision arguments of the recording pass: the production tables and
// Philox keys; the gas is a unit density so that collide_candidate accepts
// every candidate handed to it (the decision is taken against p0 before).
gasx_status record_coll(gasx_ctx* ctx, const DiffState* d, CollArgs* ca)
{
    gasx_gas g{};
    g.density = d->d_one;
    g.gnr = 1; g.gnz = 1;
    g.gdr = 2.0 * ctx->p.nr * ctx->p.dr;
    g.gdz = 2.0 * ctx->p.nz * ctx->p.dz;
    g.t_scalar = 300.0;
    g.n_max = 1.0;
    return make_coll(ctx, &g, ctx->p.ddt, d->seed, 0, d->run, ca);
}

// Step s of the model: field cycle F -> F' with the current J_s (FP32 copies
// into f32), then the particle step in -> out depositing J_{s+1}.
gasx_status diff_step(gasx_ctx* ctx, DiffState* d, int s, bool record, const CollArgs& ca0,
                      float* const* in, float* const* out, float* const* f32, int& parity)
{
    gasx_status st0 = reduce_current(ctx, d);
    if (st0 != GASX_OK) return st0;
    FdtdRun r{};
    double* const* fin = parity ? d->Fb : d->F;
    double* const* fout = parity ? d->F : d->Fb;
    r.er_in = fin[0]; r.ez_in = fin[1]; r.h_in = fin[2];
    r.er_out = fout[0]; r.ez_out = fout[1]; r.h_out = fout[2];
    r.er32 = f32[0]; r.ez32 = f32[1]; r.h32 = f32[2];
    r.er_mid = ctx->er_m; r.ez_mid = ctx->ez_m; r.h_mid = ctx->h_m;
    r.qr = d->qr; r.qz = d->qz;
    HIPCHK(launch_fdtd(ctx->fa, r, nullptr, 1, 1, ctx->stream), "field cycle");
    parity ^= 1;
    const bool dep = s + 1 < d->steps;
    if (dep) {
        gasx_status st = zero_current(ctx, d, 1);   // the other copies were zeroed by the reduce
        if (st != GASX_OK) return st;
    }
    const DiffArgs a = diff_args(ctx, d, dep);
    CollArgs ca = ca0;
    ca.step = (uint32_t)s;
    const Fld f{f32[0], f32[1], f32[2]};
    if (record)
        hipLaunchKernelGGL(k_diff_step<true>, dim3(blocks_for(d->n)), dim3(kDiffBlock), 0, ctx->stream, a, ca,
                           state_of(in), state_of(out), f, events_of(d, s), d->qr, d->qz);
    else
        hipLaunchKernelGGL(k_diff_step<false>, dim3(blocks_for(d->n)), dim3(kDiffBlock), 0, ctx->stream, a, ca,
                           state_of(in), state_of(out), f, events_of(d, s), d->qr, d->qz);
    HIPCHK(hipGetLastError(), "differentiable step");
    return GASX_OK;
}

int objective_first_cell(const gasx_ctx* ctx, double r_frac)
{
    // first Er cell with (i + 1/2) dr >= r_frac Rp
    int i = 0;
    while (i < ctx->p.nr && (i + 0.5) < r_frac * ctx->p.nr) ++i;
    return i;
}

gasx_status launch_objective(gasx_ctx* ctx, DiffState* d, const double* er, double* lEr, double* w_out,
                             unsigned long long* amax = nullptr)
{
    const double coef = ctx->p.phip * ctx->p.eps0 * ctx->p.dr * ctx->p.dz;
    hipLaunchKernelGGL(k_adj_objective, dim3(64), dim3(256), 0, ctx->stream, ctx->p.nr, ctx->p.nz, coef,
                       ctx->p.dr, objective_first_cell(ctx, d->r_frac), er, lEr, w_out, amax);
    HIPCHK(hipGetLastError(), "objective");
    return GASX_OK;
}

// Forward pass from the initial state, checkpointing every `every` steps;
// record: live collisions at the reference profile (events written).
gasx_status diff_forward_pass(gasx_ctx* ctx, DiffState* d, bool record, double* w_out)
{
    const int64_t nh = (int64_t)nh_of(ctx);
    const int64_t fs[3] = {(int64_t)ner_of(ctx), (int64_t)nez_of(ctx), nh};
    gasx_status st = GASX_OK;
    for (int c = 0; c < 3 && st == GASX_OK; ++c) st = zero64(ctx, d->F[c], fs[c]);
    if (st == GASX_OK) st = zero_current(ctx, d);
    if (st != GASX_OK) return st;
    CollArgs ca{};
    if (record && (st = record_coll(ctx, d, &ca)) != GASX_OK) return st;
    float* const* work[2] = {&d->seg_p[0], &d->seg_p[7]};
    const DiffArgs a = diff_args(ctx, d, true);
    hipLaunchKernelGGL(k_diff_prologue, dim3(blocks_for(d->n)), dim3(kDiffBlock), 0, ctx->stream, a,
                       state_of(d->init), state_of(work[0]), d->qr, d->qz);
    HIPCHK(hipGetLastError(), "differentiable prologue");
    float* tmp32[3] = {ctx->er32, ctx->ez32, ctx->h32};
    int parity = 0, cur = 0;
    for (int g = 0; g < d->nseg; ++g) {
        const int s0 = g * d->every, s1 = std::min(d->steps, s0 + d->every);
        st = copy_state(ctx, d, &d->ck_p[7 * g], work[cur]);
        if (st == GASX_OK) st = reduce_current(ctx, d);
        if (st == GASX_OK) st = copy_fields(ctx, &d->ck_f[3 * g], parity ? d->Fb : d->F);
        if (st == GASX_OK) st = copy_current(ctx, d->ck_q[2 * g], d->ck_q[2 * g + 1], d->qr, d->qz);
        for (int s = s0; s < s1 && st == GASX_OK; ++s) {
            st = diff_step(ctx, d, s, record, ca, work[cur], work[cur ^ 1], tmp32, parity);
            cur ^= 1;
        }
        if (st != GASX_OK) return st;
    }
    st = copy_fields(ctx, &d->ck_f[3 * d->nseg], parity ? d->Fb : d->F);
    if (st == GASX_OK) st = zero64(ctx, d->acc, d->nn + 2);
    if (st == GASX_OK) st = launch_objective(ctx, d, d->ck_f[3 * d->nseg], nullptr, d->acc + d->nn + 1);
    if (st != GASX_OK) return st;
    double w = 0.0;
    HIPCHK(hipMemcpyAsync(&w, d->acc + d->nn + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    if (!std::isfinite(w)) return fail(ctx, GASX_E_NUMERIC, "non-finite objective");
    *w_out = w;
    return GASX_OK;
}

// Transposed field cycle: lF <- M^T lF (through lFo, then swapped) and
// (ljr, ljz) <- B^T lF.
gasx_status adj_cycle(gasx_ctx* ctx, DiffState* d)
{
    const AdjFdtdRun r{d->lF[0], d->lF[1], d->lF[2], d->lFo[0], d->lFo[1], d->lFo[2], d->ljr, d->ljz,
                       d->lFm[0], d->lFm[1], d->lFm[2]};
    HIPCHK(launch_fdtd_adj(ctx->fa, r, ctx->stream), "transposed field cycle");
    for (int c = 0; c < 3; ++c) std::swap(d->lF[c], d->lFo[c]);
    return GASX_OK;
}

gasx_status upload_nodes(gasx_ctx* ctx, float* dst, const std::vector<double>& v)
{
    std::vector<float> f(kMaxNodes, 0.0f);
    for (size_t j = 0; j < v.size(); ++j) f[j] = (float)v[j];
    HIPCHK(hipMemcpyAsync(dst, f.data(), kMaxNodes * sizeof(float), hipMemcpyHostToDevice, ctx->stream),
           "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return GASX_OK;
}

gasx_status diff_alloc(gasx_ctx* ctx, DiffState* d)
{
    const size_t n = (size_t)d->n, S = (size_t)d->steps, K = (size_t)d->every;
    const size_t nh = nh_of(ctx), fs[5] = {ner_of(ctx), nez_of(ctx), nh, nh, nh};
    HIPCHK(own(d, &d->d_nodes, kMaxNodes), "alloc");
    HIPCHK(own(d, &d->d_nodes0, kMaxNodes), "alloc");
    HIPCHK(own(d, &d->d_one, 1), "alloc");
    for (auto& p : d->init) HIPCHK(own(d, &p, n), "alloc initial state");
    HIPCHK(own(d, &d->dvx, S * n), "alloc events");
    HIPCHK(own(d, &d->dvy, S * n), "alloc events");
    HIPCHK(own(d, &d->dvz, S * n), "alloc events");
    HIPCHK(own(d, &d->p0, S * n), "alloc events");
    HIPCHK(own(d, &d->col, S * n), "alloc events");
    d->ck_p.assign(7 * (size_t)d->nseg, nullptr);
    for (auto& p : d->ck_p) HIPCHK(own(d, &p, n), "alloc checkpoints");
    d->ck_f.assign(3 * (size_t)(d->nseg + 1), nullptr);
    for (size_t i = 0; i < d->ck_f.size(); ++i) HIPCHK(own(d, &d->ck_f[i], fs[i % 3]), "alloc checkpoints");
    d->ck_q.assign(2 * (size_t)d->nseg, nullptr);
    for (auto& p : d->ck_q) HIPCHK(own(d, &p, nh), "alloc checkpoints");
    d->seg_p.assign(7 * (K + 1), nullptr);
    for (auto& p : d->seg_p) HIPCHK(own(d, &p, n), "alloc segment states");
    d->seg_f.assign(3 * K, nullptr);
    for (size_t i = 0; i < d->seg_f.size(); ++i) HIPCHK(own(d, &d->seg_f[i], fs[i % 3]), "alloc segment fields");
    for (int c = 0; c < 3; ++c) {
        HIPCHK(own(d, &d->F[c], fs[c]), "alloc fields");
        HIPCHK(own(d, &d->Fb[c], fs[c]), "alloc fields");
        HIPCHK(own(d, &d->lF[c], fs[c]), "alloc field lambdas");
        HIPCHK(own(d, &d->lFo[c], fs[c]), "alloc field lambdas");
        HIPCHK(own(d, &d->lFm[c], fs[c]), "alloc field lambdas");
    }
    HIPCHK(own(d, &d->qr, nh * kJCopies), "alloc current");
    HIPCHK(own(d, &d->qz, nh * kJCopies), "alloc current");
    for (int c = 0; c < 3; ++c) HIPCHK(own(d, &d->lFc[c], fs[c] * kJCopies), "alloc field lambda copies");
    HIPCHK(own(d, &d->ljr, nh), "alloc current lambdas");
    HIPCHK(own(d, &d->ljz, nh), "alloc current lambdas");
    for (auto& l : d->lam)
        for (auto& p : l) HIPCHK(own(d, &p, n), "alloc particle lambdas");
    HIPCHK(own(d, &d->acc, kMaxNodes + 2), "alloc accumulators");
    return GASX_OK;
}

}  // namespace

void gasx_host::diff_release(gasx_ctx* ctx)
{
    if (!ctx || !ctx->diff) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->diff->owned) (void)hipFree(p);
    delete ctx->diff;
    ctx->diff = nullptr;
}

extern "C" gasx_status gasx_diff_record(gasx_ctx* ctx, const gasx_species* e, const gasx_diff_spec* sp)
{
    if (!ctx) return GASX_E_ARG;
    if (!e || !sp || !e->x || !e->y || !e->z || !e->vx || !e->vy || !e->vz || !e->count || !(e->mass > 0))
        return fail(ctx, GASX_E_ARG, "gasx_diff_record: incomplete species");
    if (sp->n_nodes < 2 || sp->n_nodes > kMaxNodes || !sp->nodes0)
        return fail(ctx, GASX_E_ARG, "gasx_diff_record: n_nodes must be in [2, 64]");
    if (sp->n_steps < 1 || sp->checkpoint_every < 1 || !(sp->r_frac >= 0.0 && sp->r_frac < 1.0))
        return fail(ctx, GASX_E_ARG, "gasx_diff_record: bad steps, checkpoint interval or r_frac");
    for (int j = 0; j < sp->n_nodes; ++j)
        if (!(sp->nodes0[j] > 0.0) || !std::isfinite(sp->nodes0[j]))
            return fail(ctx, GASX_E_ARG, "gasx_diff_record: reference profile must be positive");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    int64_t n64 = 0;
    gasx_status st = gasx_get_count(ctx, e, &n64);
    if (st != GASX_OK) return st;
    if (n64 < 1 || n64 > e->capacity || n64 > (int64_t)1 << 28)
        return fail(ctx, GASX_E_ARG, "gasx_diff_record: particle count out of range");
    diff_release(ctx);
    DiffState* d = new (std::nothrow) DiffState();
    if (!d) return GASX_E_NOMEM;
    ctx->diff = d;
    d->n = (int)n64;
    d->steps = sp->n_steps;
    d->every = std::min(sp->checkpoint_every, sp->n_steps);
    d->nseg = (d->steps + d->every - 1) / d->every;
    d->nn = sp->n_nodes;
    d->r_frac = sp->r_frac;
    d->seed = sp->seed;
    d->run = sp->run_id;
    d->qnum = e->qnum;
    d->mass = e->mass;
    d->nodes0.assign(sp->nodes0, sp->nodes0 + sp->n_nodes);
    d->nodes = d->nodes0;
    st = diff_alloc(ctx, d);
    if (st == GASX_OK) {
        float* src[6] = {e->x, e->y, e->z, e->vx, e->vy, e->vz};
        for (int c = 0; c < 6 && st == GASX_OK; ++c) st = copy_d(ctx, d->init[c], src[c], (size_t)d->n * sizeof(float));
    }
    if (st == GASX_OK) {
        std::vector<float> ones((size_t)d->n, 1.0f);
        const hipError_t h1 = hipMemcpyAsync(d->init[6], ones.data(), ones.size() * sizeof(float),
                                             hipMemcpyHostToDevice, ctx->stream);
        const hipError_t h2 = hipMemcpyAsync(d->d_one, ones.data(), sizeof(float), hipMemcpyHostToDevice,
                                             ctx->stream);
        const hipError_t h3 = hipStreamSynchronize(ctx->stream);
        const hipError_t he = h1 != hipSuccess ? h1 : (h2 != hipSuccess ? h2 : h3);
        if (he != hipSuccess) st = hip_fail(ctx, he, "copy");
    }
    if (st == GASX_OK) st = upload_nodes(ctx, d->d_nodes0, d->nodes0);
    if (st == GASX_OK) st = upload_nodes(ctx, d->d_nodes, d->nodes0);
    double w = 0.0;
    if (st == GASX_OK) st = diff_forward_pass(ctx, d, true, &w);
    if (st != GASX_OK) {
        diff_release(ctx);
        return st;
    }
    d->amp = 1.0;
    d->have_forward = false;
    return GASX_OK;
}

extern "C" gasx_status gasx_diff_forward(gasx_ctx* ctx, const double* nodes, double amplitude, double* objective)
{
    if (!ctx) return GASX_E_ARG;
    DiffState* d = ctx->diff;
    if (!d) return fail(ctx, GASX_E_STATE, "gasx_diff_forward: no recorded run (gasx_diff_record)");
    if (!nodes || !objective || !(amplitude > 0.0) || !std::isfinite(amplitude))
        return fail(ctx, GASX_E_ARG, "gasx_diff_forward: bad profile, amplitude or output");
    for (int j = 0; j < d->nn; ++j)
        if (!(nodes[j] >= 0.0) || !std::isfinite(nodes[j]))
            return fail(ctx, GASX_E_ARG, "gasx_diff_forward: profile must be finite and non-negative");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    d->have_forward = false;
    d->nodes.assign(nodes, nodes + d->nn);
    d->amp = amplitude;
    gasx_status st = upload_nodes(ctx, d->d_nodes, d->nodes);
    if (st == GASX_OK) st = diff_forward_pass(ctx, d, false, objective);
    d->have_forward = st == GASX_OK;
    return st;
}

extern "C" gasx_status gasx_diff_backward(gasx_ctx* ctx, double* grad_nodes, double* grad_amplitude)
{
    if (!ctx) return GASX_E_ARG;
    DiffState* d = ctx->diff;
    if (!d || !d->have_forward) return fail(ctx, GASX_E_STATE, "gasx_diff_backward: call gasx_diff_forward first");
    if (!grad_nodes || !grad_amplitude) return fail(ctx, GASX_E_ARG, "gasx_diff_backward: null output");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const int64_t nh = (int64_t)nh_of(ctx);
    gasx_status st = zero64(ctx, d->acc, d->nn + 2);
    if (st == GASX_OK) st = zero64(ctx, d->lF[1], (int64_t)nez_of(ctx));
    if (st == GASX_OK) st = zero64(ctx, d->lF[2], nh);
    if (st == GASX_OK) st = zero64(ctx, d->ljr, nh);
    if (st == GASX_OK) st = zero64(ctx, d->ljz, nh);
    const int64_t fsz[3] = {(int64_t)ner_of(ctx), (int64_t)nez_of(ctx), nh};
    for (int c = 0; c < 3 && st == GASX_OK; ++c) st = zero64(ctx, d->lFc[c], fsz[c] * kJCopies);
    // dW/dF_S (Er only); W itself is recomputed into acc[nn + 1], max |dW/dEr| into acc[nn]'s bits
    unsigned long long* amax = reinterpret_cast<unsigned long long*>(d->acc + d->nn);
    if (st == GASX_OK) st = launch_objective(ctx, d, d->ck_f[3 * d->nseg], d->lF[0], d->acc + d->nn + 1, amax);
    if (st != GASX_OK) return st;
    // The adjoint is linear in its seed: scale it by a power of two S so that
    // max |seed| is in [1, 2), which keeps the FP32 particle sensitivities
    // clear of underflow whatever the objective's size; the results are
    // divided by S (exactly) at the end.
    double seed_max = 0.0;
    HIPCHK(hipMemcpyAsync(&seed_max, d->acc + d->nn, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    const double S = (seed_max > 0.0 && std::isfinite(seed_max)) ? std::ldexp(1.0, -std::ilogb(seed_max)) : 1.0;
    st = zero64(ctx, d->acc + d->nn, 1);
    if (st != GASX_OK) return st;
    hipLaunchKernelGGL(k_scale_f64, dim3(blocks_for((int64_t)ner_of(ctx))), dim3(kDiffBlock), 0, ctx->stream,
                       d->lF[0], (int64_t)ner_of(ctx), S);
    HIPCHK(hipGetLastError(), "seed scale");
    for (int c = 0; c < 7; ++c)
        HIPCHK(hipMemsetAsync(d->lam[0][c], 0, (size_t)d->n * sizeof(float), ctx->stream), "zero");
    const Acc acc{d->acc, d->acc + d->nn};
    const CollArgs ca{};
    int cur = 0;
    for (int g = d->nseg - 1; g >= 0; --g) {
        const int s0 = g * d->every, s1 = std::min(d->steps, s0 + d->every);
        // recompute the segment from its checkpoint, keeping every state and field copy
        st = copy_state(ctx, d, &d->seg_p[0], &d->ck_p[7 * g]);
        if (st == GASX_OK) st = zero_current(ctx, d);
        if (st == GASX_OK) st = copy_fields(ctx, d->F, &d->ck_f[3 * g]);
        if (st == GASX_OK) st = copy_current(ctx, d->qr, d->qz, d->ck_q[2 * g], d->ck_q[2 * g + 1]);
        int parity = 0;
        for (int s = s0; s < s1 && st == GASX_OK; ++s) {
            const int k = s - s0;
            st = diff_step(ctx, d, s, false, ca, &d->seg_p[7 * k], &d->seg_p[7 * (k + 1)], &d->seg_f[3 * k],
                           parity);
        }
        if (st != GASX_OK) return st;
        for (int s = s1 - 1; s >= s0; --s) {
            const int k = s - s0;
            // the last step's output state reaches no later field: its adjoint is zero
            if (s + 1 < d->steps) {
                const DiffArgs a = diff_args(ctx, d, true);
                const Fld f{d->seg_f[3 * k], d->seg_f[3 * k + 1], d->seg_f[3 * k + 2]};
                hipLaunchKernelGGL(k_diff_adj_step, dim3(blocks_for(d->n)), dim3(kDiffBlock), 0, ctx->stream, a,
                                   state_of(&d->seg_p[7 * k]), f, events_of(d, s), state_of(d->lam[cur]),
                                   state_of(d->lam[cur ^ 1]), d->ljr, d->ljz, d->lFc[0], d->lFc[1], d->lFc[2],
                                   acc);
                HIPCHK(hipGetLastError(), "adjoint step");
                cur ^= 1;
                for (int c = 0; c < 3; ++c)
                    hipLaunchKernelGGL(k_fold_copies, dim3(blocks_for(fsz[c])), dim3(kDiffBlock), 0, ctx->stream,
                                       d->lF[c], d->lFc[c], fsz[c]);
                HIPCHK(hipGetLastError(), "fold field lambdas");
            }
            st = adj_cycle(ctx, d);
            if (st != GASX_OK) return st;
        }
    }
    // the prologue's deposit of J_0 at x^{1/2}_0 (checkpoint 0's state)
    const DiffArgs a = diff_args(ctx, d, true);
    hipLaunchKernelGGL(k_diff_adj_prologue, dim3(blocks_for(d->n)), dim3(kDiffBlock), 0, ctx->stream, a,
                       state_of(&d->ck_p[0]), d->ljr, d->ljz, acc);
    HIPCHK(hipGetLastError(), "adjoint prologue");
    std::vector<double> h((size_t)d->nn + 2);
    HIPCHK(hipMemcpyAsync(h.data(), d->acc, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
           "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    for (int j = 0; j < d->nn; ++j) grad_nodes[j] = h[j] / S;
    *grad_amplitude = h[d->nn] / S;
    for (size_t j = 0; j < h.size(); ++j)
        if (!std::isfinite(h[j])) return fail(ctx, GASX_E_NUMERIC, "non-finite gradient");
    return GASX_OK;
}

extern "C" gasx_status gasx_diff_events(gasx_ctx* ctx, int32_t step, float* dvx, float* dvy, float* dvz,
                                        float* p0, uint8_t* collided)
{
    if (!ctx) return GASX_E_ARG;
    DiffState* d = ctx->diff;
    if (!d) return fail(ctx, GASX_E_STATE, "gasx_diff_events: no recorded run");
    if (step < 0 || step >= d->steps || !dvx || !dvy || !dvz || !p0 || !collided)
        return fail(ctx, GASX_E_ARG, "gasx_diff_events: bad step or null output");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const DEvents ev = events_of(d, step);
    const size_t n = (size_t)d->n;
    gasx_status st = copy_d(ctx, dvx, ev.dvx, n * sizeof(float));
    if (st == GASX_OK) st = copy_d(ctx, dvy, ev.dvy, n * sizeof(float));
    if (st == GASX_OK) st = copy_d(ctx, dvz, ev.dvz, n * sizeof(float));
    if (st == GASX_OK) st = copy_d(ctx, p0, ev.p0, n * sizeof(float));
    if (st == GASX_OK) st = copy_d(ctx, collided, ev.col, n);
    if (st != GASX_OK) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return GASX_OK;
}

extern "C" gasx_status gasx_diff_release(gasx_ctx* ctx)
{
    if (!ctx) return GASX_E_ARG;
    hipSetDevice(ctx->device);
    diff_release(ctx);
    return GASX_OK;
}

extern "C" gasx_status gasx_fdtd_transpose(gasx_ctx* ctx, gasx_fields* lam, double* ljr, double* ljz)
{
    if (!ctx) return GASX_E_ARG;
    if (!lam || !lam->er || !lam->ez || !lam->hphi || !ljr || !ljz)
        return fail(ctx, GASX_E_ARG, "gasx_fdtd_transpose: null array");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const AdjFdtdRun r{lam->er, lam->ez, lam->hphi, ctx->er_b, ctx->ez_b, ctx->h_b, ljr, ljz,
                       ctx->er_m, ctx->ez_m, ctx->h_m};
    HIPCHK(launch_fdtd_adj(ctx->fa, r, ctx->stream), "transposed field cycle");
    gasx_status st = copy_d(ctx, lam->er, ctx->er_b, ner_of(ctx) * sizeof(double));
    if (st == GASX_OK) st = copy_d(ctx, lam->ez, ctx->ez_b, nez_of(ctx) * sizeof(double));
    if (st == GASX_OK) st = copy_d(ctx, lam->hphi, ctx->h_b, nh_of(ctx) * sizeof(double));
    return st;
}
