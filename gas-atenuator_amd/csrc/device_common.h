// Device helpers shared by the particle kernels: Philox4x32-10, wedge
// boundary conditions, clamped area weights, table lookups and the collision
// physics of electron_transport.cu / find_CS0.m.
//
// Every function of the particle step turns FP contraction off: the single-run
// and the batched-ensemble kernels inline them into differently shaped code,
// and only uncontracted arithmetic is guaranteed to round the same in both
// (a batched run is bit-identical to the same run alone).
#pragma once

#include "gasx_internal.h"

namespace gasx {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): published multipliers
// and Weyl increments.
// ---------------------------------------------------------------------------
struct U4 { uint32_t a, b, c, d; };

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32x32->64 multiply each (v_mad_u64_u32) instead of separate
        // low and high halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

// Counter layout (key = seed): (slot, step, run, block).  The collision
// decision uniform u1 of slot s is word s % 4 of the block (s / 4, step, run,
// kDecisionBlock), one draw for four consecutive slots; a candidate's u2..u4
// are words 1..3 of block 0 of its own slot; block 1 is the ion Maxwellian,
// blocks 2..66 the elastic-angle sampler.
constexpr uint32_t kDecisionBlock = 0xFFFFFFFFu;

// (2k+1)*2^-24 with k the top 23 bits: strictly inside (0,1), exact in FP32.
__device__ __forceinline__ float u01(uint32_t x)
{
    return (float)((x >> 8) | 1u) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Geometry helpers
// ---------------------------------------------------------------------------
// Marks an absorbed slot (reference parDelete = 1).
__device__ __forceinline__ bool is_dead(float z) { return !(z == z); }

// Hardware square root / reciprocal (about 1 ulp).  Used everywhere except
// the fixed-point deposit, whose arithmetic must be IEEE-exact.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// x <= 0 (a particle that crossed the axis): the reference's literal
// phi = atan(y/x) (not the polar angle there), shift by multiples of PHIp,
// x = r cos phi, y = r sin phi, velocity rotated by the shift.  Rare.
__device__ __forceinline__ void wedge_wrap_axis(const Geom& g, float r2, float& x, float& y,
                                             float& vx, float& vy)
{
#pragma clang fp contract(off)
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
    if (!(phi == phi)) { phi = 0.0f; rot = 0.0f; }   // x = y = 0: on the axis
    float sr, cr, sp, cp;
    __sincosf(rot, &sr, &cr);                        // |rot| <= pi/2 + PHIp
    __sincosf(phi, &sp, &cp);                        // phi in [0, PHIp]
    const float vx1 = vx * cr - vy * sr;
    const float vy1 = vx * sr + vy * cr;
    vx = vx1; vy = vy1;
    x = r * cp; y = r * sp;
}

// Wedge wrap with velocity rotation and z reflection: EM2.cu:29-56 /
// EM4.cu:150-177.  The reference takes phi = atan(y/x) and, while phi < 0 or
// phi > PHIp, shifts phi by +-PHIp and rotates the velocity by the same
// angle; x = r cos phi, y = r sin phi then rotates the position by the same
// total angle.  For x > 0 (atan is the polar angle) this is done here as
// repeated rotations by +-PHIp with precomputed cos/sin (at most 16).  For
// x <= 0, where atan(y/x) is not the polar angle, the reference's formula is
// followed literally.  Returns false when r > Rp (absorbed).
__device__ __forceinline__ bool wedge_bc(const Geom& g, float& x, float& y, float& z,
                                         float& vx, float& vy, float& vz)
{
#pragma clang fp contract(off)
    const float r2 = x * x + y * y;
    if (r2 > g.rp * g.rp) return false;
    if (x > 0.0f) {
        int guard = 0;
#pragma unroll 1
        while (y < 0.0f && guard++ < 32) {              // phi < 0: rotate by +PHIp
            const float x1 = x * g.cphip - y * g.sphip, y1 = x * g.sphip + y * g.cphip;
            const float u1 = vx * g.cphip - vy * g.sphip, w1 = vx * g.sphip + vy * g.cphip;
            x = x1; y = y1; vx = u1; vy = w1;
        }
#pragma unroll 1
        while (y * g.cphip > x * g.sphip && guard++ < 32) {   // phi > PHIp: rotate by -PHIp
            const float x1 = x * g.cphip + y * g.sphip, y1 = -x * g.sphip + y * g.cphip;
            const float u1 = vx * g.cphip + vy * g.sphip, w1 = -vx * g.sphip + vy * g.cphip;
            x = x1; y = y1; vx = u1; vy = w1;
        }
    } else {
        wedge_wrap_axis(g, r2, x, y, vx, vy);
    }
    if (z > g.lp) { z = 2.0f * g.lp - z; vz = -vz; }
    if (z < 0.0f) { z = -z; vz = -vz; }
    return true;
}

// Clamped linear weights on a cell-centred axis (cells at (i+1/2)*d,
// i in [0,n)): EM7.cu:56-162 and EM4.cu:82-122.  Out-of-range stencils
// collapse onto the edge cell with weight 1.
__device__ __forceinline__ void half_axis(float s, int n, int& i0, int& i1, float& w0, float& w1)
{
#pragma clang fp contract(off)
    // s = x/d - 0.5
    const float fl = floorf(s);
    const int a = (int)fl;
    const float f = s - fl;
    const bool lo = a < 0, hi = a >= n - 1;
    i0 = lo ? 0 : (hi ? n - 1 : a);
    i1 = lo ? 0 : (hi ? n - 1 : a + 1);
    w0 = (lo || hi) ? 1.0f : 1.0f - f;
    w1 = (lo || hi) ? 0.0f : f;
}

// p[i] for an element index below 2^30: the byte offset zero-extended onto
// the base, so a uniform base stays in scalar registers (one vector add).
__device__ __forceinline__ float ld_off(const float* p, uint32_t i)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + (uint64_t)(i << 2));
}

// Linear weights on a node axis (nodes at i*d, i in [0,n]): EM4.cu:41-80.
__device__ __forceinline__ void node_axis(float s, int n, int& i0, int& i1, float& w0, float& w1)
{
#pragma clang fp contract(off)
    const float fl = floorf(s);
    int a = (int)fl;
    float f = s - fl;
    const bool out = a < 0 || a >= n;
    f = out ? 0.0f : f;
    a = a < 0 ? 0 : (a > n ? n : a);
    i0 = a;
    i1 = a + 1 > n ? n : a + 1;
    w0 = 1.0f - f; w1 = f;
}

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------
// interp1 'linear' on a sorted table; caller guarantees x[0] <= xq <= x[n-1].
__device__ __forceinline__ float interp_in(const float* x, const float* y, int n, float xq)
{
#pragma clang fp contract(off)
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (x[mid] <= xq) lo = mid; else hi = mid;
    }
    const float s = (xq - x[lo]) * frcp(x[hi] - x[lo]);
    return y[lo] + s * (y[hi] - y[lo]);
}

// find_CS0.m:14-56 for an energy T (eV) inside [15.6, 20000].
__device__ __forceinline__ void cross_sections(const SmallTabs* t, float T, float& se, float& si, float& sx)
{
#pragma clang fp contract(off)
    if (T >= 50.0f) se = interp_in(t->el_e, t->el_s, t->n_el, T);
    else se = interp_in(t->sel_e, t->sel_s, t->n_sel, T);
    si = interp_in(t->ion_e, t->ion_s, t->n_ion, T);
    float s = 0.0f;
    const float l2T = __log2f(T);
    for (int i = 0; i < t->n_exc; ++i) {
        const float l2r = t->exc_l2w[i] - l2T;            // log2(W/T)
        const float b = exp2f(t->exc_om[i] * l2r);        // (W/T)^omega
        float c = 1.0f - exp2f(t->exc_ga[i] * l2r);       // 1 - (W/T)^gamma
        c = exp2f(t->exc_nu[i] * __log2f(c));              // ^nu
        s += t->exc_q[i] * b * c;
    }
    sx = s;
}

// Speed from kinetic energy E (eV), findV.m:5-8 in cancellation-free form.
__device__ __forceinline__ float speed_from_ev(float E, float mc2_ev, float c)
{
#pragma clang fp contract(off)
    const float g1 = E * frcp(mc2_ev);
    return c * fsqrt(g1 * (g1 + 2.0f)) * frcp(1.0f + g1);
}

// Kinetic energy (eV) of speed^2 v2: EM4.cu:138-140, /e of
// electricmagneticfield6.m:172, as m c^2 beta^2 gamma^2 / (1+gamma).
__device__ __forceinline__ float ev_from_speed2(float v2, float inv_c2, float mc2_ev)
{
#pragma clang fp contract(off)
    const float b2 = v2 * inv_c2;
    const float g2 = frcp(1.0f - b2);
    const float gm = fsqrt(g2);
    return mc2_ev * (b2 * g2 * frcp(1.0f + gm));
}

// rotate.m:4-13 with v^2 - vz^2 evaluated as vx^2 + vy^2.
__device__ __forceinline__ void rotate_v(float& vx, float& vy, float& vz, float th, float ph, float eps)
{
#pragma clang fp contract(off)
    const float x = vx, y = vy, zz = vz;
    const float vp2 = x * x + y * y;
    const float v = fsqrt(vp2 + zz * zz);
    float st, ct, sp, cp;
    __sincosf(th, &st, &ct);   // th in [0, pi], ph in [0, 2pi): no range reduction needed
    __sincosf(ph, &sp, &cp);
    if (vp2 > eps) {
        const float ivp = frsq(vp2);
        const float vp = vp2 * ivp;
        const float k = st * ivp;
        vx = x * ct + k * (x * zz * cp - v * y * sp);
        vy = y * ct + k * (y * zz * cp + v * x * sp);
        vz = zz * ct - vp * st * cp;
    } else {
        vx = v * st * cp;
        vy = v * st * sp;
        vz = v * ct;
    }
}

// Mean excitation loss, get_ImpExc2.m:19-32 (form 0) or
// electron_transport.cu:212-219 (form 1).
__device__ __forceinline__ float exc_edep(const SmallTabs* t, float T, int form)
{
#pragma clang fp contract(off)
    float k1 = 0.0f, k2 = 0.0f;
    const float l2T = __log2f(T);
    for (int i = 0; i < t->n_exc; ++i) {
        const float W = t->exc_w[i];
        const float l2r = t->exc_l2w[i] - l2T;
        float b = exp2f(t->exc_om[i] * l2r);
        float c;
        if (form == 0) {
            c = 1.0f - exp2f(t->exc_ga[i] * l2r);
            c = exp2f(t->exc_nu[i] * __log2f(c));
        } else {
            c = 1.0f - W / T;
            c = exp2f(t->exc_nu[i] * t->exc_ga[i] * __log2f(c));
        }
        float w = t->exc_a[i] * b * c;
        k1 += w * W;
        k2 += w;
    }
    return k1 / k2;
}

}  // namespace gasx
