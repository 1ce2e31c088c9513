// Device helpers shared by the particle kernels: Philox4x32-10, wedge
// boundary conditions, clamped area weights, table lookups and the collision
// physics of electron_transport.cu / find_CS0.m.
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
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

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

// Wedge wrap with velocity rotation and z reflection: EM2.cu:29-56 /
// EM4.cu:150-177.  phi = atan(y/x) (the reference's quadrant convention); a
// particle outside [0, PHIp] is moved by whole multiples of PHIp, its velocity
// rotated by the same angle.  Returns false when r > Rp (absorbed).
__device__ __forceinline__ bool wedge_bc(const Geom& g, float& x, float& y, float& z,
                                         float& vx, float& vy, float& vz)
{
    float r2 = x * x;
    r2 = r2 + y * y;
    float r = sqrtf(r2);
    if (r > g.rp) return false;
    // inside test without atan: x > 0, y >= 0, y*cos(PHIp) <= x*sin(PHIp)
    bool inside = (x > 0.0f) && (y >= 0.0f) && (y * g.cphip <= x * g.sphip);
    if (!inside) {
        float phi = atanf(y / x);
        float rot;
        if (phi < 0.0f) {
            float k = ceilf(-phi / g.phip);
            phi = phi + k * g.phip;
            rot = k * g.phip;
        } else if (phi > g.phip) {
            float k = ceilf((phi - g.phip) / g.phip);
            phi = phi - k * g.phip;
            rot = -k * g.phip;
        } else {
            rot = 0.0f;
        }
        if (!(phi == phi)) { phi = 0.0f; rot = 0.0f; }   // x = y = 0: on the axis
        float sr, cr, sp, cp;
        sincosf(rot, &sr, &cr);
        sincosf(phi, &sp, &cp);
        float vx1 = vx * cr - vy * sr;
        float vy1 = vx * sr + vy * cr;
        vx = vx1; vy = vy1;
        x = r * cp; y = r * sp;
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
    // s = x/d - 0.5
    float fl = floorf(s);
    int a = (int)fl;
    float f = s - fl;
    if (a < 0) { i0 = 0; i1 = 0; w0 = 1.0f; w1 = 0.0f; }
    else if (a >= n - 1) { i0 = n - 1; i1 = n - 1; w0 = 1.0f; w1 = 0.0f; }
    else { i0 = a; i1 = a + 1; w0 = 1.0f - f; w1 = f; }
}

// Linear weights on a node axis (nodes at i*d, i in [0,n]): EM4.cu:41-80.
__device__ __forceinline__ void node_axis(float s, int n, int& i0, int& i1, float& w0, float& w1)
{
    float fl = floorf(s);
    int a = (int)fl;
    float f = s - fl;
    if (a < 0) { a = 0; f = 0.0f; }
    if (a >= n) { a = n; f = 0.0f; }
    i0 = a; i1 = a + 1 > n ? n : a + 1;
    w0 = 1.0f - f; w1 = f;
}

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------
// interp1 'linear' on a sorted table; caller guarantees x[0] <= xq <= x[n-1].
__device__ __forceinline__ float interp_in(const float* x, const float* y, int n, float xq)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (x[mid] <= xq) lo = mid; else hi = mid;
    }
    float s = (xq - x[lo]) / (x[hi] - x[lo]);
    return y[lo] + s * (y[hi] - y[lo]);
}

// find_CS0.m:14-56 for an energy T (eV) inside [15.6, 20000].
__device__ __forceinline__ void cross_sections(const DevTables* t, float T, float& se, float& si, float& sx)
{
    if (T >= 50.0f) se = interp_in(t->el_e, t->el_s, t->n_el, T);
    else se = interp_in(t->sel_e, t->sel_s, t->n_sel, T);
    si = interp_in(t->ion_e, t->ion_s, t->n_ion, T);
    float s = 0.0f;
    float l2T = __log2f(T);
    for (int i = 0; i < t->n_exc; ++i) {
        float W = t->exc_w[i];
        float l2r = __log2f(W) - l2T;                     // log2(W/T)
        float b = exp2f(t->exc_om[i] * l2r);              // (W/T)^omega
        float c = 1.0f - exp2f(t->exc_ga[i] * l2r);       // 1 - (W/T)^gamma
        c = exp2f(t->exc_nu[i] * __log2f(c));              // ^nu
        s += t->exc_q[i] * b * c;
    }
    sx = s;
}

// Speed from kinetic energy E (eV), findV.m:5-8 in cancellation-free form.
__device__ __forceinline__ float speed_from_ev(float E, float mc2_ev, float c)
{
    float g1 = E / mc2_ev;
    return c * sqrtf(g1 * (g1 + 2.0f)) / (1.0f + g1);
}

// Kinetic energy (eV) of speed^2 v2: EM4.cu:138-140, /e of
// electricmagneticfield6.m:172, as m c^2 beta^2 gamma^2 / (1+gamma).
__device__ __forceinline__ float ev_from_speed2(float v2, float inv_c2, float mc2_ev)
{
    float b2 = v2 * inv_c2;
    float g2 = 1.0f / (1.0f - b2);
    float gm = sqrtf(g2);
    return mc2_ev * (b2 * g2 / (1.0f + gm));
}

// rotate.m:4-13 with v^2 - vz^2 evaluated as vx^2 + vy^2.
__device__ __forceinline__ void rotate_v(float& vx, float& vy, float& vz, float th, float ph, float eps)
{
    float x = vx, y = vy, zz = vz;
    float vp2 = x * x + y * y;
    float v = sqrtf(vp2 + zz * zz);
    float st, ct, sp, cp;
    sincosf(th, &st, &ct);
    sincosf(ph, &sp, &cp);
    if (vp2 > eps) {
        float vp = sqrtf(vp2);
        float k = st / vp;
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
__device__ __forceinline__ float exc_edep(const DevTables* t, float T, int form)
{
    float k1 = 0.0f, k2 = 0.0f;
    float l2T = __log2f(T);
    for (int i = 0; i < t->n_exc; ++i) {
        float W = t->exc_w[i];
        float l2r = __log2f(W) - l2T;
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
