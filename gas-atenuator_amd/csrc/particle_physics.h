// Per-particle physics of the electron step, shared by the production
// kernels (particles.hip) and the differentiable model and its adjoint
// (adjoint.hip): EM7 deposit weights in fixed point, EM4 gather + push, EM2
// half drift, and the collision of one candidate (electron_transport.cu with
// the find_CS0.m cross sections and the get_ElasticSc_sub1.m angle law).
#pragma once

#include "device_common.h"

namespace gasx {

// ---- deposit (EM7.cu:56-202 in fixed point) ---------------------------------
// The contributions are computed in FP32 with contraction off so that they
// are reproducible on the CPU; their int64 sum is order independent.
struct DepCells {
    int i0, i1, k0, k1;
    long long qr[4], qz[4];   // cells (i0,k0) (i1,k0) (i0,k1) (i1,k1)
};

__device__ __forceinline__ void deposit_cells(const Geom& g, float x, float y, float z, float vx,
                                              float vy, float vz, float kq, DepCells& d)
{
#pragma clang fp contract(off)
    float r2 = x * x;
    r2 = r2 + y * y;
    float r = sqrtf(r2);
    float vr;
    if (r > 0.0f) {
        float t = vx * x;
        t = t + vy * y;
        vr = t / r;
    } else {
        vr = vx;
    }
    float qrv = vr * kq, qzv = vz * kq;
    float sr = r * g.inv_dr;
    sr = sr - 0.5f;
    float sz = z * g.inv_dz;
    sz = sz - 0.5f;
    float wr0, wr1, wz0, wz1;
    half_axis(sr, g.nr, d.i0, d.i1, wr0, wr1);
    half_axis(sz, g.nz, d.k0, d.k1, wz0, wz1);
    const float w[4] = {wr0 * wz0, wr1 * wz0, wr0 * wz1, wr1 * wz1};
    // |w q v / c| 2^32 < 2^31 whenever |q v| < c/2 (always for the electrons
    // here): one float->int32 conversion instead of the longer int64 one
    // (a branch, not a select: the int64 conversions run only in the rare
    // wave that holds such a particle)
    const bool small = fabsf(qrv) < 2147483648.0f && fabsf(qzv) < 2147483648.0f;
    if (small) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            d.qr[a] = (long long)(int)rintf(w[a] * qrv);
            d.qz[a] = (long long)(int)rintf(w[a] * qzv);
        }
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            d.qr[a] = (long long)rintf(w[a] * qrv);
            d.qz[a] = (long long)rintf(w[a] * qzv);
        }
    }
}

__device__ __forceinline__ int dep_cell(const DepCells& d, int a, int nr)
{
    return ((a & 1) ? d.i1 : d.i0) + ((a & 2) ? d.k1 : d.k0) * nr;
}

// ---- gather + push (EM4.cu:36-148) ------------------------------------------
// In two parts so that a lane can have several particles' field reads in
// flight together: gather_issue computes the stencils and reads the 12 field
// values, push_apply interpolates and pushes.  push_one is the two in a row.
struct FieldGather {
    float er[4], h[4], ez[4];
    float a0, a1, b0, b1, c0, c1, d0, d1, cph, sph;
};

__device__ __forceinline__ void gather_issue(const Geom& g, const PartArgs& pa, float x, float y, float z,
                                             FieldGather& G)
{
#pragma clang fp contract(off)
    const float r2 = x * x + y * y;
    float cph = 1.0f, sph = 0.0f, r = 0.0f;
    if (r2 > 0.0f) {
        const float ir = frsq(r2);
        r = r2 * ir;
        cph = x * ir; sph = y * ir;
    }
    G.cph = cph; G.sph = sph;
    const float srn = r * g.inv_dr, szn = z * g.inv_dz;
    const float srh = srn - 0.5f, szh = szn - 0.5f;
    int i0, i1, k0, k1, kh0, kh1, j0, j1;
    const int nr = g.nr, ld = nr + 1;
    // Er: cell-centred in r, nodes in z (EM4.cu:41-60)
    half_axis(srh, g.nr, i0, i1, G.a0, G.a1);
    node_axis(szn, g.nz, k0, k1, G.b0, G.b1);
    // Hphi: cell-centred in both (EM4.cu:82-122); shares the r weights of Er
    half_axis(szh, g.nz, kh0, kh1, G.c0, G.c1);
    // Ez: nodes in r, cell-centred in z (EM4.cu:61-80)
    node_axis(srn, g.nr, j0, j1, G.d0, G.d1);
    G.er[0] = pa.er[i0 + k0 * nr]; G.er[1] = pa.er[i1 + k0 * nr];
    G.er[2] = pa.er[i0 + k1 * nr]; G.er[3] = pa.er[i1 + k1 * nr];
    G.h[0] = pa.h[i0 + kh0 * nr]; G.h[1] = pa.h[i1 + kh0 * nr];
    G.h[2] = pa.h[i0 + kh1 * nr]; G.h[3] = pa.h[i1 + kh1 * nr];
    G.ez[0] = pa.ez[j0 + kh0 * ld]; G.ez[1] = pa.ez[j1 + kh0 * ld];
    G.ez[2] = pa.ez[j0 + kh1 * ld]; G.ez[3] = pa.ez[j1 + kh1 * ld];
}

__device__ __forceinline__ void push_apply(const Geom& g, const PartArgs& pa, const FieldGather& G, float& vx,
                                           float& vy, float& vz)
{
#pragma clang fp contract(off)
    const float cph = G.cph, sph = G.sph;
    const float vr = vx * cph + vy * sph;
    const float Er = G.b0 * (G.a0 * G.er[0] + G.a1 * G.er[1]) + G.b1 * (G.a0 * G.er[2] + G.a1 * G.er[3]);
    const float H = G.c0 * (G.a0 * G.h[0] + G.a1 * G.h[1]) + G.c1 * (G.a0 * G.h[2] + G.a1 * G.h[3]);
    const float Ez = G.c0 * (G.d0 * G.ez[0] + G.d1 * G.ez[1]) + G.c1 * (G.d0 * G.ez[2] + G.d1 * G.ez[3]);
    // EM4.cu:126-144: magnetic term uses v^n; explicit (not Boris) update
    const float fr = pa.q_over_m * (Er - vz * g.mu0 * H);
    const float fz = pa.q_over_m * (Ez + vr * g.mu0 * H);
    const float v2 = vx * vx + vy * vy + vz * vz;
    const float gm = rsqrtf(1.0f - v2 * g.inv_c2);
    const float ux = gm * vx + (fr * cph) * g.dt;
    const float uy = gm * vy + (fr * sph) * g.dt;
    const float uz = gm * vz + fz * g.dt;
    const float igp = rsqrtf(1.0f + (ux * ux + uy * uy + uz * uz) * g.inv_c2);
    vx = ux * igp; vy = uy * igp; vz = uz * igp;
}

__device__ __forceinline__ void push_one(const Geom& g, const PartArgs& pa, float x, float y,
                                         float z, float& vx, float& vy, float& vz)
{
    FieldGather G;
    gather_issue(g, pa, x, y, z, G);
    push_apply(g, pa, G, vx, vy, vz);
}

__device__ __forceinline__ void half_drift(const Geom& g, float& x, float& y, float& z, float vx,
                                           float vy, float vz)
{
#pragma clang fp contract(off)
    const float h = 0.5f * g.dt;
    x = x + h * vx;
    y = y + h * vy;
    z = z + h * vz;
}

// ---- elastic angle ------------------------------------------------------------
// Reference semantics (get_ElasticSc_sub1.m:4-19 on the merged tables): log DCS
// linear in log E at each node angle, trapezoid CDF over theta, inverse linear
// interpolation -- i.e. interval k of 1 deg is chosen with probability
// proportional to m_k(w) = p_k(w) + p_{k+1}(w), p_k(w) = sind(k) exp((1-w) V_i[k]
// + w V_{i+1}[k]), and theta is uniform inside it.  Sampled exactly by
// rejection: exp is convex, so m_k(w) <= (1-w) m_k(i) + w m_k(i+1); propose
// node i or i+1 with weights (1-w) Z_i, w Z_{i+1}, draw k from that node's
// precomputed interval CDF (binary search), accept with m_k(w) / M_k.
__device__ inline float elastic_theta_exact(const DevTables* t, const SmallTabs* st, int i,
                                                 float w, float u)
{
#pragma clang fp contract(off)
    const float* v0 = &t->dcs_logv[i * kNAng];
    const float* v1 = &t->dcs_logv[(i + 1) * kNAng];
    float tot = 0.0f, prev = 0.0f;
#pragma unroll 1
    for (int k = 0; k < kNAng; ++k) {
        const float p = st->sind[k] * __expf((1.0f - w) * v0[k] + w * v1[k]);
        if (k > 0) tot += prev + p;
        prev = p;
    }
    const float target = u * tot;
    float cdf = 0.0f, p0 = 0.0f, s = 1.0f;
    int k = 0;
#pragma unroll 1
    for (k = 0; k < kNAng - 1; ++k) {
        const float p1 = st->sind[k + 1] * __expf((1.0f - w) * v0[k + 1] + w * v1[k + 1]);
        const float nxt = cdf + (p0 + p1);
        if (nxt >= target || k == kNAng - 2) {
            s = (nxt > cdf) ? (target - cdf) / (nxt - cdf) : 0.0f;
            break;
        }
        cdf = nxt;
        p0 = p1;
    }
    s = fminf(fmaxf(s, 0.0f), 1.0f);
    return ((float)k + s) * (3.14159265358979f / 180.0f);
}

__device__ inline float elastic_theta(const DevTables* t, const SmallTabs* st, const CollArgs& ca,
                               uint32_t slot, float e_kev)
{
#pragma clang fp contract(off)
    const float lx = logf(e_kev);
    int i = 0, hi = st->n_dcs - 1;                 // bin: dcs_loge[i] <= lx (clamped to n-2)
    while (hi - i > 1) {
        const int mid = (i + hi) >> 1;
        if (st->dcs_loge[mid] < lx) i = mid; else hi = mid;
    }
    const float w = (lx - st->dcs_loge[i]) / (st->dcs_loge[i + 1] - st->dcs_loge[i]);
    const float wa = (1.0f - w) * st->dcs_z[i], wb = w * st->dcs_z[i + 1];
    const float* v0 = &t->dcs_logv[i * kNAng];
    const float* v1 = &t->dcs_logv[(i + 1) * kNAng];
    const float* p0 = &t->dcs_p[i * kNAng];
    const float* p1 = &t->dcs_p[(i + 1) * kNAng];
#pragma unroll 1
    for (uint32_t it = 0; it < 64; ++it) {
        const U4 R = philox4x32_10(slot, ca.step, ca.run, 2u + it, ca.key0, ca.key1);
        const float a = u01(R.a), b = u01(R.b), c = u01(R.c), d = u01(R.d);
        const int j = (a * (wa + wb) < wa) ? i : i + 1;
        const float* C = &t->dcs_cdf[j * kNAng];
        // largest k with C[k] <= b lies in [guide[g], guide[g+1]] (binary search)
        const int gi = (int)(b * (float)kGuide);
        int k = st->guide[j][gi], khi = st->guide[j][gi + 1] + 1;
        if (khi > kNAng - 1) khi = kNAng - 1;
        while (khi - k > 1) {
            const int mid = (k + khi) >> 1;
            if (C[mid] <= b) k = mid; else khi = mid;
        }
        const float M = (1.0f - w) * (p0[k] + p0[k + 1]) + w * (p1[k] + p1[k + 1]);
        const float m = st->sind[k] * __expf((1.0f - w) * v0[k] + w * v1[k]) +
                        st->sind[k + 1] * __expf((1.0f - w) * v0[k + 1] + w * v1[k + 1]);
        if (c * M < m) return ((float)k + d) * (3.14159265358979f / 180.0f);
    }
    const U4 R = philox4x32_10(slot, ca.step, ca.run, 66u, ca.key0, ca.key1);
    return elastic_theta_exact(t, st, i, w, u01(R.a));
}

// ---- collision of one candidate (electron_transport.cu:64-227) ---------------
enum : int { kEvMask = 3, kEvViol = 4, kEvSat = 8, kEvSec = 16, kEvIon = 32 };

struct CollOut {
    int ev;
    float svx, svy, svz;
    float ivx, ivy, ivz;
    float pt;              // p_T of the check (0 outside the tables' energy window)
};

// find_CS0.m cross sections, interp1 NaN gating (collide only for
// 15.6 eV <= T <= 20 keV), linear p_T, branch choice by u2, and the elastic /
// ionization / excitation kinematics of electron_transport.cu with the
// MATLAB choices.  Contraction is off in the collision math (here and in
// its callees): the kernels that inline it (single run, batched ensemble)
// then cannot fuse differently, and a batched run stays bit-identical to the
// same run alone.
__device__ inline void collide_candidate(const Geom& g, const CollArgs& ca, const SmallTabs* stab,
                                  uint32_t slot, float x, float y, float z, float& vx, float& vy,
                                  float& vz, float u1, float u2, float u3, float u4, bool want_ion,
                                  CollOut& o)
{
#pragma clang fp contract(off)
    o.ev = 0;
    o.pt = 0.0f;
    const float v2 = vx * vx + vy * vy + vz * vz;
    const float T = ev_from_speed2(v2, g.inv_c2, ca.ev_per_mc2_em);
    if (!(T >= 15.6f && T <= 20000.0f)) return;
    const DevTables* t = ca.tab;
    float se, si, sx;
    cross_sections(stab, T, se, si, sx);
    const float r = fsqrt(x * x + y * y);
    int ir = (int)floorf(r * ca.inv_gdr), iz = (int)floorf(z * ca.inv_gdz);
    ir = ir < 0 ? 0 : (ir > ca.gnr - 1 ? ca.gnr - 1 : ir);
    iz = iz < 0 ? 0 : (iz > ca.gnz - 1 ? ca.gnz - 1 : iz);
    const float dens = ca.dens[iz * ca.gnr + ir];
    const float st = se + si + sx;
    const float ve = fsqrt(v2);
    const float pt = ve * ca.dt_coll * dens * st;      // electron_transport.cu:74-77
    o.pt = pt;
    if (pt > ca.pmax && ca.pmax < 1.0f) o.ev |= kEvViol;
    if (!(u1 < pt)) return;                             // :85
    if (pt > 1.0f) o.ev |= kEvSat;
    const float twopi = 6.28318530717959f;
    if (u2 * st < se) {                                 // elastic (:86-137)
        const float th = elastic_theta(t, stab, ca, slot, T * 1e-3f);
        rotate_v(vx, vy, vz, th, twopi * u4, ca.rot_eps);
        o.ev |= 1;
    } else if (u2 * st < se + si && T > 15.6f) {        // ionization (:139-205)
        const float I = 15.6f;
        const float T0 = 4.17f - 1000.0f / (T + 2.0f * I);
        const float tt = 13.8f * T / (T + I);
        const float Tm = (T - I) * 0.5f;
        const float k1 = atanf((Tm - T0) / tt);
        const float k2 = atanf(T0 / tt);
        float k0 = T0 + tt * tanf(u3 * (k1 + k2) - k2);
        k0 = fminf(fmaxf(k0, 0.0f), Tm);
        const float E1 = k0, E2 = T - E1 - I;
        const float Ep = E2 > E1 ? E2 : E1;
        const float Es = E2 > E1 ? E1 : E2;
        const float Ee = ca.mc2_mc_ev;
        const float kt = k0 / T;
        const float sp = sqrtf(kt / ((1.0f - kt) * T / (2.0f * Ee) + 1.0f));
        const float ss = sqrtf((1.0f - kt) / (1.0f + k0 / (2.0f * Ee)));
        const float thp = asinf(fminf(sp, 1.0f)), ths = asinf(fminf(ss, 1.0f));
        const float php = twopi * u4, phs = php - 3.14159265358979f;
        const float ratio = speed_from_ev(Ep, Ee, g.c) / ve;
        vx *= ratio; vy *= ratio; vz *= ratio;
        rotate_v(vx, vy, vz, thp, php, ca.rot_eps);
        const float vs = speed_from_ev(Es, Ee, g.c);
        float sts, cts, sps, cps;
        __sincosf(ths, &sts, &cts);   // |angles| <= pi: no range reduction needed
        __sincosf(phs, &sps, &cps);
        o.ev |= 2 | kEvSec;
        o.svx = vs * sts * cps; o.svy = vs * sts * sps; o.svz = vs * cts;   // :198-200
        if (want_ion) {                                  // :202-204, Box-Muller normals
            const U4 Q = philox4x32_10(slot, ca.step, ca.run, 1u, ca.key0, ca.key1);
            const float a1 = u01(Q.a), a2 = u01(Q.b), a3 = u01(Q.c), a4 = u01(Q.d);
            const float Tg = ca.temp ? ca.temp[iz * ca.gnr + ir] : ca.t_scalar;
            const float sg = ca.ion_vth_k * sqrtf(Tg);
            const float rr1 = sqrtf(-2.0f * logf(a1)), rr2 = sqrtf(-2.0f * logf(a3));
            float s1, c1, s2, c2;
            __sincosf(twopi * a2, &s1, &c1);
            __sincosf(twopi * a4, &s2, &c2);
            o.ev |= kEvIon;
            o.ivx = sg * rr1 * c1; o.ivy = sg * rr1 * s1; o.ivz = sg * rr2 * c2;
        }
    } else {                                             // excitation (:206-226)
        const float edep = exc_edep(stab, T, ca.exc_form);
        const float ratio = speed_from_ev(T - edep, ca.mc2_mc_ev, g.c) / ve;
        vx *= ratio; vy *= ratio; vz *= ratio;
        o.ev |= 3;
    }
}

}  // namespace gasx
