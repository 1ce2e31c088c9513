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
    // every index is clamped into its array: unsigned 32-bit element offsets
    // (24-bit row products) keep the address arithmetic to one add per read
    const uint32_t r0 = __umul24((uint32_t)k0, (uint32_t)nr), r1 = __umul24((uint32_t)k1, (uint32_t)nr);
    const uint32_t h0 = __umul24((uint32_t)kh0, (uint32_t)nr), h1 = __umul24((uint32_t)kh1, (uint32_t)nr);
    const uint32_t e0 = __umul24((uint32_t)kh0, (uint32_t)ld), e1 = __umul24((uint32_t)kh1, (uint32_t)ld);
    G.er[0] = ld_off(pa.er, (uint32_t)i0 + r0); G.er[1] = ld_off(pa.er, (uint32_t)i1 + r0);
    G.er[2] = ld_off(pa.er, (uint32_t)i0 + r1); G.er[3] = ld_off(pa.er, (uint32_t)i1 + r1);
    G.h[0] = ld_off(pa.h, (uint32_t)i0 + h0); G.h[1] = ld_off(pa.h, (uint32_t)i1 + h0);
    G.h[2] = ld_off(pa.h, (uint32_t)i0 + h1); G.h[3] = ld_off(pa.h, (uint32_t)i1 + h1);
    G.ez[0] = ld_off(pa.ez, (uint32_t)j0 + e0); G.ez[1] = ld_off(pa.ez, (uint32_t)j1 + e0);
    G.ez[2] = ld_off(pa.ez, (uint32_t)j0 + e1); G.ez[3] = ld_off(pa.ez, (uint32_t)j1 + e1);
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
    // element offsets of nodes i and i+1 in the angle tables (unsigned 32-bit:
    // a uniform table base stays in scalar registers)
    const uint32_t o0 = (uint32_t)i * kNAng, o1 = o0 + kNAng;
#pragma unroll 1
    for (uint32_t it = 0; it < 64; ++it) {
        const U4 R = philox4x32_10(slot, ca.step, ca.run, 2u + it, ca.key0, ca.key1);
        const float a = u01(R.a), b = u01(R.b), c = u01(R.c), d = u01(R.d);
        const bool first = a * (wa + wb) < wa;
        const int j = first ? i : i + 1;
        const uint32_t oc = first ? o0 : o1;
        // largest k with C[k] <= b lies in [guide[g], guide[g+1]] (binary search)
        const int gi = (int)(b * (float)kGuide);
        int k = st->guide[j][gi], khi = st->guide[j][gi + 1] + 1;
        if (khi > kNAng - 1) khi = kNAng - 1;
        while (khi - k > 1) {
            const int mid = (k + khi) >> 1;
            if (ld_off(t->dcs_cdf, oc + (uint32_t)mid) <= b) k = mid; else khi = mid;
        }
        const uint32_t uk = (uint32_t)k;
        const float M = (1.0f - w) * (ld_off(t->dcs_p, o0 + uk) + ld_off(t->dcs_p, o0 + uk + 1)) +
                        w * (ld_off(t->dcs_p, o1 + uk) + ld_off(t->dcs_p, o1 + uk + 1));
        const float m = st->sind[k] * __expf((1.0f - w) * ld_off(t->dcs_logv, o0 + uk) +
                                             w * ld_off(t->dcs_logv, o1 + uk)) +
                        st->sind[k + 1] * __expf((1.0f - w) * ld_off(t->dcs_logv, o0 + uk + 1) +
                                                 w * ld_off(t->dcs_logv, o1 + uk + 1));
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

// Gas cell of a position (the density / temperature grid, clamped).
__device__ __forceinline__ void gas_cell(const CollArgs& ca, float x, float y, float z, int& ir, int& iz)
{
#pragma clang fp contract(off)
    const float r = fsqrt(x * x + y * y);
    ir = (int)floorf(r * ca.inv_gdr);
    iz = (int)floorf(z * ca.inv_gdz);
    ir = ir < 0 ? 0 : (ir > ca.gnr - 1 ? ca.gnr - 1 : ir);
    iz = iz < 0 ? 0 : (iz > ca.gnz - 1 ? ca.gnz - 1 : iz);
}

// The collision check of one candidate (electron_transport.cu:64-85 with the
// find_CS0.m cross sections and interp1 NaN gating: only 15.6 eV <= T <=
// 20 keV collides): linear p_T = v dt n sigma_T, the decision u1 < p_T and
// the branch by u2 (typ 1 elastic, 2 ionization, 3 excitation, 0 none);
// ev carries kEvViol (p_T above the rejection bound) and kEvSat (p_T > 1).
// Contraction is off in the collision math (here and in the kinematics
// below): the kernels that inline it (single run, batched ensemble) then
// cannot fuse differently, and a batched run stays bit-identical to the
// same run alone.
struct CollDecision {
    int typ;
    int ev;
    float pt;              // p_T of the check (0 outside the tables' energy window)
};

__device__ inline void collide_decide(const Geom& g, const CollArgs& ca, const SmallTabs* stab, float x, float y,
                                      float z, float vx, float vy, float vz, float u1, float u2, CollDecision& d)
{
#pragma clang fp contract(off)
    d.typ = 0;
    d.ev = 0;
    d.pt = 0.0f;
    const float v2 = vx * vx + vy * vy + vz * vz;
    const float T = ev_from_speed2(v2, g.inv_c2, ca.ev_per_mc2_em);
    if (!(T >= 15.6f && T <= 20000.0f)) return;
    float se, si, sx;
    cross_sections(stab, T, se, si, sx);
    int ir, iz;
    gas_cell(ca, x, y, z, ir, iz);
    const float dens = ca.dens[iz * ca.gnr + ir];
    const float st = se + si + sx;
    const float ve = fsqrt(v2);
    const float pt = ve * ca.dt_coll * dens * st;      // electron_transport.cu:74-77
    d.pt = pt;
    if (pt > ca.pmax && ca.pmax < 1.0f) d.ev |= kEvViol;
    if (!(u1 < pt)) return;                             // :85
    if (pt > 1.0f) d.ev |= kEvSat;
    if (u2 * st < se) d.typ = 1;                        // elastic (:86-137)
    else if (u2 * st < se + si && T > 15.6f) d.typ = 2; // ionization (:139-205)
    else d.typ = 3;                                     // excitation (:206-226)
}

// The energy (eV) and speed of the push's electron, as collide_decide
// computes them.
__device__ __forceinline__ void candidate_energy(const Geom& g, const CollArgs& ca, float vx, float vy, float vz,
                                                 float& T, float& ve)
{
#pragma clang fp contract(off)
    const float v2 = vx * vx + vy * vy + vz * vz;
    T = ev_from_speed2(v2, g.inv_c2, ca.ev_per_mc2_em);
    ve = fsqrt(v2);
}

// Elastic scattering (electron_transport.cu:86-137): polar angle from the
// merged DCS tables, azimuth 2 pi u4.
__device__ inline void collide_elastic(const CollArgs& ca, const SmallTabs* stab, uint32_t slot, float T, float u4,
                                       float& vx, float& vy, float& vz)
{
#pragma clang fp contract(off)
    const float twopi = 6.28318530717959f;
    const float th = elastic_theta(ca.tab, stab, ca, slot, T * 1e-3f);
    rotate_v(vx, vy, vz, th, twopi * u4, ca.rot_eps);
}

// Ionization (electron_transport.cu:139-205): secondary energy by u3, the
// primary and the secondary directions, and with want_ion the N2+ Maxwellian
// velocity at the gas temperature of the site (:202-204).
__device__ inline void collide_ionize(const Geom& g, const CollArgs& ca, uint32_t slot, float x, float y, float z,
                                      float T, float ve, float u3, float u4, bool want_ion, float& vx, float& vy,
                                      float& vz, CollOut& o)
{
#pragma clang fp contract(off)
    const float twopi = 6.28318530717959f;
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
        int ir, iz;
        gas_cell(ca, x, y, z, ir, iz);
        const float Tg = ca.temp ? ca.temp[iz * ca.gnr + ir] : ca.t_scalar;
        const float sg = ca.ion_vth_k * sqrtf(Tg);
        const float rr1 = sqrtf(-2.0f * logf(a1)), rr2 = sqrtf(-2.0f * logf(a3));
        float s1, c1, s2, c2;
        __sincosf(twopi * a2, &s1, &c1);
        __sincosf(twopi * a4, &s2, &c2);
        o.ev |= kEvIon;
        o.ivx = sg * rr1 * c1; o.ivy = sg * rr1 * s1; o.ivz = sg * rr2 * c2;
    }
}

// Excitation (electron_transport.cu:206-226): the mean loss of the levels.
__device__ inline void collide_excite(const Geom& g, const CollArgs& ca, const SmallTabs* stab, float T, float ve,
                                      float& vx, float& vy, float& vz)
{
#pragma clang fp contract(off)
    const float edep = exc_edep(stab, T, ca.exc_form);
    const float ratio = speed_from_ev(T - edep, ca.mc2_mc_ev, g.c) / ve;
    vx *= ratio; vy *= ratio; vz *= ratio;
}

// One candidate from decision to kinematics (the collision step of a lane
// on its own; the particle kernel runs the parts in passes per type).
__device__ inline void collide_candidate(const Geom& g, const CollArgs& ca, const SmallTabs* stab,
                                         uint32_t slot, float x, float y, float z, float& vx, float& vy,
                                         float& vz, float u1, float u2, float u3, float u4, bool want_ion,
                                         CollOut& o)
{
    CollDecision d;
    collide_decide(g, ca, stab, x, y, z, vx, vy, vz, u1, u2, d);
    o.ev = d.ev | d.typ;
    o.pt = d.pt;
    if (d.typ == 0) return;
    float T, ve;
    candidate_energy(g, ca, vx, vy, vz, T, ve);
    if (d.typ == 1) collide_elastic(ca, stab, slot, T, u4, vx, vy, vz);
    else if (d.typ == 2) collide_ionize(g, ca, slot, x, y, z, T, ve, u3, u4, want_ion, vx, vy, vz, o);
    else collide_excite(g, ca, stab, T, ve, vx, vy, vz);
}

}  // namespace gasx
