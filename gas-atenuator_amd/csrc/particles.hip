// Fused particle kernel: one pass over the FP32 SoA per step.
//
//   kPush     EM4.cu:36-177   gather E,H at x^{n+1/2}, relativistic push, half drift, BC
//   kCollide  electron_transport.cu:64-227 (+ find_CS0.m cross sections)
//   kDrift    EM2.cu:23-57    next half drift + BC  (x^{n+3/2})
//   kDeposit  EM7.cu:51-202   fixed-point current at x^{n+3/2} with v^{n+1}
//   kStageSec / kStageIon     electron_collision.m:63-75 append, staged per chunk
//                             in slot order so the final order is deterministic
//
// A workgroup owns chunks of kChunk consecutive slots (grid-stride over
// chunks).  The slot count is read once at kernel start; appends land in a
// later kernel (launch_append), so the count is stable inside this kernel.
#include "device_common.h"

namespace gasx {

// ---- deposit (EM7.cu:56-202 in fixed point) ---------------------------------
// Weights are computed in FP32 with contraction off (the library is built with
// -ffp-contract=off) so that the integer contributions are reproducible on the
// CPU; the int64 sum is then order independent.
__device__ __forceinline__ void deposit_one(const Geom& g, float x, float y, float z, float vx,
                                            float vy, float vz, float kq,
                                            unsigned long long* qr, unsigned long long* qz)
{
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
    int i0, i1, k0, k1;
    float wr0, wr1, wz0, wz1;
    half_axis(sr, g.nr, i0, i1, wr0, wr1);
    half_axis(sz, g.nz, k0, k1, wz0, wz1);
    const int c00 = i0 + k0 * g.nr, c10 = i1 + k0 * g.nr, c01 = i0 + k1 * g.nr, c11 = i1 + k1 * g.nr;
    const float w00 = wr0 * wz0, w10 = wr1 * wz0, w01 = wr0 * wz1, w11 = wr1 * wz1;
    atomicAdd(&qr[c00], (unsigned long long)(long long)rintf(w00 * qrv));
    atomicAdd(&qz[c00], (unsigned long long)(long long)rintf(w00 * qzv));
    if (w10 != 0.0f) {
        atomicAdd(&qr[c10], (unsigned long long)(long long)rintf(w10 * qrv));
        atomicAdd(&qz[c10], (unsigned long long)(long long)rintf(w10 * qzv));
    }
    if (w01 != 0.0f) {
        atomicAdd(&qr[c01], (unsigned long long)(long long)rintf(w01 * qrv));
        atomicAdd(&qz[c01], (unsigned long long)(long long)rintf(w01 * qzv));
    }
    if (w11 != 0.0f) {
        atomicAdd(&qr[c11], (unsigned long long)(long long)rintf(w11 * qrv));
        atomicAdd(&qz[c11], (unsigned long long)(long long)rintf(w11 * qzv));
    }
}

// ---- gather + push (EM4.cu:36-148) ------------------------------------------
__device__ __forceinline__ void push_one(const Geom& g, const PartArgs& pa, float x, float y,
                                         float z, float& vx, float& vy, float& vz)
{
    float r2 = x * x;
    r2 = r2 + y * y;
    float r = sqrtf(r2);
    float cph = 1.0f, sph = 0.0f;
    if (r > 0.0f) { cph = x / r; sph = y / r; }
    float vr = vx * cph + vy * sph;
    const float srh = r * g.inv_dr - 0.5f, srn = r * g.inv_dr;
    const float szh = z * g.inv_dz - 0.5f, szn = z * g.inv_dz;
    int i0, i1, k0, k1;
    float a0, a1, b0, b1;
    // Er: cell-centred in r, nodes in z (EM4.cu:41-60)
    half_axis(srh, g.nr, i0, i1, a0, a1);
    node_axis(szn, g.nz, k0, k1, b0, b1);
    const int nr = g.nr;
    float Er = b0 * (a0 * pa.er[i0 + k0 * nr] + a1 * pa.er[i1 + k0 * nr]) +
               b1 * (a0 * pa.er[i0 + k1 * nr] + a1 * pa.er[i1 + k1 * nr]);
    // Ez: nodes in r, cell-centred in z (EM4.cu:61-80)
    node_axis(srn, g.nr, i0, i1, a0, a1);
    half_axis(szh, g.nz, k0, k1, b0, b1);
    const int ld = nr + 1;
    float Ez = b0 * (a0 * pa.ez[i0 + k0 * ld] + a1 * pa.ez[i1 + k0 * ld]) +
               b1 * (a0 * pa.ez[i0 + k1 * ld] + a1 * pa.ez[i1 + k1 * ld]);
    // Hphi: cell-centred in both (EM4.cu:82-122)
    half_axis(srh, g.nr, i0, i1, a0, a1);
    float H = b0 * (a0 * pa.h[i0 + k0 * nr] + a1 * pa.h[i1 + k0 * nr]) +
              b1 * (a0 * pa.h[i0 + k1 * nr] + a1 * pa.h[i1 + k1 * nr]);
    // EM4.cu:126-144: magnetic term uses v^n; explicit (not Boris) update
    float fr = pa.q_over_m * (Er - vz * g.mu0 * H);
    float fz = pa.q_over_m * (Ez + vr * g.mu0 * H);
    float ax = fr * cph, ay = fr * sph;
    float v2 = vx * vx + vy * vy + vz * vz;
    float gm = 1.0f / sqrtf(1.0f - v2 * g.inv_c2);
    float ux = gm * vx + ax * g.dt, uy = gm * vy + ay * g.dt, uz = gm * vz + fz * g.dt;
    float gp = sqrtf(1.0f + (ux * ux + uy * uy + uz * uz) * g.inv_c2);
    vx = ux / gp; vy = uy / gp; vz = uz / gp;
}

// ---- elastic angle: get_ElasticSc_sub1.m:4-19 ---------------------------------
// log-DCS linear in log E at each node angle (griddata on the grid lines),
// trapezoid CDF over theta, inverse linear interpolation of U*CDF_end.
// Constant factors (2*pi*2/TCS) cancel in the normalised inverse.
__device__ float elastic_theta(const DevTables* t, const float* sind, float e_kev, float u)
{
    float lx = logf(e_kev);
    int n = t->n_dcs, i = 0;
    while (i < n - 2 && lx > t->dcs_loge[i + 1]) ++i;
    float w = (lx - t->dcs_loge[i]) / (t->dcs_loge[i + 1] - t->dcs_loge[i]);
    const float* v0 = &t->dcs_logv[i * kNAng];
    const float* v1 = &t->dcs_logv[(i + 1) * kNAng];
    float tot = 0.0f, prev = 0.0f;
    for (int k = 0; k < kNAng; ++k) {
        float pdf = sind[k] * __expf((1.0f - w) * v0[k] + w * v1[k]);
        if (k > 0) tot += 0.5f * (prev + pdf);
        prev = pdf;
    }
    float target = u * tot;
    float cdf = 0.0f;
    float p0 = sind[0] * __expf((1.0f - w) * v0[0] + w * v1[0]);
    int k = 0;
    float s = 1.0f;
    for (k = 0; k < kNAng - 1; ++k) {
        float p1 = sind[k + 1] * __expf((1.0f - w) * v0[k + 1] + w * v1[k + 1]);
        float nxt = cdf + 0.5f * (p0 + p1);
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

struct CollOut {
    int ev;          // 0 none, 1 elastic, 2 ionization, 3 excitation
    bool sec;        // secondary electron produced
    float svx, svy, svz;
    bool ion;
    float ivx, ivy, ivz;
    bool viol, sat;
};

// electron_transport.cu:64-227 with the find_CS0.m cross sections and the
// interp1 NaN gating (collide only if 15.6 eV <= T <= 20 keV).
__device__ __forceinline__ void collide_one(const Geom& g, const CollArgs& ca, const float* sind,
                                            uint32_t slot, float x, float y, float z,
                                            float& vx, float& vy, float& vz, bool want_ion,
                                            CollOut& o)
{
    o.ev = 0; o.sec = false; o.ion = false; o.viol = false; o.sat = false;
    U4 R = philox4x32_10(slot, ca.step, ca.run, 0u, ca.key0, ca.key1);
    float u1 = u01(R.a);
    if (!(u1 < ca.pmax)) return;                 // exact rejection: p_T <= pmax for every candidate
    float v2 = vx * vx + vy * vy + vz * vz;
    float T = ev_from_speed2(v2, g.inv_c2, ca.ev_per_mc2_em);
    if (!(T >= 15.6f && T <= 20000.0f)) return;
    const DevTables* t = ca.tab;
    float se, si, sx;
    cross_sections(t, T, se, si, sx);
    float r = sqrtf(x * x + y * y);
    int ir = (int)floorf(r * ca.inv_gdr), iz = (int)floorf(z * ca.inv_gdz);
    ir = ir < 0 ? 0 : (ir > ca.gnr - 1 ? ca.gnr - 1 : ir);
    iz = iz < 0 ? 0 : (iz > ca.gnz - 1 ? ca.gnz - 1 : iz);
    float dens = ca.dens[iz * ca.gnr + ir];
    float st = se + si + sx;
    float ve = sqrtf(v2);
    float pt = ve * ca.dt_coll * dens * st;      // electron_transport.cu:74-77
    o.viol = pt > ca.pmax && ca.pmax < 1.0f;
    if (!(u1 < pt)) return;                      // :85
    o.sat = pt > 1.0f;
    float u2 = u01(R.b), u3 = u01(R.c), u4 = u01(R.d);
    const float twopi = 6.28318530717959f;
    if (u2 * st < se) {                          // elastic (:86-137)
        float th = elastic_theta(t, sind, T * 1e-3f, u3);
        rotate_v(vx, vy, vz, th, twopi * u4, ca.rot_eps);
        o.ev = 1;
    } else if (u2 * st < se + si && T > 15.6f) { // ionization (:139-205)
        const float I = 15.6f;
        float T0 = 4.17f - 1000.0f / (T + 2.0f * I);
        float tt = 13.8f * T / (T + I);
        float Tm = (T - I) * 0.5f;
        float k1 = atanf((Tm - T0) / tt);
        float k2 = atanf(T0 / tt);
        float k0 = T0 + tt * tanf(u3 * (k1 + k2) - k2);
        k0 = fminf(fmaxf(k0, 0.0f), Tm);
        float E1 = k0, E2 = T - E1 - I;
        float Ep = E2 > E1 ? E2 : E1;
        float Es = E2 > E1 ? E1 : E2;
        float Ee = ca.mc2_mc_ev;
        float kt = k0 / T;
        float sp = sqrtf(kt / ((1.0f - kt) * T / (2.0f * Ee) + 1.0f));
        float ss = sqrtf((1.0f - kt) / (1.0f + k0 / (2.0f * Ee)));
        float thp = asinf(fminf(sp, 1.0f)), ths = asinf(fminf(ss, 1.0f));
        float php = twopi * u4, phs = php - 3.14159265358979f;
        float vp = speed_from_ev(Ep, Ee, g.c);
        float ratio = vp / ve;
        vx *= ratio; vy *= ratio; vz *= ratio;
        rotate_v(vx, vy, vz, thp, php, ca.rot_eps);
        float vs = speed_from_ev(Es, Ee, g.c);
        float sts, cts, sps, cps;
        sincosf(ths, &sts, &cts);
        sincosf(phs, &sps, &cps);
        o.sec = true;
        o.svx = vs * sts * cps; o.svy = vs * sts * sps; o.svz = vs * cts;   // :198-200
        if (want_ion) {                           // :202-204, Box-Muller normals
            U4 Q = philox4x32_10(slot, ca.step, ca.run, 1u, ca.key0, ca.key1);
            float a1 = u01(Q.a), a2 = u01(Q.b), a3 = u01(Q.c), a4 = u01(Q.d);
            float Tg = ca.temp ? ca.temp[iz * ca.gnr + ir] : ca.t_scalar;
            float sg = ca.ion_vth_k * sqrtf(Tg);
            float rr1 = sqrtf(-2.0f * logf(a1)), rr2 = sqrtf(-2.0f * logf(a3));
            float s1, c1, s2, c2;
            sincosf(twopi * a2, &s1, &c1);
            sincosf(twopi * a4, &s2, &c2);
            o.ion = true;
            o.ivx = sg * rr1 * c1; o.ivy = sg * rr1 * s1; o.ivz = sg * rr2 * c2;
        }
        o.ev = 2;
    } else {                                     // excitation (:206-226)
        float edep = exc_edep(t, T, ca.exc_form);
        float vp = speed_from_ev(T - edep, ca.mc2_mc_ev, g.c);
        float ratio = vp / ve;
        vx *= ratio; vy *= ratio; vz *= ratio;
        o.ev = 3;
    }
}

__device__ __forceinline__ void half_drift(const Geom& g, float& x, float& y, float& z, float vx,
                                           float vy, float vz)
{
    x = x + 0.5f * g.dt * vx;
    y = y + 0.5f * g.dt * vy;
    z = z + 0.5f * g.dt * vz;
}

// Exclusive rank of `pred` among the workgroup's lanes (lane order) and the
// workgroup total.  Every thread of the block must call it.
__device__ __forceinline__ int block_rank(bool pred, int* s_wave, int& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long m = __ballot(pred);
    unsigned long long lt = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
    int within = __popcll(lt);
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        int c = s_wave[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + within;
}

__device__ __forceinline__ void wave_count(bool pred, unsigned long long* ctr)
{
    unsigned long long m = __ballot(pred);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

template <uint32_t F>
__global__ __launch_bounds__(kChunk) void k_particles(Geom g, PartArgs pa, CollArgs ca)
{
    __shared__ int s_wave[kChunk / 64];
    __shared__ float s_sind[kNAng];
    if (F & kCollide) {
        for (int k = threadIdx.x; k < kNAng; k += blockDim.x) {
            float s = (k == 0 || k == 180) ? 0.0f : (k == 90 ? 1.0f : sinf(k * (3.14159265358979f / 180.0f)));
            s_sind[k] = s;
        }
        __syncthreads();
    }
    const int n = *pa.count;
    const int nchunks = (n + kChunk - 1) / kChunk;
    for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const int slot = chunk * kChunk + (int)threadIdx.x;
        const bool valid = slot < n;
        float x = 0.f, y = 0.f, z = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
        bool alive = false;
        if (valid) {
            z = pa.z[slot];
            alive = !is_dead(z);
            if (alive) {
                x = pa.x[slot]; y = pa.y[slot];
                vx = pa.vx[slot]; vy = pa.vy[slot]; vz = pa.vz[slot];
            }
        }
        const bool was_alive = alive;
        bool absorbed = false;
        if ((F & kPush) && alive) {
            push_one(g, pa, x, y, z, vx, vy, vz);
            half_drift(g, x, y, z, vx, vy, vz);
            if (!wedge_bc(g, x, y, z, vx, vy, vz)) { alive = false; absorbed = true; }
        }
        CollOut co;
        co.ev = 0; co.sec = false; co.ion = false; co.viol = false; co.sat = false;
        if ((F & kCollide) && alive) {
            collide_one(g, ca, s_sind, (uint32_t)slot, x, y, z, vx, vy, vz, (F & kStageIon) != 0, co);
        }
        const float cx = x, cy = y, cz = z;   // collision site
        bool has_sec = (F & kStageSec) && co.sec;
        float sx = x, sy = y, sz = z, svx = co.svx, svy = co.svy, svz = co.svz;
        bool sec_absorbed = false;
        if (F & kDrift) {
            if (alive) {
                half_drift(g, x, y, z, vx, vy, vz);
                if (!wedge_bc(g, x, y, z, vx, vy, vz)) { alive = false; absorbed = true; }
            }
            if (has_sec) {
                half_drift(g, sx, sy, sz, svx, svy, svz);
                if (!wedge_bc(g, sx, sy, sz, svx, svy, svz)) { has_sec = false; sec_absorbed = true; }
            }
        }
        if (F & kDeposit) {
            if (alive) deposit_one(g, x, y, z, vx, vy, vz, pa.kq, pa.qr, pa.qz);
            if (has_sec) deposit_one(g, sx, sy, sz, svx, svy, svz, pa.kq, pa.qr, pa.qz);
        }
        if (valid && was_alive) {
            if (alive) {
                pa.x[slot] = x; pa.y[slot] = y; pa.z[slot] = z;
                pa.vx[slot] = vx; pa.vy[slot] = vy; pa.vz[slot] = vz;
            } else {
                pa.z[slot] = __int_as_float(0x7fc00000);   // absorbed: z = NaN
            }
        }
        if (F & kStageSec) {
            int tot;
            int rank = block_rank(has_sec, s_wave, tot);
            if (has_sec) {
                const int dst = chunk * kChunk + rank;
                pa.sec[0][dst] = sx; pa.sec[1][dst] = sy; pa.sec[2][dst] = sz;
                pa.sec[3][dst] = svx; pa.sec[4][dst] = svy; pa.sec[5][dst] = svz;
            }
            if (threadIdx.x == 0) pa.sec_cnt[chunk] = tot;
        }
        if (F & kStageIon) {
            int tot;
            int rank = block_rank(co.ion, s_wave, tot);
            if (co.ion) {
                const int dst = chunk * kChunk + rank;
                // ion at the ionization site (electron_collision.m:63)
                pa.ion[0][dst] = cx; pa.ion[1][dst] = cy; pa.ion[2][dst] = cz;
                pa.ion[3][dst] = co.ivx; pa.ion[4][dst] = co.ivy; pa.ion[5][dst] = co.ivz;
            }
            if (threadIdx.x == 0) pa.ion_cnt[chunk] = tot;
        }
        if (pa.counters) {
            if (F & kCollide) {
                wave_count(co.ev == 1, &pa.counters[kCElastic]);
                wave_count(co.ev == 2, &pa.counters[kCIonize]);
                wave_count(co.ev == 3, &pa.counters[kCExcite]);
                wave_count(co.viol, &pa.counters[kCViol]);
                wave_count(co.sat, &pa.counters[kCSat]);
            }
            wave_count(absorbed || sec_absorbed, &pa.counters[kCAbsorbed]);
        }
    }
}

// ---- append: exclusive scan of per-chunk counts, then copy ------------------
// src_count: slot count of the population whose kernel staged the entries
// (it fixes the number of chunks); dst_count: the population appended to.
// tmp2[0] <- old dst count, tmp2[1] <- number of staged chunks.
__global__ __launch_bounds__(1024) void k_append_scan(const int32_t* src_count, int32_t* dst_count,
                                                      int32_t cap, const int32_t* stage_cnt,
                                                      int32_t* offs, int32_t* tmp2,
                                                      int32_t max_chunks,
                                                      unsigned long long* dropped)
{
    __shared__ int s_part[1024];
    const int ns = *src_count;
    const int n = *dst_count;
    int nch = (ns + kChunk - 1) / kChunk;
    if (nch > max_chunks) nch = max_chunks;
    const int per = (nch + blockDim.x - 1) / blockDim.x;
    const int beg = threadIdx.x * per;
    const int end = min(beg + per, nch);
    int local = 0;
    for (int i = beg; i < end; ++i) local += stage_cnt[i];
    s_part[threadIdx.x] = local;
    __syncthreads();
    // Hillis-Steele inclusive scan over the 1024 partials
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        int v = (int)threadIdx.x >= o ? s_part[threadIdx.x - o] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = s_part[threadIdx.x] - local;
    for (int i = beg; i < end; ++i) {
        offs[i] = run;
        run += stage_cnt[i];
    }
    if (threadIdx.x == blockDim.x - 1) {
        const int total = s_part[threadIdx.x];
        const int room = cap - n;
        const int app = total < room ? total : room;
        tmp2[0] = n;
        tmp2[1] = nch;
        *dst_count = n + app;
        if (total > app && dropped) atomicAdd(dropped, (unsigned long long)(total - app));
    }
}

__global__ __launch_bounds__(kChunk) void k_append_copy(float* d0, float* d1, float* d2, float* d3,
                                                        float* d4, float* d5, int32_t cap,
                                                        const float* s0, const float* s1,
                                                        const float* s2, const float* s3,
                                                        const float* s4, const float* s5,
                                                        const int32_t* stage_cnt,
                                                        const int32_t* offs, const int32_t* tmp2)
{
    const int base = tmp2[0];
    const int nch = tmp2[1];
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const int cnt = stage_cnt[c];
        if ((int)threadIdx.x < cnt) {
            const int dst = base + offs[c] + (int)threadIdx.x;
            const int src = c * kChunk + (int)threadIdx.x;
            if (dst < cap) {
                d0[dst] = s0[src]; d1[dst] = s1[src]; d2[dst] = s2[src];
                d3[dst] = s3[src]; d4[dst] = s4[src]; d5[dst] = s5[src];
            }
        }
    }
}

hipError_t launch_append(float* const dst[6], int32_t* dst_count, int32_t cap,
                         float* const stage[6], const int32_t* stage_cnt,
                         const int32_t* src_count, int32_t* scan_tmp, int32_t* tmp2,
                         int32_t max_chunks, unsigned long long* dropped, hipStream_t s)
{
    hipLaunchKernelGGL(k_append_scan, dim3(1), dim3(1024), 0, s, src_count, dst_count, cap,
                       stage_cnt, scan_tmp, tmp2, max_chunks, dropped);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int blocks = max_chunks < 4096 ? (max_chunks > 0 ? max_chunks : 1) : 4096;
    hipLaunchKernelGGL(k_append_copy, dim3(blocks), dim3(kChunk), 0, s, dst[0], dst[1], dst[2],
                       dst[3], dst[4], dst[5], cap, stage[0], stage[1], stage[2], stage[3],
                       stage[4], stage[5], stage_cnt, scan_tmp, tmp2);
    return hipGetLastError();
}

// ---- cross sections (find_CS0.m) --------------------------------------------
__global__ void k_cross_sections(const DevTables* t, const float* e_kev, int64_t n, float* se,
                                 float* si, float* sx)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float T = e_kev[i] * 1000.0f;
        const float nanv = __int_as_float(0x7fc00000);
        float a, b, c;
        // elastic: find_CS0.m:16-25 (NaN outside 1 eV .. 20 keV)
        if (e_kev[i] >= 0.05f)
            a = (T >= t->el_e[0] && T <= t->el_e[t->n_el - 1]) ? interp_in(t->el_e, t->el_s, t->n_el, T) : nanv;
        else
            a = (T >= t->sel_e[0] && T <= t->sel_e[t->n_sel - 1]) ? interp_in(t->sel_e, t->sel_s, t->n_sel, T) : nanv;
        b = (T >= t->ion_e[0] && T <= t->ion_e[t->n_ion - 1]) ? interp_in(t->ion_e, t->ion_s, t->n_ion, T) : nanv;
        float s = 0.0f;
        for (int j = 0; j < t->n_exc; ++j) {
            float W = t->exc_w[j];
            float wr = W / T;
            float bb = powf(wr, t->exc_om[j]);
            float cc = powf(1.0f - powf(wr, t->exc_ga[j]), t->exc_nu[j]);
            s += t->exc_q[j] * bb * cc;
        }
        c = s;
        se[i] = a; si[i] = b; sx[i] = c;
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
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_particles<F>, dim3(blocks), dim3(kChunk), 0, s, g, pa, ca);
    return hipGetLastError();
}

hipError_t launch_particles(uint32_t flags, const Geom& g, const PartArgs& pa, const CollArgs& ca,
                            int max_blocks, hipStream_t s)
{
    int blocks = max_blocks > 0 ? max_blocks : 1;
    switch (flags) {
    case kDrift: return launch_t<kDrift>(g, pa, ca, blocks, s);
    case kDeposit: return launch_t<kDeposit>(g, pa, ca, blocks, s);
    case kDrift | kDeposit: return launch_t<kDrift | kDeposit>(g, pa, ca, blocks, s);
    case kPush: return launch_t<kPush>(g, pa, ca, blocks, s);
    case kPush | kDrift | kDeposit: return launch_t<kPush | kDrift | kDeposit>(g, pa, ca, blocks, s);
    case kCollide: return launch_t<kCollide>(g, pa, ca, blocks, s);
    case kCollide | kStageSec: return launch_t<kCollide | kStageSec>(g, pa, ca, blocks, s);
    case kCollide | kStageSec | kStageIon:
        return launch_t<kCollide | kStageSec | kStageIon>(g, pa, ca, blocks, s);
    case kPush | kCollide: return launch_t<kPush | kCollide>(g, pa, ca, blocks, s);
    case kPush | kCollide | kStageSec: return launch_t<kPush | kCollide | kStageSec>(g, pa, ca, blocks, s);
    case kPush | kCollide | kStageSec | kStageIon:
        return launch_t<kPush | kCollide | kStageSec | kStageIon>(g, pa, ca, blocks, s);
    case kPush | kCollide | kStageIon: return launch_t<kPush | kCollide | kStageIon>(g, pa, ca, blocks, s);
    case kPush | kCollide | kDrift | kDeposit | kStageIon:
        return launch_t<kPush | kCollide | kDrift | kDeposit | kStageIon>(g, pa, ca, blocks, s);
    case kPush | kCollide | kDrift | kDeposit:
        return launch_t<kPush | kCollide | kDrift | kDeposit>(g, pa, ca, blocks, s);
    case kPush | kCollide | kDrift | kDeposit | kStageSec:
        return launch_t<kPush | kCollide | kDrift | kDeposit | kStageSec>(g, pa, ca, blocks, s);
    case kPush | kCollide | kDrift | kDeposit | kStageSec | kStageIon:
        return launch_t<kPush | kCollide | kDrift | kDeposit | kStageSec | kStageIon>(g, pa, ca, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace gasx
