/*
 * oracle.c -- FP64 CPU restatement of the reference electron PIC step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libgasx.so.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function restates the reference file:line named in its comment.
 * Where the reference's MATLAB and CUDA versions disagree, the choice is a
 * flag in orc_params (ez_metric, axis_j_sign, exc_form, rot_eps); defaults
 * are set by the Python side.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the particle, field and collision loops (orc_set_threads).
 * 1 by default: the checker's FP64 sums then have one fixed order.  With
 * more threads every particle and every cell is still computed by the same
 * arithmetic; only the deposit's per-thread partial currents are summed in
 * thread order, which changes the last bits of J. */
static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* Tables                                                              */
/* ------------------------------------------------------------------ */

/* MATLAB interp1(x, y, xq) 'linear': NaN outside [x(1), x(end)]
 * (find_CS0.m:21,25,33 rely on this NaN to suppress collisions). */
double orc_interp1(const double* x, const double* y, int n, double xq)
{
    if (!(xq >= x[0]) || !(xq <= x[n - 1])) return NAN;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (x[mid] <= xq) lo = mid; else hi = mid;
    }
    if (xq == x[lo]) return y[lo];
    if (xq == x[hi]) return y[hi];
    double s = (xq - x[lo]) / (x[hi] - x[lo]);
    return y[lo] + s * (y[hi] - y[lo]);
}

/* Table preparation.
 *  find_CS0.m:16-33  : elastic TCS (keV, a0^2) , small-energy TCS (eV, 1e-20 m^2),
 *                      ionization TCS (eV, 1e-20 m^2)
 *  electron_collision.m:12-33 : merged TCS / log-DCS used for angle sampling
 *  electron_collision.m:36-41 : excitation fit parameters (columns j,A,omega,v,miu)
 */
int orc_tables_build(orc_tables* t,
                     const double* t_el, int n_el,
                     const double* t_sel, int n_sel,
                     const double* t_sel1, int n_sel1,
                     const double* t_del, int w_del,
                     const double* t_sdel, int w_sdel,
                     const double* t_ion, int n_ion,
                     const double* t_exc, int n_exc)
{
    const double sqr_a0 = 2.8002852e-21;  /* electron_collision.m:12 */
    memset(t, 0, sizeof(*t));
    if (n_el > ORC_MAXT || n_sel > ORC_MAXT || n_ion > ORC_MAXT || n_exc > 16) return 1;
    int n_small = n_sel1 - 1;                      /* electron_collision.m:17 drops the last row */
    int n_dcs = (w_sdel - 1) + (w_del - 1);        /* electron_collision.m:23-25 */
    if (n_small + n_el != n_dcs || n_dcs > ORC_MAXT) return 2;

    t->n_el = n_el;
    for (int i = 0; i < n_el; ++i) { t->el_e[i] = t_el[i]; t->el_s[i] = t_el[n_el + i]; }
    t->n_sel = n_sel;
    for (int i = 0; i < n_sel; ++i) { t->sel_e[i] = t_sel[2 * i]; t->sel_s[i] = t_sel[2 * i + 1]; }
    t->n_ion = n_ion;
    for (int i = 0; i < n_ion; ++i) { t->ion_e[i] = t_ion[2 * i]; t->ion_s[i] = t_ion[2 * i + 1]; }
    t->n_exc = n_exc;
    for (int i = 0; i < n_exc; ++i) {
        t->exc_w[i] = t_exc[5 * i + 0];
        t->exc_a[i] = t_exc[5 * i + 1];
        t->exc_om[i] = t_exc[5 * i + 2];
        t->exc_nu[i] = t_exc[5 * i + 3];   /* column "v"   -> exponent nu */
        t->exc_ga[i] = t_exc[5 * i + 4];   /* column "miu" -> exponent gamma */
    }

    /* merged TCS: [t_sel_1(1:end-1) converted to keV and a0^2 per atom ; t_el'] */
    t->n_dcs = n_dcs;
    for (int i = 0; i < n_small; ++i)
        t->dcs_tcs[i] = t_sel1[2 * i + 1] * 1e-20 / sqr_a0 / 2.0;   /* electron_collision.m:15 */
    for (int i = 0; i < n_el; ++i) t->dcs_tcs[n_small + i] = t->el_s[i];
    /* merged DCS: [t_sdel , t_del(:,2:end)]; row 0 energies, rows 1..181 angles 0..180 */
    for (int j = 0; j < n_dcs; ++j) {
        const double* tab; int w, col;
        if (j < w_sdel - 1) { tab = t_sdel; w = w_sdel; col = j + 1; }
        else { tab = t_del; w = w_del; col = j - (w_sdel - 1) + 1; }
        t->dcs_e[j] = tab[col];
        t->dcs_loge[j] = log(tab[col]);                              /* xInterp, :32 */
        for (int k = 0; k < ORC_NANG; ++k)
            t->dcs_logv[j * ORC_NANG + k] = log(tab[(k + 1) * w + col]); /* V_interp, :30 */
    }
    return 0;
}

/* find_CS0.m:14-56: sigma_el, sigma_ion, sigma_exc for energies in keV. */
static void cs_one(const orc_tables* t, double e, double* se, double* si, double* sx)
{
    const double sqr_a0 = 2.8002852e-21;          /* find_CS0.m:17 */
    if (e >= 0.05)
        *se = orc_interp1(t->el_e, t->el_s, t->n_el, e) * sqr_a0 * 2.0;   /* :21 */
    else
        *se = orc_interp1(t->sel_e, t->sel_s, t->n_sel, e * 1000.0) * 1e-20; /* :25 */
    double T = e * 1000.0;
    *si = orc_interp1(t->ion_e, t->ion_s, t->n_ion, T) * 1e-20;       /* :33 */
    const double q0 = 6.542e-14;                  /* :42 */
    double s = 0.0;
    for (int i = 0; i < t->n_exc; ++i) {          /* :47-55 */
        double a = q0 * t->exc_a[i] / (t->exc_w[i] * t->exc_w[i]);
        double b = pow(t->exc_w[i] / T, t->exc_om[i]);
        double c = pow(1.0 - pow(t->exc_w[i] / T, t->exc_ga[i]), t->exc_nu[i]);
        s += a * b * c;
    }
    *sx = s * 1e-4;                               /* :56 */
}

void orc_cross_sections(const orc_tables* t, const double* e_kev, int64_t n,
                        double* sig_el, double* sig_ion, double* sig_exc)
{
    for (int64_t i = 0; i < n; ++i) cs_one(t, e_kev[i], &sig_el[i], &sig_ion[i], &sig_exc[i]);
}

/* MATLAB sind() on integer degrees: exact zeros at 0 and 180. */
static double sind_int(int k)
{
    if (k == 0 || k == 180) return 0.0;
    if (k == 90) return 1.0;
    return sin(k * M_PI / 180.0);
}

/* Elastic polar angle, get_ElasticSc_sub1.m:4-19 with the merged tables of
 * electron_collision.m:22-33.  griddata at the node angles 0..180 reduces to
 * linear interpolation of log-DCS in log E along each angle row.  The CDF is
 * cumtrapz over theta in radians and theta = interp1(CDF, theta, U*CDF_end). */
double orc_elastic_theta(const orc_tables* t, double e_kev, double u)
{
    double x = log(e_kev);
    int n = t->n_dcs, i = 0;
    while (i < n - 2 && x > t->dcs_loge[i + 1]) ++i;
    double w = (x - t->dcs_loge[i]) / (t->dcs_loge[i + 1] - t->dcs_loge[i]);
    const double* v0 = &t->dcs_logv[i * ORC_NANG];
    const double* v1 = &t->dcs_logv[(i + 1) * ORC_NANG];
    double cdf[ORC_NANG];
    double dth = M_PI / 180.0, prev = 0.0;
    cdf[0] = 0.0;
    for (int k = 0; k < ORC_NANG; ++k) {
        double a = (1.0 - w) * v0[k] + w * v1[k];
        double pdf = 2.0 * M_PI * sind_int(k) * exp(a) * 2.0;   /* TCS_A cancels in U*CDF_end */
        if (k > 0) cdf[k] = cdf[k - 1] + 0.5 * (prev + pdf) * dth;
        prev = pdf;
    }
    double target = u * cdf[ORC_NANG - 1];
    int k = 0;
    while (k < ORC_NANG - 2 && cdf[k + 1] < target) ++k;
    double s = (target - cdf[k]) / (cdf[k + 1] - cdf[k]);
    return (k + s) * dth;
}

/* The distribution of orc_elastic_theta, sampled by rejection instead of by
 * inversion (the HIP kernel's method).  The inverse-CDF picks interval k with
 * probability m_k(w) = p_k(w) + p_{k+1}(w), p_k(w) = sind(k) exp((1-w)V_i[k] +
 * w V_{i+1}[k]), and theta uniformly inside it.  By convexity of exp,
 * m_k(w) <= (1-w) m_k(i) + w m_k(i+1) =: M_k, so: pick node i or i+1 with
 * weights (1-w) Z_i and w Z_{i+1} (Z = sum of m), draw k from that node's
 * interval distribution, accept with probability m_k(w)/M_k, and place
 * theta uniformly in the interval.  Uniforms come from Philox blocks 2, 3, ...
 * of the particle's counter. */
static void node_masses(const orc_tables* t, int j, double* m, double* Z)
{
    double prev = 0.0, z = 0.0;
    for (int k = 0; k < ORC_NANG; ++k) {
        double p = sind_int(k) * exp(t->dcs_logv[j * ORC_NANG + k]);
        if (k > 0) { m[k - 1] = prev + p; z += prev + p; }
        prev = p;
    }
    *Z = z;
}

double orc_elastic_theta_rej(const orc_tables* t, double e_kev, uint32_t slot, uint32_t step,
                             uint32_t run, uint64_t seed)
{
    double x = log(e_kev);
    int n = t->n_dcs, i = 0;
    while (i < n - 2 && x > t->dcs_loge[i + 1]) ++i;
    double w = (x - t->dcs_loge[i]) / (t->dcs_loge[i + 1] - t->dcs_loge[i]);
    double m0[ORC_NANG], m1[ORC_NANG], Z0, Z1;
    node_masses(t, i, m0, &Z0);
    node_masses(t, i + 1, m1, &Z1);
    const double* v0 = &t->dcs_logv[i * ORC_NANG];
    const double* v1 = &t->dcs_logv[(i + 1) * ORC_NANG];
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    for (uint32_t it = 0; it < 64; ++it) {
        uint32_t ctr[4] = { slot, step, run, 2u + it }, o[4];
        orc_philox4x32_10(ctr, key, o);
        double a = orc_u01(o[0]), b = orc_u01(o[1]), c = orc_u01(o[2]), d = orc_u01(o[3]);
        double wa = (1.0 - w) * Z0, wb = w * Z1;
        const double* mm = (a * (wa + wb) < wa) ? m0 : m1;
        double Z = (mm == m0) ? Z0 : Z1;
        double target = b * Z, acc = 0.0;
        int k = 0;
        while (k < ORC_NANG - 2 && acc + mm[k] <= target) { acc += mm[k]; ++k; }
        double M = (1.0 - w) * m0[k] + w * m1[k];
        double pk = sind_int(k) * exp((1.0 - w) * v0[k] + w * v1[k]);
        double pk1 = sind_int(k + 1) * exp((1.0 - w) * v0[k + 1] + w * v1[k + 1]);
        if (c * M < pk + pk1) return (k + d) * (M_PI / 180.0);
    }
    /* not reached in practice (acceptance >= ~0.9); keep the exact sampler */
    uint32_t ctr[4] = { slot, step, run, 66u }, o[4];
    orc_philox4x32_10(ctr, key, o);
    return orc_elastic_theta(t, e_kev, orc_u01(o[0]));
}

/* Green-Sawada secondary energy and binary-encounter angles,
 * get_ImpIonized.m:13-63 (== electron_transport.cu:141-171). */
void orc_ionize(double T, double rn, double e_e,
                double* k0o, double* ep, double* es, double* th_p, double* th_s)
{
    const double I = 15.6, T_a = 1000.0, T_b = 2 * I, t_b = I, T_s = 4.17, t_s = 13.8;
    double T0 = T_s - T_a / (T + T_b);
    double t = t_s * T / (T + t_b);
    double Tm = (T - I) / 2.0;
    double k1 = atan((Tm - T0) / t);
    double k2 = atan(T0 / t);
    double k0 = T0 + t * tan(rn * (k1 + k2) - k2);
    if (k0 < 0.0) k0 = 0.0;   /* rounding guard: k0 > 0 for rn > 0 in exact arithmetic */
    if (k0 > Tm) k0 = Tm;
    double E1 = k0, E2 = T - E1 - I;
    if (E2 > E1) { *ep = E2; *es = E1; } else { *ep = E1; *es = E2; }
    double sp = sqrt((k0 / T) / ((1.0 - k0 / T) * T / (2.0 * e_e) + 1.0));
    double ss = sqrt((1.0 - k0 / T) / (1.0 + k0 / (2.0 * e_e)));
    *th_p = asin(sp);
    *th_s = asin(ss);
    *k0o = k0;
}

/* Mean excitation loss: get_ImpExc2.m:19-32 (form 0) or the powf variant of
 * electron_transport.cu:212-219 (form 1).  q0 cancels in the ratio. */
double orc_exc_edep(const orc_tables* t, double T, int form)
{
    double k1 = 0.0, k2 = 0.0;
    for (int i = 0; i < t->n_exc; ++i) {
        double W = t->exc_w[i];
        double a = t->exc_a[i] / (W * W);
        double b = pow(W / T, t->exc_om[i]);
        double c = form == 0 ? pow(1.0 - pow(W / T, t->exc_ga[i]), t->exc_nu[i])
                             : pow(pow(1.0 - W / T, t->exc_ga[i]), t->exc_nu[i]);
        k1 += a * b * c * W;
        k2 += a * b * c;
    }
    return k1 / k2;
}

/* rotate.m:4-13; the degeneracy test v^2 - vz^2 is evaluated as vx^2+vy^2. */
void orc_rotate(double* vx, double* vy, double* vz, double th, double ph, double eps)
{
    double x = *vx, y = *vy, zz = *vz;
    double vp2 = x * x + y * y;
    double v = sqrt(vp2 + zz * zz);
    double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    if (vp2 > eps) {
        double vp = sqrt(vp2);
        *vx = x * ct + st / vp * (x * zz * cp - v * y * sp);
        *vy = y * ct + st / vp * (y * zz * cp + v * x * sp);
        *vz = zz * ct - vp * st * cp;
    } else {
        *vx = v * st * cp;
        *vy = v * st * sp;
        *vz = v * ct;
    }
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11.                                          */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* u in (0,1): (2k+1) * 2^-24 with k the top 23 bits; exact in FP32 and FP64. */
double orc_u01(uint32_t x)
{
    return (double)((x >> 8) | 1u) * (1.0 / 16777216.0);
}

/* ------------------------------------------------------------------ */
/* FDTD                                                                */
/* ------------------------------------------------------------------ */

/* EM3_1.cu:9-15 : Hphi += tau/mu * (dEz/dr - dEr/dz). */
void orc_fdtd_h(double* h, const double* er, const double* ez, int nr, int nz,
                double mu, double dr, double dz, double tau)
{
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nr; ++i) {
            double dez = ez[(i + 1) + k * (nr + 1)] - ez[i + k * (nr + 1)];
            double der = er[i + (k + 1) * nr] - er[i + k * nr];
            h[i + k * nr] = h[i + k * nr] + (tau / (mu * dr) * dez - tau / (mu * dz) * der);
        }
}

/* EM3.cu:14-42 : Jacobi E update (Er0/Ez0 are the pre-update copies,
 * electricmagneticfield6.m:149 passes the same arrays as inputs and outputs). */
void orc_fdtd_e(double* er, double* ez, const double* h, const double* jr, const double* jz,
                int nr, int nz, double eps, double dr, double dz, double tau,
                int ez_metric, int axis_j_sign)
{
    size_t ner = (size_t)nr * (nz + 1), nez = (size_t)(nr + 1) * nz;
    double* er0 = (double*)malloc(ner * sizeof(double));
    double* ez0 = (double*)malloc(nez * sizeof(double));
    memcpy(er0, er, ner * sizeof(double));
    memcpy(ez0, ez, nez * sizeof(double));
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i < nr; ++i) {
            int n = i + k * nr;
            if (k == 0) er[n] = er0[n + nr];                  /* :22 */
            else if (k == nz) er[n] = er0[n - nr];            /* :26 */
            else er[n] = er0[n] - tau / eps * (jr[n - nr] + jr[n]) / 2
                         - tau / (eps * dz) * (h[n] - h[n - nr]);   /* :29 */
        }
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nr; ++i) {                        /* Ez(nr,k) is never written */
            int nzi = i + k * (nr + 1), n = i + k * nr;
            if (i == 0)
                ez[nzi] = ez0[nzi] + axis_j_sign * tau / eps * jz[n] + 4 * tau / (eps * dr) * h[n]; /* :35 */
            else {
                double rs = (ez_metric ? (i + 1) : i) * dr;                                  /* :39 (D1) */
                ez[nzi] = ez0[nzi] - tau / eps * (jz[n - 1] + jz[n]) / 2
                          + tau / eps * (1 / (2 * rs) + 1 / dr) * h[n]
                          + tau / eps * (1 / (2 * rs) - 1 / dr) * h[n - 1];
            }
        }
    free(er0);
    free(ez0);
}

/* electricmagneticfield6.m:146-151: H back half a substep, n_sub leapfrog
 * substeps with J frozen, H forward half a substep.  Substep = ddt/n_sub. */
void orc_fdtd_push(const orc_params* p, double* er, double* ez, double* h,
                   const double* jr, const double* jz, int n_push)
{
    double dtf = p->ddt / p->n_sub;
    for (int s = 0; s < n_push; ++s) {
        orc_fdtd_h(h, er, ez, p->nr, p->nz, p->mu0, p->dr, p->dz, -dtf / 2);
        for (int i = 0; i < p->n_sub; ++i) {
            orc_fdtd_h(h, er, ez, p->nr, p->nz, p->mu0, p->dr, p->dz, dtf);
            orc_fdtd_e(er, ez, h, jr, jz, p->nr, p->nz, p->eps0, p->dr, p->dz, dtf,
                       p->ez_metric, p->axis_j_sign);
        }
        orc_fdtd_h(h, er, ez, p->nr, p->nz, p->mu0, p->dr, p->dz, dtf / 2);
    }
}

/* ------------------------------------------------------------------ */
/* Particles                                                           */
/* ------------------------------------------------------------------ */

/* wedge wrap with velocity rotation + z reflection: EM2.cu:31-56 / EM4.cu:152-177 */
static void wedge_bc(const orc_params* p, double* x, double* y, double* z,
                     double* vx, double* vy, double* vz, double* r, double* phi, int32_t* del)
{
    double Rp = p->nr * p->dr, Lp = p->nz * p->dz, PHIp = p->phip;
    *r = sqrt(*x * *x + *y * *y);
    *phi = atan(*y / *x);
    if (*r > Rp) *del = 1;
    double cP = cos(PHIp), sP = sin(PHIp);
    while (*phi < 0) {
        *phi += PHIp;
        double vx1 = *vx * cP - *vy * sP;
        *vy = *vx * sP + *vy * cP;
        *vx = vx1;
    }
    while (*phi > PHIp) {
        *phi -= PHIp;
        double vx1 = *vx * cP + *vy * sP;
        *vy = -*vx * sP + *vy * cP;
        *vx = vx1;
    }
    *x = *r * cos(*phi);
    *y = *r * sin(*phi);
    if (*z > Lp) { *z = 2 * Lp - *z; *vz = -*vz; }
    if (*z < 0) { *z = -*z; *vz = -*vz; }
}

/* EM2.cu:23-57: half drift, wedge/z boundary conditions, v_r. */
void orc_em2(const orc_params* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
             double* r, double* phi, double* vr, int32_t* del, int64_t n)
{
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t i = 0; i < n; ++i) {
        if (del[i] == 1) continue;
        x[i] += 0.5 * vx[i] * p->ddt;
        y[i] += 0.5 * vy[i] * p->ddt;
        z[i] += 0.5 * vz[i] * p->ddt;
        wedge_bc(p, &x[i], &y[i], &z[i], &vx[i], &vy[i], &vz[i], &r[i], &phi[i], &del[i]);
        vr[i] = vx[i] * cos(phi[i]) + vy[i] * sin(phi[i]);
    }
}

/* EM7.cu:51-202: clamped area weighting onto the Hphi grid,
 * J += FN*b*q*v / V(i),  V(i) = PHIp/2*((i+1)^2-i^2) dr^2 dz. */
static void em7_range(const orc_params* p, const double* r, const double* z, const double* q,
                      const double* vr, const double* vz, double* jr, double* jz,
                      const int32_t* del, int64_t m0, int64_t m1)
{
    int nr = p->nr, nz = p->nz;
    double dr = p->dr, dz = p->dz;
    for (int64_t m = m0; m < m1; ++m) {
        if (del[m] == 1) continue;       /* EM7.cu:163-173 zero weights */
        double r0 = r[m], z0 = z[m];
        int ar = (int)floor(r0 / dr - 0.5), az = (int)floor(z0 / dz - 0.5);
        int ci[2], ck[2]; double wr[2], wz[2]; int ncr, ncz;
        if (ar < 0) { ncr = 1; ci[0] = ar + 1; wr[0] = 1.0; }
        else if (ar >= nr - 1) { ncr = 1; ci[0] = ar; wr[0] = 1.0; }
        else { ncr = 2; ci[0] = ar; ci[1] = ar + 1;
               wr[0] = ((ar + 1.5) * dr - r0) / dr; wr[1] = (r0 - (ar + 0.5) * dr) / dr; }
        if (az < 0) { ncz = 1; ck[0] = az + 1; wz[0] = 1.0; }
        else if (az >= nz - 1) { ncz = 1; ck[0] = az; wz[0] = 1.0; }
        else { ncz = 2; ck[0] = az; ck[1] = az + 1;
               wz[0] = ((az + 1.5) * dz - z0) / dz; wz[1] = (z0 - (az + 0.5) * dz) / dz; }
        for (int a = 0; a < ncr; ++a) {
            int i = ci[a];
            double V = p->phip / 2 * ((((i + 1) * dr) * ((i + 1) * dr) - (i * dr) * (i * dr)) * dz);
            for (int b = 0; b < ncz; ++b) {
                double w = wr[a] * wz[b];
                jr[i + ck[b] * nr] += p->fn * w * q[m] * vr[m] / V;
                jz[i + ck[b] * nr] += p->fn * w * q[m] * vz[m] / V;
            }
        }
    }
}

void orc_em7(const orc_params* p, const double* r, const double* z, const double* q,
             const double* vr, const double* vz, double* jr, double* jz,
             const int32_t* del, int64_t n)
{
    const int T = g_threads;
    if (T <= 1 || n < 4096) {
        em7_range(p, r, z, q, vr, vz, jr, jz, del, 0, n);
        return;
    }
    /* per-thread partial currents over contiguous particle ranges, summed in
     * thread order */
    size_t nc = (size_t)p->nr * p->nz;
    double* part = (double*)calloc((size_t)T * 2 * nc, sizeof(double));
    if (!part) {
        em7_range(p, r, z, q, vr, vz, jr, jz, del, 0, n);
        return;
    }
#pragma omp parallel num_threads(T)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        int64_t m0 = n * t / nt, m1 = n * (t + 1) / nt;
        em7_range(p, r, z, q, vr, vz, part + (size_t)t * 2 * nc, part + ((size_t)t * 2 + 1) * nc, del, m0, m1);
    }
    for (int t = 0; t < T; ++t)
        for (size_t c = 0; c < nc; ++c) {
            jr[c] += part[(size_t)t * 2 * nc + c];
            jz[c] += part[((size_t)t * 2 + 1) * nc + c];
        }
    free(part);
}

static int ifloor(double v) { return (int)floor(v); }

/* EM4.cu:36-177: bilinear gather with the EM4 edge rules, Lorentz force with
 * v^n in the magnetic term, relativistic momentum update, half drift + BC. */
void orc_em4(const orc_params* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
             double* r, double* phi, const double* vr, const double* er, const double* ez,
             const double* h, const double* q, const double* m, double* ke, int32_t* del, int64_t n)
{
    int R = p->nr, Z = p->nz;
    double dr = p->dr, dz = p->dz, c = p->c, mu = p->mu0, dt = p->ddt;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t i = 0; i < n; ++i) {
        if (del[i] == 1) continue;
        double r0 = r[i], z0 = z[i];
        int ar, az;
        double E_r, E_z, Hp;
        /* Er (EM4.cu:41-60); indices past the array end carry zero weight */
        ar = ifloor(r0 / dr - 0.5); az = ifloor(z0 / dz);
        {
            int kz1 = az + 1 > Z ? Z : az + 1;
            double wz0 = (az * dz + dz - z0), wz1 = (z0 - az * dz);
            if (ar < 0) E_r = (er[(ar + 1) + az * R] * wz0 + er[(ar + 1) + kz1 * R] * wz1) / dz;
            else if (ar >= R - 1) E_r = (er[ar + az * R] * wz0 + er[ar + kz1 * R] * wz1) / dz;
            else E_r = (er[ar + az * R] * ((ar + 1.5) * dr - r0) * ((az + 1) * dz - z0)
                        + er[(ar + 1) + az * R] * (r0 - (ar + 0.5) * dr) * ((az + 1) * dz - z0)
                        + er[ar + kz1 * R] * ((ar + 1.5) * dr - r0) * (z0 - az * dz)
                        + er[(ar + 1) + kz1 * R] * (r0 - (ar + 0.5) * dr) * (z0 - az * dz)) / (dr * dz);
        }
        /* Ez (EM4.cu:61-80) */
        ar = ifloor(r0 / dr); az = ifloor(z0 / dz - 0.5);
        {
            int ir1 = ar + 1 > R ? R : ar + 1;
            double wr0 = (ar * dr + dr - r0), wr1 = (r0 - ar * dr);
            if (az < 0) E_z = (ez[ar + (az + 1) * (R + 1)] * wr0 + ez[ir1 + (az + 1) * (R + 1)] * wr1) / dr;
            else if (az >= Z - 1) E_z = (ez[ar + az * (R + 1)] * wr0 + ez[ir1 + az * (R + 1)] * wr1) / dr;
            else E_z = (ez[ar + az * (R + 1)] * ((ar + 1) * dr - r0) * ((az + 1.5) * dz - z0)
                        + ez[ir1 + az * (R + 1)] * (r0 - ar * dr) * ((az + 1.5) * dz - z0)
                        + ez[ar + (az + 1) * (R + 1)] * ((ar + 1) * dr - r0) * (z0 - (az + 0.5) * dz)
                        + ez[ir1 + (az + 1) * (R + 1)] * (r0 - ar * dr) * (z0 - (az + 0.5) * dz)) / (dr * dz);
        }
        /* Hphi (EM4.cu:82-122) */
        ar = ifloor(r0 / dr - 0.5); az = ifloor(z0 / dz - 0.5);
        if (ar < 0) {
            if (az < 0) Hp = h[(ar + 1) + (az + 1) * R];
            else if (az >= Z - 1) Hp = h[(ar + 1) + az * R];
            else Hp = (h[(ar + 1) + az * R] * ((az + 1.5) * dz - z0) + h[(ar + 1) + (az + 1) * R] * (z0 - (az + 0.5) * dz)) / dz;
        } else if (ar >= R - 1) {
            if (az < 0) Hp = h[ar + (az + 1) * R];
            else if (az >= Z - 1) Hp = h[ar + az * R];
            else Hp = (h[ar + az * R] * ((az + 1.5) * dz - z0) + h[ar + (az + 1) * R] * (z0 - (az + 0.5) * dz)) / dz;
        } else if (az < 0) {
            Hp = (h[ar + (az + 1) * R] * ((ar + 1.5) * dr - r0) + h[(ar + 1) + (az + 1) * R] * (r0 - (ar + 0.5) * dr)) / dr;
        } else if (az >= Z - 1) {
            Hp = (h[ar + az * R] * ((ar + 1.5) * dr - r0) + h[(ar + 1) + az * R] * (r0 - (ar + 0.5) * dr)) / dr;
        } else {
            Hp = (h[ar + az * R] * ((ar + 1.5) * dr - r0) * ((az + 1.5) * dz - z0)
                  + h[(ar + 1) + az * R] * (r0 - (ar + 0.5) * dr) * ((az + 1.5) * dz - z0)
                  + h[ar + (az + 1) * R] * ((ar + 1.5) * dr - r0) * (z0 - (az + 0.5) * dz)
                  + h[(ar + 1) + (az + 1) * R] * (r0 - (ar + 0.5) * dr) * (z0 - (az + 0.5) * dz)) / (dr * dz);
        }
        /* force, EM4.cu:126-129 */
        double Fr = q[i] * (E_r + (-vz[i] * Hp * mu));
        double Fz = q[i] * (E_z + (vr[i] * Hp * mu));
        double Fx = Fr * cos(phi[i]), Fy = Fr * sin(phi[i]);
        /* EM4.cu:134-144 */
        double g = 1 / sqrt(1 - (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) / (c * c));
        double ux = g * vx[i] + Fx / m[i] * dt;
        double uy = g * vy[i] + Fy / m[i] * dt;
        double uz = g * vz[i] + Fz / m[i] * dt;
        g = sqrt(1 + (ux * ux + uy * uy + uz * uz) / (c * c));
        if (ke) ke[i] = (g - 1) * m[i] * c * c;
        vx[i] = ux / g; vy[i] = uy / g; vz[i] = uz / g;
        /* EM4.cu:146-177 */
        x[i] += 0.5 * dt * vx[i];
        y[i] += 0.5 * dt * vy[i];
        z[i] += 0.5 * dt * vz[i];
        wedge_bc(p, &x[i], &y[i], &z[i], &vx[i], &vy[i], &vz[i], &r[i], &phi[i], &del[i]);
    }
}

/* electricmagneticfield6.m:138-156 for n_push pushes. */
void orc_em_step(const orc_params* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                 const double* q, const double* m, double* ke, int32_t* del, int64_t n,
                 double* er, double* ez, double* h, double* jr, double* jz, int n_push)
{
    double* r = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* phi = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* vr = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    size_t ncell = (size_t)p->nr * p->nz;
    for (int s = 0; s < n_push; ++s) {
        orc_em2(p, x, y, z, vx, vy, vz, r, phi, vr, del, n);
        memset(jr, 0, ncell * sizeof(double));
        memset(jz, 0, ncell * sizeof(double));
        orc_em7(p, r, z, q, vr, vz, jr, jz, del, n);
        orc_fdtd_push(p, er, ez, h, jr, jz, 1);
        orc_em4(p, x, y, z, vx, vy, vz, r, phi, vr, er, ez, h, q, m, ke, del, n);
    }
    free(r); free(phi); free(vr);
}

/* Fixed-point emulation of the HIP deposit.  The HIP kernel computes, in
 * FP32 with contraction off, the clamped area weights of EM7.cu:56-162 in
 * normalized coordinates and accumulates llrint(w * v * qnum * 2^32 / c)
 * into int64 cells; the physical current is FN*e*c*2^-32 * Q / V(i). */
void orc_deposit_fx32(const orc_params* p, const float* x, const float* y, const float* z,
                      const float* vx, const float* vy, const float* vz, int64_t n, float qnum,
                      int64_t* jr, int64_t* jz)
{
    const int nr = p->nr, nz = p->nz;
    const float inv_dr = (float)(1.0 / p->dr), inv_dz = (float)(1.0 / p->dz);
    const float kq = (float)(4294967296.0 / p->c) * qnum;
    for (int64_t m = 0; m < n; ++m) {
        float xx = x[m], yy = y[m], zz = z[m];
        if (!(zz == zz)) continue;                     /* absorbed particle */
        float r2 = xx * xx;
        r2 = r2 + yy * yy;
        float r = sqrtf(r2);
        float vr;
        if (r > 0.0f) { float t = vx[m] * xx; t = t + vy[m] * yy; vr = t / r; }
        else vr = vx[m];
        float qr = vr * kq, qz = vz[m] * kq;
        float sr = r * inv_dr; sr = sr - 0.5f;
        float sz = zz * inv_dz; sz = sz - 0.5f;
        int ar = (int)floorf(sr), az = (int)floorf(sz);
        float fr = sr - (float)ar, fz = sz - (float)az;
        int i0, i1, k0, k1; float wr0, wr1, wz0, wz1;
        if (ar < 0) { i0 = 0; i1 = 0; wr0 = 1.0f; wr1 = 0.0f; }
        else if (ar >= nr - 1) { i0 = nr - 1; i1 = nr - 1; wr0 = 1.0f; wr1 = 0.0f; }
        else { i0 = ar; i1 = ar + 1; wr0 = 1.0f - fr; wr1 = fr; }
        if (az < 0) { k0 = 0; k1 = 0; wz0 = 1.0f; wz1 = 0.0f; }
        else if (az >= nz - 1) { k0 = nz - 1; k1 = nz - 1; wz0 = 1.0f; wz1 = 0.0f; }
        else { k0 = az; k1 = az + 1; wz0 = 1.0f - fz; wz1 = fz; }
        int cells[4] = { i0 + k0 * nr, i1 + k0 * nr, i0 + k1 * nr, i1 + k1 * nr };
        float w[4] = { wr0 * wz0, wr1 * wz0, wr0 * wz1, wr1 * wz1 };
        for (int a = 0; a < 4; ++a) {
            float tr = w[a] * qr, tz = w[a] * qz;
            jr[cells[a]] += llrintf(tr);
            jz[cells[a]] += llrintf(tz);
        }
    }
}

void orc_j_from_fixed(const orc_params* p, const int64_t* qr, const int64_t* qz, double* jr, double* jz)
{
    double unit = p->fn * p->e_em * p->c / 4294967296.0;
    double vcoef = p->phip / 2 * p->dr * p->dr * p->dz;
    for (int k = 0; k < p->nz; ++k)
        for (int i = 0; i < p->nr; ++i) {
            double V = vcoef * (double)((i + 1) * (i + 1) - i * i);
            jr[i + k * p->nr] = unit * (double)qr[i + k * p->nr] / V;
            jz[i + k * p->nr] = unit * (double)qz[i + k * p->nr] / V;
        }
}

/* ------------------------------------------------------------------ */
/* Collisions                                                          */
/* ------------------------------------------------------------------ */

/* speed from kinetic energy in eV: findV.m:5-8 in the cancellation-free form
 * beta = sqrt(g1*(g1+2))/(1+g1), g1 = E*e/(m c^2). */
static double speed_from_ev(double E, double e, double m, double c)
{
    double g1 = E * e / (m * c * c);
    return c * sqrt(g1 * (g1 + 2.0)) / (1.0 + g1);
}

/* kinetic energy in eV from speed: EM4.cu:138-140 with /(e) of
 * electricmagneticfield6.m:172, written as m c^2 * beta^2 g^2/(1+g). */
static double ev_from_speed(double v2, double e, double m, double c)
{
    double b2 = v2 / (c * c);
    double g2 = 1.0 / (1.0 - b2);
    double g = sqrt(g2);
    return m * c * c * (b2 * g2 / (1.0 + g)) / e;
}

static void gas_lookup(const orc_gas* g, double r, double z, double* n, double* T)
{
    int ir = (int)floor(r / g->gdr), iz = (int)floor(z / g->gdz);
    if (ir < 0) ir = 0;
    if (ir > g->gnr - 1) ir = g->gnr - 1;
    if (iz < 0) iz = 0;
    if (iz > g->gnz - 1) iz = g->gnz - 1;
    *n = g->dens[iz * g->gnr + ir];
    *T = g->temp ? g->temp[iz * g->gnr + ir] : g->t_scalar;
}

/* electron_transport.cu:64-227 with the MATLAB choices (find_CS0.m cross
 * sections, NaN gating, get_ElasticSc_sub1.m angles, get_ImpExc2.m loss).
 * Random numbers: Philox4x32-10, key = seed, counter = (slot, step, run, block);
 * the decision uniform from block 0xFFFFFFFF of the slot's group of four. */
void orc_collide(const orc_params* p, const orc_tables* t, const orc_gas* g,
                 double* x, double* y, double* z, double* vx, double* vy, double* vz,
                 const int32_t* del, int64_t n, double dt_coll,
                 uint64_t seed, uint32_t step, uint32_t run, int elastic_mode,
                 int8_t* ev, double* sec_v, double* ion_v)
{
    const double I = 15.6;
    const double c = p->c;
    const double e_e = p->me_mc * c * c / p->e_mc;      /* electron_transport.cu:52 */
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1024) if (g_threads > 1)
    for (int64_t m = 0; m < n; ++m) {
        ev[m] = 0;
        if (del[m] == 1) continue;
        double v2 = vx[m] * vx[m] + vy[m] * vy[m] + vz[m] * vz[m];
        double ve = sqrt(v2);
        double T = ev_from_speed(v2, p->e_em, p->me_em, c);     /* eV */
        /* u1: word m % 4 of the decision block of slot group m / 4; u2..u4:
         * words 1..3 of block 0 of slot m */
        uint32_t dctr[4] = { (uint32_t)m >> 2, step, run, 0xFFFFFFFFu }, d[4];
        orc_philox4x32_10(dctr, key, d);
        uint32_t ctr[4] = { (uint32_t)m, step, run, 0 }, o[4];
        orc_philox4x32_10(ctr, key, o);
        double u1 = orc_u01(d[m & 3]), u2 = orc_u01(o[1]), u3 = orc_u01(o[2]), u4 = orc_u01(o[3]);
        if (!(T >= 15.6 && T <= 20000.0)) continue;    /* interp1 NaN gating */
        double se, si, sx;
        cs_one(t, T / 1000.0, &se, &si, &sx);
        double r = sqrt(x[m] * x[m] + y[m] * y[m]);
        double dens, Tg;
        gas_lookup(g, r, z[m], &dens, &Tg);
        double st = se + si + sx;
        double pt = ve * dt_coll * dens * st;         /* electron_transport.cu:74-77 */
        if (!(u1 < pt)) continue;                      /* :85 */
        if (u2 * st < se) {                            /* :86 elastic */
            double th = elastic_mode == 0 ? orc_elastic_theta(t, T / 1000.0, u3)
                                          : orc_elastic_theta_rej(t, T / 1000.0, (uint32_t)m, step, run, seed);
            double ph = 2 * M_PI * u4;
            orc_rotate(&vx[m], &vy[m], &vz[m], th, ph, p->rot_eps);
            ev[m] = 1;
        } else if (u2 * st < se + si && T > I) {      /* :139 ionization */
            double k0, ep, es, thp, ths;
            orc_ionize(T, u3, e_e, &k0, &ep, &es, &thp, &ths);
            double php = 2 * M_PI * u4, phs = php - M_PI;
            double vp = speed_from_ev(ep, p->e_mc, p->me_mc, c);
            double ratio = vp / ve;
            vx[m] *= ratio; vy[m] *= ratio; vz[m] *= ratio;
            orc_rotate(&vx[m], &vy[m], &vz[m], thp, php, p->rot_eps);
            double vs = speed_from_ev(es, p->e_mc, p->me_mc, c);
            sec_v[3 * m + 0] = vs * sin(ths) * cos(phs);   /* :198-200 */
            sec_v[3 * m + 1] = vs * sin(ths) * sin(phs);
            sec_v[3 * m + 2] = vs * cos(ths);
            if (ion_v) {                                    /* :202-204 */
                uint32_t ctr2[4] = { (uint32_t)m, step, run, 1 }, o2[4];
                orc_philox4x32_10(ctr2, key, o2);
                double a1 = orc_u01(o2[0]), a2 = orc_u01(o2[1]), a3 = orc_u01(o2[2]), a4 = orc_u01(o2[3]);
                double rr1 = sqrt(-2 * log(a1)), rr2 = sqrt(-2 * log(a3));
                double sg = sqrt(p->k_b * Tg / p->m_ion);
                ion_v[3 * m + 0] = sg * rr1 * cos(2 * M_PI * a2);
                ion_v[3 * m + 1] = sg * rr1 * sin(2 * M_PI * a2);
                ion_v[3 * m + 2] = sg * rr2 * cos(2 * M_PI * a4);
            }
            ev[m] = 2;
        } else {                                        /* :206 excitation */
            double edep = orc_exc_edep(t, T, p->exc_form);
            double vp = speed_from_ev(T - edep, p->e_mc, p->me_mc, c);
            double ratio = vp / ve;
            vx[m] *= ratio; vy[m] *= ratio; vz[m] *= ratio;
            ev[m] = 3;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Tallies: electricmagneticfield5.m:167-177                            */
/* ------------------------------------------------------------------ */
static double trapz2(const double* f, int ni, int nk, int ld, double r0, double dr, double dz, int rweight)
{
    /* trapz over i (r coordinate r0+i*dr) then over k (spacing dz) of f*r */
    double tot = 0.0;
    for (int k = 0; k < nk; ++k) {
        double row = 0.0;
        for (int i = 0; i + 1 < ni; ++i) {
            double ra = r0 + i * dr, rb = r0 + (i + 1) * dr;
            double fa = f[i + k * ld] * (rweight ? ra : 1), fb = f[i + 1 + k * ld] * (rweight ? rb : 1);
            row += 0.5 * (fa + fb) * dr;
        }
        double wk = (k == 0 || k == nk - 1) ? 0.5 : 1.0;
        if (nk == 1) wk = 0.0;
        tot += wk * row * dz;
    }
    return tot;
}

void orc_tallies(const orc_params* p, const double* er, const double* ez, const double* h,
                 const double* vx, const double* vy, const double* vz, const double* m,
                 const int32_t* del, int64_t n, double out[4])
{
    int R = p->nr, Z = p->nz;
    size_t ner = (size_t)R * (Z + 1), nez = (size_t)(R + 1) * Z, nh = (size_t)R * Z;
    double* s = (double*)malloc((ner > nez ? (ner > nh ? ner : nh) : (nez > nh ? nez : nh)) * sizeof(double));
    for (size_t i = 0; i < ner; ++i) s[i] = er[i] * er[i];
    double wer = trapz2(s, R, Z + 1, R, p->dr / 2, p->dr, p->dz, 1);
    for (size_t i = 0; i < nez; ++i) s[i] = ez[i] * ez[i];
    double wez = trapz2(s, R + 1, Z, R + 1, 0.0, p->dr, p->dz, 1);
    for (size_t i = 0; i < nh; ++i) s[i] = h[i] * h[i];
    double wh = trapz2(s, R, Z, R, p->dr / 2, p->dr, p->dz, 1);
    free(s);
    double ke = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (del[i] == 1) continue;
        double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        ke += ev_from_speed(v2, 1.0, m[i], p->c);     /* joules */
    }
    out[0] = p->phip * 0.5 * p->eps0 * (wer + wez);
    out[1] = p->phip * 0.5 * p->mu0 * wh;
    out[2] = p->fn * ke;
    out[3] = out[0] + out[1] + out[2];
}

/* ------------------------------------------------------------------ */
/* Transport loop: per step EM push (field6 sequence, 1 push) then a    */
/* collision step with dt_coll = ddt (electron_transport_field3.m:61-109 */
/* with one push per collision step); secondaries appended in slot order */
/* (electron_collision.m:68-75).  Absorbed particles keep their slot.   */
/* ------------------------------------------------------------------ */
int orc_run(const orc_params* p, const orc_tables* t, const orc_gas* g,
            double* x, double* y, double* z, double* vx, double* vy, double* vz, int32_t* del,
            int64_t* n_io, int64_t cap, double* er, double* ez, double* h,
            int n_steps, int collide, int secondaries, uint64_t seed, uint32_t run,
            uint32_t step0, int elastic_mode, orc_counts* counts)
{
    int64_t n = *n_io;
    size_t ncell = (size_t)p->nr * p->nz;
    double* jr = (double*)calloc(ncell, sizeof(double));
    double* jz = (double*)calloc(ncell, sizeof(double));
    double* q = (double*)malloc((size_t)cap * sizeof(double));
    double* m = (double*)malloc((size_t)cap * sizeof(double));
    int8_t* ev = (int8_t*)malloc((size_t)cap);
    double* sv = (double*)malloc((size_t)cap * 3 * sizeof(double));
    if (!jr || !jz || !q || !m || !ev || !sv) return 1;
    for (int64_t i = 0; i < cap; ++i) { q[i] = -p->e_em; m[i] = p->me_em; }
    memset(counts, 0, sizeof(*counts));
    for (int s = 0; s < n_steps; ++s) {
        orc_em_step(p, x, y, z, vx, vy, vz, q, m, NULL, del, n, er, ez, h, jr, jz, 1);
        if (!collide) continue;
        orc_collide(p, t, g, x, y, z, vx, vy, vz, del, n, p->ddt, seed, step0 + (uint32_t)s, run,
                    elastic_mode, ev, sv, NULL);
        int64_t n0 = n;
        for (int64_t i = 0; i < n0; ++i) {
            if (ev[i] == 1) counts->n_elastic++;
            else if (ev[i] == 3) counts->n_excite++;
            else if (ev[i] == 2) {
                counts->n_ionize++;
                if (!secondaries) continue;
                if (n >= cap) { counts->n_dropped++; continue; }
                x[n] = x[i]; y[n] = y[i]; z[n] = z[i];
                vx[n] = sv[3 * i]; vy[n] = sv[3 * i + 1]; vz[n] = sv[3 * i + 2];
                del[n] = 0;
                ++n;
            }
        }
    }
    *n_io = n;
    free(jr); free(jz); free(q); free(m); free(ev); free(sv);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Photoionization source: INCIDENCE1.m:27-158                           */
/* ------------------------------------------------------------------ */
#define ORC_SRC_BLOCK 0x40000000u

static void src_block(uint64_t seed, uint32_t e, uint32_t pulse, uint32_t run, uint32_t b, double u[4])
{
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t ctr[4] = { e, pulse, run, ORC_SRC_BLOCK + b }, o[4];
    orc_philox4x32_10(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[i] = orc_u01(o[i]);
}

/* findV.m:5-8 */
static double find_v(double ke)
{
    double c = 3e8, m = 9.109e-31;
    double gm = 1.0 + ke / (m * c * c);
    return c * sqrt(1.0 - 1.0 / (gm * gm));
}

/* findPEangle.m:6-14 constants for electron kinetic energy dE (J), m_e = 9.11e-31 */
static void sauter_consts(double dE, double* A, double* bg, double* g0)
{
    double erest = 9.11e-31 * 3e8 * 3e8;
    double gm = 1.0 + dE / erest;
    double beta = sqrt(dE * (dE + 2 * erest)) / (dE + erest);
    *A = 1.0 / beta - 1.0;
    *bg = 0.5 * beta * gm * (gm - 1.0) * (gm - 2.0);
    *g0 = 2.0 * (1.0 / *A + *bg);
}

/* findPEangle.m:16-33: proposals r1 -> v (inverse of v / (A + v)^3 on [0, 2]),
 * accepted when r2 g0 < g(v); the last of 128 proposals if none is */
static double sauter_v(double A, double bg, double g0, uint64_t seed, uint32_t e, uint32_t pulse, uint32_t run)
{
    double v = 0.0;
    for (int b = 0; b < 64; ++b) {
        double u[4];
        src_block(seed, e, pulse, run, 3u + (uint32_t)b, u);
        for (int h = 0; h < 2; ++h) {
            double r1 = u[2 * h], r2 = u[2 * h + 1];
            double d = (A + 2.0) * (A + 2.0) - 4.0 * r1;
            v = 2.0 * A / d * (2.0 * r1 + (A + 2.0) * sqrt(r1));
            double gv = (2.0 - v) * (1.0 / (A + v) + bg);
            if (r2 * g0 < gv) return v;
        }
    }
    return v;
}

double orc_sauter_v(double dE, uint64_t seed, uint32_t event, uint32_t pulse, uint32_t run)
{
    double A, bg, g0;
    sauter_consts(dE, &A, &bg, &g0);
    return sauter_v(A, bg, g0, seed, event, pulse, run);
}

static void thermal(uint64_t seed, uint32_t e, uint32_t pulse, uint32_t run, double n[3])
{
    double u[4];
    src_block(seed, e, pulse, run, 1u, u);
    double twopi = 6.283185307179586;
    double r1 = sqrt(-2.0 * log(u[0])), r2 = sqrt(-2.0 * log(u[2]));
    n[0] = r1 * cos(twopi * u[1]);
    n[1] = r1 * sin(twopi * u[1]);
    n[2] = r2 * cos(twopi * u[3]);
}

int orc_photoionize(const orc_params* p, const orc_gas* g, double* dens, double* temp, int deplete,
                    const orc_photo_source* s, int64_t max_events, double* pe, double* ae, double* ion,
                    double* vth, int32_t* cell, orc_photo_info* info)
{
    /* findEShell.m:4-12 */
    int shell = -1;
    for (int i = 0; i < s->n_shells; ++i)
        if (s->photon_kev >= s->binding_ev[i] / 1000.0) { shell = i; break; }
    if (shell < 0) return -1;
    double e_shell = s->binding_ev[shell] / 1000.0;
    int auger = shell != s->n_shells - 1;                       /* INCIDENCE1.m:124 */
    /* findSigmaPE.m:18-22 */
    double xs[256], ys[256];
    if (s->n_mac < 2 || s->n_mac > 256) return -1;
    for (int i = 0; i < s->n_mac; ++i) { xs[i] = s->mac[2 * i]; ys[i] = s->mac[2 * i + 1]; }
    double mac = orc_interp1(xs, ys, s->n_mac, s->photon_kev / 1000.0);
    if (!(mac >= 0)) return -1;
    double sigma = mac * s->atomic_mass / 6.02e23 * 1e-4;
    /* GUASS.m:3-15 with the integral of x exp(-x^2/(2 s^2)) in closed form */
    double sg = s->beam_radius / 3.0;
    int nr = (int)lround(s->beam_radius / g->gdr);
    if (nr > g->gnr) nr = g->gnr;
    if (nr > 4096) nr = 4096;
    if (nr < 1) nr = 1;
    double* N = (double*)malloc((size_t)nr * sizeof(double));
    double tot = 0.0;
    for (int i = 0; i < nr; ++i) {
        double a0 = i * g->gdr, a1 = (i + 1) * g->gdr;
        N[i] = sg * sg * (exp(-a0 * a0 / (2 * sg * sg)) - exp(-a1 * a1 / (2 * sg * sg)));
        tot += N[i];
    }
    for (int i = 0; i < nr; ++i) N[i] = s->n_photons * (N[i] / tot);
    double dE = (s->photon_kev - e_shell) * 1.6e-16;             /* INCIDENCE1.m:28 */
    double A, bg, g0;
    sauter_consts(dE, &A, &bg, &g0);
    double v_pe = find_v(dE), v_ae = find_v(s->auger_kev * 1.6e-16);
    double p_ph = s->photon_kev * 1.6e-16 / 3e8;                 /* :113 */
    double m_e = 9.11e-31, m_mol = 2.0 * s->atomic_mass * 1.6726e-27, kb = 1.38064853e-23;
    double twopi = 6.283185307179586, pi = 3.141592653589793;
    int ncell = nr * g->gnz;
    int32_t* cnt = (int32_t*)malloc((size_t)ncell * sizeof(int32_t));
    /* INCIDENCE1.m:51-60: Beer-Lambert down the slabs, per ring */
    for (int k = 0; k < g->gnz; ++k)
        for (int i = 0; i < nr; ++i) {
            double n = dens[k * g->gnr + i];
            double dN = n * N[i];
            dN = dN * sigma;
            dN = dN * g->gdz;
            N[i] = N[i] - dN;
            cnt[k * nr + i] = (int32_t)round(dN / p->fn);
        }
    int64_t total = 0;
    for (int c = 0; c < ncell; ++c) total += cnt[c];
    if (total > max_events) { free(N); free(cnt); return -2; }
    int64_t e = 0;
    for (int c = 0; c < ncell; ++c) {
        int ring = c % nr, slab = c / nr;
        double T = g->temp ? temp[slab * g->gnr + ring] : g->t_scalar;
        double sv = sqrt(kb * T / m_mol);
        double s2 = 0.0;
        for (int j = 0; j < cnt[c]; ++j, ++e) {
            double u[4], nn[3];
            src_block(s->seed, (uint32_t)e, s->pulse, s->run, 0u, u);
            double r = ((double)ring + u[0]) * g->gdr;                  /* :67 */
            double phi = u[1] * p->phip;                                /* :68 */
            double z = ((double)slab + u[2]) * g->gdz;                  /* :62 */
            double chi = twopi * u[3];                                  /* :102 */
            double x = r * cos(phi), y = r * sin(phi);
            double v = sauter_v(A, bg, g0, s->seed, (uint32_t)e, s->pulse, s->run);
            double ct = 1.0 - v, st = sqrt(v * (2.0 - v));
            double pvx = v_pe * st * cos(chi), pvy = v_pe * st * sin(chi), pvz = v_pe * ct;
            thermal(s->seed, (uint32_t)e, s->pulse, s->run, nn);
            double tx = sv * nn[0], ty = sv * nn[1], tz = sv * nn[2];
            s2 = s2 + ((tx * tx + ty * ty) + tz * tz);
            double mx = m_mol * tx - m_e * pvx;
            double my = m_mol * ty - m_e * pvy;
            double mz = p_ph + m_mol * tz - m_e * pvz;                  /* :126-128 */
            if (auger) {                                                /* :130-154 */
                double ua[4];
                src_block(s->seed, (uint32_t)e, s->pulse, s->run, 2u, ua);
                double th = pi * ua[0], ch = twopi * ua[1];
                double ax = v_ae * sin(th) * cos(ch), ay = v_ae * sin(th) * sin(ch), az = v_ae * cos(th);
                mx = mx - m_e * ax; my = my - m_e * ay; mz = mz - m_e * az;
                ae[e] = x; ae[max_events + e] = y; ae[2 * max_events + e] = z;
                ae[3 * max_events + e] = ax; ae[4 * max_events + e] = ay; ae[5 * max_events + e] = az;
            }
            pe[e] = x; pe[max_events + e] = y; pe[2 * max_events + e] = z;
            pe[3 * max_events + e] = pvx; pe[4 * max_events + e] = pvy; pe[5 * max_events + e] = pvz;
            ion[e] = x; ion[max_events + e] = y; ion[2 * max_events + e] = z;
            ion[3 * max_events + e] = mx / m_mol; ion[4 * max_events + e] = my / m_mol;
            ion[5 * max_events + e] = mz / m_mol;
            vth[e] = tx; vth[max_events + e] = ty; vth[2 * max_events + e] = tz;
            cell[e] = c;
        }
        if (deplete && cnt[c] > 0) {                                   /* :87-92 */
            double V = p->phip / 2.0 * ((double)(ring + 1) * (ring + 1) - (double)ring * ring) * g->gdr * g->gdr * g->gdz;
            int gi = slab * g->gnr + ring;
            double Nc = dens[gi] * V;
            double Nn = Nc - p->fn * (double)cnt[c];
            double E = 1.5 * kb * Nc * T - 0.5 * p->fn * m_mol * s2;
            if (Nn > 0.0) { dens[gi] = Nn / V; temp[gi] = E / (1.5 * kb * Nn); }
            else { dens[gi] = 0.0; temp[gi] = T; }
        }
    }
    double left = 0.0;
    for (int i = 0; i < nr; ++i) left += N[i];
    info->photons_left = left;
    info->sigma_pe = sigma;
    info->electron_kev = s->photon_kev - e_shell;
    info->auger = auger;
    info->n_rings = nr;
    info->n_events = total;
    free(N); free(cnt);
    return 0;
}


/* ------------------------------------------------------------------ */
/* Electrostatic ion step: ion_staticelectrofield_push.m:21-156 with
 * Poisson.m:25-68 on Poisson.m's cell-centre grid (R = (i + 1/2) delta,
 * Poisson.m:11-12).  Where the two files disagree (ion charge on the nodes
 * 0:delta:R, a clamp of the linear cell index) the cell that holds the ion,
 * clamped in r and z separately, is used; Poisson.m's outer loop repeats one
 * identical solve, which is done once. */
static const double ES_E = 1.6e-19, ES_KB = 1.38064853e-23, ES_EPS = 8.854187817e-12;   /* Poisson.m:2-8 */

static int es_cell(const orc_es_params* p, double x, double y, double z)
{
    double r = sqrt(x * x + y * y);
    int ir = (int)floor(r / p->delta), iz = (int)floor(z / p->delta);
    if (ir < 0) ir = 0;
    if (ir > p->nr - 1) ir = p->nr - 1;
    if (iz < 0) iz = 0;
    if (iz > p->nz - 1) iz = p->nz - 1;
    return iz * p->nr + ir;
}

static double es_trap_w(int i, int n) { return (i == 0 || i == n - 1) ? 0.5 : 1.0; }

void orc_es_poisson(const orc_es_params* p, const double* rho_i, double* phi, double* rho_e, orc_es_info* info)
{
    const int nr = p->nr, nz = p->nz, N = nr * nz;
    const double d = p->delta, d2 = d * d, den = 2 / d2 + 2 / d2;
    const double beta = ES_E / (ES_KB * p->t_e);
    double int_rho_i = 0.0;                                 /* Poisson.m:22 */
    for (int c = 0; c < N; ++c) {
        int i = c % nr, j = c / nr;
        int_rho_i += es_trap_w(i, nr) * es_trap_w(j, nz) * (rho_i[c] * ((i + 0.5) * d));
    }
    int_rho_i *= d2;
    double* nxt = (double*)malloc((size_t)N * sizeof(double));
    for (int c = 0; c < N; ++c) { phi[c] = 0.0; rho_e[c] = 0.0; }   /* :28-31 */
    double n_e0 = p->n_e0, maxd = 0.0, mean = 0.0;
    int k = 0;
    for (;;) {                                               /* :36 */
        ++k;
        double se = 0.0;
        for (int j = 0; j < nz; ++j)
            for (int i = 1; i < nr; ++i) {                   /* :41-56 (row 0 below) */
                int c = j * nr + i;
                double nv = 0.0;                             /* PHI(lenr,:), PHI(:,1), PHI(:,lenz) = 0 */
                if (i < nr - 1 && j > 0 && j < nz - 1) {
                    double pc = phi[c];
                    double re = p->boltzmann ? ES_E * n_e0 * exp(beta * pc) : 0.0;
                    rho_e[c] = re;
                    double RHO = rho_i[c] - re;
                    double R = (i + 0.5) * d;
                    double g = (RHO / ES_EPS + (phi[c + 1] + phi[c - 1]) / d2 +
                                1 / (R + d) * (phi[c + 1] - phi[c - 1]) / (2 * d) +
                                (phi[c - nr] + phi[c + nr]) / d2) / den;
                    nv = g > 0 ? g : pc;                     /* :58 PHI(g>0) = g(g>0) */
                    se += es_trap_w(i, nr) * es_trap_w(j, nz) * (re * R);
                }
                nxt[c] = nv;
            }
        for (int j = 0; j < nz; ++j)                         /* :59 PHI(1,:) = PHI(2,:); then :61-62 */
            nxt[j * nr] = (j == 0 || j == nz - 1) ? 0.0 : nxt[j * nr + 1];
        double md = 0.0, sp = 0.0;
        for (int c = 0; c < N; ++c) {
            double dd = fabs(phi[c] - nxt[c]);
            if (dd > md) md = dd;
            sp += nxt[c];
        }
        if (p->boltzmann) n_e0 = n_e0 * int_rho_i / (se * d2);   /* :57 */
        maxd = md;
        mean = sp / (double)N;
        memcpy(phi, nxt, (size_t)N * sizeof(double));
        if (!(maxd > p->tol * mean && k < p->max_iter)) break;
    }
    free(nxt);
    info->iterations = k;
    info->converged = maxd > p->tol * mean ? 0 : 1;
    info->n_e0 = n_e0;
    info->max_dphi = maxd;
    info->mean_phi = mean;
    info->int_rho_i = int_rho_i;
}

void orc_es_field(const orc_es_params* p, const double* phi, double* er, double* ez)
{
    const int nr = p->nr, nz = p->nz;
    const double d = p->delta;
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nr; ++i) {
            int c = j * nr + i;
            double gr = i == 0 ? (phi[c + 1] - phi[c]) / d
                      : i == nr - 1 ? (phi[c] - phi[c - 1]) / d : (phi[c + 1] - phi[c - 1]) / (2 * d);
            double gz = j == 0 ? (phi[c + nr] - phi[c]) / d
                      : j == nz - 1 ? (phi[c] - phi[c - nr]) / d : (phi[c + nr] - phi[c - nr]) / (2 * d);
            er[c] = -gr;
            ez[c] = -gz;
        }
}

void orc_es_push(const orc_es_params* p, int nsp, double* const* x, double* const* y, double* const* z,
                 double* const* vx, double* const* vy, double* const* vz, const int64_t* n, const double* q,
                 const double* m, double* phi, double* rho_i, double* rho_e, double* er, double* ez,
                 orc_es_info* info)
{
    const int nr = p->nr, N = p->nr * p->nz;
    const double d = p->delta, hdt = 0.5 * p->dt;
    int32_t* cnt = (int32_t*)calloc((size_t)nsp * N, sizeof(int32_t));
    for (int s = 0; s < nsp; ++s)                            /* :28-31 half drift, :48-66 cells */
        for (int64_t a = 0; a < n[s]; ++a) {
            if (z[s][a] != z[s][a]) continue;
            x[s][a] = x[s][a] + hdt * vx[s][a];
            y[s][a] = y[s][a] + hdt * vy[s][a];
            z[s][a] = z[s][a] + hdt * vz[s][a];
            cnt[(size_t)s * N + es_cell(p, x[s][a], y[s][a], z[s][a])]++;
        }
    for (int c = 0; c < N; ++c) {                            /* :68-72 */
        int i = c % nr;
        double qq = 0.0;
        for (int s = 0; s < nsp; ++s) qq = qq + p->weight * ES_E * q[s] * (double)cnt[(size_t)s * N + c];
        double V = 0.5 * p->phip * ((double)((i + 1) * (i + 1) - i * i) * d * d) * d;
        rho_i[c] = qq / V;
    }
    free(cnt);
    orc_es_poisson(p, rho_i, phi, rho_e, info);
    orc_es_field(p, phi, er, ez);
    for (int s = 0; s < nsp; ++s) {                          /* :112-147 */
        double k = q[s] * ES_E / m[s];
        for (int64_t a = 0; a < n[s]; ++a) {
            if (z[s][a] != z[s][a]) continue;
            double xh = x[s][a], yh = y[s][a], zh = z[s][a];
            int c = es_cell(p, xh, yh, zh);
            double r = sqrt(xh * xh + yh * yh);
            double ct = r > 0 ? xh / r : 1.0, st = r > 0 ? yh / r : 0.0;
            double fr = k * er[c], fz = k * ez[c];
            vx[s][a] = vx[s][a] + fr * ct * p->dt;
            vy[s][a] = vy[s][a] + fr * st * p->dt;
            vz[s][a] = vz[s][a] + fz * p->dt;
            x[s][a] = xh + hdt * vx[s][a];
            y[s][a] = yh + hdt * vy[s][a];
            z[s][a] = zh + hdt * vz[s][a];
        }
    }
}


/* ------------------------------------------------------------------ */
/* DSMC neutral-gas step (main2.m:247-264): free motion, specular walls,
 * the (r, phi, z) cell map, and the no-time-counter hard-sphere collisions
 * with Larsen-Borgnakke internal energy of collision_HS_internal.m.
 *
 * round32 = 1 rounds every stored position, velocity and internal energy to
 * FP32 (the HIP path's storage; its arithmetic is FP64 like this file's), so
 * the two agree bit for bit away from the last ulp of atan2/sin/cos/pow. */
/* ------------------------------------------------------------------ */

static double st32(double v, int round32) { return round32 ? (double)(float)v : v; }

/* cell_structure.m:6-22 (phi_divide, dphi, sum_cell) and calculate_VC.m:5-17
 * (one cell's volume per ring; every z layer has the same cells).  sum0 is
 * the exclusive prefix (sum0[0] = 0, sum0[n_rings] = cellnumxy). */
int64_t orc_cell_map_build(const orc_cell_map* m, int32_t* phi_divide, int32_t* sum0, double* dphi,
                           double* vc_ring)
{
    const int rmax = m->n_rings;
    const double dr = 1.0 * m->rp / rmax, dz = 1.0 * m->lp / m->n_layers;   /* calculate_VC.m:6-7 */
    int64_t acc = 0;
    for (int i = 1; i <= rmax; ++i) {
        const double l = (double)i / rmax * m->rp * m->phip;                /* cell_structure.m:10 */
        const int32_t pd = (int32_t)ceil(l / m->delta);                     /* :11 */
        const double dp = m->phip / pd;                                     /* :12 */
        const double a = (double)i * dr, b = (double)(i - 1) * dr;
        if (phi_divide) phi_divide[i - 1] = pd;
        if (sum0) sum0[i - 1] = (int32_t)acc;
        if (dphi) dphi[i - 1] = dp;
        if (vc_ring) vc_ring[i - 1] = 0.5 * (a * a - b * b) * dp * dz;      /* calculate_VC.m:15 */
        acc += pd;
    }
    if (sum0) sum0[rmax] = (int32_t)acc;
    return acc;
}

/* put_into_cells_symmetry.m:4-16 with r = sqrt(x^2 + y^2), phi = atan2(y, x)
 * of the stored position; 1-based ceil indices, clamped into the map (the
 * reference's index is 0 on the exact lower edges).  0-based result:
 * (iz - 1) cellnumxy + sum0[ir - 1] + iphi - 1; -1 for z = NaN (absorbed). */
int32_t orc_dsmc_cell(const orc_cell_map* m, const int32_t* pd, const int32_t* sum0, const double* dphi,
                      double x, double y, double z)
{
    if (z != z) return -1;
    const double dz = m->lp / m->n_layers, dr = m->rp / m->n_rings;
    const double r = sqrt(x * x + y * y), phi = atan2(y, x);
    double fz = ceil(z / dz), fr = ceil(r / dr);
    int iz = fz < 1.0 ? 1 : (fz > m->n_layers ? m->n_layers : (int)fz);
    int ir = fr < 1.0 ? 1 : (fr > m->n_rings ? m->n_rings : (int)fr);
    double fp = ceil(phi / dphi[ir - 1]);
    int ip = fp < 1.0 ? 1 : (fp > pd[ir - 1] ? pd[ir - 1] : (int)fp);
    return (int32_t)((int64_t)(iz - 1) * sum0[m->n_rings] + sum0[ir - 1] + ip - 1);
}

void orc_dsmc_cells(const orc_cell_map* m, int64_t n, const double* x, const double* y, const double* z,
                    int32_t* cell)
{
    int32_t* pd = (int32_t*)malloc(sizeof(int32_t) * m->n_rings);
    int32_t* s0 = (int32_t*)malloc(sizeof(int32_t) * (m->n_rings + 1));
    double* dp = (double*)malloc(sizeof(double) * m->n_rings);
    orc_cell_map_build(m, pd, s0, dp, NULL);
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t i = 0; i < n; ++i) cell[i] = orc_dsmc_cell(m, pd, s0, dp, x[i], y[i], z[i]);
    free(pd); free(s0); free(dp);
}

/* free_motion.m:3-7, then boundary_conditions_specular.m:6-64 for one
 * molecule: the r = Rp wall (specular_scattering_xy.m:5-40, once), the
 * periodic phi faces (boundary_phi_0.m / boundary_phi_phimax.m as intended:
 * position and velocity rotated by +-phimax; the reference's boundary_phi_0
 * names undefined x, y and takes one argument fewer than its caller
 * passes), then the z walls (specular_scattering_z.m:5-32).  counts:
 * [boundary events (N_boundary), r-wall, z-wall, phi wraps]. */
static void dsmc_move_one(const orc_cell_map* m, double dt, double* px, double* py, double* pz, double* pvx,
                          double* pvy, double* pvz, int64_t c[4])
{
    const double Rp = m->rp, Lp = m->lp, pm = m->phip;
    double vx = *pvx, vy = *pvy, vz = *pvz;
    double x = *px + vx * dt, y = *py + vy * dt, z = *pz + vz * dt;     /* free_motion.m:3-5 */
    double r = sqrt(x * x + y * y), phi = atan2(y, x);                   /* :6-7 */
    if (r > Rp) {                                                        /* specular_scattering_xy.m */
        c[0]++; c[1]++;
        const double x0 = x - vx * dt, y0 = y - vy * dt;
        const double r0 = sqrt(x0 * x0 + y0 * y0);
        const double t1 = ((r - Rp) / (r - r0)) * dt;                    /* :15 */
        const double xi = x - vx * t1, yi = y - vy * t1, zi = z - vz * t1;
        const double th = atan2(yi, xi);                                 /* :23 */
        const double ct = cos(th), s = sin(th);
        double vr = vx * ct + vy * s;
        const double vt = -vx * s + vy * ct;
        vr = -vr;                                                        /* :30 */
        vx = vr * ct - vt * s;
        vy = vr * s + vt * ct;
        x = xi + vx * t1; y = yi + vy * t1; z = zi + vz * t1;            /* :38-40 */
        r = sqrt(x * x + y * y);
        phi = atan2(y, x);                                               /* boundary_conditions_specular.m:12-13 */
    }
    const double cp = cos(pm), sp = sin(pm);
    while (phi < 0) {                                                    /* :18-29 */
        c[0]++; c[3]++;
        const double rxy = sqrt(x * x + y * y);
        phi = phi + pm;
        x = rxy * cos(phi); y = rxy * sin(phi);
        const double nvx = vx * cp - vy * sp, nvy = vx * sp + vy * cp;
        vx = nvx; vy = nvy;
    }
    while (phi > pm) {                                                   /* :31-41 */
        c[0]++; c[3]++;
        const double rxy = sqrt(x * x + y * y);
        phi = phi - pm;
        x = rxy * cos(phi); y = rxy * sin(phi);
        const double nvx = vx * cp + vy * sp, nvy = -vx * sp + vy * cp;
        vx = nvx; vy = nvy;
    }
    if (z > Lp) {                                                        /* :44-54, specular_scattering_z.m:15 */
        c[0]++; c[2]++;
        const double z0 = z - vz * dt;
        const double t1 = (z - Lp) / (z - z0) * dt;
        const double xi = x - vx * t1, yi = y - vy * t1, zi = z - vz * t1;
        vz = -vz;
        x = xi + vx * t1; y = yi + vy * t1; z = zi + vz * t1;
    }
    if (z < 0) {                                                         /* :56-64, specular_scattering_z.m:17 */
        c[0]++; c[2]++;
        const double z0 = z - vz * dt;
        const double t1 = (0 - z) / (z0 - z) * dt;
        const double xi = x - vx * t1, yi = y - vy * t1, zi = z - vz * t1;
        vz = -vz;
        x = xi + vx * t1; y = yi + vy * t1; z = zi + vz * t1;
    }
    *px = x; *py = y; *pz = z; *pvx = vx; *pvy = vy; *pvz = vz;
}

void orc_dsmc_move(const orc_cell_map* m, double dt, int64_t n, double* x, double* y, double* z, double* vx,
                   double* vy, double* vz, int round32, int64_t counts[4])
{
    int64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(+ : c0, c1, c2, c3) if (g_threads > 1)
    for (int64_t i = 0; i < n; ++i) {
        if (z[i] != z[i]) continue;
        int64_t c[4] = {0, 0, 0, 0};
        dsmc_move_one(m, dt, &x[i], &y[i], &z[i], &vx[i], &vy[i], &vz[i], c);
        c0 += c[0]; c1 += c[1]; c2 += c[2]; c3 += c[3];
        x[i] = st32(x[i], round32); y[i] = st32(y[i], round32); z[i] = st32(z[i], round32);
        vx[i] = st32(vx[i], round32); vy[i] = st32(vy[i], round32); vz[i] = st32(vz[i], round32);
    }
    counts[0] += c0; counts[1] += c1; counts[2] += c2; counts[3] += c3;
}

/* Philox counter blocks of an attempt: counter = (attempt, cell, step, block). */
#define DSMC_BLK0 0x50000000u      /* sub1, sub2, acceptance, scattering cosine */
#define DSMC_BLK1 0x50000001u      /* scattering azimuth */
#define DSMC_REDRAW 0x50000100u    /* + r: the r-th redraw of an equal pair (sub1, sub2) */
#define DSMC_ET 0x51000000u        /* + j: translational-energy rounds 2j, 2j+1 */
#define DSMC_ETRY 0x52000000u      /* + j: internal-energy split rounds 2j, 2j+1 */
#define DSMC_MAX_REDRAW 4096
#define DSMC_MAX_ROUNDS 65536

double orc_dsmc_u(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

static void dsmc_draw(uint64_t seed, uint32_t a, uint32_t cell, uint32_t step, uint32_t blk, uint32_t o[4])
{
    const uint32_t ctr[4] = {a, cell, step, blk};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, o);
}

/* x^e, with the exponents 0 and 1 of zeta = 2 exact */
static double dsmc_pow(double x, double e) { return e == 1.0 ? x : (e == 0.0 ? 1.0 : pow(x, e)); }

/* collision_HS_internal.m:26-135 for one cell: n_att attempts in order on
 * the cell's molecules list[0..N) (ascending slot).  c: [attempts,
 * collisions, SC > SCmax (n_SC)]. */
static void dsmc_cell(const orc_dsmc_params* p, const int64_t* list, int64_t N, double vc, uint64_t seed,
                      uint32_t cell, uint32_t step, double* vx, double* vy, double* vz, double* ei, int round32,
                      int64_t c[3])
{
    if (N <= 2) return;                                                  /* :26 */
    const double npair = floor(p->sc_max * p->fn * (double)N * (double)(N - 1) / vc * 0.5 * p->dt);   /* :27 */
    int64_t n_att = (int64_t)npair;
    if (n_att < p->min_attempts) n_att = p->min_attempts;               /* the do-while of :31-135 */
    const double d2 = p->diameter * p->diameter, m = p->mass, w = p->omega, zeta = p->zeta;
    const double k1 = (zeta + 0.5 - w) / (1.5 - w), a1 = 1.5 - w;      /* :65 */
    const double k2 = (zeta + 0.5 - w) / (zeta - 1), b1 = zeta - 1;
    const double c2 = pow(2.0, zeta - 2), h = zeta / 2 - 1;             /* :77 */
    for (int64_t a = 0; a < n_att; ++a) {
        uint32_t o[4], q[4];
        dsmc_draw(seed, (uint32_t)a, cell, step, DSMC_BLK0, o);
        int64_t s1 = (int64_t)(orc_dsmc_u(o[0]) * (double)N), s2 = (int64_t)(orc_dsmc_u(o[1]) * (double)N);
        int r = 0;
        while (s1 == s2 && r < DSMC_MAX_REDRAW) {                        /* :32-36 */
            ++r;
            dsmc_draw(seed, (uint32_t)a, cell, step, DSMC_REDRAW + (uint32_t)r, q);
            s1 = (int64_t)(orc_dsmc_u(q[0]) * (double)N);
            s2 = (int64_t)(orc_dsmc_u(q[1]) * (double)N);
        }
        c[0]++;
        if (s1 == s2) continue;
        const int64_t i1 = list[s1], i2 = list[s2];
        double ur = vx[i1] - vx[i2], vr = vy[i1] - vy[i2], wr = vz[i1] - vz[i2];   /* :38-40 */
        double cr = sqrt(ur * ur + vr * vr + wr * wr);
        const double SC = cr * M_PI * d2;                                /* :42 */
        if (SC > p->sc_max) c[2]++;                                      /* :48-51 */
        if (!(orc_dsmc_u(o[2]) < SC / p->sc_max)) continue;              /* :52-54 */
        const double um = (vx[i1] + vx[i2]) / 2, vm = (vy[i1] + vy[i2]) / 2, wm = (vz[i1] + vz[i2]) / 2;
        const double Ec = 0.25 * m * (cr * cr) + ei[i1] + ei[i2];        /* :59 */
        double Et = 0.0;
        for (int j = 0; j < DSMC_MAX_ROUNDS; ++j) {                       /* :63-70 */
            if ((j & 1) == 0) dsmc_draw(seed, (uint32_t)a, cell, step, DSMC_ET + (uint32_t)(j >> 1), q);
            Et = Ec * orc_dsmc_u(q[2 * (j & 1)]);
            const double pr = dsmc_pow(k1 * (Et / Ec), a1) * dsmc_pow(k2 * (1 - Et / Ec), b1);
            if (orc_dsmc_u(q[2 * (j & 1) + 1]) < pr) break;
        }
        const double Ei = Ec - Et;                                       /* :71 */
        double etry = 0.0;
        for (int j = 0; j < DSMC_MAX_ROUNDS; ++j) {                       /* :74-83 */
            if ((j & 1) == 0) dsmc_draw(seed, (uint32_t)a, cell, step, DSMC_ETRY + (uint32_t)(j >> 1), q);
            etry = Ei * orc_dsmc_u(q[2 * (j & 1)]);
            const double pr = c2 * dsmc_pow(etry / Ei, h) * dsmc_pow(1 - etry / Ei, h);
            if (orc_dsmc_u(q[2 * (j & 1) + 1]) < pr) break;
        }
        ei[i1] = st32(etry, round32);                                    /* :84-85 */
        ei[i2] = st32(Ei - etry, round32);
        const double cri = 2 * sqrt(Et / m);                             /* :86 */
        ur = ur * cri / cr; vr = vr * cri / cr; wr = wr * cri / cr;
        cr = cri;
        double urf, vrf, wrf;
        uint32_t o1[4];
        dsmc_draw(seed, (uint32_t)a, cell, step, DSMC_BLK1, o1);
        if (fabs(p->alpha - 1) < 1e-6) {                                 /* VHS, :91-97 */
            const double B = 2 * orc_dsmc_u(o[3]) - 1;
            const double A = sqrt(1 - B * B);
            const double C = 2 * M_PI * orc_dsmc_u(o1[0]);
            urf = cr * B;
            vrf = cr * A * cos(C);
            wrf = cr * A * sin(C);
        } else {   /* VSS, :98-113, with the w' term of Bird's rotation (+ u w sin e; :108 has -, which
                      changes |g'| and breaks energy conservation) */
            const double B = 2 * pow(orc_dsmc_u(o[3]), 1 / p->alpha) - 1;
            const double A = sqrt(1 - B * B);
            const double C = 2 * M_PI * orc_dsmc_u(o1[0]);
            const double S = sin(C), CC = cos(C);
            const double D = sqrt(vr * vr + wr * wr);
            if (D > 1e-6) {
                urf = B * ur + A * S * D;
                vrf = B * vr + A * (cr * wr * CC - ur * vr * S) / D;
                wrf = B * wr - A * (cr * vr * CC + ur * wr * S) / D;   /* Bird's sign; the reference's - loses energy */
            } else {
                urf = B * ur;
                vrf = A * CC * ur;
                wrf = A * S * ur;
            }
        }
        vx[i1] = st32(um + 0.5 * urf, round32); vx[i2] = st32(um - 0.5 * urf, round32);   /* :115-120 */
        vy[i1] = st32(vm + 0.5 * vrf, round32); vy[i2] = st32(vm - 0.5 * vrf, round32);
        vz[i1] = st32(wm + 0.5 * wrf, round32); vz[i2] = st32(wm - 0.5 * wrf, round32);
        c[1]++;
    }
}

/* collision_HS_internal.m:11-158: every cell's molecules (ascending slot,
 * `a = find(INDEX == indexnum)`) collide in their own sequence of attempts;
 * cells are independent (the reference's parfor).  cell[i] from
 * orc_dsmc_cells (-1 skipped).  counts: [attempts, collisions, n_SC]. */
void orc_dsmc_collide(const orc_cell_map* m, const orc_dsmc_params* p, int64_t n, const int32_t* cell,
                      double* vx, double* vy, double* vz, double* ei, uint64_t seed, uint32_t step, int round32,
                      int64_t counts[3])
{
    int32_t* pd = (int32_t*)malloc(sizeof(int32_t) * m->n_rings);
    int32_t* s0 = (int32_t*)malloc(sizeof(int32_t) * (m->n_rings + 1));
    double* vcr = (double*)malloc(sizeof(double) * m->n_rings);
    const int64_t nxy = orc_cell_map_build(m, pd, s0, NULL, vcr);
    const int64_t ncell = nxy * m->n_layers;
    int64_t* start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* list = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int32_t* ring = (int32_t*)malloc(sizeof(int32_t) * (size_t)nxy);
    for (int i = 0; i < m->n_rings; ++i)
        for (int k = s0[i]; k < s0[i + 1]; ++k) ring[k] = i;
    for (int64_t i = 0; i < n; ++i)
        if (cell[i] >= 0) start[cell[i] + 1]++;
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncell > 0 ? ncell : 1));
    for (int64_t c = 0; c < ncell; ++c) fill[c] = start[c];
    for (int64_t i = 0; i < n; ++i)
        if (cell[i] >= 0) list[fill[cell[i]]++] = i;
    int64_t tot[3] = {0, 0, 0};
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 4) if (g_threads > 1)
    for (int64_t c = 0; c < ncell; ++c) {
        int64_t cc[3] = {0, 0, 0};
        dsmc_cell(p, list + start[c], start[c + 1] - start[c], vcr[ring[c % nxy]], seed, (uint32_t)c, step, vx, vy,
                  vz, ei, round32, cc);
#pragma omp atomic
        tot[0] += cc[0];
#pragma omp atomic
        tot[1] += cc[1];
#pragma omp atomic
        tot[2] += cc[2];
    }
    for (int k = 0; k < 3; ++k) counts[k] += tot[k];
    free(pd); free(s0); free(vcr); free(start); free(list); free(ring); free(fill);
}

/* particle_density0.m:24-36 (FN count / VC per cell) and
 * count_T_numbensity.m:11-13 (T = m / 3 / kB mean(v.^2) of the cell's
 * molecules, v the speed) on the cell map; and the same on the (ring,
 * layer) grid of the map (rz arrays n_layers x n_rings, index iz n_rings +
 * ir: the gas grid of the electron step).  Empty cells: density 0, T 0. */
void orc_dsmc_sample(const orc_cell_map* m, int64_t n, const int32_t* cell, const double* vx, const double* vy,
                     const double* vz, double fn, double mass, double kb, double* cdens, double* ctemp,
                     double* rdens, double* rtemp)
{
    int32_t* pd = (int32_t*)malloc(sizeof(int32_t) * m->n_rings);
    int32_t* s0 = (int32_t*)malloc(sizeof(int32_t) * (m->n_rings + 1));
    double* vcr = (double*)malloc(sizeof(double) * m->n_rings);
    const int64_t nxy = orc_cell_map_build(m, pd, s0, NULL, vcr);
    const int64_t ncell = nxy * m->n_layers;
    int64_t* cnt = (int64_t*)calloc((size_t)ncell, sizeof(int64_t));
    double* s2 = (double*)calloc((size_t)ncell, sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        if (cell[i] < 0) continue;
        const double v = sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        cnt[cell[i]]++;
        s2[cell[i]] += v * v;
    }
    for (int iz = 0; iz < m->n_layers; ++iz)
        for (int ir = 0; ir < m->n_rings; ++ir) {
            int64_t rc = 0;
            double rs = 0.0;
            for (int k = s0[ir]; k < s0[ir + 1]; ++k) {
                const int64_t c = (int64_t)iz * nxy + k;
                if (cdens) cdens[c] = fn * (double)cnt[c] / vcr[ir];
                if (ctemp) ctemp[c] = cnt[c] ? mass / 3 / kb * (s2[c] / (double)cnt[c]) : 0.0;
                rc += cnt[c];
                rs += s2[c];
            }
            const int64_t g = (int64_t)iz * m->n_rings + ir;
            if (rdens) rdens[g] = fn * (double)rc / (vcr[ir] * pd[ir]);
            if (rtemp) rtemp[g] = rc ? mass / 3 / kb * (rs / (double)rc) : 0.0;
        }
    free(pd); free(s0); free(vcr); free(cnt); free(s2);
}
