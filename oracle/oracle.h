/*
 * oracle.h -- FP64 CPU restatement of the reference electron PIC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path (libgasx.so): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  Nothing in the product links,
 * imports or falls back to it.
 *
 * Every function cites the reference file:line it restates.  Reference
 * files are MATLAB (.m) and CUDA (.cu) sources of ShanmeiHE/GAS-ATENUATOR.
 *
 * Array conventions (reference semantics, 0-based, r fastest):
 *   Hphi  nr x nz        index i + k*nr        (EM3_1.cu:12)
 *   Er    nr x (nz+1)    index i + k*nr        (EM3.cu:18)
 *   Ez    (nr+1) x nz    index i + k*(nr+1)    (EM3.cu:19)
 *   Jr,Jz nr x nz        on the Hphi grid      (EM7.cu:56-57)
 */
#ifndef GASX_ORACLE_H
#define GASX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NANG 181
#define ORC_MAXT 64

typedef struct {
    int32_t nr, nz;
    double dr, dz;
    double phip;        /* wedge angle PHIp */
    double ddt;         /* push step (electricmagneticfield6.m:135, ddt = ratio*dt) */
    int32_t n_sub;      /* field substeps per push (electricmagneticfield6.m:147, 2*ratio) */
    double fn;          /* macro weight FN */
    double c, eps0, mu0;      /* electricmagneticfield6.m:8-10 */
    double e_em, me_em;       /* electricmagneticfield6.m:5,11 */
    double e_mc, me_mc;       /* electron_transport.cu:47,49 */
    double m_ion, k_b;        /* electron_transport.cu:50-51 */
    int32_t ez_metric;        /* 0: 1/(2r) with r = i*dr; 1: r=(i+1)*dr as EM3.cu:39 */
    int32_t axis_j_sign;      /* +1 as EM3.cu:35, -1 alternative */
    int32_t exc_form;         /* 0: (1-(W/T)^g)^nu (find_CS0.m:52); 1: ((1-W/T)^g)^nu (electron_transport.cu:215) */
    double rot_eps;           /* rotate.m:5 threshold on vx^2+vy^2 */
} orc_params;

typedef struct {
    int32_t n_el;  double el_e[ORC_MAXT], el_s[ORC_MAXT];    /* keV ; a0^2 per atom */
    int32_t n_sel; double sel_e[ORC_MAXT], sel_s[ORC_MAXT];  /* eV ; 1e-20 m^2 */
    int32_t n_ion; double ion_e[ORC_MAXT], ion_s[ORC_MAXT];  /* eV ; 1e-20 m^2 */
    int32_t n_exc; double exc_w[16], exc_a[16], exc_om[16], exc_nu[16], exc_ga[16];
    int32_t n_dcs;                       /* merged DCS energies (36) */
    double dcs_e[ORC_MAXT];              /* keV */
    double dcs_loge[ORC_MAXT];           /* log(E keV) = xInterp */
    double dcs_tcs[ORC_MAXT];            /* yInterp, a0^2 per atom */
    double dcs_logv[ORC_MAXT * ORC_NANG];/* V_interp, [e*181 + angle] */
} orc_tables;

/* ---- threads of the loops (OpenMP; 1 by default) ---- */
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- tables (electron_collision.m:12-41, find_CS0.m:14-56) ---- */
int orc_tables_build(orc_tables* t,
                     const double* t_el, int n_el,      /* 2 x n_el row-major */
                     const double* t_sel, int n_sel,    /* n_sel x 2 */
                     const double* t_sel1, int n_sel1,  /* n_sel1 x 2 */
                     const double* t_del, int w_del,    /* 182 x w_del */
                     const double* t_sdel, int w_sdel,  /* 182 x w_sdel */
                     const double* t_ion, int n_ion,    /* n_ion x 2 */
                     const double* t_exc, int n_exc);   /* n_exc x 5 */
double orc_interp1(const double* x, const double* y, int n, double xq);
void orc_cross_sections(const orc_tables* t, const double* e_kev, int64_t n,
                        double* sig_el, double* sig_ion, double* sig_exc);
double orc_elastic_theta(const orc_tables* t, double e_kev, double u);
double orc_elastic_theta_rej(const orc_tables* t, double e_kev, uint32_t slot, uint32_t step,
                             uint32_t run, uint64_t seed);
void orc_ionize(double t_ev, double rn, double e_e,
                double* k0, double* ep, double* es, double* th_p, double* th_s);
double orc_exc_edep(const orc_tables* t, double t_ev, int form);
void orc_rotate(double* vx, double* vy, double* vz, double th, double ph, double eps);

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u01(uint32_t x);

/* ---- FDTD (EM3_1.cu:2-17, EM3.cu:2-43, electricmagneticfield6.m:146-151) ---- */
void orc_fdtd_h(double* h, const double* er, const double* ez, int nr, int nz,
                double mu, double dr, double dz, double tau);
void orc_fdtd_e(double* er, double* ez, const double* h, const double* jr, const double* jz,
                int nr, int nz, double eps, double dr, double dz, double tau,
                int ez_metric, int axis_j_sign);
void orc_fdtd_push(const orc_params* p, double* er, double* ez, double* h,
                   const double* jr, const double* jz, int n_push);

/* ---- particles (EM2.cu, EM4.cu, EM7.cu) ---- */
void orc_em2(const orc_params* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
             double* r, double* phi, double* vr, int32_t* del, int64_t n);
void orc_em7(const orc_params* p, const double* r, const double* z, const double* q,
             const double* vr, const double* vz, double* jr, double* jz,
             const int32_t* del, int64_t n);
void orc_em4(const orc_params* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
             double* r, double* phi, const double* vr, const double* er, const double* ez,
             const double* h, const double* q, const double* m, double* ke, int32_t* del, int64_t n);
/* one field6 push sequence (electricmagneticfield6.m:138-156) on a mixed-species set */
void orc_em_step(const orc_params* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                 const double* q, const double* m, double* ke, int32_t* del, int64_t n,
                 double* er, double* ez, double* h, double* jr, double* jz, int n_push);

/* fixed-point emulation of the HIP deposit (bit-exact target) */
void orc_deposit_fx32(const orc_params* p, const float* x, const float* y, const float* z,
                      const float* vx, const float* vy, const float* vz, int64_t n, float qnum,
                      int64_t* jr, int64_t* jz);
void orc_j_from_fixed(const orc_params* p, const int64_t* qr, const int64_t* qz, double* jr, double* jz);

/* ---- collisions (electron_transport.cu:64-227 with find_CS0.m / MATLAB choices) ---- */
typedef struct {
    const double* dens;  /* gas density n[iz*gnr + ir] (m^-3) */
    const double* temp;  /* gas temperature (K), may be NULL -> t_scalar */
    int32_t gnr, gnz; double gdr, gdz; double t_scalar;
} orc_gas;

/* event codes written to ev[]: 0 none, 1 elastic, 2 ionization, 3 excitation */
void orc_collide(const orc_params* p, const orc_tables* t, const orc_gas* g,
                 double* x, double* y, double* z, double* vx, double* vy, double* vz,
                 const int32_t* del, int64_t n, double dt_coll,
                 uint64_t seed, uint32_t step, uint32_t run,
                 int elastic_mode /* 0: inverse CDF (reference), 1: rejection (HIP) */,
                 int8_t* ev, double* sec_v /* 3*n: secondary velocity */,
                 double* ion_v /* 3*n or NULL */);

/* ---- photoionization source (INCIDENCE1.m:27-158, GUASS.m:3-15, findEShell.m:1-12,
 *      findSigmaPE.m:8-22, findPEangle.m:10-33, findV.m:5-8) ---- */
typedef struct {
    double photon_kev, n_photons, beam_radius;
    const double* binding_ev; int32_t n_shells;   /* eV, file order */
    const double* mac; int32_t n_mac;             /* n_mac x 2: MeV ; cm^2/g */
    double atomic_mass, auger_kev;
    uint64_t seed; uint32_t pulse, run;
} orc_photo_source;

typedef struct {
    double photons_left, sigma_pe, electron_kev;
    int32_t auger, n_rings;
    int64_t n_events;
} orc_photo_info;

/* Events in emission order (slab by slab, rings outward).  pe, ae, ion: 6 x
 * max_events (x, y, z, vx, vy, vz rows of length max_events); ae unused
 * without an inner shell; cell[e] = slab * n_rings + ring; vth: 3 x max_events
 * thermal velocities of the ions (before the recoil).  With deplete, dens and
 * temp (gnz x gnr) are updated in place (INCIDENCE1.m:87-92).  Returns 0, or
 * -1 (bad input) / -2 (more events than max_events). */
int orc_photoionize(const orc_params* p, const orc_gas* g, double* dens, double* temp, int deplete,
                    const orc_photo_source* s, int64_t max_events, double* pe, double* ae, double* ion,
                    double* vth, int32_t* cell, orc_photo_info* info);
/* Sauter sample v = 1 - cos(theta) of a photoelectron of kinetic energy dE (J)
 * from the proposals (r1, r2) pairs of the source's blocks (findPEangle.m:16-33) */
double orc_sauter_v(double dE_joule, uint64_t seed, uint32_t event, uint32_t pulse, uint32_t run);

/* ---- electrostatic ion step (ion_staticelectrofield_push.m:21-156, Poisson.m:25-68) ---- */
typedef struct {
    int32_t nr, nz;
    double delta, phip, weight, t_e, dt, n_e0, tol;
    int32_t max_iter;
    int32_t boltzmann;      /* 1: Poisson.m's Boltzmann electrons; 0: none (rho_e = 0, n_e0 kept) */
} orc_es_params;
typedef struct {
    int32_t iterations, converged;
    double n_e0, max_dphi, mean_phi, int_rho_i;
} orc_es_info;
/* Poisson.m's iteration for the charge density rho_i (nr*nz, index j*nr + i):
 * phi and rho_e (outputs, nr*nz) */
void orc_es_poisson(const orc_es_params* p, const double* rho_i, double* phi, double* rho_e, orc_es_info* info);
/* E = -gradient(PHI) with MATLAB's rule (Poisson.m:69-71) */
void orc_es_field(const orc_es_params* p, const double* phi, double* er, double* ez);
/* The whole step for nsp species (x[s] .. vz[s] arrays of n[s] ions, charge
 * number q[s], mass m[s]; z = NaN marks an absorbed ion), in place. */
void orc_es_push(const orc_es_params* p, int nsp, double* const* x, double* const* y, double* const* z,
                 double* const* vx, double* const* vy, double* const* vz, const int64_t* n, const double* q,
                 const double* m, double* phi, double* rho_i, double* rho_e, double* er, double* ez,
                 orc_es_info* info);

/* ---- DSMC neutral-gas step (main2.m:247-264, collision_HS_internal.m) ---- */
typedef struct {
    int32_t n_rings, n_layers;   /* rmax (r_divide), zmax (z_divide) */
    double rp, lp, phip;         /* Rp, Lp, PHIp */
    double delta;                /* target arc length of cell_structure.m */
} orc_cell_map;
typedef struct {
    double dt, fn, mass, diameter, sc_max, omega, zeta, alpha;
    int32_t min_attempts;        /* 1: the reference's do-while (one attempt when Npair = 0) */
} orc_dsmc_params;
/* cell_structure.m + calculate_VC.m: returns cellnumxy; any output may be NULL
 * (phi_divide, dphi, vc_ring: n_rings; sum0: n_rings + 1, exclusive prefix) */
int64_t orc_cell_map_build(const orc_cell_map* m, int32_t* phi_divide, int32_t* sum0, double* dphi,
                           double* vc_ring);
int32_t orc_dsmc_cell(const orc_cell_map* m, const int32_t* pd, const int32_t* sum0, const double* dphi,
                      double x, double y, double z);
void orc_dsmc_cells(const orc_cell_map* m, int64_t n, const double* x, const double* y, const double* z,
                    int32_t* cell);
/* counts += [boundary events, r wall, z walls, phi wraps] */
void orc_dsmc_move(const orc_cell_map* m, double dt, int64_t n, double* x, double* y, double* z, double* vx,
                   double* vy, double* vz, int round32, int64_t counts[4]);
double orc_dsmc_u(uint32_t w);
/* counts += [attempts, collisions, SC > SCmax] */
void orc_dsmc_collide(const orc_cell_map* m, const orc_dsmc_params* p, int64_t n, const int32_t* cell,
                      double* vx, double* vy, double* vz, double* ei, uint64_t seed, uint32_t step, int round32,
                      int64_t counts[3]);
void orc_dsmc_sample(const orc_cell_map* m, int64_t n, const int32_t* cell, const double* vx, const double* vy,
                     const double* vz, double fn, double mass, double kb, double* cdens, double* ctemp,
                     double* rdens, double* rtemp);

/* ---- tallies (electricmagneticfield5.m:167-177) ---- */
void orc_tallies(const orc_params* p, const double* er, const double* ez, const double* h,
                 const double* vx, const double* vy, const double* vz, const double* m,
                 const int32_t* del, int64_t n, double out[4]);

/* ---- the full transport loop used as the CPU baseline ---- */
typedef struct {
    int64_t n_elastic, n_ionize, n_excite, n_dropped;
} orc_counts;
/* arrays have capacity cap; *n is updated with appended secondaries */
int orc_run(const orc_params* p, const orc_tables* t, const orc_gas* g,
            double* x, double* y, double* z, double* vx, double* vy, double* vz, int32_t* del,
            int64_t* n, int64_t cap, double* er, double* ez, double* h,
            int n_steps, int collide, int secondaries, uint64_t seed, uint32_t run,
            uint32_t step0, int elastic_mode, orc_counts* counts);

#ifdef __cplusplus
}
#endif
#endif
