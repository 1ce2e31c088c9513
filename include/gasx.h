/*
 * gasx.h -- C ABI of the MI355X electron PIC step (libgasx.so).
 *
 * Drop-in boundary for the reference's MATLAB -> CUDA launch surface
 * (parallel.gpu.CUDAKernel + feval, electricmagneticfield6.m:77-156 and
 * electron_collision.m:4-58).  Each entry point below names the reference
 * kernel or function it replaces.  All calls are extern "C", return a
 * gasx_status, never throw, and take caller-owned device buffers with
 * explicit sizes.  A gasx_ctx is bound to one device and one HIP stream; it
 * is thread-compatible, not thread-safe.  Calls are asynchronous on the
 * context's stream unless documented otherwise.
 *
 * Grid conventions (reference semantics, 0-based, r fastest):
 *   Hphi  nr x nz       index i + k*nr        (EM3_1.cu:12)
 *   Er    nr x (nz+1)   index i + k*nr        (EM3.cu:18)
 *   Ez    (nr+1) x nz   index i + k*(nr+1)    (EM3.cu:19)
 *   Jr,Jz nr x nz       on the Hphi grid      (EM7.cu:56-57)
 * Rp = nr*dr, Lp = nz*dz; particles live in the wedge 0 <= phi <= PHIp.
 *
 * Particles are FP32 structure-of-arrays (x, y, z, vx, vy, vz) in the wedge's
 * Cartesian frame.  A slot whose z is NaN holds an absorbed particle (the
 * reference's parDelete = 1, EM2.cu:31-33); gasx_sort removes them.  The
 * other five values of an absorbed slot are unspecified: the push kernels
 * run on every slot below the count without a branch and may overwrite them.
 * Field state is FP64 (the reference's double arrays); the particle push
 * gathers from FP32 copies the library refreshes after every field cycle.
 */
#ifndef GASX_H
#define GASX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GASX_ABI_VERSION 8

typedef enum gasx_status {
    GASX_OK = 0,
    GASX_E_ARG = 1,        /* bad argument, shape or size */
    GASX_E_CAPACITY = 2,   /* appends exceeded capacity; the excess was dropped and counted */
    GASX_E_HIP = 3,        /* HIP runtime error (message in gasx_last_error) */
    GASX_E_NOMEM = 4,      /* device allocation failed */
    GASX_E_STATE = 5,      /* call sequence error */
    GASX_E_NUMERIC = 6     /* non-finite state or fixed-point range exceeded */
} gasx_status;

typedef struct gasx_ctx gasx_ctx;

/* Physical and numerical parameters.  Constants are kept per reference file
 * because the reference spells them differently in different files. */
typedef struct gasx_params {
    int32_t nr, nz;       /* Hphi cells in r and z */
    double dr, dz;        /* cell size (m) */
    double phip;          /* wedge angle PHIp (rad), main3.m:9 */
    double ddt;           /* push step (s): electricmagneticfield6.m:135 ddt = ratio*dt */
    int32_t n_sub;        /* field substeps per push: electricmagneticfield6.m:147 (2*ratio) */
    double fn;            /* macro-particle weight FN */
    double c;             /* 3e8 */
    double eps0, mu0;     /* electricmagneticfield6.m:9-10 */
    double e_em, me_em;   /* 1.6e-19, 9.1096e-31: electricmagneticfield6.m:5,11 */
    double e_mc, me_mc;   /* 1.6022e-19, 9.109e-31: electron_transport.cu:47,49 */
    double m_ion, k_b;    /* 4.65e-26, 1.38064853e-23: electron_transport.cu:50-51 */
    int32_t ez_metric;    /* 0: 1/(2r) at r = i*dr; 1: r = (i+1)*dr as EM3.cu:39 */
    int32_t axis_j_sign;  /* +1 as EM3.cu:35; -1 alternative */
    int32_t exc_form;     /* 0: (1-(W/T)^g)^nu, find_CS0.m:52; 1: ((1-W/T)^g)^nu, electron_transport.cu:215 */
    double rot_eps;       /* rotate.m:5 degeneracy threshold on vx^2+vy^2 (1e-5) */
} gasx_params;

/* Raw N2 tables exactly as the reference's CSV files hold them (row-major;
 * MATLAB callers pass the transpose of their column-major matrices).  The
 * library performs the merges of electron_collision.m:12-41. */
typedef struct gasx_raw_tables {
    const double* t_el;   int32_t n_el;    /* 2 x n_el   TotalElasticCS_Nitrogen.csv (keV ; a0^2) */
    const double* t_sel;  int32_t n_sel;   /* n_sel x 2  TotalElasticCS_smalleV_Nitrogen.csv (eV ; 1e-20 m^2) */
    const double* t_sel1; int32_t n_sel1;  /* n_sel1 x 2 TotalElasticCS_smalleV_Nitrogen_1.csv */
    const double* t_del;  int32_t w_del;   /* 182 x w_del ElasticCS_Nitrogen.csv */
    const double* t_sdel; int32_t w_sdel;  /* 182 x w_sdel ElasticDCS_smalleV_Nitrogen.csv */
    const double* t_ion;  int32_t n_ion;   /* n_ion x 2  TotalIonCS_Nitrogen.csv (eV ; 1e-20 m^2) */
    const double* t_exc;  int32_t n_exc;   /* n_exc x 5  ParameterExcCS_Nitrogen.csv (j,A,omega,v,miu) */
} gasx_raw_tables;

/* One particle population (caller-owned device memory). */
typedef struct gasx_species {
    float *x, *y, *z, *vx, *vy, *vz;
    int32_t* count;       /* device int32: slots in use (live + absorbed) */
    int64_t capacity;     /* slots allocated */
    float qnum;           /* charge in units of e_em: -1 electron, +1 N2+, +2 N2++ (electricmagneticfield6.m:68) */
    double mass;          /* kg (electricmagneticfield6.m:69) */
} gasx_species;

/* Field state (caller-owned device memory, FP64, reference layout). */
typedef struct gasx_fields {
    double *er, *ez, *hphi;
} gasx_fields;

/* Neutral gas seen by the collision step: density (and temperature) on an
 * r-z grid; electron_transport.cu:67-76 looks them up per electron. */
typedef struct gasx_gas {
    const float* density;      /* device, gnr*gnz, index iz*gnr + ir (m^-3) */
    const float* temperature;  /* device, same layout, or NULL to use t_scalar (K) */
    int32_t gnr, gnz;
    double gdr, gdz;
    double t_scalar;
    double n_max;              /* >= max density; the rejection bound uses it (exactness needs an upper bound) */
} gasx_gas;

typedef struct gasx_counters {
    int64_t n_elastic, n_ionize, n_excite;
    int64_t n_dropped_sec;     /* secondaries dropped at capacity (GASX_E_CAPACITY) */
    int64_t n_dropped_ion;
    int64_t n_bound_violations;/* candidate with p_T above the rejection bound (must stay 0) */
    int64_t n_absorbed;        /* particles absorbed at r > Rp */
    int64_t n_saturated;       /* collisions with p_T > 1 (linear p_T of electron_transport.cu:74-85 saturates) */
    int64_t n_pushed;          /* particle pushes performed (particle-steps) */
} gasx_counters;

/* Kernel timing of the last gasx_run with time_kernels = 1 (HIP events on
 * the context's stream around every launch).  The counts are those of the
 * timed launches (every time_every-th launch of a kind). */
typedef struct gasx_timing {
    double particle_ms;        /* sum over particle launches (push + collide of a step, or fused passes) */
    double fdtd_ms;            /* sum over field-kernel launches */
    double append_ms;          /* sum over secondary/ion append launches */
    double sort_ms;            /* sum over cell sorts */
    int64_t n_particle_launches, n_fdtd_launches, n_append_launches, n_sorts;
    double graph_ms;           /* sum over graph launches (gasx_batch_run with use_graph) */
    int64_t n_graph_launches;
    double drift_ms;           /* sum over drift + deposit launches of the collision loop */
    int64_t n_drift_launches;
} gasx_timing;

/* Options of the fused production loop. */
typedef struct gasx_run_opts {
    int32_t n_steps;
    int32_t collide;            /* 1: one collision check per push with dt_coll = ddt */
    int32_t secondaries;        /* 1: append ionization secondaries (electron_collision.m:68-75) */
    int32_t sort_every;         /* cell sort every k steps (0 = never) */
    uint64_t seed;              /* Philox key */
    uint32_t run_id;            /* Philox counter word 2 */
    uint32_t step0;             /* Philox counter word 1 of the first step */
    int32_t use_graph;          /* 1: capture the step sequence in a hipGraph and replay */
    int32_t time_kernels;       /* 1: record HIP events around every launch (gasx_get_timing);
                                   2: the push + collision launches only (the least
                                   overhead between launches) */
    int32_t time_every;         /* with time_kernels: time launches 0, k, 2k, ... of each kind
                                   (k = time_every; 0 or 1: every launch) */
} gasx_run_opts;

/* ---- context ------------------------------------------------------------ */
/* Replaces the per-call CUDAKernel construction of electricmagneticfield6.m:77-111
 * and the per-call table reads of electron_transport_field3.m:8-37. */
gasx_status gasx_ctx_create(int device, const gasx_params* p, const gasx_raw_tables* t, gasx_ctx** out);
gasx_status gasx_ctx_destroy(gasx_ctx* ctx);
gasx_status gasx_ctx_set_stream(gasx_ctx* ctx, void* hip_stream);
gasx_status gasx_synchronize(gasx_ctx* ctx);
const char* gasx_status_string(gasx_status s);
const char* gasx_last_error(const gasx_ctx* ctx);
int gasx_abi_version(void);

/* ---- FDTD ------------------------------------------------------------------
 * Replaces EM3_1.cu / EM3.cu as sequenced by electricmagneticfield6.m:146-151:
 * n_push field cycles (H back half a substep, n_sub leapfrog substeps, H
 * forward half a substep) with the current jr, jz (A/m^2, FP64 device, nr*nz)
 * frozen over each cycle.  Also refreshes the FP32 gather copies. */
gasx_status gasx_fdtd(gasx_ctx* ctx, gasx_fields* f, const double* jr, const double* jz, int n_push);

/* ---- particle kernels -----------------------------------------------------
 * EM2.cu: half drift with the wedge / z boundary conditions. */
gasx_status gasx_half_drift(gasx_ctx* ctx, gasx_species* s);
/* EM7.cu: deposit into int64 fixed-point accumulators qr, qz (device, nr*nz;
 * the caller zeroes them).  Physical current = FN*e_em*c*2^-32 * q / V(i),
 * V(i) = PHIp/2*((i+1)^2-i^2)*dr^2*dz (EM7.cu:174-176).  The integer sum is
 * independent of particle order, so the result is bit-reproducible. */
gasx_status gasx_deposit(gasx_ctx* ctx, const gasx_species* s, int64_t* qr, int64_t* qz);
/* Converts fixed-point accumulators to A/m^2 (FP64). */
gasx_status gasx_current_from_fixed(gasx_ctx* ctx, const int64_t* qr, const int64_t* qz, double* jr, double* jz);
/* EM4.cu: gather E, H at the particle, relativistic push, half drift + BC.
 * Uses the FP32 field copies refreshed by the last gasx_fdtd / gasx_em_step. */
gasx_status gasx_push(gasx_ctx* ctx, gasx_species* s, const gasx_fields* f);
/* Refreshes the FP32 gather copies from f (needed only after the caller wrote f). */
gasx_status gasx_fields_sync(gasx_ctx* ctx, const gasx_fields* f);
/* Replaces electricmagneticfield6.m:138-166: n_push pushes of EM2 -> EM7 ->
 * field cycle -> EM4 for up to 3 species sharing one current. */
gasx_status gasx_em_step(gasx_ctx* ctx, gasx_species* sp, int n_species, gasx_fields* f, int n_push);

/* ---- collisions -------------------------------------------------------------
 * Replaces find_CS0.m:1-61 (sigma in m^2; NaN outside the tables as interp1). */
gasx_status gasx_cross_sections(gasx_ctx* ctx, const float* e_kev, int64_t n,
                                float* sig_el, float* sig_ion, float* sig_exc);
/* Replaces electron_transport.cu + electron_collision.m:58-75: one collision
 * check of duration dt_coll for every live electron; ionization secondaries
 * are appended to e (slot order of their parents) and, when ions != NULL,
 * N2+ ions with Maxwellian velocities are appended to ions.  Random numbers
 * come from Philox4x32-10 with key = seed and counter = (slot, step, run, block). */
gasx_status gasx_collide(gasx_ctx* ctx, gasx_species* e, gasx_species* ions, const gasx_gas* g,
                         double dt_coll, uint64_t seed, uint32_t step, uint32_t run);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gasx_collide with the same arguments and results, whose
 * staging of secondaries and ions and whose append scan live in a fixed
 * device array of the library (one per GPU, 48 MiB, held until the process
 * ends) instead of buffers the context allocates on first use; the call
 * allocates nothing.  e->capacity must not exceed GASX_NOALLOC_MAX_SLOTS
 * (GASX_E_ARG otherwise).  All calls on one GPU share that array: calls
 * that overlap on one GPU (two streams, two contexts) race. */
#define GASX_NOALLOC_MAX_SLOTS (1 << 20)
gasx_status gasx_collide_noalloc(gasx_ctx* ctx, gasx_species* e, gasx_species* ions, const gasx_gas* g,
                                 double dt_coll, uint64_t seed, uint32_t step, uint32_t run);

/* ---- production loop --------------------------------------------------------
 * electron_transport_field3.m:61-116 with one push per collision step, fused:
 * per step one particle kernel (gather, push, collide, append, next half drift,
 * deposit) and one field kernel.  Secondaries are appended in parent slot order.
 * ions may be NULL. */
gasx_status gasx_run(gasx_ctx* ctx, gasx_species* e, gasx_species* ions, gasx_fields* f,
                     const gasx_gas* g, const gasx_run_opts* o);

/* ---- batched ensemble -------------------------------------------------------
 * Many independent runs of the production loop on the context's grid, advanced
 * together: one particle launch and one field launch per step cover every run.
 * The reference has no batched form; each run is exactly what gasx_run with
 * the same seed, run id (opts.run_id + r), gas and ions would produce
 * (sort_every must be 0).  Particle slots of all runs live in one pooled SoA:
 * run r owns [base[r], base[r] + capacity[r]), base[r] a multiple of
 * gasx_batch_align().  With an ion pool (ix non-NULL), the N2+ ions of
 * ionizations are appended to it as gasx_run appends them to its ions
 * (electron_collision.m:63-66): run r owns [ion_base[r], ion_base[r] +
 * ion_capacity[r]) of the ion pool, ion_base[r] a multiple of
 * gasx_batch_align(). */
typedef struct gasx_batch {
    int32_t nruns;                       /* 1 .. 1024 */
    float *x, *y, *z, *vx, *vy, *vz;     /* pooled SoA, device */
    const int64_t* base;                 /* host, nruns slot offsets */
    const int64_t* capacity;             /* host, nruns */
    int32_t* counts;                     /* device int32[nruns]: slots in use per run */
    double *er, *ez, *hphi;              /* device: nruns field sets back to back, reference layout each */
    const float* density;                /* device: nruns * gnr * gnz, index r*gnr*gnz + iz*gnr + ir (m^-3) */
    const double* n_max;                 /* host, nruns: upper bound of each run's density */
    int32_t gnr, gnz;
    double gdr, gdz;
    double t_scalar;                     /* gas temperature (K) */
    float *ix, *iy, *iz, *ivx, *ivy, *ivz;   /* pooled ion SoA, device, or ix = NULL: no ions */
    const int64_t* ion_base;             /* host, nruns */
    const int64_t* ion_capacity;         /* host, nruns */
    int32_t* ion_counts;                 /* device int32[nruns]: ion slots in use per run */
} gasx_batch;

/* Launch shapes of the batched loop (picked by gasx_batch_autotune). */
typedef struct gasx_batch_tuning {
    int32_t fdtd_ti, fdtd_tk;            /* field tile in cells (r x z) */
    int32_t particle_blocks_per_cu;      /* persistent particle workgroups per CU */
} gasx_batch_tuning;

int gasx_batch_align(void);
/* Advances every run by o->n_steps steps (collide / secondaries / seed /
 * run_id / step0 / use_graph as gasx_run; use_graph = 1 captures all steps in
 * one hipGraph, instantiated once per batch shape and replayed). */
gasx_status gasx_batch_run(gasx_ctx* ctx, gasx_batch* b, const gasx_run_opts* o);
/* Times each candidate launch shape for trial_steps steps on a snapshot of the
 * batch (restored afterwards), keeps the fastest in the context and returns it. */
gasx_status gasx_batch_autotune(gasx_ctx* ctx, gasx_batch* b, const gasx_run_opts* o, int trial_steps,
                                gasx_batch_tuning* out);
gasx_status gasx_batch_set_tuning(gasx_ctx* ctx, const gasx_batch_tuning* t);
gasx_status gasx_batch_get_tuning(gasx_ctx* ctx, gasx_batch_tuning* t);
/* Per-run event counters (nruns entries; synchronizes). */
gasx_status gasx_batch_counters(gasx_ctx* ctx, int nruns, gasx_counters* out, int reset);

/* Cell sort by (iz, ir) of the particle's Hphi cell; absorbed slots are dropped
 * and count is reduced (the reference's logical-index removal,
 * electricmagneticfield6.m:176-178). */
gasx_status gasx_sort(gasx_ctx* ctx, gasx_species* s);

/* FP64 energy tallies of electricmagneticfield5.m:167-177:
 * out = {W_E, W_H, KE (= FN * sum of kinetic energies, J), total}. */
gasx_status gasx_tallies(gasx_ctx* ctx, const gasx_fields* f, const gasx_species* sp, int n_species,
                         double out[4]);
/* Event counters accumulated by gasx_collide / gasx_run (synchronizes). */
gasx_status gasx_get_counters(gasx_ctx* ctx, gasx_counters* out, int reset);
/* Host copy of a device count (synchronizes). */
gasx_status gasx_get_count(gasx_ctx* ctx, const gasx_species* s, int64_t* n);
/* Collision rejection bound in use for gas g with step dt_coll (for tests). */
gasx_status gasx_collision_bound(gasx_ctx* ctx, const gasx_gas* g, double dt_coll, double* pmax);
/* Kernel times of the last timed gasx_run (synchronizes). */
gasx_status gasx_get_timing(gasx_ctx* ctx, gasx_timing* out);

/* ---- photoionization source ---------------------------------------------------
 * Replaces INCIDENCE1.m:27-158 (with GUASS.m:3-15, findEShell.m:1-12,
 * findSigmaPE.m:8-22, findPEangle.m:10-33, findV.m:5-8) on the r-z gas grid
 * g: a Gaussian X-ray beam along +z (sigma = beam_radius / 3, rings of width
 * gdr up to round(beam_radius / gdr)) is absorbed slab by slab (Beer-Lambert,
 * dN = n N sigma_pe gdz, N -= dN); each (ring, slab) cell emits
 * round(dN / FN) events.  An event appends a photoelectron of energy
 * Ex - E_shell with a Sauter polar angle about +z, an Auger electron
 * (auger_kev, polar angle uniform in [0, pi]) when the shell is not the last
 * one of binding_ev, and an ion (to `ions`, may be NULL) with the cell's
 * thermal velocity plus the recoil of the photon and the electrons.
 * Electrons are appended as all photoelectrons (event order: slab by slab,
 * rings outward), then all Auger electrons in the same order.  With
 * density_out / temperature_out (device, gnz x gnr, may alias the gas arrays)
 * the gas loses the ionized molecules and their thermal energy
 * (INCIDENCE1.m:87-92).  The reference rounds per (r, phi, z) DSMC cell; this
 * grid has no phi cells and rounds per (r, z) cell.  Synchronizes; returns
 * GASX_E_CAPACITY (with the result filled) when electrons or ions were dropped. */
typedef struct gasx_photo_source {
    double photon_kev;          /* Ex (main3.m:26) */
    double n_photons;           /* photons entering the wedge (main3.m:28) */
    double beam_radius;         /* laser_r (main3.m:25), m */
    const double* binding_ev;   /* host: shell binding energies in file order (EB_14.csv), eV */
    int32_t n_shells;
    const double* mac;          /* host: n_mac x 2 row-major (14_MAC.csv: MeV ; cm^2/g), energies increasing */
    int32_t n_mac;
    double atomic_mass;         /* m_a (main3.m:27) */
    double auger_kev;           /* 0.383 (INCIDENCE1.m:130) */
    uint64_t seed;              /* Philox key */
    uint32_t pulse, run;        /* Philox counter words 1 and 2 */
    float* density_out;         /* device or NULL (no depletion) */
    float* temperature_out;     /* device, with density_out */
} gasx_photo_source;

typedef struct gasx_photo_result {
    double photons_left;        /* No_photon (INCIDENCE1.m:84) */
    int64_t n_events;           /* photoionized molecules (excited_molecule_number) */
    int64_t n_electrons, n_ions;                  /* appended */
    int64_t n_dropped_electrons, n_dropped_ions;  /* beyond capacity */
    double sigma_pe;            /* m^2 (findSigmaPE.m:22) */
    double electron_kev;        /* photoelectron energy Ex - E_shell */
    int32_t auger;              /* 1: inner shell, an Auger electron per event */
    int32_t n_rings;            /* beam rings on the gas grid */
} gasx_photo_result;

gasx_status gasx_photoionize(gasx_ctx* ctx, gasx_species* e, gasx_species* ions, const gasx_gas* g,
                             const gasx_photo_source* src, gasx_photo_result* out);

/* ---- electrostatic ion step ---------------------------------------------------
 * Replaces ion_staticelectrofield_push.m:21-156 with Poisson.m:25-68: the ion
 * species are half drifted (x += dt/2 v), counted in the cells of an r-z grid
 * of nr x nz cells of size delta (cell centres at (i + 1/2) delta, the grid of
 * Poisson.m:11-12; r and z cells clamped separately), their charge density
 * rho_i = W e sum_s q_s n_s / V(i) is formed (V(i) = PHIp/2 ((i+1)^2 - i^2)
 * delta^3), and the potential PHI of Poisson.m's iteration is found: Jacobi
 * sweeps of its stencil with Boltzmann electrons rho_e = e n_e0
 * exp(e PHI / kB Te), n_e0 renormalised each sweep so that the electrons'
 * charge equals the ions', only positive updates kept, PHI(0,:) = PHI(1,:)
 * and PHI = 0 on the other three edges, until max|dPHI| <= tol mean(PHI)
 * (or max_iter sweeps).  Then E = -gradient(PHI) (MATLAB's rule) and every
 * ion gets v += q e E(cell) / m dt along r-hat and z-hat of its half-drifted
 * position and x = x_half + dt/2 v.  phi, rho_i, rho_e, e_r, e_z: device,
 * nr*nz doubles each, index j*nr + i (outputs).  The reference repeats one
 * identical solve 100 times (Poisson.m:25-27); it is done once.
 * Synchronizes. */
typedef struct gasx_es_params {
    int32_t nr, nz;             /* cells in r and z, >= 3 each */
    double delta;               /* cell size (m): delta_r = delta_z (Poisson.m:9-10) */
    double phip;                /* wedge angle of the cell volumes */
    double weight;              /* W: real ions per simulated ion */
    double t_e;                 /* electron temperature (K) */
    double dt;                  /* ion step (s) */
    double n_e0;                /* starting electron density scale (Poisson.m:33: 1e15) */
    double tol;                 /* stopping rule factor (Poisson.m:36: 1/100) */
    int32_t max_iter;           /* sweep cap (the reference has none) */
} gasx_es_params;

typedef struct gasx_es_result {
    int32_t iterations;         /* Jacobi sweeps done */
    int32_t converged;          /* 1: stopped by the rule, 0: by max_iter */
    double n_e0;                /* final electron density scale */
    double max_dphi, mean_phi;  /* the stopping quantities of the last sweep */
    double int_rho_i;           /* trapz(trapz(rho_i .* R)) (Poisson.m:22) */
} gasx_es_result;

gasx_status gasx_ion_es_push(gasx_ctx* ctx, gasx_species* ions, int n_species, const gasx_es_params* p,
                             double* phi, double* rho_i, double* rho_e, double* e_r, double* e_z,
                             gasx_es_result* out);

/* ---- DSMC neutral gas -----------------------------------------------------------
 * Replaces the molecule step of main2.m:247-271: free_motion.m:3-7,
 * boundary_conditions_specular.m:6-64 (specular_scattering_xy.m,
 * boundary_phi_0.m / boundary_phi_phimax.m, specular_scattering_z.m),
 * put_into_cells_symmetry.m:4-16 and calculate_VC.m:5-17 on the cell map of
 * cell_structure.m:6-22, and collision_HS_internal.m:11-158 (no-time-counter
 * hard-sphere pairs, Larsen-Borgnakke exchange of internal energy, VHS or VSS
 * scattering).  Molecules are FP32 structure-of-arrays (x, y, z, vx, vy, vz in
 * the wedge's Cartesian frame, ei internal energy in J); the arithmetic is
 * FP64.  A molecule's cell is that of its stored position: ring
 * ceil(r / dr), phi cell ceil(phi / dphi(ring)), layer ceil(z / dz), clamped
 * into the map; its molecules are taken in ascending slot order (the
 * reference's `find(INDEX == indexnum)`).  Each cell runs
 * max(Npair, min_attempts) attempts, Npair = floor(SCmax FN N (N-1) / VC
 * 0.5 dt), when it holds more than 2 molecules; every attempt draws its
 * random numbers from Philox4x32-10 with key = seed and counter = (attempt,
 * cell, step, block), so the result does not depend on how the GPU schedules
 * the attempts.  VC is the volume of the cell's ring cell (the reference's
 * VC holds one layer; every layer has the same cells). */
typedef struct gasx_cell_map {
    int32_t n_rings, n_layers;  /* rmax (r_divide) and zmax (z_divide) */
    double rp, lp, phip;        /* Rp, Lp, PHIp */
    double delta;               /* target arc length of cell_structure.m (main2.m:42 lambda) */
} gasx_cell_map;

typedef struct gasx_molecules {
    float *x, *y, *z, *vx, *vy, *vz, *ei;   /* device, capacity each */
    int32_t* count;                          /* device int32: slots in use; z = NaN marks an unused slot */
    int64_t capacity;
} gasx_molecules;

typedef struct gasx_dsmc_params {
    double dt;                  /* main2.m:44 */
    double fn;                  /* real molecules per simulated one (main2.m:47) */
    double mass;                /* m (main2.m:15) */
    double diameter;            /* d of the hard-sphere cross section (main2.m:14) */
    double sc_max;              /* SCmax = max of cr sigma (main2.m:35) */
    double omega;               /* w (main2.m:16), < 1.5 */
    double zeta;                /* internal degrees of freedom (main2.m:17), >= 2 */
    double alpha;               /* alpha_ (main2.m:36): 1 = VHS isotropic scattering, else VSS */
    int32_t min_attempts;       /* 1: the reference's loop, which makes one attempt even when Npair = 0; 0: Npair */
    int32_t reorder;            /* 1: permute the molecules into cell order (stable) before colliding;
                                   unused slots are dropped and count reduced */
} gasx_dsmc_params;

typedef struct gasx_dsmc_counters {
    int64_t n_attempts, n_collisions;
    int64_t n_sc_exceed;        /* attempts with cr sigma > SCmax (n_SC, collision_HS_internal.m:48-51) */
    int64_t n_boundary;         /* N_boundary of boundary_conditions_specular.m */
    int64_t n_wall_r, n_wall_z, n_wrap;
} gasx_dsmc_counters;

/* Cells per layer (cellnumxy) and phi cells per ring (n_rings entries, may be NULL). */
gasx_status gasx_cell_map_info(const gasx_cell_map* m, int32_t* phi_divide, int64_t* cells_per_layer);
/* 0-based cell of every slot below the count (device int32; n_cells for an unused slot). */
gasx_status gasx_dsmc_cells(gasx_ctx* ctx, const gasx_molecules* mol, const gasx_cell_map* map, int32_t* cell);
/* free_motion.m + boundary_conditions_specular.m: one move of duration dt. */
gasx_status gasx_dsmc_move(gasx_ctx* ctx, gasx_molecules* mol, const gasx_cell_map* map, double dt);
/* collision_HS_internal.m on the current positions (synchronizes once: the sort takes the count on the host). */
gasx_status gasx_dsmc_collide(gasx_ctx* ctx, gasx_molecules* mol, const gasx_cell_map* map,
                              const gasx_dsmc_params* p, uint64_t seed, uint32_t step);
/* main2.m:250-264: move, cells, collisions (synchronizes once). */
gasx_status gasx_dsmc_step(gasx_ctx* ctx, gasx_molecules* mol, const gasx_cell_map* map,
                           const gasx_dsmc_params* p, uint64_t seed, uint32_t step);
/* The gas the electron step reads (particle_density0.m:24-36, count_T_numbensity.m:11-13):
 * per cell of the map density FN N / VC and temperature m / (3 kB) mean(v^2)
 * (cell_density, cell_temperature: device, n_cells, may be NULL), and the same
 * per (ring, layer) cell of the map (rz_density, rz_temperature: device,
 * n_layers x n_rings, index iz * n_rings + ir, the gasx_gas layout with
 * gdr = Rp / n_rings, gdz = Lp / n_layers; may be NULL).  Empty cells: 0. */
gasx_status gasx_dsmc_sample(gasx_ctx* ctx, const gasx_molecules* mol, const gasx_cell_map* map, double fn,
                             double mass, float* cell_density, float* cell_temperature, float* rz_density,
                             float* rz_temperature);
/* Counters accumulated by the DSMC calls (synchronizes). */
gasx_status gasx_dsmc_get_counters(gasx_ctx* ctx, gasx_dsmc_counters* out, int reset);

/* ---- Differentiable run and its adjoint ------------------------------------
 * No counterpart in the reference, which has no gradients: these entry points
 * serve callers that fit the gas profile or the source to an observed field
 * (a MATLAB optimizer loop over the MEX shim, or torch.autograd).
 *
 * The model is the electron step of gasx_run (EM2.cu drift and wedge
 * boundary map, EM7.cu deposit in FP64, EM3_1.cu / EM3.cu field cycle, EM4.cu
 * gather and push) with the collisions of a recording pass frozen pathwise
 * (common random numbers): each particle-step keeps its collision decision,
 * velocity kick and the probability p0 = |v| ddt n0(r) sigma_T(E) at the
 * reference profile n0; a replay at profile n applies the same kicks and
 * multiplies the particle's weight by the likelihood ratio of the decision,
 * n(r)/n0(r) after a collision and (1 - p0 n/n0)/(1 - p0) otherwise.  The
 * density profile n(r) is linear between nodes at (j + 1/2) Rp / n_nodes and
 * constant beyond the first and last.  Secondaries are not created.  The
 * source amplitude A multiplies the deposit (FN -> A FN).  The objective is
 * the electric energy of the outer shell at the last step,
 * W = PHIp eps0 / 2 sum over Er cells with (i + 1/2) dr >= r_frac Rp of Er^2 r dr dz. */
typedef struct gasx_diff_spec {
    int32_t n_nodes;          /* profile nodes, 2..64 */
    const double* nodes0;     /* host, n_nodes: reference profile n0 (m^-3), > 0 */
    int32_t n_steps;          /* pushes */
    int32_t checkpoint_every; /* steps between the adjoint's checkpoints */
    uint64_t seed;            /* Philox key of the recording pass */
    uint32_t run_id;          /* Philox counter word 2 */
    double r_frac;            /* objective shell, in [0, 1) */
} gasx_diff_spec;

/* Records a run of the electrons in `e` (weight 1, fields zero at the start)
 * at the reference profile; the species is only read.  Replaces any earlier
 * recorded run of the context. */
gasx_status gasx_diff_record(gasx_ctx* ctx, const gasx_species* e, const gasx_diff_spec* spec);
/* Replays the recorded run at profile `nodes` (host, n_nodes) and amplitude;
 * returns W (synchronizes). */
gasx_status gasx_diff_forward(gasx_ctx* ctx, const double* nodes, double amplitude, double* objective);
/* dW/dnodes (host, n_nodes) and dW/dA of the last gasx_diff_forward, by the
 * hand-written adjoint with checkpoint recomputation (synchronizes). */
gasx_status gasx_diff_backward(gasx_ctx* ctx, double* grad_nodes, double* grad_amplitude);
/* Device copies of the recorded events of one step (caller's device arrays of
 * the recorded particle count): kick, p0 and the collision flag. */
gasx_status gasx_diff_events(gasx_ctx* ctx, int32_t step, float* dvx, float* dvy, float* dvz, float* p0,
                             uint8_t* collided);
/* Frees the recorded run. */
gasx_status gasx_diff_release(gasx_ctx* ctx);
/* The adjoint's transposed field cycle on its own (for tests): with the
 * cycle F' = M F + B J of gasx_fdtd (one push), replaces lam by M^T lam and
 * writes B^T lam to ljr, ljz (device, nr*nz). */
gasx_status gasx_fdtd_transpose(gasx_ctx* ctx, gasx_fields* lam, double* ljr, double* ljz);

#ifdef __cplusplus
}
#endif
#endif /* GASX_H */
