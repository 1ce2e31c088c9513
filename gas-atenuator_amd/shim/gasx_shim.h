/*
 * gasx_shim.h -- host-array glue between the reference's MATLAB call surface
 * and the C ABI of include/gasx.h.
 *
 * The reference's GPU work is reached from four MATLAB functions, each
 * called once per transport step or pulse:
 *   INCIDENCE1(...)             main3.m:134                     -> gasx_shim_photoionize
 *   electricmagneticfield6(...) electron_transport_field3.m:68  -> gasx_shim_em_step
 *   find_CS0(...)               electron_transport_field3.m:108 -> gasx_shim_find_cs
 *   electron_collision(...)     electron_transport_field3.m:109 -> gasx_shim_collide
 * Their arguments are host double arrays: particles as N x 3 column-major
 * matrices (x, y, z columns: three arrays of N doubles), fields in the
 * reference's own layout.  The shim keeps the three species (electrons,
 * N2+ from electron-impact ionization, N2++ from photoionization;
 * electricmagneticfield6.m:65-69) and the fields resident on the device
 * between calls; the set / get calls move them to and from host arrays.
 * gasx_mex.c maps MATLAB's mxArrays onto these calls; transport_driver.c
 * replays main3.m / electron_transport_field3.m with them from C, and
 * dsmc_driver.c main2.m's molecule loop.
 *
 * Plain C, host only; every call returns a gasx_status.
 */
#ifndef GASX_SHIM_H
#define GASX_SHIM_H

#include <stdint.h>

#include "gasx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gasx_shim gasx_shim;

enum { GASX_SHIM_ELECTRONS = 0, GASX_SHIM_ION1 = 1, GASX_SHIM_ION2 = 2 };

/* Loads the N2 tables from data_dir (the reference's CSV files, read as
 * electron_transport_field3.m:8-37 reads them), creates the context and the
 * device buffers: capacity[3] slots per species, a gnr x gnz gas grid of
 * cells gdr x gdz. */
gasx_status gasx_shim_create(int device, const gasx_params* p, const char* data_dir, const int64_t capacity[3],
                             int32_t gnr, int32_t gnz, double gdr, double gdz, gasx_shim** out);
void gasx_shim_destroy(gasx_shim* sh);
gasx_ctx* gasx_shim_context(gasx_shim* sh);

/* Host -> device: n particles of a species, Cartesian positions and
 * velocities (doubles; stored as FP32). */
gasx_status gasx_shim_set_particles(gasx_shim* sh, int species, const double* x, const double* y, const double* z,
                                    const double* vx, const double* vy, const double* vz, int64_t n);
/* Device -> host: the live particles of a species in slot order (absorbed
 * ones removed, electricmagneticfield6.m:176-178; the device copy is
 * compacted the same way).  The arrays stay valid until the next call. */
gasx_status gasx_shim_get_particles(gasx_shim* sh, int species, const double** x, const double** y,
                                    const double** z, const double** vx, const double** vy, const double** vz,
                                    int64_t* n);
/* Fields in the reference layout (Er nr x (nz+1), Ez (nr+1) x nz, Hphi nr x nz, r fastest). */
gasx_status gasx_shim_set_fields(gasx_shim* sh, const double* er, const double* ez, const double* hphi);
gasx_status gasx_shim_get_fields(gasx_shim* sh, double* er, double* ez, double* hphi);
/* Gas density (m^-3) and temperature (K) per gas cell, gnz x gnr, iz * gnr + ir. */
gasx_status gasx_shim_set_gas(gasx_shim* sh, const double* density, const double* temperature);
gasx_status gasx_shim_get_gas(gasx_shim* sh, double* density, double* temperature);

/* INCIDENCE1.m:27-158: photoelectrons and Auger electrons to the
 * electrons, N2++ to ion2, the gas depleted.  ex_kev, n_photons and
 * laser_r as main3.m:25-28; the binding energies and MAC table from data_dir. */
gasx_status gasx_shim_photoionize(gasx_shim* sh, double ex_kev, double n_photons, double laser_r, uint64_t seed,
                                  uint32_t pulse, gasx_photo_result* result);
/* electricmagneticfield6.m:138-188: n_push pushes of all three species
 * with the fields. */
gasx_status gasx_shim_em_step(gasx_shim* sh, int n_push);
/* find_CS0.m:1-61 on host energies (keV). */
gasx_status gasx_shim_find_cs(gasx_shim* sh, const double* e_kev, int64_t n, double* sig_el, double* sig_ion,
                              double* sig_exc);
/* electron_collision.m:1-75: one collision check of duration dt for every
 * electron; secondaries appended to the electrons, N2+ to ion1. */
gasx_status gasx_shim_collide(gasx_shim* sh, double dt, uint64_t seed, uint32_t step);
/* Kinetic energies (keV) of the live electrons in slot order, as
 * electricmagneticfield6.m:172 returns them (the push's constants). */
gasx_status gasx_shim_electron_kev(gasx_shim* sh, const double** e_kev, int64_t* n);

/* ion_staticelectrofield_push.m:21-156 with Poisson.m:25-68 on the two ion
 * species (N2+ q = +1, N2++ q = +2) of an nr x nz grid of cells delta (the
 * wedge angle of create's parameters): the ions are pushed in place; PHI,
 * rho_i, rho_e, E_r and E_z come back as nr x nz host doubles, r fastest
 * (MATLAB's lenr x lenz layout; any may be NULL).  The reference's RHO is
 * rho_i - rho_e. */
gasx_status gasx_shim_es_push(gasx_shim* sh, int32_t nr, int32_t nz, double delta, double dt, double fn, double t_e,
                              double* phi, double* rho_i, double* rho_e, double* e_r, double* e_z,
                              gasx_es_result* res);

/* ---- DSMC neutral gas: main2.m's molecule loop (main2.m:247-271) ----------
 * main2.m:59-106's molecules (positions, velocities, internal energy, host
 * doubles; stored as FP32) and cell_structure.m's map onto the device, with
 * room for `capacity` molecules. */
gasx_status gasx_shim_dsmc_load(gasx_shim* sh, const gasx_cell_map* map, const double* x, const double* y,
                                const double* z, const double* vx, const double* vy, const double* vz,
                                const double* ei, int64_t n, int64_t capacity);
/* main2.m:250-264: free flight, specular walls, cells and collision_HS_internal.m. */
gasx_status gasx_shim_dsmc_step(gasx_shim* sh, const gasx_dsmc_params* p, uint64_t seed, uint32_t step);
/* particle_density0.m / count_T_numbensity.m: density (m^-3) and temperature
 * (K) per cell of the map and per (ring, layer), n_layers x n_rings with the
 * ring fastest; any output may be NULL.  With to_gas the (ring, layer) values
 * also become the electron step's gas (the gas grid given to create must then
 * be n_rings x n_layers). */
gasx_status gasx_shim_dsmc_sample(gasx_shim* sh, double fn, double mass, double* cell_density,
                                  double* cell_temperature, double* rz_density, double* rz_temperature, int to_gas);
/* The molecules back to host in slot order (valid until the next call). */
gasx_status gasx_shim_dsmc_get(gasx_shim* sh, const double** x, const double** y, const double** z,
                               const double** vx, const double** vy, const double** vz, const double** ei,
                               int64_t* n);
gasx_status gasx_shim_dsmc_counters(gasx_shim* sh, gasx_dsmc_counters* out, int reset);
/* Cells of the loaded map and its (ring, layer) cells: the sample's output sizes. */
gasx_status gasx_shim_dsmc_sizes(gasx_shim* sh, int64_t* n_cells, int64_t* n_rz);

#ifdef __cplusplus
}
#endif
#endif /* GASX_SHIM_H */
