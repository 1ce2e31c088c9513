/*
 * dsmc_driver.c -- main2.m's molecule loop through the host-array glue
 * (gasx_shim.h), in the reference's call order:
 *
 *   molecules and cell map                    main2.m:40-106, cell_structure.m
 *   for step = 1 : N_STEPS                    main2.m:247
 *       free_motion + boundary_conditions_specular                     :250-254
 *       put_into_cells_symmetry + calculate_VC + collision_HS_internal :258-264
 *   particle_density0 / count_T_numbensity (the gas of the electron step)
 *
 * usage: dsmc_driver DATA_DIR IN OUT N_RINGS N_LAYERS RP LP PHIP DELTA DT FN N_STEPS SEED REORDER
 *
 * IN holds the molecules (little-endian): int64 n, then x, y, z, vx, vy, vz,
 * ei as n doubles each.  N2 (m = 4.65e-26 kg, d = 3.7e-10 m), hard-sphere
 * pairs with Larsen-Borgnakke exchange (omega = 0.74, zeta = 2), SCmax of
 * main2.m:35 at 300 K.  The electron context of the shim gets an N_RINGS x
 * N_LAYERS gas grid of the map's (ring, layer) cells, so the sample becomes its
 * gas.  OUT receives: "GASXDSM1", int64 n, the seven molecule arrays, int64
 * attempts, collisions, sc_exceed, boundary, wall_r, wall_z, wraps, then the
 * (ring, layer) density and temperature of the sample and the electron step's
 * gas density and temperature read back (N_LAYERS x N_RINGS doubles each,
 * ring fastest).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gasx_shim.h"

static int fail(const char* what, gasx_status s, gasx_shim* sh)
{
    fprintf(stderr, "dsmc_driver: %s: %s (%s)\n", what, gasx_status_string(s),
            sh ? gasx_last_error(gasx_shim_context(sh)) : "");
    return 1;
}

static void put64(FILE* f, int64_t v) { fwrite(&v, sizeof v, 1, f); }
static void putd(FILE* f, const double* v, size_t n) { fwrite(v, sizeof(double), n, f); }

int main(int argc, char** argv)
{
    if (argc != 15) {
        fprintf(stderr, "usage: %s DATA_DIR IN OUT N_RINGS N_LAYERS RP LP PHIP DELTA DT FN N_STEPS SEED REORDER\n",
                argv[0]);
        return 2;
    }
    const char* dir = argv[1];
    gasx_cell_map map;
    map.n_rings = atoi(argv[4]);
    map.n_layers = atoi(argv[5]);
    map.rp = atof(argv[6]);
    map.lp = atof(argv[7]);
    map.phip = atof(argv[8]);
    map.delta = atof(argv[9]);
    const double dt = atof(argv[10]), fn = atof(argv[11]);
    const int n_steps = atoi(argv[12]);
    const uint64_t seed = strtoull(argv[13], NULL, 10);
    const int reorder = atoi(argv[14]);
    if (map.n_rings < 1 || map.n_layers < 1 || !(dt > 0) || !(fn > 0) || n_steps < 0) {
        fprintf(stderr, "dsmc_driver: bad arguments\n");
        return 2;
    }
    FILE* in = fopen(argv[2], "rb");
    if (!in) {
        fprintf(stderr, "dsmc_driver: cannot read %s\n", argv[2]);
        return 2;
    }
    int64_t n = 0;
    if (fread(&n, sizeof n, 1, in) != 1 || n < 0 || n > 100000000) {
        fprintf(stderr, "dsmc_driver: bad molecule count\n");
        fclose(in);
        return 2;
    }
    double* mol = (double*)malloc((size_t)(n > 0 ? n : 1) * 7 * sizeof(double));
    if (!mol || fread(mol, sizeof(double), (size_t)n * 7, in) != (size_t)n * 7) {
        fprintf(stderr, "dsmc_driver: short molecule file\n");
        fclose(in);
        free(mol);
        return 2;
    }
    fclose(in);

    /* a small electron grid: the shim needs one; its gas grid is the map's (ring, layer) grid */
    gasx_params p;
    memset(&p, 0, sizeof p);
    p.nr = 16; p.nz = 16; p.dr = 1e-5; p.dz = 1e-5; p.phip = map.phip;
    p.ddt = 1e-14; p.n_sub = 10; p.fn = 1.0;
    p.c = 3e8; p.eps0 = 8.854187817e-12; p.mu0 = 4e-7 * 3.14159265358979323846;  /* electricmagneticfield6.m:8-11 */
    p.e_em = 1.6e-19; p.me_em = 9.1096e-31;
    p.e_mc = 1.6022e-19; p.me_mc = 9.109e-31;                                     /* electron_transport.cu:47-51 */
    p.m_ion = 4.65e-26; p.k_b = 1.38064853e-23;
    p.ez_metric = 0; p.axis_j_sign = 1; p.exc_form = 0; p.rot_eps = 1e-5;
    const int64_t cap[3] = {1024, 1024, 1024};
    gasx_shim* sh = NULL;
    gasx_status s = gasx_shim_create(0, &p, dir, cap, map.n_rings, map.n_layers, map.rp / map.n_rings,
                                     map.lp / map.n_layers, &sh);
    if (s != GASX_OK) return fail("create", s, NULL);
    s = gasx_shim_dsmc_load(sh, &map, mol, mol + n, mol + 2 * n, mol + 3 * n, mol + 4 * n, mol + 5 * n,
                            mol + 6 * n, n, n > 0 ? n : 1);
    if (s != GASX_OK) return fail("dsmc_load", s, sh);

    const double kb = 1.38064853e-23, m = 4.65e-26;                               /* main2.m:5-17 */
    gasx_dsmc_params dp;
    memset(&dp, 0, sizeof dp);
    dp.dt = dt; dp.fn = fn; dp.mass = m; dp.diameter = 3.7e-10;
    dp.sc_max = 15.0 * (3.14159265358979323846 * 4.17e-10 * 4.17e-10) * sqrt(kb * 300.0 / m);   /* main2.m:35 */
    dp.omega = 0.74; dp.zeta = 2.0; dp.alpha = 1.0; dp.min_attempts = 1; dp.reorder = reorder;
    if ((s = gasx_shim_dsmc_counters(sh, &(gasx_dsmc_counters){0}, 1)) != GASX_OK) return fail("counters", s, sh);
    for (int k = 0; k < n_steps; ++k)
        if ((s = gasx_shim_dsmc_step(sh, &dp, seed, (uint32_t)k)) != GASX_OK) return fail("dsmc_step", s, sh);

    const size_t nrz = (size_t)map.n_rings * map.n_layers;
    double* rz = (double*)malloc(nrz * 4 * sizeof(double));
    if (!rz) return fail("alloc", GASX_E_NOMEM, sh);
    if ((s = gasx_shim_dsmc_sample(sh, fn, m, NULL, NULL, rz, rz + nrz, 1)) != GASX_OK) return fail("sample", s, sh);
    if ((s = gasx_shim_get_gas(sh, rz + 2 * nrz, rz + 3 * nrz)) != GASX_OK) return fail("get_gas", s, sh);
    gasx_dsmc_counters c;
    if ((s = gasx_shim_dsmc_counters(sh, &c, 0)) != GASX_OK) return fail("counters", s, sh);
    const double *x, *y, *z, *vx, *vy, *vz, *ei;
    int64_t nm = 0;
    if ((s = gasx_shim_dsmc_get(sh, &x, &y, &z, &vx, &vy, &vz, &ei, &nm)) != GASX_OK) return fail("dsmc_get", s, sh);

    FILE* f = fopen(argv[3], "wb");
    if (!f) {
        fprintf(stderr, "dsmc_driver: cannot write %s\n", argv[3]);
        return 2;
    }
    fwrite("GASXDSM1", 1, 8, f);
    put64(f, nm);
    const double* cols[7] = {x, y, z, vx, vy, vz, ei};
    for (int k = 0; k < 7; ++k) putd(f, cols[k], (size_t)nm);
    put64(f, c.n_attempts); put64(f, c.n_collisions); put64(f, c.n_sc_exceed); put64(f, c.n_boundary);
    put64(f, c.n_wall_r); put64(f, c.n_wall_z); put64(f, c.n_wrap);
    putd(f, rz, nrz * 4);
    fclose(f);
    printf("dsmc_driver: %lld molecules, %d steps, %lld attempts, %lld collisions\n", (long long)nm, n_steps,
           (long long)c.n_attempts, (long long)c.n_collisions);
    free(rz);
    free(mol);
    gasx_shim_destroy(sh);
    return 0;
}
