/*
 * transport_driver.c -- one pulse of main3.m through the host-array glue
 * (gasx_shim.h), in the reference's call order:
 *
 *   INCIDENCE1                               main3.m:134
 *   for time = 1 : n_transport               electron_transport_field3.m:61
 *       electricmagneticfield6 (n_push pushes)                         :68
 *       find_CS0 (on the electrons' kinetic energies)                  :108
 *       electron_collision (dt_transport = n_push * ddt)               :109
 *
 * usage: transport_driver DATA_DIR OUT NR NZ FN PHOTONS N_TRANSPORT N_PUSH SEED
 *
 * Grid: NR x NZ cells of 10 um, PHIp = 0.1, ddt = 1e-14 s with 10 field
 * substeps, uniform N2 at 2.5e25 m^-3 and 300 K on the same grid; a 1 keV
 * pulse of PHOTONS photons in a beam of radius 150 um.  OUT receives a
 * binary record (little-endian): "GASXDRV1", int64 n_e, n_ion1, n_ion2, nr,
 * nz, n_cs, elastic, ionize, excite, events; double photons_left; the three
 * species as 6 rows of doubles each (x, y, z, vx, vy, vz), Er, Ez, Hphi, the
 * gas density and temperature, and the last find_CS0 call's energies (keV)
 * and sigma_el, sigma_ion, sigma_exc.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gasx_shim.h"

static int fail(const char* what, gasx_status s, gasx_shim* sh)
{
    fprintf(stderr, "transport_driver: %s: %s (%s)\n", what, gasx_status_string(s),
            sh ? gasx_last_error(gasx_shim_context(sh)) : "");
    return 1;
}

static void put64(FILE* f, int64_t v) { fwrite(&v, sizeof v, 1, f); }
static void putd(FILE* f, const double* v, size_t n) { fwrite(v, sizeof(double), n, f); }

int main(int argc, char** argv)
{
    if (argc != 10) {
        fprintf(stderr, "usage: %s DATA_DIR OUT NR NZ FN PHOTONS N_TRANSPORT N_PUSH SEED\n", argv[0]);
        return 2;
    }
    const char* dir = argv[1];
    const char* out = argv[2];
    const int nr = atoi(argv[3]), nz = atoi(argv[4]);
    const double fn = atof(argv[5]), photons = atof(argv[6]);
    const int n_transport = atoi(argv[7]), n_push = atoi(argv[8]);
    const uint64_t seed = strtoull(argv[9], NULL, 10);
    if (nr < 2 || nz < 2 || !(fn > 0) || !(photons > 0) || n_transport < 1 || n_push < 1) {
        fprintf(stderr, "transport_driver: bad arguments\n");
        return 2;
    }
    gasx_params p;
    memset(&p, 0, sizeof p);
    p.nr = nr; p.nz = nz; p.dr = 1e-5; p.dz = 1e-5; p.phip = 0.1;
    p.ddt = 1e-14; p.n_sub = 10; p.fn = fn;
    p.c = 3e8; p.eps0 = 8.854187817e-12; p.mu0 = 4e-7 * 3.14159265358979323846;  /* electricmagneticfield6.m:8-11 */
    p.e_em = 1.6e-19; p.me_em = 9.1096e-31;
    p.e_mc = 1.6022e-19; p.me_mc = 9.109e-31;                                     /* electron_transport.cu:47-51 */
    p.m_ion = 4.65e-26; p.k_b = 1.38064853e-23;
    p.ez_metric = 0; p.axis_j_sign = 1; p.exc_form = 0; p.rot_eps = 1e-5;
    const int64_t cap[3] = {8000000, 4000000, 4000000};
    gasx_shim* sh = NULL;
    gasx_status s = gasx_shim_create(0, &p, dir, cap, nr, nz, p.dr, p.dz, &sh);
    if (s != GASX_OK) return fail("create", s, NULL);
    const size_t ng = (size_t)nr * nz;
    double* dens = (double*)malloc(ng * sizeof(double));
    double* temp = (double*)malloc(ng * sizeof(double));
    for (size_t i = 0; i < ng; ++i) { dens[i] = 2.5e25; temp[i] = 300.0; }
    if ((s = gasx_shim_set_gas(sh, dens, temp)) != GASX_OK) return fail("set_gas", s, sh);

    gasx_photo_result pr;
    s = gasx_shim_photoionize(sh, 1.0, photons, 150e-6, seed, 0, &pr);
    if (s != GASX_OK) return fail("photoionize", s, sh);

    const double* kev = NULL;
    int64_t n_cs = 0;
    double *se = NULL, *si = NULL, *sx = NULL, *ke_copy = NULL;
    for (int t = 0; t < n_transport; ++t) {
        if ((s = gasx_shim_em_step(sh, n_push)) != GASX_OK) return fail("em_step", s, sh);
        if ((s = gasx_shim_electron_kev(sh, &kev, &n_cs)) != GASX_OK) return fail("electron_kev", s, sh);
        free(se); free(si); free(sx); free(ke_copy);
        se = (double*)malloc((size_t)(n_cs + 1) * sizeof(double));
        si = (double*)malloc((size_t)(n_cs + 1) * sizeof(double));
        sx = (double*)malloc((size_t)(n_cs + 1) * sizeof(double));
        ke_copy = (double*)malloc((size_t)(n_cs + 1) * sizeof(double));
        memcpy(ke_copy, kev, (size_t)n_cs * sizeof(double));
        if ((s = gasx_shim_find_cs(sh, ke_copy, n_cs, se, si, sx)) != GASX_OK) return fail("find_cs", s, sh);
        if ((s = gasx_shim_collide(sh, n_push * p.ddt, seed, (uint32_t)t)) != GASX_OK) return fail("collide", s, sh);
    }
    gasx_counters c;
    s = gasx_get_counters(gasx_shim_context(sh), &c, 0);
    if (s != GASX_OK && s != GASX_E_CAPACITY) return fail("counters", s, sh);

    FILE* f = fopen(out, "wb");
    if (!f) {
        perror(out);
        return 1;
    }
    const double* sp[3][6];   /* each species has its own host buffers */
    int64_t n[3];
    for (int k = 0; k < 3; ++k) {
        s = gasx_shim_get_particles(sh, k, &sp[k][0], &sp[k][1], &sp[k][2], &sp[k][3], &sp[k][4], &sp[k][5], &n[k]);
        if (s != GASX_OK) return fail("get_particles", s, sh);
    }
    fwrite("GASXDRV1", 1, 8, f);
    put64(f, n[0]); put64(f, n[1]); put64(f, n[2]);
    put64(f, nr); put64(f, nz); put64(f, n_cs);
    put64(f, c.n_elastic); put64(f, c.n_ionize); put64(f, c.n_excite); put64(f, pr.n_events);
    putd(f, &pr.photons_left, 1);
    for (int k = 0; k < 3; ++k)
        for (int q = 0; q < 6; ++q) putd(f, sp[k][q], (size_t)n[k]);
    const size_t ner = (size_t)nr * (nz + 1), nez = (size_t)(nr + 1) * nz, nh = (size_t)nr * nz;
    double* er = (double*)malloc(ner * sizeof(double));
    double* ez = (double*)malloc(nez * sizeof(double));
    double* h = (double*)malloc(nh * sizeof(double));
    if ((s = gasx_shim_get_fields(sh, er, ez, h)) != GASX_OK) return fail("get_fields", s, sh);
    if ((s = gasx_shim_get_gas(sh, dens, temp)) != GASX_OK) return fail("get_gas", s, sh);
    putd(f, er, ner); putd(f, ez, nez); putd(f, h, nh);
    putd(f, dens, ng); putd(f, temp, ng);
    putd(f, ke_copy, (size_t)n_cs); putd(f, se, (size_t)n_cs); putd(f, si, (size_t)n_cs); putd(f, sx, (size_t)n_cs);
    fclose(f);
    printf("transport_driver: %lld electrons, %lld N2+, %lld N2++ after %d transport steps; "
           "events el/ion/exc %lld/%lld/%lld\n",
           (long long)n[0], (long long)n[1], (long long)n[2], n_transport, (long long)c.n_elastic,
           (long long)c.n_ionize, (long long)c.n_excite);
    free(dens); free(temp); free(er); free(ez); free(h);
    free(se); free(si); free(sx); free(ke_copy);
    gasx_shim_destroy(sh);
    return 0;
}
