/*
 * es_driver.c -- one call of ion_staticelectrofield_push.m (with Poisson.m)
 * through the host-array glue (gasx_shim.h), as electron_transport_field2.m:270
 * makes it: the two ion species in, pushed ions, PHI, rho_i, rho_e and E out.
 *
 * usage: es_driver DATA_DIR IN OUT NR NZ DELTA DT FN T_E PHIP
 *
 * IN (little-endian): int64 n1, then x, y, z, vx, vy, vz of the N2+ ions as
 * n1 doubles each; int64 n2 and the same for the N2++ ions.  OUT receives
 * "GASXESD1", int64 n1, the six N2+ arrays, int64 n2, the six N2++ arrays,
 * int64 iterations and converged, double n_e0, then PHI, rho_i, rho_e, E_r and
 * E_z as NR x NZ doubles each (r fastest).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gasx_shim.h"

static int fail(const char* what, gasx_status s, gasx_shim* sh)
{
    fprintf(stderr, "es_driver: %s: %s (%s)\n", what, gasx_status_string(s),
            sh ? gasx_last_error(gasx_shim_context(sh)) : "");
    return 1;
}

static void put64(FILE* f, int64_t v) { fwrite(&v, sizeof v, 1, f); }
static void putd(FILE* f, const double* v, size_t n) { fwrite(v, sizeof(double), n, f); }

static double* read_species(FILE* in, int64_t* n)
{
    if (fread(n, sizeof *n, 1, in) != 1 || *n < 0 || *n > 100000000) return NULL;
    double* a = (double*)malloc((size_t)(*n > 0 ? *n : 1) * 6 * sizeof(double));
    if (a && fread(a, sizeof(double), (size_t)*n * 6, in) != (size_t)*n * 6) {
        free(a);
        return NULL;
    }
    return a;
}

int main(int argc, char** argv)
{
    if (argc != 11) {
        fprintf(stderr, "usage: %s DATA_DIR IN OUT NR NZ DELTA DT FN T_E PHIP\n", argv[0]);
        return 2;
    }
    const int nr = atoi(argv[4]), nz = atoi(argv[5]);
    const double delta = atof(argv[6]), dt = atof(argv[7]), fn = atof(argv[8]), t_e = atof(argv[9]);
    const double phip = atof(argv[10]);
    if (nr < 3 || nz < 3 || !(delta > 0) || !(fn > 0) || !(t_e > 0) || !(phip > 0)) {
        fprintf(stderr, "es_driver: bad arguments\n");
        return 2;
    }
    FILE* in = fopen(argv[2], "rb");
    if (!in) {
        fprintf(stderr, "es_driver: cannot read %s\n", argv[2]);
        return 2;
    }
    int64_t n1 = 0, n2 = 0;
    double* a = read_species(in, &n1);
    double* b = a ? read_species(in, &n2) : NULL;
    fclose(in);
    if (!a || !b) {
        fprintf(stderr, "es_driver: bad ion file\n");
        free(a);
        return 2;
    }

    gasx_params p;
    memset(&p, 0, sizeof p);
    p.nr = nr; p.nz = nz; p.dr = delta; p.dz = delta; p.phip = phip;
    p.ddt = 1e-14; p.n_sub = 10; p.fn = fn;
    p.c = 3e8; p.eps0 = 8.854187817e-12; p.mu0 = 4e-7 * 3.14159265358979323846;  /* electricmagneticfield6.m:8-11 */
    p.e_em = 1.6e-19; p.me_em = 9.1096e-31;
    p.e_mc = 1.6022e-19; p.me_mc = 9.109e-31;                                     /* electron_transport.cu:47-51 */
    p.m_ion = 4.65e-26; p.k_b = 1.38064853e-23;                                   /* ion_staticelectrofield_push.m:3-5 */
    p.ez_metric = 0; p.axis_j_sign = 1; p.exc_form = 0; p.rot_eps = 1e-5;
    const int64_t cap[3] = {1024, n1 > 0 ? n1 : 1, n2 > 0 ? n2 : 1};
    gasx_shim* sh = NULL;
    gasx_status s = gasx_shim_create(0, &p, argv[1], cap, 1, 1, nr * delta, nz * delta, &sh);
    if (s != GASX_OK) return fail("create", s, NULL);
    s = gasx_shim_set_particles(sh, GASX_SHIM_ION1, a, a + n1, a + 2 * n1, a + 3 * n1, a + 4 * n1, a + 5 * n1, n1);
    if (s != GASX_OK) return fail("set ion1", s, sh);
    s = gasx_shim_set_particles(sh, GASX_SHIM_ION2, b, b + n2, b + 2 * n2, b + 3 * n2, b + 4 * n2, b + 5 * n2, n2);
    if (s != GASX_OK) return fail("set ion2", s, sh);

    const size_t ng = (size_t)nr * nz;
    double* g = (double*)malloc(5 * ng * sizeof(double));
    if (!g) return fail("alloc", GASX_E_NOMEM, sh);
    gasx_es_result r;
    s = gasx_shim_es_push(sh, nr, nz, delta, dt, fn, t_e, g, g + ng, g + 2 * ng, g + 3 * ng, g + 4 * ng, &r);
    if (s != GASX_OK) return fail("es_push", s, sh);

    FILE* f = fopen(argv[3], "wb");
    if (!f) {
        fprintf(stderr, "es_driver: cannot write %s\n", argv[3]);
        return 2;
    }
    fwrite("GASXESD1", 1, 8, f);
    for (int k = GASX_SHIM_ION1; k <= GASX_SHIM_ION2; ++k) {
        const double *x, *y, *z, *vx, *vy, *vz;
        int64_t n = 0;
        s = gasx_shim_get_particles(sh, k, &x, &y, &z, &vx, &vy, &vz, &n);
        if (s != GASX_OK) return fail("get ions", s, sh);
        put64(f, n);
        const double* cols[6] = {x, y, z, vx, vy, vz};
        for (int c = 0; c < 6; ++c) putd(f, cols[c], (size_t)n);
    }
    put64(f, r.iterations);
    put64(f, r.converged);
    putd(f, &r.n_e0, 1);
    putd(f, g, 5 * ng);
    fclose(f);
    printf("es_driver: %lld + %lld ions, %d sweeps, converged %d\n", (long long)n1, (long long)n2, r.iterations,
           r.converged);
    free(g);
    free(a);
    free(b);
    gasx_shim_destroy(sh);
    return 0;
}
