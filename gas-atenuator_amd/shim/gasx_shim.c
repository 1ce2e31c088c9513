/*
 * gasx_shim.c -- host-array glue over the C ABI (see gasx_shim.h).
 *
 * Plain C99 over the HIP runtime's C API: device buffers are allocated once
 * per shim, host arrays are converted to the ABI's FP32 structure of arrays
 * on the way in and back to doubles on the way out.
 */
#define __HIP_PLATFORM_AMD__
#include "gasx_shim.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float* d[6];          /* device x, y, z, vx, vy, vz */
    int32_t* d_count;
    int64_t cap;
    float* h32;           /* host staging, 6 x cap floats */
    double* h64;          /* host results, 6 x cap doubles */
    int64_t n;            /* host copy of the count after the last get */
} shim_species;

struct gasx_shim {
    gasx_ctx* ctx;
    gasx_params p;
    gasx_raw_tables raw;
    double* tab_mem[7];                 /* the CSV matrices behind raw */
    double* binding_ev;                 /* EB_14.csv */
    int32_t n_shells;
    double* mac;                        /* 14_MAC.csv */
    int32_t n_mac;
    shim_species sp[3];
    double *d_er, *d_ez, *d_h;
    float *d_dens, *d_temp;
    float* h_gas;
    int32_t gnr, gnz;
    double gdr, gdz;
    double n_max;                       /* upper bound of the density for the collision step */
    float* d_cs;                        /* cross-section scratch: 4 x cs_cap */
    float* h_cs;
    int64_t cs_cap;
    double* h_kev;
    /* DSMC molecules of main2.m (x, y, z, vx, vy, vz, ei) and their cell map */
    float* dm[7];
    int32_t* dm_count;
    int64_t dm_cap;
    float* h_mol;                       /* host staging, 7 x dm_cap */
    double* h_mol64;
    gasx_cell_map map;
    int64_t n_cells;
    float* d_samp;                      /* sample outputs: 2 x n_cells + 2 x (rings x layers) */
    float* h_samp;
};

/* ---- CSV (the reference's data files, any line endings, optional header) ---- */
static double* read_csv(const char* path, int* rows, int* cols)
{
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* text = (char*)malloc((size_t)len + 2);
    if (!text || fread(text, 1, (size_t)len, f) != (size_t)len) {
        fclose(f);
        free(text);
        return NULL;
    }
    fclose(f);
    text[len] = '\n';
    text[len + 1] = 0;
    size_t cap = 1024, n = 0;
    double* v = (double*)malloc(cap * sizeof(double));
    int r = 0, c0 = -1;
    char* line = text;
    while (*line) {
        char* end = line;
        while (*end && *end != '\n' && *end != '\r') ++end;
        char save = *end;
        *end = 0;
        int c = 0, numeric = 1;
        size_t n0 = n;
        for (char* tok = line; *tok;) {
            char* stop;
            double x = strtod(tok, &stop);
            if (stop == tok) { numeric = 0; break; }
            if (n == cap) { cap *= 2; v = (double*)realloc(v, cap * sizeof(double)); }
            v[n++] = x;
            ++c;
            tok = stop;
            while (*tok == ' ' || *tok == '\t') ++tok;
            if (*tok == ',') ++tok;
            else if (*tok) { numeric = 0; break; }
        }
        if (!numeric || c == 0) {
            n = n0;                       /* header or blank line */
        } else {
            if (c0 < 0) c0 = c;
            if (c != c0) { free(text); free(v); return NULL; }
            ++r;
        }
        line = save ? end + 1 : end;
    }
    free(text);
    *rows = r;
    *cols = c0 < 0 ? 0 : c0;
    return v;
}

static gasx_status load_tables(gasx_shim* sh, const char* dir)
{
    static const char* files[7] = {"TotalElasticCS_Nitrogen.csv", "TotalElasticCS_smalleV_Nitrogen.csv",
                                   "TotalElasticCS_smalleV_Nitrogen_1.csv", "ElasticCS_Nitrogen.csv",
                                   "ElasticDCS_smalleV_Nitrogen.csv", "TotalIonCS_Nitrogen.csv",
                                   "ParameterExcCS_Nitrogen.csv"};
    int rows[7], cols[7];
    char path[4096];
    for (int k = 0; k < 7; ++k) {
        snprintf(path, sizeof path, "%s/%s", dir, files[k]);
        sh->tab_mem[k] = read_csv(path, &rows[k], &cols[k]);
        if (!sh->tab_mem[k]) return GASX_E_ARG;
    }
    gasx_raw_tables* t = &sh->raw;
    t->t_el = sh->tab_mem[0];   t->n_el = cols[0];     /* 2 x n_el */
    t->t_sel = sh->tab_mem[1];  t->n_sel = rows[1];    /* n x 2 */
    t->t_sel1 = sh->tab_mem[2]; t->n_sel1 = rows[2];
    t->t_del = sh->tab_mem[3];  t->w_del = cols[3];    /* 182 x w */
    t->t_sdel = sh->tab_mem[4]; t->w_sdel = cols[4];
    t->t_ion = sh->tab_mem[5];  t->n_ion = rows[5];
    t->t_exc = sh->tab_mem[6];  t->n_exc = rows[6];    /* n x 5, header skipped */
    if (rows[0] != 2 || rows[3] != 182 || rows[4] != 182 || cols[6] != 5) return GASX_E_ARG;
    int r, c;
    snprintf(path, sizeof path, "%s/EB_14.csv", dir);
    sh->binding_ev = read_csv(path, &r, &c);
    if (!sh->binding_ev) return GASX_E_ARG;
    sh->n_shells = r * c;
    snprintf(path, sizeof path, "%s/14_MAC.csv", dir);
    sh->mac = read_csv(path, &r, &c);
    if (!sh->mac || c != 2) return GASX_E_ARG;
    sh->n_mac = r;
    return GASX_OK;
}

#define HCHK(x)                                   \
    do {                                          \
        if ((x) != hipSuccess) return GASX_E_HIP; \
    } while (0)

static gasx_species species_view(gasx_shim* sh, int k)
{
    static const float qnum[3] = {-1.0f, 1.0f, 2.0f};   /* electricmagneticfield6.m:68 */
    gasx_species s;
    shim_species* q = &sh->sp[k];
    s.x = q->d[0]; s.y = q->d[1]; s.z = q->d[2];
    s.vx = q->d[3]; s.vy = q->d[4]; s.vz = q->d[5];
    s.count = q->d_count;
    s.capacity = q->cap;
    s.qnum = qnum[k];
    s.mass = k == 0 ? sh->p.me_em : sh->p.m_ion;        /* electricmagneticfield6.m:69 */
    return s;
}

static gasx_gas gas_view(gasx_shim* sh)
{
    gasx_gas g;
    g.density = sh->d_dens;
    g.temperature = sh->d_temp;
    g.gnr = sh->gnr; g.gnz = sh->gnz;
    g.gdr = sh->gdr; g.gdz = sh->gdz;
    g.t_scalar = 300.0;
    g.n_max = sh->n_max;
    return g;
}

gasx_status gasx_shim_create(int device, const gasx_params* p, const char* data_dir, const int64_t capacity[3],
                             int32_t gnr, int32_t gnz, double gdr, double gdz, gasx_shim** out)
{
    if (!out || !p || !data_dir || !capacity || gnr < 1 || gnz < 1 || !(gdr > 0) || !(gdz > 0)) return GASX_E_ARG;
    *out = NULL;
    gasx_shim* sh = (gasx_shim*)calloc(1, sizeof(gasx_shim));
    if (!sh) return GASX_E_NOMEM;
    sh->p = *p;
    gasx_status st = load_tables(sh, data_dir);
    if (st == GASX_OK) st = gasx_ctx_create(device, p, &sh->raw, &sh->ctx);
    if (st != GASX_OK) {
        gasx_shim_destroy(sh);
        return st;
    }
    /* the library's kernels and the shim's synchronous copies all on the
     * null stream, so they are ordered without further synchronization */
    st = gasx_ctx_set_stream(sh->ctx, NULL);
    if (st != GASX_OK) {
        gasx_shim_destroy(sh);
        return st;
    }
    int ok = 1;
    for (int k = 0; k < 3 && ok; ++k) {
        shim_species* q = &sh->sp[k];
        q->cap = capacity[k] > 0 ? capacity[k] : 1;
        for (int c = 0; c < 6 && ok; ++c) ok = hipMalloc((void**)&q->d[c], (size_t)q->cap * sizeof(float)) == hipSuccess;
        ok = ok && hipMalloc((void**)&q->d_count, sizeof(int32_t)) == hipSuccess;
        ok = ok && hipMemset(q->d_count, 0, sizeof(int32_t)) == hipSuccess;
        q->h32 = (float*)malloc((size_t)q->cap * 6 * sizeof(float));
        q->h64 = (double*)malloc((size_t)q->cap * 6 * sizeof(double));
        ok = ok && q->h32 && q->h64;
    }
    const size_t nh = (size_t)p->nr * p->nz, ner = (size_t)p->nr * (p->nz + 1), nez = (size_t)(p->nr + 1) * p->nz;
    ok = ok && hipMalloc((void**)&sh->d_er, ner * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc((void**)&sh->d_ez, nez * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc((void**)&sh->d_h, nh * sizeof(double)) == hipSuccess;
    ok = ok && hipMemset(sh->d_er, 0, ner * sizeof(double)) == hipSuccess;
    ok = ok && hipMemset(sh->d_ez, 0, nez * sizeof(double)) == hipSuccess;
    ok = ok && hipMemset(sh->d_h, 0, nh * sizeof(double)) == hipSuccess;
    sh->gnr = gnr; sh->gnz = gnz; sh->gdr = gdr; sh->gdz = gdz;
    const size_t ng = (size_t)gnr * gnz;
    ok = ok && hipMalloc((void**)&sh->d_dens, ng * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void**)&sh->d_temp, ng * sizeof(float)) == hipSuccess;
    sh->h_gas = (float*)malloc(ng * 2 * sizeof(float));
    ok = ok && sh->h_gas;
    if (!ok) {
        gasx_shim_destroy(sh);
        return GASX_E_NOMEM;
    }
    *out = sh;
    return GASX_OK;
}

void gasx_shim_destroy(gasx_shim* sh)
{
    if (!sh) return;
    if (sh->ctx) gasx_synchronize(sh->ctx);
    for (int k = 0; k < 3; ++k) {
        for (int c = 0; c < 6; ++c)
            if (sh->sp[k].d[c]) hipFree(sh->sp[k].d[c]);
        if (sh->sp[k].d_count) hipFree(sh->sp[k].d_count);
        free(sh->sp[k].h32);
        free(sh->sp[k].h64);
    }
    for (int c = 0; c < 7; ++c)
        if (sh->dm[c]) hipFree(sh->dm[c]);
    if (sh->dm_count) hipFree(sh->dm_count);
    if (sh->d_samp) hipFree(sh->d_samp);
    free(sh->h_mol);
    free(sh->h_mol64);
    free(sh->h_samp);
    void* dev[] = {sh->d_er, sh->d_ez, sh->d_h, sh->d_dens, sh->d_temp, sh->d_cs};
    for (size_t i = 0; i < sizeof dev / sizeof dev[0]; ++i)
        if (dev[i]) hipFree(dev[i]);
    free(sh->h_gas);
    free(sh->h_cs);
    free(sh->h_kev);
    for (int k = 0; k < 7; ++k) free(sh->tab_mem[k]);
    free(sh->binding_ev);
    free(sh->mac);
    if (sh->ctx) gasx_ctx_destroy(sh->ctx);
    free(sh);
}

gasx_ctx* gasx_shim_context(gasx_shim* sh) { return sh ? sh->ctx : NULL; }

static gasx_status upload_species(gasx_shim* sh, int k, const double* const src[6], int64_t n)
{
    shim_species* q = &sh->sp[k];
    if (n < 0 || n > q->cap) return GASX_E_CAPACITY;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    for (int c = 0; c < 6; ++c) {
        float* h = q->h32 + (size_t)c * q->cap;
        for (int64_t i = 0; i < n; ++i) h[i] = (float)src[c][i];
        if (n) HCHK(hipMemcpy(q->d[c], h, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    }
    const int32_t cnt = (int32_t)n;
    HCHK(hipMemcpy(q->d_count, &cnt, sizeof cnt, hipMemcpyHostToDevice));
    q->n = n;
    return GASX_OK;
}

gasx_status gasx_shim_set_particles(gasx_shim* sh, int species, const double* x, const double* y, const double* z,
                                    const double* vx, const double* vy, const double* vz, int64_t n)
{
    if (!sh || species < 0 || species > 2 || (n > 0 && (!x || !y || !z || !vx || !vy || !vz))) return GASX_E_ARG;
    const double* src[6] = {x, y, z, vx, vy, vz};
    return upload_species(sh, species, src, n);
}

gasx_status gasx_shim_get_particles(gasx_shim* sh, int species, const double** x, const double** y,
                                    const double** z, const double** vx, const double** vy, const double** vz,
                                    int64_t* n)
{
    if (!sh || species < 0 || species > 2 || !n) return GASX_E_ARG;
    shim_species* q = &sh->sp[species];
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    int32_t cnt = 0;
    HCHK(hipMemcpy(&cnt, q->d_count, sizeof cnt, hipMemcpyDeviceToHost));
    for (int c = 0; c < 6; ++c)
        if (cnt) HCHK(hipMemcpy(q->h32 + (size_t)c * q->cap, q->d[c], (size_t)cnt * sizeof(float), hipMemcpyDeviceToHost));
    /* keep the live slots in order (a NaN z marks an absorbed particle) */
    int64_t m = 0;
    const float* zz = q->h32 + 2 * (size_t)q->cap;
    for (int64_t i = 0; i < cnt; ++i) {
        if (zz[i] != zz[i]) continue;
        for (int c = 0; c < 6; ++c) q->h64[(size_t)c * q->cap + m] = (double)q->h32[(size_t)c * q->cap + i];
        ++m;
    }
    if (m != cnt) {                          /* the device holds the same compacted population */
        const double* src[6];
        for (int c = 0; c < 6; ++c) src[c] = q->h64 + (size_t)c * q->cap;
        st = upload_species(sh, species, src, m);
        if (st != GASX_OK) return st;
    }
    q->n = m;
    const double* h = q->h64;
    const size_t cp = (size_t)q->cap;
    if (x) *x = h;
    if (y) *y = h + cp;
    if (z) *z = h + 2 * cp;
    if (vx) *vx = h + 3 * cp;
    if (vy) *vy = h + 4 * cp;
    if (vz) *vz = h + 5 * cp;
    *n = m;
    return GASX_OK;
}

gasx_status gasx_shim_set_fields(gasx_shim* sh, const double* er, const double* ez, const double* hphi)
{
    if (!sh || !er || !ez || !hphi) return GASX_E_ARG;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    const size_t nh = (size_t)sh->p.nr * sh->p.nz, ner = (size_t)sh->p.nr * (sh->p.nz + 1),
                 nez = (size_t)(sh->p.nr + 1) * sh->p.nz;
    HCHK(hipMemcpy(sh->d_er, er, ner * sizeof(double), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(sh->d_ez, ez, nez * sizeof(double), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(sh->d_h, hphi, nh * sizeof(double), hipMemcpyHostToDevice));
    gasx_fields f = {sh->d_er, sh->d_ez, sh->d_h};
    return gasx_fields_sync(sh->ctx, &f);
}

gasx_status gasx_shim_get_fields(gasx_shim* sh, double* er, double* ez, double* hphi)
{
    if (!sh || !er || !ez || !hphi) return GASX_E_ARG;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    const size_t nh = (size_t)sh->p.nr * sh->p.nz, ner = (size_t)sh->p.nr * (sh->p.nz + 1),
                 nez = (size_t)(sh->p.nr + 1) * sh->p.nz;
    HCHK(hipMemcpy(er, sh->d_er, ner * sizeof(double), hipMemcpyDeviceToHost));
    HCHK(hipMemcpy(ez, sh->d_ez, nez * sizeof(double), hipMemcpyDeviceToHost));
    HCHK(hipMemcpy(hphi, sh->d_h, nh * sizeof(double), hipMemcpyDeviceToHost));
    return GASX_OK;
}

gasx_status gasx_shim_set_gas(gasx_shim* sh, const double* density, const double* temperature)
{
    if (!sh || !density || !temperature) return GASX_E_ARG;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    const size_t ng = (size_t)sh->gnr * sh->gnz;
    double nmax = 0.0;
    for (size_t i = 0; i < ng; ++i) {
        sh->h_gas[i] = (float)density[i];
        sh->h_gas[ng + i] = (float)temperature[i];
        if ((double)sh->h_gas[i] > nmax) nmax = (double)sh->h_gas[i];
    }
    sh->n_max = nmax;
    HCHK(hipMemcpy(sh->d_dens, sh->h_gas, ng * sizeof(float), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(sh->d_temp, sh->h_gas + ng, ng * sizeof(float), hipMemcpyHostToDevice));
    return GASX_OK;
}

gasx_status gasx_shim_get_gas(gasx_shim* sh, double* density, double* temperature)
{
    if (!sh || !density || !temperature) return GASX_E_ARG;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    const size_t ng = (size_t)sh->gnr * sh->gnz;
    HCHK(hipMemcpy(sh->h_gas, sh->d_dens, ng * sizeof(float), hipMemcpyDeviceToHost));
    HCHK(hipMemcpy(sh->h_gas + ng, sh->d_temp, ng * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ng; ++i) {
        density[i] = sh->h_gas[i];
        temperature[i] = sh->h_gas[ng + i];
    }
    return GASX_OK;
}

gasx_status gasx_shim_photoionize(gasx_shim* sh, double ex_kev, double n_photons, double laser_r, uint64_t seed,
                                  uint32_t pulse, gasx_photo_result* result)
{
    if (!sh || !result) return GASX_E_ARG;
    gasx_photo_source s;
    memset(&s, 0, sizeof s);
    s.photon_kev = ex_kev;
    s.n_photons = n_photons;
    s.beam_radius = laser_r;
    s.binding_ev = sh->binding_ev;
    s.n_shells = sh->n_shells;
    s.mac = sh->mac;
    s.n_mac = sh->n_mac;
    s.atomic_mass = 14.0;          /* m_a, main3.m:27 */
    s.auger_kev = 0.383;           /* INCIDENCE1.m:130 */
    s.seed = seed;
    s.pulse = pulse;
    s.run = 0;
    s.density_out = sh->d_dens;    /* the gas is depleted in place (INCIDENCE1.m:87-92) */
    s.temperature_out = sh->d_temp;
    gasx_species e = species_view(sh, GASX_SHIM_ELECTRONS), i2 = species_view(sh, GASX_SHIM_ION2);
    gasx_gas g = gas_view(sh);
    return gasx_photoionize(sh->ctx, &e, &i2, &g, &s, result);
}

gasx_status gasx_shim_em_step(gasx_shim* sh, int n_push)
{
    if (!sh || n_push < 0) return GASX_E_ARG;
    gasx_species sp[3];
    for (int k = 0; k < 3; ++k) sp[k] = species_view(sh, k);
    gasx_fields f = {sh->d_er, sh->d_ez, sh->d_h};
    return gasx_em_step(sh->ctx, sp, 3, &f, n_push);
}

static gasx_status ensure_cs(gasx_shim* sh, int64_t n)
{
    if (n <= sh->cs_cap) return GASX_OK;
    if (sh->d_cs) hipFree(sh->d_cs);
    free(sh->h_cs);
    sh->d_cs = NULL;
    sh->h_cs = NULL;
    sh->cs_cap = 0;
    HCHK(hipMalloc((void**)&sh->d_cs, (size_t)n * 4 * sizeof(float)));
    sh->h_cs = (float*)malloc((size_t)n * 4 * sizeof(float));
    if (!sh->h_cs) return GASX_E_NOMEM;
    sh->cs_cap = n;
    return GASX_OK;
}

gasx_status gasx_shim_find_cs(gasx_shim* sh, const double* e_kev, int64_t n, double* sig_el, double* sig_ion,
                              double* sig_exc)
{
    if (!sh || n < 0 || (n > 0 && (!e_kev || !sig_el || !sig_ion || !sig_exc))) return GASX_E_ARG;
    if (n == 0) return GASX_OK;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st == GASX_OK) st = ensure_cs(sh, n);
    if (st != GASX_OK) return st;
    for (int64_t i = 0; i < n; ++i) sh->h_cs[i] = (float)e_kev[i];
    HCHK(hipMemcpy(sh->d_cs, sh->h_cs, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    float* d = sh->d_cs;
    st = gasx_cross_sections(sh->ctx, d, n, d + n, d + 2 * n, d + 3 * n);
    if (st == GASX_OK) st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    HCHK(hipMemcpy(sh->h_cs + n, d + n, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) {
        sig_el[i] = sh->h_cs[n + i];
        sig_ion[i] = sh->h_cs[2 * n + i];
        sig_exc[i] = sh->h_cs[3 * n + i];
    }
    return GASX_OK;
}

gasx_status gasx_shim_collide(gasx_shim* sh, double dt, uint64_t seed, uint32_t step)
{
    if (!sh || !(dt > 0)) return GASX_E_ARG;
    gasx_species e = species_view(sh, GASX_SHIM_ELECTRONS), i1 = species_view(sh, GASX_SHIM_ION1);
    gasx_gas g = gas_view(sh);
    return gasx_collide(sh->ctx, &e, &i1, &g, dt, seed, step, 0);
}

gasx_status gasx_shim_electron_kev(gasx_shim* sh, const double** e_kev, int64_t* n)
{
    if (!sh || !e_kev || !n) return GASX_E_ARG;
    const double *vx, *vy, *vz;
    int64_t m;
    gasx_status st = gasx_shim_get_particles(sh, GASX_SHIM_ELECTRONS, NULL, NULL, NULL, &vx, &vy, &vz, &m);
    if (st != GASX_OK) return st;
    free(sh->h_kev);
    sh->h_kev = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    if (!sh->h_kev) return GASX_E_NOMEM;
    const double c = sh->p.c, me = sh->p.me_em, e = sh->p.e_em;
    for (int64_t i = 0; i < m; ++i) {       /* EM4.cu:138-140, electricmagneticfield6.m:172 */
        const double b2 = (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) / (c * c);
        const double g2 = 1.0 / (1.0 - b2), g = sqrt(g2);
        sh->h_kev[i] = me * c * c * (b2 * g2 / (1.0 + g)) / (1000.0 * e);
    }
    *e_kev = sh->h_kev;
    *n = m;
    return GASX_OK;
}

/* ---- DSMC neutral gas (main2.m) --------------------------------------------- */

static gasx_molecules molecules_view(gasx_shim* sh)
{
    gasx_molecules m;
    m.x = sh->dm[0]; m.y = sh->dm[1]; m.z = sh->dm[2];
    m.vx = sh->dm[3]; m.vy = sh->dm[4]; m.vz = sh->dm[5]; m.ei = sh->dm[6];
    m.count = sh->dm_count;
    m.capacity = sh->dm_cap;
    return m;
}

static void free_molecules(gasx_shim* sh)
{
    for (int c = 0; c < 7; ++c) {
        if (sh->dm[c]) hipFree(sh->dm[c]);
        sh->dm[c] = NULL;
    }
    if (sh->dm_count) hipFree(sh->dm_count);
    if (sh->d_samp) hipFree(sh->d_samp);
    sh->dm_count = NULL;
    sh->d_samp = NULL;
    free(sh->h_mol);
    free(sh->h_mol64);
    free(sh->h_samp);
    sh->h_mol = NULL;
    sh->h_mol64 = NULL;
    sh->h_samp = NULL;
    sh->dm_cap = 0;
}

gasx_status gasx_shim_dsmc_load(gasx_shim* sh, const gasx_cell_map* map, const double* x, const double* y,
                                const double* z, const double* vx, const double* vy, const double* vz,
                                const double* ei, int64_t n, int64_t capacity)
{
    if (!sh || !map || n < 0 || capacity < 1 || n > capacity || capacity > INT32_MAX ||
        (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !ei)))
        return GASX_E_ARG;
    int64_t nxy = 0;
    gasx_status st = gasx_cell_map_info(map, NULL, &nxy);
    if (st != GASX_OK) return st;
    st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    const int64_t ncell = nxy * map->n_layers, nrz = (int64_t)map->n_rings * map->n_layers;
    if (capacity != sh->dm_cap || ncell != sh->n_cells || nrz != (int64_t)sh->map.n_rings * sh->map.n_layers) {
        free_molecules(sh);
        int ok = 1;
        for (int c = 0; c < 7 && ok; ++c) ok = hipMalloc((void**)&sh->dm[c], (size_t)capacity * sizeof(float)) == hipSuccess;
        ok = ok && hipMalloc((void**)&sh->dm_count, sizeof(int32_t)) == hipSuccess;
        ok = ok && hipMalloc((void**)&sh->d_samp, (size_t)(2 * ncell + 2 * nrz) * sizeof(float)) == hipSuccess;
        sh->h_mol = (float*)malloc((size_t)capacity * 7 * sizeof(float));
        sh->h_mol64 = (double*)malloc((size_t)capacity * 7 * sizeof(double));
        sh->h_samp = (float*)malloc((size_t)(2 * ncell + 2 * nrz) * sizeof(float));
        sh->dm_cap = capacity;
        if (!ok || !sh->h_mol || !sh->h_mol64 || !sh->h_samp) {
            free_molecules(sh);
            return GASX_E_NOMEM;
        }
    }
    sh->map = *map;
    sh->n_cells = ncell;
    const double* src[7] = {x, y, z, vx, vy, vz, ei};
    for (int c = 0; c < 7; ++c) {
        float* h = sh->h_mol + (size_t)c * capacity;
        for (int64_t i = 0; i < n; ++i) h[i] = (float)src[c][i];
        if (n) HCHK(hipMemcpy(sh->dm[c], h, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    }
    const int32_t cnt = (int32_t)n;
    HCHK(hipMemcpy(sh->dm_count, &cnt, sizeof cnt, hipMemcpyHostToDevice));
    return GASX_OK;
}

gasx_status gasx_shim_dsmc_step(gasx_shim* sh, const gasx_dsmc_params* p, uint64_t seed, uint32_t step)
{
    if (!sh || !p || !sh->dm_count) return GASX_E_ARG;
    gasx_molecules m = molecules_view(sh);
    return gasx_dsmc_step(sh->ctx, &m, &sh->map, p, seed, step);
}

gasx_status gasx_shim_dsmc_sample(gasx_shim* sh, double fn, double mass, double* cell_density,
                                  double* cell_temperature, double* rz_density, double* rz_temperature, int to_gas)
{
    if (!sh || !sh->dm_count || !(fn > 0) || !(mass > 0)) return GASX_E_ARG;
    const int64_t nc = sh->n_cells, nrz = (int64_t)sh->map.n_rings * sh->map.n_layers;
    if (to_gas && (sh->gnr != sh->map.n_rings || sh->gnz != sh->map.n_layers)) return GASX_E_ARG;
    gasx_molecules m = molecules_view(sh);
    float* d = sh->d_samp;
    gasx_status st = gasx_dsmc_sample(sh->ctx, &m, &sh->map, fn, mass, d, d + nc, d + 2 * nc, d + 2 * nc + nrz);
    if (st == GASX_OK) st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    HCHK(hipMemcpy(sh->h_samp, d, (size_t)(2 * nc + 2 * nrz) * sizeof(float), hipMemcpyDeviceToHost));
    const float* h = sh->h_samp;
    for (int64_t i = 0; i < nc; ++i) {
        if (cell_density) cell_density[i] = h[i];
        if (cell_temperature) cell_temperature[i] = h[nc + i];
    }
    for (int64_t i = 0; i < nrz; ++i) {
        if (rz_density) rz_density[i] = h[2 * nc + i];
        if (rz_temperature) rz_temperature[i] = h[2 * nc + nrz + i];
    }
    if (to_gas) {   /* the (ring, layer) gas becomes the electron step's gas (particle_density0.m) */
        double nmax = 0.0;
        for (int64_t i = 0; i < nrz; ++i)
            if ((double)h[2 * nc + i] > nmax) nmax = (double)h[2 * nc + i];
        HCHK(hipMemcpy(sh->d_dens, d + 2 * nc, (size_t)nrz * sizeof(float), hipMemcpyDeviceToDevice));
        HCHK(hipMemcpy(sh->d_temp, d + 2 * nc + nrz, (size_t)nrz * sizeof(float), hipMemcpyDeviceToDevice));
        sh->n_max = nmax;
    }
    return GASX_OK;
}

gasx_status gasx_shim_dsmc_get(gasx_shim* sh, const double** x, const double** y, const double** z,
                               const double** vx, const double** vy, const double** vz, const double** ei,
                               int64_t* n)
{
    if (!sh || !sh->dm_count || !n) return GASX_E_ARG;
    gasx_status st = gasx_synchronize(sh->ctx);
    if (st != GASX_OK) return st;
    int32_t cnt = 0;
    HCHK(hipMemcpy(&cnt, sh->dm_count, sizeof cnt, hipMemcpyDeviceToHost));
    const size_t cp = (size_t)sh->dm_cap;
    for (int c = 0; c < 7; ++c) {
        if (cnt) HCHK(hipMemcpy(sh->h_mol + c * cp, sh->dm[c], (size_t)cnt * sizeof(float), hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < cnt; ++i) sh->h_mol64[c * cp + i] = (double)sh->h_mol[c * cp + i];
    }
    const double* h = sh->h_mol64;
    const double** out[7] = {x, y, z, vx, vy, vz, ei};
    for (int c = 0; c < 7; ++c)
        if (out[c]) *out[c] = h + c * cp;
    *n = cnt;
    return GASX_OK;
}

gasx_status gasx_shim_dsmc_counters(gasx_shim* sh, gasx_dsmc_counters* out, int reset)
{
    if (!sh || !out) return GASX_E_ARG;
    return gasx_dsmc_get_counters(sh->ctx, out, reset);
}

gasx_status gasx_shim_dsmc_sizes(gasx_shim* sh, int64_t* n_cells, int64_t* n_rz)
{
    if (!sh || !sh->dm_count || !n_cells || !n_rz) return GASX_E_ARG;
    *n_cells = sh->n_cells;
    *n_rz = (int64_t)sh->map.n_rings * sh->map.n_layers;
    return GASX_OK;
}

/* ---- electrostatic ion step (ion_staticelectrofield_push.m) ----------------- */

gasx_status gasx_shim_es_push(gasx_shim* sh, int32_t nr, int32_t nz, double delta, double dt, double fn, double t_e,
                              double* phi, double* rho_i, double* rho_e, double* e_r, double* e_z,
                              gasx_es_result* res)
{
    if (!sh || !res || nr < 3 || nz < 3 || !(delta > 0) || !(dt >= 0) || !(fn > 0) || !(t_e > 0)) return GASX_E_ARG;
    const size_t ng = (size_t)nr * nz;
    double* d = NULL;
    HCHK(hipMalloc((void**)&d, 5 * ng * sizeof(double)));
    gasx_species sp[2] = {species_view(sh, GASX_SHIM_ION1), species_view(sh, GASX_SHIM_ION2)};
    gasx_es_params p;
    memset(&p, 0, sizeof p);
    p.nr = nr; p.nz = nz; p.delta = delta; p.phip = sh->p.phip; p.weight = fn; p.t_e = t_e; p.dt = dt;
    p.n_e0 = 1e15;                 /* Poisson.m:33 */
    p.tol = 0.01;                  /* Poisson.m:36 */
    p.max_iter = 1000000;
    gasx_status st = gasx_ion_es_push(sh->ctx, sp, 2, &p, d, d + ng, d + 2 * ng, d + 3 * ng, d + 4 * ng, res);
    if (st == GASX_OK) st = gasx_synchronize(sh->ctx);
    double* out[5] = {phi, rho_i, rho_e, e_r, e_z};
    for (int k = 0; k < 5 && st == GASX_OK; ++k)
        if (out[k] && hipMemcpy(out[k], d + k * ng, ng * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            st = GASX_E_HIP;
    hipFree(d);
    return st;
}
