/*
 * gasx_mex.c -- MATLAB MEX gateway: the reference's four GPU-touching
 * functions and main2.m's molecule step on top of gasx_shim.c, with their
 * argument conventions.  Build
 * where MATLAB is installed (its mex.h is needed; this repository's image
 * has no MATLAB):
 *
 *   mex -R2017b gasx_mex.c gasx_shim.c -I../../include -I/opt/rocm/include \
 *       -L../gasx -lgasx -L/opt/rocm/lib -lamdhip64
 *
 * MATLAB side (the bodies of the reference functions become one call each):
 *
 *   h = gasx_mex('create', delta, dt, Rp, Lp, PHIp, FN, data_dir, caps, gas_dr, gas_dz);
 *   [No_photon, E_position, E_velocity, ion2_position, ion2_velocity, dens, temp] = ...
 *       gasx_mex('incidence', h, Ex, No_photon_initial, laser_r, dens, temp, seed, pulse);   % INCIDENCE1
 *   [ep, ev, i1p, i1v, i2p, i2v, Er, Ez, Hphi, ke] = ...
 *       gasx_mex('em_step', h, ep, ev, i1p, i1v, i2p, i2v, Er, Ez, Hphi, cycle_time);        % electricmagneticfield6
 *   [s_el, s_ion, s_exc] = gasx_mex('find_cs', h, E_keV);                                    % find_CS0
 *   [ep, ev, i1p, i1v] = gasx_mex('collide', h, ep, ev, i1p, i1v, dens, temp, dt, seed, step); % electron_collision
 *   [i1p, i1v, i2p, i2v, RHO, rho_e, PHI, T_e, E_r] = gasx_mex('es_push', h, i1p, i1v, i2p, i2v, ...
 *       delta, delta_t, Rp, Lp, phimax, FN, T_e);                                           % ion_staticelectrofield_push
 *   gasx_mex('dsmc_load', h, [n_rings n_layers Rp Lp PHIp delta], position, velocity, ei, capacity); % main2.m:40-106
 *   gasx_mex('dsmc_step', h, dt, FN, seed, step, reorder);                                  % main2.m:250-264
 *   [rho_cell, T_cell, rho_rz, T_rz] = gasx_mex('dsmc_sample', h, FN, to_gas);               % particle_density0
 *   [position, velocity, ei] = gasx_mex('dsmc_get', h);
 *   gasx_mex('destroy', h);
 *
 * Particles are N x 3 doubles (Cartesian positions in the wedge, velocities),
 * fields the reference's NR x (NZ+1), (NR+1) x NZ and NR x NZ arrays, gas
 * density and temperature gnr x gnz (r fastest, one value per ring: the
 * mean over the ring's phi cells of cell_structure.m).
 */
#include "mex.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "gasx_shim.h"

static gasx_shim* handle_of(const mxArray* a)
{
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("gasx:handle", "invalid handle");
    uint64_t v;
    memcpy(&v, mxGetData(a), sizeof v);
    return (gasx_shim*)(uintptr_t)v;
}

static void check(gasx_status s, gasx_shim* sh)
{
    if (s != GASX_OK)
        mexErrMsgIdAndTxt("gasx:status", "%s: %s", gasx_status_string(s),
                          sh ? gasx_last_error(gasx_shim_context(sh)) : "");
}

static double scalar(const mxArray* a) { return mxGetScalar(a); }

/* N x 3 matrix -> species (x, y, z columns of positions, velocities) */
static void put_species(gasx_shim* sh, int k, const mxArray* pos, const mxArray* vel)
{
    const int64_t n = (int64_t)mxGetM(pos);
    if (mxGetN(pos) != 3 || mxGetN(vel) != 3 || (int64_t)mxGetM(vel) != n)
        mexErrMsgIdAndTxt("gasx:shape", "particles must be N x 3");
    const double *p = mxGetPr(pos), *v = mxGetPr(vel);
    check(gasx_shim_set_particles(sh, k, p, p + n, p + 2 * n, v, v + n, v + 2 * n, n), sh);
}

static void get_species(gasx_shim* sh, int k, mxArray** pos, mxArray** vel)
{
    const double *x, *y, *z, *vx, *vy, *vz;
    int64_t n;
    check(gasx_shim_get_particles(sh, k, &x, &y, &z, &vx, &vy, &vz, &n), sh);
    *pos = mxCreateDoubleMatrix((mwSize)n, 3, mxREAL);
    *vel = mxCreateDoubleMatrix((mwSize)n, 3, mxREAL);
    double *p = mxGetPr(*pos), *v = mxGetPr(*vel);
    memcpy(p, x, n * sizeof(double)); memcpy(p + n, y, n * sizeof(double)); memcpy(p + 2 * n, z, n * sizeof(double));
    memcpy(v, vx, n * sizeof(double)); memcpy(v + n, vy, n * sizeof(double)); memcpy(v + 2 * n, vz, n * sizeof(double));
}

static mxArray* copy_of(const mxArray* a)
{
    return mxDuplicateArray(a);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[])
{
    char cmd[32];
    if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd) != 0) mexErrMsgIdAndTxt("gasx:cmd", "command expected");
    if (!strcmp(cmd, "create")) {
        if (nrhs != 11) mexErrMsgIdAndTxt("gasx:args", "create: 10 arguments");
        const double delta = scalar(prhs[1]), dt = scalar(prhs[2]), Rp = scalar(prhs[3]), Lp = scalar(prhs[4]);
        char dir[4096];
        mxGetString(prhs[7], dir, sizeof dir);
        const double* caps = mxGetPr(prhs[8]);
        const int64_t cap[3] = {(int64_t)caps[0], (int64_t)caps[1], (int64_t)caps[2]};
        const double gdr = scalar(prhs[9]), gdz = scalar(prhs[10]);
        gasx_params p;
        memset(&p, 0, sizeof p);
        p.nr = (int32_t)(Rp / delta + 0.5); p.nz = (int32_t)(Lp / delta + 0.5);
        p.dr = delta; p.dz = delta; p.phip = scalar(prhs[5]);
        p.ddt = 5 * dt; p.n_sub = 10;                         /* electricmagneticfield6.m:134-147 */
        p.fn = scalar(prhs[6]);
        p.c = 3e8; p.eps0 = 8.854187817e-12; p.mu0 = 4e-7 * 3.14159265358979323846;
        p.e_em = 1.6e-19; p.me_em = 9.1096e-31; p.e_mc = 1.6022e-19; p.me_mc = 9.109e-31;
        p.m_ion = 4.65e-26; p.k_b = 1.38064853e-23;
        p.ez_metric = 0; p.axis_j_sign = 1; p.exc_form = 0; p.rot_eps = 1e-5;
        gasx_shim* sh = NULL;
        check(gasx_shim_create(0, &p, dir, cap, (int32_t)(Rp / gdr + 0.5), (int32_t)(Lp / gdz + 0.5), gdr, gdz, &sh),
              NULL);
        plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
        const uint64_t v = (uint64_t)(uintptr_t)sh;
        memcpy(mxGetData(plhs[0]), &v, sizeof v);
    } else if (!strcmp(cmd, "incidence")) {
        if (nrhs != 9 || nlhs != 7) mexErrMsgIdAndTxt("gasx:args", "incidence: 8 arguments, 7 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        check(gasx_shim_set_gas(sh, mxGetPr(prhs[5]), mxGetPr(prhs[6])), sh);
        check(gasx_shim_set_particles(sh, GASX_SHIM_ELECTRONS, NULL, NULL, NULL, NULL, NULL, NULL, 0), sh);
        check(gasx_shim_set_particles(sh, GASX_SHIM_ION2, NULL, NULL, NULL, NULL, NULL, NULL, 0), sh);
        gasx_photo_result r;
        gasx_status s = gasx_shim_photoionize(sh, scalar(prhs[2]), scalar(prhs[3]), scalar(prhs[4]),
                                              (uint64_t)scalar(prhs[7]), (uint32_t)scalar(prhs[8]), &r);
        if (s != GASX_OK && s != GASX_E_CAPACITY) check(s, sh);   /* dropped particles are counted in r */
        plhs[0] = mxCreateDoubleScalar(r.photons_left);
        get_species(sh, GASX_SHIM_ELECTRONS, &plhs[1], &plhs[2]);
        get_species(sh, GASX_SHIM_ION2, &plhs[3], &plhs[4]);
        plhs[5] = copy_of(prhs[5]);
        plhs[6] = copy_of(prhs[6]);
        check(gasx_shim_get_gas(sh, mxGetPr(plhs[5]), mxGetPr(plhs[6])), sh);
    } else if (!strcmp(cmd, "em_step")) {
        if (nrhs != 12 || nlhs != 10) mexErrMsgIdAndTxt("gasx:args", "em_step: 11 arguments, 10 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        put_species(sh, GASX_SHIM_ELECTRONS, prhs[2], prhs[3]);
        put_species(sh, GASX_SHIM_ION1, prhs[4], prhs[5]);
        put_species(sh, GASX_SHIM_ION2, prhs[6], prhs[7]);
        check(gasx_shim_set_fields(sh, mxGetPr(prhs[8]), mxGetPr(prhs[9]), mxGetPr(prhs[10])), sh);
        check(gasx_shim_em_step(sh, (int)(scalar(prhs[11]) / 5.0 + 0.5)), sh);   /* cycle / ratio pushes */
        get_species(sh, GASX_SHIM_ELECTRONS, &plhs[0], &plhs[1]);
        get_species(sh, GASX_SHIM_ION1, &plhs[2], &plhs[3]);
        get_species(sh, GASX_SHIM_ION2, &plhs[4], &plhs[5]);
        plhs[6] = copy_of(prhs[8]);
        plhs[7] = copy_of(prhs[9]);
        plhs[8] = copy_of(prhs[10]);
        check(gasx_shim_get_fields(sh, mxGetPr(plhs[6]), mxGetPr(plhs[7]), mxGetPr(plhs[8])), sh);
        const double* ke;
        int64_t n;
        check(gasx_shim_electron_kev(sh, &ke, &n), sh);
        plhs[9] = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
        memcpy(mxGetPr(plhs[9]), ke, n * sizeof(double));
    } else if (!strcmp(cmd, "find_cs")) {
        if (nrhs != 3 || nlhs != 3) mexErrMsgIdAndTxt("gasx:args", "find_cs: 2 arguments, 3 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        const mwSize n = mxGetNumberOfElements(prhs[2]);
        for (int k = 0; k < 3; ++k) plhs[k] = mxCreateDoubleMatrix(n, 1, mxREAL);
        check(gasx_shim_find_cs(sh, mxGetPr(prhs[2]), (int64_t)n, mxGetPr(plhs[0]), mxGetPr(plhs[1]),
                                mxGetPr(plhs[2])),
              sh);
    } else if (!strcmp(cmd, "collide")) {
        if (nrhs != 11 || nlhs != 4) mexErrMsgIdAndTxt("gasx:args", "collide: 10 arguments, 4 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        put_species(sh, GASX_SHIM_ELECTRONS, prhs[2], prhs[3]);
        put_species(sh, GASX_SHIM_ION1, prhs[4], prhs[5]);
        check(gasx_shim_set_gas(sh, mxGetPr(prhs[6]), mxGetPr(prhs[7])), sh);
        gasx_status s = gasx_shim_collide(sh, scalar(prhs[8]), (uint64_t)scalar(prhs[9]), (uint32_t)scalar(prhs[10]));
        check(s, sh);
        get_species(sh, GASX_SHIM_ELECTRONS, &plhs[0], &plhs[1]);
        get_species(sh, GASX_SHIM_ION1, &plhs[2], &plhs[3]);
    } else if (!strcmp(cmd, "es_push")) {
        /* [i1p, i1v, i2p, i2v, RHO, rho_e, PHI, T_e, gridE_r] = gasx_mex('es_push', h, i1p, i1v, i2p, i2v,
         *     delta, delta_t, Rp, Lp, phimax, FN, T_e): ion_staticelectrofield_push.m's signature (its
         *     PHI and RHO inputs only feed an energy printout there) */
        if (nrhs != 13 || nlhs != 9) mexErrMsgIdAndTxt("gasx:args", "es_push: 12 arguments, 9 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        put_species(sh, GASX_SHIM_ION1, prhs[2], prhs[3]);
        put_species(sh, GASX_SHIM_ION2, prhs[4], prhs[5]);
        const double delta = scalar(prhs[6]);
        const int32_t nr = (int32_t)(scalar(prhs[8]) / delta + 0.5), nz = (int32_t)(scalar(prhs[9]) / delta + 0.5);
        const double t_e = scalar(prhs[12]);
        plhs[4] = mxCreateDoubleMatrix((mwSize)nr, (mwSize)nz, mxREAL);
        plhs[5] = mxCreateDoubleMatrix((mwSize)nr, (mwSize)nz, mxREAL);
        plhs[6] = mxCreateDoubleMatrix((mwSize)nr, (mwSize)nz, mxREAL);
        plhs[8] = mxCreateDoubleMatrix((mwSize)nr, (mwSize)nz, mxREAL);
        double* rho = mxGetPr(plhs[4]);
        gasx_es_result r;
        check(gasx_shim_es_push(sh, nr, nz, delta, scalar(prhs[7]), scalar(prhs[11]), t_e, mxGetPr(plhs[6]), rho,
                                mxGetPr(plhs[5]), mxGetPr(plhs[8]), NULL, &r),
              sh);
        const double* re = mxGetPr(plhs[5]);
        for (size_t i = 0; i < (size_t)nr * nz; ++i) rho[i] -= re[i];   /* RHO = rho_i - rho_e (Poisson.m:43) */
        get_species(sh, GASX_SHIM_ION1, &plhs[0], &plhs[1]);
        get_species(sh, GASX_SHIM_ION2, &plhs[2], &plhs[3]);
        plhs[7] = mxCreateDoubleScalar(t_e);
    } else if (!strcmp(cmd, "dsmc_load")) {
        /* main2.m:40-106: molecules (N x 3 positions and velocities, N x 1
         * internal energies) on the cell map of cell_structure.m
         * (map = [n_rings n_layers Rp Lp PHIp delta]), room for capacity */
        if (nrhs != 7) mexErrMsgIdAndTxt("gasx:args", "dsmc_load: 6 arguments");
        gasx_shim* sh = handle_of(prhs[1]);
        if (mxGetNumberOfElements(prhs[2]) != 6) mexErrMsgIdAndTxt("gasx:shape", "map = [n_rings n_layers Rp Lp PHIp delta]");
        const double* mp = mxGetPr(prhs[2]);
        gasx_cell_map map;
        map.n_rings = (int32_t)mp[0]; map.n_layers = (int32_t)mp[1];
        map.rp = mp[2]; map.lp = mp[3]; map.phip = mp[4]; map.delta = mp[5];
        const int64_t n = (int64_t)mxGetM(prhs[3]);
        if (mxGetN(prhs[3]) != 3 || mxGetN(prhs[4]) != 3 || (int64_t)mxGetM(prhs[4]) != n ||
            (int64_t)mxGetNumberOfElements(prhs[5]) != n)
            mexErrMsgIdAndTxt("gasx:shape", "molecules: N x 3 positions, N x 3 velocities, N internal energies");
        const double *x = mxGetPr(prhs[3]), *v = mxGetPr(prhs[4]);
        check(gasx_shim_dsmc_load(sh, &map, x, x + n, x + 2 * n, v, v + n, v + 2 * n, mxGetPr(prhs[5]), n,
                                  (int64_t)scalar(prhs[6])),
              sh);
    } else if (!strcmp(cmd, "dsmc_step")) {
        /* main2.m:250-264 for the resident molecules:
         * gasx_mex('dsmc_step', h, dt, FN, seed, step, reorder) with N2's
         * constants and SCmax of main2.m:5-17,35 */
        if (nrhs != 7) mexErrMsgIdAndTxt("gasx:args", "dsmc_step: 6 arguments");
        gasx_shim* sh = handle_of(prhs[1]);
        const double kb = 1.38064853e-23, m = 4.65e-26;
        gasx_dsmc_params dp;
        memset(&dp, 0, sizeof dp);
        dp.dt = scalar(prhs[2]); dp.fn = scalar(prhs[3]); dp.mass = m; dp.diameter = 3.7e-10;
        dp.sc_max = 15.0 * (3.14159265358979323846 * 4.17e-10 * 4.17e-10) * sqrt(kb * 300.0 / m);
        dp.omega = 0.74; dp.zeta = 2.0; dp.alpha = 1.0; dp.min_attempts = 1; dp.reorder = (int32_t)scalar(prhs[6]);
        check(gasx_shim_dsmc_step(sh, &dp, (uint64_t)scalar(prhs[4]), (uint32_t)scalar(prhs[5])), sh);
    } else if (!strcmp(cmd, "dsmc_sample")) {
        /* [rho_cell, T_cell, rho_rz, T_rz] = gasx_mex('dsmc_sample', h, FN, to_gas):
         * particle_density0.m / count_T_numbensity.m; with to_gas the (ring,
         * layer) values become the electron step's gas */
        if (nrhs != 4 || nlhs != 4) mexErrMsgIdAndTxt("gasx:args", "dsmc_sample: 3 arguments, 4 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        int64_t ncell = 0, nrz = 0;
        check(gasx_shim_dsmc_sizes(sh, &ncell, &nrz), sh);
        plhs[0] = mxCreateDoubleMatrix((mwSize)ncell, 1, mxREAL);
        plhs[1] = mxCreateDoubleMatrix((mwSize)ncell, 1, mxREAL);
        plhs[2] = mxCreateDoubleMatrix((mwSize)nrz, 1, mxREAL);
        plhs[3] = mxCreateDoubleMatrix((mwSize)nrz, 1, mxREAL);
        check(gasx_shim_dsmc_sample(sh, scalar(prhs[2]), 4.65e-26, mxGetPr(plhs[0]), mxGetPr(plhs[1]),
                                    mxGetPr(plhs[2]), mxGetPr(plhs[3]), (int)scalar(prhs[3])),
              sh);
    } else if (!strcmp(cmd, "dsmc_get")) {
        /* [position, velocity, ei] = gasx_mex('dsmc_get', h) */
        if (nrhs != 2 || nlhs != 3) mexErrMsgIdAndTxt("gasx:args", "dsmc_get: 1 argument, 3 outputs");
        gasx_shim* sh = handle_of(prhs[1]);
        const double *x, *y, *z, *vx, *vy, *vz, *ei;
        int64_t n;
        check(gasx_shim_dsmc_get(sh, &x, &y, &z, &vx, &vy, &vz, &ei, &n), sh);
        plhs[0] = mxCreateDoubleMatrix((mwSize)n, 3, mxREAL);
        plhs[1] = mxCreateDoubleMatrix((mwSize)n, 3, mxREAL);
        plhs[2] = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
        double *p = mxGetPr(plhs[0]), *v = mxGetPr(plhs[1]);
        memcpy(p, x, n * sizeof(double)); memcpy(p + n, y, n * sizeof(double)); memcpy(p + 2 * n, z, n * sizeof(double));
        memcpy(v, vx, n * sizeof(double)); memcpy(v + n, vy, n * sizeof(double)); memcpy(v + 2 * n, vz, n * sizeof(double));
        memcpy(mxGetPr(plhs[2]), ei, n * sizeof(double));
    } else if (!strcmp(cmd, "destroy")) {
        if (nrhs != 2) mexErrMsgIdAndTxt("gasx:args", "destroy: 1 argument");
        gasx_shim_destroy(handle_of(prhs[1]));
    } else {
        mexErrMsgIdAndTxt("gasx:cmd", "unknown command %s", cmd);
    }
}
