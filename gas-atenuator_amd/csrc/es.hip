// Electrostatic ion step: the reference's ion push in the field of a
// Boltzmann-electron Poisson solve (ion_staticelectrofield_push.m:21-156 with
// Poisson.m:25-68), on an r-z grid of nr x nz cells of size delta (cell
// centres at ((i + 1/2) delta, (j + 1/2) delta), Poisson.m:11-12).
//
//   k_es_drift_count   per ion: half drift x += dt/2 v (FP64, stored FP32)
//                      and a count in the cell that holds it (NGP)
//   k_es_rho           rho_i = W e sum_s q_s count_s / V(i), V(i) =
//                      PHIp/2 ((i+1)^2 - i^2) delta^3 (integer counts: the
//                      density is independent of the ions' order)
//   k_es_poisson       one workgroup iterates Poisson.m:37-67 to its stopping
//                      rule: a Jacobi sweep of the reference's stencil with
//                      Boltzmann electrons rho_e = e n_e0 exp(e PHI / kB Te),
//                      n_e0 renormalised every sweep to the ions' charge, only
//                      positive updates kept (PHI(g>0) = g(g>0)), the
//                      boundary rows, and max|dPHI| > tol mean(PHI) to go on
//   k_es_field         E = -gradient(PHI) with MATLAB's gradient rule
//                      (central inside, one-sided at the edges)
//   k_es_kick          per ion: E of its cell, v += q e E / m dt along r-hat
//                      of the half-drifted position, x = x_half + dt/2 v
//
// Where the reference's two functions disagree (ion charge on the node grid
// 0:delta:R but the solve on cell centres; a clamp of the linear cell index)
// this file follows Poisson.m's cell-centre grid and clamps r and z cells
// separately; Poisson.m's outer loop repeats one identical solve 100 times
// (PHI restarts from zero and its stopping quantity is never updated), which
// this file does once.
#include <algorithm>
#include <cmath>
#include <string>

#include "context.h"
#include "device_common.h"

using namespace gasx;
using namespace gasx_host;

namespace {

constexpr int kEsBlock = 1024;        // the Poisson workgroup
constexpr int kEsMaxSpecies = 4;
constexpr double kEsE = 1.6e-19;      // Poisson.m:8
constexpr double kEsKb = 1.38064853e-23;   // Poisson.m:2
constexpr double kEsEps = 8.854187817e-12; // Poisson.m:5

struct EsGrid {
    int32_t nr, nz;
    double delta, phip;
};

__device__ __forceinline__ int es_cell(const EsGrid& g, double x, double y, double z)
{
    const double r = sqrt(x * x + y * y);
    int ir = (int)floor(r / g.delta), iz = (int)floor(z / g.delta);
    ir = ir < 0 ? 0 : (ir > g.nr - 1 ? g.nr - 1 : ir);
    iz = iz < 0 ? 0 : (iz > g.nz - 1 ? g.nz - 1 : iz);
    return iz * g.nr + ir;
}

// ion_staticelectrofield_push.m:28,31 (half drift) and :48-66 (cell of the ion)
__global__ void k_es_drift_count(float* x, float* y, float* z, const float* vx, const float* vy, const float* vz,
                                 const int32_t* count, double half_dt, EsGrid g, int32_t* cnt)
{
#pragma clang fp contract(off)
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (is_dead(z[i])) continue;                          // absorbed slot
        const double xh = (double)x[i] + half_dt * (double)vx[i];
        const double yh = (double)y[i] + half_dt * (double)vy[i];
        const double zh = (double)z[i] + half_dt * (double)vz[i];
        x[i] = (float)xh; y[i] = (float)yh; z[i] = (float)zh;
        atomicAdd(&cnt[es_cell(g, (double)(float)xh, (double)(float)yh, (double)(float)zh)], 1);
    }
}

struct RhoArgs {
    EsGrid g;
    int32_t nsp;
    double wq[kEsMaxSpecies];         // W e q_s
    const int32_t* cnt;               // nsp x nr*nz
    double* rho_i;
};

// :68-72 (grid_rho_i ./ (1/2 phi ((R+dr)^2 - R^2) dz))
__global__ void k_es_rho(RhoArgs a)
{
#pragma clang fp contract(off)
    const int N = a.g.nr * a.g.nz;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        const int i = c % a.g.nr;
        double q = 0.0;
        for (int s = 0; s < a.nsp; ++s) q = q + a.wq[s] * (double)a.cnt[(size_t)s * N + c];
        const double d = a.g.delta;
        const double V = 0.5 * a.g.phip * ((double)((i + 1) * (i + 1) - i * i) * d * d) * d;
        a.rho_i[c] = q / V;
    }
}

struct PoissonArgs {
    EsGrid g;
    double te, tol, n_e0;
    int32_t max_iter;
    const double* rho_i;
    double* phi;                      // result
    double* phi_tmp;                  // the other Jacobi buffer
    double* rho_e;
    double* stats;                    // [iterations, converged, n_e0, max|dPHI|, mean PHI, int rho_i R]
};

// trapz weight of index i of n uniform points (spacing 1)
__device__ __forceinline__ double trap_w1(int i, int n) { return (i == 0 || i == n - 1) ? 0.5 : 1.0; }

__device__ double es_block_sum(double v, double* s_red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kEsBlock / 64; ++w) t += s_red[w];
    __syncthreads();                   // s_red is reused by the next reduction
    return t;
}

__device__ double es_block_max(double v, double* s_red)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kEsBlock / 64; ++w) t = fmax(t, s_red[w]);
    __syncthreads();
    return t;
}

// Poisson.m:37-67 in one workgroup (the sweeps need a grid-wide result,
// n_e0 and the stopping quantity, after each).  Jacobi: every sweep reads the
// previous PHI only.  Sums run in a fixed order, so the solve is
// deterministic.
__global__ __launch_bounds__(kEsBlock) void k_es_poisson(PoissonArgs a)
{
#pragma clang fp contract(off)
    __shared__ double s_red[kEsBlock / 64];
    __shared__ int s_go;
    const int nr = a.g.nr, nz = a.g.nz, N = nr * nz;
    const double d = a.g.delta, d2 = d * d;
    const double den = 2 / d2 + 2 / d2;          // (2/delta_r^2 + 2/delta_z^2)
    const double beta = kEsE / (kEsKb * a.te);
    // int_rho_i = trapz(gridz, trapz(gridr, rho_i .* R))          :22
    double part = 0.0;
    for (int c = threadIdx.x; c < N; c += kEsBlock) {
        const int i = c % nr, j = c / nr;
        part += trap_w1(i, nr) * trap_w1(j, nz) * (a.rho_i[c] * ((i + 0.5) * d));
    }
    const double int_rho_i = es_block_sum(part, s_red) * d2;
    double* cur = a.phi;
    double* nxt = a.phi_tmp;
    for (int c = threadIdx.x; c < N; c += kEsBlock) {   // PHI = 0, rho_e = 0 (:28-31)
        cur[c] = 0.0;
        a.rho_e[c] = 0.0;
    }
    __syncthreads();
    double n_e0 = a.n_e0;
    int k = 0;
    double maxd = 0.0, mean = 0.0;
    for (;;) {
        ++k;
        // the sweep (:41-56) and PHI(g>0) = g(g>0), PHI(lenr,:) = 0,
        // PHI(:,1) = PHI(:,lenz) = 0 (:58-62); row 0 after the barrier
        double se = 0.0;
        for (int c = threadIdx.x; c < N; c += kEsBlock) {
            const int i = c % nr, j = c / nr;
            if (i == 0) continue;
            double nv = 0.0;
            if (i < nr - 1 && j > 0 && j < nz - 1) {
                const double p = cur[c];
                const double re = kEsE * n_e0 * exp(beta * p);
                a.rho_e[c] = re;
                const double RHO = a.rho_i[c] - re;
                const double R = (i + 0.5) * d;
                const double g = (RHO / kEsEps + (cur[c + 1] + cur[c - 1]) / d2 +
                                  1 / (R + d) * (cur[c + 1] - cur[c - 1]) / (2 * d) +
                                  (cur[c - nr] + cur[c + nr]) / d2) / den;
                nv = g > 0 ? g : p;
                se += trap_w1(i, nr) * trap_w1(j, nz) * (re * R);
            }
            nxt[c] = nv;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < nz; j += kEsBlock)   // PHI(1,:) = PHI(2,:), then the z ends
            nxt[j * nr] = (j == 0 || j == nz - 1) ? 0.0 : nxt[j * nr + 1];
        __syncthreads();
        double md = 0.0, sp = 0.0;
        for (int c = threadIdx.x; c < N; c += kEsBlock) {
            md = fmax(md, fabs(cur[c] - nxt[c]));
            sp += nxt[c];
        }
        const double se_tot = es_block_sum(se, s_red) * d2;
        maxd = es_block_max(md, s_red);
        mean = es_block_sum(sp, s_red) / (double)N;
        // n_e0 = n_e0 * int_rho_i / trapz(gridz, trapz(gridr, rho_e .* R))   :57
        n_e0 = n_e0 * int_rho_i / se_tot;
        if (threadIdx.x == 0) s_go = (maxd > a.tol * mean && k < a.max_iter) ? 1 : 0;
        double* t = cur; cur = nxt; nxt = t;
        __syncthreads();
        if (!s_go) break;
    }
    if (cur != a.phi)
        for (int c = threadIdx.x; c < N; c += kEsBlock) a.phi[c] = cur[c];
    if (threadIdx.x == 0) {
        a.stats[0] = (double)k;
        a.stats[1] = maxd > a.tol * mean ? 0.0 : 1.0;
        a.stats[2] = n_e0;
        a.stats[3] = maxd;
        a.stats[4] = mean;
        a.stats[5] = int_rho_i;
    }
}

// ion_staticelectrofield_push.m:70-73 via Poisson.m:69-71: [E_z, E_r] =
// gradient(PHI, dz, dr), E = -(E_r, E_z)
__global__ void k_es_field(EsGrid g, const double* phi, double* er, double* ez)
{
#pragma clang fp contract(off)
    const int nr = g.nr, nz = g.nz, N = nr * nz;
    const double d = g.delta;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        const int i = c % nr, j = c / nr;
        const double gr = i == 0 ? (phi[c + 1] - phi[c]) / d
                         : i == nr - 1 ? (phi[c] - phi[c - 1]) / d
                                       : (phi[c + 1] - phi[c - 1]) / (2 * d);
        const double gz = j == 0 ? (phi[c + nr] - phi[c]) / d
                         : j == nz - 1 ? (phi[c] - phi[c - nr]) / d
                                       : (phi[c + nr] - phi[c - nr]) / (2 * d);
        er[c] = -gr;
        ez[c] = -gz;
    }
}

// :112-147: the force q e E of the ion's cell along r-hat and z-hat of its
// half-drifted position, v += F / m dt, x = x_half + dt/2 v (an ion on the
// axis, where the reference divides by r = 0, is pushed along x)
__global__ void k_es_kick(float* x, float* y, float* z, float* vx, float* vy, float* vz, const int32_t* count,
                          EsGrid g, const double* er, const double* ez, double qe_over_m, double dt)
{
#pragma clang fp contract(off)
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (is_dead(z[i])) continue;
        const double xh = x[i], yh = y[i], zh = z[i];
        const int c = es_cell(g, xh, yh, zh);
        const double r = sqrt(xh * xh + yh * yh);
        const double ct = r > 0 ? xh / r : 1.0, st = r > 0 ? yh / r : 0.0;
        const double fr = qe_over_m * er[c], fz = qe_over_m * ez[c];
        const double nvx = (double)vx[i] + fr * ct * dt;
        const double nvy = (double)vy[i] + fr * st * dt;
        const double nvz = (double)vz[i] + fz * dt;
        vx[i] = (float)nvx; vy[i] = (float)nvy; vz[i] = (float)nvz;
        x[i] = (float)(xh + 0.5 * dt * nvx);
        y[i] = (float)(yh + 0.5 * dt * nvy);
        z[i] = (float)(zh + 0.5 * dt * nvz);
    }
}

}  // namespace

extern "C" gasx_status gasx_ion_es_push(gasx_ctx* ctx, gasx_species* ions, int n_species, const gasx_es_params* p,
                                        double* phi, double* rho_i, double* rho_e, double* e_r, double* e_z,
                                        gasx_es_result* out)
{
    if (!ctx || !ions || !p || !phi || !rho_i || !rho_e || !e_r || !e_z || !out) return GASX_E_ARG;
    if (n_species < 1 || n_species > kEsMaxSpecies)
        return fail(ctx, GASX_E_ARG, "gasx_ion_es_push: 1 to 4 ion species");
    for (int s = 0; s < n_species; ++s) {
        const gasx_species& sp = ions[s];
        if (!sp.x || !sp.y || !sp.z || !sp.vx || !sp.vy || !sp.vz || !sp.count || sp.capacity < 1 ||
            sp.capacity > (int64_t)INT32_MAX || !(sp.mass > 0) || !(sp.qnum > 0))
            return fail(ctx, GASX_E_ARG, "gasx_ion_es_push: invalid ion species " + std::to_string(s));
    }
    if (p->nr < 3 || p->nz < 3 || (int64_t)p->nr * p->nz > (int64_t)1 << 24 || !(p->delta > 0) ||
        !(p->phip > 0) || !(p->weight > 0) || !(p->t_e > 0) || !std::isfinite(p->dt) || !(p->n_e0 > 0) ||
        !(p->tol >= 0) || p->max_iter < 1)
        return fail(ctx, GASX_E_ARG, "gasx_ion_es_push: invalid parameters");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const EsGrid g{p->nr, p->nz, p->delta, p->phip};
    const size_t N = (size_t)p->nr * p->nz;
    const size_t cnt_bytes = (size_t)n_species * N * sizeof(int32_t);
    const size_t bytes = ((cnt_bytes + 255) & ~(size_t)255) + N * sizeof(double) + 8 * sizeof(double);
    char* w = nullptr;
    HIPCHK(hipMalloc((void**)&w, bytes), "alloc es buffers");
    int32_t* cnt = (int32_t*)w;
    double* phi_tmp = (double*)(w + ((cnt_bytes + 255) & ~(size_t)255));
    double* stats = phi_tmp + N;
    double hs[8] = {};
    hipError_t err = hipMemsetAsync(cnt, 0, cnt_bytes, ctx->stream);
    for (int s = 0; s < n_species && err == hipSuccess; ++s) {
        const gasx_species& sp = ions[s];
        const int blocks = (int)std::min<int64_t>(4096, (sp.capacity + 255) / 256);
        hipLaunchKernelGGL(k_es_drift_count, dim3(blocks), dim3(256), 0, ctx->stream, sp.x, sp.y, sp.z, sp.vx,
                           sp.vy, sp.vz, sp.count, 0.5 * p->dt, g, cnt + (size_t)s * N);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        RhoArgs ra{};
        ra.g = g;
        ra.nsp = n_species;
        for (int s = 0; s < n_species; ++s) ra.wq[s] = p->weight * kEsE * (double)ions[s].qnum;
        ra.cnt = cnt;
        ra.rho_i = rho_i;
        hipLaunchKernelGGL(k_es_rho, dim3((unsigned)std::min<size_t>(1024, (N + 255) / 256)), dim3(256), 0,
                           ctx->stream, ra);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        PoissonArgs pa{};
        pa.g = g;
        pa.te = p->t_e; pa.tol = p->tol; pa.n_e0 = p->n_e0; pa.max_iter = p->max_iter;
        pa.rho_i = rho_i; pa.phi = phi; pa.phi_tmp = phi_tmp; pa.rho_e = rho_e; pa.stats = stats;
        hipLaunchKernelGGL(k_es_poisson, dim3(1), dim3(kEsBlock), 0, ctx->stream, pa);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_es_field, dim3((unsigned)std::min<size_t>(1024, (N + 255) / 256)), dim3(256), 0,
                           ctx->stream, g, (const double*)phi, e_r, e_z);
        err = hipGetLastError();
    }
    for (int s = 0; s < n_species && err == hipSuccess; ++s) {
        const gasx_species& sp = ions[s];
        const int blocks = (int)std::min<int64_t>(4096, (sp.capacity + 255) / 256);
        hipLaunchKernelGGL(k_es_kick, dim3(blocks), dim3(256), 0, ctx->stream, sp.x, sp.y, sp.z, sp.vx, sp.vy,
                           sp.vz, sp.count, g, (const double*)e_r, (const double*)e_z,
                           (double)sp.qnum * kEsE / sp.mass, p->dt);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpyAsync(hs, stats, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    hipFree(w);
    if (err != hipSuccess) return hip_fail(ctx, err, "electrostatic ion step");
    out->iterations = (int32_t)hs[0];
    out->converged = (int32_t)hs[1];
    out->n_e0 = hs[2];
    out->max_dphi = hs[3];
    out->mean_phi = hs[4];
    out->int_rho_i = hs[5];
    return GASX_OK;
}
