// One field cycle per launch (electricmagneticfield6.m:146-151):
//   H -= dtf/2 curl E ; n_sub x { H += dtf curl E ; E += dtf (curl H - J)/eps } ; H += dtf/2 curl E
// with the EM3_1.cu:9-15 / EM3.cu:14-42 update rules.
//
// Temporal blocking: a workgroup owns a ti x tk tile of Hphi cells and
// carries a halo of `halo` cells on every side in LDS.  Each half update
// invalidates one more ring of the halo (H reads E at i+1,k+1; E reads H at
// i-1,k-1), so a halo of n_sub+3 keeps the interior exact for the whole
// cycle and the cycle needs no grid-wide synchronisation.  State is FP64
// (read from *_in, written to *_out: overlapping halos make in-place
// updates racy); FP32 copies for the particle gather are written alongside.
// The current comes either from the int64 fixed-point accumulators of the
// deposit (J = junit * Q / V(i)) or from FP64 arrays.
#include "gasx_internal.h"

namespace gasx {

template <int BS, int MAXP>
__global__ __launch_bounds__(BS) void k_fdtd(FdtdArgs a, FdtdRun r0, const FdtdRun* runs, int fixed_j)
{
    extern __shared__ double lds[];
    const int tiles_per_run = a.ntr * a.ntk;
    const int run = blockIdx.x / tiles_per_run;
    const int t = blockIdx.x - run * tiles_per_run;
    FdtdRun R = runs ? runs[run] : r0;
    // a cycle split over two launches: the first writes the mid buffers, the
    // second reads them (a.part: 0 whole cycle, 1 first, 2 second)
    if (a.part == 1) { R.er_out = R.er_mid; R.ez_out = R.ez_mid; R.h_out = R.h_mid; }
    if (a.part == 2) { R.er_in = R.er_mid; R.ez_in = R.ez_mid; R.h_in = R.h_mid; }
    const bool final_part = a.part != 1;
    const int nr = a.nr, nz = a.nz;
    const int ti_idx = t % a.ntr, tk_idx = t / a.ntr;
    const int i0 = ti_idx * a.ti, i1 = min(nr, i0 + a.ti);
    const int k0 = tk_idx * a.tk, k1 = min(nz, k0 + a.tk);
    const int ri0 = max(0, i0 - a.halo), ri1 = min(nr, i1 + a.halo);
    const int rk0 = max(0, k0 - a.halo), rk1 = min(nz, k1 + a.halo);
    const int W = ri1 - ri0, Hh = rk1 - rk0;
    // LDS: H  W x Hh ; Er W x (Hh+1) ; Ez (W+1) x Hh
    double* sH = lds;
    double* sEr = sH + W * Hh;
    double* sEz = sEr + W * (Hh + 1);
    const int ldz = W + 1;
    const int npos = (W + 1) * (Hh + 1);   // positions (i,k), i in [ri0,ri1], k in [rk0,rk1]

    // ---- load ------------------------------------------------------------------
    // Every global read of the thread is issued before any is used (indices
    // clamped into the arrays, so the reads need no guard; a value read for a
    // position that does not use it is dropped below).
    double vH[MAXP], vEr[MAXP], vEz[MAXP], vr0[MAXP], vr1[MAXP], vz0[MAXP], vz1[MAXP];
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
        const int p = min((int)threadIdx.x + q * BS, npos - 1);
        const int li = p % (W + 1), lk = p / (W + 1);
        const int gi = ri0 + li, gk = rk0 + lk;              // gi <= nr, gk <= nz
        const int ci = min(gi, nr - 1), ck = min(gk, nz - 1);
        vH[q] = R.h_in[ci + ck * nr];
        vEr[q] = R.er_in[ci + gk * nr];
        vEz[q] = R.ez_in[gi + ck * (nr + 1)];
        const int km = max(gk - 1, 0), im = max(gi - 1, 0);
        if (fixed_j) {
            vr0[q] = (double)(long long)R.qr[ci + km * nr];
            vr1[q] = (double)(long long)R.qr[ci + ck * nr];
            vz0[q] = (double)(long long)R.qz[min(im, nr - 1) + ck * nr];
            vz1[q] = (double)(long long)R.qz[ci + ck * nr];
        } else {
            vr0[q] = R.jr[ci + km * nr];
            vr1[q] = R.jr[ci + ck * nr];
            vz0[q] = R.jz[min(im, nr - 1) + ck * nr];
            vz1[q] = R.jz[ci + ck * nr];
        }
    }

    // ---- LDS tile, per-position sources and coefficients (registers) ---------
    // J = junit * Q / V(i) for the fixed-point current, as read for FP64 J
    auto jscale = [&](int i) -> double {
        return fixed_j ? a.junit / (a.vcoef * (double)((i + 1) * (i + 1) - i * i)) : 1.0;
    };
    double srcR[MAXP], srcZ[MAXP], cA[MAXP], cB[MAXP];
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
        const int p = threadIdx.x + q * BS;
        srcR[q] = 0.0; srcZ[q] = 0.0; cA[q] = 0.0; cB[q] = 0.0;
        if (p < npos) {
            const int li = p % (W + 1), lk = p / (W + 1);
            const int gi = ri0 + li, gk = rk0 + lk;
            if (li < W && lk < Hh) sH[li + lk * W] = vH[q];
            if (li < W) sEr[li + lk * W] = vEr[q];
            if (lk < Hh) sEz[li + lk * ldz] = vEz[q];
            if (li < W && gk >= 1 && gk <= nz - 1) {   // EM3.cu:29 (jr at k-1 and k)
                const double js = jscale(gi);
                srcR[q] = a.ce * (js * vr0[q] + js * vr1[q]) / 2;
            }
            if (lk < Hh) {
                if (gi == 0)                           // EM3.cu:35
                    srcZ[q] = ((double)a.axis_sign * a.ce) * (jscale(0) * vz1[q]);
                else if (gi <= nr - 1) {               // EM3.cu:39
                    srcZ[q] = a.ce * (jscale(gi - 1) * vz0[q] + jscale(gi) * vz1[q]) / 2;
                    const double rs = (a.ez_metric ? (gi + 1) : gi) * a.dr;
                    cA[q] = a.ce * (1 / (2 * rs) + 1 / a.dr);
                    cB[q] = a.ce * (1 / (2 * rs) - 1 / a.dr);
                }
            }
        }
    }
    __syncthreads();

    // ---- H update (EM3_1.cu:15), scaled by s in {-1/2, 1, 1/2} -------------
    auto h_update = [&](double s) {
        const double cr = s * a.ch_r, cz = s * a.ch_z;
#pragma unroll
        for (int q = 0; q < MAXP; ++q) {
            const int p = threadIdx.x + q * BS;
            if (p < npos) {
                const int li = p % (W + 1), lk = p / (W + 1);
                if (li < W && lk < Hh) {
                    const double dez = sEz[(li + 1) + lk * ldz] - sEz[li + lk * ldz];
                    const double der = sEr[li + (lk + 1) * W] - sEr[li + lk * W];
                    sH[li + lk * W] = sH[li + lk * W] + (cr * dez - cz * der);
                }
            }
        }
        __syncthreads();
    };

    // ---- E update (EM3.cu:14-42), Jacobi in the boundary copies -------------
    auto e_update = [&]() {
        double nEr[MAXP], nEz[MAXP];
#pragma unroll
        for (int q = 0; q < MAXP; ++q) {
            const int p = threadIdx.x + q * BS;
            nEr[q] = 0.0; nEz[q] = 0.0;
            if (p < npos) {
                const int li = p % (W + 1), lk = p / (W + 1);
                const int gi = ri0 + li, gk = rk0 + lk;
                if (li < W) {
                    double v = sEr[li + lk * W];
                    if (gk == 0) v = sEr[li + 1 * W];                     // :22
                    else if (gk == nz) v = sEr[li + (lk - 1) * W];        // :26
                    else if (lk > 0 && lk < Hh)                            // :29
                        v = v - srcR[q] - a.cez * (sH[li + lk * W] - sH[li + (lk - 1) * W]);
                    nEr[q] = v;
                }
                if (lk < Hh) {
                    double v = sEz[li + lk * ldz];
                    if (gi == 0)                                           // :35
                        v = v + srcZ[q] + a.c4 * sH[0 + lk * W];
                    else if (gi <= nr - 1 && li > 0 && li < W)            // :39
                        v = v - srcZ[q] + cA[q] * sH[li + lk * W] + cB[q] * sH[(li - 1) + lk * W];
                    nEz[q] = v;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXP; ++q) {
            const int p = threadIdx.x + q * BS;
            if (p < npos) {
                const int li = p % (W + 1), lk = p / (W + 1);
                if (li < W) sEr[li + lk * W] = nEr[q];
                if (lk < Hh) sEz[li + lk * ldz] = nEz[q];
            }
        }
        __syncthreads();
    };

    const int s_lo = a.part == 2 ? a.split : 0, s_hi = a.part == 1 ? a.split : a.n_sub;
    if (a.part != 2) h_update(-0.5);
    for (int s = s_lo; s < s_hi; ++s) {
        h_update(1.0);
        e_update();
    }
    if (final_part) h_update(0.5);

    // ---- write the tile interior ------------------------------------------
    for (int p = threadIdx.x; p < npos; p += BS) {
        const int li = p % (W + 1), lk = p / (W + 1);
        const int gi = ri0 + li, gk = rk0 + lk;
        const bool in_i = gi >= i0 && gi < i1;
        const bool in_k = gk >= k0 && gk < k1;
        if (li < W && lk < Hh && in_i && in_k) {
            const double v = sH[li + lk * W];
            R.h_out[gi + gk * nr] = v;
            if (final_part) R.h32[gi + gk * nr] = (float)v;
            if (final_part && R.qr_clear) {
                R.qr_clear[gi + gk * nr] = 0ull;
                R.qz_clear[gi + gk * nr] = 0ull;
            }
        }
        if (li < W && in_i && (in_k || (gk == nz && k1 == nz))) {
            const double v = sEr[li + lk * W];
            R.er_out[gi + gk * nr] = v;
            if (final_part) R.er32[gi + gk * nr] = (float)v;
        }
        if (lk < Hh && in_k && (in_i || (gi == nr && i1 == nr))) {
            const double v = sEz[li + lk * ldz];
            R.ez_out[gi + gk * (nr + 1)] = v;
            if (final_part) R.ez32[gi + gk * (nr + 1)] = (float)v;
        }
    }
}

size_t fdtd_lds_bytes(const FdtdArgs& a)
{
    const size_t W = (size_t)a.ti + 2 * a.halo, Hh = (size_t)a.tk + 2 * a.halo;
    return (W * Hh + W * (Hh + 1) + (W + 1) * Hh) * sizeof(double);
}

template <int BS, int MAXP>
static hipError_t launch_fdtd_t(const FdtdArgs& a, const FdtdRun& r0, const FdtdRun* runs_dev,
                                int nruns, int fixed_j, hipStream_t s)
{
    const size_t lds = fdtd_lds_bytes(a);
    const int blocks = nruns * a.ntr * a.ntk;
    hipLaunchKernelGGL((k_fdtd<BS, MAXP>), dim3(blocks), dim3(BS), lds, s, a, r0, runs_dev, fixed_j);
    return hipGetLastError();
}

// ---- transposed field cycle (adjoint of k_fdtd's cycle) -------------------------
// The cycle is linear: F' = M F + B J.  This kernel applies M^T to the field
// sensitivities (lEr, lEz, lH) and returns B^T of them as the current's
// sensitivities (ljr, ljz, overwritten), with the same temporal blocking as
// k_fdtd: the transposed updates run in reverse order -- H(+1/2)^T,
// n_sub x {E^T, H(1)^T}, H(-1/2)^T -- each one reading the neighbours on the
// opposite side to its forward update, so the same halo keeps the interior
// exact.
//   H(s)^T:  lEz(i,k) += s ch_r (lH(i-1,k) - lH(i,k)), lEr(i,k) += s ch_z (lH(i,k) - lH(i,k-1))
//            (terms outside the H grid are absent)
//   E^T:     lH(i,k) += -cez lEr(i,k) + cez lEr(i,k+1)   (Er rows 1..nz-1 only)
//                      + (i == 0 ? c4 : cA(i)) lEz(i,k) + cB(i+1) lEz(i+1,k)   (i+1 <= nr-1)
//            ljr(i,k) -= ce/2 (lEr(i,k) + lEr(i,k+1)) over the same rows,
//            ljz(i,k) += (i == 0 ? axis ce lEz(0,k) : -ce/2 lEz(i,k)) - ce/2 lEz(i+1,k) (i+1 <= nr-1),
//            then the Jacobi boundary copies Er(i,0) <- Er(i,1), Er(i,nz) <- Er(i,nz-1)
//            move the boundary rows' sensitivities onto rows 1 and nz-1.
template <int BS, int MAXP>
__global__ __launch_bounds__(BS) void k_fdtd_adj(FdtdArgs a, AdjFdtdRun R)
{
    extern __shared__ double lds[];
    // split as the forward cycle (a.part 1: H(+1/2)^T and n_sub - split
    // substeps into the mid buffers; 2: the other split substeps and
    // H(-1/2)^T, adding to the current sensitivities of part 1)
    if (a.part == 1) { R.ler_out = R.ler_mid; R.lez_out = R.lez_mid; R.lh_out = R.lh_mid; }
    if (a.part == 2) { R.ler_in = R.ler_mid; R.lez_in = R.lez_mid; R.lh_in = R.lh_mid; }
    const int nr = a.nr, nz = a.nz;
    const int t = blockIdx.x;
    const int ti_idx = t % a.ntr, tk_idx = t / a.ntr;
    const int i0 = ti_idx * a.ti, i1 = min(nr, i0 + a.ti);
    const int k0 = tk_idx * a.tk, k1 = min(nz, k0 + a.tk);
    const int ri0 = max(0, i0 - a.halo), ri1 = min(nr, i1 + a.halo);
    const int rk0 = max(0, k0 - a.halo), rk1 = min(nz, k1 + a.halo);
    const int W = ri1 - ri0, Hh = rk1 - rk0;
    double* sH = lds;
    double* sEr = sH + W * Hh;
    double* sEz = sEr + W * (Hh + 1);
    const int ldz = W + 1;
    const int npos = (W + 1) * (Hh + 1);

    // every global read issued before use (clamped indices, as in k_fdtd)
    double vH[MAXP], vEr[MAXP], vEz[MAXP];
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
        const int p = min((int)threadIdx.x + q * BS, npos - 1);
        const int li = p % (W + 1), lk = p / (W + 1);
        const int gi = ri0 + li, gk = rk0 + lk;              // gi <= nr, gk <= nz
        const int ci = min(gi, nr - 1), ck = min(gk, nz - 1);
        vH[q] = R.lh_in[ci + ck * nr];
        vEr[q] = R.ler_in[ci + gk * nr];
        vEz[q] = R.lez_in[gi + ck * (nr + 1)];
    }
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
        const int p = threadIdx.x + q * BS;
        if (p < npos) {
            const int li = p % (W + 1), lk = p / (W + 1);
            if (li < W && lk < Hh) sH[li + lk * W] = vH[q];
            if (li < W) sEr[li + lk * W] = vEr[q];
            if (lk < Hh) sEz[li + lk * ldz] = vEz[q];
        }
    }
    // per H position: the E^T coefficients and the current accumulators
    double cAe[MAXP], cBe[MAXP], accR[MAXP], accZ[MAXP];
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
        const int p = threadIdx.x + q * BS;
        cAe[q] = 0.0; cBe[q] = 0.0; accR[q] = 0.0; accZ[q] = 0.0;
        if (p < npos) {
            const int gi = ri0 + p % (W + 1);
            auto rs = [&](int i) { return (a.ez_metric ? (i + 1) : i) * a.dr; };
            if (gi == 0) cAe[q] = a.c4;
            else if (gi <= nr - 1) cAe[q] = a.ce * (1 / (2 * rs(gi)) + 1 / a.dr);
            if (gi + 1 <= nr - 1) cBe[q] = a.ce * (1 / (2 * rs(gi + 1)) - 1 / a.dr);
        }
    }
    __syncthreads();

    auto h_adj = [&](double s) {
        const double cr = s * a.ch_r, cz = s * a.ch_z;
#pragma unroll
        for (int q = 0; q < MAXP; ++q) {
            const int p = threadIdx.x + q * BS;
            if (p < npos) {
                const int li = p % (W + 1), lk = p / (W + 1);
                const int gi = ri0 + li, gk = rk0 + lk;
                if (li < W) {
                    double v = 0.0;
                    if (gk < nz && lk < Hh) v += sH[li + lk * W];
                    if (gk >= 1 && lk >= 1) v -= sH[li + (lk - 1) * W];
                    sEr[li + lk * W] += cz * v;
                }
                if (lk < Hh) {
                    double v = 0.0;
                    if (gi >= 1 && li >= 1) v += sH[(li - 1) + lk * W];
                    if (gi < nr && li < W) v -= sH[li + lk * W];
                    sEz[li + lk * ldz] += cr * v;
                }
            }
        }
        __syncthreads();
    };

    auto e_adj = [&]() {
#pragma unroll
        for (int q = 0; q < MAXP; ++q) {
            const int p = threadIdx.x + q * BS;
            if (p < npos) {
                const int li = p % (W + 1), lk = p / (W + 1);
                const int gi = ri0 + li, gk = rk0 + lk;
                if (li < W && lk < Hh) {
                    double h = sH[li + lk * W];
                    if (gk >= 1) {                        // Er row gk is interior (gk <= nz-1 always)
                        const double e = sEr[li + lk * W];
                        h -= a.cez * e;
                        accR[q] -= 0.5 * a.ce * e;
                    }
                    if (gk + 1 <= nz - 1) {
                        const double e = sEr[li + (lk + 1) * W];
                        h += a.cez * e;
                        accR[q] -= 0.5 * a.ce * e;
                    }
                    const double ez0 = sEz[li + lk * ldz];
                    h += cAe[q] * ez0;
                    accZ[q] += (gi == 0 ? (double)a.axis_sign * a.ce : -0.5 * a.ce) * ez0;
                    if (gi + 1 <= nr - 1) {
                        const double ez1 = sEz[(li + 1) + lk * ldz];
                        h += cBe[q] * ez1;
                        accZ[q] -= 0.5 * a.ce * ez1;
                    }
                    sH[li + lk * W] = h;
                }
            }
        }
        __syncthreads();
        // boundary rows of Er onto rows 1 and nz-1 (atomic: with nz = 2 both land on row 1)
#pragma unroll
        for (int q = 0; q < MAXP; ++q) {
            const int p = threadIdx.x + q * BS;
            if (p < npos) {
                const int li = p % (W + 1), lk = p / (W + 1);
                const int gk = rk0 + lk;
                if (li < W && gk == 0 && lk + 1 <= Hh) {
                    const double v = sEr[li + lk * W];
                    sEr[li + lk * W] = 0.0;
                    atomicAdd(&sEr[li + (lk + 1) * W], v);
                }
                if (li < W && gk == nz && lk >= 1) {
                    const double v = sEr[li + lk * W];
                    sEr[li + lk * W] = 0.0;
                    atomicAdd(&sEr[li + (lk - 1) * W], v);
                }
            }
        }
        __syncthreads();
    };

    const int n_here = a.part == 0 ? a.n_sub : (a.part == 1 ? a.n_sub - a.split : a.split);
    if (a.part != 2) h_adj(0.5);
    for (int s = 0; s < n_here; ++s) {
        e_adj();
        h_adj(1.0);
    }
    if (a.part != 1) h_adj(-0.5);

    for (int q = 0; q < MAXP; ++q) {
        const int p = threadIdx.x + q * BS;
        if (p >= npos) continue;
        const int li = p % (W + 1), lk = p / (W + 1);
        const int gi = ri0 + li, gk = rk0 + lk;
        const bool in_i = gi >= i0 && gi < i1;
        const bool in_k = gk >= k0 && gk < k1;
        if (li < W && lk < Hh && in_i && in_k) {
            R.lh_out[gi + gk * nr] = sH[li + lk * W];
            const double pr = a.part == 2 ? R.ljr[gi + gk * nr] : 0.0;   // part 1's share
            const double pz = a.part == 2 ? R.ljz[gi + gk * nr] : 0.0;
            R.ljr[gi + gk * nr] = pr + accR[q];
            R.ljz[gi + gk * nr] = pz + accZ[q];
        }
        if (li < W && in_i && (in_k || (gk == nz && k1 == nz))) R.ler_out[gi + gk * nr] = sEr[li + lk * W];
        if (lk < Hh && in_k && (in_i || (gi == nr && i1 == nr))) R.lez_out[gi + gk * (nr + 1)] = sEz[li + lk * ldz];
    }
}

template <int BS, int MAXP>
static hipError_t launch_fdtd_adj_t(const FdtdArgs& a, const AdjFdtdRun& r, hipStream_t s)
{
    hipLaunchKernelGGL((k_fdtd_adj<BS, MAXP>), dim3(a.ntr * a.ntk), dim3(BS), fdtd_lds_bytes(a), s, a, r);
    return hipGetLastError();
}

static hipError_t launch_fdtd_adj_part(const FdtdArgs& a, const AdjFdtdRun& r, hipStream_t s)
{
    const int W = a.ti + 2 * a.halo, Hh = a.tk + 2 * a.halo;
    const int npos = (W + 1) * (Hh + 1);
    if (fdtd_lds_bytes(a) > 160 * 1024) return hipErrorInvalidValue;
    if (npos <= 512 * 2) return launch_fdtd_adj_t<512, 2>(a, r, s);
    if (npos <= 1024 * 2) return launch_fdtd_adj_t<1024, 2>(a, r, s);
    if (npos <= 1024 * 4) return launch_fdtd_adj_t<1024, 4>(a, r, s);
    if (npos <= 1024 * 8) return launch_fdtd_adj_t<1024, 8>(a, r, s);
    return hipErrorInvalidValue;
}

hipError_t launch_fdtd_adj(const FdtdArgs& a, const AdjFdtdRun& r, hipStream_t s)
{
    if (a.split <= 0 || a.split >= a.n_sub) {
        FdtdArgs w = a;
        w.part = 0;
        w.halo = a.n_sub + 3;
        return launch_fdtd_adj_part(w, r, s);
    }
    FdtdArgs p1 = a, p2 = a;
    p1.part = 1; p1.halo = a.n_sub - a.split + 2;
    p2.part = 2; p2.halo = a.split + 2;
    hipError_t e = launch_fdtd_adj_part(p1, r, s);
    return e == hipSuccess ? launch_fdtd_adj_part(p2, r, s) : e;
}

// Raises the dynamic LDS limit of every field-kernel variant to the CU's
// 160 KB (once per device, at context creation: not a stream operation, so it
// must not wait for the first launch, which may be under graph capture).
hipError_t fdtd_prepare()
{
    const void* fns[] = {(const void*)k_fdtd<512, 2>,     (const void*)k_fdtd<1024, 2>,
                         (const void*)k_fdtd<1024, 4>,    (const void*)k_fdtd<1024, 8>,
                         (const void*)k_fdtd_adj<512, 2>, (const void*)k_fdtd_adj<1024, 2>,
                         (const void*)k_fdtd_adj<1024, 4>, (const void*)k_fdtd_adj<1024, 8>};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static hipError_t launch_fdtd_part(const FdtdArgs& a, const FdtdRun& r0, const FdtdRun* runs_dev, int nruns,
                                   int fixed_j, hipStream_t s)
{
    const int W = a.ti + 2 * a.halo, Hh = a.tk + 2 * a.halo;
    const int npos = (W + 1) * (Hh + 1);
    if (fdtd_lds_bytes(a) > 160 * 1024) return hipErrorInvalidValue;
    // the most threads a tile can use, a few positions each: one workgroup per
    // CU on small grids, so resident waves are what hides the latency
    if (npos <= 512 * 2) return launch_fdtd_t<512, 2>(a, r0, runs_dev, nruns, fixed_j, s);
    if (npos <= 1024 * 2) return launch_fdtd_t<1024, 2>(a, r0, runs_dev, nruns, fixed_j, s);
    if (npos <= 1024 * 4) return launch_fdtd_t<1024, 4>(a, r0, runs_dev, nruns, fixed_j, s);
    if (npos <= 1024 * 8) return launch_fdtd_t<1024, 8>(a, r0, runs_dev, nruns, fixed_j, s);
    return hipErrorInvalidValue;
}

// One field cycle; with 0 < a.split < n_sub, as two launches of split and
// n_sub - split substeps through the mid buffers, each with the halo its own
// substeps need (less redundant halo work for one more pass over the fields).
hipError_t launch_fdtd(const FdtdArgs& a, const FdtdRun& r0, const FdtdRun* runs_dev, int nruns,
                       int fixed_j, hipStream_t s)
{
    if (a.split <= 0 || a.split >= a.n_sub) {
        FdtdArgs w = a;
        w.part = 0;
        w.halo = a.n_sub + 3;
        return launch_fdtd_part(w, r0, runs_dev, nruns, fixed_j, s);
    }
    FdtdArgs p1 = a, p2 = a;
    p1.part = 1; p1.halo = a.split + 2;
    p2.part = 2; p2.halo = a.n_sub - a.split + 2;
    hipError_t e = launch_fdtd_part(p1, r0, runs_dev, nruns, fixed_j, s);
    return e == hipSuccess ? launch_fdtd_part(p2, r0, runs_dev, nruns, fixed_j, s) : e;
}

// ---- fixed point -> A/m^2 ------------------------------------------------------
__global__ void k_j_from_fixed(FdtdArgs a, const unsigned long long* qr,
                               const unsigned long long* qz, double* jr, double* jz)
{
    const int n = a.nr * a.nz;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int i = c % a.nr;
        const double V = a.vcoef * (double)((i + 1) * (i + 1) - i * i);
        jr[c] = a.junit * (double)(long long)qr[c] / V;
        jz[c] = a.junit * (double)(long long)qz[c] / V;
    }
}

hipError_t launch_j_from_fixed(const FdtdArgs& a, const unsigned long long* qr,
                               const unsigned long long* qz, double* jr, double* jz, hipStream_t s)
{
    const int n = a.nr * a.nz;
    int blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_j_from_fixed, dim3(blocks), dim3(256), 0, s, a, qr, qz, jr, jz);
    return hipGetLastError();
}

__global__ void k_to_f32(const double* src, float* dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

hipError_t launch_to_f32(const double* src, float* dst, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

__global__ void k_j_reduce(unsigned long long* qr, unsigned long long* qz, int64_t n, size_t stride)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long a = qr[c], b = qz[c];
#pragma unroll
        for (int k = 1; k < kJCopies; ++k) {
            a += qr[k * stride + c];
            b += qz[k * stride + c];
            qr[k * stride + c] = 0ull;
            qz[k * stride + c] = 0ull;
        }
        qr[c] = a;
        qz[c] = b;
    }
}

hipError_t launch_j_reduce(unsigned long long* qr, unsigned long long* qz, int64_t ncells, size_t stride,
                           hipStream_t s)
{
    if (ncells <= 0 || kJCopies == 1) return hipSuccess;   // one copy: nothing to fold
    int64_t blocks = (ncells + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_j_reduce, dim3((unsigned)blocks), dim3(256), 0, s, qr, qz, ncells, stride);
    return hipGetLastError();
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

hipError_t launch_fill_u64(unsigned long long* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace gasx
