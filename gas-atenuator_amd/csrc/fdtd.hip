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

// Update kinds of a thread's positions, decoded once per launch.
// Er: flags (update; write the Jacobi copy into the row below / above)
enum : unsigned char { kErUpdate = 1, kErCopyBelow = 2, kErCopyAbove = 4 };
enum : unsigned char { kEzKeep = 0, kEzAxis, kEzUpdate };

// Workgroup b runs on XCD b mod 8 (round-robin dispatch); tile b' of the
// grid is given to it such that each XCD takes a contiguous eighth of the
// tiles, whose overlapping halos then meet in that XCD's L2.  A bijection of
// [0, G): the last G mod 8 workgroups keep their own index.
__device__ __forceinline__ int xcd_tile(int b, int G)
{
    const int per = G >> 3;
    return b < (per << 3) ? (b & 7) * per + (b >> 3) : b;
}

// Thread layout: a thread owns a strip of K consecutive positions (li, lk) of
// one column li (positions i in [ri0, ri1], k in [rk0, rk1]; a position
// holds H, Er and Ez of that index where they exist).  It keeps their values
// in registers and publishes every update to LDS for its neighbours, so a
// half update reads from LDS only the neighbours outside the strip: Ez(i+1)
// and the Er above the strip for H; the H below the strip and H(i-1) for E.
// Strips start at multiples of K >= 2, so rows 0 and 1 are always one
// thread's (the Jacobi copy Er(0) <- Er(1) stays in its registers).
template <int BS, int K>
__global__ __launch_bounds__(BS) void k_fdtd(FdtdArgs a, FdtdRun r0, const FdtdRun* runs, int fixed_j)
{
    static_assert(K >= 2, "rows 0 and 1 share a strip");
    extern __shared__ double lds[];
    const int tiles_per_run = a.ntr * a.ntk;
    const int bt = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int run = bt / tiles_per_run;
    const int t = bt - run * tiles_per_run;
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
    // LDS: H  W x Hh ; Er W x (Hh+1) ; Ez (W+1) x Hh ; Jr, Jz  W x Hh (the
    // current as read: fixed point converted to double, or FP64)
    double* sH = lds;
    double* sEr = sH + W * Hh;
    double* sEz = sEr + W * (Hh + 1);
    double* sJr = sEz + (W + 1) * Hh;
    double* sJz = sJr + W * Hh;
    const int ldz = W + 1;
    const int ncol = W + 1, nstrip = (Hh + 1 + K - 1) / K;
    const int tid = (int)threadIdx.x;
    const bool active = tid < ncol * nstrip;
    const int li = active ? tid % ncol : 0;
    const int sk = active ? (tid / ncol) * K : 0;       // the strip's first row
    const int gi = ri0 + li;                            // <= nr
    const int ci = min(gi, nr - 1);

    // ---- load ------------------------------------------------------------------
    // Every global read of the thread is issued before any is used (indices
    // clamped into the arrays, so the reads need no guard; a value read for a
    // position that does not use it is dropped below).  The current is read
    // once per cell; the E updates take their neighbours' (k-1 for Er, i-1 for
    // Ez, inside the tile for every updated position) from LDS.
    double H[K], Er[K], Ez[K], vJr[K], vJz[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int lk = min(sk + j, Hh);
        const int gk = rk0 + lk;                        // <= nz
        const int ck = min(gk, nz - 1);
        H[j] = R.h_in[ci + ck * nr];
        Er[j] = R.er_in[ci + gk * nr];
        Ez[j] = R.ez_in[gi + ck * (nr + 1)];
        if (fixed_j) {
            vJr[j] = (double)(long long)R.qr[ci + ck * nr];
            vJz[j] = (double)(long long)R.qz[ci + ck * nr];
        } else {
            vJr[j] = R.jr[ci + ck * nr];
            vJz[j] = R.jz[ci + ck * nr];
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int lk = sk + j;
        if (!active || lk > Hh) continue;
        if (li < W && lk < Hh) {
            sH[li + lk * W] = H[j];
            sJr[li + lk * W] = vJr[j];
            sJz[li + lk * W] = vJz[j];
        }
        if (li < W) sEr[li + lk * W] = Er[j];
        if (lk < Hh) sEz[li + lk * ldz] = Ez[j];
    }
    __syncthreads();

    // ---- kinds, sources and coefficients (registers) ----------------------
    // J = junit * Q / V(i) for the fixed-point current, as read for FP64 J
    auto jscale = [&](int i) -> double {
        return fixed_j ? a.junit / (a.vcoef * (double)((i + 1) * (i + 1) - i * i)) : 1.0;
    };
    // Ez: the axis column (EM3.cu:35) and the interior columns (:39) are
    // column properties; the rest (halo edge columns, i = nr) keep their values
    const unsigned char ezcol = (gi == 0) ? kEzAxis : (gi <= nr - 1 && li > 0 && li < W) ? kEzUpdate : kEzKeep;
    double cA = 0.0, cB = 0.0;
    if (ezcol == kEzUpdate) {
        const double rs = (a.ez_metric ? (gi + 1) : gi) * a.dr;
        cA = a.ce * (1 / (2 * rs) + 1 / a.dr);
        cB = a.ce * (1 / (2 * rs) - 1 / a.dr);
    }
    // Er rows gk = 0 and nz take the Jacobi copies of rows 1 and nz-1
    // (EM3.cu:22,26): the owner of rows 1 and nz-1 writes them, from the
    // values before the update, so the E update reads no other thread's E
    // and runs in place between two barriers.
    bool hk[K];
    unsigned char erk[K], ezk[K];
    double srcR[K], srcZ[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int lk = sk + j;
        const int gk = rk0 + lk;
        const bool valid = active && lk <= Hh;
        hk[j] = valid && li < W && lk < Hh;
        erk[j] = 0; ezk[j] = kEzKeep; srcR[j] = 0.0; srcZ[j] = 0.0;
        if (!valid) continue;
        if (li < W && gk >= 1 && gk <= nz - 1 && lk > 0 && lk < Hh) {   // EM3.cu:29 (jr at k-1 and k)
            const double js = jscale(gi);
            srcR[j] = a.ce * (js * sJr[li + (lk - 1) * W] + js * sJr[li + lk * W]) / 2;
            erk[j] = kErUpdate | ((gk == 1 && lk == 1) ? kErCopyBelow : 0) |
                     ((gk == nz - 1 && lk + 1 == Hh) ? kErCopyAbove : 0);
        }
        if (lk < Hh) {
            ezk[j] = ezcol;
            if (ezcol == kEzAxis)
                srcZ[j] = ((double)a.axis_sign * a.ce) * (jscale(0) * sJz[li + lk * W]);
            else if (ezcol == kEzUpdate)
                srcZ[j] = a.ce * (jscale(gi - 1) * sJz[(li - 1) + lk * W] + jscale(gi) * sJz[li + lk * W]) / 2;
        }
    }
    // (no barrier: the first H update reads only E, in place since the load)

    // ---- H update (EM3_1.cu:15), scaled by s in {-1/2, 1, 1/2} -------------
    // reads E, writes the thread's own H cells.  Er(k+1) inside the strip is
    // the thread's register: rows it does not update are halo edges (never
    // changed) or the boundary row nz when the thread wrote its copy.
    // Each half update issues all its LDS reads first (addresses clamped into
    // the tile; a value read for a position that does not use it is
    // dropped), then computes and writes: the reads of the strip's positions
    // then wait on one LDS round trip, not one per position (the compiler
    // cannot move a read above a write to the same LDS array).
    auto h_update = [&](double s) {
        const double cr = s * a.ch_r, cz = s * a.ch_z;
        double ez1[K];
#pragma unroll
        for (int j = 0; j < K; ++j) ez1[j] = sEz[min(li + 1, W) + min(sk + j, Hh - 1) * ldz];
        const double ertop = sEr[min(li, W - 1) + min(sk + K, Hh) * W];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (hk[j]) {
                const double er1 = j + 1 < K ? Er[j + 1] : ertop;
                const double dez = ez1[j] - Ez[j];
                const double der = er1 - Er[j];
                H[j] = H[j] + (cr * dez - cz * der);
                sH[li + (sk + j) * W] = H[j];
            }
        }
        __syncthreads();
    };

    // ---- E update (EM3.cu:14-42): reads H, writes the thread's own E --------
    auto e_update = [&]() {
        double hl[K];
#pragma unroll
        for (int j = 0; j < K; ++j) hl[j] = sH[max(li - 1, 0) + min(sk + j, Hh - 1) * W];
        const double hbot = sH[min(li, W - 1) + max(sk - 1, 0) * W];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int lk = sk + j;
            if (erk[j] & kErUpdate) {                                      // :29
                const double hb = j > 0 ? H[j - 1] : hbot;
                const double old = Er[j];
                Er[j] = old - srcR[j] - a.cez * (H[j] - hb);
                sEr[li + lk * W] = Er[j];
                if (erk[j] & kErCopyBelow) {                               // :22 Er(0) <- Er(1)
                    sEr[li + (lk - 1) * W] = old;
                    if (j > 0) Er[j - 1] = old;
                }
                if (erk[j] & kErCopyAbove) {                               // :26 Er(nz) <- Er(nz-1)
                    sEr[li + (lk + 1) * W] = old;
                    if (j + 1 < K) Er[j + 1] = old;
                }
            }
            if (ezk[j] == kEzAxis) {                                       // :35 (li = 0: H(0) is the thread's)
                Ez[j] = Ez[j] + srcZ[j] + a.c4 * H[j];
                sEz[li + lk * ldz] = Ez[j];
            } else if (ezk[j] == kEzUpdate) {                              // :39
                Ez[j] = Ez[j] - srcZ[j] + cA * H[j] + cB * hl[j];
                sEz[li + lk * ldz] = Ez[j];
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

    // ---- write the tile interior (from LDS, where every update landed) ------
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int lk = sk + j;
        if (!active || lk > Hh) continue;
        const int gk = rk0 + lk;
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

// the forward kernel also holds the tile's current
size_t fdtd_fwd_lds_bytes(const FdtdArgs& a)
{
    const size_t W = (size_t)a.ti + 2 * a.halo, Hh = (size_t)a.tk + 2 * a.halo;
    return fdtd_lds_bytes(a) + 2 * W * Hh * sizeof(double);
}

template <int BS, int K>
static hipError_t launch_fdtd_t(const FdtdArgs& a, const FdtdRun& r0, const FdtdRun* runs_dev,
                                int nruns, int fixed_j, hipStream_t s)
{
    const size_t lds = fdtd_fwd_lds_bytes(a);
    const int blocks = nruns * a.ntr * a.ntk;
    hipLaunchKernelGGL((k_fdtd<BS, K>), dim3(blocks), dim3(BS), lds, s, a, r0, runs_dev, fixed_j);
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
    const void* fns[] = {(const void*)k_fdtd<256, 2>,     (const void*)k_fdtd<512, 2>,
                         (const void*)k_fdtd<768, 2>,     (const void*)k_fdtd<1024, 2>,
                         (const void*)k_fdtd<256, 3>,     (const void*)k_fdtd<512, 3>,
                         (const void*)k_fdtd<768, 3>,     (const void*)k_fdtd<1024, 3>,
                         (const void*)k_fdtd<256, 4>,     (const void*)k_fdtd<512, 4>,
                         (const void*)k_fdtd<768, 4>,     (const void*)k_fdtd<1024, 4>,
                         (const void*)k_fdtd<1024, 6>,    (const void*)k_fdtd<1024, 8>,
                         (const void*)k_fdtd_adj<512, 2>, (const void*)k_fdtd_adj<1024, 2>,
                         (const void*)k_fdtd_adj<1024, 4>, (const void*)k_fdtd_adj<1024, 8>};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Strip length K and workgroup size of the forward kernel for a's tiles:
// false when no instantiation holds them.
#ifndef GASX_FDTD_MIN_K
#define GASX_FDTD_MIN_K 2
#endif
bool fdtd_fwd_shape(const FdtdArgs& a, int* K, int* BS)
{
    const int W = a.ti + 2 * a.halo, Hh = a.tk + 2 * a.halo;
    static const int ks[] = {2, 3, 4, 6, 8};
    for (int k : ks) {
        if (k < GASX_FDTD_MIN_K) continue;
        const int threads = (W + 1) * ((Hh + 1 + k - 1) / k);
        if (threads > 1024) continue;
        const int bs = k > 4 ? 1024 : threads <= 256 ? 256 : threads <= 512 ? 512 : threads <= 768 ? 768 : 1024;
        *K = k; *BS = bs;
        return true;
    }
    return false;
}

static hipError_t launch_fdtd_part(const FdtdArgs& a, const FdtdRun& r0, const FdtdRun* runs_dev, int nruns,
                                   int fixed_j, hipStream_t s)
{
    if (fdtd_fwd_lds_bytes(a) > 160 * 1024) return hipErrorInvalidValue;
    // the shortest strips whose threads fit a workgroup (the most threads: one
    // workgroup per CU on small grids, so resident waves hide the latency),
    // in the smallest workgroup size that holds them
    int K = 0, BS = 0;
    if (!fdtd_fwd_shape(a, &K, &BS)) return hipErrorInvalidValue;
#define GASX_FDTD_CASE(bs, k) \
    if (BS == bs && K == k) return launch_fdtd_t<bs, k>(a, r0, runs_dev, nruns, fixed_j, s);
    GASX_FDTD_CASE(256, 2) GASX_FDTD_CASE(512, 2) GASX_FDTD_CASE(768, 2) GASX_FDTD_CASE(1024, 2)
    GASX_FDTD_CASE(256, 3) GASX_FDTD_CASE(512, 3) GASX_FDTD_CASE(768, 3) GASX_FDTD_CASE(1024, 3)
    GASX_FDTD_CASE(256, 4) GASX_FDTD_CASE(512, 4) GASX_FDTD_CASE(768, 4) GASX_FDTD_CASE(1024, 4)
    GASX_FDTD_CASE(1024, 6) GASX_FDTD_CASE(1024, 8)
#undef GASX_FDTD_CASE
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
