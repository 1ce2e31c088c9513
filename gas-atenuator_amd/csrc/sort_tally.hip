// Cell sort (new, no reference counterpart) and the FP64 energy tallies of
// electricmagneticfield5.m:167-177.
#include <rocprim/device/device_radix_sort.hpp>

#include "device_common.h"

namespace gasx {

// ---- cell sort ----------------------------------------------------------------
// key = k0*nr + i0, the first cell of the particle's EM7 deposit stencil
// (cell-centred axes, clamped at the edges), so that neighbouring slots add
// into the same four cells and the particle kernel can sum them before its
// atomics; absorbed slots get key nr*nz and sort last.  Stable radix sort
// keeps slot order within a key.
__device__ __forceinline__ uint32_t cell_key(const Geom& g, float x, float y, float z)
{
#pragma clang fp contract(off)
    if (is_dead(z)) return (uint32_t)(g.nr * g.nz);
    float r2 = x * x;
    r2 = r2 + y * y;
    const float r = sqrtf(r2);
    float sr = r * g.inv_dr;
    sr = sr - 0.5f;
    float sz = z * g.inv_dz;
    sz = sz - 0.5f;
    int ir = (int)floorf(sr), iz = (int)floorf(sz);
    ir = ir < 0 ? 0 : (ir >= g.nr ? g.nr - 1 : ir);
    iz = iz < 0 ? 0 : (iz >= g.nz ? g.nz - 1 : iz);
    return (uint32_t)(iz * g.nr + ir);
}

// Four consecutive slots per thread: float4 loads where the arrays allow,
// uint4 stores into the (256-byte aligned) sort scratch.  The slot count is
// read on the device (info[0] <- n).  Workgroup b writes to info[2 + b] the
// descents of the key sequence among its slots (slot i in [1, n) with
// key[i] < key[i-1]): their sum tells the host how far from cell order the
// slots are.  No atomics: every workgroup writes its own word.
constexpr int kSortKeyBlocks = kSortInfoWords - 2;

__global__ __launch_bounds__(256) void k_sort_keys(Geom g, const float* x, const float* y, const float* z,
                                                   const int32_t* count, int32_t cap, uint32_t* keys,
                                                   uint32_t* vals, int32_t* info)
{
    const int n = *count < cap ? *count : cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) info[0] = n;
    const bool aligned = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y) |
                           reinterpret_cast<uintptr_t>(z)) & 15) == 0;
    const int lane = threadIdx.x & 63;
    const int ng = (n + 3) >> 2;
    int desc = 0;
    // a wave covers 64 consecutive groups of four slots, so all its lanes
    // take the same number of iterations except in the last one, where the
    // active lanes are a prefix of the wave
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += gridDim.x * blockDim.x) {
        const int i0 = gi * 4;
        uint32_t k[4];
        if (aligned && i0 + 4 <= n) {
            const float4 a = *reinterpret_cast<const float4*>(x + i0);
            const float4 b = *reinterpret_cast<const float4*>(y + i0);
            const float4 c = *reinterpret_cast<const float4*>(z + i0);
            k[0] = cell_key(g, a.x, b.x, c.x); k[1] = cell_key(g, a.y, b.y, c.y);
            k[2] = cell_key(g, a.z, b.z, c.z); k[3] = cell_key(g, a.w, b.w, c.w);
            *reinterpret_cast<uint4*>(keys + i0) = make_uint4(k[0], k[1], k[2], k[3]);
            *reinterpret_cast<uint4*>(vals + i0) = make_uint4(i0, i0 + 1, i0 + 2, i0 + 3);
        } else {
            for (int j = 0; j < 4; ++j) {
                const int i = i0 + j < n ? i0 + j : n - 1;   // past n: the last slot again (no descent)
                k[j] = cell_key(g, x[i], y[i], z[i]);
                if (i0 + j < n) {
                    keys[i0 + j] = k[j];
                    vals[i0 + j] = (uint32_t)(i0 + j);
                }
            }
        }
        // the key before slot i0: the previous lane's last key, or for lane 0
        // slot i0-1 itself
        uint32_t prev = __shfl_up(k[3], 1, 64);
        if (lane == 0) prev = i0 > 0 ? cell_key(g, x[i0 - 1], y[i0 - 1], z[i0 - 1]) : 0u;
        desc += (int)(k[0] < prev) + (int)(k[1] < k[0]) + (int)(k[2] < k[1]) + (int)(k[3] < k[2]);
    }
    __shared__ int s_desc[4];
    for (int o = 32; o > 0; o >>= 1) desc += __shfl_xor(desc, o, 64);
    if (lane == 0) s_desc[threadIdx.x >> 6] = desc;
    __syncthreads();
    if (threadIdx.x == 0) info[2 + blockIdx.x] = s_desc[0] + s_desc[1] + s_desc[2] + s_desc[3];
}

// Sets the live count from the sorted keys: ascending, dead slots (key ==
// dead) last.
__device__ __forceinline__ void set_live_count(const uint32_t* keys, int i0, int n, uint32_t dead, int32_t* count)
{
    for (int i = i0; i < i0 + 4 && i < n; ++i) {
        const uint32_t kk = keys[i];
        if (kk < dead && (i + 1 == n || keys[i + 1] >= dead)) *count = i + 1;
        if (i == 0 && kk >= dead) *count = 0;
    }
}

// Gathers the sorted order into the scratch copies (four slots per thread,
// uint4 / float4 accesses of the scratch) and sets the live count.  For slots
// close to cell order: neighbouring destinations read neighbouring sources.
__global__ __launch_bounds__(256) void k_permute(const float* s0, const float* s1, const float* s2, const float* s3,
                                                 const float* s4, const float* s5, float* d0, float* d1, float* d2,
                                                 float* d3, float* d4, float* d5, const uint32_t* perm,
                                                 const uint32_t* keys, uint32_t dead, int32_t n, int32_t* count)
{
    const int ng = (n + 3) >> 2;
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += gridDim.x * blockDim.x) {
        const int i0 = gi * 4;
        if (i0 + 4 <= n) {
            const uint4 p4 = *reinterpret_cast<const uint4*>(perm + i0);
            const uint32_t q[4] = {p4.x, p4.y, p4.z, p4.w};
            // a permutation of [0, n): every source index is in range
            *reinterpret_cast<float4*>(d0 + i0) = make_float4(s0[q[0]], s0[q[1]], s0[q[2]], s0[q[3]]);
            *reinterpret_cast<float4*>(d1 + i0) = make_float4(s1[q[0]], s1[q[1]], s1[q[2]], s1[q[3]]);
            *reinterpret_cast<float4*>(d2 + i0) = make_float4(s2[q[0]], s2[q[1]], s2[q[2]], s2[q[3]]);
            *reinterpret_cast<float4*>(d3 + i0) = make_float4(s3[q[0]], s3[q[1]], s3[q[2]], s3[q[3]]);
            *reinterpret_cast<float4*>(d4 + i0) = make_float4(s4[q[0]], s4[q[1]], s4[q[2]], s4[q[3]]);
            *reinterpret_cast<float4*>(d5 + i0) = make_float4(s5[q[0]], s5[q[1]], s5[q[2]], s5[q[3]]);
        } else {
            for (int i = i0; i < n; ++i) {
                const uint32_t s = perm[i];
                d0[i] = s0[s]; d1[i] = s1[s]; d2[i] = s2[s];
                d3[i] = s3[s]; d4[i] = s4[s]; d5[i] = s5[s];
            }
        }
        set_live_count(keys, i0, n, dead, count);
    }
}

// Far from cell order, a gather of six 4-byte values per slot fetches six
// cache lines per slot.  The disordered path first packs each slot into a
// 32-byte record (x, y, z, vx | vy, vz, -, -; two 16-byte stores, coalesced),
// then gathers one aligned record per destination: one line, not six.
__global__ __launch_bounds__(256) void k_pack_records(const float* s0, const float* s1, const float* s2,
                                                      const float* s3, const float* s4, const float* s5,
                                                      float4* rec, int32_t n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        rec[2 * (size_t)i] = make_float4(s0[i], s1[i], s2[i], s3[i]);
        rec[2 * (size_t)i + 1] = make_float4(s4[i], s5[i], 0.0f, 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_permute_records(const float4* rec, float* d0, float* d1, float* d2,
                                                         float* d3, float* d4, float* d5, const uint32_t* perm,
                                                         const uint32_t* keys, uint32_t dead, int32_t n,
                                                         int32_t* count)
{
    const int ng = (n + 3) >> 2;
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += gridDim.x * blockDim.x) {
        const int i0 = gi * 4;
        if (i0 + 4 <= n) {
            const uint4 p4 = *reinterpret_cast<const uint4*>(perm + i0);
            const uint32_t q[4] = {p4.x, p4.y, p4.z, p4.w};
            float4 a[4], b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {   // a permutation of [0, n): every record is in range
                a[j] = rec[2 * (size_t)q[j]];
                b[j] = rec[2 * (size_t)q[j] + 1];
            }
            *reinterpret_cast<float4*>(d0 + i0) = make_float4(a[0].x, a[1].x, a[2].x, a[3].x);
            *reinterpret_cast<float4*>(d1 + i0) = make_float4(a[0].y, a[1].y, a[2].y, a[3].y);
            *reinterpret_cast<float4*>(d2 + i0) = make_float4(a[0].z, a[1].z, a[2].z, a[3].z);
            *reinterpret_cast<float4*>(d3 + i0) = make_float4(a[0].w, a[1].w, a[2].w, a[3].w);
            *reinterpret_cast<float4*>(d4 + i0) = make_float4(b[0].x, b[1].x, b[2].x, b[3].x);
            *reinterpret_cast<float4*>(d5 + i0) = make_float4(b[0].y, b[1].y, b[2].y, b[3].y);
        } else {
            for (int i = i0; i < n; ++i) {
                const float4 a = rec[2 * (size_t)perm[i]], b = rec[2 * (size_t)perm[i] + 1];
                d0[i] = a.x; d1[i] = a.y; d2[i] = a.z; d3[i] = a.w; d4[i] = b.x; d5[i] = b.y;
            }
        }
        set_live_count(keys, i0, n, dead, count);
    }
}

// Onesweep with 9-bit digits: the cell keys of grids up to 2^18 cells (dead
// slots take key nr*nz) sort in two passes instead of three of 8 bits.
using SortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;

// More descents than one slot in kDisorder: the record path.
constexpr int kDisorder = 8;

// scratch layout: [keys_in n][keys_out n][vals_in n][vals_out n][6 x float n]
// [records 8 x n][info][pad][rocprim tmp]
namespace {
struct SortLayout {
    size_t a = 0, info_bytes = 0, tmp_bytes = 0, need = 0;
    int bits = 1;
};

hipError_t sort_layout(const Geom& g, size_t cap, hipStream_t st, SortLayout& L)
{
    L.a = (cap * 4 + 255) & ~(size_t)255;
    const uint32_t nkeys = (uint32_t)(g.nr * g.nz);
    L.bits = 1;
    while ((1u << L.bits) <= nkeys && L.bits < 32) ++L.bits;
    hipError_t e = rocprim::radix_sort_pairs<SortConfig>((void*)nullptr, L.tmp_bytes, (uint32_t*)nullptr,
                                                         (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                         (size_t)(cap > 0 ? cap : 1), 0, L.bits, st);
    if (e != hipSuccess) return e;
    L.info_bytes = ((2 + kSortKeyBlocks) * sizeof(int32_t) + 255) & ~(size_t)255;
    L.need = 18 * L.a + L.info_bytes + ((L.tmp_bytes + 255) & ~(size_t)255);
    return hipSuccess;
}

int key_blocks(size_t cap)
{
    int blocks = (int)(((cap + 3) / 4 + 255) / 256);
    if (blocks > kSortKeyBlocks) blocks = kSortKeyBlocks;
    return blocks < 1 ? 1 : blocks;
}
}  // namespace

hipError_t sort_begin(const Geom& g, gasx_species* sp, void* scratch, size_t scratch_bytes, size_t* need_bytes,
                      hipStream_t st, int32_t* pinned_info, hipEvent_t ready)
{
    const size_t cap = (size_t)sp->capacity;
    SortLayout L;
    hipError_t e = sort_layout(g, cap, st, L);
    if (e != hipSuccess) return e;
    if (need_bytes) *need_bytes = L.need;
    if (!scratch || scratch_bytes < L.need || !pinned_info || !ready) return hipErrorInvalidValue;
    char* base = (char*)scratch;
    uint32_t* k_in = (uint32_t*)(base);
    uint32_t* v_in = (uint32_t*)(base + 2 * L.a);
    int32_t* info = (int32_t*)(base + 18 * L.a);
    const int blocks = key_blocks(cap);
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, st, g, sp->x, sp->y, sp->z, sp->count,
                       (int32_t)cap, k_in, v_in, info);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(pinned_info, info, (2 + (size_t)blocks) * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipEventRecord(ready, st);
}

hipError_t sort_finish(const Geom& g, gasx_species* sp, void* scratch, hipStream_t st, float** sorted_out,
                       const int32_t* pinned_info, hipEvent_t ready)
{
    const size_t cap = (size_t)sp->capacity;
    SortLayout L;
    hipError_t e = sort_layout(g, cap, st, L);
    if (e != hipSuccess) return e;
    char* base = (char*)scratch;
    uint32_t* k_in = (uint32_t*)(base);
    uint32_t* k_out = (uint32_t*)(base + L.a);
    uint32_t* v_in = (uint32_t*)(base + 2 * L.a);
    uint32_t* v_out = (uint32_t*)(base + 3 * L.a);
    float* tmp[6];
    for (int c = 0; c < 6; ++c) tmp[c] = (float*)(base + (4 + c) * L.a);
    float4* rec = (float4*)(base + 10 * L.a);
    void* rtmp = (void*)(base + 18 * L.a + L.info_bytes);
    // the one host synchronisation of the sort (rocPRIM takes the size on the
    // host): on the key pass's event only, so work queued after it keeps the
    // GPU busy meanwhile
    e = hipEventSynchronize(ready);
    if (e != hipSuccess) return e;
    const int blocks = key_blocks(cap);
    const int32_t n = pinned_info[0];
    int64_t descents = 0;
    for (int b = 0; b < blocks; ++b) descents += pinned_info[2 + b];
    if (n <= 0) {
        if (sorted_out) {   // nothing to permute: the particles stay where they are
            float* own[6] = {sp->x, sp->y, sp->z, sp->vx, sp->vy, sp->vz};
            for (int c = 0; c < 6; ++c) sorted_out[c] = own[c];
        }
        return hipSuccess;
    }
    size_t tb = L.tmp_bytes;
    e = rocprim::radix_sort_pairs<SortConfig>(rtmp, tb, k_in, k_out, v_in, v_out, (size_t)n, 0, L.bits, st);
    if (e != hipSuccess) return e;
    const uint32_t dead = (uint32_t)(g.nr * g.nz);
    int pblocks = ((n + 3) / 4 + 255) / 256;
    if (pblocks > 8192) pblocks = 8192;
    if ((int64_t)descents * kDisorder > (int64_t)n) {
        int rblocks = (n + 255) / 256;
        if (rblocks > 8192) rblocks = 8192;
        hipLaunchKernelGGL(k_pack_records, dim3(rblocks), dim3(256), 0, st, sp->x, sp->y, sp->z, sp->vx, sp->vy,
                           sp->vz, rec, n);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_permute_records, dim3(pblocks), dim3(256), 0, st, (const float4*)rec, tmp[0], tmp[1],
                           tmp[2], tmp[3], tmp[4], tmp[5], v_out, k_out, dead, n, sp->count);
    } else {
        hipLaunchKernelGGL(k_permute, dim3(pblocks), dim3(256), 0, st, sp->x, sp->y, sp->z, sp->vx,
                           sp->vy, sp->vz, tmp[0], tmp[1], tmp[2], tmp[3], tmp[4], tmp[5], v_out, k_out,
                           dead, n, sp->count);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (sorted_out) {
        for (int c = 0; c < 6; ++c) sorted_out[c] = tmp[c];
        return e;
    }
    // copy back (slots past the live count hold stale values; the count excludes them)
    float* dst[6] = {sp->x, sp->y, sp->z, sp->vx, sp->vy, sp->vz};
    for (int c = 0; c < 6 && e == hipSuccess; ++c)
        e = hipMemcpyAsync(dst[c], tmp[c], (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st);
    return e;
}

// ---- FP64 tallies (electricmagneticfield5.m:167-177) ---------------------------
// W_E = PHIp*eps/2 * (trapz trapz Er^2 r + trapz trapz Ez^2 r), W_H likewise
// with mu, on the grid coordinates of each array (electricmagneticfield6.m:25-61).
constexpr int kTallyBlocks = 256;

__device__ __forceinline__ double trap_w(int i, int n) { return (i == 0 || i == n - 1) ? 0.5 : 1.0; }

__global__ __launch_bounds__(256) void k_tally_fields(int nr, int nz, double dr, double dz,
                                                      const double* er, const double* ez,
                                                      const double* h, double* partial)
{
    double se = 0.0, sh = 0.0;
    const int ner = nr * (nz + 1), nez = (nr + 1) * nz, nh = nr * nz;
    const int stride = gridDim.x * blockDim.x;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ner; c += stride) {
        int i = c % nr, k = c / nr;
        double r = (i + 0.5) * dr;
        se += trap_w(i, nr) * trap_w(k, nz + 1) * er[c] * er[c] * r;
    }
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nez; c += stride) {
        int i = c % (nr + 1), k = c / (nr + 1);
        double r = i * dr;
        se += trap_w(i, nr + 1) * trap_w(k, nz) * ez[c] * ez[c] * r;
    }
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nh; c += stride) {
        int i = c % nr, k = c / nr;
        double r = (i + 0.5) * dr;
        sh += trap_w(i, nr) * trap_w(k, nz) * h[c] * h[c] * r;
    }
    __shared__ double s_e[256], s_h[256];
    s_e[threadIdx.x] = se * dr * dz;
    s_h[threadIdx.x] = sh * dr * dz;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_e[threadIdx.x] += s_e[threadIdx.x + o];
            s_h[threadIdx.x] += s_h[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s_e[0];
        partial[2 * blockIdx.x + 1] = s_h[0];
    }
}

__global__ __launch_bounds__(256) void k_tally_ke(const float* vx, const float* vy, const float* vz,
                                                  const float* z, const int32_t* count, double mc2,
                                                  double inv_c2, double* partial)
{
    const int n = *count;
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (is_dead(z[i])) continue;
        double a = vx[i], b = vy[i], c = vz[i];
        double b2 = (a * a + b * b + c * c) * inv_c2;
        double g2 = 1.0 / (1.0 - b2);
        double gm = sqrt(g2);
        s += mc2 * (b2 * g2 / (1.0 + gm));
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] += sm[0];
}

__global__ void k_tally_final(const double* pf, const double* pk, double phip, double eps0,
                              double mu0, double fn, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double se = 0.0, sh = 0.0, sk = 0.0;
    for (int b = 0; b < kTallyBlocks; ++b) {
        se += pf[2 * b];
        sh += pf[2 * b + 1];
        sk += pk[b];
    }
    out[0] = phip * 0.5 * eps0 * se;
    out[1] = phip * 0.5 * mu0 * sh;
    out[2] = fn * sk;
    out[3] = out[0] + out[1] + out[2];
}

__global__ void k_zero_d(double* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

hipError_t launch_tallies(const gasx_params& p, const double* er, const double* ez, const double* h,
                          const gasx_species* sp, int nsp, double* partial, double* out,
                          hipStream_t s)
{
    // partial: 2*kTallyBlocks field partials + kTallyBlocks KE partials
    double* pf = partial;
    double* pk = partial + 2 * kTallyBlocks;
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, s, pk, kTallyBlocks);
    hipLaunchKernelGGL(k_tally_fields, dim3(kTallyBlocks), dim3(256), 0, s, p.nr, p.nz, p.dr, p.dz,
                       er, ez, h, pf);
    for (int k = 0; k < nsp; ++k) {
        const double c = p.c;
        hipLaunchKernelGGL(k_tally_ke, dim3(kTallyBlocks), dim3(256), 0, s, sp[k].vx, sp[k].vy,
                           sp[k].vz, sp[k].z, sp[k].count, sp[k].mass * c * c, 1.0 / (c * c), pk);
    }
    hipLaunchKernelGGL(k_tally_final, dim3(1), dim3(64), 0, s, pf, pk, p.phip, p.eps0, p.mu0, p.fn,
                       out);
    return hipGetLastError();
}

}  // namespace gasx
