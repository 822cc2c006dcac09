// pd_setup.hip -- the mesh set-up of a context built on the device
// (pd_create_from_points): families from a cell list, the Morton order, the
// internal CSR, the reverse-bond index, then the host mirrors the ABI needs.
//
// The reference finds neighbours by an all-pairs O(N^2) scan in long double
// (findNeighbor, src/SetupParticleSystem.cpp:100-136, cutoff |xi| <
// 1.01 dx delta at :113, neighbours in ascending id order) and orders nothing.
// Here the cell side is the cutoff, so a particle's neighbours lie in its 3x3
// cells; the distance test is FP64 (|xi|^2 < cutoff^2 with the cutoff formed in
// long double on the host), which differs from the long-double test only for
// a distance within ~1e-15 relative of the cutoff.  Families come out in
// ascending caller id, as findNeighbor and pd_build_families give them, and the
// particles in the same Morton order pd_create uses, so a context made here
// and one made by pd_create from pd_build_families' families are identical.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>

#include "pd_context.h"

namespace {

__global__ void k_cell_keys(int n, const double2* __restrict__ X, double xmin, double ymin, double cs,
                            int ncx, int ncy, unsigned int* __restrict__ key, int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 x = X[i];
    const int cx = min(ncx - 1, max(0, (int)floor((x.x - xmin) / cs)));
    const int cy = min(ncy - 1, max(0, (int)floor((x.y - ymin) / cs)));
    key[i] = (unsigned int)cy * (unsigned int)ncx + (unsigned int)cx;
    idx[i] = i;
}

__global__ void k_cell_ranges(int n, const unsigned int* __restrict__ skey, int* __restrict__ cstart,
                              int* __restrict__ cend)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const unsigned int k = skey[s];
    if (s == 0 || skey[s - 1] != k) cstart[k] = s;
    if (s == n - 1 || skey[s + 1] != k) cend[k] = s + 1;
}

// Neighbours of caller particle i: count them (col == nullptr) or write them in
// ascending id order into col[row[i] ..].
__global__ void k_neighbours(int n, const double2* __restrict__ X, const int* __restrict__ sidx,
                             const int* __restrict__ cstart, const int* __restrict__ cend, double xmin,
                             double ymin, double cs, int ncx, int ncy, double rc2, int* __restrict__ cnt,
                             const int* __restrict__ row, int* __restrict__ col)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 xi = X[i];
    const int cx = min(ncx - 1, max(0, (int)floor((xi.x - xmin) / cs)));
    const int cy = min(ncy - 1, max(0, (int)floor((xi.y - ymin) / cs)));
    int m = 0;
    const int base = col ? row[i] : 0;
    for (int yy = max(0, cy - 1); yy <= min(ncy - 1, cy + 1); ++yy)
        for (int xx = max(0, cx - 1); xx <= min(ncx - 1, cx + 1); ++xx) {
            const int k = yy * ncx + xx;
            for (int s = cstart[k]; s < cend[k]; ++s) {
                const int j = sidx[s];
                if (j == i) continue;
                const double2 xj = X[j];
                const double r0 = xj.x - xi.x, r1 = xj.y - xi.y;
                if (r0 * r0 + r1 * r1 < rc2) {
                    if (col) {
                        // insertion into the sorted prefix
                        int t = m;
                        while (t > 0 && col[base + t - 1] > j) {
                            col[base + t] = col[base + t - 1];
                            --t;
                        }
                        col[base + t] = j;
                    }
                    ++m;
                }
            }
        }
    if (!col) cnt[i] = m;
}

__device__ __forceinline__ unsigned long long spread32d(unsigned long long x)
{
    x &= 0xffffffffULL;
    x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

// Morton key of the dx-grid cell, the same quantisation as pd_create.
__global__ void k_morton(int n, const double2* __restrict__ X, double xmin, double ymin, double dx,
                         unsigned long long* __restrict__ key, int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 x = X[i];
    const double qx = fmin(fmax(floor((x.x - xmin) / dx + 0.5), 0.0), 4294967295.0);
    const double qy = fmin(fmax(floor((x.y - ymin) / dx + 0.5), 0.0), 4294967295.0);
    key[i] = spread32d((unsigned long long)qx) | (spread32d((unsigned long long)qy) << 1);
    idx[i] = i;
}

// Internal order: X, V, row lengths and the inverse permutation.
__global__ void k_permute(int n, const int* __restrict__ perm, const double2* __restrict__ Xc,
                          const double* __restrict__ Vc, const int* __restrict__ cnt,
                          double2* __restrict__ X, double* __restrict__ V, int* __restrict__ len,
                          int* __restrict__ iperm)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = perm[q];
    X[q] = Xc[i];
    V[q] = Vc[i];
    len[q] = cnt[i];
    iperm[i] = q;
}

// Internal CSR: the caller's row of particle perm[q] with neighbours renamed.
__global__ void k_internal_csr(int n, const int* __restrict__ perm, const int* __restrict__ iperm,
                               const int* __restrict__ rowc, const int* __restrict__ colc,
                               const int* __restrict__ row, int* __restrict__ col, int* __restrict__ bmap)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = perm[q];
    int bi = row[q];
    for (int e = rowc[i]; e < rowc[i + 1]; ++e, ++bi) {
        col[bi] = iperm[colc[e]];
        bmap[bi] = e;
    }
}

// rev[b] for b = (q -> r): the position of q in r's row (-1 if absent).
__global__ void k_reverse(int n, const int* __restrict__ row, const int* __restrict__ col,
                          int* __restrict__ rev, int* __restrict__ missing)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    for (int b = row[q]; b < row[q + 1]; ++b) {
        const int r = col[b];
        int found = -1;
        for (int e = row[r]; e < row[r + 1]; ++e)
            if (col[e] == q) {
                found = e;
                break;
            }
        rev[b] = found;
        if (found < 0) *missing = 1;
    }
}

int grid_of(long n) { return (int)((n + 255) / 256); }

struct DevTemp {
    std::vector<void*> ptrs;
    ~DevTemp()
    {
        for (void* p : ptrs)
            if (p) hipFree(p);
    }
    template <typename T>
    hipError_t get(T** p, size_t count)
    {
        *p = nullptr;
        hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) ptrs.push_back((void*)*p);
        return e;
    }
};

}  // namespace

#define SETCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(c, PD_ERR_HIP, "set-up: %s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int build_mesh_on_device(pd_context* c, const double* Xh, const double* Vh)
{
    const int n = c->n;
    const pd_params& p = c->p;
    hipStream_t s = c->stream;
    // cutoff of findNeighbor (src/SetupParticleSystem.cpp:113), formed in long double
    const long double cut = (long double)1.01 * (long double)p.dx * (long double)p.horizon;
    const double cs = (double)cut;
    const double rc2 = (double)(cut * cut);
    double xmin = Xh[0], ymin = Xh[1], xmax = Xh[0], ymax = Xh[1];
    for (int i = 0; i < n; ++i) {
        xmin = std::min(xmin, Xh[2 * i]);
        xmax = std::max(xmax, Xh[2 * i]);
        ymin = std::min(ymin, Xh[2 * i + 1]);
        ymax = std::max(ymax, Xh[2 * i + 1]);
    }
    const double fx = std::floor((xmax - xmin) / cs) + 1, fy = std::floor((ymax - ymin) / cs) + 1;
    if (fx * fy > 1.0e9)
        return fail(c, PD_ERR_UNSUPPORTED, "set-up: %.3g cells of the cutoff size; the domain is too sparse", fx * fy);
    const int ncx = (int)std::max(1.0, fx), ncy = (int)std::max(1.0, fy);
    const long ncell = (long)ncx * ncy;

    DevTemp t;
    double2* Xc;
    double* Vc;
    unsigned int *key, *skey;
    int *idx, *sidx, *cstart, *cend, *cnt, *rowc, *colc;
    SETCHK(t.get(&Xc, n));
    SETCHK(t.get(&Vc, n));
    SETCHK(t.get(&key, n));
    SETCHK(t.get(&skey, n));
    SETCHK(t.get(&idx, n));
    SETCHK(t.get(&sidx, n));
    SETCHK(t.get(&cstart, ncell));
    SETCHK(t.get(&cend, ncell));
    SETCHK(t.get(&cnt, n));
    SETCHK(t.get(&rowc, (size_t)n + 1));
    SETCHK(hipMemcpyAsync(Xc, Xh, sizeof(double2) * n, hipMemcpyHostToDevice, s));
    SETCHK(hipMemcpyAsync(Vc, Vh, sizeof(double) * n, hipMemcpyHostToDevice, s));
    SETCHK(hipMemsetAsync(cstart, 0, sizeof(int) * ncell, s));
    SETCHK(hipMemsetAsync(cend, 0, sizeof(int) * ncell, s));

    // --- cell list ---
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_of(n)), dim3(256), 0, s, n, Xc, xmin, ymin, cs, ncx, ncy, key, idx);
    SETCHK(hipGetLastError());
    int kbits = 1;
    while (kbits < 32 && (1L << kbits) < ncell) ++kbits;
    size_t tmp_bytes = 0, tb = 0;
    SETCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, idx, sidx, n, 0, kbits, s));
    tmp_bytes = std::max(tmp_bytes, tb);
    SETCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, rowc, n, s));
    tmp_bytes = std::max(tmp_bytes, tb);
    unsigned long long *mkey, *smkey;
    int* maxm;
    SETCHK(t.get(&mkey, n));
    SETCHK(t.get(&smkey, n));
    SETCHK(t.get(&maxm, 1));
    SETCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, mkey, smkey, idx, sidx, n, 0, 64, s));
    tmp_bytes = std::max(tmp_bytes, tb);
    SETCHK(hipcub::DeviceReduce::Max(nullptr, tb, cnt, maxm, n, s));
    tmp_bytes = std::max(tmp_bytes, tb);
    void* tmp;
    SETCHK(t.get((char**)&tmp, tmp_bytes));
    SETCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, skey, idx, sidx, n, 0, kbits, s));
    hipLaunchKernelGGL(k_cell_ranges, dim3(grid_of(n)), dim3(256), 0, s, n, skey, cstart, cend);
    SETCHK(hipGetLastError());

    // --- families in caller order ---
    hipLaunchKernelGGL(k_neighbours, dim3(grid_of(n)), dim3(256), 0, s, n, Xc, sidx, cstart, cend, xmin,
                       ymin, cs, ncx, ncy, rc2, cnt, (const int*)nullptr, (int*)nullptr);
    SETCHK(hipGetLastError());
    SETCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, rowc, n, s));
    SETCHK(hipcub::DeviceReduce::Max(tmp, tmp_bytes, cnt, maxm, n, s));
    int last[2], mm = 0;
    SETCHK(hipMemcpyAsync(&last[0], rowc + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    SETCHK(hipMemcpyAsync(&last[1], cnt + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    SETCHK(hipMemcpyAsync(&mm, maxm, sizeof(int), hipMemcpyDeviceToHost, s));
    SETCHK(hipStreamSynchronize(s));
    const long nb = (long)last[0] + last[1];
    if (nb > INT_MAX) return fail(c, PD_ERR_UNSUPPORTED, "set-up: %ld bonds exceed the int32 bond index", nb);
    if (mm > 64) return fail(c, PD_ERR_UNSUPPORTED, "family of %d bonds: at most 64 are supported", mm);
    const int nbi = (int)nb;
    SETCHK(hipMemcpyAsync(rowc + n, &nbi, sizeof(int), hipMemcpyHostToDevice, s));
    SETCHK(t.get(&colc, (size_t)nb));
    hipLaunchKernelGGL(k_neighbours, dim3(grid_of(n)), dim3(256), 0, s, n, Xc, sidx, cstart, cend, xmin,
                       ymin, cs, ncx, ncy, rc2, cnt, (const int*)rowc, colc);
    SETCHK(hipGetLastError());
    c->nb = nb;
    c->maxM = mm;

    // --- Morton order, internal CSR, reverse bonds ---
    hipLaunchKernelGGL(k_morton, dim3(grid_of(n)), dim3(256), 0, s, n, Xc, xmin, ymin, p.dx, mkey, idx);
    SETCHK(hipGetLastError());
    SETCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, mkey, smkey, idx, sidx, n, 0, 64, s));
    int *perm = sidx, *iperm, *len, *bmap, *missing;
    SETCHK(t.get(&iperm, n));
    SETCHK(t.get(&len, n));
    SETCHK(t.get(&bmap, (size_t)nb));
    SETCHK(t.get(&missing, 1));
    if (int rc = alloc_mesh(c)) return rc;
    hipLaunchKernelGGL(k_permute, dim3(grid_of(n)), dim3(256), 0, s, n, perm, Xc, Vc, cnt, c->X, c->V, len, iperm);
    SETCHK(hipGetLastError());
    SETCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, len, c->row, n, s));
    SETCHK(hipMemcpyAsync(c->row + n, &nbi, sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_internal_csr, dim3(grid_of(n)), dim3(256), 0, s, n, perm, iperm, rowc, colc, c->row,
                       c->col, bmap);
    SETCHK(hipGetLastError());
    SETCHK(pd::launch_length_order(n, c->row, c->X, c->col, bmap, s));
    SETCHK(hipMemsetAsync(missing, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_reverse, dim3(grid_of(n)), dim3(256), 0, s, n, c->row, c->col, c->rev, missing);
    SETCHK(hipGetLastError());

    // --- host mirrors ---
    int miss = 0;
    c->perm.resize(n);
    c->iperm.resize(n);
    c->row_host.resize((size_t)n + 1);
    c->rev_host.resize(std::max(nb, 1L));
    c->V_int.resize(n);
    std::vector<int> bm(std::max(nb, 1L));
    SETCHK(hipMemcpyAsync(&miss, missing, sizeof(int), hipMemcpyDeviceToHost, s));
    SETCHK(hipMemcpyAsync(c->perm.data(), perm, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    SETCHK(hipMemcpyAsync(c->iperm.data(), iperm, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    SETCHK(hipMemcpyAsync(c->row_host.data(), c->row, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    SETCHK(hipMemcpyAsync(c->V_int.data(), c->V, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (nb > 0) {
        SETCHK(hipMemcpyAsync(c->rev_host.data(), c->rev, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
        SETCHK(hipMemcpyAsync(bm.data(), bmap, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    }
    SETCHK(hipStreamSynchronize(s));
    if (miss) return fail(c, PD_ERR_INVALID, "set-up: a bond has no reverse (asymmetric distance test)");
    c->rev_host.resize(nb);
    c->bmap.assign(bm.begin(), bm.begin() + nb);
    c->X_caller.assign(Xh, Xh + 2 * (size_t)n);
    return PD_OK;
}
