// pd_kernels.hip -- gfx950 kernels of the bond-associated NOSB peridynamics
// step other than the bond kernel itself (pd_bond.h, instantiated per precision
// in pd_bond_f64.hip / pd_bond_f32.hip): the per-particle gather and velocity
// update, the position update with boundary conditions, K^-1 assembly at
// set-up, small list kernels, the copy probe, and the launch dispatch.
#include "pd_bond.h"

namespace pd {

// Velocity BC, a = f/rho, u += v dt + 1/2 a dt^2, fixed BC for particle i
// (one step's prologue, src/main_seq.cpp:686-692), from its current v and f.
__device__ __forceinline__ void position_update(const PosArgs& p, int i, double2 v, double2 f)
{
#pragma clang fp contract(off)
    const int code = p.bc ? p.bc[i] : -1;
    if (code >= 0) {
        v = p.v0[code];
        p.v[i] = v;
    }
    const double2 acc = make_double2(f.x / p.rho, f.y / p.rho);
    p.a[i] = acc;
    double2 u = p.u[i];
    const double dt2 = p.dt * p.dt;
    u.x = u.x + (v.x * p.dt + (0.5 * acc.x) * dt2);
    u.y = u.y + (v.y * p.dt + (0.5 * acc.y) * dt2);
    if (code == -2) u = make_double2(0.0, 0.0);
    p.u[i] = u;
}

// f_i = sum_j T_ij V_j - sum_j T_ji V_i ; v += 1/2 (a + f/rho) dt.
// GATHER_LANES lanes per particle read its W row with unit stride (coalesced),
// then reduce across the group.  POS: the step is followed by another one in
// the same pd_step call, so the next step's position update (k_positions)
// runs here on the final f and v, with the same operations in the same order.
// LANES lanes per particle read its W row; each lane first loads UNROLL entries
// of its share at once, then loops over the rest.  The launch picks the shape
// from the largest family: FP32 rows of up to 28 entries (an interior family at
// delta = 3.015) take 4 x 7, one round; otherwise 8 lanes with 4, 6 or 8 loads
// in flight, so the row fits one round.  Measured, unfused gather: 1024 x 1024
// plate (families of 28) FP32 4 x 7 64-65 us, 4 x 8 65-66 us, 8 x 4 70-71 us,
// 2 x 16 83 us, FP64 8 x 4 101 us, 4 x 8 108 us; 4096 x 4096 plate (families
// of 48) FP32 8 x 4 1.56 ms, 4 x 7 1.89 ms.
template <typename R, bool POS, int GATHER_LANES, int GATHER_UNROLL>
__global__ void __launch_bounds__(256) k_gather(GatherArgs g, PosArgs p)
{
    typedef typename Vec2<R>::T R2;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, g.xcd);
    const int gid = blk * blockDim.x + threadIdx.x;
    const int i = gid / GATHER_LANES;
    const int t = gid % GATHER_LANES;
    const bool ok = i < g.n;
    const int ic = ok ? i : 0;
    // the particle's own state is loaded before the row reduction (the lanes
    // of a group share each address), so its latency overlaps the W reads
    const int b0 = g.row[ic], b1 = g.row[ic + 1];
    const double2 fo = g.fown[ic];
    const double2 a0 = g.a[ic];
    const double2 vi = g.v[ic];
    double2 ui = make_double2(0.0, 0.0);
    int code = -1;
    if constexpr (POS) {
        ui = p.u[ic];
        code = p.bc ? p.bc[ic] : -1;
    }
    const R2* W = reinterpret_cast<const R2*>(g.W);
    double f0 = 0.0, f1 = 0.0;
    // the first GATHER_UNROLL entries of the lane's share of the row are loaded
    // together (indices clamped to entry 0, values masked), so they are in
    // flight at once; longer rows continue in the loop.  Same summation order.
    R2 wv[GATHER_UNROLL];
#pragma unroll
    for (int q = 0; q < GATHER_UNROLL; ++q) {
        const int b = b0 + t + q * GATHER_LANES;
        wv[q] = W[ok && b < b1 ? b : 0];
    }
#pragma unroll
    for (int q = 0; q < GATHER_UNROLL; ++q) {
        if (ok && b0 + t + q * GATHER_LANES < b1) {
            f0 += double(wv[q].x);
            f1 += double(wv[q].y);
        }
    }
    if (ok) {
        for (int b = b0 + t + GATHER_UNROLL * GATHER_LANES; b < b1; b += GATHER_LANES) {
            const R2 w = W[b];
            f0 += double(w.x);
            f1 += double(w.y);
        }
    }
#pragma unroll
    for (int off = GATHER_LANES / 2; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off, 64);
        f1 += __shfl_xor(f1, off, 64);
    }
    // the guard's list is used up (k_bond_fix ran before this kernel)
    if (g.famcnt && gid == 0) *g.famcnt = 0u;
    if (!ok || t != 0) return;
    {
#pragma clang fp contract(off)
    f0 = fo.x - f0;
    f1 = fo.y - f1;
    const double an0 = f0 / g.rho, an1 = f1 / g.rho;
    double2 v;
    v.x = vi.x + (0.5 * (a0.x + an0)) * g.dt;
    v.y = vi.y + (0.5 * (a0.y + an1)) * g.dt;
    if constexpr (POS) {
        // f is not stored here: nothing reads it before the last step of the
        // pd_step call, whose gather (POS = false) stores it; the next step
        // needs only a = f / rho.
        // the next step's position update (k_positions), same operations
        if (code >= 0) v = p.v0[code];
        g.v[i] = v;
        g.a[i] = make_double2(an0, an1);
        const double dt2 = p.dt * p.dt;
        double2 u;
        u.x = ui.x + (v.x * p.dt + (0.5 * an0) * dt2);
        u.y = ui.y + (v.y * p.dt + (0.5 * an1) * dt2);
        if (code == -2) u = make_double2(0.0, 0.0);
        p.u[i] = u;
    } else {
        g.f[i] = make_double2(f0, f1);
        g.v[i] = v;
        if (g.write_a) g.a[i] = make_double2(an0, an1);
    }
    }
}

__global__ void __launch_bounds__(256) k_positions(PosArgs p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    position_update(p, i, p.v[i], p.f[i]);
}

__global__ void k_vbc_list(int count, const int* ids, double2 v0, double2* v, double2* a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = ids[t];
    v[i] = v0;
    a[i] = make_double2(0.0, 0.0);
}

__global__ void k_fixed_list(int count, const int* ids, double2* u)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    u[ids[t]] = make_double2(0.0, 0.0);
}

// Smallest internal index whose u, v or f is not finite (integer atomic).
__global__ void __launch_bounds__(256) k_check_finite(int n, const double2* __restrict__ u,
                                                      const double2* __restrict__ v,
                                                      const double2* __restrict__ f,
                                                      unsigned int* __restrict__ first)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 a = u[i], b = v[i], c = f[i];
    if (!(isfinite(a.x) && isfinite(a.y) && isfinite(b.x) && isfinite(b.y) && isfinite(c.x) && isfinite(c.y)))
        atomicMin(first, (unsigned int)i);
}

// Alive, non-singular bonds whose shape tensor has condition number >= cmin,
// read from the stored K^-1 (same eigenvectors, reciprocal eigenvalues).  For
// a symmetric 2x2 matrix with eigenvalues l1 >= l2 > 0 and c = l1 / l2,
// tr^2 / det = c + 2 + 1/c, so c >= cmin exactly when tr^2 >= r det with
// r = cmin + 2 + 1/cmin.  A stored inverse whose det is not positive (an FP32
// rounding of a nearly singular one) counts as ill-conditioned.  Broken and
// singular bonds store K^-1 = 0 and are skipped.
template <typename KT>
__global__ void __launch_bounds__(256) k_count_ill_conditioned(long nb, const KT* __restrict__ kinv,
                                                               const unsigned char* __restrict__ broken,
                                                               double r, unsigned int* __restrict__ count)
{
    const long stride = (long)gridDim.x * blockDim.x;
    unsigned int mine = 0;
    for (long b = (long)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += stride) {
        const double a = kinv[3 * b], o = kinv[3 * b + 1], c = kinv[3 * b + 2];
        if (broken[b] || (a == 0.0 && o == 0.0 && c == 0.0)) continue;
        const double det = a * c - o * o, tr = a + c;
        mine += (det <= 0.0 || tr * tr >= r * det) ? 1u : 0u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off, 64);
    if ((threadIdx.x & 63) == 0 && mine) atomicAdd(count, mine);
}

// Rank of each bond's reference length in its family (squared lengths in
// FP64, ties broken by position), one thread per family.
// Families in ascending order of |X_j - X_i|^2, ties in their present order:
// one thread per family reorders its (col, bmap) entries in place.
__global__ void __launch_bounds__(256) k_length_order(int n, const int* __restrict__ row,
                                                      const double2* __restrict__ X, int* __restrict__ col,
                                                      int* __restrict__ bmap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = row[i], m = min(row[i + 1] - b0, 64);
    int cj[64], bm[64], ord[64];
    for (int e = 0; e < m; ++e) {
        cj[e] = col[b0 + e];
        bm[e] = bmap[b0 + e];
    }
    family_length_order(m, cj, X, n, X[i], ord);
    for (int r = 0; r < m; ++r) {
        col[b0 + r] = cj[ord[r]];
        bmap[b0 + r] = bm[ord[r]];
    }
}

__global__ void k_fill_u32(unsigned int* p, unsigned int v, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = v;
}

// Plain streaming copy of n4 32-byte elements (grid-stride), the reference
// kernel of known traffic for bandwidth and HBM-counter calibration.
__global__ void __launch_bounds__(256) k_copy_probe(const double4* __restrict__ src,
                                                    double4* __restrict__ dst, long n4)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

bool tile_supported(int mixed, int lpf, int bpl, int mf)
{
    if (mf == 2) return !mixed && ((bpl == 1 && (lpf == 32 || lpf == 64)) || (lpf == 16 && bpl == 3));
    if (mf == 3) return lpf == 16 && (bpl == 2 || bpl == 3);
    if (mf == 4) return mixed && lpf == 16 && (bpl == 2 || bpl == 3);
    if (mf && !mixed) return (bpl == 1 && (lpf == 32 || lpf == 64)) || (lpf == 16 && (bpl == 2 || bpl == 3));
    if (bpl == 1) return lpf == 32 || lpf == 64;
    if (bpl == 2) return lpf == 16 || (lpf == 32 && mixed);
    if (bpl == 3) return lpf == 16;
    return false;
}

hipError_t launch_bond(const BondArgs& a, int mixed, int lpf, int bpl, int mf, int dmg, int n2_is_10,
                       int wpb, int pgrid, hipStream_t s)
{
    if (!tile_supported(mixed, lpf, bpl, mf)) return hipErrorInvalidValue;
    return mixed ? launch_bond_f32(a, lpf, bpl, mf, dmg, n2_is_10, wpb, pgrid, s)
                 : launch_bond_f64(a, lpf, bpl, mf, dmg, n2_is_10, wpb, pgrid, s);
}

template <int LPF, int BPL, int N2>
static hipError_t launch_shape_t(const BondArgs& a, int wpb, hipStream_t s)
{
    const int grid = ceil_div(a.n, (long)wpb * Tile<LPF, BPL>::FPW);
    if (grid == 0) return hipSuccess;
    const size_t lds = (size_t)wpb * Tile<LPF, BPL>::LDS_DOUBLES * sizeof(double);
    hipLaunchKernelGGL((k_shape<LPF, BPL, N2>), dim3(grid), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_shape(const BondArgs& a, int lpf, int bpl, int n2_is_10, int wpb, hipStream_t s)
{
    if (bpl == 1 && lpf == 32) return n2_is_10 ? launch_shape_t<32, 1, 1>(a, wpb, s) : launch_shape_t<32, 1, 0>(a, wpb, s);
    if (bpl == 1 && lpf == 64) return n2_is_10 ? launch_shape_t<64, 1, 1>(a, wpb, s) : launch_shape_t<64, 1, 0>(a, wpb, s);
    if (bpl == 2 && lpf == 16) return n2_is_10 ? launch_shape_t<16, 2, 1>(a, wpb, s) : launch_shape_t<16, 2, 0>(a, wpb, s);
    if (bpl == 2 && lpf == 32) return n2_is_10 ? launch_shape_t<32, 2, 1>(a, wpb, s) : launch_shape_t<32, 2, 0>(a, wpb, s);
    if (bpl == 3 && lpf == 16) return n2_is_10 ? launch_shape_t<16, 3, 1>(a, wpb, s) : launch_shape_t<16, 3, 0>(a, wpb, s);
    return hipErrorInvalidValue;
}

template <typename R, int LANES, int UNROLL>
static hipError_t launch_gather_t(const GatherArgs& g, hipStream_t s, const PosArgs* next)
{
    const int grid = ceil_div((long)g.n * LANES, 256);
    if (grid == 0) return hipSuccess;
    const PosArgs p = next ? *next : PosArgs{};
    if (next)
        hipLaunchKernelGGL((k_gather<R, true, LANES, UNROLL>), dim3(grid), dim3(256), 0, s, g, p);
    else
        hipLaunchKernelGGL((k_gather<R, false, LANES, UNROLL>), dim3(grid), dim3(256), 0, s, g, p);
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_gather_r(const GatherArgs& g, hipStream_t s, const PosArgs* next)
{
    if (sizeof(R) == 4 && g.max_family <= 28) return launch_gather_t<R, 4, 7>(g, s, next);
    if (g.max_family <= 32) return launch_gather_t<R, 8, 4>(g, s, next);
    if (g.max_family <= 48) return launch_gather_t<R, 8, 6>(g, s, next);
    return launch_gather_t<R, 8, 8>(g, s, next);
}

hipError_t launch_gather(const GatherArgs& g, int mixed, hipStream_t s, const PosArgs* next)
{
    return mixed ? launch_gather_r<float>(g, s, next) : launch_gather_r<double>(g, s, next);
}

hipError_t launch_positions(const PosArgs& p, hipStream_t s)
{
    const int grid = ceil_div(p.n, 256);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_positions, dim3(grid), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_velocity_bc_list(int count, const int* ids, double2 v0, double2* v, double2* a,
                                   hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_vbc_list, dim3(ceil_div(count, 256)), dim3(256), 0, s, count, ids, v0, v, a);
    return hipGetLastError();
}

hipError_t launch_fixed_bc_list(int count, const int* ids, double2* u, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fixed_list, dim3(ceil_div(count, 256)), dim3(256), 0, s, count, ids, u);
    return hipGetLastError();
}

hipError_t launch_check_finite(int n, const double2* u, const double2* v, const double2* f,
                               unsigned int* first, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, first, 0xffffffffu, 1);
    hipLaunchKernelGGL(k_check_finite, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, u, v, f, first);
    return hipGetLastError();
}

hipError_t launch_count_ill_conditioned(long nb, const void* kinv, bool kinv_f32, const unsigned char* broken,
                                       double cmin, unsigned int* count, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, count, 0u, 1);
    if (nb <= 0) return hipGetLastError();
    const long want = (nb + 255) / 256;
    const int grid = (int)(want < 256L * 16 ? want : 256L * 16);
    const double r = cmin + 2.0 + 1.0 / cmin;
    if (kinv_f32)
        hipLaunchKernelGGL(k_count_ill_conditioned<float>, dim3(grid), dim3(256), 0, s, nb,
                           static_cast<const float*>(kinv), broken, r, count);
    else
        hipLaunchKernelGGL(k_count_ill_conditioned<double>, dim3(grid), dim3(256), 0, s, nb,
                           static_cast<const double*>(kinv), broken, r, count);
    return hipGetLastError();
}

hipError_t launch_length_order(int n, const int* row, const double2* X, int* col, int* bmap, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_length_order, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, row, X, col, bmap);
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int* p, unsigned int v, int count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, v, count);
    return hipGetLastError();
}

hipError_t launch_copy_probe(const double4* src, double4* dst, long n4, hipStream_t s)
{
    if (n4 <= 0) return hipSuccess;
    const long want = (n4 + 255) / 256;
    const int grid = (int)(want < 256L * 64 ? want : 256L * 64);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, src, dst, n4);
    return hipGetLastError();
}

}  // namespace pd
