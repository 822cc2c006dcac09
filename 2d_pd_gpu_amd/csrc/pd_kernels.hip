// pd_kernels.hip -- gfx950 kernels of the bond-associated NOSB peridynamics step.
//
// Math (EricYJA/2D_PD_GPU, src/main_seq.cpp):
//   w_i(j,k) = exp(-n1 |l_ij - l_ik| / (delta dx)) (1/2 + 1/2 cos theta_jk)^n2      :103-118
//   K_ik     = sum_l w_i(k,l) V_l xi_l (x) xi_l                                      :135-140
//   N_ik     = K_ik + D_ik,  D_ik = sum_l w_i(k,l) V_l du_l (x) xi_l  (u = y - X)    :138
//   F_ik - I = D_ik K_ik^-1,  eps = sym(F - I),  sigma = C eps (plane stress)        :224-243
//   T_ij     = (1/V_H) [sum_k w_i(j,k) V_k sigma_ik K_ik^-1] xi_ij                   :298-371
//   f_i      = sum_j (T_ij - T_ji) V_j,  v += 1/2 (a + f/rho) dt                     :411-466
// The reference evaluates K_ik, N_ik inside the j loop (O(M^3) per particle);
// they do not depend on j, and w is symmetric, so each family is evaluated once
// here (O(M^2)).  Displacement form keeps F - I free of cancellation.
//
// Layout: one family per half-wave (families of <= 32 bonds, 2 per wave) or per
// wave (<= 64 bonds).  Lane k owns bond k of its family.  Each lane keeps its
// weight row w(k, .) in registers; the per-bond operands that every lane needs
// (unit bond vector, radial factors, V du (x) xi, later V sigma K^-1) go through
// a per-wave LDS record that the l-loop reads as a broadcast.  No float
// atomics: each family writes its own-bond sum and, into the reverse-bond slot
// of each neighbour's row, the term that neighbour must subtract; the gather
// kernel then reads one contiguous row per particle.
#include "pd_internal.h"

namespace pd {

// Order LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Consecutive blocks land on different XCDs (round-robin dispatch); remap so
// that each XCD works through one contiguous range of families and keeps their
// neighbours in its own L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks, int nxcd)
{
    if (nxcd <= 1) return b;
    int x = b % nxcd, idx = b / nxcd;
    int per = nblocks / nxcd, rem = nblocks % nxcd;
    return x * per + (x < rem ? x : rem) + idx;
}

// Bond stretch with every operation rounded on its own, so (i->j) and (j->i)
// give the same bits; matches the oracle's or_bond_stretch.
__device__ __forceinline__ double bond_stretch(double x0, double x1, double d0, double d1)
{
#pragma clang fp contract(off)
    double e0 = x0 + d0;
    double e1 = x1 + d1;
    double len = sqrt(x0 * x0 + x1 * x1);
    double cur = sqrt(e0 * e0 + e1 * e1);
    return (cur - len) / len;
}

template <typename R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }

// w(k,l) from unit bond vectors and the radial factors p = exp(A l), q = exp(-A l):
// exp(-A |l_k - l_l|) = min(q_k p_l, p_k q_l).
template <typename R, int N2_IS_10>
__device__ __forceinline__ R weight(R exk, R eyk, R pk, R qk, R exl, R eyl, R pl, R ql, R n2)
{
    R c = exk * exl + eyk * eyl;
    R h = R(0.5) + R(0.5) * c;
    h = rmin(rmax(h, R(0)), R(1));
    R hp;
    if constexpr (N2_IS_10) {
        R h2 = h * h;
        R h4 = h2 * h2;
        R h8 = h4 * h4;
        hp = h8 * h2;
    } else {
        hp = pow(h, n2);
    }
    return rmin(qk * pl, pk * ql) * hp;
}

template <> __device__ __forceinline__ float weight<float, 0>(float exk, float eyk, float pk, float qk,
                                                              float exl, float eyl, float pl, float ql,
                                                              float n2)
{
    float c = exk * exl + eyk * eyl;
    float h = 0.5f + 0.5f * c;
    h = rmin(rmax(h, 0.0f), 1.0f);
    return rmin(qk * pl, pk * ql) * powf(h, n2);
}

// Sum over the lanes of one family (LPF lanes, aligned).
template <int LPF>
__device__ __forceinline__ double family_sum(double x)
{
#pragma unroll
    for (int off = LPF / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// FP64 assembly of K_ik over the alive bonds of each family in the wave, and
// its inverse under the scale-relative singularity rule.  rec: this wave's LDS
// records (64 x 8 doubles).  Every lane of the wave must call it.
template <int LPF, int N2_IS_10>
__device__ __forceinline__ void assemble_kinv(double* rec, int lane, int half, bool act, bool alive,
                                              double ex, double ey, double pe, double qe,
                                              double Vxx, double Vxy, double Vyy, double n2,
                                              double ktol, double& k00, double& k01, double& k11,
                                              bool& singular)
{
    double* my = rec + lane * 8;
    my[0] = ex;
    my[1] = ey;
    my[2] = pe;
    my[3] = qe;
    my[4] = alive ? Vxx : 0.0;
    my[5] = alive ? Vxy : 0.0;
    my[6] = alive ? Vyy : 0.0;
    wave_sync();
    const double* base = rec + half * (32 * 8);
    double K00 = 0.0, K01 = 0.0, K11 = 0.0;
#pragma unroll 2
    for (int l = 0; l < LPF; ++l) {
        const double* r = base + l * 8;
        double w = weight<double, N2_IS_10>(ex, ey, pe, qe, r[0], r[1], r[2], r[3], n2);
        K00 += w * r[4];
        K01 += w * r[5];
        K11 += w * r[6];
    }
    wave_sync();
    double det = K00 * K11 - K01 * K01;
    double tr = K00 + K11;
    singular = false;
    if (alive && det > ktol * tr * tr) {
        k00 = K11 / det;
        k01 = -K01 / det;
        k11 = K00 / det;
    } else {
        k00 = k01 = k11 = 0.0;
        singular = alive;
    }
    (void)act;
}

template <typename R> struct Vec2;
template <> struct Vec2<double> { typedef double2 T; };
template <> struct Vec2<float> { typedef float2 T; };

// Per-lane bond setup shared by the shape and bond kernels.
struct BondLane {
    int fam, k, b, j, M;
    bool famok, act;
    double x0, x1, d0, d1, len, Vj, Vi;
};

template <int LPF>
__device__ __forceinline__ BondLane load_lane(const BondArgs& a, int lane, int half, int wave_global)
{
    constexpr int FPW = 64 / LPF;
    BondLane L;
    L.k = (LPF == 32) ? (lane & 31) : lane;
    L.fam = wave_global * FPW + half;
    L.famok = L.fam < a.n;
    const int fi = L.famok ? L.fam : 0;
    int beg = 0, M = 0;
    if (L.famok) {
        beg = a.row[fi];
        M = a.row[fi + 1] - beg;
    }
    L.M = M;
    L.act = L.famok && L.k < M;
    L.b = beg + L.k;
    L.j = L.act ? a.col[L.b] : fi;
    const double2 Xi = a.X[fi], Xj = a.X[L.j];
    const double2 ui = a.u[fi], uj = a.u[L.j];
    L.x0 = Xj.x - Xi.x;
    L.x1 = Xj.y - Xi.y;
    L.d0 = uj.x - ui.x;
    L.d1 = uj.y - ui.y;
    L.len = sqrt(L.x0 * L.x0 + L.x1 * L.x1);
    L.Vj = a.V[L.j];
    L.Vi = a.V[fi];
    return L;
}

// K_ik^-1 for every bond of every family (context creation, pd_set_broken).
template <int LPF, int N2_IS_10>
__global__ void __launch_bounds__(512) k_shape(BondArgs a)
{
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int half = (LPF == 32) ? (lane >> 5) : 0;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, a.xcd);
    BondLane L = load_lane<LPF>(a, lane, half, blk * wpb + wib);
    const bool alive = L.act && a.broken[L.b] == 0;
    const double il = L.act ? 1.0 / L.len : 0.0;
    const double ex = L.x0 * il, ey = L.x1 * il;
    const double pe = L.act ? exp(a.A * L.len) : 0.0;
    const double qe = L.act ? exp(-a.A * L.len) : 0.0;
    double k00, k01, k11;
    bool sing;
    assemble_kinv<LPF, N2_IS_10>(smem + wib * 64 * 8, lane, half, L.act, alive, ex, ey, pe, qe,
                                 L.Vj * L.x0 * L.x0, L.Vj * L.x0 * L.x1, L.Vj * L.x1 * L.x1,
                                 a.n2, a.ktol, k00, k01, k11, sing);
    if (L.act) {
        a.kinv[3 * (size_t)L.b] = k00;
        a.kinv[3 * (size_t)L.b + 1] = k01;
        a.kinv[3 * (size_t)L.b + 2] = k11;
    }
    const uint64_t sm = __ballot(sing);
    if (lane == 0 && sm) atomicAdd(a.counters + 1, (unsigned int)__popcll(sm));
}

// One explicit force evaluation: bond breaking, K re-assembly for families that
// lost a bond, then the two family contractions and the force-state outputs.
template <typename R, int LPF, int N2_IS_10>
__global__ void __launch_bounds__(512) k_bond(BondArgs a)
{
    typedef typename Vec2<R>::T R2;
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int half = (LPF == 32) ? (lane >> 5) : 0;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, a.xcd);
    BondLane L = load_lane<LPF>(a, lane, half, blk * wpb + wib);
    double* rec64 = smem + wib * 64 * 8;

    // ---- bond breaking (FP64, irreversible, both directions agree) ----
    bool alive = L.act && a.broken[L.b] == 0;
    bool newly = false;
    if (a.s0 > 0.0 && alive) {
        double s = bond_stretch(L.x0, L.x1, L.d0, L.d1);
        if (s > a.s0) {
            alive = false;
            newly = true;
            a.broken[L.b] = 1;
        }
    }
    const uint64_t nm = __ballot(newly);
    if (lane == 0 && nm) atomicAdd(a.counters, (unsigned int)__popcll(nm));

    // ---- geometry (FP64) ----
    const double il = L.act ? 1.0 / L.len : 0.0;
    const double ex = L.x0 * il, ey = L.x1 * il;
    const double pe = L.act ? exp(a.A * L.len) : 0.0;
    const double qe = L.act ? exp(-a.A * L.len) : 0.0;

    // ---- K_ik^-1: stored, or re-assembled in FP64 when the family changed ----
    double k00 = 0.0, k01 = 0.0, k11 = 0.0;
    if (nm != 0 || a.force_k) {
        bool sing;
        assemble_kinv<LPF, N2_IS_10>(rec64, lane, half, L.act, alive, ex, ey, pe, qe,
                                     L.Vj * L.x0 * L.x0, L.Vj * L.x0 * L.x1, L.Vj * L.x1 * L.x1,
                                     a.n2, a.ktol, k00, k01, k11, sing);
        if (L.act) {
            a.kinv[3 * (size_t)L.b] = k00;
            a.kinv[3 * (size_t)L.b + 1] = k01;
            a.kinv[3 * (size_t)L.b + 2] = k11;
        }
        const uint64_t sm = __ballot(sing);
        if (lane == 0 && sm) atomicAdd(a.counters + 1, (unsigned int)__popcll(sm));
    } else if (alive) {
        k00 = a.kinv[3 * (size_t)L.b];
        k01 = a.kinv[3 * (size_t)L.b + 1];
        k11 = a.kinv[3 * (size_t)L.b + 2];
    }

    // ---- records for the first contraction ----
    R* rec = reinterpret_cast<R*>(rec64);
    const R exr = R(ex), eyr = R(ey), per = R(pe), qer = R(qe);
    const R x0r = R(L.x0), x1r = R(L.x1);
    const R Vd0 = alive ? R(L.Vj * L.d0) : R(0), Vd1 = alive ? R(L.Vj * L.d1) : R(0);
    {
        R* my = rec + lane * 8;
        my[0] = exr;
        my[1] = eyr;
        my[2] = per;
        my[3] = qer;
        my[4] = Vd0 * x0r;
        my[5] = Vd0 * x1r;
        my[6] = Vd1 * x0r;
        my[7] = Vd1 * x1r;
    }
    // loop bound: largest family in the wave
    int mfam = L.M;
    if (LPF == 32) mfam = max(mfam, __shfl_xor(mfam, 32, 64));
    const int mmax = __builtin_amdgcn_readfirstlane(mfam);
    wave_sync();

    // ---- contraction 1: D_ik = sum_l w(k,l) V_l du_l (x) xi_l ----
    const R* base = rec + half * (32 * 8);
    const R n2r = R(a.n2);
    R w[LPF];
    R D00 = 0, D01 = 0, D10 = 0, D11 = 0;
#pragma unroll
    for (int l = 0; l < LPF; ++l) {
        if (l < mmax) {
            const R* r = base + l * 8;
            const R om = weight<R, N2_IS_10>(exr, eyr, per, qer, r[0], r[1], r[2], r[3], n2r);
            w[l] = om;
            D00 += om * r[4];
            D01 += om * r[5];
            D10 += om * r[6];
            D11 += om * r[7];
        } else {
            w[l] = R(0);
        }
        // keep the scheduler from hoisting every record load of the unrolled
        // loop (which would spill)
        if ((l & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }

    // ---- F - I = D K^-1, plane-stress sigma, Q_k = V_k sigma K^-1 ----
    const double Ei = a.E ? a.E[L.famok ? L.fam : 0] : a.E0;
    const double pf = Ei / (1.0 - a.nu * a.nu);
    const R c11 = R(pf), c12 = R(pf * a.nu), c33 = R(pf * (1.0 - a.nu));
    const R K00 = R(k00), K01 = R(k01), K11 = R(k11);
    const R G00 = D00 * K00 + D01 * K01;
    const R G01 = D00 * K01 + D01 * K11;
    const R G10 = D10 * K00 + D11 * K01;
    const R G11 = D10 * K01 + D11 * K11;
    const R e11 = G00, e22 = G11, e12 = R(0.5) * (G01 + G10);
    const R s11 = c11 * e11 + c12 * e22;
    const R s22 = c12 * e11 + c11 * e22;
    const R s12 = c33 * e12;
    const R Vk = alive ? R(L.Vj) : R(0);
    const R Q00 = (s11 * K00 + s12 * K01) * Vk;
    const R Q01 = (s11 * K01 + s12 * K11) * Vk;
    const R Q10 = (s12 * K00 + s22 * K01) * Vk;
    const R Q11 = (s12 * K01 + s22 * K11) * Vk;
    wave_sync();
    {
        R* my = rec + lane * 8;
        my[4] = Q00;
        my[5] = Q01;
        my[6] = Q10;
        my[7] = Q11;
    }
    wave_sync();

    // ---- contraction 2: S_ij = sum_k w(j,k) V_k sigma_ik K_ik^-1 ----
    R S00 = 0, S01 = 0, S10 = 0, S11 = 0;
#pragma unroll
    for (int l = 0; l < LPF; ++l) {
        if (l < mmax) {
            const R* r = base + l * 8;
            S00 += w[l] * r[4];
            S01 += w[l] * r[5];
            S10 += w[l] * r[6];
            S11 += w[l] * r[7];
        }
        if ((l & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    const double VH = family_sum<LPF>(alive ? L.Vj : 0.0);
    R T0 = 0, T1 = 0;
    if (alive) {
        const R ivh = R(1.0 / VH);
        T0 = (S00 * x0r + S01 * x1r) * ivh;
        T1 = (S10 * x0r + S11 * x1r) * ivh;
    }

    // ---- outputs: own-bond sum and the reverse-slot terms ----
    const double fo0 = family_sum<LPF>(double(T0) * L.Vj);
    const double fo1 = family_sum<LPF>(double(T1) * L.Vj);
    if (L.famok && L.k == 0) a.fown[L.fam] = make_double2(fo0, fo1);
    if (L.act) {
        R2 t;
        t.x = T0 * R(L.Vi);
        t.y = T1 * R(L.Vi);
        reinterpret_cast<R2*>(a.W)[a.rev[L.b]] = t;
    }
}

// f_i = sum_j T_ij V_j - sum_j T_ji V_i ; v += 1/2 (a + f/rho) dt.
template <typename R>
__global__ void __launch_bounds__(256) k_gather(GatherArgs g)
{
    typedef typename Vec2<R>::T R2;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, g.xcd);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const R2* W = reinterpret_cast<const R2*>(g.W);
    double f0 = 0.0, f1 = 0.0;
    const int e = g.row[i + 1];
    for (int b = g.row[i]; b < e; ++b) {
        const R2 t = W[b];
        f0 += double(t.x);
        f1 += double(t.y);
    }
    const double2 fo = g.fown[i];
    f0 = fo.x - f0;
    f1 = fo.y - f1;
    {
#pragma clang fp contract(off)
        const double2 a0 = g.a[i];
        double2 v = g.v[i];
        const double an0 = f0 / g.rho, an1 = f1 / g.rho;
        v.x = v.x + (0.5 * (a0.x + an0)) * g.dt;
        v.y = v.y + (0.5 * (a0.y + an1)) * g.dt;
        g.v[i] = v;
        g.f[i] = make_double2(f0, f1);
        if (g.write_a) g.a[i] = make_double2(an0, an1);
    }
}

// Velocity BC, a = f/rho, u += v dt + 1/2 a dt^2, fixed BC (one step's
// prologue, src/main_seq.cpp:686-692).
__global__ void __launch_bounds__(256) k_positions(PosArgs p)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int code = p.bc ? p.bc[i] : -1;
    double2 v = p.v[i];
    if (code >= 0) {
        v = p.v0[code];
        p.v[i] = v;
    }
    const double2 f = p.f[i];
    const double2 acc = make_double2(f.x / p.rho, f.y / p.rho);
    p.a[i] = acc;
    double2 u = p.u[i];
    const double dt2 = p.dt * p.dt;
    u.x = u.x + (v.x * p.dt + (0.5 * acc.x) * dt2);
    u.y = u.y + (v.y * p.dt + (0.5 * acc.y) * dt2);
    if (code == -2) u = make_double2(0.0, 0.0);
    p.u[i] = u;
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

__global__ void k_fill_u32(unsigned int* p, unsigned int v, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = v;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

template <typename R, int LPF, int N2>
static hipError_t launch_bond_t(const BondArgs& a, int wpb, hipStream_t s)
{
    const int fpw = 64 / LPF;
    const int grid = ceil_div(a.n, (long)wpb * fpw);
    if (grid == 0) return hipSuccess;
    const size_t lds = (size_t)wpb * 64 * 8 * sizeof(double);
    hipLaunchKernelGGL((k_bond<R, LPF, N2>), dim3(grid), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_bond(const BondArgs& a, int mixed, int lpf, int n2_is_10, int wpb, hipStream_t s)
{
    if (mixed) {
        if (lpf == 32) return n2_is_10 ? launch_bond_t<float, 32, 1>(a, wpb, s) : launch_bond_t<float, 32, 0>(a, wpb, s);
        return n2_is_10 ? launch_bond_t<float, 64, 1>(a, wpb, s) : launch_bond_t<float, 64, 0>(a, wpb, s);
    }
    if (lpf == 32) return n2_is_10 ? launch_bond_t<double, 32, 1>(a, wpb, s) : launch_bond_t<double, 32, 0>(a, wpb, s);
    return n2_is_10 ? launch_bond_t<double, 64, 1>(a, wpb, s) : launch_bond_t<double, 64, 0>(a, wpb, s);
}

template <int LPF, int N2>
static hipError_t launch_shape_t(const BondArgs& a, int wpb, hipStream_t s)
{
    const int fpw = 64 / LPF;
    const int grid = ceil_div(a.n, (long)wpb * fpw);
    if (grid == 0) return hipSuccess;
    const size_t lds = (size_t)wpb * 64 * 8 * sizeof(double);
    hipLaunchKernelGGL((k_shape<LPF, N2>), dim3(grid), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_shape(const BondArgs& a, int lpf, int n2_is_10, int wpb, hipStream_t s)
{
    if (lpf == 32) return n2_is_10 ? launch_shape_t<32, 1>(a, wpb, s) : launch_shape_t<32, 0>(a, wpb, s);
    return n2_is_10 ? launch_shape_t<64, 1>(a, wpb, s) : launch_shape_t<64, 0>(a, wpb, s);
}

hipError_t launch_gather(const GatherArgs& g, int mixed, hipStream_t s)
{
    const int grid = ceil_div(g.n, 256);
    if (grid == 0) return hipSuccess;
    if (mixed)
        hipLaunchKernelGGL(k_gather<float>, dim3(grid), dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL(k_gather<double>, dim3(grid), dim3(256), 0, s, g);
    return hipGetLastError();
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

hipError_t launch_fill_u32(unsigned int* p, unsigned int v, int count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, v, count);
    return hipGetLastError();
}

}  // namespace pd
