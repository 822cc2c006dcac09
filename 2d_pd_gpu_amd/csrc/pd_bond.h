// pd_bond.h -- device code of the bond kernel (family contractions, breaking,
// force-state outputs) shared by the per-precision instantiation units.
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
// Layout: a family of up to FS = LPF * BPL bonds is handled by LPF lanes, each
// owning BPL bonds; a wave holds 64 / LPF families.  Each lane keeps its weight
// rows w(k, .) in registers; the per-bond operands that every lane of a family
// needs (unit bond vector, radial factors, V du (x) xi, later V sigma K^-1) go
// through per-wave LDS records that the l-loop reads as broadcasts.  No float
// atomics: each family writes its own-bond sum and, into the reverse-bond slot
// of each neighbour's row, the term that neighbour must subtract; the gather
// kernel then reads one contiguous row per particle.
#pragma once

#include <type_traits>

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

// Breaking test s = (|xi + du| - |xi|) / |xi| > s0, evaluated as
// |xi + du|^2 > c |xi|^2 with c = (1 + s0)^2 and every operation rounded on its
// own, so (i->j) and (j->i) give the same bits; matches the oracle's
// or_bond_breaks.
__device__ __forceinline__ bool bond_breaks(double x0, double x1, double d0, double d1, double c)
{
#pragma clang fp contract(off)
    const double e0 = x0 + d0;
    const double e1 = x1 + d1;
    const double cur2 = e0 * e0 + e1 * e1;
    const double len2 = x0 * x0 + x1 * x1;
    return cur2 > c * len2;
}

// Mazars bond damage for one bond's strain (e11, e22, e12): restatement of
// computeDamageStatus (src/main_seq.cpp:473-598).  Closed-form principal strains
// l1 <= l2 replace Eigen's SelfAdjointEigenSolver (:523-526); eps_eq =
// sqrt(sum <l>+^2) (:529-536); past the threshold (:540) the tension and
// compression parts share eps's principal directions, so their ascending
// spectra are <l>+ and <l>- (the reference accumulates them into uninitialised
// matrices, :509, :513; here they start at zero), giving alpha_t, alpha_c
// (:571-578) and d = alpha_t d_t + alpha_c d_c (:581-584).  Mode 1 puts the
// running threshold into the law as :581-582 do; mode 2 the initial kappa0.
// Returns true when kappa and d changed.
__device__ __forceinline__ bool mazars_update(const BondArgs& a, double e11, double e22, double e12,
                                              double& kappa, double& d)
{
    const double m = 0.5 * (e11 + e22);
    const double h = 0.5 * (e11 - e22);
    const double r = sqrt(h * h + e12 * e12);
    const double l1 = m - r, l2 = m + r;
    const double t1 = l1 > 0.0 ? l1 : 0.0, t2 = l2 > 0.0 ? l2 : 0.0;
    const double c1 = l1 < 0.0 ? l1 : 0.0, c2 = l2 < 0.0 ? l2 : 0.0;
    const double s2 = t1 * t1 + t2 * t2;
    const double eq = s2 > 0.0 ? sqrt(s2) : 0.0;
    if (!(eq > kappa)) return false;
    const double e2 = eq * eq;
    double at = 0.0, ac = 0.0;
    if (t1 > 0.0) at += t1 * (t1 + c1) / e2;
    if (t2 > 0.0) at += t2 * (t2 + c2) / e2;
    if (c1 > 0.0) ac += c1 * (t1 + c1) / e2;
    if (c2 > 0.0) ac += c2 * (t2 + c2) / e2;
    const double kr = a.dmg_mode == 2 ? a.kap0 : kappa;
    const double dt = 1.0 - kr * (1.0 - a.At) / eq - a.At / exp(a.Bt * (eq - kr));
    const double dc = 1.0 - kr * (1.0 - a.Ac) / eq - a.Ac / exp(a.Bc * (eq - kr));
    d = at * dt + ac * dc;
    kappa = eq;
    return true;
}

// One v_min (fmin / fminf would first quiet an operand loaded from LDS with a
// v_max of it with itself; the operands here are finite).  Inline asm gets no
// hazard wait states from the compiler: only for operands computed by the
// VALU or loaded from memory, never for matrix-core results.
__device__ __forceinline__ double rmin(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float rmin(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double rmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double rpow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float rpow(float a, float b) { return powf(a, b); }

// w(k,l) = exp(-A |l_k - l_l|) (1/2 + 1/2 cos theta_kl)^n2 from the k-side half
// unit vector (hx, hy) = e_k / 2, the l-side unit vector, and the radial
// factors p = exp(A l), q = exp(-A l): exp(-A |l_k - l_l|) = min(q_k p_l, p_k q_l).
// 1/2 + 1/2 cos lies in [0, 1] up to rounding; for n2 = 10 the power is four
// squarings and needs no clamp (a value a few ulp outside moves w by a few ulp),
// the general power clamps at 0 (src/main_seq.cpp:115 clamps cos to [-1, 1]).
template <typename R, int N2_IS_10>
__device__ __forceinline__ R weight(R hxk, R hyk, R pk, R qk, R exl, R eyl, R pl, R ql, R n2)
{
    const R h = hxk * exl + (hyk * eyl + R(0.5));
    R hp;
    if constexpr (N2_IS_10) {
        const R h2 = h * h;
        const R h4 = h2 * h2;
        const R h8 = h4 * h4;
        hp = h8 * h2;
    } else {
        hp = rpow(rmax(h, R(0)), n2);
    }
    return rmin(qk * pl, pk * ql) * hp;
}

// Sum over the lanes of one family (LPF lanes, aligned).
template <int LPF>
__device__ __forceinline__ float family_sum(float x)
{
#pragma unroll
    for (int off = LPF / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <int LPF>
__device__ __forceinline__ double family_sum(double x)
{
#pragma unroll
    for (int off = LPF / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <typename R> struct Vec2;
template <> struct Vec2<double> { typedef double2 T; };
template <> struct Vec2<float> { typedef float2 T; };

// Launch tile: LPF lanes per family, BPL bonds per lane (lane t owns bonds
// t, t + LPF, ...), so a family has FS = LPF * BPL slots and a wave holds
// FPW = 64 / LPF families.  More bonds per lane cut the LDS broadcast traffic
// per weight evaluation (each record read feeds BPL rows) at the cost of
// registers for BPL weight rows.
template <int LPF, int BPL> struct Tile {
    static constexpr int FS = LPF * BPL;
    static constexpr int FPW = 64 / LPF;
    // LDS per wave, in doubles: FS rows of FPW FP64 records (8 values each)
    // plus one 16-byte pad chunk per row (see rec_at)
    static constexpr int LDS_DOUBLES = FS * (4 * FPW + 1) * 2;
};

// LDS record layout.  Row k holds record k of each of the wave's FPW
// families; value e (0..7) sits in 16-byte chunk e / VPC, and the chunks are
// interleaved across the families, so one broadcast read (every lane of a
// family asks for the same chunk) touches FPW consecutive 16-byte words in
// distinct banks.  A row is an odd number of chunks long (one pad chunk), so
// the record writes, where the lanes of a family write consecutive rows, fall
// into distinct 16-byte bank groups instead of all into one.
template <typename T, int FPW>
__device__ __forceinline__ int rec_at(int k, int slot, int e)
{
    constexpr int VPC = 16 / (int)sizeof(T);
    constexpr int NCH = 8 / VPC;
    constexpr int ROW = (NCH * FPW + 1) * VPC;
    return k * ROW + ((e / VPC) * FPW + slot) * VPC + e % VPC;
}

// Whole-chunk record writes (one 16-byte LDS store per chunk c).
template <int FPW>
__device__ __forceinline__ void rec_put(float* rec, int k, int slot, int c, float v0, float v1, float v2,
                                        float v3)
{
    *reinterpret_cast<float4*>(rec + rec_at<float, FPW>(k, slot, 4 * c)) = make_float4(v0, v1, v2, v3);
}

template <int FPW>
__device__ __forceinline__ void rec_put(double* rec, int k, int slot, int c, double v0, double v1)
{
    *reinterpret_cast<double2*>(rec + rec_at<double, FPW>(k, slot, 2 * c)) = make_double2(v0, v1);
}

// Symmetric weight blocks (mixed-precision rank form, 16 lanes x 3 bonds):
// lane t of a family holds w(16 S + t, 16 B + j), j = 0..15, of block (S, B),
// S < B.  Since w(k, l) = w(l, k), block (B, S) is its transpose: the lanes
// write their rows into the family's 16 x 17 LDS tile at lds + tb (rows of
// 17 floats, so the column reads fall into distinct banks), read column t back as
// w(16 B + t, 16 S + j) and add those weights times the records' values
// rec[16 S + j][4 + lc] into slot B's matrix-core accumulator.
typedef float v4f __attribute__((ext_vector_type(4)));
template <int FPW, int S, int B, int FS>
__device__ __forceinline__ void trx_block(const float (&w)[3][FS], float* lds, int tb, int t, int slot, int lc,
                                          int mmax, v4f& acc)
{
    if (16 * B >= mmax) return;   // rows 16 B .. of every family in the wave are empty
    // the lane's row and column starts in the tile, opaque to the compiler
    // (an empty asm emits nothing) so that each access is one base register
    // plus an immediate offset instead of an add per access
    int rw = tb + 17 * t, cl = tb + t;
    asm("" : "+v"(rw));
    asm("" : "+v"(cl));
#pragma unroll
    for (int j = 0; j < 16; ++j) lds[rw + j] = w[S][16 * B + j];
    wave_sync();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float om = lds[cl + 17 * j];
        const float pa = lds[rec_at<float, FPW>(16 * S + j, slot, 4 + lc)];
        acc = __builtin_amdgcn_mfma_f32_4x4x1f32(pa, om, acc, 0, 0, 0);
    }
    wave_sync();
}

// Values 4h .. 4h + 3 of record k.
template <int FPW>
__device__ __forceinline__ void rec_put4(float* rec, int k, int slot, int h, float v0, float v1, float v2,
                                         float v3)
{
    rec_put<FPW>(rec, k, slot, h, v0, v1, v2, v3);
}

template <int FPW>
__device__ __forceinline__ void rec_put4(double* rec, int k, int slot, int h, double v0, double v1,
                                         double v2, double v3)
{
    rec_put<FPW>(rec, k, slot, 2 * h, v0, v1);
    rec_put<FPW>(rec, k, slot, 2 * h + 1, v2, v3);
}

// The family a lane works on.  Loads are issued without branches (indices
// clamped to valid entries, values masked where they are used), so that a
// wave can keep a later group's loads in flight across its current work: the
// vector-memory counter is waited on in issue order, and a load skipped on one
// path makes every wait after the merge a full drain.
struct LaneFam {
    int fam, t, slot, r0, r1, beg, M;
    bool famok;
    double2 Xi, ui;
    double Vi, pf;
};

template <int LPF>
__device__ __forceinline__ LaneFam load_family(const BondArgs& a, int lane, int wave_global)
{
    LaneFam F;
    F.slot = lane / LPF;
    F.t = lane % LPF;
    F.fam = wave_global * (64 / LPF) + F.slot;
    F.famok = F.fam < a.n;
    const int fi = F.famok ? F.fam : 0;
    F.r0 = a.row[fi];
    F.r1 = a.row[fi + 1];
    F.Xi = a.X[fi];
    F.ui = a.u[fi];
    F.Vi = a.V[fi];
    F.pf = a.pf[fi & a.pf_mask];
    return F;
}

// Family extent from the row pointers (first use of the row loads).
__device__ __forceinline__ void family_extent(LaneFam& F)
{
    F.beg = F.r0;
    F.M = F.famok ? F.r1 - F.r0 : 0;
}

// One bond slot of a lane.
struct BondSlot {
    int b, k;
    bool act;
    double x0, x1, d0, d1, Vj;
};

__device__ __forceinline__ BondSlot load_slot(const BondArgs& a, const LaneFam& F, int k)
{
    BondSlot S;
    S.k = k;
    S.act = F.famok && k < F.M;
    S.b = F.beg + k;
    const int j = S.act ? a.col[S.b] : (F.famok ? F.fam : 0);
    const double2 Xj = a.X[j], uj = a.u[j];
    S.x0 = Xj.x - F.Xi.x;
    S.x1 = Xj.y - F.Xi.y;
    S.d0 = uj.x - F.ui.x;
    S.d1 = uj.y - F.ui.y;
    S.Vj = a.V[j];
    return S;
}

// Bond geometry for the weight: unit vector e, half vector e/2 and the radial
// factors exp(+-A l).  Inactive slots get all zeros, so their weight is 0.
template <typename R> struct Geo {
    R ex, ey, hx, hy, pe, qe, len;
};

// FP64 1/sqrt(x) and 1/x for positive normal x: the hardware estimate and two
// Newton steps (quadratic convergence from ~2^-23 leaves a final error of an
// ulp or two, far inside the FP64 parity budget), instead of the
// correctly-rounded sqrt, division and a second exp.
__device__ __forceinline__ double rsqrt64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = fma(y, fma(-hx * y, y, 0.5), y);
    y = fma(y, fma(-hx * y, y, 0.5), y);
    return y;
}

__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

__device__ __forceinline__ Geo<double> geometry64(const BondArgs& a, const BondSlot& S, double Ae)
{
    Geo<double> g;
    // branch-free: an inactive slot's math runs on a harmless value and is
    // selected away, so a lane's slots schedule together
    const double l2 = S.x0 * S.x0 + S.x1 * S.x1;
    const double ilr = rsqrt64(S.act ? l2 : 1.0);
    const double il = S.act ? ilr : 0.0;
    const double len = l2 * il;
    g.len = len;
    g.ex = S.x0 * il;
    g.ey = S.x1 * il;
    g.hx = 0.5 * g.ex;
    g.hy = 0.5 * g.ey;
    const double pr = exp(Ae * (S.act ? len : 0.0));
    g.pe = S.act ? pr : 0.0;
    const double qr = rcp64(pr);
    g.qe = S.act ? qr : 0.0;
    return g;
}

// FP32 geometry for the mixed-precision contractions (hardware rsq / exp / rcp).
__device__ __forceinline__ Geo<float> geometry32(const BondArgs& a, const BondSlot& S, double Ae)
{
    Geo<float> g;
    const float x0 = float(S.x0), x1 = float(S.x1);
    const float l2 = x0 * x0 + x1 * x1;
    const float il = S.act ? __builtin_amdgcn_rsqf(l2) : 0.0f;
    const float len = l2 * il;
    g.len = len;
    g.ex = x0 * il;
    g.ey = x1 * il;
    g.hx = 0.5f * g.ex;
    g.hy = 0.5f * g.ey;
    const float Al = float(Ae) * len;
    g.pe = S.act ? __expf(Al) : 0.0f;
    g.qe = S.act ? __builtin_amdgcn_rcpf(g.pe) : 0.0f;
    return g;
}

template <typename R> struct GeoOf;
template <> struct GeoOf<double> {
    static __device__ __forceinline__ Geo<double> get(const BondArgs& a, const BondSlot& S, double Ae)
    {
        return geometry64(a, S, Ae);
    }
};
template <> struct GeoOf<float> {
    static __device__ __forceinline__ Geo<float> get(const BondArgs& a, const BondSlot& S, double Ae)
    {
        return geometry32(a, S, Ae);
    }
};

// FP64 assembly of K_ik over the alive bonds of each family in the wave, and
// its inverse under the scale-relative singularity rule.  rec: this wave's LDS
// records (Tile::LDS_DOUBLES).  Every lane of the wave must call it.
template <int LPF, int BPL, int N2_IS_10>
__device__ __forceinline__ void assemble_kinv(double* rec, const LaneFam& F, const BondSlot* S,
                                              const bool* alive, const Geo<double>* g, double n2,
                                              double ktol, double (*kinv)[3], bool* singular)
{
    constexpr int FS = Tile<LPF, BPL>::FS;
    constexpr int FPW = Tile<LPF, BPL>::FPW;
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        const int k = S[s].k;
        const double V = alive[s] ? S[s].Vj : 0.0;
        rec_put<FPW>(rec, k, F.slot, 0, g[s].ex, g[s].ey);
        rec_put<FPW>(rec, k, F.slot, 1, g[s].pe, g[s].qe);
        rec_put<FPW>(rec, k, F.slot, 2, V * S[s].x0 * S[s].x0, V * S[s].x0 * S[s].x1);
        rec_put<FPW>(rec, k, F.slot, 3, V * S[s].x1 * S[s].x1, 0.0);
    }
    wave_sync();
    double K[BPL][3];
#pragma unroll
    for (int s = 0; s < BPL; ++s) K[s][0] = K[s][1] = K[s][2] = 0.0;
#pragma unroll 2
    for (int l = 0; l < FS; ++l) {
        const double ex = rec[rec_at<double, FPW>(l, F.slot, 0)];
        const double ey = rec[rec_at<double, FPW>(l, F.slot, 1)];
        const double pe = rec[rec_at<double, FPW>(l, F.slot, 2)];
        const double qe = rec[rec_at<double, FPW>(l, F.slot, 3)];
        const double kxx = rec[rec_at<double, FPW>(l, F.slot, 4)];
        const double kxy = rec[rec_at<double, FPW>(l, F.slot, 5)];
        const double kyy = rec[rec_at<double, FPW>(l, F.slot, 6)];
#pragma unroll
        for (int s = 0; s < BPL; ++s) {
            const double w = weight<double, N2_IS_10>(g[s].hx, g[s].hy, g[s].pe, g[s].qe, ex, ey, pe, qe, n2);
            K[s][0] += w * kxx;
            K[s][1] += w * kxy;
            K[s][2] += w * kyy;
        }
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        const double det = K[s][0] * K[s][2] - K[s][1] * K[s][1];
        const double tr = K[s][0] + K[s][2];
        singular[s] = false;
        if (alive[s] && det > ktol * tr * tr) {
            kinv[s][0] = K[s][2] / det;
            kinv[s][1] = -K[s][1] / det;
            kinv[s][2] = K[s][0] / det;
        } else {
            kinv[s][0] = kinv[s][1] = kinv[s][2] = 0.0;
            singular[s] = alive[s];
        }
    }
}

template <int LPF>
__device__ __forceinline__ int wave_max_family(int m)
{
#pragma unroll
    for (int off = LPF; off < 64; off <<= 1) m = max(m, __shfl_xor(m, off, 64));
    return __builtin_amdgcn_readfirstlane(m);
}

// An alive, non-singular K whose condition number reaches the guard's bound,
// read from its FP64 inverse (same eigenvectors, reciprocal eigenvalues): for
// eigenvalues l1 >= l2 > 0 and c = l1 / l2, tr^2 / det = c + 2 + 1/c.
__device__ __forceinline__ bool kinv_guarded(const double* k, double illr)
{
    const double det = k[0] * k[2] - k[1] * k[1], tr = k[0] + k[2];
    const bool zero = k[0] == 0.0 && k[1] == 0.0 && k[2] == 0.0;
    return !zero && (det <= 0.0 || tr * tr >= illr * det);
}

// Whether any lane of the caller's family (LPF lanes, aligned) has x set.
template <int LPF>
__device__ __forceinline__ bool family_any(bool x, int slot)
{
    const unsigned long long b = __ballot(x);
    const unsigned long long m = LPF == 64 ? ~0ull : ((1ull << LPF) - 1) << (slot * LPF);
    return (b & m) != 0;
}

// K^-1 storage: FP64 values rounded to the bond arithmetic's precision (the
// mixed-precision kernel rounds them to FP32 before use anyway).
template <typename KT>
__device__ __forceinline__ void store_kinv(const BondArgs& a, int b, const double* k)
{
    KT* d = reinterpret_cast<KT*>(a.kinv) + 3 * (size_t)b;
    d[0] = KT(k[0]);
    d[1] = KT(k[1]);
    d[2] = KT(k[2]);
}

// K_ik^-1 for every bond of every family (context creation, pd_set_broken).
template <int LPF, int BPL, int N2_IS_10>
__global__ void __launch_bounds__(512) k_shape(BondArgs a)
{
    extern __shared__ __attribute__((aligned(16))) double smem[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, a.xcd);
    LaneFam F = load_family<LPF>(a, lane, blk * wpb + wib);
    family_extent(F);
    BondSlot S[BPL];
    bool alive[BPL];
    Geo<double> g[BPL];
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        S[s] = load_slot(a, F, F.t + s * LPF);
        alive[s] = S[s].act && a.broken[S[s].b] == 0;
        g[s] = geometry64(a, S[s], a.A);
    }
    double kinv[BPL][3];
    bool sing[BPL];
    assemble_kinv<LPF, BPL, N2_IS_10>(smem + wib * Tile<LPF, BPL>::LDS_DOUBLES, F, S, alive, g, a.n2,
                                      a.ktol, kinv, sing);
    int nsing = 0;
    bool guard = false;
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        if (S[s].act) {
            if (a.kinv_f32)
                store_kinv<float>(a, S[s].b, kinv[s]);
            else
                store_kinv<double>(a, S[s].b, kinv[s]);
            guard = guard || kinv_guarded(kinv[s], a.illr);
        }
        nsing += __popcll(__ballot(sing[s]));
    }
    if (lane == 0 && nsing) atomicAdd(a.counters + 1, (unsigned int)nsing);
    if (a.famflag) {
        guard = family_any<LPF>(guard, F.slot);
        if (F.famok && F.t == 0) a.famflag[F.fam] = guard ? 1 : 0;
    }
}

// Global inputs of one family group for one lane, gathered in three dependent
// stages: row pointers; then everything indexed by the bond (neighbour id,
// reverse bond, broken flag, K^-1, damage state); then the neighbours' state.
// The stages only issue loads into these raw fields; the values are first
// used in finish_loads at the start of the group's body, so that a persistent
// wave can keep its next group's loads in flight across the current group.
template <typename R, int BPL> struct FamLoad {
    LaneFam F;
    BondSlot S[BPL];
    int jr[BPL];              // neighbour id as loaded (valid where S.act)
    int rv[BPL];              // reverse bond
    unsigned char br[BPL];
    R kraw[BPL][3];           // K^-1 as stored
    double2 Xj[BPL], uj[BPL];
    double Vj[BPL];
    double kap[BPL], dm[BPL];  // damage threshold and value (damage kernels)
    unsigned char ff;          // the family's guard flag (mixed precision)
    static constexpr bool GUARD = sizeof(R) == 4;
};

template <int LPF, typename FL>
__device__ __forceinline__ void stage_rows(const BondArgs& a, FL& L, int lane, int group)
{
    L.F = load_family<LPF>(a, lane, group);
    if constexpr (FL::GUARD) L.ff = a.famflag[L.F.famok ? L.F.fam : 0];
}

template <typename R, int LPF, int BPL, int DMG = 0, int FIX = 0>
__device__ __forceinline__ void stage_cols(const BondArgs& a, FamLoad<R, BPL>& L)
{
    family_extent(L.F);
    const R* kv = reinterpret_cast<const R*>(a.kinv);
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        BondSlot& S = L.S[s];
        S.k = L.F.t + s * LPF;
        S.act = L.F.famok && S.k < L.F.M;
        S.b = L.F.beg + S.k;
        const size_t b = S.act ? (size_t)S.b : 0;  // the bond arrays hold at least one entry
        L.jr[s] = a.col[b];
        L.br[s] = a.broken[b];
        L.rv[s] = a.rev[b];
        if constexpr (FIX) {
            // the FP64 fix-up assembles K itself (the stored K^-1 is FP32)
            L.kraw[s][0] = L.kraw[s][1] = L.kraw[s][2] = R(0);
        } else {
            L.kraw[s][0] = kv[3 * b];
            L.kraw[s][1] = kv[3 * b + 1];
            L.kraw[s][2] = kv[3 * b + 2];
        }
        if constexpr (DMG) {
            L.kap[s] = a.kappa[b];
            L.dm[s] = a.dmg[b];
        }
    }
}

template <typename R, int LPF, int BPL>
__device__ __forceinline__ void stage_bonds(const BondArgs& a, FamLoad<R, BPL>& L)
{
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        const int j = L.S[s].act ? L.jr[s] : (L.F.famok ? L.F.fam : 0);
        L.Xj[s] = a.X[j];
        L.uj[s] = a.u[j];
        L.Vj[s] = a.V[j];
    }
}

// First use of the loaded values: bond vector, displacement difference, volume.
template <typename R, int BPL>
__device__ __forceinline__ void finish_loads(FamLoad<R, BPL>& L)
{
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        BondSlot& S = L.S[s];
        S.x0 = L.Xj[s].x - L.F.Xi.x;
        S.x1 = L.Xj[s].y - L.F.Xi.y;
        S.d0 = L.uj[s].x - L.F.ui.x;
        S.d1 = L.uj[s].y - L.F.ui.y;
        S.Vj = L.Vj[s];
    }
}

struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

typedef float v4f __attribute__((ext_vector_type(4)));

// The two contractions are dense per-family products, D = W P and S = W Q with
// the symmetric weight matrix W (M x M) and P, Q (M x 4, one 2x2 tensor per
// bond).  MF = 0 forms them with VALU FMAs; MF = 1 hands them to the matrix
// cores while the VALU evaluates the weights:
//  * FP32, v_mfma_f32_4x4x1 (16 blocks of 4 lanes, C_b[r][t] += A[4b+r] B[4b+t]):
//    A = component (lane & 3) of record l, B = the lane's own weight w(k, l), so
//    lane k accumulates all four components of its own D_k / S_k.  The tile
//    layout is unchanged.
//  * FP64, v_mfma_f64_4x4x4 (4 blocks of 4x4x4; A[p][i][k] in lane 16k+4p+i,
//    B[p][k][j] in lane 16k+4p+j, C[p][i][j] in lane 16i+4p+j): one step takes
//    16 bonds kb = 16s + (lane & 15) against 4 neighbours l = 4q + (lane >> 4),
//    with A = component (lane & 3) of record l and B = w(kb, l); lane holds
//    component (lane >> 4) of D_kb.  Families are taken one after the other and
//    the results go back to the bond-per-lane layout through LDS.
template <typename R, int MF> struct Contraction {
    static constexpr bool MF32 = (MF == 1 || MF == 4) && sizeof(R) == 4;
    static constexpr bool MF64 = MF == 1 && sizeof(R) == 8;
    static constexpr bool MW = MF == 2;
    static constexpr bool RF = MF == 3 || MF == 4;   // rank-form weights (below)
};

// ---------------------------------------------------------------------------
// Matrix-core weights (MF = 2, "MW"): the weight of a pair comes from a rank-3
// product on the matrix cores instead of ten VALU operations.
//
// With p = exp(A l / n2), q = 1 / p and h = 1/2 + 1/2 e_k.e_l,
//   w(k,l) = exp(-A |l_k - l_l|) h^n2 = (g / 2)^n2,
//   g = U_k . L_l (l_l >= l_k)  or  L_k . U_l (l_l <= l_k),
// where U = p (1, e_x, e_y) and L = q (1, e_x, e_y): the radial factor folds
// into the base of the power, and which of the two forms applies follows from
// the order of the lengths.  Each family is stored in ascending order of its
// bonds' lengths, so its records are in rank order, and in
// a 16 x 16 block of (bond, neighbour) ranks every pair has the same form,
// except on the diagonal blocks, which take the smaller of the two (it is the
// one that applies).  Per 16 x 16 block:
//   G^T = L-or-U(neighbours) x U-or-L(bonds)^T   one v_mfma_{f32,f64}_16x16x4
//   w   = G^10 (the 1/2^n2 is folded into the P / Q records)    4 VALU per pair
//   D  += w P                      FP32: four v_mfma_f32_4x4x1 (16 blocks)
//                                  FP64: four v_mfma_f64_4x4x4 (4 blocks)
// Output layouts (checked on the device): 16x16x4 f32, lane l register r holds
// C[4 (l / 16) + r][l % 16]; 16x16x4 f64, C[l / 16 + 4 r][l % 16].  A family's
// blocks are taken one after another by the whole wave; the results go back
// to the records' P slots, from where each lane reads its bonds' sums.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename R, int LPF, int BPL> struct MWL {
    static constexpr int FS = LPF * BPL;
    static constexpr int FPW = 64 / LPF;
    static constexpr int NB = FS / 16;
    // Row strides chosen against LDS bank conflicts (64 banks of 4 bytes, a
    // wave served in 32-lane halves).  U / L rows (16 consecutive slots per
    // lane group, rows hi = 0..3 a half-wave apart): an odd multiple of 16
    // slots, so the two rows of a half land in distinct banks.  P rows (read
    // with the component as the row index): in FP64 a stride of FSP + 8
    // doubles puts the four rows a quarter of the banks apart.  Families are
    // skewed by 8 doubles so the record writes of a wave's families spread.
    static constexpr int FSP = FS % 32 == 0 ? FS + 16 : FS;
    static constexpr int PSP = sizeof(R) == 8 ? FSP + 8 : FSP;
    // per family 10 rows: U (3), L (3), P / Q (4); the fourth component of
    // the rank-3 Gram products is a zero the lanes supply themselves
    static constexpr int ROWS = 10, PROW = 6;
    static constexpr int FAM = 6 * FSP + 4 * PSP + (sizeof(R) == 8 ? 8 : 0);
    static constexpr int DOUBLES = (int)((FPW * FAM * sizeof(R) + 7) / 8);
    static __device__ __forceinline__ constexpr int row(int c) { return c < PROW ? c * FSP : PROW * FSP + (c - PROW) * PSP; }
    static __device__ __forceinline__ int at(int f, int c, int slot) { return f * FAM + row(c) + slot; }
};

template <typename R, int N2_IS_10>
__device__ __forceinline__ void mw_geometry(const BondArgs& a, const BondSlot& S, R (&U)[3], R (&Lq)[3])
{
    if constexpr (sizeof(R) == 8) {
        const double l2 = S.x0 * S.x0 + S.x1 * S.x1;
        const double il = S.act ? rsqrt64(l2) : 0.0;
        const double len = l2 * il;
        const double ex = S.x0 * il, ey = S.x1 * il;
        const double Ar = N2_IS_10 ? a.A * 0.1 : a.A / a.n2;
        const double p = S.act ? exp(Ar * len) : 0.0;
        const double q = S.act ? rcp64(p) : 0.0;
        U[0] = p; U[1] = p * ex; U[2] = p * ey;
        Lq[0] = q; Lq[1] = q * ex; Lq[2] = q * ey;
    } else {
        const float x0 = float(S.x0), x1 = float(S.x1);
        const float l2 = x0 * x0 + x1 * x1;
        const float il = S.act ? __builtin_amdgcn_rsqf(l2) : 0.0f;
        const float len = l2 * il;
        const float ex = x0 * il, ey = x1 * il;
        const float Ar = N2_IS_10 ? float(a.A * 0.1) : float(a.A / a.n2);
        const float p = S.act ? __expf(Ar * len) : 0.0f;
        const float q = S.act ? __builtin_amdgcn_rcpf(p) : 0.0f;
        U[0] = p; U[1] = p * ex; U[2] = p * ey;
        Lq[0] = q; Lq[1] = q * ex; Lq[2] = q * ey;
    }
}

// G^T block on the matrix cores: lane l gives A = component l/16 of its
// neighbour record, B = component l/16 of its bond record.
__device__ __forceinline__ v4f mw_gram(float a, float b)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
}
__device__ __forceinline__ v4d mw_gram(double a, double b)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, v4d{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
}

template <int N2_IS_10, typename R>
__device__ __forceinline__ R mw_pow(R g, R n2)
{
    if constexpr (N2_IS_10) {
        const R g2 = g * g;
        const R g4 = g2 * g2;
        const R g8 = g4 * g4;
        return g8 * g2;
    } else {
        return rpow(rmax(g, R(0)), n2);
    }
}

// One contraction X_k = sum_l w(k,l) Y_l for every family of the wave, Y in
// the P rows of the MW records; X replaces Y there.  Per family the neighbour
// blocks s2 are the outer loop: the U and L rows of every block are read
// once, the 16 x 16 Gram blocks of all bond blocks s1 against s2 are
// independent matrix-core products, and each s1 keeps its own accumulator, so
// the MFMA chains of different s1 interleave.  Every block of the tile is
// evaluated (slots past a family's end hold zero records and add nothing), so
// the loop is straight-line code the scheduler can interleave.
template <typename R, int LPF, int BPL, int N2_IS_10, typename HA, typename HB>
__device__ __forceinline__ void mw_contract(R* lds, int lane, R n2, const HA& hook_a, const HB& hook_b)
{
    typedef MWL<R, LPF, BPL> M;
    constexpr int FSP = M::FSP;
    constexpr int PSP = M::PSP;
    constexpr int NB = M::NB;
    constexpr int PR = M::PROW;
    // families taken together (their chains interleave)
    constexpr int FG = M::FPW >= 2 ? 2 : 1;
    const int lo = lane & 15, hi = lane >> 4, lc = lane & 3;
    const int hc = hi < 3 ? hi : 2;   // lanes of component 3 read a valid row, then supply 0
    typedef typename std::conditional<sizeof(R) == 4, v4f, double>::type Acc;
    typedef decltype(mw_gram(R(0), R(0))) G4;
#pragma unroll
    for (int f0 = 0; f0 < M::FPW; f0 += FG) {
        R* base[FG];
        R U[FG][NB], Lq[FG][NB];
        Acc acc[FG][NB];
#pragma unroll
        for (int e = 0; e < FG; ++e) {
            base[e] = lds + M::at(f0 + e, 0, 0);
            // component hi of the U and L records of slot 16 s + lo
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const R u = base[e][hc * FSP + 16 * b + lo];
                const R l = base[e][(3 + hc) * FSP + 16 * b + lo];
                U[e][b] = hi < 3 ? u : R(0);
                Lq[e][b] = hi < 3 ? l : R(0);
                if constexpr (sizeof(R) == 4) acc[e][b] = v4f{0.f, 0.f, 0.f, 0.f};
                else acc[e][b] = 0.0;
            }
        }
#pragma unroll
        for (int s2 = 0; s2 < NB; ++s2) {
            R w[FG][NB][4];
#pragma unroll
            for (int e = 0; e < FG; ++e)
#pragma unroll
                for (int s1 = 0; s1 < NB; ++s1) {
                    G4 g;
                    if (s2 > s1) {
                        g = mw_gram(Lq[e][s2], U[e][s1]);
                    } else if (s2 < s1) {
                        g = mw_gram(U[e][s2], Lq[e][s1]);
                    } else {
                        const G4 gu = mw_gram(Lq[e][s2], U[e][s1]), gl = mw_gram(U[e][s2], Lq[e][s1]);
                        // fmin, not the inline-asm rmin: its operands come
                        // straight from the matrix cores, and only compiled
                        // code gets the wait states an MFMA result needs
                        // before a VALU reads it
#pragma unroll
                        for (int r = 0; r < 4; ++r) g[r] = fmin(R(gu[r]), R(gl[r]));
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) w[e][s1][r] = mw_pow<N2_IS_10>(R(g[r]), n2);
                }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int e = 0; e < FG; ++e) {
                    R pa;
                    if constexpr (sizeof(R) == 4) pa = base[e][PR * FSP + lc * PSP + 16 * s2 + 4 * hi + r];
                    else pa = base[e][PR * FSP + lc * PSP + 16 * s2 + 4 * r + hi];
#pragma unroll
                    for (int s1 = 0; s1 < NB; ++s1) {
                        if constexpr (sizeof(R) == 4)
                            acc[e][s1] = __builtin_amdgcn_mfma_f32_4x4x1f32(pa, w[e][s1][r], acc[e][s1], 0, 0, 0);
                        else
                            acc[e][s1] = __builtin_amdgcn_mfma_f64_4x4x4f64(pa, w[e][s1][r], acc[e][s1], 0, 0, 0);
                    }
                }
            }
        }
        if constexpr (sizeof(R) == 4) {
            // partial sums of lanes l, l + 16, l + 32, l + 48 (the four
            // neighbour groups of each 4x4x1 block row)
#pragma unroll
            for (int e = 0; e < FG; ++e)
#pragma unroll
                for (int s1 = 0; s1 < NB; ++s1)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        acc[e][s1][c] += __shfl_xor(acc[e][s1][c], 16, 64);
                        acc[e][s1][c] += __shfl_xor(acc[e][s1][c], 32, 64);
                    }
        }
        wave_sync();
#pragma unroll
        for (int e = 0; e < FG; ++e)
#pragma unroll
            for (int s1 = 0; s1 < NB; ++s1) {
                if constexpr (sizeof(R) == 4) {
                    // lane l < 16 holds the four components of slot 16 s1 + l
                    if (lane < 16) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) base[e][PR * FSP + c * PSP + 16 * s1 + lane] = acc[e][s1][c];
                    }
                } else {
                    // lane l holds component l / 16 of slot 16 s1 + l % 16
                    base[e][PR * FSP + hi * PSP + 16 * s1 + lo] = acc[e][s1];
                }
            }
        if (f0 == 0) hook_a();
    }
    hook_b();
}

// Rank-form weights (MF = 3: VALU contractions, MF = 4: FP32 matrix-core
// contractions; tiles of 16 lanes per family).  Each family is stored in
// ascending order of its bonds' reference lengths (ties in the caller's
// order), and lane t owns the bonds t, t + 16, t + 32: slot s holds bond
// (rank) 16 s + t.  With p = exp(A l / n2), q = 1 / p, h = 1/2 + 1/2 e_k.e_l,
// U = p (1, e), L = q (1, e),
//   w(k,l) = exp(-A |l_k - l_l|) h^n2 = (g / 2)^n2,
//   g = U_k . L_l (l_l >= l_k)  or  L_k . U_l (l_l <= l_k):
// the radial factor folds into the base of the power, and which of the two
// forms applies follows from the order of the lengths.  Record l lies above
// all of slot s's bonds when l / 16 > s and below them when l / 16 < s, so
//   w = (q_l (U_k0 + U_k1 e_xl + U_k2 e_yl))^n2 / 2^n2    (l / 16 > s)
//   w = (p_l (L_k0 + L_k1 e_xl + L_k2 e_yl))^n2 / 2^n2    (l / 16 < s)
// three VALU for the base and four for the tenth power, against ten for
// exp(-A |l_k - l_l|) h^10 with its min.  In the diagonal block (l / 16 = s)
// the form follows from the pair: L_k = r_k U_k with r_k = q_k / p_k = q_k^2,
// so the L form is the U form's base times r_k p_l instead of q_l, and the
// smaller of the two applies: g = (U_k . (1, e_l)) min(q_l, r_k p_l), five
// VALU for the base.  p, q here are exp(+-A l / n2), and 1/2^n2 is folded into
// the P and Q records.  k holds U_k (0..2), L_k (3..5) and r_k (6).
//
// (Taking the bond's own radial factor out of the block sums, two VALU for the
// base, was measured 2-3 times less accurate in FP32 where the diagonal blocks
// then form their base in the scaled units; the form above is kept.)
template <typename R, int FORM>
__device__ __forceinline__ R base_rf(const R* k, R exl, R eyl, R pl, R ql)
{
    if constexpr (FORM > 0)
        return (k[0] + k[1] * exl + k[2] * eyl) * ql;
    else if constexpr (FORM < 0)
        return (k[3] + k[4] * exl + k[5] * eyl) * pl;
    else
        return (k[0] + k[1] * exl + k[2] * eyl) * rmin(ql, k[6] * pl);
}

template <typename R, int N2_IS_10, int FORM>
__device__ __forceinline__ R weight_rf(const R* k, R exl, R eyl, R pl, R ql, R n2)
{
    return mw_pow<N2_IS_10>(base_rf<R, FORM>(k, exl, eyl, pl, ql), n2);
}

// The k-side factors of the rank form from a bond's geometry.
template <typename R>
__device__ __forceinline__ void rf_kside(R ex, R ey, R pe, R qe, R* k)
{
    k[0] = pe;
    k[1] = pe * ex;
    k[2] = pe * ey;
    k[3] = qe;
    k[4] = qe * ex;
    k[5] = qe * ey;
    k[6] = qe * qe;
}

// k-side operands of the FP64 matrix-core contractions for one family: the
// half unit vector and the radial factors of the bonds 16 s + lk (s < NS).
template <int NS> struct MfK {
    double v[NS][4];

    template <typename R, int FPW>
    __device__ __forceinline__ void load(const R* rec, int f, int lk)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int kb = 16 * s + lk;
            v[s][0] = 0.5 * rec[rec_at<R, FPW>(kb, f, 0)];
            v[s][1] = 0.5 * rec[rec_at<R, FPW>(kb, f, 1)];
            v[s][2] = rec[rec_at<R, FPW>(kb, f, 2)];
            v[s][3] = rec[rec_at<R, FPW>(kb, f, 3)];
        }
    }

    template <int N2_IS_10>
    __device__ __forceinline__ double w(int s, double ex, double ey, double pe, double qe, double n2) const
    {
        return weight<double, N2_IS_10>(v[s][0], v[s][1], v[s][2], v[s][3], ex, ey, pe, qe, n2);
    }
};

// Neighbour-side operands of one step of the FP64 matrix-core contractions:
// the geometry of record row lb (read when the weights are evaluated) and
// component lc of its P / Q values.
struct MfL {
    double ex, ey, pe, qe, y;

    template <typename R, int FPW, bool GEO = true>
    __device__ __forceinline__ void load(const R* rec, int f, int lb, int lc)
    {
        if constexpr (GEO) {
            ex = rec[rec_at<R, FPW>(lb, f, 0)];
            ey = rec[rec_at<R, FPW>(lb, f, 1)];
            pe = rec[rec_at<R, FPW>(lb, f, 2)];
            qe = rec[rec_at<R, FPW>(lb, f, 3)];
        } else {
            ex = ey = pe = qe = 0.0;
        }
        y = rec[rec_at<R, FPW>(lb, f, 4 + lc)];
    }
};

// LDS per wave of a bond-kernel variant, in doubles.
template <typename R, int LPF, int BPL, int MF> struct WaveLds {
    static constexpr int DOUBLES = MF == 2 && MWL<R, LPF, BPL>::DOUBLES > Tile<LPF, BPL>::LDS_DOUBLES
                                       ? MWL<R, LPF, BPL>::DOUBLES
                                       : Tile<LPF, BPL>::LDS_DOUBLES;
};

// One explicit force evaluation of a family group: bond breaking, K
// re-assembly for families that lost a bond, then the two family contractions
// and the force-state outputs.  hook_a runs halfway through contraction 1,
// hook_b after it (a persistent wave issues its next group's loads there).
// FIX = 1: the FP64 re-evaluation of a guarded family of a mixed-precision
// context (k_bond_fix): K is assembled here (force_k), nothing is stored but the
// T terms (as FP32, the context's W type), f_own and, with DMG, the family's
// damage update (which the mixed kernel leaves to it).
template <typename R, int LPF, int BPL, int N2_IS_10, int MF, int DMG, typename HA, typename HB,
          int FIX = 0>
__device__ __forceinline__ void bond_body(const BondArgs& a, FamLoad<R, BPL>& L, double* rec64,
                                          int lane, const HA& hook_a, const HB& hook_b)
{
    static_assert(!FIX || sizeof(R) == 8, "the fix-up is FP64");
    finish_loads(L);
    constexpr bool MF32 = Contraction<R, MF>::MF32;
    constexpr bool MF64 = Contraction<R, MF>::MF64;
    constexpr bool MW = Contraction<R, MF>::MW;
    constexpr bool RF = Contraction<R, MF>::RF;
    static_assert(!(MF64 || MW) || (Tile<LPF, BPL>::FS % 16 == 0),
                  "matrix-core contractions: families of a multiple of 16 slots");
    static_assert(!RF || LPF == 16, "rank-form weights: 16 lanes per family");
    typedef typename Vec2<R>::T R2;
    constexpr int FS = Tile<LPF, BPL>::FS;
    constexpr int FPW = Tile<LPF, BPL>::FPW;
    const LaneFam& F = L.F;
    BondSlot (&S)[BPL] = L.S;
    bool alive[BPL];
    int nnew = 0;
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        // ---- bond breaking (FP64, irreversible, both directions agree) ----
        alive[s] = S[s].act && L.br[s] == 0;
        bool newly = false;
        if (a.s0 > 0.0 && alive[s] && bond_breaks(S[s].x0, S[s].x1, S[s].d0, S[s].d1, a.brk_c)) {
            alive[s] = false;
            newly = true;
            a.broken[S[s].b] = 1;
        }
        nnew += __popcll(__ballot(newly));
    }
    if (!FIX && lane == 0 && nnew) atomicAdd(a.counters, (unsigned int)nnew);

    // ---- K_ik^-1: stored, or re-assembled in FP64 when the family changed;
    // used in the bond arithmetic's precision ----
    R kin[BPL][3];
    bool guard = false;   // mixed precision: the family is evaluated again in FP64
    if constexpr (FamLoad<R, BPL>::GUARD) guard = L.ff != 0;
    if (nnew != 0 || a.force_k) {
        double kinv[BPL][3];
        Geo<double> g64[BPL];
#pragma unroll
        for (int s = 0; s < BPL; ++s) g64[s] = geometry64(a, S[s], a.A);
        bool sing[BPL];
        assemble_kinv<LPF, BPL, N2_IS_10>(rec64, F, S, alive, g64, a.n2, a.ktol, kinv, sing);
        if constexpr (!FIX) {
            int nsing = 0;
            bool g = false;
#pragma unroll
            for (int s = 0; s < BPL; ++s) {
                if (S[s].act) {
                    store_kinv<R>(a, S[s].b, kinv[s]);
                    g = g || kinv_guarded(kinv[s], a.illr);
                }
                nsing += __popcll(__ballot(sing[s]));
            }
            if (lane == 0 && nsing) atomicAdd(a.counters + 1, (unsigned int)nsing);
            if constexpr (FamLoad<R, BPL>::GUARD) {
                guard = family_any<LPF>(g, F.slot);
                if (F.famok && F.t == 0 && guard != (L.ff != 0)) a.famflag[F.fam] = guard ? 1 : 0;
            }
        }
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) kin[s][c] = R(kinv[s][c]);
    } else {
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) kin[s][c] = S[s].act ? L.kraw[s][c] : R(0);
    }

    if constexpr (FamLoad<R, BPL>::GUARD) {
        if (a.famlist && guard && F.famok && F.t == 0) {
            const unsigned int pos = atomicAdd(a.famcnt, 1u);
            if (pos < (unsigned int)a.n) a.famlist[pos] = F.fam;
        }
    }

    // ---- records for the first contraction ----
    // The VALU FP64 one-bond-per-lane contraction sends V l du and reuses the
    // record's unit vector (P = (V l du) (x) e, 48 bytes per record instead of
    // 64, two more VALU per weight); the others send the four products, which
    // is faster where the VALU is the bound (FP64 two bonds per lane: -3 %).
    constexpr bool COMPACT = sizeof(R) == 8 && !MF && BPL == 1;
    R* rec = reinterpret_cast<R*>(__builtin_assume_aligned(rec64, 16));
    Geo<R> g[BPL];
    // the 2^-n2 of (g/2)^n2 in the matrix-core and rank-form records
    const R mw_scale = N2_IS_10 ? R(1.0 / 1024.0) : R(exp2(-a.n2));
    if constexpr (MW) {
        typedef MWL<R, LPF, BPL> M;
#pragma unroll
        for (int s = 0; s < BPL; ++s) {
            R U[3], Lq[3];
            mw_geometry<R, N2_IS_10>(a, S[s], U, Lq);
            R* r0 = rec + M::at(F.slot, 0, S[s].k);
            r0[0] = U[0];
            r0[M::row(1)] = U[1];
            r0[M::row(2)] = U[2];
            r0[M::row(3)] = Lq[0];
            r0[M::row(4)] = Lq[1];
            r0[M::row(5)] = Lq[2];
            const R x0r = R(S[s].x0), x1r = R(S[s].x1);
            const R Vd0 = alive[s] ? R(S[s].Vj * S[s].d0) * mw_scale : R(0);
            const R Vd1 = alive[s] ? R(S[s].Vj * S[s].d1) * mw_scale : R(0);
            r0[M::row(6)] = Vd0 * x0r;
            r0[M::row(7)] = Vd0 * x1r;
            r0[M::row(8)] = Vd1 * x0r;
            r0[M::row(9)] = Vd1 * x1r;
        }
    }
    // rank form: the lane's own bonds are its slots (families in length
    // order); their k-side factors U = p (1, e), L = q (1, e), r = q / p
    R Kk[RF ? BPL : 1][7];
    if constexpr (!MW) {
    // rank-form records carry exp(+-A l / n2) and the P scaled by 2^-n2
    const double Ag = RF ? (N2_IS_10 ? a.A * 0.1 : a.A / a.n2) : a.A;
    if constexpr (sizeof(R) == 8) {
        // every slot's FP64 geometry first: independent chains side by side
#pragma unroll
        for (int s = 0; s < BPL; ++s) g[s] = GeoOf<R>::get(a, S[s], Ag);
    }
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        if constexpr (sizeof(R) == 4) g[s] = GeoOf<R>::get(a, S[s], Ag);
        const int k = S[s].k;
        rec_put4<FPW>(rec, k, F.slot, 0, g[s].ex, g[s].ey, g[s].pe, g[s].qe);
        if constexpr (RF) rf_kside(g[s].ex, g[s].ey, g[s].pe, g[s].qe, Kk[s]);
        if constexpr (COMPACT) {
            const double Vl = alive[s] ? S[s].Vj * double(g[s].len) : 0.0;
            rec_put<FPW>(rec, k, F.slot, 2, R(Vl * S[s].d0), R(Vl * S[s].d1));
        } else {
            const R x0r = R(S[s].x0), x1r = R(S[s].x1);
            const R sc = RF ? mw_scale : R(1);
            const R Vd0 = alive[s] ? R(S[s].Vj * S[s].d0) * sc : R(0);
            const R Vd1 = alive[s] ? R(S[s].Vj * S[s].d1) * sc : R(0);
            rec_put4<FPW>(rec, k, F.slot, 1, Vd0 * x0r, Vd0 * x1r, Vd1 * x0r, Vd1 * x1r);
        }
    }
    }
    // records past a family's end are zero (inactive slots), so the loops may
    // run to the wave's largest family rounded up to 4 without guards
    const int mmax = wave_max_family<LPF>(F.M);
    wave_sync();
    // The FP64 rank-form weights (n2 = 10) of slots s0 .. BPL - 1 against
    // record l, each step of the tenth power taken for every slot before the
    // next: the slots' chains run interleaved rather than one after the other
    // (measured -1.3 % at three bonds per lane)
    auto wt_all = [&](int l, R ex, R ey, R pe, R qe, R (&om)[BPL], int s0) {
        R gb[BPL], g2[BPL], g4[BPL];
#pragma unroll
        for (int s = s0; s < BPL; ++s) {
            const int bl = l >> 4;
            if (bl > s) gb[s] = base_rf<R, 1>(Kk[s], ex, ey, pe, qe);
            else if (bl < s) gb[s] = base_rf<R, -1>(Kk[s], ex, ey, pe, qe);
            else gb[s] = base_rf<R, 0>(Kk[s], ex, ey, pe, qe);
        }
#pragma unroll
        for (int s = s0; s < BPL; ++s) g2[s] = gb[s] * gb[s];
#pragma unroll
        for (int s = s0; s < BPL; ++s) g4[s] = g2[s] * g2[s];
#pragma unroll
        for (int s = s0; s < BPL; ++s) g4[s] = g4[s] * g4[s];
#pragma unroll
        for (int s = s0; s < BPL; ++s) om[s] = g4[s] * g2[s];
    };
    // the weight of slot s's bond against record l
    auto wt = [&](int s, int l, R ex, R ey, R pe, R qe) -> R {
        if constexpr (RF) {
            const int bl = l >> 4;
            if (bl > s) return weight_rf<R, N2_IS_10, 1>(Kk[s], ex, ey, pe, qe, R(a.n2));
            if (bl < s) return weight_rf<R, N2_IS_10, -1>(Kk[s], ex, ey, pe, qe, R(a.n2));
            return weight_rf<R, N2_IS_10, 0>(Kk[s], ex, ey, pe, qe, R(a.n2));
        } else {
            return weight<R, N2_IS_10>(g[s].hx, g[s].hy, g[s].pe, g[s].qe, ex, ey, pe, qe, R(a.n2));
        }
    };

    // ---- contraction 1: D_ik = sum_l w(k,l) V_l du_l (x) xi_l ----
    const R n2r = R(a.n2);
    // FP64 matrix-core layout: NS sets of 16 bonds, NQ steps of 4 neighbours
    constexpr int NS = MF64 ? FS / 16 : 1;
    constexpr int NQ = MF64 ? FS / 4 : 1;
    constexpr int WF = MF64 ? FPW : 1;
    const int lk = lane & 15, lq = lane >> 4, lc = lane & 3;
    // The FP64 matrix-core tile keeps its weights for contraction 2 (WF x NQ x
    // NS per lane) when they are at most 32 doubles (the (32,1) tile); larger
    // caches evaluate them again there (measured: 64-slot families 2.47 ->
    // 1.52 ms at 28-bond families, 48.6 -> 45.1 ms at 48-bond families), while
    // the (32,1) tile loses without its cache (0.92 -> 1.16 ms) although its
    // occupancy rises from 4 to 7 waves per SIMD.
    constexpr bool MFRW = MF64 && WF * NQ * NS > 32;
    double wc[MFRW ? 1 : WF][NQ][NS];   // FP64 matrix-core weight cache
    int Mf[WF];              // family sizes (wave-uniform)
    // Families of 64 slots evaluate the weights again in contraction 2 rather
    // than keep a row of 64 weights live per lane: the occupancy gained
    // outweighs the extra VALU work there (measured: FP64 -24 %, mixed -4 %),
    // not for one-bond-per-lane 32-slot families (+7 % and more).
    // NC of the lane's BPL weight rows are kept for contraction 2, the others
    // are evaluated again there (FP64 two-bonds-per-lane: one row of 32 doubles
    // fewer, measured -2.4 %; three bonds per lane, rows of 48: one row in
    // FP32, none in FP64)
    constexpr int NC = (MF64 || MW || (!MF && FS >= 64)) ? 0
                       : BPL == 3                    ? (sizeof(R) == 8 ? 1 : 3)
                       : (sizeof(R) == 8 && BPL == 2) ? 1
                                                      : BPL;
    constexpr bool RW = NC < BPL;
    // neighbours per scheduling group of the contraction loops (each group's
    // record loads are issued together; 2 frees no registers in the FP64
    // two-bonds-per-lane tile, whose live set is its two weight rows; the rank
    // form with three bonds per lane runs 8, measured -0.7 % (FP64) and -1.4 %
    // (mixed) against 4 on families of 48, where two bonds per lane lose 6 %
    // with 8 on families of up to 28)
    constexpr int GL = (RF && BPL == 3) ? 8 : 4;
    static_assert(FS % GL == 0 && (FS / 2 - GL) % GL == 0, "a load group ends halfway (hook_a)");
    R w[NC > 0 ? NC : 1][NC > 0 ? FS : 1];  // per-lane weight rows
    // mixed-precision rank form with three bonds per lane: symmetric blocks
    // (below); TRO is the transpose tile's offset in the wave's LDS, in floats,
    // after the FS rows of records
    constexpr bool TRX = MF32 && RF && BPL == 3 && NC == BPL;
    // FP64 rank form, n2 = 10: the slots' weights interleaved (wt_all)
    constexpr bool ILW = RF && sizeof(R) == 8 && N2_IS_10;
    constexpr int TRO = FS * (2 * FPW + 1) * 4;
    static_assert(!TRX || TRO + FPW * 16 * 17 <= 2 * WaveLds<R, LPF, BPL, MF>::DOUBLES,
                  "transpose tile within the wave's LDS");
    R D[BPL][4];
    v4f Dv[BPL];
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        D[s][0] = D[s][1] = D[s][2] = D[s][3] = R(0);
        Dv[s] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    if constexpr (MW) {
        mw_contract<R, LPF, BPL, N2_IS_10>(rec, lane, n2r, hook_a, hook_b);
        wave_sync();
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) D[s][c] = rec[MWL<R, LPF, BPL>::at(F.slot, MWL<R, LPF, BPL>::PROW + c, S[s].k)];
    } else if constexpr (MF64) {
        double Dacc[WF][NS];
#pragma unroll
        for (int f = 0; f < WF; ++f) {
            Mf[f] = __builtin_amdgcn_readfirstlane(__shfl(F.M, f * LPF, 64));
            MfK<NS> ks;
            ks.template load<R, FPW>(rec, f, lk);
#pragma unroll
            for (int s = 0; s < NS; ++s) Dacc[f][s] = 0.0;
            // neighbour rows 4 q + lq, loaded one step ahead of their use
            MfL nb[2];
            nb[0].template load<R, FPW>(rec, f, lq, lc);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q + 1 < NQ) nb[(q + 1) & 1].template load<R, FPW>(rec, f, 4 * (q + 1) + lq, lc);
                if (4 * q < Mf[f]) {
                    const MfL& r = nb[q & 1];
                    // every bond set, also those past the family's end (their
                    // records are zero, so are their weights): no branch
                    // between the independent weight chains
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        const double om = ks.template w<N2_IS_10>(s, r.ex, r.ey, r.pe, r.qe, a.n2);
                        if constexpr (!MFRW) wc[f][q][s] = om;
                        Dacc[f][s] = __builtin_amdgcn_mfma_f64_4x4x4f64(r.y, om, Dacc[f][s], 0, 0, 0);
                    }
                } else if constexpr (!MFRW) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) wc[f][q][s] = 0.0;
                }
                __builtin_amdgcn_sched_barrier(0);
                if (f == WF - 1 && q == NQ / 2 - 1) hook_a();
            }
        }
        hook_b();
        // back to one bond per lane: lane holds component lq of bond 16s + lk
        wave_sync();
#pragma unroll
        for (int f = 0; f < WF; ++f)
#pragma unroll
            for (int s = 0; s < NS; ++s) rec[rec_at<R, FPW>(16 * s + lk, f, 4 + lq)] = Dacc[f][s];
        wave_sync();
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) D[s][c] = rec[rec_at<R, FPW>(S[s].k, F.slot, 4 + c)];
    } else if constexpr (TRX) {
        // The weights are symmetric, w(k, l) = w(l, k): of a family's 48 x 48
        // weight matrix, block (s, b) (rows 16 s .. 16 s + 15, columns 16 b ..)
        // is evaluated for b >= s only and kept for contraction 2; block (b, s)
        // is its transpose, formed through LDS in each contraction (trx_block):
        // a third fewer weight evaluations and a third fewer kept weights.
        // Columns of blocks 1 and 2 first, then the diagonal block of slot 0,
        // then the transposed blocks.
        const int tb = TRO + F.slot * (16 * 17);   // the family's transpose tile
#pragma unroll
        for (int l0 = 16; l0 < FS; l0 += GL) {
            if (l0 < mmax) {
#pragma unroll
                for (int l = l0; l < l0 + GL; ++l) {
                    const R ex = rec[rec_at<R, FPW>(l, F.slot, 0)];
                    const R ey = rec[rec_at<R, FPW>(l, F.slot, 1)];
                    const R pe = rec[rec_at<R, FPW>(l, F.slot, 2)];
                    const R qe = rec[rec_at<R, FPW>(l, F.slot, 3)];
                    const R Pa = rec[rec_at<R, FPW>(l, F.slot, 4 + lc)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        if (s > (l >> 4)) continue;
                        const R om = wt(s, l, ex, ey, pe, qe);
                        w[s][l] = om;
                        Dv[s] = __builtin_amdgcn_mfma_f32_4x4x1f32(Pa, om, Dv[s], 0, 0, 0);
                    }
                }
            } else {
#pragma unroll
                for (int l = l0; l < l0 + GL; ++l)
#pragma unroll
                    for (int s = 0; s < BPL; ++s)
                        if (s <= (l >> 4)) w[s][l] = R(0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        hook_a();
#pragma unroll
        for (int l0 = 0; l0 < 16; l0 += GL) {
            if (l0 < mmax) {
#pragma unroll
                for (int l = l0; l < l0 + GL; ++l) {
                    const R ex = rec[rec_at<R, FPW>(l, F.slot, 0)];
                    const R ey = rec[rec_at<R, FPW>(l, F.slot, 1)];
                    const R pe = rec[rec_at<R, FPW>(l, F.slot, 2)];
                    const R qe = rec[rec_at<R, FPW>(l, F.slot, 3)];
                    const R Pa = rec[rec_at<R, FPW>(l, F.slot, 4 + lc)];
                    const R om = wt(0, l, ex, ey, pe, qe);
                    w[0][l] = om;
                    Dv[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(Pa, om, Dv[0], 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int l = l0; l < l0 + GL; ++l) w[0][l] = R(0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        trx_block<FPW, 0, 1>(w, rec, tb, F.t, F.slot, lc, mmax, Dv[1]);
        trx_block<FPW, 0, 2>(w, rec, tb, F.t, F.slot, lc, mmax, Dv[2]);
        trx_block<FPW, 1, 2>(w, rec, tb, F.t, F.slot, lc, mmax, Dv[2]);
        hook_b();
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) D[s][c] = Dv[s][c];
    } else {
#pragma unroll
    for (int l0 = 0; l0 < FS; l0 += GL) {
        if (l0 < mmax) {
#pragma unroll
            for (int l = l0; l < l0 + GL; ++l) {
                const R ex = rec[rec_at<R, FPW>(l, F.slot, 0)];
                const R ey = rec[rec_at<R, FPW>(l, F.slot, 1)];
                const R pe = rec[rec_at<R, FPW>(l, F.slot, 2)];
                const R qe = rec[rec_at<R, FPW>(l, F.slot, 3)];
                if constexpr (MF32) {
                    const R Pa = rec[rec_at<R, FPW>(l, F.slot, 4 + lc)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        const R om = wt(s, l, ex, ey, pe, qe);
                        if (s < NC) w[s][l] = om;
                        Dv[s] = __builtin_amdgcn_mfma_f32_4x4x1f32(Pa, om, Dv[s], 0, 0, 0);
                    }
                } else if constexpr (COMPACT) {
                    const R a0 = rec[rec_at<R, FPW>(l, F.slot, 4)];
                    const R a1 = rec[rec_at<R, FPW>(l, F.slot, 5)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        const R om = wt(s, l, ex, ey, pe, qe);
                        if (s < NC) w[s][l] = om;
                        const R t0 = om * a0, t1 = om * a1;
                        D[s][0] += t0 * ex;
                        D[s][1] += t0 * ey;
                        D[s][2] += t1 * ex;
                        D[s][3] += t1 * ey;
                    }
                } else if constexpr (ILW) {
                    const R P0 = rec[rec_at<R, FPW>(l, F.slot, 4)];
                    const R P1 = rec[rec_at<R, FPW>(l, F.slot, 5)];
                    const R P2 = rec[rec_at<R, FPW>(l, F.slot, 6)];
                    const R P3 = rec[rec_at<R, FPW>(l, F.slot, 7)];
                    R om[BPL];
                    wt_all(l, ex, ey, pe, qe, om, 0);
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        if (s < NC) w[s][l] = om[s];
                        D[s][0] += om[s] * P0;
                        D[s][1] += om[s] * P1;
                        D[s][2] += om[s] * P2;
                        D[s][3] += om[s] * P3;
                    }
                } else {
                    const R P0 = rec[rec_at<R, FPW>(l, F.slot, 4)];
                    const R P1 = rec[rec_at<R, FPW>(l, F.slot, 5)];
                    const R P2 = rec[rec_at<R, FPW>(l, F.slot, 6)];
                    const R P3 = rec[rec_at<R, FPW>(l, F.slot, 7)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        const R om = wt(s, l, ex, ey, pe, qe);
                        if (s < NC) w[s][l] = om;
                        D[s][0] += om * P0;
                        D[s][1] += om * P1;
                        D[s][2] += om * P2;
                        D[s][3] += om * P3;
                    }
                }
            }
        } else {
#pragma unroll
            for (int l = l0; l < l0 + GL; ++l)
#pragma unroll
                for (int s = 0; s < BPL; ++s)
                    if (s < NC) w[s][l] = R(0);
        }
        // keep the scheduler from hoisting every record load of the unrolled
        // loop (which would spill)
        __builtin_amdgcn_sched_barrier(0);
        if (l0 == FS / 2 - GL) hook_a();
    }
    hook_b();
    if constexpr (MF32) {
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) D[s][c] = Dv[s][c];
    }
    }

    // ---- F - I = D K^-1, plane-stress sigma, Q_k = V_k sigma K^-1 ----
    const double pf = F.pf;
    const R c11 = R(pf), c12 = R(pf * a.nu), c33 = R(pf * (1.0 - a.nu));
    R Q[BPL][4];
    double dsum = 0.0;
    // A guarded family's damage is updated by the FP64 fix-up that follows,
    // from its FP64 strain; the mixed kernel leaves its state alone (an update
    // from the FP32 strain first would raise kappa to a wrong value that the
    // fix-up could not lower).
    bool dmg_here = true;
    if constexpr (FamLoad<R, BPL>::GUARD) dmg_here = !(a.famlist && guard);
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        const R K00 = kin[s][0], K01 = kin[s][1], K11 = kin[s][2];
        const R G00 = D[s][0] * K00 + D[s][1] * K01;
        const R G01 = D[s][0] * K01 + D[s][1] * K11;
        const R G10 = D[s][2] * K00 + D[s][3] * K01;
        const R G11 = D[s][2] * K01 + D[s][3] * K11;
        const R e11 = G00, e22 = G11, e12 = R(0.5) * (G01 + G10);
        if constexpr (DMG) {
            // a broken bond counts as fully damaged; a singular K (K^-1 = 0)
            // gives zero strain, so its bond is left as it was
            double d = L.dm[s];
            if (S[s].act && dmg_here) {
                if (!alive[s]) {
                    if (d != 1.0) {
                        d = 1.0;
                        a.dmg[S[s].b] = 1.0;
                    }
                } else {
                    double kap = L.kap[s];
                    if (mazars_update(a, double(e11), double(e22), double(e12), kap, d)) {
                        a.kappa[S[s].b] = kap;
                        a.dmg[S[s].b] = d;
                    }
                }
                dsum += d;
            }
        }
        const R s11 = c11 * e11 + c12 * e22;
        const R s22 = c12 * e11 + c11 * e22;
        const R s12 = c33 * e12;
        // the matrix-core and rank-form records carry Q scaled by 2^-n2
        const R Vk = alive[s] ? R(S[s].Vj) * ((MW || RF) ? mw_scale : R(1)) : R(0);
        Q[s][0] = (s11 * K00 + s12 * K01) * Vk;
        Q[s][1] = (s11 * K01 + s12 * K11) * Vk;
        Q[s][2] = (s12 * K00 + s22 * K01) * Vk;
        Q[s][3] = (s12 * K01 + s22 * K11) * Vk;
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        if constexpr (MW) {
            constexpr int PSPm = MWL<R, LPF, BPL>::PSP;
            R* r8 = rec + MWL<R, LPF, BPL>::at(F.slot, MWL<R, LPF, BPL>::PROW, S[s].k);
#pragma unroll
            for (int c = 0; c < 4; ++c) r8[c * PSPm] = Q[s][c];
        } else if constexpr (RF) {
            rec_put4<FPW>(rec, S[s].k, F.slot, 1, Q[s][0], Q[s][1], Q[s][2], Q[s][3]);
        } else {
            const int k = S[s].k;
            rec_put4<FPW>(rec, k, F.slot, 1, Q[s][0], Q[s][1], Q[s][2], Q[s][3]);
        }
    }
    wave_sync();

    // ---- contraction 2: S_ij = sum_k w(j,k) V_k sigma_ik K_ik^-1 ----
    R Sm[BPL][4];
#pragma unroll
    for (int s = 0; s < BPL; ++s) Sm[s][0] = Sm[s][1] = Sm[s][2] = Sm[s][3] = R(0);
    if constexpr (MW) {
        mw_contract<R, LPF, BPL, N2_IS_10>(rec, lane, n2r, NoHook(), NoHook());
        wave_sync();
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) Sm[s][c] = rec[MWL<R, LPF, BPL>::at(F.slot, MWL<R, LPF, BPL>::PROW + c, S[s].k)];
    } else if constexpr (MF64) {
        double Sacc[WF][NS];
#pragma unroll
        for (int f = 0; f < WF; ++f) {
            // the records' geometry values (0..3) are still in place
            MfK<NS> ks;
            if constexpr (MFRW) ks.template load<R, FPW>(rec, f, lk);
#pragma unroll
            for (int s = 0; s < NS; ++s) Sacc[f][s] = 0.0;
            MfL nb[2];
            nb[0].template load<R, FPW, MFRW>(rec, f, lq, lc);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q + 1 < NQ) nb[(q + 1) & 1].template load<R, FPW, MFRW>(rec, f, 4 * (q + 1) + lq, lc);
                if (4 * q < Mf[f]) {
                    const MfL& r = nb[q & 1];
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        double om;
                        if constexpr (MFRW)
                            om = ks.template w<N2_IS_10>(s, r.ex, r.ey, r.pe, r.qe, a.n2);
                        else
                            om = wc[f][q][s];
                        Sacc[f][s] = __builtin_amdgcn_mfma_f64_4x4x4f64(r.y, om, Sacc[f][s], 0, 0, 0);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wave_sync();
#pragma unroll
        for (int f = 0; f < WF; ++f)
#pragma unroll
            for (int s = 0; s < NS; ++s) rec[rec_at<R, FPW>(16 * s + lk, f, 4 + lq)] = Sacc[f][s];
        wave_sync();
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) Sm[s][c] = rec[rec_at<R, FPW>(S[s].k, F.slot, 4 + c)];
    } else if constexpr (TRX) {
        // the kept blocks (s, b), b >= s, then the transposed ones
        const int tb = TRO + F.slot * (16 * 17);   // the family's transpose tile
        v4f Sv[BPL];
#pragma unroll
        for (int s = 0; s < BPL; ++s) Sv[s] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int l0 = 0; l0 < FS; l0 += GL) {
            if (l0 < mmax) {
#pragma unroll
                for (int l = l0; l < l0 + GL; ++l) {
                    const R Qa = rec[rec_at<R, FPW>(l, F.slot, 4 + lc)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s)
                        if (s <= (l >> 4)) Sv[s] = __builtin_amdgcn_mfma_f32_4x4x1f32(Qa, w[s][l], Sv[s], 0, 0, 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        trx_block<FPW, 0, 1>(w, rec, tb, F.t, F.slot, lc, mmax, Sv[1]);
        trx_block<FPW, 0, 2>(w, rec, tb, F.t, F.slot, lc, mmax, Sv[2]);
        trx_block<FPW, 1, 2>(w, rec, tb, F.t, F.slot, lc, mmax, Sv[2]);
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) Sm[s][c] = Sv[s][c];
    } else {
    v4f Sv[BPL];
#pragma unroll
    for (int s = 0; s < BPL; ++s) Sv[s] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int l0 = 0; l0 < FS; l0 += GL) {
        if (l0 < mmax) {
#pragma unroll
            for (int l = l0; l < l0 + GL; ++l) {
                R ex = R(0), ey = R(0), pe = R(0), qe = R(0);
                if constexpr (RW) {
                    ex = rec[rec_at<R, FPW>(l, F.slot, 0)];
                    ey = rec[rec_at<R, FPW>(l, F.slot, 1)];
                    pe = rec[rec_at<R, FPW>(l, F.slot, 2)];
                    qe = rec[rec_at<R, FPW>(l, F.slot, 3)];
                }
                if constexpr (MF32) {
                    const R Qa = rec[rec_at<R, FPW>(l, F.slot, 4 + lc)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        R om;
                        if (s >= NC)
                            om = wt(s, l, ex, ey, pe, qe);
                        else
                            om = w[s][l];
                        Sv[s] = __builtin_amdgcn_mfma_f32_4x4x1f32(Qa, om, Sv[s], 0, 0, 0);
                    }
                } else if constexpr (ILW) {
                    const R Q0 = rec[rec_at<R, FPW>(l, F.slot, 4)];
                    const R Q1 = rec[rec_at<R, FPW>(l, F.slot, 5)];
                    const R Q2 = rec[rec_at<R, FPW>(l, F.slot, 6)];
                    const R Q3 = rec[rec_at<R, FPW>(l, F.slot, 7)];
                    R om[BPL];
                    wt_all(l, ex, ey, pe, qe, om, NC);
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        const R o = s >= NC ? om[s] : w[s][l];
                        Sm[s][0] += o * Q0;
                        Sm[s][1] += o * Q1;
                        Sm[s][2] += o * Q2;
                        Sm[s][3] += o * Q3;
                    }
                } else {
                    const R Q0 = rec[rec_at<R, FPW>(l, F.slot, 4)];
                    const R Q1 = rec[rec_at<R, FPW>(l, F.slot, 5)];
                    const R Q2 = rec[rec_at<R, FPW>(l, F.slot, 6)];
                    const R Q3 = rec[rec_at<R, FPW>(l, F.slot, 7)];
#pragma unroll
                    for (int s = 0; s < BPL; ++s) {
                        R om;
                        if (s >= NC)
                            om = wt(s, l, ex, ey, pe, qe);
                        else
                            om = w[s][l];
                        Sm[s][0] += om * Q0;
                        Sm[s][1] += om * Q1;
                        Sm[s][2] += om * Q2;
                        Sm[s][3] += om * Q3;
                    }
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (MF32) {
#pragma unroll
        for (int s = 0; s < BPL; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) Sm[s][c] = Sv[s][c];
    }
    }
    // V_H and the own-bond force sum accumulate in the bond arithmetic's
    // precision (FP32 in mixed precision, like the T they sum; measured -2 %)
    R vsum = R(0);
#pragma unroll
    for (int s = 0; s < BPL; ++s) vsum += alive[s] ? R(S[s].Vj) : R(0);
    const double VH = family_sum<LPF>(vsum);
    // 1 / V_H: FP64 division, or the hardware reciprocal in FP32
    R ivh;
    if constexpr (sizeof(R) == 8)
        ivh = VH > 0.0 ? 1.0 / VH : 0.0;
    else
        ivh = VH > 0.0 ? __builtin_amdgcn_rcpf(float(VH)) : 0.0f;
    R T0[BPL], T1[BPL];
    R fo0 = R(0), fo1 = R(0);
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        T0[s] = R(0);
        T1[s] = R(0);
        if (alive[s]) {
            const R x0r = R(S[s].x0), x1r = R(S[s].x1);
            T0[s] = (Sm[s][0] * x0r + Sm[s][1] * x1r) * ivh;
            T1[s] = (Sm[s][2] * x0r + Sm[s][3] * x1r) * ivh;
        }
        fo0 += T0[s] * R(S[s].Vj);
        fo1 += T1[s] * R(S[s].Vj);
    }

    // ---- outputs: own-bond sum and the reverse-slot terms ----
    fo0 = family_sum<LPF>(fo0);
    fo1 = family_sum<LPF>(fo1);
    if (F.famok && F.t == 0) a.fown[F.fam] = make_double2(double(fo0), double(fo1));
    if constexpr (DMG) {
        // damageStatus_i = sum_k d_ik / |H_i| (src/main_seq.cpp:593)
        dsum = family_sum<LPF>(dsum);
        if (F.famok && F.t == 0 && dmg_here) a.dstat[F.fam] = F.M > 0 ? dsum / F.M : 0.0;
    }
#pragma unroll
    for (int s = 0; s < BPL; ++s) {
        if (S[s].act) {
            R2 t;
            t.x = T0[s] * R(F.Vi);
            t.y = T1[s] * R(F.Vi);
            if constexpr (FIX)
                reinterpret_cast<float2*>(a.W)[L.rv[s]] = make_float2(float(t.x), float(t.y));
            else
                reinterpret_cast<R2*>(a.W)[L.rv[s]] = t;
        }
    }
}

// Non-persistent launch: one wave per family group.
template <typename R, int LPF, int BPL, int N2_IS_10, int MF, int DMG>
__global__ void __launch_bounds__(512) k_bond(BondArgs a)
{
    extern __shared__ __attribute__((aligned(16))) double smem[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, a.xcd);
    FamLoad<R, BPL> L;
    stage_rows<LPF>(a, L, lane, blk * wpb + wib);
    stage_cols<R, LPF, BPL, DMG>(a, L);
    stage_bonds<R, LPF, BPL>(a, L);
    bond_body<R, LPF, BPL, N2_IS_10, MF, DMG>(a, L, smem + wib * WaveLds<R, LPF, BPL, MF>::DOUBLES, lane,
                                              NoHook(), NoHook());
}

// Persistent launch: each block owns a contiguous range of family groups (per
// XCD after the remap); each of its waves walks the range with stride wpb and
// issues the next group's three load stages inside the current group's
// contraction, so the dependent gathers overlap compute.  Every wave leaves
// the loop when its range is done (no grid-wide wait).
template <typename R, int LPF, int BPL, int N2_IS_10, int MF, int DMG>
__global__ void __launch_bounds__(512) k_bond_p(BondArgs a)
{
    constexpr int FPW = Tile<LPF, BPL>::FPW;
    extern __shared__ __attribute__((aligned(16))) double smem[];
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int blk = xcd_remap(blockIdx.x, gridDim.x, a.xcd);
    const int ngroups = (a.n + FPW - 1) / FPW;
    const int per = ngroups / (int)gridDim.x, rem = ngroups % (int)gridDim.x;
    const int g0 = blk * per + (blk < rem ? blk : rem);
    const int g1 = g0 + per + (blk < rem ? 1 : 0);
    double* rec64 = smem + wib * WaveLds<R, LPF, BPL, MF>::DOUBLES;
    int g = g0 + wib;
    if (g >= g1) return;
    FamLoad<R, BPL> cur, nxt;
    stage_rows<LPF>(a, cur, lane, g);
    stage_cols<R, LPF, BPL, DMG>(a, cur);
    stage_bonds<R, LPF, BPL>(a, cur);
    for (;;) {
        const int gn = g + wpb;
        const bool more = gn < g1;
        // the last group prefetches itself again (discarded): the loads stay
        // unconditional, so waits on the current group's data are counted
        // exactly instead of draining the prefetch
        const int gp = more ? gn : g;
        stage_rows<LPF>(a, nxt, lane, gp);
        bond_body<R, LPF, BPL, N2_IS_10, MF, DMG>(
            a, cur, rec64, lane,
            [&]() { stage_cols<R, LPF, BPL, DMG>(a, nxt); },
            [&]() { stage_bonds<R, LPF, BPL>(a, nxt); });
        if (!more) break;
        cur = nxt;
        g = gn;
        wave_sync();
    }
}

// FP64 re-evaluation of the guarded families a mixed-precision k_bond launch
// collected (a.famlist, *a.famcnt entries): one family per wave on the 64-lane
// tile, K assembled in FP64, the FP32 T terms and f_own overwritten, and with
// DMG the family's damage updated from its FP64 strain.  Launched right after
// that k_bond on the same stream, with a.s0 = 0 (the bonds broke there) and
// a.force_k = 1; the gather that follows clears the count.
template <int N2_IS_10, int DMG>
__global__ void __launch_bounds__(256) k_bond_fix(BondArgs a)
{
    extern __shared__ __attribute__((aligned(16))) double smem[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const unsigned int c0 = *a.famcnt;
    const unsigned int cnt = c0 < (unsigned int)a.n ? c0 : (unsigned int)a.n;
    for (unsigned int e = blockIdx.x * wpb + wib; e < cnt; e += gridDim.x * wpb) {
        FamLoad<double, 1> L;
        L.F = load_family<64>(a, lane, a.famlist[e]);
        stage_cols<double, 64, 1, DMG, 1>(a, L);
        stage_bonds<double, 64, 1>(a, L);
        bond_body<double, 64, 1, N2_IS_10, 0, DMG, NoHook, NoHook, 1>(
            a, L, smem + wib * Tile<64, 1>::LDS_DOUBLES, lane, NoHook(), NoHook());
        wave_sync();
    }
}

}  // namespace pd
