// pd_internal.h -- kernel argument blocks and launchers shared by the HIP
// kernels (pd_kernels.hip) and the context / C-ABI layer (pd_context.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pd {

// Everything the bond kernel needs for one launch.  Particle arrays are in the
// internal (space-filling-curve) order; bond arrays in the internal CSR order,
// one family (the bonds of particle i) contiguous and in ascending order of
// the bonds' reference lengths |X_j - X_i| (ties in the caller's order).
struct BondArgs {
    int n;                    // families (= particles)
    int xcd;                  // XCDs to spread consecutive blocks over (<= 1: off)
    const double2* X;         // initial positions
    const double* V;          // volumes
    const int* row;           // CSR row pointer, n+1
    const int* col;           // neighbour of each bond
    const int* rev;           // reverse bond of each bond
    const double2* u;         // displacement y - X
    unsigned char* broken;    // per bond, 1 = broken (irreversible)
    void* kinv;               // per bond K_ik^-1 (xx, xy, yy), assembled and inverted in
                              // FP64, stored in the bond arithmetic's precision
    int kinv_f32;             // 1: stored as float (mixed precision), 0: double
    void* W;                  // per bond, slot rev[b] <- T_ij V_i (double2 / float2)
    double2* fown;            // per particle sum_j T_ij V_j
    const double* pf;         // plane-stress factor E/(1-nu^2): pf[i & pf_mask]
    int pf_mask;              // ~0: one value per particle; 0: one value for all
    double nu;                // Poisson ratio
    double A;                 // n1 / (horizon * dx)
    double n2;                // angular exponent
    double s0;                // critical stretch (<= 0: none)
    double brk_c;             // (1 + s0)^2
    double ktol;              // singular-K tolerance
    int force_k;              // 1: assemble K for every family in this launch
    unsigned int* counters;   // [0] newly broken bonds, [1] singular K, [2] first non-finite,
                              // [3] ill-conditioned K (pd_get_stats)
    // Mazars bond damage (kernels built with damage only)
    int dmg_mode;             // 1: running threshold in the law (reference), 2: initial
    double kap0, At, Bt, Ac, Bc;
    double* kappa;            // per bond threshold
    double* dmg;              // per bond damage
    double* dstat;            // per particle damageStatus
    // Mixed-precision guard: a family with an alive bond whose K has a condition
    // number of at least the guard's bound is evaluated again in FP64 after the
    // mixed kernel (k_bond_fix), which overwrites its T terms and f_own.
    unsigned char* famflag;   // per family: 1 = guarded (maintained by every K assembly)
    int* famlist;             // guarded families collected this launch (nullptr: none)
    unsigned int* famcnt;     // entries in famlist (cleared by the gather that follows)
    double illr;              // cond(K) >= c  <=>  tr(K)^2 >= illr det(K), illr = c + 2 + 1/c
};

struct GatherArgs {
    int n;
    int xcd;
    const int* row;
    const double2* fown;
    const void* W;
    double2* v;
    double2* a;
    double2* f;
    double rho, dt;
    int write_a;
    int max_family;           // largest family: picks the launch shape
    unsigned int* famcnt;     // mixed-precision guard: the list count the preceding
                              // k_bond / k_bond_fix used, cleared here (nullptr: none)
};

struct PosArgs {
    int n;
    double2* u;
    double2* v;
    double2* a;
    const double2* f;
    const int* bc;            // per particle code, or nullptr
    const double2* v0;        // per velocity group
    double rho, dt;
};

// Tiles: (lanes per family, bonds per lane) in {(32,1), (64,1), (16,2),
// (16,3)} and, for mixed precision, (32,2); a family may hold up to
// lpf * bpl bonds.
// mf = 1: family contractions on the matrix cores (every tile except FP64
// (32,2), which does not exist).  mf = 2 (FP64 only): matrix-core weights and
// contractions (tiles (32,1), (64,1), (16,3); n2 > 0).  mf = 3: rank-form
// weights, VALU contractions; mf = 4 (mixed only): rank-form weights with
// matrix-core contractions (tiles (16,2), (16,3); n2 > 0).
bool tile_supported(int mixed, int lpf, int bpl, int mf);
// pgrid > 0: persistent launch with pgrid workgroups (each walks a contiguous
// range of family groups, prefetching the next group's loads).
// dmg = 1: also update the Mazars bond damage.
hipError_t launch_bond(const BondArgs& a, int mixed, int lpf, int bpl, int mf, int dmg, int n2_is_10,
                       int wpb, int pgrid, hipStream_t s);
hipError_t launch_bond_f64(const BondArgs& a, int lpf, int bpl, int mf, int dmg, int n2_is_10, int wpb,
                           int pgrid, hipStream_t s);
hipError_t launch_bond_f64_rank(const BondArgs& a, int lpf, int bpl, int mf, int dmg, int n2_is_10, int wpb,
                                int pgrid, hipStream_t s);
hipError_t launch_bond_f32(const BondArgs& a, int lpf, int bpl, int mf, int dmg, int n2_is_10, int wpb,
                           int pgrid, hipStream_t s);
hipError_t launch_shape(const BondArgs& a, int lpf, int bpl, int n2_is_10, int wpb, hipStream_t s);
// next != nullptr: also run the following step's position update (a
// pd_step of several steps), as launch_positions(*next) would after it.
hipError_t launch_gather(const GatherArgs& a, int mixed, hipStream_t s, const PosArgs* next = nullptr);
hipError_t launch_positions(const PosArgs& a, hipStream_t s);
hipError_t launch_velocity_bc_list(int count, const int* ids, double2 v0, double2* v,
                                   double2* a, hipStream_t s);
hipError_t launch_fixed_bc_list(int count, const int* ids, double2* u, hipStream_t s);
hipError_t launch_fill_u32(unsigned int* p, unsigned int v, int count, hipStream_t s);
// Each family's (col, bmap) entries in ascending order of |X_j - X_i|^2,
// ties in their present order (the order of the families' bond storage).
hipError_t launch_length_order(int n, const int* row, const double2* X, int* col, int* bmap, hipStream_t s);

// The length order of one family of m <= 64 bonds (neighbours cj, the
// particle's position xi): ord[r] = the family position of the bond of rank r,
// ascending |X_j - X_i|^2 with ties in position order.  Each operation is
// rounded on its own, as in the host set-up.  Neighbour indices outside
// [0, n) sort as zero length (callers reject such meshes).
__device__ inline void family_length_order(int m, const int* cj, const double2* X, int n, double2 xi, int* ord)
{
#pragma clang fp contract(off)
    double len2[64];
    for (int e = 0; e < m; ++e) {
        const int j = cj[e];
        const double2 xj = (j >= 0 && j < n) ? X[j] : xi;
        const double d0 = xj.x - xi.x, d1 = xj.y - xi.y;
        const double s0 = d0 * d0, s1 = d1 * d1;
        len2[e] = s0 + s1;
    }
    for (int e = 0; e < m; ++e) {
        int r = 0;
        for (int f = 0; f < m; ++f) r += (len2[f] < len2[e] || (len2[f] == len2[e] && f < e)) ? 1 : 0;
        ord[r] = e;
    }
}
hipError_t launch_bond_fix(const BondArgs& a, int n2_is_10, int dmg, int grid, hipStream_t s);
hipError_t launch_count_ill_conditioned(long nb, const void* kinv, bool kinv_f32, const unsigned char* broken,
                                       double cmin, unsigned int* count, hipStream_t s);
// *first = smallest index with a non-finite u, v or f (0xffffffff if none).
hipError_t launch_check_finite(int n, const double2* u, const double2* v, const double2* f,
                               unsigned int* first, hipStream_t s);
hipError_t launch_copy_probe(const double4* src, double4* dst, long n4, hipStream_t s);

}  // namespace pd
