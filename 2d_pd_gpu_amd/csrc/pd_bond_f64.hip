// pd_bond_f64.hip -- FP64 bond-kernel variants with records in bond order:
// tiles (32,1), (64,1), (16,2), (16,3) with VALU or matrix-core contractions,
// each with and without the damage update; the FP64 fix-up of the guarded
// families of a mixed-precision step.
#include "pd_bond_launch.h"

namespace pd {

hipError_t launch_bond_f64(const BondArgs& a, int lpf, int bpl, int mf, int dmg, int n2_is_10, int wpb,
                           int pgrid, hipStream_t s)
{
    // the variants with records in length-rank order build in their own unit
    if (mf >= 2) return launch_bond_f64_rank(a, lpf, bpl, mf, dmg, n2_is_10, wpb, pgrid, s);
    if (dmg && mf) {
        if (bpl == 1 && lpf == 32) return launch_bond_n2<double, 32, 1, 1, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 1 && lpf == 64) return launch_bond_n2<double, 64, 1, 1, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 2 && lpf == 16) return launch_bond_n2<double, 16, 2, 1, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 3 && lpf == 16) return launch_bond_n2<double, 16, 3, 1, 1>(a, n2_is_10, wpb, pgrid, s);
    } else if (dmg) {
        if (bpl == 1 && lpf == 32) return launch_bond_n2<double, 32, 1, 0, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 1 && lpf == 64) return launch_bond_n2<double, 64, 1, 0, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 2 && lpf == 16) return launch_bond_n2<double, 16, 2, 0, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 3 && lpf == 16) return launch_bond_n2<double, 16, 3, 0, 1>(a, n2_is_10, wpb, pgrid, s);
    } else if (mf) {
        if (bpl == 1 && lpf == 32) return launch_bond_n2<double, 32, 1, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 1 && lpf == 64) return launch_bond_n2<double, 64, 1, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 2 && lpf == 16) return launch_bond_n2<double, 16, 2, 1>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 3 && lpf == 16) return launch_bond_n2<double, 16, 3, 1>(a, n2_is_10, wpb, pgrid, s);
    } else {
        if (bpl == 1 && lpf == 32) return launch_bond_n2<double, 32, 1, 0>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 1 && lpf == 64) return launch_bond_n2<double, 64, 1, 0>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 2 && lpf == 16) return launch_bond_n2<double, 16, 2, 0>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 3 && lpf == 16) return launch_bond_n2<double, 16, 3, 0>(a, n2_is_10, wpb, pgrid, s);
    }
    return hipErrorInvalidValue;
}

// The FP64 fix-up of a mixed-precision step (k_bond_fix, with the damage
// update of the guarded families when dmg = 1): a grid of `grid`
// workgroups of 4 waves, sized to the device by the caller, so the launch does
// not depend on how many families the preceding k_bond collected.
hipError_t launch_bond_fix(const BondArgs& a, int n2_is_10, int dmg, int grid, hipStream_t s)
{
    if (a.n <= 0 || grid <= 0) return hipSuccess;
    const size_t lds = 4 * Tile<64, 1>::LDS_DOUBLES * sizeof(double);
    if (n2_is_10)
        hipLaunchKernelGGL((dmg ? k_bond_fix<1, 1> : k_bond_fix<1, 0>), dim3(grid), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL((dmg ? k_bond_fix<0, 1> : k_bond_fix<0, 0>), dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

}  // namespace pd
