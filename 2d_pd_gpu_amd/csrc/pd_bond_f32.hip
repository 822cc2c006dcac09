// pd_bond_f32.hip -- mixed-precision bond-kernel variants: tiles (32,1),
// (64,1), (16,2), (32,2), (16,3), each with VALU or matrix-core contractions,
// and (16,2), (16,3) with rank-form weights and either contraction, with and
// without the damage update.
#include "pd_bond_launch.h"

namespace pd {

template <int MF, int DMG>
static hipError_t launch_f32_mf(const BondArgs& a, int lpf, int bpl, int n2_is_10, int wpb, int pgrid,
                                hipStream_t s)
{
    if (bpl == 1 && lpf == 32) return launch_bond_n2<float, 32, 1, MF, DMG>(a, n2_is_10, wpb, pgrid, s);
    if (bpl == 1 && lpf == 64) return launch_bond_n2<float, 64, 1, MF, DMG>(a, n2_is_10, wpb, pgrid, s);
    if (bpl == 2 && lpf == 16) return launch_bond_n2<float, 16, 2, MF, DMG>(a, n2_is_10, wpb, pgrid, s);
    if (bpl == 2 && lpf == 32) return launch_bond_n2<float, 32, 2, MF, DMG>(a, n2_is_10, wpb, pgrid, s);
    if (bpl == 3 && lpf == 16) return launch_bond_n2<float, 16, 3, MF, DMG>(a, n2_is_10, wpb, pgrid, s);
    return hipErrorInvalidValue;
}

hipError_t launch_bond_f32(const BondArgs& a, int lpf, int bpl, int mf, int dmg, int n2_is_10, int wpb,
                           int pgrid, hipStream_t s)
{
    if (mf == 2) return hipErrorInvalidValue;   // matrix-core weights: FP64 only
    if (mf == 3 || mf == 4) {
        // rank-form weights, VALU (3) or matrix-core (4) contractions
        if (bpl == 2 && lpf == 16)
            return mf == 3 ? (dmg ? launch_bond_n2<float, 16, 2, 3, 1>(a, n2_is_10, wpb, pgrid, s)
                                  : launch_bond_n2<float, 16, 2, 3, 0>(a, n2_is_10, wpb, pgrid, s))
                           : (dmg ? launch_bond_n2<float, 16, 2, 4, 1>(a, n2_is_10, wpb, pgrid, s)
                                  : launch_bond_n2<float, 16, 2, 4, 0>(a, n2_is_10, wpb, pgrid, s));
        if (bpl == 3 && lpf == 16)
            return mf == 3 ? (dmg ? launch_bond_n2<float, 16, 3, 3, 1>(a, n2_is_10, wpb, pgrid, s)
                                  : launch_bond_n2<float, 16, 3, 3, 0>(a, n2_is_10, wpb, pgrid, s))
                           : (dmg ? launch_bond_n2<float, 16, 3, 4, 1>(a, n2_is_10, wpb, pgrid, s)
                                  : launch_bond_n2<float, 16, 3, 4, 0>(a, n2_is_10, wpb, pgrid, s));
        return hipErrorInvalidValue;
    }
    if (dmg) return mf ? launch_f32_mf<1, 1>(a, lpf, bpl, n2_is_10, wpb, pgrid, s)
                       : launch_f32_mf<0, 1>(a, lpf, bpl, n2_is_10, wpb, pgrid, s);
    return mf ? launch_f32_mf<1, 0>(a, lpf, bpl, n2_is_10, wpb, pgrid, s)
              : launch_f32_mf<0, 0>(a, lpf, bpl, n2_is_10, wpb, pgrid, s);
}

}  // namespace pd
