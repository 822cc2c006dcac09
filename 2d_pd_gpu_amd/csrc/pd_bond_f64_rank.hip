// pd_bond_f64_rank.hip -- FP64 bond-kernel variants that use the families'
// length order: matrix-core weights on tiles (32,1), (64,1), (16,3); rank-form
// weights with VALU contractions on tiles (16,2), (16,3); each with and
// without the damage update.
#include "pd_bond_launch.h"

namespace pd {

template <int MF, int DMG>
static hipError_t launch_rank(const BondArgs& a, int lpf, int bpl, int n2_is_10, int wpb, int pgrid,
                              hipStream_t s)
{
    if constexpr (MF == 2) {
        if (bpl == 1 && lpf == 32) return launch_bond_n2<double, 32, 1, 2, DMG>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 1 && lpf == 64) return launch_bond_n2<double, 64, 1, 2, DMG>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 3 && lpf == 16) return launch_bond_n2<double, 16, 3, 2, DMG>(a, n2_is_10, wpb, pgrid, s);
    } else {
        if (bpl == 2 && lpf == 16) return launch_bond_n2<double, 16, 2, 3, DMG>(a, n2_is_10, wpb, pgrid, s);
        if (bpl == 3 && lpf == 16) return launch_bond_n2<double, 16, 3, 3, DMG>(a, n2_is_10, wpb, pgrid, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_bond_f64_rank(const BondArgs& a, int lpf, int bpl, int mf, int dmg, int n2_is_10, int wpb,
                                int pgrid, hipStream_t s)
{
    switch (mf) {
    case 2: return dmg ? launch_rank<2, 1>(a, lpf, bpl, n2_is_10, wpb, pgrid, s)
                       : launch_rank<2, 0>(a, lpf, bpl, n2_is_10, wpb, pgrid, s);
    case 3: return dmg ? launch_rank<3, 1>(a, lpf, bpl, n2_is_10, wpb, pgrid, s)
                       : launch_rank<3, 0>(a, lpf, bpl, n2_is_10, wpb, pgrid, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pd
