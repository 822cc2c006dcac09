// pd_bond_launch.h -- launch of one bond-kernel variant (included by the
// per-precision instantiation units only).
#pragma once

#include "pd_bond.h"

namespace pd {

template <typename R, int LPF, int BPL, int N2, int MF, int DMG>
static hipError_t launch_bond_t(const BondArgs& a, int wpb, int pgrid, hipStream_t s)
{
    const long groups = ((long)a.n + Tile<LPF, BPL>::FPW - 1) / Tile<LPF, BPL>::FPW;
    const int grid = (int)((groups + wpb - 1) / wpb);
    if (grid == 0) return hipSuccess;
    const size_t lds = (size_t)wpb * WaveLds<R, LPF, BPL, MF>::DOUBLES * sizeof(double);
    if (pgrid > 0 && pgrid < grid)
        hipLaunchKernelGGL((k_bond_p<R, LPF, BPL, N2, MF, DMG>), dim3(pgrid), dim3(64 * wpb), lds, s, a);
    else
        hipLaunchKernelGGL((k_bond<R, LPF, BPL, N2, MF, DMG>), dim3(grid), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

template <typename R, int LPF, int BPL, int MF, int DMG = 0>
static hipError_t launch_bond_n2(const BondArgs& a, int n2_is_10, int wpb, int pgrid, hipStream_t s)
{
    return n2_is_10 ? launch_bond_t<R, LPF, BPL, 1, MF, DMG>(a, wpb, pgrid, s)
                    : launch_bond_t<R, LPF, BPL, 0, MF, DMG>(a, wpb, pgrid, s);
}

}  // namespace pd
