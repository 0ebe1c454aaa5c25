// Instances of the grouped pppp ladder for one arithmetic type and a set of V
// sources (included by the ladder_*.hip translation units).
#pragma once

#include "ladder_launch.hpp"

namespace ladder_inst {

template <typename CT, int SRC, int LO, int HI>
void launch(unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L, const LadderOps<CT> &o) {
    hipLaunchKernelGGL((k_ladder_grp<CT, SRC, LO, HI>), dim3(grid), dim3(kThreads), lds, st, L, o.tiles, o.bptr,
                       o.chans, o.vp, o.lg, o.tas, o.zero, o.dp, o.ctl);
}

template <typename CT, int SRC>
void launch_src(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L, const LadderOps<CT> &o) {
    static_assert(NTMAX == 8, "launch groups 1, 2, 3-4, 5-8, 1-8");
    switch (g) {
        case 0: launch<CT, SRC, 1, 1>(grid, lds, st, L, o); break;
        case 1: launch<CT, SRC, 2, 2>(grid, lds, st, L, o); break;
        case 2: launch<CT, SRC, 3, 4>(grid, lds, st, L, o); break;
        case 3: launch<CT, SRC, 5, 8>(grid, lds, st, L, o); break;
        default: launch<CT, SRC, 1, 8>(grid, lds, st, L, o); break;
    }
}

template <typename CT, int NT>
size_t max_class_bytes() {
    size_t b = 0;
    for (int k = LK_DEFAULT; k <= LK_BF16X3W; k++) b = b > ladder_class_bytes<CT, NT>(k) ? b : ladder_class_bytes<CT, NT>(k);
    return b;
}

template <typename CT>
size_t range_bytes(int lo, int hi) {
    const size_t c[NTMAX + 1] = {0, max_class_bytes<CT, 1>(), max_class_bytes<CT, 2>(), max_class_bytes<CT, 3>(),
                                 max_class_bytes<CT, 4>(), max_class_bytes<CT, 5>(), max_class_bytes<CT, 6>(),
                                 max_class_bytes<CT, 7>(), max_class_bytes<CT, 8>()};
    size_t b = 0;
    for (int n = lo; n <= hi; n++) b = b > c[n] ? b : c[n];
    return b;
}

template <typename CT, int SRC, int LO, int HI>
hipError_t allow() {
    const size_t lds = ladder_tab_bytes(gas_source(SRC) ? kGasTableMax : 0) + range_bytes<CT>(LO, HI);
    return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ladder_grp<CT, SRC, LO, HI>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <typename CT, int SRC>
hipError_t allow_src() {
    hipError_t e;
    if ((e = allow<CT, SRC, 1, 1>()) != hipSuccess || (e = allow<CT, SRC, 2, 2>()) != hipSuccess ||
        (e = allow<CT, SRC, 3, 4>()) != hipSuccess || (e = allow<CT, SRC, 5, 8>()) != hipSuccess ||
        (e = allow<CT, SRC, 1, 8>()) != hipSuccess)
        return e;
    return hipSuccess;
}

}  // namespace ladder_inst
