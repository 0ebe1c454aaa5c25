// Host entry points of the grouped pppp-ladder instances (k_ladder_grp),
// compiled in their own translation units (ladder_*.hip), one per
// arithmetic type and V source family, so that they build in parallel.
#pragma once

#include "kernels_ladder.hpp"

// Operands of one grouped ladder launch.
template <typename CT>
struct LadderOps {
    const DTile *tiles;
    const int32_t *bptr;
    const DChan *chans;
    const CT *vp;
    LadderGen lg;
    const CT *tas;
    const double *zero;
    CT *dp;
    const LoopCtl *ctl;
};

// Launch group g (0: classes 1, 1: class 2, 2: classes 3-4, 3: classes 5-8,
// 4: classes 1-8) of V source src (LadderSource) on `st`.
void launch_ladder_grp(int src, int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                       const LadderOps<float> &o);
void launch_ladder_grp(int src, int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                       const LadderOps<double> &o);
// Raise each instance's dynamic LDS limit to the most it can be launched
// with; returns a hipError_t.
int allow_ladder_lds_f32_gas();
int allow_ladder_lds_f32_other();
int allow_ladder_lds_f64_gas();
int allow_ladder_lds_f64_other();
int allow_ladder_lds_f32_gasb();
int allow_ladder_lds_f64_gasb();
