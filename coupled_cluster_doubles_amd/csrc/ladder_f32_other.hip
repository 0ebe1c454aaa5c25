// Grouped pppp-ladder instances: FP32 iterations, stored V and the generic generator.
#include "ladder_inst.hpp"

void launch_ladder_grp_f32_gas(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                               const LadderOps<float> &o);
void launch_ladder_grp_f32_gasb(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                                const LadderOps<float> &o);

void launch_ladder_grp(int src, int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                       const LadderOps<float> &o) {
    if (src == LV_GAS) launch_ladder_grp_f32_gas(g, grid, lds, st, L, o);
    else if (src == LV_GAS_BATCH) launch_ladder_grp_f32_gasb(g, grid, lds, st, L, o);
    else if (src == LV_GENERIC) ladder_inst::launch_src<float, LV_GENERIC>(g, grid, lds, st, L, o);
    else ladder_inst::launch_src<float, LV_STORED>(g, grid, lds, st, L, o);
}
int allow_ladder_lds_f32_other() {
    hipError_t e = ladder_inst::allow_src<float, LV_STORED>();
    return (int)(e != hipSuccess ? e : ladder_inst::allow_src<float, LV_GENERIC>());
}
