// Grouped pppp-ladder instances: FP64 iterations, stored V and the generic generator.
#include "ladder_inst.hpp"

void launch_ladder_grp_f64_gas(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                               const LadderOps<double> &o);
void launch_ladder_grp_f64_gasb(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                                const LadderOps<double> &o);

void launch_ladder_grp(int src, int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                       const LadderOps<double> &o) {
    if (src == LV_GAS) launch_ladder_grp_f64_gas(g, grid, lds, st, L, o);
    else if (src == LV_GAS_BATCH) launch_ladder_grp_f64_gasb(g, grid, lds, st, L, o);
    else if (src == LV_GENERIC) ladder_inst::launch_src<double, LV_GENERIC>(g, grid, lds, st, L, o);
    else ladder_inst::launch_src<double, LV_STORED>(g, grid, lds, st, L, o);
}
int allow_ladder_lds_f64_other() {
    hipError_t e = ladder_inst::allow_src<double, LV_STORED>();
    return (int)(e != hipSuccess ? e : ladder_inst::allow_src<double, LV_GENERIC>());
}
