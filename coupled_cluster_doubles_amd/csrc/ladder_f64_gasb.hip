// Grouped pppp-ladder instances: FP64 iterations, a matrix-free batch with
// electron-gas solves (per-solve tables, generic generator for the others).
#include "ladder_inst.hpp"

void launch_ladder_grp_f64_gasb(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                                const LadderOps<double> &o) {
    ladder_inst::launch_src<double, LV_GAS_BATCH>(g, grid, lds, st, L, o);
}
int allow_ladder_lds_f64_gasb() { return (int)ladder_inst::allow_src<double, LV_GAS_BATCH>(); }
