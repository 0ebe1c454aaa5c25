// Grouped pppp-ladder instances: FP64 iterations, generated electron gas.
#include "ladder_inst.hpp"

void launch_ladder_grp_f64_gas(int g, unsigned grid, size_t lds, hipStream_t st, const LadderLaunch &L,
                               const LadderOps<double> &o) {
    ladder_inst::launch_src<double, LV_GAS>(g, grid, lds, st, L, o);
}
int allow_ladder_lds_f64_gas() { return (int)ladder_inst::allow_src<double, LV_GAS>(); }
