// Definition of launch_stage (stage_launch.hpp), included only by the
// stage_*.hip translation units that instantiate it.
#pragma once
#include "stage_launch.hpp"

namespace nexd {

template <typename R, bool LS, int VAR, int ORDER, bool ATT, bool PML>
void launch_stage_one(const StageArgs<R> &a, hipStream_t s) {
    using S = Shape<ORDER, ATT, PML, R, VAR>;
    const unsigned grid = (unsigned)((a.nlist + S::EB - 1) / S::EB) * (unsigned)a.nshot;
    if (grid == 0) return;
    hipLaunchKernelGGL((stage_kernel<ORDER, ATT, PML, R, LS, VAR>), dim3(grid), dim3(S::NTHR), 0, s, a);
}

template <typename R, bool LS, int VAR, int ORDER>
void launch_stage_order(bool att, bool pml, const StageArgs<R> &a, hipStream_t s) {
    if (att && pml) launch_stage_one<R, LS, VAR, ORDER, true, true>(a, s);
    else if (att) launch_stage_one<R, LS, VAR, ORDER, true, false>(a, s);
    else if (pml) launch_stage_one<R, LS, VAR, ORDER, false, true>(a, s);
    else launch_stage_one<R, LS, VAR, ORDER, false, false>(a, s);
}

template <typename R, bool LS, int VAR>
void launch_stage(int order, bool att, bool pml, const StageArgs<R> &a, hipStream_t s) {
    switch (order) {
        case 1: launch_stage_order<R, LS, VAR, 1>(att, pml, a, s); break;
        case 2: launch_stage_order<R, LS, VAR, 2>(att, pml, a, s); break;
        case 3: launch_stage_order<R, LS, VAR, 3>(att, pml, a, s); break;
        case 4: launch_stage_order<R, LS, VAR, 4>(att, pml, a, s); break;
        default: launch_stage_order<R, LS, VAR, 5>(att, pml, a, s); break;
    }
}

}  // namespace nexd
