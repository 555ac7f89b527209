// stage_kernel instantiations for double, the TVD schemes (Euler, RK2, RK3).
#include "stage_impl.hpp"

template void nexd::launch_stage<double, false, 0>(int, bool, bool, const nexd::StageArgs<double> &, hipStream_t);
