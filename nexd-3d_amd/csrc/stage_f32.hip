// stage_kernel instantiations for float, the TVD schemes (Euler, RK2, RK3).
#include "stage_impl.hpp"

template void nexd::launch_stage<float, false, 0>(int, bool, bool, const nexd::StageArgs<float> &, hipStream_t);
