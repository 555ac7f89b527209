// stage_kernel instantiations for double, low-storage RK4.
#include "stage_impl.hpp"

template void nexd::launch_stage<double, true, 0>(int, bool, bool, const nexd::StageArgs<double> &, hipStream_t);
