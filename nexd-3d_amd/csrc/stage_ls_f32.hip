// stage_kernel instantiations for float, low-storage RK4.
#include "stage_impl.hpp"

template void nexd::launch_stage<float, true, 0>(int, bool, bool, const nexd::StageArgs<float> &, hipStream_t);
