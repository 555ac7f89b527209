// stage_kernel instantiations for float, the TVD schemes, block shape 1
// (two blocks per CU; the autotuner's alternative to shape 0).
#include "stage_impl.hpp"

template void nexd::launch_stage<float, false, 1>(int, bool, bool, const nexd::StageArgs<float> &, hipStream_t);
