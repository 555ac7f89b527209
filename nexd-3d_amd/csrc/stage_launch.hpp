// Launchers of the stage kernel, one per (real type, scheme class).  Each
// is compiled in its own translation unit (stage_*.hip), so the kernel's
// instantiations build in parallel and only once for the single-shot and
// the batched path.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "plain_kernels.hpp"

namespace nexd {

// Launches stage_kernel<order, att, pml, R, LS> for a.nlist elements and
// a.nshot shots on stream s (nothing for an empty list).  LS selects the
// low-storage RK4 instantiation.
template <typename R, bool LS>
void launch_stage(int order, bool att, bool pml, const StageArgs<R> &a, hipStream_t s);

// elements per block of an instantiation (block counts of the energy partials)
template <typename R>
inline int stage_eb(int order, bool att, bool pml) {
    auto pick = [&](auto ord) {
        constexpr int O = decltype(ord)::value;
        if (att && pml) return Shape<O, true, true, R>::EB;
        if (att) return Shape<O, true, false, R>::EB;
        if (pml) return Shape<O, false, true, R>::EB;
        return Shape<O, false, false, R>::EB;
    };
    switch (order) {
        case 1: return pick(std::integral_constant<int, 1>{});
        case 2: return pick(std::integral_constant<int, 2>{});
        case 3: return pick(std::integral_constant<int, 3>{});
        case 4: return pick(std::integral_constant<int, 4>{});
        default: return pick(std::integral_constant<int, 5>{});
    }
}

// element blocks of one list launch (per shot)
template <typename R>
inline int stage_blocks(int order, int n, bool att, bool pml) {
    const int eb = stage_eb<R>(order, att, pml);
    return (n + eb - 1) / eb;
}

}  // namespace nexd
