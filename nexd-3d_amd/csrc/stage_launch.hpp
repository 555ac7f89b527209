// Launchers of the stage kernel, one per (real type, scheme class).  Each
// is compiled in its own translation unit (stage_*.hip), so the kernel's
// instantiations build in parallel and only once for the single-shot and
// the batched path.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "plain_kernels.hpp"

namespace nexd {

// Launches stage_kernel<order, att, pml, R, LS, VAR> for a.nlist elements
// and a.nshot shots on stream s (nothing for an empty list).  LS selects the
// low-storage RK4 instantiation, VAR the block shape (FP32 TVD only has
// VAR 1).
template <typename R, bool LS, int VAR>
void launch_stage(int order, bool att, bool pml, const StageArgs<R> &a, hipStream_t s);

// the block shapes the autotuner chooses from, for a real type and scheme
template <typename R>
constexpr int stage_variants(bool ls) { return (sizeof(R) == 4 && !ls) ? 2 : 1; }

// elements per block of an instantiation (block counts of the energy partials)
template <typename R>
inline int stage_eb(int order, bool att, bool pml, int var = 0) {
    auto pick = [&](auto ord) {
        constexpr int O = decltype(ord)::value;
        if (var == 1) {
            if (att && pml) return Shape<O, true, true, R, 1>::EB;
            if (att) return Shape<O, true, false, R, 1>::EB;
            if (pml) return Shape<O, false, true, R, 1>::EB;
            return Shape<O, false, false, R, 1>::EB;
        }
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
inline int stage_blocks(int order, int n, bool att, bool pml, int var = 0) {
    const int eb = stage_eb<R>(order, att, pml, var);
    return (n + eb - 1) / eb;
}

}  // namespace nexd
