// ytrc_k_vr_b.hip -- VolumeRenderSampler kernels: n_samples == 10 unrolled.
#include "ytrc_kernels.h"
#include "ytrc_pick.h"

namespace {

// NFP 4 and 8 exist only without early ray termination (MODE bit 2) and
// without the n_samples == 10 unrolling (MODE bit 1); projections use at
// most 4 interleaved fields.
template <int KIND, int MODE>
ytrc::pfn_t variant_p(int nfp, bool rec)
{
    constexpr bool wide = !(MODE & 6);
    switch (nfp) {
    case 1: return rec ? k_persist<KIND, 1, true, MODE> : k_persist<KIND, 1, false, MODE>;
    case 2: return rec ? k_persist<KIND, 2, true, MODE> : k_persist<KIND, 2, false, MODE>;
    case 4:
        if constexpr (wide) return rec ? k_persist<KIND, 4, true, MODE> : k_persist<KIND, 4, false, MODE>;
        return nullptr;
    default:
        if constexpr (wide && KIND != 0) return rec ? k_persist<KIND, 8, true, MODE> : k_persist<KIND, 8, false, MODE>;
        return nullptr;
    }
}

template <int KIND, int MODE>
ytrc::kfn_t variant_k(int nfp, bool rec)
{
    constexpr bool wide = !(MODE & 6);
    switch (nfp) {
    case 1: return rec ? k_render<KIND, 1, true, MODE> : k_render<KIND, 1, false, MODE>;
    case 2: return rec ? k_render<KIND, 2, true, MODE> : k_render<KIND, 2, false, MODE>;
    case 4:
        if constexpr (wide) return rec ? k_render<KIND, 4, true, MODE> : k_render<KIND, 4, false, MODE>;
        return nullptr;
    default:
        if constexpr (wide && KIND != 0) return rec ? k_render<KIND, 8, true, MODE> : k_render<KIND, 8, false, MODE>;
        return nullptr;
    }
}

}  // namespace

namespace ytrc {

pfn_t pick_persist_vr_b(int nfp, bool rec, int mode)
{
    switch (mode) {
    case 2: return variant_p<1, 2>(nfp, rec);
    case 3: return variant_p<1, 3>(nfp, rec);
    // mode 3 with vertex blocks staged in LDS (MODE bit 4): one field, no record
    case 19: return (nfp == 1 && !rec) ? k_persist<1, 1, false, 19> : nullptr;
    default: return nullptr;
    }
}

kfn_t pick_kernel_vr_b(int nfp, bool rec, int mode)
{
    switch (mode) {
    case 2: return variant_k<1, 2>(nfp, rec);
    case 3: return variant_k<1, 3>(nfp, rec);
    default: return nullptr;
    }
}

}  // namespace ytrc
