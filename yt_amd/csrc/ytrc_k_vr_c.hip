// ytrc_k_vr_c.hip -- VolumeRenderSampler kernels with early ray termination.
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

pfn_t pick_persist_vr_c(int nfp, bool rec, int mode)
{
    switch (mode) {
    case 4: return variant_p<1, 4>(nfp, rec);
    case 5: return variant_p<1, 5>(nfp, rec);
    case 6: return variant_p<1, 6>(nfp, rec);
    case 7: return variant_p<1, 7>(nfp, rec);
    default: return nullptr;
    }
}

kfn_t pick_kernel_vr_c(int nfp, bool rec, int mode)
{
    switch (mode) {
    case 4: return variant_k<1, 4>(nfp, rec);
    case 5: return variant_k<1, 5>(nfp, rec);
    case 6: return variant_k<1, 6>(nfp, rec);
    case 7: return variant_k<1, 7>(nfp, rec);
    default: return nullptr;
    }
}

}  // namespace ytrc
