// ytrc_pick.h -- kernel-variant lookup across the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "ytrc_device.h"

namespace ytrc {

// k_persist<KIND, NFP, REC, MODE> and k_render<KIND, NFP, REC, MODE>.
typedef void (*pfn_t)(const DImg *, const int2 *, int, unsigned int *, const float4 *, int, int);
typedef void (*kfn_t)(const DImg *, int);

// Each returns the variant for (nfp, rec, mode) within its unit's family,
// or nullptr.  Families: projection, interpolated projection and light-source
// rendering (modes 0, 8); volume rendering modes
// 0-1, 2-3, 4-7 (early ray termination), 8-11 (per-pixel lens).
pfn_t pick_persist_proj(int nfp, bool rec, int mode);
kfn_t pick_kernel_proj(int nfp, bool rec, int mode);
pfn_t pick_persist_iproj(int nfp, bool rec, int mode);
kfn_t pick_kernel_iproj(int nfp, bool rec, int mode);
pfn_t pick_persist_light(int nfp, bool rec, int mode);
kfn_t pick_kernel_light(int nfp, bool rec, int mode);
pfn_t pick_persist_vr_a(int nfp, bool rec, int mode);
kfn_t pick_kernel_vr_a(int nfp, bool rec, int mode);
pfn_t pick_persist_vr_b(int nfp, bool rec, int mode);
kfn_t pick_kernel_vr_b(int nfp, bool rec, int mode);
pfn_t pick_persist_vr_c(int nfp, bool rec, int mode);
kfn_t pick_kernel_vr_c(int nfp, bool rec, int mode);
pfn_t pick_persist_vr_d(int nfp, bool rec, int mode);
kfn_t pick_kernel_vr_d(int nfp, bool rec, int mode);

}  // namespace ytrc
