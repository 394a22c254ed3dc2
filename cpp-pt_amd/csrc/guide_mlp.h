// Learned path guiding: a small MLP that maps a path vertex (position, shading
// normal, outgoing direction) to a mixture of von Mises-Fisher lobes over the
// sphere of incident directions.
//
// It plays the role of the SD-tree of practical path guiding (the reference's
// PPG renderer, src/renderer/ppg.h:41-214 with src/algorithms/sdtree.h): a
// directional distribution queried at each non-delta vertex, sampled with
// probability GUIDE_FRACTION and combined with BSDF sampling by one-sample MIS
// (balance heuristic, as ppg.h:74-103), and trained online from the radiance
// the tracer's own paths carry back along the sampled directions (the role of
// SDTree::splat/update, sdtree.h:445-538).
//
// Network: one-blob encoding of 7 inputs x 32 bins = 224 features ->
// 4 hidden layers of width 64 (ReLU) -> 48 outputs (8 lobes x [logit,
// log kappa, mu.xyz], 8 spare), or width 128 -> 80 outputs (16 lobes) for
// large scenes; the two are instances of guide_net.h in the namespaces guide
// and guide128.  BF16 weights and activations on MFMA
// (v_mfma_f32_16x16x16_bf16) with FP32 accumulation; FP32 master weights and
// Adam moments.  A block of 4 waves processes tiles of 64 samples; wave w owns
// rows 16w..16w+15 of every activation matrix, so the layers of one tile need
// no data exchange between waves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpt {

namespace guide {  // hidden width 64, 8 lobes
#define CPT_GUIDE_HID 64
#define CPT_GUIDE_LOBES 8
#include "guide_net.h"
#undef CPT_GUIDE_HID
#undef CPT_GUIDE_LOBES
}  // namespace guide

namespace guide128 {  // hidden width 128, 16 lobes
#define CPT_GUIDE_HID 128
#define CPT_GUIDE_LOBES 16
#include "guide_net.h"
#undef CPT_GUIDE_HID
#undef CPT_GUIDE_LOBES
}  // namespace guide128

}  // namespace cpt
