// batch_f32_x.hip -- the FP32 batched substep kernels of the kGeoX geometry
// compaction, compiled with -ffp-contract=fast so the mixed-precision stencil
// and substep arithmetic uses fused multiply-adds (the FP64 path stays
// uncontracted; see batch_kernels.hpp).  One file per compaction, so the three
// sets compile in parallel.
#include "batch_kernels.hpp"

VMI_BATCH_F32_SUBSTEPS(, vibe_mi::kGeoX)
