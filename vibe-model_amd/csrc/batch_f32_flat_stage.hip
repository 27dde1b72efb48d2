// batch_f32_flat_stage.hip -- the FP32 batched stage kernels of the kGeoFlat
// geometry compaction, compiled with -ffp-contract=fast like the substep
// kernels of batch_f32_flat.hip.  Stage and substep kernels sit in separate
// files so that each set gets the instruction-scheduling strategy measured
// fastest for it (Makefile).
#include "batch_kernels.hpp"

VMI_BATCH_F32_STAGES(, vibe_mi::kGeoFlat)
