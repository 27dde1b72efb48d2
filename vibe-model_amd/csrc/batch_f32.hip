// batch_f32.hip -- the FP32 batched kernels, compiled with -ffp-contract=fast
// so the mixed-precision stencil and substep arithmetic uses fused
// multiply-adds (the FP64 path stays uncontracted; see batch_kernels.hpp).
#include "batch_kernels.hpp"

VMI_BATCH_F32_ALL()
