// The spectral M2L of the cube method on the device (placeholder until the batched
// FFT kernels land): plans built with FFMM_M2L_SPECTRAL are rejected at run time.
#include "ffmm_internal.h"

namespace ffmm {

void m2l_spectral_level(Plan &plan, int l) {
  (void)plan;
  (void)l;
  fail(FFMM_ERR_ARGUMENT, "the spectral M2L is not available in this build");
}

}  // namespace ffmm
