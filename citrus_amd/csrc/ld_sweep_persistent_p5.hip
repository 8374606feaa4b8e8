// Instantiations of the persistent LD sweep kernel for 36 spherical-harmonic moments (P5), float32 and
// bfloat16 source expansion, every direction-tile count and source layout.
#include "sweep_persistent.h"

namespace citrus {

hipError_t persistent_launch_p5(const PersistentArgs& a, bool bf16, int layout, size_t lds, hipStream_t s) {
  return bf16 ? launch_persistent_src<36, true>(a, layout, lds, s) : launch_persistent_src<36, false>(a, layout, lds, s);
}

}  // namespace citrus
