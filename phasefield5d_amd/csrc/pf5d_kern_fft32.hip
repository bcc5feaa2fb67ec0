// Instances of the batched FFT kernel in float (see pf5d_instances.h)
#define PF5D_INST template
#include "pf5d_instances.h"

namespace pf5d {
PF5D_FFT32_INSTANCES
}  // namespace pf5d
