// Instances of the stencil-stage kernels (see pf5d_instances.h)
#define PF5D_INST template
#include "pf5d_instances.h"

namespace pf5d {
PF5D_STENCIL_INSTANCES
}  // namespace pf5d
