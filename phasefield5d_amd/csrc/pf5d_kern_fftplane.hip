// Instances of the plane passes of the 3D FFT pair (see pf5d_instances.h)
#define PF5D_INST template
#include "pf5d_instances.h"

namespace pf5d {
PF5D_FFTP_INSTANCES
}  // namespace pf5d
