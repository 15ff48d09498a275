// Walk kernel instantiations for geometry G_PSPH.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(psph, G_PSPH, 1, 1)
