// Walk kernel instantiations for geometry G_SPH.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(sph, G_SPH, 1, 1)
