// Walk kernel instantiations for geometry G_SLAB.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(slab, G_SLAB, 1, 1)
