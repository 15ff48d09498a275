// Walk kernel instantiations for geometry G_PCYL.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(pcyl, G_PCYL, 1, 1)
