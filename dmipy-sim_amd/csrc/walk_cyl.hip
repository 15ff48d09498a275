// Walk kernel instantiations for geometry G_CYL.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(cyl, G_CYL, 1, 1)
