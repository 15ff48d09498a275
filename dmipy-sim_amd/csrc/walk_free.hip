// Walk kernel instantiations for geometry G_FREE.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(free, G_FREE, 0, 0)
