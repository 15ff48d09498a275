// Walk kernel instantiations for geometry G_BOX1D.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(box1d, G_BOX1D, 1, 0)
