// Walk kernel instantiations for geometry G_ELL.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(ell, G_ELL, 1, 1)
