// Walk kernel instantiations for geometry G_SHELLC.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(shellc, G_SHELLC, 0, 1)
