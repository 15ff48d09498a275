// Walk kernel instantiations for geometry G_SHELL.
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(shell, G_SHELL, 0, 1)
