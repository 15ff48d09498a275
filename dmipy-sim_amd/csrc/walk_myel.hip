// Walk kernel instantiations for geometry G_MYEL (MyelinatedCylinder; its two
// membranes are handled inside the step, so only the reflect wall mode exists).
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(myel, G_MYEL, 0, 0)
