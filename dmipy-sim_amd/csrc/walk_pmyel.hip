// Walk kernel instantiations for geometry G_PMYEL (PackedMyelinatedCylinders; its
// membranes and exterior clamp are handled inside the step: reflect wall mode only).
#include "walk_kernel.h"
#include "walk_launch.h"

DMW_DEFINE_LAUNCHER(pmyel, G_PMYEL, 0, 0)
