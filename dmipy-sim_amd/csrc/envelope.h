// envelope.h -- the operand range of the walk's unscaled exact arithmetic.
#pragma once

#include "../../include/dmipy_walk.h"

namespace dmw {

// The wall routines use correctly rounded square roots and quotients without range
// scaling (rn_arith.h), exact while squared lengths and their float32 differences stay
// above 2^-96: geometric lengths of at least 1e-10 and steps of at least 1e-11 in the
// caller's length unit (an angstrom, in metres).  Returns why a plan is outside.
inline const char* outside_float_envelope(const dmw_geometry& g, float step_len) {
  if (g.kind == DMW_FREE || g.kind == DMW_BOX1D) return nullptr;
  const char* len_msg = "geometric lengths below 1e-10 (in the length unit used) are outside the "
                        "walk's float32 range; express lengths in a smaller unit";
  const char* step_msg = "step lengths below 1e-11 (in the length unit used) are outside the "
                         "walk's float32 range; express lengths in a smaller unit";
  auto small_len = [](double v) { return v > 0.0 && v < 1e-10; };
  auto small_step = [](double v) { return v > 0.0 && v < 1e-11; };
  if (small_len(g.size) || small_len(g.cell) || small_len(g.aux)) return len_msg;
  if (g.kind == DMW_ELLIPSOID)
    for (int k = 0; k < 3; ++k)
      if (g.axes[k] > 1e20f) return len_msg;                       // 1/a^2 with a < 1e-10
  if (g.radii)
    for (int k = 0; k < g.n_obj; ++k)
      if (small_len(g.radii[k])) return len_msg;
  if (small_step(step_len)) return step_msg;
  if (g.myelin)
    for (int c = 0; c < 3; ++c)
      if (small_step(g.myelin->step_len[c])) return step_msg;
  if (g.packed_myelin && g.packed_myelin->inner_radii && g.packed_myelin->step_len)
    for (int k = 0; k < g.n_obj; ++k) {
      if (small_len(g.packed_myelin->inner_radii[k])) return len_msg;
      for (int c = 0; c < 3; ++c)
        if (small_step(g.packed_myelin->step_len[k * 3 + c])) return step_msg;
    }
  return nullptr;
}

}  // namespace dmw
