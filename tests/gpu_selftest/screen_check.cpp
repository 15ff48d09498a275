// screen_check.cpp -- test infrastructure: the walk's wall screen
// (dmipy-sim_amd/csrc/screen.h) compiled for the host, evaluated the way the walk
// kernel evaluates it, so tests/test_screen.py can hold it against the oracle's
// float32 root test on many configurations.
//   screen_check_clear: for walker p[i], step s[i], object centre c[i] and radius
//   R[i] (p and p + s in the same frame as the centre image the kernel caches),
//   clear[i] = 1 when the screen clears the sub-step for that object.
//   screen_check_grid: builds the walk's neighbour grid (csrc/grid.h) with the
//   expanded cells of the screened path and counts (point, object) pairs, for
//   random points of every expanded cell box, with the point within one step of
//   the object's wall but the object missing from the cell's list.
#include <math.h>
#include <stdint.h>

#include <random>

#include "../../dmipy-sim_amd/csrc/grid.h"
#include "../../dmipy-sim_amd/csrc/screen.h"

extern "C" {

void screen_check_clear(int64_t n, const float* p, const float* s, const float* c, const float* R,
                        float margin, float step_max, int32_t dim, int32_t* clear) {
  const float m2 = 2.0f * margin;
  const float c_in = margin * margin;
  const float c_out = c_in + 0.25f * step_max * step_max * (1.0f + 0x1p-20f);
  for (int64_t i = 0; i < n; ++i) {
    const float* pi = p + 3 * i;
    const float* si = s + 3 * i;
    const float* ci = c + 3 * i;
    const float nx = pi[0] + si[0], ny = pi[1] + si[1], nz = pi[2] + si[2];
    const float f0 = dmw::screen_f(pi[0], pi[1], pi[2], ci[0], ci[1], ci[2], R[i], dim);
    const float f1 = dmw::screen_f(nx, ny, nz, ci[0], ci[1], ci[2], R[i], dim);
    clear[i] = dmw::screen_clear(f0, f1, R[i], m2, c_out, c_in) ? 1 : 0;
  }
}

double screen_check_margin(double cell, double r_max, double step) {
  return dmw::screen_margin(cell, r_max, step);
}

int64_t screen_check_grid(int32_t n_obj, const float* centers, const float* radii, float cell,
                          float step, float nudge, int32_t dim, double hyst, int32_t per_cell,
                          uint64_t seed, int64_t* pairs) {
  dmw_geometry g{};
  g.kind = dim == 2 ? DMW_PACKED_CYLINDERS : DMW_PACKED_SPHERES;
  g.n_obj = n_obj;
  g.cell = cell;
  g.nudge = nudge;
  g.centers = centers;
  g.radii = radii;
  const Grid G = build_grid(g, step, dim, 1u << 30, 16, hyst);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double L = cell, h = G.h, w = hyst * G.h;
  const int nz = dim == 3 ? G.n : 1;
  int64_t missing = 0, tested = 0;
  for (int cz = 0; cz < nz; ++cz)
    for (int cy = 0; cy < G.n; ++cy)
      for (int cx = 0; cx < G.n; ++cx) {
        const int ci[3] = {cx, cy, cz};
        const int gc = (cz * G.n + cy) * G.n + cx;
        for (int s = 0; s < per_cell; ++s) {
          double pt[3] = {0, 0, 0};
          for (int d = 0; d < dim; ++d) pt[d] = -0.5 * L + ci[d] * h - w + U(rng) * (h + 2 * w);
          for (int k = 0; k < n_obj; ++k) {
            double r2 = 0.0;
            for (int d = 0; d < dim; ++d) {
              double q = pt[d] - centers[k * dim + d];
              q -= L * floor(q / L + 0.5);
              r2 += q * q;
            }
            if (fabs(sqrt(r2) - radii[k]) > step) continue;
            ++tested;
            bool found = false;
            for (int i = G.start[gc]; i < G.start[gc + 1]; ++i) found = found || G.items[i] == k;
            if (!found) ++missing;
          }
        }
      }
  *pairs = tested;
  return missing;
}

}
