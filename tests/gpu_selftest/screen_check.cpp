// screen_check.cpp -- test infrastructure: the walk's wall screen
// (dmipy-sim_amd/csrc/screen.h) compiled for the host, evaluated the way the walk
// kernel evaluates it, so tests/test_screen.py can hold it against the oracle's
// float32 root test on many configurations.
//   screen_check_clear: for walker p[i], step s[i], object centre c[i] and radius
//   R[i] (p and p + s in the same frame as the centre image the kernel caches),
//   clear[i] = 1 when the screen state set at p passes the end point p + s.
//   screen_check_grid: builds the walk's neighbour grid (csrc/grid.h) with the
//   expanded cells of the screened path (each cell's box at the level the builder
//   chose for it, as the walk reads it from the encoded list starts) and counts
//   (point, object) pairs, for random points of every cell box, with the point
//   within one step of the object's wall but the object missing from the cell's
//   list; stats receives (cells, cells at level 1, cells at level 2, longest list,
//   lists longer than max_list).
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
    // the state at the start point, then the pass test at the end point
    float rs, b;
    dmw::screen_state(dmw::screen_d2(pi[0], pi[1], pi[2], ci[0], ci[1], ci[2], dim), R[i], m2, c_out, c_in,
                      rs, b);
    clear[i] = dmw::screen_pass(dmw::screen_d2(nx, ny, nz, ci[0], ci[1], ci[2], dim), rs, b) ? 1 : 0;
  }
}

double screen_check_margin(double cell, double r_max, double step) {
  return dmw::screen_margin(cell, r_max, step);
}

int64_t screen_check_grid(int32_t n_obj, const float* centers, const float* radii, float cell,
                          float step, float nudge, int32_t dim, double hyst, int32_t max_list,
                          int32_t per_cell, uint64_t seed, int64_t* pairs, int64_t* stats) {
  dmw_geometry g{};
  g.kind = dim == 2 ? DMW_PACKED_CYLINDERS : DMW_PACKED_SPHERES;
  g.n_obj = n_obj;
  g.cell = cell;
  g.nudge = nudge;
  g.centers = centers;
  g.radii = radii;
  const Grid G = build_grid(g, step, dim, 1u << 30, 16, hyst, max_list);
  const std::vector<int> enc = G.encoded_start();
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double L = cell, h = G.h;
  const int nz = dim == 3 ? G.n : 1;
  int64_t missing = 0, tested = 0;
  for (int i = 0; i < 5; ++i) stats[i] = 0;
  for (int cz = 0; cz < nz; ++cz)
    for (int cy = 0; cy < G.n; ++cy)
      for (int cx = 0; cx < G.n; ++cx) {
        const int ci[3] = {cx, cy, cz};
        const int gc = (cz * G.n + cy) * G.n + cx;
        // the walk's reading of the cell: list bounds and box level from the encoded starts
        const unsigned s0 = (unsigned)enc[gc], lvl = s0 >> dmw::SCR_LEVEL_SHIFT;
        const int beg = (int)(s0 & dmw::SCR_LIST_MASK);
        const int end = (int)((unsigned)enc[gc + 1] & dmw::SCR_LIST_MASK);
        if (beg != G.start[gc] || end != G.start[gc + 1]) return -1;
        const double w = (double)dmw::screen_box_width((float)(hyst * G.h), lvl);
        stats[0] += 1;
        if (lvl == 1) stats[1] += 1;
        if (lvl == 2) stats[2] += 1;
        stats[3] = std::max<int64_t>(stats[3], end - beg);
        if (max_list > 0 && end - beg > max_list) stats[4] += 1;
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
            for (int i = beg; i < end; ++i) found = found || G.items[i] == k;
            if (!found) ++missing;
          }
        }
      }
  *pairs = tested;
  return missing;
}

// The grid's lists as the builder makes them: start (n_cells + 1, without level
// bits), level (n_cells) and items; returns n_cells, or -1 when cap is too small.
int64_t screen_check_lists(int32_t n_obj, const float* centers, const float* radii, float cell,
                           float step, float nudge, int32_t dim, double hyst, int32_t max_list,
                           int32_t* start, int32_t* level, int32_t* items, int64_t cap) {
  dmw_geometry g{};
  g.kind = dim == 2 ? DMW_PACKED_CYLINDERS : DMW_PACKED_SPHERES;
  g.n_obj = n_obj;
  g.cell = cell;
  g.nudge = nudge;
  g.centers = centers;
  g.radii = radii;
  const Grid G = build_grid(g, step, dim, 1u << 30, 16, hyst, max_list);
  const int64_t n_cells = (int64_t)G.start.size() - 1;
  if ((int64_t)G.start.size() > cap || (int64_t)G.items.size() > cap) return -1;
  for (int64_t c = 0; c <= n_cells; ++c) start[c] = G.start[c];
  for (int64_t c = 0; c < n_cells; ++c) level[c] = G.level.empty() ? 0 : G.level[c];
  for (size_t i = 0; i < G.items.size(); ++i) items[i] = G.items[i];
  return n_cells;
}

}
