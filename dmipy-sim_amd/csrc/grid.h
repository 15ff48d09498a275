// grid.h -- host-side periodic neighbour grid of a packed substrate.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/dmipy_walk.h"

namespace {

// Periodic neighbour grid of a packed substrate: every cell lists, in ascending
// object order, the objects whose disc/ball inflated by the largest single-step
// reach intersects the cell expanded by `hyst` cell sides on every side (the
// walk's screened path keeps a walker on its cell's list until it leaves that
// expanded box).  The grid is coarsened until the object table (obj_bytes per
// object) and the grid fit in max_bytes, or it is a single cell.
struct Grid {
  int n = 1;
  float h = 0.0f;
  std::vector<int> start, items;
  bool exact_minimage = false;
};

Grid build_grid(const dmw_geometry& g, float step_len, int dim, size_t max_bytes = 60 * 1024,
                size_t obj_bytes = 16, double hyst = 0.0) {
  Grid G;
  const int n_obj = g.n_obj;
  const double L = g.cell;
  double rmean = 0.0, rmax = 0.0;
  for (int k = 0; k < n_obj; ++k) {
    rmean += g.radii[k];
    rmax = std::max(rmax, (double)g.radii[k]);
  }
  rmean /= std::max(1, n_obj);
  const int nmax = dim == 2 ? 64 : 24;
  // cells of 0.3 mean radii: most cells then meet at most one object boundary
  // (measured on a 64-cylinder packing: 0.44 candidates per walker-step against 1.0
  // with 0.8 mean radii, 15 % faster), within the cap and the LDS budget below
  int n = (int)std::lround(L / std::max(0.3 * rmean, 1e-12));
  n = std::max(1, std::min(nmax, n));
  for (;;) {
    const double h = L / n;
    const double tol = 1e-3 * h + 1e-12;
    const double inflate = 1.001 * step_len + 4.0 * g.nudge + tol;
    const double grow = tol + hyst * h;   // the cell box expanded on every side
    const int n_cells = dim == 2 ? n * n : n * n * n;
    std::vector<std::vector<int>> lists(n_cells);
    for (int k = 0; k < n_obj; ++k) {
      double c[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) {
        c[d] = g.centers[k * dim + d];
        c[d] -= L * std::floor(c[d] / L + 0.5);
      }
      const double reach = g.radii[k] + inflate;
      for (int cz = 0; cz < (dim == 3 ? n : 1); ++cz)
        for (int cy = 0; cy < n; ++cy)
          for (int cx = 0; cx < n; ++cx) {
            const int ci[3] = {cx, cy, cz};
            bool touch = false;
            for (int sz = (dim == 3 ? -1 : 0); sz <= (dim == 3 ? 1 : 0) && !touch; ++sz)
              for (int sy = -1; sy <= 1 && !touch; ++sy)
                for (int sx = -1; sx <= 1 && !touch; ++sx) {
                  const int sh[3] = {sx, sy, sz};
                  double d2 = 0.0;
                  for (int d = 0; d < dim; ++d) {
                    const double lo = -0.5 * L + ci[d] * h - grow, hi = lo + h + 2 * grow;
                    const double cc = c[d] + sh[d] * L;
                    const double e = cc < lo ? lo - cc : (cc > hi ? cc - hi : 0.0);
                    d2 += e * e;
                  }
                  touch = d2 <= reach * reach;
                }
            if (touch) lists[(cz * n + cy) * n + cx].push_back(k);
          }
    }
    G.n = n;
    G.h = (float)h;
    G.start.assign(n_cells + 1, 0);
    G.items.clear();
    for (int i = 0; i < n_cells; ++i) {
      G.start[i] = (int)G.items.size();
      G.items.insert(G.items.end(), lists[i].begin(), lists[i].end());
    }
    G.start[n_cells] = (int)G.items.size();
    const size_t lds = (size_t)n_obj * obj_bytes + (G.start.size() + G.items.size()) * 4;
    G.exact_minimage = (rmax + inflate + (h + 2 * grow) * std::sqrt((double)dim)) >= 0.45 * L;
    if (lds <= max_bytes || n == 1) return G;
    n = std::max(1, n * 3 / 4);
  }
}


}  // namespace
