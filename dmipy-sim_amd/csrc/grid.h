// grid.h -- host-side periodic neighbour grid of a packed substrate.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/dmipy_walk.h"
#include "screen.h"

namespace {

// Periodic neighbour grid of a packed substrate: every cell lists, in ascending
// object order, the objects whose disc/ball inflated by the largest single-step
// reach intersects the cell expanded by `hyst` cell sides on every side (the
// walk's screened path keeps a walker on its cell's list until it leaves that
// expanded box).  With `max_list` > 0, a cell whose list would exceed max_list
// objects takes a smaller box, level l expanding it by hyst * (1 - l/2) sides
// (l = 1, 2), the first level whose list fits, else level 2 (the cell itself):
// the screen caches at most that many objects per walker, and a walker on a
// longer list runs the exact test on every sub-step.  The grid is coarsened
// until the object table (obj_bytes per object) and the grid fit in max_bytes,
// or it is a single cell.
struct Grid {
  int n = 1;
  float h = 0.0f;
  std::vector<int> start, items;
  std::vector<unsigned char> level;   // per cell: box level (0 unless max_list > 0)
  bool exact_minimage = false;

  // The list starts with each cell's box level in the bits above
  // dmw::SCR_LIST_MASK, as the screened walk reads them.
  std::vector<int> encoded_start() const {
    std::vector<int> s(start);
    for (size_t c = 0; c < level.size(); ++c)
      s[c] = (int)((unsigned)s[c] | ((unsigned)level[c] << dmw::SCR_LEVEL_SHIFT));
    return s;
  }
};

Grid build_grid(const dmw_geometry& g, float step_len, int dim, size_t max_bytes = 60 * 1024,
                size_t obj_bytes = 16, double hyst = 0.0, int max_list = 0) {
  Grid G;
  const int n_obj = g.n_obj;
  const double L = g.cell;
  double rmean = 0.0, rmax = 0.0;
  for (int k = 0; k < n_obj; ++k) {
    rmean += g.radii[k];
    rmax = std::max(rmax, (double)g.radii[k]);
  }
  rmean /= std::max(1, n_obj);
  std::vector<double> cen((size_t)n_obj * 3, 0.0);
  for (int k = 0; k < n_obj; ++k)
    for (int d = 0; d < dim; ++d) {
      const double c = g.centers[k * dim + d];
      cen[k * 3 + d] = c - L * std::floor(c / L + 0.5);
    }
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
    // does object k's inflated disc/ball meet cell ci expanded by `grow` on every side?
    auto touches = [&](int k, const int* ci, double grow) {
      const double reach = g.radii[k] + inflate;
      for (int sz = (dim == 3 ? -1 : 0); sz <= (dim == 3 ? 1 : 0); ++sz)
        for (int sy = -1; sy <= 1; ++sy)
          for (int sx = -1; sx <= 1; ++sx) {
            const int sh[3] = {sx, sy, sz};
            double d2 = 0.0;
            for (int d = 0; d < dim; ++d) {
              const double lo = -0.5 * L + ci[d] * h - grow, hi = lo + h + 2 * grow;
              const double cc = cen[k * 3 + d] + sh[d] * L;
              const double e = cc < lo ? lo - cc : (cc > hi ? cc - hi : 0.0);
              d2 += e * e;
            }
            if (d2 <= reach * reach) return true;
          }
      return false;
    };
    const int n_cells = dim == 2 ? n * n : n * n * n;
    std::vector<std::vector<int>> lists(n_cells);
    G.level.assign(max_list > 0 ? n_cells : 0, 0);
    for (int c = 0; c < n_cells; ++c) {
      const int ci[3] = {c % n, (c / n) % n, dim == 3 ? c / (n * n) : 0};
      for (int lvl = 0; lvl <= (max_list > 0 ? 2 : 0); ++lvl) {
        const double grow = tol + hyst * (1.0 - 0.5 * lvl) * h;
        lists[c].clear();
        for (int k = 0; k < n_obj; ++k)
          if (touches(k, ci, grow)) lists[c].push_back(k);
        if (max_list > 0) G.level[c] = (unsigned char)lvl;
        if (max_list <= 0 || hyst <= 0.0 || (int)lists[c].size() <= max_list) break;
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
    G.exact_minimage = (rmax + inflate + (h + 2 * (tol + hyst * h)) * std::sqrt((double)dim)) >= 0.45 * L;
    if (lds <= max_bytes || n == 1) return G;
    n = std::max(1, n * 3 / 4);
  }
}


}  // namespace
