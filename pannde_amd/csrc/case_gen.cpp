// case_gen.cpp -- the array work of the synthetic case generator (pannde_amd/cases.py): a
// right-hex grid with cells removed, in the generator's numbering.  Host code only; the
// numpy implementation in cases.py is the reference it must equal array for array.
// Built with -ffp-contract=off so that coordinates are origin + k * d rounded as numpy
// rounds them (one multiply, then one add).
#include <climits>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

// node (i, j, k) of the (nx+1) x (ny+1) x (nz+1) grid is used when a kept cell has it as a
// vertex: cell (i - a, j - b, k - c) for a, b, c in {0, 1}
inline bool used(int nx, int ny, int nz, const uint8_t* keep, int i, int j, int k) {
  for (int a = 0; a < 2; a++) {
    const int ci = i - a;
    if (ci < 0 || ci >= nx) continue;
    for (int b = 0; b < 2; b++) {
      const int cj = j - b;
      if (cj < 0 || cj >= ny) continue;
      const uint8_t* row = keep + ((int64_t)ci * ny + cj) * nz;
      if ((k < nz && row[k]) || (k > 0 && row[k - 1])) return true;
    }
  }
  return false;
}

}  // namespace

extern "C" {

// Counts the used nodes and kept cells of an nx x ny x nz grid (keep: C order, nonzero =
// kept).  Returns 0, or -1 for a bad size.
int cg_mesh_count(int32_t nx, int32_t ny, int32_t nz, const uint8_t* keep, int64_t* n_nodes, int64_t* n_cells) {
  if (nx < 0 || ny < 0 || nz < 0 || !keep || !n_nodes || !n_cells) return -1;
  int64_t nn = 0, nc = 0;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++) nn += used(nx, ny, nz, keep, i, j, k);
  for (int64_t c = 0; c < (int64_t)nx * ny * nz; c++) nc += keep[c] != 0;
  *n_nodes = nn;
  *n_cells = nc;
  return nn < INT32_MAX && nc < INT32_MAX ? 0 : -1;
}

// The mesh: used nodes numbered in C order of (i, j, k) at origin + index * spacing; kept
// cells in C order, their 8 vertices in CGNS order ((0,0,0), (1,0,0), (1,1,0), (0,1,0), then
// the same at k + 1); ijk [n_cells][3] the cell indices; cell_grid (may be null) [nx][ny][nz]
// the cell number of each grid cell, -1 where removed.  Returns 0.
int cg_mesh_fill(int32_t nx, int32_t ny, int32_t nz, const uint8_t* keep, double dx, double dy, double dz,
                 const double* origin, double* x, double* y, double* z, int32_t* cells, int64_t* ijk,
                 int32_t* cell_grid) {
  const int64_t NY = ny + 1, NZ = nz + 1;
  std::vector<int32_t> nid((std::size_t)(nx + 1) * NY * NZ, -1);
  int32_t n = 0;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++)
        if (used(nx, ny, nz, keep, i, j, k)) {
          nid[((std::size_t)i * NY + j) * NZ + k] = n;
          x[n] = origin[0] + (double)i * dx;
          y[n] = origin[1] + (double)j * dy;
          z[n] = origin[2] + (double)k * dz;
          n++;
        }
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  int32_t c = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const int64_t g = ((int64_t)i * ny + j) * nz + k;
        if (!keep[g]) {
          if (cell_grid) cell_grid[g] = -1;
          continue;
        }
        if (cell_grid) cell_grid[g] = c;
        for (int v = 0; v < 8; v++)
          cells[(std::size_t)c * 8 + v] = nid[((std::size_t)(i + off[v][0]) * NY + (j + off[v][1])) * NZ + (k + off[v][2])];
        ijk[(std::size_t)c * 3] = i;
        ijk[(std::size_t)c * 3 + 1] = j;
        ijk[(std::size_t)c * 3 + 2] = k;
        c++;
      }
  return 0;
}

}  // extern "C"
