// The cell order of one sweep task: the cells of a B^3 block grouped by diagonal plane a + b + c = p, in
// sweep-local coordinates (a along x, b along y, c along z, each counted from the octant's upstream corner). Shared by
// the sweep kernels, which visit a block's cells in this order, and by the first-collision sums, which store the
// deferred source of each octant in it.
#pragma once

#include <hip/hip_runtime.h>

namespace citrus {
namespace {

template <int B>
struct PlaneTable {
  static constexpr int kPlanes = 3 * B - 2;
  int offset[3 * B - 1];
  int cell[B * B * B];  // a | b << 8 | c << 16
  int slot[B * B * B];  // inverse: the position of (a, b, c) at index a + B (b + B c)
  constexpr PlaneTable() : offset(), cell(), slot() {
    int n = 0;
    for (int p = 0; p < kPlanes; ++p) {
      offset[p] = n;
      for (int a = 0; a < B; ++a)
        for (int b = 0; b < B; ++b) {
          int c = p - a - b;
          if (c >= 0 && c < B) {
            slot[a + B * (b + B * c)] = n;
            cell[n++] = a | (b << 8) | (c << 16);
          }
        }
    }
    offset[kPlanes] = n;
  }
  constexpr int max_plane() const {
    int m = 0;
    for (int p = 0; p < kPlanes; ++p) m = (offset[p + 1] - offset[p]) > m ? (offset[p + 1] - offset[p]) : m;
    return m;
  }
};

__constant__ PlaneTable<4> kPlanes4 = PlaneTable<4>();

template <int B>
__device__ __forceinline__ const PlaneTable<B>& planes();
template <>
__device__ __forceinline__ const PlaneTable<4>& planes<4>() { return kPlanes4; }

// Position in octant o's task order of the cell (lx, ly, lz) of a 4^3 block, in grid-local coordinates (x along the
// columns, y along the rows, z along the slices): the octant's sign bits (4: -x, 2: -y, 1: -z) reflect them into
// sweep-local ones.
__device__ __forceinline__ int block_slot4(int o, int lx, int ly, int lz) {
  const int a = (o & 4) ? 3 - lx : lx, b = (o & 2) ? 3 - ly : ly, c = (o & 1) ? 3 - lz : lz;
  return kPlanes4.slot[a + 4 * (b + 4 * c)];
}

}  // namespace
}  // namespace citrus
