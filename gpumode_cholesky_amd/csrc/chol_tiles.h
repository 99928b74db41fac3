// chol_tiles.h — tile orders of the trailing updates, as plain functions
// (no HIP types) so that the host can check them too.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define CHOL_HD __host__ __device__ __forceinline__
#else
#define CHOL_HD inline
#endif

namespace chol {

// Lower-triangular tile set {(ti, tj): tj < TJ, tj <= ti < TI} in groups of
// G consecutive tile columns; inside a group, row by row (the G columns of
// row ti, those left of the diagonal only, before row ti + 1).  A run of
// consecutive indices, the tiles that one XCD's CUs hold at once, then
// spans few row blocks and G column blocks (32 tiles: 8 rows x 4 columns
// for G = 4) where a column-by-column order spans 32 row blocks and one
// column block, so each operand block a run loads serves several tiles
// from the XCD's L2.  Tiles counted: T = sum over tile columns of (TI - tj).
// Maps idx in [0, T) to (ti, tj); a bijection for TJ <= TI.
CHOL_HD void decode_tile_grouped(int idx, int TI, int TJ, int G, int& ti, int& tj) {
  int base = 0;  // first index of group g
  for (int g0 = 0; g0 < TJ; g0 += G) {
    const int w = (TJ - g0) < G ? (TJ - g0) : G;        // columns in this group
    const int tri = w * (w + 1) / 2;                     // rows g0 .. g0+w-1 (1 .. w tiles)
    const int count = tri + w * (TI - g0 - w);           // then w tiles per row
    if (idx < base + count) {
      const int r = idx - base;
      if (r < tri) {
        int t = 0;  // row t of the triangle holds t + 1 tiles, starting at t (t + 1) / 2
        while ((t + 1) * (t + 2) / 2 <= r) ++t;
        ti = g0 + t;
        tj = g0 + (r - t * (t + 1) / 2);
      } else {
        const int q = r - tri;
        ti = g0 + w + q / w;
        tj = g0 + q % w;
      }
      return;
    }
    base += count;
  }
  ti = tj = -1;  // idx out of range
}

}  // namespace chol
