// Host check of decode_tile_grouped (gpumode_cholesky_amd/csrc/chol_tiles.h):
// for each (TI, TJ, G) given on the command line triples, every index of
// [0, T) maps to a distinct valid lower tile and every lower tile is hit.
// Prints "ok" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "chol_tiles.h"

int main(int argc, char** argv) {
  for (int a = 1; a + 2 < argc; a += 3) {
    const int TI = atoi(argv[a]), TJ = atoi(argv[a + 1]), G = atoi(argv[a + 2]);
    long T = 0;
    for (int j = 0; j < TJ; ++j) T += TI - j;
    std::vector<char> seen((size_t)TI * TJ, 0);
    for (long idx = 0; idx < T; ++idx) {
      int ti, tj;
      chol::decode_tile_grouped((int)idx, TI, TJ, G, ti, tj);
      if (ti < 0 || tj < 0 || tj >= TJ || ti < tj || ti >= TI) {
        printf("bad tile (%d,%d) for idx %ld TI=%d TJ=%d G=%d\n", ti, tj, idx, TI, TJ, G);
        return 1;
      }
      char& s = seen[(size_t)ti * TJ + tj];
      if (s) {
        printf("tile (%d,%d) twice, TI=%d TJ=%d G=%d\n", ti, tj, TI, TJ, G);
        return 1;
      }
      s = 1;
    }
    int ti, tj;
    chol::decode_tile_grouped((int)T, TI, TJ, G, ti, tj);
    if (ti != -1) {
      printf("idx T maps to a tile, TI=%d TJ=%d G=%d\n", TI, TJ, G);
      return 1;
    }
  }
  printf("ok\n");
  return 0;
}
