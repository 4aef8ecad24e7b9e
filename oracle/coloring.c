/* Oracle: natural-order greedy distance-2 coloring in plain C.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates
 * tatva/sparse/_coloring.py:27-48 (rows of the boolean product A @ A) and
 * :136-153 (for i = 0..n-1 in order: used = colors of the A^2 neighbours of i
 * that are already colored; color[i] = the smallest non-negative integer not
 * in used).  Deliberately the plain algorithm, with no grouping of identical
 * rows, so it is an independent check of the product's accelerated version.
 */
#include <stdint.h>
#include <stdlib.h>

int oracle_distance2_greedy(int64_t n, const int64_t *indptr, const int32_t *indices,
                            int32_t *colors) {
  int64_t cap = 64;
  int64_t *stamp = NULL; /* stamp[c] == i+1 <=> color c used by a neighbour of i */
  int64_t *seen = NULL;  /* seen[j] == i+1 <=> j already visited for row i */
  int max_color = -1;
  stamp = (int64_t *)calloc((size_t)cap, sizeof(int64_t));
  seen = (int64_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t));
  if (!stamp || !seen) {
    free(stamp);
    free(seen);
    return -1;
  }
  for (int64_t i = 0; i < n; ++i) colors[i] = -1;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t mark = i + 1;
    for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) {
      const int64_t k = indices[p];
      for (int64_t q = indptr[k]; q < indptr[k + 1]; ++q) {
        const int64_t j = indices[q];
        if (seen[j] == mark) continue;
        seen[j] = mark;
        const int32_t c = colors[j];
        if (c < 0) continue;
        if (c >= cap) {
          int64_t ncap = cap;
          while (ncap <= c) ncap *= 2;
          int64_t *ns = (int64_t *)realloc(stamp, (size_t)ncap * sizeof(int64_t));
          if (!ns) {
            free(stamp);
            free(seen);
            return -1;
          }
          for (int64_t t = cap; t < ncap; ++t) ns[t] = 0;
          stamp = ns;
          cap = ncap;
        }
        stamp[c] = mark;
      }
    }
    int32_t c = 0;
    while (c < cap && stamp[c] == mark) ++c;
    colors[i] = c;
    if (c > max_color) max_color = c;
  }
  free(stamp);
  free(seen);
  return max_color + 1;
}
