/*
 * Oracle (test infrastructure only): the reference CPU 2-opt scan restated in plain C.
 *
 * Follows src/algorithms/strategies/local_search/two_opt.py of the reference for the
 * cyclic TSP topology (closed = 0, cyclic = 1, fixed_start = 1, two_opt.py:97-114):
 *   - exhaustive pair order (i asc, j asc), j in [i+2, n-1], (0, n-1) excluded
 *     (two_opt.py:158-201);
 *   - candidate pair order (i asc, r asc), j = pos[N[t_i, r]], kept iff j > i+1,
 *     j < n and not (i == 0 and j == n-1), duplicates kept (two_opt.py:347-367);
 *   - delta = D[a,c] + D[b,d] - D[a,b] - D[c,d] in int64 (two_opt.py:237-245);
 *   - "best": first minimum in scan order, applied iff < 0 (two_opt.py:386-440);
 *   - "first": first pair with delta < 0 (two_opt.py:442-489);
 *   - "multi_move": every improving pair of the pass, stably sorted by delta (ties in scan
 *     order), applied greedily unless its window [i, j] meets an applied one (two_opt.py:491-563);
 *   - refine: at most max_iter passes, stop on a pass without a move, n < 4 is a
 *     no-op (two_opt.py:565-595).
 * It exists so that parity tests at n = 1000 and the numpy restatement can be checked
 * in seconds; tests cross-check it against oracle/two_opt.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static void reverse_segment(int32_t *t, int lo, int hi) {
  while (lo < hi) {
    int32_t tmp = t[lo];
    t[lo] = t[hi];
    t[hi] = tmp;
    ++lo;
    --hi;
  }
}

typedef struct {
  int64_t delta;
  int64_t scan;
  int i, j;
} improving_t;

static int cmp_improving(const void *pa, const void *pb) {
  const improving_t *a = (const improving_t *)pa, *b = (const improving_t *)pb;
  if (a->delta != b->delta) return a->delta < b->delta ? -1 : 1;
  return (a->scan > b->scan) - (a->scan < b->scan);
}

/* Multi-move pass. Returns the number of windows applied. */
static int multi_pass(int32_t *t, int n, const int32_t *D, long ld, const int32_t *cand, int k, int32_t *pos,
                      int64_t *deltas) {
  size_t cap = 1024, cnt = 0;
  improving_t *list = (improving_t *)malloc(cap * sizeof(improving_t));
  if (!list) return -1;
  if (cand && k > 0)
    for (int x = 0; x < n; ++x) pos[t[x]] = x;
  for (int i = 0; i <= n - 3; ++i) {
    const int a = t[i], b = t[i + 1];
    const int64_t dab = D[(long)a * ld + b];
    const int limit = (i == 0) ? n - 1 : n;
    const int width = (cand && k > 0) ? k : limit - (i + 2);
    for (int r = 0; r < width; ++r) {
      int j;
      if (cand && k > 0) {
        const int nb = cand[(long)a * k + r];
        if (nb < 0 || nb >= n) continue;
        j = pos[nb];
        if (j <= i + 1 || j >= limit) continue;
      } else {
        j = i + 2 + r;
      }
      const int c = t[j], d = t[(j + 1 == n) ? 0 : j + 1];
      const int64_t delta = (int64_t)D[(long)a * ld + c] + D[(long)b * ld + d] - dab - D[(long)c * ld + d];
      ++*deltas;
      if (delta < 0) {
        if (cnt == cap) {
          cap *= 2;
          improving_t *grown = (improving_t *)realloc(list, cap * sizeof(improving_t));
          if (!grown) {
            free(list);
            return -1;
          }
          list = grown;
        }
        list[cnt].delta = delta;
        list[cnt].scan = (int64_t)cnt;
        list[cnt].i = i;
        list[cnt].j = j;
        ++cnt;
      }
    }
  }
  qsort(list, cnt, sizeof(improving_t), cmp_improving);
  unsigned char *occupied = (unsigned char *)calloc((size_t)n, 1);
  int applied = 0;
  for (size_t m = 0; occupied && m < cnt; ++m) {
    const int i = list[m].i, j = list[m].j;
    int clash = 0;
    for (int x = i; x <= j && !clash; ++x) clash = occupied[x];
    if (clash) continue;
    reverse_segment(t, i + 1, j);
    memset(occupied + i, 1, (size_t)(j - i + 1));
    ++applied;
  }
  free(occupied);
  free(list);
  return applied;
}

/* One pass. Returns 1 when a move was applied. */
static int one_pass(int32_t *t, int n, const int32_t *D, long ld, const int32_t *cand, int k,
                    int mode, int32_t *pos, int64_t *deltas) {
  int64_t best = 0;
  int bi = -1, bj = -1;
  if (cand && k > 0) {
    for (int x = 0; x < n; ++x) pos[t[x]] = x;
    for (int i = 0; i <= n - 3; ++i) {
      const int a = t[i], b = t[i + 1];
      const int64_t dab = D[(long)a * ld + b];
      const int limit = (i == 0) ? n - 1 : n;
      for (int r = 0; r < k; ++r) {
        const int nb = cand[(long)a * k + r];
        if (nb < 0 || nb >= n) continue;
        const int j = pos[nb];
        if (j <= i + 1 || j >= limit) continue;
        const int c = t[j], d = t[(j + 1 == n) ? 0 : j + 1];
        const int64_t delta = (int64_t)D[(long)a * ld + c] + D[(long)b * ld + d] - dab - D[(long)c * ld + d];
        ++*deltas;
        if (mode == 1) {
          if (delta < 0) {
            reverse_segment(t, i + 1, j);
            return 1;
          }
        } else if (delta < best) {
          best = delta;
          bi = i;
          bj = j;
        }
      }
    }
  } else {
    for (int i = 0; i <= n - 3; ++i) {
      const int a = t[i], b = t[i + 1];
      const int64_t dab = D[(long)a * ld + b];
      const int limit = (i == 0) ? n - 1 : n;
      const int32_t *rowa = D + (long)a * ld;
      const int32_t *rowb = D + (long)b * ld;
      for (int j = i + 2; j < limit; ++j) {
        const int c = t[j], d = t[(j + 1 == n) ? 0 : j + 1];
        const int64_t delta = (int64_t)rowa[c] + rowb[d] - dab - D[(long)c * ld + d];
        if (mode == 1) {
          if (delta < 0) {
            *deltas += (int64_t)(j - (i + 2) + 1);
            reverse_segment(t, i + 1, j);
            return 1;
          }
        } else if (delta < best) {
          best = delta;
          bi = i;
          bj = j;
        }
      }
      if (limit > i + 2) *deltas += (int64_t)(limit - (i + 2));
    }
  }
  if (mode != 1 && best < 0) {
    reverse_segment(t, bi + 1, bj);
    return 1;
  }
  return 0;
}

/*
 * tours: (P, n) int32 row-major, refined in place.
 * mode: 0 = best, 1 = first, 2 = multi_move.  max_iter < 0 means "until no improving pair".
 * stats[0] = deltas evaluated, stats[1] = moves applied, stats[2] = passes run.
 * Returns 0, or -1 on bad arguments.
 */
int oracle_two_opt(int32_t *tours, int P, int n, const int32_t *D, int ld, const int32_t *cand, int k,
                   int mode, long max_iter, int64_t *stats) {
  if (P < 0 || n < 0 || mode < 0 || mode > 2) return -1;
  int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!pos) return -1;
  int64_t deltas = 0, moves = 0, passes = 0;
  for (int r = 0; r < P; ++r) {
    int32_t *t = tours + (long)r * n;
    if (n < 4 || max_iter == 0) continue;
    for (long it = 0; max_iter < 0 || it < max_iter; ++it) {
      ++passes;
      int64_t d = 0;
      const int moved = mode == 2 ? multi_pass(t, n, D, ld, cand, k, pos, &d) : one_pass(t, n, D, ld, cand, k, mode, pos, &d);
      deltas += d;
      if (moved < 0) {
        free(pos);
        return -1;
      }
      if (!moved) break;
      moves += moved;
    }
  }
  free(pos);
  if (stats) {
    stats[0] = deltas;
    stats[1] = moves;
    stats[2] = passes;
  }
  return 0;
}
