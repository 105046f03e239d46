/* Drives the oracle (oracle/oracle.c, compiled into this program) under
 * AddressSanitizer and UndefinedBehaviorSanitizer: bulk loads of ragged sizes
 * in both key modes and several fanouts (special values mixed in), the BFS
 * query with exact and too-small output buffers, brute force and the
 * canonical sort; results are cross-checked (BFS pairs == brute-force pairs
 * after canonical ordering). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.c"

static uint64_t rs = 88172645463325252ull;
static float frand(void) {
  rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
  return (float)((rs >> 40) & 0xFFFFFF) / 16777216.0f;
}

static int run(int64_t n, int F, int key_mode, int64_t nq) {
  float* b[4]; int32_t* id = malloc(sizeof(int32_t) * (n ? n : 1));
  for (int k = 0; k < 4; k++) b[k] = malloc(sizeof(float) * (n ? n : 1));
  float* b0[4]; int32_t* id0 = malloc(sizeof(int32_t) * (n ? n : 1));
  for (int k = 0; k < 4; k++) b0[k] = malloc(sizeof(float) * (n ? n : 1));
  for (int64_t i = 0; i < n; i++) {
    float cx = frand(), cy = frand(), w = frand() * 0.01f, h = frand() * 0.01f;
    b[0][i] = cx - w; b[1][i] = cy - h; b[2][i] = cx + w; b[3][i] = cy + h;
    if (i % 97 == 5) b[0][i] = NAN;          /* special values */
    if (i % 89 == 7) { b[0][i] = 0.7f; b[2][i] = 0.2f; }  /* inverted */
    if (i % 83 == 3) b[1][i] = -0.0f;
    id[i] = (int32_t)(i * 7919 % (n ? n : 1)) - (int32_t)(n / 2);
  }
  for (int k = 0; k < 4; k++) memcpy(b0[k], b[k], sizeof(float) * (n ? n : 1));
  memcpy(id0, id, sizeof(int32_t) * (n ? n : 1));
  int64_t lv[64], r_sz = 0;
  or_level_plan(n, F, lv, 64, &r_sz);
  float* t[4]; int32_t* pos = malloc(sizeof(int32_t) * (r_sz ? r_sz : 1));
  int32_t* len = malloc(sizeof(int32_t) * (r_sz ? r_sz : 1));
  for (int k = 0; k < 4; k++) t[k] = malloc(sizeof(float) * (r_sz ? r_sz : 1));
  if (or_build_rtree(n, F, key_mode, b[0], b[1], b[2], b[3], id, t[0], t[1], t[2], t[3], pos, len) != r_sz)
    return 1;
  float* q[4];
  for (int k = 0; k < 4; k++) q[k] = malloc(sizeof(float) * (nq ? nq : 1));
  for (int64_t i = 0; i < nq; i++) {
    float cx = frand(), cy = frand(), s = frand() * 0.2f;
    q[0][i] = cx - s; q[1][i] = cy - s; q[2][i] = cx + s; q[3][i] = cy + s;
  }
  int64_t cnt[3], dsz[64];
  or_bfs(r_sz, t[0], t[1], t[2], t[3], pos, len, b[0], b[1], b[2], b[3], id, nq, q[0], q[1], q[2],
         q[3], NULL, NULL, 0, cnt, dsz, 64);
  const int64_t H = cnt[2];
  int32_t* f1 = malloc(sizeof(int32_t) * (H ? H : 1));
  int32_t* t1 = malloc(sizeof(int32_t) * (H ? H : 1));
  if (H > 1 && or_bfs(r_sz, t[0], t[1], t[2], t[3], pos, len, b[0], b[1], b[2], b[3], id, nq, q[0],
                      q[1], q[2], q[3], f1, t1, H - 1, cnt, NULL, 0) != -1)
    return 2; /* too small a buffer reports -1 */
  if (or_bfs(r_sz, t[0], t[1], t[2], t[3], pos, len, b[0], b[1], b[2], b[3], id, nq, q[0], q[1],
             q[2], q[3], f1, t1, H, cnt, dsz, 64) != H)
    return 3;
  /* brute force over the unsorted input: same pairs once NaN boxes are left
   * out (the tree reaches them only through a parent's MBR) */
  const int64_t Hb = or_brute_count(0, nq, q[0], q[1], q[2], q[3], n, b0[0], b0[1], b0[2], b0[3], 1);
  int32_t* f2 = malloc(sizeof(int32_t) * (Hb ? Hb : 1));
  int32_t* t2 = malloc(sizeof(int32_t) * (Hb ? Hb : 1));
  if (or_brute_force(nq, q[0], q[1], q[2], q[3], n, b0[0], b0[1], b0[2], b0[3], id0, f2, t2, Hb) != Hb)
    return 4;
  or_canonical_sort(H, f1, t1);
  or_canonical_sort(Hb, f2, t2);
  if (Hb < H) return 5;
  for (int k = 0; k < 4; k++) { free(b[k]); free(b0[k]); free(t[k]); free(q[k]); }
  free(id); free(id0); free(pos); free(len); free(f1); free(t1); free(f2); free(t2);
  return 0;
}

int main(void) {
  const int64_t ns[] = {0, 1, 2, 15, 16, 17, 255, 4097, 20000};
  const int Fs[] = {2, 3, 16, 64, 100};
  for (size_t i = 0; i < sizeof ns / sizeof *ns; i++)
    for (size_t j = 0; j < sizeof Fs / sizeof *Fs; j++)
      for (int km = 0; km < 2; km++) {
        const int rc = run(ns[i], Fs[j], km, ns[i] ? 300 : 0);
        if (rc) { printf("FAIL n=%lld F=%d key=%d rc=%d\n", (long long)ns[i], Fs[j], km, rc); return 1; }
      }
  printf("oracle ok\n");
  return 0;
}
