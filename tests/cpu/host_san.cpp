// Drives the library's host-side code paths that need no GPU under
// AddressSanitizer and UndefinedBehaviorSanitizer (the library's sources are
// compiled into this program with the sanitizers on the host side only): the
// CSV reader on good, short and malformed files, the tree file reader on
// corrupt files, the level plan, options, and the argument checks of the
// build, query, canonical sort and serving entry points.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "curtree/curtree_c.h"

static int fails = 0;
static void expect(bool ok, const char* what) {
  if (!ok) {
    printf("FAIL %s: %s\n", what, curtree_last_error());
    fails++;
  }
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // CSV: 3 good lines, then a request for more lines than the file holds
  const std::string csv = dir + "/edges.csv";
  FILE* f = fopen(csv.c_str(), "w");
  fprintf(f, "1.5,2.5,0.5,3.5,7,0,0\n-1,-2,-3,-4,8,1,1\n10,10,10,10,9,2,2\n");
  fclose(f);
  std::vector<float> x0(8), y0(8), x1(8), y1(8);
  std::vector<int> id(8);
  expect(curtree_read_mbr_csv(csv.c_str(), 3, 0.5f, x0.data(), y0.data(), x1.data(), y1.data(),
                              id.data()) == CURTREE_OK, "csv good");
  expect(x0[0] == 0.0f && x1[0] == 2.0f && y0[0] == 2.0f && y1[0] == 4.0f && id[2] == 9,
         "csv values");
  expect(curtree_read_mbr_csv(csv.c_str(), 5, 0.f, x0.data(), y0.data(), x1.data(), y1.data(),
                              id.data()) == CURTREE_EINVAL, "csv short");
  const std::string bad = dir + "/bad.csv";
  f = fopen(bad.c_str(), "w");
  fprintf(f, "1,2,3\n");
  fclose(f);
  expect(curtree_read_mbr_csv(bad.c_str(), 1, 0.f, x0.data(), y0.data(), x1.data(), y1.data(),
                              id.data()) == CURTREE_EINVAL, "csv malformed");
  expect(curtree_read_mbr_csv((dir + "/missing.csv").c_str(), 1, 0.f, x0.data(), y0.data(),
                              x1.data(), y1.data(), id.data()) == CURTREE_EINVAL, "csv missing");
  // tree files: junk, and a header whose sizes do not match the file
  const std::string junk = dir + "/junk.ctree";
  f = fopen(junk.c_str(), "wb");
  fprintf(f, "definitely not a tree file, but long enough to hold a header");
  fclose(f);
  curtree_rtree_f32 rt{};
  curtree_bbox_f32 bx{};
  expect(curtree_tree_load(junk.c_str(), &rt, &bx) == CURTREE_EINVAL, "tree junk");
  const std::string hdr = dir + "/hdr.ctree";
  f = fopen(hdr.c_str(), "wb");
  const char magic[8] = {'C', 'T', 'R', 'E', 'E', '0', '0', '1'};
  int64_t nb = 1000, rs = 0;
  int32_t fan = 16, pad = 0;
  curtree_tree_size(nb, fan, &rs, nullptr);
  fwrite(magic, 1, 8, f);
  fwrite(&nb, 8, 1, f);
  fwrite(&rs, 8, 1, f);
  fwrite(&fan, 4, 1, f);
  fwrite(&pad, 4, 1, f);
  fclose(f);
  expect(curtree_tree_load(hdr.c_str(), &rt, &bx) == CURTREE_EINVAL, "tree short");
  // level plan
  int64_t r = 0;
  int nl = 0;
  expect(curtree_tree_size(1000000, 16, &r, &nl) == CURTREE_OK && r == 66669 && nl == 5, "plan");
  expect(curtree_tree_size(1, 16, &r, &nl) == CURTREE_OK && r == 1 && nl == 1, "plan 1");
  expect(curtree_tree_size(-1, 16, &r, &nl) == CURTREE_EINVAL, "plan negative");
  // options
  expect(curtree_set_option("containment", 0) == CURTREE_OK, "option");
  expect(curtree_set_option("containment", 5) == CURTREE_EINVAL, "option value");
  expect(curtree_set_option("no such knob", 1) == CURTREE_EINVAL, "option name");
  expect(curtree_set_option(nullptr, 1) == CURTREE_EINVAL, "option null");
  curtree_set_option("containment", -1);
  // argument checks that return before any device work
  curtree_bbox_f32 b{10, nullptr, nullptr, nullptr, nullptr, nullptr};
  curtree_rtree_f32 t{};
  t.fanout = 1;
  expect(curtree_build_f32(&b, &t, CURTREE_KEY_REF32, nullptr) == CURTREE_EINVAL, "build fanout");
  t.fanout = 16;
  expect(curtree_build_f32(&b, &t, 7, nullptr) == CURTREE_EINVAL, "build key mode");
  expect(curtree_build_f32(nullptr, &t, 0, nullptr) == CURTREE_EINVAL, "build null");
  expect(curtree_build_f32(&b, &t, 0, nullptr) == CURTREE_EINVAL, "build null arrays");
  curtree_idpair64 out{};
  expect(curtree_query_f32(&t, &b, &b, &out, 9, nullptr) == CURTREE_EINVAL, "query algo");
  expect(curtree_query_f32(&t, &b, &b, nullptr, 0, nullptr) == CURTREE_EINVAL, "query null");
  expect(curtree_pairs_canonical_sort(nullptr, nullptr, -1, nullptr) == CURTREE_EINVAL,
         "sort negative");
  expect(curtree_pairs_canonical_sort(nullptr, nullptr, 5, nullptr) == CURTREE_EINVAL,
         "sort null");
  curtree_bbox_f32 many{(3 << 20) / 2 + 1, nullptr, nullptr, nullptr, nullptr, nullptr};
  int64_t n = 0;
  expect(curtree_query_into_noalloc_f32(&t, &b, &many, nullptr, nullptr, 0, &n, nullptr) ==
             CURTREE_EINVAL, "noalloc windows");
  curtree_server_opts so{0, 1 << 20, 1 << 10, 1};
  curtree_server* sv = nullptr;
  expect(curtree_server_create(&so, &sv) == CURTREE_EINVAL, "server slots");
  expect(strlen(curtree_version()) > 0, "version");
  if (fails) return 1;
  printf("host ok\n");
  return 0;
}
