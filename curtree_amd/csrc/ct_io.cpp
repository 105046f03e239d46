// ct_io.cpp -- box ingest and tree persistence.
//
//   curtree_read_mbr_csv: the reference's CSV reader readMBR
//     (src/utility.cpp:5-40): one box per line "x1,y1,x2,y2,id,pid,nume",
//     box = [min(x1,x2)-r, max(x1,x2)+r] x [min(y1,y2)-r, max(y1,y2)+r].
//     Open and parse failures return an error instead of exit()/throw.
//   curtree_tree_save / curtree_tree_load: a binary file holding the sorted
//     boxes and the node arrays of one tree, streamed through a pinned buffer
//     (the reference never persists a tree).
#include <stdio.h>
#include <string.h>

#include "ct_internal.hpp"

using namespace ct;

namespace {

constexpr char kMagic[8] = {'C', 'T', 'R', 'E', 'E', '0', '0', '1'};
constexpr size_t kChunk = 64u << 20;

struct FileHeader {
  char magic[8];
  int64_t n_boxes;
  int64_t r_sz;
  int32_t fanout;
  int32_t reserved;
};

struct Pinned {
  void* p = nullptr;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

int write_dev(FILE* f, const void* dev, size_t bytes, Pinned& buf) {
  const char* src = (const char*)dev;
  for (size_t off = 0; off < bytes; off += kChunk) {
    size_t b = std::min(kChunk, bytes - off);
    CT_HIP(hipMemcpy(buf.p, src + off, b, hipMemcpyDeviceToHost));
    if (fwrite(buf.p, 1, b, f) != b) return set_error(CURTREE_EINVAL, "tree_save: write failed");
  }
  return CURTREE_OK;
}

int read_dev(FILE* f, void* dev, size_t bytes, Pinned& buf) {
  char* dst = (char*)dev;
  for (size_t off = 0; off < bytes; off += kChunk) {
    size_t b = std::min(kChunk, bytes - off);
    if (fread(buf.p, 1, b, f) != b) return set_error(CURTREE_EINVAL, "tree_load: short file");
    CT_HIP(hipMemcpy(dst + off, buf.p, b, hipMemcpyHostToDevice));
  }
  return CURTREE_OK;
}

// Every node's children must lie inside the tree: node children
// [pos, pos+len) below r_sz, or box children [pos-r_sz, pos-r_sz+len) inside
// the n boxes.  The query never follows a range outside them, but a file that
// holds one is corrupt and is refused.
int check_file_layout(FILE* f, int64_t n, int64_t r_sz) {
  const long off = (long)(sizeof(FileHeader) + (size_t)n * 20 + (size_t)r_sz * 16);
  const long want = off + (long)r_sz * 8;
  if (fseek(f, 0, SEEK_END) != 0 || ftell(f) != want)
    return set_error(CURTREE_EINVAL, "tree_load: the file size does not match its header");
  std::vector<int32_t> pos((size_t)r_sz), len((size_t)r_sz);
  if (fseek(f, off, SEEK_SET) != 0 || fread(pos.data(), 4, (size_t)r_sz, f) != (size_t)r_sz ||
      fread(len.data(), 4, (size_t)r_sz, f) != (size_t)r_sz)
    return set_error(CURTREE_EINVAL, "tree_load: short file");
  for (int64_t j = 0; j < r_sz; j++) {
    const int64_t p = pos[(size_t)j], l = len[(size_t)j];
    const bool nodes = p >= 0 && l >= 0 && p + l <= r_sz;
    const bool boxes = p >= r_sz && l >= 0 && p - r_sz + l <= n;
    if (!nodes && !boxes)
      return set_error(CURTREE_EINVAL,
                       "tree_load: node %lld points outside the tree (pos %lld, len %lld)",
                       (long long)j, (long long)p, (long long)l);
  }
  if (fseek(f, (long)sizeof(FileHeader), SEEK_SET) != 0)
    return set_error(CURTREE_EINVAL, "tree_load: seek failed");
  return CURTREE_OK;
}

}  // namespace

extern "C" {

CURTREE_API int curtree_read_mbr_csv(const char* path, int size, float r, float* xmin, float* ymin,
                                     float* xmax, float* ymax, int* id) {
  if (!path || size < 0 || (size > 0 && (!xmin || !ymin || !xmax || !ymax || !id)))
    return set_error(CURTREE_EINVAL, "read_mbr_csv: bad argument");
  FILE* fp = fopen(path, "r");
  if (!fp) return set_error(CURTREE_EINVAL, "read_mbr_csv: open %s failed", path);
  for (int i = 0; i < size; i++) {
    float x1, y1, x2, y2;
    int pid, nume;
    int ret = fscanf(fp, "%f,%f,%f,%f,%d,%d,%d\n", &x1, &y1, &x2, &y2, &id[i], &pid, &nume);
    if (ret != 7) {
      fclose(fp);
      return set_error(CURTREE_EINVAL, "read_mbr_csv: line %d: expected 7 fields, got %d", i + 1,
                       ret);
    }
    xmin[i] = std::min(x1, x2) - r;
    ymin[i] = std::min(y1, y2) - r;
    xmax[i] = std::max(x1, x2) + r;
    ymax[i] = std::max(y1, y2) + r;
  }
  fclose(fp);
  return CURTREE_OK;
}

CURTREE_API int curtree_tree_save(const char* path, const curtree_rtree_f32* rt,
                                  const curtree_bbox_f32* rbox) {
  if (!path || !rt || !rbox || rbox->sz < 0 || rt->sz < 0)
    return set_error(CURTREE_EINVAL, "tree_save: bad argument");
  FILE* f = fopen(path, "wb");
  if (!f) return set_error(CURTREE_EINVAL, "tree_save: cannot open %s", path);
  FileHeader h;
  memcpy(h.magic, kMagic, 8);
  h.n_boxes = rbox->sz;
  h.r_sz = rt->sz;
  h.fanout = rt->fanout;
  h.reserved = 0;
  int rc = CURTREE_OK;
  Pinned buf;
  if (fwrite(&h, sizeof h, 1, f) != 1) rc = set_error(CURTREE_EINVAL, "tree_save: write failed");
  if (!rc && hipHostMalloc(&buf.p, kChunk, hipHostMallocDefault) != hipSuccess)
    rc = set_error(CURTREE_ENOMEM, "tree_save: cannot allocate staging buffer");
  const size_t nb = (size_t)rbox->sz * 4, nt = (size_t)rt->sz * 4;
  const void* arrs[11] = {rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax, rbox->id,
                          rt->xmin,   rt->ymin,   rt->xmax,   rt->ymax,   rt->pos,
                          rt->len};
  for (int i = 0; i < 11 && !rc; i++) rc = write_dev(f, arrs[i], i < 5 ? nb : nt, buf);
  if (fclose(f) != 0 && !rc) rc = set_error(CURTREE_EINVAL, "tree_save: close failed");
  return rc;
}

// Allocates the box and node arrays (free with curtree_box_d_free and
// curtree_rt_d_free).
CURTREE_API int curtree_tree_load(const char* path, curtree_rtree_f32* rt,
                                  curtree_bbox_f32* rbox) {
  if (!path || !rt || !rbox) return set_error(CURTREE_EINVAL, "tree_load: bad argument");
  FILE* f = fopen(path, "rb");
  if (!f) return set_error(CURTREE_EINVAL, "tree_load: cannot open %s", path);
  FileHeader h;
  int rc = CURTREE_OK;
  if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, kMagic, 8) != 0)
    rc = set_error(CURTREE_EINVAL, "tree_load: %s is not a curtree file", path);
  LevelPlan lp;
  if (!rc) {
    level_plan(h.n_boxes, h.fanout, &lp);
    if (h.n_boxes < 0 || h.n_boxes > INT32_MAX || h.fanout < 2 || lp.r_sz != h.r_sz)
      rc = set_error(CURTREE_EINVAL, "tree_load: inconsistent header");
  }
  // the node ranges are checked on the host before any device memory is touched
  if (!rc) rc = check_file_layout(f, h.n_boxes, h.r_sz);
  if (rc) {
    fclose(f);
    return rc;
  }
  Pinned buf;
  if (hipHostMalloc(&buf.p, kChunk, hipHostMallocDefault) != hipSuccess) {
    fclose(f);
    return set_error(CURTREE_ENOMEM, "tree_load: cannot allocate staging buffer");
  }
  rc = curtree_box_d_alloc(rbox, (int)h.n_boxes);
  if (!rc) {
    rc = rt_alloc(rt, h.r_sz);
    rt->fanout = h.fanout;
  }
  const size_t nb = (size_t)h.n_boxes * 4, nt = (size_t)h.r_sz * 4;
  void* arrs[11] = {rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax, rbox->id,
                    rt->xmin,   rt->ymin,   rt->xmax,   rt->ymax,   rt->pos,
                    rt->len};
  for (int i = 0; i < 11 && !rc; i++) rc = read_dev(f, arrs[i], i < 5 ? nb : nt, buf);
  fclose(f);
  if (rc) {
    curtree_box_d_free(rbox);
    rt_free(rt);
  }
  return rc;
}

}  // extern "C"
