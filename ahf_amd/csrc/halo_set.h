/*
 * halo_set.h -- the resident halo lists of a context (host-side
 * bookkeeping shared by the halo gather and the unbinding).
 *
 * One device block and one pinned host block, laid out alike (off, cen,
 * order, ip_in), are kept across calls: a caller that handles one halo at a
 * time pays one copy and no allocation per call.
 */
#ifndef AHF_HALO_SET_H
#define AHF_HALO_SET_H

#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <vector>

#include "gpu_ctx.h"

/* per-halo result of the unbinding steps (device) */
struct HaloOut {
  int64_t npart;
  double M_vir, R_vir, Phi0, ovdens;
  int32_t niter, replays;
  int32_t in_b; /* final list is in buffer b (0 = a) */
  int32_t set;  /* AHFGPU_SET_* fields written */
  int64_t particle_iterations;
  int32_t sweeps, pad;
};

struct HaloSet {
  int64_t nh = 0, ne = 0;
  int64_t *off = nullptr;     /* nh+1 */
  uint32_t *ip_in = nullptr;  /* ne: members as given */
  double *cen = nullptr;      /* nh*3 */
  int32_t *order = nullptr;   /* haloes largest first */
  std::vector<int64_t> h_off;
  std::vector<int32_t> h_ord; /* the host copy of order */
  HaloOut *out = nullptr;
  uint32_t *buf_a = nullptr, *buf_b = nullptr;
  int done = 0;
  int res_staged = 0; /* the results are in `res` already (read back with the unbinding's own synchronisation) */
  char *dev = nullptr, *host = nullptr;
  size_t cap = 0;
  char *res = nullptr; /* pinned staging of the results */
  size_t res_cap = 0;
  char *up = nullptr; /* pinned staging of per-call uploads */
  size_t up_cap = 0;
  size_t o_cen = 0, o_ord = 0, o_ip = 0, need = 0;
};

static inline void free_halos(HaloSet *s)
{
  if (!s) return;
  if (s->dev) (void)hipFree(s->dev);
  if (s->host) (void)hipHostFree(s->host);
  if (s->res) (void)hipHostFree(s->res);
  if (s->up) (void)hipHostFree(s->up);
  delete s;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/* take the context's halo set (or a new one) and size it for nh haloes and
 * ne members; the device pointers are set, nothing is copied.  The set is
 * detached from the context until the caller stores it back. */
static inline int halo_set_reserve(ahfgpu_ctx *c, int64_t nh, int64_t ne, HaloSet **out)
{
  HaloSet *s = (HaloSet *)c->halo_set;
  if (!s) {
    /* the buffers of the last released set, kept for reuse */
    s = (HaloSet *)c->halo_spare;
    c->halo_spare = nullptr;
  }
  if (!s) s = new HaloSet();
  c->halo_set = nullptr;
  s->done = 0;
  s->res_staged = 0;
  s->out = nullptr;
  s->buf_a = s->buf_b = nullptr;
  s->nh = nh;
  s->ne = ne;
  s->o_cen = align256(((size_t)nh + 1) * sizeof(int64_t));
  s->o_ord = s->o_cen + align256((3 * (size_t)nh + 1) * sizeof(double));
  s->o_ip = s->o_ord + align256(((size_t)nh + 1) * sizeof(int32_t));
  s->need = s->o_ip + align256(((size_t)ne + 1) * sizeof(uint32_t));
  /* the previous call's copy may still read the pinned block */
  (void)hipStreamSynchronize(c->st);
  if (s->need > s->cap) {
    if (s->dev) (void)hipFree(s->dev);
    if (s->host) (void)hipHostFree(s->host);
    s->dev = s->host = nullptr;
    s->cap = 0;
    const size_t cap = s->need + s->need / 2;
    if (hipMalloc((void **)&s->dev, cap) != hipSuccess || hipHostMalloc((void **)&s->host, cap) != hipSuccess) {
      (void)hipGetLastError();
      free_halos(s);
      return c->fail(AHFGPU_ENOMEM, "halo arrays");
    }
    s->cap = cap;
  }
  s->off = (int64_t *)s->dev;
  s->cen = (double *)(s->dev + s->o_cen);
  s->order = (int32_t *)(s->dev + s->o_ord);
  s->ip_in = (uint32_t *)(s->dev + s->o_ip);
  *out = s;
  return 0;
}

/* fill the host block's offsets, centres and largest-first order (stable
 * on equal sizes) and keep the offsets on the host */
static inline void halo_set_host_meta(HaloSet *s, const int64_t *off, const double *centres)
{
  const int64_t nh = s->nh;
  s->h_off.assign(off, off + nh + 1);
  memcpy(s->host, off, ((size_t)nh + 1) * sizeof(int64_t));
  if (nh) memcpy(s->host + s->o_cen, centres, 3 * (size_t)nh * sizeof(double));
  int32_t *ord = (int32_t *)(s->host + s->o_ord);
  for (int64_t h = 0; h < nh; h++) ord[h] = (int32_t)h;
  std::stable_sort(ord, ord + nh,
                   [&](int32_t x, int32_t y) { return off[x + 1] - off[x] > off[y + 1] - off[y]; });
  s->h_ord.assign(ord, ord + nh);
}

/* a pinned buffer of at least `bytes` (grown as needed), or nullptr */
static inline char *pinned_grow(char **p, size_t *cap, size_t bytes)
{
  if (bytes > *cap) {
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t want = bytes + bytes / 2 + 4096;
    if (hipHostMalloc((void **)p, want) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return nullptr;
    }
    *cap = want;
  }
  return *p;
}

/* layout of the results (device block and pinned staging alike): the
 * per-halo records, then member buffers a and b, each 256-byte aligned */
static inline size_t res_off_a(int64_t nh) { return align256((size_t)nh * sizeof(HaloOut)); }
static inline size_t res_off_b(int64_t nh, int64_t ne) { return res_off_a(nh) + align256((size_t)ne * 4); }

/* pinned staging for the results (read back after the stream is idle) */
static inline char *halo_set_res(HaloSet *s, size_t bytes) { return pinned_grow(&s->res, &s->res_cap, bytes); }

/* pinned staging for the uploads of one unbinding call (every call ends
 * with its stream synchronised, so the next one may overwrite it) */
static inline char *halo_set_up(HaloSet *s, size_t bytes) { return pinned_grow(&s->up, &s->up_cap, bytes); }

#endif
