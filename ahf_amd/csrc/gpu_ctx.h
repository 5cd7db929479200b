/*
 * gpu_ctx.h -- library context: device, stream, a bump arena for the
 * per-build buffers, the resident particle set and the levels of the last
 * AMR hierarchy.
 */
#ifndef AHF_GPU_CTX_H
#define AHF_GPU_CTX_H

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../include/ahf_gpu.h"

#define AHF_CRITMULTI_D 8.0 /* src/param.h:118 */

/* bump allocator over a list of large device chunks; reset() keeps them, so
 * repeated builds of the same workload allocate nothing */
struct Arena {
  struct Chunk {
    char *base;
    size_t size, used;
  };
  std::vector<Chunk> chunks;
  size_t cur = 0;
  size_t min_chunk = (size_t)256 << 20;

  void reset()
  {
    for (auto &c : chunks) c.used = 0;
    cur = 0;
  }
  void *raw(size_t bytes)
  {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    while (cur < chunks.size()) {
      Chunk &c = chunks[cur];
      if (c.size - c.used >= bytes) {
        void *p = c.base + c.used;
        c.used += bytes;
        return p;
      }
      cur++;
    }
    size_t sz = bytes > min_chunk ? bytes : min_chunk;
    char *p = nullptr;
    if (hipMalloc((void **)&p, sz) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    chunks.push_back(Chunk{p, sz, bytes});
    cur = chunks.size() - 1;
    return p;
  }
  template <typename T>
  T *get(int64_t n)
  {
    return (T *)raw((size_t)(n > 0 ? n : 1) * sizeof(T));
  }
  void release()
  {
    for (auto &c : chunks) (void)hipFree(c.base);
    chunks.clear();
    cur = 0;
  }
};

struct DevLevel {
  int64_t L = 0, n_oct = 0, n_row = 0, n_plane = 0, n_part = 0;
  uint64_t *oct_key = nullptr;
  uint8_t *oct_link = nullptr;
  int32_t *nbr = nullptr;
  uint64_t *row_key = nullptr;
  int32_t *row_oct0 = nullptr, *row_plane = nullptr, *plane_z = nullptr, *plane_row0 = nullptr;
  int64_t *oct_parent = nullptr;
  uint32_t *pmask = nullptr;
  float *dens = nullptr;
  uint8_t *flags = nullptr;
  int64_t *node_p0 = nullptr;
  uint32_t *plist = nullptr;
  int8_t *child = nullptr;     /* per bucket entry: fine child it moved to, or -1 */
  int32_t *node2oct = nullptr; /* per node: octet of the next level, or -1 */
  double m2d = 0, m2pd = 0, crit = 0;
};

struct TimedLaunch {
  hipEvent_t a, b;
  double bytes;
};

struct ahfgpu_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  char err[512] = {0};
  Arena arena;        /* AMR build buffers */
  Arena uarena;       /* unbinding buffers */
  int64_t N = 0;
  float *px = nullptr, *py = nullptr, *pz = nullptr;
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  int64_t cap = 0;
  ahfgpu_amr_params params{};
  std::vector<DevLevel> lev;
  DevLevel rb;        /* rolled-back level (lists appended to its mothers) */
  int rolled_back = 0;
  int built = 0;
  int32_t *d_anomaly = nullptr;
  void *halo_set = nullptr; /* resident halo lists (unbind_gpu.hip) */
  hipEvent_t ev_a0 = nullptr, ev_a1 = nullptr, ev_u0 = nullptr, ev_u1 = nullptr;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_next = 0;
  std::vector<TimedLaunch> timed;
  double amr_ms = 0, unbind_ms = 0, gather_ms = 0, gather_bytes = 0;
  int64_t gather_launches = 0;

  int fail(int code, const char *fmt, ...)
  {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, sizeof(err), fmt, ap);
    va_end(ap);
    return code;
  }
  void timers_reset()
  {
    ev_next = 0;
    timed.clear();
  }
  hipEvent_t take_event()
  {
    if (ev_next == ev_pool.size()) {
      hipEvent_t e;
      hipEventCreate(&e);
      ev_pool.push_back(e);
    }
    return ev_pool[ev_next++];
  }
  hipEvent_t timer_begin()
  {
    hipEvent_t e = take_event();
    hipEventRecord(e, st);
    return e;
  }
  void timer_end(hipEvent_t a, double bytes)
  {
    hipEvent_t b = take_event();
    hipEventRecord(b, st);
    timed.push_back(TimedLaunch{a, b, bytes});
  }
  void timers_collect()
  {
    gather_ms = 0;
    gather_bytes = 0;
    gather_launches = (int64_t)timed.size();
    for (auto &t : timed) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, t.a, t.b);
      gather_ms += ms;
      gather_bytes += t.bytes;
    }
  }
};

#endif
