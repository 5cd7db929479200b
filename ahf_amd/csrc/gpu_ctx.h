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
#include "amr_core.h"

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
  uint32_t *pnode = nullptr;   /* per bucket entry: its node */
  ahf_w9 *W = nullptr;                              /* per bucket entry: TSC weights (sparse)     */
  ahf_q4 *qp = nullptr;                             /* per bucket entry: position (dense levels)  */
  uint32_t *tord = nullptr;                         /* octets in Morton order                      */
  int8_t *child = nullptr;     /* per bucket entry: fine child it moved to, or -1 */
  int32_t *node2oct = nullptr; /* per node: octet of the next level, or -1 */
  double m2d = 0, m2pd = 0, crit = 0;
};

/* timed stages of one build / unbinding (HIP events on the library stream) */
enum {
  TIMER_BIN = 0,     /* domain binning: key + radix sort + CSR           */
  TIMER_PREP,        /* k_prep                                          */
  TIMER_GATHER,      /* k_gather_light                                  */
  TIMER_HEAVY,       /* k_gather_heavy                                  */
  TIMER_REFINE,      /* k_flags + k_refine_rows + scans                 */
  TIMER_BUILD,       /* k_compact + k_nbr_next + k_pmask of a new level */
  TIMER_RELINK,      /* k_relink_child + radix sort + CSR               */
  TIMER_USORT,       /* unbinding: radial sort                          */
  TIMER_UNBIND,      /* unbinding: iterations                           */
  TIMER_HGATHER,     /* halo gather: cells, tests, member lists         */
  TIMER_PROFILES,    /* HaloProfiles: k_profiles                        */
  TIMER_COUNT
};
static const char *const k_timer_names[TIMER_COUNT] = {"bin",    "prep",        "gather_light", "gather_heavy",
                                                       "refine", "build_level", "relink",       "unbind_sort",
                                                       "unbind", "halo_gather", "halo_profiles"};

struct TimedLaunch {
  hipEvent_t a, b;
  int cat;
  double bytes;
};

struct ahfgpu_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr; /* the split unbinding path beside st (default priority) */
  hipStream_t st3 = nullptr; /* the one-wavefront unbinding tier beside st (default priority) */
  hipEvent_t ev_s0 = nullptr, ev_s1 = nullptr, ev_s2 = nullptr, ev_s3 = nullptr;
  int64_t split_min = 0;     /* haloes of at least this many members take the split path; 0: default */
  int64_t prof_split = 0;    /* haloes of at least this many members are profiled by four wavefronts; 0: default */
  /* the rest of the schedule (ahfgpu_schedule); 0: the default */
  int32_t kmax_big = 0, kmax_small = 0; /* contributions above which a node's density is folded by a wavefront */
  int32_t march_zn = 0;                 /* target planes per march workgroup */
  int64_t big_halo = 0, wave_halo = 0;  /* unbinding tiers: 512-thread workgroups from, one wavefront below */
  char err[512] = {0};
  Arena arena;        /* AMR build buffers */
  Arena uarena;       /* unbinding buffers */
  Arena parena;       /* profile buffers */
  void *prof_result = nullptr; /* last ahfgpu_halo_profiles result (profiles_gpu.hip) */
  Arena garena;       /* gridinfo buffers */
  void *gi_result = nullptr;   /* last ahfgpu_gridinfo result (gridinfo_gpu.hip) */
  int64_t N = 0;
  float *px = nullptr, *py = nullptr, *pz = nullptr;
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  /* the same, interleaved: particle q's x, y, z, vx in pv[2q] and vy, vz in
   * pv[2q + 1] (one 32-byte sector for kernels that gather members) */
  float4 *pv = nullptr;
  int64_t cap = 0;
  ahfgpu_amr_params params{};
  std::vector<DevLevel> lev;
  DevLevel rb;        /* rolled-back level (lists appended to its mothers) */
  int rolled_back = 0;
  int built = 0;
  int32_t *d_anomaly = nullptr;
  int64_t *d_sizes = nullptr;   /* level sizes gathered for one read-back  */
  int64_t *h_sizes = nullptr;   /* pinned host copy of d_sizes             */
  int32_t *h_words = nullptr;   /* pinned host words (the split path's done flag) */
  int split_spec = 0;           /* speculative passes per split iteration; 0: 3, -1: none */
  int64_t prof_tiled = 0;       /* HaloProfiles over tiles from this many members; 0: default, -1: none */
  int split_graph_batch = 0;    /* graph replays between done checks; 0: 4 */
  int split_graph_launches = 0; /* replays of the last split unbinding */
  void *halo_set = nullptr; /* resident halo lists (halo_set.h) */
  void *halo_spare = nullptr; /* a released set's buffers, kept for reuse */
  uint64_t *sfckey = nullptr; /* SFC keys of the resident particles */
  int64_t key_cap = 0;
  int keys_ok = 0;            /* keys present and in order */
  int32_t *d_flag = nullptr;  /* one device word for checks */
  int64_t gather_candidates = 0;
  int gather_mode = 0;      /* AHFGPU_GATHER_*: deposit kernel choice */
  hipEvent_t ev_a0 = nullptr, ev_a1 = nullptr, ev_u0 = nullptr, ev_u1 = nullptr;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_next = 0;
  std::vector<TimedLaunch> timed;
  bool timing = true; /* stage timers and the build/unbind event pairs */
  double amr_ms = 0, unbind_ms = 0, gather_ms = 0, gather_bytes = 0;
  int64_t gather_launches = 0;
  double t_ms[TIMER_COUNT] = {0}, t_bytes[TIMER_COUNT] = {0};
  int64_t t_n[TIMER_COUNT] = {0};

  int fail(int code, const char *fmt, ...)
  {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, sizeof(err), fmt, ap);
    va_end(ap);
    return code;
  }
  void timers_reset(int first_cat, int last_cat)
  {
    for (int k = first_cat; k <= last_cat; k++) {
      t_ms[k] = 0;
      t_bytes[k] = 0;
      t_n[k] = 0;
    }
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
    if (!timing) return nullptr;
    hipEvent_t e = take_event();
    hipEventRecord(e, st);
    return e;
  }
  void timer_end(hipEvent_t a, int cat, double bytes)
  {
    if (!a) return;
    hipEvent_t b = take_event();
    hipEventRecord(b, st);
    timed.push_back(TimedLaunch{a, b, cat, bytes});
  }
  /* after a synchronisation: fold the pending events into the totals */
  void timers_collect()
  {
    for (auto &t : timed) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, t.a, t.b);
      t_ms[t.cat] += ms;
      t_bytes[t.cat] += t.bytes;
      t_n[t.cat] += 1;
    }
    timed.clear();
    ev_next = 0;
    gather_ms = t_ms[TIMER_GATHER];
    gather_bytes = t_bytes[TIMER_GATHER];
    gather_launches = t_n[TIMER_GATHER];
  }
};

/* profiles_gpu.hip: drop the last profile result and its buffers */
void ahfgpu_free_profiles(ahfgpu_ctx *c);

/* unbind_gpu.hip: take the device address of the static scratch of
 * ahfgpu_unbind_noalloc on `device` (current), once per device */
int ahf_noalloc_init(int device);
/* releases the last ahfgpu_gridinfo result (gridinfo_gpu.hip) */
void ahfgpu_free_gridinfo(ahfgpu_ctx *c);

#endif
