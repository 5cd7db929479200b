/*
 * api.hip -- context management and particle upload of the C ABI
 * (include/ahf_gpu.h).
 */
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/ahf_gpu.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "halo_set.h"

__global__ void k_deinterleave(const float *__restrict__ xyz, int64_t n, float *__restrict__ x,
                               float *__restrict__ y, float *__restrict__ z)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = xyz[3 * i];
  y[i] = xyz[3 * i + 1];
  z[i] = xyz[3 * i + 2];
}

/* particle q's position and velocity into pv[2q], pv[2q + 1] */
__global__ void k_pack_pv(int64_t n, const float *__restrict__ x, const float *__restrict__ y,
                          const float *__restrict__ z, const float *__restrict__ vx, const float *__restrict__ vy,
                          const float *__restrict__ vz, float4 *__restrict__ pv)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  pv[2 * q] = make_float4(x[q], y[q], z[q], vx[q]);
  pv[2 * q + 1] = make_float4(vy[q], vz[q], 0.f, 0.f);
}

/* the integer value of "key" in a flat JSON object, if present */
static bool json_int(const std::string &txt, const char *key, long long *v)
{
  const std::string k = std::string("\"") + key + "\"";
  size_t p = txt.find(k);
  if (p == std::string::npos) return false;
  p = txt.find(':', p + k.size());
  if (p == std::string::npos) return false;
  char *end = nullptr;
  const long long x = strtoll(txt.c_str() + p + 1, &end, 10);
  if (end == txt.c_str() + p + 1) return false;
  *v = x;
  return true;
}

/* the schedule file: $AHF_SCHEDULE, else schedule.json beside the
 * library's directory (the package directory); silently absent */
static void load_schedule(ahfgpu_ctx *c)
{
  std::string path;
  const char *env = getenv("AHF_SCHEDULE");
  if (env && *env) {
    path = env;
  } else {
    Dl_info info;
    if (!dladdr((void *)&ahfgpu_create, &info) || !info.dli_fname) return;
    path = info.dli_fname;
    for (int up = 0; up < 2; up++) {
      const size_t slash = path.find_last_of('/');
      if (slash == std::string::npos) return;
      path.resize(slash);
    }
    path += "/schedule.json";
  }
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return;
  std::string txt;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), f)) > 0 && txt.size() < (1u << 20)) txt.append(buf, n);
  fclose(f);
  ahfgpu_schedule sc;
  ahfgpu_get_schedule(c, &sc);
  long long v;
  if (json_int(txt, "split_min", &v)) sc.split_min = v;
  if (json_int(txt, "prof_split", &v)) sc.prof_split = v;
  if (json_int(txt, "big_halo", &v)) sc.big_halo = v;
  if (json_int(txt, "wave_halo", &v)) sc.wave_halo = v;
  if (json_int(txt, "kmax_big", &v)) sc.kmax_big = (int32_t)v;
  if (json_int(txt, "kmax_small", &v)) sc.kmax_small = (int32_t)v;
  if (json_int(txt, "march_zn", &v)) sc.march_zn = (int32_t)v;
  if (json_int(txt, "graph_batch", &v)) sc.graph_batch = (int32_t)v;
  if (json_int(txt, "split_spec", &v)) sc.split_spec = (int32_t)v;
  if (json_int(txt, "prof_tiled", &v)) sc.prof_tiled = v;
  if (ahfgpu_set_schedule(c, &sc) != 0) fprintf(stderr, "ahfgpu: schedule %s ignored: %s\n", path.c_str(), c->err);
}

extern "C" int ahfgpu_set_schedule(ahfgpu_ctx *c, const ahfgpu_schedule *s)
{
  if (!c || !s) return AHFGPU_EINVAL;
  if (s->kmax_big < 0 || s->kmax_big >= 4096 || s->kmax_small < 0 || s->kmax_small >= 4096)
    return c->fail(AHFGPU_EINVAL, "kmax must be in [0, 4096)");
  if (s->march_zn != 0 && s->march_zn != 8 && s->march_zn != 16) return c->fail(AHFGPU_EINVAL, "march_zn: 0, 8 or 16");
  if (s->big_halo < 0 || s->wave_halo < 0) return c->fail(AHFGPU_EINVAL, "tier thresholds must not be negative");
  if (s->graph_batch < 0 || s->graph_batch > 64) return c->fail(AHFGPU_EINVAL, "graph_batch must be in [0, 64]");
  if (s->split_spec < -1 || s->split_spec > 16) return c->fail(AHFGPU_EINVAL, "split_spec must be in [-1, 16]");
  c->split_min = s->split_min;
  c->prof_split = s->prof_split;
  c->big_halo = s->big_halo;
  c->wave_halo = s->wave_halo;
  c->kmax_big = s->kmax_big;
  c->kmax_small = s->kmax_small;
  c->march_zn = s->march_zn;
  c->split_graph_batch = s->graph_batch;
  c->split_spec = s->split_spec;
  c->prof_tiled = s->prof_tiled;
  return 0;
}

extern "C" int ahfgpu_get_schedule(ahfgpu_ctx *c, ahfgpu_schedule *s)
{
  if (!c || !s) return AHFGPU_EINVAL;
  s->split_min = c->split_min;
  s->prof_split = c->prof_split;
  s->big_halo = c->big_halo;
  s->wave_halo = c->wave_halo;
  s->kmax_big = c->kmax_big;
  s->kmax_small = c->kmax_small;
  s->march_zn = c->march_zn;
  s->graph_batch = c->split_graph_batch;
  s->split_spec = c->split_spec;
  s->pad = 0;
  s->prof_tiled = c->prof_tiled;
  return 0;
}

extern "C" int ahfgpu_create(ahfgpu_ctx **out, int device)
{
  if (!out) return AHFGPU_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return AHFGPU_EHIP;
  }
  if (device < 0 || device >= ndev) return AHFGPU_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return AHFGPU_EHIP;
  ahfgpu_ctx *c = new ahfgpu_ctx();
  c->device = device;
  if (hipStreamCreate(&c->st) != hipSuccess || hipStreamCreate(&c->st2) != hipSuccess ||
      hipStreamCreate(&c->st3) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_s0, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_s1, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_s2, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_s3, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&c->d_anomaly, 64) != hipSuccess || hipMalloc((void **)&c->d_sizes, 64) != hipSuccess ||
      hipHostMalloc((void **)&c->h_sizes, 64, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->h_words, 64, hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev_a0) != hipSuccess || hipEventCreate(&c->ev_a1) != hipSuccess ||
      hipEventCreate(&c->ev_u0) != hipSuccess || hipEventCreate(&c->ev_u1) != hipSuccess) {
    delete c;
    return AHFGPU_EHIP;
  }
  /* the scratch of ahfgpu_unbind_noalloc on this device; without it only
   * that entry point fails (it says so), the rest of the library works */
  (void)ahf_noalloc_init(device);
  load_schedule(c);
  *out = c;
  return 0;
}

static void free_particles(ahfgpu_ctx *c)
{
  float **ps[] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz};
  for (auto p : ps) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  if (c->pv) (void)hipFree(c->pv);
  c->pv = nullptr;
  c->cap = 0;
  if (c->sfckey) (void)hipFree(c->sfckey);
  c->sfckey = nullptr;
  c->key_cap = 0;
  c->keys_ok = 0;
}

extern "C" void ahfgpu_destroy(ahfgpu_ctx *c)
{
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->st);
  ahfgpu_release_halos(c);
  free_halos((HaloSet *)c->halo_spare);
  c->halo_spare = nullptr;
  ahfgpu_free_profiles(c);
  ahfgpu_free_gridinfo(c);
  c->arena.release();
  c->uarena.release();
  free_particles(c);
  if (c->d_anomaly) (void)hipFree(c->d_anomaly);
  if (c->d_sizes) (void)hipFree(c->d_sizes);
  if (c->h_sizes) (void)hipHostFree(c->h_sizes);
  if (c->h_words) (void)hipHostFree(c->h_words);
  if (c->d_flag) (void)hipFree(c->d_flag);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(c->st2);
  (void)hipStreamSynchronize(c->st3);
  hipEvent_t evs[] = {c->ev_a0, c->ev_a1, c->ev_u0, c->ev_u1, c->ev_s0, c->ev_s1, c->ev_s2, c->ev_s3};
  for (auto e : evs)
    if (e) (void)hipEventDestroy(e);
  if (c->st) (void)hipStreamDestroy(c->st);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st3) (void)hipStreamDestroy(c->st3);
  delete c;
}

extern "C" const char *ahfgpu_last_error(const ahfgpu_ctx *c)
{
  return c ? c->err : "no context";
}

extern "C" int ahfgpu_set_particles(ahfgpu_ctx *c, const float *pos, const float *mom, uint64_t n)
{
  if (!c || !pos || n == 0) return AHFGPU_EINVAL;
  if (n >= (uint64_t)UINT32_MAX) return c->fail(AHFGPU_EINVAL, "at most 2^32-1 particles");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  if ((int64_t)n > c->cap) {
    free_particles(c);
    float **ps[] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz};
    for (auto p : ps)
      if (hipMalloc((void **)p, n * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        free_particles(c);
        return c->fail(AHFGPU_ENOMEM, "particle arrays (%llu)", (unsigned long long)n);
      }
    if (hipMalloc((void **)&c->pv, n * 2 * sizeof(float4)) != hipSuccess) {
      (void)hipGetLastError();
      free_particles(c);
      return c->fail(AHFGPU_ENOMEM, "interleaved particle array (%llu)", (unsigned long long)n);
    }
    c->cap = (int64_t)n;
  }
  float *stage = nullptr;
  if (hipMalloc((void **)&stage, n * 3 * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    return c->fail(AHFGPU_ENOMEM, "staging buffer");
  }
  const unsigned nb = (unsigned)((n + AHF_BLOCK - 1) / AHF_BLOCK);
  hipMemcpyAsync(stage, pos, n * 3 * sizeof(float), hipMemcpyHostToDevice, c->st);
  hipLaunchKernelGGL(k_deinterleave, dim3(nb), dim3(AHF_BLOCK), 0, c->st, stage, (int64_t)n, c->px, c->py, c->pz);
  if (mom) {
    hipStreamSynchronize(c->st);
    hipMemcpyAsync(stage, mom, n * 3 * sizeof(float), hipMemcpyHostToDevice, c->st);
    hipLaunchKernelGGL(k_deinterleave, dim3(nb), dim3(AHF_BLOCK), 0, c->st, stage, (int64_t)n, c->vx, c->vy,
                       c->vz);
  } else {
    hipMemsetAsync(c->vx, 0, n * sizeof(float), c->st);
    hipMemsetAsync(c->vy, 0, n * sizeof(float), c->st);
    hipMemsetAsync(c->vz, 0, n * sizeof(float), c->st);
  }
  hipLaunchKernelGGL(k_pack_pv, dim3(nb), dim3(AHF_BLOCK), 0, c->st, (int64_t)n, c->px, c->py, c->pz, c->vx, c->vy,
                     c->vz, c->pv);
  hipError_t e = hipStreamSynchronize(c->st);
  (void)hipFree(stage);
  if (e != hipSuccess) return c->fail(AHFGPU_EHIP, "particle upload: %s", hipGetErrorString(e));
  c->N = (int64_t)n;
  c->built = 0;
  c->keys_ok = 0;
  return 0;
}

extern "C" int ahfgpu_set_gather_mode(ahfgpu_ctx *c, int mode)
{
  if (!c) return AHFGPU_EINVAL;
  if (mode != AHFGPU_GATHER_AUTO && mode != AHFGPU_GATHER_SOURCE && mode != AHFGPU_GATHER_FLAT &&
      mode != AHFGPU_GATHER_WAVE && mode != AHFGPU_GATHER_FLAT_WIDE)
    return c->fail(AHFGPU_EINVAL, "unknown gather mode %d", mode);
  c->gather_mode = mode;
  return 0;
}

extern "C" int ahfgpu_set_timing(ahfgpu_ctx *c, int on)
{
  if (!c) return AHFGPU_EINVAL;
  c->timing = on != 0;
  return 0;
}

extern "C" int ahfgpu_last_timing(ahfgpu_ctx *c, double *amr_ms, double *unbind_ms, double *gather_ms,
                                  int64_t *gather_launches)
{
  if (!c) return AHFGPU_EINVAL;
  if (amr_ms) *amr_ms = c->amr_ms;
  if (unbind_ms) *unbind_ms = c->unbind_ms;
  if (gather_ms) *gather_ms = c->gather_ms;
  if (gather_launches) *gather_launches = c->gather_launches;
  return 0;
}

extern "C" int ahfgpu_timer_count(void)
{
  return TIMER_COUNT;
}

extern "C" int ahfgpu_timer_get(ahfgpu_ctx *c, int i, const char **name, double *ms, int64_t *launches,
                                double *alg_bytes)
{
  if (!c || i < 0 || i >= TIMER_COUNT) return AHFGPU_EINVAL;
  if (name) *name = k_timer_names[i];
  if (ms) *ms = c->t_ms[i];
  if (launches) *launches = c->t_n[i];
  if (alg_bytes) *alg_bytes = c->t_bytes[i];
  return 0;
}
