/*
 * gather_gpu.hip -- AHF's halo gather on gfx950: the member lists of every
 * halo from its centre and gather radius (ahf_halos_sfc_gatherParts,
 * src/libahf/ahf_halos_sfc.c:171-367), and the SFC keys it searches
 * (src/main.c:341-356).
 *
 * The particles are ordered by key, so each cell of a halo's 3x3x3 search
 * block is a contiguous range of them.  All haloes at once:
 *   1. two threads per cell of every halo's block find the cell's particle
 *      range by binary search over the keys;
 *   2. the ranges of all haloes, in halo order and block order, are laid
 *      end to end by a scan: candidate q is one particle of one cell of one
 *      halo, and q's order is the order in which AHF adds members;
 *   3. one thread per candidate tests the distance (a table gives the cell
 *      of every run of 256 candidates); a scan of the results gives every
 *      member its place in the halo's list (CSR over haloes).
 * The lists land in the context's resident halo set, ready for the
 * unbinding pass.
 */
#include <hip/hip_runtime.h>

#include <string.h>

#include <vector>

#include "../../include/ahf_gpu.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "halo_set.h"
#include "sfc_core.h"

using namespace ahf;

static inline unsigned nblk(int64_t n) { return (unsigned)((n + AHF_BLOCK - 1) / AHF_BLOCK); }

/* key of every particle at 21 bits per axis (src/main.c:343-350) */
__global__ void k_sfc_keys(const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
                           int64_t n, uint64_t *__restrict__ key)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = ahf_sfc_key((double)px[i], (double)py[i], (double)pz[i], AHF_SFC_BITS);
}

/* count of keys smaller than their predecessor (0: sorted) */
__global__ void k_keys_unsorted(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ bad)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 < n && key[i + 1] < key[i]) atomicAdd(bad, 1);
}

/* first index whose key is at least k (n if none) */
__device__ inline int64_t lower_key(const uint64_t *__restrict__ key, int64_t n, uint64_t k)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (key[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* the particle range of one search cell: thread t handles slot t % 27 of
 * halo t / 27 and one of its two bounds (the first index at or above kmin,
 * or above kmax); skipped haloes (skip[h] != 0) and far cells get empty
 * ranges */
__global__ void k_gather_cells(const double *__restrict__ cen, const double *__restrict__ rad,
                               const uint8_t *__restrict__ skip, int64_t nslot, const uint64_t *__restrict__ key,
                               int64_t n, int64_t *__restrict__ bound /* 2 per slot */)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * nslot) return;
  const int64_t slot = t >> 1, h = slot / AHF_SHELL;
  const int s = (int)(slot % AHF_SHELL);
  int64_t b = 0;
  if (!skip[h]) {
    uint64_t kmin, kmax;
    ahf_gather_cell(cen[3 * h], cen[3 * h + 1], cen[3 * h + 2], rad[h], s, &kmin, &kmax);
    if (kmin <= kmax) b = lower_key(key, n, (t & 1) ? kmax + 1 : kmin);
  }
  bound[t] = b;
}

__global__ void k_gather_counts(const int64_t *__restrict__ bound, int64_t nslot, int64_t *__restrict__ cnt)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nslot) cnt[s] = bound[2 * s + 1] - bound[2 * s];
}

#define AHF_CAND_BLOCK 256 /* candidates per entry of the block -> slot table */

/* for every run of AHF_CAND_BLOCK candidates, the slot holding its first
 * candidate: each non-empty slot claims the runs that start inside it */
__global__ void k_gather_blockslot(const int64_t *__restrict__ start, int64_t nslot, int32_t *__restrict__ tab)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  const int64_t a = start[s], e = start[s + 1];
  if (a == e) return;
  for (int64_t b = (a + AHF_CAND_BLOCK - 1) / AHF_CAND_BLOCK; b * AHF_CAND_BLOCK < e; b++) tab[b] = (int32_t)s;
}

/* the slot holding candidate q: from the slot of q's run, forward past the
 * slots that end at or before q (empty slots share their start with the
 * next one) */
__device__ inline int64_t slot_of(const int64_t *__restrict__ start, const int32_t *__restrict__ tab, int64_t q)
{
  int64_t s = tab[q / AHF_CAND_BLOCK];
  while (start[s + 1] <= q) s++;
  return s;
}

__global__ void k_gather_test(const int64_t *__restrict__ start, const int64_t *__restrict__ bound,
                              const int32_t *__restrict__ tab, int64_t nc, const double *__restrict__ cen,
                              const double *__restrict__ rad, const float *__restrict__ px,
                              const float *__restrict__ py, const float *__restrict__ pz, uint8_t *__restrict__ flag)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  const int64_t s = slot_of(start, tab, q);
  const int64_t h = s / AHF_SHELL;
  const int64_t p = bound[2 * s] + (q - start[s]);
  const double r = rad[h];
  flag[q] = (uint8_t)ahf_gather_member(px[p], py[p], pz[p], cen[3 * h], cen[3 * h + 1], cen[3 * h + 2], r * r);
}

/* member offset of every halo: the members before its first candidate */
__global__ void k_gather_off(const int64_t *__restrict__ start, int64_t nh, int64_t nc,
                             const uint32_t *__restrict__ pos, int64_t *__restrict__ off)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h > nh) return;
  const int64_t q = h < nh ? start[AHF_SHELL * h] : nc;
  off[h] = (int64_t)pos[q < nc ? q : nc];
}

__global__ void k_gather_write(const int64_t *__restrict__ start, const int64_t *__restrict__ bound,
                               const int32_t *__restrict__ tab, int64_t nc, const uint8_t *__restrict__ flag,
                               const uint32_t *__restrict__ pos, uint32_t *__restrict__ ip)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc || !flag[q]) return;
  const int64_t s = slot_of(start, tab, q);
  ip[pos[q]] = (uint32_t)(bound[2 * s] + (q - start[s]));
}

static int ensure_key_array(ahfgpu_ctx *c)
{
  if (c->sfckey && c->key_cap >= c->N) return 0;
  if (c->sfckey) (void)hipFree(c->sfckey);
  c->sfckey = nullptr;
  c->key_cap = 0;
  if (hipMalloc((void **)&c->sfckey, (size_t)(c->N > 0 ? c->N : 1) * sizeof(uint64_t)) != hipSuccess) {
    (void)hipGetLastError();
    return c->fail(AHFGPU_ENOMEM, "key array");
  }
  c->key_cap = c->N;
  return 0;
}

/* the check uses a word of its own: the unbinding arena may still hold the
 * last unbinding's results */
static int check_sorted(ahfgpu_ctx *c)
{
  if (!c->d_flag && hipMalloc((void **)&c->d_flag, 256) != hipSuccess) {
    (void)hipGetLastError();
    return c->fail(AHFGPU_ENOMEM, "key check word");
  }
  int32_t *bad = c->d_flag;
  hipMemsetAsync(bad, 0, sizeof(int32_t), c->st);
  if (c->N > 1) hipLaunchKernelGGL(k_keys_unsorted, dim3(nblk(c->N)), dim3(AHF_BLOCK), 0, c->st, c->sfckey, c->N, bad);
  int32_t h = 0;
  hipMemcpyAsync(&h, bad, sizeof(int32_t), hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess) return c->fail(AHFGPU_EHIP, "key check");
  c->keys_ok = h == 0;
  if (h) return c->fail(AHFGPU_EINVAL, "particles are not in SFC key order (%d inversions)", h);
  return 0;
}

extern "C" int ahfgpu_set_sfc_keys(ahfgpu_ctx *c, const uint64_t *keys, uint64_t n)
{
  if (!c || (n && !keys)) return AHFGPU_EINVAL;
  if ((int64_t)n != c->N) return c->fail(AHFGPU_EINVAL, "%llu keys for %lld resident particles",
                                         (unsigned long long)n, (long long)c->N);
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  int rc = ensure_key_array(c);
  if (rc) return rc;
  c->keys_ok = 0;
  if (n) hipMemcpyAsync(c->sfckey, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, c->st);
  return check_sorted(c);
}

extern "C" int ahfgpu_sfc_keys(ahfgpu_ctx *c, uint64_t *keys_out)
{
  if (!c) return AHFGPU_EINVAL;
  if (c->N <= 0) return c->fail(AHFGPU_EINVAL, "no particles resident");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  int rc = ensure_key_array(c);
  if (rc) return rc;
  c->keys_ok = 0;
  hipLaunchKernelGGL(k_sfc_keys, dim3(nblk(c->N)), dim3(AHF_BLOCK), 0, c->st, c->px, c->py, c->pz, c->N, c->sfckey);
  if (keys_out) hipMemcpyAsync(keys_out, c->sfckey, (size_t)c->N * sizeof(uint64_t), hipMemcpyDeviceToHost, c->st);
  return check_sorted(c);
}

extern "C" int ahfgpu_gather_halos(ahfgpu_ctx *c, const double *centres, const double *gather_rad,
                                   const uint8_t *skip, uint64_t nh)
{
  if (!c || (nh && (!centres || !gather_rad))) return AHFGPU_EINVAL;
  if (!c->keys_ok) return c->fail(AHFGPU_EINVAL, "no SFC keys for the resident particles (ahfgpu_sfc_keys)");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  for (uint64_t h = 0; h < nh; h++)
    if (!(gather_rad[h] >= 0.0)) return c->fail(AHFGPU_EINVAL, "gather radius of halo %llu is %g",
                                                (unsigned long long)h, gather_rad[h]);
  Arena &A = c->uarena;
  A.reset();
  const int64_t nslot = AHF_SHELL * (int64_t)nh;
  double *d_in = A.get<double>(4 * (int64_t)nh);
  uint8_t *d_skip = A.get<uint8_t>((int64_t)nh);
  int64_t *cnt = A.get<int64_t>(nslot), *bound = A.get<int64_t>(2 * nslot), *start = A.get<int64_t>(nslot + 1);
  int64_t *stmp = A.get<int64_t>(scan_tmp_elems(nslot));
  if (!d_in || !d_skip || !cnt || !bound || !start || !stmp)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (gather)");
  double *d_cen = d_in, *d_rad = d_in + 3 * nh;
  c->timers_reset(TIMER_HGATHER, TIMER_HGATHER);
  c->timed.clear();
  c->ev_next = 0;
  hipEvent_t e0 = c->timer_begin();
  if (nh) {
    hipMemcpyAsync(d_cen, centres, 3 * nh * sizeof(double), hipMemcpyHostToDevice, c->st);
    hipMemcpyAsync(d_rad, gather_rad, nh * sizeof(double), hipMemcpyHostToDevice, c->st);
    if (skip) hipMemcpyAsync(d_skip, skip, nh, hipMemcpyHostToDevice, c->st);
    else hipMemsetAsync(d_skip, 0, nh, c->st);
    hipLaunchKernelGGL(k_gather_cells, dim3(nblk(2 * nslot)), dim3(AHF_BLOCK), 0, c->st, d_cen, d_rad, d_skip,
                       nslot, c->sfckey, c->N, bound);
    hipLaunchKernelGGL(k_gather_counts, dim3(nblk(nslot)), dim3(AHF_BLOCK), 0, c->st, bound, nslot, cnt);
  }
  exclusive_scan<int64_t, int64_t>(cnt, start, nslot, stmp, c->st);
  int64_t nc = 0;
  hipMemcpyAsync(&nc, start + nslot, sizeof(int64_t), hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess) return c->fail(AHFGPU_EHIP, "gather cells");
  if (nc >= (int64_t)UINT32_MAX) return c->fail(AHFGPU_EUNSUP, "%lld gather candidates (at most 2^32-1)", (long long)nc);
  uint8_t *flag = A.get<uint8_t>(nc);
  uint32_t *pos = A.get<uint32_t>(nc + 1);
  uint32_t *ftmp = A.get<uint32_t>(scan_tmp_elems(nc));
  int64_t *d_off = A.get<int64_t>((int64_t)nh + 1);
  int32_t *tab = A.get<int32_t>(nc / AHF_CAND_BLOCK + 1);
  if (!flag || !pos || !ftmp || !d_off || !tab) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (gather)");
  if (nc) {
    hipLaunchKernelGGL(k_gather_blockslot, dim3(nblk(nslot)), dim3(AHF_BLOCK), 0, c->st, start, nslot, tab);
    hipLaunchKernelGGL(k_gather_test, dim3(nblk(nc)), dim3(AHF_BLOCK), 0, c->st, start, bound, tab, nc, d_cen, d_rad,
                       c->px, c->py, c->pz, flag);
  }
  exclusive_scan<uint8_t, uint32_t>(flag, pos, nc, ftmp, c->st);
  hipLaunchKernelGGL(k_gather_off, dim3(nblk((int64_t)nh + 1)), dim3(AHF_BLOCK), 0, c->st, start, (int64_t)nh, nc, pos,
                     d_off);
  std::vector<int64_t> off(nh + 1, 0);
  hipMemcpyAsync(off.data(), d_off, (nh + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess) return c->fail(AHFGPU_EHIP, "gather test");
  const int64_t ne = off[nh];
  HaloSet *s = nullptr;
  int rc = halo_set_reserve(c, (int64_t)nh, ne, &s);
  if (rc) return rc;
  halo_set_host_meta(s, off.data(), centres);
  hipMemcpyAsync(s->dev, s->host, s->o_ip, hipMemcpyHostToDevice, c->st);
  if (nc) hipLaunchKernelGGL(k_gather_write, dim3(nblk(nc)), dim3(AHF_BLOCK), 0, c->st, start, bound, tab, nc, flag,
                             pos, s->ip_in);
  /* algorithmic bytes: per candidate a 12-byte position read, per member a
   * 4-byte index written */
  c->timer_end(e0, TIMER_HGATHER, 12.0 * (double)nc + 4.0 * (double)ne);
  if (hipStreamSynchronize(c->st) != hipSuccess) {
    free_halos(s);
    return c->fail(AHFGPU_EHIP, "gather: %s", hipGetErrorString(hipGetLastError()));
  }
  c->timers_collect();
  c->gather_candidates = nc;
  c->halo_set = s;
  return 0;
}

extern "C" int ahfgpu_halo_members(ahfgpu_ctx *c, uint64_t *npart, uint64_t *ipart)
{
  if (!c) return AHFGPU_EINVAL;
  HaloSet *s = (HaloSet *)c->halo_set;
  if (!s) return c->fail(AHFGPU_EINVAL, "no halos resident");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  if (npart)
    for (int64_t h = 0; h < s->nh; h++) npart[h] = (uint64_t)(s->h_off[h + 1] - s->h_off[h]);
  if (ipart && s->ne) {
    std::vector<uint32_t> ip((size_t)s->ne);
    if (hipMemcpy(ip.data(), s->ip_in, (size_t)s->ne * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "member copy");
    for (int64_t e = 0; e < s->ne; e++) ipart[e] = ip[(size_t)e];
  }
  return 0;
}
