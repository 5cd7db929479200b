/*
 * unbind_gpu.hip -- AHF's radial sort + iterative unbinding of all haloes on
 * gfx950 (src/libahf/ahf_halos.c:5786-5894 sort_halo_particles and
 * :3292-3607 rem_unbound, driven per halo as in :504-510).
 *
 * 1. Sort: every member gets the key dist^2 (signed periodic wrap, FP64);
 *    a stable LSD radix sort on the key followed by a stable sort on the
 *    halo index orders each halo's members by distance, equal distances in
 *    input order.  Haloes below NminPerHalo keep their order (AHF does not
 *    sort them).
 * 2. Unbinding: a persistent kernel, one workgroup per halo, haloes taken
 *    largest first from an atomic queue.  Per iteration:
 *      - dist_j, the mid-point terms of the integral of M(<r)/r^2 and their
 *        FP64 prefix sum Phi_j; Phi0 = Phi_{n-1} + n/dist_{n-1} (AHF's two
 *        passes add the same terms in the same order);
 *      - the escape test of member j uses the bulk velocity of the seed and
 *        of every bound member before j: a sequential recurrence.  It is
 *        solved by sweeps that assume a bound set, prefix-sum the velocities
 *        of that set and re-test; the sweeps stop at the fixed point, which
 *        is the sequential answer;
 *      - the parallel sums round differently from AHF's running sums, by at
 *        most gam_n * sum|terms| (any order of summation); carried through
 *        the escape test this bounds, per member, how far the sequential
 *        vel^2 and limit can lie from the parallel ones.  A member whose
 *        margin is inside that bound (or inside `band`, relative, if the
 *        caller asks for more) sends the halo to an exact sequential replay
 *        of the iteration in AHF's own operation order; every other
 *        decision is the sequential one;
 *      - stable compaction of the bound members.
 * 3. rem_outsideRvir (:3687-3889), on request before and after the
 *    unbinding as ahf_halos_sfc.c:143,153 call it: the overdensity of the
 *    sphere through member j depends on j and its distance only, so every
 *    member is tested at once and the cut is the first failure.
 */
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/ahf_gpu.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "halo_set.h"
#include "nr_order.h"

using namespace ahf;

#define MZERO 5e-16 /* MACHINE_ZERO, src/param.h:125 */
#define PI_AHF 3.14159265358979323846264338 /* src/param.h:143 */

struct UnbindConst {
  double v2_tune, v_fac, r_fac, phi_fac, hubble, band;
  double ovlim, rho_vir, rho_fac;
  int32_t minpart, no_vbulk;
  int32_t cut_before, unbind, cut_after, pad;
};

__device__ inline double wrap_s(double d)
{
  if (d > 0.5) d -= 1.0;
  if (d < -0.5) d += 1.0;
  return d;
}

/* ----------------------------------------------------------- sort keys */

__device__ inline int64_t halo_of(const int64_t *off, int64_t nh, int64_t e)
{
  int64_t lo = 0, hi = nh - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

/* sort key of a member: dist^2 from the centre (signed periodic wrap,
 * src/libahf/ahf_halos.c:5827-5840), as FP64 bits; non-negative, so the bit
 * order is the value order.  Positions from the interleaved particle copy
 * (x, y, z in pv[2p]) */
__device__ inline uint64_t dist_key(uint32_t p, const double *__restrict__ c, const float4 *__restrict__ pv)
{
  const float4 a = pv[2 * (size_t)p];
  const double dX = wrap_s((double)a.x - c[0]);
  const double dY = wrap_s((double)a.y - c[1]);
  const double dZ = wrap_s((double)a.z - c[2]);
  const double d2 = dX * dX + dY * dY + dZ * dZ;
  return (uint64_t)__double_as_longlong(d2);
}

#define AHF_SORT_SMALL 8192
#define AHF_SORT_TIERS 5

/* which radial sort a halo of n members takes: 0 keeps the input order (AHF
 * does not sort haloes below minpart), 1..4 sort in LDS with capacities 64,
 * 512, 2048 and AHF_SORT_SMALL, -1 goes to the radix sort */
__host__ __device__ inline int sort_tier(int64_t n, int32_t minpart)
{
  if (n < minpart || n < 2) return 0;
  if (n <= 64) return 1;
  if (n <= 512) return 2;
  if (n <= 2048) return 3;
  if (n <= AHF_SORT_SMALL) return 4;
  return -1;
}

struct SortBase {
  int64_t base[AHF_SORT_TIERS];
};

/* the haloes of each tier lie at base[t] of one list (built on the host;
 * each halo is sorted on its own) */

/*
 * Radial sort of the haloes list[0..cnt), at most CAP members each, one
 * workgroup of BS threads per halo: keys in LDS, bitonic sort on (key,
 * position in the list), so equal distances keep their input order as in the
 * radix path.  With copy_only the lists are copied unsorted.
 */
template <int CAP, int BS>
__global__ void __launch_bounds__(BS) k_sort_tier(const int32_t *__restrict__ list, int64_t cnt,
                                                  const int64_t *__restrict__ off, const uint32_t *__restrict__ ip,
                                                  const double *__restrict__ cen, const float4 *__restrict__ pv,
                                                  int copy_only, uint32_t *__restrict__ out)
{
  __shared__ uint64_t s_key[CAP];
  __shared__ uint32_t s_idx[CAP];
  const int tid = threadIdx.x;
  for (int64_t t = blockIdx.x; t < cnt; t += gridDim.x) {
    const int64_t h = list[t];
    const int64_t o0 = off[h];
    const int64_t nl = off[h + 1] - o0;
    if (copy_only) {
      for (int64_t i = tid; i < nl; i += BS) out[o0 + i] = ip[o0 + i];
      continue;
    }
    const int n = (int)nl;
    int m = 2;
    while (m < n) m <<= 1;
    for (int i = tid; i < m; i += BS) {
      if (i < n) {
        s_key[i] = dist_key(ip[o0 + i], cen + 3 * h, pv);
        s_idx[i] = (uint32_t)i;
      } else {
        s_key[i] = ~0ull; /* above every finite key: padding sorts last */
        s_idx[i] = 0xFFFFFFFFu;
      }
    }
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        /* pair q compares element i (bit j clear) with i + j */
        for (int q = tid; q < (m >> 1); q += BS) {
          const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
          const int l = i + j;
          const uint64_t ka = s_key[i], kb = s_key[l];
          const uint32_t ia = s_idx[i], ib = s_idx[l];
          const bool a_gt_b = ka > kb || (ka == kb && ia > ib);
          if (a_gt_b == ((i & k) == 0)) {
            s_key[i] = kb;
            s_key[l] = ka;
            s_idx[i] = ib;
            s_idx[l] = ia;
          }
        }
        __syncthreads();
      }
    for (int i = tid; i < n; i += BS) out[o0 + i] = ip[o0 + s_idx[i]];
    __syncthreads();
  }
}

__global__ void k_halo_keys(const int64_t *__restrict__ off, int64_t nh, int64_t ne, const uint32_t *__restrict__ v,
                            uint32_t *__restrict__ hk)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) hk[i] = (uint32_t)halo_of(off, nh, (int64_t)v[i]);
}

/* the radix path over the members of the haloes bh[0..nb) only (those too
 * big for the LDS tiers): flat index e in [0, bo[nb]) is member e - bo[k] of
 * halo bh[k] */
__global__ void k_dist_keys_sel(const int64_t *__restrict__ bo, const int32_t *__restrict__ bh, int64_t nb,
                                int64_t ne, const int64_t *__restrict__ off, const uint32_t *__restrict__ ip,
                                const double *__restrict__ cen, const float4 *__restrict__ pv,
                                uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int64_t k = halo_of(bo, nb, e);
  const int64_t h = bh[k];
  val[e] = (uint32_t)e;
  key[e] = dist_key(ip[off[h] + (e - bo[k])], cen + 3 * h, pv);
}

__global__ void k_gather_members_sel(const int64_t *__restrict__ bo, const int32_t *__restrict__ bh, int64_t nb,
                                     int64_t ne, const int64_t *__restrict__ off, const uint32_t *__restrict__ v,
                                     const uint32_t *__restrict__ ip, uint32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const int64_t k = halo_of(bo, nb, i);
  const int64_t base = off[bh[k]] - bo[k];
  out[base + i] = ip[base + (int64_t)v[i]];
}

/* ------------------------------------------------------ block helpers */

struct D4 {
  double x, y, z, m;
};

__device__ inline D4 shfl_up4(D4 a, int d)
{
  D4 r;
  r.x = __shfl_up(a.x, d, 64);
  r.y = __shfl_up(a.y, d, 64);
  r.z = __shfl_up(a.z, d, 64);
  r.m = __shfl_up(a.m, d, 64);
  return r;
}

/* exclusive block scan of D4; total in *tot (all threads) */
__device__ inline D4 block_excl_scan4(D4 v, D4 *tot, D4 *sh /* [AHF_WAVES] */)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  D4 inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    D4 u = shfl_up4(inc, d);
    if (lane >= d) {
      inc.x += u.x; inc.y += u.y; inc.z += u.z; inc.m += u.m;
    }
  }
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  D4 base = {0, 0, 0, 0}, t = {0, 0, 0, 0};
  for (int k = 0; k < AHF_WAVES; k++) {
    const D4 s = sh[k];
    if (k < w) { base.x += s.x; base.y += s.y; base.z += s.z; base.m += s.m; }
    t.x += s.x; t.y += s.y; t.z += s.z; t.m += s.m;
  }
  __syncthreads();
  *tot = t;
  D4 ex;
  ex.x = base.x + inc.x - v.x;
  ex.y = base.y + inc.y - v.y;
  ex.z = base.z + inc.z - v.z;
  ex.m = base.m + inc.m - v.m;
  return ex;
}

/* AHF's indexx on n <= 64 values (1-based), with a 24-entry stack */
__device__ void dev_indexx(int n, const double *arr1, int *indx1, int *stack)
{
  int top = 0, lo = 1, hi = n;
  for (int j = 1; j <= n; j++) indx1[j] = j;
  if (n == 0) return;
  for (;;) {
    if (hi - lo < 7) {
      for (int j = lo + 1; j <= hi; j++) {
        const int it = indx1[j];
        const double a = arr1[it];
        int i = j - 1;
        for (; i >= 1; i--) {
          if (arr1[indx1[i]] <= a) break;
          indx1[i + 1] = indx1[i];
        }
        indx1[i + 1] = it;
      }
      if (top == 0) break;
      hi = stack[top--];
      lo = stack[top--];
    } else {
      const int k = (lo + hi) >> 1;
      int t;
#define SW(a, b) { t = (a); (a) = (b); (b) = t; }
      SW(indx1[k], indx1[lo + 1]);
      if (arr1[indx1[lo + 1]] > arr1[indx1[hi]]) SW(indx1[lo + 1], indx1[hi]);
      if (arr1[indx1[lo]] > arr1[indx1[hi]]) SW(indx1[lo], indx1[hi]);
      if (arr1[indx1[lo + 1]] > arr1[indx1[lo]]) SW(indx1[lo + 1], indx1[lo]);
      int i = lo + 1, j = hi;
      const int it = indx1[lo];
      const double a = arr1[it];
      for (;;) {
        do i++; while (arr1[indx1[i]] < a);
        do j--; while (arr1[indx1[j]] > a);
        if (j < i) break;
        SW(indx1[i], indx1[j]);
      }
#undef SW
      indx1[lo] = indx1[j];
      indx1[j] = it;
      top += 2;
      if (hi - i + 1 >= j - lo) {
        stack[top] = hi;
        stack[top - 1] = i;
        hi = j - 1;
      } else {
        stack[top] = j - 1;
        stack[top - 1] = lo;
        lo = i;
      }
    }
  }
}

/* overdensity of the sphere through member j at particle p, as
 * src/libahf/ahf_halos.c:3832-3856 evaluates it (M = j+1 unit masses);
 * the position from the interleaved particle copy */
__device__ inline double sphere_ovdens(const UnbindConst &K, int64_t j, uint32_t p, const double cen[3],
                                       const float4 *__restrict__ pv, double *R_out)
{
  const float4 a = pv[2 * (size_t)p];
  const double dX = wrap_s((double)a.x - cen[0]), dY = wrap_s((double)a.y - cen[1]), dZ = wrap_s((double)a.z - cen[2]);
  const double R = sqrt(dX * dX + dY * dY + dZ * dZ);
  const double V = 4. * PI_AHF / 3. * (R * R * R);
  *R_out = R;
  return (double)(j + 1) / V * K.rho_fac / K.rho_vir;
}

/*
 * rem_outsideRvir on the first n >= 1 members of list l, one workgroup of BS
 * threads: the members kept are those up to and including the first whose
 * sphere is below ovlim (a NaN overdensity stops the walk too).  Chunks of
 * 4*BS members are tested until one holds a failure.  Returns the count
 * kept; thread 0 receives M_vir, R_vir and ovdens of the last member kept.
 */
template <int BS>
__device__ int64_t rvir_cut(const UnbindConst &K, int64_t n, const uint32_t *l, const double cen[3],
                            const float4 *__restrict__ pv, double *M, double *R, double *od)
{
  constexpr int PER = 4;
  constexpr unsigned long long NONE = ~0ull;
  __shared__ unsigned long long s_first;
  const int t = threadIdx.x;
  if (t == 0) s_first = NONE;
  __syncthreads();
  unsigned long long first = NONE;
  for (int64_t base = 0; base < n; base += (int64_t)PER * BS) {
    uint32_t pp[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int64_t j = base + (int64_t)i * BS + t;
      pp[i] = j < n ? l[j] : 0u;
    }
    unsigned long long f = NONE;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int64_t j = base + (int64_t)i * BS + t;
      if (j >= n) continue;
      double r;
      const double o = sphere_ovdens(K, j, pp[i], cen, pv, &r);
      if (!(o >= K.ovlim) && (unsigned long long)j < f) f = (unsigned long long)j;
    }
    if (f != NONE) atomicMin(&s_first, f);
    __syncthreads();
    first = s_first;
    __syncthreads();
    if (first != NONE) break;
  }
  const int64_t kept = first != NONE ? (int64_t)first + 1 : n;
  if (t == 0) {
    double r;
    *od = sphere_ovdens(K, kept - 1, l[kept - 1], cen, pv, &r);
    *R = r;
    *M = (double)kept;
  }
  return kept;
}

/* exclusive scans over a workgroup of BS threads (exclusive value taken
 * from the neighbouring lane's inclusive value, block total in *tot) */
template <int BS, typename T>
__device__ inline T blk_excl_scan(T v, T *tot, T *sh /* [BS / 64] */)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T u = __shfl_up(inc, d, 64);
    if (lane >= d) inc += u;
  }
  T ex = __shfl_up(inc, 1, 64);
  if (lane == 0) ex = 0;
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  T base = 0, t = 0;
#pragma unroll
  for (int k = 0; k < BS / 64; k++) {
    const T s = sh[k];
    if (k < w) base += s;
    t += s;
  }
  __syncthreads();
  *tot = t;
  return base + ex;
}

/* a double moved across lanes by one DPP pattern (both halves); lanes
 * without a source, or outside the row mask, read 0.0 */
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v)
{
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_add4(D4 &a)
{
  a.x += dpp_d<CTRL, ROWS>(a.x);
  a.y += dpp_d<CTRL, ROWS>(a.y);
  a.z += dpp_d<CTRL, ROWS>(a.z);
  a.m += dpp_d<CTRL, ROWS>(a.m);
}

/* inclusive and exclusive wavefront scans of D4 by DPP row shifts and row
 * broadcasts (the GFX9 scan; register-only): any summation order serves
 * the unbinding, whose decisions carry a certified rounding bound */
__device__ __forceinline__ void wave_scan4_dpp(D4 v, D4 *inc, D4 *exc)
{
  D4 a = v;
  dpp_add4<0x111, 0xf>(a); /* row_shr:1 */
  dpp_add4<0x112, 0xf>(a); /* row_shr:2 */
  dpp_add4<0x114, 0xf>(a); /* row_shr:4 */
  dpp_add4<0x118, 0xf>(a); /* row_shr:8 */
  dpp_add4<0x142, 0xa>(a); /* row_bcast:15 into rows 1, 3 */
  dpp_add4<0x143, 0xc>(a); /* row_bcast:31 into rows 2, 3 */
  *inc = a;
  /* wave_shr:1: the inclusive value of the lane below, 0 on lane 0 */
  *exc = D4{dpp_d<0x138, 0xf>(a.x), dpp_d<0x138, 0xf>(a.y), dpp_d<0x138, 0xf>(a.z), dpp_d<0x138, 0xf>(a.m)};
}

template <int BS>
__device__ inline D4 blk_excl_scan4(D4 v, D4 *tot, D4 *sh /* [BS / 64] */)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  D4 inc, ex;
  wave_scan4_dpp(v, &inc, &ex);
  (void)lane;
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  D4 base = {0, 0, 0, 0}, t = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < BS / 64; k++) {
    const D4 s = sh[k];
    if (k < w) { base.x += s.x; base.y += s.y; base.z += s.z; base.m += s.m; }
    t.x += s.x; t.y += s.y; t.z += s.z; t.m += s.m;
  }
  __syncthreads();
  *tot = t;
  return D4{base.x + ex.x, base.y + ex.y, base.z + ex.z, base.m + ex.m};
}

/* exclusive prefix sums of one halo's tile values v(t), t in [t0, t1), by
 * one workgroup of 256 threads in chunks of 256 with a running carry
 * (start: the value before the first tile); put(t, prefix) receives them;
 * returns start plus the total */
template <class Get, class Put>
__device__ inline D4 blk_scan_tiles4(int32_t t0, int32_t t1, D4 start, Get v, Put put, D4 *sh4)
{
  D4 carry = start;
  for (int32_t b = t0; b < t1; b += 256) {
    const int32_t t = b + (int32_t)threadIdx.x;
    const D4 x = t < t1 ? v(t) : D4{0, 0, 0, 0};
    D4 tot;
    const D4 ex = blk_excl_scan4<256>(x, &tot, sh4);
    if (t < t1) put(t, D4{carry.x + ex.x, carry.y + ex.y, carry.z + ex.z, carry.m + ex.m});
    carry.x += tot.x; carry.y += tot.y; carry.z += tot.z; carry.m += tot.m;
  }
  return carry;
}

/*
 * The unbinding of one halo by one workgroup of BS threads.  a/b: two member
 * buffers at the halo's offset (sorted list in a); dist/phi: FP64 scratch;
 * st: status.  Passes over the members go in chunks of BS*IPT: thread t
 * takes IPT consecutive members, sums them sequentially, and the block scans
 * the thread totals.
 */
/* member data of one iteration, gathered once into the halo's scratch in
 * list order: periodic offsets from the centre (FP64) and velocities */
struct Staged {
  double *dx, *dy, *dz;
  float *vx, *vy, *vz;
};

/* per-halo factors of the certified bound of the bound test (see rec_eval) */
struct RecBound {
  double e_lim, cvx, cvy, cvz;
};

__device__ inline RecBound rec_bound(const UnbindConst &K, double Phi0, D4 asum, double gam)
{
  const double e_phi = 2.5 * gam * fabs(Phi0);
  return RecBound{K.v2_tune * 4.0 * K.phi_fac * e_phi, fabs(K.v_fac) * 5.0 * gam * asum.x,
                  fabs(K.v_fac) * 5.0 * gam * asum.y, fabs(K.v_fac) * 5.0 * gam * asum.z};
}

/*
 * One chunk [base, base + BS*IPT) of the bound-set recurrence, settled: the
 * chunk is evaluated with cv, the bulk-velocity totals of every member
 * before it, and its own statuses of the previous evaluation (bj: the first
 * guess on entry), and evaluated again until they no longer change; then
 * bj holds the answer the sequential loop gives with that carry (the
 * dependence is strictly on earlier members, so any first guess settles).
 * Out: the chunk's bound totals (the same on every thread), whether a test
 * of the final evaluation was uncertain.  flag: two words of LDS, ev the
 * evaluation counter that alternates them.
 */
template <int BS, int IPT>
__device__ void settle_chunk(const UnbindConst &K, int64_t n, int64_t base, const double *dist, const double *phi,
                             const Staged &g, double Phi0, D4 cv, const RecBound &B, int bj[IPT], D4 *tot_out,
                             int *unc_out, int *sweeps, D4 *sh4, int *sflag, int &ev)
{
  const int t = threadIdx.x;
  const int64_t j0 = base + (int64_t)t * IPT;
  /* what a member's test needs besides the running bulk velocity, once per
   * chunk: its velocity, its Hubble term and its limit v2_tune*v_esc2
   * (no limit, lim < 0, inside MACHINE_ZERO) */
  double fx[IPT], fy[IPT], fz[IPT], hx[IPT], hy[IPT], hz[IPT], lm[IPT];
#pragma unroll
  for (int i = 0; i < IPT; i++) {
    const int64_t j = j0 + i;
    fx[i] = fy[i] = fz[i] = hx[i] = hy[i] = hz[i] = 0.0;
    lm[i] = -1.0;
    if (j < n) {
      fx[i] = (double)g.vx[j]; fy[i] = (double)g.vy[j]; fz[i] = (double)g.vz[j];
      hx[i] = K.hubble * (g.dx[j]) * K.r_fac;
      hy[i] = K.hubble * (g.dy[j]) * K.r_fac;
      hz[i] = K.hubble * (g.dz[j]) * K.r_fac;
      if (dist[j] > MZERO) lm[i] = K.v2_tune * (2 * fabs(phi[j] - Phi0) * K.phi_fac);
    }
  }
  for (;;) {
    (*sweeps)++;
    /* the flag word of this evaluation; the other one was read by every
     * thread before the previous evaluation's barrier, and is cleared for
     * the next evaluation once this one's scan has passed its barriers */
    int *flag = sflag + (ev & 1);
    D4 run = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < IPT; i++)
      if (bj[i]) { run.x += fx[i]; run.y += fy[i]; run.z += fz[i]; run.m += 1.0; }
    D4 tot;
    const D4 ex = blk_excl_scan4<BS>(run, &tot, sh4);
    /* the thread's own earlier members, summed again in the same order */
    D4 lex = {0, 0, 0, 0};
    int chg = 0, unc = 0;
#pragma unroll
    for (int i = 0; i < IPT; i++) {
      const int64_t j = j0 + i;
      if (j >= n) continue;
      const double VX = cv.x + ex.x + lex.x, VY = cv.y + ex.y + lex.y, VZ = cv.z + ex.z + lex.z,
                   M = cv.m + ex.m + lex.m;
      if (bj[i]) { lex.x += fx[i]; lex.y += fy[i]; lex.z += fz[i]; lex.m += 1.0; }
      const double dVX = (fx[i] - VX / M) * K.v_fac + hx[i];
      const double dVY = (fy[i] - VY / M) * K.v_fac + hy[i];
      const double dVZ = (fz[i] - VZ / M) * K.v_fac + hz[i];
      const double vel2 = dVX * dVX + dVY * dVY + dVZ * dVZ;
      /* v_esc2 = 1e30 inside MACHINE_ZERO: the limit is never exceeded */
      const double lim = lm[i] >= 0.0 ? lm[i] : K.v2_tune * 1e30;
      const int nb = !(vel2 > lim);
      if (lm[i] >= 0.0) {
        /* certified bound on how far the sequential vel^2 and limit can
         * lie from these: Phi_j and Phi0 are sums of non-negative terms
         * (error <= gam*Phi0 each), the velocity sums are within
         * gam*2*sum|v| (the seed is counted twice); propagated through
         * the test, plus a few ulps per operation */
        const double rM = 1.0 / M;
        const double e_vx = B.cvx * rM, e_vy = B.cvy * rM, e_vz = B.cvz * rM;
        const double e_vel2 = 2.0 * (fabs(dVX) * e_vx + fabs(dVY) * e_vy + fabs(dVZ) * e_vz) + e_vx * e_vx +
                              e_vy * e_vy + e_vz * e_vz;
        const double tol = 4.0 * (e_vel2 + B.e_lim) + 1e-14 * (vel2 + lim);
        const double gap = fabs(vel2 - lim);
        if (gap <= tol || gap <= K.band * lim) unc = 1;
      }
      if (nb != bj[i]) { bj[i] = nb; chg = 1; }
    }
    if (t == 0) sflag[(ev + 1) & 1] = 0;
    if (chg | unc) atomicOr(flag, chg | (unc << 1));
    __syncthreads();
    const int f = *flag;
    ev++;
    if (!(f & 1)) {
      /* the statuses used were the final ones: so are the totals */
      *unc_out = f >> 1;
      *tot_out = tot;
      return;
    }
  }
}

/*
 * One iteration's bound-set recurrence of a halo of n members (one
 * workgroup of BS threads), from its staged data (dist, Phi_j in phi, the
 * offsets and velocities in g), Phi0, the sums of |v| (asum) and the
 * rounding factor gam of those parallel sums: st[j] = 1 for the members that
 * stay bound, as AHF's sequential loop decides them.  seed is the list
 * position of the bulk-velocity seed.  A member whose test the certified
 * rounding bound cannot settle sends the iteration to the exact sequential
 * replay (thread 0, in AHF's operation order), which also recomputes Phi0.
 * Returns 1 when it replayed.
 */
/* where rec_eval starts when earlier members are settled already: member
 * j (a chunk boundary of the caller's), the totals before it, and whether
 * a settled test was uncertain */
struct RecResume {
  int64_t j;
  D4 cv;
  int unc;
};

template <int BS, int IPT>
__device__ int rec_eval(const UnbindConst &K, int64_t n, int64_t seed, const double *dist, const double *phi,
                        uint8_t *st, Staged g, D4 asum, double gam, double *Phi0_io, int *sweeps,
                        const RecResume *resume = nullptr)
{
  constexpr int NW = BS / 64;
  constexpr int64_t CH = (int64_t)BS * IPT;
  __shared__ D4 r_sh4[NW];
  __shared__ int sflag[2]; /* double-buffered: bit 0 a status changed, bit 1 a test was uncertain */
  __shared__ double s_phi0;
  D4 *sh4 = r_sh4;
  const int t = threadIdx.x;
  if (t == 0) sflag[0] = sflag[1] = 0;
  __syncthreads();
  int ev = 0;
  double Phi0 = *Phi0_io;
  const double sx = (double)g.vx[seed], sy = (double)g.vy[seed], sz = (double)g.vz[seed];
  /* the sequential bound-set recurrence, chunk by chunk: a chunk is
   * evaluated with the final statuses of every earlier chunk (carried in
   * cv) and its own statuses of the previous evaluation, and evaluated
   * again until they no longer change; then it holds the sequential answer
   * and its totals carry to the next chunk.  Only the final evaluation's
   * margins count for the exact replay. */
  int uncertain = resume ? resume->unc : 0;
  D4 cv = resume ? resume->cv : D4{sx, sy, sz, 1.0};
  const RecBound B = rec_bound(K, Phi0, asum, gam);
  for (int64_t base = resume ? resume->j : 0; base < n; base += CH) {
    const int64_t j0 = base + (int64_t)t * IPT;
    /* statuses of the thread's members, kept in registers while the chunk
     * settles; the first evaluation assumes all bound */
    int bj[IPT];
#pragma unroll
    for (int i = 0; i < IPT; i++) bj[i] = j0 + i < n ? 1 : 0;
    D4 tot;
    int unc = 0;
    settle_chunk<BS, IPT>(K, n, base, dist, phi, g, Phi0, cv, B, bj, &tot, &unc, sweeps, sh4, sflag, ev);
    uncertain |= unc;
    cv.x += tot.x; cv.y += tot.y; cv.z += tot.z; cv.m += tot.m;
#pragma unroll
    for (int i = 0; i < IPT; i++)
      if (j0 + i < n) st[j0 + i] = (uint8_t)bj[i];
  }
  __syncthreads();
  /* exact replay of this iteration where rounding could decide (thread 0) */
  int replayed = 0;
  if (uncertain) {
    replayed = 1;
    if (t == 0) {
      double I_now = 0, I_prev = 0, d_prev = 0, M_r = 0, P0 = 0, dd = 0;
      for (int64_t j = 0; j < n; j++) {
        M_r += 1.0;
        dd = dist[j];
        if (dd > MZERO) {
          I_now = M_r / (dd * dd);
          const double I_mid = (I_now + I_prev) / 2.;
          P0 += I_mid * (dd - d_prev);
        }
        d_prev = dd;
        I_prev = I_now;
      }
      P0 += M_r / dd;
      s_phi0 = P0;
      double VX = sx, VY = sy, VZ = sz, M = 1.0, Phi = 0;
      I_now = 0; I_prev = 0; d_prev = 0; M_r = 0;
      for (int64_t j = 0; j < n; j++) {
        M_r += 1.0;
        const double dX = g.dx[j], dY = g.dy[j], dZ = g.dz[j];
        dd = sqrt(dX * dX + dY * dY + dZ * dZ);
        double v_esc2;
        if (dd > MZERO) {
          I_now = M_r / (dd * dd);
          const double I_mid = (I_now + I_prev) / 2.;
          Phi += I_mid * (dd - d_prev);
          v_esc2 = (2 * fabs(Phi - P0) * K.phi_fac);
        } else {
          v_esc2 = 1e30;
        }
        const double fx = (double)g.vx[j], fy = (double)g.vy[j], fz = (double)g.vz[j];
        const double dVX = (fx - VX / M) * K.v_fac + K.hubble * (dX) * K.r_fac;
        const double dVY = (fy - VY / M) * K.v_fac + K.hubble * (dY) * K.r_fac;
        const double dVZ = (fz - VZ / M) * K.v_fac + K.hubble * (dZ) * K.r_fac;
        const double vel2 = dVX * dVX + dVY * dVY + dVZ * dVZ;
        if (vel2 > K.v2_tune * v_esc2) {
          st[j] = 0;
        } else {
          st[j] = 1;
          M += 1.0;
          VX += 1.0 * g.vx[j];
          VY += 1.0 * g.vy[j];
          VZ += 1.0 * g.vz[j];
        }
        I_prev = I_now;
        d_prev = dd;
      }
    }
    __syncthreads();
    Phi0 = s_phi0;
  }
  *Phi0_io = Phi0;
  __syncthreads();
  return replayed;
}

template <int BS, int IPT>
__device__ void unbind_halo(const UnbindConst &K, int64_t n0, int64_t seed0, uint32_t *a, uint32_t *b,
                            double *dist, double *phi, uint8_t *st, Staged g, const double cen[3],
                            const float4 *__restrict__ pv, HaloOut *out)
{
  constexpr int NW = BS / 64;
  constexpr int64_t CH = (int64_t)BS * IPT;
  __shared__ D4 sh4[NW];
  __shared__ double shd[NW];
  __shared__ int64_t shi[NW];
  __shared__ double sPhi0;
  __shared__ unsigned long long s_last; /* 1 + index of the last bound member, 0 = none */
  __shared__ double s_rvir;
  const int t = threadIdx.x;
  int64_t n = n0;
  uint32_t *cur = a, *nxt = b;
  int niter = 0, replays = 0, in_b = 0, rvir_set = 0, sweeps = 0, set = 0;
  double R_vir = 0.0, Phi0 = 0.0, M_vir = 0.0, ovdens = 0.0;
  int64_t pit = 0;
  /* rem_outsideRvir(halo, 0): returns below minpart, else cuts and clears
   * Phi0 (:3696, :3864-3871) */
  if (K.cut_before && n >= K.minpart) {
    n = rvir_cut<BS>(K, n, cur, cen, pv, &M_vir, &R_vir, &ovdens);
    Phi0 = 0.0;
    set = AHFGPU_SET_MVIR | AHFGPU_SET_RVIR | AHFGPU_SET_PHI0 | AHFGPU_SET_OVDENS;
  }
  /* rem_unbound returns below minpart (:3312) */
  const bool unbind = K.unbind && n >= K.minpart;
  if (unbind) {
    /* its R_vir is a local assigned by the first bound member */
    R_vir = 0.0;
    set = (set | AHFGPU_SET_MVIR | AHFGPU_SET_PHI0) & ~AHFGPU_SET_RVIR;
  }
  int64_t nremove = unbind ? 4 : 0;
  while (nremove > 3) {
    niter++;
    pit += n;
    /* pass 1: distances, and the sums of |v| that bound the rounding of the
     * bulk-velocity sums */
    D4 av = {0, 0, 0, 0};
    /* eight members per thread at a time: their particle gathers are
     * independent and overlap */
    for (int64_t base = 0; base < n; base += 8 * (int64_t)BS) {
      uint32_t pp[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int64_t j = base + (int64_t)i * BS + t;
        pp[i] = j < n ? cur[j] : 0u;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int64_t j = base + (int64_t)i * BS + t;
        if (j >= n) continue;
        const uint32_t p = pp[i];
        const float4 q0 = pv[2 * (size_t)p], q1 = pv[2 * (size_t)p + 1];
        const double dX = wrap_s((double)q0.x - cen[0]), dY = wrap_s((double)q0.y - cen[1]),
                     dZ = wrap_s((double)q0.z - cen[2]);
        const float ux = q0.w, uy = q1.x, uz = q1.y;
        dist[j] = sqrt(dX * dX + dY * dY + dZ * dZ);
        g.dx[j] = dX;
        g.dy[j] = dY;
        g.dz[j] = dZ;
        g.vx[j] = ux;
        g.vy[j] = uy;
        g.vz[j] = uz;
        av.x += fabs((double)ux);
        av.y += fabs((double)uy);
        av.z += fabs((double)uz);
      }
    }
    D4 asum;
    (void)blk_excl_scan4<BS>(av, &asum, sh4);
    /* any order of summing m terms is within gam*sum|terms| of the exact sum
     * (gam = m u / (1 - m u), u = 2^-53); a few terms more cover the adds
     * around the scans and the seed counted twice */
    const double Nd = (double)n + 8.0, U = 1.1102230246251565e-16;
    const double gam = Nd * U / (1.0 - Nd * U);
    __syncthreads();
    /* pass 2: mid-point terms and their prefix sum Phi_j */
    double carry = 0.0;
    for (int64_t base = 0; base < n; base += CH) {
      const int64_t j0 = base + (int64_t)t * IPT;
      double loc[IPT];
      double run = 0.0;
#pragma unroll
      for (int i = 0; i < IPT; i++) {
        const int64_t j = j0 + i;
        double term = 0.0;
        if (j < n) {
          const double dj = dist[j];
          if (dj > MZERO) {
            const double I_now = (double)(j + 1) / (dj * dj);
            double I_prev = 0.0, d_prev = 0.0;
            if (j > 0) {
              d_prev = dist[j - 1];
              if (d_prev > MZERO) I_prev = (double)j / (d_prev * d_prev);
            }
            const double I_mid = (I_now + I_prev) / 2.;
            term = I_mid * (dj - d_prev);
          }
        }
        run += term;
        loc[i] = run;
      }
      double tot;
      const double ex = blk_excl_scan<BS, double>(run, &tot, shd);
#pragma unroll
      for (int i = 0; i < IPT; i++)
        if (j0 + i < n) phi[j0 + i] = carry + ex + loc[i];
      carry += tot;
    }
    __syncthreads();
    if (t == 0) sPhi0 = phi[n - 1] + (double)n / dist[n - 1];
    __syncthreads();
    Phi0 = sPhi0;
    replays += rec_eval<BS, IPT>(K, n, niter == 1 ? seed0 : 0, dist, phi, st, g, asum, gam, &Phi0, &sweeps);
    /* stable compaction into the other buffer; R_vir = distance of the last
     * bound member */
    if (t == 0) s_last = 0ull;
    __syncthreads();
    int64_t w0 = 0;
    unsigned long long last = 0ull;
    for (int64_t base = 0; base < n; base += CH) {
      const int64_t j0 = base + (int64_t)t * IPT;
      int64_t cnt = 0;
#pragma unroll
      for (int i = 0; i < IPT; i++)
        if (j0 + i < n && st[j0 + i]) cnt++;
      int64_t tot;
      const int64_t ex = blk_excl_scan<BS, int64_t>(cnt, &tot, shi);
      int64_t w = w0 + ex;
#pragma unroll
      for (int i = 0; i < IPT; i++) {
        const int64_t j = j0 + i;
        if (j < n && st[j]) {
          nxt[w++] = cur[j];
          last = (unsigned long long)(j + 1);
        }
      }
      w0 += tot;
    }
    if (last) atomicMax(&s_last, last);
    __syncthreads();
    if (t == 0 && s_last > 0) s_rvir = dist[s_last - 1];
    __syncthreads();
    if (s_last > 0) {
      R_vir = s_rvir;
      rvir_set = 1;
    }
    nremove = n - w0;
    n = w0;
    M_vir = (double)w0;
    uint32_t *tmp = cur;
    cur = nxt;
    nxt = tmp;
    in_b ^= 1;
    __syncthreads();
    if ((double)n < (double)K.minpart) break;
  }
  if (rvir_set) set |= AHFGPU_SET_RVIR;
  /* rem_outsideRvir(halo, 1) on the bound list */
  if (K.cut_after && n >= K.minpart) {
    n = rvir_cut<BS>(K, n, cur, cen, pv, &M_vir, &R_vir, &ovdens);
    set |= AHFGPU_SET_MVIR | AHFGPU_SET_RVIR | AHFGPU_SET_OVDENS;
  }
  if (t == 0) {
    out->npart = n;
    out->M_vir = M_vir;
    out->R_vir = R_vir;
    out->Phi0 = Phi0;
    out->ovdens = ovdens;
    out->niter = niter;
    out->replays = replays;
    out->in_b = in_b;
    out->set = set;
    out->particle_iterations = pit;
    out->sweeps = sweeps;
  }
}

/* the first iteration's seed of every halo (:3441-3470): the median |v|^2
 * (FP32) of the first no_vbulk members of the sorted list, by AHF's indexx;
 * one thread per halo */
/* one wavefront per halo (workgroups stride over the haloes): lane j
 * loads member j's |v|^2, lane 0 runs indexx on them in LDS (no_vbulk <= 64
 * values) */
#define AHF_SEED_BS 64
__global__ void __launch_bounds__(AHF_SEED_BS) k_seed(UnbindConst K, const int64_t *__restrict__ off, int64_t nh,
                                                      const uint32_t *__restrict__ a, const float4 *__restrict__ pv,
                                                      int64_t *__restrict__ seed)
{
  __shared__ double s_m2[65];
  __shared__ int s_ix[65], s_stack[24];
  const int nv = K.no_vbulk;
  for (int64_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const int64_t o = off[h], n = off[h + 1] - o;
    if (n < K.minpart || nv < 1) {
      if (threadIdx.x == 0) seed[h] = 0;
      continue;
    }
    if ((int)threadIdx.x < nv) {
      const uint32_t p = a[o + threadIdx.x];
      const float4 q0 = pv[2 * (size_t)p], q1 = pv[2 * (size_t)p + 1];
      const float ux = q0.w, uy = q1.x, uz = q1.y;
      const float v2 = ux * ux + uy * uy + uz * uz;
      s_m2[threadIdx.x + 1] = (double)v2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      dev_indexx(nv, s_m2, s_ix, s_stack);
      seed[h] = s_ix[nv / 2] - 1;
    }
    __syncthreads();
  }
}
static inline unsigned seed_blocks(int64_t nh)
{
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nh, (int64_t)1 << 20));
}

/* haloes of at least this many members go to 512-thread workgroups */
#define AHF_BIG_HALO 2048
#define AHF_WAVE_HALO 512 /* haloes below this many members: one wavefront each */
#define AHF_BIG_BLOCK 512
#define AHF_UNBIND_IPT 4

/* persistent form: workgroups take haloes order[q0..q1) from the queue
 * (largest first) */
template <int BS>
__device__ void unbind_queue(const UnbindConst &K, const int64_t *__restrict__ off, const int32_t *__restrict__ order,
                             int64_t q0, int64_t q1, int32_t *__restrict__ next, const int64_t *__restrict__ seed,
                             uint32_t *a, uint32_t *b, double *dist, double *phi, uint8_t *st, Staged g,
                             const double *__restrict__ cen, const float4 *__restrict__ pv, HaloOut *__restrict__ out)
{
  __shared__ int64_t sh_h;
  for (;;) {
    if (threadIdx.x == 0) {
      const int64_t q = q0 + atomicAdd(next, 1);
      sh_h = q < q1 ? order[q] : -1;
    }
    __syncthreads();
    const int64_t h = sh_h;
    __syncthreads();
    if (h < 0) return;
    const int64_t o = off[h], n = off[h + 1] - o;
    const double c3[3] = {cen[3 * h], cen[3 * h + 1], cen[3 * h + 2]};
    const Staged go = {g.dx + o, g.dy + o, g.dz + o, g.vx + o, g.vy + o, g.vz + o};
    unbind_halo<BS, AHF_UNBIND_IPT>(K, n, seed[h], a + o, b + o, dist + o, phi + o, st + o, go, c3, pv, out + h);
    __syncthreads();
  }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_unbind_batched(UnbindConst K, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ order, int64_t q0, int64_t q1,
                                                       int32_t *__restrict__ next, const int64_t *__restrict__ seed,
                                                       uint32_t *a, uint32_t *b, double *dist, double *phi,
                                                       uint8_t *st, Staged g,
                                                       const double *__restrict__ cen, const float4 *__restrict__ pv, HaloOut *__restrict__ out)
{
  unbind_queue<BS>(K, off, order, q0, q1, next, seed, a, b, dist, phi, st, g, cen, pv, out);
}

/* the same with the queue range read from the device: order[qb[tier] ..
 * qb[tier + 1]) */
template <int BS>
__global__ void __launch_bounds__(BS) k_unbind_batched_dq(UnbindConst K, const int64_t *__restrict__ off,
                                                          const int32_t *__restrict__ order,
                                                          const int64_t *__restrict__ qb, int tier,
                                                          int32_t *__restrict__ next, const int64_t *__restrict__ seed,
                                                          uint32_t *a, uint32_t *b, double *dist, double *phi,
                                                          uint8_t *st, Staged g, const double *__restrict__ cen,
                                                          const float4 *__restrict__ pv,
                                                          HaloOut *__restrict__ out)
{
  unbind_queue<BS>(K, off, order, qb[tier], qb[tier + 1], next, seed, a, b, dist, phi, st, g, cen, pv, out);
}

/* plain form: one halo per launch */
template <int BS>
__global__ void __launch_bounds__(BS) k_unbind_one(UnbindConst K, int64_t h, const int64_t *__restrict__ off,
                                                   const int64_t *__restrict__ seed, uint32_t *a, uint32_t *b,
                                                   double *dist, double *phi,
                                                   uint8_t *st, Staged g, const double *__restrict__ cen,
                                                   const float4 *__restrict__ pv,
                                                   HaloOut *__restrict__ out)
{
  const int64_t o = off[h], n = off[h + 1] - o;
  const double c3[3] = {cen[3 * h], cen[3 * h + 1], cen[3 * h + 2]};
  const Staged go = {g.dx + o, g.dy + o, g.dz + o, g.vx + o, g.vy + o, g.vz + o};
  unbind_halo<BS, AHF_UNBIND_IPT>(K, n, seed[h], a + o, b + o, dist + o, phi + o, st + o, go, c3, pv, out + h);
}

__global__ void k_sorted_out(const int64_t *__restrict__ off, int64_t nh, HaloOut *__restrict__ out)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  HaloOut o;
  o.npart = off[h + 1] - off[h];
  o.M_vir = 0.0;
  o.R_vir = 0.0;
  o.Phi0 = 0.0;
  o.ovdens = 0.0;
  o.niter = 0;
  o.replays = 0;
  o.in_b = 0;
  o.set = 0;
  o.particle_iterations = 0;
  o.sweeps = 0;
  o.pad = 0;
  out[h] = o;
}

__global__ void k_copy_u32(const uint32_t *__restrict__ in, int64_t n, uint32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}


/* ------------------------------------------------------- split path */

/*
 * Haloes of at least split_min members.  On one workgroup their passes over
 * the members (staging, the Phi prefix, compaction) are bound by what one CU
 * can fetch; here they run as flat kernels over tiles of all such haloes at
 * once, and only the recurrence (rec_eval) keeps one workgroup per halo.
 * The Phi sums come out in another association than the persistent
 * kernel's; the certified bound of rec_eval holds for any order.
 */
#define AHF_SPLIT_BLOCK 512
#define AHF_SPLIT_DEFAULT 16384 /* default split_min */
#define AHF_TILE 1024 /* members per tile: 256 threads x 4 */

struct SplitHalo {
  int64_t h, o, n, pit, last, w, seed;
  double M_vir, R_vir, Phi0, ovdens;
  double asum[3], phitot;
  int32_t niter, replays, in_b, set, rvir_set, sweeps, done, active;
  /* the speculative passes: tiles [first tile, f_tile) are settled, the
   * halo's tiles end at e_tile; where the sequential walk resumes */
  int32_t f_tile, e_tile;
  int64_t rs_j;
  double rs_cv[4];
  int32_t rs_unc, rs_pad;
};

/* tile t covers members [j0, j0 + AHF_TILE) of split halo `seg` */
struct SplitTile {
  int32_t seg, first;  /* first: index of the halo's first tile */
  int64_t j0;
};

__device__ inline const uint32_t *split_cur(const SplitHalo &H, const uint32_t *a, const uint32_t *b)
{
  return H.in_b ? b : a;
}

/* pass 1 of an iteration: offsets, velocities, distance of every member */
__global__ void __launch_bounds__(256) k_sp_stage(const SplitHalo *__restrict__ sh, const SplitTile *__restrict__ tiles,
                                                  const int32_t *__restrict__ ntp,
                                                  const uint32_t *a, const uint32_t *b,
                                                  const double *__restrict__ cen, const float4 *__restrict__ pv,
                                                  double *__restrict__ dist, Staged g, uint8_t *__restrict__ st)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  const uint32_t *cur = split_cur(H, a, b);
  const double *c = cen + 3 * H.h;
  /* a member's staged values do not change between iterations (the centre
   * is fixed): after the first, the compaction has carried them along */
  const bool carried = H.niter > 1;
  for (int i = 0; i < AHF_TILE / 256; i++) {
    const int64_t j = T.j0 + (int64_t)i * 256 + threadIdx.x;
    if (j >= H.n) break;
    const int64_t e = H.o + j;
    st[e] = 1;
    if (carried) continue;
    const uint32_t p = cur[e];
    const float4 q0 = pv[2 * (size_t)p], q1 = pv[2 * (size_t)p + 1];
    const double dX = wrap_s((double)q0.x - c[0]), dY = wrap_s((double)q0.y - c[1]), dZ = wrap_s((double)q0.z - c[2]);
    dist[e] = sqrt(dX * dX + dY * dY + dZ * dZ);
    g.dx[e] = dX;
    g.dy[e] = dY;
    g.dz[e] = dZ;
    g.vx[e] = q0.w;
    g.vy[e] = q1.x;
    g.vz[e] = q1.y;
  }
}

/* the mid-point term of member j (pass 2 of unbind_halo) and its |v| */
__device__ inline D4 split_term(const double *dist, const Staged &g, int64_t e, int64_t j)
{
  D4 r = {0.0, fabs((double)g.vx[e]), fabs((double)g.vy[e]), fabs((double)g.vz[e])};
  const double dj = dist[e];
  if (dj > MZERO) {
    const double I_now = (double)(j + 1) / (dj * dj);
    double I_prev = 0.0, d_prev = 0.0;
    if (j > 0) {
      d_prev = dist[e - 1];
      if (d_prev > MZERO) I_prev = (double)j / (d_prev * d_prev);
    }
    const double I_mid = (I_now + I_prev) / 2.;
    r.x = I_mid * (dj - d_prev);
  }
  return r;
}

/* the tile's inclusive prefix of (term, |vx|, |vy|, |vz|) into phi (term
 * part) and its total into ttot */
__global__ void __launch_bounds__(256) k_sp_phi_tiles(const SplitHalo *__restrict__ sh,
                                                      const SplitTile *__restrict__ tiles,
                                                  const int32_t *__restrict__ ntp,
                                                      const double *__restrict__ dist, Staged g,
                                                      double *__restrict__ phi, D4 *__restrict__ ttot)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  __shared__ D4 sh4[4];
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  const int64_t j0 = T.j0 + (int64_t)threadIdx.x * 4;
  D4 run = {0, 0, 0, 0};
  double loc[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int64_t j = j0 + i;
    if (j < H.n) {
      const D4 r = split_term(dist, g, H.o + j, j);
      run.x += r.x; run.y += r.y; run.z += r.z; run.m += r.m;
    }
    loc[i] = run.x;
  }
  D4 tot;
  const D4 ex = blk_excl_scan4<256>(run, &tot, sh4);
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (j0 + i < H.n) phi[H.o + j0 + i] = ex.x + loc[i];
  if (threadIdx.x == 0) ttot[blockIdx.x] = tot;
}

/* per split halo (one workgroup each): the carry into each tile (exclusive
 * sum of the tiles before it) and the halo's totals */
__global__ void __launch_bounds__(256) k_sp_phi_carry(SplitHalo *sh, const int32_t *__restrict__ seg_tile0,
                                                      const D4 *__restrict__ ttot, D4 *__restrict__ tcarry)
{
  __shared__ D4 sh4[4];
  const int64_t i = blockIdx.x;
  if (!sh[i].active) return;
  const D4 acc = blk_scan_tiles4(seg_tile0[i], seg_tile0[i + 1], D4{0, 0, 0, 0}, [&](int32_t t) { return ttot[t]; },
                                 [&](int32_t t, D4 x) { tcarry[t] = x; }, sh4);
  if (threadIdx.x == 0) {
    sh[i].phitot = acc.x;
    sh[i].asum[0] = acc.y;
    sh[i].asum[1] = acc.z;
    sh[i].asum[2] = acc.m;
  }
}

__global__ void __launch_bounds__(256) k_sp_phi_add(const SplitHalo *__restrict__ sh, const SplitTile *__restrict__ tiles,
                                                  const int32_t *__restrict__ ntp,
                                                    const D4 *__restrict__ tcarry, double *__restrict__ phi)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  const double cx = tcarry[blockIdx.x].x;
  for (int i = 0; i < AHF_TILE / 256; i++) {
    const int64_t j = T.j0 + (int64_t)i * 256 + threadIdx.x;
    if (j < H.n) phi[H.o + j] = cx + phi[H.o + j];
  }
}

/*
 * The speculation of the recurrence.  Every unsettled tile of an iteration
 * is settled at once (settle_chunk), each with the carry of the tiles
 * before it under the statuses of the previous pass (all bound in the
 * first), and the pass is repeated with the statuses it produced.  Per
 * halo, the tiles before f_tile are settled: they hold the sequential
 * answer.  In a pass, tile f_tile's carry comes from settled tiles only, so
 * it settles to the sequential answer; so does every later tile up to the
 * first one whose statuses changed in the pass, since the tiles between
 * kept the statuses its carry assumed.  f_tile moves past that tile, or to
 * the end when no tile changed.  The sequential walk (k_sp_rec) resumes at
 * f_tile with the carry of the settled tiles.  The carries are sums in
 * another association, which the certified bound of the test covers.
 */
__global__ void k_sp_spec_init(SplitHalo *sh, int64_t nseg, const int32_t *__restrict__ seg_tile0)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg || !sh[i].active) return;
  sh[i].f_tile = seg_tile0[i];
  sh[i].e_tile = seg_tile0[i + 1];
  sh[i].rs_unc = 0;
}

__device__ inline bool split_settled(const SplitHalo &H) { return H.f_tile >= H.e_tile; }

/* each tile's bound totals under the current statuses */
__global__ void __launch_bounds__(256) k_sp_spec_tot(const SplitHalo *__restrict__ sh,
                                                     const SplitTile *__restrict__ tiles,
                                                     const int32_t *__restrict__ ntp, const uint8_t *__restrict__ st,
                                                     Staged g, D4 *__restrict__ ttot)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  __shared__ D4 sh4[4];
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  if (split_settled(H)) return;
  const int64_t j0 = T.j0 + (int64_t)threadIdx.x * 4;
  D4 run = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int64_t e = H.o + j0 + i;
    if (j0 + i < H.n && st[e]) {
      run.x += (double)g.vx[e]; run.y += (double)g.vy[e]; run.z += (double)g.vz[e]; run.m += 1.0;
    }
  }
  D4 tot;
  (void)blk_excl_scan4<256>(run, &tot, sh4);
  if (threadIdx.x == 0) ttot[blockIdx.x] = tot;
}

/* the seed's velocity, counted once, starts every halo's bulk-velocity sums */
__device__ inline D4 split_seed_cv(const SplitHalo &H, const Staged &g)
{
  const int64_t e = H.o + (H.niter == 1 ? H.seed : 0);
  return D4{(double)g.vx[e], (double)g.vy[e], (double)g.vz[e], 1.0};
}

/* per split halo (one workgroup each): the carry into each tile */
__global__ void __launch_bounds__(256) k_sp_spec_carry(const SplitHalo *__restrict__ sh,
                                                       const int32_t *__restrict__ seg_tile0,
                                                       const D4 *__restrict__ ttot, Staged g, D4 *__restrict__ tcv)
{
  __shared__ D4 sh4[4];
  const SplitHalo &H = sh[blockIdx.x];
  if (!H.active || split_settled(H)) return;
  (void)blk_scan_tiles4(seg_tile0[blockIdx.x], H.e_tile, split_seed_cv(H, g), [&](int32_t t) { return ttot[t]; },
                        [&](int32_t t, D4 x) { tcv[t] = x; }, sh4);
}

/* one pass: every unsettled tile settled with its carry; tchg: its
 * statuses changed, tunc: a test of its final evaluation was uncertain */
__global__ void __launch_bounds__(256) k_sp_spec_eval(UnbindConst K, const SplitHalo *__restrict__ sh,
                                                      const SplitTile *__restrict__ tiles,
                                                      const int32_t *__restrict__ ntp, const double *__restrict__ dist,
                                                      const double *__restrict__ phi, uint8_t *__restrict__ st,
                                                      Staged g, const D4 *__restrict__ tcv,
                                                      uint8_t *__restrict__ tchg, uint8_t *__restrict__ tunc)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  __shared__ D4 sh4[4];
  __shared__ int sflag[2];
  __shared__ int s_chg;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  if ((int32_t)blockIdx.x < H.f_tile || split_settled(H)) return;
  const int64_t o = H.o, n = H.n;
  const Staged go = {g.dx + o, g.dy + o, g.dz + o, g.vx + o, g.vy + o, g.vz + o};
  const double Phi0 = phi[o + n - 1] + (double)n / dist[o + n - 1];
  const double Nd = (double)n + 8.0, U = 1.1102230246251565e-16;
  const double gam = Nd * U / (1.0 - Nd * U);
  const RecBound B = rec_bound(K, Phi0, D4{H.asum[0], H.asum[1], H.asum[2], 0.0}, gam);
  if (threadIdx.x == 0) {
    sflag[0] = sflag[1] = 0;
    s_chg = 0;
  }
  __syncthreads();
  const int64_t j0 = T.j0 + (int64_t)threadIdx.x * 4;
  int bj[4], b0[4];
#pragma unroll
  for (int i = 0; i < 4; i++) bj[i] = b0[i] = j0 + i < n ? (int)st[o + j0 + i] : 0;
  D4 tot;
  int unc = 0, sweeps = 0, ev = 0;
  settle_chunk<256, 4>(K, n, T.j0, dist + o, phi + o, go, Phi0, tcv[blockIdx.x], B, bj, &tot, &unc, &sweeps, sh4,
                       sflag, ev);
  int chg = 0;
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (j0 + i < n) {
      chg |= bj[i] != b0[i];
      st[o + j0 + i] = (uint8_t)bj[i];
    }
  if (chg) atomicOr(&s_chg, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    tchg[blockIdx.x] = (uint8_t)s_chg;
    tunc[blockIdx.x] = (uint8_t)unc;
  }
}

/* after a pass, per split halo (one workgroup each): the tiles up to the
 * first one that changed are settled (their final evaluations' uncertain
 * flags are kept) */
__global__ void __launch_bounds__(256) k_sp_spec_advance(SplitHalo *sh, const uint8_t *__restrict__ tchg,
                                                         const uint8_t *__restrict__ tunc)
{
  __shared__ int s_first, s_unc;
  SplitHalo &H = sh[blockIdx.x];
  if (!H.active || split_settled(H)) return;
  const int32_t f = H.f_tile, e = H.e_tile;
  if (threadIdx.x == 0) {
    s_first = e;
    s_unc = 0;
  }
  __syncthreads();
  for (int32_t t = f + (int32_t)threadIdx.x; t < e; t += 256)
    if (tchg[t]) atomicMin(&s_first, t);
  __syncthreads();
  const int32_t last = s_first < e ? s_first : e - 1; /* the settled tiles are [f, last] */
  int u = 0;
  for (int32_t t = f + (int32_t)threadIdx.x; t <= last; t += 256) u |= tunc[t];
  if (u) atomicOr(&s_unc, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    H.f_tile = last + 1;
    H.rs_unc |= s_unc;
  }
}

/* after the passes, per split halo (one workgroup each): where the
 * sequential walk resumes and the carry of the settled tiles (ttot: their
 * final totals) */
__global__ void __launch_bounds__(256) k_sp_spec_resume(SplitHalo *sh, const int32_t *__restrict__ seg_tile0,
                                                        const SplitTile *__restrict__ tiles,
                                                        const D4 *__restrict__ ttot, Staged g)
{
  __shared__ D4 sh4[4];
  SplitHalo &H = sh[blockIdx.x];
  if (!H.active) return;
  const bool settled = split_settled(H);
  const D4 acc = blk_scan_tiles4(seg_tile0[blockIdx.x], settled ? seg_tile0[blockIdx.x] : H.f_tile,
                                 split_seed_cv(H, g), [&](int32_t t) { return ttot[t]; }, [&](int32_t, D4) {}, sh4);
  if (threadIdx.x == 0) {
    H.rs_j = settled ? H.n : tiles[H.f_tile].j0;
    H.rs_cv[0] = acc.x;
    H.rs_cv[1] = acc.y;
    H.rs_cv[2] = acc.z;
    H.rs_cv[3] = acc.m;
  }
}

/* the recurrence of one iteration, one workgroup per active split halo;
 * with spec, from where the speculative passes leave it */
__global__ void __launch_bounds__(AHF_SPLIT_BLOCK) k_sp_rec(UnbindConst K, SplitHalo *sh, const double *dist,
                                                            const double *phi, uint8_t *st, Staged g, int spec)
{
  SplitHalo &H = sh[blockIdx.x];
  if (!H.active) return;
  const int64_t o = H.o, n = H.n;
  const Staged go = {g.dx + o, g.dy + o, g.dz + o, g.vx + o, g.vy + o, g.vz + o};
  double Phi0 = phi[o + n - 1] + (double)n / dist[o + n - 1];
  const double Nd = (double)n + 8.0, U = 1.1102230246251565e-16;
  const double gam = Nd * U / (1.0 - Nd * U);
  const D4 asum = {H.asum[0], H.asum[1], H.asum[2], 0.0};
  int sweeps = 0;
  const RecResume R = {H.rs_j, D4{H.rs_cv[0], H.rs_cv[1], H.rs_cv[2], H.rs_cv[3]}, H.rs_unc};
  const int rep = rec_eval<AHF_SPLIT_BLOCK, AHF_UNBIND_IPT>(K, n, H.niter == 1 ? H.seed : 0, dist + o, phi + o,
                                                            st + o, go, asum, gam, &Phi0, &sweeps,
                                                            spec ? &R : nullptr);
  if (threadIdx.x == 0) {
    H.Phi0 = Phi0;
    H.replays += rep;
    H.sweeps += sweeps;
  }
}

/* compaction: the tile's exclusive count of bound members, its total, and
 * 1 + the index of its last bound member */
__global__ void __launch_bounds__(256) k_sp_cnt_tiles(const SplitHalo *__restrict__ sh,
                                                      const SplitTile *__restrict__ tiles,
                                                  const int32_t *__restrict__ ntp,
                                                      const uint8_t *__restrict__ st, int32_t *__restrict__ wloc,
                                                      int64_t *__restrict__ tcnt, int64_t *__restrict__ tlast)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  __shared__ int64_t shi[4];
  __shared__ unsigned long long s_last;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  if (threadIdx.x == 0) s_last = 0ull;
  const int64_t j0 = T.j0 + (int64_t)threadIdx.x * 4;
  int64_t cnt = 0;
  unsigned long long last = 0ull;
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (j0 + i < H.n && st[H.o + j0 + i]) {
      cnt++;
      last = (unsigned long long)(j0 + i + 1);
    }
  int64_t tot;
  int64_t w = blk_excl_scan<256, int64_t>(cnt, &tot, shi);
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (j0 + i < H.n && st[H.o + j0 + i]) wloc[H.o + j0 + i] = (int32_t)(w++);
  if (last) atomicMax(&s_last, last);
  __syncthreads();
  if (threadIdx.x == 0) {
    tcnt[blockIdx.x] = tot;
    tlast[blockIdx.x] = (int64_t)s_last;
  }
}

/* per split halo (one workgroup each): each tile's first output slot, the
 * bound count and 1 + the index of the last bound member */
__global__ void __launch_bounds__(256) k_sp_cnt_carry(SplitHalo *sh, const int32_t *__restrict__ seg_tile0,
                                                      const int64_t *__restrict__ tcnt,
                                                      const int64_t *__restrict__ tlast, int64_t *__restrict__ tcarry)
{
  __shared__ int64_t shi[4];
  __shared__ unsigned long long s_last;
  const int64_t i = blockIdx.x;
  if (!sh[i].active) return;
  if (threadIdx.x == 0) s_last = 0ull;
  int64_t acc = 0;
  unsigned long long last = 0ull;
  const int32_t t0 = seg_tile0[i], t1 = seg_tile0[i + 1];
  for (int32_t b = t0; b < t1; b += 256) {
    const int32_t t = b + (int32_t)threadIdx.x;
    const int64_t c = t < t1 ? tcnt[t] : 0;
    if (t < t1 && (unsigned long long)tlast[t] > last) last = (unsigned long long)tlast[t];
    int64_t tot;
    const int64_t ex = blk_excl_scan<256, int64_t>(c, &tot, shi);
    if (t < t1) tcarry[t] = acc + ex;
    acc += tot;
  }
  if (last) atomicMax(&s_last, last);
  __syncthreads();
  if (threadIdx.x == 0) {
    sh[i].w = acc;
    sh[i].last = (int64_t)s_last;
  }
}

__global__ void __launch_bounds__(256) k_sp_scatter(const SplitHalo *__restrict__ sh, const SplitTile *__restrict__ tiles,
                                                  const int32_t *__restrict__ ntp,
                                                    const int64_t *__restrict__ tcarry, const uint8_t *__restrict__ st,
                                                    const int32_t *__restrict__ wloc, uint32_t *a, uint32_t *b,
                                                    const double *__restrict__ dist, Staged g,
                                                    double *__restrict__ dist_t, Staged gt)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  const uint32_t *cur = H.in_b ? b : a;
  uint32_t *nxt = H.in_b ? a : b;
  const int64_t base = tcarry[blockIdx.x];
  for (int i = 0; i < AHF_TILE / 256; i++) {
    const int64_t j = T.j0 + (int64_t)i * 256 + threadIdx.x;
    const int64_t e = H.o + j;
    if (j < H.n && st[e]) {
      /* the member and its staged values, to their place in the next list */
      const int64_t d = H.o + base + wloc[e];
      nxt[d] = cur[e];
      dist_t[d] = dist[e];
      gt.dx[d] = g.dx[e];
      gt.dy[d] = g.dy[e];
      gt.dz[d] = g.dz[e];
      gt.vx[d] = g.vx[e];
      gt.vy[d] = g.vy[e];
      gt.vz[d] = g.vz[e];
    }
  }
}

/* after the bookkeeping: the staged values of the members that stay, back
 * in place for the next iteration (unfinished haloes only) */
__global__ void __launch_bounds__(256) k_sp_restage(const SplitHalo *__restrict__ sh,
                                                    const SplitTile *__restrict__ tiles,
                                                    const int32_t *__restrict__ ntp,
                                                    const double *__restrict__ dist_t, Staged gt,
                                                    double *__restrict__ dist, Staged g)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  if (H.done) return;
  for (int i = 0; i < AHF_TILE / 256; i++) {
    const int64_t j = T.j0 + (int64_t)i * 256 + threadIdx.x;
    if (j >= H.n) break;
    const int64_t e = H.o + j;
    dist[e] = dist_t[e];
    g.dx[e] = gt.dx[e];
    g.dy[e] = gt.dy[e];
    g.dz[e] = gt.dz[e];
    g.vx[e] = gt.vx[e];
    g.vy[e] = gt.vy[e];
    g.vz[e] = gt.vz[e];
  }
}

/* the end of an iteration: rem_unbound's bookkeeping (:3590-3600) */
__global__ void k_sp_update(UnbindConst K, SplitHalo *sh, int64_t nseg, const double *__restrict__ dist)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg || !sh[i].active) return;
  SplitHalo &H = sh[i];
  if (H.last > 0) {
    H.R_vir = dist[H.o + H.last - 1];
    H.rvir_set = 1;
  }
  const int64_t nremove = H.n - H.w;
  H.n = H.w;
  H.M_vir = (double)H.w;
  H.in_b ^= 1;
  if (!(nremove > 3) || (double)H.n < (double)K.minpart) H.done = 1;
}

/* the start of an iteration (pit, niter), and the active flags */
__global__ void k_sp_begin(SplitHalo *sh, int64_t nseg)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  SplitHalo &H = sh[i];
  H.active = !H.done;
  if (H.active) {
    H.niter++;
    H.pit += H.n;
  }
}

/* tile layout of the split haloes: one tile per AHF_TILE members of every
 * halo that takes part (mode 0: the haloes still iterating; mode 1: those
 * of at least minpart members, for the rem_outsideRvir cut); seg_tile0
 * receives each halo's first tile and *ntp the tile count.  One
 * workgroup. */
__global__ void __launch_bounds__(256) k_sp_layout(const SplitHalo *__restrict__ sh, int64_t nseg, int mode,
                                                   int32_t minpart, int32_t *__restrict__ seg_tile0,
                                                   int32_t *__restrict__ ntp)
{
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nseg; base += 256) {
    const int64_t i = base + threadIdx.x;
    int32_t t = 0;
    if (i < nseg) {
      const SplitHalo &H = sh[i];
      const bool on = mode == 0 ? H.active != 0 : H.n >= minpart;
      t = on ? (int32_t)((H.n + AHF_TILE - 1) / AHF_TILE) : 0;
    }
    int32_t tot;
    const int32_t ex = block_excl_scan<int32_t>(t, &tot);
    if (i < nseg) seg_tile0[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    seg_tile0[nseg] = carry;
    *ntp = carry;
  }
}

/* the tiles of split halo blockIdx.x */
__global__ void __launch_bounds__(256) k_sp_fill_tiles(const int32_t *__restrict__ seg_tile0,
                                                       SplitTile *__restrict__ tiles)
{
  const int32_t t0 = seg_tile0[blockIdx.x], t1 = seg_tile0[blockIdx.x + 1];
  for (int32_t t = t0 + (int32_t)threadIdx.x; t < t1; t += 256)
    tiles[t] = SplitTile{(int32_t)blockIdx.x, t0, (int64_t)(t - t0) * AHF_TILE};
}

/* rem_outsideRvir over tiles: the first member of the tile whose sphere is
 * below ovlim (a NaN overdensity counts as below), or none */
__global__ void __launch_bounds__(256) k_sp_cut_tiles(UnbindConst K, const SplitHalo *__restrict__ sh,
                                                      const SplitTile *__restrict__ tiles,
                                                      const int32_t *__restrict__ ntp, const uint32_t *a,
                                                      const uint32_t *b, const double *__restrict__ cen,
                                                      const float4 *__restrict__ pv,
                                                      unsigned long long *__restrict__ tfirst)
{
  if ((int32_t)blockIdx.x >= *ntp) return;
  __shared__ unsigned long long s_first;
  const SplitTile T = tiles[blockIdx.x];
  const SplitHalo &H = sh[T.seg];
  const uint32_t *cur = split_cur(H, a, b) + H.o;
  const double c3[3] = {cen[3 * H.h], cen[3 * H.h + 1], cen[3 * H.h + 2]};
  if (threadIdx.x == 0) s_first = ~0ull;
  __syncthreads();
  unsigned long long f = ~0ull;
  for (int i = 0; i < AHF_TILE / 256; i++) {
    const int64_t j = T.j0 + (int64_t)i * 256 + threadIdx.x;
    if (j >= H.n) break;
    double r;
    const double o = sphere_ovdens(K, j, cur[j], c3, pv, &r);
    if (!(o >= K.ovlim) && (unsigned long long)j < f) f = (unsigned long long)j;
  }
  if (f != ~0ull) atomicMin(&s_first, f);
  __syncthreads();
  if (threadIdx.x == 0) tfirst[blockIdx.x] = s_first;
}

/* per split halo: the cut is the first failure over its tiles in order;
 * the members up to and including it are kept, with M_vir, R_vir and
 * ovdens of the last one kept, and the field bookkeeping of unbind_halo
 * (call 0 before the unbinding, call 1 after) */
__global__ void k_sp_cut_fin(UnbindConst K, SplitHalo *sh, int64_t nseg, int call, const int32_t *__restrict__ seg_tile0,
                             const unsigned long long *__restrict__ tfirst, const uint32_t *a, const uint32_t *b,
                             const double *__restrict__ cen, const float4 *__restrict__ pv)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  SplitHalo &H = sh[i];
  if (H.n < K.minpart) return;
  unsigned long long first = ~0ull;
  for (int32_t t = seg_tile0[i]; t < seg_tile0[i + 1]; t++)
    if (tfirst[t] != ~0ull) {
      first = tfirst[t];
      break;
    }
  const int64_t kept = first != ~0ull ? (int64_t)first + 1 : H.n;
  const uint32_t *cur = split_cur(H, a, b) + H.o;
  const double c3[3] = {cen[3 * H.h], cen[3 * H.h + 1], cen[3 * H.h + 2]};
  double R;
  const double od = sphere_ovdens(K, kept - 1, cur[kept - 1], c3, pv, &R);
  H.n = kept;
  H.M_vir = (double)kept;
  H.R_vir = R;
  H.ovdens = od;
  if (call == 0) {
    H.Phi0 = 0.0;
    H.set = AHFGPU_SET_MVIR | AHFGPU_SET_RVIR | AHFGPU_SET_PHI0 | AHFGPU_SET_OVDENS;
  } else {
    H.set |= AHFGPU_SET_MVIR | AHFGPU_SET_RVIR | AHFGPU_SET_OVDENS;
  }
}

/* 1 in *flag while some split halo still iterates (one workgroup) */
__global__ void __launch_bounds__(256) k_sp_any(const SplitHalo *__restrict__ sh, int64_t nseg, int32_t *flag)
{
  __shared__ int32_t any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int64_t i = threadIdx.x; i < nseg; i += 256)
    if (!sh[i].done) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) *flag = any;
}

/* entry into the unbinding proper (:3312): done below minpart, else R_vir
 * becomes the local of rem_unbound */
__global__ void k_sp_enter(UnbindConst K, SplitHalo *sh, int64_t nseg)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  SplitHalo &H = sh[i];
  if (K.unbind && H.n >= K.minpart) {
    H.R_vir = 0.0;
    H.set = (H.set | AHFGPU_SET_MVIR | AHFGPU_SET_PHI0) & ~AHFGPU_SET_RVIR;
    H.done = 0;
  } else {
    H.done = 1;
  }
}

__global__ void k_sp_out(const SplitHalo *__restrict__ sh, int64_t nseg, HaloOut *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  const SplitHalo &H = sh[i];
  HaloOut o;
  o.npart = H.n;
  o.M_vir = H.M_vir;
  o.R_vir = H.R_vir;
  o.Phi0 = H.Phi0;
  o.ovdens = H.ovdens;
  o.niter = H.niter;
  o.replays = H.replays;
  o.in_b = H.in_b;
  o.set = H.set | (H.rvir_set ? AHFGPU_SET_RVIR : 0);
  o.particle_iterations = H.pit;
  o.sweeps = H.sweeps;
  o.pad = 0;
  out[H.h] = o;
}

/* ------------------------------------------ indexx order of equal keys */

/*
 * AHF sorts a halo's members with Numerical Recipes' indexx
 * (src/libutility/general.c:1093-1153, called at ahf_halos.c:5846).  For
 * distinct distances any sort gives its order; on equal distances indexx
 * leaves a permutation that depends on the whole array, while the sorts
 * above keep list order.  Haloes whose sorted keys hold a tie are therefore
 * sorted again in indexx's own order: all their members laid end to end,
 * every segment of every such halo partitioned at once, one level per
 * round of kernels (the per-segment rules are in nr_order.h, where the
 * partition is derived).  A halo without a tie costs one comparison per
 * member.
 */

/* per member of the haloes with ties, and per segment of the replay */
struct TieWork {
  int32_t *th;            /* T: the haloes with ties                      */
  int64_t *tsz, *toff;    /* T, T+1: their sizes and offsets in the replay */
  int64_t *tmp64;
  double *key;            /* m: dist^2 in list order                      */
  uint32_t *idx;          /* m: indexx's indx (1-based, per halo)         */
  int32_t *seg;           /* m: segment of each position, -1 when settled  */
  uint8_t *fge, *fle;     /* m: membership in P and Q                     */
  int32_t *sGE, *sLE;     /* m+1: their exclusive prefix counts           */
  int32_t *tmp32;
  uint32_t *P, *Q;        /* m: the lists, at their segment's first position */
  int32_t *l[2], *ir[2], *b[2]; /* segment bounds and halo start (two levels) */
  double *a;              /* per segment: pivot key, pivot, element for l  */
  uint32_t *pv, *cur;
  int32_t *K, *I, *J, *cnt, *coff;
  uint8_t *act;
  int32_t *err;
};

__global__ void k_fill_u8(uint8_t *__restrict__ p, int64_t n, uint8_t v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

/* tie[h] = 1 where two neighbours of halo h's sorted list have equal keys
 * (haloes that AHF sorts only).  The halo of the workgroup's first member
 * comes from one binary search, each thread's from a walk forward from it
 * (a workgroup's consecutive members span few haloes) */
__global__ void k_tie_detect(const int64_t *__restrict__ off, int64_t nh, int64_t ne, int32_t minpart,
                             const uint32_t *__restrict__ sorted, const double *__restrict__ cen,
                             const float4 *__restrict__ pv, uint8_t *__restrict__ tie)
{
  __shared__ int64_t s_h0;
  const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
  if (threadIdx.x == 0) s_h0 = halo_of(off, nh, e0);
  __syncthreads();
  const int64_t e = e0 + threadIdx.x;
  if (e >= ne) return;
  int64_t h = s_h0;
  while (off[h + 1] <= e) h++;
  const int64_t o = off[h], n = off[h + 1] - o;
  if (sort_tier(n, minpart) == 0 || e + 1 >= o + n) return;
  const double *c = cen + 3 * h;
  if (dist_key(sorted[e], c, pv) == dist_key(sorted[e + 1], c, pv)) tie[h] = 1;
}

__global__ void k_tie_list(const uint8_t *__restrict__ tie, const int32_t *__restrict__ tpos, int64_t nh,
                           const int64_t *__restrict__ off, int32_t *__restrict__ th, int64_t *__restrict__ tsz)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh || !tie[h]) return;
  th[tpos[h]] = (int32_t)h;
  tsz[tpos[h]] = off[h + 1] - off[h];
}

/* the keys in list order (indexx's arr[1..n]), indx = 1..n, and one
 * segment per halo */
__global__ void k_tie_init(const int64_t *__restrict__ toff, int64_t T, int64_t m, const int32_t *__restrict__ th,
                           const int64_t *__restrict__ off, const uint32_t *__restrict__ ip_in,
                           const double *__restrict__ cen, const float4 *__restrict__ pv, TieWork W)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m) return;
  const int64_t k = halo_of(toff, T, g);
  const int64_t h = th[k];
  const int64_t loc = g - toff[k];
  W.key[g] = __longlong_as_double((long long)dist_key(ip_in[off[h] + loc], cen + 3 * h, pv));
  W.idx[g] = (uint32_t)(loc + 1);
  W.seg[g] = (int32_t)k;
  if (loc == 0) {
    W.l[0][k] = (int32_t)toff[k];
    W.ir[0][k] = (int32_t)(toff[k + 1] - 1);
    W.b[0][k] = (int32_t)toff[k];
  }
  if (g == 0) *W.err = 0;
}

/* per segment: straight insertion below 8 elements (then settled), else
 * the median of three */
__global__ void k_nr_prep(int64_t S, const int32_t *__restrict__ l, const int32_t *__restrict__ ir,
                          const int32_t *__restrict__ b, TieWork W)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int64_t L = l[s], R = ir[s], B = b[s];
  if (R - L < NR_SMALL) {
    nr_insertion(W.key, W.idx, B, L, R);
    W.act[s] = 0;
    return;
  }
  W.act[s] = 1;
  uint32_t pv;
  W.a[s] = nr_median3(W.key, W.idx, B, L, R, &pv);
  W.pv[s] = pv;
}

__global__ void k_nr_flags(int64_t m, const int32_t *__restrict__ l, const int32_t *__restrict__ ir,
                           const int32_t *__restrict__ b, TieWork W)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m) return;
  const int32_t s = W.seg[g];
  unsigned f = 0;
  if (s >= 0 && W.act[s]) f = nr_flags(nr_key(W.key, W.idx, b[s], g), W.a[s], g, l[s], ir[s]);
  W.fge[g] = (uint8_t)(f & 1u);
  W.fle[g] = (uint8_t)(f >> 1);
}

/* P and Q of every segment, at the segment's first position */
__global__ void k_nr_lists(int64_t m, const int32_t *__restrict__ l, const int32_t *__restrict__ ir, TieWork W)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m) return;
  const int32_t s = W.seg[g];
  if (s < 0 || !W.act[s]) return;
  const int64_t L = l[s], R = ir[s];
  if (W.fge[g]) W.P[L + (W.sGE[g] - W.sGE[L + 2])] = (uint32_t)g;
  if (W.fle[g]) W.Q[L + (W.sLE[R] - W.sLE[g + 1])] = (uint32_t)g;
}

/* per segment: where indexx's scans stop, and the element that goes to l
 * (read before the swaps) */
__global__ void k_nr_meet(int64_t S, const int32_t *__restrict__ l, const int32_t *__restrict__ ir, TieWork W)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || !W.act[s]) return;
  const int64_t L = l[s], R = ir[s];
  const int64_t nP = W.sGE[R + 1] - W.sGE[L + 2], nQ = W.sLE[R] - W.sLE[L];
  int64_t K, i, j;
  int js;
  if (nr_meet(W.P + L, W.Q + L, nP, nQ, &K, &i, &j, &js)) {
    atomicOr(W.err, 1);
    K = 0;
    i = R + 1;
    j = L;
    js = 0;
  }
  W.K[s] = (int32_t)K;
  W.I[s] = (int32_t)i;
  W.J[s] = (int32_t)j;
  W.cur[s] = js ? W.idx[W.Q[L + K - 1]] : W.idx[j];
}

/* the K swaps of every segment: position L + k swaps P[k] with Q[k] (the
 * pairs are disjoint) */
__global__ void k_nr_swap(int64_t m, const int32_t *__restrict__ l, TieWork W)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m) return;
  const int32_t s = W.seg[g];
  if (s < 0 || !W.act[s] || g - l[s] >= W.K[s]) return;
  const uint32_t p = W.P[g], q = W.Q[g];
  const uint32_t t = W.idx[p];
  W.idx[p] = W.idx[q];
  W.idx[q] = t;
}

/* per segment: the pivot to j, and how many segments follow */
__global__ void k_nr_place(int64_t S, const int32_t *__restrict__ l, const int32_t *__restrict__ ir, TieWork W)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  if (!W.act[s]) {
    W.cnt[s] = 0;
    return;
  }
  const int64_t L = l[s], R = ir[s];
  W.idx[L] = W.cur[s];
  W.idx[W.J[s]] = W.pv[s];
  W.cnt[s] = (W.J[s] - 1 - L >= 1) + (R - W.I[s] >= 1);
}

__global__ void k_nr_children(int64_t S, const int32_t *__restrict__ l, const int32_t *__restrict__ ir,
                              const int32_t *__restrict__ b, int32_t *__restrict__ l2, int32_t *__restrict__ ir2,
                              int32_t *__restrict__ b2, TieWork W)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || !W.act[s]) return;
  int64_t c = W.coff[s];
  const int32_t L = l[s], R = ir[s], B = b[s], I = W.I[s], J = W.J[s];
  if (J - 1 - L >= 1) {
    l2[c] = L;
    ir2[c] = J - 1;
    b2[c] = B;
    c++;
  }
  if (R - I >= 1) {
    l2[c] = I;
    ir2[c] = R;
    b2[c] = B;
  }
}

__global__ void k_nr_reseg(int64_t m, const int32_t *__restrict__ l, const int32_t *__restrict__ ir, TieWork W)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m) return;
  const int32_t s = W.seg[g];
  if (s < 0) return;
  if (!W.act[s]) {
    W.seg[g] = -1;
    return;
  }
  const int64_t I = W.I[s], J = W.J[s];
  const int left = J - 1 - l[s] >= 1, right = ir[s] - I >= 1;
  if (g < J) W.seg[g] = left ? W.coff[s] : -1;
  else if (g >= I) W.seg[g] = right ? W.coff[s] + left : -1;
  else W.seg[g] = -1;
}

/* the lists in indexx's order replace the sorted lists of those haloes */
__global__ void k_tie_out(int64_t m, const int64_t *__restrict__ toff, int64_t T, const int32_t *__restrict__ th,
                          const int64_t *__restrict__ off, const uint32_t *__restrict__ ip_in,
                          const uint32_t *__restrict__ idx, uint32_t *__restrict__ out)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m) return;
  const int64_t k = halo_of(toff, T, g);
  const int64_t o = off[th[k]];
  out[o + (g - toff[k])] = ip_in[o + (int64_t)idx[g] - 1];
}

static inline unsigned nb_of(int64_t n);

/* the bytes the replay takes for T haloes and m members (taken in this
 * order by nr_tie_order) */
static size_t tie_work_bytes(int64_t T, int64_t m)
{
  const size_t cs = (size_t)(m / 2 + T + 1), M = (size_t)m + 1, t = (size_t)T + 1;
  size_t z = 0;
  auto add = [&](size_t b) { z += (b + 255) & ~(size_t)255; };
  add(t * 4); add(t * 8); add(t * 8); add((size_t)scan_tmp_elems(T) * 8 + 64);
  add(M * 8); add(M * 4); add(M * 4); add(M); add(M); add(M * 4); add(M * 4);
  add((size_t)scan_tmp_elems(m > (int64_t)cs ? m : (int64_t)cs) * 4 + 64); add(M * 4); add(M * 4);
  for (int k = 0; k < 6; k++) add(cs * 4);
  add(cs * 8); add(cs * 4); add(cs * 4);
  for (int k = 0; k < 4; k++) add(cs * 4);
  add((cs + 1) * 4); add(cs); add(64);
  return z;
}

/*
 * Re-sort in indexx's order the haloes whose sorted list (in `sorted`, the
 * members as given in ip_in) holds equal keys.  tie/tpos/tmp32h: nh, nh+1
 * and scan_tmp_elems(nh) words of scratch; take(bytes) hands out device
 * memory for the replay (nullptr when out).  *n_tie: the haloes re-sorted.
 * Reads back the halo count, the member count and one segment count per
 * level.  With defer, only the detection runs: the count of haloes with
 * ties goes to c->h_words[2] once the stream reaches it, *n_tie is -1, and
 * the caller runs this again without defer if that count is not zero.
 */
template <class Take>
static int nr_tie_order(ahfgpu_ctx *c, hipStream_t st, const int64_t *off, int64_t nh, int64_t ne, int32_t minpart,
                        const uint32_t *ip_in, const double *cen, uint32_t *sorted, uint8_t *tie, int32_t *tpos,
                        int32_t *tmp32h, Take take, int64_t *n_tie, bool defer = false)
{
  *n_tie = 0;
  if (ne < 2 || nh < 1) return 0;
  hipLaunchKernelGGL(k_fill_u8, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, st, tie, nh, (uint8_t)0);
  hipLaunchKernelGGL(k_tie_detect, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, st, off, nh, ne, minpart, sorted, cen, c->pv,
                     tie);
  exclusive_scan<uint8_t, int32_t>(tie, tpos, nh, tmp32h, st);
  /* the counts come back through the context's pinned words */
  int32_t *hT = c->h_words + 2;
  if (hipMemcpyAsync(hT, tpos + nh, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "radial sort: tie count: %s", hipGetErrorString(hipGetLastError()));
  if (defer) {
    *n_tie = -1;
    return 0;
  }
  if (hipStreamSynchronize(st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "radial sort: tie count: %s", hipGetErrorString(hipGetLastError()));
  const int32_t T = *hT;
  if (T == 0) return 0;
  TieWork W;
  W.th = (int32_t *)take((size_t)(T + 1) * 4);
  W.tsz = (int64_t *)take((size_t)(T + 1) * 8);
  W.toff = (int64_t *)take((size_t)(T + 1) * 8);
  W.tmp64 = (int64_t *)take((size_t)scan_tmp_elems(T) * 8 + 64);
  if (!W.th || !W.tsz || !W.toff || !W.tmp64) return c->fail(AHFGPU_ENOMEM, "radial sort: tie buffers");
  hipLaunchKernelGGL(k_tie_list, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, st, tie, tpos, nh, off, W.th, W.tsz);
  exclusive_scan<int64_t, int64_t>(W.tsz, W.toff, T, W.tmp64, st);
  int64_t *hm = (int64_t *)(c->h_words + 4);
  if (hipMemcpyAsync(hm, W.toff + T, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "radial sort: tie members: %s", hipGetErrorString(hipGetLastError()));
  const int64_t m = *hm;
  if (m >= ((int64_t)1 << 31) - 64) return c->fail(AHFGPU_EUNSUP, "radial sort: over 2^31 members in haloes with ties");
  const int64_t cs = m / 2 + T + 1, M = m + 1;
  W.key = (double *)take((size_t)M * 8);
  W.idx = (uint32_t *)take((size_t)M * 4);
  W.seg = (int32_t *)take((size_t)M * 4);
  W.fge = (uint8_t *)take((size_t)M);
  W.fle = (uint8_t *)take((size_t)M);
  W.sGE = (int32_t *)take((size_t)M * 4);
  W.sLE = (int32_t *)take((size_t)M * 4);
  W.tmp32 = (int32_t *)take((size_t)scan_tmp_elems(std::max(m, cs)) * 4 + 64);
  W.P = (uint32_t *)take((size_t)M * 4);
  W.Q = (uint32_t *)take((size_t)M * 4);
  for (int k = 0; k < 2; k++) {
    W.l[k] = (int32_t *)take((size_t)cs * 4);
    W.ir[k] = (int32_t *)take((size_t)cs * 4);
    W.b[k] = (int32_t *)take((size_t)cs * 4);
  }
  W.a = (double *)take((size_t)cs * 8);
  W.pv = (uint32_t *)take((size_t)cs * 4);
  W.cur = (uint32_t *)take((size_t)cs * 4);
  W.K = (int32_t *)take((size_t)cs * 4);
  W.I = (int32_t *)take((size_t)cs * 4);
  W.J = (int32_t *)take((size_t)cs * 4);
  W.cnt = (int32_t *)take((size_t)cs * 4);
  W.coff = (int32_t *)take((size_t)(cs + 1) * 4);
  W.act = (uint8_t *)take((size_t)cs);
  W.err = (int32_t *)take(64);
  if (!W.key || !W.idx || !W.seg || !W.fge || !W.fle || !W.sGE || !W.sLE || !W.tmp32 || !W.P || !W.Q || !W.l[0] ||
      !W.ir[0] || !W.b[0] || !W.l[1] || !W.ir[1] || !W.b[1] || !W.a || !W.pv || !W.cur || !W.K || !W.I || !W.J ||
      !W.cnt || !W.coff || !W.act || !W.err)
    return c->fail(AHFGPU_ENOMEM, "radial sort: tie buffers");
  hipLaunchKernelGGL(k_tie_init, dim3(nb_of(m)), dim3(AHF_BLOCK), 0, st, W.toff, (int64_t)T, m, W.th, off, ip_in, cen,
                     c->pv, W);
  int64_t S = T;
  int cur = 0;
  while (S > 0) {
    const int32_t *l = W.l[cur], *ir = W.ir[cur], *b = W.b[cur];
    hipLaunchKernelGGL(k_nr_prep, dim3(nb_of(S)), dim3(AHF_BLOCK), 0, st, S, l, ir, b, W);
    hipLaunchKernelGGL(k_nr_flags, dim3(nb_of(m)), dim3(AHF_BLOCK), 0, st, m, l, ir, b, W);
    exclusive_scan<uint8_t, int32_t>(W.fge, W.sGE, m, W.tmp32, st);
    exclusive_scan<uint8_t, int32_t>(W.fle, W.sLE, m, W.tmp32, st);
    hipLaunchKernelGGL(k_nr_lists, dim3(nb_of(m)), dim3(AHF_BLOCK), 0, st, m, l, ir, W);
    hipLaunchKernelGGL(k_nr_meet, dim3(nb_of(S)), dim3(AHF_BLOCK), 0, st, S, l, ir, W);
    hipLaunchKernelGGL(k_nr_swap, dim3(nb_of(m)), dim3(AHF_BLOCK), 0, st, m, l, W);
    hipLaunchKernelGGL(k_nr_place, dim3(nb_of(S)), dim3(AHF_BLOCK), 0, st, S, l, ir, W);
    exclusive_scan<int32_t, int32_t>(W.cnt, W.coff, S, W.tmp32, st);
    hipLaunchKernelGGL(k_nr_children, dim3(nb_of(S)), dim3(AHF_BLOCK), 0, st, S, l, ir, b, W.l[cur ^ 1],
                       W.ir[cur ^ 1], W.b[cur ^ 1], W);
    hipLaunchKernelGGL(k_nr_reseg, dim3(nb_of(m)), dim3(AHF_BLOCK), 0, st, m, l, ir, W);
    int32_t S2 = 0;
    if (hipMemcpyAsync(&S2, W.coff + S, sizeof(S2), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "radial sort: tie replay: %s", hipGetErrorString(hipGetLastError()));
    S = S2;
    cur ^= 1;
  }
  int32_t err = 0;
  if (hipMemcpyAsync(&err, W.err, sizeof(err), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "radial sort: tie replay: %s", hipGetErrorString(hipGetLastError()));
  if (err) return c->fail(AHFGPU_EUNSUP, "radial sort: indexx replay found an inconsistent partition");
  hipLaunchKernelGGL(k_tie_out, dim3(nb_of(m)), dim3(AHF_BLOCK), 0, st, m, W.toff, (int64_t)T, W.th, off, ip_in, W.idx,
                     sorted);
  *n_tie = T;
  return 0;
}

/* ============================================================ host side */

static inline unsigned nb_of(int64_t n)
{
  return (unsigned)((n + AHF_BLOCK - 1) / AHF_BLOCK);
}

extern "C" int ahfgpu_set_halos(ahfgpu_ctx *c, const double *centres, const uint64_t *npart, const uint64_t *ipart,
                                uint64_t nh)
{
  if (!c || (nh && (!centres || !npart))) return AHFGPU_EINVAL;
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  std::vector<int64_t> off(nh + 1, 0);
  for (uint64_t h = 0; h < nh; h++) off[h + 1] = off[h] + (int64_t)npart[h];
  const int64_t ne = off[nh];
  if (ne >= (int64_t)UINT32_MAX) return c->fail(AHFGPU_EINVAL, "at most 2^32-1 halo members in total");
  if (ne && !ipart) return AHFGPU_EINVAL;
  for (int64_t e = 0; e < ne; e++)
    if (ipart[e] >= (uint64_t)c->N)
      return c->fail(AHFGPU_EINVAL, "member %lld refers to particle %llu beyond the resident set", (long long)e,
                     (unsigned long long)ipart[e]);
  HaloSet *s = nullptr;
  int rc = halo_set_reserve(c, (int64_t)nh, ne, &s);
  if (rc) return rc;
  halo_set_host_meta(s, off.data(), centres);
  uint32_t *h_ip = (uint32_t *)(s->host + s->o_ip);
  for (int64_t e = 0; e < ne; e++) h_ip[e] = (uint32_t)ipart[e];
  if (hipMemcpyAsync(s->dev, s->host, s->need, hipMemcpyHostToDevice, c->st) != hipSuccess) {
    free_halos(s);
    return c->fail(AHFGPU_EHIP, "halo upload");
  }
  c->halo_set = s;
  return 0;
}

/*
 * The split path for the haloes ord[0..nsplit) (the largest), on the second
 * stream, after the sort and the seeds on the first: rem_outsideRvir(0),
 * then iterations of staging (gathered once, then carried by the
 * compaction) / Phi prefix / recurrence / compaction over tiles of all
 * unfinished haloes until each one stops as rem_unbound does,
 * then rem_outsideRvir(1).  Each iteration lays out its own tiles on the
 * device and is one replay of a captured graph; the host only reads a done
 * flag between batches of replays.  The first stream waits for the result
 * before the timing ends.
 */
__global__ void k_sp_init(SplitHalo *sh, int64_t nseg, const int32_t *__restrict__ hid, const int64_t *__restrict__ off,
                          const int64_t *__restrict__ seed)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  SplitHalo H;
  memset(&H, 0, sizeof(H));
  H.h = hid[i];
  H.o = off[H.h];
  H.n = off[H.h + 1] - H.o;
  H.seed = seed[H.h];
  sh[i] = H;
}

static int split_unbind(ahfgpu_ctx *c, const UnbindConst &K, HaloSet *s, const int32_t *ord, int64_t nsplit,
                        const int64_t *seed, double *dist, double *phi, uint8_t *st, Staged g)
{
  Arena &A = c->uarena;
  SplitHalo *sh = A.get<SplitHalo>(nsplit);
  int32_t *hid = A.get<int32_t>(nsplit);
  int32_t *wloc = A.get<int32_t>(s->ne);
  int32_t *seg_tile0 = A.get<int32_t>(nsplit + 1);
  int32_t *words = A.get<int32_t>(8); /* [0] tile count, [1] some halo still iterates */
  /* tiles only shrink: the first layout bounds every later one */
  int64_t max_tiles = 0;
  for (int64_t i = 0; i < nsplit; i++) max_tiles += (s->h_off[ord[i] + 1] - s->h_off[ord[i]] + AHF_TILE - 1) / AHF_TILE;
  SplitTile *tiles = A.get<SplitTile>(max_tiles);
  D4 *ttot = A.get<D4>(max_tiles), *tcar = A.get<D4>(max_tiles);
  int64_t *tcnt = A.get<int64_t>(max_tiles), *tlast = A.get<int64_t>(max_tiles), *tcarry = A.get<int64_t>(max_tiles);
  unsigned long long *tfirst = A.get<unsigned long long>(max_tiles);
  uint8_t *tchg = A.get<uint8_t>(max_tiles), *tunc = A.get<uint8_t>(max_tiles);
  /* the staged values of the next list, written by the compaction */
  double *dist_t = A.get<double>(s->ne);
  const Staged gt = {A.get<double>(s->ne), A.get<double>(s->ne), A.get<double>(s->ne),
                     A.get<float>(s->ne),  A.get<float>(s->ne),  A.get<float>(s->ne)};
  if (!sh || !hid || !wloc || !seg_tile0 || !words || !tiles || !ttot || !tcar || !tcnt || !tlast || !tcarry || !tfirst ||
      !tchg || !tunc || !dist_t || !gt.dx || !gt.dy || !gt.dz || !gt.vx || !gt.vy || !gt.vz)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (split unbinding)");
  /* speculative passes before the sequential walk (ttot and tcar serve
   * them once the potential's prefix sums are done) */
  const int spec = c->split_spec < 0 ? 0 : c->split_spec > 0 ? c->split_spec : 3;
  hipStream_t s2 = c->st2;
  int32_t *ntp = words, *any = words + 1;
  const unsigned gs = nb_of(nsplit), nt = (unsigned)max_tiles, ns = (unsigned)nsplit;
  hipStreamWaitEvent(s2, c->ev_s0, 0);
  hipMemcpyAsync(hid, ord, (size_t)nsplit * sizeof(int32_t), hipMemcpyHostToDevice, s2);
  hipLaunchKernelGGL(k_sp_init, dim3(gs), dim3(AHF_BLOCK), 0, s2, sh, nsplit, hid, s->off, seed);
  auto cut = [&](int call) {
    hipLaunchKernelGGL(k_sp_layout, dim3(1), dim3(256), 0, s2, sh, nsplit, 1, K.minpart, seg_tile0, ntp);
    hipLaunchKernelGGL(k_sp_fill_tiles, dim3(ns), dim3(256), 0, s2, seg_tile0, tiles);
    hipLaunchKernelGGL(k_sp_cut_tiles, dim3(nt), dim3(256), 0, s2, K, sh, tiles, ntp, s->buf_a, s->buf_b, s->cen, c->pv,
                       tfirst);
    hipLaunchKernelGGL(k_sp_cut_fin, dim3(gs), dim3(AHF_BLOCK), 0, s2, K, sh, nsplit, call, seg_tile0, tfirst, s->buf_a,
                       s->buf_b, s->cen, c->pv);
  };
  if (K.cut_before) cut(0);
  hipLaunchKernelGGL(k_sp_enter, dim3(gs), dim3(AHF_BLOCK), 0, s2, K, sh, nsplit);
  /* one iteration of every unfinished halo, laid out on the device and
   * captured once as a graph; an iteration of finished haloes does
   * nothing, so the graph is replayed until the device says none is left */
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(s2, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "split unbinding: graph capture: %s", hipGetErrorString(hipGetLastError()));
  hipLaunchKernelGGL(k_sp_begin, dim3(gs), dim3(AHF_BLOCK), 0, s2, sh, nsplit);
  hipLaunchKernelGGL(k_sp_layout, dim3(1), dim3(256), 0, s2, sh, nsplit, 0, K.minpart, seg_tile0, ntp);
  hipLaunchKernelGGL(k_sp_fill_tiles, dim3(ns), dim3(256), 0, s2, seg_tile0, tiles);
  hipLaunchKernelGGL(k_sp_stage, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, s->buf_a, s->buf_b, s->cen, c->pv, dist, g,
                     st);
  hipLaunchKernelGGL(k_sp_phi_tiles, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, dist, g, phi, ttot);
  hipLaunchKernelGGL(k_sp_phi_carry, dim3(ns), dim3(256), 0, s2, sh, seg_tile0, ttot, tcar);
  hipLaunchKernelGGL(k_sp_phi_add, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, tcar, phi);
  if (spec) hipLaunchKernelGGL(k_sp_spec_init, dim3(gs), dim3(AHF_BLOCK), 0, s2, sh, nsplit, seg_tile0);
  for (int pass = 0; pass < spec; pass++) {
    hipLaunchKernelGGL(k_sp_spec_tot, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, st, g, ttot);
    hipLaunchKernelGGL(k_sp_spec_carry, dim3(ns), dim3(256), 0, s2, sh, seg_tile0, ttot, g, tcar);
    hipLaunchKernelGGL(k_sp_spec_eval, dim3(nt), dim3(256), 0, s2, K, sh, tiles, ntp, dist, phi, st, g, tcar, tchg,
                       tunc);
    hipLaunchKernelGGL(k_sp_spec_advance, dim3(ns), dim3(256), 0, s2, sh, tchg, tunc);
  }
  if (spec) {
    hipLaunchKernelGGL(k_sp_spec_tot, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, st, g, ttot);
    hipLaunchKernelGGL(k_sp_spec_resume, dim3(ns), dim3(256), 0, s2, sh, seg_tile0, tiles, ttot, g);
  }
  hipLaunchKernelGGL(k_sp_rec, dim3(ns), dim3(AHF_SPLIT_BLOCK), 0, s2, K, sh, dist, phi, st, g, spec > 0 ? 1 : 0);
  hipLaunchKernelGGL(k_sp_cnt_tiles, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, st, wloc, tcnt, tlast);
  hipLaunchKernelGGL(k_sp_cnt_carry, dim3(ns), dim3(256), 0, s2, sh, seg_tile0, tcnt, tlast, tcarry);
  hipLaunchKernelGGL(k_sp_scatter, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, tcarry, st, wloc, s->buf_a, s->buf_b,
                     dist, g, dist_t, gt);
  hipLaunchKernelGGL(k_sp_update, dim3(gs), dim3(AHF_BLOCK), 0, s2, K, sh, nsplit, dist);
  hipLaunchKernelGGL(k_sp_restage, dim3(nt), dim3(256), 0, s2, sh, tiles, ntp, dist_t, gt, dist, g);
  hipLaunchKernelGGL(k_sp_any, dim3(1), dim3(256), 0, s2, sh, nsplit, any);
  const hipError_t ec = hipStreamEndCapture(s2, &graph);
  if (ec != hipSuccess || !graph) return c->fail(AHFGPU_EHIP, "split unbinding: graph capture: %s", hipGetErrorString(ec));
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return c->fail(AHFGPU_EHIP, "split unbinding: graph instantiation: %s", hipGetErrorString(hipGetLastError()));
  }
  int rc = 0, launched = 0;
  for (int batch = c->split_graph_batch > 0 ? c->split_graph_batch : 4;;) {
    for (int k = 0; k < batch; k++) hipGraphLaunch(exec, s2);
    launched += batch;
    hipMemcpyAsync(c->h_words, any, sizeof(int32_t), hipMemcpyDeviceToHost, s2);
    if (hipStreamSynchronize(s2) != hipSuccess) {
      rc = c->fail(AHFGPU_EHIP, "split unbinding: %s", hipGetErrorString(hipGetLastError()));
      break;
    }
    if (!c->h_words[0]) break;
    /* rem_unbound stops when 3 or fewer members leave, so every iteration
     * removes at least 4: a halo cannot iterate more than n/4 times */
    if (launched > (int)(s->ne / 4) + 8) {
      rc = c->fail(AHFGPU_EHIP, "split unbinding did not finish after %d iterations", launched);
      break;
    }
    batch = 2;
  }
  c->split_graph_launches = launched;
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (rc) return rc;
  if (K.cut_after) cut(1);
  hipLaunchKernelGGL(k_sp_out, dim3(gs), dim3(AHF_BLOCK), 0, s2, sh, nsplit, s->out);
  hipEventRecord(c->ev_s1, s2);
  hipStreamWaitEvent(c->st, c->ev_s1, 0);
  return 0;
}

extern "C" int ahfgpu_set_unbind_split(ahfgpu_ctx *c, int64_t min_members)
{
  if (!c) return AHFGPU_EINVAL;
  c->split_min = min_members;
  return 0;
}

/* NminPerHalo within what the seed median supports, and a mode of the
 * documented combinations (max_run: the highest run form allowed) */
static int check_unbind_args(ahfgpu_ctx *c, const ahfgpu_unbind_params *p, int mode, int max_run)
{
  if (p->minpart < 2 || p->minpart / 2 > 64) return c->fail(AHFGPU_EINVAL, "NminPerHalo must be in [2,129]");
  const int steps = mode & ~3;
  const bool ok = (mode & 3) <= max_run &&
                  (steps == 0 || steps == AHFGPU_UNBIND_SORT_ONLY || steps == AHFGPU_UNBIND_NO_SORT ||
                   steps == AHFGPU_UNBIND_WITH_RVIR || steps == (AHFGPU_UNBIND_WITH_RVIR | AHFGPU_UNBIND_NO_SORT) ||
                   steps == AHFGPU_UNBIND_RVIR_ONLY);
  if (!ok) return c->fail(AHFGPU_EINVAL, "bad unbinding mode %d", mode);
  return 0;
}

static UnbindConst unbind_const(const ahfgpu_unbind_params *p, int mode, double band)
{
  UnbindConst K;
  K.v2_tune = p->vtune * p->vtune;
  K.v_fac = p->v_fac;
  K.r_fac = p->r_fac;
  K.phi_fac = p->phi_fac;
  K.hubble = p->hubble;
  K.band = band;
  K.ovlim = p->ovlim;
  K.rho_vir = p->rho_vir;
  K.rho_fac = p->rho_fac;
  K.minpart = p->minpart;
  K.no_vbulk = p->minpart / 2;
  K.cut_before = (mode & (AHFGPU_UNBIND_WITH_RVIR | AHFGPU_UNBIND_RVIR_ONLY)) != 0;
  K.unbind = (mode & AHFGPU_UNBIND_RVIR_ONLY) == 0;
  K.cut_after = (mode & AHFGPU_UNBIND_WITH_RVIR) != 0;
  K.pad = 0;
  return K;
}

extern "C" int ahfgpu_unbind(ahfgpu_ctx *c, const ahfgpu_unbind_params *p, int mode, double band,
                             ahfgpu_unbind_info *info)
{
  if (!c || !p) return AHFGPU_EINVAL;
  HaloSet *s = (HaloSet *)c->halo_set;
  if (!s) return c->fail(AHFGPU_EINVAL, "no halos resident (call ahfgpu_set_halos first)");
  {
    const int rc = check_unbind_args(c, p, mode, 1);
    if (rc) return rc;
  }
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  s->done = 0;
  s->res_staged = 0;
  Arena &A = c->uarena;
  A.reset();
  const int64_t ne = s->ne, nh = s->nh;
  const UnbindConst K = unbind_const(p, mode, band);
  uint64_t *k0 = A.get<uint64_t>(ne), *k1 = A.get<uint64_t>(ne);
  uint32_t *v0 = A.get<uint32_t>(ne), *v1 = A.get<uint32_t>(ne);
  uint32_t *hk0 = A.get<uint32_t>(ne), *hk1 = A.get<uint32_t>(ne);
  uint32_t *rtmp = A.get<uint32_t>(radix_tmp_u32<uint64_t>(ne));
  /* the per-halo records and both member buffers in one block, laid out as
   * the results' pinned staging is, so that one copy reads them back */
  {
    const size_t o_a = res_off_a(nh), o_b = res_off_b(nh, ne);
    char *blk = (char *)A.raw(o_b + (size_t)ne * 4 + 16);
    s->out = blk ? (HaloOut *)blk : nullptr;
    s->buf_a = blk ? (uint32_t *)(blk + o_a) : nullptr;
    s->buf_b = blk ? (uint32_t *)(blk + o_b) : nullptr;
  }
  double *dist = A.get<double>(ne), *phi = A.get<double>(ne);
  uint8_t *st = A.get<uint8_t>(ne);
  Staged g = {A.get<double>(ne), A.get<double>(ne), A.get<double>(ne),
              A.get<float>(ne), A.get<float>(ne), A.get<float>(ne)};
  int32_t *next = A.get<int32_t>(3);
  int64_t *seed = A.get<int64_t>(nh);
  if (!k0 || !k1 || !v0 || !v1 || !hk0 || !hk1 || !rtmp || !s->buf_a || !s->buf_b || !dist || !phi || !st ||
      !s->out || !next || !seed || !g.dx || !g.dy || !g.dz || !g.vx || !g.vy || !g.vz)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (unbinding)");
  c->timers_reset(TIMER_USORT, TIMER_UNBIND);
  c->timed.clear();
  c->ev_next = 0;
  if (c->timing) hipEventRecord(c->ev_u0, c->st);
  hipEvent_t es = c->timer_begin();
  const int run = mode & 3;
  int dev_cu = 256;
  hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device);
  /* haloes of at most AHF_SORT_SMALL members are sorted in LDS, one
   * workgroup each, in tiers by size; the bigger ones by the radix sort of
   * their members */
  std::vector<int64_t> bo(1, 0);
  std::vector<int32_t> bh;
  int64_t tcount[AHF_SORT_TIERS] = {0, 0, 0, 0, 0};
  for (int64_t h = 0; h < nh; h++) {
    const int64_t n = s->h_off[h + 1] - s->h_off[h];
    const int t = sort_tier(n, p->minpart);
    if (t >= 0) {
      tcount[t]++;
    } else {
      bh.push_back((int32_t)h);
      bo.push_back(bo.back() + n);
    }
  }
  const int64_t nsel = (int64_t)bh.size(), nes = bo.back();
  const bool sort_only = (mode & AHFGPU_UNBIND_SORT_ONLY) != 0;
  const bool no_sort = (mode & (AHFGPU_UNBIND_NO_SORT | AHFGPU_UNBIND_RVIR_ONLY)) != 0;
  int64_t n_tie = 0;
  uint8_t *tie = nullptr;
  int32_t *tpos = nullptr, *tmp32h = nullptr;
  bool tiers_forked = false;
  if (ne > 0 && no_sort) {
    hipLaunchKernelGGL(k_copy_u32, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, c->st, s->ip_in, ne, s->buf_a);
  } else if (ne > 0) {
    if (nh > nsel) {
      SortBase sb;
      int64_t acc = 0;
      for (int t = 0; t < AHF_SORT_TIERS; t++) {
        sb.base[t] = acc;
        acc += tcount[t];
      }
      /* the tier lists, built on the host from the sizes it holds, go up
       * through the set's pinned staging */
      int32_t *tlist = A.get<int32_t>(acc);
      int32_t *hl = (int32_t *)halo_set_up(s, (size_t)acc * sizeof(int32_t) + 16);
      if (!tlist || !hl) return c->fail(AHFGPU_ENOMEM, "sort tier lists");
      {
        int64_t fill[AHF_SORT_TIERS];
        for (int t = 0; t < AHF_SORT_TIERS; t++) fill[t] = sb.base[t];
        for (int64_t h = 0; h < nh; h++) {
          const int t = sort_tier(s->h_off[h + 1] - s->h_off[h], p->minpart);
          if (t >= 0) hl[fill[t]++] = (int32_t)h;
        }
      }
      hipMemcpyAsync(tlist, hl, (size_t)acc * sizeof(int32_t), hipMemcpyHostToDevice, c->st);
      /* the tiers sort disjoint haloes into disjoint ranges of buf_a, as
       * the radix sort below does: they run on the second and third
       * streams beside it, and the first stream waits for both before the
       * tie pass reads buf_a.  (The events are recorded again, for other
       * hand-overs, after these waits are issued.) */
      hipEventRecord(c->ev_s0, c->st);
      hipStreamWaitEvent(c->st2, c->ev_s0, 0);
      hipStreamWaitEvent(c->st3, c->ev_s0, 0);
      auto grid_of = [](int64_t n) { return dim3((unsigned)std::min<int64_t>(n, (int64_t)1 << 20)); };
#define AHF_SORT_LAUNCH(T, CAP, BS, COPY, STREAM)                                                                  \
  if (tcount[T] > 0)                                                                                               \
    hipLaunchKernelGGL((k_sort_tier<CAP, BS>), grid_of(tcount[T]), dim3(BS), 0, STREAM, tlist + sb.base[T],        \
                       tcount[T], s->off, s->ip_in, s->cen, c->pv, COPY, s->buf_a);
      AHF_SORT_LAUNCH(4, AHF_SORT_SMALL, 1024, 0, c->st2)
      AHF_SORT_LAUNCH(3, 2048, 512, 0, c->st3)
      AHF_SORT_LAUNCH(2, 512, 256, 0, c->st3)
      AHF_SORT_LAUNCH(1, 64, 64, 0, c->st2)
      AHF_SORT_LAUNCH(0, 64, 64, 1, c->st2)
#undef AHF_SORT_LAUNCH
      hipEventRecord(c->ev_s1, c->st2);
      hipEventRecord(c->ev_s2, c->st3);
      tiers_forked = true;
    }
    if (nsel > 0) {
      int64_t *d_bo = A.get<int64_t>(nsel + 1);
      int32_t *d_bh = A.get<int32_t>(nsel);
      if (!d_bo || !d_bh) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (sort)");
      hipMemcpyAsync(d_bo, bo.data(), (size_t)(nsel + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c->st);
      hipMemcpyAsync(d_bh, bh.data(), (size_t)nsel * sizeof(int32_t), hipMemcpyHostToDevice, c->st);
      hipLaunchKernelGGL(k_dist_keys_sel, dim3(nb_of(nes)), dim3(AHF_BLOCK), 0, c->st, d_bo, d_bh, nsel, nes, s->off,
                         s->ip_in, s->cen, c->pv, k0, v0);
      int w = radix_sort<uint64_t, uint32_t>(k0, v0, k1, v1, nes, 64, (uint32_t *)rtmp, c->st);
      uint32_t *vs = w ? v1 : v0;
      uint32_t *vo = w ? v0 : v1;
      hipLaunchKernelGGL(k_halo_keys, dim3(nb_of(nes)), dim3(AHF_BLOCK), 0, c->st, d_bo, nsel, nes, vs, hk0);
      int hb = 1;
      while ((1ll << hb) < nsel) hb++;
      int w2 = radix_sort<uint32_t, uint32_t>(hk0, vs, hk1, vo, nes, hb, (uint32_t *)rtmp, c->st);
      uint32_t *vsorted = w2 ? vo : vs;
      hipLaunchKernelGGL(k_gather_members_sel, dim3(nb_of(nes)), dim3(AHF_BLOCK), 0, c->st, d_bo, d_bh, nsel, nes,
                         s->off, vsorted, s->ip_in, s->buf_a);
    }
    if (tiers_forked) {
      hipStreamWaitEvent(c->st, c->ev_s1, 0);
      hipStreamWaitEvent(c->st, c->ev_s2, 0);
    }
    /* haloes with equal distances: indexx's order of them (the sort alone
     * reads the count with its results and re-sorts only when it is not
     * zero) */
    tie = A.get<uint8_t>(nh);
    tpos = A.get<int32_t>(nh + 1);
    tmp32h = A.get<int32_t>(scan_tmp_elems(nh) + 16);
    if (!tie || !tpos || !tmp32h) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (sort)");
    const int rc = nr_tie_order(c, c->st, s->off, nh, ne, p->minpart, s->ip_in, s->cen, s->buf_a, tie, tpos, tmp32h,
                                [&](size_t bytes) { return A.raw(bytes); }, &n_tie, sort_only);
    if (rc) return rc;
  }
  c->timer_end(es, TIMER_USORT, 2.0 * 12.0 * (double)ne + 12.0 * (double)ne);
  hipEvent_t eu = c->timer_begin();
  if (sort_only) {
    if (nh) hipLaunchKernelGGL(k_sorted_out, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, c->st, s->off, nh, s->out);
  } else if (nh && K.unbind) {
    hipLaunchKernelGGL(k_seed, dim3(seed_blocks(nh)), dim3(AHF_SEED_BS), 0, c->st, K, s->off, nh, s->buf_a, c->pv, seed);
  }
  if (sort_only) {
    /* nothing more */
  } else if (run == 0) {
    /* the order is largest first: haloes of split_min members or more take
     * the split path on the second stream; of the rest, haloes of
     * AHF_BIG_HALO members or more get AHF_BIG_BLOCK-thread workgroups, the
     * others AHF_BLOCK */
    int64_t nbig = 0, nsplit = 0;
    const std::vector<int32_t> &ord = s->h_ord;
    const int64_t split_min = c->split_min > 0 ? c->split_min : AHF_SPLIT_DEFAULT;
    auto size_of = [&](int64_t q) { return s->h_off[ord[q] + 1] - s->h_off[ord[q]]; };
    if (c->split_min >= 0)
      while (nsplit < nh && size_of(nsplit) >= split_min && size_of(nsplit) >= p->minpart) nsplit++;
    const int64_t big_halo = c->big_halo > 0 ? c->big_halo : AHF_BIG_HALO;
    const int64_t wave_halo = c->wave_halo > 0 ? std::min<int64_t>(c->wave_halo, big_halo) : AHF_WAVE_HALO;
    while (nbig < nh && size_of(nbig) >= big_halo) nbig++;
    if (nbig < nsplit) nbig = nsplit;
    /* haloes below AHF_WAVE_HALO members: one wavefront each */
    int64_t nmid = nbig;
    while (nmid < nh && size_of(nmid) >= wave_halo) nmid++;
    hipMemsetAsync(next, 0, 3 * sizeof(int32_t), c->st);
    /* the second and third streams start after the sort, the seeds and the
     * queue counters */
    hipEventRecord(c->ev_s0, c->st);
    if (nbig > nsplit) {
      const int64_t grid = std::min<int64_t>(nbig - nsplit, (int64_t)dev_cu * 2);
      hipLaunchKernelGGL(k_unbind_batched<AHF_BIG_BLOCK>, dim3((unsigned)grid), dim3(AHF_BIG_BLOCK), 0, c->st, K,
                         s->off, s->order, nsplit, nbig, next, seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->pv, s->out);
    }
    if (nmid > nbig) {
      const int64_t grid = std::min<int64_t>(nmid - nbig, (int64_t)dev_cu * 8);
      hipLaunchKernelGGL(k_unbind_batched<AHF_BLOCK>, dim3((unsigned)grid), dim3(AHF_BLOCK), 0, c->st, K, s->off,
                         s->order, nbig, nmid, next + 1, seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->pv, s->out);
    }
    if (nh > nmid && nmid == 0) {
      /* only one-wavefront haloes: on the first stream, no hand-over */
      const int64_t grid = std::min<int64_t>(nh, (int64_t)dev_cu * 32);
      hipLaunchKernelGGL(k_unbind_batched<64>, dim3((unsigned)grid), dim3(64), 0, c->st, K, s->off, s->order, 0, nh,
                         next + 2, seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->pv, s->out);
    } else if (nh > nmid) {
      const int64_t grid = std::min<int64_t>(nh - nmid, (int64_t)dev_cu * 32);
      hipStreamWaitEvent(c->st3, c->ev_s0, 0);
      hipLaunchKernelGGL(k_unbind_batched<64>, dim3((unsigned)grid), dim3(64), 0, c->st3, K, s->off, s->order, nmid,
                         nh, next + 2, seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->pv, s->out);
      hipEventRecord(c->ev_s2, c->st3);
      hipStreamWaitEvent(c->st, c->ev_s2, 0);
    }
    if (nsplit > 0) {
      int rc = split_unbind(c, K, s, ord.data(), nsplit, seed, dist, phi, st, g);
      if (rc) return rc;
    }
  } else {
    /* plain form: one halo per launch, results synchronised per halo as a
     * per-call rem_unbound() would need them */
    for (int64_t h = 0; h < nh; h++) {
      if (s->h_off[h + 1] - s->h_off[h] >= (c->big_halo > 0 ? c->big_halo : AHF_BIG_HALO))
        hipLaunchKernelGGL(k_unbind_one<AHF_BIG_BLOCK>, dim3(1), dim3(AHF_BIG_BLOCK), 0, c->st, K, h, s->off,
                           seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->pv,
                           s->out);
      else
        hipLaunchKernelGGL(k_unbind_one<AHF_BLOCK>, dim3(1), dim3(AHF_BLOCK), 0, c->st, K, h, s->off, seed,
                           s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->pv, s->out);
      hipStreamSynchronize(c->st);
    }
  }
  c->timer_end(eu, TIMER_UNBIND, 0.0);
  if (c->timing) hipEventRecord(c->ev_u1, c->st);
  /* the results go to their pinned staging before the same
   * synchronisation */
  {
    char *R = halo_set_res(s, res_off_b(nh, ne) + (size_t)ne * 4 + 16);
    if (R && hipMemcpyAsync(R, s->out, res_off_b(nh, ne) + (size_t)ne * 4, hipMemcpyDeviceToHost, c->st) == hipSuccess)
      s->res_staged = 1;
  }
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "unbinding: %s", hipGetErrorString(hipGetLastError()));
  if (n_tie < 0) {
    n_tie = 0;
    if (c->h_words[2] > 0) {
      s->res_staged = 0; /* the lists are re-sorted below */
      /* the records k_sorted_out wrote do not depend on the order */
      const int rc = nr_tie_order(c, c->st, s->off, nh, ne, p->minpart, s->ip_in, s->cen, s->buf_a, tie, tpos, tmp32h,
                                  [&](size_t bytes) { return A.raw(bytes); }, &n_tie);
      if (rc) return rc;
      if (hipStreamSynchronize(c->st) != hipSuccess)
        return c->fail(AHFGPU_EHIP, "radial sort: %s", hipGetErrorString(hipGetLastError()));
    }
  }
  c->timers_collect();
  float ms = 0.f;
  if (c->timing) hipEventElapsedTime(&ms, c->ev_u0, c->ev_u1);
  c->unbind_ms = ms;
  s->done = 1;
  if (info) {
    std::vector<HaloOut> ho((size_t)nh);
    if (nh) hipMemcpy(ho.data(), s->out, (size_t)nh * sizeof(HaloOut), hipMemcpyDeviceToHost);
    memset(info, 0, sizeof(*info));
    info->n_halos = nh;
    info->n_member_in = ne;
    info->n_tie_halos = n_tie;
    for (auto &o : ho) {
      info->n_member_out += o.npart;
      info->particle_iterations += o.particle_iterations;
      info->n_exact_replays += o.replays;
      if (o.sweeps > info->max_sweeps) info->max_sweeps = o.sweeps;
      if (o.niter > info->max_iterations) info->max_iterations = o.niter;
    }
  }
  return 0;
}

extern "C" int ahfgpu_unbind_results(ahfgpu_ctx *c, uint64_t *npart, double *props, uint64_t *ipart)
{
  if (!c) return AHFGPU_EINVAL;
  HaloSet *s = (HaloSet *)c->halo_set;
  if (!s || !s->done) return c->fail(AHFGPU_EINVAL, "no unbinding result");
  const int64_t nh = s->nh, ne = s->ne;
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  /* one pinned block: the per-halo records, then both member buffers,
   * read back in one copy when the device holds them the same way */
  const size_t o_a = res_off_a(nh), o_b = res_off_b(nh, ne);
  char *R = halo_set_res(s, o_b + (size_t)ne * 4 + 16);
  if (!R) return c->fail(AHFGPU_ENOMEM, "pinned staging of the results");
  if (s->res_staged) {
    /* read back already */
  } else if ((char *)s->buf_a == (char *)s->out + o_a && (char *)s->buf_b == (char *)s->out + o_b) {
    hipMemcpyAsync(R, s->out, o_b + (size_t)ne * 4, hipMemcpyDeviceToHost, c->st);
  } else {
    if (nh) hipMemcpyAsync(R, s->out, (size_t)nh * sizeof(HaloOut), hipMemcpyDeviceToHost, c->st);
    if (ne) {
      hipMemcpyAsync(R + o_a, s->buf_a, (size_t)ne * 4, hipMemcpyDeviceToHost, c->st);
      hipMemcpyAsync(R + o_b, s->buf_b, (size_t)ne * 4, hipMemcpyDeviceToHost, c->st);
    }
  }
  if (!s->res_staged && hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "unbinding results: %s", hipGetErrorString(hipGetLastError()));
  const HaloOut *ho = (const HaloOut *)R;
  const uint32_t *a = (const uint32_t *)(R + o_a), *b = (const uint32_t *)(R + o_b);
  for (int64_t h = 0; h < nh; h++) {
    const HaloOut &o = ho[h];
    if (npart) npart[h] = (uint64_t)o.npart;
    if (props) {
      double *q = props + AHFGPU_PROPS * h;
      q[0] = o.M_vir;
      q[1] = o.R_vir;
      q[2] = o.Phi0;
      q[3] = o.ovdens;
      q[4] = (double)o.niter;
      q[5] = (double)o.set;
    }
    if (ipart) {
      const uint32_t *src = o.in_b ? b : a;
      const int64_t base = s->h_off[(size_t)h];
      for (int64_t j = 0; j < o.npart; j++) ipart[base + j] = src[base + j];
    }
  }
  return 0;
}

extern "C" void ahfgpu_release_halos(ahfgpu_ctx *c)
{
  if (!c) return;
  /* the set's buffers stay with the context for the next set (the larger
   * of two when one is kept already); ahfgpu_destroy frees them */
  HaloSet *s = (HaloSet *)c->halo_set, *k = (HaloSet *)c->halo_spare;
  c->halo_set = nullptr;
  if (!s) return;
  if (k && k->cap >= s->cap) {
    free_halos(s);
    return;
  }
  free_halos(k);
  c->halo_spare = s;
}

/* ================================== unbinding without device allocation */

/*
 * ahfgpu_unbind_noalloc: the batched unbinding of haloes handed over in
 * device memory, for callers that cannot allocate.  Every intermediate
 * lives in `scratch`, one static array per GPU that the library runs on,
 * held until the process ends; its address on each device is taken once,
 * when a context is created there (ahf_noalloc_init).  Nothing here calls
 * hipMalloc or hipFree.  The steps are those of ahfgpu_unbind in the
 * batched mode with every halo in the persistent kernel, all laid out on
 * the device: the largest-first queue order (a stable radix sort of the
 * sizes), its tier bounds, the radial sort (stable radix sorts on the
 * distance key and then the halo), the seeds, the three persistent tiers,
 * and the copy of the results into the caller's arrays.
 */
#define AHF_NA_SCRATCH_BYTES ((size_t)192 << 20)
#define AHF_NA_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[AHF_NA_SCRATCH_BYTES];

static void *g_na_scratch[AHF_NA_MAX_DEVICES]; /* scratch's device address, per device */
static std::mutex g_na_mu;

/* byte offsets of the arrays of one call inside scratch */
struct NaLayout {
  size_t k0, k1, v0, v1, hk0, hk1, rtmp, a, b, dist, phi, st, dx, dy, dz, vx, vy, vz;
  size_t ok0, ok1, ov0, ov1, ortmp, seed, out, qb, words;
  size_t tie, tpos, tmp32h, tiew, tiew_bytes; /* the indexx replay of haloes with equal distances */
  size_t total;
};

static NaLayout na_layout(int64_t nh, int64_t ne)
{
  NaLayout L;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += align256(bytes > 0 ? bytes : 1);
    return r;
  };
  const size_t m = (size_t)(ne > 0 ? ne : 1), h = (size_t)(nh > 0 ? nh : 1);
  L.k0 = take(m * 8);
  L.k1 = take(m * 8);
  L.v0 = take(m * 4);
  L.v1 = take(m * 4);
  L.hk0 = take(m * 4);
  L.hk1 = take(m * 4);
  L.rtmp = take((size_t)radix_tmp_u32<uint64_t>((int64_t)m) * 4);
  L.a = take(m * 4);
  L.b = take(m * 4);
  L.dist = take(m * 8);
  L.phi = take(m * 8);
  L.st = take(m);
  L.dx = take(m * 8);
  L.dy = take(m * 8);
  L.dz = take(m * 8);
  L.vx = take(m * 4);
  L.vy = take(m * 4);
  L.vz = take(m * 4);
  L.ok0 = take(h * 4);
  L.ok1 = take(h * 4);
  L.ov0 = take(h * 4);
  L.ov1 = take(h * 4);
  L.ortmp = take((size_t)radix_tmp_u32<uint32_t>((int64_t)h) * 4);
  L.seed = take(h * 8);
  L.out = take(h * sizeof(HaloOut));
  L.qb = take(4 * sizeof(int64_t));
  L.words = take(8 * sizeof(int32_t));
  L.tie = take(h);
  L.tpos = take((h + 1) * 4);
  L.tmp32h = take((size_t)scan_tmp_elems((int64_t)h) * 4 + 64);
  L.tiew_bytes = tie_work_bytes((int64_t)h, (int64_t)m);
  L.tiew = take(L.tiew_bytes);
  L.total = o;
  return L;
}

/* once per device, with that device current: the address of scratch */
int ahf_noalloc_init(int device)
{
  if (device < 0 || device >= AHF_NA_MAX_DEVICES) return AHFGPU_EINVAL;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_scratch[device]) return 0;
  if (na_layout(AHFGPU_NOALLOC_MAX_HALOS, AHFGPU_NOALLOC_MAX_MEMBERS).total > AHF_NA_SCRATCH_BYTES)
    return AHFGPU_EINVAL;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return AHFGPU_EHIP;
  }
  g_na_scratch[device] = ptr;
  return 0;
}

static void *na_scratch_of(int device)
{
  if (device < 0 || device >= AHF_NA_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_na_mu);
  return g_na_scratch[device];
}

__global__ void k_na_zero(int32_t *w, int n)
{
  if ((int)threadIdx.x < n) w[threadIdx.x] = 0;
}

/* the caller's device inputs, before any kernel relies on them: offsets
 * that start at 0, do not decrease and end at ne; members below N.  Bit 0
 * of *bad: offsets, bit 1: members */
__global__ void k_na_check(const int64_t *__restrict__ off, int64_t nh, int64_t ne, const uint32_t *__restrict__ ip,
                           int64_t N, int32_t *__restrict__ bad)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= nh) {
    const int64_t o = off[i];
    const bool ok = o >= 0 && o <= ne && (i > 0 || o == 0) && (i < nh || o == ne) && (i == 0 || off[i - 1] <= o);
    if (!ok) atomicOr(bad, 1);
  }
  if (i < ne && (int64_t)ip[i] >= N) atomicOr(bad, 2);
}

/* queue order: sizes as descending keys, a stable sort keeps equal sizes in
 * index order (as the host's stable sort does) */
__global__ void k_na_order_keys(const int64_t *__restrict__ off, int64_t nh, uint32_t *__restrict__ key,
                                uint32_t *__restrict__ val)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  key[h] = 0xFFFFFFFFu - (uint32_t)(off[h + 1] - off[h]);
  val[h] = (uint32_t)h;
}

/* the first queue position whose halo has fewer than `lim` members (sizes do
 * not increase along the queue) */
__device__ inline int64_t na_first_below(const int64_t *off, const int32_t *order, int64_t nh, int64_t lim)
{
  int64_t lo = 0, hi = nh;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int32_t h = order[mid];
    if (off[h + 1] - off[h] >= lim) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* tier bounds of the persistent kernels and their queue counters */
__global__ void k_na_bounds(const int64_t *__restrict__ off, const int32_t *__restrict__ order, int64_t nh,
                            int64_t big_halo, int64_t wave_halo, int64_t *__restrict__ qb, int32_t *__restrict__ next)
{
  if (threadIdx.x != 0) return;
  const int64_t nbig = na_first_below(off, order, nh, big_halo);
  int64_t nmid = na_first_below(off, order, nh, wave_halo);
  if (nmid < nbig) nmid = nbig;
  qb[0] = 0;
  qb[1] = nbig;
  qb[2] = nmid;
  qb[3] = nh;
  next[0] = next[1] = next[2] = 0;
}

/* radial sort keys of all members: dist^2 where the halo is sorted, 0 where
 * AHF leaves the list as it is; the value is the list position */
__global__ void k_na_member_keys(const int64_t *__restrict__ off, int64_t nh, int64_t ne, int32_t minpart,
                                 const uint32_t *__restrict__ ip, const double *__restrict__ cen,
                                 const float4 *__restrict__ pv, uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int64_t h = halo_of(off, nh, e);
  const int64_t n = off[h + 1] - off[h];
  key[e] = sort_tier(n, minpart) != 0 ? dist_key(ip[e], cen + 3 * h, pv) : 0ull;
  val[e] = (uint32_t)e;
}

__global__ void k_na_take(const uint32_t *__restrict__ pos, int64_t ne, const uint32_t *__restrict__ ip,
                          uint32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) out[i] = ip[pos[i]];
}

__global__ void k_na_out_halo(const HaloOut *__restrict__ out, int64_t nh, uint64_t *__restrict__ npart,
                              double *__restrict__ props)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const HaloOut o = out[h];
  npart[h] = (uint64_t)o.npart;
  double *q = props + AHFGPU_PROPS * h;
  q[0] = o.M_vir;
  q[1] = o.R_vir;
  q[2] = o.Phi0;
  q[3] = o.ovdens;
  q[4] = (double)o.niter;
  q[5] = (double)o.set;
}

__global__ void k_na_out_members(const int64_t *__restrict__ off, int64_t nh, int64_t ne,
                                 const HaloOut *__restrict__ out, const uint32_t *__restrict__ a,
                                 const uint32_t *__restrict__ b, uint32_t *__restrict__ dst)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int64_t h = halo_of(off, nh, e);
  if (e - off[h] < out[h].npart) dst[e] = out[h].in_b ? b[e] : a[e];
}

extern "C" int ahfgpu_unbind_noalloc(ahfgpu_ctx *c, const ahfgpu_unbind_params *p, int mode, double band,
                                     const double *centres, const int64_t *off, const uint32_t *members, int64_t nh,
                                     uint64_t *npart_out, double *props_out, uint32_t *members_out)
{
  if (!c || !p) return AHFGPU_EINVAL;
  {
    const int rc = check_unbind_args(c, p, mode, 0);
    if (rc) return rc;
  }
  if (nh < 0 || nh > AHFGPU_NOALLOC_MAX_HALOS)
    return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc takes at most %d haloes (got %lld)",
                   AHFGPU_NOALLOC_MAX_HALOS, (long long)nh);
  if (nh == 0) return 0;
  if (!centres || !off || !npart_out || !props_out || !members_out)
    return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc: null device array");
  if (c->N <= 0) return c->fail(AHFGPU_EINVAL, "no particles resident");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  char *base = (char *)na_scratch_of(c->device);
  if (!base) {
    /* set up when the context was created; retried here, as that call may
     * have failed */
    if (ahf_noalloc_init(c->device) != 0 || !(base = (char *)na_scratch_of(c->device)))
      return c->fail(AHFGPU_EHIP, "ahfgpu_unbind_noalloc: no device address for the scratch array of device %d",
                     c->device);
  }
  hipStream_t st = c->st;
  /* the member count: the last offset, after checking the offsets' ends */
  int64_t ne = 0;
  if (hipMemcpyAsync(&ne, off + nh, sizeof(ne), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "ahfgpu_unbind_noalloc: reading off[nh]: %s", hipGetErrorString(hipGetLastError()));
  if (ne < 0 || ne > AHFGPU_NOALLOC_MAX_MEMBERS)
    return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc takes at most %d members (off[nh] = %lld)",
                   AHFGPU_NOALLOC_MAX_MEMBERS, (long long)ne);
  if (ne > 0 && !members) return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc: null member list");
  const NaLayout L = na_layout(nh, ne);
  if (L.total > AHF_NA_SCRATCH_BYTES) return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc: scratch too small");
  uint64_t *k0 = (uint64_t *)(base + L.k0), *k1 = (uint64_t *)(base + L.k1);
  uint32_t *v0 = (uint32_t *)(base + L.v0), *v1 = (uint32_t *)(base + L.v1);
  uint32_t *hk0 = (uint32_t *)(base + L.hk0), *hk1 = (uint32_t *)(base + L.hk1);
  uint32_t *rtmp = (uint32_t *)(base + L.rtmp);
  uint32_t *buf_a = (uint32_t *)(base + L.a), *buf_b = (uint32_t *)(base + L.b);
  double *dist = (double *)(base + L.dist), *phi = (double *)(base + L.phi);
  uint8_t *stt = (uint8_t *)(base + L.st);
  const Staged g = {(double *)(base + L.dx), (double *)(base + L.dy), (double *)(base + L.dz),
                    (float *)(base + L.vx),  (float *)(base + L.vy),  (float *)(base + L.vz)};
  uint32_t *ok0 = (uint32_t *)(base + L.ok0), *ok1 = (uint32_t *)(base + L.ok1);
  uint32_t *ov0 = (uint32_t *)(base + L.ov0), *ov1 = (uint32_t *)(base + L.ov1);
  uint32_t *ortmp = (uint32_t *)(base + L.ortmp);
  int64_t *seed = (int64_t *)(base + L.seed);
  HaloOut *hout = (HaloOut *)(base + L.out);
  int64_t *qb = (int64_t *)(base + L.qb);
  int32_t *words = (int32_t *)(base + L.words); /* [0..2] queue counters, [3] input check */
  /* the caller's offsets and members are checked before anything reads
   * through them; scratch keeps the last call's contents, so the check
   * word is cleared first */
  hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(64), 0, st, words, 8);
  const int64_t nchk = std::max<int64_t>(nh + 1, ne);
  hipLaunchKernelGGL(k_na_check, dim3(nb_of(nchk)), dim3(AHF_BLOCK), 0, st, off, nh, ne, members, c->N, words + 3);
  int32_t bad = 0;
  if (hipMemcpyAsync(&bad, words + 3, sizeof(bad), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "ahfgpu_unbind_noalloc: input check: %s", hipGetErrorString(hipGetLastError()));
  if (bad & 1) return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc: offsets must start at 0 and not decrease");
  if (bad & 2) return c->fail(AHFGPU_EINVAL, "ahfgpu_unbind_noalloc: a member refers to a particle beyond the resident set");
  const UnbindConst K = unbind_const(p, mode, band);
  /* largest first, and the tier bounds */
  hipLaunchKernelGGL(k_na_order_keys, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, st, off, nh, ok0, ov0);
  const int wo = radix_sort<uint32_t, uint32_t>(ok0, ov0, ok1, ov1, nh, 32, ortmp, st);
  const int32_t *order = (const int32_t *)(wo ? ov1 : ov0);
  const int64_t big_halo = c->big_halo > 0 ? c->big_halo : AHF_BIG_HALO;
  const int64_t wave_halo = c->wave_halo > 0 ? std::min<int64_t>(c->wave_halo, big_halo) : AHF_WAVE_HALO;
  hipLaunchKernelGGL(k_na_bounds, dim3(1), dim3(64), 0, st, off, order, nh, big_halo, wave_halo, qb, words);
  /* radial sort (or the lists as given) into buf_a */
  const bool sort_only = (mode & AHFGPU_UNBIND_SORT_ONLY) != 0;
  const bool no_sort = (mode & (AHFGPU_UNBIND_NO_SORT | AHFGPU_UNBIND_RVIR_ONLY)) != 0;
  if (ne > 0 && no_sort) {
    hipLaunchKernelGGL(k_copy_u32, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, st, members, ne, buf_a);
  } else if (ne > 0) {
    hipLaunchKernelGGL(k_na_member_keys, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, st, off, nh, ne, p->minpart, members,
                       centres, c->pv, k0, v0);
    const int w = radix_sort<uint64_t, uint32_t>(k0, v0, k1, v1, ne, 64, rtmp, st);
    uint32_t *vs = w ? v1 : v0, *vo = w ? v0 : v1;
    hipLaunchKernelGGL(k_halo_keys, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, st, off, nh, ne, vs, hk0);
    int hb = 1;
    while ((1ll << hb) < nh) hb++;
    const int w2 = radix_sort<uint32_t, uint32_t>(hk0, vs, hk1, vo, ne, hb, rtmp, st);
    hipLaunchKernelGGL(k_na_take, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, st, w2 ? vo : vs, ne, members, buf_a);
    /* haloes with equal distances: indexx's order of them, in the region
     * reserved for the largest replay these sizes allow */
    char *tw = base + L.tiew;
    size_t tw_left = L.tiew_bytes;
    auto take_tw = [&](size_t bytes) -> void * {
      bytes = align256(bytes > 0 ? bytes : 1);
      if (bytes > tw_left) return nullptr;
      void *r = tw;
      tw += bytes;
      tw_left -= bytes;
      return r;
    };
    int64_t n_tie = 0;
    const int rc = nr_tie_order(c, st, off, nh, ne, p->minpart, members, centres, buf_a, (uint8_t *)(base + L.tie),
                                (int32_t *)(base + L.tpos), (int32_t *)(base + L.tmp32h), take_tw, &n_tie);
    if (rc) return rc;
  }
  if (sort_only) {
    hipLaunchKernelGGL(k_sorted_out, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, st, off, nh, hout);
  } else {
    if (K.unbind)
      hipLaunchKernelGGL(k_seed, dim3(seed_blocks(nh)), dim3(AHF_SEED_BS), 0, st, K, off, nh, buf_a, c->pv,
                         seed);
    int dev_cu = 256;
    hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device);
    /* the one-wavefront tier on the third stream, as in the batched form */
    hipEventRecord(c->ev_s0, st);
    hipStreamWaitEvent(c->st3, c->ev_s0, 0);
    hipLaunchKernelGGL(k_unbind_batched_dq<AHF_BIG_BLOCK>, dim3((unsigned)std::min<int64_t>(nh, 2 * dev_cu)),
                       dim3(AHF_BIG_BLOCK), 0, st, K, off, order, qb, 0, words, seed, buf_a, buf_b, dist, phi, stt, g,
                       centres, c->pv, hout);
    hipLaunchKernelGGL(k_unbind_batched_dq<AHF_BLOCK>, dim3((unsigned)std::min<int64_t>(nh, 8 * dev_cu)),
                       dim3(AHF_BLOCK), 0, st, K, off, order, qb, 1, words + 1, seed, buf_a, buf_b, dist, phi, stt, g,
                       centres, c->pv, hout);
    hipLaunchKernelGGL(k_unbind_batched_dq<64>, dim3((unsigned)std::min<int64_t>(nh, 32 * dev_cu)), dim3(64), 0,
                       c->st3, K, off, order, qb, 2, words + 2, seed, buf_a, buf_b, dist, phi, stt, g, centres, c->pv, hout);
    hipEventRecord(c->ev_s2, c->st3);
    hipStreamWaitEvent(st, c->ev_s2, 0);
  }
  hipLaunchKernelGGL(k_na_out_halo, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, st, hout, nh, npart_out, props_out);
  if (ne > 0)
    hipLaunchKernelGGL(k_na_out_members, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, st, off, nh, ne, hout, buf_a, buf_b,
                       members_out);
  const hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return c->fail(AHFGPU_EHIP, "ahfgpu_unbind_noalloc: %s", hipGetErrorString(e));
  const hipError_t e2 = hipGetLastError();
  if (e2 != hipSuccess) return c->fail(AHFGPU_EHIP, "ahfgpu_unbind_noalloc: %s", hipGetErrorString(e2));
  return 0;
}
