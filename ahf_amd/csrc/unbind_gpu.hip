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
#include <vector>

#include "../../include/ahf_gpu.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "halo_set.h"

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
 * order is the value order */
__device__ inline uint64_t dist_key(uint32_t p, const double *__restrict__ c, const float *__restrict__ px,
                                    const float *__restrict__ py, const float *__restrict__ pz)
{
  const double dX = wrap_s((double)px[p] - c[0]);
  const double dY = wrap_s((double)py[p] - c[1]);
  const double dZ = wrap_s((double)pz[p] - c[2]);
  const double d2 = dX * dX + dY * dY + dZ * dZ;
  return (uint64_t)__double_as_longlong(d2);
}

#define AHF_SORT_SMALL 8192
#define AHF_SORT_THREADS 1024

/*
 * Radial sort of haloes of at most AHF_SORT_SMALL members, one workgroup per
 * halo: keys in LDS, bitonic sort on (key, position in the list), so equal
 * distances keep their input order as in the radix path.  Haloes below
 * minpart keep their order (AHF does not sort them).
 */
__global__ void __launch_bounds__(AHF_SORT_THREADS) k_sort_small(const int64_t *__restrict__ off, int64_t nh,
                                                                 const uint32_t *__restrict__ ip,
                                                                 const double *__restrict__ cen,
                                                                 const float *__restrict__ px,
                                                                 const float *__restrict__ py,
                                                                 const float *__restrict__ pz, int32_t minpart,
                                                                 uint32_t *__restrict__ out)
{
  __shared__ uint64_t s_key[AHF_SORT_SMALL];
  __shared__ uint32_t s_idx[AHF_SORT_SMALL];
  const int tid = threadIdx.x, bs = blockDim.x;
  for (int64_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const int64_t o0 = off[h];
    const int n = (int)(off[h + 1] - o0);
    if (n < minpart || n < 2) {
      for (int i = tid; i < n; i += bs) out[o0 + i] = ip[o0 + i];
      continue;
    }
    int m = 2;
    while (m < n) m <<= 1;
    for (int i = tid; i < m; i += bs) {
      if (i < n) {
        s_key[i] = dist_key(ip[o0 + i], cen + 3 * h, px, py, pz);
        s_idx[i] = (uint32_t)i;
      } else {
        s_key[i] = ~0ull; /* above every finite key: padding sorts last */
        s_idx[i] = 0xFFFFFFFFu;
      }
    }
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < m; i += bs) {
          const int l = i ^ j;
          if (l <= i) continue;
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
    for (int i = tid; i < n; i += bs) out[o0 + i] = ip[o0 + s_idx[i]];
    __syncthreads();
  }
}

__global__ void k_dist_keys(const int64_t *__restrict__ off, int64_t nh, int64_t ne, const uint32_t *__restrict__ ip,
                            const double *__restrict__ cen, const float *__restrict__ px,
                            const float *__restrict__ py, const float *__restrict__ pz, int32_t minpart,
                            uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int64_t h = halo_of(off, nh, e);
  val[e] = (uint32_t)e;
  if (off[h + 1] - off[h] < minpart) {
    key[e] = 0; /* not sorted by AHF: keep input order */
    return;
  }
  key[e] = dist_key(ip[e], cen + 3 * h, px, py, pz);
}

__global__ void k_halo_keys(const int64_t *__restrict__ off, int64_t nh, int64_t ne, const uint32_t *__restrict__ v,
                            uint32_t *__restrict__ hk)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) hk[i] = (uint32_t)halo_of(off, nh, (int64_t)v[i]);
}

__global__ void k_gather_members(const uint32_t *__restrict__ v, const uint32_t *__restrict__ ip, int64_t ne,
                                  uint32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) out[i] = ip[v[i]];
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

/* exact replay helper: AHF's indexx on n <= 64 values (1-based) */
__device__ void dev_indexx(int n, const double *arr1, int *indx1)
{
  int stack[24];
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

/* per-halo output record */
struct HaloOut {
  int64_t npart;
  double M_vir, R_vir, Phi0, ovdens;
  int32_t niter, replays;
  int32_t in_b; /* final list is in buffer b (0 = a) */
  int32_t set;  /* AHFGPU_SET_* fields written */
  int64_t particle_iterations;
  int32_t sweeps, pad;
};

/* overdensity of the sphere through member j at particle p, as
 * src/libahf/ahf_halos.c:3832-3856 evaluates it (M = j+1 unit masses) */
__device__ inline double sphere_ovdens(const UnbindConst &K, int64_t j, uint32_t p, const double cen[3],
                                       const float *__restrict__ px, const float *__restrict__ py,
                                       const float *__restrict__ pz, double *R_out)
{
  const double dX = wrap_s((double)px[p] - cen[0]), dY = wrap_s((double)py[p] - cen[1]),
               dZ = wrap_s((double)pz[p] - cen[2]);
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
                            const float *__restrict__ px, const float *__restrict__ py,
                            const float *__restrict__ pz, double *M, double *R, double *od)
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
      const double o = sphere_ovdens(K, j, pp[i], cen, px, py, pz, &r);
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
    *od = sphere_ovdens(K, kept - 1, l[kept - 1], cen, px, py, pz, &r);
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

template <int BS>
__device__ inline D4 blk_excl_scan4(D4 v, D4 *tot, D4 *sh /* [BS / 64] */)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  D4 inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const D4 u = shfl_up4(inc, d);
    if (lane >= d) {
      inc.x += u.x; inc.y += u.y; inc.z += u.z; inc.m += u.m;
    }
  }
  D4 ex = shfl_up4(inc, 1);
  if (lane == 0) ex = D4{0, 0, 0, 0};
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

template <int BS, int IPT>
__device__ void unbind_halo(const UnbindConst &K, int64_t n0, int64_t seed0, uint32_t *a, uint32_t *b,
                            double *dist, double *phi, uint8_t *st, Staged g, const double cen[3],
                            const float *__restrict__ px,
                            const float *__restrict__ py, const float *__restrict__ pz,
                            const float *__restrict__ vx, const float *__restrict__ vy,
                            const float *__restrict__ vz, HaloOut *out)
{
  constexpr int NW = BS / 64;
  constexpr int64_t CH = (int64_t)BS * IPT;
  __shared__ D4 sh4[NW];
  __shared__ double shd[NW];
  __shared__ int64_t shi[NW];
  __shared__ int sflag[4];
  __shared__ double sPhi0;
  __shared__ int64_t s_seed;
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
    n = rvir_cut<BS>(K, n, cur, cen, px, py, pz, &M_vir, &R_vir, &ovdens);
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
        const double dX = wrap_s((double)px[p] - cen[0]), dY = wrap_s((double)py[p] - cen[1]),
                     dZ = wrap_s((double)pz[p] - cen[2]);
        const float ux = vx[p], uy = vy[p], uz = vz[p];
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
    /* seed of the bulk velocity (:3441-3470) */
    if (t == 0) {
      if (niter == 1) {
        s_seed = seed0;
      } else {
        s_seed = 0;
      }
    }
    for (int64_t j = t; j < n; j += BS) st[j] = 1;
    __syncthreads();
    Phi0 = sPhi0;
    const double sx = (double)g.vx[s_seed], sy = (double)g.vy[s_seed], sz = (double)g.vz[s_seed];
    /* the sequential bound-set recurrence, chunk by chunk: a chunk is
     * evaluated with the final statuses of every earlier chunk (carried in
     * cv) and its own statuses of the previous evaluation, and evaluated
     * again until they no longer change; then it holds the sequential answer
     * and its totals carry to the next chunk.  Only the final evaluation's
     * margins count for the exact replay. */
    int uncertain = 0;
    D4 cv = {sx, sy, sz, 1.0};
    for (int64_t base = 0; base < n; base += CH) {
      const int64_t j0 = base + (int64_t)t * IPT;
      double fx[IPT], fy[IPT], fz[IPT];
#pragma unroll
      for (int i = 0; i < IPT; i++) {
        const int64_t j = j0 + i;
        fx[i] = fy[i] = fz[i] = 0.0;
        if (j < n) {
          fx[i] = (double)g.vx[j]; fy[i] = (double)g.vy[j]; fz[i] = (double)g.vz[j];
        }
      }
      for (;;) {
        sweeps++;
        if (t == 0) { sflag[0] = 0; sflag[1] = 0; }
        __syncthreads();
        int bj[IPT];
        D4 run = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < IPT; i++) {
          const int64_t j = j0 + i;
          bj[i] = j < n ? st[j] : 0;
          if (bj[i]) { run.x += fx[i]; run.y += fy[i]; run.z += fz[i]; run.m += 1.0; }
        }
        D4 tot;
        const D4 ex = blk_excl_scan4<BS>(run, &tot, sh4);
        /* the thread's own earlier members, summed again in the same order */
        D4 lex = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < IPT; i++) {
          const int64_t j = j0 + i;
          if (j >= n) continue;
          const double VX = cv.x + ex.x + lex.x, VY = cv.y + ex.y + lex.y, VZ = cv.z + ex.z + lex.z,
                       M = cv.m + ex.m + lex.m;
          if (bj[i]) { lex.x += fx[i]; lex.y += fy[i]; lex.z += fz[i]; lex.m += 1.0; }
          const double dj = dist[j];
          double v_esc2;
          if (dj > MZERO) v_esc2 = (2 * fabs(phi[j] - Phi0) * K.phi_fac);
          else v_esc2 = 1e30;
          const double dX = g.dx[j], dY = g.dy[j], dZ = g.dz[j];
          const double dVX = (fx[i] - VX / M) * K.v_fac + K.hubble * (dX) * K.r_fac;
          const double dVY = (fy[i] - VY / M) * K.v_fac + K.hubble * (dY) * K.r_fac;
          const double dVZ = (fz[i] - VZ / M) * K.v_fac + K.hubble * (dZ) * K.r_fac;
          const double vel2 = dVX * dVX + dVY * dVY + dVZ * dVZ;
          const double lim = K.v2_tune * v_esc2;
          const int nb = !(vel2 > lim);
          if (v_esc2 < 1e30) {
            /* certified bound on how far the sequential vel^2 and limit can
             * lie from these: Phi_j and Phi0 are sums of non-negative terms
             * (error <= gam*Phi0 each), the velocity sums are within
             * gam*2*sum|v| (the seed is counted twice); propagated through
             * the test, plus a few ulps per operation */
            const double e_phi = 2.5 * gam * fabs(Phi0);
            const double e_lim = K.v2_tune * 4.0 * K.phi_fac * e_phi;
            const double e_vx = fabs(K.v_fac) * 5.0 * gam * asum.x / M;
            const double e_vy = fabs(K.v_fac) * 5.0 * gam * asum.y / M;
            const double e_vz = fabs(K.v_fac) * 5.0 * gam * asum.z / M;
            const double e_vel2 = 2.0 * (fabs(dVX) * e_vx + fabs(dVY) * e_vy + fabs(dVZ) * e_vz) + e_vx * e_vx +
                                  e_vy * e_vy + e_vz * e_vz;
            const double tol = 4.0 * (e_vel2 + e_lim) + 1e-14 * (vel2 + lim);
            const double gap = fabs(vel2 - lim);
            if (gap <= tol || gap <= K.band * lim) sflag[1] = 1;
          }
          if (nb != bj[i]) { st[j] = (uint8_t)nb; sflag[0] = 1; }
        }
        __syncthreads();
        const int changed = sflag[0], unc = sflag[1];
        __syncthreads();
        if (!changed) {
          /* the statuses used were the final ones: so are the totals */
          uncertain |= unc;
          cv.x += tot.x; cv.y += tot.y; cv.z += tot.z; cv.m += tot.m;
          break;
        }
      }
    }
    /* exact replay of this iteration where rounding could decide (thread 0) */
    if (uncertain) {
      replays++;
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
        sPhi0 = P0;
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
      Phi0 = sPhi0;
    }
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
    n = rvir_cut<BS>(K, n, cur, cen, px, py, pz, &M_vir, &R_vir, &ovdens);
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
__global__ void k_seed(UnbindConst K, const int64_t *__restrict__ off, int64_t nh, const uint32_t *__restrict__ a,
                       const float *__restrict__ vx, const float *__restrict__ vy, const float *__restrict__ vz,
                       int64_t *__restrict__ seed)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const int64_t o = off[h], n = off[h + 1] - o;
  seed[h] = 0;
  if (n < K.minpart || K.no_vbulk < 1) return;
  double m2[65];
  int ix[65];
  const int nv = K.no_vbulk;
  for (int j = 0; j < nv; j++) {
    const uint32_t p = a[o + j];
    const float v2 = vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p];
    m2[j + 1] = (double)v2;
  }
  dev_indexx(nv, m2, ix);
  seed[h] = ix[nv / 2] - 1;
}

/* haloes of at least this many members go to 512-thread workgroups */
#define AHF_BIG_HALO 2048
#define AHF_BIG_BLOCK 512
#define AHF_UNBIND_IPT 4

/* persistent form: workgroups take haloes order[q0..q1) from the queue
 * (largest first) */
template <int BS>
__global__ void __launch_bounds__(BS) k_unbind_batched(UnbindConst K, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ order, int64_t q0, int64_t q1,
                                                       int32_t *__restrict__ next, const int64_t *__restrict__ seed,
                                                       uint32_t *a, uint32_t *b, double *dist, double *phi,
                                                       uint8_t *st, Staged g,
                                                       const double *__restrict__ cen, const float *__restrict__ px,
                                                       const float *__restrict__ py, const float *__restrict__ pz,
                                                       const float *__restrict__ vx, const float *__restrict__ vy,
                                                       const float *__restrict__ vz, HaloOut *__restrict__ out)
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
    unbind_halo<BS, AHF_UNBIND_IPT>(K, n, seed[h], a + o, b + o, dist + o, phi + o, st + o, go, c3, px, py, pz, vx,
                                    vy, vz, out + h);
    __syncthreads();
  }
}

/* plain form: one halo per launch */
template <int BS>
__global__ void __launch_bounds__(BS) k_unbind_one(UnbindConst K, int64_t h, const int64_t *__restrict__ off,
                                                   const int64_t *__restrict__ seed, uint32_t *a, uint32_t *b,
                                                   double *dist, double *phi,
                                                   uint8_t *st, Staged g, const double *__restrict__ cen,
                                                   const float *__restrict__ px, const float *__restrict__ py,
                                                   const float *__restrict__ pz, const float *__restrict__ vx,
                                                   const float *__restrict__ vy, const float *__restrict__ vz,
                                                   HaloOut *__restrict__ out)
{
  const int64_t o = off[h], n = off[h + 1] - o;
  const double c3[3] = {cen[3 * h], cen[3 * h + 1], cen[3 * h + 2]};
  const Staged go = {g.dx + o, g.dy + o, g.dz + o, g.vx + o, g.vy + o, g.vz + o};
  unbind_halo<BS, AHF_UNBIND_IPT>(K, n, seed[h], a + o, b + o, dist + o, phi + o, st + o, go, c3, px, py, pz, vx,
                                  vy, vz, out + h);
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

extern "C" int ahfgpu_unbind(ahfgpu_ctx *c, const ahfgpu_unbind_params *p, int mode, double band,
                             ahfgpu_unbind_info *info)
{
  if (!c || !p) return AHFGPU_EINVAL;
  HaloSet *s = (HaloSet *)c->halo_set;
  if (!s) return c->fail(AHFGPU_EINVAL, "no halos resident (call ahfgpu_set_halos first)");
  if (p->minpart < 2 || p->minpart / 2 > 64) return c->fail(AHFGPU_EINVAL, "NminPerHalo must be in [2,129]");
  {
    const int steps = mode & ~3;
    const bool ok = (mode & 3) <= 1 &&
                    (steps == 0 || steps == AHFGPU_UNBIND_SORT_ONLY || steps == AHFGPU_UNBIND_NO_SORT ||
                     steps == AHFGPU_UNBIND_WITH_RVIR || steps == (AHFGPU_UNBIND_WITH_RVIR | AHFGPU_UNBIND_NO_SORT) ||
                     steps == AHFGPU_UNBIND_RVIR_ONLY);
    if (!ok) return c->fail(AHFGPU_EINVAL, "bad unbinding mode %d", mode);
  }
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  Arena &A = c->uarena;
  A.reset();
  const int64_t ne = s->ne, nh = s->nh;
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
  uint64_t *k0 = A.get<uint64_t>(ne), *k1 = A.get<uint64_t>(ne);
  uint32_t *v0 = A.get<uint32_t>(ne), *v1 = A.get<uint32_t>(ne);
  uint32_t *hk0 = A.get<uint32_t>(ne), *hk1 = A.get<uint32_t>(ne);
  uint32_t *rtmp = A.get<uint32_t>(radix_tmp_u32<uint64_t>(ne));
  s->buf_a = A.get<uint32_t>(ne);
  s->buf_b = A.get<uint32_t>(ne);
  double *dist = A.get<double>(ne), *phi = A.get<double>(ne);
  uint8_t *st = A.get<uint8_t>(ne);
  Staged g = {A.get<double>(ne), A.get<double>(ne), A.get<double>(ne),
              A.get<float>(ne), A.get<float>(ne), A.get<float>(ne)};
  s->out = A.get<HaloOut>(nh);
  int32_t *next = A.get<int32_t>(2);
  int64_t *seed = A.get<int64_t>(nh);
  if (!k0 || !k1 || !v0 || !v1 || !hk0 || !hk1 || !rtmp || !s->buf_a || !s->buf_b || !dist || !phi || !st ||
      !s->out || !next || !seed || !g.dx || !g.dy || !g.dz || !g.vx || !g.vy || !g.vz)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (unbinding)");
  c->timers_reset(TIMER_USORT, TIMER_UNBIND);
  c->timed.clear();
  c->ev_next = 0;
  hipEventRecord(c->ev_u0, c->st);
  hipEvent_t es = c->timer_begin();
  const int run = mode & 3;
  int dev_cu = 256;
  hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device);
  /* the largest halo that is sorted decides the sort: LDS per halo, or the
   * radix sort of all members */
  int64_t max_sorted = 0;
  for (int64_t h = 0; h < nh; h++) {
    const int64_t n = s->h_off[h + 1] - s->h_off[h];
    if (n >= p->minpart && n > max_sorted) max_sorted = n;
  }
  const bool small_sort = max_sorted <= AHF_SORT_SMALL;
  const bool sort_only = (mode & AHFGPU_UNBIND_SORT_ONLY) != 0;
  const bool no_sort = (mode & (AHFGPU_UNBIND_NO_SORT | AHFGPU_UNBIND_RVIR_ONLY)) != 0;
  if (ne > 0 && no_sort) {
    hipLaunchKernelGGL(k_copy_u32, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, c->st, s->ip_in, ne, s->buf_a);
  } else if (ne > 0 && small_sort) {
    const int64_t grid = std::min<int64_t>(nh, (int64_t)dev_cu * 2);
    hipLaunchKernelGGL(k_sort_small, dim3((unsigned)grid), dim3(AHF_SORT_THREADS), 0, c->st, s->off, nh, s->ip_in,
                       s->cen, c->px, c->py, c->pz, p->minpart, s->buf_a);
  } else if (ne > 0) {
    hipLaunchKernelGGL(k_dist_keys, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, c->st, s->off, nh, ne, s->ip_in, s->cen,
                       c->px, c->py, c->pz, p->minpart, k0, v0);
    int w = radix_sort<uint64_t, uint32_t>(k0, v0, k1, v1, ne, 64, (uint32_t *)rtmp, c->st);
    uint32_t *vs = w ? v1 : v0;
    uint32_t *vo = w ? v0 : v1;
    hipLaunchKernelGGL(k_halo_keys, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, c->st, s->off, nh, ne, vs, hk0);
    int hb = 1;
    while ((1ll << hb) < nh) hb++;
    int w2 = radix_sort<uint32_t, uint32_t>(hk0, vs, hk1, vo, ne, hb, (uint32_t *)rtmp, c->st);
    uint32_t *vsorted = w2 ? vo : vs;
    hipLaunchKernelGGL(k_gather_members, dim3(nb_of(ne)), dim3(AHF_BLOCK), 0, c->st, vsorted, s->ip_in, ne,
                       s->buf_a);
  }
  c->timer_end(es, TIMER_USORT, 2.0 * 12.0 * (double)ne + 12.0 * (double)ne);
  hipEvent_t eu = c->timer_begin();
  if (sort_only) {
    if (nh) hipLaunchKernelGGL(k_sorted_out, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, c->st, s->off, nh, s->out);
  } else if (nh && K.unbind) {
    hipLaunchKernelGGL(k_seed, dim3(nb_of(nh)), dim3(AHF_BLOCK), 0, c->st, K, s->off, nh, s->buf_a, c->vx, c->vy,
                       c->vz, seed);
  }
  if (sort_only) {
    /* nothing more */
  } else if (run == 0) {
    /* the order is largest first: haloes of AHF_BIG_HALO members or more
     * get AHF_BIG_BLOCK-thread workgroups, the rest AHF_BLOCK */
    int64_t nbig = 0;
    {
      std::vector<int32_t> ord((size_t)nh);
      if (nh) hipMemcpy(ord.data(), s->order, (size_t)nh * sizeof(int32_t), hipMemcpyDeviceToHost);
      while (nbig < nh && s->h_off[ord[nbig] + 1] - s->h_off[ord[nbig]] >= AHF_BIG_HALO) nbig++;
    }
    hipMemsetAsync(next, 0, 2 * sizeof(int32_t), c->st);
    if (nbig > 0) {
      const int64_t grid = std::min<int64_t>(nbig, (int64_t)dev_cu * 2);
      hipLaunchKernelGGL(k_unbind_batched<AHF_BIG_BLOCK>, dim3((unsigned)grid), dim3(AHF_BIG_BLOCK), 0, c->st, K,
                         s->off, s->order, (int64_t)0, nbig, next, seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->px,
                         c->py, c->pz, c->vx, c->vy, c->vz, s->out);
    }
    if (nh > nbig) {
      const int64_t grid = std::min<int64_t>(nh - nbig, (int64_t)dev_cu * 8);
      hipLaunchKernelGGL(k_unbind_batched<AHF_BLOCK>, dim3((unsigned)grid), dim3(AHF_BLOCK), 0, c->st, K, s->off,
                         s->order, nbig, nh, next + 1, seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->px,
                         c->py,
                         c->pz, c->vx, c->vy, c->vz, s->out);
    }
  } else {
    /* plain form: one halo per launch, results synchronised per halo as a
     * per-call rem_unbound() would need them */
    for (int64_t h = 0; h < nh; h++) {
      if (s->h_off[h + 1] - s->h_off[h] >= AHF_BIG_HALO)
        hipLaunchKernelGGL(k_unbind_one<AHF_BIG_BLOCK>, dim3(1), dim3(AHF_BIG_BLOCK), 0, c->st, K, h, s->off,
                           seed, s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->px, c->py, c->pz, c->vx, c->vy, c->vz,
                           s->out);
      else
        hipLaunchKernelGGL(k_unbind_one<AHF_BLOCK>, dim3(1), dim3(AHF_BLOCK), 0, c->st, K, h, s->off, seed,
                           s->buf_a, s->buf_b, dist, phi, st, g, s->cen, c->px, c->py, c->pz, c->vx, c->vy, c->vz, s->out);
      hipStreamSynchronize(c->st);
    }
  }
  c->timer_end(eu, TIMER_UNBIND, 0.0);
  hipEventRecord(c->ev_u1, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "unbinding: %s", hipGetErrorString(hipGetLastError()));
  c->timers_collect();
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev_u0, c->ev_u1);
  c->unbind_ms = ms;
  s->done = 1;
  if (info) {
    std::vector<HaloOut> ho((size_t)nh);
    if (nh) hipMemcpy(ho.data(), s->out, (size_t)nh * sizeof(HaloOut), hipMemcpyDeviceToHost);
    memset(info, 0, sizeof(*info));
    info->n_halos = nh;
    info->n_member_in = ne;
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
  std::vector<HaloOut> ho((size_t)nh);
  std::vector<uint32_t> a((size_t)ne), b((size_t)ne);
  if (nh) hipMemcpy(ho.data(), s->out, (size_t)nh * sizeof(HaloOut), hipMemcpyDeviceToHost);
  if (ne) {
    hipMemcpy(a.data(), s->buf_a, (size_t)ne * 4, hipMemcpyDeviceToHost);
    hipMemcpy(b.data(), s->buf_b, (size_t)ne * 4, hipMemcpyDeviceToHost);
  }
  for (int64_t h = 0; h < nh; h++) {
    const HaloOut &o = ho[(size_t)h];
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
      const std::vector<uint32_t> &src = o.in_b ? b : a;
      const int64_t base = s->h_off[(size_t)h];
      for (int64_t j = 0; j < o.npart; j++) ipart[base + j] = src[(size_t)(base + j)];
    }
  }
  return 0;
}

extern "C" void ahfgpu_release_halos(ahfgpu_ctx *c)
{
  if (!c) return;
  free_halos((HaloSet *)c->halo_set);
  c->halo_set = nullptr;
}
