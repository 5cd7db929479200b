/*
 * amr_gpu.hip -- device pipeline of AHF's AMR hierarchy on gfx950.
 *
 * One call builds what src/main.c:614-648 builds on the CPU (gen_domgrids,
 * ll, zero_dens + assign_npart, the recursive gen_AMRhierarchy with
 * refine_grid / relink / relink_back), level by level, with everything
 * resident in HBM.  Per level the kernels are:
 *
 *   k_pmask        presence mask of every node (get_TSCnodes rules)
 *   k_prep         particle-node separations of every bucket entry
 *   k_gather_light exact-order TSC density, one lane per node (assign_npart)
 *   k_gather_heavy the same, one wavefront per node with many contributions
 *   k_flags        interior + refinement criterion (test_node)
 *   k_refine_rows  ref_pquad/ref_cquad/ref_nquad block tests and run state
 *                  machine, one thread per level row
 *   scans +
 *   k_compact      octets of the next level, sorted (z,y,x)
 *   k_nbr_next     27-octet neighbour table of the next level
 *   k_relink_child particles to the first interior child containing them
 *   k_fine_count,  the fine buckets, each filled from its mother's bucket
 *   k_fine_fill    in bucket order
 *
 * The per-thread rules live in amr_core.h.  The domain level is binned by a
 * stable LSD radix sort (devprims.h) so every node's bucket is in particle
 * order, which fixes AHF's list order and hence its FP32 summation order.
 */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/ahf_gpu.h"
#include "amr_core.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "level_host.h"

using namespace ahf;

/* ================================================================ kernels */

__global__ void k_domain_octets(int64_t half, uint64_t *__restrict__ key, uint8_t *__restrict__ link,
                                int32_t *__restrict__ nbr)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t no = half * half * half;
  if (o >= no) return;
  const int64_t ox = o % half, oy = (o / half) % half, oz = o / (half * half);
  key[o] = ahf_pack3((uint64_t)ox, (uint64_t)oy, (uint64_t)oz);
  link[o] = ox < half - 1 ? 1 : 0;
  for (int ez = -1; ez <= 1; ez++)
    for (int ey = -1; ey <= 1; ey++)
      for (int ex = -1; ex <= 1; ex++) {
        const int64_t nx = ahf_wrap(ox + ex, half), ny = ahf_wrap(oy + ey, half), nz = ahf_wrap(oz + ez, half);
        nbr[27 * o + ahf_e27(ex, ey, ez)] = (int32_t)((nz * half + ny) * half + nx);
      }
}

__global__ void k_domain_rows(int64_t half, uint64_t *__restrict__ row_key, int32_t *__restrict__ row_oct0,
                              int32_t *__restrict__ row_plane, int32_t *__restrict__ plane_z,
                              int32_t *__restrict__ plane_row0)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r <= half * half) row_oct0[r] = (int32_t)(r * half);
  if (r < half * half) {
    row_key[r] = (uint64_t)(r % half) | ((uint64_t)(r / half) << AHF_KEY_BITS);
    row_plane[r] = (int32_t)(r / half);
  }
  if (r <= half) plane_row0[r] = (int32_t)(r * half);
  if (r < half) plane_z[r] = (int32_t)r;
}

/* ll(): cell = (unsigned long)(L*pos), cells beyond L-1 go to 0
 * (src/libamr_serial/lltools.c:59-65); key = octet-major node index */
__global__ void k_bin_domain(const float *__restrict__ px, const float *__restrict__ py,
                             const float *__restrict__ pz, int64_t n, int64_t L, uint32_t *__restrict__ key,
                             uint32_t *__restrict__ val)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double dl = (double)L;
  const float pp[3] = {px[p], py[p], pz[p]};
  int64_t c[3];
  for (int d = 0; d < 3; d++) {
    uint64_t v = (uint64_t)(dl * (double)pp[d]);
    if (v > (uint64_t)(L - 1)) v = 0;
    c[d] = (int64_t)v;
  }
  const int64_t half = L / 2;
  const int64_t o = ((c[2] >> 1) * half + (c[1] >> 1)) * half + (c[0] >> 1);
  key[p] = (uint32_t)(8 * o + (c[2] & 1) * 4 + (c[1] & 1) * 2 + (c[0] & 1));
  val[p] = (uint32_t)p;
}

__global__ void k_iota(uint32_t *__restrict__ v, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

/* CSR start of every node from the sorted keys: first index with key >= m */
__global__ void k_csr_lower_bound(const uint32_t *__restrict__ skey, int64_t n, int64_t nn,
                                  int64_t *__restrict__ p0)
{
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m > nn) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)skey[mid] < m) lo = mid + 1; else hi = mid;
  }
  p0[m] = lo;
}

/* one thread per octet: its neighbour row and the run links it needs are
 * read once for the 8 children (ahf_presence_mask_oct); imask, when given,
 * receives the octet's interior children as bits (all 27 present) */
__global__ void k_pmask(level_view v, uint32_t *__restrict__ pmask, uint8_t *__restrict__ imask)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= v.n_oct) return;
  int32_t nb[27];
#pragma unroll
  for (int i = 0; i < 27; i++) nb[i] = v.nbr[27 * o + i];
  uint8_t lk0[9], lkm[9];
#pragma unroll
  for (int ez = -1; ez <= 1; ez++)
#pragma unroll
    for (int ey = -1; ey <= 1; ey++) {
      const int li = (ez + 1) * 3 + (ey + 1);
      const int32_t a = nb[ahf_e27(0, ey, ez)], b = nb[ahf_e27(-1, ey, ez)];
      lk0[li] = a >= 0 ? v.oct_link[a] : 0;
      lkm[li] = b >= 0 ? v.oct_link[b] : 0;
    }
  const int64_t ox = ahf_kx(v.oct_key[o]), half = v.L / 2;
  uint4 r0, r1;
  r0.x = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 0);
  r0.y = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 1);
  r0.z = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 2);
  r0.w = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 3);
  r1.x = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 4);
  r1.y = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 5);
  r1.z = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 6);
  r1.w = ahf_presence_mask_oct(nb, lk0, lkm, ox, half, 7);
  uint4 *out = reinterpret_cast<uint4 *>(pmask + 8 * o);
  out[0] = r0;
  out[1] = r1;
  if (imask) {
    const uint32_t pm[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    uint32_t im = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) im |= (pm[c] == AHF_FULL27) ? 1u << c : 0u;
    imask[o] = (uint8_t)im;
  }
}

/* weight record of every bucket entry (entries [0, *n)): its particle's
 * TSC weights against the entry's own node */
__global__ void k_prep(level_view v, const int64_t *__restrict__ n_ptr, int64_t n_max,
                       const uint32_t *__restrict__ pnode, const uint32_t *__restrict__ plist,
                       const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
                       double m2pd, ahf_w9 *__restrict__ W)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_max || q >= *n_ptr) return;
  const uint32_t p = plist[q];
  ahf_w9 w;
  ahf_entry_weights(&v, (int64_t)pnode[q], px[p], py[p], pz[p], m2pd, &w);
  W[q] = w;
}

/* position record of every bucket entry (entries [0, *n)), for the
 * per-source walk of dense levels */
__global__ void k_prep_q(const int64_t *__restrict__ n_ptr, int64_t n_max, const uint32_t *__restrict__ plist,
                       const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
                       ahf_q4 *__restrict__ qp)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_max || q >= *n_ptr) return;
  const uint32_t p = plist[q];
  ahf_q4 e;
  e.x = px[p];
  e.y = py[p];
  e.z = pz[p];
  e.w = 0.0f;
  qp[q] = e;
}


/* Morton order of a full grid of half^3 octets, half a power of two: the
 * i-th octet in Morton order is the one whose interleaved coordinates are i */
__global__ void k_morton_full(int64_t half, uint32_t *__restrict__ tord)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half * half * half) return;
  const int64_t x = (int64_t)ahf_compact3((uint64_t)i), y = (int64_t)ahf_compact3((uint64_t)i >> 1),
                z = (int64_t)ahf_compact3((uint64_t)i >> 2);
  tord[i] = (uint32_t)((z * half + y) * half + x);
}

/* Morton order of the next level from this one's: a fine octet is a
 * refined node, its Morton code is the coarse octet's followed by the child
 * index cz*4+cy*2+cx, so the fine octets in Morton order are the refined
 * nodes taken octet by octet in Morton order, children in index order.
 * Position k = 8*(rank of the octet in Morton order) + child. */
__global__ void k_morton_next_flags(const uint32_t *__restrict__ tord, const int32_t *__restrict__ node2oct,
                                    int64_t nn, int32_t *__restrict__ flag)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  flag[k] = node2oct[8 * (int64_t)tord[k >> 3] + (k & 7)] >= 0 ? 1 : 0;
}

__global__ void k_morton_next_emit(const uint32_t *__restrict__ tord, const int32_t *__restrict__ node2oct,
                                   int64_t nn, const int32_t *__restrict__ pos, uint32_t *__restrict__ ftord)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const int32_t o2 = node2oct[8 * (int64_t)tord[k >> 3] + (k & 7)];
  if (o2 >= 0) ftord[pos[k]] = (uint32_t)o2;
}

/* Morton key of every octet, for the target order of the gather */
__global__ void k_morton_keys(level_view v, uint64_t *__restrict__ mk, uint32_t *__restrict__ idx)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= v.n_oct) return;
  mk[o] = ahf_morton(v.oct_key[o]);
  idx[o] = (uint32_t)o;
}

#define AHF_XCDS 8
#define AHF_GL_NB 64 /* source nodes around one octet: 4x4x4 */
/* k_gather_light keeps the records of an octet's 4x4x4 neighbourhood as
 * three word arrays (bucket start, entry count, presence mask) at index
 * i + 4j + 18k, octets 72 words apart: the 8 children of 8 octets reading
 * the same relative source then hit 64 distinct banks ({0,1,4,5,18,19,22,23}
 * + 8 * octet modulo 64; the first 32 lanes also cover 32 banks) */
#define AHF_GL_PS 18
#define AHF_GL_OS 72

/* traversal position t of a target's sources -> offset along one axis, for
 * the axis rotation r (0 inside, 1 at X = 0, 2 at X = L-1; ahf_axis_order) */
__device__ __forceinline__ int rot_offset(int i, int r)
{
  const int v = i + r;
  return (v >= 3 ? v - 3 : v) - 1;
}

/* the three weights of record q for a source at offset (ex,ey,ez) */
__device__ __forceinline__ double w9_contrib(const ahf_w9 *__restrict__ W, uint32_t q, int iz, int iy, int ix)
{
  const double *w = reinterpret_cast<const double *>(W + q);
  return w[iz] * w[3 + iy] * w[6 + ix];
}

/*
 * Exact-order density, one lane per target node.
 *
 * Block b is remapped so that each XCD (blocks are dealt to the 8 XCDs in
 * turn) walks one contiguous stretch of the Morton order of the octets
 * (tord); the 8 children of an octet are 8 adjacent lanes.  The 8 lanes of
 * an octet first load the records of the 64 source nodes of the octet's
 * 4x4x4 neighbourhood (bucket start, entry count, presence mask) into LDS.
 * Each lane then marks, in a 27-bit mask in AHF's traversal order, the
 * sources that contribute (present as get_TSCnodes finds them, non-empty),
 * and walks them as one flat sequence of entries: the wavefront runs as long
 * as its longest lane, not as the sum over sources of the longest bucket.
 * Every contribution is two products of the entry's precomputed weights
 * (k_prep) and the FP32 rounding.  Targets with more than kmax
 * contributions are left to k_gather_heavy.
 */
template <bool OFF32>
__global__ void __launch_bounds__(AHF_BLOCK) k_gather_light(level_view v, const uint32_t *__restrict__ tord,
                                                            int64_t per_xcd, const uint32_t *__restrict__ pmask,
                                                            const int64_t *__restrict__ p0,
                                                            const ahf_w9 *__restrict__ W, int descending,
                                                            uint32_t kmax, float *__restrict__ dens,
                                                            uint32_t *__restrict__ heavy,
                                                            uint32_t *__restrict__ n_heavy)
{
  __shared__ uint32_t s_st[AHF_BLOCK / 8 * AHF_GL_OS], s_ct[AHF_BLOCK / 8 * AHF_GL_OS],
      s_pk[AHF_BLOCK / 8 * AHF_GL_OS];
  const int64_t b = blockIdx.x;
  const int64_t t = ((b % AHF_XCDS) * per_xcd + b / AHF_XCDS) * blockDim.x + threadIdx.x;
  const bool active = t < 8 * v.n_oct;
  const int c = threadIdx.x & 7;
  const int nb0 = (threadIdx.x >> 3) * AHF_GL_OS;
  const int64_t o = active ? (int64_t)tord[t >> 3] : 0;
  if (active) {
    const int32_t *nrow = v.nbr + 27 * o;
    for (int m = 0; m < 8; m++) {
      const int u = c * 8 + m;
      const int i = u & 3, j = (u >> 2) & 3, k = u >> 4;
      const int on = nrow[ahf_e27(i == 0 ? -1 : (i == 3), j == 0 ? -1 : (j == 3), k == 0 ? -1 : (k == 3))];
      uint32_t st = 0u, ct = 0u, pk = 0u;
      if (on >= 0) {
        const int64_t sn = 8 * (int64_t)on + (((k + 1) & 1) << 2 | ((j + 1) & 1) << 1 | ((i + 1) & 1));
        const int64_t a = p0[sn];
        st = (uint32_t)a;
        ct = (uint32_t)(p0[sn + 1] - a);
        pk = pmask[sn];
      }
      const int x = nb0 + i + 4 * j + AHF_GL_PS * k;
      s_st[x] = st;
      s_ct[x] = ct;
      s_pk[x] = pk;
    }
  }
  __syncthreads();
  if (!active) return;
  const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
  /* neighbourhood record of the source at offset (ex,ey,ez): index
   * own + (ez+1)*18 + (ey+1)*4 + (ex+1) with own the target's corner */
  const int rec0 = nb0 + cz * AHF_GL_PS + cy * 4 + cx;
  const int64_t n = 8 * o + c;
  int64_t X, Y, Z;
  ahf_node_xyz(&v, o, c, &X, &Y, &Z);
  const int rx = X == 0 ? 1 : (X == v.L - 1 ? 2 : 0);
  const int ry = Y == 0 ? 1 : (Y == v.L - 1 ? 2 : 0);
  const int rz = Z == 0 ? 1 : (Z == v.L - 1 ? 2 : 0);
  /* per axis, offset+1 of the source at traversal rank 0, 1, 2 as 2-bit
   * fields (identity 0|1<<2|2<<4 inside the box) */
  const uint32_t px = (uint32_t)(rot_offset(0, rx) + 1) | (uint32_t)(rot_offset(1, rx) + 1) << 2 |
                      (uint32_t)(rot_offset(2, rx) + 1) << 4;
  const uint32_t py = (uint32_t)(rot_offset(0, ry) + 1) | (uint32_t)(rot_offset(1, ry) + 1) << 2 |
                      (uint32_t)(rot_offset(2, ry) + 1) << 4;
  const uint32_t pz = (uint32_t)(rot_offset(0, rz) + 1) | (uint32_t)(rot_offset(1, rz) + 1) << 2 |
                      (uint32_t)(rot_offset(2, rz) + 1) << 4;
  /* contributing sources as bits tz*16 + ty*4 + tx of their traversal ranks */
  uint64_t mask = 0;
  uint64_t total = 0;
#pragma unroll 1
  for (int tz = 0; tz < 3; tz++)
#pragma unroll
    for (int ty = 0; ty < 3; ty++)
#pragma unroll
      for (int tx = 0; tx < 3; tx++) {
        const int ez = rot_offset(tz, rz), ey = rot_offset(ty, ry), ex = rot_offset(tx, rx);
        const int x = rec0 + (ez + 1) * AHF_GL_PS + (ey + 1) * 4 + (ex + 1);
        const uint32_t cnt = s_ct[x], pm = s_pk[x];
        if (((pm >> ((1 - ez) * 9 + (1 - ey) * 3 + (1 - ex))) & 1u) && cnt) {
          mask |= 1ull << (tz * 16 + ty * 4 + tx);
          total += cnt;
        }
      }
  if (total > kmax) {
    heavy[atomicAdd(n_heavy, 1u)] = (uint32_t)n;
    return;
  }
  const uint32_t step = descending ? 0xFFFFFFFFu : 1u;
  const char *Wb = reinterpret_cast<const char *>(W);
  float d = -1.0f;
  uint32_t rem = 0, q = 0, oz = 0, oy = 0, ox = 0;
  for (uint32_t it = 0; it < (uint32_t)total; it++) {
    if (rem == 0) {
      const int pos = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      const uint32_t ez1 = (pz >> ((pos >> 4) << 1)) & 3u, ey1 = (py >> (((pos >> 2) & 3) << 1)) & 3u,
                     ex1 = (px >> ((pos & 3) << 1)) & 3u;
      const int x = rec0 + (int)(ez1 * AHF_GL_PS + ey1 * 4 + ex1);
      const uint32_t st = s_st[x];
      rem = s_ct[x];
      q = step == 1u ? st : st + rem - 1u;
      /* byte offsets of a[1-ez], wy[1-ey], wx[1-ex] inside a record */
      oz = (2u - ez1) * 8u;
      oy = 24u + (2u - ey1) * 8u;
      ox = 48u + (2u - ex1) * 8u;
    }
    double w;
    if constexpr (OFF32) {
      const uint32_t base = q * (uint32_t)sizeof(ahf_w9);
      w = *reinterpret_cast<const double *>(Wb + (base + oz)) * *reinterpret_cast<const double *>(Wb + (base + oy)) *
          *reinterpret_cast<const double *>(Wb + (base + ox));
    } else {
      const char *rb = Wb + (uint64_t)q * sizeof(ahf_w9);
      w = *reinterpret_cast<const double *>(rb + oz) * *reinterpret_cast<const double *>(rb + oy) *
          *reinterpret_cast<const double *>(rb + ox);
    }
    d = (float)((double)d + w);
    q += step;
    rem--;
  }
  dens[n] = d;
}

/* contribution of one entry record to a target at offset (-ex,-ey,-ez) from
 * its source node, whose centre is (cx,cy,cz) */
__device__ inline double entry_contrib(const ahf_q4 &e, double cx, double cy, double cz, double L, double m2pd,
                                       int ex, int ey, int ez)
{
  return ahf_contrib(m2pd, ahf_pn_sep(e.x, cx, L), ahf_pn_sep(e.y, cy, L), ahf_pn_sep(e.z, cz, L), -ex, -ey, -ez);
}

/* ahf_tsc_w with the offset known at compile time */
template <int D>
__device__ __forceinline__ double tsc_w_c(double s)
{
  if constexpr (D == 0) return 0.75 - s * s;
  else if constexpr (D < 0) return ((0.5 - s) * (0.5 - s)) / 2;
  else return ((0.5 + s) * (0.5 + s)) / 2;
}

template <int EX, int EY, int EZ>
__device__ __forceinline__ double entry_contrib_c(const ahf_q4 &e, double cx, double cy, double cz, double L,
                                                  double m2pd)
{
  return m2pd * tsc_w_c<-EZ>(ahf_pn_sep(e.z, cz, L)) * tsc_w_c<-EY>(ahf_pn_sep(e.y, cy, L)) *
         tsc_w_c<-EX>(ahf_pn_sep(e.x, cx, L));
}

/*
 * Source (EX,EY,EZ) of a target whose 27 neighbours need no periodic wrap:
 * its entries in walk order (step +1, or -1 as a 32-bit wrap), loaded four
 * at a time ahead of the sequential FP32 rounding.  rec points at the
 * target's own node in the octet's 4x4x4 neighbourhood records.
 */
template <int EX, int EY, int EZ>
__device__ __forceinline__ void gather_source_c(float &d, const uint4 *rec, const ahf_q4 *__restrict__ qp,
                                                uint32_t step, double L, double m2pd, const double *cxv,
                                                const double *cyv, const double *czv)
{
  constexpr int OFF = EZ * 16 + EY * 4 + EX;
  constexpr int BIT = (1 - EZ) * 9 + (1 - EY) * 3 + (1 - EX);
  const uint4 r = rec[OFF];
  if (!((r.z >> BIT) & 1u) || r.y == 0) return;
  const double cx = cxv[EX + 1], cy = cyv[EY + 1], cz = czv[EZ + 1];
  const uint32_t cnt = r.y;
  uint32_t q = step == 1u ? r.x : r.x + cnt - 1u;
  uint32_t k = 0;
  for (; k + 4 <= cnt; k += 4) {
    const ahf_q4 e0 = qp[q], e1 = qp[q + step], e2 = qp[q + 2u * step], e3 = qp[q + 3u * step];
    const double w0 = entry_contrib_c<EX, EY, EZ>(e0, cx, cy, cz, L, m2pd);
    const double w1 = entry_contrib_c<EX, EY, EZ>(e1, cx, cy, cz, L, m2pd);
    const double w2 = entry_contrib_c<EX, EY, EZ>(e2, cx, cy, cz, L, m2pd);
    const double w3 = entry_contrib_c<EX, EY, EZ>(e3, cx, cy, cz, L, m2pd);
    d = (float)((double)d + w0);
    d = (float)((double)d + w1);
    d = (float)((double)d + w2);
    d = (float)((double)d + w3);
    q += 4u * step;
  }
  for (; k < cnt; k++, q += step) d = (float)((double)d + entry_contrib_c<EX, EY, EZ>(qp[q], cx, cy, cz, L, m2pd));
}

#define AHF_GS_ROW(EY, EZ)                                                           \
  gather_source_c<-1, EY, EZ>(d, rec, qp, step, L, m2pd, cxv, cyv, czv);            \
  gather_source_c<0, EY, EZ>(d, rec, qp, step, L, m2pd, cxv, cyv, czv);             \
  gather_source_c<1, EY, EZ>(d, rec, qp, step, L, m2pd, cxv, cyv, czv)
#define AHF_GS_PLANE(EZ) \
  AHF_GS_ROW(-1, EZ);    \
  AHF_GS_ROW(0, EZ);     \
  AHF_GS_ROW(1, EZ)

/*
 * Exact-order density for dense levels, one lane per target node, walking
 * the 27 sources one by one over the entries' position records.
 *
 * Block b is remapped so that each XCD (blocks are dealt to the 8 XCDs in
 * turn) walks one contiguous stretch of the Morton order of the octets
 * (tord); the 8 children of an octet are 8 adjacent lanes.  The 8 lanes of
 * an octet first load the records of the 64 source nodes of the octet's
 * 4x4x4 neighbourhood (bucket start, entry count, presence mask) into LDS;
 * every child then counts and gathers its 27 sources from there.  Targets
 * with more than kmax contributions are left to k_gather_heavy.  Targets
 * away from the periodic faces walk their sources in the fixed order
 * z,y,x = -1,0,1 with compile-time offsets; the others follow AHF's rotated
 * order at the faces (ahf_axis_order).
 */
__global__ void __launch_bounds__(AHF_BLOCK) k_gather_src(level_view v, const uint32_t *__restrict__ tord,
                                                            int64_t per_xcd, const uint32_t *__restrict__ pmask,
                                                            const int64_t *__restrict__ p0,
                                                            const ahf_q4 *__restrict__ qp, int descending,
                                                            double m2pd, uint32_t kmax, float *__restrict__ dens,
                                                            uint32_t *__restrict__ heavy,
                                                            uint32_t *__restrict__ n_heavy)
{
  __shared__ uint4 s_rec[AHF_BLOCK / 8 * AHF_GL_NB];
  const int64_t b = blockIdx.x;
  const int64_t t = ((b % AHF_XCDS) * per_xcd + b / AHF_XCDS) * blockDim.x + threadIdx.x;
  const bool active = t < 8 * v.n_oct;
  const int c = threadIdx.x & 7;
  const int nb0 = (threadIdx.x >> 3) * AHF_GL_NB;
  const int64_t o = active ? (int64_t)tord[t >> 3] : 0;
  if (active) {
    const int32_t *nrow = v.nbr + 27 * o;
    for (int m = 0; m < 8; m++) {
      const int u = c * 8 + m;
      const int i = u & 3, j = (u >> 2) & 3, k = u >> 4;
      const int on = nrow[ahf_e27(i == 0 ? -1 : (i == 3), j == 0 ? -1 : (j == 3), k == 0 ? -1 : (k == 3))];
      uint4 r = make_uint4(0u, 0u, 0u, 0u);
      if (on >= 0) {
        const int64_t sn = 8 * (int64_t)on + (((k + 1) & 1) << 2 | ((j + 1) & 1) << 1 | ((i + 1) & 1));
        const int64_t a = p0[sn];
        r.x = (uint32_t)a;
        r.y = (uint32_t)(p0[sn + 1] - a);
        r.z = pmask[sn];
      }
      s_rec[nb0 + u] = r;
    }
  }
  __syncthreads();
  if (!active) return;
  const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
  /* the target's own node in the neighbourhood: offsets (dx,dy,dz) add
   * dz*16 + dy*4 + dx */
  const uint4 *rec = s_rec + nb0 + (cz + 1) * 16 + (cy + 1) * 4 + (cx + 1);
  const int64_t n = 8 * o + c;
  uint64_t total = 0;
  for (int ez = -1; ez <= 1; ez++)
    for (int ey = -1; ey <= 1; ey++)
      for (int ex = -1; ex <= 1; ex++) {
        const uint4 r = rec[ez * 16 + ey * 4 + ex];
        if ((r.z >> ((1 - ez) * 9 + (1 - ey) * 3 + (1 - ex))) & 1u) total += r.y;
      }
  if (total > kmax) {
    heavy[atomicAdd(n_heavy, 1u)] = (uint32_t)n;
    return;
  }
  const double L = (double)v.L;
  int64_t X, Y, Z;
  ahf_node_xyz(&v, o, c, &X, &Y, &Z);
  double cxv[3], cyv[3], czv[3];
  ahf_centres3(X, v.L, cxv);
  ahf_centres3(Y, v.L, cyv);
  ahf_centres3(Z, v.L, czv);
  float d = -1.0f;
  const bool inner = X > 0 && X < v.L - 1 && Y > 0 && Y < v.L - 1 && Z > 0 && Z < v.L - 1;
  if (inner) {
    const uint32_t step = descending ? 0xFFFFFFFFu : 1u;
    AHF_GS_PLANE(-1);
    AHF_GS_PLANE(0);
    AHF_GS_PLANE(1);
  } else {
    int oz[3], oy[3], ox[3];
    ahf_axis_order(Z, v.L, oz);
    ahf_axis_order(Y, v.L, oy);
    ahf_axis_order(X, v.L, ox);
    for (int a3 = 0; a3 < 3; a3++)
      for (int b3 = 0; b3 < 3; b3++)
        for (int e3 = 0; e3 < 3; e3++) {
          const int ez = oz[a3], ey = oy[b3], ex = ox[e3];
          const uint4 r = rec[ez * 16 + ey * 4 + ex];
          if (!((r.z >> ((1 - ez) * 9 + (1 - ey) * 3 + (1 - ex))) & 1u)) continue;
          const int64_t cnt = r.y;
          if (cnt == 0) continue;
          const int64_t q0 = r.x;
          const double scx = cxv[ex + 1], scy = cyv[ey + 1], scz = czv[ez + 1];
          for (int64_t k = 0; k < cnt; k++)
            d = (float)((double)d +
                        entry_contrib(qp[descending ? q0 + cnt - 1 - k : q0 + k], scx, scy, scz, L, m2pd, ex, ey, ez));
        }
  }
  dens[n] = d;
}

#define AHF_SLAB 16                          /* target tile: 16x16 nodes of one plane */
#define AHF_SLAB_W (AHF_SLAB + 2)            /* its sources in one plane: 18x18 nodes */
#define AHF_SLAB_N (AHF_SLAB_W * AHF_SLAB_W)
#define AHF_SLAB_CAP_BIG 2752                /* staged entries per plane, dense levels */
#define AHF_SLAB_CAP_SMALL 512               /* the same for about one entry per node */
#define AHF_MARCH_CAP_SMALL 448              /* the marching walk at about one entry per node */

/* node (X,Y,Z) of a level that holds every octet: its octets are all
 * (L/2)^3 keys in (z,y,x) order */
__device__ __forceinline__ int64_t full_node(int64_t X, int64_t Y, int64_t Z, int64_t H)
{
  const int64_t o = ((Z >> 1) * H + (Y >> 1)) * H + (X >> 1);
  return 8 * o + ((Z & 1) << 2 | (Y & 1) << 1 | (X & 1));
}

/* ahf_tsc_w(s, d) for an offset d known only at run time, as
 * A + B*(C + D*s)^2 with
 *   d =  0: C = 0,   D = 1,  A = 0.75, B = -1   -> 0.75 - s*s
 *   d = +1: C = 0.5, D = 1,  A = 0,    B = 0.5  -> ((0.5 + s)^2)/2
 *   d = -1: C = 0.5, D = -1, A = 0,    B = 0.5  -> ((0.5 - s)^2)/2
 * Every step is exact or rounds exactly as AHF's formula (products with
 * +-1 and 0.5 and sums with 0 are exact, x + (-y) is x - y), so the bits
 * are those of ahf_tsc_w. */
struct TscRt {
  double A, B, C, D;
  __device__ __forceinline__ explicit TscRt(int d)
      : A(d == 0 ? 0.75 : 0.0), B(d == 0 ? -1.0 : 0.5), C(d == 0 ? 0.0 : 0.5), D(d < 0 ? -1.0 : 1.0)
  {
  }
  __device__ __forceinline__ double operator()(double s) const
  {
    const double q = C + D * s;
    return A + B * (q * q);
  }
};

/*
 * Exact-order density of a level that holds every octet (the domain grid,
 * and any level refined everywhere), one workgroup per 16x16 tile of target
 * nodes in one z-plane, one lane per target; with faces set only the planes
 * Z = 0 and Z = L-1, whose z order wraps (the others march, k_gather_march).
 *
 * All targets of the tile share Z, so they visit the three source planes in
 * the same order (ahf_axis_order of Z).  The workgroup stages one source
 * plane at a time: the entries of its 18x18 source nodes are copied into
 * LDS, each as m2pd*wz (wz the TSC weight along z for this plane's offset)
 * and either its separations sy, sx (W9 false) or its three weights along
 * y and along x (W9 true).  Every lane then walks its (up to) nine sources
 * of the plane in AHF's order, each source's entries in list order, and
 * rounds ((m2pd*wz)*wy)*wx into its FP32 node value as AHF does.  Entries of
 * nodes that do not fit in the CAP staged slots are read from the position
 * records instead, with the same arithmetic.  Targets with more than kmax
 * contributions are left to k_gather_heavy_q.  Tiles are dealt so that each
 * XCD walks a contiguous run of them (consecutive planes share two source
 * planes in its L2).
 */
template <int CAP, bool W9>
__global__ void __launch_bounds__(AHF_BLOCK) k_gather_slab(level_view v, int64_t per_xcd,
                                                           const uint32_t *__restrict__ pmask,
                                                           const int64_t *__restrict__ p0,
                                                           const ahf_q4 *__restrict__ qp, int descending,
                                                           double m2pd, uint32_t kmax, float *__restrict__ dens,
                                                           uint32_t *__restrict__ heavy,
                                                           uint32_t *__restrict__ n_heavy, int faces)
{
  constexpr int NV = W9 ? 7 : 3; /* doubles per staged entry */
  __shared__ uint32_t s_start[3 * AHF_SLAB_N], s_cnt[3 * AHF_SLAB_N], s_pm[3 * AHF_SLAB_N];
  __shared__ uint32_t s_off[AHF_SLAB_N + 1];
  __shared__ double s_cx[AHF_SLAB_W], s_cy[AHF_SLAB_W];
  __shared__ double s_e[NV * CAP];
  const int64_t L = v.L, H = L >> 1, T = L / AHF_SLAB;
  const int64_t b = blockIdx.x;
  const int64_t tile = (b % AHF_XCDS) * per_xcd + b / AHF_XCDS;
  if (tile >= T * T * (faces ? 2 : L)) return; /* the whole workgroup */
  const int64_t zi = tile / (T * T), ty = (tile / T) % T, tx = tile % T;
  const int64_t Z = faces ? (zi == 0 ? 0 : L - 1) : zi;
  const int tid = threadIdx.x, lx = tid & (AHF_SLAB - 1), ly = tid / AHF_SLAB;
  const int64_t X = tx * AHF_SLAB + lx, Y = ty * AHF_SLAB + ly;
  const double Ld = (double)L;
  int oz[3], oy[3], ox[3];
  ahf_axis_order(Z, L, oz);
  ahf_axis_order(Y, L, oy);
  ahf_axis_order(X, L, ox);
  /* bucket start, entry count and presence mask of the 3 x 18 x 18 source
   * nodes, planes in traversal order; centres of the 18 source columns and
   * rows */
  for (int i = tid; i < 3 * AHF_SLAB_N; i += AHF_BLOCK) {
    const int k = i / AHF_SLAB_N, u = i - k * AHF_SLAB_N;
    const int64_t Xs = ahf_wrap(tx * AHF_SLAB - 1 + u % AHF_SLAB_W, L);
    const int64_t Ys = ahf_wrap(ty * AHF_SLAB - 1 + u / AHF_SLAB_W, L);
    const int64_t sn = full_node(Xs, Ys, ahf_wrap(Z + oz[k], L), H);
    const int64_t a = p0[sn];
    s_start[i] = (uint32_t)a;
    s_cnt[i] = (uint32_t)(p0[sn + 1] - a);
    s_pm[i] = pmask[sn];
  }
  if (tid < AHF_SLAB_W) s_cx[tid] = ahf_centre(ahf_wrap(tx * AHF_SLAB - 1 + tid, L), Ld);
  else if (tid < 2 * AHF_SLAB_W)
    s_cy[tid - AHF_SLAB_W] = ahf_centre(ahf_wrap(ty * AHF_SLAB - 1 + (tid - AHF_SLAB_W), L), Ld);
  __syncthreads();
  /* contributing sources as bits k*9 + ty*3 + tx of their traversal ranks */
  uint32_t mask = 0;
  uint64_t total = 0;
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int bb = 0; bb < 3; bb++)
#pragma unroll
      for (int e = 0; e < 3; e++) {
        const int ez = oz[k], ey = oy[bb], ex = ox[e];
        const int u = k * AHF_SLAB_N + (ly + 1 + ey) * AHF_SLAB_W + (lx + 1 + ex);
        if (!((s_pm[u] >> ((1 - ez) * 9 + (1 - ey) * 3 + (1 - ex))) & 1u)) continue;
        total += s_cnt[u];
        if (s_cnt[u]) mask |= 1u << (k * 9 + bb * 3 + e);
      }
  const int64_t n = full_node(X, Y, Z, H);
  const bool active = total <= kmax;
  if (!active) heavy[atomicAdd(n_heavy, 1u)] = (uint32_t)n;
  const uint32_t step = descending ? 0xFFFFFFFFu : 1u;
  float d = -1.0f;
  for (int k = 0; k < 3; k++) {
    const int ez = oz[k];
    const uint32_t *cnt_k = s_cnt + k * AHF_SLAB_N;
    const uint32_t *start_k = s_start + k * AHF_SLAB_N;
    /* LDS slot of each node's first entry: node pairs per thread */
    uint32_t c0 = 0, c1 = 0;
    if (tid < AHF_SLAB_N / 2) {
      c0 = cnt_k[2 * tid];
      c1 = cnt_k[2 * tid + 1];
    }
    uint32_t tot = 0;
    const uint32_t ex0 = ahf::block_excl_scan<uint32_t>(c0 + c1, &tot);
    if (tid < AHF_SLAB_N / 2) {
      s_off[2 * tid] = ex0;
      s_off[2 * tid + 1] = ex0 + c0;
    }
    if (tid == 0) s_off[AHF_SLAB_N] = tot;
    __syncthreads();
    const double czs = ahf_centre(ahf_wrap(Z + ez, L), Ld);
    const TscRt tz(-ez);
    const uint32_t lim = tot < (uint32_t)CAP ? tot : (uint32_t)CAP;
    for (uint32_t e = tid; e < lim; e += AHF_BLOCK) {
      /* the node holding entry e: s_off[lo] <= e < s_off[lo + 1] */
      int lo = 0, hi = AHF_SLAB_N;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid;
        else hi = mid;
      }
      if (s_off[lo] + cnt_k[lo] > (uint32_t)CAP) continue; /* read from the records */
      const ahf_q4 q = qp[start_k[lo] + (e - s_off[lo])];
      const int row = lo / AHF_SLAB_W, col = lo - row * AHF_SLAB_W;
      const double sx = ahf_pn_sep(q.x, s_cx[col], Ld);
      const double sy = ahf_pn_sep(q.y, s_cy[row], Ld);
      s_e[e] = m2pd * tz(ahf_pn_sep(q.z, czs, Ld));
      if constexpr (W9) {
        s_e[1 * CAP + e] = tsc_w_c<-1>(sy);
        s_e[2 * CAP + e] = tsc_w_c<0>(sy);
        s_e[3 * CAP + e] = tsc_w_c<1>(sy);
        s_e[4 * CAP + e] = tsc_w_c<-1>(sx);
        s_e[5 * CAP + e] = tsc_w_c<0>(sx);
        s_e[6 * CAP + e] = tsc_w_c<1>(sx);
      } else {
        s_e[CAP + e] = sy;
        s_e[2 * CAP + e] = sx;
      }
    }
    __syncthreads();
    if (active) {
      const uint32_t m9 = (mask >> (9 * k)) & 0x1FFu;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        if (!((m9 >> t) & 1u)) continue;
        const int ey = oy[t / 3], ex = ox[t % 3];
        const int u = (ly + 1 + ey) * AHF_SLAB_W + (lx + 1 + ex);
        const uint32_t cnt = cnt_k[u], o = s_off[u];
        if (o + cnt <= (uint32_t)CAP) {
          uint32_t q = descending ? o + cnt - 1u : o;
          if constexpr (W9) {
            /* weight index 1 - e for a source at offset e */
            const double *wy = s_e + (2 - ey) * CAP;
            const double *wx = s_e + (5 - ex) * CAP;
            for (uint32_t j = 0; j < cnt; j++, q += step) d = (float)((double)d + (s_e[q] * wy[q]) * wx[q]);
          } else {
            const TscRt fy(-ey), fx(-ex);
            for (uint32_t j = 0; j < cnt; j++, q += step)
              d = (float)((double)d + (s_e[q] * fy(s_e[CAP + q])) * fx(s_e[2 * CAP + q]));
          }
        } else {
          const TscRt fy(-ey), fx(-ex);
          const double cxs = s_cx[lx + 1 + ex], cys = s_cy[ly + 1 + ey];
          uint32_t q = descending ? start_k[u] + cnt - 1u : start_k[u];
          for (uint32_t j = 0; j < cnt; j++, q += step) {
            const ahf_q4 r = qp[q];
            const double az = m2pd * tz(ahf_pn_sep(r.z, czs, Ld));
            d = (float)((double)d + (az * fy(ahf_pn_sep(r.y, cys, Ld))) * fx(ahf_pn_sep(r.x, cxs, Ld)));
          }
        }
      }
    }
    __syncthreads();
  }
  if (active) dens[n] = d;
}

/* shared state of one marching workgroup (k_gather_march) */
struct MarchView {
  const uint32_t *cnt, *start, *off; /* the walked plane's slot */
  const double *e;                   /* its staged entries */
  const double *cx, *cy;             /* source column and row centres */
  const ahf_q4 *qp;
  double m2pd, L, czs;
  uint32_t step;
  int lx, ly;
};

/* the contribution of a staged or recorded entry to the three targets above,
 * at and below the walked plane, in the FP32 values d0 (plane P+1, which sees
 * the source at z offset -1), d1 (plane P) and d2 (plane P-1), each only
 * where its use flag is set */
__device__ __forceinline__ void fold3(float &d0, float &d1, float &d2, bool u0, bool u1, bool u2, double a0, double a1,
                                      double a2, double wy, double wx)
{
  const float n0 = (float)((double)d0 + (a0 * wy) * wx);
  const float n1 = (float)((double)d1 + (a1 * wy) * wx);
  const float n2 = (float)((double)d2 + (a2 * wy) * wx);
  d0 = u0 ? n0 : d0;
  d1 = u1 ? n1 : d1;
  d2 = u2 ? n2 : d2;
}

/*
 * Source (B,E) (traversal ranks along y and x) of the walked plane for the
 * lane's three targets.  Inside the box (FACE false) the offsets are B-1,
 * E-1 and the TSC weights take compile-time forms; at the periodic faces
 * they follow the lane's rotated order.  The z weights of the three targets
 * are those of offsets -1, 0, +1 (TSC index d = 1, 0, -1).
 */
template <int CAP, bool FULL, bool FACE, int B, int E>
__device__ __forceinline__ void march_source(const MarchView &m, const int *oy, const int *ox, uint32_t m0, uint32_t m1,
                                             uint32_t m2, float &d0, float &d1, float &d2)
{
  constexpr int BIT = B * 3 + E;
  const bool u0 = (m0 >> BIT) & 1u, u1 = (m1 >> BIT) & 1u, u2 = (m2 >> BIT) & 1u;
  if (!(u0 || u1 || u2)) return;
  const int ey = FACE ? oy[B] : B - 1, ex = FACE ? ox[E] : E - 1;
  const int u = (m.ly + 1 + ey) * AHF_SLAB_W + (m.lx + 1 + ex);
  const uint32_t cnt = m.cnt[u], o = m.off[u];
  const TscRt fy(-ey), fx(-ex);
  if (o + cnt <= (uint32_t)CAP) {
    uint32_t q = m.step == 1u ? o : o + cnt - 1u;
    for (uint32_t j = 0; j < cnt; j++, q += m.step) {
      double a0, a1, a2, wy, wx;
      if constexpr (FULL) {
        a0 = m.e[q];
        a1 = m.e[CAP + q];
        a2 = m.e[2 * CAP + q];
        wy = m.e[(4 - ey) * CAP + q];
        wx = m.e[(7 - ex) * CAP + q];
      } else {
        const double sz = m.e[q], sy = m.e[CAP + q], sx = m.e[2 * CAP + q];
        a0 = m.m2pd * tsc_w_c<1>(sz);
        a1 = m.m2pd * tsc_w_c<0>(sz);
        a2 = m.m2pd * tsc_w_c<-1>(sz);
        if constexpr (FACE) {
          wy = fy(sy);
          wx = fx(sx);
        } else {
          wy = tsc_w_c<1 - B>(sy);
          wx = tsc_w_c<1 - E>(sx);
        }
      }
      fold3(d0, d1, d2, u0, u1, u2, a0, a1, a2, wy, wx);
    }
  } else {
    const double cxs = m.cx[m.lx + 1 + ex], cys = m.cy[m.ly + 1 + ey];
    uint32_t q = m.step == 1u ? m.start[u] : m.start[u] + cnt - 1u;
    for (uint32_t j = 0; j < cnt; j++, q += m.step) {
      const ahf_q4 r = m.qp[q];
      const double sz = ahf_pn_sep(r.z, m.czs, m.L);
      fold3(d0, d1, d2, u0, u1, u2, m.m2pd * tsc_w_c<1>(sz), m.m2pd * tsc_w_c<0>(sz), m.m2pd * tsc_w_c<-1>(sz),
            fy(ahf_pn_sep(r.y, cys, m.L)), fx(ahf_pn_sep(r.x, cxs, m.L)));
    }
  }
}

template <int CAP, bool FULL, bool FACE>
__device__ __forceinline__ void march_plane(const MarchView &m, const int *oy, const int *ox, uint32_t m0, uint32_t m1,
                                            uint32_t m2, float &d0, float &d1, float &d2)
{
  march_source<CAP, FULL, FACE, 0, 0>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 0, 1>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 0, 2>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 1, 0>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 1, 1>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 1, 2>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 2, 0>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 2, 1>(m, oy, ox, m0, m1, m2, d0, d1, d2);
  march_source<CAP, FULL, FACE, 2, 2>(m, oy, ox, m0, m1, m2, d0, d1, d2);
}

/* presence and entry counts of a lane's nine sources in one staged source
 * plane S, for the three targets that visit it: role 0 is the target S+1
 * (S is its z offset -1), role 1 the target S, role 2 the target S-1.
 * Bit r*9 + B*3 + E of m is set when the source at traversal ranks (B,E)
 * is present for role r (get_TSCnodes) and holds entries; c[r] sums them. */
struct PlaneSum {
  uint32_t m, c0, c1, c2;
};

__device__ __forceinline__ PlaneSum plane_sum(const uint32_t *cnt, const uint32_t *pm, const int *oy, const int *ox,
                                              int lx, int ly)
{
  PlaneSum r = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int bb = 0; bb < 3; bb++)
#pragma unroll
    for (int e = 0; e < 3; e++) {
      const int ey = oy[bb], ex = ox[e];
      const int u = (ly + 1 + ey) * AHF_SLAB_W + (lx + 1 + ex);
      const uint32_t c = cnt[u];
      if (!c) continue;
      const uint32_t w = pm[u] >> ((1 - ey) * 3 + (1 - ex));
      const int t = bb * 3 + e;
      /* role r sees the plane at z offset r-1: presence bit (2-r)*9 + ... */
      if ((w >> 18) & 1u) {
        r.m |= 1u << t;
        r.c0 += c;
      }
      if ((w >> 9) & 1u) {
        r.m |= 1u << (9 + t);
        r.c1 += c;
      }
      if (w & 1u) {
        r.m |= 1u << (18 + t);
        r.c2 += c;
      }
    }
  return r;
}

/*
 * Exact-order density of the planes 1 <= Z <= L-2 of a level that holds
 * every octet, whose sources lie in the planes Z-1, Z, Z+1 in that order.
 * One workgroup per 16x16 column tile and run of zn target planes, one lane
 * per (x,y) column.  The workgroup marches up through the source planes,
 * staging each one in LDS once (as in k_gather_slab), and every lane folds
 * the staged plane P into its three open targets: plane P+1 (its first
 * source plane), P (its second) and P-1 (its third, after which the value
 * is final).  All three visit the plane's sources in the same order, so one
 * walk serves them and the TSC weights along y and x are shared.  Bucket
 * records of the planes P, P+1, P+2 are kept in a ring of three slots, and
 * each lane summarises a plane's sources once when it arrives, so that
 * target P+1's contribution count (for the kmax test) is known when it
 * opens.  FULL stages the nine weights of each entry and walks the plane's
 * sources as one flat sequence from a per-lane queue in LDS (few entries
 * per source: the wavefront runs as long as its longest lane over the
 * plane, not over each source); otherwise the separations are staged and
 * the sources are walked one by one with compile-time weights.
 */
template <int CAP, bool FULL>
__global__ void __launch_bounds__(AHF_BLOCK) k_gather_march(level_view v, int64_t per_xcd, int64_t zn,
                                                            const uint32_t *__restrict__ pmask,
                                                            const int64_t *__restrict__ p0,
                                                            const ahf_q4 *__restrict__ qp, int descending,
                                                            double m2pd, uint32_t kmax, float *__restrict__ dens,
                                                            uint32_t *__restrict__ heavy,
                                                            uint32_t *__restrict__ n_heavy)
{
  constexpr int NV = FULL ? 9 : 3; /* doubles per staged entry */
  __shared__ uint32_t s_start[3 * AHF_SLAB_N], s_cnt[3 * AHF_SLAB_N], s_pm[AHF_SLAB_N];
  __shared__ uint32_t s_off[AHF_SLAB_N + 1];
  __shared__ double s_cx[AHF_SLAB_W], s_cy[AHF_SLAB_W];
  __shared__ double s_e[NV * CAP];
  __shared__ uint32_t s_q[FULL ? 9 * AHF_BLOCK : 1];
  const int64_t L = v.L, H = L >> 1, T = L / AHF_SLAB;
  const int64_t nch = (L - 2 + zn - 1) / zn;
  const int64_t b = blockIdx.x;
  const int64_t w = (b % AHF_XCDS) * per_xcd + b / AHF_XCDS;
  if (w >= T * T * nch) return; /* the whole workgroup */
  const int64_t ch = w / (T * T), ty = (w / T) % T, tx = w % T;
  const int64_t Z0 = 1 + ch * zn, Z1 = Z0 + zn < L - 1 ? Z0 + zn : L - 1;
  const int tid = threadIdx.x, lx = tid & (AHF_SLAB - 1), ly = tid / AHF_SLAB;
  const int64_t X = tx * AHF_SLAB + lx, Y = ty * AHF_SLAB + ly;
  const double Ld = (double)L;
  const bool face = tx == 0 || tx == T - 1 || ty == 0 || ty == T - 1;
  int oy[3], ox[3];
  ahf_axis_order(Y, L, oy);
  ahf_axis_order(X, L, ox);
  if (tid < AHF_SLAB_W) s_cx[tid] = ahf_centre(ahf_wrap(tx * AHF_SLAB - 1 + tid, L), Ld);
  else if (tid < 2 * AHF_SLAB_W)
    s_cy[tid - AHF_SLAB_W] = ahf_centre(ahf_wrap(ty * AHF_SLAB - 1 + (tid - AHF_SLAB_W), L), Ld);
  /* bucket records of source plane Zs into ring slot Zs mod 3 (presence
   * masks into the one slot that plane_sum reads) */
  auto load_plane = [&](int64_t Zs) {
    const int slot = (int)(Zs % 3);
    const int64_t Zw = ahf_wrap(Zs, L);
    for (int i = tid; i < AHF_SLAB_N; i += AHF_BLOCK) {
      const int64_t Xs = ahf_wrap(tx * AHF_SLAB - 1 + i % AHF_SLAB_W, L);
      const int64_t Ys = ahf_wrap(ty * AHF_SLAB - 1 + i / AHF_SLAB_W, L);
      const int64_t sn = full_node(Xs, Ys, Zw, H);
      const int64_t a = p0[sn];
      s_start[slot * AHF_SLAB_N + i] = (uint32_t)a;
      s_cnt[slot * AHF_SLAB_N + i] = (uint32_t)(p0[sn + 1] - a);
      s_pm[i] = pmask[sn]; /* summarised before the next load */
    }
  };
  auto sum_of = [&](int64_t Zs) {
    const int slot = (int)(Zs % 3);
    return plane_sum(s_cnt + slot * AHF_SLAB_N, s_pm, oy, ox, lx, ly);
  };
  load_plane(Z0 - 1);
  __syncthreads();
  PlaneSum sP = sum_of(Z0 - 1);
  __syncthreads();
  load_plane(Z0);
  __syncthreads();
  PlaneSum sP1 = sum_of(Z0);
  __syncthreads();
  MarchView m;
  m.off = s_off;
  m.e = s_e;
  m.cx = s_cx;
  m.cy = s_cy;
  m.qp = qp;
  m.m2pd = m2pd;
  m.L = Ld;
  m.step = descending ? 0xFFFFFFFFu : 1u;
  m.lx = lx;
  m.ly = ly;
  /* the open targets: P+1 (n), P (c), P-1 (p); plane summaries of P, P+1 */
  float dn = -1.0f, dc = -1.0f, dp = -1.0f;
  uint32_t mn = 0, mc = 0, mp = 0;
  bool an = false, ac = false, ap = false;
  for (int64_t P = Z0 - 1; P <= Z1; P++) {
    load_plane(P + 2);
    __syncthreads();
    const PlaneSum sP2 = sum_of(P + 2);
    const int sp = (int)(P % 3);
    const uint32_t *cnt_p = s_cnt + sp * AHF_SLAB_N, *start_p = s_start + sp * AHF_SLAB_N;
    uint32_t c0 = 0, c1 = 0;
    if (tid < AHF_SLAB_N / 2) {
      c0 = cnt_p[2 * tid];
      c1 = cnt_p[2 * tid + 1];
    }
    uint32_t tot = 0;
    const uint32_t ex0 = ahf::block_excl_scan<uint32_t>(c0 + c1, &tot);
    if (tid < AHF_SLAB_N / 2) {
      s_off[2 * tid] = ex0;
      s_off[2 * tid + 1] = ex0 + c0;
    }
    if (tid == 0) s_off[AHF_SLAB_N] = tot;
    __syncthreads();
    const double czs = ahf_centre(ahf_wrap(P, L), Ld);
    if constexpr (FULL) {
      /* one thread per staged entry (the node holding entry e found by a
       * binary search of the offsets): the plane's work spread evenly,
       * whatever the nodes' counts */
      const uint32_t lim = tot < (uint32_t)CAP ? tot : (uint32_t)CAP;
      for (uint32_t e = tid; e < lim; e += AHF_BLOCK) {
        int lo = 0, hi = AHF_SLAB_N;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_off[mid] <= e) lo = mid;
          else hi = mid;
        }
        const int u = lo;
        if (s_off[u] + cnt_p[u] > (uint32_t)CAP) continue; /* read from the records */
        const int row = u / AHF_SLAB_W, col = u - row * AHF_SLAB_W;
        const double cxs = s_cx[col], cys = s_cy[row];
        {
          const ahf_q4 q = qp[start_p[u] + (e - s_off[u])];
          const double sz = ahf_pn_sep(q.z, czs, Ld);
          const double sy = ahf_pn_sep(q.y, cys, Ld);
          const double sx = ahf_pn_sep(q.x, cxs, Ld);
          s_e[e] = m2pd * tsc_w_c<1>(sz);
          s_e[CAP + e] = m2pd * tsc_w_c<0>(sz);
          s_e[2 * CAP + e] = m2pd * tsc_w_c<-1>(sz);
          s_e[3 * CAP + e] = tsc_w_c<-1>(sy);
          s_e[4 * CAP + e] = tsc_w_c<0>(sy);
          s_e[5 * CAP + e] = tsc_w_c<1>(sy);
          s_e[6 * CAP + e] = tsc_w_c<-1>(sx);
          s_e[7 * CAP + e] = tsc_w_c<0>(sx);
          s_e[8 * CAP + e] = tsc_w_c<1>(sx);
        }
      }
    } else {
      const uint32_t lim = tot < (uint32_t)CAP ? tot : (uint32_t)CAP;
      for (uint32_t e = tid; e < lim; e += AHF_BLOCK) {
        /* the node holding entry e: s_off[lo] <= e < s_off[lo + 1] */
        int lo = 0, hi = AHF_SLAB_N;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_off[mid] <= e) lo = mid;
          else hi = mid;
        }
        if (s_off[lo] + cnt_p[lo] > (uint32_t)CAP) continue; /* read from the records */
        const ahf_q4 q = qp[start_p[lo] + (e - s_off[lo])];
        const int row = lo / AHF_SLAB_W, col = lo - row * AHF_SLAB_W;
        s_e[e] = ahf_pn_sep(q.z, czs, Ld);
        s_e[CAP + e] = ahf_pn_sep(q.y, s_cy[row], Ld);
        s_e[2 * CAP + e] = ahf_pn_sep(q.x, s_cx[col], Ld);
      }
    }
    /* target P+1 opens: planes P (its z offset -1), P+1 (0), P+2 (+1) */
    mn = 0;
    an = false;
    dn = -1.0f;
    if (P + 1 >= Z0 && P + 1 < Z1) {
      const uint64_t total = (uint64_t)sP.c0 + sP1.c1 + sP2.c2;
      an = total <= kmax;
      if (an) mn = (sP.m & 0x1FFu) | (sP1.m & (0x1FFu << 9)) | (sP2.m & (0x1FFu << 18));
      else heavy[atomicAdd(n_heavy, 1u)] = (uint32_t)full_node(X, Y, P + 1, H);
    }
    __syncthreads();
    m.cnt = cnt_p;
    m.start = start_p;
    m.czs = czs;
    /* plane P is the first source plane of P+1, the second of P, the third of P-1 */
    const uint32_t m0 = mn & 0x1FFu, m1 = ac ? (mc >> 9) & 0x1FFu : 0u, m2 = ap ? (mp >> 18) & 0x1FFu : 0u;
    if constexpr (FULL) {
      /* Rows fast path (ascending lists, every source staged): the three
       * sources of one source row are adjacent nodes of the staged plane, so
       * their entries form one range of the staged order, walked from the
       * first source in the lane's x order and wrapping round the range
       * (the order at the periodic faces is a rotation); the source of an
       * entry, and so its x weight, follows from the range's inner
       * boundaries.  A lane qualifies when each of its three targets takes
       * either none or all of the plane's non-empty sources, and the
       * wavefront takes the path only if all its lanes qualify. */
      bool rows = false;
      const int rot = ox[0] + 1; /* first source in x order: 0, 1 or 2 along a row */
      if (tot <= (uint32_t)CAP && m.step == 1u) {
        uint32_t ne9 = 0;
#pragma unroll
        for (int bb = 0; bb < 3; bb++) {
          const int u = (ly + 1 + oy[bb]) * AHF_SLAB_W + lx;
          const uint32_t o0 = s_off[u], o1 = s_off[u + 1], o2 = s_off[u + 2], o3 = s_off[u + 3];
#pragma unroll
          for (int e = 0; e < 3; e++) {
            const int ex = ox[e];
            const uint32_t ce = ex < 0 ? o1 - o0 : (ex == 0 ? o2 - o1 : o3 - o2);
            if (ce) ne9 |= 1u << (bb * 3 + e);
          }
        }
        const bool ok = (m0 == 0u || m0 == ne9) && (m1 == 0u || m1 == ne9) && (m2 == 0u || m2 == ne9);
        rows = __ballot(!ok) == 0ull;
      }
      if (rows) {
        const bool r0 = m0 != 0u, r1 = m1 != 0u, r2 = m2 != 0u;
        const int oy0 = oy[0], oy1 = oy[1], oy2 = oy[2];
        uint32_t ntot = 0;
#pragma unroll
        for (int bb = 0; bb < 3; bb++) {
          const int u = (ly + 1 + oy[bb]) * AHF_SLAB_W + lx;
          ntot += s_off[u + 3] - s_off[u];
        }
        /* the current row: next entry q, entries left, range [lo, hi) with
         * inner boundaries b1, b2, and the offset of its y weight */
        uint32_t q = 0, rem = 0, lo = 0, hi = 0, b1 = 0, b2 = 0, y3 = 0;
        int br = -1;
        for (uint32_t it = 0; it < ntot; it++) {
          while (rem == 0u) {
            br++;
            const int ey = br == 0 ? oy0 : (br == 1 ? oy1 : oy2);
            const int u = (ly + 1 + ey) * AHF_SLAB_W + lx;
            lo = s_off[u];
            b1 = s_off[u + 1];
            b2 = s_off[u + 2];
            hi = s_off[u + 3];
            rem = hi - lo;
            q = ahf_row_start(lo, b1, b2, hi, rot);
            y3 = (3u + (uint32_t)(1 - ey)) * (uint32_t)CAP;
          }
          /* x weight index 1 - ex of the entry's source: 8, 7, 6 along the row */
          const uint32_t wx = (8u - ahf_row_src(q, b1, b2)) * (uint32_t)CAP;
          fold3(dn, dc, dp, r0, r1, r2, s_e[q], s_e[CAP + q], s_e[2 * CAP + q], s_e[y3 + q], s_e[wx + q]);
          q = ahf_row_next(q, lo, hi);
          rem--;
        }
      } else {
      /* the lane's sources of this plane in order, as one flat sequence */
      uint32_t *qq = s_q + tid * 9;
      if (tot <= (uint32_t)CAP) {
        /* every source of the plane is staged (the common case): LDS slot
         * (10 bits), count (12 bits; a light target's sources hold at most
         * kmax < 4096 entries), roles (3), weight indices 1-ey and 1-ex;
         * the next source's word is read one switch ahead */
        uint32_t nq = 0, ntot = 0;
#pragma unroll
        for (int t = 0; t < 9; t++) {
          const uint32_t roles = ((m0 >> t) & 1u) | ((m1 >> t) & 1u) << 1 | ((m2 >> t) & 1u) << 2;
          if (!roles) continue;
          const int ey = oy[t / 3], ex = ox[t % 3];
          const int u = (ly + 1 + ey) * AHF_SLAB_W + (lx + 1 + ex);
          const uint32_t c = cnt_p[u];
          qq[nq++] = s_off[u] | c << 10 | roles << 22 | (uint32_t)(1 - ey) << 25 | (uint32_t)(1 - ex) << 27;
          ntot += c;
        }
        uint32_t k = 0, rem = 0, q = 0, nxt = qq[0];
        uint32_t oy3 = 0, ox3 = 0;
        bool r0 = false, r1 = false, r2 = false;
        for (uint32_t it = 0; it < ntot; it++) {
          if (rem == 0) {
            const uint32_t wq = nxt;
            k++;
            nxt = qq[k < 9u ? k : 8u];
            rem = (wq >> 10) & 0xFFFu;
            q = m.step == 1u ? (wq & 0x3FFu) : (wq & 0x3FFu) + rem - 1u;
            r0 = (wq >> 22) & 1u;
            r1 = (wq >> 23) & 1u;
            r2 = (wq >> 24) & 1u;
            oy3 = (3u + ((wq >> 25) & 3u)) * (uint32_t)CAP;
            ox3 = (6u + ((wq >> 27) & 3u)) * (uint32_t)CAP;
          }
          fold3(dn, dc, dp, r0, r1, r2, s_e[q], s_e[CAP + q], s_e[2 * CAP + q], s_e[oy3 + q], s_e[ox3 + q]);
          q += m.step;
          rem--;
        }
      } else {
      uint32_t nq = 0, ntot = 0;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const uint32_t roles = ((m0 >> t) & 1u) | ((m1 >> t) & 1u) << 1 | ((m2 >> t) & 1u) << 2;
        if (!roles) continue;
        const int ey = oy[t / 3], ex = ox[t % 3];
        const uint32_t u = (uint32_t)((ly + 1 + ey) * AHF_SLAB_W + (lx + 1 + ex));
        const uint32_t c = cnt_p[u], o = s_off[u];
        const uint32_t staged = o + c <= (uint32_t)CAP;
        /* LDS slot or node (10 bits), count (12), roles (3), weight indices
         * 1-ey and 1-ex (2 + 2), staged (1) */
        qq[nq++] = (staged ? o : u) | c << 10 | roles << 22 | (uint32_t)(1 - ey) << 25 | (uint32_t)(1 - ex) << 27 |
                   staged << 29;
        ntot += c;
      }
      uint32_t k = 0, rem = 0, q = 0;
      int iy = 0, ix = 0;
      bool r0 = false, r1 = false, r2 = false, stg = true;
      double cys = 0.0, cxs = 0.0;
      for (uint32_t it = 0; it < ntot; it++) {
        if (rem == 0) {
          const uint32_t wq = qq[k++];
          rem = (wq >> 10) & 0xFFFu;
          r0 = (wq >> 22) & 1u;
          r1 = (wq >> 23) & 1u;
          r2 = (wq >> 24) & 1u;
          iy = (int)((wq >> 25) & 3u);
          ix = (int)((wq >> 27) & 3u);
          stg = (wq >> 29) & 1u;
          const uint32_t first = stg ? (wq & 0x3FFu) : start_p[wq & 0x3FFu];
          q = m.step == 1u ? first : first + rem - 1u;
          if (!stg) {
            cys = s_cy[ly + 2 - iy];
            cxs = s_cx[lx + 2 - ix];
          }
        }
        if (stg) {
          fold3(dn, dc, dp, r0, r1, r2, s_e[q], s_e[CAP + q], s_e[2 * CAP + q], s_e[(3 + iy) * CAP + q],
                s_e[(6 + ix) * CAP + q]);
        } else {
          const ahf_q4 r = qp[q];
          const double sz = ahf_pn_sep(r.z, czs, Ld);
          fold3(dn, dc, dp, r0, r1, r2, m2pd * tsc_w_c<1>(sz), m2pd * tsc_w_c<0>(sz), m2pd * tsc_w_c<-1>(sz),
                TscRt(iy - 1)(ahf_pn_sep(r.y, cys, Ld)), TscRt(ix - 1)(ahf_pn_sep(r.x, cxs, Ld)));
        }
        q += m.step;
        rem--;
      }
      }
      }
    } else {
      if (face) march_plane<CAP, FULL, true>(m, oy, ox, m0, m1, m2, dn, dc, dp);
      else march_plane<CAP, FULL, false>(m, oy, ox, m0, m1, m2, dn, dc, dp);
    }
    if (ap) dens[full_node(X, Y, P - 1, H)] = dp;
    dp = dc;
    dc = dn;
    mp = mc;
    mc = mn;
    ap = ac;
    ac = an;
    sP = sP1;
    sP1 = sP2;
    __syncthreads();
  }
}

/* inclusive prefix sum over the 64 lanes of a wavefront by DPP row shifts
 * (within each row of 16) and row broadcasts of lanes 15 and 31 into the
 * rows above (the GFX9 scan; register-only) */
__device__ __forceinline__ int32_t wave_incl_scan_dpp(int32_t v)
{
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); /* row_shr:1 */
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); /* row_shr:2 */
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); /* row_shr:4 */
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); /* row_shr:8 */
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); /* row_bcast:15 into rows 1, 3 */
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); /* row_bcast:31 into rows 2, 3 */
  return v;
}

__device__ inline double readlane_d(double x, int l)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

__device__ inline int64_t readlane_i64(int64_t x, int l)
{
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)x, l);
  const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), l);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

/*
 * AHF's d = (float)((double)d + c) over the contributions cv of lanes
 * [0, mm), in lane order, for the whole wavefront (d and mm uniform).
 * Inside one binade of d the steps are integer increments of d/ulp that
 * depend on c alone (ahf_fold_lane), so a run of them is one prefix sum over
 * the lanes; exact ties take the parity of the running value, lane by lane.
 * The first lane whose step leaves the binade (or whose start is zero or
 * subnormal) takes the plain FP32 step, and the fold resumes after it.
 * Every round settles at least one lane.
 */
__device__ float fold_wave(float d, double cv, int mm)
{
  const int lane = threadIdx.x & 63;
  int start = 0;
  while (start < mm) {
    int ok;
    const int e = ahf_fold_exponent(d, &ok);
    if (!ok) {
      d = (float)((double)d + readlane_d(cv, start));
      start++;
      continue;
    }
    const int neg = d < 0.0f;
    const int64_t K0 = ahf_fold_K(d, e);
    const bool act = lane >= start && lane < mm;
    int64_t R = 0, r = 0;
    int tie = 0, lok = 0;
    if (act) lok = ahf_fold_lane(cv, e, &R, &r, &tie);
    uint64_t tmask = __ballot(act && lok && tie);
    while (tmask) {
      const int l = __ffsll((unsigned long long)tmask) - 1;
      const uint64_t below = __ballot(act && (r & 1)) & ((1ull << l) - 1);
      const int par = (int)((K0 & 1) ^ (int64_t)(__popcll(below) & 1));
      const int64_t rl = readlane_i64(r, l);
      if (lane == l && ((par + (int)(rl & 1)) & 1)) r += 1;
      tmask &= tmask - 1;
    }
    /* a step that stays in the binade has |r| < 2^23, so r clamped to
     * +-2^24 keeps every prefix up to the first step that leaves it exact,
     * and 64 clamped steps fit in 32 bits */
    const int32_t rc = (int32_t)(r > ((int64_t)1 << 24) ? ((int64_t)1 << 24) : (r < -((int64_t)1 << 24) ? -((int64_t)1 << 24) : r));
    const int32_t Pi = wave_incl_scan_dpp(act ? rc : 0);
    const int64_t Kin = K0 + (int64_t)(Pi - (act ? rc : 0));
    const int valid = act && lok && ahf_fold_K_in(Kin, neg) && ahf_fold_T_in(Kin, R, neg);
    const uint64_t badm = __ballot(act && !valid);
    const int bad = badm ? __ffsll((unsigned long long)badm) - 1 : mm;
    if (bad > start) d = (float)ldexp((double)(K0 + (int64_t)__builtin_amdgcn_readlane(Pi, bad - 1)), e - 23);
    if (bad < mm) {
      d = (float)((double)d + readlane_d(cv, bad));
      start = bad + 1;
    } else {
      start = mm;
    }
  }
  return d;
}

/*
 * The contributing sources of heavy target n, flattened: lane j < 27 holds
 * source j of AHF's visiting order (ahf_axis_order on each axis, z outer),
 * its first entry q0, its entry count (0 when absent or masked out) and the
 * exclusive prefix of the counts; the return value is the total.
 */
struct HeavySrc {
  int64_t q0, cnt, off;
  int ex, ey, ez;
};

__device__ inline int64_t heavy_sources(const level_view &v, const uint32_t *__restrict__ pmask,
                                        const int64_t *__restrict__ p0, int64_t n, HeavySrc *hs)
{
  const int lane = threadIdx.x & 63;
  const int64_t o = n >> 3;
  const int c = (int)(n & 7);
  hs->q0 = 0;
  hs->cnt = 0;
  hs->ex = hs->ey = hs->ez = 0;
  if (lane < 27) {
    int64_t X, Y, Z;
    ahf_node_xyz(&v, o, c, &X, &Y, &Z);
    int oz[3], oy[3], ox[3];
    ahf_axis_order(Z, v.L, oz);
    ahf_axis_order(Y, v.L, oy);
    ahf_axis_order(X, v.L, ox);
    hs->ez = oz[lane / 9];
    hs->ey = oy[(lane / 3) % 3];
    hs->ex = ox[lane % 3];
    const int64_t s = ahf_nbr_node(v.nbr, o, c, hs->ex, hs->ey, hs->ez);
    if (s >= 0 && ((pmask[s] >> ((1 - hs->ez) * 9 + (1 - hs->ey) * 3 + (1 - hs->ex))) & 1u)) {
      hs->q0 = p0[s];
      hs->cnt = p0[s + 1] - hs->q0;
    }
  }
  int64_t inc = hs->cnt;
#pragma unroll
  for (int d = 1; d < 32; d <<= 1) {
    const int64_t t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  hs->off = inc - hs->cnt;
  return readlane_i64(inc, 26);
}

/*
 * The source (lane of heavy_sources) and entry of flat contribution k =
 * base + lane, for one round of 64.  j0, uniform, is the source holding
 * contribution base and is advanced round by round; the few sources that
 * start inside the round are taken from their lanes with readlane (scalar
 * reads, no cross-lane shuffles).  A source holds k when it is the last
 * whose prefix is at most k (empty sources share their prefix with the next).
 */
struct HeavyLane {
  int64_t q;
  int ex, ey, ez;
  bool in;
};

__device__ inline HeavyLane heavy_lane(const HeavySrc &hs, int64_t total, int64_t base, int &j0, int descending)
{
  const int lane = threadIdx.x & 63;
  const int64_t k = base + lane;
  while (j0 < 26 && readlane_i64(hs.off, j0 + 1) <= base) j0++;
  const int64_t kend = (base + 63 < total - 1) ? base + 63 : total - 1;
  const int code = hs.ex + 1 + 3 * (hs.ey + 1) + 9 * (hs.ez + 1);
  int64_t q0 = readlane_i64(hs.q0, j0), cnt = readlane_i64(hs.cnt, j0), off = readlane_i64(hs.off, j0);
  int cd = __builtin_amdgcn_readlane(code, j0);
  for (int j = j0 + 1; j < 27; j++) {
    const int64_t oj = readlane_i64(hs.off, j);
    if (oj > kend) break;
    const int64_t qj = readlane_i64(hs.q0, j), cj = readlane_i64(hs.cnt, j);
    const int dj = __builtin_amdgcn_readlane(code, j);
    const bool mine = k >= oj;
    q0 = mine ? qj : q0;
    cnt = mine ? cj : cnt;
    off = mine ? oj : off;
    cd = mine ? dj : cd;
  }
  HeavyLane r;
  const int64_t kk = k - off;
  r.in = k < total;
  r.q = descending ? q0 + cnt - 1 - kk : q0 + kk;
  r.ex = cd % 3 - 1;
  r.ey = (cd / 3) % 3 - 1;
  r.ez = cd / 9 - 1;
  return r;
}

/* exact-order density of the heavy targets, one wavefront per target: the
 * node's contributions, source by source in AHF's order and each source's
 * entries in list order, are taken 64 at a time (one per lane) and folded
 * into the FP32 node value (fold_wave) */
__global__ void __launch_bounds__(AHF_BLOCK) k_gather_heavy(level_view v, const uint32_t *__restrict__ pmask,
                                                            const int64_t *__restrict__ p0,
                                                            const ahf_w9 *__restrict__ W, int descending,
                                                            float *__restrict__ dens,
                                                            const uint32_t *__restrict__ heavy,
                                                            const uint32_t *__restrict__ n_heavy)
{
  const int lane = threadIdx.x & 63;
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t nh = (int64_t)*n_heavy;
  for (int64_t i = gw; i < nh; i += nw) {
    const int64_t n = heavy[i];
    HeavySrc hs;
    const int64_t total = heavy_sources(v, pmask, p0, n, &hs);
    float d = -1.0f;
    /* the next round's weights are loaded before this round is folded */
    int j0 = 0;
    auto fetch = [&](int64_t base, double &b0, double &b1, double &b2) {
      const HeavyLane hl = heavy_lane(hs, total, base, j0, descending);
      b0 = b1 = b2 = 0.0;
      if (hl.in) {
        const double *w = reinterpret_cast<const double *>(W + hl.q);
        b0 = w[1 - hl.ez];
        b1 = w[3 + (1 - hl.ey)];
        b2 = w[6 + (1 - hl.ex)];
      }
    };
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (total > 0) fetch(0, a0, a1, a2);
    for (int64_t base = 0; base < total; base += 64) {
      /* lanes past the end fold nothing (mm below) */
      const double cv = (a0 * a1) * a2;
      double b0 = 0.0, b1 = 0.0, b2 = 0.0;
      if (base + 64 < total) fetch(base + 64, b0, b1, b2);
      d = fold_wave(d, cv, (int)(total - base < 64 ? total - base : 64));
      a0 = b0;
      a1 = b1;
      a2 = b2;
    }
    if (lane == 0) dens[n] = d;
  }
}

/* k_gather_heavy over position records (dense levels) */
__global__ void __launch_bounds__(AHF_BLOCK) k_gather_heavy_q(level_view v, const uint32_t *__restrict__ pmask,
                                                            const int64_t *__restrict__ p0,
                                                            const ahf_q4 *__restrict__ qp, int descending,
                                                            double m2pd, float *__restrict__ dens,
                                                            const uint32_t *__restrict__ heavy,
                                                            const uint32_t *__restrict__ n_heavy)
{
  const int lane = threadIdx.x & 63;
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t nh = (int64_t)*n_heavy;
  const double L = (double)v.L;
  for (int64_t i = gw; i < nh; i += nw) {
    const int64_t n = heavy[i];
    int64_t X, Y, Z;
    ahf_node_xyz(&v, n >> 3, (int)(n & 7), &X, &Y, &Z);
    double cxv[3], cyv[3], czv[3];
    ahf_centres3(X, v.L, cxv);
    ahf_centres3(Y, v.L, cyv);
    ahf_centres3(Z, v.L, czv);
    HeavySrc hs;
    const int64_t total = heavy_sources(v, pmask, p0, n, &hs);
    float d = -1.0f;
    /* the next round's records are loaded before this round is folded */
    int ex = 0, ey = 0, ez = 0;
    bool in = false;
    ahf_q4 e4 = {0.0f, 0.0f, 0.0f, 0.0f};
    int j0 = 0;
    auto fetch = [&](int64_t base, ahf_q4 &rec, int &fx, int &fy, int &fz, bool &ok) {
      const HeavyLane hl = heavy_lane(hs, total, base, j0, descending);
      fx = hl.ex;
      fy = hl.ey;
      fz = hl.ez;
      ok = hl.in;
      if (ok) rec = qp[hl.q];
    };
    if (total > 0) fetch(0, e4, ex, ey, ez, in);
    for (int64_t base = 0; base < total; base += 64) {
      const double cv = in ? ahf_contrib(m2pd, ahf_pn_sep(e4.x, cxv[ex + 1], L), ahf_pn_sep(e4.y, cyv[ey + 1], L),
                                         ahf_pn_sep(e4.z, czv[ez + 1], L), -ex, -ey, -ez)
                           : 0.0;
      ahf_q4 n4 = {0.0f, 0.0f, 0.0f, 0.0f};
      int nx = 0, ny = 0, nz = 0;
      bool nin = false;
      if (base + 64 < total) fetch(base + 64, n4, nx, ny, nz, nin);
      d = fold_wave(d, cv, (int)(total - base < 64 ? total - base : 64));
      e4 = n4;
      ex = nx;
      ey = ny;
      ez = nz;
      in = nin;
    }
    if (lane == 0) dens[n] = d;
  }
}

/*
 * ahf_node_flags for the 8 children of one octet per thread: a child with
 * all 27 neighbours present is interior, and tested when one of its 18
 * neighbours at x offsets 0, +1 reaches the density threshold.  The
 * children's neighbours lie in the 3x4x4 nodes from (X0, Y0-1, Z0) (X0,
 * Y0, Z0 the octet's first node), each density read once.  Missing
 * neighbour octets only occur for children that are not interior.
 */
__global__ void k_flags(level_view v, const uint32_t *__restrict__ pmask, const float *__restrict__ dens,
                        double thr, uint8_t *__restrict__ flags)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= v.n_oct) return;
  const uint4 *pm4 = reinterpret_cast<const uint4 *>(pmask + 8 * o);
  const uint4 pa = pm4[0], pb = pm4[1];
  const uint32_t pm[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
  uint32_t any = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) any |= (pm[c] == AHF_FULL27) ? 1u << c : 0u;
  uint64_t hit = 0; /* bit xi + 3*(yi + 4*zi) of the 3x4x4 neighbourhood */
  if (any) {
    int32_t nb[27];
#pragma unroll
    for (int i = 0; i < 27; i++) nb[i] = v.nbr[27 * o + i];
    /* every load issued at once: a missing octet reads this octet's node
     * and its bit is dropped */
    float dv[48];
#pragma unroll
    for (int zi = 0; zi < 4; zi++)
#pragma unroll
      for (int yi = 0; yi < 4; yi++)
#pragma unroll
        for (int xi = 0; xi < 3; xi++) {
          const int zr = zi - 1, yr = yi - 1, xr = xi;
          const int oz = zr < 0 ? -1 : (zr > 1 ? 1 : 0), oy = yr < 0 ? -1 : (yr > 1 ? 1 : 0), oxo = xr > 1 ? 1 : 0;
          const int32_t on = nb[ahf_e27(oxo, oy, oz)];
          const int64_t m = 8 * (int64_t)(on >= 0 ? on : (int32_t)o) + ((zr & 1) << 2 | (yr & 1) << 1 | (xr & 1));
          dv[xi + 3 * (yi + 4 * zi)] = dens[m];
        }
#pragma unroll
    for (int zi = 0; zi < 4; zi++)
#pragma unroll
      for (int yi = 0; yi < 4; yi++)
#pragma unroll
        for (int xi = 0; xi < 3; xi++) {
          const int zr = zi - 1, yr = yi - 1, xr = xi;
          const int oz = zr < 0 ? -1 : (zr > 1 ? 1 : 0), oy = yr < 0 ? -1 : (yr > 1 ? 1 : 0), oxo = xr > 1 ? 1 : 0;
          const int bit = xi + 3 * (yi + 4 * zi);
          if (nb[ahf_e27(oxo, oy, oz)] >= 0 && (double)dv[bit] >= thr) hit |= 1ull << bit;
        }
  }
  uint8_t f[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
    /* xi in {cx, cx+1}, yi in {cy..cy+2}, zi in {cz..cz+2} */
    uint64_t box = 0;
#pragma unroll
    for (int zi = 0; zi < 3; zi++)
#pragma unroll
      for (int yi = 0; yi < 3; yi++) box |= 3ull << (cx + 3 * (cy + yi + 4 * (cz + zi)));
    f[c] = !((any >> c) & 1u) ? (uint8_t)0 : (uint8_t)(NF_INTERIOR | ((hit & box) ? NF_TEST : 0u));
  }
  uint2 w;
  w.x = (uint32_t)f[0] | (uint32_t)f[1] << 8 | (uint32_t)f[2] << 16 | (uint32_t)f[3] << 24;
  w.y = (uint32_t)f[4] | (uint32_t)f[5] << 8 | (uint32_t)f[6] << 16 | (uint32_t)f[7] << 24;
  *reinterpret_cast<uint2 *>(flags + 8 * o) = w;
}

/* one thread per level row (octet row r, cy, cz) */
__global__ void k_refine_rows(level_view v, uint8_t *__restrict__ flags, int64_t *__restrict__ cnt,
                              int32_t *__restrict__ anomaly)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * v.n_row) return;
  const int64_t r = t >> 2;
  const int cy = (int)(t & 1), cz = (int)((t >> 1) & 1);
  int64_t c = 0;
  if (ahf_plane_tested(&v, v.row_plane[r], cz) && ahf_row_tested(&v, r, cy)) {
    int an = 0;
    c = ahf_refine_row(&v, flags, r, cy, cz, &an);
    if (an) atomicOr(anomaly, 1);
  }
  cnt[ahf_row_rank(&v, r, cy, cz)] = c;
}

/* has-flag of every level plane (octet plane p, cz) */
__global__ void k_plane_has(level_view v, const int64_t *__restrict__ cnt, int32_t *__restrict__ has)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * v.n_plane) return;
  const int64_t p = t >> 1;
  const int cz = (int)(t & 1);
  const int64_t q0 = v.plane_row0[p], nr = v.plane_row0[p + 1] - q0;
  const int64_t r0 = 4 * q0 + cz * 2 * nr;
  int32_t h = 0;
  for (int64_t k = 0; k < 2 * nr && !h; k++) h = cnt[r0 + k] > 0;
  has[t] = h;
}

__global__ void k_nonzero(const int64_t *__restrict__ cnt, int64_t n, int32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cnt[i] > 0;
}

/* writes the next level's octets, octet rows and node2oct of this level */
__global__ void k_compact(level_view v, const uint8_t *__restrict__ flags, const int64_t *__restrict__ cnt,
                          const int64_t *__restrict__ oct_off, const int32_t *__restrict__ row_off,
                          const int32_t *__restrict__ plane_off, uint64_t *__restrict__ key2,
                          uint8_t *__restrict__ link2, int64_t *__restrict__ parent2, uint64_t *__restrict__ rkey2,
                          int32_t *__restrict__ roct2, int32_t *__restrict__ rplane2, int32_t *__restrict__ node2oct)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * v.n_row) return;
  const int64_t r = t >> 2;
  const int cy = (int)(t & 1), cz = (int)((t >> 1) & 1);
  const int64_t rank = ahf_row_rank(&v, r, cy, cz);
  if (cnt[rank] == 0) return;
  const int64_t p = v.row_plane[r];
  const int64_t Y = 2 * ahf_row_y(v.row_key[r]) + cy, Z = 2 * (int64_t)v.plane_z[p] + cz;
  const int64_t nr = row_off[rank];
  int64_t w = oct_off[rank];
  rkey2[nr] = (uint64_t)Y | ((uint64_t)Z << AHF_KEY_BITS);
  roct2[nr] = (int32_t)w;
  rplane2[nr] = plane_off[2 * p + cz];
  for (int64_t o = v.row_oct0[r]; o < v.row_oct0[r + 1]; o++)
    for (int cx = 0; cx < 2; cx++) {
      const int64_t nd = 8 * o + cz * 4 + cy * 2 + cx;
      const uint8_t f = flags[nd];
      if (!(f & NF_REFINED)) continue;
      const int64_t X = 2 * ahf_kx(v.oct_key[o]) + cx;
      key2[w] = ahf_pack3((uint64_t)X, (uint64_t)Y, (uint64_t)Z);
      link2[w] = (f & NF_LINK) ? 1 : 0;
      parent2[w] = nd;
      node2oct[nd] = (int32_t)w;
      w++;
    }
}

__global__ void k_plane_fill(level_view v, const int32_t *__restrict__ has, const int32_t *__restrict__ plane_off,
                             const int32_t *__restrict__ row_off, int32_t *__restrict__ pz2,
                             int32_t *__restrict__ prow2)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * v.n_plane || !has[t]) return;
  const int64_t p = t >> 1;
  const int cz = (int)(t & 1);
  const int64_t q0 = v.plane_row0[p], nr = v.plane_row0[p + 1] - q0;
  const int32_t np = plane_off[t];
  pz2[np] = 2 * v.plane_z[p] + cz;
  prow2[np] = row_off[4 * q0 + cz * 2 * nr];
}

/* the refine step's sizes gathered for one read-back */
__global__ void k_sizes(const int64_t *__restrict__ n_oct, const int64_t *__restrict__ n_part,
                        const int32_t *__restrict__ n_row, const int32_t *__restrict__ n_plane,
                        const int32_t *__restrict__ anomaly, int64_t *__restrict__ out)
{
  out[0] = *n_oct;
  out[1] = *n_part;
  out[2] = *n_row;
  out[3] = *n_plane;
  out[4] = *anomaly;
}

__global__ void k_set2(int32_t *__restrict__ a, int64_t ia, int32_t va, int32_t *__restrict__ b, int64_t ib,
                       int32_t vb)
{
  a[ia] = va;
  b[ib] = vb;
}

__global__ void k_nbr_next(const int32_t *__restrict__ nbr, const int64_t *__restrict__ parent2,
                           const int32_t *__restrict__ node2oct, int64_t n2, int32_t *__restrict__ nbr2)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 27 * n2) return;
  const int64_t o2 = t < (int64_t)UINT32_MAX ? (int64_t)((uint32_t)t / 27u) : t / 27;
  const int e = (int)(t - 27 * o2);
  const int ex = e % 3 - 1, ey = (e / 3) % 3 - 1, ez = e / 9 - 1;
  const int64_t pn = parent2[o2];
  const int64_t m = ahf_nbr_node(nbr, pn >> 3, (int)(pn & 7), ex, ey, ez);
  nbr2[t] = m < 0 ? -1 : node2oct[m];
}

/* relink (src/libamr_serial/relink.c:31-288): per bucket entry, the fine
 * child its particle moves to (first interior child, in fine traversal
 * order, whose closed cell holds it); the key for the stable sort that
 * builds the fine buckets is the fine node, or all-ones if it stays */
/* imask2: interior children of each fine octet (k_pmask); qp: the
 * entries' position records when the level has them, else the particle
 * arrays are read */
__global__ void k_relink_child(int64_t n_ent, const uint32_t *__restrict__ pnode,
                               const uint32_t *__restrict__ plist, const int32_t *__restrict__ node2oct,
                               const uint64_t *__restrict__ key, const uint8_t *__restrict__ imask2,
                               const ahf_q4 *__restrict__ qp, const float *__restrict__ px,
                               const float *__restrict__ py, const float *__restrict__ pz, double Lf,
                               int8_t *__restrict__ child)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_ent) return;
  const uint32_t m = pnode[q];
  const int32_t o2 = node2oct[m];
  int ch = -1;
  if (o2 >= 0) {
    const uint8_t imask = imask2[o2];
    if (imask) {
      const uint64_t mk = key[m >> 3];
      const int cm = (int)(m & 7);
      const int64_t MX = 2 * ahf_kx(mk) + (cm & 1), MY = 2 * ahf_ky(mk) + ((cm >> 1) & 1),
                    MZ = 2 * ahf_kz(mk) + (cm >> 2);
      float x, y, z;
      if (qp) {
        const ahf_q4 r = qp[q];
        x = r.x;
        y = r.y;
        z = r.z;
      } else {
        const uint32_t p = plist[q];
        x = px[p];
        y = py[p];
        z = pz[p];
      }
      ch = ahf_relink_child_h(x, y, z, MX, MY, MZ, Lf, imask);
    }
  }
  child[q] = (int8_t)ch;
}

/*
 * Fine buckets without a sort: a particle only moves to a child of its own
 * node, so fine node t = 8*o2 + c holds the entries of its mother's bucket
 * whose child is c, in bucket order (AHF's head insertion during relink
 * leaves them in this order relative to the walk).  One thread per fine
 * octet walks the mother's bucket once to count its 8 children; after a scan,
 * each entry finds its place from the earlier entries of its bucket.  Buckets of more
 * than AHF_BIG_BUCKET entries are listed for k_fine_big, which walks them
 * one wavefront at a time.
 */
#define AHF_BIG_BUCKET 32

__global__ void k_fine_count(const int64_t *__restrict__ oct_parent, int64_t n2, const int64_t *__restrict__ p0c,
                             const int8_t *__restrict__ child, int64_t *__restrict__ cnt,
                             uint32_t *__restrict__ big, uint32_t *__restrict__ n_big)
{
  const int64_t t8 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t8 >= n2) return;
  const int64_t m = oct_parent[t8];
  const int64_t a = p0c[m], b = p0c[m + 1];
  if (b - a > AHF_BIG_BUCKET) {
    big[atomicAdd(n_big, 1u)] = (uint32_t)t8;
    return;
  }
  /* the 8 children's counts as 16-bit fields (at most AHF_BIG_BUCKET) */
  uint64_t lo = 0, hi = 0;
  for (int64_t q = a; q < b; q++) {
    const int c = child[q];
    if (c < 0) continue;
    if (c < 4) lo += 1ull << (16 * c);
    else hi += 1ull << (16 * (c - 4));
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    cnt[8 * t8 + c] = (int64_t)((lo >> (16 * c)) & 0xFFFFu);
    cnt[8 * t8 + 4 + c] = (int64_t)((hi >> (16 * c)) & 0xFFFFu);
  }
}

/* one thread per coarse bucket entry: its place in the fine bucket of its
 * child is the fine bucket's start plus the number of earlier entries of
 * the mother's bucket with the same child (buckets over AHF_BIG_BUCKET
 * entries are k_fine_big's) */
__global__ void k_fine_fill(int64_t n_ent, const uint32_t *__restrict__ pnode, const int32_t *__restrict__ node2oct,
                            const int64_t *__restrict__ p0c, const int8_t *__restrict__ child,
                            const uint32_t *__restrict__ plistc, const int64_t *__restrict__ p0f,
                            uint32_t *__restrict__ plistf, uint32_t *__restrict__ pnodef)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_ent) return;
  const int c = child[q];
  if (c < 0) return;
  const uint32_t m = pnode[q];
  const int64_t a = p0c[m];
  if (p0c[m + 1] - a > AHF_BIG_BUCKET) return;
  const int64_t t = 8 * (int64_t)node2oct[m] + c;
  int64_t rank = 0;
  for (int64_t k = a; k < q; k++) rank += child[k] == c;
  const int64_t w = p0f[t] + rank;
  plistf[w] = plistc[q];
  pnodef[w] = (uint32_t)t;
}

/* the listed big buckets, one wavefront per bucket: 64 entries per step,
 * each ranked among the earlier entries of its child by ballots.  FILL =
 * false writes the 8 counts, FILL = true the entries. */
template <bool FILL>
__global__ void __launch_bounds__(AHF_BLOCK) k_fine_big(const int64_t *__restrict__ oct_parent,
                                                        const int64_t *__restrict__ p0c,
                                                        const int8_t *__restrict__ child,
                                                        const uint32_t *__restrict__ plistc,
                                                        const uint32_t *__restrict__ big,
                                                        const uint32_t *__restrict__ n_big,
                                                        int64_t *__restrict__ cnt, const int64_t *__restrict__ p0f,
                                                        uint32_t *__restrict__ plistf,
                                                        uint32_t *__restrict__ pnodef)
{
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t nb = (int64_t)*n_big;
  for (int64_t i = gw; i < nb; i += nw) {
    const int64_t o2 = big[i];
    const int64_t m = oct_parent[o2];
    const int64_t a = p0c[m], b = p0c[m + 1];
    /* lane c < 8 keeps the running count of child c */
    int64_t run = 0, base = 0;
    if (FILL && lane < 8) base = p0f[8 * o2 + lane];
    for (int64_t q0 = a; q0 < b; q0 += 64) {
      const int64_t q = q0 + lane;
      const int ch = q < b ? (int)child[q] : -1;
      uint64_t bal[8];
#pragma unroll
      for (int k = 0; k < 8; k++) bal[k] = __ballot(ch == k);
      /* every lane takes part in the shuffle: the first free slot of its child */
      const int64_t slot = __shfl(base + run, ch >= 0 ? ch : 0, 64);
      if (FILL && ch >= 0) {
        uint64_t mine = bal[0];
#pragma unroll
        for (int k = 1; k < 8; k++) mine = ch == k ? bal[k] : mine;
        const int64_t pos = slot + (int64_t)__popcll(mine & lt);
        plistf[pos] = plistc[q];
        pnodef[pos] = (uint32_t)(8 * o2 + ch);
      }
      uint64_t my = bal[0];
#pragma unroll
      for (int k = 1; k < 8; k++) my = lane == k ? bal[k] : my;
      if (lane < 8) run += (int64_t)__popcll(my);
    }
    if (!FILL && lane < 8) cnt[8 * o2 + lane] = run;
  }
}

/* neighbour table of a standalone level: binary search of the periodic
 * neighbour key in the sorted octet keys */
__global__ void k_nbr_search(const uint64_t *__restrict__ key, int64_t n_oct, int64_t half,
                             int32_t *__restrict__ nbr)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 27 * n_oct) return;
  const int64_t o = t / 27;
  const int e = (int)(t % 27);
  const int ex = e % 3 - 1, ey = (e / 3) % 3 - 1, ez = e / 9 - 1;
  const uint64_t k = key[o];
  const uint64_t want = ahf_pack3((uint64_t)ahf_wrap(ahf_kx(k) + ex, half), (uint64_t)ahf_wrap(ahf_ky(k) + ey, half),
                                  (uint64_t)ahf_wrap(ahf_kz(k) + ez, half));
  int64_t lo = 0, hi = n_oct - 1, found = -1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const uint64_t km = key[mid];
    if (km == want) { found = mid; break; }
    if (km < want) lo = mid + 1; else hi = mid - 1;
  }
  nbr[t] = (int32_t)found;
}

/* ============================================================ host side */

static inline unsigned nblk(int64_t n)
{
  return (unsigned)((n + AHF_BLOCK - 1) / AHF_BLOCK);
}

#define LAUNCH(kern, n, ...)                                                                 \
  do {                                                                                       \
    if ((n) > 0) hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(AHF_BLOCK), 0, c->st, __VA_ARGS__); \
  } while (0)

static level_view make_view(const DevLevel &l)
{
  level_view v;
  v.L = l.L;
  v.n_oct = l.n_oct;
  v.oct_key = l.oct_key;
  v.oct_link = l.oct_link;
  v.nbr = l.nbr;
  v.n_row = l.n_row;
  v.row_key = l.row_key;
  v.row_oct0 = l.row_oct0;
  v.row_plane = l.row_plane;
  v.n_plane = l.n_plane;
  v.plane_z = l.plane_z;
  v.plane_row0 = l.plane_row0;
  return v;
}

/* bucket entries per node from which a level counts as dense */
#define AHF_DENSE_ENTRIES_PER_NODE 3
/* heavy-node workgroups: 7 wavefronts per SIMD on 256 CUs, each wavefront
 * folding one node at a time */
#define AHF_HEAVY_BLOCKS 1792
/* nodes with more contributions than this go to the heavy kernels */
#define AHF_KMAX_BIG 384u   /* levels of 2^21 nodes or more */
#define AHF_KMAX_SMALL 96u

/*
 * Exact-order gather of one level over its n_ent bucket entries: the octets
 * sorted into Morton order (the target order), then either
 *  - dense levels (many entries per node): a position record per entry and
 *    the per-source walk (k_gather_src), whose records stay in L2; or
 *  - sparse levels: a TSC weight record per entry and the flat walk
 *    (k_gather_light), which keeps a wavefront's lanes busy when buckets
 *    hold a particle or two;
 * and one wavefront per node with more than kmax contributions.  Every
 * choice gives the same bits.  Needs pmask.
 */
static int gather_level(ahfgpu_ctx *c, DevLevel &l, int64_t n_ent, int descending, bool timed)
{
  const int64_t nn = 8 * l.n_oct, no = l.n_oct;
  level_view v = make_view(l);
  Arena &A = c->arena;
  /* a level that holds every octet takes the staged plane walk */
  const bool slab = c->gather_mode == AHFGPU_GATHER_AUTO && l.L % AHF_SLAB == 0 &&
                    no == (l.L / 2) * (l.L / 2) * (l.L / 2);
  const bool dense = slab || c->gather_mode == AHFGPU_GATHER_SOURCE ||
                     ((c->gather_mode == AHFGPU_GATHER_AUTO || c->gather_mode == AHFGPU_GATHER_WAVE) &&
                      n_ent >= AHF_DENSE_ENTRIES_PER_NODE * nn);
  l.dens = A.get<float>(nn);
  l.qp = nullptr;
  l.W = nullptr;
  if (dense)
    l.qp = A.get<ahf_q4>(n_ent > 0 ? n_ent : 1);
  else
    l.W = A.get<ahf_w9>(n_ent > 0 ? n_ent : 1);
  uint32_t *heavy = A.get<uint32_t>(nn);
  if (!l.dens || !(dense ? (void *)l.qp : (void *)l.W) || !heavy) return AHFGPU_ENOMEM;
  uint32_t *n_heavy = (uint32_t *)(c->d_anomaly + 1);
  hipMemsetAsync(n_heavy, 0, sizeof(uint32_t), c->st);
  hipEvent_t e0 = timed ? c->timer_begin() : nullptr;
  if (dense)
    LAUNCH(k_prep_q, n_ent, l.node_p0 + nn, n_ent, l.plist, c->px, c->py, c->pz, l.qp);
  else
    LAUNCH(k_prep, n_ent, v, l.node_p0 + nn, n_ent, l.pnode, l.plist, c->px, c->py, c->pz, l.m2pd, l.W);
  if (!l.tord) {
    /* no order inherited from a coarser level: sort the octets' Morton keys */
    uint64_t *mk0 = A.get<uint64_t>(no), *mk1 = A.get<uint64_t>(no);
    uint32_t *ix0 = A.get<uint32_t>(no), *ix1 = A.get<uint32_t>(no);
    uint32_t *rtmp = A.get<uint32_t>(radix_tmp_u32<uint64_t>(no));
    if (!mk0 || !mk1 || !ix0 || !ix1 || !rtmp) return AHFGPU_ENOMEM;
    int axis_bits = 1;
    while ((1ll << axis_bits) < l.L / 2) axis_bits++;
    LAUNCH(k_morton_keys, no, v, mk0, ix0);
    const int w = radix_sort<uint64_t, uint32_t>(mk0, ix0, mk1, ix1, no, 3 * axis_bits, rtmp, c->st);
    l.tord = w ? ix1 : ix0;
  }
  const uint32_t *tord = l.tord;
  if (timed)
    c->timer_end(e0, TIMER_PREP, (12.0 + 8.0 + (dense ? 16.0 : 80.0)) * (double)n_ent + 24.0 * (double)no);
  hipEvent_t e1 = timed ? c->timer_begin() : nullptr;
  const int64_t per_xcd = (nblk(nn) + AHF_XCDS - 1) / AHF_XCDS;
  const dim3 grid((unsigned)(AHF_XCDS * per_xcd));
  /* long lanes in a small launch leave the chip idle: hand them to the
   * wavefront-per-node kernel sooner */
  const uint32_t kmax_big = c->kmax_big > 0 ? (uint32_t)c->kmax_big : AHF_KMAX_BIG;
  const uint32_t kmax_small = c->kmax_small > 0 ? (uint32_t)c->kmax_small : AHF_KMAX_SMALL;
  const uint32_t kmax = c->gather_mode == AHFGPU_GATHER_WAVE ? 0u : nn >= ((int64_t)1 << 21) ? kmax_big : kmax_small;
  static_assert(AHF_KMAX_BIG < 4096 && AHF_KMAX_SMALL < 4096, "the march's queue words hold counts below 4096");
  if (slab) {
    /* the planes 1..L-2 march, the two face planes take the per-plane walk */
    const int64_t side = l.L / AHF_SLAB;
    const int64_t zn = c->march_zn > 0 ? c->march_zn : l.L <= 128 ? 8 : 16, nch = (l.L - 2 + zn - 1) / zn;
    const int64_t per_m = (side * side * nch + AHF_XCDS - 1) / AHF_XCDS;
    const int64_t per_f = (side * side * 2 + AHF_XCDS - 1) / AHF_XCDS;
    const dim3 mgrid((unsigned)(AHF_XCDS * per_m)), fgrid((unsigned)(AHF_XCDS * per_f));
    if (n_ent >= 2 * nn) {
      hipLaunchKernelGGL((k_gather_march<AHF_SLAB_CAP_BIG, false>), mgrid, dim3(AHF_BLOCK), 0, c->st, v, per_m, zn,
                         l.pmask, l.node_p0, l.qp, descending, l.m2pd, kmax, l.dens, heavy, n_heavy);
      hipLaunchKernelGGL((k_gather_slab<AHF_SLAB_CAP_BIG, false>), fgrid, dim3(AHF_BLOCK), 0, c->st, v, per_f, l.pmask,
                         l.node_p0, l.qp, descending, l.m2pd, kmax, l.dens, heavy, n_heavy, 1);
    } else {
      hipLaunchKernelGGL((k_gather_march<AHF_MARCH_CAP_SMALL, true>), mgrid, dim3(AHF_BLOCK), 0, c->st, v, per_m, zn,
                         l.pmask, l.node_p0, l.qp, descending, l.m2pd, kmax, l.dens, heavy, n_heavy);
      hipLaunchKernelGGL((k_gather_slab<AHF_SLAB_CAP_SMALL, true>), fgrid, dim3(AHF_BLOCK), 0, c->st, v, per_f, l.pmask,
                         l.node_p0, l.qp, descending, l.m2pd, kmax, l.dens, heavy, n_heavy, 1);
    }
  } else if (dense)
    hipLaunchKernelGGL(k_gather_src, grid, dim3(AHF_BLOCK), 0, c->st, v, tord, per_xcd, l.pmask, l.node_p0, l.qp,
                       descending, l.m2pd, kmax, l.dens, heavy, n_heavy);
  else if ((uint64_t)(n_ent > 0 ? n_ent : 1) * sizeof(ahf_w9) < ((uint64_t)1 << 32) &&
           c->gather_mode != AHFGPU_GATHER_FLAT_WIDE)
    hipLaunchKernelGGL(k_gather_light<true>, grid, dim3(AHF_BLOCK), 0, c->st, v, tord, per_xcd, l.pmask, l.node_p0,
                       l.W, descending, kmax, l.dens, heavy, n_heavy);
  else
    hipLaunchKernelGGL(k_gather_light<false>, grid, dim3(AHF_BLOCK), 0, c->st, v, tord, per_xcd, l.pmask, l.node_p0,
                       l.W, descending, kmax, l.dens, heavy, n_heavy);
  if (timed) c->timer_end(e1, TIMER_GATHER, 12.0 * (double)n_ent + 13.0 * (double)nn);
  hipEvent_t e2 = timed ? c->timer_begin() : nullptr;
  if (dense)
    hipLaunchKernelGGL(k_gather_heavy_q, dim3(AHF_HEAVY_BLOCKS), dim3(AHF_BLOCK), 0, c->st, v, l.pmask, l.node_p0, l.qp,
                       descending, l.m2pd, l.dens, heavy, n_heavy);
  else
    hipLaunchKernelGGL(k_gather_heavy, dim3(AHF_HEAVY_BLOCKS), dim3(AHF_BLOCK), 0, c->st, v, l.pmask, l.node_p0, l.W,
                       descending, l.dens, heavy, n_heavy);
  if (timed) c->timer_end(e2, TIMER_HEAVY, 0.0);
  return 0;
}

/* TSC deposit of one level over its n_ent bucket entries in AHF's exact
 * order (even levels walk their lists from the tail of each bucket, odd
 * levels from the head) */
static int deposit_level(ahfgpu_ctx *c, DevLevel &l, int level_index, int64_t n_ent)
{
  const int64_t nn = 8 * l.n_oct;
  if (!l.pmask) {
    l.pmask = c->arena.get<uint32_t>(nn);
    if (!l.pmask) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (pmask, level %d)", level_index);
    level_view v = make_view(l);
    LAUNCH(k_pmask, l.n_oct, v, l.pmask, (uint8_t *)nullptr);
  }
  if (gather_level(c, l, n_ent, (level_index % 2) == 0, true))
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (deposit, level %d)", level_index);
  return 0;
}

static int build_domain(ahfgpu_ctx *c, DevLevel &d)
{
  const int64_t L = c->params.L_dom, half = L / 2, no = half * half * half, N = c->N;
  d.L = L;
  d.n_oct = no;
  d.n_row = half * half;
  d.n_plane = half;
  Arena &A = c->arena;
  d.oct_key = A.get<uint64_t>(no);
  d.oct_link = A.get<uint8_t>(no);
  d.nbr = A.get<int32_t>(27 * no);
  d.row_key = A.get<uint64_t>(d.n_row);
  d.row_oct0 = A.get<int32_t>(d.n_row + 1);
  d.row_plane = A.get<int32_t>(d.n_row);
  d.plane_z = A.get<int32_t>(d.n_plane);
  d.plane_row0 = A.get<int32_t>(d.n_plane + 1);
  d.node_p0 = A.get<int64_t>(8 * no + 1);
  uint32_t *k0 = A.get<uint32_t>(N), *v0 = A.get<uint32_t>(N), *k1 = A.get<uint32_t>(N), *v1 = A.get<uint32_t>(N);
  uint32_t *rtmp = A.get<uint32_t>(radix_tmp_u32<uint32_t>(N));
  if (!d.oct_key || !d.oct_link || !d.nbr || !d.row_key || !d.row_oct0 || !d.row_plane || !d.plane_z ||
      !d.plane_row0 || !d.node_p0 || !k0 || !v0 || !k1 || !v1 || !rtmp)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (domain)");
  LAUNCH(k_domain_octets, no, half, d.oct_key, d.oct_link, d.nbr);
  if ((half & (half - 1)) == 0) {
    d.tord = A.get<uint32_t>(no);
    if (!d.tord) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (domain)");
    LAUNCH(k_morton_full, no, half, d.tord);
  }
  LAUNCH(k_domain_rows, half * half + 1, half, d.row_key, d.row_oct0, d.row_plane, d.plane_z, d.plane_row0);
  hipEvent_t e0 = c->timer_begin();
  LAUNCH(k_bin_domain, N, c->px, c->py, c->pz, N, L, k0, v0);
  int bits = 1;
  while ((1ll << bits) < 8 * no) bits++;
  const int where = radix_sort<uint32_t, uint32_t>(k0, v0, k1, v1, N, bits, rtmp, c->st);
  d.pnode = where ? k1 : k0;
  d.plist = where ? v1 : v0;
  LAUNCH(k_csr_lower_bound, 8 * no + 1, d.pnode, N, 8 * no, d.node_p0);
  c->timer_end(e0, TIMER_BIN, 28.0 * (double)N);
  d.n_part = N;
  const double dl = (double)L;
  d.m2d = (dl * dl * dl) / (double)N;
  d.m2pd = (dl * dl * dl) / (double)N;
  d.crit = c->params.Nth_dom * d.m2d;
  return deposit_level(c, d, 0, N);
}

/* refine level `ci` into a new level; returns 1 if a level was added,
 * 0 when nothing refines or the new level is rolled back, <0 on error */
static int refine_level(ahfgpu_ctx *c, int ci)
{
  DevLevel &l = c->lev[ci];
  Arena &A = c->arena;
  const int64_t nn = 8 * l.n_oct, nrank = 4 * l.n_row, nplr = 2 * l.n_plane;
  level_view v = make_view(l);
  l.flags = A.get<uint8_t>(nn);
  int64_t *cnt = A.get<int64_t>(nrank + 1);
  int64_t *oct_off = A.get<int64_t>(nrank + 1);
  int32_t *rhas = A.get<int32_t>(nrank + 1);
  int32_t *row_off = A.get<int32_t>(nrank + 1);
  int32_t *phas = A.get<int32_t>(nplr + 1);
  int32_t *plane_off = A.get<int32_t>(nplr + 1);
  int64_t *stmp64 = A.get<int64_t>(scan_tmp_elems(nrank));
  int32_t *stmp32 = A.get<int32_t>(scan_tmp_elems(nrank > nplr ? nrank : nplr));
  if (!l.flags || !cnt || !oct_off || !rhas || !row_off || !phas || !plane_off || !stmp64 || !stmp32)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (flags, level %d)", ci);
  hipEvent_t e0 = c->timer_begin();
  LAUNCH(k_flags, l.n_oct, v, l.pmask, l.dens, l.crit - 1.0, l.flags);
  hipMemsetAsync(c->d_anomaly, 0, sizeof(int32_t), c->st);
  LAUNCH(k_refine_rows, nrank, v, l.flags, cnt, c->d_anomaly);
  exclusive_scan<int64_t, int64_t>(cnt, oct_off, nrank, stmp64, c->st);
  LAUNCH(k_nonzero, nrank, cnt, nrank, rhas);
  exclusive_scan<int32_t, int32_t>(rhas, row_off, nrank, stmp32, c->st);
  LAUNCH(k_plane_has, nplr, v, cnt, phas);
  exclusive_scan<int32_t, int32_t>(phas, plane_off, nplr, stmp32, c->st);
  c->timer_end(e0, TIMER_REFINE, 4.0 * (double)nn);
  /* sizes of the new level, and the particle count of this level, in one
   * read-back */
  struct {
    int64_t n_oct, n_part, n_row, n_plane, anomaly;
  } sz;
  hipLaunchKernelGGL(k_sizes, dim3(1), dim3(1), 0, c->st, oct_off + nrank, l.node_p0 + nn, row_off + nrank,
                     plane_off + nplr, c->d_anomaly, c->d_sizes);
  hipMemcpyAsync(c->h_sizes, c->d_sizes, 5 * sizeof(int64_t), hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "refine sizes: %s", hipGetErrorString(hipGetLastError()));
  memcpy(&sz, c->h_sizes, sizeof(sz));
  l.n_part = sz.n_part;
  if (sz.anomaly)
    return c->fail(AHFGPU_EUNSUP, "level %d: AHF would create an empty x-run here (undefined behaviour)", ci);
  if (sz.n_oct == 0) return 0;
  if (ci + 1 >= AHFGPU_MAX_LEVELS) return c->fail(AHFGPU_EINVAL, "too many levels");
  if (8 * sz.n_oct >= (int64_t)UINT32_MAX) return c->fail(AHFGPU_EINVAL, "level %d: too many nodes", ci + 1);
  DevLevel f;
  f.L = 2 * l.L;
  f.n_oct = sz.n_oct;
  f.n_row = sz.n_row;
  f.n_plane = sz.n_plane;
  const int64_t n2 = f.n_oct, nn2 = 8 * n2, nent = l.n_part;
  f.oct_key = A.get<uint64_t>(n2);
  f.oct_link = A.get<uint8_t>(n2);
  f.oct_parent = A.get<int64_t>(n2);
  f.nbr = A.get<int32_t>(27 * n2);
  f.row_key = A.get<uint64_t>(f.n_row);
  f.row_oct0 = A.get<int32_t>(f.n_row + 1);
  f.row_plane = A.get<int32_t>(f.n_row);
  f.plane_z = A.get<int32_t>(f.n_plane);
  f.plane_row0 = A.get<int32_t>(f.n_plane + 1);
  f.pmask = A.get<uint32_t>(nn2);
  f.node_p0 = A.get<int64_t>(nn2 + 1);
  l.node2oct = A.get<int32_t>(nn);
  l.child = A.get<int8_t>(nent);
  int64_t *fcnt = A.get<int64_t>(nn2);
  uint32_t *fbig = A.get<uint32_t>(n2);
  int64_t *ftmp = A.get<int64_t>(scan_tmp_elems(nn2));
  if (!f.oct_key || !f.oct_link || !f.oct_parent || !f.nbr || !f.row_key || !f.row_oct0 || !f.row_plane ||
      !f.plane_z || !f.plane_row0 || !f.pmask || !f.node_p0 || !l.node2oct || !l.child || !fcnt || !fbig || !ftmp)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (level %d)", ci + 1);
  hipEvent_t e1 = c->timer_begin();
  hipMemsetAsync(l.node2oct, 0xFF, (size_t)nn * sizeof(int32_t), c->st);
  LAUNCH(k_compact, nrank, v, l.flags, cnt, oct_off, row_off, plane_off, f.oct_key, f.oct_link, f.oct_parent,
         f.row_key, f.row_oct0, f.row_plane, l.node2oct);
  LAUNCH(k_plane_fill, nplr, v, phas, plane_off, row_off, f.plane_z, f.plane_row0);
  if (l.tord) {
    int32_t *mflag = A.get<int32_t>(nn), *mpos = A.get<int32_t>(nn + 1);
    int32_t *mtmp = A.get<int32_t>(scan_tmp_elems(nn));
    f.tord = A.get<uint32_t>(n2);
    if (!mflag || !mpos || !mtmp || !f.tord) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (level %d)", ci + 1);
    LAUNCH(k_morton_next_flags, nn, l.tord, l.node2oct, nn, mflag);
    exclusive_scan<int32_t, int32_t>(mflag, mpos, nn, mtmp, c->st);
    LAUNCH(k_morton_next_emit, nn, l.tord, l.node2oct, nn, mpos, f.tord);
  }
  hipLaunchKernelGGL(k_set2, dim3(1), dim3(1), 0, c->st, f.row_oct0, f.n_row, (int32_t)f.n_oct, f.plane_row0,
                     f.n_plane, (int32_t)f.n_row);
  LAUNCH(k_nbr_next, 27 * n2, l.nbr, f.oct_parent, l.node2oct, n2, f.nbr);
  level_view fv = make_view(f);
  uint8_t *imask2 = A.get<uint8_t>(n2);
  if (!imask2) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (level %d)", ci + 1);
  LAUNCH(k_pmask, fv.n_oct, fv, f.pmask, imask2);
  c->timer_end(e1, TIMER_BUILD, 150.0 * (double)n2);
  /* relink: child of every entry, then the fine buckets */
  hipEvent_t e2 = c->timer_begin();
  LAUNCH(k_relink_child, nent, nent, l.pnode, l.plist, l.node2oct, l.oct_key, imask2, l.qp, c->px, c->py, c->pz,
         (double)f.L, l.child);
  uint32_t *n_big = (uint32_t *)(c->d_anomaly + 2);
  hipMemsetAsync(n_big, 0, sizeof(uint32_t), c->st);
  LAUNCH(k_fine_count, n2, f.oct_parent, n2, l.node_p0, l.child, fcnt, fbig, n_big);
  hipLaunchKernelGGL(k_fine_big<false>, dim3(512), dim3(AHF_BLOCK), 0, c->st, f.oct_parent, l.node_p0, l.child,
                     l.plist, fbig, n_big, fcnt, nullptr, nullptr, nullptr);
  exclusive_scan<int64_t, int64_t>(fcnt, f.node_p0, nn2, ftmp, c->st);
  /* the fine level keeps at most the coarse level's entries */
  f.plist = A.get<uint32_t>(nent);
  f.pnode = A.get<uint32_t>(nent);
  if (!f.plist || !f.pnode) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (level %d)", ci + 1);
  LAUNCH(k_fine_fill, nent, nent, l.pnode, l.node2oct, l.node_p0, l.child, l.plist, f.node_p0, f.plist, f.pnode);
  hipLaunchKernelGGL(k_fine_big<true>, dim3(512), dim3(AHF_BLOCK), 0, c->st, f.oct_parent, l.node_p0, l.child,
                     l.plist, fbig, n_big, nullptr, f.node_p0, f.plist, f.pnode);
  c->timer_end(e2, TIMER_RELINK, 29.0 * (double)nent + 24.0 * (double)nn2);
  /* the particles the fine level keeps: its deposit kernel depends on how
   * many entries each node holds, which is read back only where the coarse
   * level's count (an upper bound) leaves the choice open; the bound then
   * sizes the deposit's records and launches, and the count is read with
   * the next level's sizes */
  const int gm = c->gather_mode;
  const bool full_f = gm == AHFGPU_GATHER_AUTO && f.L % AHF_SLAB == 0 && f.n_oct == (f.L / 2) * (f.L / 2) * (f.L / 2);
  const bool decided = gm == AHFGPU_GATHER_FLAT || gm == AHFGPU_GATHER_FLAT_WIDE || gm == AHFGPU_GATHER_SOURCE || full_f ||
                       nent < AHF_DENSE_ENTRIES_PER_NODE * nn2;
  int64_t nent_f = nent;
  f.n_part = -1;
  if (!decided) {
    hipMemcpyAsync(c->h_sizes + 7, f.node_p0 + nn2, sizeof(int64_t), hipMemcpyDeviceToHost, c->st);
    if (hipStreamSynchronize(c->st) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "level %d relink: %s", ci + 1, hipGetErrorString(hipGetLastError()));
    nent_f = c->h_sizes[7];
    f.n_part = nent_f;
  }
  f.m2d = l.m2d * AHF_CRITMULTI_D;
  f.m2pd = l.m2pd * AHF_CRITMULTI_D;
  f.crit = c->params.Nth_ref * f.m2d;
  int rc = deposit_level(c, f, ci + 1, nent_f);
  if (rc) return rc;
  if (nn2 < AHF_MIN_NNODES) {
    /* relink_back (src/libamr_serial/generate_grids.c:231-238): the fine
     * lists go back to their mothers; keep them for the final lists */
    c->rb = f;
    c->rolled_back = 1;
    return 0;
  }
  c->lev.push_back(f);
  return 1;
}

/* ============================================================== C ABI */

extern "C" int ahfgpu_amr_build(ahfgpu_ctx *c, const ahfgpu_amr_params *p, ahfgpu_amr_info *info)
{
  if (!c || !p) return AHFGPU_EINVAL;
  if (c->N <= 0) return c->fail(AHFGPU_EINVAL, "no particles resident (call ahfgpu_set_particles first)");
  /* octets need an even domain grid; refinement doubles it up to L_max */
  bool lmax_ok = p->L_dom >= 4 && (p->L_dom & 1) == 0 && p->L_max >= p->L_dom && p->L_max % p->L_dom == 0;
  if (lmax_ok) {
    const int64_t r = p->L_max / p->L_dom;
    lmax_ok = (r & (r - 1)) == 0;
  }
  if (!lmax_ok)
    return c->fail(AHFGPU_EINVAL, "need an even L_dom >= 4 and L_max = L_dom * 2^k (got %lld, %lld)",
                   (long long)p->L_dom, (long long)p->L_max);
  if (p->L_max > (1 << 22)) return c->fail(AHFGPU_EINVAL, "L_max above 2^22 is not supported");
  if (c->N >= (int64_t)UINT32_MAX) return c->fail(AHFGPU_EINVAL, "at most 2^32-1 particles");
  if ((int64_t)p->L_dom * p->L_dom * p->L_dom >= ((int64_t)1 << 31))
    return c->fail(AHFGPU_EINVAL, "domain grid too large");
  c->params = *p;
  c->arena.reset();
  c->lev.clear();
  c->rolled_back = 0;
  c->rb = DevLevel();
  c->timers_reset(TIMER_BIN, TIMER_RELINK);
  c->timed.clear();
  c->ev_next = 0;
  hipEventRecord(c->ev_a0, c->st);
  c->lev.push_back(DevLevel());
  int rc = build_domain(c, c->lev[0]);
  if (rc) return rc;
  for (;;) {
    const int ci = (int)c->lev.size() - 1;
    if (c->lev[ci].L == p->L_max) break;
    rc = refine_level(c, ci);
    if (rc < 0) return rc;
    if (rc == 0) break;
  }
  hipEventRecord(c->ev_a1, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "amr build: %s", hipGetErrorString(hipGetLastError()));
  for (auto *lv : {&c->lev.back(), &c->rb}) {
    if (lv->n_part < 0 && lv->node_p0) {
      hipMemcpy(&lv->n_part, lv->node_p0 + 8 * lv->n_oct, sizeof(int64_t), hipMemcpyDeviceToHost);
    }
  }
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev_a0, c->ev_a1);
  c->amr_ms = ms;
  c->timers_collect();
  c->built = 1;
  if (info) {
    memset(info, 0, sizeof(*info));
    info->n_levels = (int32_t)c->lev.size();
    info->rolled_back = c->rolled_back;
    for (size_t i = 0; i < c->lev.size() && i < AHFGPU_MAX_LEVELS; i++) {
      info->n_nodes[i] = 8 * c->lev[i].n_oct;
      info->n_part[i] = c->lev[i].n_part;
      info->deposits += c->lev[i].n_part;
    }
  }
  return 0;
}

/* ---------------------------------------------------- read back levels */

template <typename T>
static std::vector<T> d2h(const T *d, int64_t n, hipStream_t st)
{
  std::vector<T> h((size_t)(n > 0 ? n : 0));
  if (n > 0) hipMemcpyAsync(h.data(), d, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, st);
  return h;
}

/* final per-node lists of level `lev` in AHF's list order, on the host */
static int final_lists(ahfgpu_ctx *c, int lev, std::vector<int64_t> &list0, std::vector<uint32_t> &list)
{
  const DevLevel &l = c->lev[lev];
  const int64_t nn = 8 * l.n_oct;
  const bool last = lev == (int)c->lev.size() - 1;
  std::vector<int64_t> p0 = d2h(l.node_p0, nn + 1, c->st);
  std::vector<uint32_t> pl = d2h(l.plist, l.n_part, c->st);
  std::vector<int8_t> ch;
  std::vector<int32_t> n2o;
  std::vector<int64_t> fp0;
  std::vector<uint32_t> fpl;
  const bool moved_info = l.child != nullptr;
  const bool rb = last && c->rolled_back;
  if (moved_info) ch = d2h(l.child, l.n_part, c->st);
  if (rb) {
    n2o = d2h(l.node2oct, nn, c->st);
    fp0 = d2h(c->rb.node_p0, 8 * c->rb.n_oct + 1, c->st);
    fpl = d2h(c->rb.plist, c->rb.n_part, c->st);
  }
  if (hipStreamSynchronize(c->st) != hipSuccess) return c->fail(AHFGPU_EHIP, "read back lists");
  list0.assign((size_t)nn + 1, 0);
  list.resize((size_t)l.n_part);
  const bool desc = (lev % 2) == 0, fdesc = ((lev + 1) % 2) == 0;
  int64_t w = 0;
  for (int64_t n = 0; n < nn; n++) {
    list0[(size_t)n] = w;
    const int64_t a = p0[(size_t)n], b = p0[(size_t)n + 1];
    for (int64_t k = 0; k < b - a; k++) {
      const int64_t q = desc ? b - 1 - k : a + k;
      if (!moved_info || ch[(size_t)q] < 0) list[(size_t)w++] = pl[(size_t)q];
    }
    if (rb && n2o[(size_t)n] >= 0) {
      for (int k = 0; k < 8; k++) {
        const int64_t fn = 8 * (int64_t)n2o[(size_t)n] + k;
        const int64_t f0 = fp0[(size_t)fn], f1 = fp0[(size_t)fn + 1];
        for (int64_t j = 0; j < f1 - f0; j++) list[(size_t)w++] = fpl[(size_t)(fdesc ? f1 - 1 - j : f0 + j)];
      }
    }
  }
  list0[(size_t)nn] = w;
  return 0;
}

extern "C" int ahfgpu_amr_level_sizes(ahfgpu_ctx *c, int lev, int64_t *L, int64_t *n_oct, int64_t *n_listed,
                                      double *critdens, double *m2pd)
{
  if (!c || !c->built || lev < 0 || lev >= (int)c->lev.size()) return AHFGPU_EINVAL;
  const DevLevel &l = c->lev[lev];
  if (L) *L = l.L;
  if (n_oct) *n_oct = l.n_oct;
  if (n_listed) {
    /* particles that stay on this level (plus any rolled back into it) */
    int64_t n = l.n_part;
    if (l.child && !(lev == (int)c->lev.size() - 1 && c->rolled_back)) {
      int64_t moved = lev + 1 < (int)c->lev.size() ? c->lev[lev + 1].n_part : 0;
      n -= moved;
    }
    *n_listed = n;
  }
  if (critdens) *critdens = l.crit;
  if (m2pd) *m2pd = l.m2pd;
  return 0;
}

extern "C" int ahfgpu_amr_level_copy(ahfgpu_ctx *c, int lev, uint64_t *oct_key, uint8_t *oct_link, float *dens,
                                     int64_t *list0, uint32_t *list)
{
  if (!c || !c->built || lev < 0 || lev >= (int)c->lev.size()) return AHFGPU_EINVAL;
  const DevLevel &l = c->lev[lev];
  if (oct_key) hipMemcpyAsync(oct_key, l.oct_key, (size_t)l.n_oct * 8, hipMemcpyDeviceToHost, c->st);
  if (oct_link) hipMemcpyAsync(oct_link, l.oct_link, (size_t)l.n_oct, hipMemcpyDeviceToHost, c->st);
  if (dens) hipMemcpyAsync(dens, l.dens, (size_t)l.n_oct * 8 * 4, hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess) return c->fail(AHFGPU_EHIP, "level copy");
  if (list0 || list) {
    std::vector<int64_t> l0;
    std::vector<uint32_t> li;
    int rc = final_lists(c, lev, l0, li);
    if (rc) return rc;
    if (list0) memcpy(list0, l0.data(), l0.size() * sizeof(int64_t));
    if (list) memcpy(list, li.data(), (size_t)l0.back() * sizeof(uint32_t));
  }
  return 0;
}

extern "C" int ahfgpu_amr_dump(ahfgpu_ctx *c, const char *path)
{
  if (!c || !c->built || !path) return AHFGPU_EINVAL;
  const int nl = (int)c->lev.size();
  std::vector<host_level> hl((size_t)nl);
  std::vector<std::vector<uint64_t>> keys((size_t)nl);
  std::vector<std::vector<uint8_t>> links((size_t)nl);
  std::vector<std::vector<float>> dens((size_t)nl);
  std::vector<std::vector<int64_t>> l0((size_t)nl);
  std::vector<std::vector<uint32_t>> li((size_t)nl);
  for (int i = 0; i < nl; i++) {
    const DevLevel &l = c->lev[i];
    keys[(size_t)i] = d2h(l.oct_key, l.n_oct, c->st);
    links[(size_t)i] = d2h(l.oct_link, l.n_oct, c->st);
    dens[(size_t)i] = d2h(l.dens, 8 * l.n_oct, c->st);
    int rc = final_lists(c, i, l0[(size_t)i], li[(size_t)i]);
    if (rc) return rc;
    host_level &h = hl[(size_t)i];
    memset(&h, 0, sizeof(h));
    h.L = l.L;
    h.n_oct = l.n_oct;
    h.critdens = l.crit;
    h.masstopartdens = l.m2pd;
    h.size_no_nodes = (uint64_t)(8 * l.n_oct);
    h.size_no_part = (uint64_t)l.n_part;
    h.oct_key = keys[(size_t)i].data();
    h.oct_link = links[(size_t)i].data();
    h.dens = dens[(size_t)i].data();
    h.list0 = l0[(size_t)i].data();
    h.list = li[(size_t)i].data();
  }
  return hierarchy_dump(path, hl.data(), nl) ? c->fail(AHFGPU_EIO, "cannot write %s", path) : 0;
}

/* ------------------------------------------------ standalone deposit */

extern "C" int ahfgpu_deposit_level(ahfgpu_ctx *c, int64_t L, const uint64_t *oct_key, const uint8_t *oct_link,
                                    int64_t n_oct, const int64_t *list0, const uint32_t *list, double m2pd,
                                    float *dens_out)
{
  if (!c || !oct_key || !oct_link || !list0 || n_oct <= 0 || L < 4 || (L & 1)) return AHFGPU_EINVAL;
  if (c->N <= 0) return c->fail(AHFGPU_EINVAL, "no particles resident");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  const int64_t nn = 8 * n_oct, ne = list0[nn], half = L / 2;
  for (int64_t i = 1; i < n_oct; i++)
    if (oct_key[i] <= oct_key[i - 1]) return c->fail(AHFGPU_EINVAL, "octet keys must be strictly increasing");
  for (int64_t q = 0; q < ne; q++)
    if ((int64_t)list[q] >= c->N) return c->fail(AHFGPU_EINVAL, "list entry beyond the resident particles");
  /* rows and planes of the octets */
  std::vector<uint64_t> rkey;
  std::vector<int32_t> roct0, rplane, pz, prow0;
  for (int64_t i = 0; i < n_oct; i++) {
    const uint64_t k = oct_key[i];
    const bool new_row = i == 0 || ahf_ky(k) != ahf_ky(oct_key[i - 1]) || ahf_kz(k) != ahf_kz(oct_key[i - 1]);
    if (!new_row) continue;
    if (i == 0 || ahf_kz(k) != ahf_kz(oct_key[i - 1])) {
      pz.push_back((int32_t)ahf_kz(k));
      prow0.push_back((int32_t)rkey.size());
    }
    rkey.push_back((uint64_t)ahf_ky(k) | ((uint64_t)ahf_kz(k) << AHF_KEY_BITS));
    roct0.push_back((int32_t)i);
    rplane.push_back((int32_t)pz.size() - 1);
  }
  roct0.push_back((int32_t)n_oct);
  prow0.push_back((int32_t)rkey.size());
  std::vector<uint32_t> pnode((size_t)(ne > 0 ? ne : 1));
  for (int64_t n = 0; n < nn; n++)
    for (int64_t q = list0[n]; q < list0[n + 1]; q++) pnode[(size_t)q] = (uint32_t)n;
  c->arena.reset();
  c->lev.clear();
  c->built = 0;
  Arena &A = c->arena;
  DevLevel l;
  l.L = L;
  l.n_oct = n_oct;
  l.n_row = (int64_t)rkey.size();
  l.n_plane = (int64_t)pz.size();
  l.oct_key = A.get<uint64_t>(n_oct);
  l.oct_link = A.get<uint8_t>(n_oct);
  l.nbr = A.get<int32_t>(27 * n_oct);
  l.row_key = A.get<uint64_t>(l.n_row);
  l.row_oct0 = A.get<int32_t>(l.n_row + 1);
  l.row_plane = A.get<int32_t>(l.n_row);
  l.plane_z = A.get<int32_t>(l.n_plane);
  l.plane_row0 = A.get<int32_t>(l.n_plane + 1);
  l.node_p0 = A.get<int64_t>(nn + 1);
  l.plist = A.get<uint32_t>(ne);
  l.pnode = A.get<uint32_t>(ne);
  if (!l.oct_key || !l.oct_link || !l.nbr || !l.row_key || !l.row_oct0 || !l.row_plane || !l.plane_z ||
      !l.plane_row0 || !l.node_p0 || !l.plist || !l.pnode)
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (deposit)");
  hipMemcpyAsync(l.oct_key, oct_key, (size_t)n_oct * 8, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.oct_link, oct_link, (size_t)n_oct, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.row_key, rkey.data(), rkey.size() * 8, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.row_oct0, roct0.data(), roct0.size() * 4, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.row_plane, rplane.data(), rplane.size() * 4, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.plane_z, pz.data(), pz.size() * 4, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.plane_row0, prow0.data(), prow0.size() * 4, hipMemcpyHostToDevice, c->st);
  hipMemcpyAsync(l.node_p0, list0, (size_t)(nn + 1) * 8, hipMemcpyHostToDevice, c->st);
  if (ne > 0) {
    hipMemcpyAsync(l.plist, list, (size_t)ne * 4, hipMemcpyHostToDevice, c->st);
    hipMemcpyAsync(l.pnode, pnode.data(), (size_t)ne * 4, hipMemcpyHostToDevice, c->st);
  }
  LAUNCH(k_nbr_search, 27 * n_oct, l.oct_key, n_oct, half, l.nbr);
  l.m2pd = m2pd;
  /* buckets hold the lists in list order: walk them forwards */
  level_view v = make_view(l);
  l.pmask = A.get<uint32_t>(nn);
  if (!l.pmask) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (deposit)");
  LAUNCH(k_pmask, l.n_oct, v, l.pmask, (uint8_t *)nullptr);
  if (gather_level(c, l, ne, 0, false)) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (deposit)");
  if (dens_out) hipMemcpyAsync(dens_out, l.dens, (size_t)nn * 4, hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "deposit: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}
