/*
 * devprims.h -- wave64 device primitives used by the AMR and unbinding
 * pipelines: exclusive prefix sums and a stable LSD radix sort.
 *
 * Written for gfx950 (CDNA4): 64-lane wavefronts, ballots are 64-bit,
 * 256-thread workgroups (4 waves).
 */
#ifndef AHF_DEVPRIMS_H
#define AHF_DEVPRIMS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define AHF_BLOCK 256
#define AHF_WAVES (AHF_BLOCK / 64)

namespace ahf {

/* inclusive wave scan of v (wave64, shuffle-up) */
template <typename T>
__device__ inline T wave_incl_scan(T v)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

/* exclusive block scan of one value per thread; returns the block total in *total */
template <typename T>
__device__ inline T block_excl_scan(T v, T *total)
{
  __shared__ T wsum[AHF_WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < AHF_WAVES; k++) {
    T s = wsum[k];
    if (k < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

/* ---------------------------------------------------------- exclusive scan */

#define AHF_SCAN_ITEMS 8
#define AHF_SCAN_TILE (AHF_BLOCK * AHF_SCAN_ITEMS)

/* per tile: the exclusive scan and the tile's sum (into tile_sum, or into
 * *total when the whole input is one tile) */
template <typename Tin, typename Tout>
__global__ void k_scan_tiles(const Tin *__restrict__ in, Tout *__restrict__ out, Tout *__restrict__ tile_sum,
                             int64_t n, Tout *__restrict__ total)
{
  const int64_t base = (int64_t)blockIdx.x * AHF_SCAN_TILE + (int64_t)threadIdx.x * AHF_SCAN_ITEMS;
  Tout v[AHF_SCAN_ITEMS];
  Tout s = 0;
#pragma unroll
  for (int i = 0; i < AHF_SCAN_ITEMS; i++) {
    const int64_t k = base + i;
    v[i] = k < n ? (Tout)in[k] : (Tout)0;
    s += v[i];
  }
  Tout tot;
  Tout ex = block_excl_scan<Tout>(s, &tot);
#pragma unroll
  for (int i = 0; i < AHF_SCAN_ITEMS; i++) {
    const int64_t k = base + i;
    if (k < n) out[k] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) {
    if (total)
      *total = tot;
    else
      tile_sum[blockIdx.x] = tot;
  }
}

template <typename T>
__global__ void k_scan_add(T *__restrict__ out, const T *__restrict__ tile_off, int64_t n)
{
  const int64_t base = (int64_t)blockIdx.x * AHF_SCAN_TILE;
  const T add = tile_off[blockIdx.x];
  for (int i = threadIdx.x; i < AHF_SCAN_TILE; i += AHF_BLOCK) {
    const int64_t k = base + i;
    if (k < n) out[k] += add;
  }
}

/* one workgroup scans up to any length sequentially by tiles */
template <typename T>
__global__ void k_scan_single(T *__restrict__ a, int64_t n, T *__restrict__ total)
{
  __shared__ T carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += AHF_BLOCK) {
    const int64_t k = base + threadIdx.x;
    T v = k < n ? a[k] : (T)0;
    T tot;
    T ex = block_excl_scan<T>(v, &tot);
    const T c = carry_s;
    if (k < n) a[k] = c + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = c + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry_s;
}

/*
 * out[i] = sum_{j<i} in[i]; out[n] = total (out must hold n+1 entries).
 * tmp must hold at least scan_tmp_elems(n) elements of Tout.
 */
inline int64_t scan_tmp_elems(int64_t n)
{
  const int64_t tiles = (n + AHF_SCAN_TILE - 1) / AHF_SCAN_TILE;
  return tiles + 1;
}

template <typename Tin, typename Tout>
inline void exclusive_scan(const Tin *in, Tout *out, int64_t n, Tout *tmp, hipStream_t st)
{
  if (n <= 0) {
    hipMemsetAsync(out, 0, sizeof(Tout), st);
    return;
  }
  const int64_t tiles = (n + AHF_SCAN_TILE - 1) / AHF_SCAN_TILE;
  if (tiles == 1) {
    hipLaunchKernelGGL((k_scan_tiles<Tin, Tout>), dim3(1), dim3(AHF_BLOCK), 0, st, in, out, tmp, n, out + n);
    return;
  }
  hipLaunchKernelGGL((k_scan_tiles<Tin, Tout>), dim3((unsigned)tiles), dim3(AHF_BLOCK), 0, st, in, out, tmp, n,
                     (Tout *)nullptr);
  hipLaunchKernelGGL((k_scan_single<Tout>), dim3(1), dim3(AHF_BLOCK), 0, st, tmp, tiles, out + n);
  if (tiles > 1) hipLaunchKernelGGL((k_scan_add<Tout>), dim3((unsigned)tiles), dim3(AHF_BLOCK), 0, st, out, tmp, n);
}

/* ------------------------------------------------------ stable radix sort */

#define AHF_RS_ITEMS 16
#define AHF_RS_TILE (AHF_BLOCK * AHF_RS_ITEMS)

template <typename K>
__device__ inline uint32_t rs_digit(K k, int shift)
{
  return (uint32_t)((k >> shift) & (K)0xFF);
}

template <typename K>
__global__ void k_rs_hist(const K *__restrict__ keys, int64_t n, int shift, uint32_t *__restrict__ hist,
                          int64_t nblk)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * AHF_RS_TILE;
  for (int i = 0; i < AHF_RS_ITEMS; i++) {
    const int64_t e = base + (int64_t)i * AHF_BLOCK + threadIdx.x;
    if (e < n) atomicAdd(&h[rs_digit<K>(keys[e], shift)], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

/* stable scatter: element order inside a tile is (item, thread) = index order */
template <typename K, typename V>
__global__ void k_rs_scatter(const K *__restrict__ kin, const V *__restrict__ vin, K *__restrict__ kout,
                             V *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ off,
                             int64_t nblk)
{
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[AHF_WAVES][256];
  __shared__ uint32_t base_s[256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  run[t] = 0;
  base_s[t] = off[(int64_t)t * nblk + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int64_t tbase = (int64_t)blockIdx.x * AHF_RS_TILE;
  for (int i = 0; i < AHF_RS_ITEMS; i++) {
    for (int k = 0; k < AHF_WAVES; k++) wcnt[k][t] = 0;
    __syncthreads();
    const int64_t e = tbase + (int64_t)i * AHF_BLOCK + t;
    const bool valid = e < n;
    K key = valid ? kin[e] : (K)0;
    const uint32_t d = rs_digit<K>(key, shift);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot(valid && ((d >> b) & 1u));
      match &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t wrank = (uint32_t)__popcll(match & lt);
    if (valid && wrank == 0) wcnt[w][d] = (uint32_t)__popcll(match);
    __syncthreads();
    {
      uint32_t s = run[t];
      for (int k = 0; k < AHF_WAVES; k++) {
        const uint32_t c = wcnt[k][t];
        wcnt[k][t] = s;
        s += c;
      }
      run[t] = s;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = base_s[d] + wcnt[w][d] + wrank;
      kout[pos] = key;
      vout[pos] = vin[e];
    }
    __syncthreads();
  }
}

template <typename K>
inline int64_t radix_tmp_u32(int64_t n)
{
  const int64_t nblk = (n + AHF_RS_TILE - 1) / AHF_RS_TILE;
  return 2 * (256 * nblk + 1) + scan_tmp_elems(256 * nblk) + 8;
}

/*
 * Stable sort of (keys, vals) on the low `bits` key bits.  The result ends
 * in (k0, v0) or (k1, v1); returns 0 if it is in (k0, v0), 1 otherwise.
 * tmp: radix_tmp_u32<K>(n) uint32 elements.
 */
template <typename K, typename V>
inline int radix_sort(K *k0, V *v0, K *k1, V *v1, int64_t n, int bits, uint32_t *tmp, hipStream_t st)
{
  if (n <= 1) return 0;
  const int64_t nblk = (n + AHF_RS_TILE - 1) / AHF_RS_TILE;
  uint32_t *hist = tmp, *off = tmp + (256 * nblk + 1), *stmp = off + (256 * nblk + 1);
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    K *ki = cur ? k1 : k0;
    V *vi = cur ? v1 : v0;
    K *ko = cur ? k0 : k1;
    V *vo = cur ? v0 : v1;
    hipLaunchKernelGGL((k_rs_hist<K>), dim3((unsigned)nblk), dim3(AHF_BLOCK), 0, st, ki, n, shift, hist, nblk);
    exclusive_scan<uint32_t, uint32_t>(hist, off, 256 * nblk, stmp, st);
    hipLaunchKernelGGL((k_rs_scatter<K, V>), dim3((unsigned)nblk), dim3(AHF_BLOCK), 0, st, ki, vi, ko, vo, n,
                       shift, off, nblk);
    cur ^= 1;
  }
  return cur;
}

} // namespace ahf

#endif
