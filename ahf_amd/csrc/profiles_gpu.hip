/*
 * profiles_gpu.hip -- HaloProfiles (src/libahf/ahf_halos.c:3961-5060) of
 * every resident halo for AHF's default flags (unit masses, R_max and r2
 * from the sorted particle list, Hubble drag in the kinetic energy).
 *
 * One wavefront per halo walks the members in their radial order, 64 at a
 * time.  AHF's walk is one running sum after another (centre of mass,
 * inertia tensor, bulk velocity, angular momentum, kinetic energy, the
 * potential integral, potential energy), each rounded in member order.  The
 * wavefront computes every per-member term in parallel (a lane per member),
 * stages the terms in LDS, and then lets each running sum advance over the
 * 64 staged terms on a lane of its own (a lane per sum): the additions are
 * AHF's, in AHF's order, so the sums carry AHF's roundings, while the
 * divisions, square roots and products around them run 64 wide.  A sum
 * whose running value feeds later terms (the bulk velocity into the
 * angular momentum and kinetic energy, the potential into the energies)
 * leaves its running value in place of the terms.
 *
 * Radial bins: member j is taken into the current bin and closes every bin
 * whose radius it reaches (ahf_halos.c:4209-4620: the loop condition tests
 * the distance of the member taken last).  With bin radii that do not
 * decrease (the sorted lists always give these), the bins closed after
 * member j are a running maximum over the members, a wavefront scan; other
 * radii take the same automaton sequentially.  A bin's values are the
 * running sums at the member that closes it.
 *
 * Haloes of 2048 members or more take a workgroup of four wavefronts (256
 * members per chunk) on a second stream, with two chunks in flight: the
 * second wavefront runs the next chunk's bulk-velocity sums while the first
 * runs the current chunk's other sums.
 *
 * After the walk the lanes turn to the bins (one lane per bin, at most 64
 * bins below 2^32 members): profile values and the inertia eigen
 * decomposition per bin; then R_max and r2 from the smoothed per-member
 * curves (find_max, src/libutility/general.c:590-665) and the integral
 * properties.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "../../include/ahf_gpu.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "halo_set.h"

#define PRF_PI 3.14159265358979323846264338 /* src/param.h:39 */
#define PRF_GRAV 4.3006485e-9              /* src/param.h:45 */
#define PRF_ZERO 1e-6                       /* src/param.h:124 (no -DDOUBLE) */
#define PRF_MZERO 5e-16                     /* src/param.h:125 */
#define PRF_MINSHELL 10                     /* AHF_MINPART_SHELL, src/param.h:16 */
#define PRF_NIGNORE 5                       /* AHF_Rmax_r2_NIGNORE, src/param.h:25 */
#define AHF_PROF_BIG 2048                   /* default: haloes this large take a workgroup of */
#define AHF_PROF_BIG_NW 4                   /* ... this many wavefronts */

/* per-bin record columns: the running values at the member that closes a bin */
enum { BT_A11 = 0, BT_A22, BT_A33, BT_A12, BT_A13, BT_A23, BT_LX, BT_LY, BT_LZ, BT_T, BT_EP, BT_VESC, BT_N };

/* profile values per bin and integral properties, in the order of
 * include/ahf_gpu.h */
enum {
  PQ_R, PQ_NPART, PQ_NVPART, PQ_OVDENS, PQ_DENS, PQ_V2CIRC, PQ_VESC2, PQ_SIGV, PQ_EKIN, PQ_EPOT, PQ_LX, PQ_LY, PQ_LZ,
  PQ_E1X, PQ_E1Y, PQ_E1Z, PQ_E2X, PQ_E2Y, PQ_E2Z, PQ_E3X, PQ_E3Y, PQ_E3Z, PQ_AX1, PQ_AX2, PQ_AX3, PQ_N
};
enum {
  PP_NBINS, PP_MVIR, PP_VX, PP_VY, PP_VZ, PP_RMAX, PP_V2MAX, PP_R2, PP_SIGV, PP_VESC2, PP_EKIN, PP_EPOT,
  PP_AX1, PP_AX2, PP_AX3, PP_E1X, PP_E1Y, PP_E1Z, PP_E2X, PP_E2Y, PP_E2Z, PP_E3X, PP_E3Y, PP_E3Z,
  PP_LX, PP_LY, PP_LZ, PP_LAMBDA, PP_LAMBDAE, PP_CNFW, PP_R1, PP_CR1, PP_SURFP, PP_FMHIRES,
  PP_MBP_OFF, PP_MBP_X, PP_MBP_Y, PP_MBP_Z, PP_MBP_VX, PP_MBP_VY, PP_MBP_VZ,
  PP_COM_OFF, PP_COM_X, PP_COM_Y, PP_COM_Z, PP_N
};
static_assert(PQ_N == AHFGPU_PROF_NQ && PP_N == AHFGPU_PROF_NP, "profile value counts");

struct ProfConst {
  double r_fac, v_fac, m_fac, phi_fac, hubble;
  int64_t minpart;
};

/* LDS of one halo's workgroup of NW wavefronts: a chunk is 64*NW members */
/* member q's position and velocity from the interleaved particle copy (one
 * 32-byte sector instead of six scattered words) */
__device__ __forceinline__ void pv_member(const float4 *__restrict__ pv, uint32_t q, float &x, float &y, float &z,
                                          float &vx, float &vy, float &vz)
{
  const float4 a = pv[2 * (size_t)q], b = pv[2 * (size_t)q + 1];
  x = a.x;
  y = a.y;
  z = a.z;
  vx = a.w;
  vy = b.x;
  vz = b.y;
}

template <int NW>
struct ProfShared {
  static constexpr int C = 64 * NW;
  double ta[12][C + 1]; /* centre of mass x3, inertia x6, bulk velocity x3 (rows padded: bank spread) */
  double tc[5][C + 1];  /* angular momentum x3, kinetic energy, potential */
  double d_s[C];        /* each member's distance */
  double i_s[C];        /* I_now after each member; first the own M/d^2 */
  double rad[64];       /* bin radii */
  double *bt;           /* per-bin records, bin b at bt[b * PQ_N]: the halo's own profile rows, written
                           over at the end (out of LDS, for occupancy) */
  int64_t bclose[64];   /* member index that closed each bin */
  int32_t bprev[C], bcur[C];
  uint64_t cm[NW], fm[NW], hm[NW]; /* per-wavefront masks: closing members, I_now steps, hits */
  int32_t wv[NW];                  /* per-wavefront maxima / counts */
  double rv[NW];                   /* per-wavefront reduction values */
  int64_t ri[NW];
  double fin[16];
  int32_t flag;
  /* the pipelined walk of NW > 1: two chunks in flight */
  static constexpr int C2 = NW > 1 ? C : 1;
  double tg[NW > 1 ? 2 : 1][9][C2 + 1]; /* centre of mass x3, inertia x6 of each chunk */
  int32_t bp2[NW > 1 ? 2 : 1][C2], bc2[NW > 1 ? 2 : 1][C2];
  uint64_t cm2[2][NW];
};

/* ------------------------------------------------------- per-bin maths */

/* a(i,j) -= s*(a(k,l) + a(i,j)*tau), a(k,l) += s*(a(i,j) - a(k,l)*tau),
 * both from the values before the step */
__device__ inline void prf_rotate(double a[4][4], int i, int j, int k, int l, double s, double tau)
{
  const double g = a[i][j], h = a[k][l];
  a[i][j] = g - s * (h + g * tau);
  a[k][l] = h + s * (g - h * tau);
}

/*
 * Eigen decomposition of a symmetric 3x3 matrix by cyclic Jacobi sweeps
 * (Rutishauser's method in the form Numerical Recipes gives it, as
 * src/libutility/general.c:1163-1240 uses it): rotations above a threshold
 * of 0.2*S/9 in the first three sweeps, off-diagonal elements negligible
 * against both diagonal ones dropped after the fourth, at most 50 sweeps.
 * 1-based indices as there, so the rotations and the signs of the
 * eigenvectors come out the same.
 */
__device__ void prf_jacobi3(double a[4][4], double d[4], double v[4][4])
{
  double b[4], z[4];
  for (int p = 1; p <= 3; p++) {
    for (int q = 1; q <= 3; q++) v[p][q] = 0.0;
    v[p][p] = 1.0;
    b[p] = d[p] = a[p][p];
    z[p] = 0.0;
  }
  for (int sweep = 1; sweep <= 50; sweep++) {
    const double sm = (fabs(a[1][2]) + fabs(a[1][3])) + fabs(a[2][3]);
    if (sm == 0.0) return;
    const double tresh = sweep < 4 ? 0.2 * sm / 9 : 0.0;
    for (int p = 1; p <= 2; p++) {
      for (int q = p + 1; q <= 3; q++) {
        const double g = 100.0 * fabs(a[p][q]);
        if (sweep > 4 && fabs(d[p]) + g == fabs(d[p]) && fabs(d[q]) + g == fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if (fabs(h) + g == fabs(h)) {
            t = a[p][q] / h;
          } else {
            const double theta = 0.5 * h / a[p][q];
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          const double c = 1.0 / sqrt(1 + t * t), s = t * c, tau = s / (1.0 + c);
          h = t * a[p][q];
          z[p] -= h;
          z[q] += h;
          d[p] -= h;
          d[q] += h;
          a[p][q] = 0.0;
          for (int j = 1; j <= p - 1; j++) prf_rotate(a, j, p, j, q, s, tau);
          for (int j = p + 1; j <= q - 1; j++) prf_rotate(a, p, j, j, q, s, tau);
          for (int j = q + 1; j <= 3; j++) prf_rotate(a, p, j, q, j, s, tau);
          for (int j = 1; j <= 3; j++) prf_rotate(v, j, p, j, q, s, tau);
        }
      }
    }
    for (int p = 1; p <= 3; p++) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0;
    }
  }
}

/* get_axes (src/libutility/specific.c:135-170): eigenvalues largest first,
 * eigenvectors in the columns of e (ties keep their index order, as the
 * insertion pass of indexx does for three values) */
__device__ void prf_axes(const double t[6], double ax[3], double e[3][3])
{
  double a[4][4] = {}, d[4], v[4][4];
  a[1][1] = t[0];
  a[2][2] = t[1];
  a[3][3] = t[2];
  a[1][2] = a[2][1] = t[3];
  a[1][3] = a[3][1] = t[4];
  a[2][3] = a[3][2] = t[5];
  prf_jacobi3(a, d, v);
  int idx[4] = {0, 1, 2, 3};
  for (int j = 2; j <= 3; j++) {
    const int ix = idx[j];
    int i = j - 1;
    for (; i >= 1 && !(d[idx[i]] <= d[ix]); i--) idx[i + 1] = idx[i];
    idx[i + 1] = ix;
  }
  for (int k = 0; k < 3; k++) {
    ax[k] = d[idx[3 - k]];
    for (int r = 0; r < 3; r++) e[r][k] = v[r + 1][idx[3 - k]];
  }
}

/* calc_cNFW (specific.c:2024-2060): bisection on Prada et al. (2012) Eq. 9 */
__device__ double prf_cnfw(double V2_max, double V2_vir)
{
  const double ratio = V2_max / V2_vir;
  if (ratio <= 1 || ratio > 5.9) return -1;
  double a = 2.2, b = 100;
  while (b - a > 1e-3) {
    const double c = (a + b) / 2;
    const double fa = 0.216 * a / (log(1 + a) - a / (1 + a)) - ratio;
    const double fc = 0.216 * c / (log(1 + c) - c / (1 + c)) - ratio;
    if (fa * fc > 0) a = c;
    else b = c;
  }
  return (a + b) / 2.0;
}

/* calc_cR1 (specific.c:1990-2016): bisection on Wang et al.'s R1(c) */
__device__ double prf_cr1(double R1)
{
  if (R1 <= 0.19 || 0.585 <= R1) return -1.0;
  double a = 1.0, b = 500.0;
  while (b - a > 1e-3) {
    const double c = (a + b) / 2;
    const double fa = (a - 2.0 * log(1.0 + a) + a / (1.0 + a)) / (a * (log(1.0 + a) - a / (1.0 + a))) - R1;
    const double fc = (c - 2.0 * log(1.0 + c) + c / (1.0 + c)) / (c * (log(1.0 + c) - c / (1.0 + c))) - R1;
    if (fa * fc > 0) a = c;
    else b = c;
  }
  return (a + b) / 2.0;
}

/* ---------------------------------------------------- wavefront helpers */

/* a running sum over row[i0..i1): each term replaced by the running value.
 * Sixteen terms are read at a time so that one LDS round trip serves
 * sixteen dependent additions.  WITH_E: a second sum over the running
 * values less phi0 rides along (the potential energy over the potential,
 * ahf_halos.c:4383-4395: Upart = Phi - Phi0 with unit weights). */
template <bool WITH_E>
__device__ inline void chain_run(double *row, int i0, int i1, double &acc, double &acc_e, double phi0)
{
  double a = acc, e = acc_e;
  int i = i0;
  for (; i + 16 <= i1; i += 16) {
    double t[16];
#pragma unroll
    for (int k = 0; k < 16; k++) t[k] = row[i + k];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      a += t[k];
      t[k] = a;
      if (WITH_E) e += a - phi0;
    }
#pragma unroll
    for (int k = 0; k < 16; k++) row[i + k] = t[k];
  }
  for (; i < i1; i++) {
    a += row[i];
    row[i] = a;
    if (WITH_E) e += a - phi0;
  }
  acc = a;
  acc_e = e;
}

/* one running sum per lane that has a row (row != nullptr) over the cnt
 * staged terms: each term is replaced by the running value; at every member
 * of the masks cm (one per wavefront) the value is recorded for the bins it
 * closes, in column col (col < 0: not recorded), and with WITH_E the
 * potential energy in column ecol */
template <bool WITH_E, int NW>
__device__ inline void chain_pass(double *row, int cnt, ProfShared<NW> &S, double &acc, double &acc_e, double phi0,
                                  int col, int ecol)
{
  if (row) {
    int i = 0;
    for (int w = 0; w < NW; w++) {
      uint64_t m = S.cm[w];
      while (m) {
        const int e = 64 * w + __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        chain_run<WITH_E>(row, i, e + 1, acc, acc_e, phi0);
        i = e + 1;
        for (int b = S.bprev[e]; b < S.bcur[e]; b++) {
          if (col >= 0) S.bt[b * PQ_N + col] = acc;
          if (WITH_E && ecol >= 0) S.bt[b * PQ_N + ecol] = acc_e;
        }
      }
    }
    chain_run<WITH_E>(row, i, cnt, acc, acc_e, phi0);
  }
  __syncthreads();
}

/* chain_pass over explicit event arrays and without the barrier: two
 * passes of different chunks run on different wavefronts at once */
template <bool WITH_E, int NW>
__device__ inline void chain_events(double *row, int cnt, const uint64_t *cm, const int32_t *bprev,
                                    const int32_t *bcur, ProfShared<NW> &S, double &acc, double &acc_e, double phi0,
                                    int col, int ecol)
{
  if (!row) return;
  int i = 0;
  for (int w = 0; w < NW; w++) {
    uint64_t m = cm[w];
    while (m) {
      const int e = 64 * w + __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      chain_run<WITH_E>(row, i, e + 1, acc, acc_e, phi0);
      i = e + 1;
      for (int b = bprev[e]; b < bcur[e]; b++) {
        if (col >= 0) S.bt[b * PQ_N + col] = acc;
        if (WITH_E && ecol >= 0) S.bt[b * PQ_N + ecol] = acc_e;
      }
    }
  }
  chain_run<WITH_E>(row, i, cnt, acc, acc_e, phi0);
}

__device__ inline int wave_prefix_max(int v)
{
  const int lane = threadIdx.x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v = max(v, u);
  }
  return v;
}

/* y[t] after PASSES passes of smooth3 (general.c:548-566) over y[0..m),
 * m >= 3, from the raw values y[t-PASSES..t+PASSES]: each pass averages
 * three neighbours, two at either end, with smooth3's operand order */
template <int PASSES>
__device__ inline double smoothed_at(const double *__restrict__ y, int64_t m, int64_t t)
{
  double w[2 * PASSES + 1];
#pragma unroll
  for (int k = -PASSES; k <= PASSES; k++) {
    const int64_t i = t + k;
    w[k + PASSES] = (i >= 0 && i < m) ? y[i] : 0.0;
  }
#pragma unroll
  for (int p = 1; p <= PASSES; p++) {
    double nw[2 * PASSES + 1];
#pragma unroll
    for (int k = -(PASSES - p); k <= PASSES - p; k++) {
      const int64_t i = t + k;
      const int c = k + PASSES;
      double v = 0.0;
      if (i == 0) v = (w[c] + w[c + 1]) / 2.;
      else if (i == m - 1) v = (w[c] + w[c - 1]) / 2.;
      else if (i > 0 && i < m - 1) v = (w[c - 1] + w[c] + w[c + 1]) / 3.;
      nw[c] = v;
    }
#pragma unroll
    for (int k = -(PASSES - p); k <= PASSES - p; k++) w[k + PASSES] = nw[k + PASSES];
  }
  return w[PASSES];
}

/* (value, index) of a lane against another: find_max's left pass keeps the
 * first maximum above -10 (general.c:628-637) */
__device__ inline void first_max_merge(double &bv, int64_t &bi, double ov, int64_t oi)
{
  if (oi >= 0 && (bi < 0 || ov > bv || (!(bv > ov) && oi < bi))) {
    bv = ov;
    bi = oi;
  }
}

/* r2's and R_max's peaks in one pass over the members from NIGNORE on:
 * the first maxima over [0, m-2] of rho r^2 after three smoothing passes
 * and of V_c^2 after one (find_max's right pass cannot move a maximum of
 * its left pass: the range it scans lies inside the left pass's) */
template <int NW>
__device__ void find_peaks(const double *__restrict__ dr2, const double *__restrict__ vc2, int64_t m,
                           ProfShared<NW> &S, int64_t *i_r2, int64_t *i_rmax)
{
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  constexpr int C = 64 * NW;
  double bv3 = -10.0, bv1 = -10.0;
  int64_t bi3 = -1, bi1 = -1;
  for (int64_t t0 = 0; t0 < m - 1; t0 += 4 * C) {
    double v3[4], v1[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t t = t0 + u * C + tid;
      const bool ok = t < m - 1;
      const int64_t tt = ok ? t : 0;
      if (m >= 3) {
        v3[u] = smoothed_at<3>(dr2, m, tt);
        v1[u] = smoothed_at<1>(vc2, m, tt);
      } else {
        v3[u] = dr2[tt];
        v1[u] = vc2[tt];
      }
      if (!ok) v3[u] = v1[u] = -10.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t t = t0 + u * C + tid;
      if (v3[u] > bv3) {
        bv3 = v3[u];
        bi3 = t;
      }
      if (v1[u] > bv1) {
        bv1 = v1[u];
        bi1 = t;
      }
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    first_max_merge(bv3, bi3, __shfl_xor(bv3, d, 64), __shfl_xor(bi3, d, 64));
    first_max_merge(bv1, bi1, __shfl_xor(bv1, d, 64), __shfl_xor(bi1, d, 64));
  }
  __syncthreads();
  if (lane == 0) {
    S.rv[wid] = bv3;
    S.ri[wid] = bi3;
  }
  __syncthreads();
  double v = S.rv[0];
  int64_t i = S.ri[0];
  for (int w = 1; w < NW; w++) first_max_merge(v, i, S.rv[w], S.ri[w]);
  *i_r2 = i >= 0 ? i : m - 1;
  __syncthreads();
  if (lane == 0) {
    S.rv[wid] = bv1;
    S.ri[wid] = bi1;
  }
  __syncthreads();
  v = S.rv[0];
  i = S.ri[0];
  for (int w = 1; w < NW; w++) first_max_merge(v, i, S.rv[w], S.ri[w]);
  *i_rmax = i >= 0 ? i : m - 1;
  __syncthreads();
}

/* binning_parameter (specific.c:259-319) by one wavefront: the first
 * member from MINPART/10 on (below the last) whose distance is not below
 * MACHINE_ZERO, else half the last member's distance; distances with |d|
 * folded to 1-|d| beyond half the box; rad[0..nbins) the bin radii */
__device__ void prf_radii(const ProfConst &K, int64_t n, const uint32_t *__restrict__ ip, double Xc, double Yc,
                          double Zc, int nbins, const float *__restrict__ px, const float *__restrict__ py,
                          const float *__restrict__ pz, double *rad)
{
  const int lane = threadIdx.x & 63;
  auto dist_fold = [&](int64_t j) {
    const uint32_t p = ip[j];
    double dX = fabs((double)px[p] - Xc), dY = fabs((double)py[p] - Yc), dZ = fabs((double)pz[p] - Zc);
    if (dX > 0.5) dX = 1.0 - dX;
    if (dY > 0.5) dY = 1.0 - dY;
    if (dZ > 0.5) dZ = 1.0 - dZ;
    return sqrt(dX * dX + dY * dY + dZ * dZ);
  };
  const double dist_max = dist_fold(n - 1);
  double dist_min = dist_max / 2.;
  for (int64_t j0 = (int64_t)floor(((double)K.minpart / 10.) + 0.5); j0 < n - 1; j0 += 64) {
    const int64_t j = j0 + lane;
    const double d = j < n - 1 ? dist_fold(j) : 0.0;
    const uint64_t m = __ballot(j < n - 1 && !(d < PRF_MZERO));
    if (m) {
      dist_min = __shfl(d, __ffsll((unsigned long long)m) - 1, 64);
      break;
    }
  }
  const double ldmin = log10(dist_min), ldmax = log10(dist_max);
  const double ldr = (ldmax - ldmin) / (double)nbins;
  if (lane < nbins) rad[lane] = lane == nbins - 1 ? dist_max + PRF_ZERO : pow(10., ldmin + ((double)lane + 1) * ldr);
}

/* ------------------------------------------- the walk over the device */

/*
 * Haloes of prof_tiled members or more are walked by kernels over tiles of
 * 1024 members spread over the device, not by one workgroup.  Everything
 * of the walk but its running sums is a function of one member or a
 * running maximum over the members before it (the bins a member closes,
 * the last member beyond MACHINE_ZERO), so it runs per member after
 * prefix maxima over the tiles.  The running sums keep AHF's order and
 * roundings: each is one lane of a wavefront that advances all of them at
 * once, their terms staged in LDS by other wavefronts (k_prf_walk); the
 * sums whose terms depend on the bulk velocity's and the potential's
 * running values run two batches behind those.  k_profiles then takes the bins,
 * totals and most bound member from here and does the rest.  Bin radii
 * that decrease keep the sequential walk.
 */
#define PRF_TILE 1024
#define AHF_PROF_TILED 32768

struct PrfTiled {
  int64_t h, n, P, to; /* halo, members, members the walk reaches, offset in the tiled scratch */
  int64_t base;        /* the halo's offset in the member arrays */
  int64_t mb;          /* most bound member, -1: none */
  int64_t i_r2, i_rm;  /* r2's and R_max's members after NIGNORE (find_peaks) */
  int64_t ib;          /* the first member at or beyond R_max */
  double Xc, Yc, Zc, Phi0, Emin, Rmax;
  int32_t nbins, ok, t0, t1; /* ok: radii that do not decrease; tiles [t0, t1) */
  const uint32_t *ip;
  double rad[64];
  int64_t bclose[64];
  double bt[64][BT_N];
  double fin[16];
};

struct PrfTile {
  int32_t seg, pad;
  int64_t j0;
};

/* the per-member rows of the tiled walk (each Nt doubles or ints, Nt the
 * tiled haloes' members end to end) */
struct PrfRows {
  double *dd, *gx, *gy, *gz, *v[3], *tt, *phi, *phir, *com[3], *ine[6];
  int32_t *c, *bp, *bi;
};

/* a running sum of the tiled walk: terms in (minus Phi0 when sub_phi0),
 * running values to out (or none), recorded at bin closings in column col
 * (or none), the total to fin[fin] (or none) */
struct PrfChain {
  const double *in;
  double *out;
  int32_t col, fin, sub_phi0, pad;
};

/* per tiled halo (one wavefront): its members, centre, Phi0, bin radii */
__global__ void __launch_bounds__(64) k_prf_bin(ProfConst K, PrfTiled *T, const int64_t *__restrict__ off,
                                                const HaloOut *__restrict__ out, const double *__restrict__ phi_rvir,
                                                const uint32_t *__restrict__ ip_in, const uint32_t *__restrict__ buf_a,
                                                const uint32_t *__restrict__ buf_b, const double *__restrict__ cen,
                                                const int32_t *__restrict__ nbins_h, const float *__restrict__ px,
                                                const float *__restrict__ py, const float *__restrict__ pz)
{
  PrfTiled &H = T[blockIdx.x];
  const int64_t h = H.h, base = off[h];
  const uint32_t *ip = out ? (out[h].in_b ? buf_b : buf_a) + base : ip_in + base;
  const double Phi0 = out ? out[h].Phi0 : phi_rvir[2 * h];
  const int nbins = nbins_h[h];
  const double Xc = cen[3 * h], Yc = cen[3 * h + 1], Zc = cen[3 * h + 2];
  __shared__ double rad[64];
  prf_radii(K, H.n, ip, Xc, Yc, Zc, nbins, px, py, pz, rad);
  __syncthreads();
  const int lane = threadIdx.x;
  const bool mono = __ballot(lane < nbins - 1 && !(rad[lane] <= rad[lane + 1])) == 0;
  if (lane < nbins) H.rad[lane] = rad[lane];
  if (lane == 0) {
    H.ip = ip;
    H.base = base;
    H.Xc = Xc;
    H.Yc = Yc;
    H.Zc = Zc;
    H.Phi0 = Phi0;
    H.nbins = nbins;
    H.ok = mono ? 1 : 0;
    H.P = 0;
    H.mb = -1;
    H.Emin = 1e30;
  }
}

/* per member: offsets, distance, velocity, the bins its distance reaches;
 * per tile the largest of those and the last member beyond MACHINE_ZERO */
__global__ void __launch_bounds__(256) k_prf_geom(const PrfTiled *__restrict__ T, const PrfTile *__restrict__ tiles,
                                                  PrfRows R, const float4 *__restrict__ pv,
                                                  int32_t *__restrict__ tc_max, int64_t *__restrict__ tfl_max)
{
  __shared__ int32_t s_c;
  __shared__ unsigned long long s_fl;
  __shared__ double s_rad[64];
  const PrfTile Tl = tiles[blockIdx.x];
  const PrfTiled &H = T[Tl.seg];
  if (!H.ok) return;
  if (threadIdx.x == 0) {
    s_c = 0;
    s_fl = 0ull;
  }
  const int nbins = H.nbins;
  if (threadIdx.x < nbins) s_rad[threadIdx.x] = H.rad[threadIdx.x];
  __syncthreads();
  int cmax = 0;
  int64_t flmax = -1;
  for (int i = 0; i < PRF_TILE / 256; i++) {
    const int64_t j = Tl.j0 + (int64_t)i * 256 + threadIdx.x;
    if (j >= H.n) break;
    const int64_t e = H.to + j;
    const uint32_t p = H.ip[j];
    float fx, fy, fz, fvx, fvy, fvz;
    pv_member(pv, p, fx, fy, fz, fvx, fvy, fvz);
    double dX = (double)fx - H.Xc, dY = (double)fy - H.Yc, dZ = (double)fz - H.Zc;
    if (dX > 0.5) dX -= 1.0;
    if (dY > 0.5) dY -= 1.0;
    if (dZ > 0.5) dZ -= 1.0;
    if (dX < -0.5) dX += 1.0;
    if (dY < -0.5) dY += 1.0;
    if (dZ < -0.5) dZ += 1.0;
    const double d = sqrt(dX * dX + dY * dY + dZ * dZ);
    R.dd[e] = d;
    R.gx[e] = dX;
    R.gy[e] = dY;
    R.gz[e] = dZ;
    R.v[0][e] = (double)fvx;
    R.v[1][e] = (double)fvy;
    R.v[2][e] = (double)fvz;
    /* bins whose radius is not above d: the radii of a tiled halo do not
     * decrease, so AHF's walk up the bins stops where a binary search does */
    int lo = 0, hi = nbins;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_rad[mid] <= d)
        lo = mid + 1;
      else
        hi = mid;
    }
    int c = lo;
    if (j == H.n - 1) c = nbins;
    R.c[e] = c;
    cmax = max(cmax, c);
    if (d > PRF_MZERO) flmax = j;
  }
  if (cmax) atomicMax(&s_c, cmax);
  if (flmax >= 0) atomicMax(&s_fl, (unsigned long long)(flmax + 1));
  __syncthreads();
  if (threadIdx.x == 0) {
    tc_max[blockIdx.x] = s_c;
    tfl_max[blockIdx.x] = (int64_t)s_fl - 1;
  }
}

/* per tiled halo (one wavefront): the largest bin count and the last
 * member beyond MACHINE_ZERO before each tile (exclusive prefix maxima over
 * the halo's tiles, in place) */
__global__ void __launch_bounds__(64) k_prf_tscan(const PrfTiled *__restrict__ T, int32_t *__restrict__ tc_max,
                                                  int64_t *__restrict__ tfl_max)
{
  const PrfTiled &H = T[blockIdx.x];
  if (!H.ok) return;
  const int lane = threadIdx.x;
  int32_t c = 0;   /* over the tiles of earlier rounds */
  int64_t fl = -1;
  for (int32_t t0 = H.t0; t0 < H.t1; t0 += 64) {
    const int32_t t = t0 + lane;
    const bool in = t < H.t1;
    int32_t ci = in ? tc_max[t] : 0;
    int64_t fi = in ? tfl_max[t] : -1;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t cu = __shfl_up(ci, d, 64);
      const int64_t fu = __shfl_up(fi, d, 64);
      if (lane >= d) {
        ci = max(ci, cu);
        fi = max(fi, fu);
      }
    }
    int32_t ce = __shfl_up(ci, 1, 64);
    int64_t fe = __shfl_up(fi, 1, 64);
    if (lane == 0) {
      ce = 0;
      fe = -1;
    }
    if (in) {
      tc_max[t] = max(ce, c);
      tfl_max[t] = max(fe, fl);
    }
    c = max(c, __shfl(ci, 63, 64));
    fl = max(fl, __shfl(fi, 63, 64));
  }
}

/* per member: the bins it closes, whether the walk reaches it, and its
 * terms that do not depend on the bulk velocity (the potential's term,
 * centre of mass, inertia tensor, V_c^2, rho r^2, ahf_halos.c:4333-4395);
 * rounds of 256 consecutive members, the carries (bins closed and the last
 * member beyond MACHINE_ZERO before each member) as workgroup prefix maxima */
__global__ void __launch_bounds__(256) k_prf_terms1(PrfTiled *T, const PrfTile *__restrict__ tiles, PrfRows R,
                                                    const int32_t *__restrict__ tc_before,
                                                    const int64_t *__restrict__ tfl_before,
                                                    double *__restrict__ rad_m, double *__restrict__ vc2,
                                                    double *__restrict__ dr2)
{
  __shared__ int32_t s_wc[4];
  __shared__ long long s_wf[4];
  __shared__ unsigned long long s_P;
  const PrfTile Tl = tiles[blockIdx.x];
  PrfTiled &H = T[Tl.seg];
  if (!H.ok) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_P = 0ull;
  const double FourPiThird = 4. * PRF_PI / 3.;
  auto I_of = [&](int64_t k) { return k >= 0 ? (double)(k + 1) / (R.dd[H.to + k] * R.dd[H.to + k]) : 0.0; };
  /* over the rounds before: bins closed, last member beyond MACHINE_ZERO */
  int cx = tc_before[blockIdx.x];
  long long fx = tfl_before[blockIdx.x];
  int64_t last_proc = -1;
  for (int i = 0; i < PRF_TILE / 256; i++) {
    const int64_t j = Tl.j0 + (int64_t)i * 256 + tid;
    const bool inn = j < H.n;
    const int64_t e = H.to + (inn ? j : 0);
    const int c = inn ? R.c[e] : 0;
    const double d = inn ? R.dd[e] : 0.0;
    const long long f = inn && d > PRF_MZERO ? (long long)j : -1;
    /* inclusive prefix maxima over the workgroup's members of this round */
    int ci = c;
    long long fi = f;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
      const int cu = __shfl_up(ci, dd, 64);
      const long long fu = __shfl_up(fi, dd, 64);
      if (lane >= dd) {
        ci = max(ci, cu);
        fi = max(fi, fu);
      }
    }
    __syncthreads(); /* s_wc / s_wf of the round before are read */
    if (lane == 63) {
      s_wc[wid] = ci;
      s_wf[wid] = fi;
    }
    __syncthreads();
    int cw = cx;
    long long fw = fx;
    for (int w = 0; w < wid; w++) {
      cw = max(cw, s_wc[w]);
      fw = max(fw, s_wf[w]);
    }
    /* before this member (exclusive) and with it (inclusive) */
    int ce = __shfl_up(ci, 1, 64);
    long long fe = __shfl_up(fi, 1, 64);
    if (lane == 0) {
      ce = 0;
      fe = -1;
    }
    const int bp = max(cw, ce), bi = max(max(cw, ci), bp);
    const long long lf_prev = max(fw, fe), lf = max(fw, fi);
    for (int w = 0; w < 4; w++) {
      cx = max(cx, s_wc[w]);
      fx = max(fx, s_wf[w]);
    }
    if (!inn) continue;
    const double I_now = lf == j ? (double)(j + 1) / (d * d) : I_of(lf);
    const double I_prev = I_of(lf_prev);
    const bool proc = bp < H.nbins;
    R.bp[e] = bp;
    R.bi[e] = bi;
    if (proc) {
      last_proc = j;
      const double dX = R.gx[e], dY = R.gy[e], dZ = R.gz[e];
      const double pre = j > 0 ? R.dd[e - 1] : 0.0;
      const double M = (double)(j + 1);
      R.com[0][e] = H.Xc + dX;
      R.com[1][e] = H.Yc + dY;
      R.com[2][e] = H.Zc + dZ;
      R.ine[0][e] = dX * dX;
      R.ine[1][e] = dY * dY;
      R.ine[2][e] = dZ * dZ;
      R.ine[3][e] = dX * dY;
      R.ine[4][e] = dX * dZ;
      R.ine[5][e] = dY * dZ;
      R.phi[e] = d > PRF_MZERO ? ((I_now + I_prev) / 2.) * (d - pre) : -0.0; /* -0.0: an exact no-op */
      rad_m[H.base + j] = d;
      vc2[H.base + j] = M / d;
      const double dV = FourPiThird * ((d * d * d) - (pre * pre * pre));
      const double hm = (d + pre) / 2.;
      dr2[H.base + j] = 1.0 / dV * (hm * hm);
      for (int b = bp; b < bi; b++) H.bclose[b] = j;
    } else {
      /* members the walk does not reach read as AHF's zeroed arrays */
      rad_m[H.base + j] = vc2[H.base + j] = dr2[H.base + j] = 0.0;
    }
  }
  if (last_proc >= 0) atomicMax(&s_P, (unsigned long long)(last_proc + 1));
  __syncthreads();
  if (tid == 0 && s_P) atomicMax((unsigned long long *)&H.P, s_P);
}

/* the running sums over the members the walk reaches, in AHF's order
 * (ahf_halos.c:4333-4395): one lane of one wavefront per sum, so that one
 * FP64 addition advances all of them.
 *
 * The wavefront's lanes only read.  They take their terms, up to 32
 * members ahead, from columns in LDS (column c: sum c, PRF_FB members of a
 * batch), and keep the running value before every PRF_FS-th member as a
 * checkpoint.  Sums 0..3 (bulk velocity, potential) sum batch b while sums
 * 4..17 sum batch b - 2: their terms (angular momentum and kinetic energy
 * over the bulk velocity's running values, the potential energy over the
 * potential's) need the first four's running values of the batch.
 *
 * PRF_FH more wavefronts, each owning PRF_FQ rows of every batch (a row a
 * lane), finish and stage while the batches are summed.  They replay the
 * first four sums over batch b - 1 from their checkpoints (the same
 * additions from the same values) in place, the potential's also into its
 * rows.  They compute batch b - 1's terms of the other sums and
 * stage them, with the centre of mass, inertia and potential energy terms.
 * They stage the first four's terms of batch b + 1.  And they replay the
 * other sums over batch b - 3 up to the members that close bins, for the
 * bin records.
 *
 * Members past the last take -0.0, which leaves every sum as it is.  Sum c
 * of chains[c]; sum 17 (sub_phi0) adds the potential's running value less
 * Phi0. */
#define PRF_FB 448       /* members per batch */
#define PRF_FQ 64        /* rows a helper wavefront owns: a row a lane */
#define PRF_FH 7         /* helper wavefronts */
#define PRF_FROWS 482    /* a column: the batch, 34 rows read ahead; 2 * 482 = 4 mod 64 dwords between columns */
#define PRF_FS 16        /* members per checkpoint */
#define PRF_NV 4         /* sums to rows: bulk velocity, potential */
#define PRF_NL 18        /* sums */
static_assert(PRF_FB == PRF_FQ * PRF_FH && PRF_FQ == 64, "a helper lane owns one row");

/* one batch of one lane's sum over its column; ck[s * PRF_NL]: the value
 * before member s * PRF_FS of the batch, kept in registers until the batch
 * ends (a write to LDS between the reads would hold the reads behind it) */
__device__ __forceinline__ void prf_lane_batch(const double *col, double &a, double *ck)
{
  static_assert(PRF_FS == 16, "two checkpoints per 32 members");
  double ckr[PRF_FB / PRF_FS];
  double x[16], y[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = col[k];
#pragma unroll
  for (int i = 0; i < PRF_FB; i += 32) {
    /* rows i+16 .. i+47 are read ahead: at most PRF_FB + 15 < PRF_FROWS */
    ckr[i / PRF_FS] = a;
    /* the reads of the next sixteen issue before these sixteen additions
     * (the scheduler would otherwise gather them behind the additions) */
#pragma unroll
    for (int k = 0; k < 16; k++) y[k] = col[i + 16 + k];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 16; k++) a += x[k];
    __builtin_amdgcn_sched_barrier(0);
    ckr[i / PRF_FS + 1] = a;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = col[i + 32 + k];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 16; k++) a += y[k];
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int u = 0; u < PRF_FB / PRF_FS; u++) ck[u * PRF_NL] = ckr[u];
}

#define PRF_FW (1 + PRF_FH) /* wavefronts: the sums, then the helpers */
__global__ void __launch_bounds__(64 * PRF_FW) k_prf_walk(PrfTiled *T, const PrfChain *__restrict__ chains,
                                                          PrfRows R, ProfConst K)
{
  __shared__ __align__(16) double s_v[2][PRF_NL][PRF_FROWS];
  __shared__ double s_ck[2][PRF_FB / PRF_FS][PRF_NL];
  __shared__ int2 s_bins[2][PRF_FB];
  PrfTiled &H = T[blockIdx.x];
  if (!H.ok) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t P = H.P, to = H.to;
  const double Phi0 = H.Phi0;
  const int64_t nb = (P + PRF_FB - 1) / PRF_FB;
  /* step b sums slot b & 1: the first four over batch b, the others over
   * batch b - 2 (-0.0 where there is none) */
  const int64_t nit = nb + 3;
  if (w == 0) {
    double a = 0.0;
    __syncthreads(); /* the first batch staged */
    for (int64_t b = 0; b < nit; b++) {
      if (lane < PRF_NL) prf_lane_batch(&s_v[b & 1][lane][0], a, &s_ck[b & 1][0][lane]);
      __syncthreads();
    }
    if (lane < PRF_NL && chains[lane].fin >= 0) H.fin[chains[lane].fin] = a;
    return;
  }
  /* helper wavefront hw owns rows [hw PRF_FQ, (hw + 1) PRF_FQ) of every
   * batch, so that no other wavefront touches them between barriers; lane
   * l owns row r1 */
  const int hw = w - 1;
  const int r1 = hw * PRF_FQ + lane;
  /* global, not generic, accesses: generic ones count against the LDS
   * counter too, and the barrier's wait on that would stall on the loads
   * of later batches */
  typedef __attribute__((address_space(1))) double gdouble;
  typedef const __attribute__((address_space(1))) double gcdouble;
  typedef const __attribute__((address_space(1))) int32_t gcint;
  auto ld = [](const double *p, int64_t ix) { return *(gcdouble *)(p + ix); };
  const double hub = K.hubble, rf = K.r_fac, vf = K.v_fac;
  /* the first four's terms of one batch and the member data of another,
   * two sets in flight */
  struct Regs {
    double vt[PRF_NV]; /* terms of sums 0..3 */
    double g[3], v[3], com[3], ine[6];
    int32_t p, q; /* bins closed: [p, q) */
  };
  Regs R0, R1;
  /* set of step b: the first four's terms of batch b + 1 and the member
   * data of batch b - 1; load(Rg, bq) fills it for step bq + 1 */
  auto load = [&](Regs &Rg, int64_t bq) {
    if (bq + 2 >= 0 && bq + 2 < nb) {
      const int64_t j = (bq + 2) * PRF_FB + r1;
      const int64_t ix = to + (j < P ? j : P - 1);
#pragma unroll
      for (int c = 0; c < PRF_NV; c++) Rg.vt[c] = ld(chains[c].in, ix);
    }
    if (bq >= 0 && bq < nb) {
      const int64_t j = bq * PRF_FB + r1;
      const int64_t ix = to + (j < P ? j : P - 1);
      Rg.g[0] = ld(R.gx, ix);
      Rg.g[1] = ld(R.gy, ix);
      Rg.g[2] = ld(R.gz, ix);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        Rg.v[k] = ld(R.v[k], ix);
        Rg.com[k] = ld(R.com[k], ix);
      }
#pragma unroll
      for (int k = 0; k < 6; k++) Rg.ine[k] = ld(R.ine[k], ix);
      Rg.p = *(gcint *)(R.bp + ix);
      Rg.q = *(gcint *)(R.bi + ix);
    }
  };
  /* the first four's terms of batch B into slot s */
  auto stage_v = [&](const Regs &Rg, int s, int64_t B) {
    const bool in = B < nb && B * PRF_FB + r1 < P;
#pragma unroll
    for (int c = 0; c < PRF_NV; c++) s_v[s][c][r1] = in ? Rg.vt[c] : -0.0;
  };
  /* the other sums' terms of batch B into slot s, from the member data and
   * the first four's running values there (ahf_halos.c:4333-4395) */
  auto stage_c = [&](const Regs &Rg, int s, int64_t B) {
    const int64_t j = B * PRF_FB + r1;
    const bool in = B >= 0 && B < nb && j < P;
    double col[PRF_NL - PRF_NV];
    const double M = (double)(j + 1);
    const double dX = Rg.g[0], dY = Rg.g[1], dZ = Rg.g[2];
    double dVX = Rg.v[0] - s_v[s][0][r1] / M, dVY = Rg.v[1] - s_v[s][1][r1] / M, dVZ = Rg.v[2] - s_v[s][2][r1] / M;
    col[0] = dY * dVZ - dZ * dVY;
    col[1] = dZ * dVX - dX * dVZ;
    col[2] = dX * dVY - dY * dVX;
    dVX += hub * dX * rf / vf;
    dVY += hub * dY * rf / vf;
    dVZ += hub * dZ * rf / vf;
    col[3] = dVX * dVX + dVY * dVY + dVZ * dVZ;
    if (in) R.tt[to + j] = col[3];
#pragma unroll
    for (int c = 0; c < 3; c++) col[4 + c] = Rg.com[c];
#pragma unroll
    for (int c = 0; c < 6; c++) col[7 + c] = Rg.ine[c];
    col[13] = s_v[s][3][r1] - Phi0;
#pragma unroll
    for (int c = 0; c < PRF_NL - PRF_NV; c++) s_v[s][PRF_NV + c][r1] = in ? col[c] : -0.0;
    s_bins[s][r1] = in ? make_int2(Rg.p, Rg.q) : make_int2(0, 0);
  };
  /* the first four over batch B (slot s) in place, the potential's also
   * into its rows: lane (u, c) replays sum c over checkpoint span u of the
   * wavefront's rows */
  constexpr int NU = PRF_FQ / PRF_FS;
  static_assert(NU * PRF_NV <= 64, "one lane per sum and span");
  const bool vact = lane < NU * PRF_NV;
  const int vc = vact ? lane % PRF_NV : 0, vu = vact ? lane / PRF_NV : 0;
  gdouble *const vout = (gdouble *)chains[vc].out; /* the potential's rows; the others in LDS only */
  const bool vst = vout != nullptr;
  auto replay_v = [&](int s, int64_t B) {
    const int64_t r0 = B * PRF_FB + hw * PRF_FQ + vu * PRF_FS;
    const int64_t mq = P - r0 < PRF_FS ? P - r0 : PRF_FS;
    if (!vact || mq <= 0) return;
    double *colp = &s_v[s][vc][hw * PRF_FQ + vu * PRF_FS];
    double a = s_ck[s][hw * NU + vu][vc];
    gdouble *o = vst ? vout + to + r0 : nullptr;
    if (mq == PRF_FS) {
      double t[PRF_FS];
#pragma unroll
      for (int k = 0; k < PRF_FS; k++) t[k] = colp[k];
#pragma unroll
      for (int k = 0; k < PRF_FS; k++) {
        a += t[k];
        colp[k] = a;
        if (vst) o[k] = a;
      }
    } else {
      for (int k = 0; k < mq; k++) {
        a += colp[k];
        colp[k] = a;
        if (vst) o[k] = a;
      }
    }
  };
  /* the other sums over batch B (slot s) up to each of the wavefront's
   * members that closes bins, in order, from the checkpoint of its span */
  const int cc = lane < PRF_NL - PRF_NV ? PRF_NV + lane : PRF_NV;
  const int ccol = chains[cc].col;
  auto replay_bins = [&](int s) {
    const int2 q1 = s_bins[s][r1];
    uint64_t m = __ballot(q1.x < q1.y);
    const double *colp = &s_v[s][cc][hw * PRF_FQ];
    double a = 0.0;
    int pos = 0, span = -1;
    while (m) {
      const int rr = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      if (rr / PRF_FS != span) {
        span = rr / PRF_FS;
        pos = span * PRF_FS;
        a = s_ck[s][hw * NU + span][cc];
      }
      for (; pos <= rr; pos++) a += colp[pos];
      const int2 q = s_bins[s][hw * PRF_FQ + rr];
      if (lane < PRF_NL - PRF_NV && ccol >= 0)
        for (int x = q.x; x < q.y; x++) H.bt[x][ccol] = a;
    }
  };
  /* step b: slot (b + 1) & 1 holds the first four over batch b - 1 and
   * the others over batch b - 3, summed at step b - 1 */
  auto step = [&](int64_t b, Regs &Rg) {
    const int s = (int)((b + 1) & 1);
    if (b - 3 >= 0 && b - 3 < nb) replay_bins(s);
    if (b - 1 >= 0 && b - 1 < nb) replay_v(s, b - 1);
    stage_c(Rg, s, b - 1);
    stage_v(Rg, s, b + 1);
    load(Rg, b + 1);
    __syncthreads();
  };
  /* slot 0 for step 0: the first four over batch 0, the others -0.0 */
  load(R0, -2);
  stage_v(R0, 0, 0);
  stage_c(R0, 0, -1);
  load(R1, -1);
  load(R0, 0);
  __syncthreads();
  int64_t b = 0;
  for (; b + 2 <= nit; b += 2) {
    step(b, R1);
    step(b + 1, R0);
  }
  if (b < nit) step(b, R1);
}

/* per member the walk reaches: escape speed at the bins it closes, and
 * per tile the first member at the lowest energy below 1e30 */
__global__ void __launch_bounds__(256) k_prf_finish(PrfTiled *T, const PrfTile *__restrict__ tiles, PrfRows R,
                                                    double *__restrict__ tEmin, int64_t *__restrict__ tmb)
{
  __shared__ double s_v[4];
  __shared__ int64_t s_i[4];
  const PrfTile Tl = tiles[blockIdx.x];
  PrfTiled &H = T[Tl.seg];
  if (!H.ok) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double bv = 1e30;
  int64_t bl = -1;
  for (int i = 0; i < PRF_TILE / 256; i++) {
    const int64_t j = Tl.j0 + (int64_t)i * 256 + tid;
    if (j >= H.P) break;
    const int64_t e = H.to + j;
    const double U = R.phir[e] - H.Phi0;
    const double Ep = 0.5 * R.tt[e] + U;
    for (int b = R.bp[e]; b < R.bi[e]; b++) H.bt[b][BT_VESC] = 2 * fabs(U);
    if (Ep < bv) {
      bv = Ep;
      bl = j;
    }
  }
  /* the first at the lowest: a lower index wins a tie */
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const double ov = __shfl_xor(bv, d, 64);
    const int64_t ol = __shfl_xor(bl, d, 64);
    if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
      bv = ov;
      bl = ol;
    }
  }
  if (lane == 0) {
    s_v[wid] = bv;
    s_i[wid] = bl;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; w++) {
      const double ov = s_v[w];
      const int64_t ol = s_i[w];
      if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
        bv = ov;
        bl = ol;
      }
    }
    tEmin[blockIdx.x] = bv;
    tmb[blockIdx.x] = bl;
  }
}

/* per tiled halo (one wavefront): the most bound member over its tiles,
 * the lower index on equal energies */
__global__ void __launch_bounds__(64) k_prf_done(PrfTiled *T, const double *__restrict__ tEmin,
                                                 const int64_t *__restrict__ tmb)
{
  PrfTiled &H = T[blockIdx.x];
  if (!H.ok) return;
  const int lane = threadIdx.x;
  double bv = 1e30;
  int64_t bl = -1;
  auto merge = [&](double ov, int64_t ol) {
    if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
      bv = ov;
      bl = ol;
    }
  };
  for (int32_t t = H.t0 + lane; t < H.t1; t += 64) merge(tEmin[t], tmb[t]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) merge(__shfl_xor(bv, d, 64), __shfl_xor(bl, d, 64));
  if (lane == 0) {
    H.Emin = bv;
    H.mb = bl;
  }
}

/* r2's and R_max's peaks (find_peaks) over one tile of a halo walked over
 * the device: the first maxima of the tile's t = member - NIGNORE in
 * [j0, j0 + PRF_TILE), t < m - 1; tv/ti[2 tile + 0]: rho r^2 after three
 * smoothing passes, [2 tile + 1]: V_c^2 after one */
__global__ void __launch_bounds__(256) k_prf_peaks(const PrfTiled *__restrict__ T, const PrfTile *__restrict__ tiles,
                                                   const double *__restrict__ vc2, const double *__restrict__ dr2,
                                                   double *__restrict__ tv, int64_t *__restrict__ ti)
{
  __shared__ double s_v[2][4];
  __shared__ int64_t s_i[2][4];
  const PrfTile Tl = tiles[blockIdx.x];
  const PrfTiled &H = T[Tl.seg];
  if (!H.ok) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t m = H.n - PRF_NIGNORE;
  const double *y3 = dr2 + H.base + PRF_NIGNORE, *y1 = vc2 + H.base + PRF_NIGNORE;
  double bv3 = -10.0, bv1 = -10.0;
  int64_t bi3 = -1, bi1 = -1;
  for (int u = 0; u < PRF_TILE / 256; u++) {
    const int64_t t = Tl.j0 + u * 256 + tid;
    if (t >= m - 1) break;
    const double v3 = m >= 3 ? smoothed_at<3>(y3, m, t) : y3[t];
    const double v1 = m >= 3 ? smoothed_at<1>(y1, m, t) : y1[t];
    if (v3 > bv3) {
      bv3 = v3;
      bi3 = t;
    }
    if (v1 > bv1) {
      bv1 = v1;
      bi1 = t;
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    first_max_merge(bv3, bi3, __shfl_xor(bv3, d, 64), __shfl_xor(bi3, d, 64));
    first_max_merge(bv1, bi1, __shfl_xor(bv1, d, 64), __shfl_xor(bi1, d, 64));
  }
  if (lane == 0) {
    s_v[0][wid] = bv3;
    s_i[0][wid] = bi3;
    s_v[1][wid] = bv1;
    s_i[1][wid] = bi1;
  }
  __syncthreads();
  if (tid < 2) {
    double v = s_v[tid][0];
    int64_t i = s_i[tid][0];
    for (int w = 1; w < 4; w++) first_max_merge(v, i, s_v[tid][w], s_i[tid][w]);
    tv[2 * blockIdx.x + tid] = v;
    ti[2 * blockIdx.x + tid] = i;
  }
}

/* per tiled halo (one wavefront): the peaks over its tiles, R_max */
__global__ void __launch_bounds__(64) k_prf_peaks_done(PrfTiled *T, const double *__restrict__ rad_m,
                                                       const double *__restrict__ tv, const int64_t *__restrict__ ti)
{
  PrfTiled &H = T[blockIdx.x];
  if (!H.ok) return;
  const int lane = threadIdx.x;
  double bv3 = -10.0, bv1 = -10.0;
  int64_t bi3 = -1, bi1 = -1;
  for (int32_t t = H.t0 + lane; t < H.t1; t += 64) {
    first_max_merge(bv3, bi3, tv[2 * t], ti[2 * t]);
    first_max_merge(bv1, bi1, tv[2 * t + 1], ti[2 * t + 1]);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    first_max_merge(bv3, bi3, __shfl_xor(bv3, d, 64), __shfl_xor(bi3, d, 64));
    first_max_merge(bv1, bi1, __shfl_xor(bv1, d, 64), __shfl_xor(bi1, d, 64));
  }
  if (lane == 0) {
    const int64_t m = H.n - PRF_NIGNORE;
    H.i_r2 = bi3 >= 0 ? bi3 : m - 1;
    H.i_rm = bi1 >= 0 ? bi1 : m - 1;
    const double *x = rad_m + H.base + PRF_NIGNORE;
    H.Rmax = (x[H.i_rm] + x[H.i_rm]) / 2;
  }
}

/* per tile: the first member below the last at or beyond R_max
 * (ahf_halos.c:4892-4894), INT64_MAX: none */
__global__ void __launch_bounds__(256) k_prf_rmax(const PrfTiled *__restrict__ T, const PrfTile *__restrict__ tiles,
                                                  const double *__restrict__ rad_m, int64_t *__restrict__ tib)
{
  __shared__ int64_t s_i[4];
  const PrfTile Tl = tiles[blockIdx.x];
  const PrfTiled &H = T[Tl.seg];
  if (!H.ok) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t f = INT64_MAX;
  for (int u = PRF_TILE / 256 - 1; u >= 0; u--) {
    const int64_t j = Tl.j0 + u * 256 + tid;
    if (j < H.n - 1 && !(rad_m[H.base + j] < H.Rmax)) f = j;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) f = min(f, (int64_t)__shfl_xor(f, d, 64));
  if (lane == 0) s_i[wid] = f;
  __syncthreads();
  if (tid == 0) tib[blockIdx.x] = min(min(s_i[0], s_i[1]), min(s_i[2], s_i[3]));
}

/* per tiled halo (one wavefront): the first member at or beyond R_max over
 * its tiles, else the last */
__global__ void __launch_bounds__(64) k_prf_rmax_done(PrfTiled *T, const int64_t *__restrict__ tib)
{
  PrfTiled &H = T[blockIdx.x];
  if (!H.ok) return;
  const int lane = threadIdx.x;
  int64_t f = INT64_MAX;
  for (int32_t t = H.t0 + lane; t < H.t1; t += 64) f = min(f, tib[t]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) f = min(f, (int64_t)__shfl_xor(f, d, 64));
  if (lane == 0) H.ib = f < INT64_MAX ? f : H.n - 1;
}

/* ------------------------------------------------------------- kernel */

/* one workgroup of NW wavefronts per halo: haloes list[blockIdx.x] (list:
 * the haloes this launch takes; NW = 1 skips haloes of big or more members,
 * which a launch with more wavefronts takes) */
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_profiles(
    ProfConst K, const int64_t *__restrict__ off, const int32_t *__restrict__ list, int64_t big,
    const HaloOut *__restrict__ out, const double *__restrict__ phi_rvir, const uint32_t *__restrict__ ip_in,
    const uint32_t *__restrict__ buf_a, const uint32_t *__restrict__ buf_b, const double *__restrict__ cen,
    const int32_t *__restrict__ nbins_h, const int64_t *__restrict__ pofs, const float *__restrict__ px,
    const float *__restrict__ py, const float *__restrict__ pz, const float *__restrict__ vx,
    const float *__restrict__ vy, const float *__restrict__ vz, double *__restrict__ s_rad,
    double *__restrict__ s_vc2, double *__restrict__ s_dr2, double *__restrict__ props, double *__restrict__ prof,
    const PrfTiled *__restrict__ tl, const int32_t *__restrict__ tmap, const float4 *__restrict__ pv)
{
  constexpr int C = 64 * NW;
  __shared__ ProfShared<NW> S;
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const int64_t h = list[blockIdx.x];
  const int nbins = nbins_h[h];
  if (nbins <= 0) return;
  const int64_t base = off[h];
  int64_t n;
  const uint32_t *ip;
  double Phi0, R_vir;
  if (out) {
    n = out[h].npart;
    ip = (out[h].in_b ? buf_b : buf_a) + base;
    Phi0 = out[h].Phi0;
    R_vir = out[h].R_vir;
  } else {
    n = off[h + 1] - base;
    ip = ip_in + base;
    Phi0 = phi_rvir[2 * h];
    R_vir = phi_rvir[2 * h + 1];
  }
  if (NW == 1 && n >= big) return;
  const double Xc = cen[3 * h], Yc = cen[3 * h + 1], Zc = cen[3 * h + 2];
  double *rad_m = s_rad + base, *vc2 = s_vc2 + base, *dr2 = s_dr2 + base;
  const double FourPiThird = 4. * PRF_PI / 3.;

  /* binning_parameter (specific.c:259-319): the first member from
   * MINPART/10 on (below the last) whose distance is not below MACHINE_ZERO,
   * else half the last member's distance; distances with |d| folded to
   * 1-|d| beyond half the box (the first wavefront; LDS to the others) */
  /* a halo walked over the device (k_prf_*) brings its radii, bins,
   * totals and most bound member */
  const PrfTiled *TP = (tmap && tmap[h] >= 0 && tl[tmap[h]].ok) ? &tl[tmap[h]] : nullptr;
  if (tid == 0) S.bt = prof + pofs[h] * PQ_N;
  if (TP) {
    if (tid < nbins) S.rad[tid] = TP->rad[tid];
  } else if (wid == 0) {
    prf_radii(K, n, ip, Xc, Yc, Zc, nbins, px, py, pz, S.rad);
  }
  __syncthreads();
  const bool mono = __ballot(lane < nbins - 1 && !(S.rad[lane] <= S.rad[lane + 1])) == 0;

  /* the walk: per chunk, the bulk velocity's running sums (lanes 0-2),
   * then the terms about it, then every other running sum in one pass:
   * angular momentum (lanes 0-2), kinetic energy (3), the potential and
   * with it the potential energy (4), centre of mass (5-7), inertia tensor
   * (8-13); the sums run on the first wavefront */
  double accV = 0.0, accC = 0.0, accE = 0.0, dummy = 0.0;
  double I_car = 0.0, d_car = 0.0, Emin = 1e30;
  int64_t mb = -1, P = 0;
  int bcar = 0;
  const double hub = K.hubble, rf = K.r_fac, vf = K.v_fac;
  double *const rowV = tid < 3 ? S.ta[9 + tid] : nullptr;
  double *const rowC = tid < 5 ? S.tc[tid] : tid < 14 ? S.ta[tid - 5] : nullptr;
  const int colC = tid < 4 ? BT_LX + tid : (tid >= 8 && tid < 14) ? BT_A11 + tid - 8 : -1;
  const int ecolC = tid == 4 ? BT_EP : -1;
  /* the members of the next chunk are read while the current one is
   * summed, their indices one chunk earlier still */
  float nx = 0.f, ny = 0.f, nz = 0.f, nvx = 0.f, nvy = 0.f, nvz = 0.f;
  if (tid < n) pv_member(pv, ip[tid], nx, ny, nz, nvx, nvy, nvz);
  uint32_t qn = C + tid < n ? ip[C + tid] : 0;
  if (TP) {
    P = TP->P;
    mb = TP->mb;
    for (int k = tid; k < nbins * BT_N; k += C) S.bt[(k / BT_N) * PQ_N + k % BT_N] = TP->bt[k / BT_N][k % BT_N];
    if (tid < 64) S.bclose[tid] = TP->bclose[tid];
    if (tid < 16) S.fin[tid] = TP->fin[tid];
  } else if constexpr (NW > 1) {
    /* two chunks in flight: the bulk velocity's sums of chunk k+1 run on the
     * second wavefront while the first runs chunk k's other sums */
    struct Chunk {
      int64_t j0;
      int cnt, bp, bi;
      bool proc;
      double dX, dY, dZ, d, vxp, vyp, vzp, T;
    };
    double *const rowV2 = (tid >= 64 && tid < 67) ? S.ta[9 + tid - 64] : nullptr;
    /* stage chunk j0 into buffer par: geometry, bins, terms */
    auto stage = [&](int64_t j0, int par, Chunk &r) {
      const int64_t j = j0 + tid;
      const bool inb = j < n;
      r.j0 = j0;
      r.dX = r.dY = r.dZ = r.d = r.vxp = r.vyp = r.vzp = r.T = 0.0;
      const float cx = nx, cy = ny, cz = nz, cvx = nvx, cvy = nvy, cvz = nvz;
      if (j + C < n) pv_member(pv, qn, nx, ny, nz, nvx, nvy, nvz);
      qn = j + 2 * C < n ? ip[j + 2 * C] : 0;
      if (inb) {
        double dX = (double)cx - Xc, dY = (double)cy - Yc, dZ = (double)cz - Zc;
        if (dX > 0.5) dX -= 1.0;
        if (dY > 0.5) dY -= 1.0;
        if (dZ > 0.5) dZ -= 1.0;
        if (dX < -0.5) dX += 1.0;
        if (dY < -0.5) dY += 1.0;
        if (dZ < -0.5) dZ += 1.0;
        r.dX = dX;
        r.dY = dY;
        r.dZ = dZ;
        r.d = sqrt(dX * dX + dY * dY + dZ * dZ);
        r.vxp = (double)cvx;
        r.vyp = (double)cvy;
        r.vzp = (double)cvz;
      }
      const double d = r.d;
      S.d_s[tid] = d;
      int bp, bi;
      if (mono) {
        int c = bcar;
        if (inb) {
          while (c < nbins && S.rad[c] <= d) c++;
          if (j == n - 1) c = nbins;
        }
        const uint64_t any = __ballot(c > bcar);
        if (lane == 0) S.hm[wid] = any;
        __syncthreads();
        bool some = false;
        for (int w = 0; w < NW; w++) some |= S.hm[w] != 0;
        bi = bcar;
        if (some) {
          bi = wave_prefix_max(c);
          if (lane == 63) S.wv[wid] = bi;
          __syncthreads();
          for (int w = 0; w < wid; w++) bi = max(bi, S.wv[w]);
          bi = max(bi, bcar);
        }
        S.bc2[par][tid] = bi;
        __syncthreads();
        bp = tid ? S.bc2[par][tid - 1] : bcar;
      } else {
        __syncthreads();
        if (tid == 0) {
          int b = bcar;
          for (int i = 0; i < C; i++) {
            S.bp2[par][i] = b;
            if (b < nbins && j0 + i < n)
              while (b < nbins && (S.d_s[i] >= S.rad[b] || j0 + i == n - 1)) b++;
            S.bc2[par][i] = b;
          }
        }
        __syncthreads();
        bp = S.bp2[par][tid];
        bi = S.bc2[par][tid];
      }
      const bool proc = inb && bp < nbins;
      {
        const uint64_t pm = __ballot(proc), cmw = __ballot(proc && bi > bp);
        if (lane == 0) {
          S.wv[wid] = __popcll(pm);
          S.cm2[par][wid] = cmw;
        }
      }
      __syncthreads();
      int cnt = 0;
      for (int w = 0; w < NW; w++) cnt += S.wv[w];
      if (proc) {
        S.tg[par][0][tid] = Xc + r.dX;
        S.tg[par][1][tid] = Yc + r.dY;
        S.tg[par][2][tid] = Zc + r.dZ;
        S.tg[par][3][tid] = r.dX * r.dX;
        S.tg[par][4][tid] = r.dY * r.dY;
        S.tg[par][5][tid] = r.dZ * r.dZ;
        S.tg[par][6][tid] = r.dX * r.dY;
        S.tg[par][7][tid] = r.dX * r.dZ;
        S.tg[par][8][tid] = r.dY * r.dZ;
        S.ta[9][tid] = r.vxp;
        S.ta[10][tid] = r.vyp;
        S.ta[11][tid] = r.vzp;
        S.bp2[par][tid] = bp;
        S.bc2[par][tid] = bi;
        for (int b = bp; b < bi; b++) S.bclose[b] = j;
        rad_m[j] = d;
      }
      __syncthreads();
      bcar = S.bc2[par][cnt - 1];
      r.cnt = cnt;
      r.bp = bp;
      r.bi = bi;
      r.proc = proc;
    };
    /* the terms about the running bulk velocity (ahf_halos.c:4333-4395) */
    auto terms = [&](Chunk &r) {
      const int64_t j = r.j0 + tid;
      const double M = (double)(j + 1), d = r.d;
      double dVX = r.vxp - S.ta[9][tid] / M, dVY = r.vyp - S.ta[10][tid] / M, dVZ = r.vzp - S.ta[11][tid] / M;
      const double lx = r.dY * dVZ - r.dZ * dVY, ly = r.dZ * dVX - r.dX * dVZ, lz = r.dX * dVY - r.dY * dVX;
      dVX += hub * r.dX * rf / vf;
      dVY += hub * r.dY * rf / vf;
      dVZ += hub * r.dZ * rf / vf;
      r.T = dVX * dVX + dVY * dVY + dVZ * dVZ;
      const bool fl = r.proc && d > PRF_MZERO;
      S.i_s[tid] = M / (d * d);
      {
        const uint64_t fmw = __ballot(fl);
        if (lane == 0) S.fm[wid] = fmw;
      }
      __syncthreads();
      int64_t src = -1;
      {
        const uint64_t le = S.fm[wid] & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
        if (le) src = 64 * wid + 63 - __clzll((long long)le);
        for (int w = wid - 1; w >= 0 && src < 0; w--)
          if (S.fm[w]) src = 64 * w + 63 - __clzll((long long)S.fm[w]);
      }
      const double I_now = src >= 0 ? S.i_s[src] : I_car;
      __syncthreads();
      S.i_s[tid] = I_now;
      __syncthreads();
      const double I_prev = tid ? S.i_s[tid - 1] : I_car, pre = tid ? S.d_s[tid - 1] : d_car;
      if (r.proc) {
        S.tc[0][tid] = lx;
        S.tc[1][tid] = ly;
        S.tc[2][tid] = lz;
        S.tc[3][tid] = r.T;
        S.tc[4][tid] = fl ? ((I_now + I_prev) / 2.) * (d - pre) : -0.0; /* -0.0: an exact no-op */
        vc2[j] = M / d;
        const double dV = FourPiThird * ((d * d * d) - (pre * pre * pre));
        const double hm = (d + pre) / 2.;
        dr2[j] = 1.0 / dV * (hm * hm);
      }
      I_car = S.i_s[r.cnt - 1];
      d_car = S.d_s[r.cnt - 1];
      __syncthreads();
    };
    /* energies, escape speeds, the most bound member */
    auto finish = [&](const Chunk &r) {
      const double U = S.tc[4][tid] - Phi0;
      const double Ep = 0.5 * r.T + U;
      if (r.proc)
        for (int b = r.bp; b < r.bi; b++) S.bt[b * PQ_N + BT_VESC] = 2 * fabs(U);
      const bool cand = r.proc && Ep < Emin;
      const uint64_t cw = __ballot(cand);
      if (lane == 0) S.hm[wid] = cw;
      __syncthreads();
      bool some = false;
      for (int w = 0; w < NW; w++) some |= S.hm[w] != 0;
      if (some) {
        double bv = cand ? Ep : 0.0;
        int64_t bl = cand ? tid : -1;
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) {
          const double ov = __shfl_xor(bv, dd, 64);
          const int64_t ol = __shfl_xor(bl, dd, 64);
          if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
            bv = ov;
            bl = ol;
          }
        }
        if (lane == 0) {
          S.rv[wid] = bv;
          S.ri[wid] = bl;
        }
        __syncthreads();
        bv = S.rv[0];
        bl = S.ri[0];
        for (int w = 1; w < NW; w++) {
          const double ov = S.rv[w];
          const int64_t ol = S.ri[w];
          if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
            bv = ov;
            bl = ol;
          }
        }
        Emin = bv;
        mb = r.j0 + bl;
      }
      __syncthreads();
    };
    Chunk cur, nxt;
    int par = 0;
    stage(0, 0, cur);
    chain_events<false, NW>(rowV2, cur.cnt, S.cm2[0], S.bp2[0], S.bc2[0], S, accV, dummy, 0.0, -1, -1);
    __syncthreads();
    terms(cur);
    for (;;) {
      const bool next = cur.cnt == C && cur.j0 + C < n && bcar < nbins;
      if (next) stage(cur.j0 + C, par ^ 1, nxt);
      /* chunk k+1's bulk velocity beside chunk k's other sums */
      if (next) chain_events<false, NW>(rowV2, nxt.cnt, S.cm2[par ^ 1], S.bp2[par ^ 1], S.bc2[par ^ 1], S, accV,
                                        dummy, 0.0, -1, -1);
      double *const rowCk = tid < 5 ? S.tc[tid] : tid < 14 ? S.tg[par][tid - 5] : nullptr;
      chain_events<true, NW>(rowCk, cur.cnt, S.cm2[par], S.bp2[par], S.bc2[par], S, accC, accE, Phi0, colC, ecolC);
      __syncthreads();
      finish(cur);
      P += cur.cnt;
      if (!next) break;
      terms(nxt);
      cur = nxt;
      par ^= 1;
    }
    if (tid >= 64 && tid < 67) S.fin[9 + tid - 64] = accV;
  } else {
  for (int64_t j0 = 0; j0 < n && bcar < nbins; j0 += C) {
    const int64_t j = j0 + tid;
    const bool inb = j < n;
    double dX = 0.0, dY = 0.0, dZ = 0.0, d = 0.0, vxp = 0.0, vyp = 0.0, vzp = 0.0;
    const float cx = nx, cy = ny, cz = nz, cvx = nvx, cvy = nvy, cvz = nvz;
    if (j + C < n) pv_member(pv, qn, nx, ny, nz, nvx, nvy, nvz);
    qn = j + 2 * C < n ? ip[j + 2 * C] : 0;
    if (inb) {
      dX = (double)cx - Xc;
      dY = (double)cy - Yc;
      dZ = (double)cz - Zc;
      if (dX > 0.5) dX -= 1.0;
      if (dY > 0.5) dY -= 1.0;
      if (dZ > 0.5) dZ -= 1.0;
      if (dX < -0.5) dX += 1.0;
      if (dY < -0.5) dY += 1.0;
      if (dZ < -0.5) dZ += 1.0;
      d = sqrt(dX * dX + dY * dY + dZ * dZ);
      vxp = (double)cvx;
      vyp = (double)cvy;
      vzp = (double)cvz;
    }
    S.d_s[tid] = d;
    /* bins closed after each member */
    int bp, bi;
    if (mono) {
      /* the bins from the current one on whose radius d reaches (with
       * radii that do not decrease, those below it are closed already) */
      int c = bcar;
      if (inb) {
        while (c < nbins && S.rad[c] <= d) c++;
        if (j == n - 1) c = nbins;
      }
      const uint64_t any = __ballot(c > bcar);
      if (lane == 0) S.hm[wid] = any;
      __syncthreads();
      bool some = false;
      for (int w = 0; w < NW; w++) some |= S.hm[w] != 0;
      bi = bcar;
      if (some) {
        bi = wave_prefix_max(c);
        if (lane == 63) S.wv[wid] = bi;
        __syncthreads();
        for (int w = 0; w < wid; w++) bi = max(bi, S.wv[w]);
        bi = max(bi, bcar);
      }
      S.bcur[tid] = bi;
      __syncthreads();
      bp = tid ? S.bcur[tid - 1] : bcar;
    } else {
      __syncthreads();
      if (tid == 0) {
        int b = bcar;
        for (int i = 0; i < C; i++) {
          S.bprev[i] = b;
          if (b < nbins && j0 + i < n)
            while (b < nbins && (S.d_s[i] >= S.rad[b] || j0 + i == n - 1)) b++;
          S.bcur[i] = b;
        }
      }
      __syncthreads();
      bp = S.bprev[tid];
      bi = S.bcur[tid];
    }
    const bool proc = inb && bp < nbins;
    {
      const uint64_t pm = __ballot(proc), cmw = __ballot(proc && bi > bp);
      if (lane == 0) {
        S.wv[wid] = __popcll(pm);
        S.cm[wid] = cmw;
      }
    }
    __syncthreads();
    int cnt = 0;
    for (int w = 0; w < NW; w++) cnt += S.wv[w];
    if (proc) {
      S.ta[0][tid] = Xc + dX;
      S.ta[1][tid] = Yc + dY;
      S.ta[2][tid] = Zc + dZ;
      S.ta[3][tid] = dX * dX;
      S.ta[4][tid] = dY * dY;
      S.ta[5][tid] = dZ * dZ;
      S.ta[6][tid] = dX * dY;
      S.ta[7][tid] = dX * dZ;
      S.ta[8][tid] = dY * dZ;
      S.ta[9][tid] = vxp;
      S.ta[10][tid] = vyp;
      S.ta[11][tid] = vzp;
      S.bprev[tid] = bp;
      S.bcur[tid] = bi;
      for (int b = bp; b < bi; b++) S.bclose[b] = j;
      rad_m[j] = d;
    }
    __syncthreads();
    /* bulk velocity */
    chain_pass<false, NW>(rowV, cnt, S, accV, dummy, 0.0, -1, -1);

    /* angular momentum and kinetic energy terms about the running bulk
     * velocity; the potential integral's term (ahf_halos.c:4333-4395) */
    const double M = (double)(j + 1);
    double dVX = vxp - S.ta[9][tid] / M, dVY = vyp - S.ta[10][tid] / M, dVZ = vzp - S.ta[11][tid] / M;
    const double lx = dY * dVZ - dZ * dVY, ly = dZ * dVX - dX * dVZ, lz = dX * dVY - dY * dVX;
    dVX += hub * dX * rf / vf;
    dVY += hub * dY * rf / vf;
    dVZ += hub * dZ * rf / vf;
    const double T = dVX * dVX + dVY * dVY + dVZ * dVZ;
    /* I_now: M/d^2 of the last member so far with d > MACHINE_ZERO */
    const bool fl = proc && d > PRF_MZERO;
    S.i_s[tid] = M / (d * d);
    {
      const uint64_t fmw = __ballot(fl);
      if (lane == 0) S.fm[wid] = fmw;
    }
    __syncthreads();
    int64_t src = -1;
    {
      const uint64_t le = S.fm[wid] & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
      if (le) src = 64 * wid + 63 - __clzll((long long)le);
      for (int w = wid - 1; w >= 0 && src < 0; w--)
        if (S.fm[w]) src = 64 * w + 63 - __clzll((long long)S.fm[w]);
    }
    const double I_now = src >= 0 ? S.i_s[src] : I_car;
    __syncthreads();
    S.i_s[tid] = I_now;
    __syncthreads();
    const double I_prev = tid ? S.i_s[tid - 1] : I_car, pre = tid ? S.d_s[tid - 1] : d_car;
    if (proc) {
      S.tc[0][tid] = lx;
      S.tc[1][tid] = ly;
      S.tc[2][tid] = lz;
      S.tc[3][tid] = T;
      S.tc[4][tid] = fl ? ((I_now + I_prev) / 2.) * (d - pre) : -0.0; /* -0.0: an exact no-op */
      vc2[j] = M / d;
      const double dV = FourPiThird * ((d * d * d) - (pre * pre * pre));
      const double hm = (d + pre) / 2.;
      dr2[j] = 1.0 / dV * (hm * hm);
    }
    I_car = S.i_s[cnt - 1];
    d_car = S.d_s[cnt - 1];
    __syncthreads();
    /* angular momentum, kinetic energy, potential and potential energy,
     * centre of mass, inertia tensor */
    chain_pass<true, NW>(rowC, cnt, S, accC, accE, Phi0, colC, ecolC);

    const double U = S.tc[4][tid] - Phi0;
    const double Ep = 0.5 * T + U;
    if (proc)
      for (int b = bp; b < bi; b++) S.bt[b * PQ_N + BT_VESC] = 2 * fabs(U);
    /* most bound member: the first at the lowest energy below the one so far */
    {
      const bool cand = proc && Ep < Emin;
      const uint64_t cw = __ballot(cand);
      if (lane == 0) S.hm[wid] = cw;
      __syncthreads();
      bool some = false;
      for (int w = 0; w < NW; w++) some |= S.hm[w] != 0;
      if (some) {
        double bv = cand ? Ep : 0.0;
        int64_t bl = cand ? tid : -1;
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) {
          const double ov = __shfl_xor(bv, dd, 64);
          const int64_t ol = __shfl_xor(bl, dd, 64);
          if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
            bv = ov;
            bl = ol;
          }
        }
        if (lane == 0) {
          S.rv[wid] = bv;
          S.ri[wid] = bl;
        }
        __syncthreads();
        bv = S.rv[0];
        bl = S.ri[0];
        for (int w = 1; w < NW; w++) {
          const double ov = S.rv[w];
          const int64_t ol = S.ri[w];
          if (ol >= 0 && (bl < 0 || ov < bv || (!(bv < ov) && ol < bl))) {
            bv = ov;
            bl = ol;
          }
        }
        Emin = bv;
        mb = j0 + bl;
      }
    }
    bcar = S.bcur[cnt - 1];
    __syncthreads();
    P += cnt;
    if (cnt < C) break;
  }
  if (tid < 3) S.fin[9 + tid] = accV;
  }
  /* members the walk did not reach read as AHF's zeroed arrays */
  if (!TP)
    for (int64_t j = P + tid; j < n; j += C) rad_m[j] = vc2[j] = dr2[j] = 0.0;
  if (!TP && tid >= 5 && tid < 8) S.fin[tid - 5] = accC;
  __syncthreads();

  /* the bins */
  double *pr = prof + pofs[h] * PQ_N;
  if (tid < nbins) {
    const int b = tid;
    const double M = (double)(S.bclose[b] + 1), Mp = b ? (double)(S.bclose[b - 1] + 1) : 0.0;
    const double r = S.rad[b], rp = b ? S.rad[b - 1] : 0.0;
    const double Vol = FourPiThird * (r * r * r), Vp = b ? FourPiThird * (rp * rp * rp) : 0.0;
    const double dM = M - Mp, dV = Vol - Vp;
    double *q = pr + b * PQ_N;
    /* the bin's record, read before the row is written over it */
    double bt[BT_N];
    for (int k = 0; k < BT_N; k++) bt[k] = q[k];
    double ax[3] = {1.0, 0.0, 0.0}, e[3][3] = {};
    if (S.bclose[b] + 1 > PRF_MINSHELL) prf_axes(bt, ax, e);
    q[PQ_R] = r;
    q[PQ_NPART] = M;
    q[PQ_NVPART] = M;
    q[PQ_OVDENS] = M / Vol;
    q[PQ_DENS] = dV > 0 ? dM / dV : 0.0;
    q[PQ_V2CIRC] = M / r;
    q[PQ_VESC2] = bt[BT_VESC];
    q[PQ_SIGV] = sqrt(bt[BT_T] / M);
    q[PQ_EKIN] = 0.5 * bt[BT_T];
    q[PQ_EPOT] = 0.5 * bt[BT_EP];
    q[PQ_LX] = bt[BT_LX];
    q[PQ_LY] = bt[BT_LY];
    q[PQ_LZ] = bt[BT_LZ];
    for (int k = 0; k < 3; k++)
      for (int r3 = 0; r3 < 3; r3++) q[PQ_E1X + 3 * k + r3] = e[r3][k];
    q[PQ_AX1] = 1.0;
    q[PQ_AX2] = ax[0] > 0. ? sqrt(ax[1] / ax[0]) : 0.0;
    q[PQ_AX3] = ax[0] > 0. ? sqrt(ax[2] / ax[0]) : 0.0;
  }
  __syncthreads();

  /* r2 and R_max: find_max over members NIGNORE.. of rho r^2 (three
   * smoothing passes) and of V_c^2 (one pass) */
  const int64_t m = n - PRF_NIGNORE;
  const double *x = rad_m + PRF_NIGNORE;
  int64_t i_r2, i_rm;
  if (TP) {
    i_r2 = TP->i_r2;
    i_rm = TP->i_rm;
  } else {
    find_peaks<NW>(dr2 + PRF_NIGNORE, vc2 + PRF_NIGNORE, m, S, &i_r2, &i_rm);
  }
  const double r2 = (x[i_r2] + x[i_r2]) / 2;
  const double R_max = (x[i_rm] + x[i_rm]) / 2;
  /* the first member at or beyond R_max (ahf_halos.c:4892-4894) */
  int64_t ib = TP ? TP->ib : n - 1;
  for (int64_t j0 = 0; !TP && j0 < n - 1; j0 += C) {
    const int64_t j = j0 + tid;
    const uint64_t mk = __ballot(j < n - 1 && !(rad_m[j] < R_max));
    if (lane == 0) S.hm[wid] = mk;
    __syncthreads();
    int64_t found = -1;
    for (int w = NW - 1; w >= 0; w--)
      if (S.hm[w]) found = j0 + 64 * w + __ffsll((unsigned long long)S.hm[w]) - 1;
    __syncthreads();
    if (found >= 0) {
      ib = found;
      break;
    }
  }
  if (tid != 0) return;

  const double rib = rad_m[ib];
  const double ovd = ib < P ? (double)(ib + 1) / (FourPiThird * (rib * rib * rib)) : 0.0;
  const double M_max = ovd * FourPiThird * (rib * rib * rib);
  const double V_max = M_max / R_max;
  const double Ms = (double)P;
  const double *last = pr + (nbins - 1) * PQ_N, *prev = pr + (nbins - 2) * PQ_N;
  double *Pp = props + h * PP_N;
  const double Lx = last[PQ_LX], Ly = last[PQ_LY], Lz = last[PQ_LZ];
  const double absL = sqrt(Lx * Lx + Ly * Ly + Lz * Lz);
  Pp[PP_NBINS] = (double)nbins;
  Pp[PP_MVIR] = Ms;
  Pp[PP_VX] = S.fin[9] / Ms;
  Pp[PP_VY] = S.fin[10] / Ms;
  Pp[PP_VZ] = S.fin[11] / Ms;
  Pp[PP_RMAX] = R_max;
  Pp[PP_V2MAX] = V_max;
  Pp[PP_R2] = r2;
  Pp[PP_SIGV] = last[PQ_SIGV];
  Pp[PP_VESC2] = last[PQ_VESC2];
  Pp[PP_EKIN] = last[PQ_EKIN];
  Pp[PP_EPOT] = last[PQ_EPOT];
  Pp[PP_AX1] = last[PQ_AX1];
  Pp[PP_AX2] = last[PQ_AX2];
  Pp[PP_AX3] = last[PQ_AX3];
  for (int k = 0; k < 9; k++) Pp[PP_E1X + k] = last[PQ_E1X + k];
  if (absL > 0) {
    Pp[PP_LX] = Lx / absL;
    Pp[PP_LY] = Ly / absL;
    Pp[PP_LZ] = Lz / absL;
    double lam = absL / Ms / sqrt(2. * Ms * R_vir);
    lam *= K.v_fac * sqrt(K.r_fac / (PRF_GRAV * K.m_fac));
    Pp[PP_LAMBDA] = lam;
    /* calc_lambdaE (ahf_halos.c:3936-3951) */
    double t1 = sqrt(K.m_fac * Ms);
    t1 = t1 * t1 * t1;
    double t2 = Pp[PP_EKIN] * K.m_fac * (K.v_fac * K.v_fac);
    double t3 = Pp[PP_EPOT] * K.m_fac * K.phi_fac;
    t2 = sqrt(fabs(t2 + t3));
    t1 = t2 / t1;
    t2 = K.m_fac * K.r_fac * K.v_fac * absL;
    t2 = t2 / (K.m_fac * Ms);
    Pp[PP_LAMBDAE] = t1 * t2 / PRF_GRAV;
  } else {
    Pp[PP_LX] = Pp[PP_LY] = Pp[PP_LZ] = Pp[PP_LAMBDA] = Pp[PP_LAMBDAE] = 0.0;
  }
  Pp[PP_CNFW] = prf_cnfw(V_max, Ms / R_vir);
  /* calc_R1 (specific.c:1977-1988) */
  {
    const double r0 = pr[PQ_R];
    double R1 = ((r0 / 2.0) * (r0 / 2.0) * (r0 / 2.0)) * pr[PQ_DENS] * r0;
    for (int b = 1; b < nbins; b++) {
      const double rb = pr[b * PQ_N + PQ_R], rq = pr[(b - 1) * PQ_N + PQ_R];
      const double rmid = (rb + rq) / 2.0, drr = rb - rq;
      R1 += (rmid * rmid * rmid) * pr[b * PQ_N + PQ_DENS] * drr;
    }
    Pp[PP_R1] = 4 * PRF_PI * R1 / Ms / R_vir;
  }
  Pp[PP_CR1] = prf_cr1(Pp[PP_R1]);
  {
    const double Ts = 2.0 * (last[PQ_EKIN] - prev[PQ_EKIN]);
    const double frad = fabs(prev[PQ_R] / last[PQ_R]);
    Pp[PP_SURFP] = -0.125 * ((1. + frad) * (1. + frad) * (1. + frad)) / (1. - (frad * frad * frad)) * Ts;
  }
  Pp[PP_FMHIRES] = Ms > 0 ? Ms / (Ms + 0.0) : 0.0;
  if (mb >= 0) {
    const uint32_t q = ip[mb];
    double dX = fabs(px[q] - Xc), dY = fabs(py[q] - Yc), dZ = fabs(pz[q] - Zc);
    if (dX > 0.5) dX -= 1.0;
    if (dY > 0.5) dY -= 1.0;
    if (dZ > 0.5) dZ -= 1.0;
    Pp[PP_MBP_OFF] = sqrt(dX * dX + dY * dY + dZ * dZ);
    Pp[PP_MBP_X] = px[q];
    Pp[PP_MBP_Y] = py[q];
    Pp[PP_MBP_Z] = pz[q];
    Pp[PP_MBP_VX] = vx[q];
    Pp[PP_MBP_VY] = vy[q];
    Pp[PP_MBP_VZ] = vz[q];
  } else {
    Pp[PP_MBP_OFF] = -1.0;
    Pp[PP_MBP_X] = Pp[PP_MBP_Y] = Pp[PP_MBP_Z] = Pp[PP_MBP_VX] = Pp[PP_MBP_VY] = Pp[PP_MBP_VZ] = 0.0;
  }
  {
    const double cx = fmod(S.fin[0] / Ms + 1., 1.), cy = fmod(S.fin[1] / Ms + 1., 1.),
                 cz = fmod(S.fin[2] / Ms + 1., 1.);
    double dX = fabs(cx - Xc), dY = fabs(cy - Yc), dZ = fabs(cz - Zc);
    if (dX > 0.5) dX -= 1.0;
    if (dY > 0.5) dY -= 1.0;
    if (dZ > 0.5) dZ -= 1.0;
    Pp[PP_COM_OFF] = sqrt(dX * dX + dY * dY + dZ * dZ);
    Pp[PP_COM_X] = cx;
    Pp[PP_COM_Y] = cy;
    Pp[PP_COM_Z] = cz;
  }
}

/* member count of every halo after the construction steps (or as given) */
__global__ void k_prof_count(const int64_t *__restrict__ off, const HaloOut *__restrict__ out, int64_t nh,
                             int64_t *__restrict__ n)
{
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  n[h] = out ? out[h].npart : off[h + 1] - off[h];
}

/* --------------------------------------------------------------- host */

struct ProfResult {
  int64_t nh = 0, nbins_tot = 0;
  double *props = nullptr, *prof = nullptr;
};

extern "C" int ahfgpu_halo_profiles(ahfgpu_ctx *c, const ahfgpu_profile_params *p, const double *phi0_rvir,
                                    int32_t *nbins_out)
{
  if (!c || !p) return AHFGPU_EINVAL;
  HaloSet *s = (HaloSet *)c->halo_set;
  if (!s) return c->fail(AHFGPU_EINVAL, "no halos resident (call ahfgpu_set_halos first)");
  if (!phi0_rvir && !s->done) return c->fail(AHFGPU_EINVAL, "no construction result (run ahfgpu_unbind or pass Phi0/R_vir)");
  /* find_max reads members NIGNORE.. of each list, so a halo needs more */
  if (p->minpart <= PRF_NIGNORE) return c->fail(AHFGPU_EINVAL, "NminPerHalo must exceed %d", PRF_NIGNORE);
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  const int64_t nh = s->nh, ne = s->ne;
  const HaloOut *out = phi0_rvir ? nullptr : s->out;
  Arena &A = c->parena;
  A.reset();
  int64_t *d_n = A.get<int64_t>(nh);
  int32_t *d_nb = A.get<int32_t>(nh);
  int64_t *d_pofs = A.get<int64_t>(nh);
  double *d_pr = phi0_rvir ? A.get<double>(2 * nh) : nullptr;
  double *sc = A.get<double>(3 * ne);
  if (!d_n || !d_nb || !d_pofs || !sc || (phi0_rvir && !d_pr))
    return c->fail(AHFGPU_ENOMEM, "device arena exhausted (profiles)");
  std::vector<int64_t> hn((size_t)nh), pofs((size_t)nh);
  std::vector<int32_t> nb((size_t)nh);
  if (nh) {
    hipLaunchKernelGGL(k_prof_count, dim3((unsigned)((nh + AHF_BLOCK - 1) / AHF_BLOCK)), dim3(AHF_BLOCK), 0, c->st,
                       s->off, out, nh, d_n);
    hipMemcpyAsync(hn.data(), d_n, (size_t)nh * 8, hipMemcpyDeviceToHost, c->st);
    if (hipStreamSynchronize(c->st) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "profiles: %s", hipGetErrorString(hipGetLastError()));
  }
  /* binning_parameter's bin count (specific.c:266-270), on the host's libm
   * as AHF computes it */
  int64_t tot = 0;
  for (int64_t h = 0; h < nh; h++) {
    const int64_t n = hn[(size_t)h];
    int k = 0;
    if (n >= p->minpart) {
      k = (int)(6.2 * (log10((double)n)) - 3.5);
      k = k > 2 ? k : 2;
    }
    nb[(size_t)h] = k;
    pofs[(size_t)h] = tot;
    tot += k;
  }
  /* haloes of big members or more go to workgroups of AHF_PROF_BIG_NW
   * wavefronts on the second stream, largest first; the rest to one
   * wavefront each */
  const int64_t big = c->prof_split == 0 ? AHF_PROF_BIG : c->prof_split < 0 ? INT64_MAX : c->prof_split;
  std::vector<int32_t> bigl;
  for (int64_t h = 0; h < nh; h++)
    if (nb[(size_t)h] > 0 && hn[(size_t)h] >= big) bigl.push_back((int32_t)h);
  std::stable_sort(bigl.begin(), bigl.end(), [&](int32_t x, int32_t y) { return hn[(size_t)x] > hn[(size_t)y]; });
  const int64_t nbig = (int64_t)bigl.size();
  int32_t *d_big = nbig ? A.get<int32_t>(nbig) : nullptr;
  /* of those, haloes of prof_tiled members or more are walked over the
   * device (k_prf_*); tiles of PRF_TILE members, rows end to end */
  const int64_t tiled = c->prof_tiled == 0 ? AHF_PROF_TILED : c->prof_tiled < 0 ? INT64_MAX : c->prof_tiled;
  std::vector<PrfTiled> htl;
  std::vector<PrfTile> htile;
  std::vector<int32_t> tmap_h;
  int64_t Nt = 0;
  for (int32_t h : bigl) {
    const int64_t n = hn[(size_t)h];
    if (n < tiled) continue;
    if (tmap_h.empty()) tmap_h.assign((size_t)nh, -1);
    PrfTiled T;
    memset(&T, 0, sizeof(T));
    T.h = h;
    T.n = n;
    T.to = Nt;
    T.t0 = (int32_t)htile.size();
    for (int64_t j0 = 0; j0 < n; j0 += PRF_TILE) htile.push_back(PrfTile{(int32_t)htl.size(), 0, j0});
    T.t1 = (int32_t)htile.size();
    tmap_h[(size_t)h] = (int32_t)htl.size();
    htl.push_back(T);
    Nt += n + (n & 1); /* rows at even offsets: member pairs are 16-byte aligned */
  }
  const int64_t ntl = (int64_t)htl.size(), ntile = (int64_t)htile.size();
  PrfTiled *d_tl = nullptr;
  PrfTile *d_tile = nullptr;
  int32_t *d_tmap = nullptr, *tcmax = nullptr;
  int64_t *tflmax = nullptr, *tmb = nullptr;
  double *tEmin = nullptr, *tpv = nullptr;
  int64_t *tpi = nullptr, *tib = nullptr;
  PrfChain *d_chains = nullptr;
  PrfRows R;
  memset(&R, 0, sizeof(R));
  if (ntl) {
    d_tl = A.get<PrfTiled>(ntl);
    d_tile = A.get<PrfTile>(ntile);
    d_tmap = A.get<int32_t>(nh);
    tcmax = A.get<int32_t>(ntile);
    tflmax = A.get<int64_t>(ntile);
    tmb = A.get<int64_t>(ntile);
    tEmin = A.get<double>(ntile);
    tpv = A.get<double>(2 * ntile);
    tpi = A.get<int64_t>(2 * ntile);
    tib = A.get<int64_t>(ntile);
    d_chains = A.get<PrfChain>(18);
    double *rows = A.get<double>(19 * Nt);
    int32_t *irows = A.get<int32_t>(3 * Nt);
    if (!d_tl || !d_tile || !d_tmap || !tcmax || !tflmax || !tmb || !tEmin || !tpv || !tpi || !tib || !d_chains ||
        !rows || !irows)
      return c->fail(AHFGPU_ENOMEM, "device arena exhausted (profiles over the device)");
    double *w = rows;
    auto take = [&]() { double *r = w; w += Nt; return r; };
    R.dd = take(); R.gx = take(); R.gy = take(); R.gz = take();
    for (int k = 0; k < 3; k++) { R.v[k] = take(); R.com[k] = take(); }
    R.tt = take(); R.phi = take(); R.phir = take();
    for (int k = 0; k < 6; k++) R.ine[k] = take();
    R.c = irows; R.bp = irows + Nt; R.bi = irows + 2 * Nt;
  }
  ProfResult *r = (ProfResult *)c->prof_result;
  if (!r) c->prof_result = r = new ProfResult();
  r->nh = nh;
  r->nbins_tot = tot;
  r->props = A.get<double>(nh * PP_N);
  r->prof = A.get<double>(tot * PQ_N);
  if (!r->props || !r->prof || (nbig && !d_big)) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (profiles)");
  if (nh) {
    hipMemcpyAsync(d_nb, nb.data(), (size_t)nh * 4, hipMemcpyHostToDevice, c->st);
    hipMemcpyAsync(d_pofs, pofs.data(), (size_t)nh * 8, hipMemcpyHostToDevice, c->st);
    if (nbig) hipMemcpyAsync(d_big, bigl.data(), (size_t)nbig * 4, hipMemcpyHostToDevice, c->st);
    if (ntl) {
      /* the running sums: bulk velocity (3) and potential (to its row);
       * then angular momentum (3) and kinetic energy (their terms made in
       * the walk), centre of mass (3), inertia tensor (6) and potential
       * energy (over the potential's running values less Phi0) */
      std::vector<PrfChain> ch;
      for (int k = 0; k < 3; k++) ch.push_back(PrfChain{R.v[k], nullptr, -1, 9 + k, 0, 0});
      ch.push_back(PrfChain{R.phi, R.phir, -1, -1, 0, 0});
      for (int k = 0; k < 3; k++) ch.push_back(PrfChain{nullptr, nullptr, BT_LX + k, -1, 0, 0});
      ch.push_back(PrfChain{nullptr, nullptr, BT_T, -1, 0, 0});
      for (int k = 0; k < 3; k++) ch.push_back(PrfChain{R.com[k], nullptr, -1, k, 0, 0});
      for (int k = 0; k < 6; k++) ch.push_back(PrfChain{R.ine[k], nullptr, BT_A11 + k, -1, 0, 0});
      ch.push_back(PrfChain{R.phir, nullptr, BT_EP, -1, 1, 0});
      hipMemcpyAsync(d_chains, ch.data(), ch.size() * sizeof(PrfChain), hipMemcpyHostToDevice, c->st);
      hipMemcpyAsync(d_tl, htl.data(), (size_t)ntl * sizeof(PrfTiled), hipMemcpyHostToDevice, c->st);
      hipMemcpyAsync(d_tile, htile.data(), (size_t)ntile * sizeof(PrfTile), hipMemcpyHostToDevice, c->st);
      hipMemcpyAsync(d_tmap, tmap_h.data(), (size_t)nh * 4, hipMemcpyHostToDevice, c->st);
      /* the staging buffers are read before the host vectors go */
      if (hipStreamSynchronize(c->st) != hipSuccess)
        return c->fail(AHFGPU_EHIP, "profiles: %s", hipGetErrorString(hipGetLastError()));
    }
    if (phi0_rvir) hipMemcpyAsync(d_pr, phi0_rvir, (size_t)nh * 16, hipMemcpyHostToDevice, c->st);
    hipMemsetAsync(r->props, 0, (size_t)nh * PP_N * 8, c->st);
    ProfConst K{p->r_fac, p->v_fac, p->m_fac, p->phi_fac, p->hubble, p->minpart};
    hipEvent_t e0 = c->timer_begin();
    if (nbig) {
      hipEventRecord(c->ev_s0, c->st);
      hipStreamWaitEvent(c->st2, c->ev_s0, 0);
      if (ntl) {
        hipStream_t s2 = c->st2;
        const dim3 gt((unsigned)ntile);
        hipLaunchKernelGGL(k_prf_bin, dim3((unsigned)ntl), dim3(64), 0, s2, K, d_tl, s->off, out, d_pr, s->ip_in,
                           s->buf_a, s->buf_b, s->cen, d_nb, c->px, c->py, c->pz);
        hipLaunchKernelGGL(k_prf_geom, gt, dim3(256), 0, s2, d_tl, d_tile, R, c->pv, tcmax, tflmax);
        hipLaunchKernelGGL(k_prf_tscan, dim3((unsigned)ntl), dim3(64), 0, s2, d_tl, tcmax, tflmax);
        hipLaunchKernelGGL(k_prf_terms1, gt, dim3(256), 0, s2, d_tl, d_tile, R, tcmax, tflmax, sc, sc + ne,
                           sc + 2 * ne);
        /* r2, R_max and the first member beyond it on the third stream,
         * beside the running sums */
        hipEventRecord(c->ev_s2, s2);
        hipStreamWaitEvent(c->st3, c->ev_s2, 0);
        hipLaunchKernelGGL(k_prf_peaks, gt, dim3(256), 0, c->st3, d_tl, d_tile, sc + ne, sc + 2 * ne, tpv, tpi);
        hipLaunchKernelGGL(k_prf_peaks_done, dim3((unsigned)ntl), dim3(64), 0, c->st3, d_tl, sc, tpv, tpi);
        hipLaunchKernelGGL(k_prf_rmax, gt, dim3(256), 0, c->st3, d_tl, d_tile, sc, tib);
        hipLaunchKernelGGL(k_prf_rmax_done, dim3((unsigned)ntl), dim3(64), 0, c->st3, d_tl, tib);
        hipEventRecord(c->ev_s3, c->st3);
        hipLaunchKernelGGL(k_prf_walk, dim3((unsigned)ntl), dim3(64 * PRF_FW), 0, s2, d_tl, d_chains, R, K);
        hipLaunchKernelGGL(k_prf_finish, gt, dim3(256), 0, s2, d_tl, d_tile, R, tEmin, tmb);
        hipLaunchKernelGGL(k_prf_done, dim3((unsigned)ntl), dim3(64), 0, s2, d_tl, tEmin, tmb);
        hipStreamWaitEvent(s2, c->ev_s3, 0);
      }
      /* the haloes walked over the device (the first ntl of the list) after
       * their walk; the others on the first stream meanwhile */
      auto big_launch = [&](hipStream_t st, int64_t i0, int64_t cnt) {
        hipLaunchKernelGGL(k_profiles<AHF_PROF_BIG_NW>, dim3((unsigned)cnt), dim3(64 * AHF_PROF_BIG_NW), 0, st, K,
                           s->off, d_big + i0, big, out, d_pr, s->ip_in, s->buf_a, s->buf_b, s->cen, d_nb, d_pofs,
                           c->px, c->py, c->pz, c->vx, c->vy, c->vz, sc, sc + ne, sc + 2 * ne, r->props, r->prof,
                           d_tl, d_tmap, c->pv);
      };
      if (ntl) big_launch(c->st2, 0, ntl);
      if (nbig > ntl) big_launch(ntl ? c->st : c->st2, ntl, nbig - ntl);
      hipEventRecord(c->ev_s1, c->st2);
    }
    hipLaunchKernelGGL(k_profiles<1>, dim3((unsigned)nh), dim3(64), 0, c->st, K, s->off, s->order, big, out, d_pr,
                       s->ip_in, s->buf_a, s->buf_b, s->cen, d_nb, d_pofs, c->px, c->py, c->pz, c->vx, c->vy, c->vz,
                       sc, sc + ne, sc + 2 * ne, r->props, r->prof, (const PrfTiled *)nullptr,
                       (const int32_t *)nullptr, c->pv);
    if (nbig) hipStreamWaitEvent(c->st, c->ev_s1, 0);
    c->timer_end(e0, TIMER_PROFILES, 0.0);
  }
  c->timers_reset(TIMER_PROFILES, TIMER_PROFILES);
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "profiles: %s", hipGetErrorString(hipGetLastError()));
  c->timers_collect();
  if (nbins_out && nh) memcpy(nbins_out, nb.data(), (size_t)nh * 4);
  return 0;
}

extern "C" int ahfgpu_profile_results(ahfgpu_ctx *c, double *props, double *prof)
{
  if (!c) return AHFGPU_EINVAL;
  ProfResult *r = (ProfResult *)c->prof_result;
  if (!r) return c->fail(AHFGPU_EINVAL, "no profile result (call ahfgpu_halo_profiles first)");
  if (props && r->nh) hipMemcpy(props, r->props, (size_t)r->nh * PP_N * 8, hipMemcpyDeviceToHost);
  if (prof && r->nbins_tot) hipMemcpy(prof, r->prof, (size_t)r->nbins_tot * PQ_N * 8, hipMemcpyDeviceToHost);
  if (hipDeviceSynchronize() != hipSuccess) return c->fail(AHFGPU_EHIP, "profile results");
  return 0;
}

void ahfgpu_free_profiles(ahfgpu_ctx *c)
{
  delete (ProfResult *)c->prof_result;
  c->prof_result = nullptr;
  c->parena.release();
}

extern "C" int ahfgpu_set_profile_split(ahfgpu_ctx *c, int64_t min_members)
{
  if (!c) return AHFGPU_EINVAL;
  c->prof_split = min_members;
  return 0;
}
