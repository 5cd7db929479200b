/*
 * oracle_profiles.h -- test infrastructure only: the CPU oracle's
 * restatement of AHF's HaloProfiles (src/libahf/ahf_halos.c:3961-5060)
 * for the default build (dark matter only, equal masses, the R_max / r2
 * search over the sorted particle list of AHFparticle_Rmax_r2, Hubble drag
 * in the kinetic energy), restated per halo as one sequential walk over its
 * members in their (radially sorted) order: every sum is AHF's running sum
 * in AHF's order, so the results follow its roundings.  The helpers restate
 * binning_parameter (src/libutility/specific.c:259-320), smooth3 and find_max
 * (src/libutility/general.c:548-665), the Jacobi eigen solver and get_axes
 * (general.c:1163-1240, specific.c:135-180), calc_lambdaE
 * (ahf_halos.c:3936-3951), calc_R1, calc_cR1 and calc_cNFW
 * (specific.c:1968-2060).
 */
#ifndef ORACLE_PROFILES_H
#define ORACLE_PROFILES_H

#include <math.h>
#include <stdint.h>

#define AHF_PHD static inline

#define AHF_PROF_GRAV 4.3006485e-9 /* [Mpc km^2]/[Msun s^2], src/param.h:45 */
#define AHF_PROF_PI 3.14159265358979323846264338
#define AHF_PROF_ZERO 1e-6          /* FP32 build, src/param.h:124 */
#define AHF_PROF_MZERO 5e-16        /* MACHINE_ZERO, src/param.h:125 */
#define AHF_PROF_MINSHELL 10        /* AHF_MINPART_SHELL, src/param.h:16 */
#define AHF_PROF_NIGNORE 5          /* AHF_Rmax_r2_NIGNORE, src/param.h:25 */
#define AHF_PROF_MAXBINS 64         /* 6.2*log10(2^32) - 3.5 < 64 */

/* per-bin profile quantities (halo->prof.*), AHF_PROF_NQ per bin */
enum {
  PQ_R, PQ_NPART, PQ_NVPART, PQ_OVDENS, PQ_DENS, PQ_V2CIRC, PQ_VESC2, PQ_SIGV, PQ_EKIN, PQ_EPOT, PQ_LX, PQ_LY, PQ_LZ,
  PQ_E1X, PQ_E1Y, PQ_E1Z, PQ_E2X, PQ_E2Y, PQ_E2Z, PQ_E3X, PQ_E3Y, PQ_E3Z, PQ_AX1, PQ_AX2, PQ_AX3, AHF_PROF_NQ
};

/* integral properties, AHF_PROF_NP per halo (the HALO fields HaloProfiles
 * sets; PP_NBINS is an integer stored as a double) */
enum {
  PP_NBINS, PP_MVIR, PP_VX, PP_VY, PP_VZ, PP_RMAX, PP_V2MAX, PP_R2, PP_SIGV, PP_VESC2, PP_EKIN, PP_EPOT,
  PP_AX1, PP_AX2, PP_AX3, PP_E1X, PP_E1Y, PP_E1Z, PP_E2X, PP_E2Y, PP_E2Z, PP_E3X, PP_E3Y, PP_E3Z,
  PP_LX, PP_LY, PP_LZ, PP_LAMBDA, PP_LAMBDAE, PP_CNFW, PP_R1, PP_CR1, PP_SURFP, PP_FMHIRES,
  PP_MBP_OFF, PP_MBP_X, PP_MBP_Y, PP_MBP_Z, PP_MBP_VX, PP_MBP_VY, PP_MBP_VZ,
  PP_COM_OFF, PP_COM_X, PP_COM_Y, PP_COM_Z, AHF_PROF_NP
};

typedef struct {
  int64_t minpart;
  double r_fac, v_fac, m_fac, phi_fac, hubble;
} ahf_prof_consts;

/* radial bin count of a halo of n members (binning_parameter) */
AHF_PHD int ahf_prof_nbins(int64_t n)
{
  int nb = (int)(6.2 * (log10((double)n)) - 3.5);
  return nb > 2 ? nb : 2;
}

/* three smoothing passes of nsmooth (smooth3): ends average two values */
AHF_PHD void ahf_prof_smooth3(double *y, double *tmp, int64_t n, int nsmooth)
{
  if (nsmooth == 0 || n < 3) return;
  for (int j = 0; j < nsmooth; j++) {
    tmp[0] = (y[0] + y[1]) / 2.;
    for (int64_t i = 1; i < n - 1; i++) tmp[i] = (y[i - 1] + y[i] + y[i + 1]) / 3.;
    tmp[n - 1] = (y[n - 1] + y[n - 2]) / 2.;
    for (int64_t i = 0; i < n; i++) y[i] = tmp[i];
  }
}

/* the peak of y over x (find_max): the smoothed maximum from the left, then
 * a search back from halfway to the end; x and y_raw averaged at both */
AHF_PHD void ahf_prof_find_max(const double *x, const double *y_raw, int64_t n, int ismooth, double *y, double *tmp,
                               double *x_max, double *y_max)
{
  for (int64_t i = 0; i < n; i++) y[i] = y_raw[i];
  ahf_prof_smooth3(y, tmp, n, ismooth);
  double ymax = -10.0;
  int64_t i0 = n - 1;
  for (int64_t i = 0; i < n - 1; i++)
    if (y[i] > ymax) {
      ymax = y[i];
      i0 = i;
    }
  int64_t i1 = i0;
  for (int64_t i = (n - 1 + i0) / 2; i > i0; i--)
    if (y[i] > ymax) {
      ymax = y[i];
      i1 = i;
    }
  *x_max = (x[i0] + x[i1]) / 2;
  *y_max = (y_raw[i0] + y_raw[i1]) / 2;
}

/*
 * Eigen decomposition of a symmetric 3x3 matrix by cyclic Jacobi rotations
 * (the published method of Rutishauser as given in Numerical Recipes:
 * a threshold of 0.2*S/n^2 in the first three sweeps, small off-diagonal
 * elements dropped after four, at most 50 sweeps), 1-based like the
 * original so that the rotation sequence, and with it the eigenvector
 * signs, are the same.  d[1..3] eigenvalues, v[1..3][j] eigenvector j.
 */
AHF_PHD void ahf_prof_rotate(double a[4][4], int i, int j, int k, int l, double s, double tau)
{
  const double g = a[i][j], h = a[k][l];
  a[i][j] = g - s * (h + g * tau);
  a[k][l] = h + s * (g - h * tau);
}

AHF_PHD void ahf_prof_jacobi3(double a[4][4], double d[4], double v[4][4])
{
  const int n = 3;
  double b[4], z[4];
  for (int p = 1; p <= n; p++) {
    for (int q = 1; q <= n; q++) v[p][q] = 0.0;
    v[p][p] = 1.0;
  }
  for (int p = 1; p <= n; p++) {
    b[p] = d[p] = a[p][p];
    z[p] = 0.0;
  }
  for (int sweep = 1; sweep <= 50; sweep++) {
    double sm = 0.0;
    for (int p = 1; p <= n - 1; p++)
      for (int q = p + 1; q <= n; q++) sm += fabs(a[p][q]);
    if (sm == 0.0) return;
    const double tresh = sweep < 4 ? 0.2 * sm / (n * n) : 0.0;
    for (int p = 1; p <= n - 1; p++) {
      for (int q = p + 1; q <= n; q++) {
        const double g = 100.0 * fabs(a[p][q]);
        if (sweep > 4 && (double)(fabs(d[p]) + g) == (double)fabs(d[p]) &&
            (double)(fabs(d[q]) + g) == (double)fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if ((double)(fabs(h) + g) == (double)fabs(h)) {
            t = (a[p][q]) / h;
          } else {
            const double theta = 0.5 * h / (a[p][q]);
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
          for (int j = 1; j <= p - 1; j++) ahf_prof_rotate(a, j, p, j, q, s, tau);
          for (int j = p + 1; j <= q - 1; j++) ahf_prof_rotate(a, p, j, j, q, s, tau);
          for (int j = q + 1; j <= n; j++) ahf_prof_rotate(a, p, j, q, j, s, tau);
          for (int j = 1; j <= n; j++) ahf_prof_rotate(v, j, p, j, q, s, tau);
        }
      }
    }
    for (int p = 1; p <= n; p++) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0;
    }
  }
}

/* get_axes: eigenvalues largest first into axis1..3, the matching unit
 * vectors into the columns of it (the order of three values by a stable
 * insertion sort, as NR's indexx does below seven elements) */
AHF_PHD void ahf_prof_get_axes(double it[3][3], double *axis1, double *axis2, double *axis3)
{
  double a[4][4], d[4], v[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) a[i][j] = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i + 1][j + 1] = it[i][j];
  ahf_prof_jacobi3(a, d, v);
  int idx[4] = {0, 1, 2, 3};
  for (int j = 2; j <= 3; j++) {
    const int ix = idx[j];
    const double key = d[ix];
    int i = j - 1;
    for (; i >= 1; i--) {
      if (d[idx[i]] <= key) break;
      idx[i + 1] = idx[i];
    }
    idx[i + 1] = ix;
  }
  *axis1 = d[idx[3]];
  *axis2 = d[idx[2]];
  *axis3 = d[idx[1]];
  for (int r = 0; r < 3; r++) {
    it[r][0] = v[r + 1][idx[3]];
    it[r][1] = v[r + 1][idx[2]];
    it[r][2] = v[r + 1][idx[1]];
  }
}

AHF_PHD double ahf_prof_lambdaE(const ahf_prof_consts *K, double absL, double Mhalo, double Mass, double Ekin,
                                double Epot)
{
  double t1 = sqrt(K->m_fac * Mhalo);
  t1 = t1 * t1 * t1;
  double t2 = Ekin * K->m_fac * (K->v_fac * K->v_fac);
  double t3 = Epot * K->m_fac * K->phi_fac;
  t2 = fabs(t2 + t3);
  t2 = sqrt(t2);
  t1 = t2 / t1;
  t2 = K->m_fac * K->r_fac * K->v_fac * absL;
  t2 = t2 / (K->m_fac * Mass);
  t3 = t1 * t2;
  return t3 / AHF_PROF_GRAV;
}

/* bisection for the NFW concentration of a V2_max/V2_vir ratio (Prada et
 * al. 2012, Eq. 9), -1 outside [1, 5.9] */
AHF_PHD double ahf_prof_cnfw(double V2_max, double V2_vir)
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

AHF_PHD double ahf_prof_r1(const double *prof, int nbins)
{
  const double r0 = prof[PQ_R];
  double R1 = ((r0 / 2.0) * (r0 / 2.0) * (r0 / 2.0)) * prof[PQ_DENS] * (r0);
  for (int ib = 1; ib < nbins; ib++) {
    const double r = prof[ib * AHF_PROF_NQ + PQ_R], rp = prof[(ib - 1) * AHF_PROF_NQ + PQ_R];
    const double rmid = (r + rp) / 2.0, dr = r - rp;
    R1 += (rmid * rmid * rmid) * prof[ib * AHF_PROF_NQ + PQ_DENS] * dr;
  }
  return 4 * AHF_PROF_PI * R1;
}

/* bisection for the concentration of the R1 measure (Wang et al.), -1 when
 * R1 lies outside (0.19, 0.585) */
AHF_PHD double ahf_prof_cr1(double R1)
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

/*
 * HaloProfiles of one halo of n members, in their sorted order.  Member j
 * is particle ip[j], with position (px, py, pz)[ip[j]*ps] and velocity
 * (vx, vy, vz)[ip[j]*ps] (ps 3 and px = pos, py = pos + 1, ... for
 * interleaved arrays).  cen, Phi0 and R_vir are the halo's (set by the
 * construction steps).  scratch holds 6*n doubles.  Writes props[AHF_PROF_NP] and prof[nbins*AHF_PROF_NQ];
 * returns nbins, or 0 (nothing written) below minpart members.
 */
AHF_PHD int ahf_halo_profiles(int64_t n, const uint64_t *ip, const float *px, const float *py, const float *pz,
                              const float *vx, const float *vy, const float *vz, int64_t ps, const double cen[3],
                              double Phi0, double R_vir, const ahf_prof_consts *K, double *scratch, double *props,
                              double *prof)
{
  if (n < K->minpart) return 0;
  const double Xc = cen[0], Yc = cen[1], Zc = cen[2];
  const double FourPiThird = 4. * AHF_PROF_PI / 3.;
  const int nbins = ahf_prof_nbins(n);
  /* binning_parameter: the first member from MINPART/10 on with a distance
   * above MACHINE_ZERO, and the last member */
  int64_t jb = (int64_t)floor(((double)K->minpart / 10.) + 0.5);
  double dist_min = -1.0, dist_max;
  while (jb < n - 1 && dist_min < AHF_PROF_MZERO) {
    const int64_t p = ip[jb] * ps;
    double dX = fabs((double)px[p] - Xc), dY = fabs((double)py[p] - Yc), dZ = fabs((double)pz[p] - Zc);
    if (dX > 0.5) dX = 1.0 - dX;
    if (dY > 0.5) dY = 1.0 - dY;
    if (dZ > 0.5) dZ = 1.0 - dZ;
    dist_min = sqrt(dX * dX + dY * dY + dZ * dZ);
    jb++;
  }
  {
    const int64_t p = ip[n - 1] * ps;
    double dX = fabs((double)px[p] - Xc), dY = fabs((double)py[p] - Yc), dZ = fabs((double)pz[p] - Zc);
    if (dX > 0.5) dX = 1.0 - dX;
    if (dY > 0.5) dY = 1.0 - dY;
    if (dZ > 0.5) dZ = 1.0 - dZ;
    dist_max = sqrt(dX * dX + dY * dY + dZ * dZ);
  }
  if (dist_min < AHF_PROF_MZERO) dist_min = dist_max / 2.;
  const double ldist_min = log10(dist_min), ldist_max = log10(dist_max);
  const double ldr = (ldist_max - ldist_min) / (double)nbins;
  double *radius = scratch, *ovdens = scratch + n, *Vcirc2 = scratch + 2 * n, *dens_r2 = scratch + 3 * n;
  for (int64_t j = 0; j < n; j++) radius[j] = ovdens[j] = Vcirc2[j] = dens_r2[j] = 0.0;
  double Phi = 0.0, pre_dist = 0.0, I_prev = 0.0, I_now = 0.0, M_prev = 0, V_prev = 0.0, rad_prev = dist_min;
  double M_sphere_prev = 0.0, prev_dist = 0.0, M_sphere = 0.0, VXc = 0.0, VYc = 0.0, VZc = 0.0;
  double a11 = 0.0, a22 = 0.0, a33 = 0.0, a12 = 0.0, a13 = 0.0, a23 = 0.0, sig_v = 0.0, Lx = 0.0, Ly = 0.0, Lz = 0.0;
  double CoM[3] = {0.0, 0.0, 0.0}, Epot = 0.0, Ekin = 0.0, Emin = 1e30, M_hires = 0.0, M_lores = 0.0, v_esc2 = 0.0;
  int64_t npart = 0, mb = -1, jpart = 0;
  double cur_dist = -1.0;
  for (int ibin = 0; ibin < nbins; ibin++) {
    const double lcur_rad = ldist_min + ((double)ibin + 1) * ldr;
    double cur_rad = pow(10., lcur_rad);
    if (ibin == nbins - 1) cur_rad = dist_max + AHF_PROF_ZERO;
    while (cur_dist < cur_rad && jpart < n) {
      const int64_t p = ip[jpart] * ps;
      npart++;
      const double weight = 1.0;
      M_hires += weight;
      M_sphere += weight;
      double dX = ((double)px[p] - Xc), dY = ((double)py[p] - Yc), dZ = ((double)pz[p] - Zc);
      if (dX > 0.5) dX -= 1.0;
      if (dY > 0.5) dY -= 1.0;
      if (dZ > 0.5) dZ -= 1.0;
      if (dX < -0.5) dX += 1.0;
      if (dY < -0.5) dY += 1.0;
      if (dZ < -0.5) dZ += 1.0;
      cur_dist = sqrt(dX * dX + dY * dY + dZ * dZ);
      CoM[0] += weight * (Xc + dX);
      CoM[1] += weight * (Yc + dY);
      CoM[2] += weight * (Zc + dZ);
      a11 += weight * dX * dX;
      a22 += weight * dY * dY;
      a33 += weight * dZ * dZ;
      a12 += weight * dX * dY;
      a13 += weight * dX * dZ;
      a23 += weight * dY * dZ;
      VXc += weight * vx[p];
      VYc += weight * vy[p];
      VZc += weight * vz[p];
      const double VXp = (double)vx[p], VYp = (double)vy[p], VZp = (double)vz[p];
      double dVX = (VXp - VXc / M_sphere), dVY = (VYp - VYc / M_sphere), dVZ = (VZp - VZc / M_sphere);
      Lx += weight * (dY * dVZ - dZ * dVY);
      Ly += weight * (dZ * dVX - dX * dVZ);
      Lz += weight * (dX * dVY - dY * dVX);
      dVX += K->hubble * dX * K->r_fac / K->v_fac;
      dVY += K->hubble * dY * K->r_fac / K->v_fac;
      dVZ += K->hubble * dZ * K->r_fac / K->v_fac;
      const double Tpart = weight * (dVX * dVX + dVY * dVY + dVZ * dVZ);
      double Upart;
      if (cur_dist > AHF_PROF_MZERO) {
        I_now = M_sphere / (cur_dist * cur_dist);
        const double I_mid = (I_now + I_prev) / 2.;
        Phi += I_mid * (cur_dist - pre_dist);
        Upart = (Phi - Phi0) * weight;
      } else {
        Upart = (Phi - Phi0) * weight;
      }
      v_esc2 = 2 * fabs(Upart) / weight;
      I_prev = I_now;
      pre_dist = cur_dist;
      Epot += Upart;
      sig_v += Tpart;
      Ekin += Tpart;
      const double Epart = (0.5 * Tpart + Upart);
      if (Epart < Emin) {
        Emin = Epart;
        mb = jpart;
      }
      radius[jpart] = cur_dist;
      ovdens[jpart] = M_sphere / (FourPiThird * (cur_dist * cur_dist * cur_dist));
      const double dV = FourPiThird * ((cur_dist * cur_dist * cur_dist) - (prev_dist * prev_dist * prev_dist));
      const double dM = M_sphere - M_sphere_prev;
      Vcirc2[jpart] = M_sphere / cur_dist;
      M_sphere_prev = M_sphere;
      dens_r2[jpart] = dM / dV * (((cur_dist + prev_dist) / 2.) * ((cur_dist + prev_dist) / 2.));
      prev_dist = cur_dist;
      jpart++;
    }
    const double Volume = FourPiThird * (cur_rad * cur_rad * cur_rad);
    const double dM = M_sphere - M_prev, dV = Volume - V_prev;
    double it[3][3], axis1, axis2, axis3;
    if (npart > AHF_PROF_MINSHELL) {
      it[0][0] = a11;
      it[1][1] = a22;
      it[2][2] = a33;
      it[0][1] = it[1][0] = a12;
      it[0][2] = it[2][0] = a13;
      it[1][2] = it[2][1] = a23;
      ahf_prof_get_axes(it, &axis1, &axis2, &axis3);
    } else {
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) it[i][k] = 0;
      axis1 = 1;
      axis2 = 0;
      axis3 = 0;
    }
    double *q = prof + ibin * AHF_PROF_NQ;
    q[PQ_R] = cur_rad;
    q[PQ_NPART] = (double)npart;
    q[PQ_NVPART] = M_sphere;
    q[PQ_OVDENS] = M_sphere / Volume;
    q[PQ_DENS] = dV > 0 ? dM / dV : 0.0;
    q[PQ_V2CIRC] = M_sphere / cur_rad;
    q[PQ_VESC2] = v_esc2;
    q[PQ_SIGV] = sqrt(sig_v / M_sphere);
    q[PQ_EKIN] = 0.5 * Ekin;
    q[PQ_EPOT] = 0.5 * Epot;
    q[PQ_LX] = Lx;
    q[PQ_LY] = Ly;
    q[PQ_LZ] = Lz;
    q[PQ_E1X] = it[0][0];
    q[PQ_E1Y] = it[1][0];
    q[PQ_E1Z] = it[2][0];
    q[PQ_E2X] = it[0][1];
    q[PQ_E2Y] = it[1][1];
    q[PQ_E2Z] = it[2][1];
    q[PQ_E3X] = it[0][2];
    q[PQ_E3Y] = it[1][2];
    q[PQ_E3Z] = it[2][2];
    q[PQ_AX1] = 1.0;
    q[PQ_AX2] = axis1 > 0. ? sqrt(axis2 / axis1) : 0.0;
    q[PQ_AX3] = axis1 > 0. ? sqrt(axis3 / axis1) : 0.0;
    M_prev = M_sphere;
    V_prev = Volume;
    rad_prev = cur_rad;
  }
  (void)rad_prev;
  double *P = props;
  const double *last = prof + (nbins - 1) * AHF_PROF_NQ, *prev = prof + (nbins - 2) * AHF_PROF_NQ;
  CoM[0] = fmod(CoM[0] / M_sphere + 1., 1.);
  CoM[1] = fmod(CoM[1] / M_sphere + 1., 1.);
  CoM[2] = fmod(CoM[2] / M_sphere + 1., 1.);
  /* r2 and R_max over the sorted particle list, the first members ignored */
  double *y = scratch + 4 * n, *tmp = scratch + 5 * n, x_max, y_max;
  const int64_t ni = AHF_PROF_NIGNORE, np = n - ni;
  ahf_prof_find_max(radius + ni, dens_r2 + ni, np, 3, y, tmp, &x_max, &y_max);
  const double r2 = x_max;
  ahf_prof_find_max(radius + ni, Vcirc2 + ni, np, 1, y, tmp, &x_max, &y_max);
  const double R_max = x_max;
  int64_t ib = 0;
  while (radius[ib] < x_max && ib < n - 1) ib++;
  const double M_max = ovdens[ib] * FourPiThird * (radius[ib] * radius[ib] * radius[ib]);
  const double V_max = M_max / R_max;
  Lx = last[PQ_LX];
  Ly = last[PQ_LY];
  Lz = last[PQ_LZ];
  const double absL = sqrt(Lx * Lx + Ly * Ly + Lz * Lz);
  P[PP_NBINS] = (double)nbins;
  P[PP_MVIR] = M_sphere;
  P[PP_VX] = VXc / M_sphere;
  P[PP_VY] = VYc / M_sphere;
  P[PP_VZ] = VZc / M_sphere;
  P[PP_RMAX] = R_max;
  P[PP_V2MAX] = V_max;
  P[PP_R2] = r2;
  P[PP_SIGV] = last[PQ_SIGV];
  P[PP_VESC2] = last[PQ_VESC2];
  P[PP_EKIN] = last[PQ_EKIN];
  P[PP_EPOT] = last[PQ_EPOT];
  P[PP_AX1] = last[PQ_AX1];
  P[PP_AX2] = last[PQ_AX2];
  P[PP_AX3] = last[PQ_AX3];
  for (int k = 0; k < 9; k++) P[PP_E1X + k] = last[PQ_E1X + k];
  if (absL > 0) {
    P[PP_LX] = Lx / absL;
    P[PP_LY] = Ly / absL;
    P[PP_LZ] = Lz / absL;
    double lam = absL / M_sphere / sqrt(2. * M_sphere * R_vir);
    lam *= K->v_fac * sqrt(K->r_fac / (AHF_PROF_GRAV * K->m_fac));
    P[PP_LAMBDA] = lam;
    P[PP_LAMBDAE] = ahf_prof_lambdaE(K, absL, M_sphere, M_sphere, P[PP_EKIN], P[PP_EPOT]);
  } else {
    P[PP_LX] = P[PP_LY] = P[PP_LZ] = P[PP_LAMBDA] = P[PP_LAMBDAE] = 0.0;
  }
  P[PP_CNFW] = ahf_prof_cnfw(V_max, M_sphere / R_vir);
  P[PP_R1] = ahf_prof_r1(prof, nbins) / M_sphere / R_vir;
  P[PP_CR1] = ahf_prof_cr1(P[PP_R1]);
  const double Ts = 2.0 * (last[PQ_EKIN] - prev[PQ_EKIN]);
  const double frad = fabs(prev[PQ_R] / last[PQ_R]);
  P[PP_SURFP] = -0.125 * ((1. + frad) * (1. + frad) * (1. + frad)) / (1. - (frad * frad * frad)) * Ts;
  P[PP_FMHIRES] = M_hires > 0 ? M_hires / (M_hires + M_lores) : 0.0;
  if (mb >= 0) {
    const int64_t p = ip[mb] * ps;
    double dX = fabs(px[p] - Xc), dY = fabs(py[p] - Yc), dZ = fabs(pz[p] - Zc);
    if (dX > 0.5) dX -= 1.0;
    if (dY > 0.5) dY -= 1.0;
    if (dZ > 0.5) dZ -= 1.0;
    P[PP_MBP_OFF] = sqrt(dX * dX + dY * dY + dZ * dZ);
    P[PP_MBP_X] = px[p];
    P[PP_MBP_Y] = py[p];
    P[PP_MBP_Z] = pz[p];
    P[PP_MBP_VX] = vx[p];
    P[PP_MBP_VY] = vy[p];
    P[PP_MBP_VZ] = vz[p];
  } else {
    P[PP_MBP_OFF] = -1.0;
    P[PP_MBP_X] = P[PP_MBP_Y] = P[PP_MBP_Z] = P[PP_MBP_VX] = P[PP_MBP_VY] = P[PP_MBP_VZ] = 0.0;
  }
  {
    double dX = fabs(CoM[0] - Xc), dY = fabs(CoM[1] - Yc), dZ = fabs(CoM[2] - Zc);
    if (dX > 0.5) dX -= 1.0;
    if (dY > 0.5) dY -= 1.0;
    if (dZ > 0.5) dZ -= 1.0;
    P[PP_COM_OFF] = sqrt(dX * dX + dY * dY + dZ * dZ);
  }
  P[PP_COM_X] = CoM[0];
  P[PP_COM_Y] = CoM[1];
  P[PP_COM_Z] = CoM[2];
  return nbins;
}

#endif /* ORACLE_PROFILES_H */
