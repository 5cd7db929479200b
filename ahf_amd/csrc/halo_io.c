/*
 * halo_io.c -- AHF's ASCII catalogue writers for the haloes the library
 * constructs and profiles: .AHF_halos, .AHF_profiles and .AHF_particles
 * (src/libahf/ahf_io.c:79-210 and the legacy formats at :637-826,
 * :828-969, :1094-1175, :2209-2345), for AHF's default flags (no
 * AHFnewHaloIDs, AHFcR1, AHFvmbp, MULTIMASS or GAS_PARTICLES).
 *
 * The column names, their order and each column's printf conversion are
 * the file format, so they are AHF's; the code that emits them is a table
 * of columns per file.  Host C: the values come from the HALO structs that
 * HaloProfiles (on the GPU) filled.
 */
#define _POSIX_C_SOURCE 200809L
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ahf_dropin.h"

#define PDM_ID 1 /* (int)(-PDM), PDM = -1.0 (src/param.h:27) */

/* ------------------------------------------------------------- columns */

typedef double (*halo_val)(const HALO *h);
typedef struct {
  const char *name, *fmt; /* fmt: the conversion, with its leading tab */
  int kind;               /* 0 double, 1 long, 2 int */
  halo_val get;
} halo_col;

/* the index in the written order of the halo idx[] points at
 * (idx_inv, src/libutility/specific.c:183-197) */
static int idx_inv(const unsigned long *idx, int numHalos, int i)
{
  if (i < 0) return -1;
  int j = 0;
  while (j < numHalos && idx[j] != (unsigned long)i) j++;
  return j;
}

#define HV(NAME, EXPR) \
  static double NAME(const HALO *h) { return (EXPR); }
HV(hv_mvir, h->M_vir * m_fac)
HV(hv_npart, (double)h->npart)
HV(hv_xc, h->pos.x * x_fac * 1000.)
HV(hv_yc, h->pos.y * x_fac * 1000.)
HV(hv_zc, h->pos.z * x_fac * 1000.)
HV(hv_vxc, h->vel.x * v_fac)
HV(hv_vyc, h->vel.y * v_fac)
HV(hv_vzc, h->vel.z * v_fac)
HV(hv_rvir, h->R_vir * x_fac * 1000.)
HV(hv_rmax, h->R_max * x_fac * 1000.)
HV(hv_r2, h->r2 * x_fac * 1000.)
HV(hv_mbp, h->mbp_offset * x_fac * 1000.)
HV(hv_com, h->com_offset * x_fac * 1000.)
HV(hv_vmax, sqrt(h->V2_max * phi_fac))
HV(hv_vesc, sqrt(h->v_esc2 * phi_fac))
HV(hv_sigv, h->sigV * v_fac)
HV(hv_lambda, h->lambda)
HV(hv_lambdae, h->lambdaE)
HV(hv_lx, h->AngMom.x)
HV(hv_ly, h->AngMom.y)
HV(hv_lz, h->AngMom.z)
HV(hv_b, h->axis.y)
HV(hv_c, h->axis.z)
HV(hv_eax, h->E1.x)
HV(hv_eay, h->E1.y)
HV(hv_eaz, h->E1.z)
HV(hv_ebx, h->E2.x)
HV(hv_eby, h->E2.y)
HV(hv_ebz, h->E2.z)
HV(hv_ecx, h->E3.x)
HV(hv_ecy, h->E3.y)
HV(hv_ecz, h->E3.z)
HV(hv_ovdens, h->ovdens)
HV(hv_nbins, (double)h->prof.nbins)
HV(hv_fmhires, h->fMhires)
HV(hv_ekin, h->Ekin * m_fac * (v_fac * v_fac))
HV(hv_epot, h->Epot * m_fac * phi_fac)
HV(hv_surfp, h->SurfP * m_fac * (v_fac * v_fac))
HV(hv_phi0, h->Phi0 * phi_fac)
HV(hv_cnfw, h->cNFW)
HV(hv_nsub, (double)h->numSubStruct)
#undef HV

/* .AHF_halos after the two leading identifier columns (write_halos_line,
 * src/libahf/ahf_io.c:2234-2280) */
static const halo_col k_halo_cols[] = {
    {"numSubStruct", "\t%10d", 2, hv_nsub}, {"Mhalo", "\t%12.6g", 0, hv_mvir},
    {"npart", "\t%10ld", 1, hv_npart},      {"Xc", "\t%16.8f", 0, hv_xc},
    {"Yc", "\t%16.8f", 0, hv_yc},           {"Zc", "\t%16.8f", 0, hv_zc},
    {"VXc", "\t%8.2f", 0, hv_vxc},          {"VYc", "\t%8.2f", 0, hv_vyc},
    {"VZc", "\t%8.2f", 0, hv_vzc},          {"Rhalo", "\t%10.2f", 0, hv_rvir},
    {"Rmax", "\t%10.2f", 0, hv_rmax},       {"r2", "\t%10.5f", 0, hv_r2},
    {"mbp_offset", "\t%10.5f", 0, hv_mbp},  {"com_offset", "\t%10.5f", 0, hv_com},
    {"Vmax", "\t%8.2f", 0, hv_vmax},        {"v_esc", "\t%12.6f", 0, hv_vesc},
    {"sigV", "\t%8.2f", 0, hv_sigv},        {"lambda", "\t%10.6f", 0, hv_lambda},
    {"lambdaE", "\t%12.6f", 0, hv_lambdae}, {"Lx", "\t%12.4g", 0, hv_lx},
    {"Ly", "\t%12.4g", 0, hv_ly},           {"Lz", "\t%12.4g", 0, hv_lz},
    {"b", "\t%10.6f", 0, hv_b},             {"c", "\t%10.6f", 0, hv_c},
    {"Eax", "\t%10.6f", 0, hv_eax},         {"Eay", "\t%10.6f", 0, hv_eay},
    {"Eaz", "\t%10.6f", 0, hv_eaz},         {"Ebx", "\t%10.6f", 0, hv_ebx},
    {"Eby", "\t%10.6f", 0, hv_eby},         {"Ebz", "\t%10.6f", 0, hv_ebz},
    {"Ecx", "\t%10.6f", 0, hv_ecx},         {"Ecy", "\t%10.6f", 0, hv_ecy},
    {"Ecz", "\t%10.6f", 0, hv_ecz},         {"ovdens", "\t%8.2f", 0, hv_ovdens},
    {"nbins", "\t%6d", 2, hv_nbins},        {"fMhires", "\t%8.6f", 0, hv_fmhires},
    {"Ekin", "\t%12.6g", 0, hv_ekin},       {"Epot", "\t%12.6g", 0, hv_epot},
    {"SurfP", "\t%12.6g", 0, hv_surfp},     {"Phi0", "\t%12.6g", 0, hv_phi0},
    {"cNFW", "\t%12.6g", 0, hv_cnfw},
};

/* per-bin columns of .AHF_profiles (src/libahf/ahf_io.c:657-680, 760-783) */
typedef double (*bin_val)(const HALO *h, int b);
typedef struct {
  const char *name, *fmt;
  int kind; /* 0 double, 1 long */
  bin_val get;
} bin_col;

#define BV(NAME, EXPR) \
  static double NAME(const HALO *h, int b) { return (EXPR); }
BV(bv_npart, (double)h->prof.npart[b])
BV(bv_m, h->prof.nvpart[b] * m_fac)
BV(bv_ovdens, h->prof.ovdens[b] * rho_fac / global.rho_vir)
BV(bv_dens, h->prof.dens[b] * rho_fac / global.rho_vir)
BV(bv_vcirc, sqrt(h->prof.v2_circ[b] * phi_fac))
BV(bv_vesc, sqrt(h->prof.v_esc2[b] * phi_fac))
BV(bv_sigv, h->prof.sig_v[b] * v_fac)
BV(bv_lx, h->prof.Lx[b] * m_fac * r_fac * v_fac)
BV(bv_ly, h->prof.Ly[b] * m_fac * r_fac * v_fac)
BV(bv_lz, h->prof.Lz[b] * m_fac * r_fac * v_fac)
BV(bv_b, h->prof.axis2[b])
BV(bv_c, h->prof.axis3[b])
BV(bv_eax, h->prof.E1x[b])
BV(bv_eay, h->prof.E1y[b])
BV(bv_eaz, h->prof.E1z[b])
BV(bv_ebx, h->prof.E2x[b])
BV(bv_eby, h->prof.E2y[b])
BV(bv_ebz, h->prof.E2z[b])
BV(bv_ecx, h->prof.E3x[b])
BV(bv_ecy, h->prof.E3y[b])
BV(bv_ecz, h->prof.E3z[b])
BV(bv_ekin, h->prof.Ekin[b] * m_fac * (v_fac * v_fac))
BV(bv_epot, h->prof.Epot[b] * m_fac * phi_fac)
#undef BV

static const bin_col k_bin_cols[] = {
    {"npart", "\t%10ld", 1, bv_npart},  {"M_in_r", "\t%e", 0, bv_m},     {"ovdens", "\t%12.2f", 0, bv_ovdens},
    {"dens", "\t%12.2f", 0, bv_dens},   {"vcirc", "\t%8.2f", 0, bv_vcirc}, {"vesc", "\t%10.6f", 0, bv_vesc},
    {"sigv", "\t%8.2f", 0, bv_sigv},    {"Lx", "\t%12.4g", 0, bv_lx},    {"Ly", "\t%12.4g", 0, bv_ly},
    {"Lz", "\t%12.4g", 0, bv_lz},       {"b", "\t%10.6f", 0, bv_b},      {"c", "\t%10.6f", 0, bv_c},
    {"Eax", "\t%10.6f", 0, bv_eax},     {"Eay", "\t%10.6f", 0, bv_eay},  {"Eaz", "\t%10.6f", 0, bv_eaz},
    {"Ebx", "\t%10.6f", 0, bv_ebx},     {"Eby", "\t%10.6f", 0, bv_eby},  {"Ebz", "\t%10.6f", 0, bv_ebz},
    {"Ecx", "\t%10.6f", 0, bv_ecx},     {"Ecy", "\t%10.6f", 0, bv_ecy},  {"Ecz", "\t%10.6f", 0, bv_ecz},
    {"Ekin", "\t%e", 0, bv_ekin},       {"Epot", "\t%e", 0, bv_epot},
};

#define NCOLS(t) ((int)(sizeof(t) / sizeof((t)[0])))

/* "#" and each column as "name(k)\t", numbered from 1 */
static void header(FILE *f, const char *const *names, int n)
{
  fputc('#', f);
  for (int k = 0; k < n; k++) fprintf(f, "%s(%d)\t", names[k], k + 1);
  fputc('\n', f);
}

static FILE *open_out(const char *fprefix, const char *suffix)
{
  const size_t n = strlen(fprefix) + strlen(suffix) + 1;
  char *name = (char *)malloc(n);
  if (!name) return NULL;
  snprintf(name, n, "%s%s", fprefix, suffix);
  /* AHF's default build names each file on stderr (VERBOSE, src/define.h:20) */
  fprintf(stderr, "%s\n", name);
  FILE *f = fopen(name, "w");
  if (!f) {
    fprintf(stderr, "could not open %s\n", name);
    free(name);
    exit(1);
  }
  free(name);
  return f;
}

static void put_value(FILE *f, const char *fmt, int kind, double v)
{
  if (kind == 1) fprintf(f, fmt, (long)v);
  else if (kind == 2) fprintf(f, fmt, (int)v);
  else fprintf(f, fmt, v);
}

/* -------------------------------------------------------------- writers */

void ahf_io_WriteHalos(const char *fprefix, HALO *halos, unsigned long *idx, int numHalos)
{
  FILE *f = open_out(fprefix, ".AHF_halos");
  if (!f) return;
  const char *names[2 + NCOLS(k_halo_cols)];
  names[0] = "ID";
  names[1] = "hostHalo";
  for (int k = 0; k < NCOLS(k_halo_cols); k++) names[2 + k] = k_halo_cols[k].name;
  header(f, names, 2 + NCOLS(k_halo_cols));
  for (int j = 0; j < numHalos; j++) {
    const HALO *h = halos + idx[j];
    if (h->npart < (long unsigned)simu.AHF_MINPART) continue;
    fprintf(f, "%10lu", (long unsigned)j);
    fprintf(f, "\t%10d", idx_inv(idx, numHalos, h->hostHalo));
    for (int k = 0; k < NCOLS(k_halo_cols); k++)
      put_value(f, k_halo_cols[k].fmt, k_halo_cols[k].kind, k_halo_cols[k].get(h));
    fputc('\n', f);
  }
  fclose(f);
}

void ahf_io_WriteProfiles(const char *fprefix, HALO *halos, unsigned long *idx, int numHalos)
{
  FILE *f = open_out(fprefix, ".AHF_profiles");
  if (!f) return;
  const char *names[1 + NCOLS(k_bin_cols)];
  names[0] = "r";
  for (int k = 0; k < NCOLS(k_bin_cols); k++) names[1 + k] = k_bin_cols[k].name;
  header(f, names, 1 + NCOLS(k_bin_cols));
  for (int j = 0; j < numHalos; j++) {
    const HALO *h = halos + idx[j];
    if (h->npart < (long unsigned)simu.AHF_MINPART) continue;
    const int nb = h->prof.nbins;
    /* bins inside the convergence radius of Power et al. (2003), Eq. 20,
     * are written with a negative radius (ahf_io.c:722-759) */
    int r_conv = 0;
    for (int b = 0; b < nb; b++) {
      const double nDM = (double)h->prof.npart[b];
      double t_relax = nDM / log(nDM) / 8. * h->prof.r[b] / sqrt(h->prof.v2_circ[b]);
      const double t0 =
          0.6 * h->prof.r[nb - 1] / sqrt(h->prof.v2_circ[nb - 1]) * r_fac / sqrt(phi_fac) * 1E3;
      t_relax *= r_fac / sqrt(phi_fac);
      t_relax *= 1E3;
      if (t_relax < t0) r_conv = b + 1;
    }
    for (int b = 0; b < nb; b++) {
      const double rad = b < r_conv ? -h->prof.r[b] : h->prof.r[b];
      fprintf(f, "%8.4f", rad * x_fac * 1000.);
      for (int k = 0; k < NCOLS(k_bin_cols); k++)
        put_value(f, k_bin_cols[k].fmt, k_bin_cols[k].kind, k_bin_cols[k].get(h, b));
      fputc('\n', f);
    }
  }
  fclose(f);
}

void ahf_io_WriteParticles(const char *fprefix, HALO *halos, unsigned long *idx, int numHalos)
{
  FILE *f = open_out(fprefix, ".AHF_particles");
  if (!f) return;
  /* the halo count leads the file: a placeholder of the same width is
   * overwritten once the count is known (ahf_io.c:1106, 1168-1171) */
  unsigned long good = 0;
  fprintf(f, "%15lu\n", good);
  for (int j = 0; j < numHalos; j++) {
    const HALO *h = halos + idx[j];
    if (h->npart < (long unsigned)simu.AHF_MINPART) continue;
    good++;
    fprintf(f, "%lu %10ld\n", (unsigned long)h->npart, (long)j);
    for (long unsigned k = 0; k < h->npart; k++)
      fprintf(f, "%" PRIu64 "\t%d\n", (uint64_t)global.fst_part[h->ipart[k]].id, PDM_ID);
  }
  rewind(f);
  fprintf(f, "%15lu", good);
  fclose(f);
}
