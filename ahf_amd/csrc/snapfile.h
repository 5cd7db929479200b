/*
 * snapfile.h -- the project's binary snapshot container (test and bench
 * infrastructure, not an AHF format).
 *
 * One file holds a synthetic particle set in AHF's internal units together
 * with the parameters of the hot path:
 *   - particles: FP32 positions in box units [0,1) and FP32 velocities in
 *     AHF units (velocity x v_fac = peculiar km/s), uint64 ids, in the order
 *     AHF's ll() would receive them (src/main.c:341-356 sorts them first);
 *   - AMR parameters: LgridDomain, LgridMax, NperDomCell, NperRefCell
 *     (src/startrun.c:499-517);
 *   - unbinding parameters: NminPerHalo, VescTune and the conversion factors
 *     r_fac, v_fac, phi_fac, Hubble of src/libahf/ahf_halos.c:199-213;
 *   - the virial overdensity of rem_outsideRvir(): global.ovlim and
 *     global.rho_vir (src/libahf/ahf_halos.c:204-221; rho_fac there is
 *     pmass/boxsize^3 and is not stored);
 *   - a halo list: centre (box units, FP64), gather radius (box units, the
 *     HALO.gatherRad that ahf_halos_sfc_gatherParts() receives) and member
 *     indices (CSR), the input rem_unbound() sees after
 *     sort_halo_particles().
 *
 * Generated snapshots hold their particles in Hilbert-key order, as AHF
 * sorts them (src/main.c:341-356).
 *
 * Layout (little endian, no padding between sections):
 *   snap_header
 *   float    pos[npart][3]
 *   float    mom[npart][3]
 *   uint64_t id[npart]
 *   double   centre[nhalos][3]
 *   double   gather_rad[nhalos]
 *   uint64_t halo_npart[nhalos]
 *   uint64_t ipart[n_ipart]
 */
#ifndef AHF_SNAPFILE_H
#define AHF_SNAPFILE_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SNAP_MAGIC "AHFSNAP3"

typedef struct {
  char     magic[8];
  uint64_t npart;
  double   boxsize;     /* Mpc/h */
  double   omega0, lambda0, a;
  double   pmass;       /* Msun/h */
  int32_t  L_dom;       /* LgridDomain */
  int32_t  L_max;       /* LgridMax */
  double   Nth_dom;     /* NperDomCell */
  double   Nth_ref;     /* NperRefCell */
  int32_t  minpart;     /* NminPerHalo */
  int32_t  pad0;
  double   vtune;       /* VescTune */
  double   r_fac, v_fac, phi_fac, hubble;
  uint64_t nhalos;
  uint64_t n_ipart;
  uint64_t seed;
  int32_t  config;      /* generator configuration, -1 if hand-made */
  int32_t  pad1;
  double   ovlim;       /* virial overdensity (global.ovlim) */
  double   rho_vir;     /* reference density it is relative to (global.rho_vir) */
} snap_header;

typedef struct {
  snap_header h;
  float    *pos;        /* npart*3 */
  float    *mom;        /* npart*3 */
  uint64_t *id;         /* npart */
  double   *centre;     /* nhalos*3 */
  double   *gather_rad; /* nhalos */
  uint64_t *halo_npart; /* nhalos */
  uint64_t *ipart;      /* n_ipart */
} snap_data;

static inline int snap_read(const char *path, snap_data *s)
{
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  memset(s, 0, sizeof(*s));
  if (fread(&s->h, sizeof(s->h), 1, f) != 1 || memcmp(s->h.magic, SNAP_MAGIC, 8)) {
    fclose(f);
    return -2;
  }
  uint64_t n = s->h.npart, nh = s->h.nhalos, ni = s->h.n_ipart;
  s->pos = (float *)malloc(n * 3 * sizeof(float) + 1);
  s->mom = (float *)malloc(n * 3 * sizeof(float) + 1);
  s->id = (uint64_t *)malloc(n * sizeof(uint64_t) + 1);
  s->centre = (double *)malloc(nh * 3 * sizeof(double) + 1);
  s->gather_rad = (double *)malloc(nh * sizeof(double) + 1);
  s->halo_npart = (uint64_t *)malloc(nh * sizeof(uint64_t) + 1);
  s->ipart = (uint64_t *)malloc(ni * sizeof(uint64_t) + 1);
  int bad = !s->pos || !s->mom || !s->id || !s->centre || !s->gather_rad || !s->halo_npart || !s->ipart;
  bad = bad || fread(s->pos, sizeof(float), n * 3, f) != n * 3;
  bad = bad || fread(s->mom, sizeof(float), n * 3, f) != n * 3;
  bad = bad || fread(s->id, sizeof(uint64_t), n, f) != n;
  bad = bad || fread(s->centre, sizeof(double), nh * 3, f) != nh * 3;
  bad = bad || fread(s->gather_rad, sizeof(double), nh, f) != nh;
  bad = bad || fread(s->halo_npart, sizeof(uint64_t), nh, f) != nh;
  bad = bad || fread(s->ipart, sizeof(uint64_t), ni, f) != ni;
  fclose(f);
  return bad ? -3 : 0;
}

static inline void snap_free(snap_data *s)
{
  free(s->pos); free(s->mom); free(s->id);
  free(s->centre); free(s->gather_rad); free(s->halo_npart); free(s->ipart);
  memset(s, 0, sizeof(*s));
}

#endif
