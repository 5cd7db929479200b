/*
 * emu_device.c -- test-only host stand-in for the device API
 * (include/ahf_gpu.h) that the drop-in's grid functions call: the AMR
 * build is the host replay of the device pipeline (amr_emu.c), and the
 * level read-back copies its levels out.  It lets ahf_amd/csrc/dropin.c
 * build AHF's grids on a machine without a GPU, for the ownership check in
 * dropin_free_check.c; every other entry point reports AHFGPU_EUNSUP.  It
 * is never linked into the library.
 */
#define AMR_EMU_NO_MAIN
#include "amr_emu.c"

#include "../../include/ahf_gpu.h"

struct ahfgpu_ctx {
  float *pos;
  uint64_t n;
  lev_t lev[EMU_MAX_LEVELS];
  int nlev;
  host_level *hl;
  char err[128];
};

static void emu_drop(ahfgpu_ctx *c)
{
  if (c->hl) emu_free(c->lev, c->nlev, c->hl);
  c->hl = NULL;
  c->nlev = 0;
}

static int unsup(ahfgpu_ctx *c, const char *what)
{
  if (c) snprintf(c->err, sizeof(c->err), "%s: not available in the host stand-in", what);
  return AHFGPU_EUNSUP;
}

int ahfgpu_create(ahfgpu_ctx **ctx, int device)
{
  (void)device;
  *ctx = (ahfgpu_ctx *)calloc(1, sizeof(ahfgpu_ctx));
  return *ctx ? 0 : AHFGPU_ENOMEM;
}

void ahfgpu_destroy(ahfgpu_ctx *c)
{
  if (!c) return;
  emu_drop(c);
  free(c->pos);
  free(c);
}

const char *ahfgpu_last_error(const ahfgpu_ctx *c) { return c ? c->err : "no context"; }

int ahfgpu_set_timing(ahfgpu_ctx *c, int on)
{
  (void)c;
  (void)on;
  return 0;
}

int ahfgpu_set_particles(ahfgpu_ctx *c, const float *pos_xyz, const float *mom_xyz, uint64_t n)
{
  (void)mom_xyz;
  emu_drop(c);
  free(c->pos);
  c->pos = (float *)malloc(3 * (size_t)n * sizeof(float) + 4);
  if (!c->pos) return AHFGPU_ENOMEM;
  memcpy(c->pos, pos_xyz, 3 * (size_t)n * sizeof(float));
  c->n = n;
  return 0;
}

int ahfgpu_amr_build(ahfgpu_ctx *c, const ahfgpu_amr_params *p, ahfgpu_amr_info *info)
{
  emu_drop(c);
  snap_data s;
  memset(&s, 0, sizeof(s));
  s.h.npart = c->n;
  s.h.L_dom = p->L_dom;
  s.h.L_max = p->L_max;
  s.h.Nth_dom = p->Nth_dom;
  s.h.Nth_ref = p->Nth_ref;
  s.pos = c->pos;
  if (emu_build(&s, c->lev, &c->nlev, &c->hl)) snprintf(c->err, sizeof(c->err), "refinement anomaly");
  memset(info, 0, sizeof(*info));
  info->n_levels = c->nlev;
  info->rolled_back = c->lev[c->nlev - 1].rolled_back_child;
  for (int k = 0; k < c->nlev && k < AHFGPU_MAX_LEVELS; k++) {
    info->n_nodes[k] = (int64_t)c->hl[k].size_no_nodes;
    info->n_part[k] = (int64_t)c->hl[k].size_no_part;
    info->deposits += info->n_part[k];
  }
  return 0;
}

int ahfgpu_amr_level_sizes(ahfgpu_ctx *c, int lev, int64_t *L, int64_t *n_oct, int64_t *n_listed, double *critdens,
                           double *masstopartdens)
{
  if (!c->hl || lev < 0 || lev >= c->nlev) return AHFGPU_EINVAL;
  const host_level *h = &c->hl[lev];
  if (L) *L = h->L;
  if (n_oct) *n_oct = h->n_oct;
  if (n_listed) *n_listed = h->list0[8 * h->n_oct];
  if (critdens) *critdens = h->critdens;
  if (masstopartdens) *masstopartdens = h->masstopartdens;
  return 0;
}

int ahfgpu_amr_level_copy(ahfgpu_ctx *c, int lev, uint64_t *oct_key, uint8_t *oct_link, float *dens, int64_t *list0,
                          uint32_t *list)
{
  if (!c->hl || lev < 0 || lev >= c->nlev) return AHFGPU_EINVAL;
  const host_level *h = &c->hl[lev];
  const int64_t nn = 8 * h->n_oct;
  if (oct_key) memcpy(oct_key, h->oct_key, (size_t)h->n_oct * 8);
  if (oct_link) memcpy(oct_link, h->oct_link, (size_t)h->n_oct);
  if (dens) memcpy(dens, h->dens, (size_t)nn * 4);
  if (list0) memcpy(list0, h->list0, (size_t)(nn + 1) * 8);
  if (list) memcpy(list, h->list, (size_t)h->list0[nn] * 4);
  return 0;
}

int ahfgpu_deposit_level(ahfgpu_ctx *c, int64_t L, const uint64_t *oct_key, const uint8_t *oct_link, int64_t n_oct,
                         const int64_t *list0, const uint32_t *list, double masstopartdens, float *dens_out)
{
  (void)L; (void)oct_key; (void)oct_link; (void)n_oct; (void)list0; (void)list; (void)masstopartdens;
  (void)dens_out;
  return unsup(c, "deposit");
}

int ahfgpu_set_halos(ahfgpu_ctx *c, const double *centres, const uint64_t *npart, const uint64_t *ipart, uint64_t nh)
{
  (void)centres; (void)npart; (void)ipart; (void)nh;
  return unsup(c, "haloes");
}

int ahfgpu_set_sfc_keys(ahfgpu_ctx *c, const uint64_t *keys, uint64_t n)
{
  (void)keys; (void)n;
  return unsup(c, "SFC keys");
}

int ahfgpu_gather_halos(ahfgpu_ctx *c, const double *centres, const double *gather_rad, const uint8_t *skip,
                        uint64_t nh)
{
  (void)centres; (void)gather_rad; (void)skip; (void)nh;
  return unsup(c, "gather");
}

int ahfgpu_halo_members(ahfgpu_ctx *c, uint64_t *npart, uint64_t *ipart)
{
  (void)npart; (void)ipart;
  return unsup(c, "halo members");
}

int ahfgpu_unbind(ahfgpu_ctx *c, const ahfgpu_unbind_params *p, int mode, double band, ahfgpu_unbind_info *info)
{
  (void)p; (void)mode; (void)band; (void)info;
  return unsup(c, "unbinding");
}

int ahfgpu_unbind_results(ahfgpu_ctx *c, uint64_t *npart, double *props, uint64_t *ipart)
{
  (void)npart; (void)props; (void)ipart;
  return unsup(c, "unbinding results");
}

void ahfgpu_release_halos(ahfgpu_ctx *c) { (void)c; }

int ahfgpu_halo_profiles(ahfgpu_ctx *c, const ahfgpu_profile_params *p, const double *phi0_rvir, int32_t *nbins)
{
  (void)p; (void)phi0_rvir; (void)nbins;
  return unsup(c, "profiles");
}

int ahfgpu_profile_results(ahfgpu_ctx *c, double *props, double *prof)
{
  (void)props; (void)prof;
  return unsup(c, "profile results");
}

int ahfgpu_gridinfo(ahfgpu_ctx *c, const ahfgpu_gridinfo_params *p, ahfgpu_gridinfo_info *info)
{
  (void)p; (void)info;
  return unsup(c, "gridinfo");
}

int ahfgpu_gridinfo_results(ahfgpu_ctx *c, int grid, int32_t *colour, int32_t *refs)
{
  (void)grid; (void)colour; (void)refs;
  return unsup(c, "gridinfo results");
}
