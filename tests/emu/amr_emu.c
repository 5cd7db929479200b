/*
 * amr_emu.c -- test-only host replay of the device AMR pipeline.
 *
 * Runs the same per-thread functions as the HIP kernels
 * (ahf_amd/csrc/amr_core.h), one loop iteration per GPU thread and in the
 * same kernel order as ahf_amd/csrc/amr_gpu.hip, then writes the AHFGRID1
 * dump.  It exists so the level representation and every per-thread rule can
 * be checked against the CPU oracle without a GPU; it is never linked into
 * the library.
 *
 *   amr_emu <snapshot> <grids.out> [<ginfo.out> <ovlim> <rho_vir> <refine_mass>]
 *
 * With the last four arguments it also replays the ahf_gridinfo pass with
 * the per-octet rules of the kernels (ahf_amd/csrc/gridinfo_core.h: octet
 * edges, start and periodic flags; components under their lowest octet,
 * colours from the starts, refinements at the roots) and writes the
 * AHFGINF1 dump of the oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../ahf_amd/csrc/amr_core.h"
#include "../../ahf_amd/csrc/gridinfo_core.h"
#include "../../ahf_amd/csrc/level_host.h"
#include "../../ahf_amd/csrc/snapfile.h"

typedef struct {
  level_view v;
  uint64_t *oct_key;
  uint8_t *oct_link;
  int32_t *nbr;
  uint64_t *row_key;
  int32_t *row_oct0, *row_plane, *plane_z, *plane_row0;
  int64_t *oct_parent;
  uint32_t *pmask;
  float *dens;
  uint8_t *flags;
  int64_t *node_p0;
  uint32_t *plist;
  uint8_t *moved;
  int32_t *node2oct;
  double m2d, m2pd, crit;
  int rolled_back_child; /* next level was built and rolled back */
  int64_t fine_n_oct;    /* octets of the rolled-back level */
  int64_t *fine_p0;      /* its CSR (kept for the appended lists) */
  uint32_t *fine_list;
} lev_t;

static void *xm(size_t n)
{
  void *p = calloc(n ? n : 1, 1);
  if (!p) { fprintf(stderr, "emu: out of memory\n"); exit(3); }
  return p;
}

static void set_view(lev_t *l, int64_t L, int64_t n_oct, int64_t n_row, int64_t n_plane)
{
  l->v.L = L; l->v.n_oct = n_oct; l->v.n_row = n_row; l->v.n_plane = n_plane;
  l->v.oct_key = l->oct_key; l->v.oct_link = l->oct_link; l->v.nbr = l->nbr;
  l->v.row_key = l->row_key; l->v.row_oct0 = l->row_oct0; l->v.row_plane = l->row_plane;
  l->v.plane_z = l->plane_z; l->v.plane_row0 = l->plane_row0;
}

/* deposit as the device does it: one weight record per bucket entry
 * (k_prep), then the exact-order flat walk of every node over them
 * (k_gather_light/_heavy) */
static void deposit_q(lev_t *l, const float *px, const float *py, const float *pz, int descending)
{
  int64_t nn = 8 * l->v.n_oct, ne = l->node_p0[nn];
  if (!l->pmask) {
    l->pmask = (uint32_t *)xm((size_t)nn * 4);
    for (int64_t n = 0; n < nn; n++) l->pmask[n] = ahf_presence_mask(&l->v, n / 8, (int)(n % 8));
  }
  ahf_w9 *W = (ahf_w9 *)xm((size_t)(ne > 0 ? ne : 1) * sizeof(ahf_w9));
  for (int64_t n = 0; n < nn; n++)
    for (int64_t q = l->node_p0[n]; q < l->node_p0[n + 1]; q++) {
      uint32_t p = l->plist[q];
      ahf_entry_weights(&l->v, n, px[p], py[p], pz[p], l->m2pd, &W[q]);
    }
  l->dens = (float *)xm((size_t)nn * 4);
  for (int64_t n = 0; n < nn; n++)
    l->dens[n] = ahf_gather_dens_w(&l->v, l->pmask, l->node_p0, W, descending, n / 8, (int)(n % 8));
  free(W);
}

static int64_t gi_find(int64_t *par, int64_t x)
{
  while (par[x] != x) x = par[x];
  return x;
}

static void gi_join(int64_t *par, int64_t a, int64_t b)
{
  a = gi_find(par, a);
  b = gi_find(par, b);
  if (a < b) par[b] = a;
  else if (b < a) par[a] = b;
}

/* traversal rank (z, y, x) of node n = 8*o + c, from the octet rows and
 * planes: the order in which the AHFGINF1 dump lists a grid's nodes */
static int64_t node_rank(const level_view *v, int64_t n)
{
  const int64_t o = n >> 3;
  const int c = (int)(n & 7), cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
  int64_t lo = 0, hi = v->n_row - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (v->row_oct0[mid] <= o) lo = mid;
    else hi = mid - 1;
  }
  const int64_t r = lo, p = v->row_plane[r];
  const int64_t pb = v->row_oct0[v->plane_row0[p]], pn = v->row_oct0[v->plane_row0[p + 1]] - pb;
  const int64_t rb = v->row_oct0[r], rn = v->row_oct0[r + 1] - rb;
  return 8 * pb + cz * 4 * pn + 4 * (rb - pb) + cy * 2 * rn + 2 * (o - rb) + cx;
}

/* the gridinfo pass on levels [0, nlev) with the kernels' per-octet rules
 * (gridinfo_core.h), components joined sequentially */
static int gridinfo_replay(lev_t *lev, int nlev, const snap_data *s, const char *out, double ovlim, double rho_vir,
                           double refine_mass)
{
  int startg = 1;
  double ovd = 0.0;
  for (int i = 0; i < nlev; i++) {
    const double len = s->h.boxsize / (double)lev[i].v.L;
    const double vol = len * len * len;
    ovd = (s->h.Nth_ref * refine_mass / vol) / rho_vir;
    if (ovd < ovlim) startg = i;
  }
  const int min_ref = startg, no_grids = nlev - 1 - min_ref + 1;
  FILE *f = fopen(out, "wb");
  if (!f) return 1;
  int32_t *num_iso = (int32_t *)xm(sizeof(int32_t) * (size_t)(no_grids > 0 ? no_grids : 1));
  int32_t **colour = (int32_t **)xm(sizeof(int32_t *) * (size_t)(no_grids > 0 ? no_grids : 1));
  int64_t *nnode = (int64_t *)xm(sizeof(int64_t) * (size_t)(no_grids > 0 ? no_grids : 1));
  int32_t *refs = NULL;
  int64_t nrefs = 0, base = 0;
  int32_t tot = 0;
  for (int g = 0; g < no_grids; g++) {
    lev_t *l = &lev[min_ref + g];
    const int64_t no = l->v.n_oct, nn = 8 * no;
    int64_t *par = (int64_t *)xm((size_t)no * 8), *sname = (int64_t *)xm((size_t)no * 8);
    int64_t *rpos = (int64_t *)xm((size_t)no * 8);
    uint8_t *start = (uint8_t *)xm((size_t)no), *seen = (uint8_t *)xm((size_t)nn);
    uint32_t *pf = (uint32_t *)xm((size_t)no * 4);
    for (int64_t o = 0; o < no; o++) par[o] = o;
    for (int64_t o = 0; o < no; o++) {
      int64_t e[3];
      ahf_gi_oct_edges(&l->v, o, e);
      start[o] = (uint8_t)ahf_gi_oct_start(&l->v, o, e);
      pf[o] = ahf_gi_oct_periodic(&l->v, o, e);
      for (int d = 0; d < 3; d++)
        if (e[d] >= 0) gi_join(par, e[d], o);
    }
    for (int64_t o = 0; o < no; o++)
      if (par[o] != o) pf[gi_find(par, o)] |= pf[o];
    int64_t ns = 0, nr = 0;
    for (int64_t o = 0; o < no; o++) {
      sname[o] = ns;
      ns += start[o];
      rpos[o] = nr;
      nr += par[o] == o;
    }
    colour[g] = (int32_t *)xm((size_t)nn * 4);
    nnode[g] = nn;
    for (int64_t n = 0; n < nn; n++) {
      const int64_t r = node_rank(&l->v, n);
      if (r < 0 || r >= nn || seen[r]) {
        fprintf(stderr, "emu: traversal ranks of grid %d are not a permutation\n", g);
        return 4;
      }
      seen[r] = 1;
      colour[g][r] = (int32_t)(base + sname[gi_find(par, n >> 3)] + 1);
    }
    refs = (int32_t *)realloc(refs, sizeof(int32_t) * 6 * (size_t)(nrefs + nr + 1));
    for (int64_t o = 0; o < no; o++)
      if (par[o] == o) {
        int32_t *e = refs + 6 * (nrefs + rpos[o]);
        e[0] = (int32_t)(base + sname[o] + 1);
        e[1] = g;
        e[2] = (int32_t)rpos[o];
        e[3] = (int32_t)(pf[o] & 1u);
        e[4] = (int32_t)((pf[o] >> 1) & 1u);
        e[5] = (int32_t)((pf[o] >> 2) & 1u);
      }
    nrefs += nr;
    num_iso[g] = (int32_t)nr;
    tot += (int32_t)nr;
    base += ns;
    free(par); free(sname); free(rpos); free(start); free(seen); free(pf);
  }
  const int32_t hdr[4] = {min_ref, no_grids, tot, (int32_t)nrefs};
  fwrite("AHFGINF1", 1, 8, f);
  fwrite(hdr, 4, 4, f);
  fwrite(&ovd, 8, 1, f);
  if (no_grids > 0) fwrite(num_iso, 4, (size_t)no_grids, f);
  for (int g = 0; g < no_grids; g++) {
    const uint64_t n = (uint64_t)nnode[g];
    fwrite(&n, 8, 1, f);
    fwrite(colour[g], 4, (size_t)n, f);
    free(colour[g]);
  }
  if (nrefs) fwrite(refs, 4, 6 * (size_t)nrefs, f);
  fclose(f);
  free(refs); free(colour); free(nnode); free(num_iso);
  return 0;
}

/* the device pipeline replayed on snapshot s (only its header's grid
 * parameters and its positions are read): the levels in lev[0..*nlev_out)
 * (lev[*nlev_out] holds a level built and rolled back, if any) and their
 * final host levels in *hl_out, whose key, link and density arrays are the
 * levels' own.  lev must hold EMU_MAX_LEVELS zeroed entries.  Returns the
 * anomaly flag of the refinement rules. */
#define EMU_MAX_LEVELS 64
static int emu_build(const snap_data *sp, lev_t *lev, int *nlev_out, host_level **hl_out)
{
  const snap_data s = *sp;
  const int64_t N = (int64_t)s.h.npart, L0 = s.h.L_dom, half = L0 / 2;
  float *px = (float *)xm((size_t)N * 4), *py = (float *)xm((size_t)N * 4), *pz = (float *)xm((size_t)N * 4);
  for (int64_t p = 0; p < N; p++) { px[p] = s.pos[3 * p]; py[p] = s.pos[3 * p + 1]; pz[p] = s.pos[3 * p + 2]; }

  int nlev = 1;
  /* ---- domain level as octets of a virtual L/2 grid ---- */
  lev_t *d = &lev[0];
  int64_t no = half * half * half;
  d->oct_key = (uint64_t *)xm((size_t)no * 8);
  d->oct_link = (uint8_t *)xm((size_t)no);
  d->nbr = (int32_t *)xm((size_t)no * 27 * 4);
  d->row_key = (uint64_t *)xm((size_t)half * half * 8);
  d->row_oct0 = (int32_t *)xm((size_t)(half * half + 1) * 4);
  d->row_plane = (int32_t *)xm((size_t)half * half * 4);
  d->plane_z = (int32_t *)xm((size_t)half * 4);
  d->plane_row0 = (int32_t *)xm((size_t)(half + 1) * 4);
  for (int64_t o = 0; o < no; o++) {
    int64_t ox = o % half, oy = (o / half) % half, oz = o / (half * half);
    d->oct_key[o] = ahf_pack3((uint64_t)ox, (uint64_t)oy, (uint64_t)oz);
    d->oct_link[o] = ox < half - 1;
    for (int ez = -1; ez <= 1; ez++)
      for (int ey = -1; ey <= 1; ey++)
        for (int ex = -1; ex <= 1; ex++) {
          int64_t nx = (ox + ex + half) % half, ny = (oy + ey + half) % half, nz = (oz + ez + half) % half;
          d->nbr[27 * o + ahf_e27(ex, ey, ez)] = (int32_t)((nz * half + ny) * half + nx);
        }
  }
  for (int64_t r = 0; r < half * half; r++) {
    d->row_key[r] = (uint64_t)(r % half) | ((uint64_t)(r / half) << 21);
    d->row_oct0[r] = (int32_t)(r * half);
    d->row_plane[r] = (int32_t)(r / half);
  }
  d->row_oct0[half * half] = (int32_t)no;
  for (int64_t p = 0; p <= half; p++) {
    if (p < half) d->plane_z[p] = (int32_t)p;
    d->plane_row0[p] = (int32_t)(p * half);
  }
  set_view(d, L0, no, half * half, half);
  double dl = (double)L0;
  d->m2d = (dl * dl * dl) / (double)N;
  d->m2pd = (dl * dl * dl) / (double)N;
  d->crit = s.h.Nth_dom * d->m2d;
  /* ll(): floor binning, stable bucket in array order */
  {
    int64_t nn = 8 * no;
    int64_t *key = (int64_t *)xm((size_t)N * 8);
    d->node_p0 = (int64_t *)xm((size_t)(nn + 1) * 8);
    for (int64_t p = 0; p < N; p++) {
      int64_t c[3];
      for (int k = 0; k < 3; k++) {
        uint64_t n = (uint64_t)(dl * (double)s.pos[3 * p + k]);
        if (n > (uint64_t)(L0 - 1)) n = 0;
        c[k] = (int64_t)n;
      }
      int64_t o = ((c[2] >> 1) * half + (c[1] >> 1)) * half + (c[0] >> 1);
      key[p] = 8 * o + (c[2] & 1) * 4 + (c[1] & 1) * 2 + (c[0] & 1);
      d->node_p0[key[p] + 1]++;
    }
    for (int64_t n = 0; n < nn; n++) d->node_p0[n + 1] += d->node_p0[n];
    int64_t *fill = (int64_t *)xm((size_t)nn * 8);
    d->plist = (uint32_t *)xm((size_t)N * 4);
    for (int64_t p = 0; p < N; p++) d->plist[d->node_p0[key[p]] + fill[key[p]]++] = (uint32_t)p;
    free(fill);
    free(key);
  }
  deposit_q(d, px, py, pz, 1);

  int anomaly = 0;
  for (;;) {
    lev_t *c = &lev[nlev - 1];
    const int64_t L = c->v.L, nn = 8 * c->v.n_oct;
    c->moved = (uint8_t *)xm((size_t)c->node_p0[nn]);
    if (L == s.h.L_max) break;
    /* flags (test_node) */
    c->flags = (uint8_t *)xm((size_t)nn);
    for (int64_t n = 0; n < nn; n++)
      c->flags[n] = ahf_node_flags(&c->v, c->pmask, c->dens, c->crit - 1.0, n / 8, (int)(n % 8));
    /* refine rows (tested planes/rows only) */
    int64_t nrank = 4 * c->v.n_row;
    int64_t *cnt = (int64_t *)xm((size_t)(nrank + 1) * 8);
    for (int64_t r = 0; r < c->v.n_row; r++)
      for (int cz = 0; cz < 2; cz++)
        for (int cy = 0; cy < 2; cy++) {
          if (!ahf_plane_tested(&c->v, c->v.row_plane[r], cz)) continue;
          if (!ahf_row_tested(&c->v, r, cy)) continue;
          cnt[ahf_row_rank(&c->v, r, cy, cz)] = ahf_refine_row(&c->v, c->flags, r, cy, cz, &anomaly);
        }
    int64_t tot = 0;
    for (int64_t k = 0; k < nrank; k++) tot += cnt[k];
    if (tot == 0) { free(cnt); break; }
    /* compaction: rows in rank order */
    lev_t *f = &lev[nlev];
    memset(f, 0, sizeof(*f));
    f->oct_key = (uint64_t *)xm((size_t)tot * 8);
    f->oct_link = (uint8_t *)xm((size_t)tot);
    f->oct_parent = (int64_t *)xm((size_t)tot * 8);
    f->row_key = (uint64_t *)xm((size_t)nrank * 8);
    f->row_oct0 = (int32_t *)xm((size_t)(nrank + 1) * 4);
    f->row_plane = (int32_t *)xm((size_t)nrank * 4);
    f->plane_z = (int32_t *)xm((size_t)(2 * c->v.n_plane) * 4);
    f->plane_row0 = (int32_t *)xm((size_t)(2 * c->v.n_plane + 1) * 4);
    c->node2oct = (int32_t *)xm((size_t)nn * 4);
    for (int64_t n = 0; n < nn; n++) c->node2oct[n] = -1;
    int64_t no2 = 0, nr2 = 0, np2 = 0;
    for (int64_t p = 0; p < c->v.n_plane; p++)
      for (int cz = 0; cz < 2; cz++) {
        int64_t Z = 2 * c->v.plane_z[p] + cz;
        int plane_open = 0;
        for (int64_t r = c->v.plane_row0[p]; r < c->v.plane_row0[p + 1]; r++)
          for (int cy = 0; cy < 2; cy++) {
            if (!cnt[ahf_row_rank(&c->v, r, cy, cz)]) continue;
            int64_t Y = 2 * ahf_row_y(c->v.row_key[r]) + cy;
            if (!plane_open) { f->plane_z[np2] = (int32_t)Z; f->plane_row0[np2] = (int32_t)nr2; np2++; plane_open = 1; }
            f->row_key[nr2] = (uint64_t)Y | ((uint64_t)Z << 21);
            f->row_oct0[nr2] = (int32_t)no2;
            f->row_plane[nr2] = (int32_t)(np2 - 1);
            nr2++;
            for (int64_t o = c->v.row_oct0[r]; o < c->v.row_oct0[r + 1]; o++)
              for (int cx = 0; cx < 2; cx++) {
                int64_t n = 8 * o + cz * 4 + cy * 2 + cx;
                if (!(c->flags[n] & NF_REFINED)) continue;
                int64_t X = 2 * ahf_kx(c->v.oct_key[o]) + cx;
                f->oct_key[no2] = ahf_pack3((uint64_t)X, (uint64_t)Y, (uint64_t)Z);
                f->oct_link[no2] = (c->flags[n] & NF_LINK) ? 1 : 0;
                f->oct_parent[no2] = n;
                c->node2oct[n] = (int32_t)no2;
                no2++;
              }
          }
      }
    f->row_oct0[nr2] = (int32_t)no2;
    f->plane_row0[np2] = (int32_t)nr2;
    free(cnt);
    /* neighbour table */
    f->nbr = (int32_t *)xm((size_t)no2 * 27 * 4);
    for (int64_t o2 = 0; o2 < no2; o2++) {
      int64_t pn = f->oct_parent[o2];
      for (int ez = -1; ez <= 1; ez++)
        for (int ey = -1; ey <= 1; ey++)
          for (int ex = -1; ex <= 1; ex++) {
            int64_t m = ahf_nbr_node(c->v.nbr, pn / 8, (int)(pn % 8), ex, ey, ez);
            f->nbr[27 * o2 + ahf_e27(ex, ey, ez)] = m < 0 ? -1 : c->node2oct[m];
          }
    }
    set_view(f, 2 * L, no2, nr2, np2);
    f->m2d = c->m2d * 8.0;
    f->m2pd = c->m2pd * 8.0;
    f->crit = s.h.Nth_ref * f->m2d;
    /* presence masks of the fine level (needed by relink) */
    int64_t nnf = 8 * no2;
    f->pmask = (uint32_t *)xm((size_t)nnf * 4);
    for (int64_t n = 0; n < nnf; n++) f->pmask[n] = ahf_presence_mask(&f->v, n / 8, (int)(n % 8));
    /* relink: counts then stable fill, mothers in node order */
    f->node_p0 = (int64_t *)xm((size_t)(nnf + 1) * 8);
    int8_t *child = (int8_t *)xm((size_t)c->node_p0[nn]);
    for (int64_t m = 0; m < nn; m++) {
      int32_t o2 = c->node2oct[m];
      if (o2 < 0) continue;
      uint8_t imask = 0;
      for (int k = 0; k < 8; k++)
        if (f->pmask[8 * (int64_t)o2 + k] == AHF_FULL27) imask |= (uint8_t)(1u << k);
      uint64_t mk = c->v.oct_key[m / 8];
      int cm = (int)(m % 8);
      int64_t MX = 2 * ahf_kx(mk) + (cm & 1), MY = 2 * ahf_ky(mk) + ((cm >> 1) & 1), MZ = 2 * ahf_kz(mk) + (cm >> 2);
      for (int64_t q = c->node_p0[m]; q < c->node_p0[m + 1]; q++) {
        uint32_t p = c->plist[q];
        int ch = ahf_relink_child(px[p], py[p], pz[p], MX, MY, MZ, (double)(2 * L), imask);
        child[q] = (int8_t)ch;
        if (ch >= 0) { f->node_p0[8 * (int64_t)o2 + ch + 1]++; c->moved[q] = 1; }
      }
    }
    for (int64_t n = 0; n < nnf; n++) f->node_p0[n + 1] += f->node_p0[n];
    f->plist = (uint32_t *)xm((size_t)f->node_p0[nnf] * 4);
    int64_t *fill = (int64_t *)xm((size_t)nnf * 8);
    for (int64_t m = 0; m < nn; m++) {
      int32_t o2 = c->node2oct[m];
      if (o2 < 0) continue;
      for (int64_t q = c->node_p0[m]; q < c->node_p0[m + 1]; q++) {
        if (child[q] < 0) continue;
        int64_t fn = 8 * (int64_t)o2 + child[q];
        f->plist[f->node_p0[fn] + fill[fn]++] = c->plist[q];
      }
    }
    free(fill);
    free(child);
    /* deposit on the fine level: list order flips each level */
    deposit_q(f, px, py, pz, (nlev % 2) == 0);
    if (nnf < AHF_MIN_NNODES) {
      /* relink_back: keep the fine lists to append them to the mothers */
      c->rolled_back_child = 1;
      c->fine_n_oct = no2;
      c->fine_p0 = f->node_p0;
      c->fine_list = f->plist;
      break;
    }
    nlev++;
  }

  /* final per-node lists and dump */
  host_level *hl = (host_level *)xm(sizeof(host_level) * (size_t)nlev);
  for (int i = 0; i < nlev; i++) {
    lev_t *c = &lev[i];
    int64_t nn = 8 * c->v.n_oct;
    host_level *h = &hl[i];
    h->L = c->v.L;
    h->n_oct = c->v.n_oct;
    h->critdens = c->crit;
    h->masstopartdens = c->m2pd;
    h->size_no_nodes = (uint64_t)nn;
    h->size_no_part = (uint64_t)c->node_p0[nn];
    h->oct_key = c->oct_key;
    h->oct_link = c->oct_link;
    h->dens = c->dens;
    h->list0 = (int64_t *)xm((size_t)(nn + 1) * 8);
    h->list = (uint32_t *)xm((size_t)c->node_p0[nn] * 4 + 4);
    int desc = (i % 2) == 0, fine_desc = ((i + 1) % 2) == 0;
    int64_t w = 0;
    for (int64_t n = 0; n < nn; n++) {
      h->list0[n] = w;
      int64_t p0 = c->node_p0[n], p1 = c->node_p0[n + 1];
      for (int64_t k = 0; k < p1 - p0; k++) {
        int64_t q = desc ? p1 - 1 - k : p0 + k;
        if (!c->moved || !c->moved[q] || c->rolled_back_child) h->list[w++] = c->plist[q];
      }
      if (c->rolled_back_child && c->node2oct && c->node2oct[n] >= 0) {
        /* moved particles were appended back child by child: undo the
         * removal above by dropping them first, then append in fine order */
        w = h->list0[n];
        for (int64_t k = 0; k < p1 - p0; k++) {
          int64_t q = desc ? p1 - 1 - k : p0 + k;
          if (!c->moved[q]) h->list[w++] = c->plist[q];
        }
        for (int ch = 0; ch < 8; ch++) {
          int64_t fn = 8 * (int64_t)c->node2oct[n] + ch;
          int64_t f0 = c->fine_p0[fn], f1 = c->fine_p0[fn + 1];
          for (int64_t k = 0; k < f1 - f0; k++) h->list[w++] = c->fine_list[fine_desc ? f1 - 1 - k : f0 + k];
        }
      }
    }
    h->list0[nn] = w;
  }
  free(px);
  free(py);
  free(pz);
  *nlev_out = nlev;
  *hl_out = hl;
  return anomaly;
}

/* everything emu_build allocated */
static void emu_free(lev_t *lev, int nlev, host_level *hl)
{
  for (int i = 0; i < nlev; i++) {
    free(hl[i].list0);
    free(hl[i].list);
  }
  free(hl);
  for (int i = 0; i < EMU_MAX_LEVELS; i++) {
    lev_t *l = &lev[i];
    free(l->oct_key); free(l->oct_link); free(l->nbr); free(l->row_key); free(l->row_oct0); free(l->row_plane);
    free(l->plane_z); free(l->plane_row0); free(l->oct_parent); free(l->pmask); free(l->dens); free(l->flags);
    free(l->node_p0); free(l->plist); free(l->moved); free(l->node2oct);
    memset(l, 0, sizeof(*l)); /* fine_p0, fine_list: the next level's node_p0, plist */
  }
}

#ifndef AMR_EMU_NO_MAIN
int main(int argc, char **argv)
{
  if (argc < 3) { fprintf(stderr, "usage: amr_emu snap out\n"); return 2; }
  snap_data s;
  if (snap_read(argv[1], &s)) { fprintf(stderr, "emu: bad snapshot\n"); return 2; }
  static lev_t lev[EMU_MAX_LEVELS];
  int nlev = 0;
  host_level *hl = NULL;
  const int anomaly = emu_build(&s, lev, &nlev, &hl);
  int rc = hierarchy_dump(argv[2], hl, nlev);
  if (!rc && argc >= 7) {
    rc = gridinfo_replay(lev, nlev, &s, argv[3], atof(argv[4]), atof(argv[5]), atof(argv[6]));
    if (rc == 4) return 4;
  }
  printf("{\"levels\": %d, \"anomaly\": %d}\n", nlev, anomaly);
  return rc ? 1 : (anomaly ? 3 : 0);
}
#endif
