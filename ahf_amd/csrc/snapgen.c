/*
 * snapgen.c -- seeded synthetic snapshots for the deposit + unbinding path.
 *
 * Writes the project's snapshot container (snapfile.h) holding
 * particles in AHF's internal units (positions in box units, velocities in
 * units where v * v_fac = peculiar km/s, v_fac = boxsize/t_unit/a with
 * t_unit = 1/H0, src/startrun.c:547 and src/libahf/ahf_halos.c:201) and the
 * halo list that sort_halo_particles()/rem_unbound() receive.
 *
 * Workloads (all sizes scale with `side`, the cube root of the particle
 * count; the defaults are the full sizes):
 *   config 0: 128^3 particles, 50 Mpc/h, 40 NFW halos, M log-uniform in
 *             [1e11,1e13] Msun/h, c ~ U(5,15), Maxwellian velocities at the
 *             virial dispersion, 10% of each halo's particles at 2 v_esc;
 *             uniform background with sigma_v = 50 km/s.  LgridDomain 64.
 *   config 1: 256^3, 100 Mpc/h, 2000 hosts with dN/dM ~ M^-1.9 on
 *             [1e10,1e14]; hosts above 1e12 carry 5-20 NFW subhaloes holding
 *             10% of the host mass; 10% interlopers.  LgridDomain 128.
 *   config 2: 100000 isolated NFW haloes with N ~ N^-2 on [20, 2e6] particles,
 *             c ~ U(4,20), 15% of particles above escape speed and 5% within
 *             1e-4 (relative) of VescTune * v_esc.  No background, no grid.
 *   config 3: 512^3, 300 Mpc/h, 70/30 background/haloes, 20000 hosts from the
 *             config-1 mass function.  LgridDomain 256, LgridMax
 *             LgridDomain * 2^10 (up to 10 refinement levels).
 *   config 4: 896^3, 300 Mpc/h, 200000 hosts from the config-1 mass
 *             function; hosts above 1e12 carry 5-20 subhaloes holding 10% of
 *             the host mass, and every subhalo again carries subhaloes
 *             holding 10% of its own mass, three nested levels deep (3-8 at
 *             the second level, 2-4 at the third, where each would hold at
 *             least NminPerHalo particles); 10% interlopers.  A halo's list
 *             holds its own particles and those of all its subhaloes.
 *             LgridDomain 448, LgridMax LgridDomain * 2^12 (12 levels).
 * All use Omega_m = 0.3, Omega_L = 0.7, a = 1, VescTune 1.5, NminPerHalo 20,
 * NperDomCell = NperRefCell = 5, and unless stated LgridMax = LgridDomain *
 * 2^8.
 *
 * Randomness is a counter-based splitmix64 stream per (seed, object), so the
 * file is byte-identical for a given (config, seed, side) whatever the
 * thread count.  Particles are finally ordered by their 21-bit Hilbert key,
 * as AHF orders them (src/main.c:341-356); equal keys keep the generation
 * order.  Each halo gets AHF's gather radius (src/libahf/ahf_halos.c:
 * 2987-3050): half the distance to the nearest halo with more particles,
 * at least its virial radius, at most MaxGatherRad = 3 Mpc/h (and 1/4 box).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "snapfile.h"
#include "sfc_core.h"

#define RHOC0 2.7755397e11   /* h^2 Msun / Mpc^3, src/param.h:38 */
#define GRAV 4.3006485e-9    /* Mpc km^2 / (Msun s^2), src/param.h:39 */
#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------ random ---- */

static inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct { uint64_t key, ctr; } rng_t;

static inline rng_t rng_make(uint64_t seed, uint64_t stream)
{
  rng_t r = {mix64(seed * 0x632BE59BD9B4E019ull + mix64(stream)), 0};
  return r;
}
static inline uint64_t rng_u64(rng_t *r) { return mix64(r->key + 0x9E3779B97F4A7C15ull * ++r->ctr); }
static inline double rng_uniform(rng_t *r) { return (double)(rng_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline double rng_gauss(rng_t *r)
{
  double u1 = rng_uniform(r), u2 = rng_uniform(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * PI_D * u2);
}
static inline void rng_dir(rng_t *r, double v[3])
{
  double ct = 2.0 * rng_uniform(r) - 1.0, ph = 2.0 * PI_D * rng_uniform(r);
  double st = sqrt(fmax(0.0, 1.0 - ct * ct));
  v[0] = st * cos(ph); v[1] = st * sin(ph); v[2] = ct;
}

/* --------------------------------------------------------------- NFW ---- */

typedef struct {
  double M;          /* Msun/h */
  double c;
  double cen[3];     /* Mpc/h */
  double vbulk[3];   /* km/s */
  int64_t n;         /* particles of this halo (incl. interlopers) */
  int64_t first;     /* first particle slot */
  double frac_unb;   /* fraction boosted above escape */
  double frac_edge;  /* fraction placed at the escape threshold */
  int host;          /* index of host halo, -1 for hosts */
} halo_spec;

static inline double nfw_mu(double x) { return log1p(x) - x / (1.0 + x); }

static double r_vir_of(double M) { return cbrt(3.0 * M / (4.0 * PI_D * 200.0 * RHOC0)); }

/* radius (Mpc/h) with M(<r)/M(<Rvir) = u */
static double nfw_radius(double u, double c, double rs)
{
  double target = u * nfw_mu(c), lo = 0.0, hi = c;
  for (int it = 0; it < 60; it++) {
    double mid = 0.5 * (lo + hi);
    if (nfw_mu(mid) < target) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi) * rs;
}

/* escape speed (km/s) of an untruncated NFW halo at radius r */
static double nfw_vesc(double r, double M, double c, double rs)
{
  double x = fmax(r / rs, 1e-6);
  double phi = GRAV * M / nfw_mu(c) * log1p(x) / (x * rs);
  return sqrt(2.0 * phi);
}

/* -------------------------------------------------------------- build ---- */

typedef struct {
  int config;
  int64_t npart;
  double box, pmass, vtune;
  int L_dom, L_max;
  double Nth_dom, Nth_ref;
  int isolated;
  double sigma_bg;
} gen_params;

static int64_t n_of_mass(double M, double mp) { return (int64_t)floor(M / mp + 0.5); }

/* draw from p(M) ~ M^-alpha on [m1, m2] */
static double powerlaw(rng_t *r, double alpha, double m1, double m2)
{
  double a = 1.0 - alpha, u = rng_uniform(r);
  return pow(pow(m1, a) + u * (pow(m2, a) - pow(m1, a)), 1.0 / a);
}

static void place_halo(halo_spec *h, const gen_params *g, rng_t *r)
{
  for (int d = 0; d < 3; d++) {
    h->cen[d] = (g->isolated ? 0.1 + 0.8 * rng_uniform(r) : rng_uniform(r)) * g->box;
    h->vbulk[d] = 150.0 * rng_gauss(r);
  }
}

/* room for one more halo spec */
static halo_spec *halo_push(halo_spec **hs, int64_t *nh, int64_t *cap)
{
  if (*nh == *cap) {
    const int64_t nc = *cap * 2 + 64;
    halo_spec *p = (halo_spec *)realloc(*hs, (size_t)nc * sizeof(halo_spec));
    if (!p) return NULL;
    memset(p + *cap, 0, (size_t)(nc - *cap) * sizeof(halo_spec));
    *hs = p;
    *cap = nc;
  }
  return &(*hs)[(*nh)++];
}

/*
 * ns subhaloes of halo `parent` (an index: the array may move) holding 10%
 * of its mass, placed inside its NFW profile (mass fraction 0.1-0.9) with
 * its bulk velocity plus its virial dispersion; the parent keeps 90%.
 * depth > 1 gives each new subhalo subhaloes of its own (nested levels).
 */
static int add_subs(const gen_params *g, rng_t *r, halo_spec **hs, int64_t *nh, int64_t *cap, int64_t parent, int ns,
                    int depth)
{
  static const int lo_n[4] = {0, 5, 3, 2}, span_n[4] = {0, 16, 6, 3};
  double w[20], wsum = 0.0;
  for (int k = 0; k < ns; k++) { w[k] = 0.2 + rng_uniform(r); wsum += w[k]; }
  const double Mp = (*hs)[parent].M, cp = (*hs)[parent].c, rv = r_vir_of(Mp), rs = rv / cp;
  const double sig = sqrt(GRAV * Mp / (2.0 * rv));
  (*hs)[parent].M = 0.9 * Mp;
  const int64_t first = *nh;
  for (int k = 0; k < ns; k++) {
    halo_spec *s = halo_push(hs, nh, cap);
    if (!s) return -1;
    const halo_spec *P = &(*hs)[parent];
    s->M = 0.1 * Mp * w[k] / wsum;
    s->c = 5.0 + 10.0 * rng_uniform(r);
    s->frac_unb = 0.10;
    s->host = (int)parent;
    double rr = nfw_radius(0.1 + 0.8 * rng_uniform(r), P->c, rs), dir[3];
    rng_dir(r, dir);
    for (int d = 0; d < 3; d++) {
      s->cen[d] = P->cen[d] + rr * dir[d];
      s->vbulk[d] = P->vbulk[d] + sig * rng_gauss(r);
    }
  }
  if (depth > 1) {
    const int level = 4 - depth + 1; /* nesting level of the grandchildren: 2 or 3 */
    for (int64_t i = first; i < first + ns; i++) {
      int nc = lo_n[level] + (int)(rng_uniform(r) * (double)span_n[level]);
      if (nc > lo_n[level] + span_n[level] - 1) nc = lo_n[level] + span_n[level] - 1;
      /* only where every grandchild would hold NminPerHalo particles */
      while (nc > 0 && 0.1 * (*hs)[i].M / (double)nc < 20.0 * g->pmass) nc--;
      if (nc > 0 && add_subs(g, r, hs, nh, cap, i, nc, depth - 1)) return -1;
    }
  }
  return 0;
}

/* hosts and their subhaloes (sub_depth nested levels below hosts above
 * 1e12 Msun/h: 0 none, 1 config 1, 3 config 4); returns the count, -1 when
 * out of memory */
static int64_t make_halos(const gen_params *g, uint64_t seed, int64_t nhost, int sub_depth,
                          double mlo, double mhi, int logflat, halo_spec **out)
{
  int64_t cap = nhost * 22 + 16, nh = 0;
  halo_spec *hs = (halo_spec *)calloc((size_t)cap, sizeof(halo_spec));
  if (!hs) return -1;
  rng_t r = rng_make(seed, 1);
  for (int64_t i = 0; i < nhost; i++) {
    halo_spec *h = halo_push(&hs, &nh, &cap);
    if (!h) { free(hs); return -1; }
    h->M = logflat ? exp(log(mlo) + rng_uniform(&r) * (log(mhi) - log(mlo)))
                   : powerlaw(&r, 1.9, mlo, mhi);
    h->c = 5.0 + 10.0 * rng_uniform(&r);
    h->frac_unb = 0.10;
    h->frac_edge = 0.0;
    h->host = -1;
    place_halo(h, g, &r);
    if (sub_depth > 0 && h->M > 1e12) {
      int ns = 5 + (int)(rng_uniform(&r) * 16.0);
      if (ns > 20) ns = 20;
      if (add_subs(g, &r, &hs, &nh, &cap, nh - 1, ns, sub_depth)) { free(hs); return -1; }
    }
  }
  *out = hs;
  return nh;
}

static int cmp_u64pair(const void *a, const void *b)
{
  const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

/*
 * Sort n (key, index) pairs by key, then index (the same total order as one
 * qsort with cmp_u64pair): a parallel scatter into 2^16 buckets by the top
 * 16 of the 63 key bits, then each bucket sorted on its own.  Returns -1
 * when out of memory.
 */
static int sort_pairs(uint64_t *kv, int64_t n)
{
  enum { BB = 16, NB = 1 << BB };
  const int shift = 3 * AHF_SFC_BITS - BB;
  uint64_t *tmp = (uint64_t *)malloc((size_t)(n ? n : 1) * 2 * sizeof(uint64_t));
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  int64_t *cnt = (int64_t *)calloc((size_t)nt * NB + NB + 1, sizeof(int64_t));
  if (!tmp || !cnt) { free(tmp); free(cnt); return -1; }
  int64_t *boff = cnt + (size_t)nt * NB;
  const int64_t chunk = (n + nt - 1) / nt;
#pragma omp parallel num_threads(nt)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    const int64_t a = t * chunk, b = a + chunk < n ? a + chunk : n;
    int64_t *c = cnt + (size_t)t * NB;
    for (int64_t i = a; i < b; i++) c[kv[2 * i] >> shift]++;
#pragma omp barrier
#pragma omp single
    {
      int64_t acc = 0;
      for (int64_t k = 0; k < NB; k++) {
        boff[k] = acc;
        for (int u = 0; u < nt; u++) {
          const int64_t v = cnt[(size_t)u * NB + k];
          cnt[(size_t)u * NB + k] = acc;
          acc += v;
        }
      }
      boff[NB] = acc;
    }
    for (int64_t i = a; i < b; i++) {
      const int64_t q = c[kv[2 * i] >> shift]++;
      tmp[2 * q] = kv[2 * i];
      tmp[2 * q + 1] = kv[2 * i + 1];
    }
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < NB; k++)
    if (boff[k + 1] - boff[k] > 1)
      qsort(tmp + 2 * boff[k], (size_t)(boff[k + 1] - boff[k]), 2 * sizeof(uint64_t), cmp_u64pair);
  memcpy(kv, tmp, (size_t)n * 2 * sizeof(uint64_t));
  free(tmp);
  free(cnt);
  return 0;
}


static inline float to_unit(double x_mpc, double box, int periodic)
{
  double u = x_mpc / box;
  if (periodic) u -= floor(u);
  float f = (float)u;
  if (f >= 1.0f) f = 0.0f;
  if (f < 0.0f) f = 0.0f;
  return f;
}

/*
 * AHF's gather radius per halo (src/libahf/ahf_halos.c:2987-3050) in box
 * units: half the periodic distance to the nearest halo with more
 * particles, or maxGathRad when there is none; then at least the virial
 * radius and at most maxGathRad = min(MaxGatherRad/box, 1/4) with
 * MaxGatherRad = 3 Mpc/h.  Only haloes within 2 maxGathRad can set a
 * radius below the cap, so a cell grid of that width finds them.
 */
static double *gather_radii(const gen_params *g, const halo_spec *hs, const int64_t *np, const double *cen, int64_t nh)
{
  double *grad = (double *)malloc((size_t)(nh ? nh : 1) * sizeof(double));
  if (!grad) return NULL;
  const double gmax = fmin(3.0 / g->box, 0.25);
  int nc = (int)floor(1.0 / (2.0 * gmax));
  if (nc < 3) nc = 1; /* fewer than 3 cells a side: one cell, the whole box */
  const int64_t ncell = (int64_t)nc * nc * nc;
  int64_t *head = (int64_t *)malloc((size_t)ncell * sizeof(int64_t));
  int64_t *next = (int64_t *)malloc((size_t)(nh ? nh : 1) * sizeof(int64_t));
  if (!head || !next) { free(grad); free(head); free(next); return NULL; }
  for (int64_t c = 0; c < ncell; c++) head[c] = -1;
  int *cc = (int *)malloc((size_t)(nh ? nh : 1) * 3 * sizeof(int));
  if (!cc) { free(grad); free(head); free(next); return NULL; }
  for (int64_t i = 0; i < nh; i++) {
    for (int d = 0; d < 3; d++) {
      int c = (int)floor(cen[3 * i + d] * nc);
      cc[3 * i + d] = c < 0 ? 0 : c >= nc ? nc - 1 : c;
    }
    const int64_t c = ((int64_t)cc[3 * i + 2] * nc + cc[3 * i + 1]) * nc + cc[3 * i];
    next[i] = head[c];
    head[c] = i;
  }
  const int reach = nc >= 3 ? 1 : 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nh; i++) {
    double best = 1e300;
    int found = 0;
    for (int a = -reach; a <= reach; a++)
      for (int b = -reach; b <= reach; b++)
        for (int e = -reach; e <= reach; e++) {
          const int x = (cc[3 * i] + e + nc) % nc, y = (cc[3 * i + 1] + b + nc) % nc,
                    z = (cc[3 * i + 2] + a + nc) % nc;
          for (int64_t j = head[((int64_t)z * nc + y) * nc + x]; j >= 0; j = next[j]) {
            if (j == i || np[j] <= np[i]) continue;
            double d2 = 0.0;
            for (int d = 0; d < 3; d++) {
              double dd = fabs(cen[3 * i + d] - cen[3 * j + d]);
              if (dd > 0.5) dd = 1.0 - dd;
              d2 += dd * dd;
            }
            if (d2 < best) best = d2;
            found = 1;
          }
        }
    double r = found ? 0.5 * sqrt(best) : gmax;
    const double rv = r_vir_of(hs[i].M) / g->box;
    if (r < rv) r = rv;
    if (r > gmax) r = gmax;
    grad[i] = r;
  }
  free(head); free(next); free(cc);
  return grad;
}

/* the snapshot of the haloes hs[0..nh) plus n_bg background particles, in
 * memory (arrays malloc'd into *out; release with snap_free) */
static int build_snap(const gen_params *g, uint64_t seed, halo_spec *hs, int64_t nh, int64_t n_bg, snap_data *out)
{
  const double a = 1.0, v_fac = g->box / 0.01 / a;
  int64_t nhalo_part = 0;
  for (int64_t i = 0; i < nh; i++) { hs[i].first = nhalo_part; nhalo_part += hs[i].n; }
  int64_t N = nhalo_part + n_bg;
  float *pos = (float *)malloc((size_t)N * 3 * sizeof(float));
  float *mom = (float *)malloc((size_t)N * 3 * sizeof(float));
  if (!pos || !mom) return -1;

  /* haloes */
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < nh; i++) {
    halo_spec *h = &hs[i];
    rng_t r = rng_make(seed, 1000 + (uint64_t)i);
    double rv = r_vir_of(h->M), rs = rv / h->c, sig = sqrt(GRAV * h->M / (2.0 * rv));
    int64_t n_unb = (int64_t)floor(h->frac_unb * (double)h->n + 0.5);
    int64_t n_edge = (int64_t)floor(h->frac_edge * (double)h->n + 0.5);
    for (int64_t j = 0; j < h->n; j++) {
      double rr = nfw_radius(rng_uniform(&r), h->c, rs), dir[3], v[3];
      rng_dir(&r, dir);
      if (j < n_unb || j < n_unb + n_edge) {
        double vesc = nfw_vesc(rr, h->M, h->c, rs);
        double speed = (j < n_unb) ? 2.0 * vesc : g->vtune * vesc * (1.0 + 1e-4 * (2.0 * rng_uniform(&r) - 1.0));
        double vd[3];
        rng_dir(&r, vd);
        for (int d = 0; d < 3; d++) v[d] = speed * vd[d];
      } else {
        for (int d = 0; d < 3; d++) v[d] = sig * rng_gauss(&r);
      }
      int64_t p = h->first + j;
      for (int d = 0; d < 3; d++) {
        pos[3 * p + d] = to_unit(h->cen[d] + rr * dir[d], g->box, 1);
        mom[3 * p + d] = (float)((h->vbulk[d] + v[d]) / v_fac);
      }
    }
  }
  /* uniform background in blocks of 65536 */
  int64_t nblk = (n_bg + 65535) / 65536;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nblk; b++) {
    rng_t r = rng_make(seed, 0x100000000ull + (uint64_t)b);
    int64_t j1 = (b + 1) * 65536 < n_bg ? (b + 1) * 65536 : n_bg;
    for (int64_t j = b * 65536; j < j1; j++) {
      int64_t p = nhalo_part + j;
      for (int d = 0; d < 3; d++) {
        pos[3 * p + d] = to_unit(rng_uniform(&r) * g->box, g->box, 1);
        mom[3 * p + d] = (float)(g->sigma_bg * rng_gauss(&r) / v_fac);
      }
    }
  }

  /* Hilbert order */
  uint64_t *kv = (uint64_t *)malloc((size_t)N * 2 * sizeof(uint64_t));
  if (!kv) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < N; p++) {
    kv[2 * p] = ahf_sfc_key((double)pos[3 * p + 0], (double)pos[3 * p + 1], (double)pos[3 * p + 2], AHF_SFC_BITS);
    kv[2 * p + 1] = (uint64_t)p;
  }
  if (sort_pairs(kv, N)) return -1;
  uint64_t *newidx = (uint64_t *)malloc((size_t)N * sizeof(uint64_t));
  float *pos2 = (float *)malloc((size_t)N * 3 * sizeof(float));
  float *mom2 = (float *)malloc((size_t)N * 3 * sizeof(float));
  uint64_t *id = (uint64_t *)malloc((size_t)N * sizeof(uint64_t));
  if (!newidx || !pos2 || !mom2 || !id) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < N; q++) {
    uint64_t p = kv[2 * q + 1];
    newidx[p] = (uint64_t)q;
    memcpy(pos2 + 3 * q, pos + 3 * p, 3 * sizeof(float));
    memcpy(mom2 + 3 * q, mom + 3 * p, 3 * sizeof(float));
    id[q] = p;
  }
  free(kv); free(pos); free(mom);

  /* halo lists: a halo's list holds its own particles and those of every
   * subhalo nested in it (a subhalo always comes after its parent) */
  int64_t n_ip = 0;
  int64_t *hostn = (int64_t *)calloc((size_t)(nh ? nh : 1), sizeof(int64_t));
  if (!hostn) return -1;
  for (int64_t i = 0; i < nh; i++) hostn[i] = hs[i].n;
  for (int64_t i = nh - 1; i >= 0; i--)
    if (hs[i].host >= 0) hostn[hs[i].host] += hostn[i];
  for (int64_t i = 0; i < nh; i++) n_ip += hostn[i];
  uint64_t *ipart = (uint64_t *)malloc((size_t)(n_ip ? n_ip : 1) * sizeof(uint64_t));
  uint64_t *hnp = (uint64_t *)malloc((size_t)(nh ? nh : 1) * sizeof(uint64_t));
  double *cen = (double *)malloc((size_t)(nh ? nh : 1) * 3 * sizeof(double));
  int64_t *fill = (int64_t *)calloc((size_t)(nh ? nh : 1), sizeof(int64_t));
  int64_t off = 0;
  for (int64_t i = 0; i < nh; i++) { fill[i] = off; off += hostn[i]; }
  for (int64_t i = 0; i < nh; i++)
    for (int64_t a = i; a >= 0; a = hs[a].host)
      for (int64_t j = 0; j < hs[i].n; j++) ipart[fill[a]++] = newidx[hs[i].first + j];
  for (int64_t i = 0; i < nh; i++) {
    hnp[i] = (uint64_t)hostn[i];
    for (int d = 0; d < 3; d++) {
      double u = hs[i].cen[d] / g->box;
      if (!g->isolated) u -= floor(u);
      cen[3 * i + d] = u;
    }
  }
  double *grad = gather_radii(g, hs, hostn, cen, nh);
  if (!grad) return -1;

  snap_header hd;
  memset(&hd, 0, sizeof(hd));
  memcpy(hd.magic, SNAP_MAGIC, 8);
  hd.npart = (uint64_t)N;
  hd.boxsize = g->box;
  hd.omega0 = 0.3;
  hd.lambda0 = 0.7;
  hd.a = a;
  hd.pmass = g->pmass;
  hd.L_dom = g->L_dom;
  hd.L_max = g->L_max;
  hd.Nth_dom = g->Nth_dom;
  hd.Nth_ref = g->Nth_ref;
  hd.minpart = 20;
  hd.vtune = g->vtune;
  hd.r_fac = g->box * a;
  hd.v_fac = v_fac;
  hd.phi_fac = GRAV * g->pmass / (g->box * a);
  hd.hubble = 100.0 * sqrt(0.7 * (1.0 - 1.0 / (a * a)) + 0.3 * (1.0 / (a * a * a) - 1.0 / (a * a)) + 1.0 / (a * a));
  /* haloes are sized at 200 times the critical density, so rem_outsideRvir
   * cuts near the generated virial radius: Delta = 200 relative to
   * a^3 rho_crit(a), the comoving form of src/libahf/ahf_halos.c:204-221 */
  hd.ovlim = 200.0;
  hd.rho_vir = RHOC0 * a * a * a * (hd.hubble / 100.0) * (hd.hubble / 100.0);
  hd.nhalos = (uint64_t)nh;
  hd.n_ipart = (uint64_t)n_ip;
  hd.seed = seed;
  hd.config = g->config;

  out->h = hd;
  out->pos = pos2;
  out->mom = mom2;
  out->id = id;
  out->centre = cen;
  out->gather_rad = grad;
  out->halo_npart = hnp;
  out->ipart = ipart;
  free(newidx); free(hostn); free(fill);
  return 0;
}

static int write_snap(const char *path, const snap_data *d)
{
  const snap_header *hd = &d->h;
  const size_t N = (size_t)hd->npart, nh = (size_t)hd->nhalos, n_ip = (size_t)hd->n_ipart;
  FILE *f = fopen(path, "wb");
  int bad = !f;
  if (!bad) {
    bad |= fwrite(hd, sizeof(*hd), 1, f) != 1;
    bad |= fwrite(d->pos, sizeof(float), N * 3, f) != N * 3;
    bad |= fwrite(d->mom, sizeof(float), N * 3, f) != N * 3;
    bad |= fwrite(d->id, sizeof(uint64_t), N, f) != N;
    bad |= fwrite(d->centre, sizeof(double), nh * 3, f) != nh * 3;
    bad |= fwrite(d->gather_rad, sizeof(double), nh, f) != nh;
    bad |= fwrite(d->halo_npart, sizeof(uint64_t), nh, f) != nh;
    bad |= fwrite(d->ipart, sizeof(uint64_t), n_ip, f) != n_ip;
    bad |= fclose(f) != 0;
  }
  return bad ? -2 : 0;
}

/*
 * ahf_snapgen_make: generate workload `config` at `side`^3 particles
 * (side <= 0: the full size) with `seed`, scaling the halo count by
 * (side/side_full)^3, into *out (release with ahf_snapgen_release).
 * Returns 0 on success.
 */
int ahf_snapgen_make(int config, int side, uint64_t seed, snap_data *out)
{
  memset(out, 0, sizeof(*out));
  gen_params g;
  memset(&g, 0, sizeof(g));
  g.config = config;
  g.vtune = 1.5;
  g.Nth_dom = 5.0;
  g.Nth_ref = 5.0;
  g.sigma_bg = 50.0;
  int full_side = config == 0 ? 128 : config == 1 ? 256 : config == 3 ? 512 : config == 4 ? 896 : 0;
  int dom_full = config == 0 ? 64 : config == 1 ? 128 : config == 3 ? 256 : config == 4 ? 448 : 0;
  halo_spec *hs = NULL;
  int64_t nh = 0, n_bg = 0;
  if (config == 2) {
    g.isolated = 1;
    g.box = 100.0;
    g.pmass = 1e8;
    g.L_dom = 64;
    g.L_max = 64 << 8;
    int64_t nhalo = side > 0 ? (int64_t)side : 100000;
    hs = (halo_spec *)calloc((size_t)nhalo, sizeof(halo_spec));
    rng_t r = rng_make(seed, 2);
    for (int64_t i = 0; i < nhalo; i++) {
      halo_spec *h = &hs[i];
      double u = rng_uniform(&r);
      /* N ~ N^-2 on [20, 2e6] */
      double n = 1.0 / (1.0 / 20.0 - u * (1.0 / 20.0 - 1.0 / 2e6));
      h->n = (int64_t)floor(n);
      h->M = (double)h->n * g.pmass;
      h->c = 4.0 + 16.0 * rng_uniform(&r);
      h->frac_unb = 0.15;
      h->frac_edge = 0.05;
      h->host = -1;
      place_halo(h, &g, &r);
    }
    nh = nhalo;
    n_bg = 0;
  } else if (config == 0 || config == 1 || config == 3 || config == 4) {
    if (side <= 0) side = full_side;
    double scale = (double)side / (double)full_side;
    g.npart = (int64_t)side * side * side;
    g.box = (config == 0 ? 50.0 : config == 1 ? 100.0 : 300.0) * scale;
    g.pmass = 0.3 * RHOC0 * g.box * g.box * g.box / (double)g.npart;
    g.L_dom = (int)(dom_full * scale);
    if (g.L_dom < 4) g.L_dom = 4;
    g.L_max = g.L_dom << (config == 3 ? 10 : config == 4 ? 12 : 8);
    int64_t nhost = config == 0 ? 40 : config == 1 ? 2000 : config == 3 ? 20000 : 200000;
    nhost = (int64_t)ceil((double)nhost * scale * scale * scale);
    if (config == 0) nh = make_halos(&g, seed, nhost, 0, 1e11, 1e13, 1, &hs);
    else nh = make_halos(&g, seed, nhost, config == 1 ? 1 : config == 4 ? 3 : 0, 1e10, 1e14, 0, &hs);
    if (nh < 0) return -1;
    int64_t tot = 0;
    for (int64_t i = 0; i < nh; i++) { hs[i].n = n_of_mass(hs[i].M, g.pmass); tot += hs[i].n; }
    int64_t budget = config == 3 ? (int64_t)(0.3 * (double)g.npart) : g.npart;
    if (config == 3 && tot > 0) {
      /* rescale halo masses so haloes hold 30% of the particles */
      double f = (double)budget / (double)tot;
      tot = 0;
      for (int64_t i = 0; i < nh; i++) { hs[i].M *= f; hs[i].n = n_of_mass(hs[i].M, g.pmass); tot += hs[i].n; }
    }
    while (tot > g.npart && nh > 0) { tot -= hs[nh - 1].n; nh--; }
    n_bg = g.npart - tot;
  } else {
    return -10;
  }
  int rc = build_snap(&g, seed, hs, nh, n_bg, out);
  free(hs);
  return rc;
}

void ahf_snapgen_release(snap_data *d)
{
  snap_free(d);
}

/* ahf_snapgen_make, written to `path` */
int ahf_snapgen_write(int config, int side, uint64_t seed, const char *path)
{
  snap_data d;
  int rc = ahf_snapgen_make(config, side, seed, &d);
  if (rc == 0) rc = write_snap(path, &d);
  snap_free(&d);
  return rc;
}
