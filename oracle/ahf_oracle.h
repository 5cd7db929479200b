/*
 * ahf_oracle.h -- CPU restatement of AHF's particle->grid and unbinding hot
 * path.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU leg of
 * the benchmark may call it, as the checker; the product path never does.
 *
 * Every function restates the reference function named in its comment,
 * with the same floating-point operation order (FP32 density accumulated as
 * float += double, FP64 everywhere in the unbinding except the FP32 |v|^2 of
 * the bulk-velocity seed), compiled with -ffp-contract=off.
 */
#ifndef AHF_ORACLE_H
#define AHF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AMR hierarchy on a snapshot; writes the AHFGRID1 dump (same layout as
 * oracle/ref_driver.c).  Returns number of grids, <0 on error. */
int oracle_amr_dump(const char *snap_path, const char *out_path);

/* sort_halo_particles + rem_unbound for every halo of the snapshot; writes the
 * AHFHALO1 dump.  stable_ties != 0 orders equal distances by list position
 * instead of reproducing indexx's quicksort permutation.  threads <= 0: all. */
int oracle_unbind_dump(const char *snap_path, const char *out_path, int threads, int stable_ties);

/* lower-level pieces, used by the known-answer tests */

/* TSC weights of one coordinate (src/libamr_serial/density.c:356-360) */
void oracle_tsc_weights(double pn_sep, double w[3]);

/* indexx permutation (1-based, src/libutility/general.c:1093-1153) */
void oracle_indexx(uint64_t n, const double *arr1, uint64_t *indx1);

/* rem_unbound on one halo given by member indices (modified in place) */
typedef struct {
  int32_t minpart;
  double vtune, r_fac, v_fac, phi_fac, hubble;
} oracle_unbind_params;

typedef struct {
  uint64_t npart;     /* bound members after unbinding */
  double M_vir, R_vir, Phi0;
  int32_t niter;
  int32_t rvir_set;   /* 0: no particle was ever bound (R_vir undefined in AHF) */
} oracle_unbind_result;

/* rem_outsideRvir on one radially sorted member list: returns the members
 * kept; M_vir, R_vir, ovdens are written unless n < minpart */
typedef struct {
  int32_t minpart;
  double ovlim, rho_vir, rho_fac;
} oracle_rvir_params;

uint64_t oracle_rem_outsideRvir(const float *pos, const uint64_t *ipart, uint64_t n, const double centre[3],
                                const oracle_rvir_params *p, double *M_vir, double *R_vir, double *ovdens);

/* a halo's fields after the construction steps, as a zeroed HALO would hold
 * them; `set` says which fields some step wrote */
#define ORACLE_SET_MVIR 1
#define ORACLE_SET_RVIR 2
#define ORACLE_SET_PHI0 4
#define ORACLE_SET_OVDENS 8
typedef struct {
  uint64_t npart;
  double M_vir, R_vir, Phi0, ovdens;
  int32_t niter;
  int32_t set;
} oracle_halo_state;

void oracle_construct_halo(const float *pos, const float *mom, uint64_t *ipart, uint64_t n, const double centre[3],
                           const oracle_unbind_params *u, const oracle_rvir_params *r, int stable_ties,
                           oracle_halo_state *st);

/* sort + rem_outsideRvir(0) + rem_unbound + rem_outsideRvir(1) for every halo
 * of the snapshot; AHFHALO1 dump with ovdens as the fourth value */
int oracle_construct_dump(const char *snap_path, const char *out_path, int threads, int stable_ties);

/* Hilbert index of a grid cell (src/libsfc/hilbert.c:197) and the SFC key
 * of a position (src/libsfc/hilbert_util.c:68) */
uint64_t oracle_hilbert_c2i(const uint64_t c[3], unsigned bits);
uint64_t oracle_sfc_key(double x, double y, double z, unsigned bits);

/* ahf_halos_sfc_gatherParts for one halo over key-sorted particles: the
 * members are written to *out (realloc'ed), their count returned */
uint64_t oracle_gather_halo(const float *pos, const uint64_t *key, uint64_t n, const double centre[3],
                            double gather_rad, uint64_t **out, uint64_t *n_cand);

/* gatherParts (construct = 0), or gatherParts + sort + rem_outsideRvir(0) +
 * rem_unbound + rem_outsideRvir(1) (construct = 1), for every halo of the
 * snapshot from its centre and gather radius; AHFHALO1 dump.  -3 when the
 * particles are not in key order. */
int oracle_gather_dump(const char *snap_path, const char *out_path, int threads, int stable_ties, int construct,
                       uint64_t *n_cand);

void oracle_sort_halo(const float *pos, uint64_t n, uint64_t *ipart, const double centre[3],
                      int32_t minpart, int stable_ties);
void oracle_rem_unbound(const float *pos, const float *mom, uint64_t *ipart, uint64_t n,
                        const double centre[3], const oracle_unbind_params *p,
                        oracle_unbind_result *res);

/* HaloProfiles (src/libahf/ahf_halos.c:3961-5060) for AHF's default flags:
 * props[ORACLE_PROF_NP] and prof[nbins * ORACLE_PROF_NQ] of one constructed
 * halo (members in radial order); returns nbins, 0 below minpart.  The
 * order of the values is the enum of oracle_profiles.h. */
#define ORACLE_PROF_NQ 25
#define ORACLE_PROF_NP 45
typedef struct {
  int64_t minpart;
  double r_fac, v_fac, m_fac, phi_fac, hubble;
} oracle_prof_params;
int oracle_halo_profiles(const float *pos, const float *mom, const uint64_t *ipart, uint64_t n,
                         const double centre[3], double Phi0, double R_vir, const oracle_prof_params *p,
                         double *props, double *prof);
int oracle_profiles_dump(const char *snap_path, const char *out_path, int threads, int gather);

/* ahf_gridinfo (src/libahf/ahf_gridinfo.c:59-860) after the AMR build:
 * min_ref from ovlim and rho_vir (the comoving density AHF's overdensities
 * are relative to) and refine_mass (pmass * med_weight); writes the
 * AHFGINF1 dump of oracle/ref_driver.c's gridinfo command.  Returns the
 * number of AHF grids, or < 0. */
int oracle_gridinfo_dump(const char *snap_path, const char *out_path, double ovlim, double rho_vir,
                         double refine_mass);
int oracle_prof_nbins(uint64_t n);
void oracle_prof_get_axes(double it[9], double axes[3]);
double oracle_prof_cnfw(double V2_max, double V2_vir);
double oracle_prof_cr1(double R1);
void oracle_prof_find_max(const double *x, const double *y, int64_t n, int ismooth, double *x_max, double *y_max);

#ifdef __cplusplus
}
#endif
#endif
