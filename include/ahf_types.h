/*
 * ahf_types.h -- memory layouts of the AHF data structures that cross the
 * drop-in boundary, for AHF's default flag set:
 *
 *   PERIODIC, AHFlean, TSC, no DOUBLE, no MULTIMASS, no GAS_PARTICLES,
 *   no METALHACK, no WITH_MPI, no AHFrestart, no AHFnewHaloIDs,
 *   no AHFphspdens, no AHFexciseSubhaloStars, no WITH_HDF5, no AHF_LRSI.
 *
 * These are the layouts AHF's own structs have under those flags
 * (src/tdef.h:33-77 particle, :107-137 node, :144-183 quads, :189-234
 * gridlist, :240-292 info_global, :296-309 tline, :315-411 param_simu,
 * :547-549 XYZ, :586-720 HALOPROFILE, :723-828 HALO, :796-818 intXYZ, SRINDEX,
 * info_ahf).  A library built
 * against this header can be linked into an AHF built with the same flags;
 * tests/test_layout.py compares every size and offset below with a probe
 * compiled against the reference headers.
 *
 * Nothing here allocates or owns memory: AHF owns particles, grids and
 * halos, exactly as in the reference.
 */
#ifndef AHF_TYPES_H
#define AHF_TYPES_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHF_NDIM 3
#define AHF_MAXTIME 10000          /* src/param.h:70 */
#define AHF_MIN_NNODES 125         /* src/param.h:54 */
#define AHF_CRITMULTI 8.0          /* src/param.h:118 */
#define AHF_MACHINE_ZERO 5e-16     /* src/param.h:125 */

typedef char boolean;             /* src/tdef.h:20 */
typedef float flouble;            /* src/tdef.h:24 (no -DDOUBLE) */
typedef uint64_t part_id_t;       /* src/tdef.h:33 */
typedef uint64_t sfc_key_t;       /* src/libsfc/sfc_curve.h:40 */

/* particle: 48 bytes (ll 8, pos 12, mom 12, sfckey 8, id 8) */
typedef struct particle *partptr;
typedef struct particle {
  partptr ll;                     /* must stay first: linked-list link */
  flouble pos[AHF_NDIM];          /* box units, [0,1) */
  flouble mom[AHF_NDIM];          /* AHF velocity units (x v_fac = km/s) */
  sfc_key_t sfckey;
  part_id_t id;
} part;

typedef struct gas  { flouble u; } gas, *gasptr;
typedef struct star { flouble dummy; } star, *starptr;

/* grid node under AHFlean: 24 bytes */
typedef struct node *nptr;
typedef struct node {
  flouble dens;
  partptr ll;
  union forcestorage {
    flouble tempdens;
    partptr new_ll;
    int colour;
  } force;
} node;

typedef struct nquad *nqptr;
typedef struct nquad {
  nptr  loc;      /* first node of this x-run */
  int   x;        /* x of first node */
  int   length;   /* nodes in the run */
  nqptr next;     /* next run on the same row, or NULL */
} nquad;

typedef struct cquad *cqptr;
typedef struct cquad {
  nqptr loc;      /* one nquad per row, rows y .. y+length-1 */
  int   y;
  int   length;
  cqptr next;     /* next block of rows on the same plane, or NULL */
} cquad;

typedef struct pquad *pqptr;
typedef struct pquad {
  cqptr loc;      /* one cquad per plane, planes z .. z+length-1 */
  int   z;
  int   length;
  pqptr next;
} pquad;

typedef struct gridlist {
  pqptr  pquad;
  long   no_pquad;
  pqptr *pquad_array;
  long unsigned l1dim;
  double spacing;
  double spacing2;
  double critdens;
  double masstodens;
  double masstopartdens;
  double old_resid;
  double cur_resid;
  double trunc_err;
  long   no_sweeps;
  double timecounter;
  double timestep;
  int    multistep;
  struct grid_size { long unsigned no_part; long unsigned no_nodes; } size;
  struct grid_time { time_t potential, density, DK, grid, hydro; } time;
  boolean next;
} gridls;

typedef struct info_global {
  gridls       *dom_grid;
  int           domgrid_no;
  partptr       fst_part;
  long unsigned no_part;
  gasptr        fst_gas;
  long unsigned no_gas;
  long          offset_gas;
  starptr       fst_star;
  long unsigned no_stars;
  long          offset_stars;
  double        a, t, z, super_t;
  int           no_timestep;
  long          fin_l1dim;
  boolean       fst_cycle;
  int           output_count;
  boolean       restart;
  double        total_time;
  double        bytes_node;
  double        bytes_part;
  boolean       terminate;
  char         *termfile_name;
  int           architecture;
  long unsigned ndummy;
  double        fdummy;
  double        ovlim, rho_b, rho_vir, max_ovdens;
  int           ioflag;
} info_global;

/* src/common.h:77-92 without WITH_MPI: the particle array as the SFC halo
 * gather sees it (the same array as info_global.fst_part, src/main.c:624),
 * ordered by sfckey */
typedef struct global_info_struct {
  partptr   fst_part;
  uint64_t  no_part;
  sfc_key_t minkey;
  sfc_key_t maxkey;
  int       ctype;                /* sfc_curve_t: SFC_CURVE_HILBERT = 0 */
  int       level;
  int       rank;
  int       size;
} global_info_t;

typedef struct tline {
  double age[AHF_MAXTIME], hubble[AHF_MAXTIME], omega[AHF_MAXTIME],
         lambda[AHF_MAXTIME], virial[AHF_MAXTIME], growth[AHF_MAXTIME],
         a[AHF_MAXTIME], t[AHF_MAXTIME], super_t[AHF_MAXTIME],
         rhoc[AHF_MAXTIME];
} tline;

typedef struct param_simu {
  int           NGRID_DOM;
  int           NGRID_MAX;
  double        Nth_dom;
  double        Nth_ref;
  int           UseRhoBack;
  double        UserDvir;
  double        MaxGatherRad;
  int           lb_level;
  double        AHF_VTUNE;
  int           AHF_MINPART;
  double        GADGET_m2Msunh;
  double        GADGET_l2Mpch;
  double        omega0;
  double        lambda0;
  tline         timeline;
  double        a_initial, a_final, z_initial, z_final;
  double        t_initial, t_final, super_t_initial, super_t_final;
  double        mean_dens;
  double        FourPiG;
  double        boxsize;
  double        pmass;
  double        t_unit;
  double        l_unit;
  double        m_unit;
  double        pos_shift[3];
  double        pos_scale;
  double        gamma, omegab, omegaDM, f_b, H_frac, T_init, B_init, e_init;
  double        no_vpart;
  long unsigned no_part;
  long unsigned no_species;
  long unsigned no_gas;
  long unsigned no_stars;
  long unsigned no_halos;
  int           NGRID_MIN;
  double        SHIFT;
  double        min_weight;
  double        max_weight;
  double        med_weight;
  int           np_limit;
  int           mmfocus;
  int           multi_mass;
  int           double_precision;
  int           hydro;
  int           magneto;
} param_simu;

typedef struct { double x, y, z; } XYZ;

typedef struct {
  int            nbins;
  double        *r;
  long unsigned *npart;
  double *nvpart, *ovdens, *dens, *v2_circ, *v_esc2, *sig_v, *Ekin, *Epot,
         *Lx, *Ly, *Lz, *axis1, *E1x, *E1y, *E1z, *axis2, *E2x, *E2y, *E2z,
         *axis3, *E3x, *E3y, *E3z;
} HALOPROFILE;

typedef struct {
  long unsigned  npart;
  long unsigned *ipart;       /* indices into global.fst_part (libc heap) */
  long unsigned  nll;
  partptr        ll;
  XYZ    pos;                 /* centre, box units */
  XYZ    vel;
  double M_vir, R_vir, sigV, v_esc2, V2_max, R_max, r2, ovdens, lambda,
         lambdaE, Ekin, Epot, SurfP, Phi0;
  XYZ    pos_com;
  double com_offset;
  XYZ    pos_mbp;
  XYZ    vel_mbp;
  double mbp_offset;
  XYZ    AngMom, axis, E1, E2, E3;
  double fMhires, cNFW, cR1, R1, R_edge;
  HALOPROFILE prof;
  int    hostHalo;
  int    hostHaloLevel;
  int    numSubStruct;
  int   *subStruct;
  double gatherRad;
  double spaRes;
  int    refLev;
  int    numNodes;
} HALO;

/* ahf_gridinfo's outputs (src/tdef.h:796-818): the isolated refinement a
 * colour names, and AHF's grid range */
typedef struct { int x, y, z; } intXYZ;
typedef struct {
  int    refLevel;
  int    isoRefIndex;
  intXYZ periodic;
} SRINDEX;
typedef struct info_ahf {
  int    no_grids;
  int    min_ref;
  time_t time;
} info_ahf;

#ifdef __cplusplus
#define AHF_STATIC_ASSERT static_assert
#else
#define AHF_STATIC_ASSERT _Static_assert
#endif
AHF_STATIC_ASSERT(sizeof(part) == 48, "part must be 48 bytes");
AHF_STATIC_ASSERT(offsetof(part, pos) == 8, "part.pos offset");
AHF_STATIC_ASSERT(offsetof(part, mom) == 20, "part.mom offset");
AHF_STATIC_ASSERT(offsetof(part, id) == 40, "part.id offset");
AHF_STATIC_ASSERT(sizeof(node) == 24, "node must be 24 bytes under AHFlean");
AHF_STATIC_ASSERT(sizeof(nquad) == 24, "nquad size");

#ifdef __cplusplus
}
#endif
#endif /* AHF_TYPES_H */
