/*
 * ahf_gpu.h -- C ABI of the MI355X (gfx950) implementation of AHF's
 * particle->grid and unbinding hot path.
 *
 * Two layers share this library:
 *
 *  1. The device-resident API below (ahfgpu_*): particles are uploaded once
 *     and stay in HBM; the whole AMR hierarchy (TSC deposit, refinement
 *     flagging, fine-grid construction, relinking on every level) and the
 *     unbinding of every halo run on the GPU.  Results are read back on
 *     demand.  This replaces, as one call each:
 *       gen_domgrids + ll + zero_dens/assign_npart + gen_AMRhierarchy
 *         (src/main.c:614-648, src/libamr_serial/generate_grids.c:24-432,
 *          lltools.c:32-92, density.c:238-492, refine_grid.c:113-985,
 *          relink.c:31-423)
 *       the per-halo sort_halo_particles + rem_unbound loop
 *         (src/libahf/ahf_halos.c:504-510, 3292-3607, 5786-5894), with
 *         rem_outsideRvir before and after rem_unbound on request
 *         (ahf_halos.c:3687-3889, called at ahf_halos_sfc.c:143,153).
 *
 *  2. Drop-in versions of AHF's own entry points with AHF's prototypes and
 *     data structures (ahf_dropin.h), implemented on top of layer 1.
 *
 * All functions return 0 on success and a negative AHFGPU_E* code on
 * failure; ahfgpu_last_error() returns a message for the last failure of a
 * context.  No function exits the process.  A context is bound to the
 * device it was created on and must not be used by two threads at once.
 */
#ifndef AHF_GPU_H
#define AHF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHFGPU_OK 0
#define AHFGPU_EINVAL (-1)   /* bad argument or call order               */
#define AHFGPU_EHIP (-2)     /* HIP runtime error                         */
#define AHFGPU_ENOMEM (-3)   /* device or host allocation failed          */
#define AHFGPU_EUNSUP (-4)   /* input reaches behaviour AHF leaves undefined */
#define AHFGPU_EIO (-5)      /* file could not be read or written          */

#define AHFGPU_MAX_LEVELS 40

typedef struct ahfgpu_ctx ahfgpu_ctx;

/* AMR parameters: AHF.input's LgridDomain, LgridMax, NperDomCell,
 * NperRefCell (src/startrun.c:499-517) */
typedef struct {
  int32_t L_dom;
  int32_t L_max;
  double Nth_dom;
  double Nth_ref;
} ahfgpu_amr_params;

typedef struct {
  int32_t n_levels;                    /* grids kept (domain = level 0)        */
  int32_t rolled_back;                 /* last attempted level had <125 nodes  */
  int64_t n_nodes[AHFGPU_MAX_LEVELS];  /* nodes per level                      */
  int64_t n_part[AHFGPU_MAX_LEVELS];   /* particles deposited per level        */
  int64_t deposits;                    /* sum of n_part over levels            */
} ahfgpu_amr_info;

/* unbinding parameters: NminPerHalo, VescTune and the conversion factors of
 * src/libahf/ahf_halos.c:199-213; for the rem_outsideRvir cut also
 * global.ovlim, global.rho_vir and rho_fac (= pmass/boxsize^3) of
 * ahf_halos.c:204-221 */
typedef struct {
  int32_t minpart;
  int32_t pad;
  double vtune;
  double r_fac;
  double v_fac;
  double phi_fac;
  double hubble;
  double ovlim;
  double rho_vir;
  double rho_fac;
} ahfgpu_unbind_params;

typedef struct {
  int64_t n_halos;
  int64_t n_member_in;                 /* members before unbinding             */
  int64_t n_member_out;                /* bound members after                  */
  int64_t particle_iterations;         /* sum over halos and iterations        */
  int64_t n_exact_replays;             /* halo iterations re-run sequentially  */
  int32_t max_iterations;
  int32_t max_sweeps;                  /* most fixed-point sweeps of one halo  */
  int64_t n_tie_halos;                 /* haloes with equal distances, sorted
                                          again in indexx's order of them      */
} ahfgpu_unbind_info;

int ahfgpu_create(ahfgpu_ctx **ctx, int device);
void ahfgpu_destroy(ahfgpu_ctx *ctx);
const char *ahfgpu_last_error(const ahfgpu_ctx *ctx);

/* copy n particles (positions in box units and AHF velocities, both as
 * interleaved xyz FP32) from host memory to the device; the device keeps
 * them as arrays per component and once more interleaved per particle
 * (56 bytes a particle in all) */
int ahfgpu_set_particles(ahfgpu_ctx *ctx, const float *pos_xyz, const float *mom_xyz, uint64_t n);

/* build the AMR hierarchy of the resident particles (device only; the last
 * hierarchy stays resident until the next build) */
int ahfgpu_amr_build(ahfgpu_ctx *ctx, const ahfgpu_amr_params *p, ahfgpu_amr_info *info);

/* write the last hierarchy in the AHFGRID1 dump layout (tests, tools) */
int ahfgpu_amr_dump(ahfgpu_ctx *ctx, const char *path);

/* read back level `lev` of the last hierarchy.  Octets are the 2x2x2 node
 * blocks of the level at resolution L/2 sorted by (z,y,x): key = ox |
 * oy<<21 | oz<<42, link bit 0 = the x-run continues into the next octet.
 * Nodes are numbered 8*octet + (cz*4+cy*2+cx).  list0/list are the final
 * per-node particle lists in AHF's list order.  Pass NULL for unwanted
 * arrays; sizes: n_oct, n_oct, 8*n_oct, 8*n_oct+1, list0[8*n_oct]. */
int ahfgpu_amr_level_sizes(ahfgpu_ctx *ctx, int lev, int64_t *L, int64_t *n_oct, int64_t *n_listed,
                           double *critdens, double *masstopartdens);
int ahfgpu_amr_level_copy(ahfgpu_ctx *ctx, int lev, uint64_t *oct_key, uint8_t *oct_link, float *dens,
                          int64_t *list0, uint32_t *list);

/* TSC deposit (assign_npart) on one standalone level given as octets (as
 * above) and per-node particle lists in list order (CSR over octet-major
 * nodes, list0 has 8*n_oct+1 entries, indices into the resident
 * particles); dens_out receives 8*n_oct FP32 densities (octet-major),
 * accumulated in AHF's order.  Replaces one assign_npart() call
 * (src/libamr_serial/density.c:238-438) on any grid. */
int ahfgpu_deposit_level(ahfgpu_ctx *ctx, int64_t L, const uint64_t *oct_key, const uint8_t *oct_link,
                         int64_t n_oct, const int64_t *list0, const uint32_t *list, double masstopartdens,
                         float *dens_out);

#define AHFGPU_GATHER_AUTO 0   /* per level: SOURCE when the level holds at least 3 entries per node */
#define AHFGPU_GATHER_SOURCE 1 /* walk the 27 sources over per-entry positions */
#define AHFGPU_GATHER_FLAT 2   /* one flat walk per node over per-entry TSC weights */
#define AHFGPU_GATHER_WAVE 3   /* AUTO's walks, but every node with a contribution folded by a wavefront */
#define AHFGPU_GATHER_FLAT_WIDE 4 /* FLAT with 64-bit record offsets, the form a level takes when its
                                     weight records pass 4 GiB */

/* No counterpart in the reference: choose the deposit kernel of the
 * following builds and deposits.  Every mode gives the same bits; AUTO
 * (the default) picks the faster one per level. */
int ahfgpu_set_gather_mode(ahfgpu_ctx *ctx, int mode);

/* copy nh haloes to the device: centres (box units, nh x 3), member counts
 * and member particle indices (CSR, indices into the resident particles) --
 * the HALO.pos / npart / ipart that sort_halo_particles() receives */
int ahfgpu_set_halos(ahfgpu_ctx *ctx, const double *centres, const uint64_t *npart, const uint64_t *ipart,
                     uint64_t nh);

/* SFC keys of the resident particles (AHF's part.sfckey, 21 bits per axis
 * on AHF's Hilbert curve, src/main.c:341-356): uploaded from the caller, or
 * computed on the device.  The particles must be in key order, as AHF
 * sorts them; both calls check it and fail with AHFGPU_EINVAL otherwise.
 * ahfgpu_set_particles drops the keys.  keys_out may be NULL. */
int ahfgpu_set_sfc_keys(ahfgpu_ctx *ctx, const uint64_t *keys, uint64_t n);
int ahfgpu_sfc_keys(ahfgpu_ctx *ctx, uint64_t *keys_out);

/* ahf_halos_sfc_gatherParts (src/libahf/ahf_halos_sfc.c:171-367) for nh
 * haloes at once: the members of halo h are the resident particles within
 * gather_rad[h] (box units) of centres[h] (periodic), in AHF's order (the
 * cells of the 3x3x3 search block in block order, each cell in key order).
 * skip[h] != 0 gives halo h no members (AHF skips a halo whose grid count
 * is 0, ahf_halos_sfc.c:122-123); skip may be NULL.  The lists replace the
 * resident halo lists, as ahfgpu_set_halos would, ready for ahfgpu_unbind;
 * ahfgpu_halo_members reads them back. */
int ahfgpu_gather_halos(ahfgpu_ctx *ctx, const double *centres, const double *gather_rad, const uint8_t *skip,
                        uint64_t nh);
int ahfgpu_halo_members(ahfgpu_ctx *ctx, uint64_t *npart, uint64_t *ipart);

#define AHFGPU_UNBIND_BATCHED 0   /* all haloes in one persistent launch   */
#define AHFGPU_UNBIND_PLAIN 1     /* one launch per halo, synchronised     */
#define AHFGPU_UNBIND_SORT_ONLY 4 /* radial sort only (sort_halo_particles) */
#define AHFGPU_UNBIND_NO_SORT 8   /* unbind the lists in the given order   */
#define AHFGPU_UNBIND_WITH_RVIR 16 /* rem_outsideRvir before and after     */
#define AHFGPU_UNBIND_RVIR_ONLY 32 /* one rem_outsideRvir cut, no sort     */

/* radial sort + unbinding of every resident halo.  mode 0: all haloes in one
 * persistent launch (size-ordered work queue); mode 1: one launch per halo,
 * synchronised per halo (the per-call form); or-ed with one of SORT_ONLY,
 * NO_SORT, WITH_RVIR (the steps of ahf_halos_sfc.c:140-153: sort,
 * rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1)) or RVIR_ONLY (the
 * cut alone on the lists as given, as one rem_outsideRvir call; the caller
 * clears Phi0 for icall 0).  A member whose bound test could go the other way under AHF's
 * sequential rounding (certified error bound of the parallel sums) is
 * decided by an exact sequential replay of the iteration; band > 0 widens
 * that to every member within band*VescTune^2*v_esc^2 of the threshold (0:
 * the certified bound only).  The resident input lists are not modified, so
 * repeated calls redo the same work. */
int ahfgpu_unbind(ahfgpu_ctx *ctx, const ahfgpu_unbind_params *p, int mode, double band,
                  ahfgpu_unbind_info *info);

/* No counterpart in the reference: in the batched mode, haloes of at least
 * min_members members (and NminPerHalo) are unbound on a second path that
 * runs their passes over the members as flat kernels over all such haloes
 * and keeps one workgroup per halo only for the recurrence; the others stay
 * in the persistent kernel.  0 restores the default (16384), a negative
 * value sends every halo to the persistent kernel.  Both paths give the
 * same members and fields. */
int ahfgpu_set_unbind_split(ahfgpu_ctx *ctx, int64_t min_members);

/* results of the last ahfgpu_unbind: per halo npart and AHFGPU_PROPS numbers
 * (M_vir, R_vir, Phi0, ovdens, unbinding iterations, and a mask of the
 * AHFGPU_SET_* fields that some step wrote -- the others keep the caller's
 * values; R_vir is not set when rem_unbound bound no member, where AHF
 * stores an unassigned local), and the members kept of halo h at
 * ipart[offset_h .. offset_h + npart_h) (offsets of the input CSR) */
#define AHFGPU_PROPS 6
#define AHFGPU_SET_MVIR 1
#define AHFGPU_SET_RVIR 2
#define AHFGPU_SET_PHI0 4
#define AHFGPU_SET_OVDENS 8
int ahfgpu_unbind_results(ahfgpu_ctx *ctx, uint64_t *npart, double *props, uint64_t *ipart);
/* drop the resident haloes; their buffers stay with the context for the
 * next ahfgpu_set_halos or gather (no allocation per call for a caller that
 * handles few haloes at a time) until ahfgpu_destroy */
void ahfgpu_release_halos(ahfgpu_ctx *ctx);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ahfgpu_set_halos + ahfgpu_unbind in the batched mode
 * (mode 0 or-ed with one step selection; every halo in the persistent
 * kernel) + ahfgpu_unbind_results in one call, with the haloes and the
 * results in the caller's DEVICE memory and every intermediate in one static
 * device array per GPU of 192 MiB, held until the process ends.  The call
 * allocates nothing.
 *   centres: nh x 3 FP64 (box units); off: nh+1 int64 CSR offsets (off[0] =
 *   0, not decreasing); members: off[nh] indices into the resident particles
 *   of ctx; npart_out: nh; props_out: nh x AHFGPU_PROPS (as
 *   ahfgpu_unbind_results); members_out: off[nh] entries, halo h's kept
 *   members at members_out[off[h] .. off[h] + npart_out[h]) (the entries
 *   after them are left as they were).
 * At most AHFGPU_NOALLOC_MAX_HALOS haloes and AHFGPU_NOALLOC_MAX_MEMBERS
 * members; larger inputs, bad offsets or members fail with AHFGPU_EINVAL
 * before any result is written.  The work runs on the context's stream and
 * the call returns when it is done; the inputs must be ready before the
 * call.  Every call on one GPU shares that GPU's array and leaves its
 * contents behind: calls that overlap on one GPU (two threads, two
 * contexts) race. */
#define AHFGPU_NOALLOC_MAX_HALOS 65536
#define AHFGPU_NOALLOC_MAX_MEMBERS 1048576
int ahfgpu_unbind_noalloc(ahfgpu_ctx *ctx, const ahfgpu_unbind_params *p, int mode, double band,
                          const double *centres, const int64_t *off, const uint32_t *members, int64_t nh,
                          uint64_t *npart_out, double *props_out, uint32_t *members_out);

/* HaloProfiles (src/libahf/ahf_halos.c:3961-5060, AHF's default flags:
 * unit masses, R_max and r2 from the sorted particle list, Hubble drag) of
 * every resident halo, one wavefront per halo.  With phi0_rvir NULL the
 * haloes are the result of the last ahfgpu_unbind (members, Phi0, R_vir);
 * otherwise the resident lists as given, in their order, with Phi0 and
 * R_vir from phi0_rvir[2*h], [2*h+1].  nbins[h] (may be NULL) receives
 * each halo's bin count, 0 where HaloProfiles skips the halo (fewer than
 * minpart members); minpart must exceed 5.  ahfgpu_profile_results then
 * copies AHFGPU_PROF_NP properties per halo (zeros where nbins is 0) and
 * AHFGPU_PROF_NQ values per bin, the bins of all haloes back to back in
 * halo order.  The value orders are listed below. */
#define AHFGPU_PROF_NQ 25 /* r npart nvpart ovdens dens v2_circ v_esc2 sig_v Ekin Epot Lx Ly Lz
                             E1xyz E2xyz E3xyz axis1 axis2 axis3 */
#define AHFGPU_PROF_NP 45 /* nbins M_vir vel.xyz R_max V2_max r2 sigV v_esc2 Ekin Epot axis.xyz
                             E1.xyz E2.xyz E3.xyz AngMom.xyz lambda lambdaE cNFW R1 cR1 SurfP
                             fMhires mbp_offset pos_mbp.xyz vel_mbp.xyz com_offset pos_com.xyz */
typedef struct {
  int32_t minpart, pad;                    /* simu.AHF_MINPART */
  double r_fac, v_fac, m_fac, phi_fac, hubble; /* ahf_halos.c:199-206 */
} ahfgpu_profile_params;
int ahfgpu_halo_profiles(ahfgpu_ctx *ctx, const ahfgpu_profile_params *p, const double *phi0_rvir, int32_t *nbins);
int ahfgpu_profile_results(ahfgpu_ctx *ctx, double *props, double *prof);
/* No counterpart in the reference: haloes of at least min_members members
 * are profiled by workgroups of four wavefronts on a second stream (the
 * per-member work spread over 256 lanes, two chunks in flight), the
 * others by one wavefront each.  0 restores the default (2048), a negative
 * value keeps every halo on one wavefront.  Both give the same values. */
int ahfgpu_set_profile_split(ahfgpu_ctx *ctx, int64_t min_members);

/* ahf_gridinfo (src/libahf/ahf_gridinfo.c:59-860, called at src/main.c:657)
 * on the last hierarchy: the spatially connected refinements of every grid
 * from min_ref on.  min_ref is the last grid whose refinement overdensity
 * (Nth_ref * refine_mass / (boxsize/L)^3) / rho_vir lies below ovlim
 * (ahf_gridinfo.c:147-175; Nth_ref is the build's); max_ovdens is the
 * finest grid's.  Per AHF grid g (hierarchy level min_ref + g) every node
 * gets the colour of its connected component over the six face neighbours
 * get_TSCnodes finds, colours numbered 1, 2, ... across the grids in the
 * order AHF starts them (traversal order (z, y, x), grid by grid).
 * ahfgpu_gridinfo_results copies grid g's node colours (8*n_oct, octet-major
 * node order as ahfgpu_amr_level_copy) and its refinements in AHF's list
 * order, 6 int32 each: colour, grid, index among the grid's refinements
 * (spatialRefIndex.isoRefIndex), periodic x, y, z (:604-760).  A node that
 * meets four or more distinct colours overflows AHF's three-colour holder
 * (undefined behaviour there); the connected components are reported. */
typedef struct {
  double ovlim;       /* virial overdensity, calc_virial(a) */
  double rho_vir;     /* a^3 calc_rho_vir(a), the density it is relative to */
  double refine_mass; /* simu.pmass * simu.med_weight */
  double boxsize;     /* simu.boxsize */
} ahfgpu_gridinfo_params;
typedef struct {
  int32_t min_ref;    /* ahf.min_ref: the hierarchy level of AHF grid 0 */
  int32_t no_grids;   /* ahf.no_grids */
  int32_t tot_iso;    /* totnumIsoRef */
  int32_t n_colours;  /* colours started over all grids */
  double max_ovdens;  /* global.max_ovdens */
  int32_t num_iso[AHFGPU_MAX_LEVELS]; /* numIsoRef per AHF grid */
  double device_ms;   /* the pass on the device (HIP events on the library's stream) */
} ahfgpu_gridinfo_info;
int ahfgpu_gridinfo(ahfgpu_ctx *ctx, const ahfgpu_gridinfo_params *p, ahfgpu_gridinfo_info *info);
int ahfgpu_gridinfo_results(ahfgpu_ctx *ctx, int grid, int32_t *colour, int32_t *refs);

/* No counterpart in the reference: the schedule of the kernels, the
 * thresholds that choose between equivalent kernels (every choice gives the
 * same results).  0 in a field keeps its default; ahfgpu_create applies the
 * JSON file named by the environment variable AHF_SCHEDULE, or else
 * schedule.json beside the library's directory (written by the autotuner,
 * python -m ahf_amd.autotune), when it exists. */
typedef struct {
  int64_t split_min;   /* haloes of this many members or more take the split path (-1: none); 16384 */
  int64_t prof_split;  /* HaloProfiles: haloes of this many members or more on four wavefronts (-1: none); 2048 */
  int64_t big_halo;    /* unbinding: haloes of this many members or more on 512-thread workgroups; 2048 */
  int64_t wave_halo;   /* unbinding: haloes below this many members on one wavefront; 512 */
  int32_t kmax_big;    /* deposit: nodes with more contributions folded by a wavefront (levels of 2^21 nodes or more); 384 */
  int32_t kmax_small;  /* ... on smaller levels; 96 */
  int32_t march_zn;    /* deposit: target planes per march workgroup (8 or 16); by level size */
  int32_t graph_batch; /* split unbinding: iteration graph replays between done checks; 4 */
  int32_t split_spec;  /* split unbinding: passes settling all tiles at once before the sequential walk
                          resumes (-1: none); 3 */
  int32_t pad;
  int64_t prof_tiled;  /* HaloProfiles: haloes of this many members or more walked by kernels over the
                          whole device, only their running sums sequential (-1: none); 32768 */
} ahfgpu_schedule;
int ahfgpu_set_schedule(ahfgpu_ctx *ctx, const ahfgpu_schedule *s);
int ahfgpu_get_schedule(ahfgpu_ctx *ctx, ahfgpu_schedule *s);

/* No counterpart in the reference: the stage timers (HIP events around
 * each stage, read by ahfgpu_last_timing and ahfgpu_timer_get) on (1, the
 * default) or off (0: the timings read 0), for callers that make many
 * small calls and do not read them */
int ahfgpu_set_timing(ahfgpu_ctx *ctx, int on);

/* device milliseconds of the last build / unbind (HIP events on the
 * library's stream), and of the dominant kernel of the last build */
int ahfgpu_last_timing(ahfgpu_ctx *ctx, double *amr_ms, double *unbind_ms, double *gather_ms,
                       int64_t *gather_launches);

/* per-stage device timing of the last build and unbinding: number of stages,
 * and for stage i its name, total milliseconds, launches and algorithmic
 * bytes (HIP events on the library's stream) */
int ahfgpu_timer_count(void);
int ahfgpu_timer_get(ahfgpu_ctx *ctx, int i, const char **name, double *ms, int64_t *launches,
                     double *alg_bytes);

#ifdef __cplusplus
}
#endif
#endif
