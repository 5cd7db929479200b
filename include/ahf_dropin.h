/*
 * ahf_dropin.h -- AHF's own entry points for the deposit / refinement /
 * unbinding path, with AHF's prototypes, structs and globals, implemented on
 * the MI355X library (ahf_gpu.h).  Link this library in place of the
 * reference objects that define them; every function keeps the reference's
 * argument meaning, ownership and return convention.
 *
 * Globals read, as the reference functions read them (they are defined by
 * AHF itself, src/common.c:23-34 and src/libahf/ahf_halos.c:163; this library
 * carries weak definitions only so that it also links on its own):
 *   simu   NGRID_DOM, NGRID_MAX, Nth_dom, Nth_ref, AHF_MINPART, AHF_VTUNE,
 *          mean_dens, no_part, no_vpart
 *   global fst_part, no_part, dom_grid, domgrid_no, fst_cycle, fin_l1dim,
 *          ovlim, rho_vir, a (ahf_gridinfo)
 *   global_info  fst_part, no_part (the same array as global's, ordered by
 *          part.sfckey, src/main.c:341-356, 624-625)
 *   r_fac, v_fac, phi_fac, rho_fac, Hubble, m_fac, x_fac (the writers)
 *
 * Errors: like the reference, the grid functions terminate the process with
 * a message on stderr when they cannot continue (device missing, allocation
 * failure); the message names the GPU failure.
 */
#ifndef AHF_DROPIN_H
#define AHF_DROPIN_H

#include "ahf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

extern param_simu simu;
extern info_global global;
extern global_info_t global_info;
extern double r_fac, v_fac, phi_fac, rho_fac, Hubble, m_fac, x_fac;

/* src/libamr_serial/generate_grids.h:7 -- the domain grid, with AHF's quads */
gridls *gen_domgrids(int *no_grids);

/* src/libamr_serial/lltools.h -- link every particle to its domain node
 * (head insertion, src/libamr_serial/lltools.c:32-92) */
void ll(long unsigned npart, partptr fst_part, gridls *cur_grid);

/* src/libamr_serial/density.h:7 -- dens = -mean_dens on every node */
void zero_dens(gridls *cur_grid);

/* src/libamr_serial/density.h:18 -- TSC number density of the particles
 * linked to the grid, in AHF's accumulation order, on the GPU; FALSE if the
 * grid has fewer than 125 nodes */
boolean assign_npart(gridls *cur_grid);

/* src/libamr_serial/generate_grids.h:9 -- the whole refinement recursion on
 * the GPU (gen_AMRhierarchy + gen_refgrid + refine_grid + relink +
 * relink_back + assign_npart on every level, generate_grids.c:126-432);
 * returns the refined grids in *grid_list with AHF's quads, densities and
 * particle lists.  Precondition as in src/main.c:623-648: every particle of
 * global.fst_part is linked to the domain grid (ll() was called). */
boolean gen_AMRhierarchy(gridls **grid_list, int *no_grids);

/* src/libahf/ahf.h:12 -- the spatially isolated refinements of the grids
 * gen_AMRhierarchy built last (src/libahf/ahf_gridinfo.c:59-860), on the
 * GPU: sets ahf.min_ref, ahf.no_grids and global.max_ovdens, colours every
 * node of the AHF grids (node.force.colour), and allocates spatialRefIndex
 * (one entry per colour started, plus one), numIsoRef and totnumIsoRef as
 * the reference does.  Calls AHF's own calc_virial and calc_rho_vir
 * (src/libutility/cosmology.c:47, 117; `double f(double a)`, which the
 * program links) with global.a; reads simu.pmass, simu.med_weight,
 * simu.boxsize, simu.Nth_ref.  AHF's log lines are not written. */
extern info_ahf ahf;
extern SRINDEX *spatialRefIndex;
extern int *numIsoRef;
extern int totnumIsoRef;
void ahf_gridinfo(gridls *grid_list, int curgrid_no);

/* src/libahf/ahf_halos.h:30 -- members ordered by distance from halo->pos */
void sort_halo_particles(HALO *halo);

/* src/libahf/ahf_halos.h:21 -- iterative removal of unbound members;
 * halo->ipart is replaced by a new malloc'd array, as in the reference.
 * Safe to call from several threads: calls made together are combined into
 * one batched GPU pass, and each returns when its own halo is done. */
void rem_unbound(HALO *halo);

/* src/libahf/ahf_halos.h:24 -- cut the radially sorted members at the
 * first sphere below global.ovlim (overdensity relative to global.rho_vir);
 * sets npart, M_vir, R_vir, ovdens, and Phi0 = 0 when icall == 0;
 * halo->ipart is realloc()ed as in the reference.  Thread-safe like
 * rem_unbound. */
void rem_outsideRvir(HALO *halo, int icall);

/* src/libahf/ahf_halos_sfc.h:28 -- append to halo->ipart the particles
 * within halo->gatherRad of halo->pos, in the reference's order (search
 * cells in block order, particles in key order); halo->ipart is realloc()ed
 * and halo->npart grows.  Reads part.sfckey of global_info's particles,
 * which must be in key order.  Thread-safe like rem_unbound. */
void ahf_halos_sfc_gatherParts(HALO *halo);

/* src/libahf/ahf_halos.h:15 -- radial profiles and integral properties of
 * a constructed halo (members in radial order, Phi0 and R_vir set), for
 * AHF's default flags; halo->prof's arrays are calloc()ed as c_profile
 * does.  Does nothing below simu.AHF_MINPART members; returns TRUE.
 * Thread-safe like rem_unbound. */
int HaloProfiles(HALO *halo);
/* No counterpart in the reference: HaloProfiles for all nhalos haloes in one
 * GPU pass.  Returns 0, or a negative AHFGPU_E* code. */
int ahf_halo_profiles(HALO *halos, long unsigned nhalos);
/* No counterpart in the reference: the whole of
 * ahf_halos_sfc_constructHalo (src/libahf/ahf_halos_sfc.c:119-157) for all
 * nhalos haloes in one batched GPU pass, HaloProfiles included: what
 * ahf_gather_construct_halos does, then the profiles of the lists the
 * device holds.  Returns 0, or a negative AHFGPU_E* code. */
int ahf_construct_profile_halos(HALO *halos, long unsigned nhalos);
/* No counterpart in the reference: the whole of
 * ahf_halos_sfc_constructHalo (src/libahf/ahf_halos_sfc.c:119-157) but
 * HaloProfiles, for all nhalos haloes in one batched GPU pass: haloes with
 * npart == 0 are left alone; the others start from an empty list and get
 * gatherParts, sort_halo_particles, rem_outsideRvir(0), rem_unbound and
 * rem_outsideRvir(1).  Returns 0, or a negative AHFGPU_E* code. */
int ahf_gather_construct_halos(HALO *halos, long unsigned nhalos);

/* No counterpart in the reference: the construction steps of
 * src/libahf/ahf_halos_sfc.c:140-153 after the gather -- sort_halo_particles,
 * rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1) -- for all nhalos
 * haloes in one batched GPU pass.  Returns 0, or a negative AHFGPU_E*
 * code. */
int ahf_construct_halos(HALO *halos, long unsigned nhalos);

/* No counterpart in the reference: sort_halo_particles + rem_unbound for
 * all nhalos haloes in one batched GPU pass, replacing the per-halo calls
 * of the OpenMP loop of src/libahf/ahf_halos.c:504-510.  Returns 0, or a
 * negative AHFGPU_E* code. */
int ahf_unbind_halos(HALO *halos, long unsigned nhalos);

/* src/libahf/ahf_io.h -- the ASCII catalogues of AHF's default build:
 * <fprefix>.AHF_halos (one line per halo of at least simu.AHF_MINPART
 * members, in the order idx[0..numHalos) gives), .AHF_profiles (its
 * bins, radii inside the convergence radius written negative) and
 * .AHF_particles (member ids), byte for byte as AHF writes them.  Host
 * code; the process exits when a file cannot be opened, as AHF's does. */
void ahf_io_WriteHalos(const char *fprefix, HALO *halos, unsigned long *idx, int numHalos);
void ahf_io_WriteProfiles(const char *fprefix, HALO *halos, unsigned long *idx, int numHalos);
void ahf_io_WriteParticles(const char *fprefix, HALO *halos, unsigned long *idx, int numHalos);
/* No counterpart in the reference: choose the device (default 0, first use)
 * and release the library's device state. */
int ahf_gpu_init(int device);
void ahf_gpu_finalize(void);
/* No counterpart in the reference: per-halo calls (sort_halo_particles,
 * rem_outsideRvir, rem_unbound, HaloProfiles, ahf_halos_sfc_gatherParts)
 * made at the same time from AHF's OpenMP loop are combined into batched
 * passes; the number of such passes and of the calls they served so far. */
void ahf_gpu_halo_call_stats(uint64_t *batches, uint64_t *requests);

#ifdef __cplusplus
}
#endif
#endif
