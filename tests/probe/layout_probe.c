/* prints sizes and offsets of include/ahf_types.h in the same JSON layout as
 * `ahf_ref layout` prints them for the reference headers */
#include <stddef.h>
#include <stdio.h>

#include "ahf_types.h"

#define OFF(T, F) printf("\"%s.%s\": %zu, ", #T, #F, offsetof(T, F))
int main(void)
{
  printf("{");
  printf("\"sizeof.part\": %zu, \"sizeof.node\": %zu, \"sizeof.nquad\": %zu, ", sizeof(part), sizeof(node),
         sizeof(nquad));
  printf("\"sizeof.cquad\": %zu, \"sizeof.pquad\": %zu, \"sizeof.gridls\": %zu, ", sizeof(cquad), sizeof(pquad),
         sizeof(gridls));
  printf("\"sizeof.HALO\": %zu, \"sizeof.HALOPROFILE\": %zu, ", sizeof(HALO), sizeof(HALOPROFILE));
  printf("\"sizeof.param_simu\": %zu, \"sizeof.info_global\": %zu, \"sizeof.global_info_t\": %zu, ", sizeof(param_simu), sizeof(info_global),
         sizeof(global_info_t));
  OFF(part, pos); OFF(part, mom); OFF(part, sfckey); OFF(part, id);
  OFF(node, dens); OFF(node, ll); OFF(node, force);
  OFF(nquad, loc); OFF(nquad, x); OFF(nquad, length); OFF(nquad, next);
  OFF(cquad, y); OFF(cquad, next); OFF(pquad, z); OFF(pquad, next);
  OFF(gridls, pquad); OFF(gridls, no_pquad); OFF(gridls, pquad_array); OFF(gridls, l1dim);
  OFF(gridls, critdens); OFF(gridls, masstodens); OFF(gridls, masstopartdens);
  OFF(gridls, multistep); OFF(gridls, size); OFF(gridls, time); OFF(gridls, next);
  OFF(HALO, npart); OFF(HALO, ipart); OFF(HALO, pos); OFF(HALO, M_vir); OFF(HALO, R_vir);
  OFF(HALO, Phi0); OFF(HALO, prof); OFF(HALO, hostHalo); OFF(HALO, gatherRad);
  OFF(HALO, numNodes);
  OFF(info_global, dom_grid); OFF(info_global, domgrid_no); OFF(info_global, fst_part);
  OFF(info_global, no_part); OFF(info_global, a); OFF(info_global, fin_l1dim);
  OFF(info_global, fst_cycle); OFF(info_global, ioflag);
  OFF(global_info_t, fst_part); OFF(global_info_t, no_part); OFF(global_info_t, ctype);
  OFF(global_info_t, size);
  OFF(param_simu, NGRID_DOM); OFF(param_simu, NGRID_MAX); OFF(param_simu, Nth_dom);
  OFF(param_simu, Nth_ref); OFF(param_simu, AHF_VTUNE); OFF(param_simu, AHF_MINPART);
  OFF(param_simu, timeline); OFF(param_simu, mean_dens); OFF(param_simu, boxsize);
  OFF(param_simu, pmass); OFF(param_simu, no_vpart); OFF(param_simu, no_part);
  OFF(param_simu, NGRID_MIN); OFF(param_simu, np_limit);
  OFF(param_simu, UseRhoBack); OFF(param_simu, UserDvir); OFF(param_simu, omega0); OFF(param_simu, lambda0);
  OFF(param_simu, med_weight); OFF(info_global, max_ovdens);
  printf("\"sizeof.SRINDEX\": %zu, \"sizeof.info_ahf\": %zu, ", sizeof(SRINDEX), sizeof(info_ahf));
  OFF(SRINDEX, refLevel); OFF(SRINDEX, isoRefIndex); OFF(SRINDEX, periodic);
  OFF(info_ahf, no_grids); OFF(info_ahf, min_ref); OFF(info_ahf, time);
  printf("\"end\": 0}\n");
  return 0;
}
