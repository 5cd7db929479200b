/*
 * gridinfo_gpu.hip -- AHF's ahf_gridinfo (src/libahf/ahf_gridinfo.c:59-860)
 * on the resident AMR hierarchy: the spatially connected refinements of
 * every grid from min_ref on.
 *
 * AHF colours the nodes of a grid in traversal order (z, y, x): a node
 * takes the lowest colour among its six face neighbours (as get_TSCnodes
 * finds them) that are already coloured, merging the others into it by
 * recolouring, or starts a new colour when none is.  The neighbours
 * already coloured are exactly those before it in traversal order, so the
 * result is the connected components of the face-neighbour graph, each
 * carrying the colour its first node started (merges keep the lowest
 * colour, and colours are numbered in the order they start).  The
 * components are unions of whole octets and colours start only at the
 * first nodes of octets, in octet order (gridinfo_core.h), so per grid:
 *   - per octet, its edges to the octets before it in x, y and z, its start
 *     flag and its periodic flags;
 *   - the parents seeded with the first octet of each x-run segment (a
 *     scan of the segment starts), then union-find over the y edges and,
 *     after a compression, over the z edges and the x wraps, every root the
 *     component's lowest octet (a larger root always hooks under a smaller
 *     one, by compare-and-swap);
 *   - a scan of the start flags names the colours; every node of an octet
 *     takes its root's colour, and the roots, in octet order, are the
 *     grid's refinements in AHF's list order, with the periodic flags of
 *     their octets (testBound).
 * AHF's colour holder has room for three colours; a node that would see
 * four or more distinct ones overflows it, which AHF leaves undefined and
 * this pass does not reproduce.
 */
#include <hip/hip_runtime.h>

#include <string.h>

#include <vector>

#include "../../include/ahf_gpu.h"
#include "amr_core.h"
#include "devprims.h"
#include "gpu_ctx.h"
#include "gridinfo_core.h"

using namespace ahf;

struct GiLevel {
  int64_t n_oct = 0, n_roots = 0, n_starts = 0;
  int32_t *colour = nullptr; /* per node (8 per octet): the colour */
  int32_t *refs = nullptr;   /* per root: name, grid, index, periodic x, y, z */
};

struct GiState {
  std::vector<GiLevel> lev;
  ahfgpu_gridinfo_info info;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~GiState()
  {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

/* per octet: x-run segment start, colour start and periodic flags */
__global__ void k_gi_flags(level_view v, uint8_t *__restrict__ seg, uint8_t *__restrict__ start,
                           uint8_t *__restrict__ pflag)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= v.n_oct) return;
  int64_t e[3];
  ahf_gi_oct_edges(&v, o, e);
  seg[o] = (uint8_t)ahf_gi_oct_seg_start(&v, o, e);
  start[o] = (uint8_t)ahf_gi_oct_start(&v, o, e);
  pflag[o] = (uint8_t)ahf_gi_oct_periodic(&v, o, e);
}

/* the first octet of every x-run segment, by segment (sid: exclusive scan
 * of the segment starts) */
__global__ void k_gi_seg_first(int64_t n, const uint8_t *__restrict__ seg, const int32_t *__restrict__ sid,
                               int32_t *__restrict__ first)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n && seg[o]) first[sid[o]] = (int32_t)o;
}

/* every octet's parent: the first octet of its segment */
__global__ void k_gi_seed(int64_t n, const uint8_t *__restrict__ seg, const int32_t *__restrict__ sid,
                          const int32_t *__restrict__ first, int32_t *__restrict__ parent)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n) parent[o] = first[sid[o] + seg[o] - 1];
}

/* the root of x: parents only ever point to lower octets */
__device__ inline int32_t gi_find(int32_t *par, int32_t x)
{
  volatile int32_t *p = par;
  int32_t cur = p[x];
  if (cur == x) return cur;
  int32_t prev = x, nxt;
  while (cur > (nxt = p[cur])) {
    p[prev] = nxt; /* path halving: still a lower octet of the same component */
    prev = cur;
    cur = nxt;
  }
  return cur;
}

__device__ inline void gi_union(int32_t *par, int32_t a, int32_t b)
{
  int32_t ra = gi_find(par, a), rb = gi_find(par, b);
  while (ra != rb) {
    if (ra > rb) {
      const int32_t t = ra;
      ra = rb;
      rb = t;
    }
    const int32_t old = atomicCAS(&par[rb], rb, ra);
    if (old == rb) return;
    rb = gi_find(par, old);
    ra = gi_find(par, ra);
  }
}

/* phase 0: the y edges; phase 1: the z edges and the x wraps */
__global__ void k_gi_union(level_view v, int phase, int32_t *parent)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= v.n_oct) return;
  int64_t e[3];
  ahf_gi_oct_edges(&v, o, e);
  if (phase == 0) {
    if (e[1] >= 0) gi_union(parent, (int32_t)e[1], (int32_t)o);
  } else {
    if (e[2] >= 0) gi_union(parent, (int32_t)e[2], (int32_t)o);
    if (e[0] >= 0 && ahf_kx(v.oct_key[o]) == 0) gi_union(parent, (int32_t)e[0], (int32_t)o);
  }
}

/* path compression between the phases: every parent becomes the root or,
 * where another thread's halving lands after it, an ancestor */
__global__ void k_gi_compress(int64_t n, int32_t *parent)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int32_t q = gi_find(parent, (int32_t)o);
  if (q != (int32_t)o) parent[o] = q;
}

/* final roots (the parents are read only here, so every walk ends at the
 * root), root flags and the component's periodic flags */
__global__ void k_gi_roots(int64_t n, const int32_t *__restrict__ parent, const uint8_t *__restrict__ pflag,
                           int32_t *__restrict__ rootof, uint8_t *__restrict__ root, uint32_t *__restrict__ pflag_root)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  int32_t q = (int32_t)o, nxt;
  while ((nxt = parent[q]) != q) q = nxt;
  rootof[o] = q;
  root[o] = q == (int32_t)o ? 1 : 0;
  if (pflag[o]) atomicOr(&pflag_root[q], (uint32_t)pflag[o]);
}

/* the colour of the eight nodes of every octet */
__global__ void k_gi_colour(int64_t n, const int32_t *__restrict__ rootof, const int32_t *__restrict__ sname,
                            int32_t base, int4 *__restrict__ colour)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int32_t c = base + sname[rootof[o]] + 1;
  colour[2 * o] = make_int4(c, c, c, c);
  colour[2 * o + 1] = make_int4(c, c, c, c);
}

__global__ void k_gi_refs(int64_t n, const uint8_t *__restrict__ root, const int32_t *__restrict__ rpos,
                          const int32_t *__restrict__ sname, const uint32_t *__restrict__ pflag_root, int32_t base,
                          int32_t grid, int32_t *__restrict__ refs)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n || !root[o]) return;
  int32_t *e = refs + 6 * (int64_t)rpos[o];
  const uint32_t f = pflag_root[o];
  e[0] = base + sname[o] + 1;
  e[1] = grid;
  e[2] = rpos[o];
  e[3] = (int32_t)(f & 1u);
  e[4] = (int32_t)((f >> 1) & 1u);
  e[5] = (int32_t)((f >> 2) & 1u);
}

__global__ void k_gi_zero_u32(uint32_t *__restrict__ p, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

static level_view gi_view(const DevLevel &l)
{
  level_view v;
  v.L = l.L;
  v.n_oct = l.n_oct;
  v.oct_key = l.oct_key;
  v.oct_link = l.oct_link;
  v.nbr = l.nbr;
  v.n_row = l.n_row;
  v.row_key = l.row_key;
  v.row_oct0 = l.row_oct0;
  v.row_plane = l.row_plane;
  v.n_plane = l.n_plane;
  v.plane_z = l.plane_z;
  v.plane_row0 = l.plane_row0;
  return v;
}

static inline unsigned gi_nb(int64_t n)
{
  return (unsigned)((n + AHF_BLOCK - 1) / AHF_BLOCK);
}

extern "C" int ahfgpu_gridinfo(ahfgpu_ctx *c, const ahfgpu_gridinfo_params *p, ahfgpu_gridinfo_info *info)
{
  if (!c || !p) return AHFGPU_EINVAL;
  if (!c->built) return c->fail(AHFGPU_EINVAL, "no AMR hierarchy (call ahfgpu_amr_build first)");
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(AHFGPU_EHIP, "hipSetDevice");
  GiState *S = (GiState *)c->gi_result;
  if (!S) {
    c->gi_result = S = new GiState();
    if (hipEventCreate(&S->e0) != hipSuccess || hipEventCreate(&S->e1) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "gridinfo: hipEventCreate");
  }
  S->lev.clear();
  memset(&S->info, 0, sizeof(S->info));
  Arena &A = c->garena;
  A.reset();
  /* min_ref: the last grid whose refinement overdensity lies below the
   * virial overdensity (ahf_gridinfo.c:147-175) */
  const int nl = (int)c->lev.size(), nref = nl - 1;
  int startg = 1;
  double ovd = 0.0;
  for (int i = 0; i <= nref; i++) {
    const double fl1dim = (double)c->lev[(size_t)i].L;
    const double len = p->boxsize / fl1dim;
    const double vol = len * len * len;
    ovd = (c->params.Nth_ref * p->refine_mass / vol) / p->rho_vir;
    if (ovd < p->ovlim) startg = i;
  }
  const int min_ref = startg, no_grids = nref - min_ref + 1;
  S->info.min_ref = min_ref;
  S->info.no_grids = no_grids;
  S->info.max_ovdens = ovd;
  int64_t base = 0;
  hipStream_t st = c->st;
  hipEventRecord(S->e0, st);
  for (int g = 0; g < no_grids && g < AHFGPU_MAX_LEVELS; g++) {
    const DevLevel &l = c->lev[(size_t)(min_ref + g)];
    const int64_t n = l.n_oct;
    if (8 * n >= ((int64_t)1 << 31)) return c->fail(AHFGPU_EUNSUP, "gridinfo: a grid of 2^31 nodes or more");
    GiLevel G;
    G.n_oct = n;
    uint8_t *seg = A.get<uint8_t>(n), *start = A.get<uint8_t>(n), *pflag = A.get<uint8_t>(n),
            *root = A.get<uint8_t>(n);
    int32_t *sid = A.get<int32_t>(n + 1), *first = A.get<int32_t>(n), *parent = A.get<int32_t>(n),
            *rootof = A.get<int32_t>(n), *sname = A.get<int32_t>(n + 1), *rpos = A.get<int32_t>(n + 1);
    uint32_t *pf_root = A.get<uint32_t>(n);
    int32_t *tmp = A.get<int32_t>(scan_tmp_elems(n) + 16);
    G.colour = A.get<int32_t>(8 * n);
    if (!seg || !start || !pflag || !root || !sid || !first || !parent || !rootof || !sname || !rpos || !pf_root ||
        !tmp || !G.colour)
      return c->fail(AHFGPU_ENOMEM, "device arena exhausted (gridinfo)");
    const level_view v = gi_view(l);
    const unsigned nb = gi_nb(n);
    hipLaunchKernelGGL(k_gi_flags, dim3(nb), dim3(AHF_BLOCK), 0, st, v, seg, start, pflag);
    exclusive_scan<uint8_t, int32_t>(seg, sid, n, tmp, st);
    hipLaunchKernelGGL(k_gi_seg_first, dim3(nb), dim3(AHF_BLOCK), 0, st, n, seg, sid, first);
    hipLaunchKernelGGL(k_gi_seed, dim3(nb), dim3(AHF_BLOCK), 0, st, n, seg, sid, first, parent);
    hipLaunchKernelGGL(k_gi_union, dim3(nb), dim3(AHF_BLOCK), 0, st, v, 0, parent);
    hipLaunchKernelGGL(k_gi_compress, dim3(nb), dim3(AHF_BLOCK), 0, st, n, parent);
    hipLaunchKernelGGL(k_gi_union, dim3(nb), dim3(AHF_BLOCK), 0, st, v, 1, parent);
    hipLaunchKernelGGL(k_gi_compress, dim3(nb), dim3(AHF_BLOCK), 0, st, n, parent);
    hipLaunchKernelGGL(k_gi_zero_u32, dim3(nb), dim3(AHF_BLOCK), 0, st, pf_root, n);
    hipLaunchKernelGGL(k_gi_roots, dim3(nb), dim3(AHF_BLOCK), 0, st, n, parent, pflag, rootof, root, pf_root);
    exclusive_scan<uint8_t, int32_t>(start, sname, n, tmp, st);
    exclusive_scan<uint8_t, int32_t>(root, rpos, n, tmp, st);
    hipLaunchKernelGGL(k_gi_colour, dim3(nb), dim3(AHF_BLOCK), 0, st, n, rootof, sname, (int32_t)base,
                       reinterpret_cast<int4 *>(G.colour));
    int32_t cnt[2] = {0, 0};
    hipMemcpyAsync(&cnt[0], sname + n, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(&cnt[1], rpos + n, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "gridinfo: %s", hipGetErrorString(hipGetLastError()));
    G.n_starts = cnt[0];
    G.n_roots = cnt[1];
    G.refs = A.get<int32_t>(6 * (G.n_roots > 0 ? G.n_roots : 1));
    if (!G.refs) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (gridinfo)");
    hipLaunchKernelGGL(k_gi_refs, dim3(nb), dim3(AHF_BLOCK), 0, st, n, root, rpos, sname, pf_root, (int32_t)base,
                       (int32_t)g, G.refs);
    base += G.n_starts;
    if (base >= INT32_MAX) return c->fail(AHFGPU_EUNSUP, "gridinfo: colours beyond 2^31");
    S->info.num_iso[g] = (int32_t)G.n_roots;
    S->info.tot_iso += (int32_t)G.n_roots;
    S->lev.push_back(G);
  }
  S->info.n_colours = (int32_t)base;
  hipEventRecord(S->e1, st);
  if (hipStreamSynchronize(st) != hipSuccess)
    return c->fail(AHFGPU_EHIP, "gridinfo: %s", hipGetErrorString(hipGetLastError()));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, S->e0, S->e1);
  S->info.device_ms = ms;
  if (info) *info = S->info;
  return 0;
}

extern "C" int ahfgpu_gridinfo_results(ahfgpu_ctx *c, int g, int32_t *colour_out, int32_t *refs_out)
{
  if (!c) return AHFGPU_EINVAL;
  GiState *S = (GiState *)c->gi_result;
  if (!S || g < 0 || g >= (int)S->lev.size()) return c->fail(AHFGPU_EINVAL, "no gridinfo result for grid %d", g);
  const GiLevel &G = S->lev[(size_t)g];
  if (colour_out) hipMemcpyAsync(colour_out, G.colour, (size_t)G.n_oct * 32, hipMemcpyDeviceToHost, c->st);
  if (refs_out && G.n_roots)
    hipMemcpyAsync(refs_out, G.refs, (size_t)G.n_roots * 6 * 4, hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess) return c->fail(AHFGPU_EHIP, "gridinfo results");
  return 0;
}

void ahfgpu_free_gridinfo(ahfgpu_ctx *c)
{
  if (!c) return;
  delete (GiState *)c->gi_result;
  c->gi_result = nullptr;
  c->garena.release();
}
