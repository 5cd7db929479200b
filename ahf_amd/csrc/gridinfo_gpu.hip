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
 * result is the connected components of the graph whose edges join each
 * node to its found face neighbours before it; a component's colour is the
 * one its first node started (merges keep the lowest colour, and colours
 * are numbered in the order they start).  Here, per grid:
 *   - each node's rank in traversal order, from the octet rows and planes;
 *   - a start flag (no found face neighbour before it) and, by a scan of
 *     the flags in rank order, the colour each start begins;
 *   - union-find over the edges, every root the component's lowest rank
 *     (a larger root always hooks under a smaller one, by compare-and-swap);
 *   - each node's colour is its root's start colour; the roots, in rank
 *     order, are the grid's refinements in AHF's list order, with the
 *     periodic flags of their nodes on the low faces (testBound).
 * AHF's colour holder has room for three colours; a node that would see
 * four or more distinct ones overflows it, which AHF leaves undefined and
 * this pass does not reproduce.
 */
#include <hip/hip_runtime.h>

#include <string.h>

#include <vector>

#include "../../include/ahf_gpu.h"
#include "amr_core.h"
#include "gridinfo_core.h"
#include "devprims.h"
#include "gpu_ctx.h"

using namespace ahf;

struct GiLevel {
  int32_t *rank;   /* per node: traversal rank */
  int32_t *parent; /* per rank: union-find parent, always a lower rank */
  int32_t *rootof; /* per rank: the component's root (its lowest rank) */
  uint8_t *start;  /* per rank: a colour starts here */
  int32_t *sname;  /* per rank (n+1): starts before it */
  uint8_t *pflag;  /* per rank: periodic flags of the component (at its root) */
  uint8_t *root;   /* per rank: a root */
  int32_t *rpos;   /* per rank (n+1): roots before it */
  int32_t *colour; /* per node: the colour */
  int32_t *refs;   /* per root: name, grid, index, periodic x, y, z */
  int64_t n_nodes = 0, n_roots = 0, n_starts = 0;
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

/* traversal ranks of the eight nodes of every octet */
__global__ void k_gi_rank(level_view v, int32_t *__restrict__ rank)
{
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= v.n_oct) return;
  const int64_t r = ahf_gi_row_of(&v, o);
  for (int c = 0; c < 8; c++) rank[8 * o + c] = (int32_t)ahf_gi_rank_in(&v, o, r, c);
}

/* every rank its own parent */
__global__ void k_gi_iota(int32_t *__restrict__ parent, int64_t nn)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nn) parent[r] = (int32_t)r;
}

/* start flags (no found face neighbour before the node) and periodic
 * flags, in rank order; each rank's parent is its lowest found face
 * neighbour before it (itself if none), which leaves the unions only the
 * other edges */
__global__ void k_gi_init(level_view v, const uint32_t *__restrict__ pmask, const int32_t *__restrict__ rank,
                          int32_t *__restrict__ parent, uint8_t *__restrict__ start, uint8_t *__restrict__ pflag)
{
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= 8 * v.n_oct) return;
  const uint32_t m = pmask[n];
  const int32_t rn = rank[n];
  int64_t u[6];
  ahf_gi_faces(&v, m, n, u);
  int32_t lo = rn;
  for (int d = 0; d < 6; d++)
    if (u[d] >= 0) {
      const int32_t ru = rank[u[d]];
      lo = ru < lo ? ru : lo;
    }
  parent[rn] = lo;
  start[rn] = lo == rn ? 1 : 0;
  pflag[rn] = (uint8_t)ahf_gi_periodic(&v, m, n);
}

/* the root of rank x: parents only ever point to lower ranks */
__device__ inline int32_t gi_find(int32_t *par, int32_t x)
{
  volatile int32_t *p = par;
  int32_t cur = p[x];
  if (cur == x) return cur;
  int32_t prev = x, nxt;
  while (cur > (nxt = p[cur])) {
    p[prev] = nxt; /* path halving: still a lower rank of the same component */
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

/* the edges: every node to its found face neighbours before it */
__global__ void k_gi_union(level_view v, const uint32_t *__restrict__ pmask, const int32_t *__restrict__ rank,
                           int32_t *parent)
{
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= 8 * v.n_oct) return;
  const int32_t rn = rank[n];
  int64_t u[6];
  ahf_gi_faces(&v, pmask[n], n, u);
  int32_t ru[6], lo = rn;
  for (int d = 0; d < 6; d++) {
    ru[d] = u[d] >= 0 ? rank[u[d]] : INT32_MAX;
    lo = ru[d] < lo ? ru[d] : lo;
  }
  /* the edge to the lowest is the initial parent */
  for (int d = 0; d < 6; d++)
    if (ru[d] < rn && ru[d] != lo) gi_union(parent, ru[d], rn);
}

/* path compression once the unions are complete: every parent becomes the
 * root or, where another thread's halving lands after it, an ancestor */
__global__ void k_gi_compress(int64_t nn, int32_t *parent)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nn) return;
  const int32_t q = gi_find(parent, (int32_t)r);
  if (q != (int32_t)r) parent[r] = q;
}

/* final roots (the parents are read only here, so every walk ends at the
 * root), root flags and the component's periodic flags */
__global__ void k_gi_roots(int64_t nn, const int32_t *__restrict__ parent, const uint8_t *__restrict__ pflag_in,
                           int32_t *__restrict__ rootof, uint8_t *__restrict__ root, uint32_t *__restrict__ pflag_root)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nn) return;
  int32_t q = (int32_t)r, nxt;
  while ((nxt = parent[q]) != q) q = nxt;
  rootof[r] = q;
  root[r] = q == (int32_t)r ? 1 : 0;
  if (pflag_in[r]) atomicOr(&pflag_root[q], (uint32_t)pflag_in[r]);
}

__global__ void k_gi_colour(int64_t nn, const int32_t *__restrict__ rank, const int32_t *__restrict__ rootof,
                            const int32_t *__restrict__ sname, int32_t base, int32_t *__restrict__ colour)
{
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  colour[n] = base + sname[rootof[rank[n]]] + 1;
}

__global__ void k_gi_refs(int64_t nn, const uint8_t *__restrict__ root, const int32_t *__restrict__ rpos,
                          const int32_t *__restrict__ sname, const uint32_t *__restrict__ pflag_root, int32_t base,
                          int32_t grid, int32_t *__restrict__ refs)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nn || !root[r]) return;
  int32_t *e = refs + 6 * (int64_t)rpos[r];
  const uint32_t f = pflag_root[r];
  e[0] = base + sname[r] + 1;
  e[1] = grid;
  e[2] = rpos[r];
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
    const int64_t nn = 8 * l.n_oct;
    if (nn >= ((int64_t)1 << 31)) return c->fail(AHFGPU_EUNSUP, "gridinfo: a grid of 2^31 nodes or more");
    if (!l.pmask) return c->fail(AHFGPU_EINVAL, "gridinfo: level %d has no presence masks", min_ref + g);
    GiLevel G;
    G.n_nodes = nn;
    G.rank = A.get<int32_t>(nn);
    G.parent = A.get<int32_t>(nn);
    G.rootof = A.get<int32_t>(nn);
    G.start = A.get<uint8_t>(nn);
    G.sname = A.get<int32_t>(nn + 1);
    G.pflag = A.get<uint8_t>(nn);
    G.root = A.get<uint8_t>(nn);
    G.rpos = A.get<int32_t>(nn + 1);
    G.colour = A.get<int32_t>(nn);
    uint32_t *pf_root = A.get<uint32_t>(nn);
    int32_t *tmp = A.get<int32_t>(scan_tmp_elems(nn) + 16);
    if (!G.rank || !G.parent || !G.rootof || !G.start || !G.sname || !G.pflag || !G.root || !G.rpos || !G.colour || !pf_root ||
        !tmp)
      return c->fail(AHFGPU_ENOMEM, "device arena exhausted (gridinfo)");
    const level_view v = gi_view(l);
    const unsigned nb = gi_nb(nn);
    hipLaunchKernelGGL(k_gi_rank, dim3(gi_nb(l.n_oct)), dim3(AHF_BLOCK), 0, st, v, G.rank);
    hipLaunchKernelGGL(k_gi_iota, dim3(nb), dim3(AHF_BLOCK), 0, st, G.parent, nn);
    hipLaunchKernelGGL(k_gi_init, dim3(nb), dim3(AHF_BLOCK), 0, st, v, l.pmask, G.rank, G.parent, G.start, G.pflag);
    hipLaunchKernelGGL(k_gi_union, dim3(nb), dim3(AHF_BLOCK), 0, st, v, l.pmask, G.rank, G.parent);
    hipLaunchKernelGGL(k_gi_compress, dim3(nb), dim3(AHF_BLOCK), 0, st, nn, G.parent);
    hipLaunchKernelGGL(k_gi_zero_u32, dim3(nb), dim3(AHF_BLOCK), 0, st, pf_root, nn);
    hipLaunchKernelGGL(k_gi_roots, dim3(nb), dim3(AHF_BLOCK), 0, st, nn, G.parent, G.pflag, G.rootof, G.root, pf_root);
    exclusive_scan<uint8_t, int32_t>(G.start, G.sname, nn, tmp, st);
    exclusive_scan<uint8_t, int32_t>(G.root, G.rpos, nn, tmp, st);
    hipLaunchKernelGGL(k_gi_colour, dim3(nb), dim3(AHF_BLOCK), 0, st, nn, G.rank, G.rootof, G.sname, (int32_t)base,
                       G.colour);
    int32_t cnt[2] = {0, 0};
    hipMemcpyAsync(&cnt[0], G.sname + nn, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(&cnt[1], G.rpos + nn, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess)
      return c->fail(AHFGPU_EHIP, "gridinfo: %s", hipGetErrorString(hipGetLastError()));
    G.n_starts = cnt[0];
    G.n_roots = cnt[1];
    G.refs = A.get<int32_t>(6 * (G.n_roots > 0 ? G.n_roots : 1));
    if (!G.refs) return c->fail(AHFGPU_ENOMEM, "device arena exhausted (gridinfo)");
    hipLaunchKernelGGL(k_gi_refs, dim3(nb), dim3(AHF_BLOCK), 0, st, nn, G.root, G.rpos, G.sname, pf_root,
                       (int32_t)base, (int32_t)g, G.refs);
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
  if (colour_out) hipMemcpyAsync(colour_out, G.colour, (size_t)G.n_nodes * 4, hipMemcpyDeviceToHost, c->st);
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
