"""Properties of an AMR hierarchy that hold for AHF's (gen_AMRhierarchy,
src/libamr_serial/generate_grids.c:278-432), checked where the oracle is
too slow to build the hierarchy itself (the 512^3 and 896^3 workloads):

  - the domain lists every particle exactly once, each in the cell
    ll() puts it in (src/libamr_serial/lltools.c:57-64), and a finer level's
    particles are a subset of the coarser level's;
  - the domain holds the whole mass: sum(dens + mean_dens) = L^3;
  - the TSC density (src/libamr_serial/density.c:293-409) of sampled nodes,
    recomputed in FP64 from every particle within 1.5 cells, matches the
    FP32 value within the rounding of a sequential FP32 sum;
  - a sampled node whose test_node criterion holds (refine_grid.c:113-138:
    interior and a density of its 3x3x2 block at or above the threshold)
    is refined: the next level holds its octet;
  - the particles of sampled fine nodes lie in the node's closed cell
    (relink.c:140-153, 196-201).

Samples are nodes whose 5x5x5 neighbourhood is present and away from the
periodic faces, so every neighbour is interior as AHF's search finds it and
every particle within 1.5 cells has been relinked down to the level.

A level is a dict as ahf_amd.Context.amr_level returns it (octets sorted by
key, nodes 8*octet + (cz*4 + cy*2 + cx), final per-node lists).  The
checker is itself checked on the CPU against the oracle's hierarchies, and
against mistakes planted in them (tests/test_amr_props.py)."""
import numpy as np

KEY_MASK = (1 << 21) - 1


def levels_from_dump(grids):
    """oracle/reference AHFGRID1 levels (dumps.GridLevel) in the octet layout"""
    out = []
    for g in grids:
        xyz = g.node_xyz()
        L = int(g.l1dim)
        okey = (xyz[:, 0] // 2) | ((xyz[:, 1] // 2) << 21) | ((xyz[:, 2] // 2) << 42)
        child = (xyz[:, 2] & 1) * 4 + (xyz[:, 1] & 1) * 2 + (xyz[:, 0] & 1)
        keys = np.unique(okey)
        o = np.searchsorted(keys, okey)
        node = 8 * o + child
        n_nodes = 8 * keys.size
        dens = np.zeros(n_nodes, np.float32)
        dens[node] = g.dens
        counts = np.zeros(n_nodes, np.int64)
        counts[node] = g.counts
        list0 = np.zeros(n_nodes + 1, np.int64)
        np.cumsum(counts, out=list0[1:])
        lst = np.zeros(int(list0[-1]), np.uint32)
        src0 = np.zeros(len(g.counts) + 1, np.int64)
        np.cumsum(g.counts.astype(np.int64), out=src0[1:])
        for t in range(len(g.counts)):
            c = int(g.counts[t])
            if c:
                d = int(list0[node[t]])
                lst[d:d + c] = g.lists[src0[t]:src0[t] + c]
        out.append({"L": L, "critdens": g.critdens, "masstopartdens": g.masstopartdens,
                    "oct_key": keys.astype(np.uint64), "dens": dens, "list0": list0, "list": lst,
                    "n_part": int(g.size_no_part)})
    return out


def _oct_xyz(keys):
    k = keys.astype(np.int64)
    return k & KEY_MASK, (k >> 21) & KEY_MASK, (k >> 42) & KEY_MASK


def node_index(lev, x, y, z):
    """node index of the nodes at (x, y, z) of the level, -1 where absent"""
    keys = lev["oct_key"]
    ok = (x // 2) | ((y // 2) << 21) | ((z // 2) << 42)
    o = np.searchsorted(keys, ok.astype(np.uint64))
    o = np.minimum(o, max(keys.size - 1, 0))
    hit = keys.size > 0
    hit = hit & (keys[o] == ok.astype(np.uint64)) if keys.size else np.zeros(x.shape, bool)
    n = 8 * o.astype(np.int64) + (z & 1) * 4 + (y & 1) * 2 + (x & 1)
    return np.where(hit, n, -1)


def _nodes_xyz(lev, nodes):
    o = nodes // 8
    c = nodes % 8
    ox, oy, oz = _oct_xyz(lev["oct_key"][o])
    return 2 * ox + (c & 1), 2 * oy + ((c >> 1) & 1), 2 * oz + (c >> 2)


def _pick(rng, n, k=200000):
    """every index below n, or k of them at random when n is larger"""
    return np.arange(n) if n <= k else rng.integers(0, n, k)


def _tsc(s):
    a = np.abs(s)
    return np.where(a <= 0.5, 0.75 - s * s, np.where(a <= 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _sample_nodes(lev, rng, k):
    """up to k nodes with a present 5x5x5 neighbourhood, 2 or more cells from
    the periodic faces"""
    n_nodes = 8 * lev["oct_key"].size
    if n_nodes == 0:
        return np.zeros(0, np.int64)
    L = lev["L"]
    cand = rng.integers(0, n_nodes, min(n_nodes, 20 * k))
    x, y, z = _nodes_xyz(lev, cand)
    ok = (x >= 3) & (x <= L - 4) & (y >= 3) & (y <= L - 4) & (z >= 3) & (z <= L - 4)
    cand, x, y, z = cand[ok], x[ok], y[ok], z[ok]
    if cand.size == 0:
        return cand
    full = np.ones(cand.size, bool)
    for dz in range(-2, 3):
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                full &= node_index(lev, x + dx, y + dy, z + dz) >= 0
    return np.unique(cand[full])[:k]


def _near_particles(parts, pos, L, want, chunk=1 << 25):
    """the particles of `parts` whose cell at resolution L (linear key) is in
    the sorted array want, and those keys; in chunks, so that a level of
    hundreds of millions of particles fits in memory"""
    marked = None
    if L ** 3 <= (1 << 31):
        marked = np.zeros(L ** 3, bool)
        marked[want] = True
    out_p, out_k = [], []
    for a in range(0, parts.size, chunk):
        p = parts[a:a + chunk].astype(np.int64)
        cell = np.floor(L * pos[p].astype(np.float64)).astype(np.int64)
        np.clip(cell, 0, L - 1, out=cell)
        ck = (cell[:, 2] * L + cell[:, 1]) * L + cell[:, 0]
        near = marked[ck] if marked is not None else np.isin(ck, want)
        out_p.append(p[near])
        out_k.append(ck[near])
    if not out_p:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    return np.concatenate(out_p), np.concatenate(out_k)


def _level_particles(levels, l):
    """particles deposited on level l: its final lists and every finer level's"""
    parts = [levels[j]["list"] for j in range(l, len(levels))]
    return np.concatenate(parts) if parts else np.zeros(0, np.uint32)


def check_hierarchy(levels, pos, rng, sample=4000, dens_levels=None):
    """list of failed properties (empty when all hold).  levels: dicts as
    above, coarsest first; pos: (N, 3) float32 box units."""
    fails = []
    N = pos.shape[0]
    # the domain lists every particle once, in ll()'s cell
    dom = levels[0]
    allp = _level_particles(levels, 0)
    if allp.size != N or np.bincount(allp.astype(np.int64), minlength=N).max(initial=0) != 1:
        fails.append("the hierarchy does not list every particle exactly once")
    L0 = dom["L"]
    lst0 = dom["list"].astype(np.int64)
    cnt0 = np.diff(dom["list0"])
    if lst0.size:
        node0 = np.repeat(np.arange(cnt0.size, dtype=np.int64), cnt0)
        pick = _pick(rng, lst0.size)
        p = lst0[pick]
        cell = np.floor(L0 * pos[p].astype(np.float64)).astype(np.int64)
        cell[cell > L0 - 1] = 0
        want = node_index(dom, cell[:, 0], cell[:, 1], cell[:, 2])
        if not np.array_equal(want, node0[pick]):
            fails.append("domain: a particle is listed outside ll()'s cell")
    # mass on the domain (every node present, periodic), to the rounding of
    # FP32 running sums of up to thousands of terms in the densest nodes
    tot = float(np.sum(dom["dens"].astype(np.float64) + 1.0))
    if abs(tot - float(L0) ** 3) > 1e-4 * float(L0) ** 3:
        fails.append(f"domain mass {tot!r} != L^3 = {float(L0) ** 3!r}")
    # per level: densities, refinement, relinked cells
    for l, lev in enumerate(levels):
        if dens_levels is not None and l not in dens_levels:
            continue
        L = lev["L"]
        nodes = _sample_nodes(lev, rng, sample)
        if nodes.size == 0:
            continue
        x, y, z = _nodes_xyz(lev, nodes)
        # particles within 1.5 cells of a sampled node: the lists of its 27
        # neighbours, from every level at or below this one
        want = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    want.append(((z + dz) * L + (y + dy)) * L + (x + dx))
        want = np.unique(np.concatenate(want))
        pp, pk = _near_particles(_level_particles(levels, l), pos, L, want)
        order = np.argsort(pk, kind="stable")
        pp, pk = pp[order], pk[order]
        m2pd = lev["masstopartdens"]
        ref = np.empty(nodes.size)
        tol = np.empty(nodes.size)
        for i in range(nodes.size):
            ks = []
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        ks.append(((z[i] + dz) * L + (y[i] + dy)) * L + (x[i] + dx))
            ks = np.array(ks)
            a = np.searchsorted(pk, ks)
            b = np.searchsorted(pk, ks, side="right")
            sel = np.concatenate([pp[u:v] for u, v in zip(a, b)])
            q = pos[sel].astype(np.float64)
            c = (np.array([x[i], y[i], z[i]], np.float64) + 0.5) / L
            w = m2pd * _tsc((q[:, 0] - c[0]) * L) * _tsc((q[:, 1] - c[1]) * L) * _tsc((q[:, 2] - c[2]) * L)
            ref[i] = -1.0 + float(np.sum(w))
            tol[i] = 1.2e-7 * (sel.size + 2) * (1.0 + float(np.sum(np.abs(w))))
        got = lev["dens"][nodes].astype(np.float64)
        bad = np.abs(got - ref) > tol
        if bad.any():
            j = int(np.flatnonzero(bad)[0])
            fails.append(f"level {l}: {int(bad.sum())} of {nodes.size} sampled densities differ from the FP64 "
                         f"TSC sum (node {int(nodes[j])}: {got[j]!r} vs {ref[j]!r})")
        # test_node true -> refined
        if l + 1 < len(levels):
            thr = lev["critdens"] - 1.0
            test = np.zeros(nodes.size, bool)
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (0, 1):
                        nb = node_index(lev, x + dx, y + dy, z + dz)
                        test |= lev["dens"][nb].astype(np.float64) >= thr
            fin = levels[l + 1]
            refined = np.isin((x | (y << 21) | (z << 42)).astype(np.uint64), fin["oct_key"])
            if np.any(test & ~refined):
                fails.append(f"level {l}: {int(np.sum(test & ~refined))} sampled nodes pass test_node "
                             f"but are not refined")
        # relink: a fine node's particles lie in its closed cell
        if l > 0 and lev["list"].size:
            cnt = np.diff(lev["list0"])
            owner = np.repeat(np.arange(cnt.size, dtype=np.int64), cnt)
            pick = _pick(rng, lev["list"].size)
            ox, oy, oz = _nodes_xyz(lev, owner[pick])
            q = pos[lev["list"][pick].astype(np.int64)].astype(np.float64)
            # AHF's centre fin_x/L + 0.5/L and this one round differently by
            # an ulp or two, which decides for particles exactly on a face
            h = 0.5 / L + 1e-15
            inside = ((np.abs(q[:, 0] - (ox + 0.5) / L) <= h) & (np.abs(q[:, 1] - (oy + 0.5) / L) <= h) &
                      (np.abs(q[:, 2] - (oz + 0.5) / L) <= h))
            if not inside.all():
                fails.append(f"level {l}: {int((~inside).sum())} sampled particles lie outside their node's cell")
    return fails
