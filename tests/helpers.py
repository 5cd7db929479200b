"""Test inputs: seeded hand-made snapshots that stress the AMR edge cases."""
import numpy as np

from ahf_amd import dumps


def fuzz_snapshot(path, seed, non_pow2=False):
    """Random small snapshot: background + Gaussian clusters, some placed on
    the periodic faces, some particles exactly on cell faces; random
    thresholds and depth.  non_pow2 draws domain grids that are even but not
    powers of two (AHF takes any LgridDomain).  Returns a dict of the
    parameters."""
    rng = np.random.default_rng(seed)
    L = int(rng.choice([6, 10, 12, 20, 24] if non_pow2 else [4, 8, 8, 16, 16, 32]))
    n_bg = int(rng.integers(0, L ** 3 + 1))
    parts = [rng.random((n_bg, 3))]
    for _ in range(int(rng.integers(1, 6))):
        c = rng.random(3)
        if rng.random() < 0.5:
            c[rng.integers(0, 3)] = rng.choice([0.0, 0.999, 0.5 / L, 1 - 0.5 / L])
        n = int(rng.integers(10, 3000))
        s = rng.choice([0.002, 0.01, 0.03, 0.1])
        parts.append((c + s * rng.standard_normal((n, 3))) % 1.0)
    if rng.random() < 0.3:
        k = int(rng.integers(1, 200))
        parts.append(rng.integers(0, 4 * L, (k, 3)) / (4.0 * L))
    pos = np.concatenate(parts).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    rng.shuffle(pos)
    nth = float(rng.choice([1.0, 2.0, 3.0, 5.0, 8.0]))
    Lmax = L << int(rng.integers(1, 7))
    nref = float(rng.choice([nth, 2.5, 5.0]))
    dumps.write_snapshot(path, pos, L_dom=L, L_max=Lmax, Nth_dom=nth, Nth_ref=nref)
    return dict(L=L, n=len(pos), Nth_dom=nth, Nth_ref=nref, L_max=Lmax)


# HaloProfiles comparisons: exact where the inputs fix the bits
# values computed from the bin radii or through log(): device libm
LIBM_BINS = {"r", "ovdens", "dens", "v2_circ"}
LIBM_PROPS = {"R1", "SurfP"}
BISECT_PROPS = {"cNFW", "cR1"}
# values built on Phi0 (end to end only)
PHI_BINS = {"v_esc2", "Epot"}
PHI_PROPS = {"v_esc2", "Epot", "lambdaE"}


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    with np.errstate(invalid="ignore"):
        near = np.abs(a - b) <= rtol * np.maximum(np.abs(a), np.abs(b))
    return bool(np.all(same | near))


def profiles_match(ref, nb, props, prof, tag, end_to_end=False, exact=False):
    """exact where the inputs fix the bits; the tolerances above elsewhere
    (exact: every value bit for bit, for host results against the reference)"""
    assert np.array_equal(ref.nbins, nb), tag
    assert int((nb > 0).sum()) > 0, tag
    exact_bits, total = 0, 0
    for j, name in enumerate(dumps.PROF_PROPS):
        a, b = ref.props[:, j], props[:, j]
        total += a.size
        exact_bits += int(np.sum((a == b) | (np.isnan(a) & np.isnan(b))))
        if exact:
            assert _close(a, b, 0.0), f"{tag}: {name}"
        elif name in BISECT_PROPS:
            assert np.all((a == b) | (np.abs(a - b) <= 1e-3)), f"{tag}: {name}"
        elif name in LIBM_PROPS:
            assert _close(a, b, 1e-12), f"{tag}: {name}"
        elif end_to_end and (name in PHI_PROPS or name.startswith(("pos_mbp", "vel_mbp", "mbp"))):
            if name in PHI_PROPS:
                assert _close(a, b, 1e-8), f"{tag}: {name}"
        else:
            assert _close(a, b, 0.0), f"{tag}: {name} differs in {int(np.sum(a != b))} haloes"
    for h in range(len(nb)):
        x, y = ref.prof[h], prof[h]
        assert x.shape == y.shape, (tag, h)
        for j, name in enumerate(dumps.PROF_BINS):
            total += x.shape[0]
            exact_bits += int(np.sum((x[:, j] == y[:, j]) | (np.isnan(x[:, j]) & np.isnan(y[:, j]))))
            if exact:
                assert _close(x[:, j], y[:, j], 0.0), f"{tag}: halo {h} {name}"
            elif name in LIBM_BINS:
                assert _close(x[:, j], y[:, j], 1e-12), f"{tag}: halo {h} {name}"
            elif end_to_end and name in PHI_BINS:
                assert _close(x[:, j], y[:, j], 1e-8), f"{tag}: halo {h} {name}"
            else:
                assert _close(x[:, j], y[:, j], 0.0), f"{tag}: halo {h} bin value {name}"
    print(f"[profiles-gpu] {tag}: {exact_bits}/{total} values bit-exact", flush=True)




def tie_snapshot(path, seed, big=False):
    """Haloes whose members lie at exactly equal distances from the centre, in
    shuffled list order, so the radial sort's permutation of equal keys is
    indexx's to decide (src/libutility/general.c:1093-1153):
      - cubic lattices around a centre on a lattice point (spacing 2^-10, all
        coordinates dyadic, so shells of equal dist^2 are exact);
      - the same across the periodic corner (centre at 0);
      - random members with a third of them duplicated at the position of
        another member;
      - a lattice shell mixed with random members;
      - a single position repeated (every key equal);
      - with big, also a 23^3 lattice (12,167 members: the radial sort's
        path for haloes over 8192 members).
    Velocities are Gaussian around a halo bulk velocity, with a tail above
    the escape speed, so unbinding removes some members.  Returns the number
    of haloes."""
    rng = np.random.default_rng(seed)
    dx = 2.0 ** -10
    pos_parts, lists, centres, n0 = [], [], [], 0

    def add(p, c):
        nonlocal n0
        p = np.asarray(p, np.float64) % 1.0
        pos_parts.append(p)
        order = rng.permutation(len(p)) + n0
        lists.append(order.astype(np.uint64))
        centres.append(c)
        n0 += len(p)

    for k in range(int(rng.integers(3, 6))):
        r = int(rng.integers(2, 6))
        g = np.arange(-r, r + 1) * dx
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        c = np.round(rng.uniform(0.1, 0.9, 3) / dx) * dx
        add(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + c, c)
    r = 3
    g = np.arange(-r, r + 1) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    add(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), np.zeros(3))
    for k in range(3):
        c = np.round(rng.uniform(0.1, 0.9, 3) / dx) * dx
        n = int(rng.integers(60, 600))
        p = c + np.round(rng.normal(0.0, 8.0, (n, 3))) * dx
        dup = rng.integers(0, n, n // 3)
        p[rng.integers(0, n, n // 3)] = p[dup]
        add(p, c)
    c = np.round(rng.uniform(0.1, 0.9, 3) / dx) * dx
    shell = np.array([[i, j, k] for i in (-2, 0, 2) for j in (-2, 0, 2) for k in (-2, 0, 2)], np.float64) * dx
    add(np.concatenate([c + shell, c + rng.normal(0.0, 4 * dx, (150, 3))]), c)
    c = np.round(rng.uniform(0.1, 0.9, 3) / dx) * dx
    add(np.repeat((c + 3 * dx)[None, :], 64, 0), c)
    if big:
        g = np.arange(-11, 12) * dx
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        c = np.round(rng.uniform(0.1, 0.9, 3) / dx) * dx
        add(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + c, c)
    pos = np.concatenate(pos_parts).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    mom = np.zeros_like(pos)
    off = 0
    for lst in lists:
        n = len(lst)
        v = rng.normal(0.0, 150.0, (n, 3)) + rng.normal(0.0, 60.0, 3)
        fast = rng.random(n) < 0.2
        v[fast] *= 4.0
        mom[lst.astype(np.int64)] = v
        off += n
    bg = rng.random((2000, 3)).astype(np.float32)
    pos = np.concatenate([pos, bg])
    mom = np.concatenate([mom, rng.normal(0.0, 0.1, (2000, 3)).astype(np.float32)])
    # a small box and particle mass so that these haloes bind most members
    dumps.write_snapshot(path, pos, mom, boxsize=1.0, pmass=1e8, L_dom=16, centres=np.array(centres),
                         halo_lists=lists, r_fac=1.0, v_fac=1.0, phi_fac=4.3e-9 * 1e8, hubble=0.0,
                         gather_rad=np.full(len(lists), 0.02), seed=seed)
    return len(lists)


def gridinfo_setup(path, how):
    """ahf_gridinfo's parameters for a snapshot file (ovlim, rho_vir,
    refine_mass, boxsize), the snapshot's virial overdensity rewritten so
    that the AHF grids start where `how` says: "hdr" keeps the snapshot's
    value, "dom" puts the first AHF grid on the domain grid (an ovlim
    between the domain grid's refinement overdensity and the next grid's),
    "one" on the first refinement (an ovlim below every grid's: AHF's
    default start, ahf_gridinfo.c:158)"""
    s = dumps.read_snapshot(path)
    rho = dumps.ahf_rho_vir(s.omega0, s.lambda0, s.a)
    ovd0 = (s.Nth_ref * s.pmass / (s.boxsize / float(s.L_dom)) ** 3) / rho
    ovlim = {"hdr": s.ovlim, "dom": 2.0 * ovd0, "one": 0.5 * ovd0}[how]
    if how != "hdr":
        off = dumps.SNAP_HEADER.size - 16
        with open(path, "r+b") as f:
            f.seek(off)
            f.write(np.float64(ovlim).tobytes())
    return ovlim, rho, s.pmass, s.boxsize
