"""GPU parity of HaloProfiles (src/libahf/ahf_halos.c:3961-5060) against the
CPU oracle, which itself matches the reference's HaloProfiles bit for bit.

Given the same constructed haloes (members in order, Phi0, R_vir), every
value comes out bit for bit, except those that pass through the device's
log10/pow/log (the bin radii and what is computed from them, and the
bisections of cNFW and cR1): those agree to 1e-12 relative (the
concentrations to the 1e-3 of their bisection), and the member counts per
bin stay exact.  End to end (construction on the GPU too), Phi0 carries the
rounding of the parallel unbinding sums (relative 1e-9), and so do the
energies built on it."""
import ctypes

import numpy as np
import pytest

from ahf_amd import dumps

from helpers import profiles_match as _compare

pytestmark = pytest.mark.gpu

@pytest.fixture(scope="module")
def ctx():
    import ahf_amd
    c = ahf_amd.Context(0)
    yield c
    c.close()


class ProfParams(ctypes.Structure):
    _fields_ = [("minpart", ctypes.c_int64), ("r_fac", ctypes.c_double), ("v_fac", ctypes.c_double),
                ("m_fac", ctypes.c_double), ("phi_fac", ctypes.c_double), ("hubble", ctypes.c_double)]


def _oracle_profile(oracle, s, lists, phi_rvir, minpart=None):
    """the oracle's HaloProfiles of given lists (as ProfileResult)"""
    oracle.oracle_halo_profiles.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                            ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(ProfParams), ctypes.c_void_p, ctypes.c_void_p]
    pp = ProfParams(s.minpart if minpart is None else minpart, s.r_fac, s.v_fac, s.pmass, s.phi_fac, s.hubble)
    pos = np.ascontiguousarray(s.pos, np.float32)
    mom = np.ascontiguousarray(s.mom, np.float32)
    nh = len(lists)
    nb = np.zeros(nh, np.int64)
    props = np.zeros((nh, len(dumps.PROF_PROPS)))
    prof = []
    for h in range(nh):
        ip = np.ascontiguousarray(lists[h], np.uint64)
        c = np.ascontiguousarray(s.centre[h], np.float64)
        pr = np.zeros((64, len(dumps.PROF_BINS)))
        k = oracle.oracle_halo_profiles(pos.ctypes.data, mom.ctypes.data, ip.ctypes.data, ip.size, c.ctypes.data,
                                        float(phi_rvir[h, 0]), float(phi_rvir[h, 1]), ctypes.byref(pp),
                                        props[h].ctypes.data, pr.ctypes.data)
        assert k >= 0
        nb[h] = k
        prof.append(pr[:k].copy())
    return dumps.ProfileResult(np.array([len(x) for x in lists], np.uint64), nb, props, prof)


def _constructed(oracle, snap, tmp_path, gather):
    """the oracle's construction (lists in order, Phi0, R_vir) of every halo"""
    out = str(tmp_path / "c.halo")
    if gather:
        assert oracle.oracle_gather_dump(snap.encode(), out.encode(), 8, 0, 1, None) == 0
    else:
        assert oracle.oracle_construct_dump(snap.encode(), out.encode(), 8, 0) == 0
    return dumps.read_halos(out)


def _set_split(ctx, split):
    """split: an int, the four-wavefront threshold (0 default, -1 none);
    "tN": haloes of N members or more walked over the device (k_prf_*)"""
    if isinstance(split, str):
        n = int(split[1:])
        ctx.set_profile_split(n)
        ctx.set_schedule(prof_tiled=n)
    else:
        ctx.set_profile_split(split)


def _reset_split(ctx):
    ctx.set_profile_split(0)
    ctx.set_schedule(prof_tiled=0)


CASES = [(2, 800, 1002, 0), (0, 64, 1000, 0), (3, 48, 5, 0), (2, 2000, 1002, 0), (2, 400, 1002, 1),
         (1, 64, 1001, 0), (4, 96, 1004, 0)]


# split: 0 the default (haloes of 2048 members or more on four wavefronts),
# 64 nearly every halo on four wavefronts, -1 every halo on one, t64 nearly
# every halo walked over the device
@pytest.mark.parametrize("split", [0, 64, -1, "t64"])
@pytest.mark.parametrize("cfg,side,seed,gather", CASES)
def test_profiles_of_given_haloes(oracle, snapgen, ctx, tmp_path, cfg, side, seed, gather, split):
    """the kernel on the oracle's constructed haloes against the oracle's
    HaloProfiles dump of the same snapshot"""
    snap = str(tmp_path / "p.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    s = dumps.read_snapshot(snap)
    c = _constructed(oracle, snap, tmp_path, gather)
    o = str(tmp_path / "o.prof")
    assert oracle.oracle_profiles_dump(snap.encode(), o.encode(), 8, gather) == 0
    ref = dumps.read_profiles(o)
    print(f"[profiles-gpu] cfg{cfg} side {side} gather {gather} split {split}: {s.nhalos} haloes", flush=True)
    _set_split(ctx, split)
    try:
        ctx.set_particles(s.pos, s.mom)
        ip = np.concatenate([x for x in c.ipart] + [np.zeros(0, np.uint64)])
        ctx.set_halos(s.centre, c.npart, ip)
        nb, props, prof = ctx.halo_profiles(s.minpart, s.r_fac, s.v_fac, s.pmass, s.phi_fac, s.hubble,
                                            phi0_rvir=np.stack([c.Phi0, c.R_vir], axis=1))
    finally:
        _reset_split(ctx)
    _compare(ref, nb, props, prof, f"cfg{cfg}/{side}/g{gather}/split{split}")


@pytest.mark.parametrize("cfg,side,seed,gather", [(2, 800, 1002, 0), (3, 48, 5, 0), (2, 400, 1002, 1)])
def test_profiles_end_to_end(oracle, snapgen, ctx, tmp_path, cfg, side, seed, gather):
    """construction and profiles on the GPU, against the oracle's dump"""
    snap = str(tmp_path / "p.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    s = dumps.read_snapshot(snap)
    o = str(tmp_path / "o.prof")
    assert oracle.oracle_profiles_dump(snap.encode(), o.encode(), 8, gather) == 0
    ref = dumps.read_profiles(o)
    print(f"[profiles-gpu] end to end cfg{cfg} side {side} gather {gather}", flush=True)
    ctx.set_particles(s.pos, s.mom)
    b = s.boxsize
    if gather:
        ctx.sfc_keys()
        ctx.gather_halos(s.centre, s.gather_rad, skip=(s.halo_npart == 0).astype(np.uint8))
    else:
        ctx.set_halos(s.centre, s.halo_npart, s.ipart)
    ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=ctx.UNBIND_WITH_RVIR,
               ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.pmass / (b * b * b))
    nb, props, prof = ctx.halo_profiles(s.minpart, s.r_fac, s.v_fac, s.pmass, s.phi_fac, s.hubble)
    npart, _, _ = ctx.unbind_results()
    assert np.array_equal(npart, ref.npart)
    _compare(ref, nb, props, prof, f"e2e cfg{cfg}/{side}/g{gather}", end_to_end=True)


def _sorted_halo(oracle, s, h):
    off = s.halo_offsets
    ip = np.ascontiguousarray(s.ipart[off[h]:off[h + 1]], np.uint64).copy()
    c = np.ascontiguousarray(s.centre[h], np.float64)
    oracle.oracle_sort_halo(np.ascontiguousarray(s.pos, np.float32).ctypes.data, ip.size, ip.ctypes.data,
                            c.ctypes.data, 2, 0)
    return ip


@pytest.mark.parametrize("split", [0, 6, "t6"])
def test_profiles_edge_cases(oracle, ctx, tmp_path, split):
    """lists in reverse order (bin radii that decrease: the sequential bin
    automaton), a member on the centre (no potential step), members at
    equal distances, the smallest halo find_max accepts, and a halo below
    minpart"""
    rng = np.random.default_rng(11)
    n = 3000
    c = np.array([0.25, 0.5, 0.75])
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = rng.uniform(1e-3, 3e-2, n)
    r[100:140] = r[100]          # a shell of equal distances
    pos = (c + u * r[:, None]).astype(np.float32)
    pos[7] = c.astype(np.float32)
    cen = pos[7].astype(np.float64)  # the centre sits on member 7
    mom = rng.standard_normal((n, 3)).astype(np.float32) * 3e-4
    snap = str(tmp_path / "edge.snap")
    sizes = [n, 1200, 6, 5, 800]
    lists, off = [], 0
    for k in sizes:
        lists.append(np.arange(off, off + k, dtype=np.uint64) % n)
        off += 37
    cents = np.array([cen, cen, cen, cen, c])
    dumps.write_snapshot(snap, pos, mom, centres=cents, halo_lists=lists, minpart=6, L_dom=16, L_max=16,
                         Nth_dom=1e9, Nth_ref=1e9)
    s = dumps.read_snapshot(snap)
    srt = [_sorted_halo(oracle, s, h) for h in range(len(sizes))]
    given = [srt[0], srt[1][::-1].copy(), srt[2], srt[3], rng.permutation(srt[4])]
    phi_rvir = np.array([[-3.0e-3, 0.03], [-1e-3, 0.02], [0.0, 0.01], [0.0, 0.01], [-2e-3, 0.02]])
    ref = _oracle_profile(oracle, s, given, phi_rvir, minpart=6)
    assert list(ref.nbins[2:4]) == [2, 0]
    print(f"[profiles-gpu] edge cases, split {split}", flush=True)
    _set_split(ctx, split)
    try:
        ctx.set_particles(s.pos, s.mom)
        ctx.set_halos(s.centre, np.array(sizes, np.uint64), np.concatenate(given))
        nb, props, prof = ctx.halo_profiles(6, s.r_fac, s.v_fac, s.pmass, s.phi_fac, s.hubble, phi0_rvir=phi_rvir)
    finally:
        _reset_split(ctx)
    _compare(ref, nb, props, prof, f"edge cases, split {split}")


def test_profiles_rejects_tiny_minpart(ctx):
    import ahf_amd
    ctx.set_particles(np.random.default_rng(0).random((100, 3)).astype(np.float32))
    ctx.set_halos(np.zeros((1, 3)), np.array([10], np.uint64), np.arange(10, dtype=np.uint64))
    with pytest.raises(ahf_amd.AhfGpuError):
        ctx.halo_profiles(5, 1.0, 1.0, 1.0, 1.0, 1.0, phi0_rvir=np.zeros((1, 2)))
