"""GPU parity of ahf_gridinfo (src/libahf/ahf_gridinfo.c:59-860): the HIP
pass over the resident hierarchy against the CPU oracle -- AHF's first grid
and grid count, every node's colour, the isolated refinements per grid and
their spatialRefIndex entries, bit for bit."""
import numpy as np
import pytest

from ahf_amd import dumps

from helpers import fuzz_snapshot, gridinfo_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ahf_amd
    c = ahf_amd.Context(0)
    yield c
    c.close()


def traversal_order(lv):
    """node indices (octet-major) in AHF's traversal order (z, y, x)"""
    k = lv["oct_key"].astype(np.int64)
    m = (1 << 21) - 1
    c = np.arange(8)
    X = (2 * (k & m))[:, None] + (c & 1)
    Y = (2 * ((k >> 21) & m))[:, None] + ((c >> 1) & 1)
    Z = (2 * ((k >> 42) & m))[:, None] + (c >> 2)
    return np.lexsort((X.ravel(), Y.ravel(), Z.ravel()))


def _check(oracle, ctx, snap, tmp_path, how, tag):
    ovlim, rho, pmass, box = gridinfo_setup(snap, how)
    out = str(tmp_path / f"{tag}.ginf")
    rc = oracle.oracle_gridinfo_dump(snap.encode(), out.encode(), ovlim, rho, pmass)
    if rc == -100:
        pytest.skip("the reference meets undefined behaviour on this input")
    assert rc >= 0
    want = dumps.read_gridinfo(out)
    s = dumps.read_snapshot(snap)
    print(f"[gridinfo-gpu] case {tag} ({how}): {s.npart} particles", flush=True)
    ctx.set_particles(s.pos, s.mom)
    ctx.amr_build(s.L_dom, s.L_max, s.Nth_dom, s.Nth_ref)
    got = ctx.gridinfo(ovlim, rho, pmass, box)
    assert (got.min_ref, got.no_grids, got.tot_iso) == (want.min_ref, want.no_grids, want.tot_iso)
    assert got.max_ovdens == want.max_ovdens
    assert np.array_equal(got.num_iso, want.num_iso)
    for g in range(want.no_grids):
        lv = ctx.amr_level(want.min_ref + g, lists=False)
        col = got.colours[g][traversal_order(lv)]
        assert np.array_equal(col, want.colours[g]), f"{tag}: grid {g} node colours differ"
    assert np.array_equal(got.refs, want.refs), f"{tag}: spatialRefIndex entries differ"
    return got


@pytest.mark.parametrize("cfg,side,seed,how", [(1, 0, 1001, "one"), (1, 64, 1001, "hdr"), (0, 64, 1000, "dom"), (3, 64, 1003, "one"),
                                              (4, 96, 1004, "hdr"), (1, 96, 5, "one"), (2, 1500, 1002, "dom")])
def test_gridinfo_generated(oracle, snapgen, ctx, tmp_path, cfg, side, seed, how):
    snap = str(tmp_path / f"cfg{cfg}.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    _check(oracle, ctx, snap, tmp_path, how, f"cfg{cfg}_{side}_{seed}")


@pytest.mark.parametrize("seed", range(8))
def test_gridinfo_fuzz(oracle, ctx, tmp_path, seed):
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, 7000 + seed, non_pow2=seed % 3 == 2)
    _check(oracle, ctx, snap, tmp_path, ("dom", "one")[seed % 2], f"fuzz{seed}")


def test_gridinfo_repeat(oracle, snapgen, ctx, tmp_path):
    # a second pass on the same hierarchy, and one after a smaller build,
    # give the same result as a first
    snap = str(tmp_path / "r.snap")
    assert snapgen.ahf_snapgen_write(1, 64, 1001, snap.encode()) == 0
    a = _check(oracle, ctx, snap, tmp_path, "one", "rep_a")
    b = ctx.gridinfo(*gridinfo_setup(snap, "one"))
    assert all(np.array_equal(x, y) for x, y in zip(a.colours, b.colours))
    small = str(tmp_path / "s.snap")
    assert snapgen.ahf_snapgen_write(1, 32, 11, small.encode()) == 0
    _check(oracle, ctx, small, tmp_path, "one", "rep_small")
    _check(oracle, ctx, snap, tmp_path, "one", "rep_b")
