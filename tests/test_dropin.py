"""The drop-in: AHF's own entry points (gen_domgrids, ll, zero_dens,
assign_npart, gen_AMRhierarchy, sort_halo_particles, rem_outsideRvir,
rem_unbound) served by
libahf_amd.so, driven by the harness that runs the reference's functions
(oracle/ref_driver.c), linked two ways:
  - oracle/_build/ahf_dropin_drv: compiled against include/ahf_dropin.h;
  - oracle/_ref/ahf_ref_linked (where the reference tree was present at build
    time): the reference's own objects, with the replaced functions made
    local by objcopy, linked with the library -- the integration recipe.
The grids returned through AHF's structs must equal the oracle's bit for bit;
the haloes must have the oracle's membership in the oracle's order."""
import os
import subprocess

import numpy as np
import pytest

from ahf_amd import dumps

from helpers import fuzz_snapshot, profiles_match

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "oracle", "_build", "ahf_dropin_drv")
LINKED = os.path.join(ROOT, "oracle", "_ref", "ahf_ref_linked")


@pytest.fixture(scope="module", params=["dropin", "linked"])
def drv(request):
    if request.param == "dropin":
        if not os.path.exists(DRV):
            pytest.fail("oracle/_build/ahf_dropin_drv missing: run `make -C oracle dropin` after building the library")
        return DRV
    if not os.path.exists(LINKED):
        pytest.skip("oracle/_ref/ahf_ref_linked not built (reference tree absent at build time)")
    return LINKED


def _run(cmd):
    print("[dropin] " + " ".join(os.path.basename(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"driver failed ({r.returncode}): {r.stderr[-2000:]}"
    return r


def _amr_case(oracle, drv, snap, tmp_path, tag):
    o_out, d_out = str(tmp_path / f"{tag}_o.grid"), str(tmp_path / f"{tag}_d.grid")
    assert oracle.oracle_amr_dump(snap.encode(), o_out.encode()) > 0
    _run([drv, "amr", snap, d_out])
    diffs = dumps.compare_grids(dumps.read_grids(d_out), dumps.read_grids(o_out))
    assert not diffs, f"{tag}: {diffs[:5]}"


@pytest.mark.parametrize("cfg,side,seed", [(1, 48, 1001), (0, 40, 1000), (3, 48, 5)])
def test_dropin_amr_configs(oracle, drv, snapgen, tmp_path, cfg, side, seed):
    snap = str(tmp_path / "c.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    _amr_case(oracle, drv, snap, tmp_path, f"cfg{cfg}_{side}_{seed}")


@pytest.mark.parametrize("seed", range(300, 306))
def test_dropin_amr_fuzz(oracle, drv, tmp_path, seed):
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, seed)
    _amr_case(oracle, drv, snap, tmp_path, f"fuzz{seed}")


@pytest.mark.parametrize("cfg,side,seed,how", [(1, 64, 1001, "one"), (0, 48, 1000, "dom"), (3, 48, 5, "hdr"),
                                              (4, 48, 1004, "one")])
def test_dropin_gridinfo(oracle, drv, snapgen, tmp_path, cfg, side, seed, how):
    """ahf_gridinfo through AHF's prototype after gen_AMRhierarchy: min_ref,
    no_grids, max_ovdens, numIsoRef, totnumIsoRef, every node's colour in
    AHF's traversal order and the spatialRefIndex entry of every colour the
    nodes carry, as the oracle (pinned to the reference) gives them"""
    from helpers import gridinfo_setup
    snap = str(tmp_path / "gi.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    ovlim, rho, pmass, _ = gridinfo_setup(snap, how)
    o_out, d_out = str(tmp_path / "o.ginf"), str(tmp_path / "d.ginf")
    assert oracle.oracle_gridinfo_dump(snap.encode(), o_out.encode(), ovlim, rho, pmass) >= 0
    _run([drv, "gridinfo", snap, d_out])
    a, b = dumps.read_gridinfo(d_out), dumps.read_gridinfo(o_out)
    assert (a.min_ref, a.no_grids, a.tot_iso, a.max_ovdens) == (b.min_ref, b.no_grids, b.tot_iso, b.max_ovdens)
    assert a.no_grids >= 1
    assert np.array_equal(a.num_iso, b.num_iso)
    for g, (x, y) in enumerate(zip(a.colours, b.colours)):
        assert np.array_equal(x, y), f"grid {g}: node colours differ"
    assert np.array_equal(a.refs, b.refs)


@pytest.mark.parametrize("seed", range(310, 314))
def test_dropin_gridinfo_fuzz(oracle, drv, tmp_path, seed):
    from helpers import gridinfo_setup
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, seed)
    ovlim, rho, pmass, _ = gridinfo_setup(snap, ("dom", "one")[seed % 2])
    o_out, d_out = str(tmp_path / "o.ginf"), str(tmp_path / "d.ginf")
    rc = oracle.oracle_gridinfo_dump(snap.encode(), o_out.encode(), ovlim, rho, pmass)
    if rc == -100:
        pytest.skip("the reference meets undefined behaviour on this input")
    assert rc >= 0
    _run([drv, "gridinfo", snap, d_out])
    a, b = dumps.read_gridinfo(d_out), dumps.read_gridinfo(o_out)
    assert (a.min_ref, a.no_grids, a.tot_iso, a.max_ovdens) == (b.min_ref, b.no_grids, b.tot_iso, b.max_ovdens)
    assert np.array_equal(a.num_iso, b.num_iso)
    assert all(np.array_equal(x, y) for x, y in zip(a.colours, b.colours))
    assert np.array_equal(a.refs, b.refs)


def _halo_match(ref, got, tag):
    assert np.array_equal(ref.npart, got.npart), f"{tag}: member counts differ"
    bad = [h for h in range(len(ref.ipart)) if not np.array_equal(ref.ipart[h], got.ipart[h])]
    assert not bad, f"{tag}: {len(bad)} member lists differ, first {bad[:5]}"
    assert np.array_equal(ref.M_vir, got.M_vir)
    ok = ref.npart >= 1
    assert np.array_equal(ref.R_vir[ok], got.R_vir[ok])
    big = ref.npart > 0
    # Phi0 comes from a parallel prefix sum on the GPU: FP64 rounding only
    rel = np.abs(ref.Phi0[big] - got.Phi0[big]) / np.maximum(np.abs(ref.Phi0[big]), 1e-300)
    assert rel.size == 0 or rel.max() < 1e-9, f"{tag}: Phi0 rel err {rel.max():.2e}"
    assert np.array_equal(ref.ovdens, got.ovdens)


@pytest.mark.parametrize("cmd", ["unbind", "unbind-batch"])
@pytest.mark.parametrize("cfg,side,seed", [(1, 64, 1001), (2, 400, 1002)])
def test_dropin_unbind(oracle, drv, snapgen, tmp_path, cmd, cfg, side, seed):
    if cmd == "unbind-batch" and drv == LINKED:
        pytest.skip("the reference's driver has no batched command")
    snap = str(tmp_path / "h.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    o_out, d_out = str(tmp_path / "o.halo"), str(tmp_path / "d.halo")
    assert oracle.oracle_unbind_dump(snap.encode(), o_out.encode(), 4, 0) == 0
    extra = ["--threads", "4"] if cmd == "unbind" else []
    _run([drv, cmd, snap, d_out, *extra])
    _halo_match(dumps.read_halos(o_out), dumps.read_halos(d_out), f"{cmd} cfg{cfg}/{side}")


@pytest.mark.parametrize("cmd", ["construct", "construct-batch"])
@pytest.mark.parametrize("cfg,side,seed", [(1, 64, 1001), (2, 400, 1002), (3, 48, 5)])
def test_dropin_construct(oracle, drv, snapgen, tmp_path, cmd, cfg, side, seed):
    """sort_halo_particles, rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1)
    per halo through AHF's prototypes, or batched"""
    if cmd == "construct-batch" and drv == LINKED:
        pytest.skip("the reference's driver has no batched command")
    snap = str(tmp_path / "c.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    o_out, d_out = str(tmp_path / "o.halo"), str(tmp_path / "d.halo")
    assert oracle.oracle_construct_dump(snap.encode(), o_out.encode(), 4, 0) == 0
    extra = ["--threads", "4"] if cmd == "construct" else []
    _run([drv, cmd, snap, d_out, *extra])
    _halo_match(dumps.read_halos(o_out), dumps.read_halos(d_out), f"{cmd} cfg{cfg}/{side}")


@pytest.mark.parametrize("cmd", ["gather", "construct-sfc", "construct-sfc-batch"])
@pytest.mark.parametrize("cfg,side,seed", [(1, 48, 1001), (3, 48, 5), (2, 400, 1002)])
def test_dropin_gather(oracle, drv, snapgen, tmp_path, cmd, cfg, side, seed):
    """ahf_halos_sfc_gatherParts per halo through AHF's prototype (alone, and
    followed by the construction steps), or the whole construction batched"""
    if cmd == "construct-sfc-batch" and drv == LINKED:
        pytest.skip("the reference's driver has no batched command")
    snap = str(tmp_path / "g.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    o_out, d_out = str(tmp_path / "o.halo"), str(tmp_path / "d.halo")
    construct = 0 if cmd == "gather" else 1
    assert oracle.oracle_gather_dump(snap.encode(), o_out.encode(), 4, 0, construct, None) == 0
    extra = [] if cmd.endswith("batch") else ["--threads", "4"]
    _run([drv, cmd, snap, d_out, *extra])
    _halo_match(dumps.read_halos(o_out), dumps.read_halos(d_out), f"{cmd} cfg{cfg}/{side}")


@pytest.mark.parametrize("cmd", ["profiles", "profiles-sfc", "profiles-batch", "profiles-sfc-batch"])
@pytest.mark.parametrize("cfg,side,seed", [(2, 400, 1002), (3, 48, 5)])
def test_dropin_profiles(oracle, drv, snapgen, tmp_path, cmd, cfg, side, seed):
    """HaloProfiles per halo through AHF's prototype after the construction
    steps (from the given lists, or gathered), or all of it batched"""
    if cmd.endswith("batch") and drv == LINKED:
        pytest.skip("the reference's driver has no batched command")
    snap = str(tmp_path / "p.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    o_out, d_out = str(tmp_path / "o.prof"), str(tmp_path / "d.prof")
    gather = 1 if "sfc" in cmd else 0
    assert oracle.oracle_profiles_dump(snap.encode(), o_out.encode(), 4, gather) == 0
    extra = [] if cmd.endswith("batch") else ["--threads", "4"]
    _run([drv, cmd, snap, d_out, *extra])
    ref, got = dumps.read_profiles(o_out), dumps.read_profiles(d_out)
    assert np.array_equal(ref.npart, got.npart)
    profiles_match(ref, got.nbins, got.props, got.prof, f"{cmd} cfg{cfg}/{side}", end_to_end=True)


def _catalogue_tokens(path):
    return [line.split("\t") for line in open(path).read().splitlines()]


def _last_digit_apart(x, y):
    """two printed numbers that differ by at most one unit in the last
    printed digit (integers must agree exactly)"""
    t = x.strip()
    if "e" in t:
        mant, ex = t.split("e")
        k = len(mant.split(".")[1]) if "." in mant else 0
        unit = 10.0 ** (int(ex) - k)
    elif "." in t:
        unit = 10.0 ** -len(t.split(".")[1])
    else:
        return False
    return abs(float(x) - float(y)) <= 1.01 * unit


@pytest.mark.parametrize("cmd", ["catalogue-sfc", "catalogue-sfc-batch"])
@pytest.mark.parametrize("cfg,side,seed", [(2, 400, 1002), (3, 48, 5)])
def test_dropin_catalogue(ref_bin, drv, snapgen, tmp_path, cmd, cfg, side, seed):
    """ahf_halos_sfc_constructHalo (gather, construction, HaloProfiles) on the
    GPU and the three catalogues, against the reference's whole pipeline:
    the same lines, columns and member ids; a printed value may differ in
    its last digit only where Phi0's parallel sums (relative 1e-9) or the
    device's log/pow (an ulp) sit on a rounding edge"""
    if cmd.endswith("batch") and drv == LINKED:
        pytest.skip("the reference's driver has no batched command")
    snap = str(tmp_path / "c.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    res = subprocess.run([ref_bin, "catalogue-sfc", snap, str(tmp_path / "r"), "--threads", "4"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    extra = [] if cmd.endswith("batch") else ["--threads", "4"]
    _run([drv, cmd, snap, str(tmp_path / "d"), *extra])
    assert (tmp_path / "r.AHF_particles").read_bytes() == (tmp_path / "d.AHF_particles").read_bytes()
    for suf in (".AHF_halos", ".AHF_profiles"):
        a, b = _catalogue_tokens(str(tmp_path / ("r" + suf))), _catalogue_tokens(str(tmp_path / ("d" + suf)))
        assert len(a) == len(b) and a[0] == b[0], suf
        differ = 0
        for la, lb in zip(a[1:], b[1:]):
            assert len(la) == len(lb), suf
            for x, y in zip(la, lb):
                if x != y:
                    differ += 1
                    assert _last_digit_apart(x, y), (suf, x, y)
        print(f"[dropin] {cmd} {suf}: {differ} printed values differ in the last digit", flush=True)
