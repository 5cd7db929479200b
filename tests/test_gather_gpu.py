"""GPU parity of the SFC keys and the halo gather (ahf_halos_sfc_gatherParts)
against the CPU oracle: identical keys, identical member lists in AHF's order,
and, followed by the construction steps on the device, identical haloes."""
import ctypes

import numpy as np
import pytest

from ahf_amd import dumps

pytestmark = pytest.mark.gpu

WITH_RVIR = 16


@pytest.fixture(scope="module")
def ctx():
    import ahf_amd
    c = ahf_amd.Context(0)
    yield c
    c.close()


def _snap(snapgen, tmp_path, cfg, side, seed):
    snap = str(tmp_path / f"g{cfg}_{side}_{seed}.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    return snap, dumps.read_snapshot(snap)


def test_sfc_keys_match_oracle(oracle, snapgen, ctx, tmp_path):
    snap, s = _snap(snapgen, tmp_path, 1, 48, 21)
    print(f"[gather-gpu] keys of {s.npart} particles", flush=True)
    ctx.set_particles(s.pos, s.mom)
    got = ctx.sfc_keys(copy=True)
    idx = np.random.default_rng(1).integers(0, s.npart, 20000)
    ref = np.array([oracle.oracle_sfc_key(*s.pos[i].astype(np.float64), 21) for i in idx], np.uint64)
    assert np.array_equal(got[idx], ref)
    assert np.all(got[1:] >= got[:-1])


def test_unsorted_keys_rejected(snapgen, ctx, tmp_path):
    import ahf_amd
    _, s = _snap(snapgen, tmp_path, 1, 24, 2)
    ctx.set_particles(s.pos, s.mom)
    keys = ctx.sfc_keys(copy=True)
    with pytest.raises(ahf_amd.AhfGpuError):
        ctx.set_sfc_keys(keys[::-1].copy())
    with pytest.raises(ahf_amd.AhfGpuError):
        ctx.gather_halos(s.centre, s.gather_rad)
    ctx.set_sfc_keys(keys)
    ctx.gather_halos(s.centre, s.gather_rad)


CASES = [(1, 48, 1001), (1, 64, 7), (0, 40, 1000), (3, 48, 5), (2, 400, 1002), (2, 2000, 77), (4, 64, 1004)]


@pytest.mark.parametrize("cfg,side,seed", CASES)
def test_gather_matches_oracle(oracle, snapgen, ctx, tmp_path, cfg, side, seed):
    snap, s = _snap(snapgen, tmp_path, cfg, side, seed)
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_gather_dump(snap.encode(), out.encode(), 8, 0, 0, None) == 0
    ref = dumps.read_halos(out)
    print(f"[gather-gpu] cfg{cfg} side {side}: {s.nhalos} haloes, {int(ref.npart.sum())} members", flush=True)
    ctx.set_particles(s.pos, s.mom)
    ctx.sfc_keys()
    ctx.gather_halos(s.centre, s.gather_rad)
    npart, lists = ctx.halo_members()
    assert np.array_equal(npart, ref.npart)
    bad = [h for h in range(s.nhalos) if not np.array_equal(lists[h], ref.ipart[h])]
    assert not bad, f"{len(bad)} lists differ, first {bad[:5]}"


@pytest.mark.parametrize("cfg,side,seed,plain", [(1, 64, 7, 0), (3, 48, 5, 0), (2, 2000, 77, 0), (2, 300, 9, 1)])
def test_gather_construct_matches_oracle(oracle, snapgen, ctx, tmp_path, cfg, side, seed, plain):
    """gather, sort, rem_outsideRvir, rem_unbound, rem_outsideRvir, all on
    the device with the lists never leaving it"""
    snap, s = _snap(snapgen, tmp_path, cfg, side, seed)
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_gather_dump(snap.encode(), out.encode(), 8, 0, 1, None) == 0
    ref = dumps.read_halos(out)
    print(f"[gather-gpu] construct cfg{cfg} side {side}", flush=True)
    ctx.set_particles(s.pos, s.mom)
    ctx.sfc_keys()
    ctx.gather_halos(s.centre, s.gather_rad, skip=(s.halo_npart == 0))
    ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=plain | WITH_RVIR,
               ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
    npart, props, lists = ctx.unbind_results()
    assert np.array_equal(npart, ref.npart)
    bad = [h for h in range(s.nhalos) if not np.array_equal(lists[h], ref.ipart[h])]
    assert not bad, f"{len(bad)} lists differ, first {bad[:5]}"
    assert np.array_equal(ref.M_vir, props[:, 0])
    assert np.array_equal(ref.R_vir, props[:, 1])
    assert np.array_equal(ref.ovdens, props[:, 3])
    big = ref.npart > 0
    rel = np.abs(ref.Phi0[big] - props[big, 2]) / np.maximum(np.abs(ref.Phi0[big]), 1e-300)
    assert rel.size == 0 or rel.max() < 1e-9


def test_gather_edge_cases(oracle, ctx, tmp_path):
    """haloes on the box faces and corners, a radius on the 2-cell grid (the
    whole box searched), a zero radius, and a skipped halo"""
    rng = np.random.default_rng(8)
    pos = rng.random((30000, 3)).astype(np.float32)
    pos[:64] = np.float32(0.0)                           # particles on the origin corner
    key = np.array([oracle.oracle_sfc_key(*p, 21) for p in pos.astype(np.float64)], np.uint64)
    order = np.argsort(key, kind="stable")
    pos = pos[order]
    cen = np.array([[0.0, 0.0, 0.0], [0.999, 0.5, 0.001], [0.5, 0.5, 0.5], [0.25, 0.75, 0.5],
                    [0.1, 0.2, 0.3], [0.7, 0.7, 0.7]])
    rad = np.array([0.02, 0.05, 0.3, 0.0, 0.01, 0.04])
    snap = str(tmp_path / "e.snap")
    dumps.write_snapshot(snap, pos, centres=cen, halo_lists=[[0]] * 5 + [[]], gather_rad=rad)
    out = str(tmp_path / "e.halo")
    assert oracle.oracle_gather_dump(snap.encode(), out.encode(), 1, 0, 0, None) == 0
    ref = dumps.read_halos(out)
    ctx.set_particles(pos)
    ctx.sfc_keys()
    ctx.gather_halos(cen, rad, skip=np.array([0, 0, 0, 0, 0, 1], np.uint8))
    npart, lists = ctx.halo_members()
    assert int(npart[5]) == 0 and int(ref.npart[5]) == 0
    assert int(npart[0]) >= 64 and int(npart[2]) > 1000
    for h in range(5):
        assert np.array_equal(lists[h], ref.ipart[h]), h
