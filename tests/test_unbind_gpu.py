"""GPU parity of sort + unbinding (and of the rem_outsideRvir cuts around
it) against the CPU oracle: identical membership and order for every halo,
identical M_vir, R_vir, ovdens; Phi0 within the FP64 rounding of a parallel
prefix sum."""
import ctypes

import numpy as np
import pytest

from ahf_amd import dumps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ahf_amd
    c = ahf_amd.Context(0)
    yield c
    c.close()


def _run_gpu(ctx, s, mode, band=0.0, split=0):
    ctx.set_unbind_split(split)
    ctx.set_particles(s.pos, s.mom)
    ctx.set_halos(s.centre, s.halo_npart, s.ipart)
    info = ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=mode, band=band)
    return info, ctx.unbind_results()


def _compare(ref, npart, props, lists, tag):
    bad = []
    for h in range(len(lists)):
        if int(ref.npart[h]) != int(npart[h]) or not np.array_equal(ref.ipart[h], lists[h]):
            bad.append(h)
    assert not bad, f"{tag}: {len(bad)} haloes differ, first {bad[:5]}"
    assert np.array_equal(ref.M_vir, props[:, 0])
    set_ = (props[:, 5].astype(np.int64) & 2) != 0
    assert np.array_equal(ref.R_vir[set_], props[set_, 1])
    big = ref.npart > 0
    rel = np.abs(ref.Phi0[big] - props[big, 2]) / np.maximum(np.abs(ref.Phi0[big]), 1e-300)
    assert rel.size == 0 or rel.max() < 1e-9, f"{tag}: Phi0 rel err {rel.max():.2e}"


# split: 0 the default threshold, 32 every halo of 32 members or more through
# the split path, -1 none
@pytest.mark.parametrize("split", [0, 32, -1])
@pytest.mark.parametrize("cfg,side,seed,mode", [(1, 64, 1001, 0), (1, 64, 1001, 1), (0, 64, 1000, 0),
                                                (2, 2000, 1002, 0), (2, 300, 7, 1), (4, 96, 1004, 0)])
def test_unbind_matches_oracle(oracle, snapgen, ctx, tmp_path, cfg, side, seed, mode, split):
    snap = str(tmp_path / "h.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_unbind_dump(snap.encode(), out.encode(), 8, 0) == 0
    ref = dumps.read_halos(out)
    s = dumps.read_snapshot(snap)
    print(f"[unbind-gpu] cfg{cfg} side {side} mode {mode} split {split}: {s.nhalos} haloes {s.n_ipart} members",
          flush=True)
    info, (npart, props, lists) = _run_gpu(ctx, s, mode, split=split)
    _compare(ref, npart, props, lists, f"cfg{cfg}/{side}/mode{mode}")
    assert info.n_member_out == int(ref.npart.sum())


@pytest.mark.parametrize("mode,split", [(0, 0), (1, 0), (0, 32)])
def test_unbind_replay_band(oracle, snapgen, ctx, tmp_path, mode, split):
    """a wide band sends haloes with members near the threshold to the exact
    sequential replay: the results must still be the oracle's"""
    snap = str(tmp_path / "b.snap")
    assert snapgen.ahf_snapgen_write(2, 2000, 1002, snap.encode()) == 0
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_unbind_dump(snap.encode(), out.encode(), 8, 0) == 0
    ref = dumps.read_halos(out)
    s = dumps.read_snapshot(snap)
    info, (npart, props, lists) = _run_gpu(ctx, s, mode, band=1e-2, split=split)
    assert info.n_exact_replays > 0
    _compare(ref, npart, props, lists, f"band/mode{mode}")


def test_unbind_repeatable(snapgen, ctx, tmp_path):
    snap = str(tmp_path / "r.snap")
    assert snapgen.ahf_snapgen_write(2, 500, 99, snap.encode()) == 0
    s = dumps.read_snapshot(snap)
    _, r1 = _run_gpu(ctx, s, 0)
    info2 = ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=0)
    r2 = ctx.unbind_results()
    assert np.array_equal(r1[0], r2[0])
    for a, b in zip(r1[2], r2[2]):
        assert np.array_equal(a, b)


WITH_RVIR, RVIR_ONLY = 16, 32


def _construct_gpu(ctx, s, mode, centre=None, npart=None, ipart=None, split=0):
    ctx.set_unbind_split(split)
    ctx.set_particles(s.pos, s.mom)
    ctx.set_halos(s.centre if centre is None else centre, s.halo_npart if npart is None else npart,
                  s.ipart if ipart is None else ipart)
    info = ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=mode,
                      ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
    return info, ctx.unbind_results()


@pytest.mark.parametrize("split", [0, 32])
@pytest.mark.parametrize("cfg,side,seed,plain", [(1, 64, 1001, 0), (1, 64, 1001, 1), (0, 64, 1000, 0),
                                                 (4, 96, 1004, 0), (2, 2000, 1002, 0), (2, 300, 7, 1), (3, 48, 5, 0)])
def test_construct_matches_oracle(oracle, snapgen, ctx, tmp_path, cfg, side, seed, plain, split):
    """sort, rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1) in one pass"""
    snap = str(tmp_path / "c.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_construct_dump(snap.encode(), out.encode(), 8, 0) == 0
    ref = dumps.read_halos(out)
    s = dumps.read_snapshot(snap)
    print(f"[construct-gpu] cfg{cfg} side {side} plain {plain}: {s.nhalos} haloes {s.n_ipart} members", flush=True)
    info, (npart, props, lists) = _construct_gpu(ctx, s, plain | WITH_RVIR, split=split)
    _compare(ref, npart, props, lists, f"construct cfg{cfg}/{side}/{plain}")
    assert np.array_equal(ref.R_vir, props[:, 1])
    assert np.array_equal(ref.ovdens, props[:, 3])
    small = s.halo_npart < s.minpart
    assert np.all(props[small, 5] == 0) and np.all(props[~small, 5] != 0)
    assert info.n_member_out == int(ref.npart.sum())


class _RvirParams(ctypes.Structure):
    _fields_ = [("minpart", ctypes.c_int32), ("ovlim", ctypes.c_double), ("rho_vir", ctypes.c_double),
                ("rho_fac", ctypes.c_double)]


@pytest.mark.parametrize("plain", [0, 1])
def test_rvir_only_matches_oracle(oracle, snapgen, ctx, tmp_path, plain):
    """one rem_outsideRvir call on radially sorted lists"""
    snap = str(tmp_path / "r.snap")
    assert snapgen.ahf_snapgen_write(2, 1000, 31, snap.encode()) == 0
    s = dumps.read_snapshot(snap)
    off = s.halo_offsets
    ip = s.ipart.copy()
    for h in range(s.nhalos):
        seg = ip[off[h]:off[h + 1]]
        c = np.ascontiguousarray(s.centre[h])
        oracle.oracle_sort_halo(s.pos.ctypes.data, len(seg), seg.ctypes.data, c.ctypes.data, s.minpart, 1)
    info, (npart, props, lists) = _construct_gpu(ctx, s, plain | RVIR_ONLY, ipart=ip, split=32)
    P = _RvirParams(s.minpart, s.ovlim, s.rho_vir, s.rho_fac)
    cut = 0
    for h in range(s.nhalos):
        seg = np.ascontiguousarray(ip[off[h]:off[h + 1]])
        c = np.ascontiguousarray(s.centre[h])
        o = [ctypes.c_double(0.0) for _ in range(3)]
        k = oracle.oracle_rem_outsideRvir(s.pos.ctypes.data, seg.ctypes.data, len(seg), c.ctypes.data,
                                          ctypes.byref(P), *[ctypes.byref(x) for x in o])
        assert int(npart[h]) == k and np.array_equal(lists[h], seg[:k]), f"halo {h}"
        if len(seg) >= s.minpart:
            assert (props[h, 0], props[h, 1], props[h, 3]) == (o[0].value, o[1].value, o[2].value), f"halo {h}"
            assert props[h, 2] == 0.0 and int(props[h, 5]) == 15
        else:
            assert int(props[h, 5]) == 0
        cut += len(seg) - k
    assert cut > 0
    assert info.max_iterations == 0


@pytest.mark.parametrize("seed,big", [(1, False), (2, True), (3, False), (4, True)])
def test_ties_match_oracle(oracle, ctx, tmp_path, seed, big):
    """haloes with members at exactly equal distances (lattices, duplicated
    positions, a halo across the periodic corner, one repeated position,
    with big a lattice over 8192 members): the device sorts them in
    indexx's order of equal keys, so the sort alone, the unbinding (batched,
    per halo, split path) and the construction give the oracle's haloes in
    its reference-pinned indexx mode"""
    from helpers import tie_snapshot
    snap = str(tmp_path / "t.snap")
    nh = tie_snapshot(snap, seed, big=big)
    s = dumps.read_snapshot(snap)
    print(f"[unbind-gpu] ties seed {seed}: {nh} haloes, {s.n_ipart} members", flush=True)
    ctx.set_particles(s.pos, s.mom)
    ctx.set_halos(s.centre, s.halo_npart, s.ipart)
    info = ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=ctx.UNBIND_SORT_ONLY)
    assert info.n_tie_halos == nh
    _, _, lists = ctx.unbind_results()
    off = s.halo_offsets
    for h in range(nh):
        ip = s.ipart[off[h]:off[h + 1]].copy()
        oracle.oracle_sort_halo(s.pos.ctypes.data, len(ip), ip.ctypes.data, s.centre[h].ctypes.data, s.minpart, 0)
        assert np.array_equal(ip, lists[h]), f"sort: halo {h}"
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_unbind_dump(snap.encode(), out.encode(), 4, 0) == 0
    ref = dumps.read_halos(out)
    for mode, split in ((0, 0), (1, 0), (0, 32)):
        info, (npart, props, lists) = _run_gpu(ctx, s, mode, split=split)
        assert info.n_tie_halos == nh
        _compare(ref, npart, props, lists, f"ties {seed} mode {mode} split {split}")
    ctx.set_unbind_split(0)
    assert oracle.oracle_construct_dump(snap.encode(), out.encode(), 4, 0) == 0
    ref = dumps.read_halos(out)
    ctx.set_halos(s.centre, s.halo_npart, s.ipart)
    ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=ctx.UNBIND_WITH_RVIR,
               ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
    npart, props, lists = ctx.unbind_results()
    _compare(ref, npart, props, lists, f"ties {seed} construct")
    assert np.array_equal(ref.ovdens, props[:, 3])


@pytest.mark.parametrize("sched", [dict(split_min=4096, big_halo=1024, wave_halo=256, graph_batch=1),
                                   dict(split_min=-1, big_halo=8192, wave_halo=1024),
                                   dict(split_min=64, graph_batch=8),
                                   dict(split_min=64, split_spec=-1), dict(split_min=64, split_spec=1),
                                   dict(split_min=32, split_spec=16)])
def test_schedule_variants(oracle, snapgen, ctx, tmp_path, sched):
    """every schedule (split path threshold, workgroup tiers, graph replays
    between done checks, speculative passes of the split recurrence: none,
    one, sixteen) gives the oracle's haloes"""
    snap = str(tmp_path / "s.snap")
    assert snapgen.ahf_snapgen_write(2, 2000, 1002, snap.encode()) == 0
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_construct_dump(snap.encode(), out.encode(), 8, 0) == 0
    ref = dumps.read_halos(out)
    s = dumps.read_snapshot(snap)
    before = ctx.schedule()
    try:
        ctx.set_schedule(**sched)
        ctx.set_particles(s.pos, s.mom)
        ctx.set_halos(s.centre, s.halo_npart, s.ipart)
        ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=ctx.UNBIND_WITH_RVIR,
                   ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
        npart, props, lists = ctx.unbind_results()
        _compare(ref, npart, props, lists, f"schedule {sched}")
        assert np.array_equal(ref.ovdens, props[:, 3])
    finally:
        ctx.set_schedule(**before)


def test_schedule_file(tmp_path, monkeypatch):
    """a context applies the schedule file named by AHF_SCHEDULE"""
    import json

    import ahf_amd
    path = tmp_path / "sched.json"
    path.write_text(json.dumps({"split_min": 12345, "kmax_big": 500, "march_zn": 8, "graph_batch": 3,
                                "tuning": {"fields": {"split_min": {"chosen": 1}}}}))
    monkeypatch.setenv("AHF_SCHEDULE", str(path))
    c = ahf_amd.Context(0)
    try:
        sc = c.schedule()
        assert sc["split_min"] == 12345 and sc["kmax_big"] == 500 and sc["march_zn"] == 8 and sc["graph_batch"] == 3
    finally:
        c.close()


def test_timing_off(oracle, snapgen, ctx, tmp_path):
    """with the stage timers off the haloes are the same and the timings
    read 0; on again, they are timed"""
    snap = str(tmp_path / "t.snap")
    assert snapgen.ahf_snapgen_write(1, 64, 1001, snap.encode()) == 0
    out = str(tmp_path / "o.halo")
    assert oracle.oracle_unbind_dump(snap.encode(), out.encode(), 4, 0) == 0
    ref = dumps.read_halos(out)
    s = dumps.read_snapshot(snap)
    try:
        ctx.set_timing(False)
        _, (npart, props, lists) = _run_gpu(ctx, s, 0)
        _compare(ref, npart, props, lists, "timing off")
        assert ctx.last_timing()["unbind_ms"] == 0.0
    finally:
        ctx.set_timing(True)
    _, (npart, props, lists) = _run_gpu(ctx, s, 0)
    _compare(ref, npart, props, lists, "timing on")
    assert ctx.last_timing()["unbind_ms"] > 0.0
