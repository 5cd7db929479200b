"""GPU parity of the AMR hierarchy: the HIP path against the CPU oracle,
bit for bit (node sets, x-run/row/plane blocks, every FP32 density, every
per-node particle list in AHF's order)."""
import os
import sys

import numpy as np
import pytest

from ahf_amd import dumps

from helpers import fuzz_snapshot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ahf_amd
    c = ahf_amd.Context(0)
    yield c
    c.close()


def _gpu_dump(ctx, snap_path, out_path, mode=0):
    s = dumps.read_snapshot(snap_path)
    ctx.set_gather_mode(mode)
    ctx.set_particles(s.pos, s.mom)
    info = ctx.amr_build(s.L_dom, s.L_max, s.Nth_dom, s.Nth_ref)
    ctx.amr_dump(out_path)
    return info


# every deposit kernel (per-source walk, flat walk, and the per-level choice)
# must give the oracle's bits
# AUTO, SOURCE, FLAT, WAVE (every node folded by a wavefront), FLAT_WIDE (the
# flat walk with the 64-bit record offsets of levels over 4 GiB of records)
MODES = [0, 1, 2, 3, 4]


def _check(oracle, ctx, snap, tmp_path, tag, mode=0):
    o_out = str(tmp_path / f"{tag}_oracle.grid")
    g_out = str(tmp_path / f"{tag}_gpu.grid")
    n = oracle.oracle_amr_dump(snap.encode(), o_out.encode())
    assert n > 0, f"oracle failed ({n})"
    print(f"[amr-gpu] case {tag} gather mode {mode}: {os.path.getsize(snap)} B snapshot", flush=True)
    info = _gpu_dump(ctx, snap, g_out, mode)
    a, b = dumps.read_grids(g_out), dumps.read_grids(o_out)
    diffs = dumps.compare_grids(a, b)
    assert not diffs, f"{tag}: " + "; ".join(diffs[:6])
    assert info.n_levels == len(b)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("cfg,side,seed", [(1, 32, 11), (1, 64, 1001), (0, 64, 1000), (1, 48, 7), (3, 48, 5),
                                          (3, 64, 1003), (4, 64, 1004)])
def test_generated_configs(oracle, snapgen, ctx, tmp_path, cfg, side, seed, mode):
    snap = str(tmp_path / f"cfg{cfg}_{side}.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    _check(oracle, ctx, snap, tmp_path, f"cfg{cfg}_{side}_{seed}", mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", list(range(0, 40)))
def test_fuzz(oracle, ctx, tmp_path, seed, mode):
    snap = str(tmp_path / f"fuzz{seed}.snap")
    fuzz_snapshot(snap, seed)
    _check(oracle, ctx, snap, tmp_path, f"fuzz{seed}", mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", list(range(40, 52)))
def test_fuzz_non_pow2(oracle, ctx, tmp_path, seed, mode):
    snap = str(tmp_path / f"fuzz{seed}.snap")
    fuzz_snapshot(snap, seed, non_pow2=True)
    _check(oracle, ctx, snap, tmp_path, f"fuzz_np2_{seed}", mode)


@pytest.mark.parametrize("n_bg,clump", [(2000, 2500), (14000, 4000), (40000, 9000)])
def test_staged_plane_overflow(oracle, ctx, tmp_path, n_bg, clump):
    """levels that hold every octet are deposited plane by plane from LDS;
    a tight clump fills a plane's staged slots, so the nodes after it in the
    plane are walked from the position records instead (both slot counts:
    under and over two entries per node)"""
    rng = np.random.default_rng(n_bg + clump)
    bg = rng.random((n_bg, 3))
    c = np.array([0.02, 0.5, 0.97])
    cl = (c + 0.004 * rng.standard_normal((clump, 3))) % 1.0
    pos = np.concatenate([bg, cl]).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    snap = str(tmp_path / f"clump{n_bg}.snap")
    dumps.write_snapshot(snap, pos, L_dom=16, L_max=128, Nth_dom=2.0, Nth_ref=3.0)
    _check(oracle, ctx, snap, tmp_path, f"clump{n_bg}_{clump}", 0)


@pytest.mark.parametrize("sched", [dict(kmax_big=8, kmax_small=8, march_zn=8), dict(kmax_big=1000, kmax_small=300, march_zn=16)])
@pytest.mark.parametrize("cfg,side,seed", [(1, 64, 1001), (3, 64, 1003)])
def test_schedule_variants(oracle, snapgen, ctx, tmp_path, cfg, side, seed, sched):
    """the deposit's fold threshold and the march's planes per workgroup
    change nothing in the hierarchy"""
    snap = str(tmp_path / f"cfg{cfg}_{side}.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    before = ctx.schedule()
    try:
        ctx.set_schedule(**sched)
        _check(oracle, ctx, snap, tmp_path, f"cfg{cfg}_{side}_{seed}_sched", 0)
    finally:
        ctx.set_schedule(**before)
