"""The per-thread rules of the HIP kernels (ahf_amd/csrc/amr_core.h),
replayed on the host by tests/emu/amr_emu.c, against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from ahf_amd import dumps

from helpers import fuzz_snapshot

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "amr_emu")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-Wno-unused-function", "-o", exe,
                    os.path.join(HERE, "emu", "amr_emu.c"), "-lm"], check=True)
    return exe


def _cmp(oracle, emu, snap, tmp_path):
    o, e = str(tmp_path / "o.grid"), str(tmp_path / "e.grid")
    assert oracle.oracle_amr_dump(snap.encode(), o.encode()) > 0
    r = subprocess.run([emu, snap, e], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    diffs = dumps.compare_grids(dumps.read_grids(e), dumps.read_grids(o))
    assert not diffs, diffs[:5]


@pytest.mark.parametrize("seed", range(200, 260))
def test_emu_fuzz(oracle, emu, tmp_path, seed):
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, seed)
    _cmp(oracle, emu, snap, tmp_path)


@pytest.mark.parametrize("seed", range(260, 272))
def test_emu_fuzz_non_pow2(oracle, emu, tmp_path, seed):
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, seed, non_pow2=True)
    _cmp(oracle, emu, snap, tmp_path)


@pytest.mark.parametrize("cfg,side,seed", [(1, 32, 21), (1, 64, 1001), (0, 32, 1000), (1, 48, 7)])
def test_emu_configs(oracle, emu, snapgen, tmp_path, cfg, side, seed):
    snap = str(tmp_path / "g.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    _cmp(oracle, emu, snap, tmp_path)


def test_wavefront_fold_replay(tmp_path):
    """the heavy kernels' parallel fold (integer prefix sums inside a binade,
    exact ties by parity, plain steps across binades) replayed on the host
    against AHF's sequential FP32 accumulation, on random and adversarial
    runs; any differing bit fails"""
    exe = str(tmp_path / "fold_check")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-Wno-unused-function", "-o", exe,
                    os.path.join(HERE, "emu", "fold_check.c"), "-lm"], check=True)
    for seed in (1, 2):
        r = subprocess.run([exe, "300000", str(seed)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:] + r.stdout
        assert '"mismatches": 0' in r.stdout


def test_fold64_replay(tmp_path):
    """an FP64 running-sum fold (integer steps on a binade's grid,
    exact ties by parity, plain steps across binades and from zero or
    subnormals) replayed on the host against sequential FP64 additions, on
    random and adversarial runs, terms aimed at binade boundaries among
    them; any differing bit fails"""
    exe = str(tmp_path / "fold64_check")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-Wno-unused-function", "-o", exe,
                    os.path.join(HERE, "emu", "fold64_check.c"), "-lm"], check=True)
    for seed in (1, 2):
        r = subprocess.run([exe, "300000", str(seed)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:] + r.stdout
        assert '"mismatches": 0' in r.stdout


def test_indexx_level_replay(oracle, tmp_path):
    """the radial sort's level-synchronous indexx (every segment of every
    halo with ties partitioned at once, ahf_amd/csrc/nr_order.h) replayed on
    the host against the oracle's sequential indexx on tie-heavy arrays (a
    few values, all equal, runs, reversed runs, organ pipes, lattice shells,
    duplicates); any permutation that differs fails"""
    exe = str(tmp_path / "nr_check")
    lib = os.path.join(os.path.dirname(HERE), "oracle", "liboracle.so")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-Wno-unused-function", "-o", exe,
                    os.path.join(HERE, "emu", "nr_check.c"), lib, "-Wl,-rpath," + os.path.dirname(lib)], check=True)
    for seed in (1, 2):
        r = subprocess.run([exe, "2000", str(seed)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:] + r.stdout
        assert '"mismatches": 0' in r.stdout


def _cmp_gridinfo(oracle, emu, snap, tmp_path, how):
    from helpers import gridinfo_setup
    ovlim, rho, pmass, _ = gridinfo_setup(snap, how)
    o, e = str(tmp_path / "o.ginf"), str(tmp_path / "e.ginf")
    rc = oracle.oracle_gridinfo_dump(snap.encode(), o.encode(), ovlim, rho, pmass)
    if rc == -100:
        pytest.skip("the reference meets undefined behaviour on this input")
    assert rc >= 0
    r = subprocess.run([emu, snap, str(tmp_path / "e.grid"), e, repr(ovlim), repr(rho), repr(pmass)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    a, b = dumps.read_gridinfo(e), dumps.read_gridinfo(o)
    assert (a.min_ref, a.no_grids, a.tot_iso, a.max_ovdens) == (b.min_ref, b.no_grids, b.tot_iso, b.max_ovdens)
    assert np.array_equal(a.num_iso, b.num_iso)
    for g, (x, y) in enumerate(zip(a.colours, b.colours)):
        assert np.array_equal(x, y), f"grid {g}: node colours differ"
    assert np.array_equal(a.refs, b.refs)


@pytest.mark.parametrize("seed", range(300, 340))
def test_emu_gridinfo_fuzz(oracle, emu, tmp_path, seed):
    """the gridinfo kernels' rules (traversal ranks from the octet rows and
    planes, starts, unions under the lowest rank) replayed on the host
    against the oracle's sequential colouring"""
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, seed, non_pow2=seed % 4 == 3)
    _cmp_gridinfo(oracle, emu, snap, tmp_path, ("dom", "one")[seed % 2])


@pytest.mark.parametrize("cfg,side,seed,how", [(1, 64, 1001, "one"), (0, 48, 1000, "dom"), (3, 48, 5, "hdr")])
def test_emu_gridinfo_configs(oracle, emu, snapgen, tmp_path, cfg, side, seed, how):
    snap = str(tmp_path / "g.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    _cmp_gridinfo(oracle, emu, snap, tmp_path, how)


def test_march_rows_walk(tmp_path):
    """the march's rows walk (the three sources of a source row as one
    wrapped range of the staged plane) visits each lane's sources in AHF's
    traversal order, at the periodic faces too, with each entry's x offset"""
    exe = str(tmp_path / "march_rows_check")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wno-unused-function", "-o", exe,
                    os.path.join(HERE, "emu", "march_rows_check.c"), "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr
