"""AHF's own free_grid (src/libutility/alloc_struct.c:584-605) releases the
grids the drop-in builds: ahf_amd/csrc/dropin.c compiled with
AddressSanitizer, on the host, with a host stand-in for the device API that
serves the replay of the device pipeline (tests/emu/emu_device.c), linked
with the reference's own objects as in the integration recipe.  Any invalid
free and any block left behind fails the run (tests/emu/dropin_free_check.c).
Needs the reference build (oracle/_ref); no GPU."""
import os
import subprocess

import pytest

from helpers import fuzz_snapshot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
LIBS = ["libamr_serial", "libahf", "libutility", "libsfc", "libio", "libio_serial", "libgravity"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    objs = [os.path.join(REF, "obj", "common.o"), os.path.join(REF, "obj", "comm.o")]
    archives = [os.path.join(REF, "linked", f"libref_{d}.a") for d in LIBS]
    if not all(os.path.exists(p) for p in objs + archives):
        pytest.skip("reference objects not built (oracle/_ref, `make -C oracle ref-linked`)")
    d = tmp_path_factory.mktemp("dfc")
    flags = ["-std=c11", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-fopenmp",
             "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "ahf_amd", "csrc")]
    srcs = {"dropin.o": os.path.join(ROOT, "ahf_amd", "csrc", "dropin.c"),
            "emu_device.o": os.path.join(ROOT, "tests", "emu", "emu_device.c"),
            "main.o": os.path.join(ROOT, "tests", "emu", "dropin_free_check.c")}
    for o, src in srcs.items():
        subprocess.run(["gcc", *flags, *inc, "-c", src, "-o", str(d / o)], check=True)
    exe = str(d / "dropin_free_check")
    subprocess.run(["gcc", "-fsanitize=address", "-fopenmp", "-o", exe, *(str(d / o) for o in srcs), *objs,
                    *archives, *archives, "-lm"], check=True)
    return exe


def _run(checker, snap):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="4")
    r = subprocess.run([checker, snap], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert '"grids"' in r.stdout
    return r.stdout


@pytest.mark.parametrize("cfg,side,seed", [(1, 48, 7), (3, 48, 5), (0, 40, 1000)])
def test_free_grid_generated(checker, snapgen, tmp_path, cfg, side, seed):
    snap = str(tmp_path / "g.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    _run(checker, snap)


@pytest.mark.parametrize("seed,np2", [(401, False), (402, True), (403, False), (404, True)])
def test_free_grid_fuzz(checker, tmp_path, seed, np2):
    """fuzz boxes: runs across the periodic faces, levels rolled back"""
    snap = str(tmp_path / "f.snap")
    fuzz_snapshot(snap, seed, non_pow2=np2)
    _run(checker, snap)
