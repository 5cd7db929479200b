"""HaloProfiles (src/libahf/ahf_halos.c:3961-5060): the CPU oracle pinned
against the reference's own HaloProfiles and against closed forms."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.optimize import brentq

from ahf_amd import dumps


def _same(a, b):
    """bit-equal float arrays (NaN equal to NaN)"""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def profiles_equal(a, b):
    assert np.array_equal(a.npart, b.npart)
    assert np.array_equal(a.nbins, b.nbins)
    bad = [dumps.PROF_PROPS[j] for j in range(a.props.shape[1]) if not _same(a.props[:, j], b.props[:, j])]
    assert not bad, f"properties differ: {bad}"
    for h, (x, y) in enumerate(zip(a.prof, b.prof)):
        assert _same(x, y), f"halo {h}: profile differs"


@pytest.mark.parametrize("cfg,side,seed,gather", [(2, 800, 1002, 0), (0, 64, 1000, 0), (3, 48, 5, 0),
                                                  (2, 400, 1002, 1), (1, 48, 1001, 1)])
def test_profiles_match_reference(oracle, ref_bin, snapgen, tmp_path, cfg, side, seed, gather):
    """the construction steps and HaloProfiles per halo, bit for bit"""
    snap = str(tmp_path / "p.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    o, r = str(tmp_path / "o.prof"), str(tmp_path / "r.prof")
    assert oracle.oracle_profiles_dump(snap.encode(), o.encode(), 4, gather) == 0
    cmd = "profiles-sfc" if gather else "profiles"
    res = subprocess.run([ref_bin, cmd, snap, r, "--threads", "4"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    a, b = dumps.read_profiles(o), dumps.read_profiles(r)
    assert int((a.nbins > 0).sum()) > 0
    profiles_equal(a, b)


def test_nbins_table(oracle):
    """(int)(6.2 log10 n - 3.5), at least 2 (specific.c:266-270)"""
    for n, want in [(1, 2), (10, 2), (20, 4), (100, 8), (1000, 15), (10 ** 6, 33), (2 ** 32 - 1, 56)]:
        assert oracle.oracle_prof_nbins(n) == want, n
    for n in range(1, 5000, 7):
        assert oracle.oracle_prof_nbins(n) == max(2, int(6.2 * math.log10(n) - 3.5))


def test_axes_are_the_eigen_decomposition(oracle):
    """get_axes: eigenvalues largest first, unit eigenvectors in the columns"""
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = rng.standard_normal((3, 3))
        t = m @ m.T
        it = t.ravel().copy()
        ax = np.zeros(3)
        oracle.oracle_prof_get_axes(it.ctypes.data, ax.ctypes.data)
        w = np.linalg.eigvalsh(t)[::-1]
        assert np.allclose(ax, w, rtol=1e-10, atol=1e-12)
        v = it.reshape(3, 3)
        assert np.allclose(t @ v, v * ax, atol=1e-9)
        assert np.allclose(v.T @ v, np.eye(3), atol=1e-10)
    # a diagonal tensor needs no rotation: the columns are the axes, sorted
    it = np.array([1.0, 0, 0, 0, 3.0, 0, 0, 0, 2.0])
    ax = np.zeros(3)
    oracle.oracle_prof_get_axes(it.ctypes.data, ax.ctypes.data)
    assert list(ax) == [3.0, 2.0, 1.0]
    assert list(it.reshape(3, 3)[:, 0]) == [0, 1, 0]


def _cnfw_ratio(c):
    return 0.216 * c / (math.log(1 + c) - c / (1 + c))


def _r1_of_c(c):
    return (c - 2 * math.log(1 + c) + c / (1 + c)) / (c * (math.log(1 + c) - c / (1 + c)))


def test_concentrations_are_roots(oracle):
    """calc_cNFW and calc_cR1 bisect to within 1e-3 of the root of Prada
    et al.'s Eq. 9 and of Wang et al.'s R1(c); -1 outside their ranges"""
    for c in (2.5, 4.0, 7.3, 12.0, 30.0, 80.0):
        got = oracle.oracle_prof_cnfw(_cnfw_ratio(c), 1.0)
        assert abs(got - c) < 1e-3, (c, got)
    assert oracle.oracle_prof_cnfw(0.9, 1.0) == -1 and oracle.oracle_prof_cnfw(6.0, 1.0) == -1
    for r1 in (0.2, 0.3, 0.45, 0.55):
        root = brentq(lambda c: _r1_of_c(c) - r1, 1.0, 500.0)
        assert abs(oracle.oracle_prof_cr1(r1) - root) < 1e-3
    assert oracle.oracle_prof_cr1(0.19) == -1 and oracle.oracle_prof_cr1(0.585) == -1


def test_find_max_peak(oracle):
    """find_max: the peak of a smoothed curve; a single spike is found where
    it is after no smoothing, and a plateau keeps its first point"""
    x = np.linspace(0.1, 2.0, 50)
    y = x * np.exp(-x)              # peaks at x = 1
    xm, ym = ctypes.c_double(), ctypes.c_double()
    oracle.oracle_prof_find_max(x.ctypes.data, y.ctypes.data, 50, 3, ctypes.byref(xm), ctypes.byref(ym))
    assert abs(xm.value - 1.0) < 0.05 and ym.value == y[np.argmin(abs(x - xm.value))]
    y = np.zeros(50)
    y[17] = 5.0
    oracle.oracle_prof_find_max(x.ctypes.data, y.ctypes.data, 50, 0, ctypes.byref(xm), ctypes.byref(ym))
    assert xm.value == x[17] and ym.value == 5.0
    y = np.ones(50)
    oracle.oracle_prof_find_max(x.ctypes.data, y.ctypes.data, 50, 1, ctypes.byref(xm), ctypes.byref(ym))
    assert xm.value == x[0]


def test_profile_closed_forms(oracle, tmp_path):
    """a halo of members on shells: the cumulative counts, mass, and the
    inertia axes of an isotropic shell set follow from the geometry"""
    rng = np.random.default_rng(5)
    n = 2000
    c = np.array([0.5, 0.5, 0.5])
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = np.sort(rng.uniform(1e-3, 2e-2, n))
    pos = (c + u * r[:, None]).astype(np.float32)
    snap = str(tmp_path / "shell.snap")
    dumps.write_snapshot(snap, pos, np.zeros_like(pos), centres=[c], halo_lists=[np.arange(n)], L_dom=16,
                         L_max=16, Nth_dom=1e9, Nth_ref=1e9, minpart=20, ovlim=1e-30)
    out = str(tmp_path / "o.prof")
    assert oracle.oracle_profiles_dump(snap.encode(), out.encode(), 1, 0) == 0
    p = dumps.read_profiles(out)
    nb = int(p.nbins[0])
    m = int(p.npart[0])
    assert nb == max(2, int(6.2 * math.log10(m) - 3.5))
    prof = p.prof[0]
    k = dumps.PROF_BINS.index
    # every member counted once, the last bin holds them all, mass == count
    assert prof[-1, k("npart")] == m and prof[-1, k("nvpart")] == m
    assert np.all(np.diff(prof[:, k("npart")]) >= 0)
    assert p.props[0, dumps.PROF_PROPS.index("M_vir")] == m
    # no velocities: no angular momentum, no kinetic energy but the Hubble flow
    assert np.all(prof[:, k("Lx")] == 0) and p.props[0, dumps.PROF_PROPS.index("lambda")] == 0
    # isotropic: the axis ratios of the whole halo are close to one
    assert p.props[0, dumps.PROF_PROPS.index("axis_y")] > 0.85


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WRITER_DRIVERS = {"dropin": os.path.join(ROOT, "oracle", "_build", "ahf_dropin_drv"),
                  "linked": os.path.join(ROOT, "oracle", "_ref", "ahf_ref_linked")}


@pytest.mark.parametrize("drv", ["dropin", "linked"])
@pytest.mark.parametrize("cfg,side,seed", [(2, 800, 1002), (3, 48, 5), (0, 64, 1000)])
def test_writers_match_reference(oracle, ref_bin, snapgen, tmp_path, cfg, side, seed, drv):
    """.AHF_halos, .AHF_profiles, .AHF_particles written by the library
    (host code) from the same haloes as AHF's own writers, byte for byte:
    through the harness built on the library's header, and through AHF's
    own objects linked with the library"""
    exe = WRITER_DRIVERS[drv]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    snap = str(tmp_path / "w.snap")
    assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    h, p = str(tmp_path / "w.halo"), str(tmp_path / "w.prof")
    assert oracle.oracle_construct_dump(snap.encode(), h.encode(), 4, 0) == 0
    assert oracle.oracle_profiles_dump(snap.encode(), p.encode(), 4, 0) == 0
    for b, pre in ((ref_bin, "r"), (exe, "d")):
        res = subprocess.run([b, "write-from", snap, h, p, str(tmp_path / pre)], capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-2000:]
    for suf in (".AHF_halos", ".AHF_profiles", ".AHF_particles"):
        a = (tmp_path / ("r" + suf)).read_bytes()
        b = (tmp_path / ("d" + suf)).read_bytes()
        assert len(a) > 100 and a == b, suf
