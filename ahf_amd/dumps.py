"""Readers for the grid-hierarchy and halo dumps (AHFGRID1 / AHFHALO1) and the
snapshot container (AHFSNAP3).

The same dump layout is written by the reference harness
(oracle/ref_driver.c), the CPU restatement (oracle/ahf_oracle.c) and this
package's GPU path, so any two of them can be compared field by field.
"""
from __future__ import annotations

import ctypes
import dataclasses
import struct

import numpy as np

SNAP_HEADER = struct.Struct("<8sQddddd ii dd ii ddddd QQQ ii dd")


@dataclasses.dataclass
class Snapshot:
    npart: int
    boxsize: float
    omega0: float
    lambda0: float
    a: float
    pmass: float
    L_dom: int
    L_max: int
    Nth_dom: float
    Nth_ref: float
    minpart: int
    vtune: float
    r_fac: float
    v_fac: float
    phi_fac: float
    hubble: float
    nhalos: int
    n_ipart: int
    seed: int
    config: int
    ovlim: float             # virial overdensity of rem_outsideRvir
    rho_vir: float           # the density it is relative to
    pos: np.ndarray          # (npart, 3) float32, box units
    mom: np.ndarray          # (npart, 3) float32
    ids: np.ndarray          # (npart,) uint64
    centre: np.ndarray       # (nhalos, 3) float64
    gather_rad: np.ndarray   # (nhalos,) float64, box units
    halo_npart: np.ndarray   # (nhalos,) uint64
    ipart: np.ndarray        # (n_ipart,) uint64

    @property
    def halo_offsets(self) -> np.ndarray:
        off = np.zeros(self.nhalos + 1, dtype=np.int64)
        np.cumsum(self.halo_npart.astype(np.int64), out=off[1:])
        return off

    @property
    def rho_fac(self) -> float:
        """pmass / boxsize^3 (src/libahf/ahf_halos.c:204), evaluated as AHF does."""
        b = self.boxsize
        return self.pmass / (b * b * b)


def read_snapshot(path: str) -> Snapshot:
    with open(path, "rb") as f:
        raw = f.read(SNAP_HEADER.size)
        v = SNAP_HEADER.unpack(raw)
        if v[0] != b"AHFSNAP3":
            raise ValueError(f"{path}: not an AHFSNAP3 file")
        (npart, box, om, ol, a, pm, Ld, Lm, nd, nr, mp, _p0, vt, rf, vf, pf, hb,
         nh, ni, seed, cfg, _p1, ovl, rhv) = v[1:]
        pos = np.fromfile(f, dtype=np.float32, count=3 * npart).reshape(npart, 3)
        mom = np.fromfile(f, dtype=np.float32, count=3 * npart).reshape(npart, 3)
        ids = np.fromfile(f, dtype=np.uint64, count=npart)
        cen = np.fromfile(f, dtype=np.float64, count=3 * nh).reshape(nh, 3)
        grad = np.fromfile(f, dtype=np.float64, count=nh)
        hnp = np.fromfile(f, dtype=np.uint64, count=nh)
        ip = np.fromfile(f, dtype=np.uint64, count=ni)
    return Snapshot(npart, box, om, ol, a, pm, Ld, Lm, nd, nr, mp, vt, rf, vf, pf, hb,
                    nh, ni, seed, cfg, ovl, rhv, pos, mom, ids, cen, grad, hnp, ip)


class _SnapData(ctypes.Structure):
    """snap_data of ahf_amd/csrc/snapfile.h: the header bytes, then the
    seven array pointers"""
    _fields_ = [("h", ctypes.c_uint8 * SNAP_HEADER.size)] + [
        (n, ctypes.c_void_p) for n in ("pos", "mom", "id", "centre", "gather_rad", "halo_npart", "ipart")]


class _GenOwner:
    """keeps the generator's arrays alive while a numpy view of them exists"""

    def __init__(self, lib, data):
        self.lib, self.data = lib, data

    def __del__(self):
        self.lib.ahf_snapgen_release(ctypes.byref(self.data))


def _view(owner, addr, ctype, count, dtype, shape):
    buf = (ctype * max(int(count), 1)).from_address(addr)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype)[:int(count)].reshape(shape)


def generate(config: int, side: int = 0, seed: int = 0, lib_path: str | None = None) -> Snapshot:
    """Workload `config` made in memory by the seeded generator
    (ahf_amd/csrc/snapgen.c, the same arrays ahf_snapgen_write puts in a
    file), for sizes where a file is too large to pass around (the 896^3
    workload).  The arrays are views of the generator's memory, released
    when no view is left."""
    import os
    if lib_path is None:
        lib_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libahf_snapgen.so")
    lib = ctypes.CDLL(lib_path)
    lib.ahf_snapgen_make.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
    lib.ahf_snapgen_release.argtypes = [ctypes.c_void_p]
    lib.ahf_snapgen_release.restype = None
    d = _SnapData()
    rc = lib.ahf_snapgen_make(config, side, seed, ctypes.byref(d))
    if rc:
        lib.ahf_snapgen_release(ctypes.byref(d))
        raise RuntimeError(f"snapshot generation failed ({rc})")
    owner = _GenOwner(lib, d)
    v = SNAP_HEADER.unpack(bytes(d.h))
    (npart, box, om, ol, a, pm, Ld, Lm, nd, nr, mp, _p0, vt, rf, vf, pf, hb,
     nh, ni, sd, cfg, _p1, ovl, rhv) = v[1:]
    f32, f64, u64 = ctypes.c_float, ctypes.c_double, ctypes.c_uint64
    pos = _view(owner, d.pos, f32, 3 * npart, np.float32, (npart, 3))
    mom = _view(owner, d.mom, f32, 3 * npart, np.float32, (npart, 3))
    ids = _view(owner, d.id, u64, npart, np.uint64, (npart,))
    cen = _view(owner, d.centre, f64, 3 * nh, np.float64, (nh, 3))
    grad = _view(owner, d.gather_rad, f64, nh, np.float64, (nh,))
    hnp = _view(owner, d.halo_npart, u64, nh, np.uint64, (nh,))
    ip = _view(owner, d.ipart, u64, ni, np.uint64, (ni,))
    return Snapshot(npart, box, om, ol, a, pm, Ld, Lm, nd, nr, mp, vt, rf, vf, pf, hb,
                    nh, ni, sd, cfg, ovl, rhv, pos, mom, ids, cen, grad, hnp, ip)


def write_snapshot(path: str, pos, mom=None, ids=None, *, boxsize=100.0, pmass=1e10,
                   L_dom=16, L_max=None, Nth_dom=5.0, Nth_ref=5.0, minpart=20, vtune=1.5,
                   r_fac=None, v_fac=None, phi_fac=None, hubble=100.0, centres=None,
                   halo_lists=None, seed=0, config=-1, omega0=0.3, lambda0=0.7, a=1.0,
                   ovlim=200.0, rho_vir=2.7755397e11, gather_rad=None):
    """Write a hand-made snapshot (tests).  pos in box units; gather_rad per
    halo (box units, default 0.05)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    n = pos.shape[0]
    mom = np.zeros((n, 3), np.float32) if mom is None else np.ascontiguousarray(mom, np.float32)
    ids = np.arange(n, dtype=np.uint64) if ids is None else np.ascontiguousarray(ids, np.uint64)
    if L_max is None:
        L_max = L_dom << 8
    r_fac = boxsize * a if r_fac is None else r_fac
    v_fac = boxsize / 0.01 / a if v_fac is None else v_fac
    phi_fac = 4.3006485e-9 * pmass / (boxsize * a) if phi_fac is None else phi_fac
    centres = np.zeros((0, 3)) if centres is None else np.asarray(centres, np.float64).reshape(-1, 3)
    halo_lists = [] if halo_lists is None else halo_lists
    hnp = np.array([len(h) for h in halo_lists], dtype=np.uint64)
    grad = (np.full(len(halo_lists), 0.05) if gather_rad is None
            else np.asarray(gather_rad, np.float64).reshape(len(halo_lists)))
    ip = (np.concatenate([np.asarray(h, np.uint64) for h in halo_lists])
          if halo_lists else np.zeros(0, np.uint64))
    hdr = SNAP_HEADER.pack(b"AHFSNAP3", n, boxsize, omega0, lambda0, a, pmass, L_dom, L_max,
                           Nth_dom, Nth_ref, minpart, 0, vtune, r_fac, v_fac, phi_fac, hubble,
                           len(halo_lists), ip.size, seed, config, 0, ovlim, rho_vir)
    with open(path, "wb") as f:
        f.write(hdr)
        pos.tofile(f)
        mom.tofile(f)
        ids.tofile(f)
        centres.astype(np.float64).tofile(f)
        grad.tofile(f)
        hnp.tofile(f)
        ip.tofile(f)


@dataclasses.dataclass
class GridLevel:
    l1dim: int
    size_no_nodes: int
    size_no_part: int
    critdens: float
    masstopartdens: float
    nquads: np.ndarray      # (n, 4) int32: z, y, x, length (traversal order)
    pquads: np.ndarray      # (n, 2) int32: z, length
    cquads: np.ndarray      # (n, 3) int32: z, y, length
    dens: np.ndarray        # (nodes,) float32, traversal order
    counts: np.ndarray      # (nodes,) uint32 list lengths
    lists: np.ndarray       # (sum counts,) uint64 particle indices in list order

    def node_xyz(self) -> np.ndarray:
        """(nodes, 3) int64 coordinates x, y, z in traversal order."""
        nq = self.nquads.astype(np.int64)
        lens = nq[:, 3]
        tot = int(lens.sum())
        rep = np.repeat(np.arange(nq.shape[0]), lens)
        start = np.zeros(nq.shape[0], np.int64)
        np.cumsum(lens[:-1], out=start[1:]) if nq.shape[0] > 1 else None
        off = np.arange(tot, dtype=np.int64) - start[rep]
        out = np.empty((tot, 3), np.int64)
        out[:, 0] = nq[rep, 2] + off
        out[:, 1] = nq[rep, 1]
        out[:, 2] = nq[rep, 0]
        return out


def read_grids(path: str) -> list[GridLevel]:
    levels = []
    with open(path, "rb") as f:
        if f.read(8) != b"AHFGRID1":
            raise ValueError(f"{path}: not an AHFGRID1 dump")
        ng, _dom = struct.unpack("<ii", f.read(8))
        for _ in range(ng):
            l1dim, nn, npt = struct.unpack("<QQQ", f.read(24))
            crit, m2pd = struct.unpack("<dd", f.read(16))
            n_nq, n_nodes, n_listed, n_pq, n_cq = struct.unpack("<QQQQQ", f.read(40))
            nquads = np.fromfile(f, np.int32, 4 * n_nq).reshape(n_nq, 4)
            pquads = np.fromfile(f, np.int32, 2 * n_pq).reshape(n_pq, 2)
            cquads = np.fromfile(f, np.int32, 3 * n_cq).reshape(n_cq, 3)
            dens = np.fromfile(f, np.float32, n_nodes)
            counts = np.fromfile(f, np.uint32, n_nodes)
            lists = np.fromfile(f, np.uint64, n_listed)
            levels.append(GridLevel(l1dim, nn, npt, crit, m2pd, nquads, pquads, cquads,
                                    dens, counts, lists))
    return levels


@dataclasses.dataclass
class HaloResult:
    npart: np.ndarray     # (nhalos,) uint64
    M_vir: np.ndarray
    R_vir: np.ndarray
    Phi0: np.ndarray
    ipart: list           # per-halo uint64 arrays (bound members, in order)
    ovdens: np.ndarray = None   # overdensity at R_vir (0 where never set)


def read_halos(path: str) -> HaloResult:
    with open(path, "rb") as f:
        if f.read(8) != b"AHFHALO1":
            raise ValueError(f"{path}: not an AHFHALO1 dump")
        (nh,) = struct.unpack("<Q", f.read(8))
        npart = np.zeros(nh, np.uint64)
        mv = np.zeros(nh)
        rv = np.zeros(nh)
        p0 = np.zeros(nh)
        od = np.zeros(nh)
        lists = []
        for i in range(nh):
            (n,) = struct.unpack("<Q", f.read(8))
            vals = struct.unpack("<dddd", f.read(32))
            npart[i] = n
            mv[i], rv[i], p0[i], od[i] = vals
            lists.append(np.fromfile(f, np.uint64, n))
    return HaloResult(npart, mv, rv, p0, lists, od)


# HaloProfiles output: the integral properties and the per-bin profile
# values in the order the AHFPROF1 dump writes them
PROF_PROPS = ("nbins", "M_vir", "vel_x", "vel_y", "vel_z", "R_max", "V2_max", "r2", "sigV", "v_esc2", "Ekin",
              "Epot", "axis_x", "axis_y", "axis_z", "E1_x", "E1_y", "E1_z", "E2_x", "E2_y", "E2_z", "E3_x",
              "E3_y", "E3_z", "AngMom_x", "AngMom_y", "AngMom_z", "lambda", "lambdaE", "cNFW", "R1", "cR1",
              "SurfP", "fMhires", "mbp_offset", "pos_mbp_x", "pos_mbp_y", "pos_mbp_z", "vel_mbp_x",
              "vel_mbp_y", "vel_mbp_z", "com_offset", "pos_com_x", "pos_com_y", "pos_com_z")
PROF_BINS = ("r", "npart", "nvpart", "ovdens", "dens", "v2_circ", "v_esc2", "sig_v", "Ekin", "Epot", "Lx", "Ly",
             "Lz", "E1x", "E1y", "E1z", "E2x", "E2y", "E2z", "E3x", "E3y", "E3z", "axis1", "axis2", "axis3")


@dataclasses.dataclass
class ProfileResult:
    npart: np.ndarray       # (nh,) members after the construction
    nbins: np.ndarray       # (nh,) 0 where HaloProfiles skipped the halo
    props: np.ndarray       # (nh, len(PROF_PROPS)), zero rows where nbins == 0
    prof: list              # per halo (nbins, len(PROF_BINS))


def read_profiles(path: str) -> ProfileResult:
    with open(path, "rb") as f:
        if f.read(8) != b"AHFPROF1":
            raise ValueError(f"{path}: not an AHFPROF1 dump")
        (nh,) = struct.unpack("<Q", f.read(8))
        npart = np.zeros(nh, np.uint64)
        nbins = np.zeros(nh, np.int64)
        props = np.zeros((nh, len(PROF_PROPS)))
        prof = []
        for i in range(nh):
            n, nb = struct.unpack("<QQ", f.read(16))
            npart[i], nbins[i] = n, nb
            if nb:
                props[i] = np.fromfile(f, np.float64, len(PROF_PROPS))
                prof.append(np.fromfile(f, np.float64, nb * len(PROF_BINS)).reshape(nb, len(PROF_BINS)))
            else:
                prof.append(np.zeros((0, len(PROF_BINS))))
    return ProfileResult(npart, nbins, props, prof)


def write_profiles(path: str, npart, nbins, props, prof) -> None:
    with open(path, "wb") as f:
        f.write(b"AHFPROF1")
        f.write(struct.pack("<Q", len(npart)))
        for i in range(len(npart)):
            f.write(struct.pack("<QQ", int(npart[i]), int(nbins[i])))
            if nbins[i]:
                f.write(np.ascontiguousarray(props[i], np.float64).tobytes())
                f.write(np.ascontiguousarray(prof[i], np.float64).tobytes())


def compare_grids(a: list[GridLevel], b: list[GridLevel], *, dens_exact=True, rtol=0.0):
    """Return a list of human-readable differences (empty when identical)."""
    diffs = []
    if len(a) != len(b):
        diffs.append(f"number of grids {len(a)} != {len(b)}")
    for li, (x, y) in enumerate(zip(a, b)):
        tag = f"level {li} (l1dim {x.l1dim})"
        for fld in ("l1dim", "size_no_nodes", "size_no_part"):
            if getattr(x, fld) != getattr(y, fld):
                diffs.append(f"{tag}: {fld} {getattr(x, fld)} != {getattr(y, fld)}")
        for fld in ("critdens", "masstopartdens"):
            if getattr(x, fld) != getattr(y, fld):
                diffs.append(f"{tag}: {fld} {getattr(x, fld)!r} != {getattr(y, fld)!r}")
        for fld in ("nquads", "pquads", "cquads"):
            u, v = getattr(x, fld), getattr(y, fld)
            if u.shape != v.shape or not np.array_equal(u, v):
                diffs.append(f"{tag}: {fld} differ ({u.shape} vs {v.shape})")
        if x.dens.shape != y.dens.shape:
            diffs.append(f"{tag}: node count {x.dens.shape[0]} != {y.dens.shape[0]}")
            continue
        if dens_exact:
            bad = np.flatnonzero(x.dens.view(np.uint32) != y.dens.view(np.uint32))
            if bad.size:
                diffs.append(f"{tag}: {bad.size} dens values differ bitwise (first at node {bad[0]}: "
                             f"{x.dens[bad[0]]!r} vs {y.dens[bad[0]]!r})")
        else:
            d = np.abs(x.dens.astype(np.float64) - y.dens.astype(np.float64))
            scale = np.abs(y.dens.astype(np.float64)) + 1.0
            worst = float(np.max(d / scale)) if d.size else 0.0
            if worst > rtol:
                diffs.append(f"{tag}: dens relative error {worst:.3e} > {rtol:.1e}")
        if not np.array_equal(x.counts, y.counts):
            diffs.append(f"{tag}: per-node particle counts differ")
        elif not np.array_equal(x.lists, y.lists):
            diffs.append(f"{tag}: per-node particle lists differ in content or order")
    return diffs


@dataclasses.dataclass
class GridInfo:
    """ahf_gridinfo's result (AHFGINF1 dump): the first AHF grid, the number
    of AHF grids, the spatially isolated refinements per grid and in
    total, the largest refinement overdensity, per AHF grid the node colours
    in traversal order, and per surviving colour (ascending) its
    spatialRefIndex entry: name, grid, index among the grid's refinements,
    periodic x, y, z"""
    min_ref: int
    no_grids: int
    tot_iso: int
    max_ovdens: float
    num_iso: np.ndarray
    colours: list
    refs: np.ndarray


def read_gridinfo(path: str) -> GridInfo:
    with open(path, "rb") as f:
        if f.read(8) != b"AHFGINF1":
            raise ValueError(f"{path}: not an AHFGINF1 dump")
        min_ref, ng, tot, nn = struct.unpack("<iiii", f.read(16))
        (mo,) = struct.unpack("<d", f.read(8))
        num_iso = np.fromfile(f, np.int32, max(ng, 0))
        cols = []
        for _ in range(max(ng, 0)):
            (n,) = struct.unpack("<Q", f.read(8))
            cols.append(np.fromfile(f, np.int32, n))
        refs = np.fromfile(f, np.int32, 6 * nn).reshape(nn, 6)
    return GridInfo(min_ref, ng, tot, mo, num_iso, cols, refs)


def ahf_rho_vir(omega0: float, lambda0: float, a: float) -> float:
    """a^3 calc_rho_vir(a) with the critical density (src/libutility/
    cosmology.c:29-53, 97-112; H0 = 100, Grav = 4.3006485e-9), in AHF's
    operation order"""
    H = 100.0 * np.sqrt(lambda0 * (1. - 1. / (a * a)) + omega0 * (1. / (a * a * a) - 1. / (a * a)) + 1. / (a * a))
    rho_crit = 3. * (H * H) / 8. / 3.14159265358979323846264338 / 4.3006485e-9
    return float((a * a * a) * rho_crit)
