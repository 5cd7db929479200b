"""ahf_amd -- MI355X (gfx950) implementation of AHF's particle->grid deposit,
AMR refinement and halo unbinding hot path.

The compute lives in ``lib/libahf_amd.so`` (hand-written HIP kernels behind
the C ABI of ``include/ahf_gpu.h``).  This module is a thin ctypes host
mirror of that ABI; it never computes anything itself and raises
``NativeLibraryMissing`` when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
MAX_LEVELS = 40
# limits of ahfgpu_unbind_noalloc (include/ahf_gpu.h)
NOALLOC_MAX_HALOS = 65536
NOALLOC_MAX_MEMBERS = 1048576


class NativeLibraryMissing(RuntimeError):
    pass


class AhfGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ahfgpu error {code}: {msg}")
        self.code = code


class AmrParams(ctypes.Structure):
    _fields_ = [("L_dom", ctypes.c_int32), ("L_max", ctypes.c_int32),
                ("Nth_dom", ctypes.c_double), ("Nth_ref", ctypes.c_double)]


class AmrInfo(ctypes.Structure):
    _fields_ = [("n_levels", ctypes.c_int32), ("rolled_back", ctypes.c_int32),
                ("n_nodes", ctypes.c_int64 * MAX_LEVELS), ("n_part", ctypes.c_int64 * MAX_LEVELS),
                ("deposits", ctypes.c_int64)]


class UnbindParams(ctypes.Structure):
    _fields_ = [("minpart", ctypes.c_int32), ("pad", ctypes.c_int32), ("vtune", ctypes.c_double),
                ("r_fac", ctypes.c_double), ("v_fac", ctypes.c_double), ("phi_fac", ctypes.c_double),
                ("hubble", ctypes.c_double), ("ovlim", ctypes.c_double), ("rho_vir", ctypes.c_double),
                ("rho_fac", ctypes.c_double)]


class UnbindInfo(ctypes.Structure):
    _fields_ = [("n_halos", ctypes.c_int64), ("n_member_in", ctypes.c_int64),
                ("n_member_out", ctypes.c_int64), ("particle_iterations", ctypes.c_int64),
                ("n_exact_replays", ctypes.c_int64), ("max_iterations", ctypes.c_int32),
                ("max_sweeps", ctypes.c_int32), ("n_tie_halos", ctypes.c_int64)]


class ProfileParams(ctypes.Structure):
    _fields_ = [("minpart", ctypes.c_int32), ("pad", ctypes.c_int32), ("r_fac", ctypes.c_double),
                ("v_fac", ctypes.c_double), ("m_fac", ctypes.c_double), ("phi_fac", ctypes.c_double),
                ("hubble", ctypes.c_double)]


class Schedule(ctypes.Structure):
    """ahfgpu_schedule: thresholds between equivalent kernels (0: default)"""
    _fields_ = [("split_min", ctypes.c_int64), ("prof_split", ctypes.c_int64), ("big_halo", ctypes.c_int64),
                ("wave_halo", ctypes.c_int64), ("kmax_big", ctypes.c_int32), ("kmax_small", ctypes.c_int32),
                ("march_zn", ctypes.c_int32), ("graph_batch", ctypes.c_int32), ("split_spec", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("prof_tiled", ctypes.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class GridinfoParams(ctypes.Structure):
    _fields_ = [("ovlim", ctypes.c_double), ("rho_vir", ctypes.c_double), ("refine_mass", ctypes.c_double),
                ("boxsize", ctypes.c_double)]


class GridinfoInfo(ctypes.Structure):
    _fields_ = [("min_ref", ctypes.c_int32), ("no_grids", ctypes.c_int32), ("tot_iso", ctypes.c_int32),
                ("n_colours", ctypes.c_int32), ("max_ovdens", ctypes.c_double),
                ("num_iso", ctypes.c_int32 * MAX_LEVELS), ("device_ms", ctypes.c_double)]


_lib = None


def lib_path(name="libahf_amd.so"):
    return os.path.join(LIBDIR, name)


def load():
    """Load the HIP library (raises NativeLibraryMissing if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise NativeLibraryMissing(f"{path} is not built; run __graft_entry__.build() or "
                                   f"python ahf_amd/build.py")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.ahfgpu_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    lib.ahfgpu_destroy.argtypes = [P]
    lib.ahfgpu_destroy.restype = None
    lib.ahfgpu_last_error.argtypes = [P]
    lib.ahfgpu_last_error.restype = ctypes.c_char_p
    lib.ahfgpu_set_particles.argtypes = [P, P, P, ctypes.c_uint64]
    lib.ahfgpu_amr_build.argtypes = [P, ctypes.POINTER(AmrParams), ctypes.POINTER(AmrInfo)]
    lib.ahfgpu_amr_dump.argtypes = [P, ctypes.c_char_p]
    lib.ahfgpu_set_gather_mode.argtypes = [P, ctypes.c_int]
    lib.ahfgpu_amr_level_sizes.argtypes = [P, ctypes.c_int] + [P] * 5
    lib.ahfgpu_amr_level_copy.argtypes = [P, ctypes.c_int] + [P] * 5
    lib.ahfgpu_last_timing.argtypes = [P, P, P, P, P]
    lib.ahfgpu_set_halos.argtypes = [P, P, P, P, ctypes.c_uint64]
    lib.ahfgpu_unbind.argtypes = [P, ctypes.POINTER(UnbindParams), ctypes.c_int, ctypes.c_double,
                                  ctypes.POINTER(UnbindInfo)]
    lib.ahfgpu_unbind_results.argtypes = [P, P, P, P]
    lib.ahfgpu_release_halos.argtypes = [P]
    lib.ahfgpu_release_halos.restype = None
    lib.ahfgpu_set_unbind_split.argtypes = [P, ctypes.c_int64]
    lib.ahfgpu_unbind_noalloc.argtypes = [P, ctypes.POINTER(UnbindParams), ctypes.c_int, ctypes.c_double,
                                          P, P, P, ctypes.c_int64, P, P, P]
    lib.ahfgpu_set_sfc_keys.argtypes = [P, P, ctypes.c_uint64]
    lib.ahfgpu_sfc_keys.argtypes = [P, P]
    lib.ahfgpu_gather_halos.argtypes = [P, P, P, P, ctypes.c_uint64]
    lib.ahfgpu_halo_members.argtypes = [P, P, P]
    lib.ahfgpu_timer_count.argtypes = []
    lib.ahfgpu_timer_get.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), P, P, P]
    lib.ahfgpu_halo_profiles.argtypes = [P, ctypes.POINTER(ProfileParams), P, P]
    lib.ahfgpu_profile_results.argtypes = [P, P, P]
    lib.ahfgpu_set_profile_split.argtypes = [P, ctypes.c_int64]
    lib.ahfgpu_set_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
    lib.ahfgpu_get_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
    lib.ahfgpu_set_timing.argtypes = [P, ctypes.c_int]
    lib.ahfgpu_gridinfo.argtypes = [P, ctypes.POINTER(GridinfoParams), ctypes.POINTER(GridinfoInfo)]
    lib.ahfgpu_gridinfo_results.argtypes = [P, ctypes.c_int, P, P]
    _lib = lib
    return lib


class Context:
    """One device context of the HIP library (resident particles + results)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.ahfgpu_create(ctypes.byref(h), device)
        if rc:
            raise AhfGpuError(rc, f"cannot create a context on device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ahfgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise AhfGpuError(rc, self.lib.ahfgpu_last_error(self.h).decode(errors="replace"))

    def set_particles(self, pos, mom=None):
        import numpy as np
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        momp = None
        if mom is not None:
            mom = np.ascontiguousarray(mom, dtype=np.float32)
            momp = mom.ctypes.data
        self._check(self.lib.ahfgpu_set_particles(self.h, pos.ctypes.data, momp, pos.shape[0]))
        self._n = pos.shape[0]

    def amr_build(self, L_dom, L_max, Nth_dom, Nth_ref):
        p = AmrParams(L_dom, L_max, Nth_dom, Nth_ref)
        info = AmrInfo()
        self._check(self.lib.ahfgpu_amr_build(self.h, ctypes.byref(p), ctypes.byref(info)))
        return info

    GATHER_AUTO, GATHER_SOURCE, GATHER_FLAT, GATHER_WAVE, GATHER_FLAT_WIDE = 0, 1, 2, 3, 4

    def set_unbind_split(self, min_members):
        """haloes of at least min_members go through the split unbinding path
        (0: default, negative: never); same results either way"""
        self._check(self.lib.ahfgpu_set_unbind_split(self.h, int(min_members)))

    def schedule(self):
        """the kernel schedule in force (ahfgpu_schedule fields, 0 = default)"""
        sc = Schedule()
        self._check(self.lib.ahfgpu_get_schedule(self.h, ctypes.byref(sc)))
        return sc.as_dict()

    def set_schedule(self, **fields):
        """change some schedule fields (the others keep their value)"""
        sc = Schedule()
        self._check(self.lib.ahfgpu_get_schedule(self.h, ctypes.byref(sc)))
        for k, v in fields.items():
            setattr(sc, k, int(v))
        self._check(self.lib.ahfgpu_set_schedule(self.h, ctypes.byref(sc)))

    def set_gather_mode(self, mode):
        """deposit kernel choice (same results in every mode)"""
        self._check(self.lib.ahfgpu_set_gather_mode(self.h, int(mode)))

    def amr_level(self, lev, lists=True):
        """level `lev` of the last hierarchy: dict of L, critdens,
        masstopartdens, oct_key (n_oct), oct_link, dens (8 n_oct, octet-major
        nodes) and, with lists, list0 (8 n_oct + 1) and list (particle
        indices in AHF's list order)"""
        import numpy as np
        L, no, nl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        crit, m2pd = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.ahfgpu_amr_level_sizes(self.h, lev, ctypes.byref(L), ctypes.byref(no), ctypes.byref(nl),
                                                    ctypes.byref(crit), ctypes.byref(m2pd)))
        n_oct = no.value
        key = np.zeros(max(n_oct, 1), np.uint64)
        link = np.zeros(max(n_oct, 1), np.uint8)
        dens = np.zeros(max(8 * n_oct, 1), np.float32)
        out = {"L": L.value, "critdens": crit.value, "masstopartdens": m2pd.value}
        if lists:
            list0 = np.zeros(8 * n_oct + 1, np.int64)
            lst = np.zeros(max(nl.value, 1), np.uint32)
            self._check(self.lib.ahfgpu_amr_level_copy(self.h, lev, key.ctypes.data, link.ctypes.data,
                                                       dens.ctypes.data, list0.ctypes.data, lst.ctypes.data))
            out.update(list0=list0, list=lst[:int(list0[-1])])
        else:
            self._check(self.lib.ahfgpu_amr_level_copy(self.h, lev, key.ctypes.data, link.ctypes.data,
                                                       dens.ctypes.data, None, None))
        out.update(oct_key=key[:n_oct], oct_link=link[:n_oct], dens=dens[:8 * n_oct])
        return out

    def amr_dump(self, path):
        self._check(self.lib.ahfgpu_amr_dump(self.h, path.encode()))

    def set_timing(self, on):
        """stage timers on (the default) or off (the timings then read 0)"""
        self._check(self.lib.ahfgpu_set_timing(self.h, 1 if on else 0))

    def last_timing(self):
        a, u, g = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.lib.ahfgpu_last_timing(self.h, ctypes.byref(a), ctypes.byref(u), ctypes.byref(g),
                                                ctypes.byref(n)))
        return {"amr_ms": a.value, "unbind_ms": u.value, "gather_ms": g.value, "gather_launches": n.value}

    def set_halos(self, centres, halo_npart, ipart):
        import numpy as np
        self._cen = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        self._hnp = np.ascontiguousarray(halo_npart, dtype=np.uint64)
        ip = np.ascontiguousarray(ipart, dtype=np.uint64)
        self._check(self.lib.ahfgpu_set_halos(self.h, self._cen.ctypes.data, self._hnp.ctypes.data,
                                              ip.ctypes.data, self._hnp.shape[0]))

    def sfc_keys(self, copy=False):
        """SFC keys of the resident particles, computed on the device (they
        must come out in order); returned when copy is set"""
        import numpy as np
        out = np.zeros(self._n, np.uint64) if copy else None
        self._check(self.lib.ahfgpu_sfc_keys(self.h, out.ctypes.data if copy else None))
        return out

    def set_sfc_keys(self, keys):
        import numpy as np
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        self._check(self.lib.ahfgpu_set_sfc_keys(self.h, k.ctypes.data, k.shape[0]))

    def gather_halos(self, centres, gather_rad, skip=None):
        """member lists of every halo from its centre and gather radius (box units)"""
        import numpy as np
        self._cen = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        rad = np.ascontiguousarray(gather_rad, dtype=np.float64)
        nh = self._cen.shape[0]
        sk = None if skip is None else np.ascontiguousarray(skip, dtype=np.uint8)
        self._check(self.lib.ahfgpu_gather_halos(self.h, self._cen.ctypes.data, rad.ctypes.data,
                                                 None if sk is None else sk.ctypes.data, nh))
        hnp = np.zeros(nh, np.uint64)
        self._check(self.lib.ahfgpu_halo_members(self.h, hnp.ctypes.data, None))
        self._hnp = hnp

    def halo_members(self):
        """(npart, list of member arrays) of the resident halo lists"""
        import numpy as np
        nh = self._hnp.shape[0]
        npart = np.zeros(nh, np.uint64)
        ip = np.zeros(int(self._hnp.sum()) + 1, np.uint64)
        self._check(self.lib.ahfgpu_halo_members(self.h, npart.ctypes.data, ip.ctypes.data))
        off = np.concatenate([[0], np.cumsum(npart.astype(np.int64))])
        return npart, [ip[off[h]:off[h + 1]].copy() for h in range(nh)]

    # ahfgpu_unbind modes (or-ed: BATCHED/PLAIN with one step selection)
    UNBIND_BATCHED, UNBIND_PLAIN = 0, 1
    UNBIND_SORT_ONLY, UNBIND_NO_SORT, UNBIND_WITH_RVIR, UNBIND_RVIR_ONLY = 4, 8, 16, 32
    # bits of props[:, 5]: the fields some step wrote
    SET_MVIR, SET_RVIR, SET_PHI0, SET_OVDENS = 1, 2, 4, 8
    PROPS = 6

    def unbind(self, minpart, vtune, r_fac, v_fac, phi_fac, hubble, mode=0, band=0.0,
               ovlim=0.0, rho_vir=1.0, rho_fac=0.0):
        """ovlim, rho_vir, rho_fac: the rem_outsideRvir cut (modes WITH_RVIR, RVIR_ONLY)"""
        p = UnbindParams(minpart, 0, vtune, r_fac, v_fac, phi_fac, hubble, ovlim, rho_vir, rho_fac)
        info = UnbindInfo()
        self._check(self.lib.ahfgpu_unbind(self.h, ctypes.byref(p), mode, band, ctypes.byref(info)))
        return info

    def unbind_results(self):
        """(npart, props[nh,6] = M_vir, R_vir, Phi0, ovdens, iterations, SET_* mask, list of member arrays)"""
        import numpy as np
        nh = self._hnp.shape[0]
        npart = np.zeros(nh, np.uint64)
        props = np.zeros((nh, self.PROPS), np.float64)
        ip = np.zeros(int(self._hnp.sum()) + 1, np.uint64)
        self._check(self.lib.ahfgpu_unbind_results(self.h, npart.ctypes.data, props.ctypes.data, ip.ctypes.data))
        off = np.concatenate([[0], np.cumsum(self._hnp.astype(np.int64))])
        lists = [ip[off[h]:off[h] + int(npart[h])].copy() for h in range(nh)]
        return npart, props, lists

    PROF_NQ, PROF_NP = 25, 45

    def set_profile_split(self, min_members):
        """haloes of at least min_members are profiled by four wavefronts
        (0: default, negative: never); same values either way"""
        self._check(self.lib.ahfgpu_set_profile_split(self.h, int(min_members)))

    def halo_profiles(self, minpart, r_fac, v_fac, m_fac, phi_fac, hubble, phi0_rvir=None):
        """HaloProfiles of every resident halo: of the last unbinding result,
        or (phi0_rvir given, shape (nh, 2)) of the lists as given.  Returns
        (nbins[nh], props[nh, PROF_NP], list of per-halo (nbins, PROF_NQ))"""
        import numpy as np
        nh = self._hnp.shape[0]
        p = ProfileParams(minpart, 0, r_fac, v_fac, m_fac, phi_fac, hubble)
        nb = np.zeros(nh, np.int32)
        pr = None if phi0_rvir is None else np.ascontiguousarray(phi0_rvir, np.float64).reshape(nh, 2)
        self._check(self.lib.ahfgpu_halo_profiles(self.h, ctypes.byref(p), None if pr is None else pr.ctypes.data,
                                                  nb.ctypes.data))
        props = np.zeros((nh, self.PROF_NP))
        prof = np.zeros((int(nb.sum()) + 1, self.PROF_NQ))
        self._check(self.lib.ahfgpu_profile_results(self.h, props.ctypes.data, prof.ctypes.data))
        off = np.concatenate([[0], np.cumsum(nb.astype(np.int64))])
        return nb, props, [prof[off[h]:off[h + 1]].copy() for h in range(nh)]

    def gridinfo(self, ovlim, rho_vir, refine_mass, boxsize):
        """ahf_gridinfo on the last hierarchy (src/libahf/ahf_gridinfo.c:59-860):
        a dumps.GridInfo whose colours are per AHF grid in octet-major node
        order (the order of amr_level(min_ref + g)), refs the surviving
        refinements (colour, grid, isoRefIndex, periodic x, y, z) in AHF's
        list order; the device milliseconds in .device_ms"""
        import numpy as np
        from . import dumps
        info = GridinfoInfo()
        self._check(self.lib.ahfgpu_gridinfo(self.h, ctypes.byref(GridinfoParams(ovlim, rho_vir, refine_mass, boxsize)),
                                             ctypes.byref(info)))
        cols, refs = [], []
        for g in range(info.no_grids):
            n_oct = ctypes.c_int64()
            self._check(self.lib.ahfgpu_amr_level_sizes(self.h, info.min_ref + g, None, ctypes.byref(n_oct), None,
                                                        None, None))
            col = np.zeros(max(8 * n_oct.value, 1), np.int32)
            ref = np.zeros((max(int(info.num_iso[g]), 1), 6), np.int32)
            self._check(self.lib.ahfgpu_gridinfo_results(self.h, g, col.ctypes.data, ref.ctypes.data))
            cols.append(col[:8 * n_oct.value])
            refs.append(ref[:int(info.num_iso[g])])
        out = dumps.GridInfo(info.min_ref, info.no_grids, info.tot_iso, info.max_ovdens,
                             np.array(info.num_iso[:info.no_grids], np.int32), cols,
                             np.concatenate(refs) if refs else np.zeros((0, 6), np.int32))
        out.device_ms = info.device_ms
        out.n_colours = info.n_colours
        return out

    def timers(self):
        """{stage: (ms, launches, algorithmic bytes)} of the last build + unbinding"""
        out = {}
        for i in range(self.lib.ahfgpu_timer_count()):
            name = ctypes.c_char_p()
            ms, b = ctypes.c_double(), ctypes.c_double()
            n = ctypes.c_int64()
            self._check(self.lib.ahfgpu_timer_get(self.h, i, ctypes.byref(name), ctypes.byref(ms),
                                                  ctypes.byref(n), ctypes.byref(b)))
            out[name.value.decode()] = (ms.value, n.value, b.value)
        return out
