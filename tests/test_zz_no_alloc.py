"""ahfgpu_unbind_noalloc (the unbinding entry point that allocates no device
memory) against ahfgpu_unbind + ahfgpu_unbind_results on the same haloes:
identical members, in order, and identical per-halo fields, on two calls in
a row, the second with a smaller input (the static scratch keeps the first
call's contents)."""
import ctypes

import numpy as np
import pytest

from ahf_amd import dumps

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """device buffers for the test, through the HIP runtime the library
    itself loaded"""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)) == 0
        self.bufs.append(p)
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        if arr.nbytes:
            assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def download(self, p, dtype, count):
        out = np.zeros(max(int(count), 1), dtype)
        if count:
            assert self.rt.hipMemcpy(out.ctypes.data, p, out.itemsize * int(count), D2H) == 0
        return out[:int(count)]

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def ctx():
    import ahf_amd
    c = ahf_amd.Context(0)
    yield c
    c.close()


def _prefix(s, max_members, max_halos):
    """the first haloes of the snapshot within the entry point's limits"""
    off = s.halo_offsets
    nh = int(min(s.nhalos, max_halos))
    while nh > 0 and off[nh] > max_members:
        nh -= 1
    return nh


def _existing(ctx, s, nh, mode):
    """the existing entry point on haloes [0, nh), every halo in the
    persistent kernel (the split path off, as in the no-alloc call)"""
    ctx.set_unbind_split(-1)
    ctx.set_halos(s.centre[:nh], s.halo_npart[:nh], s.ipart[:int(s.halo_offsets[nh])])
    ctx.unbind(s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, mode=mode,
               ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
    out = ctx.unbind_results()
    ctx.set_unbind_split(0)
    return out


def _noalloc(ctx, hip, s, nh, mode):
    import ahf_amd
    lib = ctx.lib
    lib.ahfgpu_unbind_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ahf_amd.UnbindParams), ctypes.c_int,
                                          ctypes.c_double] + [ctypes.c_void_p] * 3 + [ctypes.c_int64] + \
                                         [ctypes.c_void_p] * 3
    off = s.halo_offsets[:nh + 1].astype(np.int64)
    ne = int(off[nh])
    d_cen = hip.upload(s.centre[:nh].astype(np.float64))
    d_off = hip.upload(off)
    d_mem = hip.upload(s.ipart[:ne].astype(np.uint32))
    d_np = hip.alloc(8 * nh)
    d_props = hip.alloc(8 * ctx.PROPS * nh)
    d_out = hip.alloc(4 * ne)
    p = ahf_amd.UnbindParams(s.minpart, 0, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble, s.ovlim, s.rho_vir,
                             s.rho_fac)
    print(f"[no-alloc] call: {nh} haloes, {ne} members, mode {mode}", flush=True)
    rc = lib.ahfgpu_unbind_noalloc(ctx.h, ctypes.byref(p), mode, 0.0, d_cen, d_off, d_mem, nh, d_np, d_props, d_out)
    assert rc == 0, lib.ahfgpu_last_error(ctx.h)
    npart = hip.download(d_np, np.uint64, nh)
    props = hip.download(d_props, np.float64, ctx.PROPS * nh).reshape(nh, ctx.PROPS)
    mem = hip.download(d_out, np.uint32, ne)
    lists = [mem[off[h]:off[h] + int(npart[h])].astype(np.uint64) for h in range(nh)]
    return npart, props, lists


def _same(a, b, tag):
    npart_a, props_a, lists_a = a
    npart_b, props_b, lists_b = b
    assert np.array_equal(npart_a, npart_b), tag
    # every field bit for bit (R_vir only where a step set it)
    set_ = (props_a[:, 5].astype(np.int64) & 2) != 0
    assert np.array_equal(props_a[:, 5], props_b[:, 5]), tag
    for j in (0, 2, 3, 4):
        assert np.array_equal(props_a[:, j], props_b[:, j]), f"{tag}: field {j}"
    assert np.array_equal(props_a[set_, 1], props_b[set_, 1]), tag
    bad = [h for h in range(len(lists_a)) if not np.array_equal(lists_a[h], lists_b[h])]
    assert not bad, f"{tag}: {len(bad)} haloes differ, first {bad[:5]}"


# cfg -1: haloes with members at equal distances (the indexx order of ties)
@pytest.mark.parametrize("cfg,side,seed,mode", [(2, 1500, 1002, 0), (1, 64, 1001, 16), (0, 64, 1000, 4),
                                                (-1, 0, 2, 0), (-1, 0, 4, 16)])
def test_noalloc_matches_unbind(snapgen, ctx, tmp_path, cfg, side, seed, mode):
    import ahf_amd
    snap = str(tmp_path / "n.snap")
    if cfg < 0:
        from helpers import tie_snapshot
        tie_snapshot(snap, seed, big=True)
    else:
        assert snapgen.ahf_snapgen_write(cfg, side, seed, snap.encode()) == 0
    s = dumps.read_snapshot(snap)
    ctx.set_particles(s.pos, s.mom)
    hip = Hip()
    try:
        nh1 = _prefix(s, ahf_amd.NOALLOC_MAX_MEMBERS, ahf_amd.NOALLOC_MAX_HALOS)
        assert nh1 >= 4
        nh2 = nh1 // 2
        for nh in (nh1, nh2):
            ref = _existing(ctx, s, nh, mode)
            got = _noalloc(ctx, hip, s, nh, mode)
            _same(ref, got, f"cfg{cfg} {nh} haloes mode {mode}")
    finally:
        hip.close()


def test_noalloc_rejects_too_many_haloes(ctx):
    import ahf_amd
    lib = ctx.lib
    lib.ahfgpu_unbind_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ahf_amd.UnbindParams), ctypes.c_int,
                                          ctypes.c_double] + [ctypes.c_void_p] * 3 + [ctypes.c_int64] + \
                                         [ctypes.c_void_p] * 3
    p = ahf_amd.UnbindParams(20, 0, 1.5, 1.0, 1.0, 1.0, 0.0, 200.0, 1.0, 1.0)
    rc = lib.ahfgpu_unbind_noalloc(ctx.h, ctypes.byref(p), 0, 0.0, None, None, None,
                                   ahf_amd.NOALLOC_MAX_HALOS + 1, None, None, None)
    assert rc == -1
    assert b"at most" in lib.ahfgpu_last_error(ctx.h)
