#!/usr/bin/env python3
"""Benchmark of the MI355X deposit + unbinding path on one GPU.

One step = one pass of the hot path over the resident snapshot:
  the full AMR hierarchy (binning, TSC deposit, refinement flagging,
  fine-grid construction and relinking on every level) followed by the
  radial sort and unbinding of every halo (batched persistent form).
Inputs (particles, halo member lists) are resident in HBM before timing.

Default workload: 256^3 = 16,777,216 particles in a 100 Mpc/h box with 2000
NFW hosts (dN/dM ~ M^-1.9 on [1e10,1e14] Msun/h, hosts above 1e12 carrying
5-20 subhaloes with 10% of their mass, 10% interlopers), LgridDomain 128,
LgridMax 32768, NperDomCell = NperRefCell = 5, VescTune 1.5, NminPerHalo 20.

Prints one JSON line (rank 0).  Run `python bench.py --help` for options.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import tempfile
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s (TSC deposit + unbinding) at 256^3; batched vs plain GPU vs CPU"
HBM_PEAK_GBS = 8000.0
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ahf_ref")


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def snapshot_path(cfg, side, seed):
    d = os.environ.get("AHF_BENCH_DIR", "/tmp")
    return os.path.join(d, f"ahf_bench_s4_cfg{cfg}_{side}_{seed}.snap")


def make_snapshot(cfg, side, seed):
    import ctypes
    path = snapshot_path(cfg, side, seed)
    if os.path.exists(path):
        return path
    lib_path = os.path.join(ROOT, "ahf_amd", "lib", "libahf_snapgen.so")
    if not os.path.exists(lib_path):
        from ahf_amd import build
        build.build_snapgen()
    lib = ctypes.CDLL(lib_path)
    lib.ahf_snapgen_write.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p]
    tmp = path + f".{os.getpid()}.tmp"
    t0 = time.time()
    rc = lib.ahf_snapgen_write(cfg, side, seed, tmp.encode())
    if rc:
        raise RuntimeError(f"snapshot generation failed ({rc})")
    os.replace(tmp, path)
    log(f"generated {path} in {time.time() - t0:.1f}s")
    return path


def cpu_baseline(snap, threads):
    """The reference's own compiled functions (oracle/_ref/ahf_ref) on the
    same snapshot: AMR build single-threaded as in the reference's default
    build, unbinding with OpenMP over haloes."""
    if not os.path.exists(REF_BIN):
        return None
    r = subprocess.run([REF_BIN, "time", snap, "--threads", str(threads), "--reps", "1"],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"reference timing failed: {r.stderr[-500:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


LINKED_BIN = os.path.join(ROOT, "oracle", "_ref", "ahf_ref_linked")


def dropin_figures(snap, threads):
    """AHF's own objects linked with the library by the integration recipe
    (oracle/_ref/ahf_ref_linked) on the same snapshot file: gen_domgrids,
    ll, zero_dens, assign_npart and gen_AMRhierarchy end to end from and
    into AHF's structs, and the per-halo sort_halo_particles + rem_unbound
    calls of AHF's OpenMP halo loop on `threads` threads (best of 2 runs;
    the first run also creates the device context)"""
    if not snap or not os.path.exists(LINKED_BIN):
        return None
    r = subprocess.run([LINKED_BIN, "time", snap, "--threads", str(threads), "--reps", "2", "--skip-1thread"],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"drop-in timing failed: {r.stderr[-500:]}")
        return None
    t = json.loads(r.stdout.strip().splitlines()[-1])
    return {"amr_seconds": t["amr_seconds"], "unbind_seconds": t["unbind_seconds"], "threads": threads,
            "harness": "oracle/_ref/ahf_ref_linked time (the reference's objects + libahf_amd.so)"}


def gridinfo_figures(ctx, s, reps=5):
    """ahf_gridinfo (src/libahf/ahf_gridinfo.c:59-860) on the resident
    hierarchy, with the snapshot's virial overdensity: device milliseconds
    (median of reps) and nodes coloured per second"""
    import numpy as np

    from ahf_amd import dumps
    rho = dumps.ahf_rho_vir(s.omega0, s.lambda0, s.a)
    ms = []
    for _ in range(reps):
        gi = ctx.gridinfo(s.ovlim, rho, s.pmass, s.boxsize)
        ms.append(gi.device_ms)
    nodes = int(sum(len(c) for c in gi.colours))
    out = {"device_ms": float(np.median(ms)), "nodes": nodes, "no_grids": gi.no_grids, "min_ref": gi.min_ref,
           "refinements": gi.tot_iso, "nodes_per_s": nodes / (float(np.median(ms)) * 1e-3)}
    return out


def gridinfo_cpu(out, cpu_snap):
    """the reference's own ahf_gridinfo on the CPU sample (after its AMR
    build, which is not timed), added to the device figures"""
    from ahf_amd import dumps
    out = dict(out)
    if cpu_snap and os.path.exists(REF_BIN):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "g.ginf")
            r = subprocess.run([REF_BIN, "gridinfo", cpu_snap, path], capture_output=True, text=True, timeout=600)
            if r.returncode == 0:
                t = json.loads(r.stdout.strip().splitlines()[-1])["gridinfo_seconds"]
                g = dumps.read_gridinfo(path)
                n = int(sum(len(c) for c in g.colours))
                out.update(cpu_seconds=t, cpu_nodes=n, cpu_nodes_per_s=n / max(t, 1e-9),
                           cpu_sample=f"{os.path.basename(cpu_snap)}, {n} nodes of {g.no_grids} grids, 1 thread")
            else:
                log(f"reference gridinfo failed: {r.stderr[-300:]}")
    return out


def copy_probe(nbytes=1 << 30, reps=5):
    """Device-to-device copies of nbytes through the HIP runtime: the copy
    ceiling of this GPU (read + write bytes / time) and, under
    rocprofv3 --pmc, the reference dispatches that calibrate FETCH_SIZE and
    WRITE_SIZE (each copy is one runtime copy kernel moving nbytes each way)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    P = ctypes.c_void_p
    a, b = P(), P()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        return None
    try:
        hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
        e0, e1 = P(), P()
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        return {"bytes_per_copy": nbytes, "copies": reps + 1,
                "GBs": 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9}
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def load_traffic():
    """HBM bytes per deposit launch (one grid level) of this workload,
    measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over the deposit
    kernels and scaled by the copy-kernel calibration
    (profiles/deposit_traffic.json, written by
    python -m ahf_amd.rocprof_summary traffic)."""
    path = os.path.join(ROOT, "profiles", "deposit_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def unbinding_workload(seed=1002, reps=3, cpu=True, threads=16):
    """The unbinding-throughput workload: 100,000 isolated NFW
    haloes, N ~ N^-2 on [20, 2e6] members, 15% of members above escape speed
    and 5% within 1e-4 of it; sort + unbinding of all haloes, batched, with
    the reference's own sort_halo_particles + rem_unbound (OpenMP over
    haloes) timed on the same file."""
    import numpy as np

    import ahf_amd
    from ahf_amd import dumps
    snap = make_snapshot(2, 0, seed)
    s = dumps.read_snapshot(snap)
    ctx = ahf_amd.Context(0)
    ctx.set_particles(s.pos, s.mom)
    ctx.set_halos(s.centre, s.halo_npart, s.ipart)
    args = (s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble)
    ctx.unbind(*args)
    ms = []
    for _ in range(reps):
        info = ctx.unbind(*args)
        ms.append(ctx.last_timing()["unbind_ms"])
    # the construction steps of ahf_halos_sfc_constructHalo on the same lists
    # (sort, rem_outsideRvir, rem_unbound, rem_outsideRvir), then HaloProfiles
    pargs = (s.minpart, s.r_fac, s.v_fac, s.pmass, s.phi_fac, s.hubble)
    kw = dict(mode=ctx.UNBIND_WITH_RVIR, ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
    cms, pms = [], []
    for _ in range(reps):
        ctx.unbind(*args, **kw)
        cms.append(ctx.last_timing()["unbind_ms"])
        nb, _, _ = ctx.halo_profiles(*pargs)
        pms.append(ctx.timers()["halo_profiles"][0])
    ctx.close()
    t = float(np.median(ms))
    out = {"workload": f"{s.nhalos} isolated NFW haloes, {s.n_ipart} members (largest {int(s.halo_npart.max())}), "
                       f"seed {seed}",
           "gpu_ms": t, "members_per_s": s.n_ipart / (t * 1e-3),
           "particle_iterations_per_s": info.particle_iterations / (t * 1e-3),
           "halos_per_s": s.nhalos / (t * 1e-3), "exact_replays": int(info.n_exact_replays),
           "max_iterations": int(info.max_iterations),
           "construct_ms": float(np.median(cms)), "profiles_ms": float(np.median(pms)),
           "profiled_halos": int((nb > 0).sum())}
    if cpu and os.path.exists(REF_BIN):
        r = subprocess.run([REF_BIN, "time", snap, "--threads", str(threads), "--skip-amr"],
                           capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            rb = json.loads(r.stdout.strip().splitlines()[-1])
            out["cpu_reference_ms"] = 1e3 * rb["unbind_seconds"]
            out["cpu_reference_1thread_ms"] = 1e3 * rb.get("unbind_seconds_1thread", 0.0)
            out["cpu_threads"] = threads
            out["vs_cpu"] = out["cpu_reference_ms"] / t
        with tempfile.TemporaryDirectory() as d:
            r = subprocess.run([REF_BIN, "profiles", snap, os.path.join(d, "h.prof"), "--threads", str(threads)],
                               capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            rb = json.loads(r.stdout.strip().splitlines()[-1])
            out["cpu_reference_construct_ms"] = 1e3 * rb["construct_seconds"]
            out["cpu_reference_profiles_ms"] = 1e3 * rb["profiles_seconds"]
    return out


def construction_workload(snap, reps=3, cpu=True, threads=16):
    """The halo construction of ahf_halos_sfc_constructHalo for every halo of
    the box workload: gatherParts from each centre and gather radius, sort,
    rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1), then HaloProfiles,
    with the member lists resident on the device throughout.  The reference's
    own six functions per halo (OpenMP over haloes) are timed on the same
    file, the profiles apart."""
    import numpy as np

    import ahf_amd
    from ahf_amd import dumps
    s = dumps.read_snapshot(snap)
    ctx = ahf_amd.Context(0)
    ctx.set_particles(s.pos, s.mom)
    tk = time.perf_counter()
    ctx.sfc_keys()
    keys_ms = 1e3 * (time.perf_counter() - tk)
    skip = (s.halo_npart == 0).astype(np.uint8)
    kw = dict(mode=ctx.UNBIND_WITH_RVIR, ovlim=s.ovlim, rho_vir=s.rho_vir, rho_fac=s.rho_fac)
    args = (s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble)

    pargs = (s.minpart, s.r_fac, s.v_fac, s.pmass, s.phi_fac, s.hubble)

    def run():
        ctx.gather_halos(s.centre, s.gather_rad, skip)
        gathered = int(ctx._hnp.sum())
        info = ctx.unbind(*args, **kw)
        t_u = ctx.last_timing()["unbind_ms"]
        nb, _, _ = ctx.halo_profiles(*pargs)
        return gathered, info, t_u, nb

    run()
    g_ms, u_ms, p_ms, wall = [], [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        gathered, info, t_u, nb = run()
        wall.append(1e3 * (time.perf_counter() - t0))
        g_ms.append(ctx.timers()["halo_gather"][0])
        u_ms.append(t_u)
        p_ms.append(ctx.timers()["halo_profiles"][0])
    ctx.close()
    g, u, pm = float(np.median(g_ms)), float(np.median(u_ms)), float(np.median(p_ms))
    out = {"workload": f"{s.nhalos} haloes of the box workload gathered from their centres and gather radii",
           "gathered_members": gathered, "final_members": int(info.n_member_out),
           "profiled_halos": int((nb > 0).sum()), "profile_bins": int(nb.sum()),
           "gather_ms": g, "construct_ms": u, "profiles_ms": pm, "gpu_ms": g + u + pm,
           "wall_ms": float(np.median(wall)), "keys_ms": keys_ms, "halos_per_s": s.nhalos / ((g + u + pm) * 1e-3)}
    if cpu and os.path.exists(REF_BIN):
        with tempfile.TemporaryDirectory() as d:
            r = subprocess.run([REF_BIN, "profiles-sfc", snap, os.path.join(d, "h.prof"), "--threads", str(threads)],
                               capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            rb = json.loads(r.stdout.strip().splitlines()[-1])
            out["cpu_reference_construct_ms"] = 1e3 * rb["construct_seconds"]
            out["cpu_reference_profiles_ms"] = 1e3 * rb["profiles_seconds"]
            out["cpu_reference_ms"] = out["cpu_reference_construct_ms"] + out["cpu_reference_profiles_ms"]
            out["cpu_threads"] = threads
            out["vs_cpu"] = out["cpu_reference_ms"] / (g + u + pm)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1, help="0, 1 (default), 3 or 4: box workloads")
    ap.add_argument("--side", type=int, default=0, help="particles per dimension (0: full size)")
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--plain-steps", type=int, default=2, help="steps of the one-halo-per-launch form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference CPU timing")
    ap.add_argument("--no-unbind-workload", action="store_true",
                    help="skip the isolated-halo unbinding workload (100k haloes)")
    ap.add_argument("--no-construct-workload", action="store_true",
                    help="skip the halo construction workload (gather + cuts + unbinding)")
    ap.add_argument("--band", type=float, default=0.0,
                    help="extra relative band sent to the exact replay (0: the certified bound only)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (asked for {args.gpus}); "
              f"the unbinding would shard by halo and the deposit by slab of the domain grid, "
              f"but multi-GPU runs are not implemented.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import ahf_amd
    from ahf_amd import dumps

    # the 512^3 and 896^3 boxes are made in memory (a file of the 896^3 one
    # would hold 30 GB); the others go through a cached snapshot file, which
    # the reference's timing reads as well
    in_memory = args.config in (3, 4)
    snap = None if in_memory else make_snapshot(args.config, args.side, args.seed)
    tg = time.perf_counter()
    s = dumps.generate(args.config, args.side, args.seed) if in_memory else dumps.read_snapshot(snap)
    log(f"snapshot: {s.npart} particles, {s.nhalos} haloes, {s.n_ipart} members, L_dom {s.L_dom}, "
        f"L_max {s.L_max} ({time.perf_counter() - tg:.1f}s)")

    ctx = ahf_amd.Context(0)
    # the autotuner's schedule for this workload, where it tuned one apart
    sched_path = os.environ.get("AHF_SCHEDULE") or os.path.join(ROOT, "ahf_amd", "schedule.json")
    if os.path.exists(sched_path):
        with open(sched_path) as f:
            per = json.load(f).get("per_workload", {}).get(f"cfg{args.config}")
        if per:
            ctx.set_schedule(**per)
            log(f"schedule for cfg{args.config}: {per}")
    copy = copy_probe()
    if copy:
        log(f"device copy ceiling: {copy['GBs']:.0f} GB/s (read + write)")
    ctx.set_particles(s.pos, s.mom)
    ctx.set_halos(s.centre, s.halo_npart, s.ipart)
    uargs = (s.minpart, s.vtune, s.r_fac, s.v_fac, s.phi_fac, s.hubble)

    def step(mode=0):
        info = ctx.amr_build(s.L_dom, s.L_max, s.Nth_dom, s.Nth_ref)
        uinfo = ctx.unbind(*uargs, mode=mode, band=args.band)
        return info, uinfo

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")

    amr_ms, unb_ms, stages = [], [], {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        info, uinfo = step()
        t = ctx.last_timing()
        amr_ms.append(t["amr_ms"])
        unb_ms.append(t["unbind_ms"])
        for k, (ms, n, b) in ctx.timers().items():
            acc = stages.setdefault(k, [0.0, 0, 0.0])
            acc[0] += ms
            acc[1] += n
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = s.npart * args.steps / wall
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step; amr {np.mean(amr_ms):.3f} ms, "
        f"unbind {np.mean(unb_ms):.3f} ms (device events)")

    # plain form: one halo per launch, synchronised per halo
    plain = None
    plain_unb_ms = []
    if args.plain_steps > 0:
        tp0 = time.perf_counter()
        for i in range(args.plain_steps):
            step(mode=1)
            plain_unb_ms.append(ctx.last_timing()["unbind_ms"])
        tp1 = time.perf_counter()
        plain = s.npart * args.plain_steps / (tp1 - tp0)
        log(f"plain form: {1e3 * (tp1 - tp0) / args.plain_steps:.3f} ms/step, unbinding "
            f"{np.mean(plain_unb_ms):.3f} ms (batched {np.mean(unb_ms):.3f} ms)")

    # ahf_gridinfo on the hierarchy just built
    gi_fig = gridinfo_figures(ctx, s)

    # the drop-in boundary hands over host arrays: time their upload (AoS
    # pageable host memory -> SoA device arrays) separately; not part of value
    tu0 = time.perf_counter()
    ctx.set_particles(s.pos, s.mom)
    upload_ms = 1e3 * (time.perf_counter() - tu0)
    log(f"particle upload (host -> HBM, pos + mom): {upload_ms:.2f} ms")

    deposits = int(info.deposits)
    log("levels (nodes, particles): " + ", ".join(f"({int(info.n_nodes[i])}, {int(info.n_part[i])})"
                                                   for i in range(info.n_levels)))
    stage_ms = {k: v[0] / args.steps for k, v in stages.items()}
    stage_n = {k: v[1] / args.steps for k, v in stages.items()}
    log("stages (ms/step): " + ", ".join(f"{k} {v:.3f} ({stage_n[k]:.0f})" for k, v in stage_ms.items()))
    # roofline of the dominant stage.  Algorithmic bytes of the deposit per
    # level: 12 B (FP32 position) per deposited particle + 13 B per node
    # (8 B density accumulator, 4 B count, 1 B flag), summed over the levels
    # of one build; the gather launches one light kernel per level.
    dom = max(stage_ms, key=stage_ms.get)
    levels = [(int(info.n_part[i]), int(info.n_nodes[i])) for i in range(info.n_levels)]
    dep_bytes = sum(12 * p + 13 * n for p, n in levels)
    gather_avg_ms = stage_ms["gather_light"] / max(1.0, stage_n["gather_light"])
    achieved = (dep_bytes / max(1.0, stage_n["gather_light"])) / (gather_avg_ms * 1e-3) / 1e9
    # the whole deposit per level: the position/weight records and the
    # Morton order (prep), the lane-per-node walks (gather_light) and the
    # wavefront-per-node folds of the heaviest nodes (gather_heavy)
    dep_ms = stage_ms["prep"] + stage_ms["gather_light"] + stage_ms["gather_heavy"]
    achieved_incl = dep_bytes / (dep_ms * 1e-3) / 1e9
    traffic = load_traffic()
    cfg_tag = f"cfg{args.config}_{args.side}_{args.seed}"
    if traffic and traffic.get("workload") != cfg_tag:
        traffic = None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, ahf_amd/csrc/snapgen.c)",
        "config": {
            "workload": (f"{s.npart} particles ({round(s.npart ** (1 / 3))}^3), {s.boxsize:g} Mpc/h box, "
                         f"{s.nhalos} NFW haloes incl. subhaloes, LgridDomain {s.L_dom}, LgridMax {s.L_max}, "
                         f"NperDomCell {s.Nth_dom:g}, NperRefCell {s.Nth_ref:g}, VescTune {s.vtune:g}, "
                         f"NminPerHalo {s.minpart}"),
            "levels": int(info.n_levels),
            "deposits_per_step": deposits,
            "halo_members": int(s.n_ipart),
            "seed": args.seed,
        },
        "amr_ms": float(np.mean(amr_ms)),
        "unbind_ms": float(np.mean(unb_ms)),
        "deposits_per_s": deposits / (np.mean(amr_ms) * 1e-3),
        "halos_per_s": s.nhalos / (np.mean(unb_ms) * 1e-3),
        "particle_iterations_per_s": uinfo.particle_iterations / (np.mean(unb_ms) * 1e-3),
        "exact_replays": int(uinfo.n_exact_replays),
        "plain_value": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "unbind_plain_ms": float(np.mean(plain_unb_ms)) if plain_unb_ms else None,
        "unbind_batched_over_plain": (float(np.mean(plain_unb_ms)) / float(np.mean(unb_ms))) if plain_unb_ms else None,
        "schedule": ctx.schedule(),
        "stage_ms": stage_ms,
        "dominant_stage": dom,
        "roofline": {
            "kernel": ("exact-order TSC deposit, one timed launch per grid level: k_gather_march + "
                       "k_gather_slab (face planes) on levels holding every octet, k_gather_light on the others"),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "avg_launch_ms": gather_avg_ms,
            "launches_per_step": stage_n["gather_light"],
            "algorithmic_bytes_per_launch": dep_bytes / max(1.0, stage_n["gather_light"]),
            "achieved_deposit_inclusive": achieved_incl,
            "frac_deposit_inclusive": achieved_incl / HBM_PEAK_GBS,
            "deposit_inclusive_ms_per_step": dep_ms,
        },
        "copy_ceiling_GBs": copy["GBs"] if copy else None,
        "upload_ms": upload_ms,
        "pcie_inclusive_value": s.npart / ((ms_per_step + upload_ms) * 1e-3),
        "cpu_baseline": None,
    }
    cpu_snap = snap
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        threads = min(threads, 16)
        # the 512^3 and 896^3 boxes: the same generator at 1/64 of the
        # particles (side / 4), a bounded sample of the same workload
        cpu_snap, n_cpu, what = snap, s.npart, "whole workload once"
        if in_memory:
            side = (args.side or {3: 512, 4: 896}[args.config]) // 4
            cpu_snap = make_snapshot(args.config, side, args.seed)
            n_cpu = side ** 3
            what = f"the same generator at {side}^3 ({n_cpu} particles, 1/64 of the workload)"
        rb = cpu_baseline(cpu_snap, threads)
        if rb:
            cpu_s = rb["amr_seconds"] + rb["unbind_seconds"]
            out["cpu_baseline"] = {
                "value": n_cpu / cpu_s,
                "unit": "particles/s",
                "cores": threads,
                "kind": "reference",
                "sample": (f"{what}: AMR build {rb['amr_seconds']:.2f}s on 1 thread "
                           f"(reference default: deposit is serial), unbinding {rb['unbind_seconds']:.3f}s "
                           f"with OpenMP over haloes on {threads} threads "
                           f"({rb.get('unbind_seconds_1thread', 0.0):.3f}s on 1 thread); host {cpu_model()}"),
            }
            out["vs_cpu"] = value / out["cpu_baseline"]["value"]
    out["gridinfo"] = g = gridinfo_cpu(gi_fig, None if args.no_cpu else cpu_snap)
    ctx.close()
    if not args.no_cpu and not in_memory:
        dr = dropin_figures(snap, min(os.cpu_count() or 1, 16))
        if dr and out["cpu_baseline"]:
            dr["reference_amr_seconds"] = rb["amr_seconds"]
            dr["reference_unbind_seconds"] = rb["unbind_seconds"]
        out["dropin"] = dr
        if dr:
            log(f"drop-in (AHF's objects + the library): AMR {dr['amr_seconds']:.3f} s, per-halo unbinding "
                f"{1e3 * dr['unbind_seconds']:.2f} ms on {dr['threads']} threads"
                + (f"; reference {dr['reference_amr_seconds']:.3f} s, {1e3 * dr['reference_unbind_seconds']:.2f} ms"
                   if "reference_amr_seconds" in dr else ""))
    log(f"gridinfo: {g['device_ms']:.3f} ms on the GPU for {g['nodes']} nodes of {g['no_grids']} grids, "
        f"{g['refinements']} refinements"
        + (f"; reference CPU {g['cpu_nodes_per_s']:.3g} nodes/s ({g['cpu_sample']})" if g.get("cpu_nodes_per_s") else ""))
    if in_memory:
        args.no_construct_workload = True
    if not args.no_construct_workload:
        cw = construction_workload(snap, cpu=not args.no_cpu, threads=min(os.cpu_count() or 1, 16))
        log(f"construction workload: gather {cw['gather_ms']:.2f} ms + construct {cw['construct_ms']:.2f} ms "
            f"+ profiles {cw['profiles_ms']:.2f} ms "
            f"on the GPU ({cw['gathered_members']} gathered, {cw['final_members']} kept)"
            + (f", reference CPU {cw['cpu_reference_ms']:.1f} ms" if "cpu_reference_ms" in cw else ""))
        out["construction_workload"] = cw
    if not args.no_unbind_workload and args.side == 0 and args.config == 1:
        u = unbinding_workload(cpu=not args.no_cpu, threads=min(os.cpu_count() or 1, 16))
        log(f"unbinding workload: {u['gpu_ms']:.2f} ms on the GPU"
            + (f", reference CPU {u['cpu_reference_ms']:.1f} ms" if "cpu_reference_ms" in u else ""))
        log(f"  constructed (cuts + unbinding) {u['construct_ms']:.2f} ms, profiles {u['profiles_ms']:.2f} ms on the GPU"
            + (f"; reference CPU {u['cpu_reference_construct_ms']:.1f} ms + {u['cpu_reference_profiles_ms']:.1f} ms"
               if "cpu_reference_profiles_ms" in u else ""))
        out["unbinding_workload"] = u
    print(json.dumps(out))


if __name__ == "__main__":
    main()
