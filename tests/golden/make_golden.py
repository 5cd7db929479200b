"""Write the golden fixtures of tests/golden/ from the reference's own compiled
functions (oracle/_ref/ahf_ref, built from the reference sources by
`make -C oracle ref`).

Each fixture is data only: the input snapshot's arrays and parameters, and
the reference's output dump (AHFGRID1, AHFHALO1 or AHFPROF1 bytes) for one command:
  amr            gen_domgrids, ll, zero_dens + assign_npart, gen_AMRhierarchy
  sort           sort_halo_particles per halo
  unbind         sort_halo_particles + rem_unbound per halo
  construct      sort, rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1)
  gather         ahf_halos_sfc_gatherParts per halo
  construct-sfc  gather followed by the four construction steps
  profiles       the four construction steps, then HaloProfiles (AHFPROF1)
  profiles-sfc   gather, the four construction steps, then HaloProfiles
tests/test_golden.py replays the inputs through the oracle (and, on a GPU,
through the library) and compares with the stored outputs.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from ahf_amd import dumps  # noqa: E402
from helpers import fuzz_snapshot, tie_snapshot  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ahf_ref")
HEADER = ("boxsize", "omega0", "lambda0", "a", "pmass", "L_dom", "L_max", "Nth_dom", "Nth_ref", "minpart", "vtune",
          "r_fac", "v_fac", "phi_fac", "hubble", "seed", "config", "ovlim", "rho_vir")

# (name, how the input is made, reference commands whose outputs are stored)
CASES = [
    ("amr_fuzz142", ("fuzz", 142, True), ["amr"]),
    ("amr_fuzz305", ("fuzz", 305, False), ["amr"]),
    ("halo_cfg2_h40", ("gen", 2, 40, 12), ["unbind", "construct", "profiles"]),
    ("sfc_cfg1_side24", ("gen+centres", 1, 24, 11), ["gather", "construct-sfc"]),
    ("halo_ties_s5", ("ties", 5), ["sort", "unbind", "construct"]),
]


def make_input(spec, path):
    if spec[0] == "fuzz":
        fuzz_snapshot(path, spec[1], non_pow2=spec[2])
        return
    if spec[0] == "ties":
        tie_snapshot(path, spec[1])
        return
    lib = ctypes.CDLL(os.path.join(ROOT, "ahf_amd", "lib", "libahf_snapgen.so"))
    lib.ahf_snapgen_write.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p]
    if lib.ahf_snapgen_write(spec[1], spec[2], spec[3], path.encode()):
        raise RuntimeError(f"snapshot generation failed for {spec}")
    if spec[0] == "gen+centres":
        # the generated haloes plus 30 gather centres: on particles, on the box
        # corner and faces, with radii from the finest search grid to the
        # whole box
        s = dumps.read_snapshot(path)
        rng = np.random.default_rng(spec[3])
        extra = s.pos[rng.integers(0, s.npart, 24)].astype(np.float64)
        extra = np.concatenate([extra, [[0.0, 0.0, 0.0], [0.9999, 0.5, 0.0001], [0.5, 0.0, 0.5],
                                        [0.25, 0.25, 0.999], [0.75, 0.75, 0.75], [0.1, 0.9, 0.4]]])
        rad = np.concatenate([10 ** rng.uniform(-3.5, -1.0, 24), [0.05, 0.2, 0.3, 1e-6, 0.0, 0.12]])
        cen = np.concatenate([s.centre, extra])
        grad = np.concatenate([s.gather_rad, rad])
        off = s.halo_offsets
        lists = [s.ipart[off[h]:off[h + 1]] for h in range(s.nhalos)] + [[0]] * len(extra)
        dumps.write_snapshot(path, s.pos, s.mom, s.ids, boxsize=s.boxsize, pmass=s.pmass, L_dom=s.L_dom,
                             L_max=s.L_max, Nth_dom=s.Nth_dom, Nth_ref=s.Nth_ref, minpart=s.minpart, vtune=s.vtune,
                             r_fac=s.r_fac, v_fac=s.v_fac, phi_fac=s.phi_fac, hubble=s.hubble, centres=cen,
                             halo_lists=lists, seed=s.seed, config=s.config, omega0=s.omega0, lambda0=s.lambda0,
                             a=s.a, ovlim=s.ovlim, rho_vir=s.rho_vir, gather_rad=grad)


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/ahf_ref is not built (make -C oracle ref, with the reference tree present)")
    with tempfile.TemporaryDirectory() as d:
        for name, spec, cmds in CASES:
            snap = os.path.join(d, "in.snap")
            make_input(spec, snap)
            s = dumps.read_snapshot(snap)
            arrays = {k: np.array(getattr(s, k)) for k in HEADER}
            arrays.update(pos=s.pos, centre=s.centre, gather_rad=s.gather_rad, halo_npart=s.halo_npart,
                          ipart=s.ipart, commands=np.array(cmds))
            if cmds != ["amr"]:
                arrays["mom"] = s.mom
            for cmd in cmds:
                out = os.path.join(d, "out.dump")
                args = [REF, cmd, snap, out] + ([] if cmd == "amr" else ["--threads", "1"])
                subprocess.run(args, check=True, capture_output=True)
                arrays["ref_" + cmd] = np.fromfile(out, np.uint8)
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **arrays)
            print(f"{name}: {cmds}, {s.npart} particles, {s.nhalos} haloes, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
