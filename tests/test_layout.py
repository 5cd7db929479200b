"""include/ahf_types.h reproduces the memory layout of AHF's structs for the
default flag set: every size and offset equals the reference's, as printed
by the reference harness compiled against the reference headers."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_layout_matches_reference(ref_bin, tmp_path):
    exe = str(tmp_path / "probe")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(ROOT, "tests", "probe", "layout_probe.c")], check=True)
    ours = json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    ref = json.loads(subprocess.run([ref_bin, "layout"], capture_output=True, text=True, check=True).stdout)
    assert ours == ref
