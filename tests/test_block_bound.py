"""The meshlet pass's block test (vg::block_out_of_band, csrc/vgcut_kernels.h) checked on the CPU with the
kernel's own code compiled for the host (tests/block_bound_check.hip): over 2M random and adversarial
blocks (instances up to 10 km from the origin at scales 0.01-100, uniform and non-uniform; cameras 5 cm
to 20 km away; bands placed within 4 ulp of a run's exact own or parent value), no dropped block holds a
run that the exact binary32 LOD test puts in the band, while about 30% of blocks are dropped. The same
test with a planted mistake (ignoring how far the runs' centres spread) drops blocks that hold such runs,
so the check can fail.
"""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is missing")
def test_block_test_never_drops_a_run_in_the_band(tmp_path):
    exe = str(tmp_path / "block_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unused-command-line-argument",
                    "-o", exe, os.path.join(ROOT, "tests", "block_bound_check.hip")], check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["violations"] == 0, (out, r.stderr)
    assert out["dropped"] > 0.2 * out["cases"], out
    assert out["blocks_in_band"] > 0.3 * out["cases"], out
    assert out["planted_violations"] > 0, out
