"""The cell pass's whole-cell frustum answers (vg::cell_inside_bits, csrc/vgcut_kernels.h, with the per-cell
bounds of vgi::cell_bound, csrc/vgcut_cells.h) checked on the CPU with the kernel's and the host's own code
compiled for the host (tests/inside_bound_check.hip): over 400k random and adversarial cells of 1-128
instances (up to 30 km out, scales 0.01-100, uniform and not, swept and plain planes, planes placed within
4 ulp of the answers' thresholds facing the outermost instance, zero radii), no instance of a cell the
answers settle fails its own binary32 frustum or fast-path test, while a large share of cells is settled.
Two planted mistakes in the bounds (R_in one part in 10^3 too small; the bounding spheres' extent taken for
the culling spheres') each let instances through that fail, so the check can fail.
"""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is missing")
def test_whole_cell_answers_never_pass_a_failing_instance(tmp_path):
    exe = str(tmp_path / "inside_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unused-command-line-argument",
                    "-o", exe, os.path.join(ROOT, "tests", "inside_bound_check.hip")], check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=300)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["violations"] == 0, (out, r.stderr)
    for key in ("in_camera", "in_shadow", "cull_camera", "cull_shadow"):
        assert out[key] > 0.1 * out["cases"], out
    assert out["planted_shrunk"] > 0 and out["planted_cull"] > 0, out
