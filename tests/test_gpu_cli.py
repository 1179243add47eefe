"""pfxh-render, the C++ caller of the C ABI: a scene written by
save_scene and rendered by the command-line tool (fixed spp, a-trous and
learned denoise) gives the same colour, albedo, normal and denoised bytes as
the Python trace of the same scene and seed."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pfx_amd import Bounces, Detail, Trace, pfm, scenes
from pfx_amd import learn as Lm
from pfx_amd.scene import save_scene

pytestmark = pytest.mark.gpu

CLI = Path(__file__).resolve().parent.parent / "pfx_amd" / "pfxh-render"


@pytest.mark.parametrize("denoise", ["atrous", "learned"])
def test_cli_renders_like_the_trace(tmp_path, denoise):
    print(f"cli {denoise}", flush=True)
    assert CLI.exists(), "build pfxh-render first (python -m pfx_amd.build)"
    scene = scenes.box_spheres(seed=2, spheres=8, width=96, height=64)
    path = tmp_path / "s.pfxscn"
    save_scene(scene, path)
    prefix = tmp_path / "still"
    proc = subprocess.run([str(CLI), str(path), "--width", "96", "--height", "64", "--spp", "4", "--bounces", "4",
                           "--seed", "9", "--denoise", denoise, "--out", str(prefix)],
                          capture_output=True, text=True, timeout=120)
    print(proc.stdout, proc.stderr, flush=True)
    assert proc.returncode == 0
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 96, 64) as t:
        t.sample(4, 9)
        ref = t.readback()
        if denoise == "atrous":
            den = t.atrous()
        else:
            with Lm.Learner(weights=Lm.WEIGHTS) as L:
                den = t.learned_denoise(L)
    for name, img in (("color", ref.color), ("albedo", ref.albedo), ("normal", ref.normal), ("denoised", den)):
        got = pfm.read((tmp_path / f"still.{name}.pfm").read_bytes())
        assert got[..., :3].tobytes() == np.ascontiguousarray(img[..., :3]).tobytes(), name
