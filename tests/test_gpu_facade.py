"""The C++ facade on the device: oracle/_ref/facade_check (tests/
facade_check.cpp) sets up the reference's material fixtures with the
reference's own fixture code, extracts them with its SceneExtraction and
drives Lotus::OffscreenRenderer (UpdateScene, RenderScene, two
RenderTemporalScene frames, ReadBackScene) through the facade. Its frame must
equal, bit for bit, the frame the C ABI renders from the same fixture as
tests/golden/material-*.npz holds it (flattened by an independent path,
tests/material_fixtures.py), with the same ray-queue counts."""
import json
import os
import subprocess

import numpy as np
import pytest

import material_fixtures as mf
from lotus_amd import hip

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "facade_check")
OUT = os.path.join(ROOT, "gpurun_out")


@pytest.fixture(scope="module")
def renderer():
    status, detail = hip.probe()
    if status != hip.PASS:
        pytest.skip(detail)
    r = hip.Renderer(0)
    yield r
    r.close()


@pytest.mark.parametrize("fixture", ["material-basics-v1", "material-primvars-v1", "material-textures-v1",
                                     "material-normals-v1"])
def test_facade_frames_equal_the_c_abi(renderer, fixture):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref/facade_check is not built (the reference sources were absent)")
    os.makedirs(OUT, exist_ok=True)
    prefix = os.path.join(OUT, f"facade-{fixture}")
    print(f"facade_check {fixture}", flush=True)
    r = subprocess.run([CHECK, fixture, prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["scene"] == "pass" and line["render"] == "pass", line
    assert line["empty_update"] == "pass" and line["empty_update_builds"] == line["blas_builds"]
    assert line["temporal"] == ["pass", "pass"], line
    assert line["reset_reasons"][0] == 1 and line["reset_reasons"][1] == 0  # FirstFrame, then None
    assert line["history"] == [0, 1]
    assert line["readback"] == "pass"
    size = line["size"]
    color = np.fromfile(prefix + ".color", np.float32).reshape(size, size, 4)
    clip = np.fromfile(prefix + ".clip", np.float32)
    sc, fixture_size = mf.load(fixture)
    assert fixture_size == size
    assert line["instances"] == len(sc.instances) == line["readback_instances"]
    assert line["materials"] == len(sc.materials) == line["readback_materials"]
    renderer.apply(sc)
    out = renderer.render(clip, size, size, frame_count=2, max_bounces=8, light_sampling=hip.LIGHT_MIS)
    assert list(out["counts"]) == line["counts"]
    assert np.array_equal(out["color"].view(np.uint32), color.view(np.uint32))
