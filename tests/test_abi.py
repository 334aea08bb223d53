"""CPU checks of the C-ABI boundary: the library builds for gfx950, loads,
and exports every entry point include/lotus_hip.h declares; record layouts
match the header."""
import ctypes as C
import os
import re
import subprocess

from conftest import ROOT
from lotus_amd import hip

HEADER = os.path.join(ROOT, "include", "lotus_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lotus_hip_[a-z_]+)\s*\(", text)))


def test_header_lists_every_binding():
    assert _declared() == sorted(hip.SYMBOLS)


def test_library_exports_declared_symbols():
    L = hip.lib()  # loads without a GPU
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (lotus_hip_\w+)", out))
    for name in _declared():
        assert name in exported, name
        assert hasattr(L, name)


def test_library_holds_gfx950_code_object():
    # the embedded offload bundle names its target
    data = open(hip.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_probe_without_device_is_skip_or_pass():
    status, detail = hip.probe()
    assert status in (hip.PASS, hip.SKIP), detail


def test_struct_sizes_match_header():
    # compile a tiny C program against the header and compare offsets
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "lotus_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lotus_hip_scene_plan),
         sizeof(lotus_hip_scene_evidence),
         sizeof(lotus_hip_frame_params), sizeof(lotus_hip_frame_result),
         offsetof(lotus_hip_frame_result, traversal_nodes), offsetof(lotus_hip_scene_plan, bvh_leaf_size),
         sizeof(lotus_hip_cache_config), offsetof(lotus_hip_cache_config, resolution),
         offsetof(lotus_hip_frame_result, cache_loss), sizeof(lotus_hip_temporal_settings),
         offsetof(lotus_hip_temporal_settings, reset_reason), sizeof(lotus_hip_tuning),
         offsetof(lotus_hip_tuning, tail_divisor), offsetof(lotus_hip_scene_evidence, update),
         sizeof(lotus_hip_scene_contents), offsetof(lotus_hip_scene_contents, environment),
         offsetof(lotus_hip_scene_contents, bvh_triangles));
  return 0;
}'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        sizes = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert sizes[0] == C.sizeof(hip.ScenePlan)
    assert sizes[1] == C.sizeof(hip.SceneEvidence)
    assert sizes[2] == C.sizeof(hip.FrameParams)
    assert sizes[3] == C.sizeof(hip.FrameResult)
    assert sizes[4] == hip.FrameResult.traversal_nodes.offset
    assert sizes[5] == hip.ScenePlan.bvh_leaf_size.offset
    assert sizes[6] == C.sizeof(hip.CacheConfig)
    assert sizes[7] == hip.CacheConfig.resolution.offset
    assert sizes[8] == hip.FrameResult.cache_loss.offset
    assert sizes[9] == C.sizeof(hip.TemporalSettings)
    assert sizes[10] == hip.TemporalSettings.reset_reason.offset
    assert sizes[11] == C.sizeof(hip.Tuning)
    assert sizes[12] == hip.Tuning.tail_divisor.offset
    assert sizes[13] == hip.SceneEvidence.update.offset
    assert sizes[14] == C.sizeof(hip.SceneContents)
    assert sizes[15] == hip.SceneContents.environment.offset
    assert sizes[16] == hip.SceneContents.bvh_triangles.offset
    assert hip.TEMPORAL_PIXEL.itemsize == 64 and hip.TEMPORAL_RADIANCE.itemsize == 32
