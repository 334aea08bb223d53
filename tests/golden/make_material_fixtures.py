"""Writes tests/golden/material-*.npz: each material fixture of the
reference (adapters/common/material_fixtures.cpp, MaterialFixtures) as the
SceneUpdate the reference's own SceneExtraction makes of it, in the word
layout of oracle/ref_driver.cpp's ref_material_fixture, with its negative
control; and copies the fixture's committed reference images
(validation/reference/<id>-{mean,variance}.pfm) beside them. Needs
oracle/_ref, built from the reference sources by `make -C oracle ref`.

    python tests/golden/make_material_fixtures.py
"""
import ctypes as C
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

FIXTURES = ["material-basics-v1", "material-transmission-v1", "material-primvars-v1", "material-normals-v1",
            "material-textures-v1", "material-emission-v1", "material-alpha-v1", "material-transforms-v1",
            "material-subsets-v1", "material-mtlx-v1", "material-graphs-v1", "material-gltf-v1",
            "material-gltf-extensions-v1"]


def words(R, fixture, control):
    n = R.ref_material_fixture(fixture.encode(), control, None, 0)
    if n == 0:
        raise KeyError(fixture)
    out = np.zeros(n, np.uint32)
    assert R.ref_material_fixture(fixture.encode(), control, out.ctypes.data_as(C.POINTER(C.c_uint32)), n) == n
    return out


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built")
    for fixture in FIXTURES:
        np.savez_compressed(os.path.join(HERE, f"{fixture}.npz"), scene=words(R, fixture, 0),
                            control=words(R, fixture, 1))
        for kind in ("mean", "variance"):
            shutil.copyfile(os.path.join(oracle.REFERENCE_ROOT, "validation", "reference", f"{fixture}-{kind}.pfm"),
                            os.path.join(HERE, f"{fixture}-{kind}.pfm"))
        print(fixture)


if __name__ == "__main__":
    main()
