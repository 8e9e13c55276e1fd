"""Writes tests/golden/scenes/cornell: the procedural Cornell box (scenes.cornell,
k=2, 64x64) as reference-format scene files (config.xml, scene.xml,
materials.xml, one .obj per object), read back by tests/test_sceneio.py.

Usage: python tests/golden/make_scene_files.py
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from walkray_amd import scenes as S  # noqa: E402

if __name__ == "__main__":
    objs, cam = S.cornell(k=2, width=64, height=64)
    out = os.path.join(ROOT, "tests", "golden", "scenes", "cornell")
    print(S.write_reference_scene(out, objs, cam, spp=16, name="Cornell"))
