"""Copies the reference's small input meshes and its sample model input (data files) into
tests/golden/ (the sample model input as ``modelinput_sample``).

Run in a container where the reference tree is mounted:
    python tests/golden/make_fixtures.py /root/reference
The meshes are inputs only (the reference holds no expected outputs for them).
"""
import os
import shutil
import sys

MESHES = ["OnePrism.mesh", "TwoPrisms.mesh", "DoublePrism.mesh", "SinglePrism.mesh", "Wind48hr_02.mesh"]

if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "TissueFolding", "ToolBox", "MeshGeneration", "2DEllipse")
    here = os.path.dirname(os.path.abspath(__file__))
    for m in MESHES:
        shutil.copyfile(os.path.join(src, m), os.path.join(here, m))
        print("copied", m)
    shutil.copyfile(os.path.join(ref, "TissueFolding", "TissueFolding", "modelinput"),
                    os.path.join(here, "modelinput_sample"))
    print("copied modelinput -> modelinput_sample")
