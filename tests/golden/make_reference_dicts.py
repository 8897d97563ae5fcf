"""Copies the input dictionaries of the reference's example cases into
tests/golden/reference_examples/ as fixtures for the case reader.

These are data files the reference ships with its examples (the inputs its
CInputReader / dictReaderOF / runTimeManagerOF read: system/controlDict,
system/fvSchemes, constant/thermophysicalProperties, constant/spongeDict,
constant/turbulenceProperties); the examples hold no result files, and their
polyMesh lists are not in the repository (large blobs), so the tests pair
these dictionaries with meshes of their own.  Run in a container that has
the reference at /root/reference:

    python tests/golden/make_reference_dicts.py
"""
import os
import shutil

SRC = "/root/reference/examples"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_examples")
CASES = {
    "SBS_antiphase": "SBS_antiphase",
    "SBS_inphase": "SBS_inphase",
    "cylinder_vortex": "cylinder_vortex",
    "cylinder_vortex_unstructured": "cylinder_vortex_unstructured",
    "tandem_vortex": "tandem_vortex",
    "3D_Cylinder_Re3900_S": "3D_Cylinder_Re3900/S",
    "3D_Cylinder_Re3900_L": "3D_Cylinder_Re3900/L",
}
FILES = ["system/controlDict", "system/fvSchemes", "constant/thermophysicalProperties", "constant/spongeDict",
         "constant/turbulenceProperties"]


def main():
    for name, rel in CASES.items():
        for f in FILES:
            src = os.path.join(SRC, rel, f)
            dst = os.path.join(DST, name, f)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            shutil.copyfile(src, dst)
    print(f"wrote {len(CASES)} cases x {len(FILES)} dictionaries under {DST}")


if __name__ == "__main__":
    main()
