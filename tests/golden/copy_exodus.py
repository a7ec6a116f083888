"""Copies two of the reference's sample wall meshes (Exodus II Tri3, netCDF 64-bit
offset) into tests/golden/exodus/ as fixtures of the mesh reader's tests: the straight
tube new_cyl_D2_L13_33.e (diameter 2, length 13.33) and the constricted tube
squeeze_D8_D6_L10.e.  They are data files of the reference
(sample_files/sample_walls/wall_meshes/), copied unchanged."""
import os
import shutil

SRC = "/root/reference/sample_files/sample_walls/wall_meshes"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "exodus")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for name in ("new_cyl_D2_L13_33.e", "squeeze_D8_D6_L10.e"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
        print("copied", name)
