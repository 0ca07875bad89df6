"""Copy the data files of the reference's own tests that the host tests use
as fixtures (TELEMAC result and boundary files, the triangulation exported by
the reference's reader, a particle result file, the mesh tables the reference
saved to HDF5 from the TELEMAC flume) into tests/golden/reference_data.

    python tests/golden/copy_reference_data.py /root/reference/tests/data
"""
import os
import shutil
import sys

FILES = ["r2d_flume.slf", "geo_flume.cli", "test_triangulation.slf", "test_triangulation_meshx.txt",
         "test_triangulation_meshy.txt", "test_triangulation_meshtri.txt", "particles_2d.txt", "geo_flume.h5"]

if __name__ == "__main__":
    src = sys.argv[1]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data")
    os.makedirs(dst, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        print(f)
