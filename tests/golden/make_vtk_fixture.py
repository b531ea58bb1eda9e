"""Writes tests/golden/vtk/solo_cilindro_tetra_1_istep_0.vtk.xz: the one
results file the reference ships (output/poro_mec/elastico/tetra, written by
its pres macro at step 0 in the legacy ASCII encoding), compressed as is.
It is data: the reference's output, no code.  The results-writer test
reads its number formats and section headers.  Run from the repository
root:
    python tests/golden/make_vtk_fixture.py
"""
import lzma
import os

SRC = "/root/reference/output/poro_mec/elastico/tetra/solo_cilindro_tetra_1_istep_0.vtk"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vtk",
                   "solo_cilindro_tetra_1_istep_0.vtk.xz")

if __name__ == "__main__":
    with open(SRC, "rb") as f:
        data = f.read()
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with lzma.open(OUT, "wb", preset=9) as f:
        f.write(data)
    print(OUT, len(data), "bytes ->", os.path.getsize(OUT))
