"""Copies the data fixtures the tests use out of the SpecFWAT tree (run in the
container that holds it; the GPU box only sees the copies).

  proc000000_vp.bin  tests/toy_example_rf/model_target/proc000000_vp.bin: a
                     (5,5,5,NSPEC=1152) single-precision array in one Fortran
                     unformatted sequential record, the layout of every
                     proc%06d_<name>.bin model and kernel file
"""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/tests/toy_example_rf/model_target"

if __name__ == "__main__":
    shutil.copyfile(os.path.join(REF, "proc000000_vp.bin"), os.path.join(HERE, "proc000000_vp.bin"))
    print("copied proc000000_vp.bin")
