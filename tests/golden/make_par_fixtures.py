"""Copy the parameter-file fixtures from the reference's benchmark configuration files and extract
the halo table its A2744 run recorded (Benchmarks/ConfigFiles/hff_slmodels/A2744/Output.txt: the
reference's debug print of every SOA halo after reading A2744_SLspec.par, 6 decimals).

Also copies the bayes-map configuration of utils/maps (its m1931.par with limit blocks and two
potfiles, bayes.dat, the catalogues) into tests/golden/maps/, and the FITS header of a mass map
that Lenstool wrote (Benchmarks/ConfigFiles/hff_slmodels/A2744/ellipticite/mass.fits, the first
2880-byte block) into tests/golden/fits/. The .par of utils/maps names its galaxy catalogue
m1931_CM_zCLASH-MUSE.cat; the directory holds it as m1931_CM_zCLASH-MUSE_lenstool-asec.cat, which
is copied under the name the .par reads.

Run once where the reference is mounted:  python tests/golden/make_par_fixtures.py
Writes tests/golden/par/{m1931.par, m1931.cat, immul_ongoing.cat, A2744_SLspec.par, the three
A2744 potfiles, img_mul_constraints.cat, a2744_recorded_halos.npz}, tests/golden/maps/* and
tests/golden/fits/a2744_mass_header.bin. These are data files: model parameters, catalogues,
MCMC samples and recorded output values."""
import os
import re
import shutil

import numpy as np

REF = "/root/reference/Benchmarks/ConfigFiles"
MAPS = "/root/reference/utils/maps"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "par")
MAPS_FILES = {"m1931.par": "m1931.par", "bayes.dat": "bayes.dat", "immul_ongoing.cat": "immul_ongoing.cat",
              "potfile_cl_N-NW-S3.txt": "potfile_cl_N-NW-S3.txt",
              "m1931_CM_zCLASH-MUSE_lenstool-asec.cat": "m1931_CM_zCLASH-MUSE.cat"}
FILES = ["m1931.par", "m1931.cat", "immul_ongoing.cat"]
A2744 = ["A2744_SLspec.par", "potfile_cl_N-NW-S3.txt", "new_potfile_with_ASTRODEEP_data.txt",
         "potfile_cl_S1-S2-S4.txt", "img_mul_constraints.cat"]


def main():
    os.makedirs(OUT, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(REF, f), os.path.join(OUT, f))
    for f in A2744:
        shutil.copyfile(os.path.join(REF, "hff_slmodels", "A2744", f), os.path.join(OUT, f))
    txt = open(os.path.join(REF, "hff_slmodels", "A2744", "Output.txt")).read()
    pat = (r"Potential\[(\d+)\]: x = (\S+), y = (\S+), b0 = (\S+), vdisp = (\S+), type = (\d+), \s*"
           r"ellipticity = (\S+), ellipticity_pot = (\S+), ellipticity angle \(radians\) = (\S+), "
           r"rcore = (\S+), rcut = (\S+),\s*z = (\S+)")
    rows = np.array([[float(v) for v in m.groups()] for m in re.finditer(pat, txt)])
    names = ["index", "x", "y", "b0", "vdisp", "type", "ellipticity", "epot", "angle", "rcore", "rcut", "z"]
    np.savez(os.path.join(OUT, "a2744_recorded_halos.npz"), **{n: rows[:, k] for k, n in enumerate(names)})
    os.makedirs(os.path.join(HERE, "maps"), exist_ok=True)
    for src, dst in MAPS_FILES.items():
        shutil.copyfile(os.path.join(MAPS, src), os.path.join(HERE, "maps", dst))
    os.makedirs(os.path.join(HERE, "fits"), exist_ok=True)
    with open(os.path.join(REF, "hff_slmodels", "A2744", "ellipticite", "mass.fits"), "rb") as f:
        header = f.read(2880)
    with open(os.path.join(HERE, "fits", "a2744_mass_header.bin"), "wb") as f:
        f.write(header)
    print("fixtures written:", len(rows), "recorded halos")


if __name__ == "__main__":
    main()
