"""Copy the data files of the reference regtests that pin this path into
tests/golden/ (trajectories and *.reference outputs only).

Run in a tree where the reference sources are available:
    python tests/golden/make_fixtures.py [/root/reference]
"""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    "rt42": ["basic/rt42/trajectory.xyz", "basic/rt42/check_c.reference", "basic/rt42/check_e.reference",
             "basic/rt42/check_g.reference", "basic/rt42/ff.reference"],
    "rt42c": ["basic/rt42c/trajectory.xyz", "basic/rt42c/colvar.reference"],
    "rt42-cells": ["basic/rt42-cells/trajectory.xyz", "basic/rt42-cells/check_diff.reference",
                   "basic/rt42-cells/ff.reference"],
    "switchtraj": ["trajectories/switchtraj.xyz"],
}
for name in ["rational", "rational-MMeq2NN", "fastrational", "fastrational-MMeq2NN", "exp", "gaussian",
             "fastgaussian", "smap", "cubic", "tanh", "cosinus", "q"]:
    FILES["rt20-switch-" + name] = [f"basic/rt20-switch-{name}/COLVAR.reference",
                                    f"basic/rt20-switch-{name}/deriv.reference"]
FILES["rt11"] = ["basic/rt11/COLVAR.reference", "basic/rt11/HILLS.reference", "basic/rt11/forces.reference"]
FILES["rt67"] = ["basic/rt67/COLVAR.reference", "basic/rt67/traj.gro"]
FILES["rt-backup-grid"] = ["basic/rt-backup-grid/trajectory.xyz", "basic/rt-backup-grid/grid.reference",
                            "basic/rt-backup-grid/bck.last.grid.reference"]

# DISTANCE between a centre of mass (made whole) and an atom, forces of a RESTRAINT on it
FILES["rt18"] = ["basic/rt18/COLVAR.reference", "basic/rt18/forces.reference", "basic/rt18/mc",
                 "trajectories/trajectory.xyz"]

# per-atom coordination numbers (COORDINATIONNUMBER): SPECIESA/SPECIESB and SPECIES + MORE_THAN
FILES["symfunc-rt-single-groupb"] = ["trajectories/trajectory.xyz", "symfunc/rt-single-groupb/colvar.reference"]
FILES["symfunc-rt20"] = ["symfunc/rt20/colv2.reference"]

for d, files in FILES.items():
    os.makedirs(os.path.join(HERE, d), exist_ok=True)
    for f in files:
        shutil.copy(os.path.join(REF, "regtest", f), os.path.join(HERE, d, os.path.basename(f)))
print("copied", sum(len(v) for v in FILES.values()), "files")
