"""Writes the data fixtures of this directory from the reference's shipped
data (run in a container where the reference is mounted; the fixtures are
committed, the GPU machines never read the reference).

- naca_2D_shape_0.msh.gz      mesh_opti/naca_2D_shape_0.msh (25,274 quads)
- naca_2D_flow_1.gz           mesh_opti/naca_2D_flow_1 (state produced on shape 0)
- naca_2D_flow_0.gz           mesh_opti/naca_2D_flow_0, the state the reference's
                              run started from on shape 0 (config.txt:62-63):
                              flow_1 is flow_0 integrated for final_time =
                              0.005 s on shape 0 with config.txt's parameters
- naca_2D_shape_K.msh.gz, naca_2D_flow_K.gz, naca_2D_flow_{K+1}.gz for K in
                              CHAINED: more (mesh, start, end) triples of the
                              same chain (run_opti.sh:211-254 bumps the three
                              paths together)
- naca_2D_shape_params_K.txt  mesh_opti/naca_2D_shape_params_K.txt for K = 0 and the
                              CHAINED runs (the shape-parameter file run K read)
- trajectory.json             (m, p) of mesh_opti/naca_2D_shape_params_{0..30}.txt:
                              the sign of each step is minus the sign of the
                              reference's shape gradient (run_opti.sh:91-120)
"""
import gzip
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
CHAINED = (15, 29)


def gz(src, dst):
    with open(src, "rb") as f, gzip.GzipFile(dst, "wb", compresslevel=9, mtime=0) as g:
        shutil.copyfileobj(f, g)


def main():
    mo = os.path.join(REF, "mesh_opti")
    gz(os.path.join(mo, "naca_2D_shape_0.msh"), os.path.join(HERE, "naca_2D_shape_0.msh.gz"))
    for k in (0, 1):
        gz(os.path.join(mo, "naca_2D_flow_%d" % k), os.path.join(HERE, "naca_2D_flow_%d.gz" % k))
    for k in CHAINED:
        gz(os.path.join(mo, "naca_2D_shape_%d.msh" % k), os.path.join(HERE, "naca_2D_shape_%d.msh.gz" % k))
        for j in (k, k + 1):
            gz(os.path.join(mo, "naca_2D_flow_%d" % j), os.path.join(HERE, "naca_2D_flow_%d.gz" % j))
    for k in (0,) + CHAINED:
        name = "naca_2D_shape_params_%d.txt" % k
        shutil.copy(os.path.join(mo, name), os.path.join(HERE, name))
    traj = []
    for k in range(31):
        vals = {}
        with open(os.path.join(mo, "naca_2D_shape_params_%d.txt" % k)) as f:
            for line in f:
                mm = re.match(r"\s*(\w+)\s*\*?=\s*([-+0-9.eE]+)", line)
                if mm:
                    vals[mm.group(1)] = float(mm.group(2))
        traj.append({"m": vals["m"], "p": vals["p"], "t": vals.get("t"), "beta": vals.get("beta")})
    with open(os.path.join(HERE, "trajectory.json"), "w") as f:
        json.dump(traj, f, indent=1)


if __name__ == "__main__":
    main()
