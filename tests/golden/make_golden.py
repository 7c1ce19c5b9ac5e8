"""Write the golden vectors under tests/golden/ from the reference itself.

Runs oracle/_ref/ref_harness (the reference iModel SWM modules compiled from
their Fortran sources by oracle/Makefile.ref; build it first with
``make -f oracle/Makefile.ref``) in a scratch directory and keeps its IMXBIN
outputs as fixtures:

  ref_mesh_g{2,3}.bin       meshbuild icos/eqs/nopt + initialize_gridprop
  ref_mesh_g2_pol.bin       same, icosahedron with nodes on the poles
  ref_run_g3_tc2.bin        TC2: tendency intermediates, diag, 5 RK4 steps dt=900
  ref_run_g3_spec.bin       seeded spectral state (seed 7): same
  ref_run_g2_spec96.bin     seeded spectral state (seed 3): 96 steps dt=900

The seeded states are made by libimx (imx_ic_spectral) on the reference
mesh and handed to the reference as input (the reference has no such case).
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from imodel_amd import imxbin  # noqa: E402
from imodel_amd.swm import Mesh  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(args, cwd):
    r = subprocess.run([HARNESS, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(" ".join(args) + "\n" + r.stdout[-2000:])


def main():
    if not os.access(HARNESS, os.X_OK):
        sys.exit("build the reference harness first: make -f oracle/Makefile.ref")
    with tempfile.TemporaryDirectory() as td:
        for g, pos, name in ((2, "eqs", "ref_mesh_g2.bin"), (3, "eqs", "ref_mesh_g3.bin"),
                             (2, "pol", "ref_mesh_g2_pol.bin")):
            run(["build", str(g), pos, "out.bin"], td)
            shutil.copy(os.path.join(td, "out.bin"), os.path.join(HERE, name))
        run(["run", "build:3:eqs", "tc2", "5", "900", "1", "out.bin"], td)
        shutil.copy(os.path.join(td, "out.bin"), os.path.join(HERE, "ref_run_g3_tc2.bin"))
        for g, seed, steps, name in ((3, 7, 5, "ref_run_g3_spec.bin"), (2, 3, 96, "ref_run_g2_spec96.bin")):
            mesh = Mesh(imxbin.read(os.path.join(HERE, f"ref_mesh_g{g}.bin")))
            h, u = mesh.spectral(seed)
            imxbin.write(os.path.join(td, "state.bin"), {"h": h, "u": u})
            run(["run", f"build:{g}:eqs", "state.bin", str(steps), "900", "1", "out.bin"], td)
            shutil.copy(os.path.join(td, "out.bin"), os.path.join(HERE, name))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".bin"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
