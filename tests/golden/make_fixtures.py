"""Copy the reference's own mesh data files used as fixtures.

* examples/data/sweptwing-{left,right}000.vtu of FLOWPanel.jl are the two
  halves of the swept wing its tests run (test/runtests_meshes.jl:195-257): VTK
  unstructured grids (zlib-compressed appended binary, 768 points and 1,440
  triangles each).
* examples/data/sweptwing-openvsp.msh is the OpenVSP/Gmsh mesh of the same wing
  that test/runtests_meshes2.jl:47 solves with the watertight least-squares
  formulation (Gmsh 4.1 ASCII).
* examples/data/zeroebwb.msh and zeroebwb-TE.msh are the blended-wing-body mesh
  and its trailing-edge line of examples/blendedwingbody.jl.
* examples/data/naca662015.csv is the duct contour of test/runtests_solvers2.jl:98
  (x, y columns of the NACA 66-2015 section).

They are data, copied byte for byte; run this script in a checkout that has the
reference at /root/reference.
"""
import hashlib
import os
import shutil

SRC = "/root/reference/examples/data"
HERE = os.path.dirname(os.path.abspath(__file__))

for name in ("sweptwing-left000.vtu", "sweptwing-right000.vtu", "sweptwing-openvsp.msh", "zeroebwb.msh",
             "zeroebwb-TE.msh", "naca662015.csv"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(HERE, name))
    print(name, hashlib.sha256(open(os.path.join(HERE, name), "rb").read()).hexdigest())
