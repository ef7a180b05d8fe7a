"""Helpers to unpack the golden fixtures of tests/golden/ into a work dir."""
import json
import lzma
import os
import shutil

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

GENERATED_CASES = ["c1_lc1", "c1_lc2", "refmesh40x120_lc2"]
MATRIX_CASES = sorted(d for d in (os.listdir(GOLDEN) if os.path.isdir(GOLDEN) else []) if d.startswith("m_"))
THREE_D_CASES = sorted(d for d in (os.listdir(GOLDEN) if os.path.isdir(GOLDEN) else []) if d.startswith("r3d_"))
# subsonic outlet (25) and isothermal walls (11)
BC_CASES = sorted(d for d in (os.listdir(GOLDEN) if os.path.isdir(GOLDEN) else []) if d.startswith("b_"))
# restart from the reference's own .sav (.inp init mode 1)
RESTART_CASES = sorted(d for d in (os.listdir(GOLDEN) if os.path.isdir(GOLDEN) else []) if d.startswith("restart_"))
# adapted (hanging-node) meshes of the reference: parent zones, face/cell trees
ADAPTED_CASES = sorted(d for d in (os.listdir(GOLDEN) if os.path.isdir(GOLDEN) else []) if d.startswith("a_"))
# triangles: the wedge channel with every quadrilateral split in two
TRIANGLE_CASES = sorted(d for d in (os.listdir(GOLDEN) if os.path.isdir(GOLDEN) else []) if d.startswith("t_"))


def unpack(name, workdir):
    """Copies fixture `name` into workdir; returns (mesh_path, case_prefix, expected, meta)."""
    src = os.path.join(GOLDEN, name)
    os.makedirs(workdir, exist_ok=True)
    mesh = None
    prefix = None
    for fn in os.listdir(src):
        p = os.path.join(src, fn)
        if fn.endswith(".msh.xz"):
            mesh = os.path.join(workdir, fn[:-3])
            with lzma.open(p) as f, open(mesh, "wb") as g:
                g.write(f.read())
        elif fn.endswith(".xz"):
            with lzma.open(p) as f, open(os.path.join(workdir, fn[:-3]), "wb") as g:
                g.write(f.read())
        elif fn.split(".")[-1] in ("inp", "bvp", "btq", "par", "prop", "sav"):
            shutil.copy(p, workdir)
            prefix = os.path.join(workdir, fn.rsplit(".", 1)[0])
    exp = np.load(os.path.join(src, "expected.npz"), allow_pickle=False)
    with open(os.path.join(src, "meta.json")) as f:
        meta = json.load(f)
    if mesh is None and "mesh" in meta:  # shared mesh of several fixtures
        mesh = os.path.join(workdir, meta["mesh"] + ".msh")
        with lzma.open(os.path.join(GOLDEN, "_meshes", meta["mesh"] + ".msh.xz")) as f, open(mesh, "wb") as g:
            g.write(f.read())
    return mesh, prefix, {"prim": exp["prim"], "residual": [str(x) for x in exp["residual"]]}, meta


def residual_lines(res, first=1):
    """Residual rows formatted like the reference's residual file (time column
    dropped), numbered from iteration `first`."""
    out = []
    for i, row in enumerate(res):
        out.append("%d  " % (i + first) + " ".join("%.3e" % v for v in row))
    return out
