"""Reader for PyLith's ASCII mesh format (meshio/MeshIOAscii.cc), e.g. the reference's
test fixture tests/libtests/materials/data/hex.mesh: vertices, cells (num-corners 8 for
hexahedra in PyLith vertex order), material ids and vertex groups."""
import re

import numpy as np

from .meshgen import HEX_Q1, Mesh


def _block(text, name):
    m = re.search(name + r"\s*=\s*\{", text)
    if not m:
        return None
    depth, i = 1, m.end()
    while depth:
        depth += {"{": 1, "}": -1}.get(text[i], 0)
        i += 1
    return text[m.end():i - 1]


def read(path):
    text = open(path).read()
    index_zero = re.search(r"use-index-zero\s*=\s*(\w+)", text).group(1) == "true"
    base = 0 if index_zero else 1
    verts = _block(text, "vertices")
    coords = np.array([row.split()[1:] for row in _block(verts, "coordinates").strip().splitlines()],
                      float)
    cells = _block(text, "cells")
    ncorners = int(re.search(r"num-corners\s*=\s*(\d+)", cells).group(1))
    if ncorners != 8:
        raise ValueError("only hexahedral meshes are read")
    cl = np.array([row.split()[1:] for row in _block(cells, "simplices").strip().splitlines()],
                  np.int64) - base
    mat = np.array([row.split()[1] for row in _block(cells, "material-ids").strip().splitlines()],
                   np.int32)
    groups = {}
    for m in re.finditer(r"group\s*=\s*\{", text):
        body = _block(text[m.start():], "group")
        name = re.search(r"name\s*=\s*(\S+)", body).group(1)
        idx = np.array(_block(body, "indices").split(), np.int64) - base
        groups[name] = idx
    closure = cl.astype(np.int32)
    mesh = Mesh(HEX_Q1, closure, closure.copy(), coords, coords, None, mat)
    return mesh, groups
