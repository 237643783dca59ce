"""Polyhedral test meshes: adjacent cells of a generated mesh merged into one cell (two
hexahedra sharing a face become a 10-face cell), relabelled so that each interior face's
owner is its lower-numbered cell with the vertex order turned outward from it."""
from __future__ import annotations

import dataclasses

import numpy as np


def merge_cells(data, pairs):
    """MeshData with each (a, b) of `pairs` (disjoint, adjacent cells) merged into one
    cell: the faces between a and b are dropped, b's other faces join a's list after a's
    own, and cells after each removed b are renumbered down."""
    nc = data.n_cells
    rep = np.arange(nc)
    for a, b in pairs:
        rep[b] = a
    keep_c = rep == np.arange(nc)
    cmap = (np.cumsum(keep_c) - 1)[rep]
    own = cmap[data.face_owner]
    nbr = data.face_neighbor.copy()
    inner = nbr >= 0
    nbr[inner] = cmap[nbr[inner]]
    keep_f = ~(inner & (own == nbr))
    assert (~keep_f).sum() >= len(pairs), "a pair is not adjacent"
    fmap = np.cumsum(keep_f) - 1
    flip = inner & (own > nbr)
    off, vtx = [0], []
    for f in np.nonzero(keep_f)[0]:
        ids = data.face_vtx[data.face_vtx_off[f]:data.face_vtx_off[f + 1]]
        vtx.extend((ids[::-1] if flip[f] else ids).tolist())
        off.append(len(vtx))
    new_own = np.where(flip, nbr, own)[keep_f]
    new_nbr = np.where(flip, own, nbr)[keep_f]
    lists = [[] for _ in range(int(keep_c.sum()))]
    for c in range(nc):
        for f in data.cell_faces[data.cell_face_off[c]:data.cell_face_off[c + 1]]:
            if keep_f[f]:
                lists[cmap[c]].append(int(fmap[f]))
    cf_off = np.cumsum([0] + [len(x) for x in lists])
    return dataclasses.replace(
        data, face_vtx_off=np.asarray(off, data.face_vtx_off.dtype), face_vtx=np.asarray(vtx, data.face_vtx.dtype),
        face_owner=new_own.astype(data.face_owner.dtype), face_neighbor=new_nbr.astype(data.face_neighbor.dtype),
        face_patch=data.face_patch[keep_f], cell_face_off=cf_off.astype(data.cell_face_off.dtype),
        cell_faces=np.concatenate([np.asarray(x, data.cell_faces.dtype) for x in lists]),
        cell_vtx_off=None, cell_vtx=None)


def merged_box(P, nx, ny, nz, every=3, **kw):
    """An nx x ny x nz box with cells (i, j, k), (i + 1, j, k) merged for i = 0, every, ...
    in the rows with j + k even, so that no two merged cells touch across two faces (each
    pair of neighbouring cells keeps one face)."""
    data = P.box(nx, ny, nz, **kw)
    pairs = []
    for k in range(nz):
        for j in range(ny):
            if (j + k) % 2:
                continue
            for i in range(0, nx - 1, every):
                a = i + nx * (j + ny * k)
                pairs.append((a, a + 1))
    return merge_cells(data, pairs)
