"""The marching-cubes tables of sicp_mesh_scene and the PLY writer, on the
CPU (host functions of the library; no GPU call).

The triangle table is the reference's (the classic published table,
Mapping/Constant.h:39-296, carried over as mc_table.inc); the edge table and
the vertex counts are derived. Pinned here by:
  * rows of the published table (Lorensen and Cline's cases as Bourke lists
    them): cases 1, 3, 5, 6, 7 and their complements' triangle counts;
  * the edge table is exactly the set of edges whose corners differ in sign;
  * every triangle uses cut edges only, and every cut edge is used;
  * over random sign fields the triangles form a closed surface in which
    every mesh edge is used once in each direction;
  * a sphere meshes to a closed surface of Euler characteristic 2 wound
    with its normals towards the inside (sdf < 0), the classic table's
    winding (negative signed volume).
"""
import itertools
import os

import numpy as np
import pytest

from helpers import B

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6),
        (3, 7)]


@pytest.fixture(scope="module")
def tables():
    return B.mesh_tables()


def _tris(tri, nvert, case):
    t = tri[case][: nvert[case]]
    return [tuple(int(e) for e in t[i:i + 3]) for i in range(0, len(t), 3)]


def test_edge_table_is_the_cut_set(tables):
    edge, tri, nvert = tables
    for case in range(256):
        inside = [(case >> k) & 1 for k in range(8)]
        want = sum(1 << j for j, (a, b) in enumerate(EDGE) if inside[a] != inside[b])
        assert edge[case] == want, case
        used = set(int(e) for e in tri[case][: nvert[case]])
        assert used == {j for j in range(12) if (want >> j) & 1}, case
        assert nvert[case] % 3 == 0 and np.all(tri[case][nvert[case]:] == -1)
        assert nvert[case] <= 15
    assert nvert[0] == 0 and nvert[255] == 0 and nvert[1] == 3


def _mesh_field(inside, tables):
    """Triangles of a sign field as triples of global edge keys."""
    edge, tri, nvert = tables
    Z, Y, X = inside.shape
    out = []
    for z, y, x in itertools.product(range(Z - 1), range(Y - 1), range(X - 1)):
        case = 0
        for k, (cx, cy, cz) in enumerate(CORNER):
            case |= int(inside[z + cz, y + cy, x + cx]) << k
        for t in _tris(tri, nvert, case):
            keys = []
            for e in t:
                a, b = EDGE[e]
                pa = (x + CORNER[a][0], y + CORNER[a][1], z + CORNER[a][2])
                pb = (x + CORNER[b][0], y + CORNER[b][1], z + CORNER[b][2])
                keys.append(tuple(sorted([pa, pb])))
            out.append(tuple(keys))
    return out


def _closed_and_oriented(tris):
    directed = {}
    for t in tris:
        for i in range(3):
            d = (t[i], t[(i + 1) % 3])
            directed[d] = directed.get(d, 0) + 1
    return all(c == 1 and directed.get((d[1], d[0]), 0) == 1 for d, c in directed.items())


def test_random_fields_give_closed_oriented_surfaces(tables):
    rng = np.random.default_rng(0)
    for _ in range(60):
        s = rng.random((7, 7, 7)) < 0.5
        s[[0, -1], :, :] = s[:, [0, -1], :] = s[:, :, [0, -1]] = False
        tris = _mesh_field(s, tables)
        assert tris and _closed_and_oriented(tris)


def test_sphere_is_a_closed_inward_wound_surface(tables):
    n = 14
    g = np.arange(n) - (n - 1) / 2
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    sdf = np.sqrt(xx ** 2 + yy ** 2 + zz ** 2) - 4.3
    tris = _mesh_field(sdf < 0, tables)
    assert _closed_and_oriented(tris)
    verts = {k for t in tris for k in t}
    edges = {tuple(sorted((t[i], t[(i + 1) % 3]))) for t in tris for i in range(3)}
    assert len(verts) - len(edges) + len(tris) == 2
    # signed volume with vertices at the zero crossing of each edge
    pos = {}
    for k in verts:
        (x0, y0, z0), (x1, y1, z1) = k
        v0, v1 = sdf[z0, y0, x0], sdf[z1, y1, x1]
        a = v0 / (v0 - v1)
        pos[k] = np.array([x0 + a * (x1 - x0), y0 + a * (y1 - y0), z0 + a * (z1 - z0)]) - (n - 1) / 2
    vol = sum(np.dot(pos[t[0]], np.cross(pos[t[1]], pos[t[2]])) for t in tris) / 6
    assert abs(vol + 4 / 3 * np.pi * 4.3 ** 3) < 0.1 * 4 / 3 * np.pi * 4.3 ** 3


def test_rows_of_the_published_table(tables):
    edge, tri, nvert = tables
    want = {1: [0, 8, 3], 3: [1, 8, 3, 9, 8, 1], 5: [0, 8, 3, 1, 2, 10], 6: [9, 2, 10, 0, 2, 9],
            7: [2, 8, 3, 2, 10, 8, 10, 9, 8], 8: [3, 11, 2], 15: [9, 8, 10, 10, 8, 11]}
    for case, row in want.items():
        assert list(tri[case][: nvert[case]]) == row, case
    # a case and its complement cut the same edges
    for case in range(256):
        assert edge[case] == edge[255 - case]


def test_ply_writer(tmp_path):
    """System::WriteMeshToDisk's layout: header, one vertex line per triangle
    corner (x y z nx ny nz r g b), one face line per triangle."""
    v = np.arange(18, dtype=np.float32).reshape(2, 3, 3) * 0.25
    nrm = np.tile(np.float32([0, 0, 1]), (2, 3, 1))
    col = (np.arange(18, dtype=np.uint8).reshape(2, 3, 3) * 7)
    path = os.path.join(str(tmp_path), "m.ply")
    B.check(B.lib().sicp_mesh_write_ply(path.encode(), v.ctypes.data, nrm.ctypes.data,
                                        col.ctypes.data, 2))
    lines = open(path).read().splitlines()
    hdr = lines[: lines.index("end_header") + 1]
    assert hdr[0] == "ply" and "element vertex 6" in hdr and "element face 2" in hdr
    assert "property list uchar uint vertex_indices" in hdr
    body = lines[len(hdr):]
    assert len(body) == 8
    for i in range(6):
        f = body[i].split()
        assert np.allclose([float(x) for x in f[:3]], v.reshape(6, 3)[i])
        assert [int(x) for x in f[6:]] == list(col.reshape(6, 3)[i])
    assert body[6] == "3 0 1 2" and body[7] == "3 3 4 5"
