"""Host-side IO of the drop-in (py_sors.IO, sors_amd/cpp/sors_io.cpp) against
the reference's definitions (cpp/src/io.cpp:3-780, plane3D.h:44-63).

Fixtures: four Gmsh meshes from the reference's cpp/demos/meshes (data
files, copied to tests/golden/meshes): sphere1.msh (4.1 ASCII), arm.msh (4.1
binary), duck.msh and softcon_cylinder.msh (2.2 binary, duck with an
$ElementData section).  MshIO is not available, so the readers are pinned by
what the files fix: the node counts of their headers, nonzero tetrahedra,
closed boundary surfaces whose outward orientation gives the tetrahedra's
volume by the divergence theorem, and exact round trips through files
written in the other formats (2.2 ASCII, 4.1 ASCII with several blocks and
sparse tags).  The surface grouping and face order are checked against a
literal restatement of the reference's O(F^2) loops; the VTU / PVD files are
parsed back.  Byte-for-byte vtu11 output is parity unpinned (no vtu11).
"""
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from sors_amd import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
PKG = os.path.join(os.path.dirname(HERE), "sors_amd")


@pytest.fixture(scope="module")
def ps():
    import importlib
    import sys
    sys.path.insert(0, PKG)
    try:
        return importlib.import_module("py_sors")
    finally:
        sys.path.remove(PKG)


def header_node_count(path):
    data = open(path, "rb").read()
    i = data.index(b"$Nodes\n") + len(b"$Nodes\n")
    if b"\n4.1 1" in data[:40]:
        return struct.unpack("<4Q", data[i:i + 32])[1]
    return int(data[i:data.index(b"\n", i)].split()[1 if b"\n4.1" in data[:40] else 0])


def tet_volumes(V, T):
    P = V[T]
    return np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])) / 6


def enclosed_volume(V, faces):
    F = np.asarray(faces)
    if F.shape[1] == 4:  # quads: two triangles (0,1,2), (0,2,3)
        F = np.concatenate([F[:, [0, 1, 2]], F[:, [0, 2, 3]]])
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    return np.einsum("ij,ij->i", A, np.cross(B, C)).sum() / 6


def edge_counts(faces):
    cnt = {}
    for f in faces:
        for a, b in zip(f, list(f[1:]) + [f[0]]):
            k = (min(a, b), max(a, b))
            cnt[k] = cnt.get(k, 0) + 1
    return cnt


@pytest.mark.parametrize("name", ["sphere1", "arm", "duck", "softcon_cylinder"])
def test_read_reference_meshes(ps, name):
    path = os.path.join(MESHES, name + ".msh")
    V, T = ps.IO.read_tet_MSH(path)
    assert V.shape == (header_node_count(path), 3)
    assert T.min() >= 0 and T.max() < len(V)
    vol = tet_volumes(V, T)
    assert (vol > 0).all() or (vol < 0).all()
    groups, fele = ps.IO.extract_tet_surfaces(V, T)
    faces = [f for g in groups for f in g]
    # closed, manifold surfaces: every edge in exactly two faces, V - E + F = 2
    # per component of genus 0 (the sphere)
    cnt = edge_counts(faces)
    assert set(cnt.values()) == {2}
    # outward orientation: the divergence theorem gives the tetrahedra's volume
    assert enclosed_volume(V, faces) == pytest.approx(np.abs(vol).sum(), rel=1e-12)
    for g, fe in zip(groups, fele):
        assert len(g) == len(fe)
        for f, e in zip(g, fe):
            assert set(f) <= set(T[e])  # a face of the element it names
    if name == "sphere1":
        nv = len({v for f in faces for v in f})
        assert nv - len(cnt) + len(faces) == 2


def write_msh22_ascii(path, V, T, tags, extra_lines=True):
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        if extra_lines:
            f.write('$PhysicalNames\n1\n3 1 "body"\n$EndPhysicalNames\n')
        f.write(f"$Nodes\n{len(V)}\n")
        for t, x in zip(tags, V):
            f.write(f"{t} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}\n")
        f.write("$EndNodes\n$Elements\n")
        f.write(f"{len(T) + 1}\n1 15 2 0 1 {tags[0]}\n")  # a point element first, skipped
        for k, e in enumerate(T):
            f.write(f"{k + 2} 4 2 1 1 " + " ".join(str(tags[i]) for i in e) + "\n")
        f.write("$EndElements\n")


def write_msh41_ascii(path, V, E, tags, etype):
    """Nodes in two entity blocks, elements split over two blocks, plus a
    block of triangles (type 2) that the reader skips."""
    n, h = len(V), len(V) // 2
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        f.write("$Entities\n0 0 0 1\n1 0 0 0 1 1 1 0 0\n$EndEntities\n")
        f.write(f"$Nodes\n2 {n} {min(tags)} {max(tags)}\n")
        for lo, hi in ((0, h), (h, n)):
            f.write(f"3 1 0 {hi - lo}\n")
            for i in range(lo, hi):
                f.write(f"{tags[i]}\n")
            for i in range(lo, hi):
                f.write(f"{float(V[i][0])!r} {float(V[i][1])!r} {float(V[i][2])!r}\n")
        f.write("$EndNodes\n")
        m = len(E) // 2
        f.write(f"$Elements\n3 {len(E) + 1} 1 {len(E) + 1}\n")
        f.write(f"2 1 2 1\n{len(E) + 1} {tags[0]} {tags[1]} {tags[2]}\n")
        for lo, hi in ((0, m), (m, len(E))):
            f.write(f"3 1 {etype} {hi - lo}\n")
            for k in range(lo, hi):
                f.write(f"{k + 1} " + " ".join(str(tags[i]) for i in E[k]) + "\n")
        f.write("$EndElements\n")


def test_round_trip_through_other_formats(ps, tmp_path):
    V, T = ps.IO.read_tet_MSH(os.path.join(MESHES, "duck.msh"))
    rng = np.random.default_rng(0)
    tags = list(rng.choice(10 * len(V), size=len(V), replace=False) + 1)  # sparse, unordered node tags
    p22 = str(tmp_path / "duck22.msh")
    write_msh22_ascii(p22, V, T, tags)
    V2, T2 = ps.IO.read_tet_MSH(p22)
    assert np.array_equal(V2, V) and np.array_equal(T2, T)
    p41 = str(tmp_path / "duck41.msh")
    write_msh41_ascii(p41, V, T, tags, 4)
    V4, T4 = ps.IO.read_tet_MSH(p41)
    assert np.array_equal(V4, V) and np.array_equal(T4, T)


def test_hex_reader_and_surfaces(ps, tmp_path):
    V, H = ps.IO.make_voxel_hex_grid(3, 2, 2, 0.01, 0.02, 0.015)
    tags = list(range(5, 5 + len(V)))
    p = str(tmp_path / "grid.msh")
    write_msh41_ascii(p, V, H, tags, 5)
    V2, H2 = ps.IO.read_hex_MSH(p)
    assert np.array_equal(V2, V) and np.array_equal(H2, H)
    groups, fele = ps.IO.extract_hex_surfaces(V, H)
    assert len(groups) == 1 and len(groups[0]) == 2 * (3 * 2 + 2 * 2 + 3 * 2)
    assert set(edge_counts(groups[0]).values()) == {2}
    assert enclosed_volume(V, groups[0]) == pytest.approx(0.03 * 0.04 * 0.03, rel=1e-12)
    with pytest.raises(RuntimeError):
        ps.IO.read_tet_MSH(str(tmp_path / "missing.msh"))


# ---- the reference's surface extraction, restated literally (io.cpp:570-780)
TET_FACES = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
HEX_FACES = [[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]


def reference_surfaces(V, E):
    ed = E.shape[1]
    table = TET_FACES if ed == 4 else HEX_FACES
    surf, com, fele, seen = [], [], [], set()
    for i, e in enumerate(E):
        for fl in table:
            f = [int(e[k]) for k in fl]
            key = tuple(sorted(f))
            if key not in seen:
                seen.add(key)
                surf.append(list(key) if ed == 4 else f)
                c = np.zeros(3)
                for k in range(ed):
                    c = c + V[e[k]]
                com.append(c / ed)
                fele.append(i)
            else:
                seen.discard(key)
                idx = next(m for m, s in enumerate(surf) if tuple(sorted(s)) == key)
                del surf[idx], com[idx], fele[idx]
    groups, egroups = [], []
    while surf:
        g, ge, verts = [], [], set()
        i = 0
        while i < len(surf):
            f = surf[i]
            if not verts or any(v in verts for v in f):
                a, b, c = V[f[0]], V[f[1]], V[f[2]]
                n = np.cross(b - a, c - a)
                face = list(f)
                if np.dot(n, com[i] - a) > 0:
                    face = [f[0], f[2], f[1]] if ed == 4 else [f[0], f[3], f[2], f[1]]
                g.append(face)
                ge.append(fele[i])
                verts.update(f)
                del surf[i], com[i], fele[i]
                i = 0
                continue
            i += 1
        groups.append(g)
        egroups.append(ge)
    return groups, egroups


def two_bodies(kind):
    sc1 = mesh.beam(nx=4, ny=2, nz=2, element=kind)
    sc2 = mesh.beam(nx=2, ny=2, nz=3, element=kind)
    V = np.concatenate([sc1.X, sc2.X + np.array([0.5, 0.0, 0.0])])
    E = np.concatenate([sc2.conn + len(sc1.X), sc1.conn])  # the second body's elements first
    return V, E


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_surface_groups_match_reference_loops(ps, kind):
    V, E = two_bodies(kind)
    if kind == "tet":
        got = ps.IO.extract_tet_surfaces(V, E)
        # the reference's 5-tet split too (nonconforming faces stay inside as pairs)
        Vh, Hh = ps.IO.make_voxel_hex_grid(2, 2, 1, 1.0, 1.0, 1.0)
        T5 = ps.IO.hex_to_5_tets(Hh)
        assert ps.IO.extract_tet_surfaces(Vh, T5) == reference_surfaces(Vh, T5)
    else:
        got = ps.IO.extract_hex_surfaces(V, E)
    ref = reference_surfaces(V, E)
    assert got == ref
    assert len(got[0]) == 2
    # closed-form face counts of the two boxes (Kuhn: two triangles per quad)
    per = 2 if kind == "tet" else 1
    sizes = sorted(len(g) for g in got[0])
    assert sizes == sorted([per * 2 * (2 * 2 + 2 * 3 + 2 * 3), per * 2 * (4 * 2 + 2 * 2 + 4 * 2)])


def test_plane_square_points(ps):
    for n in ([0, 0, 1], [1, 0, 0], [-1, 0, 0], [0.3, -0.4, 0.5], [0, -1, 0]):
        p = np.array([0.2, -0.1, 0.4])
        pl = ps.Plane3D(p, np.array(n, float))
        nn = np.array(n, float) / np.linalg.norm(n)
        assert np.allclose(pl.get_normal(), nn)
        sq = np.array(pl.get_square_points()).reshape(4, 3)
        assert np.allclose(sq.mean(0), p)                       # centred on the point
        assert np.allclose((sq - p) @ nn, 0.0, atol=1e-15)      # on the plane
        sides = np.linalg.norm(np.roll(sq, -1, axis=0) - sq, axis=1)
        assert np.allclose(sides, 1.0)                          # a 1 m square
        assert abs(np.dot(sq[1] - sq[0], sq[3] - sq[0])) < 1e-14
        assert pl.is_point_on_plane(sq[2]) and not pl.is_point_on_plane(p + 2e-6 * nn)
        assert np.allclose(pl.create_square_points(2.0), (p + 2 * (sq - p)).reshape(-1))
    # x-hat normal: u = y-hat x n (plane3D.h:47-49), so the first corner is p + (-z + y)/2 rotated
    pl = ps.Plane3D(np.zeros(3), np.array([1.0, 0.0, 0.0]))
    assert np.allclose(np.array(pl.get_square_points())[:3], [0.0, 0.5, -0.5])


def arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {}
    for sec in ("PointData", "CellData", "Points", "Cells"):
        for da in piece.find(sec):
            name = da.get("Name") or sec
            out[(sec, name)] = np.array(da.text.split(), dtype=float)
    return piece, out


def test_vtu_and_pvd(ps, tmp_path):
    V, T = ps.IO.read_tet_MSH(os.path.join(MESHES, "sphere1.msh"))
    vis = ps.VisualOption()
    vis.name = "speed"
    vis.type = ps.VisualType.ScalarVertex
    vis.data = np.linspace(-1, 1, len(V)).reshape(1, -1)
    vis.threshold = 0.0
    vec = ps.VisualOption()
    vec.name = "force"
    vec.type = ps.VisualType.VectorElement
    vec.data = np.tile([[0.1], [0.2], [0.3]], (1, len(T)))
    ground = ps.Plane3D(np.array([0.0, 0.0, -0.1]), np.array([0.0, 0.0, 1.0]))
    out = str(tmp_path / "sphere_0")
    assert ps.IO.save_tet_VTU(out, V, T, [vis, vec], [ground])
    piece, a = arrays(out + ".vtu")
    assert int(piece.get("NumberOfPoints")) == len(V) + 4 and int(piece.get("NumberOfCells")) == len(T) + 1
    pts = a[("Points", "Points")].reshape(-1, 3)
    assert np.array_equal(pts[:len(V)], V)
    assert np.allclose(pts[len(V):], np.array(ground.get_square_points()).reshape(4, 3))
    conn = a[("Cells", "connectivity")].astype(int)
    assert np.array_equal(conn[:-4], T.reshape(-1)) and list(conn[-4:]) == list(range(len(V), len(V) + 4))
    assert np.array_equal(a[("Cells", "offsets")], 4.0 * np.arange(1, len(T) + 2))
    assert list(a[("Cells", "types")][-2:]) == [10, 9]
    speed = a[("PointData", "speed")]
    assert np.array_equal(speed[:len(V)], np.where(vis.data[0] > 0.0, vis.data[0], 0.0))
    assert (speed[len(V):] == 0).all()
    assert np.allclose(a[("PointData", "Displacement")][:len(V)], V.sum(1))
    assert np.array_equal(a[("CellData", "defaultElementColor")], np.r_[np.ones(len(T)), 0.0])
    assert np.allclose(a[("CellData", "force")].reshape(-1, 3)[:-1], [0.1, 0.2, 0.3])
    bad = ps.VisualOption()
    bad.name = "short"
    bad.type = ps.VisualType.ScalarElement
    bad.data = np.zeros((1, 3))
    with pytest.raises(RuntimeError, match="short"):
        ps.IO.save_tet_VTU(out, V, T, [bad])
    Vh, H = ps.IO.make_voxel_hex_grid(2, 1, 1, 1.0, 1.0, 1.0)
    assert ps.IO.save_hex_VTU(str(tmp_path / "grid_0"), Vh, H, [], [ground])
    _, b = arrays(str(tmp_path / "grid_0.vtu"))
    assert np.array_equal(b[("Cells", "offsets")], [8, 16, 20])
    assert list(b[("Cells", "types")]) == [12, 12, 9]
    assert ps.IO.save_pvd("sphere", str(tmp_path), 3, 0.01)
    col = ET.parse(str(tmp_path / "sphere.pvd")).getroot().find("Collection")
    ds = col.findall("DataSet")
    assert [d.get("file") for d in ds] == ["sphere_0.vtu", "sphere_1.vtu", "sphere_2.vtu"]
    assert [float(d.get("timestep")) for d in ds] == [0.0, 0.01, 0.02]
