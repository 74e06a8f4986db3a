"""Scene ingestion (include/rs_import.h, rs_amd.importers, rs_amd.pyscene), pinned by what
the formats and the mathematics fix rather than by restating the importers:

* the library exports every symbol the header declares;
* Falcor's TriangleMesh generators: vertex counts, sphere radius, disk area, outward faces
  (the flattened soup's face normals agree with the vertex normals everywhere);
* instance transforms: positions by the matrix, normals by its inverse transpose (a
  non-uniformly scaled slope keeps its normal perpendicular), a mirroring transform keeps
  the faces outward;
* PLY in its three encodings and OBJ read the same mesh that was written;
* pbrt-v4: the transform directives against numpy products of the pbrt conventions, the
  camera from LookAt and fov, conductor reflectance r reproducing r (eta 1, k = 2 sqrt(r) /
  sqrt(1 - r) has normal-incidence Fresnel r), the importer's roughness, disk and
  area-light-scale behaviours, instancing, Include, errors with file:line;
* Loop subdivision: four times the triangles per level and exact reproduction of a
  planar height field (the scheme reproduces linear functions);
* Mitsuba 3: camera and film, $default substitution, BSDF conversions, to_world order;
* .pyscene: Falcor's Euler convention against scipy's extrinsic xyz, T * R * S, and the
  interpreter refusing anything outside the scene API;
* the same Cornell box written in the three formats gives the same camera, bounds,
  per-material surface area and emitted power.
"""
import math
import os
import re
import struct

import numpy as np
import pytest

from rs_amd import abi, importers, pyscene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def _tri_normals(s):
    P = s.positions[s.indices]
    return np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])


def _areas(s):
    return 0.5 * np.linalg.norm(_tri_normals(s), axis=1)


def _faces_follow_normals(s, tol=0.0):
    fn = _tri_normals(s)
    vn = s.normals[s.indices].mean(axis=1)
    return np.all(np.einsum("ij,ij->i", fn, vn) > tol)


def _builder_with(kind, params, transform=None, base=(0.5, 0.5, 0.5)):
    b = importers.SceneBuilder()
    m = b.add_material(base=base, roughness=1.0)
    b.add_primitive(kind, params, m, transform)
    return b.build()


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "rs_import.h")).read()
    declared = set(re.findall(r"\b(rs_builder_\w+)\s*\(", hdr))
    assert declared == set(abi.IMPORT_SYMBOLS)
    lib = importers.load_import_library()
    for name in abi.IMPORT_SYMBOLS:
        assert hasattr(lib, name), name


def test_primitives():
    q = _builder_with("quad", [2.0, 3.0])
    assert q.num_triangles == 2 and np.isclose(_areas(q).sum(), 6.0)
    assert np.allclose(q.normals, [0, 1, 0]) and _faces_follow_normals(q)
    c = _builder_with("cube", [1.0, 2.0, 3.0])
    assert c.num_triangles == 12 and np.isclose(_areas(c).sum(), 2 * (2 + 3 + 6))
    assert _faces_follow_normals(c)
    s = _builder_with("sphere", [0.7, 32, 16])
    assert s.positions.shape[0] == 33 * 17 and s.num_triangles == 2 * 32 * 16
    assert np.allclose(np.linalg.norm(s.positions, axis=1), 0.7, atol=1e-6)
    fn = _tri_normals(s)
    good = np.linalg.norm(fn, axis=1) > 1e-4 * np.linalg.norm(fn, axis=1).max()  # pole rows are degenerate
    centroid = s.positions[s.indices].mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", fn[good], centroid[good]) > 0)
    n = 24
    d = _builder_with("disk", [1.5, n])
    assert d.num_triangles == n and np.isclose(_areas(d).sum(), 0.5 * n * 1.5**2 * math.sin(2 * math.pi / n))
    assert _faces_follow_normals(d)


def test_instance_transforms():
    # a slope z = x/2 in a quad, scaled non-uniformly: the normal stays perpendicular
    b = importers.SceneBuilder()
    m = b.add_material()
    P = np.array([[0, 0, 0], [1, 0, 0.5], [1, 1, 0.5], [0, 1, 0]], np.float32)
    n = np.cross(P[1] - P[0], P[3] - P[0])
    n /= np.linalg.norm(n)
    N = np.tile(n, (4, 1))
    T = np.diag([3.0, 0.5, 2.0, 1.0])
    T[:3, 3] = [1, 2, 3]
    b.add_mesh(P, [[0, 1, 2], [0, 2, 3]], m, normals=N, transform=T)
    s = b.build()
    assert np.allclose(s.positions, (P @ T[:3, :3].T) + T[:3, 3], atol=1e-6)
    e1, e2 = s.positions[1] - s.positions[0], s.positions[3] - s.positions[0]
    assert abs(np.dot(s.normals[0], e1)) < 1e-6 and abs(np.dot(s.normals[0], e2)) < 1e-6
    assert _faces_follow_normals(s)
    # a mirroring transform keeps the cube's faces outward
    M = np.diag([-1.0, 1.0, 1.0, 1.0])
    c = _builder_with("cube", [1, 1, 1], M)
    assert _faces_follow_normals(c)


def _write_ply(path, P, faces, fmt):
    head = [b"ply", b"format " + fmt.encode() + b" 1.0", b"comment test", b"element vertex %d" % len(P),
            b"property float x", b"property float y", b"property float z", b"property uchar red",
            b"element face %d" % len(faces), b"property list uchar int vertex_indices",
            b"element extra 2", b"property double value", b"end_header"]
    with open(path, "wb") as f:
        f.write(b"\n".join(head) + b"\n")
        if fmt == "ascii":
            for p in P:
                f.write(b"%r %r %r 7\n" % tuple(float(x) for x in p))
            for face in faces:
                f.write((" ".join([str(len(face))] + [str(i) for i in face]) + "\n").encode())
            f.write(b"1.5\n2.5\n")
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            for p in P:
                f.write(struct.pack(e + "fffB", *[float(x) for x in p], 7))
            for face in faces:
                f.write(struct.pack(e + "B" + "i" * len(face), len(face), *face))
            f.write(struct.pack(e + "dd", 1.5, 2.5))


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_encodings(tmp_path, fmt):
    rng = np.random.default_rng(1)
    P = rng.uniform(-1, 1, (6, 3)).astype(np.float32)
    faces = [(0, 1, 2), (2, 3, 4, 5), (0, 2, 5)]  # a quad is triangulated as a fan
    path = str(tmp_path / f"m_{fmt}.ply")
    _write_ply(path, P, faces, fmt)
    b = importers.SceneBuilder()
    m = b.add_material()
    b.add_mesh_file(path, m, smooth_normals=True)
    s = b.build()
    assert np.array_equal(s.positions, P)
    assert s.indices.tolist() == [[0, 1, 2], [2, 3, 4], [2, 4, 5], [0, 2, 5]]
    assert np.allclose(np.linalg.norm(s.normals, axis=1), 1.0, atol=1e-6)
    # flat normals: each triangle keeps its own three vertices and its face normal
    b2 = importers.SceneBuilder()
    b2.add_mesh_file(path, b2.add_material(), smooth_normals=False)
    f = b2.build()
    assert f.positions.shape[0] == 12 and _faces_follow_normals(f)


def test_obj(tmp_path):
    path = tmp_path / "m.obj"
    path.write_text("# quad with normals\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                    "f 1/1/1 2/1/1 3/1/1 4/1/1\nf -4//1 -2//1 -1//1\n")
    b = importers.SceneBuilder()
    m = b.add_material()
    b.add_mesh_file(str(path), m)
    s = b.build()
    assert s.num_triangles == 3
    assert np.allclose(s.normals, [0, 0, 1]) and _faces_follow_normals(s)
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    b2 = importers.SceneBuilder()
    with pytest.raises(importers.ImportError_, match="index out of range"):
        b2.add_mesh_file(str(bad), b2.add_material())


def _pbrt(tmp_path, body, name="s.pbrt"):
    p = tmp_path / name
    p.write_text(body)
    return importers.import_scene(str(p))


def _rot(deg, axis):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    t = math.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(4)
    R[:3, :3] = np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K
    return R


def test_pbrt_transforms_and_camera(tmp_path):
    tri = '"point3 P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2] "normal N" [0 0 1 0 0 1 0 0 1]'
    s = _pbrt(tmp_path, f"""
        Scale -1 1 1
        LookAt 1 2 3  0 0 0  0 1 0
        Camera "perspective" "float fov" 30 "float lensradius" 0.1 "float focaldistance" 4
        Film "rgb" "integer xresolution" 320 "integer yresolution" 200
        WorldBegin
        AttributeBegin
          Translate 1 2 3
          Rotate 30 0 0 1
          Scale 2 2 2
          Shape "trianglemesh" {tri}
        AttributeEnd
        AttributeBegin
          Transform [1 0 0 0  0 1 0 0  0 0 1 0  5 6 7 1]
          ConcatTransform [2 0 0 0  0 1 0 0  0 0 1 0  0 0 0 1]
          Shape "trianglemesh" {tri}
        AttributeEnd
        Shape "trianglemesh" {tri}
    """)
    P0 = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], float)
    T = np.eye(4)
    T[:3, 3] = [1, 2, 3]
    M1 = T @ _rot(30, [0, 0, 1]) @ np.diag([2, 2, 2, 1])
    M2 = np.eye(4)
    M2[:3, 3] = [5, 6, 7]
    M2 = M2 @ np.diag([2, 1, 1, 1])  # pbrt writes matrices for row vectors: translation last
    want = np.concatenate([(P0 @ M1.T)[:, :3], (P0 @ M2.T)[:, :3], P0[:, :3]])
    assert np.allclose(s.positions, want, atol=1e-5)
    c = s.camera_info
    assert np.allclose(c.position, [1, 2, 3], atol=1e-5)
    d = np.array(c.target) - np.array(c.position)
    assert np.allclose(d / np.linalg.norm(d), -np.array([1, 2, 3]) / math.sqrt(14), atol=1e-5)
    assert math.isclose(c.focalLength, 12.0 / math.tan(math.radians(15)), rel_tol=1e-5)
    assert c.apertureRadius == pytest.approx(0.1) and c.focalDistance == pytest.approx(4)
    assert s.film == (320, 200)
    cam = s.camera()
    assert cam.aspectRatio == pytest.approx(1.6)


def test_pbrt_materials_and_lights(tmp_path):
    quad = ('"point3 P" [0 0 0 1 0 0 1 1 0 0 1 0] "integer indices" [0 1 2 0 2 3] '
            '"normal N" [0 0 1 0 0 1 0 0 1 0 0 1]')
    s = _pbrt(tmp_path, f"""
        Camera "perspective"
        WorldBegin
        LightSource "infinite" "rgb L" [0.5 0.6 0.7]
        LightSource "distant" "point3 to" [0 -1 0]
        Texture "tc" "spectrum" "constant" "rgb value" [0.2 0.3 0.4]
        Texture "rc" "float" "constant" "float value" 0.09
        MakeNamedMaterial "copper" "string type" "conductor" "rgb reflectance" [0.9 0.5 0.25] "float roughness" 0.1
        AttributeBegin
          Material "diffuse" "texture reflectance" "tc"
          Shape "trianglemesh" {quad}
        AttributeEnd
        AttributeBegin
          NamedMaterial "copper"
          Shape "trianglemesh" {quad}
        AttributeEnd
        AttributeBegin
          Material "coateddiffuse" "rgb reflectance" [0.1 0.2 0.3] "texture roughness" "rc" "bool remaproughness" false
          Shape "trianglemesh" {quad}
        AttributeEnd
        AttributeBegin
          AreaLightSource "diffuse" "rgb L" [4 5 6] "float scale" 10
          Shape "trianglemesh" {quad}
        AttributeEnd
        AttributeBegin
          Material "interface"
          Translate 5 0 0
          Shape "disk" "float radius" 2 "float height" 0.5
        AttributeEnd
        AttributeBegin
          ReverseOrientation
          Material "dielectric" "float eta" 1.33
          Shape "sphere" "float radius" 0.5
        AttributeEnd
    """)
    mats = s.materials
    tm = s.tri_material
    first = {int(tm[t]) for t in range(0, s.num_triangles)}
    assert len(first) == 6
    diffuse = mats[tm[0]]
    assert np.allclose(diffuse[0], [0.2, 0.3, 0.4]) and diffuse[1] == 1.0 and diffuse[3] == 0.0 and diffuse[6] == 1
    copper = mats[tm[2]]
    # eta 1, k = 2 sqrt(r)/sqrt(1-r): the normal-incidence Fresnel reflectance is r
    assert np.allclose(copper[0], [0.9, 0.5, 0.25], atol=2e-6) and copper[3] == 1.0
    # getRoughness broadcasts u + v: alpha = sqrt(0.1 + 0.1); StandardMaterial roughness sqrt(alpha)
    assert copper[1] == pytest.approx(math.sqrt(math.sqrt(0.2)), rel=1e-6)
    coated = mats[tm[4]]
    assert coated[1] == pytest.approx(math.sqrt(0.09 + 0.09), rel=1e-6)  # not remapped
    light = mats[tm[6]]
    assert np.allclose(light[2], [4, 5, 6]) and np.allclose(light[0], 0.0)  # 'scale' not applied
    # the disk ignores its Translate (placed by height * YtoZ only) and has radius 2
    disk_tris = np.where(tm == tm[8])[0]
    dp = s.positions[s.indices[disk_tris]].reshape(-1, 3)
    assert np.allclose(dp[:, 2], 0.5) and np.isclose(np.abs(dp[:, :2]).max(), 2.0, atol=1e-5)
    interface = mats[tm[8]]
    assert np.allclose(interface[0], 1.0) and interface[1] == 0.0 and interface[6] == 1  # the default material
    glass = mats[tm[-1]]
    assert glass[4] == pytest.approx(1.33) and np.allclose(glass[0], 0.0)
    # ReverseOrientation: the sphere's faces point inwards while its normals stay outward
    sph = np.where(tm == tm[-1])[0]
    fn = _tri_normals(s)[sph]
    good = np.linalg.norm(fn, axis=1) > 1e-4 * np.linalg.norm(fn, axis=1).max()
    cen = s.positions[s.indices[sph]].mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", fn[good], cen[good]) < 0)
    assert s.env_map is not None and s.env_map["texels"].shape == (1, 1, 4)
    assert np.allclose(s.env_map["texels"][0, 0, :3], [0.5, 0.6, 0.7])
    assert "analytic light" in s.log and "specular transmission is not modelled" in s.log


def test_pbrt_errors_name_the_line(tmp_path):
    with pytest.raises(importers.ImportError_, match=r"bad\.pbrt:3: unknown directive 'Frobnicate'"):
        _pbrt(tmp_path, "Camera \"perspective\"\nWorldBegin\nFrobnicate 1 2 3\n", "bad.pbrt")
    with pytest.raises(importers.ImportError_, match="before 'WorldBegin'"):
        _pbrt(tmp_path, "Camera \"perspective\"\n", "nowb.pbrt")
    with pytest.raises(importers.ImportError_, match="not defined"):
        _pbrt(tmp_path, "WorldBegin\nObjectInstance \"nope\"\n", "noobj.pbrt")


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_loop_subdivision_reproduces_a_plane(tmp_path, levels):
    # a 3x3-vertex grid on the plane z = 0.3 x - 0.2 y + 0.1
    xs, ys = np.meshgrid(np.linspace(0, 1, 3), np.linspace(0, 1, 3))
    P = np.stack([xs.ravel(), ys.ravel(), 0.3 * xs.ravel() - 0.2 * ys.ravel() + 0.1], 1)
    idx = []
    for j in range(2):
        for i in range(2):
            a, b, c, d = j * 3 + i, j * 3 + i + 1, (j + 1) * 3 + i + 1, (j + 1) * 3 + i
            idx += [a, b, c, a, c, d]
    s = _pbrt(tmp_path, "WorldBegin\nShape \"loopsubdiv\" \"integer levels\" %d \"integer indices\" [%s] "
              "\"point3 P\" [%s]\n" % (levels, " ".join(map(str, idx)), " ".join(repr(float(v)) for v in P.ravel())))
    assert s.num_triangles == 8 * 4**levels
    X = s.positions.astype(np.float64)
    assert np.allclose(X[:, 2], 0.3 * X[:, 0] - 0.2 * X[:, 1] + 0.1, atol=2e-6)
    n = np.array([-0.3, 0.2, 1.0]) / np.linalg.norm([-0.3, 0.2, 1.0])
    assert np.allclose(s.normals, n, atol=1e-5)
    assert X[:, :2].min() >= -1e-6 and X[:, :2].max() <= 1 + 1e-6


def test_mitsuba(tmp_path):
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    p = tmp_path / "s.xml"
    p.write_text("""<scene version="3.0.0">
      <default name="w" value="400"/>
      <sensor type="thinlens">
        <float name="fov" value="50"/><float name="aperture_radius" value="0.05"/>
        <float name="focus_distance" value="3"/>
        <transform name="to_world"><lookat origin="0, 1, 5" target="0, 1, 0" up="0, 1, 0"/></transform>
        <film type="hdrfilm"><integer name="width" value="$w"/><integer name="height" value="300"/></film>
      </sensor>
      <bsdf type="roughplastic" id="p"><rgb name="diffuse_reflectance" value="0.2, 0.4, 0.6"/>
        <float name="alpha" value="0.25"/></bsdf>
      <bsdf type="conductor" id="au"><rgb name="eta" value="0.2, 0.5, 1.5"/><rgb name="k" value="3.0, 2.3, 1.9"/></bsdf>
      <shape type="obj"><string name="filename" value="tri.obj"/><ref id="p"/>
        <transform name="to_world"><scale value="2"/><translate x="1"/></transform></shape>
      <shape type="sphere"><float name="radius" value="0.5"/><point name="center" x="0" y="3" z="0"/>
        <ref id="au"/></shape>
      <emitter type="constant"><rgb name="radiance" value="0.25"/></emitter>
    </scene>""")
    s = importers.import_scene(str(p))
    assert s.film == (400, 300)
    c = s.camera_info
    assert np.allclose(c.position, [0, 1, 5]) and np.allclose(c.target, [0, 1, 4])
    assert math.isclose(c.focalLength, (24.0 / 300 * 400) / (2 * math.tan(math.radians(25))), rel_tol=1e-5)
    assert c.apertureRadius == pytest.approx(0.05) and c.focalDistance == pytest.approx(3)
    tm = s.tri_material
    plastic = s.materials[tm[0]]
    assert np.allclose(plastic[0], [0.2, 0.4, 0.6]) and plastic[1] == pytest.approx(0.5)
    assert plastic[4] == pytest.approx(1.49 / 1.000277)
    assert np.allclose(s.positions[:3], [[1, 0, 0], [3, 0, 0], [1, 2, 0]])  # scale, then translate
    gold = s.materials[tm[-1]]
    eta, k = np.array([0.2, 0.5, 1.5]), np.array([3.0, 2.3, 1.9])
    assert np.allclose(gold[0], ((eta - 1) ** 2 + k**2) / ((eta + 1) ** 2 + k**2), atol=1e-6) and gold[3] == 1.0
    sph = s.positions[s.indices[tm == tm[-1]]].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(sph - [0, 3, 0], axis=1), 0.5, atol=1e-5)
    assert np.allclose(s.env_map["texels"][0, 0, :3], 0.25)


def test_pyscene_euler_convention_and_composition():
    from scipy.spatial.transform import Rotation
    e = [20.0, -35.0, 110.0]
    t = pyscene.Transform(translation=pyscene.Vec((1, 2, 3)), scaling=pyscene.Vec((2, 3, 4)), rotationEulerDeg=e)
    R = Rotation.from_euler("xyz", e, degrees=True).as_matrix()  # extrinsic x, then y, then z
    want = np.eye(4)
    want[:3, :3] = R @ np.diag([2, 3, 4])
    want[:3, 3] = [1, 2, 3]
    assert np.allclose(t.matrix, want, atol=1e-12)
    la = pyscene.Transform(position=(0, 0, 5), target=(0, 0, 0), up=(0, 1, 0))
    assert np.allclose(la.matrix[:3, 2], [0, 0, 1]) and np.allclose(la.matrix[:3, 3], [0, 0, 5])


@pytest.mark.parametrize("body,msg", [
    ("import os\n", "not allowed"),
    ("x = float3(1, 2, 3).__class__\n", "private attributes"),
    ("def f():\n    pass\n", "not supported"),
    ("m = StandardMaterial('a')\nm.flux = 3\n", "unsupported property"),
    ("open('x')\n", "not defined"),
    ("sceneBuilder.foo = 1\n", "cannot be set"),
])
def test_pyscene_refuses_non_scene_code(tmp_path, body, msg):
    p = tmp_path / "bad.pyscene"
    p.write_text(body)
    with pytest.raises(importers.ImportError_, match=msg):
        importers.import_scene(str(p))


def _stats(s):
    areas = _areas(s)
    per = {}
    for m, a in zip(s.tri_material, areas):
        key = tuple(np.round(s.materials[m][0], 4)) + tuple(np.round(s.materials[m][2], 4))
        per[key] = per.get(key, 0.0) + a
    power = sum(a * np.array(s.materials[m][2]) for m, a in zip(s.tri_material, areas))
    return per, power


def test_cornell_box_same_in_three_formats():
    scenes = {f: importers.import_scene(os.path.join(SCENES, "cornell." + f)) for f in ("pyscene", "pbrt", "xml")}
    ref_per, ref_power = _stats(scenes["pyscene"])
    for fmt, s in scenes.items():
        per, power = _stats(s)
        assert per.keys() == ref_per.keys(), fmt
        for k in per:
            assert per[k] == pytest.approx(ref_per[k], rel=1e-5), (fmt, k)
        assert np.allclose(power, ref_power, rtol=1e-5), fmt
        assert np.allclose(power, 0.25 * np.array([17, 12, 4]), rtol=1e-5)
        assert np.allclose(s.positions.min(0), [-1, -1, -1], atol=1e-5)
        assert np.allclose(s.positions.max(0), [1, 1, 1], atol=1e-5)
        c = s.camera_info
        d = np.array(c.target) - np.array(c.position)
        assert np.allclose(c.position, [0, 0, 3.5]) and np.allclose(d / np.linalg.norm(d), [0, 0, -1], atol=1e-6)
        assert math.isclose(c.focalLength, 12.0 / math.tan(math.radians(20)), rel_tol=1e-5), fmt
        assert _faces_follow_normals(s), fmt
        # the emitter faces down into the room
        lt = np.array([np.any(np.array(s.materials[m][2]) > 0) for m in s.tri_material])
        assert np.all(_tri_normals(s)[lt][:, 1] < 0), fmt
