"""Falcor Python scene scripts (.pyscene) read without executing them.

The reference's PythonImporter (Source/plugins/importers/PythonImporter/PythonImporter.cpp:
140-190) runs the script with `from falcor import *` and a `sceneBuilder` object.  Here the
script's syntax tree is interpreted over a fixed subset of that API, so a scene file can
only build a scene: assignments, attribute stores, for loops over ranges and lists,
if/else, arithmetic and the calls below.  Anything else is an error naming the line.

API subset (Falcor's Python bindings, Source/Falcor/Scene/*.cpp `FALCOR_SCRIPT_BINDING`):
  * float2/float3/float4 vectors with elementwise arithmetic;
  * Transform(translation=, scaling=, rotationEuler=, rotationEulerDeg=, position=,
    target=, up=) composed T * R * S (Transform.cpp:80-110, quatFromEulerAngles and
    quatFromLookAt of Utils/Math/QuaternionMath.h);
  * StandardMaterial(name) with baseColor, roughness, metallic, emissiveColor,
    emissiveFactor, indexOfRefraction, doubleSided (Falcor defaults: base colour 1,
    roughness 0, metallic 0, IoR 1.5, single-sided); specular and diffuse transmission are
    accepted and reported (the modelled BSDF is opaque);
  * TriangleMesh.createQuad/createDisk/createCube/createSphere/createFromFile, and
    TriangleMesh() with addVertex/addTriangle;
  * Camera() with position, target, up, focalLength, frameHeight, focalDistance,
    apertureRadius (the pass orients cameras with up = +y);
  * EnvMap.createFromFile(path) (PFM images here) with intensity, tint, rotation;
  * sceneBuilder.addTriangleMesh, addNode(name, transform, parent), addMeshInstance,
    addMaterial, addCamera, importScene (.pbrt, .xml, .pyscene), the envMap property;
    analytic lights (addLight) are reported and skipped (the pass has none).
"""
import ast
import math
import operator
import os

import numpy as np

from . import abi
from .importers import ImportError_, SceneBuilder


class Vec(tuple):
    """float2/3/4 with elementwise arithmetic and .x .y .z .w."""

    def _op(self, other, f):
        if isinstance(other, (tuple, list)):
            if len(other) != len(self):
                raise TypeError("vector sizes differ")
            return Vec(f(a, b) for a, b in zip(self, other))
        return Vec(f(a, other) for a in self)

    def _rop(self, other, f):
        return Vec(f(other, a) for a in self)

    __add__ = lambda s, o: s._op(o, operator.add)  # noqa: E731
    __sub__ = lambda s, o: s._op(o, operator.sub)  # noqa: E731
    __mul__ = lambda s, o: s._op(o, operator.mul)  # noqa: E731
    __truediv__ = lambda s, o: s._op(o, operator.truediv)  # noqa: E731
    __radd__ = lambda s, o: s._rop(o, operator.add)  # noqa: E731
    __rsub__ = lambda s, o: s._rop(o, operator.sub)  # noqa: E731
    __rmul__ = lambda s, o: s._rop(o, operator.mul)  # noqa: E731
    __rtruediv__ = lambda s, o: s._rop(o, operator.truediv)  # noqa: E731

    def __neg__(self):
        return Vec(-a for a in self)

    @property
    def x(self):
        return self[0]

    @property
    def y(self):
        return self[1]

    @property
    def z(self):
        return self[2]

    @property
    def w(self):
        return self[3]


def _vec(n):
    def make(*args):
        if len(args) == 1 and isinstance(args[0], (int, float)):
            return Vec([float(args[0])] * n)
        flat = []
        for a in args:
            flat.extend(a if isinstance(a, (tuple, list)) else [a])
        if len(flat) != n:
            raise TypeError(f"float{n} needs {n} components")
        return Vec(float(v) for v in flat)
    return make


def _quat_from_euler(e):
    c = [math.cos(a * 0.5) for a in e]
    s = [math.sin(a * 0.5) for a in e]
    x = s[0] * c[1] * c[2] - c[0] * s[1] * s[2]
    y = c[0] * s[1] * c[2] + s[0] * c[1] * s[2]
    z = c[0] * c[1] * s[2] - s[0] * s[1] * c[2]
    w = c[0] * c[1] * c[2] + s[0] * s[1] * s[2]
    return x, y, z, w


def _matrix_from_quat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


class Transform:
    def __init__(self, **kw):
        self.translation = Vec((0.0, 0.0, 0.0))
        self.scaling = Vec((1.0, 1.0, 1.0))
        self._rot = np.eye(3)
        look = {}
        for k, v in kw.items():
            if k == "translation":
                self.translation = Vec(float(a) for a in v)
            elif k == "scaling":
                self.scaling = Vec([float(v)] * 3) if isinstance(v, (int, float)) else Vec(float(a) for a in v)
            elif k == "rotationEuler":
                self.rotationEuler = v
            elif k == "rotationEulerDeg":
                self.rotationEulerDeg = v
            elif k in ("position", "target", "up"):
                look[k] = Vec(float(a) for a in v)
            else:
                raise TypeError(f"Transform: unsupported argument '{k}'")
        if len(look) == 3:
            self.lookAt(look["position"], look["target"], look["up"])

    def __setattr__(self, k, v):
        if k == "rotationEuler":
            self._rot = _matrix_from_quat(_quat_from_euler([float(a) for a in v]))
        elif k == "rotationEulerDeg":
            self._rot = _matrix_from_quat(_quat_from_euler([math.radians(float(a)) for a in v]))
        elif k in ("translation", "scaling", "_rot"):
            object.__setattr__(self, k, v)
        else:
            raise AttributeError(f"Transform has no settable attribute '{k}'")

    def lookAt(self, position, target, up):
        d = np.asarray(target, float) - np.asarray(position, float)
        d /= np.linalg.norm(d)
        z = -d
        r = np.cross(np.asarray(up, float), z)
        r /= np.linalg.norm(r)
        u = np.cross(z, r)
        object.__setattr__(self, "translation", Vec(float(a) for a in position))
        object.__setattr__(self, "_rot", np.stack([r, u, z], 1))

    @property
    def matrix(self):
        T = np.eye(4)
        T[:3, 3] = self.translation
        R = np.eye(4)
        R[:3, :3] = self._rot
        S = np.diag([*self.scaling, 1.0])
        return T @ R @ S


class StandardMaterial:
    _ATTRS = {"baseColor", "roughness", "metallic", "emissiveColor", "emissiveFactor", "indexOfRefraction",
              "doubleSided", "specularTransmission", "diffuseTransmission", "name", "thinSurface",
              "transmissionColor"}

    def __init__(self, name="", **kw):
        self.name = name
        self.baseColor = Vec((1.0, 1.0, 1.0, 1.0))
        self.roughness = 0.0
        self.metallic = 0.0
        self.emissiveColor = Vec((0.0, 0.0, 0.0))
        self.emissiveFactor = 1.0
        self.indexOfRefraction = 1.5
        self.doubleSided = False
        self.specularTransmission = 0.0
        self.diffuseTransmission = 0.0
        self.thinSurface = False
        self.transmissionColor = Vec((1.0, 1.0, 1.0))
        for k, v in kw.items():
            setattr(self, k, v)

    def __setattr__(self, k, v):
        if k not in self._ATTRS:
            raise AttributeError(f"StandardMaterial: unsupported property '{k}'")
        object.__setattr__(self, k, v)


class TriangleMesh:
    def __init__(self):
        self.kind = "arrays"
        self.params = None
        self.path = None
        self.smooth = False
        self.vertices = []
        self.triangles = []
        self.frontFaceCW = False
        self.name = ""

    @staticmethod
    def createQuad(size=Vec((1.0, 1.0))):
        m = TriangleMesh()
        m.kind, m.params = "quad", [float(size[0]), float(size[1])]
        return m

    @staticmethod
    def createDisk(radius=1.0, segments=32):
        m = TriangleMesh()
        m.kind, m.params = "disk", [float(radius), float(segments)]
        return m

    @staticmethod
    def createCube(size=Vec((1.0, 1.0, 1.0))):
        m = TriangleMesh()
        m.kind, m.params = "cube", [float(a) for a in size]
        return m

    @staticmethod
    def createSphere(radius=1.0, segmentsU=32, segmentsV=32):
        m = TriangleMesh()
        m.kind, m.params = "sphere", [float(radius), float(segmentsU), float(segmentsV)]
        return m

    @staticmethod
    def createFromFile(path, smoothNormals=False):
        m = TriangleMesh()
        m.kind, m.path, m.smooth = "file", path, bool(smoothNormals)
        return m

    def addVertex(self, position, normal, texCoord=None):
        self.vertices.append((tuple(position), tuple(normal)))
        return len(self.vertices) - 1

    def addTriangle(self, i0, i1, i2):
        self.triangles.append((int(i0), int(i1), int(i2)))


class Camera:
    _ATTRS = {"position", "target", "up", "focalLength", "frameHeight", "focalDistance", "apertureRadius",
              "nearPlane", "farPlane", "name", "shutterSpeed"}

    def __init__(self, name=""):
        self.name = name
        self.position = Vec((0.0, 0.0, 0.0))
        self.target = Vec((0.0, 0.0, -1.0))
        self.up = Vec((0.0, 1.0, 0.0))
        self.focalLength = 21.0
        self.frameHeight = 24.0
        self.focalDistance = 10000.0
        self.apertureRadius = 0.0
        self.nearPlane = 0.1
        self.farPlane = 1000.0
        self.shutterSpeed = 0.004

    def __setattr__(self, k, v):
        if k not in self._ATTRS:
            raise AttributeError(f"Camera: unsupported property '{k}'")
        object.__setattr__(self, k, v)


class EnvMap:
    def __init__(self, path):
        self.path = path
        self.intensity = 1.0
        self.tint = Vec((1.0, 1.0, 1.0))
        self.rotation = Vec((0.0, 0.0, 0.0))

    @staticmethod
    def createFromFile(path):
        return EnvMap(path)


class _AnalyticLight:
    def __init__(self, name=""):
        object.__setattr__(self, "name", name)

    def __setattr__(self, k, v):
        object.__setattr__(self, k, v)


class PySceneBuilder:
    """The `sceneBuilder` object a script sees; records the scene and fills a builder."""

    def __init__(self, path, log):
        self._dir = os.path.dirname(os.path.abspath(path))
        self._log = log
        self._meshes = []      # (TriangleMesh, StandardMaterial)
        self._nodes = []       # world matrices
        self._instances = []   # (node, mesh id)
        self._materials = []
        self._camera = None
        self._imports = []
        self.envMap = None

    def addTriangleMesh(self, mesh, material, isAnimated=False):
        if not isinstance(mesh, TriangleMesh) or not isinstance(material, StandardMaterial):
            raise TypeError("addTriangleMesh(mesh, material)")
        self._meshes.append((mesh, material))
        return len(self._meshes) - 1

    def addNode(self, name, transform=None, parent=None):
        m = (transform or Transform()).matrix
        if parent is not None:
            m = self._nodes[int(parent)] @ m
        self._nodes.append(m)
        return len(self._nodes) - 1

    def addMeshInstance(self, nodeID, meshID):
        self._instances.append((int(nodeID), int(meshID)))

    def addMaterial(self, material):
        self._materials.append(material)
        return len(self._materials) - 1

    def addCamera(self, camera):
        if self._camera is None:
            self._camera = camera
        return 0

    def addLight(self, light):
        self._log.append(f"light '{getattr(light, 'name', '')}': analytic lights are not modelled by the pass; skipped")
        return 0

    def importScene(self, path, *args, **kw):
        self._imports.append(os.path.join(self._dir, path))

    def loadLightProfile(self, *args, **kw):
        self._log.append("light profiles are not supported and ignored")

    def fill(self, b):
        mat_ids = {}

        def mat_id(m):
            if id(m) not in mat_ids:
                if m.specularTransmission or m.diffuseTransmission:
                    self._log.append(f"material '{m.name}': transmission is not modelled")
                emissive = [float(c) * float(m.emissiveFactor) for c in m.emissiveColor[:3]]
                mat_ids[id(m)] = b.add_material(base=m.baseColor[:3], roughness=m.roughness, emissive=emissive,
                                                metallic=m.metallic, ior=m.indexOfRefraction,
                                                double_sided=1 if m.doubleSided else 0)
            return mat_ids[id(m)]

        for m in self._materials:
            mat_id(m)
        native = {}
        for node, mesh_id in self._instances:
            mesh, material = self._meshes[mesh_id]
            mid = mat_id(material)
            T = self._nodes[node]
            if id(mesh) in native:
                b.add_instance(native[id(mesh)], mid, T)
                continue
            if mesh.kind == "file":
                nid = b.add_mesh_file(os.path.join(self._dir, mesh.path), mid, T, mesh.smooth)
            elif mesh.kind == "arrays":
                if not mesh.triangles:
                    raise ImportError_("TriangleMesh without triangles")
                P = np.array([v[0] for v in mesh.vertices], np.float32)
                N = np.array([v[1] for v in mesh.vertices], np.float32)
                nid = b.add_mesh(P, np.array(mesh.triangles, np.uint32), mid, normals=N, transform=T,
                                 front_face_cw=mesh.frontFaceCW)
            else:
                nid = b.add_primitive(mesh.kind, mesh.params, mid, T)
            native[id(mesh)] = nid
        if self._camera is not None:
            c = self._camera
            b.set_camera(c.position, c.target, c.up, c.focalLength, c.frameHeight, c.apertureRadius, c.focalDistance)
        if self.envMap is not None:
            e = self.envMap
            b.set_env_map_file(os.path.join(self._dir, e.path), e.intensity, e.tint, e.rotation)
        for path in self._imports:
            ext = os.path.splitext(path)[1].lower()
            if ext == ".pyscene":
                load(path, builder=b)
            else:
                b.import_file(path)


_SAFE_MATH = {k: getattr(math, k) for k in ("sin", "cos", "tan", "asin", "acos", "atan", "atan2", "sqrt", "pi",
                                            "radians", "degrees", "floor", "ceil", "exp", "log", "pow", "fabs")}
_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.FloorDiv: operator.floordiv, ast.Mod: operator.mod, ast.Pow: operator.pow}
_CMPOPS = {ast.Eq: operator.eq, ast.NotEq: operator.ne, ast.Lt: operator.lt, ast.LtE: operator.le,
           ast.Gt: operator.gt, ast.GtE: operator.ge}
_CALLABLE_ATTRS = {
    TriangleMesh: {"createQuad", "createDisk", "createCube", "createSphere", "createFromFile", "addVertex",
                   "addTriangle"},
    EnvMap: {"createFromFile"},
    Transform: {"lookAt"},
    PySceneBuilder: {"addTriangleMesh", "addNode", "addMeshInstance", "addMaterial", "addCamera", "addLight",
                     "importScene", "loadLightProfile"},
}


class _Interpreter:
    def __init__(self, path, sb, log):
        self.path = path
        self.log = log
        self.env = {
            "sceneBuilder": sb, "float2": _vec(2), "float3": _vec(3), "float4": _vec(4), "Transform": Transform,
            "StandardMaterial": StandardMaterial, "TriangleMesh": TriangleMesh, "Camera": Camera, "EnvMap": EnvMap,
            "DirectionalLight": _AnalyticLight, "PointLight": _AnalyticLight, "DistantLight": _AnalyticLight,
            "range": range, "len": len, "min": min, "max": max, "abs": abs, "float": float, "int": int,
            "True": True, "False": False, "None": None, "print": lambda *a, **k: None,
        }
        self.sb = sb
        self._steps = 0

    def error(self, node, msg):
        raise ImportError_(f"{self.path}:{getattr(node, 'lineno', 0)}: {msg}")

    def run(self, tree):
        for st in tree.body:
            self.stmt(st)

    def stmt(self, st):
        self._steps += 1
        if self._steps > 10_000_000:
            self.error(st, "script too long")
        if isinstance(st, ast.Expr):
            self.expr(st.value)
        elif isinstance(st, ast.Assign):
            v = self.expr(st.value)
            for t in st.targets:
                self.assign(t, v)
        elif isinstance(st, ast.AugAssign):
            cur = self.expr(st.target if not isinstance(st.target, ast.Name) else ast.Name(st.target.id, ast.Load()))
            self.assign(st.target, _BINOPS[type(st.op)](cur, self.expr(st.value)))
        elif isinstance(st, ast.For):
            for v in self.expr(st.iter):
                self.assign(st.target, v)
                for s in st.body:
                    self.stmt(s)
        elif isinstance(st, ast.If):
            for s in (st.body if self.expr(st.test) else st.orelse):
                self.stmt(s)
        elif isinstance(st, ast.ImportFrom):
            if st.module != "falcor":
                self.error(st, f"import from '{st.module}' is not allowed")
        elif isinstance(st, ast.Import):
            for a in st.names:
                if a.name == "math":
                    self.env[a.asname or "math"] = _SafeModule(_SAFE_MATH)
                elif a.name != "falcor":
                    self.error(st, f"import of '{a.name}' is not allowed")
        elif isinstance(st, ast.Pass):
            pass
        else:
            self.error(st, f"statement '{type(st).__name__}' is not supported in scene scripts")

    def assign(self, t, v):
        if isinstance(t, ast.Name):
            self.env[t.id] = v
        elif isinstance(t, ast.Attribute):
            obj = self.expr(t.value)
            if not isinstance(obj, (StandardMaterial, Camera, Transform, EnvMap, TriangleMesh, PySceneBuilder,
                                    _AnalyticLight)):
                self.error(t, f"cannot set attribute '{t.attr}' here")
            if isinstance(obj, PySceneBuilder) and t.attr != "envMap":
                self.error(t, f"sceneBuilder.{t.attr} cannot be set")
            if isinstance(obj, TriangleMesh) and t.attr not in ("frontFaceCW", "name"):
                self.error(t, f"TriangleMesh.{t.attr} cannot be set")
            try:
                setattr(obj, t.attr, v)
            except (AttributeError, TypeError) as e:
                self.error(t, str(e))
        elif isinstance(t, (ast.Tuple, ast.List)):
            vals = list(v)
            if len(vals) != len(t.elts):
                self.error(t, "unpacking mismatch")
            for e, x in zip(t.elts, vals):
                self.assign(e, x)
        else:
            self.error(t, "unsupported assignment target")

    def expr(self, e):
        if isinstance(e, ast.Constant):
            return e.value
        if isinstance(e, ast.Name):
            if e.id not in self.env:
                self.error(e, f"name '{e.id}' is not defined")
            return self.env[e.id]
        if isinstance(e, (ast.Tuple, ast.List)):
            vals = [self.expr(x) for x in e.elts]
            return tuple(vals) if isinstance(e, ast.Tuple) else vals
        if isinstance(e, ast.BinOp):
            return _BINOPS[type(e.op)](self.expr(e.left), self.expr(e.right))
        if isinstance(e, ast.UnaryOp):
            v = self.expr(e.operand)
            if isinstance(e.op, ast.USub):
                return -v
            if isinstance(e.op, ast.UAdd):
                return v
            if isinstance(e.op, ast.Not):
                return not v
        if isinstance(e, ast.BoolOp):
            vals = [self.expr(x) for x in e.values]
            return all(vals) if isinstance(e.op, ast.And) else any(vals)
        if isinstance(e, ast.Compare):
            left = self.expr(e.left)
            for op, r in zip(e.ops, e.comparators):
                right = self.expr(r)
                if type(op) not in _CMPOPS or not _CMPOPS[type(op)](left, right):
                    return False
                left = right
            return True
        if isinstance(e, ast.IfExp):
            return self.expr(e.body) if self.expr(e.test) else self.expr(e.orelse)
        if isinstance(e, ast.Subscript):
            return self.expr(e.value)[self.expr(e.slice)]
        if isinstance(e, ast.JoinedStr):
            return "".join(str(self.expr(v.value)) if isinstance(v, ast.FormattedValue) else v.value
                           for v in e.values)
        if isinstance(e, ast.Attribute):
            obj = self.expr(e.value)
            if e.attr.startswith("_"):
                self.error(e, "private attributes are not accessible")
            if isinstance(obj, type):
                if e.attr not in _CALLABLE_ATTRS.get(obj, set()):
                    self.error(e, f"{obj.__name__}.{e.attr} is not supported")
                return getattr(obj, e.attr)
            allowed = _CALLABLE_ATTRS.get(type(obj), set())
            data = getattr(type(obj), "_ATTRS", set()) | {"x", "y", "z", "w", "translation", "scaling", "matrix",
                                                          "envMap", "path", "intensity", "tint", "rotation",
                                                          "frontFaceCW", "name"}
            if e.attr not in allowed and e.attr not in data and not isinstance(obj, _SafeModule):
                self.error(e, f"attribute '{e.attr}' is not supported")
            try:
                return getattr(obj, e.attr)
            except AttributeError as ex:
                self.error(e, str(ex))
        if isinstance(e, ast.Call):
            f = self.expr(e.func)
            args = [self.expr(a) for a in e.args]
            kw = {k.arg: self.expr(k.value) for k in e.keywords}
            try:
                return f(*args, **kw)
            except ImportError_:
                raise
            except Exception as ex:  # report, with the line
                self.error(e, f"{type(ex).__name__}: {ex}")
        self.error(e, f"expression '{type(e).__name__}' is not supported in scene scripts")


class _SafeModule:
    def __init__(self, names):
        self.__dict__.update(names)


_ACTIVE = set()


def load(path, builder=None):
    """Interpret a .pyscene file into `builder` (a new SceneBuilder by default) and return it."""
    path = os.path.abspath(path)
    if path in _ACTIVE:
        raise ImportError_(f"{path}: scene is imported recursively")
    with open(path) as f:
        src = f.read()
    try:
        tree = ast.parse(src, path)
    except SyntaxError as e:
        raise ImportError_(f"{path}:{e.lineno}: {e.msg}")
    b = builder or SceneBuilder()
    log = []
    sb = PySceneBuilder(path, log)
    _ACTIVE.add(path)
    try:
        _Interpreter(path, sb, log).run(tree)
        sb.fill(b)
    finally:
        _ACTIVE.discard(path)
    b.pyscene_log = getattr(b, "pyscene_log", []) + [f"{path}: {m}" for m in log]
    return b


__all__ = ["load", "Vec", "Transform", "StandardMaterial", "TriangleMesh", "Camera", "EnvMap", "abi"]
