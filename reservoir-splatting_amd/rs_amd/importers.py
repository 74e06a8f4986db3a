"""Scene ingestion: Falcor's SceneBuilder and importer plugins in front of the pass.

`import_scene(path)` reads a pbrt-v4 file (.pbrt), a Mitsuba 3 file (.xml) or a Falcor
Python scene script (.pyscene) and returns an `ImportedScene`: the flattened triangle soup
`ReservoirSplatting.setScene` takes, plus the file's camera, environment map and film size.
pbrt and Mitsuba files are parsed and converted by the native builder (include/rs_import.h,
csrc/ingest); a .pyscene script is read by `rs_amd.pyscene`, which interprets the
SceneBuilder calls of its syntax tree (nothing from the file is executed) and drives the
same builder.  Reference: Source/Falcor/Scene/SceneBuilder.cpp, Source/plugins/importers/
{PBRTImporter,MitsubaImporter,PythonImporter}.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class ImportError_(RuntimeError):
    """A scene file could not be read, parsed or converted (the message names file and line)."""


def library_path():
    return os.environ.get("RS_AMD_IMPORT_LIBRARY") or os.path.join(os.path.dirname(_HERE), "lib",
                                                                    "librs_import.so")


def load_import_library():
    """The native ingestion library; raises if it has not been built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError_(f"ingestion library missing: {path} (run __graft_entry__.build())")
        _LIB = abi.bind_import(C.CDLL(path))
    return _LIB


def _mat4(m):
    if m is None:
        return None
    a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(4, 4))
    return a


class SceneBuilder:
    """Thin wrapper of rs_builder: materials, meshes with instance transforms, camera,
    environment map and film, flattened on `build()`."""

    def __init__(self):
        self._lib = load_import_library()
        self._h = self._lib.rs_builder_create()
        if not self._h:
            raise MemoryError("rs_builder_create failed")
        self._keep = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and self._lib is not None:
            self._lib.rs_builder_destroy(h)
            self._h = None

    @property
    def log(self):
        return self._lib.rs_builder_log(self._h).decode()

    def _check(self, rc, what):
        if rc < 0:
            msg = [line for line in self.log.splitlines() if line.startswith("error") or ": " in line]
            raise ImportError_(f"{what} failed ({rc}): {msg[-1] if msg else ''}")
        return rc

    def import_file(self, path):
        return self._check(self._lib.rs_builder_import(self._h, os.path.abspath(path).encode()), f"import {path}")

    def add_material(self, base=(1.0, 1.0, 1.0), roughness=0.0, emissive=(0.0, 0.0, 0.0), metallic=0.0, ior=1.5,
                     lobes=abi.LOBE_ALL, double_sided=0):
        m = abi.Material()
        m.baseColor[:] = [float(x) for x in base[:3]]
        m.roughness = float(roughness)
        m.emissive[:] = [float(x) for x in emissive[:3]]
        m.metallic = float(metallic)
        m.ior = float(ior)
        m.activeLobes = int(lobes)
        m.doubleSided = int(double_sided)
        return self._check(self._lib.rs_builder_add_material(self._h, C.byref(m)), "add_material")

    def add_mesh(self, positions, indices, material, normals=None, transform=None, front_face_cw=False):
        P = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        I = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1, 3))
        N = None if normals is None else np.ascontiguousarray(np.asarray(normals, np.float32).reshape(-1, 3))
        T = _mat4(transform)
        rc = self._lib.rs_builder_add_mesh(self._h, P.ctypes.data, None if N is None else N.ctypes.data, P.shape[0],
                                           I.ctypes.data, I.shape[0], None if T is None else T.ctypes.data,
                                           int(material), int(bool(front_face_cw)))
        return self._check(rc, "add_mesh")

    def add_primitive(self, kind, params, material, transform=None):
        p = np.zeros(4, np.float32)
        p[:len(params)] = params
        T = _mat4(transform)
        rc = self._lib.rs_builder_add_primitive(self._h, abi.PRIM[kind], p.ctypes.data,
                                                None if T is None else T.ctypes.data, int(material))
        return self._check(rc, f"add_primitive {kind}")

    def add_mesh_file(self, path, material, transform=None, smooth_normals=False):
        T = _mat4(transform)
        rc = self._lib.rs_builder_add_mesh_file(self._h, os.path.abspath(path).encode(),
                                                None if T is None else T.ctypes.data, int(material),
                                                int(bool(smooth_normals)))
        return self._check(rc, f"add_mesh_file {path}")

    def add_instance(self, mesh, material, transform=None):
        T = _mat4(transform)
        rc = self._lib.rs_builder_add_instance(self._h, int(mesh), None if T is None else T.ctypes.data,
                                               int(material))
        return self._check(rc, "add_instance")

    def set_camera(self, position, target, up=(0.0, 1.0, 0.0), focal_length=21.0, frame_height=24.0,
                   aperture_radius=0.0, focal_distance=10000.0):
        c = abi.ImportCamera()
        c.position[:] = [float(x) for x in position[:3]]
        c.target[:] = [float(x) for x in target[:3]]
        c.up[:] = [float(x) for x in up[:3]]
        c.focalLength = float(focal_length)
        c.frameHeight = float(frame_height)
        c.apertureRadius = float(aperture_radius)
        c.focalDistance = float(focal_distance)
        c.present = 1
        return self._check(self._lib.rs_builder_set_camera(self._h, C.byref(c)), "set_camera")

    def set_env_map(self, texels, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0)):
        d, keep = abi.env_map_desc(texels, intensity, tint, rotation)
        return self._check(self._lib.rs_builder_set_env_map(self._h, C.byref(d)), "set_env_map")

    def set_env_map_file(self, path, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0),
                         octahedral=False):
        t = (C.c_float * 3)(*[float(x) for x in tint])
        r = (C.c_float * 3)(*[float(x) for x in rotation])
        rc = self._lib.rs_builder_set_env_map_file(self._h, os.path.abspath(path).encode(), float(intensity), t, r,
                                                   int(bool(octahedral)))
        return self._check(rc, f"set_env_map_file {path}")

    def set_film(self, width, height):
        return self._check(self._lib.rs_builder_set_film(self._h, int(width), int(height)), "set_film")

    def set_diffuse_model(self, model):
        return self._check(self._lib.rs_builder_set_diffuse_model(self._h, int(model)), "set_diffuse_model")

    def build(self):
        return ImportedScene(self)


class ImportedScene:
    """The flattened scene of a builder, usable as `ReservoirSplatting.setScene(scene)`."""

    def __init__(self, builder):
        self._builder = builder
        lib, h = builder._lib, builder._h
        d = abi.SceneDesc()
        builder._check(lib.rs_builder_get_scene(h, C.byref(d)), "get_scene")
        self._desc = d
        nv, nt, nm = d.numVertices, d.numTriangles, d.numMaterials

        def view(ptr, n, dtype, cols):
            if n == 0:
                return np.zeros((0, cols), dtype)
            buf = (C.c_char * (n * cols * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype).reshape(n, cols).copy()

        self.positions = view(d.positions, nv, np.float32, 3)
        self.normals = view(d.normals, nv, np.float32, 3)
        self.indices = view(d.indices, nt, np.uint32, 3)
        self.tri_material = view(d.triMaterial, nt, np.uint32, 1).reshape(-1)
        arr = (abi.Material * nm).from_address(d.materials) if nm else (abi.Material * 0)()
        self.material_array = (abi.Material * nm)(*[arr[i] for i in range(nm)])
        self.materials = [(tuple(m.baseColor), m.roughness, tuple(m.emissive), m.metallic, m.ior, m.activeLobes,
                           m.doubleSided) for m in self.material_array]
        self.diffuse_model = d.diffuseModel
        ni = d.numInstances
        src = (abi.InstanceDesc * ni).from_address(d.instances) if ni else (abi.InstanceDesc * 0)()
        self.instance_array = (abi.InstanceDesc * ni)(*[src[i] for i in range(ni)])
        cam = abi.ImportCamera()
        lib.rs_builder_get_camera(h, C.byref(cam))
        self.camera_info = cam if cam.present else None
        env = abi.EnvMapDesc()
        if lib.rs_builder_get_env_map(h, C.byref(env)) == 0:
            n = env.width * env.height * 4
            tex = np.frombuffer((C.c_float * n).from_address(env.texels), np.float32).reshape(
                env.height, env.width, 4).copy()
            self.env_map = {"texels": tex, "intensity": env.intensity, "tint": tuple(env.tint),
                            "rotation": tuple(env.rotation)}
        else:
            self.env_map = None
        w, h_ = C.c_uint32(), C.c_uint32()
        lib.rs_builder_get_film(h, C.byref(w), C.byref(h_))
        self.film = (w.value, h_.value) if w.value and h_.value else None
        self.log = builder.log + "".join(m + "\n" for m in getattr(builder, "pyscene_log", []))

    @property
    def num_triangles(self):
        return int(self.indices.shape[0])

    def desc(self):
        d = abi.SceneDesc()
        d.positions = self.positions.ctypes.data
        d.normals = self.normals.ctypes.data
        d.numVertices = self.positions.shape[0]
        d.indices = self.indices.ctypes.data
        d.triMaterial = self.tri_material.ctypes.data
        d.numTriangles = self.indices.shape[0]
        d.materials = C.addressof(self.material_array)
        d.numMaterials = len(self.materials)
        d.diffuseModel = self.diffuse_model
        d.instances = C.addressof(self.instance_array) if len(self.instance_array) else None
        d.numInstances = len(self.instance_array)
        return d

    def camera(self, width=None, height=None, shutter=0.0):
        """The file's camera as the pass's rs_camera for a width x height frame (the film
        size by default)."""
        if self.camera_info is None:
            raise ImportError_("the scene defines no camera")
        if width is None or height is None:
            width, height = self.film or (1920, 1080)
        c = self.camera_info
        cam = abi.Camera()
        cam.position[:] = list(c.position)
        cam.target[:] = list(c.target)
        cam.focalLength = c.focalLength
        cam.frameHeight = c.frameHeight
        cam.aspectRatio = float(width) / float(height)
        cam.focalDistance = c.focalDistance
        cam.apertureRadius = c.apertureRadius
        cam.shutterSpeed = float(shutter)
        return cam

    def apply(self, rs_pass):
        """Set this scene, and its environment map if it has one, on a ReservoirSplatting pass."""
        rs_pass.setScene(self)
        if self.env_map is not None:
            e = self.env_map
            rs_pass.setEnvMap(e["texels"], e["intensity"], e["tint"], e["rotation"])


def import_scene(path):
    """Read a .pbrt, .xml (Mitsuba 3) or .pyscene file into an ImportedScene."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".pyscene":
        from . import pyscene
        return pyscene.load(path).build()
    if ext not in (".pbrt", ".xml"):
        raise ImportError_(f"{path}: unsupported scene format '{ext}' (.pbrt, .xml, .pyscene)")
    b = SceneBuilder()
    b.import_file(path)
    return b.build()
