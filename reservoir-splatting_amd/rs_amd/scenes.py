"""Procedural scenes and camera paths of the benchmark workloads.

All random draws come from SplitMix64 seeded with the workload's seed, so a scene is a
pure function of its parameters.  Geometry is emitted as a flattened world-space
triangle soup (positions, normals, indices, per-triangle material) ready for
`rs_set_scene`.

Workloads (sizes as stated):
  * box_room            -- 5 Lambertian walls of 16x16 quads, two boxes of 4x4 quads per
                           face, one 0.5x0.5 quad light of radiance 15 (2,946 triangles).
  * instanced_field     -- ~1M triangles: 700 icospheres (subdivision 3) and 8,000 boxes
                           over a ground plane, GGX roughness U[0.1,0.8], albedo
                           U[0.1,0.9], 64 downward-facing emissive triangles of radiance
                           U[5,50].
  * moving_field        -- ~2M triangles: 256 rigidly moving icospheres (subdivision 4)
                           with seeded velocities up to 0.05 per frame, 0.7M static
                           triangles, 64 emissive triangles.
  * patch_room          -- the box room with 64 GGX + Frostbite patch materials (glossy checks).
  * differentiable_room -- the box room with exactly 64 seeded materials (differentiable workload).
  * batch_workload      -- 16 scenes of uneven resolution and size (sized_field).
  * furnace_cube, sky_env_map -- environment-map lighting checks.
"""
import math

import numpy as np

from . import abi

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform doubles from the top 53 bits."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))


class Scene:
    """World-space triangle soup plus materials."""

    def __init__(self):
        self.pos = []
        self.nrm = []
        self.idx = []
        self.mat_of_tri = []
        self.materials = []
        self.nverts = 0
        self.mesh_vertices = []  # (first vertex, vertex count) of each add_mesh call
        self.mesh_triangles = []  # (first triangle, triangle count) of each add_mesh call
        self.ntris = 0
        self.diffuse_model = abi.DIFFUSE_FROSTBITE

    def add_material(self, base=(0.5, 0.5, 0.5), roughness=0.5, emissive=(0.0, 0.0, 0.0), metallic=0.0,
                     ior=1.5, lobes=abi.LOBE_ALL, double_sided=1):
        self.materials.append((tuple(base), roughness, tuple(emissive), metallic, ior, lobes, double_sided))
        return len(self.materials) - 1

    def add_mesh(self, P, N, I, material):
        P = np.asarray(P, np.float32).reshape(-1, 3)
        N = np.asarray(N, np.float32).reshape(-1, 3)
        I = np.asarray(I, np.uint32).reshape(-1, 3)
        self.pos.append(P)
        self.nrm.append(N)
        self.idx.append(I + np.uint32(self.nverts))
        mat = np.asarray(material, np.uint32)
        self.mat_of_tri.append(np.broadcast_to(mat, (I.shape[0],)).astype(np.uint32))
        self.mesh_vertices.append((self.nverts, P.shape[0]))
        self.mesh_triangles.append((self.ntris, I.shape[0]))
        self.nverts += P.shape[0]
        self.ntris += I.shape[0]

    def finalize(self):
        # a scene without meshes is empty (lit only by an environment map, if any)
        cat = lambda parts, shape, dt: np.ascontiguousarray(  # noqa: E731
            np.concatenate(parts).astype(dt) if parts else np.zeros(shape, dt))
        self.positions = cat(self.pos, (0, 3), np.float32)
        self.normals = cat(self.nrm, (0, 3), np.float32)
        self.indices = cat(self.idx, (0, 3), np.uint32)
        self.tri_material = cat(self.mat_of_tri, (0,), np.uint32)
        arr = (abi.Material * len(self.materials))()
        for i, (b, r, e, m, ior, lobes, ds) in enumerate(self.materials):
            arr[i].baseColor[:] = [float(x) for x in b]
            arr[i].roughness = float(r)
            arr[i].emissive[:] = [float(x) for x in e]
            arr[i].metallic = float(m)
            arr[i].ior = float(ior)
            arr[i].activeLobes = int(lobes)
            arr[i].doubleSided = int(ds)
        self.material_array = arr
        # every add_mesh call is one instance (rs_instance_desc), placed with the identity
        inst = (abi.InstanceDesc * len(self.mesh_vertices))()
        for i, ((fv, nv), (ft, nt)) in enumerate(zip(self.mesh_vertices, self.mesh_triangles)):
            inst[i].firstVertex, inst[i].numVertices, inst[i].firstTriangle, inst[i].numTriangles = fv, nv, ft, nt
        self.instance_array = inst
        del self.pos, self.nrm, self.idx, self.mat_of_tri
        return self

    @property
    def num_instances(self):
        return len(self.mesh_vertices)

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
        d.materials = C_addr(self.material_array)
        d.numMaterials = len(self.materials)
        d.diffuseModel = self.diffuse_model
        d.instances = C_addr(self.instance_array)
        d.numInstances = len(self.instance_array)
        return d


def C_addr(obj):
    import ctypes
    return ctypes.addressof(obj)


# ------------------------------------------------------------------ primitives

def quad_grid(origin, du, dv, n):
    """n x n grid of quads spanning origin + [0,1]du + [0,1]dv, normal = du x dv."""
    origin, du, dv = (np.asarray(x, np.float64) for x in (origin, du, dv))
    g = np.linspace(0.0, 1.0, n + 1)
    uu, vv = np.meshgrid(g, g, indexing="ij")
    P = origin + uu[..., None] * du + vv[..., None] * dv
    P = P.reshape(-1, 3)
    nrm = np.cross(du, dv)
    nrm = nrm / np.linalg.norm(nrm)
    N = np.broadcast_to(nrm, P.shape)
    I = []
    for i in range(n):
        for j in range(n):
            a = i * (n + 1) + j
            b = (i + 1) * (n + 1) + j
            c = (i + 1) * (n + 1) + j + 1
            d = i * (n + 1) + j + 1
            I += [a, b, c, a, c, d]
    return P, N, np.array(I, np.uint32)


def box_mesh(center, size, rot_y_deg, n):
    """Axis-aligned box (outward faces, n x n quads per face) rotated about y."""
    hx, hy, hz = (0.5 * s for s in size)
    faces = [
        ((hx, -hy, hz), (0, 0, -2 * hz), (0, 2 * hy, 0)),    # +x
        ((-hx, -hy, -hz), (0, 0, 2 * hz), (0, 2 * hy, 0)),   # -x
        ((-hx, hy, hz), (2 * hx, 0, 0), (0, 0, -2 * hz)),    # +y
        ((-hx, -hy, -hz), (2 * hx, 0, 0), (0, 0, 2 * hz)),   # -y
        ((-hx, -hy, hz), (2 * hx, 0, 0), (0, 2 * hy, 0)),    # +z
        ((hx, -hy, -hz), (-2 * hx, 0, 0), (0, 2 * hy, 0)),   # -z
    ]
    a = math.radians(rot_y_deg)
    R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    Ps, Ns, Is, off = [], [], [], 0
    for o, du, dv in faces:
        P, N, I = quad_grid(o, du, dv, n)
        Ps.append(P @ R.T + np.asarray(center))
        Ns.append(N @ R.T)
        Is.append(I + off)
        off += P.shape[0]
    return np.concatenate(Ps), np.concatenate(Ns), np.concatenate(Is)


def icosphere(subdiv):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        NF = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            NF += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = NF
    P = np.array(V, np.float64)
    return P, P.copy(), np.array(F, np.uint32).reshape(-1)


# ------------------------------------------------------------------ workloads

def box_room(seed=0):
    """Box room [-1,1]^3 (front open), two boxes, one quad light of radiance 15."""
    rng = SplitMix64(seed)
    s = Scene()
    s.diffuse_model = abi.DIFFUSE_LAMBERT
    lam = abi.LOBE_DIFFUSE_REFLECTION
    walls = [
        ((-1, -1, 1), (0, 0, -2), (2, 0, 0)),    # floor y=-1, normal +y
        ((-1, 1, -1), (0, 0, 2), (2, 0, 0)),     # ceiling y=1, normal -y
        ((-1, -1, -1), (0, 2, 0), (2, 0, 0)),    # back z=-1 ... normal du x dv
        ((-1, -1, 1), (0, 2, 0), (0, 0, -2)),    # left x=-1
        ((1, -1, -1), (0, 2, 0), (0, 0, 2)),     # right x=1
    ]
    for o, du, dv in walls:
        P, N, I = quad_grid(o, du, dv, 16)
        # orient the normal towards the room's interior
        if np.dot(N[0], -np.asarray(o, np.float64) - 0.5 * (np.asarray(du) + np.asarray(dv))) < 0:
            I = I.reshape(-1, 3)[:, ::-1].reshape(-1)
            N = -N
        albedo = tuple(rng.uniform(0.2, 0.8) for _ in range(3))
        m = s.add_material(albedo, roughness=1.0, lobes=lam)
        s.add_mesh(P, N, I, m)
    for center, size, rot in (((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0), ((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0)):
        P, N, I = box_mesh(center, size, rot, 4)
        albedo = tuple(rng.uniform(0.2, 0.8) for _ in range(3))
        m = s.add_material(albedo, roughness=1.0, lobes=lam)
        s.add_mesh(P, N, I, m)
    P, N, I = quad_grid((-0.25, 0.999, -0.25), (0.5, 0, 0), (0, 0, 0.5), 1)
    m = s.add_material((0.5, 0.5, 0.5), roughness=1.0, emissive=(15.0, 15.0, 15.0), lobes=lam)
    s.add_mesh(P, N, I, m)
    return s.finalize()


def patch_room(seed=5, patches_per_wall=16):
    """The box room with the floor and the back, left and right walls split into
    `patches_per_wall` patches each (64 by default), every patch its own StandardMaterial
    (base colour U[0.2,0.8]^3, roughness U[0.2,0.9], GGX specular and Frostbite diffuse); the
    ceiling, the two boxes and the light as in box_room.  The differentiable-rendering
    workload: its patch materials are the parameters."""
    rng = SplitMix64(seed)
    s = Scene()
    s.diffuse_model = abi.DIFFUSE_FROSTBITE
    walls = [
        ((-1, -1, 1), (0, 0, -2), (2, 0, 0), True),     # floor
        ((-1, 1, -1), (0, 0, 2), (2, 0, 0), False),     # ceiling
        ((-1, -1, -1), (0, 2, 0), (2, 0, 0), True),     # back
        ((-1, -1, 1), (0, 2, 0), (0, 0, -2), True),     # left
        ((1, -1, -1), (0, 2, 0), (0, 0, 2), True),      # right
    ]
    n = 16
    side = int(round(patches_per_wall ** 0.5))
    for o, du, dv, patched in walls:
        P, N, I = quad_grid(o, du, dv, n)
        if np.dot(N[0], -np.asarray(o, np.float64) - 0.5 * (np.asarray(du) + np.asarray(dv))) < 0:
            I = I.reshape(-1, 3)[:, ::-1].reshape(-1)
            N = -N
        if patched:
            ids = []
            for _ in range(side * side):
                ids.append(s.add_material(tuple(rng.uniform(0.2, 0.8) for _ in range(3)),
                                          roughness=rng.uniform(0.2, 0.9)))
            q = np.arange(n * n)
            qi, qj = q // n, q % n
            patch = (qi * side // n) * side + (qj * side // n)
            mats = np.repeat(np.asarray(ids, np.uint32)[patch], 2)   # two triangles per quad
        else:
            mats = s.add_material((0.7, 0.7, 0.7), roughness=0.6)
        s.add_mesh(P, N, I, mats)
    for center, size, rot in (((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0), ((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0)):
        P, N, I = box_mesh(center, size, rot, 4)
        m = s.add_material(tuple(rng.uniform(0.2, 0.8) for _ in range(3)), roughness=rng.uniform(0.2, 0.9))
        s.add_mesh(P, N, I, m)
    P, N, I = quad_grid((-0.25, 0.999, -0.25), (0.5, 0, 0), (0, 0, 0.5), 1)
    m = s.add_material((0.5, 0.5, 0.5), roughness=1.0, emissive=(15.0, 15.0, 15.0), lobes=abi.LOBE_DIFFUSE_REFLECTION)
    s.add_mesh(P, N, I, m)
    return s.finalize()


def differentiable_room(seed=5):
    """The differentiable workload's scene: the box room with exactly 64 materials, each a
    StandardMaterial (GGX specular, Frostbite diffuse) with base colour U[0.2,0.8]^3 and
    roughness U[0.2,0.9] drawn from `seed`: the floor and the back and left walls split into
    4x4 patches (48 materials), the right wall into 15 (its last two cells share one), and the
    quad light's material (radiance 15; its base colour and roughness shape what it reflects).
    The ceiling and the two boxes take patch materials, so the geometry and the material ids
    are the same for every seed and only the 64 parameter sets change: seed 5 gives the target,
    seed 6 the starting point."""
    rng = SplitMix64(seed)
    s = Scene()
    s.diffuse_model = abi.DIFFUSE_FROSTBITE
    walls = [
        ((-1, -1, 1), (0, 0, -2), (2, 0, 0), 16),     # floor
        ((-1, -1, -1), (0, 2, 0), (2, 0, 0), 16),     # back
        ((-1, -1, 1), (0, 2, 0), (0, 0, -2), 16),     # left
        ((1, -1, -1), (0, 2, 0), (0, 0, 2), 15),      # right
    ]
    n = 16
    ids = []
    for o, du, dv, count in walls:
        P, N, I = quad_grid(o, du, dv, n)
        if np.dot(N[0], -np.asarray(o, np.float64) - 0.5 * (np.asarray(du) + np.asarray(dv))) < 0:
            I = I.reshape(-1, 3)[:, ::-1].reshape(-1)
            N = -N
        wall = [s.add_material(tuple(rng.uniform(0.2, 0.8) for _ in range(3)), roughness=rng.uniform(0.2, 0.9))
                for _ in range(count)]
        ids.append(wall)
        q = np.arange(n * n)
        patch = np.minimum((q // n * 4 // n) * 4 + (q % n * 4 // n), count - 1)
        s.add_mesh(P, N, I, np.repeat(np.asarray(wall, np.uint32)[patch], 2))
    P, N, I = quad_grid((-1, 1, -1), (0, 0, 2), (2, 0, 0), n)   # ceiling, facing down
    if np.dot(N[0], (0.0, -1.0, 0.0)) < 0:
        I = I.reshape(-1, 3)[:, ::-1].reshape(-1)
        N = -N
    s.add_mesh(P, N, I, ids[0][5])
    for (center, size, rot), m in zip((((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0),
                                       ((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0)), (ids[1][6], ids[2][9])):
        P, N, I = box_mesh(center, size, rot, 4)
        s.add_mesh(P, N, I, m)
    P, N, I = quad_grid((-0.25, 0.999, -0.25), (0.5, 0, 0), (0, 0, 0.5), 1)
    s.add_mesh(P, N, I, s.add_material(tuple(rng.uniform(0.2, 0.8) for _ in range(3)), roughness=rng.uniform(0.2, 0.9),
                                       emissive=(15.0, 15.0, 15.0)))
    assert len(s.materials) == 64
    return s.finalize()


def instanced_field(seed=2, num_spheres=700, num_boxes=8000, num_lights=64, sphere_subdiv=3):
    """About 1M triangles of instanced icospheres and boxes, GGX materials, emissive triangles."""
    rng = SplitMix64(seed)
    s = Scene()
    s.diffuse_model = abi.DIFFUSE_FROSTBITE
    # ground plane
    P, N, I = quad_grid((-25, 0, 25), (50, 0, 0), (0, 0, -50), 8)
    m = s.add_material((0.5, 0.5, 0.5), roughness=0.6)
    s.add_mesh(P, N, I, m)
    sp, sn, si = icosphere(sphere_subdiv)
    Ps, Ns, Is, Ms = [], [], [], []
    off = s.nverts
    for _ in range(num_spheres):
        sc = rng.uniform(0.2, 1.0)
        c = (rng.uniform(-20, 20), rng.uniform(0, 5), rng.uniform(-20, 20))
        m = s.add_material(tuple(rng.uniform(0.1, 0.9) for _ in range(3)), roughness=rng.uniform(0.1, 0.8))
        Ps.append(sp * sc + np.asarray(c))
        Ns.append(sn)
        Is.append(si + off)
        Ms.append(np.full(si.shape[0] // 3, m, np.uint32))
        off += sp.shape[0]
    bp, bn, bi = box_mesh((0, 0, 0), (1, 1, 1), 0.0, 1)
    for _ in range(num_boxes):
        sc = rng.uniform(0.2, 1.0)
        c = np.array((rng.uniform(-20, 20), rng.uniform(0, 5), rng.uniform(-20, 20)))
        a = math.radians(rng.uniform(0, 360))
        R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
        m = s.add_material(tuple(rng.uniform(0.1, 0.9) for _ in range(3)), roughness=rng.uniform(0.1, 0.8))
        Ps.append((bp * sc) @ R.T + c)
        Ns.append(bn @ R.T)
        Is.append(bi + off)
        Ms.append(np.full(bi.shape[0] // 3, m, np.uint32))
        off += bp.shape[0]
    # the spheres and boxes are one mesh (one instance) of pre-offset indices
    idx = np.concatenate(Is).astype(np.uint32).reshape(-1, 3)
    s.pos.append(np.concatenate(Ps).astype(np.float32))
    s.nrm.append(np.concatenate(Ns).astype(np.float32))
    s.idx.append(idx)
    s.mat_of_tri.append(np.concatenate(Ms))
    s.mesh_vertices.append((s.nverts, off - s.nverts))
    s.mesh_triangles.append((s.ntris, idx.shape[0]))
    s.nverts = off
    s.ntris += idx.shape[0]
    for _ in range(num_lights):
        c = np.array((rng.uniform(-18, 18), rng.uniform(8, 10), rng.uniform(-18, 18)))
        r = rng.uniform(0.5, 1.5)
        a = rng.uniform(0, 2 * math.pi)
        tri = [c + r * np.array((math.cos(a + k * 2 * math.pi / 3), 0.0, math.sin(a + k * 2 * math.pi / 3)))
               for k in range(3)]
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        if n[1] > 0:  # face down
            tri = [tri[0], tri[2], tri[1]]
        L = rng.uniform(5, 50)
        m = s.add_material((0.0, 0.0, 0.0), roughness=1.0, emissive=(L, L, L))
        s.add_mesh(np.array(tri), np.tile((0.0, -1.0, 0.0), (3, 1)), (0, 1, 2), m)
    return s.finalize()


def moving_field(seed=3, num_moving=256, num_static_spheres=500, num_static_boxes=5000, num_lights=64,
                 max_speed=0.05):
    """About 2M triangles: 256 rigidly moving icospheres at subdivision level 4 (5,120
    triangles each, 1.31M) and 0.7M static triangles (500 icospheres at level 3 and 5,000
    boxes), GGX materials, a ground plane and 64 static emissive triangles.

    Returns (scene, velocities): velocities maps each moving instance's mesh index to its
    per-frame displacement (direction uniform on the sphere, length U[0, max_speed]), for
    moved_vertices."""
    rng = SplitMix64(seed)
    s = Scene()
    s.diffuse_model = abi.DIFFUSE_FROSTBITE
    P, N, I = quad_grid((-25, 0, 25), (50, 0, 0), (0, 0, -50), 8)
    s.add_mesh(P, N, I, s.add_material((0.5, 0.5, 0.5), roughness=0.6))

    def material():
        return s.add_material(tuple(rng.uniform(0.1, 0.9) for _ in range(3)), roughness=rng.uniform(0.1, 0.8))

    def place():
        return rng.uniform(0.2, 1.0), np.array((rng.uniform(-20, 20), rng.uniform(0, 5), rng.uniform(-20, 20)))

    # static instances, one mesh
    sp3, sn3, si3 = icosphere(3)
    bp, bn, bi = box_mesh((0, 0, 0), (1, 1, 1), 0.0, 1)
    Ps, Ns, Is, Ms = [], [], [], []
    off = 0
    for k in range(num_static_spheres + num_static_boxes):
        sc, c = place()
        m = material()
        if k < num_static_spheres:
            P, N, I = sp3 * sc + c, sn3, si3
        else:
            a = math.radians(rng.uniform(0, 360))
            R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
            P, N, I = (bp * sc) @ R.T + c, bn @ R.T, bi
        Ps.append(P)
        Ns.append(N)
        Is.append(I + off)
        Ms.append(np.full(I.shape[0] // 3, m, np.uint32))
        off += P.shape[0]
    s.add_mesh(np.concatenate(Ps), np.concatenate(Ns), np.concatenate(Is), np.concatenate(Ms))
    # moving instances, one mesh each
    sp4, sn4, si4 = icosphere(4)
    velocities = {}
    for _ in range(num_moving):
        sc, c = place()
        s.add_mesh(sp4 * sc + c, sn4, si4, material())
        z = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0.0, 2.0 * math.pi)
        rxy = math.sqrt(max(0.0, 1.0 - z * z))
        speed = rng.uniform(0.0, max_speed)
        velocities[len(s.mesh_vertices) - 1] = (speed * rxy * math.cos(phi), speed * z, speed * rxy * math.sin(phi))
    for _ in range(num_lights):
        c = np.array((rng.uniform(-18, 18), rng.uniform(8, 10), rng.uniform(-18, 18)))
        r = rng.uniform(0.5, 1.5)
        a = rng.uniform(0, 2 * math.pi)
        tri = [c + r * np.array((math.cos(a + k * 2 * math.pi / 3), 0.0, math.sin(a + k * 2 * math.pi / 3)))
               for k in range(3)]
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        if n[1] > 0:  # face down
            tri = [tri[0], tri[2], tri[1]]
        L = rng.uniform(5, 50)
        m = s.add_material((0.0, 0.0, 0.0), roughness=1.0, emissive=(L, L, L))
        s.add_mesh(np.array(tri), np.tile((0.0, -1.0, 0.0), (3, 1)), (0, 1, 2), m)
    return s.finalize(), velocities


def furnace_cube(albedo=0.5):
    """A unit cube with flat face normals and a Lambertian material of `albedo`, and no
    emissive triangles: lit only by an environment map.  The cube is convex, so under a
    constant environment radiance L every visible point has radiance albedo * L."""
    s = Scene()
    s.diffuse_model = abi.DIFFUSE_LAMBERT
    P, N, I = box_mesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 30.0, 2)
    s.add_mesh(P, N, I, s.add_material((albedo, albedo, albedo), roughness=1.0, lobes=abi.LOBE_DIFFUSE_REFLECTION))
    return s.finalize()


def sky_env_map(width=128, height=64, sun=(0.35, 0.6), sun_radius=0.04, sun_radiance=200.0):
    """A lat-long sky: a vertical gradient from a bluish zenith to a warm horizon, a grey
    ground below the horizon, and a small bright sun disc at lat-long position `sun`
    (u, v).  Returns an (H, W, 3) float32 image, row 0 at the top (the zenith)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    uu, vv = np.meshgrid(u, v)
    t = np.clip(vv * 2.0, 0.0, 1.0)[..., None]
    sky = (1.0 - t) * np.array([0.3, 0.5, 1.2]) + t * np.array([1.0, 0.8, 0.6])
    ground = np.broadcast_to(np.array([0.25, 0.22, 0.2]), sky.shape)
    img = np.where((vv > 0.5)[..., None], ground, sky)
    du = np.minimum(np.abs(uu - sun[0]), 1.0 - np.abs(uu - sun[0])) * 2.0
    d = np.hypot(du, vv - sun[1])
    img = img + (d < sun_radius)[..., None] * sun_radiance
    return img.astype(np.float32)


def sized_field(seed, triangles):
    """instanced_field at about `triangles` triangles (90% icosphere, 10% box triangles, 64
    lights), for the batched workload's scenes of uneven size."""
    spheres = max(1, int(round(0.9 * triangles / 1280)))
    boxes = max(1, int(round(0.1 * triangles / 12)))
    return instanced_field(seed=seed, num_spheres=spheres, num_boxes=boxes)


def batch_workload(count=16, seed=1000):
    """The batched workload: `count` scenes with resolutions drawn from {320x240, 640x480,
    1280x720, 1920x1080} and triangle counts log-uniform in [1e4, 2e6], scene i seeded
    with seed + i.  Returns [(scene, width, height)]."""
    sizes = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]
    out = []
    for i in range(count):
        rng = SplitMix64(seed + i)
        w, h = sizes[min(3, int(rng.uniform(0.0, 4.0)))]
        tris = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2e6)))))
        out.append((sized_field(seed + i, tris), w, h))
    return out


# ------------------------------------------------------------------ cameras

def focal_length_mm(fov_y_deg, frame_height=24.0):
    return 0.5 * frame_height / math.tan(math.radians(fov_y_deg) * 0.5)


def make_camera(pos, target, fov_y_deg, aspect, aperture=0.0, focal_distance=10000.0, shutter=0.0):
    c = abi.Camera()
    c.position[:] = [float(x) for x in pos]
    c.target[:] = [float(x) for x in target]
    c.frameHeight = 24.0
    c.focalLength = focal_length_mm(fov_y_deg)
    c.aspectRatio = float(aspect)
    c.focalDistance = float(focal_distance)
    c.apertureRadius = float(aperture)
    c.shutterSpeed = float(shutter)
    return c


def moved_vertices(scene, frame, velocities):
    """Positions after `frame` frames of rigid translation: velocities maps a mesh index
    (order of add_mesh calls) to a per-frame displacement."""
    P = scene.positions.copy()
    for mesh, v in velocities.items():
        first, count = scene.mesh_vertices[mesh]
        P[first:first + count] += np.float32(frame) * np.asarray(v, np.float32)
    return np.ascontiguousarray(P)


def instance_transforms(scene, frame, velocities):
    """Row-major 3x4 transforms of every instance after `frame` frames of rigid translation
    (identity for the instances velocities does not name): the transforms that place the
    vertices moved_vertices computes, for rs_update_instances."""
    X = np.zeros((scene.num_instances, 12), np.float32)
    X[:, 0] = X[:, 5] = X[:, 10] = 1.0
    for mesh, v in velocities.items():
        X[mesh, 3::4] = np.float32(frame) * np.asarray(v, np.float32)
    return np.ascontiguousarray(X)


def box_room_camera(frame, width, height, aperture=0.0, focal_distance=10000.0, shutter=0.0, speed=0.01):
    """Pinhole (or thin lens) at (0,0,3.5) looking at the origin, translating +`speed` x per frame, fovY 40."""
    dx = speed * frame
    return make_camera((dx, 0.0, 3.5), (dx, 0.0, 0.0), 40.0, width / height, aperture=aperture,
                       focal_distance=focal_distance, shutter=shutter)


def orbit_camera(frame, width, height, radius=25.0, height_=6.0, deg_per_frame=0.5, fov=60.0, **kw):
    """Orbit around the origin at `radius`, `height_` above ground, fixed degrees per frame."""
    a = math.radians(deg_per_frame * frame)
    return make_camera((radius * math.sin(a), height_, radius * math.cos(a)), (0.0, 1.0, 0.0), fov, width / height, **kw)
