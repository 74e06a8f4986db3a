"""Finite-difference pins of the primary-hit Jacobians recorded by the oracle's path tracer.

The subpixel Jacobian (RS/PathTracer.slang:821-835) is |cos x1| / d^2 / |cos sensor|^3.
For a pinhole camera whose film lies L pixels from the eye, the film-to-surface area
ratio is dA_x1 / dA_film = d^2 cos^3(sensor) / (L^2 |cos x1|), so the recorded value times
the area ratio is 1 / L^2 at every pixel.  The test measures the area ratio by finite
differences (a local affine fit of the hit position against the film position of the
pixels' actual samples) on a tilted emissive plane, so a wrong power of a cosine, a wrong
distance power or a wrong normal shows up as a pixel-dependent product.
"""
import ctypes as C
import math

import numpy as np

import oracle_harness as oh
import rs_amd
from rs_amd import abi, scenes


def tilted_plane(tilt_deg):
    s = scenes.Scene()
    s.diffuse_model = abi.DIFFUSE_LAMBERT
    m = s.add_material((0.5, 0.5, 0.5), roughness=1.0, emissive=(1.0, 1.0, 1.0), lobes=abi.LOBE_DIFFUSE_REFLECTION)
    a = math.radians(tilt_deg)
    du = np.array([40.0 * math.cos(a), 0.0, 40.0 * math.sin(a)])
    dv = np.array([0.0, 40.0, 0.0])
    origin = np.array([0.0, 0.0, -6.0]) - 0.5 * du - 0.5 * dv
    P, N, I = scenes.quad_grid(origin, du, dv, 1)
    s.add_mesh(P, N, I, m)
    return s.finalize()


def options(**kw):
    o = abi.Options()
    rs_amd.load_library().rs_default_options(C.byref(o))
    o.maxSurfaceBounces = o.maxDiffuseBounces = o.maxSpecularBounces = o.maxTransmissionBounces = 1
    o.enableTemporalResampling = 0
    o.enableSpatialResampling = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_subpixel_jacobian_matches_finite_difference_area_ratio():
    W = H = 320
    fov = 50.0
    scene = tilted_plane(45.0)
    p = oh.OraclePass(options())
    p.setScene(scene)
    p.execute(scenes.make_camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), fov, 1.0), W, H)
    rec = oh.decode_recon(p.read_buffer("history_recon", W, H))
    prim = rec["firstHitPrim"].astype(np.int64)
    assert (prim < scene.num_triangles).all(), "the plane must fill the frame"
    assert (rec["pathLength"] == 1).all()
    tri = scene.positions[scene.indices[prim]].astype(np.float64)  # (P, 3 vertices, 3)
    uv = rec["firstHitUV"].astype(np.float64)
    x1 = (1 - uv[:, :1] - uv[:, 1:]) * tri[:, 0] + uv[:, :1] * tri[:, 1] + uv[:, 1:] * tri[:, 2]
    ys, xs = np.divmod(np.arange(W * H), W)
    film = np.stack([xs + rec["subPixel"][:, 0], ys + rec["subPixel"][:, 1]], 1).astype(np.float64)
    J = rec["subPixelJacobian"].astype(np.float64)
    L = (H / 2) / math.tan(math.radians(fov) / 2)
    x1 = x1.reshape(H, W, 3)
    film = film.reshape(H, W, 2)
    J = J.reshape(H, W)
    products = []
    for y in range(3, H - 3, 23):
        for x in range(3, W - 3, 23):
            F = film[y - 2:y + 3, x - 2:x + 3].reshape(-1, 2)
            X = x1[y - 2:y + 3, x - 2:x + 3].reshape(-1, 3)
            Jw = J[y - 2:y + 3, x - 2:x + 3].reshape(-1)
            A = np.concatenate([F - F.mean(0), np.ones((25, 1))], 1)
            coef, *_ = np.linalg.lstsq(A, X, rcond=None)  # rows: d/dfx, d/dfy, value at the mean
            area = np.linalg.norm(np.cross(coef[0], coef[1]))
            # the recorded Jacobian at the same point (the window's mean film position)
            jc, *_ = np.linalg.lstsq(A, Jw, rcond=None)
            products.append(jc[2] * area)
    products = np.array(products)
    assert np.allclose(products, 1.0 / L ** 2, rtol=5e-3), (products.min() * L * L, products.max() * L * L)
    # the plane is tilted, so the recorded values vary strongly across the frame
    assert J.max() / J.min() > 2.0


def test_lens_vertex_jacobian_matches_finite_difference_area_ratio():
    """Thin lens: with the film position fixed, the lens point maps to the primary hit
    through the in-focus point, and the lens-to-surface area ratio is
    (d1/d0)^2 cos(sensor) / |cos x1|; the recorded lens-vertex Jacobian
    (RS/PathTracer.slang:821-835) is its inverse, so their product is 1.  The area ratio
    is measured by a joint affine fit of the hit against (film, lens) over a window."""
    W = H = 256
    fov, r, focus = 50.0, 0.05, 1.5
    scene = tilted_plane(35.0)
    p = oh.OraclePass(options())
    p.setScene(scene)
    p.execute(scenes.make_camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), fov, 1.0, aperture=r, focal_distance=focus), W, H)
    rec = oh.decode_recon(p.read_buffer("history_recon", W, H))
    prim = rec["firstHitPrim"].astype(np.int64)
    assert (prim < scene.num_triangles).all()
    tri = scene.positions[scene.indices[prim]].astype(np.float64)
    uv = rec["firstHitUV"].astype(np.float64)
    x1 = ((1 - uv[:, :1] - uv[:, 1:]) * tri[:, 0] + uv[:, :1] * tri[:, 1] + uv[:, 1:] * tri[:, 2]).reshape(H, W, 3)
    ys, xs = np.divmod(np.arange(W * H), W)
    film = np.stack([xs + rec["subPixel"][:, 0], ys + rec["subPixel"][:, 1]], 1).astype(np.float64).reshape(H, W, 2)
    lens = rec["lensSample"].astype(np.float64).reshape(H, W, 2)
    J = rec["lensVertexJacobian"].astype(np.float64).reshape(H, W)
    products = []
    for y in range(4, H - 4, 19):
        for x in range(4, W - 4, 19):
            F = film[y - 3:y + 4, x - 3:x + 4].reshape(-1, 2)
            Ls = lens[y - 3:y + 4, x - 3:x + 4].reshape(-1, 2)
            X = x1[y - 3:y + 4, x - 3:x + 4].reshape(-1, 3)
            Jw = J[y - 3:y + 4, x - 3:x + 4].reshape(-1)
            A = np.concatenate([F - F.mean(0), Ls - Ls.mean(0), np.ones((49, 1))], 1)
            coef, *_ = np.linalg.lstsq(A, X, rcond=None)
            area_per_lens_area = np.linalg.norm(np.cross(coef[2], coef[3])) / (r * r)
            jc, *_ = np.linalg.lstsq(A, Jw, rcond=None)
            products.append(jc[4] * area_per_lens_area)
    products = np.array(products)
    assert np.allclose(products, 1.0, rtol=1.5e-2), (products.min(), products.max())


def _hit_pos_normal(scene, prim, uv):
    tri = scene.positions[scene.indices[prim]].astype(np.float64)
    x = (1 - uv[:, :1] - uv[:, 1:]) * tri[:, 0] + uv[:, :1] * tri[:, 1] + uv[:, 1:] * tri[:, 2]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    return x, n / np.linalg.norm(n, axis=1, keepdims=True)


def test_secondary_path_jacobian_closed_forms():
    """The secondary-path Jacobian is the area-measure density of the reconnection vertex
    (RS/PathTracer.slang:837-854, NEE :1079-1133).  In the Lambertian box room (cosine
    sampling, pdf = cos/pi) it has closed forms per path shape:
      NEE at v1:             p_select / area of the light triangle (= 0.5 / 0.125 = 4)
      path to v2 (+NEE):     (cos1 / pi) * |cos2| / d12^2
      path through v3:       (cos1 / pi) * |cos2| / d12^2 * (cos(n2, wo2) / pi)"""
    W, H = 48, 40
    room = scenes.box_room(seed=0)
    p = oh.OraclePass(options(maxSurfaceBounces=4, maxDiffuseBounces=4, maxSpecularBounces=4,
                              maxTransmissionBounces=4))
    p.setScene(room)
    p.execute(scenes.box_room_camera(0, W, H), W, H)
    rec = oh.decode_recon(p.read_buffer("history_recon", W, H))
    L = rec["pathLength"]
    nee = rec["lightIsNEE"] == 1
    J = rec["secondaryPathJacobian"].astype(np.float64)
    # NEE at the primary hit: the light's area density
    m = (L == 2) & nee
    assert m.sum() > 50
    assert np.allclose(J[m], 4.0, rtol=1e-5)
    # BSDF-sampled second vertex
    k = (L >= 3)
    assert k.sum() > 200
    x1, n1 = _hit_pos_normal(room, rec["firstHitPrim"][k].astype(np.int64), rec["firstHitUV"][k].astype(np.float64))
    x2, n2 = _hit_pos_normal(room, rec["secondHitPrim"][k].astype(np.int64), rec["secondHitUV"][k].astype(np.float64))
    # The continuation ray leaves from x1 offset along the face normal (computeRayOrigin),
    # so its direction is the one from that origin to x2.
    n1o = np.where(((x2 - x1) * n1).sum(1, keepdims=True) >= 0, n1, -n1)
    pn = np.ascontiguousarray(np.concatenate([x1, n1o], 1), np.float32)
    org = np.empty((pn.shape[0], 3), np.float32)
    oh.oracle_lib().rs_prim_ray_origin(pn.ctypes.data_as(C.c_void_p), C.c_uint32(pn.shape[0]),
                                       org.ctypes.data_as(C.c_void_p))
    w = x2 - org.astype(np.float64)
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    d = x2 - x1
    dist2 = (d * d).sum(1)
    pdf1 = np.abs((n1 * w).sum(1)) / math.pi
    G = np.abs((n2 * w).sum(1)) / dist2
    expect = pdf1 * G
    through_v3 = (L[k] >= 4) | ~nee[k]
    wo2 = rec["secondWo"][k].astype(np.float64)
    pdf2 = np.abs((n2 * wo2).sum(1)) / math.pi
    expect = np.where(through_v3, expect * pdf2, expect)
    assert np.allclose(J[k], expect, rtol=2e-4), np.abs(J[k] / expect - 1).max()


def _masked_pdf(mats, normals, wi, wo, mask):
    """StandardBSDF evalPdf with ReSTIR lobe mask `mask` (1 diffuse reflection, 4 specular
    reflection) at world directions: the isotropic BSDF's pdf depends only on the angles to
    the normal and between the directions, so any tangent frame around the normal serves."""
    lib = oh.oracle_lib()
    t = np.cross(normals, np.where(np.abs(normals[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]]))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(normals, t)
    loc = lambda v: np.ascontiguousarray(np.stack([(v * t).sum(1), (v * b).sum(1), (v * normals).sum(1)], 1),
                                         np.float32)  # noqa: E731
    wi_l, wo_l = loc(wi), loc(wo)
    out = np.empty(wi.shape[0], np.float64)
    f = np.empty(3, np.float32)
    p = np.empty(1, np.float32)
    for i in range(wi.shape[0]):
        m = np.ascontiguousarray(mats[i], np.float32)
        lib.rs_prim_bsdf_eval(m.ctypes.data_as(C.c_void_p), C.c_uint32(0xFF), C.c_uint32(abi.DIFFUSE_FROSTBITE),
                              C.c_uint32(mask[i]), wi_l[i].ctypes.data_as(C.c_void_p), wo_l[i].ctypes.data_as(C.c_void_p),
                              C.c_uint32(1), f.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p))
        out[i] = p[0]
    return out


def test_secondary_path_jacobian_lobe_masked_on_glossy_room():
    """On the GGX + Frostbite patch room the secondary-path Jacobian is the lobe-specific
    density of the sampled component at each scatter vertex (mi.sample with the lobe mask set,
    RS/PathTracer.slang:335-338) times the geometry term of the reconnection vertex
    (:837-854): J = evalPdf_{1 << c1}(x1) |cos2| / d12^2 [* evalPdf_{1 << c2}(x2)], the mask
    of component type c (0 diffuse, 2 specular reflection) being the one the reconnection
    evaluates (RS/ShiftMapping.slang:454).  Specular (c = 2) and diffuse components at both
    vertices, against the recorded Jacobian."""
    W, H = 64, 48
    room = scenes.patch_room()
    p = oh.OraclePass(options(maxSurfaceBounces=4, maxDiffuseBounces=4, maxSpecularBounces=4,
                              maxTransmissionBounces=4))
    p.setScene(room)
    p.execute(scenes.box_room_camera(0, W, H), W, H)
    rec = oh.decode_recon(p.read_buffer("history_recon", W, H))
    L = rec["pathLength"]
    nee = rec["lightIsNEE"] == 1
    c1, c2 = rec["firstComp"], rec["secondComp"]
    k = (L >= 3)
    assert (c1[k] == 2).sum() > 50 and (c1[k] == 0).sum() > 50
    through_v3 = (L >= 4) | ~nee
    assert ((c2 == 2) & k & through_v3).sum() > 20
    J = rec["secondaryPathJacobian"].astype(np.float64)[k]
    p1, p2 = rec["firstHitPrim"][k].astype(np.int64), rec["secondHitPrim"][k].astype(np.int64)
    x1, n1 = _hit_pos_normal(room, p1, rec["firstHitUV"][k].astype(np.float64))
    x2, n2 = _hit_pos_normal(room, p2, rec["secondHitUV"][k].astype(np.float64))
    mats = np.array([list(m[0]) + [m[1], m[3], m[4]] for m in room.materials], np.float64)
    mat1, mat2 = mats[room.tri_material[p1]], mats[room.tri_material[p2]]
    wi1 = rec["firstWi"][k].astype(np.float64)
    n1v = np.where((n1 * wi1).sum(1, keepdims=True) >= 0, n1, -n1)   # double-sided: the normal faces the viewer
    pn = np.ascontiguousarray(np.concatenate([x1, np.where(((x2 - x1) * n1).sum(1, keepdims=True) >= 0, n1, -n1)], 1),
                              np.float32)
    org = np.empty((pn.shape[0], 3), np.float32)
    oh.oracle_lib().rs_prim_ray_origin(pn.ctypes.data_as(C.c_void_p), C.c_uint32(pn.shape[0]),
                                       org.ctypes.data_as(C.c_void_p))
    w12 = x2 - org.astype(np.float64)
    w12 /= np.linalg.norm(w12, axis=1, keepdims=True)
    d = x2 - x1
    G = np.abs((n2 * w12).sum(1)) / (d * d).sum(1)
    pdf1 = _masked_pdf(mat1, n1v, wi1, w12, (1 << c1[k]).astype(np.uint32))
    expect = pdf1 * G
    wo2 = rec["secondWo"][k].astype(np.float64)
    n2v = np.where((n2 * -w12).sum(1, keepdims=True) >= 0, n2, -n2)
    pdf2 = _masked_pdf(mat2, n2v, -w12, wo2, (1 << c2[k]).astype(np.uint32))
    tv = through_v3[k]
    expect = np.where(tv, expect * pdf2, expect)
    rel = np.abs(J / expect - 1)
    # float32 positions and directions feed a sharp GGX lobe (alpha down to 0.04): the bulk
    # agrees to float rounding, the worst grazing configuration to about 1e-4
    assert np.median(rel) < 1e-5, np.median(rel)
    assert rel.max() < 1e-3, rel.max()
    # the wrong mask (total pdf, or the other lobe) would miss by far more than that
    tot = _masked_pdf(mat1, n1v, wi1, w12, np.zeros(k.sum(), np.uint32)) * G
    tot = np.where(tv, tot * pdf2, tot)
    spec = c1[k] == 2
    assert np.median(np.abs(J[spec] / tot[spec] - 1)) > 1e-2
