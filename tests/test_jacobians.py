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
