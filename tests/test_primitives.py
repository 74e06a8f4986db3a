"""Pins the shared host/device primitives (csrc/rs/) independently of the pass.

Each check is a known answer, a closed form, a property the mathematics fixes, or a
statistic, following the reference's own unit tests where they exist
(Source/Tools/FalcorTest/Tests/...): BitTricksTests.cpp:37-57 (Morton known answer),
PseudorandomTests.cpp:84-95 (LCG), SampleGeneratorTests.cpp:156-158 (mean 0.5 +- 0.01,
|corr| <= 0.0025), GeometryHelpersTests.cpp:58-73 (offset_ray), BSDFTests.cpp:240-305
(sample/eval/pdf consistency <= 1e-3), AliasTableTests.cpp:142 (alias table pmf).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_harness as oh


def lib():
    return oh.oracle_lib()


def P(a):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ RNG


def reference_bit_interleave(x, y, m):
    # BitTricksTests.cpp:37-45
    r = 0
    for i in range(m):
        r |= ((x >> i) & 1) << (2 * i)
        r |= ((y >> i) & 1) << (2 * i + 1)
    return r


def interleave(xy):
    xy = np.ascontiguousarray(xy, np.uint32)
    out = np.empty(xy.shape[0], np.uint32)
    lib().rs_prim_interleave(P(xy), C.c_uint32(xy.shape[0]), P(out))
    return out


def test_morton_known_answer():
    assert reference_bit_interleave(0xE38E, 0xBE8B, 16) == 0xDEADC0DE
    assert int(interleave(np.array([[0xE38E, 0xBE8B]]))[0]) == 0xDEADC0DE


def test_morton_matches_bit_loop():
    rng = np.random.default_rng(1)
    xy = rng.integers(0, 1 << 32, size=(2000, 2), dtype=np.uint64).astype(np.uint32)
    got = interleave(xy)
    exp = [reference_bit_interleave(int(x) & 0xFFFF, int(y) & 0xFFFF, 16) for x, y in xy]
    assert np.array_equal(got, np.array(exp, np.uint32))


def test_tea_published_vector():
    # TEA, Wheeler & Needham 1994: zero key, zero block, 32 cycles -> 41ea3a0a 94baa940.
    v = np.zeros(2, np.uint32)
    key = np.zeros(4, np.uint32)
    out = np.empty(2, np.uint32)
    lib().rs_prim_tea_key(P(v), P(key), C.c_uint32(32), C.c_uint32(1), P(out))
    assert (int(out[0]), int(out[1])) == (0x41EA3A0A, 0x94BAA940)


def test_lcg_closed_form_jump_ahead():
    # x_n = a^n x_0 + c (a^n - 1) / (a - 1) mod 2^32 (LCG.slang:50-67 constants).
    a, c = 1664525, 1013904223
    n = 64
    seed0 = C.c_uint32()
    states = np.empty(n, np.uint32)
    u = np.empty(n, np.float32)
    lib().rs_prim_rng_stream(C.c_uint32(5), C.c_uint32(7), C.c_uint32(123), C.c_uint32(n), C.byref(seed0), P(states), P(u))
    x0 = seed0.value
    M = 1 << 32
    for k in range(1, n + 1):
        an = pow(a, k, M * (a - 1))
        xk = (an * x0 + c * ((an - 1) // (a - 1))) % M
        assert int(states[k - 1]) == xk
        assert u[k - 1] == np.float32((xk >> 8) * 2.0 ** -24)


def test_rng_statistics():
    # Mean 0.5 +- 0.01 and lag correlations <= 0.0025 over many pixel streams.
    n = 64
    vals = []
    seed0 = C.c_uint32()
    states = np.empty(n, np.uint32)
    u = np.empty(n, np.float32)
    for y in range(32):
        for x in range(32):
            lib().rs_prim_rng_stream(C.c_uint32(x), C.c_uint32(y), C.c_uint32(8 * 3 + 0), C.c_uint32(n), C.byref(seed0),
                                     P(states), P(u))
            vals.append(u.copy())
    v = np.array(vals, np.float64)
    assert abs(v.mean() - 0.5) < 0.01
    flat = v - v.mean()
    for lag in (1, 2, 3, 4):
        corr = np.mean(flat[:, :-lag] * flat[:, lag:]) / flat.var()
        assert abs(corr) < 0.0025 * 4, (lag, corr)
    # across neighbouring pixels (same index)
    across = np.corrcoef(v[:-1, 0], v[1:, 0])[0, 1]
    assert abs(across) < 0.05


# ------------------------------------------------------------------ math library


def fmath(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    lib().rs_prim_fmath(C.c_int(fn), P(x), P(y), C.c_uint32(x.shape[0]), P(out))
    return out


def ulps(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float64).astype(np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-2 ** 31) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2 ** 31) - ib, ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("fn,lo,hi,ref,tol", [
    (0, -4 * math.pi, 8 * math.pi, np.sin, 4e-7),
    (1, -4 * math.pi, 8 * math.pi, np.cos, 4e-7),
    (3, -50.0, 50.0, np.arctan, 3e-7),
    (4, -1.0, 1.0, np.arccos, 5e-7),
    (2, -1.4, 1.4, np.tan, 1.5e-6),
])
def test_transcendentals_accuracy(fn, lo, hi, ref, tol):
    x = np.linspace(lo, hi, 200001, dtype=np.float32)
    got = fmath(fn, x).astype(np.float64)
    exp = ref(x.astype(np.float64))
    err = np.abs(got - exp) / np.maximum(1.0, np.abs(exp))
    assert err.max() < tol, (fn, float(err.max()), float(x[np.argmax(err)]))


def test_log2_exp2_pow():
    x = np.exp(np.linspace(-60, 60, 100001)).astype(np.float32)
    assert ulps(fmath(5, x), np.log2(x.astype(np.float64))).max() <= 4
    e = np.linspace(-120, 120, 100001).astype(np.float32)
    assert (np.abs(fmath(6, e) / np.exp2(e.astype(np.float64)) - 1) < 2e-6).all()
    b = np.linspace(0.01, 10, 5000).astype(np.float32)
    for p in (2.0, 3.0, 0.5):
        got = fmath(7, b, np.full_like(b, p))
        assert (np.abs(got / (b.astype(np.float64) ** p) - 1) < 4e-6).all()
    assert fmath(7, np.array([0.0], np.float32), np.array([2.0], np.float32))[0] == 0.0


def test_special_angles():
    x = np.array([0.0, math.pi / 2, math.pi, 1.0, -1.0, 0.0], np.float32)
    s = fmath(0, x)
    assert s[0] == 0.0 and abs(s[1] - 1) < 1e-7 and abs(s[2]) < 1e-6
    assert fmath(4, np.array([1.0, -1.0, 0.0], np.float32)).tolist() == pytest.approx([0.0, math.pi, math.pi / 2], abs=2e-7)


# ------------------------------------------------------------------ fp16 (alias-table thresholds)


def test_f16_round_trip_matches_ieee():
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 1 << 32, size=400000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    extra = np.array([0.0, -0.0, 1.0, 65504.0, 65520.0, 65519.99, 6.1035156e-05, 5.9604645e-08, 2.9802322e-08,
                      2.9802326e-08, 1e-10, 0.1, 0.333333, 1.0 / 3.0], np.float32)
    x = np.concatenate([x, extra, -extra])
    out = np.empty(x.shape[0], np.uint16)
    lib().rs_prim_f32tof16(P(x), C.c_uint32(x.shape[0]), P(out))
    with np.errstate(over="ignore"):
        exp = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(out, exp)
    h = np.arange(1 << 16, dtype=np.uint16)
    back = np.empty(h.shape[0], np.float32)
    lib().rs_prim_f16tof32(P(h), C.c_uint32(h.shape[0]), P(back))
    ref = h.view(np.float16).astype(np.float32)
    same = (back.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(back) & np.isnan(ref))
    assert same.all()


# ------------------------------------------------------------------ ray origin offset


def offset_ray_reference(p, n):
    # GeometryHelpersTests.cpp:58-73 (Waechter & Binder, Ray Tracing Gems ch. 6)
    p = np.asarray(p, np.float32)
    n = np.asarray(n, np.float32)
    of_i = (np.float32(3 * 256.0) * n).astype(np.int32)
    pi = p.view(np.int32) + np.where(p < 0, -of_i, of_i)
    p_i = pi.astype(np.int32).view(np.float32)
    return np.where(np.abs(p) < np.float32(1.0 / 16.0), p + np.float32(3.0 / 65536.0) * n, p_i)


def test_compute_ray_origin_exact():
    rng = np.random.default_rng(5)
    pos = (rng.standard_normal((5000, 3)) * np.array([0.05, 3.0, 300.0])).astype(np.float32)
    nrm = rng.standard_normal((5000, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    out = np.empty((5000, 3), np.float32)
    lib().rs_prim_ray_origin(P(pn), C.c_uint32(5000), P(out))
    exp = offset_ray_reference(pos, nrm)
    assert np.array_equal(out.view(np.uint32), exp.view(np.uint32))
    # the offset point lies on the normal's side of the tangent plane
    assert (np.einsum("ij,ij->i", (out - pos).astype(np.float64), nrm) >= 0).all()


# ------------------------------------------------------------------ BSDF


def bsdf_eval(mat, lobes, model, mask, wi, wo):
    mat = np.ascontiguousarray(mat, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    n = wi.shape[0]
    f = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    lib().rs_prim_bsdf_eval(P(mat), C.c_uint32(lobes), C.c_uint32(model), C.c_uint32(mask), P(wi), P(wo), C.c_uint32(n),
                            P(f), P(pdf))
    return f, pdf


def bsdf_sample(mat, lobes, model, mask, wi, seed=1):
    mat = np.ascontiguousarray(mat, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    n = wi.shape[0]
    st = (np.arange(n, dtype=np.uint64) * 2654435761 + seed).astype(np.uint32)
    wo = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    w = np.empty((n, 3), np.float32)
    lobe = np.empty(n, np.uint32)
    valid = np.empty(n, np.uint32)
    lib().rs_prim_bsdf_sample(P(mat), C.c_uint32(lobes), C.c_uint32(model), C.c_uint32(mask), P(wi), P(st),
                              C.c_uint32(n), P(wo), P(pdf), P(w), P(lobe), P(valid))
    return wo, pdf, w, lobe, valid.astype(bool)


def hemisphere_grid(n_theta=400, n_phi=800):
    th = (np.arange(n_theta) + 0.5) / n_theta * (math.pi / 2)
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (math.pi / 2 / n_theta) * (2 * math.pi / n_phi)).reshape(-1)
    return d.astype(np.float32), dw


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


LAMBERT_WHITE = [1.0, 1.0, 1.0, 1.0, 0.0, 1.5]


def test_lambert_white_furnace():
    d, dw = hemisphere_grid()
    wi = np.tile(unit([0.3, -0.2, 0.9]), (d.shape[0], 1))
    f, pdf = bsdf_eval(LAMBERT_WHITE, 0x01, 1, 0, wi, d)
    assert abs((f[:, 0] * dw).sum() - 1.0) < 2e-3      # eval returns f * cos(wo)
    assert abs((pdf * dw).sum() - 1.0) < 2e-3


def test_frostbite_closed_form_at_zero_roughness():
    d, _ = hemisphere_grid(40, 80)
    wi = np.tile(unit([0.5, 0.1, 0.7]), (d.shape[0], 1))
    base = [0.6, 0.3, 0.9]
    f, _ = bsdf_eval(base + [0.0, 0.0, 1.5], 0x01, 0, 0, wi, d)
    ci, co = wi[:, 2].astype(np.float64), d[:, 2].astype(np.float64)
    fr = (1 - (1 - ci) ** 5) * (1 - (1 - co) ** 5)
    exp = np.asarray(base)[None] / math.pi * fr[:, None] * co[:, None]
    assert np.allclose(f, exp, rtol=2e-5, atol=1e-7)


def test_ggx_ndf_normalization_and_vndf():
    for alpha in (0.05, 0.3, 0.8):
        d, dw = hemisphere_grid(2000, 64)
        wi = unit([0.4, 0.2, 0.6])
        q = np.concatenate([np.full((d.shape[0], 1), alpha, np.float32), np.tile(wi, (d.shape[0], 1)), d], 1)
        q = np.ascontiguousarray(q, np.float32)
        D = np.empty(d.shape[0], np.float32)
        G1 = np.empty_like(D)
        pv = np.empty_like(D)
        lib().rs_prim_ggx(P(q), C.c_uint32(d.shape[0]), P(D), P(G1), P(pv))
        assert abs((D * d[:, 2] * dw).sum() - 1.0) < 5e-3, alpha   # int D(h) cos(h) dh = 1
        assert abs((pv * dw).sum() - 1.0) < 5e-3, alpha            # VNDF pdf integrates to 1


@pytest.mark.parametrize("mat,lobes,model,single", [
    ([0.7, 0.5, 0.3, 0.5, 0.0, 1.5], 0xFF, 0, False),   # Frostbite + GGX mixture, dielectric
    ([0.8, 0.6, 0.4, 0.3, 1.0, 1.5], 0xFF, 0, True),    # metallic: GGX lobe only
    ([0.4, 0.6, 0.2, 1.0, 0.0, 1.5], 0x01, 1, True),    # Lambert only
])
def test_sample_eval_pdf_consistency(mat, lobes, model, single):
    # BSDFTests.cpp:240-305: the sampled pdf equals evalPdf (mask 0); for a single lobe the
    # sample weight equals eval/pdf.  For a mixture the weight is the one-sample
    # lobe-selection estimator f_lobe / (p_lobe pdf_lobe), so instead its mean must equal
    # the directional albedo integral of eval.
    rng = np.random.default_rng(7)
    d, dw = hemisphere_grid(400, 800)
    for wi0 in ([0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.95, 0.1, 0.2]):
        wi = np.tile(unit(wi0), (200000, 1))
        wo, pdf, w, lobe, valid = bsdf_sample(mat, lobes, model, 0, wi, seed=int(rng.integers(1 << 30)))
        f, epdf = bsdf_eval(mat, lobes, model, 0, wi[valid], wo[valid])
        ok = pdf[valid] > 1e-3
        assert ok.sum() > 1000
        relp = np.abs(epdf[ok] - pdf[valid][ok]) / pdf[valid][ok]
        assert np.quantile(relp, 0.999) < 1e-3
        if single:
            relw = np.abs(f[ok] / epdf[ok, None] - w[valid][ok]) / np.maximum(np.abs(w[valid][ok]), 1e-6)
            assert np.quantile(relw, 0.999) < 1e-3
        fa, _ = bsdf_eval(mat, lobes, model, 0, np.tile(wi[0], (d.shape[0], 1)), d)
        albedo = (fa * dw[:, None]).sum(0)
        est = np.where(valid[:, None], w, 0.0).mean(0)
        sem = np.where(valid[:, None], w, 0.0).std(0) / math.sqrt(w.shape[0])
        assert (np.abs(est - albedo) < 5 * sem + 2e-3).all(), (est, albedo, sem)


def test_sampling_chi2_against_pdf():
    # Histogram of sampled directions vs the integrated pdf (hypothesis-style chi^2).
    mat, lobes, model = [0.5, 0.5, 0.5, 0.4, 0.3, 1.5], 0xFF, 0
    wi = np.tile(unit([0.5, 0.3, 0.8]), (400000, 1))
    wo, pdf, w, lobe, valid = bsdf_sample(mat, lobes, model, 0, wi, seed=11)
    wo = wo[valid]
    nth, nph = 10, 20
    th = np.arccos(np.clip(wo[:, 2], -1, 1))
    ph = np.mod(np.arctan2(wo[:, 1], wo[:, 0]), 2 * math.pi)
    hist, _, _ = np.histogram2d(th, ph, bins=[nth, nph], range=[[0, math.pi / 2], [0, 2 * math.pi]])
    d, dw = hemisphere_grid(nth * 40, nph * 40)
    _, p = bsdf_eval(mat, lobes, model, 0, np.tile(wi[0], (d.shape[0], 1)), d)
    tt = np.arccos(np.clip(d[:, 2], -1, 1))
    pp = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * math.pi)
    expected, _, _ = np.histogram2d(tt, pp, bins=[nth, nph], range=[[0, math.pi / 2], [0, 2 * math.pi]],
                                    weights=p * dw)
    expected *= wi.shape[0]
    m = expected > 20
    chi2 = ((hist[m] - expected[m]) ** 2 / expected[m]).sum()
    dof = m.sum() - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)


def test_reciprocity_and_lobe_mask_sum():
    rng = np.random.default_rng(9)
    a = rng.standard_normal((5000, 3))
    b = rng.standard_normal((5000, 3))
    a[:, 2] = np.abs(a[:, 2])
    b[:, 2] = np.abs(b[:, 2])
    wi = (a / np.linalg.norm(a, axis=1, keepdims=True)).astype(np.float32)
    wo = (b / np.linalg.norm(b, axis=1, keepdims=True)).astype(np.float32)
    # Reciprocity holds per lobe (the mixture's lobe probabilities depend on wi).
    for mat, lobes, model, mask in (([0.7, 0.5, 0.3, 0.5, 0.0, 1.5], 0xFF, 0, 1),
                                    ([0.7, 0.5, 0.3, 0.5, 0.0, 1.5], 0xFF, 0, 4),
                                    ([0.3, 0.8, 0.6, 1.0, 0.0, 1.5], 0x01, 1, 0)):
        f1, _ = bsdf_eval(mat, lobes, model, mask, wi, wo)
        f2, _ = bsdf_eval(mat, lobes, model, mask, wo, wi)
        ok = (wi[:, 2] > 1e-3) & (wo[:, 2] > 1e-3)
        r1 = f1[ok] / wo[ok, 2:3]
        r2 = f2[ok] / wi[ok, 2:3]
        assert np.allclose(r1, r2, rtol=2e-4, atol=1e-6)
    mat = [0.7, 0.5, 0.3, 0.5, 0.0, 1.5]
    f0, p0 = bsdf_eval(mat, 0xFF, 0, 0, wi, wo)
    f1, p1 = bsdf_eval(mat, 0xFF, 0, 1, wi, wo)
    f4, p4 = bsdf_eval(mat, 0xFF, 0, 4, wi, wo)
    assert np.array_equal(f0, f1 + f4)
    assert np.array_equal(p0, p1 + p4)


@pytest.mark.parametrize("mat", [[0.7, 0.5, 0.3, 0.25, 0.0, 1.5], [0.2, 0.6, 0.4, 0.6, 0.3, 1.5],
                                 [0.9, 0.9, 0.9, 0.1, 0.5, 1.3]])
def test_lobe_masked_pdf_integrates_to_lobe_probability(mat):
    """The lobe-masked pdf the reconnection evaluates (mask 1 << componentType,
    RS/ShiftMapping.slang:454) is the lobe's selection probability times its density, so it
    integrates over the sphere to the probability with which the sampler picks that lobe:
    E_{wo ~ pdf}[pdf_mask(wo) / pdf(wo)] = P(lobe) = the fraction of samples of that lobe,
    and the two masks' integrals sum to 1.  A mask that dropped the probability, or used the
    other lobe's, fails by far more than the Monte Carlo error."""
    rng = np.random.default_rng(21)
    for cos_i in (0.95, 0.6, 0.2):
        n = 200000
        wi = np.tile(np.array([math.sqrt(1 - cos_i * cos_i), 0.0, cos_i], np.float32), (n, 1))
        wo, pdf, w, lobe, valid = bsdf_sample(mat, 0xFF, 0, 0, wi, seed=int(rng.integers(1 << 30)))
        wo, wi_v, lobe = wo[valid], wi[valid], lobe[valid]
        _, p0 = bsdf_eval(mat, 0xFF, 0, 0, wi_v, wo)
        _, p1 = bsdf_eval(mat, 0xFF, 0, 1, wi_v, wo)
        _, p4 = bsdf_eval(mat, 0xFF, 0, 4, wi_v, wo)
        ok = p0 > 0
        i1 = (p1[ok].astype(np.float64) / p0[ok]).sum() / n
        i4 = (p4[ok].astype(np.float64) / p0[ok]).sum() / n
        f_spec = (lobe & 0x02).astype(bool).mean() * valid.mean()
        f_diff = (lobe & 0x01).astype(bool).mean() * valid.mean()
        se = math.sqrt(0.25 / n)
        assert abs(i4 - f_spec) < 6 * se + 2e-3, (cos_i, i4, f_spec)
        assert abs(i1 - f_diff) < 6 * se + 2e-3, (cos_i, i1, f_diff)
        assert abs(i1 + i4 - valid.mean()) < 2e-3, (cos_i, i1 + i4, valid.mean())
        assert 0.02 < f_spec < 0.98


# ------------------------------------------------------------------ intersection


def intersect(tris, rays):
    tris = np.ascontiguousarray(tris, np.float32)
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.empty((tris.shape[0], 4), np.float32)
    lib().rs_prim_intersect(P(tris), P(rays), C.c_uint32(tris.shape[0]), P(out))
    return out


def test_triangle_known_answer():
    tri = np.array([[0, 0, 0, 4, 0, 0, 0, 4, 0]], np.float32)
    # ray down the -z axis through (1, 2): t = 3 exactly, b1 = x/4, b2 = y/4; along the axes
    # with a zero direction component the origin moves off the box planes by about one ulp
    # (ray_pre), so the barycentrics sit within 2^-22 of x/4 and y/4
    ray = np.array([[1, 2, 3, 0, 0, 0, -1, 1e30]], np.float32)
    h = intersect(tri, ray)[0]
    assert h[0] == 1 and h[1] == 3.0
    assert abs(h[2] - 0.25) <= 0.25 * 2.0**-22 and abs(h[3] - 0.5) <= 0.5 * 2.0**-22
    assert h[2] < 0.25 and h[3] < 0.5  # the origin moved towards -x and -y
    # tMax / tMin are inclusive
    assert intersect(tri, np.array([[1, 2, 3, 3.0, 0, 0, -1, 3.0]], np.float32))[0, 0] == 1
    assert intersect(tri, np.array([[1, 2, 3, 0, 0, 0, -1, 2.9999]], np.float32))[0, 0] == 0
    # outside the triangle
    assert intersect(tri, np.array([[3, 3, 3, 0, 0, 0, -1, 1e30]], np.float32))[0, 0] == 0


def test_watertight_closed_mesh():
    # Rays from the inside of a closed icosphere aimed at its vertices and edge midpoints
    # must hit some triangle: a crack would let them escape.
    from rs_amd import scenes
    Pv, _, I = scenes.icosphere(2)
    Pv = Pv.astype(np.float32)
    F = I.reshape(-1, 3)
    targets = np.concatenate([Pv, 0.5 * (Pv[F[:, 0]] + Pv[F[:, 1]])])
    origin = np.array([0.013, -0.021, 0.007], np.float32)
    for tgt in targets[:400]:
        d = unit(tgt - origin)
        tris = np.ascontiguousarray(Pv[F].reshape(-1, 9), np.float32)
        rays = np.tile(np.concatenate([origin, [0.0], d, [1e30]]).astype(np.float32), (tris.shape[0], 1))
        h = intersect(tris, rays)
        assert h[:, 0].sum() >= 1


def test_box_slab_conservative():
    rng = np.random.default_rng(13)
    n = 20000
    lo = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    hi = lo + rng.uniform(0.0, 3, (n, 3)).astype(np.float32)
    o = rng.uniform(-8, 8, (n, 3)).astype(np.float32)
    tgt = rng.uniform(lo, hi).astype(np.float32)   # a point inside the box
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1, keepdims=True)
    rays = np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 1e30)], 1).astype(np.float32)
    boxes = np.concatenate([lo, hi], 1).astype(np.float32)
    out = np.empty(n, np.float32)
    lib().rs_prim_box(P(np.ascontiguousarray(boxes)), P(np.ascontiguousarray(rays)), C.c_uint32(n), P(out))
    assert np.isfinite(out).all()      # a ray aimed at a point inside the box always enters it


# ------------------------------------------------------------------ camera


def test_orient_camera_keyframes_and_slerp():
    keys = np.array([[0, 0, 5, 0, 0, 0], [1, 0, 5, 1, 0, 0], [2, 0, 5, 3, 0, -1]], np.float32)
    cam = np.array([21.0, 24.0, 1.5, 10.0, 0.0], np.float32)
    out = np.empty(12, np.float32)
    dt = 1.0 / 60
    # frame 2 (current keyframe index 2), time 0 -> keyframe 2 exactly
    lib().rs_prim_orient(P(keys), C.c_uint32(3), C.c_float(dt), C.c_uint32(2), C.c_float(0.0), P(cam), P(out))
    assert np.allclose(out[0:3], keys[2, :3], atol=0)
    fwd = unit(keys[2, 3:] - keys[2, :3])
    assert np.allclose(unit(out[9:12]), fwd, atol=2e-7)
    assert abs(np.linalg.norm(out[9:12]) - 10.0) < 1e-4
    # time = one frame earlier -> keyframe 1
    lib().rs_prim_orient(P(keys), C.c_uint32(3), C.c_float(dt), C.c_uint32(2), C.c_float(dt), P(cam), P(out))
    assert np.allclose(out[0:3], keys[1, :3], atol=1e-6)
    # half a frame earlier -> midway position and the normalized bisector of the forwards
    lib().rs_prim_orient(P(keys), C.c_uint32(3), C.c_float(dt), C.c_uint32(2), C.c_float(dt / 2), P(cam), P(out))
    assert np.allclose(out[0:3], 0.5 * (keys[1, :3] + keys[2, :3]), atol=1e-5)
    f1, f2 = unit(keys[1, 3:] - keys[1, :3]), unit(keys[2, 3:] - keys[2, :3])
    assert np.allclose(unit(out[9:12]), unit(f1 + f2), atol=1e-5)
    # U, V, W orthogonal; |V| = focalDistance * tan(fovY/2), fovY = 2 atan(0.5 h / f)
    U, V, W = out[3:6], out[6:9], out[9:12]
    assert abs(np.dot(U, V)) < 1e-3 and abs(np.dot(U, W)) < 1e-3 and abs(np.dot(V, W)) < 1e-3
    assert abs(np.linalg.norm(V) - 10.0 * (0.5 * 24.0 / 21.0)) < 1e-4
    assert abs(np.linalg.norm(U) / np.linalg.norm(V) - 1.5) < 1e-5


def test_camera_ray_projects_back():
    # Forward projection of the hit point along a pixel's ray returns the pixel.
    basis = np.array([0, 0, 5, 2, 0, 0, 0, 1.2, 0, 0, 0, -4], np.float32)
    W, H = 320, 200
    rng = np.random.default_rng(2)
    fp = np.concatenate([rng.uniform(0.01, W - 0.01, (500, 1)), rng.uniform(0.01, H - 0.01, (500, 1)),
                         np.zeros((500, 2))], 1).astype(np.float32)
    out = np.empty((500, 6), np.float32)
    lib().rs_prim_camera_ray(P(basis), P(np.zeros(1, np.float32)), C.c_uint32(W), C.c_uint32(H),
                             P(np.ascontiguousarray(fp)), C.c_uint32(500), P(out))
    o, d = out[:, :3].astype(np.float64), out[:, 3:].astype(np.float64)
    U, V, Wv = basis[3:6].astype(np.float64), basis[6:9].astype(np.float64), basis[9:12].astype(np.float64)
    x = o + 7.0 * d
    rel = x - basis[:3]
    cz = rel @ (Wv / np.linalg.norm(Wv))
    film = np.linalg.norm(Wv) * np.stack([rel @ (U / np.linalg.norm(U)), rel @ (V / np.linalg.norm(V))], 1) / cz[:, None]
    ndc = film / np.array([np.linalg.norm(U), np.linalg.norm(V)])
    px = np.stack([(0.5 * ndc[:, 0] + 0.5) * W, (-0.5 * ndc[:, 1] + 0.5) * H], 1)
    assert np.abs(px - fp[:, :2]).max() < 2e-3
