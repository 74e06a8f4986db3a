"""C-ABI checks that need no GPU: every declared symbol is exported, option validation
and error codes follow the reference pass (ReservoirSplatting.cpp:192-301, 356-375,
1571-1592), and the host-side alias table builder is exact."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_harness as oh
import rs_amd
from rs_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def product():
    return rs_amd.load_library()


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "rs.h")).read()
    declared = set(re.findall(r"\b(rs_[a-z_]+)\s*\(", hdr))
    assert declared == set(abi.PRODUCT_SYMBOLS)
    lib = C.CDLL(rs_amd.library_path())
    for name in declared:
        assert hasattr(lib, name), name


def _create(opts):
    lib = product()
    ctx = C.c_void_p()
    code = lib.rs_create(C.byref(opts), C.byref(ctx))
    return lib, ctx, code


def defaults():
    o = abi.Options()
    product().rs_default_options(C.byref(o))
    return o


def test_default_options_match_reference_header():
    o = defaults()
    # RS/ReservoirSplatting.h:124-184 and RS/Params.slang:83-86
    assert o.samplesPerPixel == 1 and o.maxSurfaceBounces == 9 and o.maxDiffuseBounces == 9
    assert o.useRussianRoulette == 1 and o.misHeuristic == 0 and o.misPowerExponent == 2.0
    assert abs(o.specularRoughnessThreshold - 0.25) < 1e-7
    assert o.temporalReuse == abi.TEMPORAL["GatherOnly"] and o.gatherMechanism == abi.GATHER["Fast"]
    assert o.numTimePartitions == 2 and o.neighborCount == 1 and o.numSpatialIterations == 1
    assert o.spatialGatherRadius == 30.0 and abs(o.artificialFrameTime - 1 / 60) < 1e-9
    assert o.useFixedSeed == 0 and o.fixedSeed == 1


@pytest.mark.parametrize("field,value,clamped", [
    ("samplesPerPixel", 0, 1), ("samplesPerPixel", 40, 16), ("maxSurfaceBounces", 1000, 254),
    ("specularRoughnessThreshold", 1.5, 1.0), ("specularRoughnessThreshold", -0.5, 0.0),
    ("numTimePartitions", 0, 1), ("numTimePartitions", 100, 64),
])
def test_option_clamping(field, value, clamped):
    o = defaults()
    setattr(o, field, value)
    lib, ctx, code = _create(o)
    try:
        assert code == abi.RS_E_BAD_OPTION
        assert lib.rs_last_error(ctx).decode()
        got = abi.Options()
        assert lib.rs_get_options(ctx, C.byref(got)) == 0
        assert getattr(got, field) == pytest.approx(clamped)
    finally:
        lib.rs_destroy(ctx)


def test_surface_bounces_raised_to_max_of_categories():
    o = defaults()
    o.maxSurfaceBounces, o.maxDiffuseBounces, o.maxSpecularBounces, o.maxTransmissionBounces = 3, 7, 2, 1
    lib, ctx, code = _create(o)
    got = abi.Options()
    lib.rs_get_options(ctx, C.byref(got))
    assert code == 0 and got.maxSurfaceBounces == 7
    lib.rs_destroy(ctx)


def test_invalid_enum_rejected():
    o = defaults()
    o.temporalReuse = 9
    lib, ctx, code = _create(o)
    assert code == abi.RS_E_INVALID_ARG and not ctx.value


def test_frame_errors_without_gpu():
    lib, ctx, code = _create(defaults())
    assert code == 0
    cam = rs_amd.scenes.box_room_camera(0, 64, 64)
    io = abi.FrameIO()
    io.width, io.height = 5000, 64
    assert lib.rs_render_frame(ctx, C.byref(cam), C.byref(io), None) == abi.RS_E_FRAME_TOO_LARGE
    io.width, io.height = 0, 64
    assert lib.rs_render_frame(ctx, C.byref(cam), C.byref(io), None) == abi.RS_E_SIZE_MISMATCH
    io.width, io.height = 64, 64
    assert lib.rs_render_frame(ctx, C.byref(cam), C.byref(io), None) == abi.RS_E_NO_SCENE
    assert lib.rs_render_frame(ctx, None, C.byref(io), None) == abi.RS_E_INVALID_ARG
    assert lib.rs_reset(ctx) == 0
    lib.rs_destroy(ctx)


def test_python_properties_mirror():
    rs = rs_amd.ReservoirSplatting({"temporalReuse": "ScatterOnly", "maxSurfaceBounces": 4, "fixedSeed": 7,
                                    "gatherOption": "Clamped", "hybridMISOption": "Balance", "colorFormat": 0})
    p = rs.getProperties()
    assert p["temporalReuse"] == "ScatterOnly" and p["gatherOption"] == "Clamped"
    assert p["maxDiffuseBounces"] == 4 and p["maxSpecularBounces"] == 4
    assert rs.useFixedSeed and rs.fixedSeed == 7
    # the graph script's 'hybridMISOption' is not a pass property (warned and ignored)
    assert any("hybridMISOption" in w for w in rs.warnings)
    rs.useFixedSeed = False
    assert not rs.getProperties().get("fixedSeed")
    with pytest.raises(rs_amd.RSError):
        rs.setProperties({"temporalReuse": "Nope"})
    rs.close()


# ------------------------------------------------------------------ alias table


def alias_pmf(table):
    """Exact pmf of a packed alias table by enumerating its slots."""
    t = table.reshape(-1, 2).astype(np.uint64)
    n = t.shape[0]
    thr = (t[:, 0] >> np.uint64(16)).astype(np.uint16).view(np.float16).astype(np.float64)
    above = (((t[:, 0] & np.uint64(0xFFFF)) << np.uint64(8)) | ((t[:, 1] >> np.uint64(24)) & np.uint64(0xFF))).astype(np.int64)
    below = (t[:, 1] & np.uint64(0xFFFFFF)).astype(np.int64)
    pmf = np.zeros(n)
    np.add.at(pmf, below, thr / n)          # u < threshold -> selectBelow
    np.add.at(pmf, above, (1 - thr) / n)    # u >= threshold -> selectAbove
    return pmf


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (64, 3), (1000, 4)])
def test_alias_table_exact_pmf_and_oracle_identical(n, seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.1, 50.0, n).astype(np.float32)
    tab = np.empty(2 * n, np.uint32)
    inv = C.c_float()
    assert product().rs_build_alias_table(w.ctypes.data, n, tab.ctypes.data, C.byref(inv)) == 0
    pmf = alias_pmf(tab)
    exp = w.astype(np.float64) / w.astype(np.float64).sum()
    # thresholds are fp16: relative precision 2^-11 per slot
    assert np.abs(pmf - exp).max() < 2.0 / n * 2 ** -10 + 1e-12
    assert abs(inv.value - 1.0 / w.astype(np.float64).sum()) / inv.value < 1e-6
    tab2 = np.empty(2 * n, np.uint32)
    inv2 = C.c_float()
    lib = oh.oracle_lib()
    lib.rs_cpu_alias_table.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_float)]
    assert lib.rs_cpu_alias_table(w.ctypes.data, n, tab2.ctypes.data, C.byref(inv2)) == 0
    assert np.array_equal(tab, tab2) and inv.value == inv2.value


def test_scene_beyond_32bit_offsets_refused():
    """Vertex and index arrays are addressed by 32-bit byte offsets on the GPU, so a scene of
    more than 357,913,941 vertices or triangles is refused before any array is read."""
    import numpy as np

    lib, ctx, code = _create(defaults())
    assert code == 0
    f = np.zeros(9, np.float32)
    u = np.zeros(3, np.uint32)
    mats = (abi.Material * 1)()
    for nv, nt in ((400_000_000, 1), (3, 400_000_000)):
        sd = abi.SceneDesc()
        sd.positions = sd.normals = f.ctypes.data
        sd.indices = u.ctypes.data
        sd.triMaterial = u.ctypes.data
        sd.materials = C.cast(mats, C.c_void_p)
        sd.numVertices, sd.numTriangles, sd.numMaterials = nv, nt, 1
        assert lib.rs_set_scene(ctx, C.byref(sd)) == abi.RS_E_UNSUPPORTED
        assert b"357,913,941" in lib.rs_last_error(ctx)
    lib.rs_destroy(ctx)
