"""The oracle (and the product's shared polygon test, ray_tri.h) against the
reference's OWN code, where that code compiles from its sources alone:
oracle/ref.mk builds oracle/_ref/libref.so from include/geometry/shape/
shape_polygon.h, include/sampler/{sample,halton}.h, src/sampler/halton.cc,
include/math/math.h (FAST_MATH, FAST_TRIG on, as the reference's default
build) and include/math/filter.h.  Bit-exact comparisons over many inputs.
Skipped where the reference sources are not mounted (the library is built by
__graft_entry__.build() only then)."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

REF_SO = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libref.so"


@pytest.fixture(scope="module")
def ref():
    if not REF_SO.exists():
        pytest.skip("oracle/_ref/libref.so is not built (the reference sources are not mounted)")
    L = C.CDLL(str(REF_SO))
    fp = C.POINTER(C.c_float)
    L.ref_polygon_intersect.argtypes = [C.c_int] + [fp] * 6 + [fp] * 3
    L.ref_polygon_intersect.restype = C.c_int
    L.ymx_polygon_test.argtypes = [C.c_int] + [fp] * 6 + [fp] * 2
    L.ymx_polygon_test.restype = C.c_float
    for name in ("ref_ri_vdc", "ref_ri_s", "ref_ri_lp"):
        getattr(L, name).argtypes = [C.c_uint32, C.c_uint32]
        getattr(L, name).restype = C.c_float
    L.ref_fnv32a.argtypes = [C.c_uint32]
    L.ref_fnv32a.restype = C.c_uint32
    L.ref_low_discrepancy.argtypes = [C.c_int, C.c_uint32]
    L.ref_low_discrepancy.restype = C.c_double
    L.ref_halton_next.argtypes = [C.c_int, C.c_uint32]
    L.ref_halton_next.restype = C.c_float
    L.ref_cos_hemisphere.argtypes = [fp, C.c_float, C.c_float, fp]
    for name in ("ref_fast_log2", "ref_fast_exp2", "ref_fast_log", "ref_fast_sin", "ref_fast_cos"):
        getattr(L, name).argtypes = [C.c_float]
        getattr(L, name).restype = C.c_float
    L.ref_fast_pow.argtypes = [C.c_float, C.c_float]
    L.ref_fast_pow.restype = C.c_float
    L.ref_gauss.argtypes = [C.c_float, C.c_float]
    L.ref_gauss.restype = C.c_float
    return L


def f3(a):
    return (C.c_float * 3)(*[float(x) for x in a])


def bits(x):
    return np.float32(x).view(np.uint32)


def polygon_cases(seed, n):
    """Random triangles and quads (planar and bent), rays aimed near them so
    that hits, edge misses and both quad triangles all occur."""
    rng = np.random.default_rng(seed)
    for _ in range(n):
        v = rng.uniform(-1, 1, (4, 3)).astype(np.float32)
        if rng.random() < 0.5:
            v[3] = (v[0] + v[2] - v[1]).astype(np.float32)  # parallelogram
        o = rng.uniform(-2, 2, 3).astype(np.float32)
        target = v[0] + rng.uniform(-0.2, 1.1) * (v[1] - v[0]) + rng.uniform(-0.2, 1.1) * (v[3] - v[0])
        d = (target - o).astype(np.float32)
        if rng.random() < 0.05:
            d[rng.integers(0, 3)] = 0.0
        yield v, o, d


def test_triangle_and_quad_tests_are_the_references(ref, oracle):
    """ora_triangle_intersect / ora_quad_intersect and the kernels' ray_tri.h
    against ShapePolygon<float, 3 / 4>::intersect itself: hit flag, t, u, v
    bit for bit."""
    hits = {3: 0, 4: 0}
    second = 0
    for v, o, d in polygon_cases(1, 20000):
        for nv in (3, 4):
            t, u, w = C.c_float(), C.c_float(), C.c_float()
            hit = bool(ref.ref_polygon_intersect(nv, f3(v[0]), f3(v[1]), f3(v[2]), f3(v[3]), f3(o), f3(d),
                                                 C.byref(t), C.byref(u), C.byref(w)))
            if nv == 3:
                got = oracle.triangle_intersect(v[0], v[1], v[2], o, d)
            else:
                got = oracle.quad_intersect(v[0], v[1], v[2], v[3], o, d)
            assert got[0] == hit
            if hit:
                hits[nv] += 1
                assert (bits(got[1]), bits(got[2]), bits(got[3])) == (bits(t.value), bits(u.value), bits(w.value))
                second += nv == 4 and w.value > u.value
            pu, pv = C.c_float(), C.c_float()
            th = ref.ymx_polygon_test(nv, f3(v[0]), f3(v[1]), f3(v[2]), f3(v[3]), f3(o), f3(d), C.byref(pu),
                                      C.byref(pv))
            assert (th > 0) == hit
            if hit:
                assert (bits(th), bits(pu.value), bits(pv.value)) == (bits(t.value), bits(u.value), bits(w.value))
    assert hits[3] > 3000 and hits[4] > 5000 and second > 1000


def test_samplers_are_the_references(ref, oracle):
    L = oracle._pt_lib()
    rng = np.random.default_rng(2)
    for i in [*range(0, 4096), *rng.integers(0, 2**32, 4000).tolist()]:
        r = int(rng.integers(0, 2**32))
        for name in ("ri_vdc", "ri_s", "ri_lp"):
            a = getattr(L, "ora_" + name)(int(i), r)
            b = getattr(ref, "ref_" + name)(int(i), r)
            assert bits(a) == bits(b), (name, i, r)
        assert L.ora_fnv32a(int(i)) == ref.ref_fnv32a(int(i))


def test_low_discrepancy_and_halton_are_the_references(ref, oracle):
    L = oracle._pt_lib()
    for dim in range(1, 50):
        for n in [*range(0, 600), 4567, 65535, 1 << 20, 123456789, 2**32 - 1]:
            assert L.ora_low_discrepancy(dim, n) == ref.ref_low_discrepancy(dim, n), (dim, n)
    for base in (2, 3, 5, 7):
        for s in [*range(0, 3000), 4567 * 3, 2**24 + 5]:
            assert bits(L.ora_halton_next(base, s)) == bits(ref.ref_halton_next(base, s)), (base, s)


def test_fast_math_is_the_references(ref, oracle):
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-40, 40, 20000), rng.uniform(0, 2 * math.pi, 20000),
                         [0.0, 0.5, 1.0, 2.0, -1.0, 1e-30, 1e30]]).astype(np.float32)
    lib = oracle.lib()
    for x in xs:
        x = float(x)
        assert bits(lib.ora_fast_sin(x)) == bits(ref.ref_fast_sin(x))
        assert bits(lib.ora_fast_cos(x)) == bits(ref.ref_fast_cos(x))
        assert bits(lib.ora_fast_exp2(x)) == bits(ref.ref_fast_exp2(x))
        if x > 0:
            assert bits(lib.ora_fast_log2(x)) == bits(ref.ref_fast_log2(x))
            assert bits(lib.ora_fast_log(x)) == bits(ref.ref_fast_log(x))
    for a, b in rng.uniform(0, 5, (20000, 2)).astype(np.float32):
        assert bits(lib.ora_fast_pow(float(a), float(b))) == bits(ref.ref_fast_pow(float(a), float(b)))


def test_gauss_filter_table_is_the_references(ref, oracle):
    t, _, _ = oracle.filter_table(1.5)
    for y in range(16):
        for x in range(16):
            e = ref.ref_gauss(np.float32((x + 0.5) * 0.0625), np.float32((y + 0.5) * 0.0625))
            assert bits(t[y, x]) == bits(e), (x, y)


def test_cosine_hemisphere_is_the_references(ref, P_mat):
    """The shinydiffuse sampler's direction is cosHemisphere(N, createCoordsSystem(N), s1, s2)."""
    rng = np.random.default_rng(4)
    m = P_mat.Material(type="shinydiffuse", color=(0.5, 0.5, 0.5))
    for _ in range(3000):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        if rng.random() < 0.1:
            n = np.float32([0, 0, rng.choice([-1, 1])])
        wo = (n + rng.normal(scale=0.3, size=3)).astype(np.float32)
        if float(np.dot(wo, n)) <= 0:
            continue
        s1, s2 = (float(np.float32(x)) for x in rng.random(2))
        out = (C.c_float * 3)()
        ref.ref_cos_hemisphere(f3(n), s1, s2, out)
        got = m.sample(n, wo, s1, s2)["wi"]
        assert np.array_equal(got.view(np.uint32), np.float32(out[:]).view(np.uint32)), (n, s1, s2)


@pytest.fixture(scope="module")
def P_mat(oracle):
    return oracle
