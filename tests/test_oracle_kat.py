"""Pins the oracle's primitive functions against published vectors, known answers and closed forms
(CPU only). Each check is chosen so that a plausible slip (wrong increment, byte order, dropped
term, swapped operand) fails it."""
import math

import numpy as np
import pytest

from sol_amd import _ffi as F
from tests import oracle_lib as O

M64 = (1 << 64) - 1


def pcg32_reference(state, inc, n):
    """pcg32_random_r of the PCG paper's minimal C implementation (pcg_basic.c)."""
    out = []
    for _ in range(n):
        old = state
        state = (old * 6364136223846793005 + inc) & M64
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        out.append(((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF)
    return out, state


def pcg32_srandom(initstate, initseq):
    inc = ((initseq << 1) | 1) & M64
    _, state = pcg32_reference(0, inc, 1)
    state = (state + initstate) & M64
    _, state = pcg32_reference(state, inc, 1)
    return state, inc


def test_pcg_reference_matches_published_demo_vector():
    # pcg32-demo output for pcg32_srandom(42, 54), printed in the PCG distribution.
    state, inc = pcg32_srandom(42, 54)
    got, _ = pcg32_reference(state, inc, 6)
    assert got == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


@pytest.mark.parametrize("seed", [0, 1, 42, 0xDEADBEEF, (1 << 32) - 1])
def test_oracle_pcg_is_pcg32_with_increment_one(seed):
    # samplers.h:26-35: state = 0; next(); state += seed; next() with inc = 1.
    state, inc = pcg32_srandom(seed, 0)
    assert inc == 1
    want, _ = pcg32_reference(state, inc, 64)
    got, _ = O.pcg(seed, 64)
    assert got.tolist() == want


def test_pcg_known_answers():
    u0, f0 = O.pcg(0, 4)
    assert u0.tolist() == [0xe4c14788, 0x379c6516, 0x5c4ab3bb, 0x601d23e0]
    np.testing.assert_array_equal(f0, np.float32([0.8935742378234863, 0.21723014116287231, 0.3605148494243622,
                                                  0.3754446506500244]))
    u1, f1 = O.pcg(1, 4)
    assert u1.tolist() == [0xe2393051, 0x01112f35, 0xd3509d35, 0x0b932f4a]
    np.testing.assert_array_equal(f1, np.float32([0.8836851119995117, 0.004168462939560413, 0.8254488110542297,
                                                  0.045214612036943436]))


def test_float_conversion_is_generate_canonical():
    # generate_canonical<float,24> over one 32-bit draw: fl(x) * 2^-32, clamped below 1.
    u, f = O.pcg(12345, 4096)
    want = (u.astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float32)
    want = np.minimum(want, np.nextafter(np.float32(1), np.float32(0)))
    np.testing.assert_array_equal(f, want)
    assert np.all(f < 1.0) and np.all(f >= 0.0)
    # Edge: x = 0xFFFFFFFF rounds to 2^32 in fp32 -> clamped to the largest float below 1.
    assert np.float32(np.uint32(0xFFFFFFFF)) * np.float32(2.0 ** -32) == np.float32(1.0)


def fnv1a32(data):
    h = 0x811C9DC5
    for b in data:
        h ^= b
        h = (h * 0x01000193) & 0xFFFFFFFF
    return h


def test_fnv1a_published_vectors():
    assert fnv1a32(b"") == 0x811C9DC5
    assert fnv1a32(b"a") == 0xE40C292C
    assert fnv1a32(b"foobar") == 0xBF9CF968


@pytest.mark.parametrize("frame,x,y", [(0, 0, 0), (3, 17, 5), (1 << 40, 1919, 1079), (63, 255, 0)])
def test_pixel_seed_is_fnv1a_over_little_endian_size_t(frame, x, y):
    data = x.to_bytes(8, "little") + y.to_bytes(8, "little") + frame.to_bytes(8, "little")
    assert O.lib.oracle_pixel_seed(frame, x, y) == fnv1a32(data)


def _cos_hemi(u, v):
    d, p = np.zeros(3, np.float32), np.zeros(1, np.float32)
    O.lib.oracle_sample_cosine_hemisphere(u, v, d.ctypes.data, p.ctypes.data)
    return d, float(p[0])


def _pow_hemi(n, u, v):
    d, p = np.zeros(3, np.float32), np.zeros(1, np.float32)
    O.lib.oracle_sample_cosine_power_hemisphere(n, u, v, d.ctypes.data, p.ctypes.data)
    return d, float(p[0])


def test_cosine_hemisphere_closed_form():
    rng = np.random.default_rng(1)
    zs = []
    for u, v in rng.random((4000, 2)):
        d, pdf = _cos_hemi(u, v)
        assert abs(np.linalg.norm(d) - 1) < 1e-6 and d[2] > 0
        assert abs(pdf - d[2] / math.pi) < 1e-7
        assert abs(math.atan2(d[1], d[0]) % (2 * math.pi) - 2 * math.pi * np.float32(u)) < 1e-4 or u < 1e-6
        zs.append(d[2])
    # cos^2(theta) of a cosine-weighted direction is uniform on [0, 1].
    z2 = np.sort(np.square(zs))
    ks = np.max(np.abs(z2 - (np.arange(1, len(z2) + 1) / len(z2))))
    assert ks < 0.03


@pytest.mark.parametrize("n", [1.0, 10.0, 250.0])
def test_power_cosine_closed_form(n):
    rng = np.random.default_rng(2)
    for u, v in rng.random((500, 2)):
        d, pdf = _pow_hemi(n, u, v)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        c = max(d[2], 0.0)
        assert abs(c - v ** (1 / (n + 1))) < 1e-5
        want = (n + 1) / (2 * math.pi) * c ** n
        assert abs(pdf - want) <= 1e-5 * max(want, 1)
    # The pdf integrates to one over the hemisphere: int (n+1) c^n dc over [0,1].
    c = np.linspace(0, 1, 200001)
    assert abs(np.trapezoid((n + 1) * c ** n, c) - 1) < 1e-3


@pytest.mark.parametrize("n", [(0, 0, 1), (0, 0, -1), (1, 0, 0), (0.6, -0.8, 0), (0.267, 0.534, -0.802)])
def test_ortho_basis_is_orthonormal_right_handed(n):
    v = np.array(n, np.float32)
    v /= np.linalg.norm(v)
    t, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
    O.lib.oracle_ortho_basis(v.ctypes.data, t.ctypes.data, b.ctypes.data)
    m = np.stack([t, b, v], 1).astype(np.float64)
    np.testing.assert_allclose(m.T @ m, np.eye(3), atol=2e-6)
    np.testing.assert_allclose(np.cross(t, b), v, atol=2e-6)


def test_ray_triangle_hand_computed():
    hit, t, u, v = O.ray_triangle((0.25, 0.25, 1), (0, 0, -1), (0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert hit and t == 1.0 and u == 0.25 and v == 0.25
    hit, t, u, v = O.ray_triangle((0.5, 0.25, 3), (0, 0, -2), (0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert hit and t == 1.5 and u == 0.5 and v == 0.25
    # outside the triangle
    assert not O.ray_triangle((0.75, 0.75, 1), (0, 0, -1), (0, 0, 0), (1, 0, 0), (0, 1, 0))[0]
    # both sides are hit (no culling); the back-side t is negative for a ray going away
    hit, t, _, _ = O.ray_triangle((0.25, 0.25, -1), (0, 0, -1), (0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert hit and t == -1.0
    # a vertex hit has barycentrics (1, 0)
    hit, t, u, v = O.ray_triangle((1, 0, 1), (0, 0, -1), (0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert hit and t == 1.0 and u == 1.0 and v == 0.0
    # parallel ray
    assert not O.ray_triangle((0.25, 0.25, 1), (1, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0))[0]


def test_ray_triangle_watertight_on_shared_edges():
    # Two triangles sharing the diagonal of a unit quad; rays aimed exactly at the diagonal must
    # hit at least one of them (a non-watertight test lets some through).
    rng = np.random.default_rng(3)
    a, b, c, d = (0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)
    misses = 0
    for s in rng.random(2000).astype(np.float32):
        org = (float(s) + 0.3, float(s) - 0.2, 2.0)
        target = np.float32([s, s, 0])
        dr = tuple((target - np.float32(org)).tolist())
        h1 = O.ray_triangle(org, dr, a, b, c)[0]
        h2 = O.ray_triangle(org, dr, a, c, d)[0]
        misses += not (h1 or h2)
    assert misses == 0


def test_fresnel_closed_forms():
    eta = np.float32(1 / 1.5)
    assert abs(O.lib.oracle_fresnel(eta, 1.0, 1.0) - 0.04) < 1e-6
    # grazing incidence reflects everything
    cos_t = math.sqrt(1 - eta * eta)
    assert abs(O.lib.oracle_fresnel(eta, 0.0, cos_t) - 1.0) < 1e-6


def _mat(kind, **kw):
    m = F.Material()
    m.type = kind
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    m.group = -1
    return m


def test_glass_total_internal_reflection_draws_nothing():
    m = _mat(F.BSDF_GLASS, ks=(1, 1, 1), kt=(1, 1, 1), eta=1 / 1.5)
    # back side (inside the glass): eta becomes 1.5; a grazing exit direction is totally reflected
    out = np.float32([0.9, 0.0, math.sqrt(1 - 0.81)])
    ok, d, pdf, cos, rgb = O.bsdf_sample(m, (0, 0, 1), out, [], front=False)
    assert ok and pdf == 1.0 and cos == 1.0
    np.testing.assert_allclose(d, [-0.9, 0.0, out[2]], atol=1e-6)   # mirror direction


def test_glass_refraction_obeys_snell():
    m = _mat(F.BSDF_GLASS, ks=(1, 1, 1), kt=(0.5, 0.6, 0.7), eta=1 / 1.5)
    out = np.float32([0.6, 0.0, 0.8])
    ok, d, pdf, cos, rgb = O.bsdf_sample(m, (0, 0, 1), out, [0.99], front=True)
    assert ok and d[2] < 0
    np.testing.assert_allclose(rgb, [0.5, 0.6, 0.7])
    sin_i, sin_t = 0.6, math.hypot(d[0], d[1]) / np.linalg.norm(d)
    assert abs(sin_i - 1.5 * sin_t) < 1e-5   # n1 sin_i = n2 sin_t
    # a small random number selects reflection (probability F)
    ok, d, _, _, rgb = O.bsdf_sample(m, (0, 0, 1), out, [0.0], front=True)
    assert ok and np.allclose(d, [-0.6, 0, 0.8], atol=1e-6) and np.allclose(rgb, 1)


def test_diffuse_sample_matches_pdf_and_eval():
    m = _mat(F.BSDF_DIFFUSE, kd=(0.2, 0.4, 0.8))
    n = np.float32([0.0, 0.6, 0.8])
    ok, d, pdf, cos, rgb = O.bsdf_sample(m, n, (0, 0, 1), [0.3, 0.7])
    assert ok
    assert abs(float(np.dot(d, n)) - cos) < 1e-6 and abs(pdf - cos / math.pi) < 1e-6
    np.testing.assert_allclose(rgb, np.float32([0.2, 0.4, 0.8]) / np.float32(math.pi), rtol=1e-6)


def test_phong_sample_is_energy_normalised():
    ns = 20.0
    m = _mat(F.BSDF_PHONG, ks=(1, 1, 1), ns=ns, ns_sample=ns)
    out = np.float32([0.0, 0.0, 1.0])
    # Along the mirror direction the lobe has its maximum: eval = (n+2)/(2 pi).
    ok, d, pdf, cos, rgb = O.bsdf_sample(m, (0, 0, 1), out, [0.0, 0.999999])
    assert ok and abs(rgb[0] - (ns + 2) / (2 * math.pi) * (d[2] ** ns)) < 1e-3 * rgb[0]
    # The estimator f cos / pdf for normal incidence equals ((n+2)/(n+1)) c for c = cos theta.
    assert abs(rgb[0] * cos / pdf - (ns + 2) / (ns + 1) * d[2]) < 1e-4


# ---- InterpBsdf: mix of a diffuse lobe (weight 1 - k) and a Phong lobe (weight k) --------------

def _interp(k, kd=(0.2, 0.4, 0.8), ks=(0.9, 0.7, 0.5), ns=12.0):
    return _mat(F.BSDF_INTERP, kd=kd, ks=ks, ns=ns, ns_sample=ns, mix=k)


@pytest.mark.parametrize("u,v", [(0.3, 0.7), (0.81, 0.05), (0.5, 0.5)])
def test_interp_with_weight_zero_or_one_is_the_single_lobe(u, v):
    n, out = (0.0, 0.6, 0.8), np.float32([0.0, 0.0, 1.0])
    # weight 0: the pick draw (any value > 0) selects the diffuse lobe, whose sample is unchanged
    ok, d, pdf, cos, rgb = O.bsdf_sample(_interp(0.0), n, out, [0.5, u, v])
    ok1, d1, pdf1, cos1, rgb1 = O.bsdf_sample(_mat(F.BSDF_DIFFUSE, kd=(0.2, 0.4, 0.8)), n, out, [u, v])
    assert ok == ok1 and np.array_equal(d, d1) and pdf == pdf1 and cos == cos1 and np.array_equal(rgb, rgb1)
    # weight 1: every pick draw (< 1) selects the Phong lobe
    ok, d, pdf, cos, rgb = O.bsdf_sample(_interp(1.0), n, out, [0.999, u, v])
    ok1, d1, pdf1, cos1, rgb1 = O.bsdf_sample(_mat(F.BSDF_PHONG, ks=(0.9, 0.7, 0.5), ns=12.0, ns_sample=12.0),
                                              n, out, [u, v])
    assert ok == ok1 and np.array_equal(d, d1) and pdf == pdf1 and cos == cos1 and np.array_equal(rgb, rgb1)


def test_interp_pick_threshold_and_mixture_density():
    """A pick draw at or below k samples the Phong lobe, above k the diffuse lobe; either way the
    density is (1 - k) cos / pi + k p_phong and the value (1 - k) kd / pi + k f_phong."""
    k, ns = 0.3, 12.0
    n, out = np.float32([0.0, 0.0, 1.0]), np.float32([0.6, 0.0, 0.8])
    kd, ks = np.float32([0.2, 0.4, 0.8]), np.float32([0.9, 0.7, 0.5])
    refl = np.float32([-0.6, 0.0, 0.8])   # mirror direction of `out` about n
    for r, lobe in ((0.3, F.BSDF_PHONG), (0.2999, F.BSDF_PHONG), (0.3001, F.BSDF_DIFFUSE), (0.9, F.BSDF_DIFFUSE)):
        ok, d, pdf, cos, rgb = O.bsdf_sample(_interp(k, ns=ns), n, out, [r, 0.4, 0.6])
        single = _mat(lobe, kd=tuple(kd), ks=tuple(ks), ns=ns, ns_sample=ns)
        ok1, d1, _, cos1, _ = O.bsdf_sample(single, n, out, [0.4, 0.6])
        assert ok and ok1 and np.array_equal(d, d1) and cos == cos1, (r, lobe)
        c_n = float(np.dot(d, n))
        c_r = max(float(np.dot(d, refl)), 0.0)
        p_phong = (ns + 1) / (2 * math.pi) * c_r ** ns
        f_phong = ks * (ns + 2) / (2 * math.pi) * c_r ** ns
        assert abs(pdf - ((1 - k) * c_n / math.pi + k * p_phong)) < 1e-5 * max(pdf, 1)
        np.testing.assert_allclose(rgb, (1 - k) * kd / math.pi + k * f_phong, rtol=1e-5, atol=1e-7)
