"""The shade kernel's solid-angle sampler of triangle emitters (Arvo 1995),
run on the host through tn_sample_triangle_host.

No reference counterpart (turner has point lights only).  Pinned by what the
geometry fixes: the solid angle against the triple-product closed form (Van
Oosterom and Strackee 1983) and against counting uniform directions on the
sphere; every sample inside the triangle; uniform density checked on the four
midpoint sub-triangles (their solid-angle shares) and against Lambert's
closed form for the cosine-weighted integral over a polygon.
"""
import ctypes as C

import numpy as np
import pytest

from turner_amd._lib import lib, ptr

pytestmark = pytest.mark.filterwarnings("ignore")


def _sample(p, tri, u):
    p = np.ascontiguousarray(p, np.float32)
    tri = np.ascontiguousarray(tri, np.float32).reshape(9)
    u = np.ascontiguousarray(u, np.float32)
    n = len(u)
    d = np.zeros((n, 3), np.float32)
    om, om_vos = C.c_float(), C.c_float()
    rc = lib().tn_sample_triangle_host(ptr(p), ptr(tri), ptr(u), n, ptr(d), C.byref(om), C.byref(om_vos))
    return rc, d, om.value, om_vos.value


def _vos(a, b, c):
    a, b, c = (x / np.linalg.norm(x) for x in (a, b, c))
    return 2 * np.arctan2(abs(np.dot(a, np.cross(b, c))), 1 + a @ b + b @ c + c @ a)


def _inside(o, d, tri, eps=1e-4):
    """Ray o + t d (t > 0) hits the triangle (FP64 Moeller-Trumbore, small margin)."""
    v0, v1, v2 = (np.asarray(v, np.float64) for v in tri)
    e1, e2 = v1 - v0, v2 - v0
    d = d.astype(np.float64)
    pv = np.cross(d, e2)
    det = pv @ e1
    inv = 1.0 / det
    tv = np.asarray(o, np.float64) - v0
    u = (pv @ tv) * inv
    q = np.cross(tv, e1)
    v = (d @ q) * inv
    t = (q @ e2) * inv
    return (u > -eps) & (v > -eps) & (u + v < 1 + eps) & (t > 0)


CASES = [
    # point, triangle: near and far, oblique, large and small solid angles
    ((0.5, 0.9, 0.5), [(0.35, 0.999, 0.35), (0.65, 0.999, 0.35), (0.65, 0.999, 0.65)]),
    ((0.3, 0.2, 0.7), [(0.35, 0.999, 0.35), (0.65, 0.999, 0.65), (0.35, 0.999, 0.65)]),
    ((0.0, 0.0, 0.0), [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)]),
    ((0.1, -0.2, 0.05), [(2.0, 0.3, 0.1), (0.2, 3.0, -0.4), (0.5, 0.2, 2.5)]),
    ((0.0, 0.0, 0.0), [(-0.05, -0.05, 1.0), (0.05, -0.05, 1.0), (0.0, 0.08, 1.0)]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_solid_angle_closed_form(case):
    p, tri = CASES[case]
    rc, _, om, om_vos = _sample(p, tri, np.zeros((0, 2)))
    assert rc == 0
    rel = [np.asarray(v, np.float64) - np.asarray(p) for v in tri]
    exact = _vos(*rel)
    assert abs(om_vos - exact) <= 2e-6 * exact + 1e-7
    assert abs(om - exact) <= 1e-4 * exact + 2e-6


def test_solid_angle_by_counting_directions():
    p, tri = CASES[3]
    rng = np.random.default_rng(4)
    n = 400_000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    frac = _inside(p, d, tri).mean()
    _, _, om, _ = _sample(p, tri, np.zeros((0, 2)))
    se = np.sqrt(frac * (1 - frac) / n) * 4 * np.pi
    assert abs(frac * 4 * np.pi - om) < 5 * se


@pytest.mark.parametrize("case", range(len(CASES)))
def test_samples_inside_and_uniform(case):
    p, tri = CASES[case]
    rng = np.random.default_rng(10 + case)
    n = 120_000
    u = rng.random((n, 2), dtype=np.float32)
    rc, d, om, _ = _sample(p, tri, u)
    assert rc == 0
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5)
    assert _inside(p, d, tri).all()
    # midpoint subdivision: the share of samples in each sub-triangle is its
    # share of the solid angle (chi-square, 3 degrees of freedom)
    v = [np.asarray(x, np.float64) for x in tri]
    m01, m12, m20 = (v[0] + v[1]) / 2, (v[1] + v[2]) / 2, (v[2] + v[0]) / 2
    subs = [(v[0], m01, m20), (m01, v[1], m12), (m20, m12, v[2]), (m01, m12, m20)]
    pc = np.asarray(p, np.float64)
    shares = np.array([_vos(*(x - pc for x in s)) for s in subs]) / om
    counts = np.array([_inside(p, d, s, eps=1e-6).sum() for s in subs])
    assert abs(counts.sum() - n) < 1e-3 * n  # only rounding on the shared edges
    expect = shares * n
    chi2 = (((counts - expect) ** 2) / expect).sum()
    assert chi2 < 25.0, (counts, expect)


def test_cosine_integral_lambert():
    """Mean of n.w over uniform samples times Omega = Lambert's polygon form
    factor integral  1/2 sum_i theta_i n . unit(p_i x p_(i+1))."""
    p, tri = CASES[1]
    rel = [np.asarray(x, np.float64) - np.asarray(p) for x in tri]
    q = [x / np.linalg.norm(x) for x in rel]
    nrm = sum(q)
    nrm /= np.linalg.norm(nrm)
    phi = np.zeros(3)
    for i in range(3):
        a, b = q[i], q[(i + 1) % 3]
        th = np.arccos(np.clip(a @ b, -1, 1))
        c = np.cross(a, b)
        phi += 0.5 * th * c / np.linalg.norm(c)
    exact = abs(nrm @ phi)
    rng = np.random.default_rng(3)
    u = rng.random((200_000, 2), dtype=np.float32)
    _, d, om, _ = _sample(p, tri, u)
    c = d.astype(np.float64) @ nrm
    est = c.mean() * om
    se = c.std() / np.sqrt(len(c)) * om
    assert abs(est - exact) < 5 * se + 1e-5 * exact


def test_degenerate_triangle_is_an_error():
    rc, _, _, _ = _sample((0, 0, 0), [(1, 0, 0), (2, 0, 0), (3, 0, 0)], np.zeros((1, 2)))
    assert rc != 0
