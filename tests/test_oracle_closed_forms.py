"""Pins the CPU oracle (oracle/lima_oracle.cpp) with mathematics independent of its code.

The reference has no tests or golden vectors (and cannot be compiled here: its
sources need the CUDA toolkit headers), so each routine is checked against
closed forms, hand-derived known answers, finite differences and invariants,
chosen so that a dropped term, a wrong sign/index or a transposed operand fails.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.oracle import lib

F32 = np.float32


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lj32(p0, p1, sigma, eps):
    out = np.zeros(3, F32)
    lib().oracle_lj_force_f32(fp(np.asarray(p0, F32)), fp(np.asarray(p1, F32)), sigma, eps, fp(out))
    return out


def lj64(p0, p1, sigma, eps):
    out = np.zeros(3)
    lib().oracle_lj_force_f64(dp(np.asarray(p0, float)), dp(np.asarray(p1, float)), sigma, eps, dp(out))
    return out


# ---- calcLJForce (Engine.cu:475-522) ---------------------------------------------------------
def test_lj_zero_at_potential_minimum():
    sigma, eps = 0.3923, 598.6          # VirtualPathMaker.h:31-32
    r = 2 ** (1 / 6) * sigma
    f = lj64([0, 0, 0], [r, 0, 0], sigma, eps)
    assert np.max(np.abs(f)) < 1e-9
    f32 = lj32([0, 0, 0], [r, 0, 0], sigma, eps)
    assert np.max(np.abs(f32)) < 1e-2     # FP32 rounding of r^6 around the zero crossing


def vpm_calc_force(dist, sigma, eps):
    """The FP64 LJ scalar of VirtualPathMaker::calcForce (VirtualPathMaker.cpp:160-170)."""
    fraction = sigma / dist
    f2 = fraction * fraction
    f6 = f2 * f2 * f2
    return -24.0 * (eps / (dist * dist)) * f6 * (1.0 - 2.0 * f6)


@pytest.mark.parametrize("r", [0.28, 0.35, 0.44, 0.8, 1.7, 3.9])
def test_lj_matches_reference_fp64_scalar(r):
    sigma, eps = 0.3923, 598.6
    d = np.array([0.6, -0.48, 0.64]) * r       # |d| = r
    f = lj64([0.1, 0.2, 0.3], np.array([0.1, 0.2, 0.3]) + d, sigma, eps)
    want = -d * vpm_calc_force(r, sigma, eps)
    np.testing.assert_allclose(f, want, rtol=1e-12, atol=0)
    f32 = lj32([0.1, 0.2, 0.3], np.array([0.1, 0.2, 0.3]) + d, sigma, eps)
    # FP32 error is relative to the uncancelled term 24 eps s/r (1 - 2s cancels near the minimum)
    scale = 24 * eps * (sigma / r) ** 6 / r
    np.testing.assert_allclose(f32, want, rtol=4e-6, atol=2e-6 * scale)


def test_lj_is_minus_gradient_of_potential():
    sigma, eps = 0.32, 300.0
    p0 = np.array([1.0, 2.0, 3.0])
    p1 = np.array([1.21, 2.27, 2.88])

    def U(x):
        r = np.linalg.norm(p1 - x)
        s6 = (sigma / r) ** 6
        return 4 * eps * (s6 * s6 - s6)

    h = 1e-7
    grad = np.array([(U(p0 + h * e) - U(p0 - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(lj64(p0, p1, sigma, eps), -grad, rtol=1e-6)


def test_lj_newton_third_law_and_direction():
    sigma, eps = 0.31, 250.0
    a, b = np.array([0.5, 0.5, 0.5]), np.array([0.75, 0.42, 0.61])
    fa, fb = lj64(a, b, sigma, eps), lj64(b, a, sigma, eps)
    np.testing.assert_allclose(fa, -fb, rtol=1e-14)
    # repulsive inside the minimum: force on a points away from b
    assert np.dot(fa, b - a) < 0


# ---- calcPairbondForces (Engine.cu:525-548) -------------------------------------------------------
def bond64(a, b, b0, kb):
    out = np.zeros(6)
    e = C.c_double(0)
    lib().oracle_pairbond_f64(dp(np.asarray(a, float)), dp(np.asarray(b, float)), b0, kb, dp(out), C.byref(e))
    return out[:3], out[3:], e.value


def test_bond_zero_at_rest_length_and_gradient():
    b0, kb = 0.15, 2.5e8
    fa, fb, e = bond64([0, 0, 0], [0.15, 0, 0], b0, kb)
    assert np.max(np.abs(fa)) < 1e-6 and e < 1e-20
    a = np.array([0.2, 0.3, 0.1])
    b = np.array([0.31, 0.37, 0.02])
    fa, fb, e = bond64(a, b, b0, kb)

    def U(x):
        return 0.5 * kb * (np.linalg.norm(x - b) - b0) ** 2

    h = 1e-8
    grad = np.array([(U(a + h * v) - U(a - h * v)) / (2 * h) for v in np.eye(3)])
    np.testing.assert_allclose(fa, -grad, rtol=1e-5)
    np.testing.assert_allclose(fb, -fa, rtol=1e-15)
    assert abs(e - U(a)) / U(a) < 1e-12


# ---- calcAnglebondForces (Engine.cu:552-578) -------------------------------------------------------
def angle64(l, m, r, t0, k):
    out = np.zeros(9)
    e = C.c_double(0)
    lib().oracle_anglebond_f64(dp(np.asarray(l, float)), dp(np.asarray(m, float)), dp(np.asarray(r, float)),
                               t0, k, dp(out), C.byref(e))
    return out.reshape(3, 3), e.value


def test_angle_known_answer_right_angle():
    """L=(1,0,0), M=0, R=(0,1,0), theta0=60 deg: err = pi/6, F_L = k*err*(0,1,0),
    F_R = k*err*(1,0,0) (each pulls its arm toward the other), F_M = -(F_L+F_R),
    energy = 0.25 k err^2 (the reference's extra 0.5, Engine.cu:568)."""
    k = 4e5
    f, e = angle64([1, 0, 0], [0, 0, 0], [0, 1, 0], math.pi / 3, k)
    err = math.pi / 2 - math.pi / 3
    np.testing.assert_allclose(f[0], [0, k * err, 0], atol=1e-9 * k)
    np.testing.assert_allclose(f[2], [k * err, 0, 0], atol=1e-9 * k)
    np.testing.assert_allclose(f[1], [-k * err, -k * err, 0], atol=1e-9 * k)
    assert abs(e - 0.25 * k * err * err) < 1e-9 * k


def test_angle_zero_at_rest_and_invariants():
    k, t0 = 4e5, math.radians(109.5)
    m = np.array([1.0, 1.0, 1.0])
    l = m + 0.15 * np.array([1.0, 0, 0])
    r = m + 0.15 * np.array([math.cos(t0), math.sin(t0), 0])
    f, e = angle64(l, m, r, t0, k)
    assert np.max(np.abs(f)) < 1e-6 * k and e < 1e-20
    rng = np.random.default_rng(3)
    for _ in range(5):
        l, m, r = rng.normal(size=(3, 3))
        f, _ = angle64(l, m, r, 1.9, k)
        np.testing.assert_allclose(f.sum(0), 0, atol=1e-9 * np.abs(f).max())
        assert abs(np.dot(f[0], l - m)) < 1e-9 * np.abs(f[0]).max() * np.linalg.norm(l - m)   # perpendicular to arm
        assert abs(np.dot(f[2], r - m)) < 1e-9 * np.abs(f[2]).max() * np.linalg.norm(r - m)


# ---- calcDihedralbondForces (Engine.cu:580-637) ----------------------------------------------------
def dih64(l, lm, rm, r, phi0, k, n):
    out = np.zeros(12)
    e = C.c_double(0)
    lib().oracle_dihedral_f64(*[dp(np.asarray(v, float)) for v in (l, lm, rm, r)], phi0, k, n, dp(out), C.byref(e))
    return out.reshape(4, 3), e.value


@pytest.mark.parametrize("phi", [0.3, 1.1, 2.0, 2.9])
def test_dihedral_known_torsion(phi):
    """Axis along z; L off LM along x, R off RM at azimuth phi: torsion = phi, so
    E = k(1 - cos(n phi - phi0)) and |F_L| = |F_R| = |k sin(n phi - phi0)|,
    F_LM = F_RM = -(F_L + F_R)/2."""
    k, n, phi0 = 1000.0, 3, 0.2
    lm, rm = np.array([0, 0, 0.0]), np.array([0, 0, 1.0])
    l = lm + [1, 0, 0]
    r = rm + [math.cos(phi), math.sin(phi), 0]
    f, e = dih64(l, lm, rm, r, phi0, k, n)
    assert abs(e - k * (1 - math.cos(n * phi - phi0))) < 1e-9 * k
    fs = abs(k * math.sin(n * phi - phi0))
    assert abs(np.linalg.norm(f[0]) - fs) < 1e-9 * k
    assert abs(np.linalg.norm(f[3]) - fs) < 1e-9 * k
    np.testing.assert_allclose(f[1], -(f[0] + f[3]) / 2, atol=1e-12 * k)
    np.testing.assert_allclose(f[2], f[1], atol=0)
    # the normals: F_L is perpendicular to the L-LM-RM plane, F_R to the LM-RM-R plane
    assert abs(np.dot(f[0], [0, 1, 0])) == pytest.approx(fs, abs=1e-9 * k)
    assert abs(np.dot(f[0], l - lm)) < 1e-9 * k


def test_dihedral_zero_force_at_minimum():
    k, n = 1000.0, 3
    phi = 2 * math.pi / 3
    lm, rm = np.array([0, 0, 0.0]), np.array([0, 0, 1.0])
    f, e = dih64(lm + [1, 0, 0], lm, rm, rm + [math.cos(phi), math.sin(phi), 0], 0.0, k, n)
    assert np.max(np.abs(f)) < 1e-9 * k and e < 1e-9 * k


# ---- integratePosition / applyHyperpos / applyPBC ---------------------------------------------------
def integrate32(pos, prev, force, mass, dt, lam, L):
    p, q = np.asarray(pos, F32).copy(), np.asarray(prev, F32).copy()
    lib().oracle_integrate_f32(fp(p), fp(q), fp(np.asarray(force, F32)), mass, dt, lam, L)
    return p, q


def test_integrator_free_flight_and_rescale():
    x, xm = np.array([1.5, 2.25, 3.0], F32), np.array([1.499, 2.2505, 3.0001], F32)
    p, q = integrate32(x, xm, [0, 0, 0], 0.018, 2e-6, 1.0, 10.0)
    np.testing.assert_array_equal(q, x)                       # prev <- old position
    np.testing.assert_allclose(p - x, x - xm, atol=2e-6)      # constant velocity
    p2, _ = integrate32(x, xm, [0, 0, 0], 0.018, 2e-6, 0.5, 10.0)
    np.testing.assert_allclose(p2 - x, 0.5 * (x - xm), atol=2e-6)   # lambda scales the displacement
    f = np.array([3e5, -1e5, 2e4], F32)
    p3, _ = integrate32(x, xm, f, 0.018, 2e-6, 1.0, 10.0)
    want = 2 * x.astype(float) - xm + f.astype(float) * (2e-6 / 0.018 * 2e-6)
    np.testing.assert_allclose(p3, want, atol=2e-6)


def test_integrator_across_the_boundary():
    """prev stored one box length away is re-imaged before the Verlet update."""
    x, xm = np.array([0.001, 5.0, 5.0], F32), np.array([9.9995, 5.0, 5.0], F32)
    p, _ = integrate32(x, xm, [0, 0, 0], 0.018, 2e-6, 1.0, 10.0)
    assert abs((p[0] - x[0]) - (x[0] - (xm[0] - 10.0))) < 2e-6


def test_hyperpos_is_minimum_image_and_pbc_wraps():
    L = np.float32(7.0)
    rng = np.random.default_rng(0)
    for _ in range(300):
        s = rng.uniform(0, L, 3).astype(F32)
        m = rng.uniform(-L * 0.4, L * 1.4, 3).astype(F32)
        mm = m.copy()
        lib().oracle_hyperpos_f32(fp(s), fp(mm), L)
        cand = np.stack([m - L, m, m + L])
        best = cand[np.argmin(np.abs(cand - s), axis=0), [0, 1, 2]]
        ok = np.abs(np.abs(s - m) % L - L / 2) > 1e-4     # away from the ambiguous half-box point
        np.testing.assert_allclose(mm[ok], best[ok], atol=1e-5)
        p = rng.uniform(-L * 0.9, L * 1.9, 3).astype(F32)
        lib().oracle_pbc_f32(fp(p), L)
        assert np.all(p >= 0) and np.all(p <= L)


def test_get_angle_clamps():
    v = np.array([1.0, 0, 0], F32)
    assert lib().oracle_get_angle_f32(fp(v), fp(v * F32(3))) == 0.0
    assert abs(lib().oracle_get_angle_f32(fp(v), fp(-v)) - math.pi) < 1e-6
