"""Pins the oracle's Roe eigenvalue schemes (Roe_M1..M8, src/Euler/Gas.hpp:878-1003) and the
HLLE+ variants (HLLEP, HLLEP_V1, Gas.hpp:601-714) against closed forms that do not share the
oracle's wave decomposition (CPU only).

Independent ingredients:
  * the physical flux F(U).n and the textbook eigenvectors of the Euler flux Jacobian at the
    Roe average (acoustic r = (1, v -+ a n, H -+ a v.n), entropy (1, v, |v|^2 / 2), shear
    (0, t, v.t)); that these diagonalise the Roe matrix is itself checked (Roe property
    R Lambda R^-1 dU = F_R - F_L);
  * the plain Roe flux with fixScale = 0, pinned by the reference golden vector
    (test/cpp/Euler/test_RiemannSolvers.cpp:257-267, test_oracle_goldens.py);
  * the HLL flux with Einfeldt speeds, (S+ F_L - S- F_R)/(S+ - S-) + S+ S-/(S+ - S-) dU.

Every scheme is checked (a) where its fix is inactive, against the golden-pinned plain Roe,
and (b) at states where the fix is active, against the eigenvalue worked out by hand for that
state, assembled with the textbook eigenvectors.  A wrong wave (sign, index, a dropped
incFScale) changes F beyond the tolerance in one of them.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po
from helpers import RS

G = 1.4


def p2c(r, v, p):
    v = np.asarray(v, float)
    return np.array([r, *(r * v), p / (G - 1) + 0.5 * r * v @ v])


def prim(U):
    v = U[1:4] / U[0]
    p = (G - 1) * (U[4] - 0.5 * U[0] * v @ v)
    return U[0], v, p


def phys_flux(U, n):
    r, v, p = prim(U)
    vn = v @ n
    return np.concatenate([[r * vn], U[1:4] * vn + p * n, [(U[4] + p) * vn]])


def roe_avg(UL, UR):
    rL, vL, pL = prim(UL)
    rR, vR, pR = prim(UR)
    HL, HR = (UL[4] + pL) / rL, (UR[4] + pR) / rR
    sL, sR = np.sqrt(rL), np.sqrt(rR)
    v = (sL * vL + sR * vR) / (sL + sR)
    H = (sL * HL + sR * HR) / (sL + sR)
    a = np.sqrt((G - 1) * (H - 0.5 * v @ v))
    return v, H, a


def tangents(n):
    t1 = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(n, t1)


def eigvecs(v, H, a, n):
    """Columns: u-a, entropy, shear 1, shear 2, u+a."""
    t1, t2 = tangents(n)
    vn = v @ n
    return np.column_stack([
        np.concatenate([[1.0], v - a * n, [H - a * vn]]),
        np.concatenate([[1.0], v, [0.5 * v @ v]]),
        np.concatenate([[0.0], t1, [v @ t1]]),
        np.concatenate([[0.0], t2, [v @ t2]]),
        np.concatenate([[1.0], v + a * n, [H + a * vn]]),
    ])


def roe_form(UL, UR, n, lam0, lam123, lam4, ULm=None, URm=None):
    """0.5 (F_L + F_R) - 0.5 R diag(lam) R^-1 (U_R - U_L), R at the Roe average of the means."""
    ULm = UL if ULm is None else ULm
    URm = UR if URm is None else URm
    v, H, a = roe_avg(ULm, URm)
    R = eigvecs(v, H, a, n)
    lam = np.array([lam0, lam123, lam123, lam123, lam4])
    return 0.5 * (phys_flux(UL, n) + phys_flux(UR, n)) - 0.5 * R @ (lam * np.linalg.solve(R, UR - UL))


def hll(UL, UR, n):
    """HLL with Einfeldt speeds S- = min(v.n - a, u_L - a_L), S+ = max(v.n + a, u_R + a_R)."""
    v, H, a = roe_avg(UL, UR)
    rL, vL, pL = prim(UL)
    rR, vR, pR = prim(UR)
    Sm = min(v @ n - a, vL @ n - np.sqrt(G * pL / rL), 0.0)
    Sp = max(v @ n + a, vR @ n + np.sqrt(G * pR / rR), 0.0)
    F = (Sp * phys_flux(UL, n) - Sm * phys_flux(UR, n)) / (Sp - Sm) + Sp * Sm / (Sp - Sm) * (UR - UL)
    return F, Sm, Sp, v, H, a


def unit(x):
    x = np.asarray(x, float)
    return x / np.linalg.norm(x)


N3 = unit([0.36, -0.48, 0.8])
STATES = [
    # (UL, UR): subsonic, both directions, with tangential jumps
    (p2c(1.0, [0.3, 0.1, -0.2], 1.0), p2c(0.8, [0.1, -0.2, 0.25], 0.7)),
    (p2c(0.7, [-0.4, 0.3, 0.1], 0.6), p2c(1.3, [0.2, 0.05, -0.3], 1.4)),
    (p2c(1.1, [0.9, -0.5, 0.4], 1.2), p2c(0.9, [1.1, -0.2, 0.1], 0.9)),
]


@pytest.mark.parametrize("k", range(len(STATES)))
def test_textbook_eigenvectors_have_roe_property(k):
    """R diag(v.n - a, v.n, v.n, v.n, v.n + a) R^-1 dU = F_R - F_L at the Roe average: the
    decomposition the closed forms below use is the Roe matrix."""
    UL, UR = STATES[k]
    v, H, a = roe_avg(UL, UR)
    R = eigvecs(v, H, a, N3)
    vn = v @ N3
    lam = np.array([vn - a, vn, vn, vn, vn + a])
    np.testing.assert_allclose(R @ (lam * np.linalg.solve(R, UR - UL)), phys_flux(UR, N3) - phys_flux(UL, N3),
                               rtol=1e-12, atol=1e-13)
    # and the plain Roe flux at fixScale 0 (golden-pinned) is this decomposition with |lambda|
    F, lamo = po.riemann(RS["Roe"], UL, UR, N3, G, fixScale=0.0, incFScale=1.0)
    np.testing.assert_allclose(F, roe_form(UL, UR, N3, abs(vn - a), abs(vn), abs(vn + a)), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(lamo, [abs(vn - a), abs(vn), abs(vn + a)], rtol=1e-13)


# ------------------------------------------------------------- (a) inactive fix == plain Roe
@pytest.mark.parametrize("rs,kw", [
    ("Roe_M3", dict(fixScale=1e-9)),                  # LD: a* = min(|u| / (0.2 fs), a) = a
    ("Roe_M4", dict(fixScale=1e-9)),                  # ID: u* = sign(u) max(0.2 a fs, |u|) = u
    ("Roe_M6", dict(fixScale=1.0, dLambda=0.0)),      # H-correction with dLambda 0
    ("Roe_M8", dict(fixScale=1e-9, dLambda=0.0)),     # H-correction + Harten-Yee, threshold ~0
])
@pytest.mark.parametrize("k", range(len(STATES)))
def test_inactive_fix_is_plain_roe(rs, kw, k):
    UL, UR = STATES[k]
    F, lam = po.riemann(RS[rs], UL, UR, N3, G, incFScale=1.0, **kw)
    Fr, lamr = po.riemann(RS["Roe"], UL, UR, N3, G, fixScale=0.0, incFScale=1.0)
    np.testing.assert_allclose(F, Fr, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(lam, lamr, rtol=1e-13)


@pytest.mark.parametrize("rs", ["Roe_M1", "Roe_M5"])
def test_local_lf_schemes_with_equal_means_are_plain_roe(rs):
    """cLLF (eig 1) and cLLF_M (eig 5) take max over the two means of |u_m -+ a_m|, |u_m|;
    with both sides' mean state equal (and the low-dissipation switch off) that is the plain
    Roe eigenvalue at that state, applied to the point-state jump."""
    Um = p2c(0.9, [0.35, -0.1, 0.2], 0.8)
    for UL, UR in STATES:
        F, lam = po.riemann(RS[rs], UL, UR, N3, G, ULm=Um, URm=Um, fixScale=1e-9, incFScale=1.0)
        Fr, lamr = po.riemann(RS["Roe"], UL, UR, N3, G, ULm=Um, URm=Um, fixScale=0.0, incFScale=1.0)
        np.testing.assert_allclose(F, Fr, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(lam, lamr, rtol=1e-13)


# ------------------------------------------------------------- (b) active fix, hand-derived
def test_h_correction_dominant_is_rusanov():
    """eig 6 and 8 with dLambda * 0.25 * fixScale above every |lambda| (and above the
    Harten-Yee threshold 0.05 fixScale (|v| + a)): all three eigenvalues equal
    c = 0.25 dLambda fixScale, so R diag(c) R^-1 = c I and F = (F_L + F_R)/2 - c (U_R - U_L)/2,
    with no eigenvector in it."""
    for UL, UR in STATES:
        for rs in ("Roe_M6", "Roe_M8"):
            for dLambda, fs in ((40.0, 1.0), (100.0, 0.5)):
                c = 0.25 * dLambda * fs
                F, lam = po.riemann(RS[rs], UL, UR, N3, G, dLambda=dLambda, fixScale=fs, incFScale=1.0)
                ref = 0.5 * (phys_flux(UL, N3) + phys_flux(UR, N3)) - 0.5 * c * (UR - UL)
                np.testing.assert_allclose(F, ref, rtol=1e-13, atol=1e-13)
                np.testing.assert_allclose(lam, [c, c, c], rtol=1e-14)


def test_lax_friedrichs_is_rusanov():
    """eig 2: one eigenvalue max(|u_L| + a_L, |u_R| + a_R) from the means, F Rusanov."""
    for UL, UR in STATES:
        F, lam = po.riemann(RS["Roe_M2"], UL, UR, N3, G)
        rL, vL, pL = prim(UL)
        rR, vR, pR = prim(UR)
        c = max(abs(vL @ N3) + np.sqrt(G * pL / rL), abs(vR @ N3) + np.sqrt(G * pR / rR))
        ref = 0.5 * (phys_flux(UL, N3) + phys_flux(UR, N3)) - 0.5 * c * (UR - UL)
        np.testing.assert_allclose(F, ref, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(lam, [c, c, c], rtol=1e-14)


def _tangential_pair():
    """Both states move tangentially to n: the Roe average has v.n = 0."""
    t1, t2 = tangents(N3)
    UL = p2c(1.0, 0.3 * t1 - 0.2 * t2, 1.0)
    UR = p2c(0.6, -0.1 * t1 + 0.4 * t2, 0.8)
    return UL, UR


def test_low_dissipation_at_zero_normal_velocity_is_central():
    """eig 3 (LD) with v.n = 0: the threshold |v.n| / (0.2 fs) is 0, so a* = 0 and every
    eigenvalue vanishes: F = (F_L + F_R) / 2."""
    UL, UR = _tangential_pair()
    F, lam = po.riemann(RS["Roe_M3"], UL, UR, N3, G, fixScale=1.0, incFScale=1.0)
    np.testing.assert_allclose(F, 0.5 * (phys_flux(UL, N3) + phys_flux(UR, N3)), rtol=1e-13, atol=1e-14)
    assert np.abs(lam).max() < 1e-14


@pytest.mark.parametrize("mach,fs,incF", [(0.05, 1.0, 1.0), (0.1, 1.0, 0.5), (0.02, 2.0, 1.0)])
def test_low_mach_fixes_by_hand(mach, fs, incF):
    """A face-normal Roe velocity u = mach * a (u > 0, |u| < 0.2 a fs):
      eig 3 (LD):  a* = min(|u| / (0.2 fs), a) = 5 u / fs  ->  lam = (|incF u - a*|, |u|, incF u + a*)
                   (lam123 keeps its plain value |u|, Gas.hpp:921-925);
      eig 4 (ID):  u* = max(0.2 a fs, incF |u|) = 0.2 a fs  ->  lam = (|u* - a|, u*, u* + a);
      eig 5 (cLLF_M) with equal means: a_m* = min(a, incF |u| / (0.2 fs)) -> (|incF u - a_m*|,
                   incF u, incF u + a_m*);
      eig 7 (HY on the plain eigenvalues squared, the reference's form, Gas.hpp:968-981):
                   lam_k = |u_k| |u_k| below/above the threshold 0.05 fs (|v| + a)."""
    n = N3
    t1, _ = tangents(n)
    a0 = np.sqrt(G * 1.0 / 1.0)
    # equal point states and means: Roe average = the state, jump in the point states only
    Um = p2c(1.0, mach * a0 * n + 0.1 * t1, 1.0)
    v, H, a = roe_avg(Um, Um)
    u = v @ n
    assert u > 0 and abs(u) < 0.2 * a * fs
    UL = p2c(1.02, mach * a0 * n + 0.12 * t1, 1.03)
    UR = p2c(0.97, mach * a0 * n + 0.05 * t1, 0.98)
    kw = dict(ULm=Um, URm=Um, fixScale=fs, incFScale=incF)

    aStar = 5 * u / fs
    F, lam = po.riemann(RS["Roe_M3"], UL, UR, n, G, **kw)
    l3 = (abs(incF * u - aStar), u, abs(incF * u + aStar))
    np.testing.assert_allclose(lam, l3, rtol=1e-13)
    np.testing.assert_allclose(F, roe_form(UL, UR, n, *l3, Um, Um), rtol=1e-12, atol=1e-13)

    uStar = 0.2 * a * fs
    F, lam = po.riemann(RS["Roe_M4"], UL, UR, n, G, **kw)
    l4 = (abs(uStar - a), uStar, uStar + a)
    np.testing.assert_allclose(lam, l4, rtol=1e-13)
    np.testing.assert_allclose(F, roe_form(UL, UR, n, *l4, Um, Um), rtol=1e-12, atol=1e-13)

    am = min(a, incF * u / (0.2 * fs))
    F, lam = po.riemann(RS["Roe_M5"], UL, UR, n, G, **kw)
    l5 = (abs(incF * u - am), incF * u, incF * u + am)
    np.testing.assert_allclose(lam, l5, rtol=1e-13)
    np.testing.assert_allclose(F, roe_form(UL, UR, n, *l5, Um, Um), rtol=1e-12, atol=1e-13)

    thr = 0.05 * fs * (np.linalg.norm(v) + a)

    def hy(x):
        return (x * x + thr * thr) / (2 * thr) if x < thr else x
    F, lam = po.riemann(RS["Roe_M7"], UL, UR, n, G, **kw)
    l7 = (hy(abs(u - a) * abs(incF * u - a)), hy(u * abs(incF * u)), hy((u + a) * (incF * u + a)))
    assert l7[1] != u * abs(incF * u)       # the threshold is active on the contact wave
    np.testing.assert_allclose(lam, l7, rtol=1e-13)
    np.testing.assert_allclose(F, roe_form(UL, UR, n, *l7, Um, Um), rtol=1e-12, atol=1e-13)


def test_id_fix_at_exactly_zero_normal_velocity_is_plain():
    """eig 4 multiplies by signTol(u, a 1e-10), which is 0 at u = 0: u* = 0 and the
    eigenvalues stay (a, 0, a)."""
    UL, UR = _tangential_pair()
    v, H, a = roe_avg(UL, UR)
    F, lam = po.riemann(RS["Roe_M4"], UL, UR, N3, G, fixScale=1.0, incFScale=1.0)
    np.testing.assert_allclose(lam, [a, 0.0, a], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(F, roe_form(UL, UR, N3, a, 0.0, a), rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------- HLLE+ variants
def _contact_shear_parts(UL, UR, v, H, a):
    R = eigvecs(v, H, a, N3)
    al = np.linalg.solve(R, UR - UL)
    return R[:, 1] * al[1], R[:, 1:4] @ al[1:4]


@pytest.mark.parametrize("dLambda,fs", [(0.0, 1.0), (0.3, 1.0), (2.0, 0.5)])
@pytest.mark.parametrize("k", range(len(STATES)))
def test_hllep_is_hll_minus_contact_antidiffusion(dLambda, fs, k):
    """HLLEP (type 0): F = F_HLL - S+ S- / (S+ - S-) * delta * (entropy-wave part of dU), and
    HLLEP_V1: the same with the entropy and both shear waves, where
    delta = a / (a + max(|v.n|, dLambda fs)) at the Roe average.  V1's eigenvalues are the HLL
    ones, |((S+ + S-) lambda - 2 (1 - delta_k) S+ S-) / (S+ - S-)| with delta_k = 0 on the
    acoustic waves."""
    UL, UR = STATES[k]
    Fh, Sm, Sp, v, H, a = hll(UL, UR, N3)
    vn = v @ N3
    delta = a / (a + max(abs(vn), dLambda * fs))
    ent, cs = _contact_shear_parts(UL, UR, v, H, a)
    F0, _ = po.riemann(RS["HLLEP"], UL, UR, N3, G, dLambda=dLambda, fixScale=fs)
    np.testing.assert_allclose(F0, Fh - Sp * Sm / (Sp - Sm) * delta * ent, rtol=1e-12, atol=1e-13)
    F1, lam1 = po.riemann(RS["HLLEP_V1"], UL, UR, N3, G, dLambda=dLambda, fixScale=fs)
    np.testing.assert_allclose(F1, Fh - Sp * Sm / (Sp - Sm) * delta * cs, rtol=1e-12, atol=1e-13)

    def lhll(lk, dk):
        return abs(((Sp + Sm) * lk - 2 * (1 - dk) * Sp * Sm) / (Sp - Sm))
    np.testing.assert_allclose(lam1, [lhll(vn - a, 0), lhll(vn, delta), lhll(vn + a, 0)], rtol=1e-13)


def test_hllep_v1_large_dlambda_is_hll():
    """delta -> 0 as dLambda fs grows: both HLLE+ variants become the plain HLL flux."""
    for UL, UR in STATES:
        Fh = hll(UL, UR, N3)[0]
        for rs in ("HLLEP", "HLLEP_V1"):
            F, _ = po.riemann(RS[rs], UL, UR, N3, G, dLambda=1e13, fixScale=1.0)
            np.testing.assert_allclose(F, Fh, rtol=1e-11, atol=1e-12)
