"""The certified-elastic test of IPF_RESIDUAL_CERTIFIED_ELASTIC against the
oracle's return map (the reference's algorithm, FunctionsSigmaProjection.h:
337-379 after FunctionsSpectralDecomp.h:261-268).

Bars, over random, near-yield and near-degenerate trial states of several
materials, with and without committed plastic strain:
  * soundness: every certified point is elastic in the oracle (mtype 0,
    branch 0);
  * its stress: the trial stress it returns equals the oracle's
    eigen-reconstructed stress to 2e-8 of the point's largest stress entry
    M (the reference's Newton eigenvalues, tolerance 1e-11 on the normalised
    cubic, and the reconstruction's rounding: measured <= 1e-8 at gaps of
    1e-2 M, ~1e-13 M typical; at coincident eigenvalues the reference's
    stress is off by up to M / 2, so those points are not certified);
  * margin: no point whose closed-form main-plane value lies above -1.9e-3
    M, or whose principal stresses come closer than 0.99e-2 M, is certified;
  * not vacuous: most elastic points of a random sample are certified.
"""
import math

import numpy as np
import pytest

import intpointsfem_amd as ipf
from oracle import oracle as O

MATERIALS = [(5000.0, 0.2, 10.0, 30.0, 30.0), (2000.0, 0.15, 1.0, 20.0, 20.0), (20000.0, 0.3, 20.0, 40.0, 40.0),
             (8000.0, 0.25, 3.0, 35.0, 10.0)]


def lame(E, nu):
    return nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def trial(e3, ep4, lam, mu):
    exx, exy, eyy, ezz = e3[:, 0] - ep4[:, 0], e3[:, 1] - ep4[:, 1], e3[:, 2] - ep4[:, 2], -ep4[:, 3]
    sxx = (lam + 2 * mu) * exx + lam * eyy + lam * ezz
    syy = lam * exx + (lam + 2 * mu) * eyy + lam * ezz
    szz = lam * exx + lam * eyy + (lam + 2 * mu) * ezz
    sxy = mu * exy
    return sxx, syy, szz, sxy


def plane(e3, ep4, lam, mu, phi, c):
    """closed-form main-plane value and largest stress entry (input generation and the margin check)"""
    sxx, syy, szz, sxy = trial(e3, ep4, lam, mu)
    cc, d = 0.5 * (sxx + syy), 0.5 * (sxx - syy)
    R = np.sqrt(d * d + sxy * sxy)
    s1 = np.maximum(cc + R, szz)
    s3 = np.minimum(cc - R, szz)
    M = np.max(np.abs(np.stack([sxx, syy, szz, sxy])), axis=0)
    return (s1 - s3) + (s1 + s3) * math.sin(phi) - 2 * c * math.cos(phi), M


def samples(rng, lam, mu, phi, c, n=6000):
    """random strains, then the same directions rescaled to sit just outside
    and just inside the margin, then near-degenerate tensors (szz close to an
    in-plane principal stress, sxx ~ syy with tiny shear)"""
    e = rng.standard_normal((n, 3)) * 10 ** rng.uniform(-5, -2, (n, 1))
    ep = np.zeros((n, 4))
    ep[n // 2:] = rng.standard_normal((n - n // 2, 4)) * 1e-3
    ep[n // 2:, 3] = -(ep[n // 2:, 0] + ep[n // 2:, 2])  # plastic flow keeps volume roughly
    out_e, out_ep = [e], [ep]
    # directions rescaled: phi(s e) = s A(e) - 2 c cos(phi) for zero plastic strain
    d = rng.standard_normal((n, 3))
    f1, M1 = plane(d, np.zeros((n, 4)), lam, mu, phi, 0.0)
    ok = f1 > 1e-12
    d, f1, M1 = d[ok], f1[ok], M1[ok]
    for k in (-3e-3, -2.2e-3, -2.05e-3, -1.95e-3, -1e-3, -1e-6, 0.0, 1e-6, 1e-3):
        # s A = 2 c cos(phi) + k s M  ->  s = 2 c cos(phi) / (A - k M)
        s = 2 * c * math.cos(phi) / (f1 - k * M1)
        s = s[s > 0]
        out_e.append(d[:len(s)] * s[:, None])
        out_ep.append(np.zeros((len(s), 4)))
    # near-degenerate: exx ~ eyy, tiny shear, so sxx ~ syy ~ szz up to small splits
    m = n // 2
    base = rng.uniform(-3e-3, 3e-3, (m, 1))
    split = 10 ** rng.uniform(-14, -4, (m, 3)) * rng.choice([-1, 1], (m, 3))
    e_deg = np.concatenate([base + split[:, :1] * base, split[:, 1:2] * base, base + split[:, 2:3] * base], axis=1)
    ep_deg = np.zeros((m, 4))
    ep_deg[:, 3] = -base[:, 0] * (1 + 10 ** rng.uniform(-12, -3, m)) * 2 * lam / (lam + 2 * mu)  # szz ~ sxx
    out_e.append(e_deg)
    out_ep.append(ep_deg)
    return np.concatenate(out_e), np.concatenate(out_ep)


@pytest.mark.parametrize("imat", range(len(MATERIALS)))
def test_certified_points_are_elastic_in_the_oracle(imat):
    E, nu, c, phid, psid = MATERIALS[imat]
    lam, mu = lame(E, nu)
    phi, psi = math.radians(phid), math.radians(psid)
    rng = np.random.default_rng(100 + imat)
    e3, ep4 = samples(rng, lam, mu, phi, c)
    n = len(e3)
    cert, sig = ipf.certify_elastic(e3, ep4, lam, mu, phi, psi, c)
    ep6 = np.zeros((n, 6))
    ep6[:, 0], ep6[:, 1], ep6[:, 3], ep6[:, 5] = ep4[:, 0], ep4[:, 1], ep4[:, 2], ep4[:, 3]
    ws, _, _, mt, br = O.compute_sigma(e3.reshape(-1), np.ones(n), lam, mu, phi, psi, c, plastic_strain6=ep6,
                                       update_mem=True)
    ws = ws.reshape(-1, 3)
    f, M = plane(e3, ep4, lam, mu, phi, c)
    elastic = mt == 0
    print(f"material {imat}: {n} points, {int(elastic.sum())} elastic in the oracle, {int(cert.sum())} certified, "
          f"{int((cert & ~elastic).sum())} certified but plastic")
    assert not (cert & ~elastic).any()
    assert (br[cert] == 0).all()
    dev = np.abs(ws[cert] - sig[cert]).max(axis=1) / np.maximum(M[cert], 1e-300)
    print(f"  certified stress vs oracle: max {dev.max():.2e} of the point's largest stress entry")
    assert dev.max() <= 2e-8
    assert not (cert & (f > -1.9e-3 * M)).any()
    sxx, syy, szz, sxy = trial(e3, ep4, lam, mu)
    cc, d = 0.5 * (sxx + syy), 0.5 * (sxx - syy)
    R = np.sqrt(d * d + sxy * sxy)
    gap = np.minimum(2 * R, np.minimum(np.abs(cc + R - szz), np.abs(cc - R - szz)))
    assert not (cert & (gap < 0.99e-2 * M)).any()
    rand = slice(0, 6000)
    assert cert[rand].sum() >= 0.5 * elastic[rand].sum()


def test_nan_is_not_certified():
    lam, mu = lame(5000.0, 0.2)
    e3 = np.array([[np.nan, 0.0, 0.0], [0.0, np.nan, 0.0], [1e-4, 0.0, np.inf]])
    cert, _ = ipf.certify_elastic(e3, np.zeros((3, 4)), lam, mu, 0.5, 0.5, 10.0)
    assert not cert.any()
