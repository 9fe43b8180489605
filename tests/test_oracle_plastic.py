"""Pins the CPU restatement of J2 elasto-plasticity with linear isotropic hardening
(oracle.plastic_refresh, the law of GMatElastoPlastic.Cartesian3d.LinearHardening2d that
docs/examples/statics/periodic_elastoplastic.py:71-104 calls; python-gmatelastoplastic >= 0.3.1,
environment.yaml:16).  GMat is not part of the reference and not installed, so no reference output
exists: parity with GMat's own code is unpinned, and these tests pin the restatement by the
identities the law fixes, each chosen so that a plausible slip (a missing 3/2, a wrong sign, the
current instead of the committed plastic state, a wrong tangent term) fails one of them."""

import numpy as np

import oracle

K, G, SY, H = 10.0, 1.0, 0.05, 0.4


def sym(rng, n, scale=1e-1):
    a = rng.standard_normal((n, 3, 3)) * scale
    return 0.5 * (a + a.transpose(0, 2, 1))


def dev(A):
    return A - np.trace(A, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3.0


def sigeq(S):
    d = dev(S)
    return np.sqrt(1.5 * np.einsum("...ij,...ij->...", d, d))


def elastic(E, k=K, g=G):
    return k * np.trace(E, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) + 2.0 * g * dev(E)


def test_elastic_limit():
    # below the yield stress the law is the elastic law, and its tangent the elastic tangent (bit-exact)
    rng = np.random.default_rng(0)
    E = sym(rng, 50, 1e-3)
    Z = np.zeros((50, 3, 3))
    S, C, Ep, ep = oracle.plastic_refresh(E, K, G, 1e3, H, Z, np.zeros(50))
    assert np.allclose(S, elastic(E), rtol=1e-14, atol=1e-17)
    assert np.array_equal(C, oracle.elastic_tangent(np.full(50, K), np.full(50, G)))
    assert np.all(Ep == 0.0) and np.all(ep == 0.0)


def test_return_map_identities():
    rng = np.random.default_rng(1)
    E = sym(rng, 200)
    Z = np.zeros((200, 3, 3))
    S, C, Ep, ep = oracle.plastic_refresh(E, K, G, SY, H, Z, np.zeros(200))
    plastic = ep > 0
    assert plastic.sum() > 150
    # consistency: the returned stress lies on the hardened yield surface
    assert np.allclose(sigeq(S)[plastic], SY + H * ep[plastic], rtol=1e-12)
    # incompressible plastic flow, symmetric plastic strain
    assert np.allclose(np.trace(Ep, axis1=1, axis2=2), 0.0, atol=1e-15)
    assert np.allclose(Ep, Ep.transpose(0, 2, 1), atol=1e-17)
    # the stress is the elastic law of the elastic strain Eps - Epsp (pins the flow factor 3/2 dg/q
    # against the stress scale 1 - 3G dg/q)
    assert np.allclose(S, elastic(E - Ep), rtol=1e-12, atol=1e-15)
    # equivalent plastic strain increment = sqrt(2/3 Epsp:Epsp) from a virgin state
    assert np.allclose(ep, np.sqrt(2.0 / 3.0 * np.einsum("pij,pij->p", Ep, Ep)), rtol=1e-12)
    # the hydrostatic stress is elastic
    assert np.allclose(np.trace(S, axis1=1, axis2=2), 3.0 * K * np.trace(E, axis1=1, axis2=2), rtol=1e-13)


def test_proportional_loading_is_path_independent():
    # radial return is exact for proportional loading with linear hardening: one increment and 40
    # increments to the same strain give the same stress and plastic strain.  Using the current
    # instead of the committed plastic state, or dropping the hardening term, breaks this.
    rng = np.random.default_rng(2)
    E = sym(rng, 20)
    one = oracle.LinearHardening(np.full(20, K), np.full(20, G), np.full(20, SY), np.full(20, H))
    one.Eps = E
    one.refresh()
    many = oracle.LinearHardening(np.full(20, K), np.full(20, G), np.full(20, SY), np.full(20, H))
    for t in np.linspace(0.0, 1.0, 41)[1:]:
        many.increment()
        many.Eps = t * E
        many.refresh()
    assert np.allclose(many.Sig, one.Sig, rtol=1e-10, atol=1e-14)
    assert np.allclose(many.epsp, one.epsp, rtol=1e-10)
    assert np.allclose(many.Epsp, one.Epsp, rtol=1e-10, atol=1e-14)


def test_tangent_is_the_derivative_of_the_stress():
    # consistent tangent (de Souza Neto et al. eq. 7.120) = d Sig / d Eps at a fixed committed state,
    # by central differences, on plastic and elastic points and from a hardened state
    rng = np.random.default_rng(3)
    n = 30
    E = sym(rng, n)
    E[::3] *= 0.01  # some points stay elastic
    Ept = dev(sym(rng, n, 1e-2))
    Ept[::3] *= 0.01
    ept = rng.uniform(0.0, 0.05, n)
    S, C, _, ep = oracle.plastic_refresh(E, K, G, SY, H, Ept, ept)
    assert 5 < np.sum(ep > ept) < n
    h = 1e-7
    for a in range(3):
        for b in range(a, 3):
            d = np.zeros((3, 3))
            d[a, b] = d[b, a] = 1.0
            Sp = oracle.plastic_refresh(E + h * d, K, G, SY, H, Ept, ept, False)[0]
            Sm = oracle.plastic_refresh(E - h * d, K, G, SY, H, Ept, ept, False)[0]
            fd = (Sp - Sm) / (2 * h)
            an = np.einsum("pijkl,kl->pij", C, d)
            assert np.allclose(an, fd, rtol=1e-5, atol=1e-6 * np.abs(C).max())
    # major and minor symmetries
    assert np.allclose(C, C.transpose(0, 3, 4, 1, 2), atol=1e-14)
    assert np.allclose(C, C.transpose(0, 2, 1, 3, 4), atol=1e-14)


def test_elastic_unloading():
    # after plastic loading, a strain decrement answers with the elastic law and leaves the plastic
    # state alone
    rng = np.random.default_rng(4)
    E = sym(rng, 10)
    mat = oracle.LinearHardening(np.full(10, K), np.full(10, G), np.full(10, SY), np.full(10, H))
    mat.Eps = E
    mat.refresh()
    mat.increment()
    S1, ep1 = mat.Sig.copy(), mat.epsp.copy()
    mat.Eps = 0.999 * E
    mat.refresh()
    assert np.array_equal(mat.epsp, ep1)
    assert np.allclose(mat.Sig - S1, elastic(-0.001 * E), rtol=1e-9, atol=1e-15)


def test_simple_shear_curve():
    # monotonic simple shear eps_xy = gamma / 2: tau = G gamma up to tau_y = sigy0 / sqrt(3), then the
    # closed-form hardening branch tau = (sigy0 + H epsp) / sqrt(3) with
    # epsp = (sqrt(3) G gamma - sigy0) / (3G + H)
    gam = np.linspace(0.0, 0.3, 61)
    E = np.zeros((gam.size, 3, 3))
    E[:, 0, 1] = E[:, 1, 0] = gam / 2
    mat = oracle.LinearHardening(np.full(gam.size, K), np.full(gam.size, G), np.full(gam.size, SY), np.full(gam.size, H))
    mat.Eps = E
    mat.refresh()
    epsp = np.maximum(0.0, (np.sqrt(3.0) * G * gam - SY) / (3 * G + H))
    tau = np.where(epsp > 0, (SY + H * epsp) / np.sqrt(3.0), G * gam)
    assert np.allclose(mat.epsp, epsp, rtol=1e-12, atol=1e-16)
    assert np.allclose(mat.Sig[:, 0, 1], tau, rtol=1e-12, atol=1e-16)
