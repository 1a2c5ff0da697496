"""Particle-filter kernels against the oracle and closed forms (needs gfx950).

Resampling with injected u0 must give the oracle's indices exactly and its
weights to 1e-12 relative; the rotation algebra of the perturbation is
checked through rotation matrices (oracle.rotmat, pinned independently);
the random parts are checked statistically against numpy draws of the same
distributions with tolerances stated per test.
"""
import numpy as np
import pytest

import oracle
import thx

pytestmark = pytest.mark.gpu


def test_resample_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(0)
    P = 17
    nIn = rng.integers(1, 3000, P).astype(np.int32)
    nIn[0] = 8192
    nOut = rng.integers(1, 2000, P).astype(np.int32)
    w = rng.uniform(0.1, 1.0, int(nIn.sum()))
    u = np.exp(rng.normal(0, 3, int(nIn.sum())))
    u0 = rng.uniform(0, 1, P) / nOut
    idx, wn, ess = gpu_ctx.pf_resample(w, u, nIn, nOut, u0=u0)
    oi, oo = np.r_[0, np.cumsum(nIn)], np.r_[0, np.cumsum(nOut)]
    for p in range(P):
        ri, rw = oracle.resample(w[oi[p]:oi[p + 1]], u[oi[p]:oi[p + 1]], int(nOut[p]), u0[p])
        np.testing.assert_array_equal(idx[oo[p]:oo[p + 1]], ri)
        np.testing.assert_allclose(wn[oo[p]:oo[p + 1]], rw, rtol=1e-12)
        W = w[oi[p]:oi[p + 1]] * u[oi[p]:oi[p + 1]]
        W /= W.sum()
        assert ess[p] == pytest.approx(1.0 / np.sum(W * W), rel=1e-10)


def test_resample_random_offset_counts(gpu_ctx):
    # without u0 each draw still satisfies the systematic-resampling bounds
    rng = np.random.default_rng(1)
    n, m = 500, 300
    w = rng.uniform(0, 1, n)
    u = rng.uniform(0.1, 1, n)
    idx, wn, _ = gpu_ctx.pf_resample(w, u, [n], [m], seed=7)
    W = w * u / np.sum(w * u)
    cnt = np.bincount(idx, minlength=n)
    assert np.all(cnt >= np.floor(m * W) - 1) and np.all(cnt <= np.ceil(m * W) + 1)
    assert np.all(np.diff(idx) >= 0)


def test_rot_mean_is_principal_eigenvector(gpu_ctx):
    rng = np.random.default_rng(2)
    P = 9
    cnt = rng.integers(5, 400, P).astype(np.int32)
    qs, ws = [], []
    for p in range(P):
        m = rng.normal(size=4)
        m /= np.linalg.norm(m)
        q = m + 0.2 * rng.normal(size=(cnt[p], 4))
        q *= np.sign(rng.normal(size=(cnt[p], 1)))     # antipodal copies
        qs.append(q / np.linalg.norm(q, axis=1, keepdims=True))
        ws.append(rng.uniform(0.1, 1, cnt[p]))
    quat, w = np.concatenate(qs), np.concatenate(ws)
    mean, k = gpu_ctx.pf_rot_mean(quat, cnt, w)
    off = np.r_[0, np.cumsum(cnt)]
    for p in range(P):
        q, wp = quat[off[p]:off[p + 1]], w[off[p]:off[p + 1]]
        A = (q * wp[:, None]).T @ q / wp.sum()
        ev, V = np.linalg.eigh(A)
        v = V[:, -1] * np.sign(V[0, -1])
        np.testing.assert_allclose(mean[p], v, atol=1e-9)
        np.testing.assert_allclose(k[p], ev[2::-1] / ev[-1], atol=1e-9)


def test_perturb_rotation_algebra(gpu_ctx):
    rng = np.random.default_rng(3)
    P = 5
    nR = np.array([3, 7, 1, 12, 4], np.int32)
    n = int(nR.sum())
    quat = rng.normal(size=(n, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    d = rng.normal(size=(n, 4))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    mean = rng.normal(size=(P, 4))
    mean /= np.linalg.norm(mean, axis=1, keepdims=True)
    out, _ = gpu_ctx.pf_perturb(quat=quat, nR=nR, mean=mean, k=np.ones((P, 3)), d=d)
    off = np.r_[0, np.cumsum(nR)]
    for p in range(P):
        Rm = oracle.rotmat(mean[p])
        for i in range(off[p], off[p + 1]):
            expect = Rm @ oracle.rotmat(d[i]) @ Rm.T @ oracle.rotmat(quat[i])
            np.testing.assert_allclose(oracle.rotmat(out[i]), expect, atol=1e-12)


def _angles(q1, q2):
    return 2 * np.arccos(np.clip(np.abs(np.sum(q1 * q2, 1)), 0, 1))


def test_perturb_rotation_distribution(gpu_ctx):
    # d ~ ACG(1, s k, s k, s k): compare the perturbation angle distribution
    # with numpy draws of the same ACG (mean within 3%)
    rng = np.random.default_rng(4)
    n, kk, pf = 20000, 0.01, 1.5
    quat = np.tile([1.0, 0, 0, 0], (n, 1))
    out, _ = gpu_ctx.pf_perturb(quat=quat, nR=[n], mean=[[1.0, 0, 0, 0]], k=[[kk, kk, kk]],
                                pf=pf, seed=11)
    g = rng.normal(size=(n, 4)) * np.sqrt([1, pf * pf * kk, pf * pf * kk, pf * pf * kk])
    ref = g / np.linalg.norm(g, axis=1, keepdims=True)
    a, b = _angles(out, quat), _angles(ref, quat)
    assert abs(a.mean() / b.mean() - 1) < 0.03
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1, atol=1e-12)


def test_perturb_translation_distribution(gpu_ctx):
    n = 40000
    t = np.zeros((n, 2))
    _, out = gpu_ctx.pf_perturb(trans=t, nT=[n], s01rho=[[1.5, 0.5, 0.3]], pf=2.0, seed=5)
    cov = np.cov(out.T)
    expect = 4.0 * np.array([[1.5 ** 2, 0.3 * 1.5 * 0.5], [0.3 * 1.5 * 0.5, 0.5 ** 2]])
    np.testing.assert_allclose(cov, expect, rtol=0.05, atol=0.02)
    assert np.abs(out.mean(0)).max() < 0.05


def test_proposal_sample_density_and_frequencies(gpu_ctx):
    rng = np.random.default_rng(6)
    C, P, nRot, nTr = 3, 4, 6000, 500
    mix = []
    for p in range(P):
        comps = []
        pis = rng.dirichlet(np.ones(C))
        for c in range(C):
            m = rng.normal(size=4)
            m /= np.linalg.norm(m)
            comps += list(m) + [rng.uniform(0.002, 0.05), pis[c]]
        mix.append(comps + [rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0.5, 2)])
    mix = np.array(mix)
    q, dens, t = gpu_ctx.proposal_sample(mix, C, nRot, nTr, seed=3)
    for p in range(P):
        mp = mix[p]
        qs = q[p * nRot:(p + 1) * nRot]
        # density: sum_c pi_c k^-1.5 ((m.x)^2 + (1-(m.x)^2)/k)^-2 / (2 pi^2)
        ref = np.zeros(nRot)
        comp = np.zeros((nRot, C))
        for c in range(C):
            m, k, pi = mp[6 * c:6 * c + 4], mp[6 * c + 4], mp[6 * c + 5]
            cdot = qs @ m
            s = cdot ** 2 + (1 - cdot ** 2) / k
            comp[:, c] = pi * k ** -1.5 / s ** 2 / (2 * np.pi ** 2)
            ref += comp[:, c]
        np.testing.assert_allclose(dens[p * nRot:(p + 1) * nRot], ref, rtol=1e-10)
        # for draws from the mixture the mean responsibility of c is pi_c
        resp = (comp / ref[:, None]).mean(0)
        pis = mp[5:6 * C:6]
        assert np.abs(resp - pis).max() < 0.02
        ts = t[p * nTr:(p + 1) * nTr]
        sh = mp[6 * C:]
        assert np.abs(ts.mean(0) - sh[:2]).max() < 5 * sh[2] / np.sqrt(nTr)
        assert abs((ts - ts.mean(0)).std() / sh[2] - 1) < 0.1


def test_more_particles_than_a_y_grid(gpu_ctx):
    """70,000 particles (more than 65,535; the particle-filter kernels put
    particles on the grid's x dimension): every particle's samples are
    perturbed, resampled and drawn."""
    rng = np.random.default_rng(70)
    P = 70_000
    print(f"\n[pf] P={P}: perturb, resample, proposal draws", flush=True)
    nR = np.full(P, 2, np.int32)
    quat = rng.normal(size=(2 * P, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    d = np.tile([1.0, 0, 0, 0], (2 * P, 1))          # identity perturbation
    mean = np.tile([1.0, 0, 0, 0], (P, 1))
    out, _ = gpu_ctx.pf_perturb(quat=quat, nR=nR, mean=mean, k=np.ones((P, 3)), d=d)
    np.testing.assert_allclose(out, quat, atol=1e-15)
    t = np.zeros((P, 2))
    _, tout = gpu_ctx.pf_perturb(trans=t, nT=np.ones(P, np.int32),
                                 s01rho=np.tile([1.0, 1.0, 0.0], (P, 1)), pf=1.0, seed=9)
    assert np.all(tout != 0.0)           # every particle's shift was drawn
    nIn = np.full(P, 4, np.int32)
    nOut = np.full(P, 3, np.int32)
    w = np.ones(4 * P)
    u = rng.uniform(0.1, 1.0, 4 * P)
    u0 = rng.uniform(0, 1, P) / 3
    idx, wn, _ = gpu_ctx.pf_resample(w, u, nIn, nOut, u0=u0)
    for p in list(range(0, P, 9973)) + [P - 1]:
        ri, rw = oracle.resample(w[4 * p:4 * p + 4], u[4 * p:4 * p + 4], 3, u0[p])
        np.testing.assert_array_equal(idx[3 * p:3 * p + 3], ri)
        np.testing.assert_allclose(wn[3 * p:3 * p + 3], rw, rtol=1e-12)
    np.testing.assert_allclose(wn.reshape(P, 3).sum(1), 1.0, rtol=1e-12)
    mix = np.tile([1.0, 0, 0, 0, 0.01, 1.0, 0.0, 0.0, 1.0], (P, 1))
    q, dens, tt = gpu_ctx.proposal_sample(mix, 1, 2, 1, seed=4)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    assert np.all(dens > 0) and np.all(tt != 0.0)
