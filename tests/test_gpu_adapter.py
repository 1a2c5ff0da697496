"""The C++ adapter (include/thunder_expect.hpp: Projector, Optimiser::expectation,
ExpectLocalRTD / PreI3D / M) driven by a compiled C++ program
(tests/cpp/adapter_check, built with the library) through
libthunder_expect.so, against the oracle and against the C ABI called
directly (needs gfx950).

  * Projector::project (batched and one rotation) against oracle.project
    (FP64 coordinates, FP32 trilinear), 1e-5 of the largest magnitude; the
    one-rotation call equals the batched row bit for bit;
  * Projector::setProjectee from a 32^3 map's spectrum against the FP64
    projectee oracle, 2e-5 of the largest magnitude, maxRadius = N/2 - 1;
  * ExpectLocalRTD + PreI3D + M for single particles (ragged: 7 to 300
    rotations, 1 to 40 shifts) against the FP64 local-scan oracle, with the
    reference's factors oldC and oldD on the marginals: relative 3e-3 (the
    bar of the FP32 local scorer);
  * Expectation::expectation (global scan, post-scan update, local phases)
    equal, bit for bit, to the same sequence of C-ABI calls with the same
    seed.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import thx
import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "adapter_check")


def _write(d, name, a, dtype):
    np.ascontiguousarray(a, dtype=dtype).tofile(os.path.join(d, name))


def _read(d, name, dtype, shape):
    return np.fromfile(os.path.join(d, f"out_{name}.bin"), dtype=dtype).reshape(shape)


def _colmajor(M):
    return M.T.reshape(9)


def test_adapter_against_oracle_and_c_abi(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: build it with __graft_entry__.build()")
    d = str(tmp_path)
    print("\n[adapter] box64: 16 particles, 300-rotation grid", flush=True)
    ctx = thx.Context(0)
    try:
        info = wl.setup_context(ctx, "box64", n_particles=16, n_rot=300)
        N, pf = 64, 2
        V = N * pf
        vol = ctx.get_volume(0, V)
        dat, ctf, sig = ctx.get_particles()
        iCol, iRow, iSig = info["iCol"], info["iRow"], info["iSig"]
        K, P = len(iCol), dat.shape[0]
        quat, trans = info["quat"], info["trans"]
        R, T = len(quat), len(trans)
        rng = np.random.default_rng(3)
        q = rng.normal(size=(6, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        mats = np.array([_colmajor(oracle.rotmat(x)) for x in q])
        Ns = 32
        vft = np.fft.rfftn(rng.normal(size=(Ns, Ns, Ns)), axes=(0, 1, 2))
        # ragged single-particle local lists around each particle's true pose
        cases = [(3, 7, 1), (0, 130, 25), (11, 300, 40)]
        lrot, ltr, loR, loT = [], [], [], []
        for l, nR, nT in cases:
            g = rng.normal(size=(nR, 4)) * [1, 0.05, 0.05, 0.05]
            g /= np.linalg.norm(g, axis=1, keepdims=True)
            tq = info["true_quat"][l]
            # g (x) true pose: rotations within a few degrees of it
            w0, x0, y0, z0 = tq
            Lm = np.array([[w0, -x0, -y0, -z0], [x0, w0, -z0, y0], [y0, z0, w0, -x0],
                           [z0, -y0, x0, w0]])
            lrot.append(g @ Lm.T)
            ltr.append(info["true_shift"][l] + rng.uniform(-2, 2, size=(nT, 2)))
            loR.append(rng.uniform(0.2, 1.0, nR))
            loT.append(rng.uniform(0.2, 1.0, nT))
        loC = rng.uniform(0.5, 1.5, len(cases))
        meta = dict(N=N, pf=pf, K=K, P=P, R=R, T=T, nM=len(mats), Ns=Ns, nL=len(cases),
                    local_nrot=sum(c[1] for c in cases), local_ntrans=sum(c[2] for c in cases),
                    mLR=100, mLT=16, maxPhase=15, precision=thx.FP32, seed=5)
        with open(os.path.join(d, "meta.txt"), "w") as f:
            for k, v in meta.items():
                f.write(f"{k} {v}\n")
        for name, a, dt in [("iCol", iCol, np.int32), ("iRow", iRow, np.int32),
                            ("iSig", iSig, np.int32), ("vol", vol, np.float32),
                            ("mats", mats, np.float64),
                            ("volft", np.stack([vft.real, vft.imag], -1), np.float32),
                            ("dat", dat, np.float32), ("ctf", ctf, np.float32),
                            ("sig", sig, np.float32),
                            ("local_idx", [[l, nR, nT, 0] for l, nR, nT in cases], np.int32),
                            ("local_rot", np.concatenate(lrot), np.float64),
                            ("local_trans", np.concatenate(ltr), np.float64),
                            ("local_oldR", np.concatenate(loR), np.float64),
                            ("local_oldT", np.concatenate(loT), np.float64),
                            ("local_oldC", loC, np.float64), ("grid_quat", quat, np.float64),
                            ("grid_trans", trans, np.float64)]:
            _write(d, name, a, dt)
        print("[adapter] running the C++ driver", flush=True)
        r = subprocess.run([BIN, d], capture_output=True, text=True, timeout=300)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        assert "error as expected" in r.stdout
        # Projector::project
        got = _read(d, "project", np.float32, (len(mats), K, 2))
        for i, x in enumerate(q):
            ref = oracle.project(vol, V, oracle.rotmat(x), iCol, iRow, pf)
            err = np.abs(got[i] - ref).max() / np.abs(ref).max()
            assert err <= 1e-5, (i, err)
        np.testing.assert_array_equal(_read(d, "project1", np.float32, (K, 2)), got[0])
        # Projector::setProjectee
        pad = _read(d, "projectee", np.float32, (Ns * pf, Ns * pf, Ns * pf // 2 + 1, 2))
        pad = pad[..., 0].astype(np.float64) + 1j * pad[..., 1]
        ref = oracle.projectee(vft, Ns, pf)
        assert np.abs(pad - ref).max() / np.abs(ref).max() <= 2e-5
        assert _read(d, "maxradius", np.int32, (1,))[0] == Ns // 2 - 1
        # ExpectLocalRTD / PreI3D / M
        for i, (l, nR, nT) in enumerate(cases):
            w = _read(d, f"local_{i}", np.float32, (2 + nR + nT,)).astype(np.float64)
            wC, wR, wT, base, _, _ = oracle.local_scan(vol, V, pf, iCol, iRow, N, lrot[i], [nR],
                                                       ltr[i], [nT], loR[i], loT[i],
                                                       dat[l:l + 1], ctf[l:l + 1], sig[l:l + 1])
            np.testing.assert_allclose(w[0], wC[0], rtol=3e-3)
            np.testing.assert_allclose(w[1], loC[i] * wC[0], rtol=3e-3)
            np.testing.assert_allclose(w[2:2 + nR], loC[i] * wR, rtol=3e-3,
                                       atol=3e-3 * loC[i] * wR.max())
            np.testing.assert_allclose(w[2 + nR:], loC[i] * wT, rtol=3e-3,
                                       atol=3e-3 * loC[i] * wT.max())
        # Expectation::expectation == the C-ABI sequence with the same seed
        ctx.score_global(0, thx.FP32)
        ctx.pf_from_global(0, 100, 16, R ** (-1.0 / 3), 1.0 / (2 * np.log(2)) / np.sqrt(np.pi),
                           0.5, 1.0, seed=5)
        phases = ctx.pf_refine(np.zeros(P, np.int32), thx.FP32, after_global=True, pf_l=2.0,
                               pf_s=0.5, max_phase=15, seed=5)
        st = ctx.pf_get_state()
        assert _read(d, "phases", np.int32, (1,))[0] == phases
        for key, shape in [("quat", (P, 100, 4)), ("trans", (P, 16, 2)), ("k", (P, 3)),
                           ("s01", (P, 2)), ("topR", (P, 4)), ("topT", (P, 2))]:
            np.testing.assert_array_equal(_read(d, key, np.float64, shape), st[key])
        np.testing.assert_array_equal(_read(d, "nPhase", np.int32, (P,)), st["nPhase"])
        print(f"[adapter] expectation: {phases} phases, converged "
              f"{np.mean(st['nPhase'] >= 0):.2f}")
    finally:
        ctx.close()
