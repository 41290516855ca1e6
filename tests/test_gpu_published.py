"""The reference's trained weights on the exact chains of its published run.

The weights data/rbm-weights/ionh0.595train1N50M150K2.mat (tests/golden/, the N=50
ion chain at h = 0.595) were loaded by the run logged at
data/energy/ion_N50_M150_K2_h0.595_c128_m200-1:207-208: `-c 128 -m 200` (128 chains,
200 samples each, the default 64 burn-in sweeps), seed 4202681539.  Its first energy
line (:209), E/N = -1.37841 -1.36371, is the e_mat of the loaded weights on chains
pcg64(4202681539 + c), c < 128 (rbm.cc:569-572, 1179).

* The FP64 path on those chains equals the oracle bit for bit (accepts, spins, H_loc)
  and its e_mat equals the oracle's mean of H_loc.
* The published line itself is not reproduced: on the same chains the oracle gives
  E/N = -1.37313, -1.35743, 0.38 % and 0.46 % above the trace, about 30 standard
  errors of the per-chain spread.  Sampling does not depend on J or h, and splitting
  e_mat into its J and h parts shows that a field about 2.1 % stronger (h = 0.607)
  reproduces both published numbers to 6e-5, while no rescaling of J does.  The
  local energy at N = 50 matches an independent numpy evaluation to 1e-15 and the J
  fixture equals the reference's Ising2d_N50_alpha1.0.mat bit for bit, so the
  published run was not made by the shipped source with these inputs.  The check at
  the trace's print precision is kept as an expected failure with that reason.
* The mixed path on the same chain seeds agrees with the FP64 path statistically.
"""
import os

import numpy as np
import pytest
import scipy.io as sio

from helpers import load_J, rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N, M, K = 50, 150, 2
H_FIELD = 0.595
SEED, C, m, B = 4202681539, 128, 200, 64
TRACE = np.array([-1.37841, -1.36371])  # ion_N50_M150_K2_h0.595_c128_m200-1:209


def trained_weights():
    d = sio.loadmat(os.path.join(GOLDEN, "ionh0.595train1N50M150K2.mat"))
    a, b, w = d["a"].reshape(K, N), d["b"].reshape(K, M), d["w"].reshape(K, N, M)
    return np.concatenate([np.concatenate([a[k], b[k], w[k].reshape(-1)]) for k in range(K)]).astype(np.complex128)


@pytest.fixture(scope="module")
def published(oracle):
    import nqes
    W = trained_weights()
    J = load_J(N)
    print(f"case published N={N} M={M} K={K} C={C} m={m} B={B} seed={SEED}", flush=True)
    ref = oracle.sample_chains(N, M, K, H_FIELD, J, W, SEED, C, m, B)
    with nqes.Solver(N, M, K, H_FIELD, J) as s:
        s.sample(W, SEED, C, m, B, 1.0)
        sp, H, acc = s.samples()
        emat = s.moments()["emat"]
        _, ties = s.ulp_ties()
    return dict(ref=ref, sp=sp, H=H, acc=acc, emat=emat, ties=ties, W=W, J=J)


def test_published_chains_bitexact(published):
    ref = published["ref"]
    assert np.array_equal(published["acc"], ref["accepts"])
    assert np.array_equal(published["sp"], ref["spins"])
    assert rel_err(published["H"], ref["H"]) == 0.0
    assert rel_err(published["emat"], ref["H"].mean(axis=0)) < 1e-12
    assert published["ties"] == 0
    ev = np.sort(np.linalg.eigvals(published["emat"] / N).real)
    print("E/N on the published chains:", ev, flush=True)
    assert np.allclose(ev, [-1.37313, -1.35743], atol=1e-5)


@pytest.mark.xfail(reason="the published trace line sits 0.38 % / 0.46 % below the shipped source's energy on "
                          "the same chains (about 30 standard errors; consistent with h = 0.607 instead of 0.595)",
                   strict=False)
def test_published_trace_line(published):
    ev = np.sort(np.linalg.eigvals(published["emat"] / N).real)
    assert np.all(np.abs(ev - TRACE) < 1e-5)


def test_published_mixed_vs_fp64(published):
    import nqes
    with nqes.Solver(N, M, K, H_FIELD, published["J"], precision=nqes.PREC_MIXED) as s:
        s.sample(published["W"], SEED, C, m, B, 1.0)
        emat = s.moments()["emat"]
    ev_m = np.sort(np.linalg.eigvals(emat / N).real)
    ev_f = np.sort(np.linalg.eigvals(published["emat"] / N).real)
    # the FP32 amplitude ratios decide a few proposals differently, after which those
    # chains differ; the estimate's standard error on 25,600 samples is ~2e-4
    assert np.all(np.abs(ev_m - ev_f) < 1e-3), (ev_m, ev_f)
