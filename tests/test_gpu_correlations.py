"""Correlation estimators on the GPU vs the CPU oracle (FP64 path: the chains are
bit-identical; the flipped amplitudes use the incremental update on the GPU and
state_fcn from scratch in the oracle, so the means agree to rounding, 1e-10)."""
import numpy as np
import pytest

from helpers import load_J, random_weights, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,M,K,C,m,B,cplx", [(8, 16, 2, 6, 5, 3, True), (6, 12, 1, 4, 4, 2, True),
                                              (4, 8, 3, 4, 3, 2, False), (8, 64, 2, 3, 3, 2, True)])
def test_correlations_match_oracle(oracle, N, M, K, C, m, B, cplx):
    import nqes
    print(f"case N={N} M={M} K={K} C={C} m={m}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=9 + N, scale=0.15, complex_=cplx)
    ref = oracle.correlations(N, M, K, 0.595, J, W, 31, C, m, B)
    with nqes.Solver(N, M, K, 0.595, J) as s:
        got = s.correlations(W, 31, C, m, B)
    for g, r in zip(got, ref):
        assert rel_err(g, r) < 1e-10, rel_err(g, r)
    emat = np.diag(np.arange(1.0, K + 1)) + 0.1 * np.ones((K, K)) + 0j
    pr = nqes.project_correlations(ref[0], emat)
    assert pr.shape == (N, K)
