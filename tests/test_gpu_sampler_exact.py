"""The sampler draws from |det Phi|^2: long Monte Carlo runs reproduce the exact
|det Phi|^2-weighted means of H_loc obtained by enumerating all 2^(K N) multi-spin
configurations with the oracle (N = 6, K = 2: 4096 configurations).  Tolerance:
6 standard errors, the standard error taken from the spread of per-chain means."""
import numpy as np
import pytest

from helpers import load_J, random_weights, synthetic_J

pytestmark = pytest.mark.gpu


def exact_emat(oracle, N, M, K, h, J, W):
    tot, acc = 0.0, np.zeros((K, K), complex)
    for cfg in range(1 << (K * N)):
        bits = np.array([[(cfg >> (k * N + i)) & 1 for i in range(N)] for k in range(K)])
        sp = oracle.pack_spins(bits)
        phi, tanh = oracle.state_fcn(N, M, K, W, sp)
        phi_inv, ok = oracle.inv(phi)
        if not ok:
            continue
        wgt = abs(np.linalg.det(phi)) ** 2
        tot += wgt
        acc += wgt * oracle.hloc(N, M, K, h, J, W, sp, phi, phi_inv, tanh)
    return acc / tot


@pytest.mark.parametrize("precision,cplx", [("fp64", True), ("mixed", True), ("mixed", False)])
def test_sampler_matches_exact_expectation(oracle, precision, cplx):
    import nqes
    N, M, K, h = 6, 12, 2, 0.8
    J = synthetic_J(N) * 1.1
    W = random_weights(N, M, K, seed=101, scale=0.25, complex_=cplx)
    ex = exact_emat(oracle, N, M, K, h, J, W)
    C, m, B = 8192, 64, 32
    prec = nqes.PREC_FP64 if precision == "fp64" else nqes.PREC_MIXED
    with nqes.Solver(N, M, K, h, J, precision=prec) as s:
        s.sample(W, 2024, C, m, B, 1.0)
        _, H, _ = s.samples()
    per_chain = H.reshape(C, m, K, K).mean(axis=1)
    mc = per_chain.mean(axis=0)
    se_re = per_chain.real.std(axis=0) / np.sqrt(C)
    se_im = per_chain.imag.std(axis=0) / np.sqrt(C)
    print(precision, "exact", np.round(ex, 5).tolist(), "mc", np.round(mc, 5).tolist(), flush=True)
    assert np.all(np.abs(mc.real - ex.real) <= 6 * se_re + 1e-9)
    assert np.all(np.abs(mc.imag - ex.imag) <= 6 * se_im + 1e-9)
