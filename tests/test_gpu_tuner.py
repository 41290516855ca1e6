"""The sampler's launch-shape tuner (mixed path).

The first sampling call of a problem shape runs the default launch shapes untimed; the next
calls try candidate shapes in turn (sweep chains per block, sweep tables in LDS or L2,
measurement samples per block), timed on the device; later calls use the fastest.  No lane computes anything else under another shape,
so every call must give the same spins, accept counts, local energies, factor/O rows and
moments bit for bit as the default shapes (NQES_TUNE=0).
"""
import numpy as np
import pytest

from helpers import load_J, random_weights

pytestmark = pytest.mark.gpu

# shapes no other test samples, so that the tuner starts on them here
CASES = [
    # N, M, K, C, m, B, complex weights
    (12, 24, 3, 96, 4, 2, False),
    (12, 24, 2, 80, 5, 2, True),
    (20, 40, 4, 40, 3, 2, False),
]


@pytest.fixture(scope="module")
def nqes():
    import nqes as mod
    return mod


@pytest.mark.parametrize("N,M,K,C,m,B,cplx", CASES)
def test_tuner_trials_same_bits(nqes, monkeypatch, N, M, K, C, m, B, cplx):
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B} complex={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=5 * N + M + K, scale=0.1, complex_=cplx)

    def run(s):
        s.sample(W, 9090, C, m, B, 0.5)
        sp, H, acc = s.samples()
        return (sp, H, acc, s.O_rows(0, C * m), s.moments()), s.sampler_launch()

    monkeypatch.setenv("NQES_TUNE", "0")
    with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
        ref, launch0 = run(s)
    assert launch0["tuning"] == -1
    monkeypatch.delenv("NQES_TUNE")
    launches = []
    with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
        for _ in range(11):  # the untimed first call, at most 8 candidates, then the tuned choice
            out, launch = run(s)
            launches.append(launch)
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[2], ref[2]), launch
            assert np.array_equal(out[1], ref[1]), launch
            assert np.array_equal(out[3], ref[3]), launch
            for key in ("o", "skk", "f", "emat"):
                assert np.array_equal(out[4][key], ref[4][key]), (key, launch)
    assert launches[0] == dict(launch0, tuning=1)  # the untimed first call: default shapes
    trials = [l for l in launches[1:] if l["tuning"] == 1]
    shapes = {(l["chains_per_block"], l["tables_in_lds"], l["samples_per_block"]) for l in trials}
    assert trials[0] == launches[0]  # the default shapes are the first timed candidate
    assert len(trials) >= 2 and len(shapes) == len(trials)
    assert launches[-1]["tuning"] == 0 and launches[-2] == launches[-1]
    fixed = (launches[-1]["chains_per_block"], launches[-1]["tables_in_lds"], launches[-1]["samples_per_block"])
    assert fixed in shapes
