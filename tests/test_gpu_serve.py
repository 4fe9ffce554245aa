"""Served form: continuous batching gives every job the same bytes as the job
run alone, and matches the oracle (small sizes)."""
import random
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import serve

pytestmark = pytest.mark.gpu
R = oracle.R_ORDER


@pytest.mark.parametrize("table_options", ["", "precompute=auto"])
def test_served_mix_equals_alone_and_oracle(table_options):
    """Also over tables loaded with their window-shifted copies, where the
    MSM jobs take the one shared bucket set."""
    print(f"serve mix [{table_options}]", flush=True)
    sys.stdout.flush()
    rng = random.Random(5)
    nrng = np.random.default_rng(5)
    tables, ks_all = [], []
    for t in range(2):
        ks = [rng.randrange(1, R) for _ in range(2048)]
        ks_all.append(ks)
        pts = zmsm.generate_points(np.array([oracle.int_to_words(k, 4) for k in ks]))
        tables.append(zmsm.MsmContext(pts, table_options))
    specs = []
    for i in range(40):
        if i % 3:
            n = rng.choice([1, 64, 500, 2048])
            sc = [rng.randrange(R) for _ in range(n)]
            specs.append(("m", i % 2, sc))
        else:
            k = rng.choice([10, 12, 13, 15])
            specs.append(("n", k, nrng.integers(0, oracle.P, size=1 << k, dtype=np.uint64), i % 2))

    def make():
        jobs = []
        for i, sp in enumerate(specs):
            if sp[0] == "m":
                words = np.array([oracle.int_to_words(v, 4) for v in sp[2]], dtype=np.uint64)
                jobs.append(serve.Job(serve.MSM, id=i, log_n=max(len(sp[2]) - 1, 1).bit_length(), table=sp[1],
                                      scalars=words))
            else:
                jobs.append(serve.Job(serve.NTT, id=i, log_n=sp[1], data=sp[2], direction=sp[3]))
        return jobs

    alone = serve.GpuServer(tables, workers=1, batching=False)
    ja = make()
    serve.run_stream(alone, ja, [0.0] * len(ja))
    alone.close()
    mixed = serve.GpuServer(tables, workers=2, batching=True, hbm_budget=2e9)
    jm = make()
    # a burst (everything queued at once, so jobs must share batches) then a Poisson tail
    arrivals = [0.0] * 24 + list(0.001 + np.cumsum(nrng.exponential(0.002, size=len(jm) - 24)))
    serve.run_stream(mixed, jm, arrivals)
    assert mixed.batches < len(jm)  # some jobs really shared a batch
    mixed.close()
    g = oracle.generator()
    for a, b, sp in zip(ja, jm, specs):
        assert a.result.tobytes() == b.result.tobytes(), a.id
        if sp[0] == "m":
            e = sum(s * k for s, k in zip(sp[2], ks_all[sp[1]])) % R
            assert oracle.affine_equal(oracle.projective_to_affine(a.result), oracle.g1_scalar_mul(g, e))
        else:
            assert np.array_equal(a.result, oracle.ntt(sp[2], inverse=bool(sp[3])))
    for t in tables:
        t.close()
