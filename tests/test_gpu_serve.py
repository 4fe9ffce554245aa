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


def test_served_workload_at_size():
    """The served workload at its size: three resident 2^24-point tables,
    a 200 GB HBM budget, 4 workers, and 36 jobs drawn as the benchmark draws
    them (70 % MSM with log2 N ~ U[14, 24] over a slice of one scalar pool,
    30 % NTT with log2 n ~ U[12, 24], forward or inverse) arriving as a burst
    and then a Poisson tail.  Every result equals the job's alone-run bytes
    (plain server) and the oracle: (sum s_i k_i mod r) G for the MSMs, the
    CPU transform for the NTTs."""
    print("serve workload: three 2^24 tables", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(3131)
    from hardcaml_zprize_amd import workload

    tables, ks_all = [], []
    for t in range(3):
        ctx, ks, _, _ = workload.generate_table(rng, 24)
        tables.append(ctx)
        ks_all.append(ks)
    pool = workload.rand_scalars(rng, 1 << 24)
    ntt_pool = workload.rand_ntt_input(rng, 1 << 24)
    specs = workload.served_specs(rng, 36)
    assert any(sp["kind"] == "ntt" for sp in specs) and any(sp["kind"] == "msm" for sp in specs)

    def make():
        jobs = []
        for i, sp in enumerate(specs):
            if sp["kind"] == "msm":
                n = 1 << sp["log_n"]
                jobs.append(serve.Job(serve.MSM, id=i, log_n=sp["log_n"], table=sp["table"],
                                      scalars=pool[sp["off"]:sp["off"] + n]))
            else:
                jobs.append(serve.Job(serve.NTT, id=i, log_n=sp["log_n"], direction=sp["direction"],
                                      data=ntt_pool[:1 << sp["log_n"]]))
        return jobs

    try:
        alone = serve.GpuServer(tables, hbm_budget=200e9, workers=1, batching=False)
        ja = make()
        serve.run_stream(alone, ja, [0.0] * len(ja))
        alone.close()
        print("serve workload: served run", flush=True)
        served = serve.GpuServer(tables, hbm_budget=200e9, workers=4, batching=True, max_batch_s=0.1)
        jm = make()
        arrivals = [0.0] * 12 + list(0.001 + np.cumsum(rng.exponential(0.004, size=len(jm) - 12)))
        serve.run_stream(served, jm, arrivals)
        served.close()
        g = oracle.generator()
        print("serve workload: oracle checks", flush=True)
        for a, b, sp in zip(ja, jm, specs):
            assert a.result.tobytes() == b.result.tobytes(), (a.id, sp)
            if sp["kind"] == "msm":
                n = 1 << sp["log_n"]
                e = oracle.dot_mod_r(pool[sp["off"]:sp["off"] + n], ks_all[sp["table"]][:n])
                assert oracle.affine_equal(oracle.projective_to_affine(a.result), oracle.g1_scalar_mul(g, e)), sp
            else:
                want = oracle.ntt(ntt_pool[:1 << sp["log_n"]], inverse=bool(sp["direction"]))
                assert np.array_equal(a.result, want), sp
    finally:
        for t in tables:
            t.close()
