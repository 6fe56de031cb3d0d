"""GPU parity of the batch scheduler (lbx_polymat_mul_batch).

* A mixed batch (shapes, degrees and the four word primes, plus jobs the batch
  kernels hand to the single-product routes: constant matrices, a non-FFT
  prime, a 2^14-point product) against the oracle, job by job.
* Graph replays: the same batch shape with new data, with the program's own
  eager products and synchronisations between replays; every replay is checked.
* The 1024-product batch: 1024 products, m,k,n in {2,...,64}, degree log-uniform in
  [8,4096], prime uniform from the four; every job against one
  lbx_polymat_mul_u32 call (itself checked against the oracle elsewhere), and a
  sample of jobs against the oracle.
"""
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixMulDomain
from linbox_amd.hip import DeviceBuffer, synchronize
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
PRIMES = [998244353, 2013265921, 469762049, 167772161]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def make_jobs(rng, specs):
    """specs: list of (p, m, k, n, sa, sb) -> host arrays + device buffers + job dicts."""
    out = []
    for (p, m, k, n, sa, sb) in specs:
        a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
        da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
        dc = DeviceBuffer(m * n * (sa + sb - 1) * 4)
        dc.zero()
        out.append(dict(p=p, m=m, k=k, n=n, sa=sa, sb=sb, a=da.data_ptr(), b=db.data_ptr(), c=dc.data_ptr(),
                        ha=a, hb=b, da=da, db=db, dc=dc))
    return out


def refill(rng, jobs):
    for j in jobs:
        j["ha"] = rng.integers(0, j["p"], j["ha"].size, dtype=np.uint64).astype(np.uint32)
        j["hb"] = rng.integers(0, j["p"], j["hb"].size, dtype=np.uint64).astype(np.uint32)
        j["da"].upload(j["ha"])
        j["db"].upload(j["hb"])
        j["dc"].zero()


def oracle_ref(j):
    if j["sa"] + j["sb"] > 2 and (j["p"] - 1) % (1 << max(1, (j["sa"] + j["sb"] - 2).bit_length())) == 0:
        st, ref = orc.mul_fft_prime(j["p"], j["m"], j["k"], j["n"], j["ha"], j["sa"], j["hb"], j["sb"])
        assert st == 0
        return ref
    return orc.mul_naive(j["p"], j["m"], j["k"], j["n"], j["ha"], j["sa"], j["hb"], j["sb"])


def mixed_specs(rng, count):
    specs = []
    dims = [2, 4, 8, 16, 32, 64]
    for _ in range(count):
        m, k, n = (int(rng.choice(dims)) for _ in range(3))
        d = int(round(np.exp(rng.uniform(np.log(8), np.log(600)))))
        specs.append((int(rng.choice(PRIMES)), m, k, n, d + 1, d + 1 - int(rng.integers(0, 3))))
    # jobs outside the batch kernels: constants, non-FFT prime, 2^14 points, 2^2 points
    specs += [(998244353, 3, 4, 5, 1, 1), (1000000007, 4, 3, 2, 20, 30), (2013265921, 2, 3, 2, 5000, 4000),
              (469762049, 2, 2, 3, 2, 3)]
    return specs


def test_batch_mixed_vs_oracle(ctx):
    rng = np.random.default_rng(77)
    jobs = make_jobs(rng, mixed_specs(rng, 40))
    say("batch mixed", len(jobs))
    lbx.mul_batch(jobs, ctx=ctx)
    synchronize()
    for i, j in enumerate(jobs):
        got = j["dc"].download()
        assert np.array_equal(got, oracle_ref(j)), (i, j["p"], j["m"], j["k"], j["n"], j["sa"], j["sb"])


def test_batch_graph_replays(ctx):
    """Same shapes, new data each time, eager single products between replays."""
    rng = np.random.default_rng(78)
    jobs = make_jobs(rng, mixed_specs(rng, 16))
    p = 469762049
    for rep in range(3):
        refill(rng, jobs)
        say("batch replay", rep)
        lbx.mul_batch(jobs, ctx=ctx)
        # eager work of the program between replays
        A = PolynomialMatrix.from_polfirst(rng.integers(0, p, (6, 40), dtype=np.uint64).astype(np.uint32), 2, 3)
        B = PolynomialMatrix.from_polfirst(rng.integers(0, p, (6, 33), dtype=np.uint64).astype(np.uint32), 3, 2)
        C = PolynomialMatrix(2, 2, 1)
        PolynomialMatrixMulDomain(p, ctx).mul(C, A, B)
        synchronize()
        for i, j in enumerate(jobs):
            assert np.array_equal(j["dc"].download(), oracle_ref(j)), (rep, i)


def test_batch_1024_products(ctx):
    """The 1024-product mix as one batch against one product call per job."""
    rng = np.random.default_rng(0x4C494E424F58 + 4)
    dims = [2, 4, 8, 16, 32, 64]
    specs = []
    for _ in range(1024):
        m, k, n = (int(rng.choice(dims)) for _ in range(3))
        d = int(round(np.exp(rng.uniform(np.log(8), np.log(4096)))))
        specs.append((int(rng.choice(PRIMES)), m, k, n, d + 1, d + 1))
    jobs = make_jobs(rng, specs)
    say("1024-product batch")
    lbx.mul_batch(jobs, ctx=ctx)
    synchronize()
    got = [j["dc"].download() for j in jobs]
    say("1024-product per-product calls")
    for i, j in enumerate(jobs):
        j["dc"].zero()
        C = PolynomialMatrix(j["m"], j["n"], j["sa"] + j["sb"] - 1, data=j["dc"])
        PolynomialMatrixMulDomain(j["p"], ctx, lbx.METHOD_PLAIN_U64).mul(
            C, PolynomialMatrix(j["m"], j["k"], j["sa"], data=j["da"]),
            PolynomialMatrix(j["k"], j["n"], j["sb"], data=j["db"]))
    synchronize()
    for i, j in enumerate(jobs):
        assert np.array_equal(got[i], j["dc"].download()), i
    say("1024-product oracle sample")
    for i in range(0, 1024, 97):
        if jobs[i]["m"] * jobs[i]["k"] * jobs[i]["n"] * jobs[i]["sa"] <= 64 * 64 * 64 * 600:
            assert np.array_equal(got[i], oracle_ref(jobs[i])), i


@pytest.mark.parametrize("form", ["int8", "plain", None])
def test_batch_int8_jobs(form, monkeypatch):
    """Jobs whose three dimensions are all >= 16 (16..64, several inner
    dimensions, lengths 2^4..2^13) take the INT8 MFMA stage of the batch
    (limb conversion, MFMA products, transpose) when forced to or when the
    tuner measures it faster; every job against the oracle.  A fresh context
    each time, so the form is decided anew."""
    if form:
        monkeypatch.setenv("LBX_BATCH_FORM", form)
    rng = np.random.default_rng({"int8": 1, "plain": 2, None: 3}[form])
    specs = []
    for (m, k, n, d) in [(16, 16, 16, 7), (32, 16, 64, 100), (64, 64, 64, 300), (16, 48, 32, 2000), (64, 32, 16, 4000),
                         (32, 32, 32, 20), (17, 33, 20, 60)]:
        specs.append((int(rng.choice(PRIMES)), m, k, n, d + 1, d + 1 - int(rng.integers(0, 3))))
    specs += [(998244353, 4, 8, 2, 30, 30), (469762049, 8, 16, 32, 100, 90)]  # plain-form jobs beside them
    jobs = make_jobs(rng, specs)
    c = lbx.Context(0)
    try:
        say("batch int8 jobs", form)
        lbx.mul_batch(jobs, ctx=c)
        synchronize()
        tuning = lbx.lib().lbx_batch_tuning(c.handle).decode()
        print("tuning:", tuning)
        for j in jobs:
            assert np.array_equal(j["dc"].download(), oracle_ref(j)), (j["p"], j["m"], j["k"], j["n"], j["sa"])
        # a replay with new data through the cached plan and graph
        refill(rng, jobs)
        lbx.mul_batch(jobs, ctx=c)
        synchronize()
        for j in jobs:
            assert np.array_equal(j["dc"].download(), oracle_ref(j))
    finally:
        c.close()
