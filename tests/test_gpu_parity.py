"""GPU parity tests: the HIP path through the C ABI against the oracle and fixtures.

Bit-exact equality is required everywhere (integer arithmetic).  Small and
mid sizes are compared word for word with the oracle (oracle/lbx_oracle.c) and
the committed golden fixtures; full-size configurations are checked through
size-independent properties (random-point evaluation, agreement of the two
pointwise forms).  Every step prints its case first so a failing run names it.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixMulDomain, PolynomialMatrixFFTMulDomain, LinboxError
from oracle import oracle as orc
from polycheck import check_by_evaluation

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PRIMES = [998244353, 2013265921, 469762049, 167772161]
METHODS = [lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def rand(rng, p, n):
    return rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)


def gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method, layouts=("polfirst",) * 3, max_rowdeg=0):
    A = PolynomialMatrix.from_polfirst(np.asarray(a).reshape(m * k, sa), m, k, layouts[0])
    B = PolynomialMatrix.from_polfirst(np.asarray(b).reshape(k * n, sb), k, n, layouts[1])
    C = PolynomialMatrix(m, n, 1, layouts[2])
    PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B, max_rowdeg)
    return C.to_polfirst()


# ------------------------------------------------------------------ NTT ---
def test_ntt_golden_kat(ctx):
    for case in load("ntt_kat.json"):
        p, L = case["p"], case["L"]
        say("ntt kat", p, L)
        x = np.array(case["in"], dtype=np.uint32)
        w = lbx.ntt_forward(p, L, x, ctx=ctx)
        assert w == case["root"]
        assert x.tolist() == case["out_br"]
        y = np.array(case["in_br"], dtype=np.uint32)
        lbx.ntt_inverse(p, L, y, root=case["root"], ctx=ctx)
        assert y.tolist() == case["out"]


@pytest.mark.parametrize("p", PRIMES)
def test_ntt_vs_oracle_all_sizes(ctx, p):
    rng = np.random.default_rng(p + 1)
    for L in range(1, 16):
        count = max(1, min(37, (1 << 18) >> L))
        say("ntt vs oracle", p, L, count)
        x = rand(rng, p, count << L)
        y = x.copy()
        w = lbx.ntt_forward(p, L, y, ctx=ctx)
        ref = np.concatenate([orc.fft_direct(p, L, x[i << L:(i + 1) << L])[0] for i in range(count)])
        assert np.array_equal(y, ref), (p, L)
        wi = orc.invroot(p, L, w)
        z = y.copy()
        lbx.ntt_inverse(p, L, z, root=wi, ctx=ctx)
        ref = np.concatenate([orc.fft_inverse(p, L, y[i << L:(i + 1) << L], root=wi)[0] for i in range(count)])
        assert np.array_equal(z, ref), (p, L)
        assert np.array_equal(z.astype(np.uint64), x.astype(np.uint64) * (1 << L) % p)


def test_ntt_errors(ctx):
    x = np.zeros(8, dtype=np.uint32)
    with pytest.raises(LinboxError) as e:
        lbx.ntt_forward(998244353, 0, x, count=1, ctx=ctx)
    assert e.value.status == 3
    with pytest.raises(LinboxError) as e:
        lbx.ntt_forward(998244353, 3, x, root=998244352, ctx=ctx)
    assert e.value.status == 4
    with pytest.raises(LinboxError) as e:
        lbx.ntt_forward(998244353, 24, x, count=0, ctx=ctx)
    assert e.value.status == 2


# -------------------------------------------------------------- product ---
@pytest.mark.parametrize("method", METHODS)
def test_mul_golden_fixtures(ctx, method):
    for cs in load("polymat_mul.json"):
        say("golden", cs["p"], cs["m"], cs["k"], cs["n"], cs["sa"], cs["sb"], cs["max_rowdeg"], method)
        got = gpu_mul(ctx, cs["p"], cs["m"], cs["k"], cs["n"], cs["a"], cs["sa"], cs["b"], cs["sb"], method,
                      max_rowdeg=cs["max_rowdeg"])
        assert got.reshape(-1).tolist() == cs["c"]


SHAPES = [(16, 16, 16, 256, 256), (5, 7, 3, 100, 37), (33, 65, 17, 40, 90), (1, 1, 1, 3000, 5000),
          (64, 3, 48, 17, 17), (2, 130, 2, 9, 8)]


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("method", METHODS)
def test_mul_vs_oracle(ctx, p, method):
    rng = np.random.default_rng(p ^ method)
    for (m, k, n, sa, sb) in SHAPES:
        say("mul vs oracle", p, m, k, n, sa, sb, method)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method)
        assert np.array_equal(got.reshape(-1), ref), (p, m, k, n, sa, sb, method)


@pytest.mark.parametrize("p", [2013265921, 998244353])
def test_mul_every_transform_length(ctx, p):
    """Product lengths 2^1..2^15 (every transform kernel instantiation, split and whole), both forms."""
    rng = np.random.default_rng(p + 77)
    for L in range(1, 16):
        pts = 1 << L
        sa = max(1, pts // 2 - (L % 3))
        sb = pts - sa  # deg = sa+sb-2 = pts-2 -> length 2^L
        if sa + sb < 3:
            sa, sb = 1, 2
        m, k, n = (3, 5, 2) if L < 13 else (2, 2, 2)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        for method in METHODS:
            say("length", L, p, m, k, n, sa, sb, method)
            got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method)
            assert np.array_equal(got.reshape(-1), ref), (L, method)


@pytest.mark.parametrize("pipe", ["1", "0"])
@pytest.mark.parametrize("chunk_bytes", [4096, 1 << 20])
def test_mul_chunked_pipeline(ctx, chunk_bytes, pipe, monkeypatch):
    """Tiny intermediate chunks force many head/tail chunks and padding rows
    inside chunks: with the two-stream hand-offs over both scratch slots, and
    (LBX_CHUNK_PIPE=0) in order on one stream through a single slot."""
    monkeypatch.setenv("LBX_CHUNK_BYTES", str(chunk_bytes))
    monkeypatch.setenv("LBX_CHUNK_PIPE", pipe)
    for (p, m, k, n, sa, sb) in [(2013265921, 37, 20, 45, 300, 200), (998244353, 64, 64, 33, 129, 128),
                                 (469762049, 5, 130, 7, 1000, 1100)]:
        rng = np.random.default_rng(m * n + chunk_bytes)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        say("chunked", chunk_bytes, pipe, p, m, k, n, sa, sb)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_INT8_MFMA)
        assert np.array_equal(got.reshape(-1), ref)


def test_mul_split_all_lengths(ctx, monkeypatch):
    """Split transforms (forward head + blockwise tail, blockwise inverse head +
    per-entry tail) at every transform length from 2^8 to 2^15, with the
    head -> tail chunking off and on."""
    cases = [(998244353, 16, 16, 16, 129, 128), (2013265921, 17, 33, 19, 300, 213),
             (469762049, 32, 20, 40, 600, 500), (2013265921, 16, 17, 18, 1025, 1000),
             (998244353, 20, 16, 16, 2048, 2049), (2013265921, 16, 16, 16, 4096, 4000),
             (2013265921, 16, 16, 17, 8192, 8192), (469762049, 16, 16, 16, 16384, 16384)]
    for (p, m, k, n, sa, sb) in cases:
        rng = np.random.default_rng(m * n * sa)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        for chunk in (None, 1 << 20):
            if chunk:
                monkeypatch.setenv("LBX_CHUNK_BYTES", str(chunk))
            else:
                monkeypatch.delenv("LBX_CHUNK_BYTES", raising=False)
            say("split", p, m, k, n, sa, sb, chunk)
            got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_INT8_MFMA)
            assert np.array_equal(got.reshape(-1), ref)


def test_mul_mfma_shapes(ctx):
    """INT8 pointwise stage on shapes with partial row bands and column blocks,
    several inner-dimension chunks, and fewer points than workgroups."""
    # register-A form: row bands of 32/64/128, k <= 128, B block <= 64 KiB;
    # LDS-A form: 96 rows, k in (128, ...] or wide B; global form: A block > 64 KiB
    for (p, m, k, n, sa, sb) in [(2013265921, 33, 100, 70, 200, 57), (998244353, 128, 128, 128, 40, 41),
                                 (469762049, 100, 64, 130, 300, 300), (2013265921, 16, 200, 16, 130, 127),
                                 (998244353, 97, 17, 33, 5, 4), (2013265921, 70, 40, 50, 90, 60),
                                 (469762049, 20, 130, 40, 64, 64), (998244353, 32, 64, 300, 33, 32),
                                 (2013265921, 150, 150, 20, 20, 20)]:
        rng = np.random.default_rng(m * k + n)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        say("mfma shape", p, m, k, n, sa, sb)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_INT8_MFMA)
        assert np.array_equal(got.reshape(-1), ref)


def test_mul_config1_vs_naive(ctx):
    """16x16.16x16 over 998244353, degree 255: bit-exact vs naive convolution."""
    p, m, k, n, sa, sb = 998244353, 16, 16, 16, 256, 256
    a = orc.fill_uniform(0x4C494E424F58 + 0, p, m * k * sa)
    b = orc.fill_uniform(0x4C494E424F58 + 100, p, k * n * sb)
    ref = orc.mul_naive(p, m, k, n, a, sa, b, sb)
    for method in METHODS:
        say("config1", method)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method)
        assert np.array_equal(got.reshape(-1), ref)


@pytest.mark.parametrize("layouts", [("matfirst", "matrowfirst", "polfirst"), ("matrowfirst", "matfirst", "matfirst"),
                                     ("polfirst", "polfirst", "matrowfirst")])
def test_mul_layouts(ctx, layouts):
    p = 469762049
    rng = np.random.default_rng(3)
    m, k, n, sa, sb = 6, 5, 7, 33, 20
    a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
    _, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
    for method in METHODS:
        say("layouts", layouts, method)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method, layouts=layouts)
        assert np.array_equal(got.reshape(-1), ref)


def test_mul_views(ctx):
    """Operands and result as SubPolynomialMatrix windows of larger storage."""
    p = 167772161
    rng = np.random.default_rng(4)
    m, k, n = 3, 4, 2
    Abig = PolynomialMatrix.from_polfirst(rand(rng, p, m * k * 50).reshape(m * k, 50), m, k, "matfirst")
    Bbig = PolynomialMatrix.from_polfirst(rand(rng, p, k * n * 40).reshape(k * n, 40), k, n, "polfirst")
    A, B = Abig.at(5, 24), Bbig.at(3, 33)  # sizes 20 and 31
    Cbig = PolynomialMatrix.from_polfirst(np.full((m * n, 80), 7, dtype=np.uint32), m, n, "matrowfirst")
    C = Cbig.at(10, 79)  # 70 >= 50 coefficients
    _, ref = orc.mul_fft_prime(p, m, k, n, A.to_polfirst().reshape(-1), 20, B.to_polfirst().reshape(-1), 31)
    say("views")
    PolynomialMatrixMulDomain(p, ctx).mul(C, A, B)
    assert C.size() == 50
    full = Cbig.to_polfirst()
    assert np.array_equal(full[:, 10:60].reshape(-1), ref)
    assert (full[:, :10] == 7).all() and (full[:, 60:] == 7).all()
    small = Cbig.at(0, 20)
    with pytest.raises(LinboxError) as e:
        PolynomialMatrixMulDomain(p, ctx).mul(small, A, B)
    assert e.value.status == 5


def test_mul_errors(ctx):
    p = 998244353
    A = PolynomialMatrix(2, 3, 4)
    B = PolynomialMatrix(2, 2, 4)
    C = PolynomialMatrix(2, 2, 7)
    with pytest.raises(LinboxError) as e:
        PolynomialMatrixMulDomain(p, ctx).mul(C, A, B)
    assert e.value.status == 5
    a = np.ones(8, dtype=np.uint32)
    c = np.zeros(8, dtype=np.uint32)
    st = lbx.lib().lbx_polymat_mul_fft_u32(ctx.handle, 7, 3, 1, 1, 1, a.ctypes.data, a.ctypes.data,
                                           c.ctypes.data, 0, None)
    assert st == 1  # 8 does not divide 7-1: bad FFT prime


def test_constant_matrices(ctx):
    """a.size()+b.size() <= 2 takes the naive route of PolynomialMatrixMulDomain::mul."""
    p = 2013265921
    rng = np.random.default_rng(9)
    m, k, n = 7, 9, 5
    a, b = rand(rng, p, m * k), rand(rng, p, k * n)
    ref = orc.mul_naive(p, m, k, n, a, 1, b, 1)
    say("constant")
    got = gpu_mul(ctx, p, m, k, n, a, 1, b, 1, lbx.METHOD_AUTO)
    assert np.array_equal(got.reshape(-1), ref)


def test_device_pointers(ctx):
    from linbox_amd.hip import DeviceBuffer, synchronize
    p, m, k, n, sa, sb = 2013265921, 8, 8, 8, 64, 64
    rng = np.random.default_rng(12)
    a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
    _, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
    da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
    dc = DeviceBuffer(m * n * (sa + sb - 1) * 4)
    for method in METHODS:
        say("device pointers", method)
        dc.zero()
        A = PolynomialMatrix(m, k, sa, data=da)
        B = PolynomialMatrix(k, n, sb, data=db)
        C = PolynomialMatrix(m, n, sa + sb - 1, data=dc)
        PolynomialMatrixMulDomain(p, ctx, method).mul(C, A, B)
        synchronize()
        assert np.array_equal(dc.download(), ref)


TORCH_SCRIPT = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixMulDomain
from oracle import oracle as orc
p, m, k, n, sa, sb = 998244353, 9, 17, 33, 70, 31
rng = np.random.default_rng(5)
a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
_, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
ta = torch.from_numpy(a.view(np.int32)).cuda()
tb = torch.from_numpy(b.view(np.int32)).cuda()
tc = torch.zeros(m * n * (sa + sb - 1), dtype=torch.int32, device="cuda")
s = torch.cuda.current_stream().cuda_stream
for method in (1, 2):
    tc.zero_()
    PolynomialMatrixMulDomain(p, method=method, stream=s).mul(
        PolynomialMatrix(m, n, sa + sb - 1, data=tc), PolynomialMatrix(m, k, sa, data=ta),
        PolynomialMatrix(k, n, sb, data=tb))
    torch.cuda.synchronize()
    assert np.array_equal(tc.cpu().numpy().view(np.uint32), ref), method
print("torch interop ok")
"""


def test_torch_tensor_interop():
    """With torch imported first, liblbx binds torch's HIP runtime and accepts its tensors."""
    import subprocess
    pytest.importorskip("torch")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    say("torch interop subprocess")
    r = subprocess.run([sys.executable, "-c", TORCH_SCRIPT, root], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "torch interop ok" in r.stdout


# ----------------------------------------------------------- full sizes ---
def full_size_check(ctx, p, m, sa, seed):
    from linbox_amd.hip import DeviceBuffer, synchronize
    a = orc.fill_uniform(seed, p, m * m * sa)
    b = orc.fill_uniform(seed + 100, p, m * m * sa)
    da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
    sc = 2 * sa - 1
    dc = DeviceBuffer(m * m * sc * 4)
    outs = []
    for method in METHODS:
        say("full size", m, sa, p, method)
        dc.zero()
        A, B, C = (PolynomialMatrix(m, m, sa, data=da), PolynomialMatrix(m, m, sa, data=db),
                   PolynomialMatrix(m, m, sc, data=dc))
        PolynomialMatrixMulDomain(p, ctx, method).mul(C, A, B)
        synchronize()
        outs.append(dc.download())
    dc.free()
    da.free()
    db.free()
    lbx.lib().lbx_trim(ctx.handle)
    assert np.array_equal(outs[0], outs[1])
    rng = np.random.default_rng(seed)
    xs = [int(x) for x in rng.integers(2, p, 2)]
    say("full size evaluation check", m, sa)
    assert check_by_evaluation(a, b, outs[0], p, m, m, m, sa, sa, sc, xs)
    # one whole row of C word for word against the oracle (row m-1: the last
    # row band, padding rows of the limb blocks next to it)
    i = m - 1
    say("full size oracle row", m, sa, i)
    st, ref = orc.mul_fft_prime(p, 1, m, m, a[i * m * sa:(i + 1) * m * sa], sa, b, sa)
    assert st == 0
    assert np.array_equal(outs[0][i * m * sc:(i + 1) * m * sc], ref)


def test_config2_full_size(ctx):
    """64x64.64x64, degree 4095, p=2013265921: plain == MFMA, C(x)=A(x)B(x) at
    random points, and the last row of C against the oracle."""
    full_size_check(ctx, 2013265921, 64, 4096, 0x4C494E424F58 + 1)


def test_config3_full_size(ctx):
    """128x128.128x128, degree 8191, p=2013265921 (the headline shape)."""
    full_size_check(ctx, 2013265921, 128, 8192, 0x4C494E424F58 + 2)


# ------------------------------------------- edge cases of the product ---
def test_constant_matrices_with_max_rowdeg(ctx):
    """sa = sb = 1 with max_rowdeg = 3: the FFT domain's result, coefficient 0
    the product of the constants and coefficients 1..3 zero, whatever c held
    (host and device storage, c pre-filled with nonzero words)."""
    from linbox_amd.hip import DeviceBuffer, synchronize
    p, m, k, n, deg = 2013265921, 5, 6, 4, 3
    rng = np.random.default_rng(31)
    a, b = rand(rng, p, m * k), rand(rng, p, k * n)
    st, ref = orc.mul_fft_prime(p, m, k, n, a, 1, b, 1, max_rowdeg=deg)
    assert st == 0
    ref = ref.reshape(m * n, deg + 1)
    assert (ref[:, 1:] == 0).all()
    assert np.array_equal(ref[:, 0], orc.mul_naive(p, m, k, n, a, 1, b, 1))
    for method in (lbx.METHOD_AUTO, lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA):
        say("constant max_rowdeg host", method)
        C = PolynomialMatrix.from_polfirst(np.full((m * n, deg + 1), 0x5A5A5A5, dtype=np.uint32), m, n)
        PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, PolynomialMatrix.from_polfirst(a.reshape(-1, 1), m, k),
                                                         PolynomialMatrix.from_polfirst(b.reshape(-1, 1), k, n), deg)
        assert C.size() == deg + 1
        assert np.array_equal(C.to_polfirst(), ref)
        say("constant max_rowdeg device", method)
        da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
        dc = DeviceBuffer.from_numpy(np.full(m * n * (deg + 1), 0x77777777, dtype=np.uint32))
        C = PolynomialMatrix(m, n, deg + 1, data=dc)
        PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, PolynomialMatrix(m, k, 1, data=da),
                                                         PolynomialMatrix(k, n, 1, data=db), deg)
        synchronize()
        assert np.array_equal(dc.download().reshape(m * n, deg + 1), ref)


def test_two_streams_one_context(ctx):
    """Products queued back to back on two streams through one context (device
    pointers, no host synchronisation between them) share the context's
    workspace; the second stream waits for the first call's work, so both
    results are exact."""
    from linbox_amd.hip import DeviceBuffer, Stream
    p = 998244353
    cases = [(32, 32, 32, 1024, 1000), (32, 48, 16, 700, 1024), (40, 20, 24, 300, 500)]
    for method in METHODS:
        streams = [Stream(), Stream(), Stream()]
        bufs, refs = [], []
        for i, (m, k, n, sa, sb) in enumerate(cases):
            rng = np.random.default_rng(100 + i)
            a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
            _, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
            da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
            dc = DeviceBuffer(m * n * (sa + sb - 1) * 4)
            dc.zero()
            bufs.append((da, db, dc))
            refs.append(ref)
        say("two streams", method)
        for i, (m, k, n, sa, sb) in enumerate(cases):
            da, db, dc = bufs[i]
            PolynomialMatrixMulDomain(p, ctx, method, stream=streams[i].value).mul(
                PolynomialMatrix(m, n, sa + sb - 1, data=dc), PolynomialMatrix(m, k, sa, data=da),
                PolynomialMatrix(k, n, sb, data=db))
        for s in streams:
            s.synchronize()
        for i in range(len(cases)):
            assert np.array_equal(bufs[i][2].download(), refs[i]), (method, i)


def test_mfma_int32_chunk_boundary(ctx):
    """Inner dimensions above 4096 run the INT8 pointwise stage in several
    exact int32 chunks of 4096 inner products.  Random residues, and constant
    operands whose centred digits are at the extremes (-128 in the three low
    digits: the anti-diagonal D_2 grows by 3 * 2^14 = 49152 per inner product,
    so from k = 43691 on one int32 accumulation of the whole inner product
    would wrap), against the oracle."""
    p = 2013265921
    for (m, k, n, sa, sb) in [(2, 5000, 3, 16, 15), (16, 4200, 16, 8, 9), (2, 4500, 2, 200, 200),
                              (3, 8193, 2, 3, 3)]:
        rng = np.random.default_rng(k + m)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        say("int32 chunks random", m, k, n, sa, sb)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_INT8_MFMA)
        assert np.array_equal(got.reshape(-1), ref)
    # centred value with digits (-128, -128, -128, -59): c = -998277248, |c| < p/2
    c = -128 - 128 * 256 - 128 * 65536 - 59 * (1 << 24)
    assert abs(c) < p // 2
    x = c % p
    digits = (x + 0x80808080) ^ 0x80808080  # the kernels' digit split (of the centred value)
    for (m, k, n) in [(2, 48000, 2), (1, 45000, 1)]:
        # a(x) = x_const, b(x) = x_const + 0*x: both transforms are constant at every point
        a = np.full(m * k, x, dtype=np.uint32)
        b = np.zeros(k * n * 2, dtype=np.uint32)
        b[0::2] = x
        st, ref = orc.mul_fft_prime(p, m, k, n, a, 1, b, 2)
        assert st == 0
        assert ref.reshape(m * n, 2)[0, 0] == k * x * x % p
        say("int32 chunks extreme digits", m, k, n, hex(digits))
        got = gpu_mul(ctx, p, m, k, n, a, 1, b, 2, lbx.METHOD_INT8_MFMA)
        assert np.array_equal(got.reshape(-1), ref)


def test_mul_fft_seam_values(ctx):
    """lbx_polymat_mul_fft_u32 (mul_fft, wordsize-fast.inl:86-190) on padded
    polfirst buffers of pts coefficients: with operands filling all pts
    coefficients the result is the cyclic product mod x^pts - 1 (the oracle's
    product truncated by max_rowdeg = pts-1); with operands padded by zeros it
    is the plain product followed by zeros.  Host and device buffers, both forms."""
    from linbox_amd.hip import DeviceBuffer, synchronize
    L_ = lbx.lib()
    for (p, m, k, n, lpts) in [(998244353, 3, 4, 5, 6), (2013265921, 17, 16, 18, 10), (469762049, 2, 3, 2, 13)]:
        pts = 1 << lpts
        rng = np.random.default_rng(lpts)
        a, b = rand(rng, p, m * k * pts), rand(rng, p, k * n * pts)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, pts, b, pts, max_rowdeg=pts - 1)
        assert st == 0
        ha = a.reshape(m * k, pts).copy()
        hb = b.reshape(k * n, pts).copy()
        ha[:, pts // 2:] = 0
        hb[:, pts // 2:] = 0
        _, ref_pad = orc.mul_fft_prime(p, m, k, n, ha[:, :pts // 2].reshape(-1), pts // 2,
                                       hb[:, :pts // 2].reshape(-1), pts // 2)
        ref_pad = np.concatenate([ref_pad.reshape(m * n, pts - 1), np.zeros((m * n, 1), np.uint32)], axis=1)
        for method in METHODS:
            for (xa, xb, want) in [(a, b, ref), (ha.reshape(-1), hb.reshape(-1), ref_pad.reshape(-1))]:
                say("mul_fft seam host", p, m, k, n, lpts, method)
                c = np.full(m * n * pts, 3, dtype=np.uint32)
                st = L_.lbx_polymat_mul_fft_u32(ctx.handle, p, lpts, m, k, n, xa.ctypes.data, xb.ctypes.data,
                                                c.ctypes.data, method, None)
                assert st == 0
                assert np.array_equal(c, want)
                say("mul_fft seam device", p, m, k, n, lpts, method)
                da, db = DeviceBuffer.from_numpy(xa), DeviceBuffer.from_numpy(xb)
                dc = DeviceBuffer(m * n * pts * 4)
                st = L_.lbx_polymat_mul_fft_u32(ctx.handle, p, lpts, m, k, n, da.data_ptr(), db.data_ptr(),
                                                dc.data_ptr(), method, None)
                assert st == 0
                synchronize()
                assert np.array_equal(dc.download(), want)
                # the operands are left intact (the reference transforms its copies in place)
                assert np.array_equal(da.download(), xa) and np.array_equal(db.download(), xb)


def test_mul_plain_point_major(ctx, monkeypatch):
    """The plain form on point-major operands (split transforms writing
    canonical words, LDS-tiled 64-bit accumulation): output tiles of
    32..128 rows and columns, several tiles per point, inner dimensions that
    are not multiples of the 64-deep stage, head/tail chunking, and both
    operand layouts forced on the same shapes."""
    cases = [(2013265921, 33, 100, 70, 200, 57), (998244353, 128, 128, 128, 40, 41),
             (469762049, 200, 70, 40, 300, 260), (2013265921, 32, 16, 32, 1025, 1000),
             (167772161, 40, 130, 260, 129, 128), (2013265921, 64, 64, 64, 4096, 4000)]
    for (p, m, k, n, sa, sb) in cases:
        rng = np.random.default_rng(m * k * n + sa)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        for layout in ("point", "entry"):
            monkeypatch.setenv("LBX_PLAIN_LAYOUT", layout)
            say("plain point-major", layout, p, m, k, n, sa, sb)
            got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_PLAIN_U64)
            assert np.array_equal(got.reshape(-1), ref), layout
    monkeypatch.setenv("LBX_PLAIN_LAYOUT", "point")
    monkeypatch.setenv("LBX_CHUNK_BYTES", str(1 << 20))
    for (p, m, k, n, sa, sb) in [(2013265921, 37, 20, 45, 300, 200), (998244353, 5, 3, 7, 500, 400)]:
        rng = np.random.default_rng(m + n)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        say("plain point-major chunked", p, m, k, n, sa, sb)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_PLAIN_U64)
        assert np.array_equal(got.reshape(-1), ref)
