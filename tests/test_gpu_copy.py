"""GPU parity of the layout conversion (lbx_polymat_copy_u32,
PolynomialMatrix::copy(M, beg, end, start), polynomial-matrix.h:206-243,
508-532, 763-790, 988-992).  A copy has one right answer, written here with
numpy fancy indexing of the three storage formulas (checked word by word in
tests/test_abi.py); at the 128 x 128, 8192-coefficient operand size the
round trip polfirst -> matfirst -> matrowfirst -> polfirst must return the input.
"""
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, LinboxError
from linbox_amd.hip import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu
LAYOUTS = ["polfirst", "matfirst", "matrowfirst"]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def rand(rng, n):
    return rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("src_lay", LAYOUTS)
@pytest.mark.parametrize("dst_lay", LAYOUTS)
def test_all_layout_pairs(ctx, src_lay, dst_lay):
    rng = np.random.default_rng(LAYOUTS.index(src_lay) * 3 + LAYOUTS.index(dst_lay))
    for (r, c, ss, ds, beg, end, start) in [(3, 5, 100, 100, 0, 99, 0), (7, 2, 130, 70, 17, 80, 3),
                                            (1, 1, 5, 5, 2, 2, 4), (65, 66, 70, 200, 0, 69, 130),
                                            (9, 130, 64, 64, 0, 63, 0),
                                            # entry counts and polfirst offsets that are multiples of four take the
                                            # 16-byte kernel for polfirst <-> matfirst (with a coefficient tail here)
                                            (4, 6, 100, 120, 8, 97, 4), (8, 8, 64, 64, 0, 63, 0),
                                            (16, 20, 200, 136, 4, 134, 0), (4, 4, 70, 70, 2, 40, 6)]:
        say("copy", src_lay, dst_lay, r, c, ss, ds, beg, end, start)
        src = PolynomialMatrix.from_polfirst(rand(rng, r * c * ss).reshape(r * c, ss), r, c, src_lay)
        fill = rand(rng, r * c * ds).reshape(r * c, ds)
        dst = PolynomialMatrix.from_polfirst(fill, r, c, dst_lay)
        dst.copy(src, beg, end, start, ctx=ctx)
        want = fill.copy()
        want[:, start:start + end - beg + 1] = src.to_polfirst()[:, beg:end + 1]
        assert np.array_equal(dst.to_polfirst(), want)


def test_views(ctx):
    rng = np.random.default_rng(7)
    r, c = 6, 5
    big = PolynomialMatrix.from_polfirst(rand(rng, r * c * 90).reshape(r * c, 90), r, c, "matrowfirst")
    src_view = big.at(10, 59)  # 50 coefficients
    dst = PolynomialMatrix.from_polfirst(np.zeros((r * c, 80), np.uint32), r, c, "matfirst")
    dst_view = dst.at(20, 79)  # 60 coefficients
    say("views")
    dst_view.copy(src_view, ctx=ctx)  # copy(M): all 50 coefficients into the view's 0..49
    want = np.zeros((r * c, 80), np.uint32)
    want[:, 20:70] = big.to_polfirst()[:, 10:60]
    assert np.array_equal(dst.to_polfirst(), want)


def test_device_pointers_and_round_trip_at_operand_size(ctx):
    """128 x 128 entries of 8192 coefficients (one 512 MiB operand of the 128^2, degree 8191 product)."""
    r = c = 128
    s = 8192
    rng = np.random.default_rng(8192)
    host = rand(rng, r * c * s)
    say("round trip", r, c, s)
    bufs = [DeviceBuffer.from_numpy(host)] + [DeviceBuffer(r * c * s * 4) for _ in range(3)]
    mats = [PolynomialMatrix(r, c, s, lay, data=b) for lay, b in
            zip(["polfirst", "matfirst", "matrowfirst", "polfirst"], bufs)]
    for i in range(3):
        mats[i + 1].copy(mats[i], ctx=ctx)
    synchronize()
    mf = bufs[1].download()
    # matfirst word t*r*c + e holds coefficient t of entry e
    idx = rng.integers(0, r * c * s, 4096)
    e, t = idx // s, idx % s
    assert np.array_equal(mf[t * r * c + e], host[idx])
    mrf = bufs[2].download()
    i, j = e // c, e % c
    assert np.array_equal(mrf[i * c * s + j + t * c], host[idx])
    assert np.array_equal(bufs[3].download(), host)
    # matfirst -> polfirst directly (the 16-byte kernel's other direction)
    back = PolynomialMatrix(r, c, s, "polfirst", data=bufs[3])
    bufs[3].zero()
    back.copy(mats[1], ctx=ctx)
    synchronize()
    assert np.array_equal(bufs[3].download(), host)


def test_errors(ctx):
    a = PolynomialMatrix(3, 4, 10)
    b = PolynomialMatrix(3, 4, 10, "matfirst")
    cases = [
        (PolynomialMatrix(4, 3, 10), a, 0, 9, 0, 5),  # other dimensions
        (b, a, 0, 10, 0, 7),                          # end beyond the source
        (b, a, 5, 4, 0, 7),                           # beg > end
        (b, a, 0, 9, 1, 5),                           # destination too short
    ]
    for dst, src, beg, end, start, code in cases:
        with pytest.raises(LinboxError) as e:
            dst.copy(src, beg, end, start, ctx=ctx)
        assert e.value.status == code
    with pytest.raises(LinboxError) as e:  # one storage for both
        a.at(0, 4).copy(a.at(5, 9), ctx=ctx)
    assert e.value.status == 7


def test_many_entries(ctx):
    """300 x 300 entries (more than 2^16), both kernels: matrowfirst -> polfirst and polfirst -> matfirst."""
    rng = np.random.default_rng(300)
    r = c = 300
    s = 6
    host = rand(rng, r * c * s).reshape(r * c, s)
    say("many entries", r, c, s)
    src = PolynomialMatrix.from_polfirst(host, r, c, "matrowfirst")
    mid = PolynomialMatrix(r, c, s, "polfirst")
    mid.copy(src, ctx=ctx)
    assert np.array_equal(mid.to_polfirst(), host)
    out = PolynomialMatrix(r, c, s, "matfirst")
    out.copy(mid, 1, 4, 2, ctx=ctx)
    want = np.zeros_like(host)
    want[:, 2:6] = host[:, 1:5]
    assert np.array_equal(out.to_polfirst(), want)
