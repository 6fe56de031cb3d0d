"""Benchmark: bit-exact FFT polynomial-matrix products per second on one MI355X.

Workload (default): 128x128 . 128x128 polynomial matrices over p = 2013265921
(15*2^27+1), entries of degree 8191 with coefficients uniform in [0,p),
product of degree 16382 through length-16384 NTTs.  A "step" is one whole
product c = a*b (forward NTTs, pointwise stage, inverse NTTs) with a, b and c
resident in HBM.  `value` is products/s of the INT8-MFMA form; the plain
uint64 vector-ALU form is timed the same way for the ratio.

The JSON line also carries:
  roofline      for the kernel with the largest time per product: its algorithmic
                HBM bytes per launch divided by its average launch time, measured
                live with HIP events recorded on the launch stream during the
                timed steps; `traffic` is the measured HBM bytes per launch from
                the committed rocprofv3 PMC summary (profiles/hbm_traffic_128x128_deg8191.txt);
  kernels       the same figures for every kernel of the product;
  batch_1024_products 1024 uneven products as one batch call vs one call per product;
  order_basis   PM_Basis of a 64 x 32 matrix power series at order 4096;
  other_configs 16x16 deg 255, 64x64 deg 4095 (both forms), 32x32 deg 1023 over a
                192-bit modulus, a PM_Basis-level middle product, 64x64 deg 4095
                over a non-FFT prime below 2^29 (few-primes route, and the
                multiprecision route for comparison), and one 128x128 x 8192
                layout conversion;
  ntt_passes    the forward and inverse NTT phases against the workload's
                algorithmic bytes, 4(mk+kn)(sa+pts) and 4mn(pts+D+1), as a
                fraction of 8 TB/s;
  int8_mfma_utilisation  32 pts mkn INT8 operations of the pointwise kernel over
                its time, as a fraction of the 5.0 POPS dense INT8 peak;
  self_check    the last timed product (each form) checked at two random points
                and on 64 random coefficients computed exactly; a mismatch stops
                the run with an error;
  host_operands the same product with numpy operands through the C ABI, PCIe
                copies included (never `value`);
  cpu_baseline  the oracle's restatement of the reference's single-prime
                mul_fft timed on this host on a bounded sample of the same
                workload and extrapolated to one product, per stage, on one
                thread and on up to 16; the host's nproc and CPU model; and the
                16x16 degree-255 configuration timed whole on one thread.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--skip-batch]
                       [--skip-order-basis] [--skip-configs] [--skip-host]
                       [--only order_basis|configs|batch]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P = 2013265921
M = K = N = 128
SA = SB = 8192
HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0  # dense INT8 MFMA, MI355X


def kernel_bytes(m, k, n, sa, sb, pts):
    """Algorithmic HBM bytes of one launch of each kernel of the product: every
    word the launch must read or write once (4 bytes per coefficient, residue or
    4-digit limb group).  With the default chunk size each kernel runs once per
    operand per product at this shape."""
    deg = sa + sb - 2
    return {
        "kernel dif_head a": 4 * m * k * (sa + pts),        # polynomial in, 2^14 words after 8 stages out
        "kernel dif_head b": 4 * k * n * (sb + pts),
        "kernel dif_tail_limbs a": 4 * m * k * pts * 2,     # 2^14 words in, 2^14 limb groups out
        "kernel dif_tail_limbs b": 4 * k * n * pts * 2,
        "kernel pw_mfma": 4 * (m * k + k * n + m * n) * pts,  # per point: A, B limb blocks in, C out
        "kernel dit_head": 4 * m * n * pts * 2,             # point-major C in, 6 stages, entry blocks out
        "kernel dit_tail": 4 * m * n * (pts + deg + 1),     # 2^14 words in, deg+1 coefficients out
    }


def traffic_table():
    """Measured HBM bytes per launch, from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a device-to-device copy of known size)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_128x128_deg8191.txt")
    table = {}
    if os.path.exists(path):
        for line in open(path):  # rows "| kernel <label> | ... | <bytes per launch> |"
            if line.startswith("| kernel "):
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                try:
                    table[cells[0]] = int(cells[-1])
                except ValueError:
                    pass
    return table


def kernel_roofline(times, bytes_per_launch):
    """Roofline of the kernel with the largest time per product (one launch per
    label per product): achieved = algorithmic bytes / average launch time."""
    t = {label: ms for label, ms in times if label in bytes_per_launch}
    if not t:
        return None, {}
    traffic = traffic_table()
    kernels = {label: {"ms": round(ms, 4),
                       "algorithmic_bytes": bytes_per_launch[label],
                       "achieved_gbs": round(bytes_per_launch[label] / (ms * 1e-3) / 1e9, 1),
                       "traffic": traffic.get(label)} for label, ms in t.items()}
    dom = max(t, key=lambda label: t[label])
    achieved = bytes_per_launch[dom] / (t[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic.get(dom),
                "algorithmic_bytes": bytes_per_launch[dom], "launch_ms": round(t[dom], 4)}
    return roofline, kernels


def run_gpu(args):
    import linbox_amd as lbx
    from linbox_amd import hip, PolynomialMatrix, PolynomialMatrixMulDomain

    ctx = lbx.Context(0)
    rng = np.random.default_rng(0x4C494E424F58 + 2)
    a = rng.integers(0, P, M * K * SA, dtype=np.uint32)
    b = rng.integers(0, P, K * N * SB, dtype=np.uint32)
    sc = SA + SB - 1
    da, db = hip.DeviceBuffer.from_numpy(a), hip.DeviceBuffer.from_numpy(b)
    dc = hip.DeviceBuffer(M * N * sc * 4)
    stream = hip.Stream()
    A, B = PolynomialMatrix(M, K, SA, data=da), PolynomialMatrix(K, N, SB, data=db)
    C = PolynomialMatrix(M, N, sc, data=dc)

    def timed(method, steps, warmup):
        dom = PolynomialMatrixMulDomain(P, ctx, method, stream=stream.value)
        for _ in range(warmup):
            dom.mul(C, A, B)
        stream.synchronize()
        hip.synchronize()
        ctx.profile(2)  # stage and per-kernel events on the launch streams
        ctx.stage_times()  # drop anything recorded before the timed region
        e0, e1 = hip.Event(), hip.Event()
        t0 = time.perf_counter()
        e0.record(stream.value)
        for _ in range(steps):
            dom.mul(C, A, B)
        e1.record(stream.value)
        e1.synchronize()
        wall = time.perf_counter() - t0
        hip.synchronize()
        ctx.profile(False)
        stages = ctx.stage_times()
        return e0.elapsed_ms(e1) / steps, wall, stages

    ms_int8, wall_int8, stages_int8 = timed(lbx.METHOD_INT8_MFMA, args.steps, args.warmup)
    print("int8 form: %.3f ms/product" % ms_int8, file=sys.stderr, flush=True)
    check = self_check(a, b, dc.download(), rng)
    print("self-check of the last timed product: %s" % check["result"], file=sys.stderr, flush=True)
    if check["result"] != "exact":
        raise SystemExit("bench: the timed product is wrong (%s)" % check)
    plain_steps = max(2, args.steps // 4)
    ms_plain, _, stages_plain = timed(lbx.METHOD_PLAIN_U64, plain_steps, 1)
    print("plain form: %.3f ms/product" % ms_plain, file=sys.stderr, flush=True)
    plain_ok = self_check(a, b, dc.download(), rng)
    if plain_ok["result"] != "exact":
        raise SystemExit("bench: the plain-form product is wrong (%s)" % plain_ok)
    da.free(), db.free(), dc.free()
    host = None if args.skip_host else host_leg(ctx, lbx, a, b)
    stream.destroy()
    ctx.close()
    return ms_int8, wall_int8, stages_int8, ms_plain, stages_plain, check, host


def self_check(a, b, c, rng, points=2, coeffs=64):
    """Check the 128x128 . 128x128 product c = a.b left by the timed loop:
    C(x) = A(x) B(x) mod p at `points` random x (Schwartz-Zippel: a wrong
    coefficient passes with probability <= deg/p per point), and `coeffs`
    random coefficients c[i][j][t] = sum_q sum_u a[i][q][u] b[q][j][t-u] mod p
    computed exactly (16-bit split of a so no 64-bit sum overflows)."""
    sc = SA + SB - 1
    A = a.reshape(M * K, SA)
    B = b.reshape(K * N, SB)
    C = c.reshape(M * N, sc)
    pp = np.uint64(P)

    def horner(X, x):
        acc = np.zeros(X.shape[0], dtype=np.uint64)
        xx = np.uint64(x)
        for t in range(X.shape[1] - 1, -1, -1):
            acc = (acc * xx + X[:, t]) % pp
        return acc

    def matmul_mod(X, Y):
        lo, hi = Y & np.uint64(0xFFFF), Y >> np.uint64(16)
        return ((X @ hi) % pp * np.uint64(1 << 16) % pp + (X @ lo) % pp) % pp

    for x in rng.integers(2, P, points):
        Ax = horner(A, x).reshape(M, K)
        Bx = horner(B, x).reshape(K, N)
        Cx = horner(C, x).reshape(M, N)
        if not np.array_equal(matmul_mod(Ax, Bx), Cx):
            return {"result": "mismatch at x=%d" % int(x)}
    alo = (A & 0xFFFF).astype(np.uint64)
    ahi = (A >> 16).astype(np.uint64)
    Bu = B.astype(np.uint64)
    for _ in range(coeffs):
        i, j, t = int(rng.integers(M)), int(rng.integers(N)), int(rng.integers(sc))
        u0, u1 = max(0, t - (SB - 1)), min(t, SA - 1)
        total = 0
        for q in range(K):
            av_lo, av_hi = alo[i * K + q, u0:u1 + 1], ahi[i * K + q, u0:u1 + 1]
            bv = Bu[q * N + j, t - u1:t - u0 + 1][::-1]
            total += int(av_lo @ bv) + (int(av_hi @ bv) << 16)
        if total % P != int(C[i * N + j, t]):
            return {"result": "mismatch at c[%d][%d][%d]" % (i, j, t)}
    return {"result": "exact", "points": points, "coefficients": coeffs}


def host_leg(ctx, lbx, a, b, reps=2):
    """The same product with host (numpy) operands through the C ABI, as a
    LinBox caller's matrices are: the call stages a and b into device buffers
    kept in the context, computes, and copies c back (1 GiB in, 1 GiB out over
    PCIe).  Wall time per call, PCIe included; never the headline value."""
    from linbox_amd import PolynomialMatrix, PolynomialMatrixMulDomain
    A = PolynomialMatrix.from_polfirst(a.reshape(M * K, SA), M, K)
    B = PolynomialMatrix.from_polfirst(b.reshape(K * N, SB), K, N)
    C = PolynomialMatrix(M, N, SA + SB - 1)
    dom = PolynomialMatrixMulDomain(P, ctx, lbx.METHOD_INT8_MFMA)
    dom.mul(C, A, B)  # staging buffers
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        dom.mul(C, A, B)
        times.append(time.perf_counter() - t0)
    ms = min(times) * 1e3
    print("host operands: %.1f ms per product (PCIe included)" % ms, file=sys.stderr, flush=True)
    return {"ms_per_product": round(ms, 2), "products_per_s": round(1e3 / ms, 2), "reps": reps,
            "bytes_over_pcie": int(a.nbytes + b.nbytes + C.data.nbytes),
            "note": "numpy operands through lbx_polymat_mul_u32 (INT8 form): copies in and out included"}


def run_batch(steps=5):
    """The 1024-product batch: independent products (m,k,n in {2,...,64}, degree log-uniform in
    [8,4096], prime uniform from four FFT primes) as one lbx_polymat_mul_batch call
    (replayed hipGraph of persistent work-queue kernels) vs one lbx_polymat_mul_u32 per product."""
    import linbox_amd as lbx
    from linbox_amd import hip, PolynomialMatrix, PolynomialMatrixMulDomain

    primes = [998244353, 2013265921, 469762049, 167772161]
    rng = np.random.default_rng(0x4C494E424F58 + 4)
    dims = [2, 4, 8, 16, 32, 64]
    ctx = lbx.Context(0)
    stream = hip.Stream()
    jobs, keep = [], []
    for _ in range(1024):
        m, k, n = (int(rng.choice(dims)) for _ in range(3))
        d = int(round(np.exp(rng.uniform(np.log(8), np.log(4096)))))
        p = int(rng.choice(primes))
        sa = sb = d + 1
        da = hip.DeviceBuffer.from_numpy(rng.integers(0, p, m * k * sa, dtype=np.uint32))
        db = hip.DeviceBuffer.from_numpy(rng.integers(0, p, k * n * sb, dtype=np.uint32))
        dc = hip.DeviceBuffer(m * n * (sa + sb - 1) * 4)
        keep += [da, db, dc]
        jobs.append(dict(p=p, m=m, k=k, n=n, sa=sa, sb=sb, a=da.data_ptr(), b=db.data_ptr(), c=dc.data_ptr()))

    def timed(fn, reps):
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream.value)
        for _ in range(reps):
            fn()
        e1.record(stream.value)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    batch = lambda: lbx.mul_batch(jobs, ctx=ctx, stream=stream.value)
    batch()  # builds the plan, tunes tiles and forms, captures the graph
    stream.synchronize()
    ms_batch = timed(batch, steps)
    tuning = json.loads(lbx.lib().lbx_batch_tuning(ctx.handle).decode())
    # share of the batch's multiply-adds (pts m k n per job) done on INT8 MFMA
    macs, macs_mfma = 0, 0
    for j in jobs:
        deg = j["sa"] + j["sb"] - 2
        pts = 1 << max(1, deg.bit_length())
        mac = pts * j["m"] * j["k"] * j["n"]
        macs += mac
        d = min(j["m"], j["k"], j["n"])
        cls = None if d < 16 or j["k"] > 4096 else ("min(m,k,n) in [16,32)" if d < 32 else "min(m,k,n)>=32")
        choice = tuning["int8_capable"].get(cls) if cls else None
        if choice and choice.get("form") == "int8_mfma" and 4 <= deg.bit_length() <= 13:
            macs_mfma += mac

    def singles():
        for j in jobs:
            C = PolynomialMatrix(j["m"], j["n"], j["sa"] + j["sb"] - 1, data=_Ptr(j["c"]))
            PolynomialMatrixMulDomain(j["p"], ctx, lbx.METHOD_AUTO, stream=stream.value).mul(
                C, PolynomialMatrix(j["m"], j["k"], j["sa"], data=_Ptr(j["a"])),
                PolynomialMatrix(j["k"], j["n"], j["sb"], data=_Ptr(j["b"])))
    singles()
    stream.synchronize()
    ms_single = timed(singles, 1)
    stream.destroy()
    ctx.close()
    print("batch: %.3f ms per 1024 products (graph), %.3f ms with one call per product" % (ms_batch, ms_single),
          file=sys.stderr, flush=True)
    return {"workload": "1024 products, m,k,n in {2..64}, degree log-uniform [8,4096], 4 primes",
            "graph_ms": round(ms_batch, 3), "graph_products_per_s": round(1024 / (ms_batch * 1e-3), 1),
            "per_product_ms": round(ms_single, 3),
            "per_product_products_per_s": round(1024 / (ms_single * 1e-3), 1),
            "speedup": round(ms_single / ms_batch, 2), "tuning": tuning,
            "mac_share_on_int8_mfma": round(macs_mfma / macs, 4)}


def staged_bytes(m, k, n, sa, sb):
    """Algorithmic HBM bytes of one staged product (forward, pointwise, inverse passes)."""
    D = sa + sb - 2
    pts = 1
    while pts < D + 1:
        pts <<= 1
    return 4 * (m * k + k * n) * (sa + pts) + 4 * (m * k + k * n + m * n) * pts + 4 * m * n * (pts + D + 1)


def run_configs(reps=10):
    """The other workloads, timed with HIP events on one stream:
    16x16 deg 255 (p = 998244353), 64x64 deg 4095 (p = 2013265921, both pointwise forms),
    32x32 deg 1023 over a 192-bit P (14 FFT primes and CRT), the middle product of a
    PM_Basis level (64x64 sigma of degree 2048 times a 64x32 serie of 4096 terms), 64x64
    deg 4095 over the non-FFT prime 268435399 through the few-primes route and, for
    comparison, through the multiprecision route (the same modulus given as two limbs), and
    the layout conversion of one 128x128 operand of 8192 coefficients (polfirst -> matfirst)."""
    import linbox_amd as lbx
    from linbox_amd import hip, PolynomialMatrix, PolynomialMatrixMulDomain, PolynomialMatrixFFTMulDomain

    ctx = lbx.Context(0)
    stream = hip.Stream()
    rng = np.random.default_rng(0x4C494E424F58 + 7)

    def timed(fn):
        fn()
        stream.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream.value)
        for _ in range(reps):
            fn()
        e1.record(stream.value)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    def dev(words):
        return hip.DeviceBuffer.from_numpy(words)

    out = {}
    for name, p, dim, deg, methods in [("16x16_deg255", 998244353, 16, 255, [lbx.METHOD_AUTO]),
                                        ("64x64_deg4095", 2013265921, 64, 4095,
                                         [lbx.METHOD_INT8_MFMA, lbx.METHOD_PLAIN_U64])]:
        sa = deg + 1
        da = dev(rng.integers(0, p, dim * dim * sa, dtype=np.uint64).astype(np.uint32))
        db = dev(rng.integers(0, p, dim * dim * sa, dtype=np.uint64).astype(np.uint32))
        dc = hip.DeviceBuffer(dim * dim * (2 * sa - 1) * 4)
        A, B = PolynomialMatrix(dim, dim, sa, data=da), PolynomialMatrix(dim, dim, sa, data=db)
        C = PolynomialMatrix(dim, dim, 2 * sa - 1, data=dc)
        for meth in methods:
            dom = PolynomialMatrixMulDomain(p, ctx, meth, stream=stream.value)
            ms = timed(lambda: dom.mul(C, A, B))
            label = {lbx.METHOD_AUTO: "auto", lbx.METHOD_INT8_MFMA: "int8_mfma", lbx.METHOD_PLAIN_U64: "plain_u64"}[meth]
            gbs = staged_bytes(dim, dim, dim, sa, sa) / (ms * 1e-3) / 1e9
            out["%s_%s" % (name, label)] = {"ms": round(ms, 4), "products_per_s": round(1e3 / ms, 1),
                                            "staged_gbs": round(gbs, 1), "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 3)}
    # 192-bit modulus through 14 FFT primes (fixture P of the CRT tests)
    with open(os.path.join(ROOT, "tests", "golden", "crt_mul.json")) as f:
        crt = json.load(f)
    Pm, nl = int(crt["P_hex"], 16), crt["nlimbs"]
    top = (Pm >> (32 * (nl - 1))) & 0xFFFFFFFF

    def rand_mp(count):
        w = rng.integers(0, 1 << 32, (count, nl), dtype=np.uint64)
        w[:, nl - 1] = rng.integers(0, top, count, dtype=np.uint64)  # below P
        return w.astype(np.uint32).reshape(-1)
    dim, sa = 32, 1024
    da, db = dev(rand_mp(dim * dim * sa)), dev(rand_mp(dim * dim * sa))
    dc = hip.DeviceBuffer(dim * dim * (2 * sa - 1) * nl * 4)
    A = PolynomialMatrix(dim, dim, sa, data=da, nlimbs=nl)
    B = PolynomialMatrix(dim, dim, sa, data=db, nlimbs=nl)
    C = PolynomialMatrix(dim, dim, 2 * sa - 1, data=dc, nlimbs=nl)
    ms = timed(lambda: lbx.mul_mp(Pm, C, A, B, ctx=ctx, stream=stream.value))
    out["32x32_deg1023_P192"] = {"ms": round(ms, 4), "products_per_s": round(1e3 / ms, 1), "fft_primes": 14,
                                 "pointwise": "auto: plain uint64 form, the 14 per-prime products as one call "
                                              "of the batch kernels"}
    ms = timed(lambda: lbx.mul_mp(Pm, C, A, B, ctx=ctx, method=lbx.METHOD_INT8_MFMA, stream=stream.value))
    out["32x32_deg1023_P192_int8"] = {"ms": round(ms, 4), "products_per_s": round(1e3 / ms, 1), "fft_primes": 14,
                                      "pointwise": "int8_mfma per prime (14 single-prime products)"}
    # middle product of a PM_Basis level: coefficients 2048..4095 of sigma1.serie
    p, m, n, o1, order = 998244353, 64, 32, 2048, 4096
    ds = dev(rng.integers(0, p, m * m * (o1 + 1), dtype=np.uint64).astype(np.uint32))
    dz = dev(rng.integers(0, p, m * n * order, dtype=np.uint64).astype(np.uint32))
    dc = hip.DeviceBuffer(m * n * (order - o1) * 4)
    S, Z = PolynomialMatrix(m, m, o1 + 1, data=ds), PolynomialMatrix(m, n, order, data=dz)
    C = PolynomialMatrix(m, n, order - o1, data=dc)
    dom = PolynomialMatrixFFTMulDomain(p, ctx, lbx.METHOD_AUTO, stream=stream.value)
    ms = timed(lambda: dom.midproduct(C, S, Z, True, o1 + 1, order))
    pts = order
    mbytes = (4 * (m * m * (o1 + 1 + pts) + m * n * (order + pts)) + 4 * (m * m + 2 * m * n) * pts +
              4 * m * n * (pts + order - o1))
    out["midproduct_64x64.64x32_order4096"] = {"ms": round(ms, 4), "staged_gbs": round(mbytes / (ms * 1e-3) / 1e9, 1)}
    # few-primes route (p < 2^29, not an FFT prime): three 31-bit FFT primes, mixed radix mod p
    p, dim, sa = 268435399, 64, 4096
    ha = rng.integers(0, p, dim * dim * sa, dtype=np.uint64).astype(np.uint32)
    hb = rng.integers(0, p, dim * dim * sa, dtype=np.uint64).astype(np.uint32)
    da, db, dc = dev(ha), dev(hb), hip.DeviceBuffer(dim * dim * (2 * sa - 1) * 4)
    A, B = PolynomialMatrix(dim, dim, sa, data=da), PolynomialMatrix(dim, dim, sa, data=db)
    C = PolynomialMatrix(dim, dim, 2 * sa - 1, data=dc)
    dom = PolynomialMatrixFFTMulDomain(p, ctx, lbx.METHOD_AUTO, stream=stream.value)
    ms = timed(lambda: dom.mul(C, A, B))
    out["64x64_deg4095_p268435399_few_primes"] = {"ms": round(ms, 4), "products_per_s": round(1e3 / ms, 1),
                                                  "fft_primes": 3}
    del da, db, dc
    two = np.zeros((dim * dim * sa, 2), np.uint32)
    two[:, 0] = ha
    da = dev(two.reshape(-1))
    two[:, 0] = hb
    db = dev(two.reshape(-1))
    del two
    dc = hip.DeviceBuffer(dim * dim * (2 * sa - 1) * 2 * 4)
    A = PolynomialMatrix(dim, dim, sa, data=da, nlimbs=2)
    B = PolynomialMatrix(dim, dim, sa, data=db, nlimbs=2)
    C = PolynomialMatrix(dim, dim, 2 * sa - 1, data=dc, nlimbs=2)
    ms = timed(lambda: lbx.mul_mp(p, C, A, B, ctx=ctx, stream=stream.value))
    out["64x64_deg4095_p268435399_multiprecision_route"] = {"ms": round(ms, 4), "products_per_s": round(1e3 / ms, 1)}
    del da, db, dc, ha, hb
    # layout conversion (PolynomialMatrix::copy) of one 128x128 operand of 8192 coefficients
    dim, sa = 128, 8192
    da = dev(rng.integers(0, 1 << 32, dim * dim * sa, dtype=np.uint64).astype(np.uint32))
    db = hip.DeviceBuffer(dim * dim * sa * 4)
    A = PolynomialMatrix(dim, dim, sa, data=da)
    B = PolynomialMatrix(dim, dim, sa, "matfirst", data=db)
    ms = timed(lambda: B.copy(A, ctx=ctx, stream=stream.value))
    gbs = 2 * 4 * dim * dim * sa / (ms * 1e-3) / 1e9
    out["copy_polfirst_to_matfirst_128x128x8192"] = {"ms": round(ms, 4), "gbs": round(gbs, 1),
                                                     "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 3)}
    del da, db
    stream.destroy()
    ctx.close()
    print("configs: " + ", ".join("%s %.3f ms" % (k, v["ms"]) for k, v in out.items()), file=sys.stderr, flush=True)
    return out


def run_order_basis(skip_cpu, reps=3):
    """PM_Basis of a 64 x 32 serie (the shape of benchmark-order-basis.C:340-346) at order
    4096 over p = 998244353, device-resident; the call returns when the basis is done, so
    wall time per call is the figure.  One M_Basis base case (order 32) is timed too, for
    the share of the 128 base cases.  With the CPU leg, the oracle's PM_Basis and the
    device's are timed at order 128, where the oracle finishes in seconds."""
    import linbox_amd as lbx
    from linbox_amd import hip, PolynomialMatrix, OrderBasis

    p, m, n, order = 998244353, 64, 32, 4096
    rng = np.random.default_rng(0x4C494E424F58 + 6)
    serie = rng.integers(0, p, m * n * order, dtype=np.uint64).astype(np.uint32)
    ctx = lbx.Context(0)
    dser = hip.DeviceBuffer.from_numpy(serie)
    dsig = hip.DeviceBuffer(m * m * (order + 1) * 4)
    ob = OrderBasis(p, ctx)

    def call(fn, o):
        sh = [0] * m
        t0 = time.perf_counter()
        d = fn(PolynomialMatrix(m, m, o + 1, data=dsig), PolynomialMatrix(m, n, o, data=dser), o, sh)
        return time.perf_counter() - t0, d, sh

    def call_sub(ob, o, dsub):
        t0 = time.perf_counter()
        ob.PM_Basis(PolynomialMatrix(m, m, o + 1, data=dsig), PolynomialMatrix(m, n, o, data=dsub), o, [0] * m)
        return time.perf_counter() - t0

    call(ob.PM_Basis, order)  # workspace, twiddles
    runs = [call(ob.PM_Basis, order) for _ in range(reps)]
    t_pm = min(r[0] for r in runs)
    d, sh = runs[0][1], runs[0][2]
    call(ob.M_Basis, 32)
    t_mb = min(call(ob.M_Basis, 32)[0] for _ in range(reps))
    out = {"workload": "PM_Basis, 64x32 serie, order 4096, p=998244353 (M_Basis below order 32)",
           "ms": round(t_pm * 1e3, 2), "bases_per_s": round(1.0 / t_pm, 2), "degree": d,
           "shift_sum": int(sum(sh)), "m_basis_order32_ms": round(t_mb * 1e3, 3),
           "base_cases": order // 32, "base_case_share": round(order // 32 * t_mb / t_pm, 3)}
    if not skip_cpu:
        from oracle import oracle as orc
        o = 128
        sub = serie.reshape(m * n, order)[:, :o].copy().reshape(-1)
        dsub = hip.DeviceBuffer.from_numpy(sub)
        t_gpu = min(call_sub(ob, o, dsub) for _ in range(reps))
        t0 = time.perf_counter()
        st = orc.pm_basis(p, m, n, sub, o, o, [0] * m)[0]
        t_cpu = time.perf_counter() - t0
        assert st == 0
        out["order128"] = {"gpu_ms": round(t_gpu * 1e3, 2), "cpu_oracle_s": round(t_cpu, 2), "cores": 1,
                           "kind": "port"}
    ctx.close()
    print("order basis: %.2f ms per PM_Basis (64x32, order 4096); M_Basis order 32: %.3f ms"
          % (t_pm * 1e3, t_mb * 1e3), file=sys.stderr, flush=True)
    return out


class _Ptr:
    """A device address usable as PolynomialMatrix storage."""

    def __init__(self, ptr):
        self.ptr = ptr

    def data_ptr(self):
        return self.ptr


def copy_probe(mib):
    """Three device-to-device copies of a known size: the runtime's blit kernel
    reads and writes exactly mib MiB each, which calibrates the HBM byte counters."""
    from linbox_amd import hip
    nbytes = mib << 20
    src, dst = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
    src.zero()
    for _ in range(3):
        hip._ok(hip.runtime().hipMemcpy(dst.ptr, src.ptr, nbytes, hip.D2D), "hipMemcpy D2D")
    hip.synchronize()
    print("copied %d MiB x3" % mib, file=sys.stderr, flush=True)


def cpu_baseline(budget_s):
    """Oracle (C restatement of mul_fft, one thread) on a bounded sample of the workload.

    Times forward DIF transforms of 2^14 points, the pointwise 128x128x128
    product at single points and inverse DIT transforms, each repeated until
    budget_s/3 seconds pass, then extrapolates to one product:
      t = (m*k + k*n) t_fwd + pts t_point + m*n t_inv.
    """
    from oracle import oracle as orc
    L = 14
    pts = 1 << L
    w = orc.primitive_root(P, L)
    pw, _ = orc.init_powers(P, L, w)
    wi = orc.invroot(P, L, w)
    pwi, _ = orc.init_powers(P, L, wi)
    rng = np.random.default_rng(7)

    def per_item(fn, make, chunk):
        done, t = 0, 0.0
        while t < budget_s / 3:
            data = make(chunk)
            t0 = time.perf_counter()
            fn(data, chunk)
            t += time.perf_counter() - t0
            done += chunk
        return t / done, done

    polys = lambda c: rng.integers(0, P, c * pts, dtype=np.uint32)
    t_fwd, n_fwd = per_item(lambda d, c: orc.stage_forward(P, L, pw, d, c), polys, 64)
    t_inv, n_inv = per_item(lambda d, c: orc.stage_inverse(P, L, pwi, d, c), polys, 64)
    npt = 8  # points per timing chunk; operands hold all 2^14 points, only npt are multiplied
    ahat = rng.integers(0, P, M * K * npt, dtype=np.uint32)
    bhat = rng.integers(0, P, K * N * npt, dtype=np.uint32)
    chat = np.zeros(M * N * npt, dtype=np.uint32)
    t_pt, n_pt = per_item(lambda d, c: orc.stage_pointwise(P, M, K, N, npt, ahat, bhat, chat, 0, c),
                          lambda c: None, npt)
    t_prod = (M * K + K * N) * t_fwd + pts * t_pt + M * N * t_inv
    sample = ("oracle mul_fft stages on one thread: %d forward and %d inverse 2^14-point NTTs, "
              "%d pointwise 128x128x128 products mod p; extrapolated to 32768 forward NTTs, "
              "16384 points and 16384 inverse NTTs (one product)" % (n_fwd, n_inv, n_pt))
    # the same stages over the host's cores (OpenMP over polynomials and points), half the budget;
    # at most 16 threads, the host CPU share of one GPU on the measurement boxes
    threads = max(1, min(16, os.cpu_count() or 1))
    mt_budget = budget_s / 2

    def per_item_mt(fn, make, chunk):
        done, t = 0, 0.0
        while t < mt_budget / 3:
            data = make(chunk)
            t0 = time.perf_counter()
            fn(data, chunk)
            t += time.perf_counter() - t0
            done += chunk
        return t / done, done

    mt_f, nf = per_item_mt(lambda d, c: orc.stage_forward_mt(P, L, pw, d, c, threads), polys, 16 * threads)
    mt_i, ni = per_item_mt(lambda d, c: orc.stage_inverse_mt(P, L, pwi, d, c, threads), polys, 16 * threads)
    npm = 4 * threads
    ahm = rng.integers(0, P, M * K * npm, dtype=np.uint32)
    bhm = rng.integers(0, P, K * N * npm, dtype=np.uint32)
    chm = np.zeros(M * N * npm, dtype=np.uint32)
    mt_p, npx = per_item_mt(lambda d, c: orc.stage_pointwise_mt(P, M, K, N, npm, ahm, bhm, chm, 0, c, threads),
                            lambda c: None, npm)
    t_mt = (M * K + K * N) * mt_f + pts * mt_p + M * N * mt_i
    mt = {"value": 1.0 / t_mt, "unit": "products/s", "cores": threads, "kind": "port", "s_per_product": round(t_mt, 2),
          "sample": "the same stages on %d OpenMP threads: %d forward and %d inverse NTTs, %d points, extrapolated "
                    "to one product" % (threads, nf, ni, npx),
          "thread_cap": "16 (the host CPU share of one GPU on the measurement boxes)",
          "stages_s": {"direct FFT_DIF": round((M * K + K * N) * mt_f, 3), "Pointwise mult": round(pts * mt_p, 3),
                       "inverse FFT_DIT": round(M * N * mt_i, 3)}}
    stages = {"direct FFT_DIF": round((M * K + K * N) * t_fwd, 3), "Pointwise mult": round(pts * t_pt, 3),
              "inverse FFT_DIT": round(M * N * t_inv, 3)}
    return 1.0 / t_prod, sample, t_prod, mt, stages


def host_info():
    """nproc, the CPU model and the thread count the CPU legs may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_config1():
    """The 16x16 . 16x16, degree 255 product over 998244353 (the reference's own
    CPU-runnable configuration) through the oracle's mul_fft restatement on one
    thread, repeated until at least one second has passed, mean time per product
    (the timing loop of benchmark-polynomial-matrix-mul-fft.C:94-97)."""
    from oracle import oracle as orc
    p, m, sa = 998244353, 16, 256
    a = orc.fill_uniform(0x4C494E424F58 + 0, p, m * m * sa)
    b = orc.fill_uniform(0x4C494E424F58 + 100, p, m * m * sa)
    reps, t = 0, 0.0
    while t < 1.0:
        t0 = time.perf_counter()
        st, _ = orc.mul_fft_prime(p, m, m, m, a, sa, b, sa)
        t += time.perf_counter() - t0
        reps += 1
        assert st == 0
    return {"workload": "16x16.16x16 deg 255 p=998244353", "ms_per_product": round(t / reps * 1e3, 3),
            "products_per_s": round(reps / t, 2), "reps": reps, "cores": 1, "kind": "port"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the first products of a process run slower while the clocks ramp up (about
    # five products at 128x128, degree 8191); ten untimed ones reach the steady state
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--skip-batch", action="store_true", help="omit the 1024-product batch measurement")
    ap.add_argument("--skip-order-basis", action="store_true", help="omit the PM_Basis measurement")
    ap.add_argument("--skip-configs", action="store_true", help="omit the other workload configurations")
    ap.add_argument("--skip-host", action="store_true", help="omit the host-operand (PCIe-inclusive) leg")
    ap.add_argument("--only", choices=["order_basis", "configs", "batch"],
                    help="run only this leg and print its JSON object (for profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run three device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: products are not sharded across GPUs, so --gpus must be 1",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.only:
        leg = {"order_basis": lambda: run_order_basis(args.skip_cpu), "configs": run_configs,
               "batch": run_batch}[args.only]
        print(json.dumps({args.only: leg()}))
        return
    ms_int8, wall, stages, ms_plain, stages_plain, check, host = run_gpu(args)
    batch = None if args.skip_batch else run_batch()
    order_basis = None if args.skip_order_basis else run_order_basis(args.skip_cpu)
    configs = None if args.skip_configs else run_configs()
    pts = 1 << 14
    roofline, kernels = kernel_roofline(stages, kernel_bytes(M, K, N, SA, SB, pts))
    if args.skip_cpu:
        cpu = None
    else:
        v, sample, t_prod, mt, cstages = cpu_baseline(args.cpu_budget)
        cpu = {"value": v, "unit": "products/s", "cores": 1, "kind": "port", "sample": sample,
               "s_per_product": round(t_prod, 2), "stages_s": cstages, "host": host_info(), "all_cores": mt,
               "config_16x16_deg255": cpu_config1()}
    value = 1000.0 / ms_int8
    stage = {k: v for k, v in stages}
    # the NTT passes against the workload table's bytes: forward 4 (mk+kn)(sa+pts) over the
    # "direct FFT_DIF" stage, inverse 4 mn (pts+D+1) over "inverse FFT_DIT"
    deg = SA + SB - 2
    ntt = {}
    for name, label, nbytes in [("forward", "direct FFT_DIF", 4 * (M * K + K * N) * (SA + pts)),
                                ("inverse", "inverse FFT_DIT", 4 * M * N * (pts + deg + 1))]:
        if label in stage:
            gbs = nbytes / (stage[label] * 1e-3) / 1e9
            ntt[name] = {"stage": label, "bytes": nbytes, "ms": round(stage[label], 4), "gbs": round(gbs, 1),
                         "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 4)}
    # INT8 MFMA utilisation of the pointwise stage: 32 pts mkn int8 ops (16 digit-pair products
    # of k-deep dot products, 2 ops each) over its time, against the dense INT8 peak
    mfma = None
    if "kernel pw_mfma" in stage:
        ops = 32 * pts * M * K * N
        tops = ops / (stage["kernel pw_mfma"] * 1e-3) / 1e12
        mfma = {"kernel": "pw_mfma", "int8_ops": ops, "ms": round(stage["kernel pw_mfma"], 4),
                "tops": round(tops, 1), "peak_tops": INT8_PEAK_TOPS, "frac": round(tops / INT8_PEAK_TOPS, 4)}
    out = {
        "metric": "bit-exact polymatrix mul/s, 128^2 deg 8191: INT8-MFMA vs uint64-ALU form",
        "value": round(value, 3),
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_int8, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "polymat_mul 128x128.128x128 deg 8191 p=2013265921 (length-16384 NTT)",
                   "m": M, "k": K, "n": N, "degree": SA - 1, "p": P, "ntt_points": pts,
                   "pointwise": "int8_mfma", "parallelism": "single_gpu",
                   "inputs": "a then b in polfirst order, uniform in [0,p) from numpy's PCG64 seeded "
                             "0x4C494E424F58 + 2 (other legs: + 4 batch, + 6 order basis, + 7 configs)"},
        "plain_u64": {"value": round(1000.0 / ms_plain, 3), "ms_per_step": round(ms_plain, 4),
                      "stages_ms": {k: round(v, 4) for k, v in stages_plain if not k.startswith("kernel ")}},
        "int8_vs_plain": round(ms_plain / ms_int8, 3),
        "stages_ms": {k: round(v, 4) for k, v in stages if not k.startswith("kernel ")},
        "kernels": kernels,
        "gmulmod_per_s": round(40265318400 / (ms_int8 * 1e-3) / 1e9, 2),
        "ntt_passes": ntt,
        "int8_mfma_utilisation": mfma,
        "self_check": check,
        "host_operands": host,
        "wall_s_timed": round(wall, 4),
        "roofline": roofline,
        "batch_1024_products": batch,
        "order_basis": order_basis,
        "other_configs": configs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
