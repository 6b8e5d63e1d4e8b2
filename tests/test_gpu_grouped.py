"""GPU tests of the persistent grouped skinny SGEMM (hmlSgemmGrouped*).

Integer-valued inputs make every result exact, so the grouped kernel must
equal the exact product of every problem bit for bit, directly and from the
captured hipGraph.  Graph replays are interleaved with eager work and host
synchronisation, and each replay is checked (beta = 1 accumulates, so a
skipped or doubled replay shows)."""
import numpy as np
import pytest

from util import int_matrix

pytestmark = pytest.mark.gpu


def make_problems(hml, rng, count, mmax=4000, kmax=700, nmax=64):
    probs, host = [], []
    for i in range(count):
        M = int(np.exp(rng.uniform(np.log(64), np.log(mmax))))
        K = int(np.exp(rng.uniform(np.log(16), np.log(kmax))))
        N = int(rng.integers(1, nmax + 1))
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        C = int_matrix(rng, (M, N))
        dA, dB, dC = (hml.DeviceBuffer.from_numpy(x) for x in (A, B, C))
        probs.append((M, N, K, dA, dB, dC))
        host.append((A, B, C))
    return probs, host


def exact(A, B):
    return A.astype(np.int64) @ B.astype(np.int64)


@pytest.mark.parametrize("kernel", ["default", "shared", "wave", "wave1"])
def test_grouped_exact_direct(hml, kernel, monkeypatch):
    """Every grouped kernel kind.  Shared-B workgroups take blocks of 4 row
    tiles and share B through LDS (128 rows per block for N <= 32, 64
    above); per-wave kernels take one 32-row (wave) or 16-row (wave1) tile
    per wave.  The default uses the shared-B kernel for N <= 32 and the
    32-row per-wave kernel above; HML_GROUPED_KERNEL forces one kind for
    every class."""
    if kernel != "default":
        monkeypatch.setenv("HML_GROUPED_KERNEL", kernel)
    rng = np.random.default_rng(10)
    probs, host = make_problems(hml, rng, 40)
    plan = hml.GroupedPlan(probs, 1.0, 0.0)
    info = plan.info()
    if kernel == "wave":
        unit = lambda n: 32
    elif kernel == "wave1":
        unit = lambda n: 16
    elif kernel == "shared":
        unit = lambda n: 64 if n > 32 else 128
    else:
        unit = lambda n: 32 if n > 32 else 128
    assert info["tiles"] == sum((p[0] + unit(p[1]) - 1) // unit(p[1]) for p in probs)
    plan.run(graph=False)
    hml.synchronize()
    for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
        np.testing.assert_array_equal(dC.to_numpy((M, N)), exact(A, B).astype(np.float32))
    plan.destroy()


def test_grouped_set_kind_every_kernel(hml):
    """All twelve (class, kind) kernels through hmlSgemmGroupedSetKind, on
    integer-valued problems of every class: exact results, and the graph
    is re-captured after a kind changes."""
    rng = np.random.default_rng(13)
    probs, host = make_problems(hml, rng, 30, mmax=2500, kmax=600)
    assert {0, 1, 2, 3} <= {hml.GroupedPlan.class_of(p[1]) for p in probs}
    plan = hml.GroupedPlan(probs, 1.0, 0.0)
    for kind in hml.GroupedPlan.KINDS:
        for c in range(hml.GroupedPlan.CLASSES):
            plan.set_kind(c, kind)
        assert plan.kinds() == [kind] * hml.GroupedPlan.CLASSES
        for graph in (False, True):
            for M, N, K, dA, dB, dC in probs:
                hml.check(hml.lib.hmlMemsetD8(dC.ptr, 0xFF, 4 * M * N))
            plan.run(graph=graph)
            hml.synchronize()
            for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
                np.testing.assert_array_equal(dC.to_numpy((M, N)), exact(A, B).astype(np.float32),
                                              err_msg="%s graph=%s" % (kind, graph))
    plan.destroy()


def test_grouped_autotune(hml):
    """hmlSgemmGroupedAutotune times every kind of every class, keeps the
    fastest, and leaves C as it was (beta = 1 here, so a trial run that
    leaked into C would show); the tuned plan then computes C + A B
    exactly."""
    rng = np.random.default_rng(14)
    probs, host = make_problems(hml, rng, 30, mmax=3000, kmax=600)
    plan = hml.GroupedPlan(probs, 1.0, 1.0)
    res = plan.autotune(reps=2)
    print("grouped autotune:", res)
    hml.synchronize()
    for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
        np.testing.assert_array_equal(dC.to_numpy((M, N)), C)
    assert sorted(res["ms"]) == [0, 1, 2, 3]
    for c, times in res["ms"].items():
        assert all(t > 0 for t in times.values())
        assert res["kinds"][c] == min(times, key=times.get)
    plan.run(graph=True)
    hml.synchronize()
    for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
        want = C.astype(np.int64) + exact(A, B)
        np.testing.assert_array_equal(dC.to_numpy((M, N)), want.astype(np.float32))
    plan.destroy()


def test_grouped_graph_replays(hml):
    rng = np.random.default_rng(11)
    probs, host = make_problems(hml, rng, 25, mmax=3000, kmax=500)
    plan = hml.GroupedPlan(probs, 1.0, 1.0)   # C += A B on every run
    stream = hml.Stream()
    scratch = hml.DeviceBuffer(1 << 20)
    for rep in range(1, 4):
        plan.run(graph=True, stream=stream.ptr)
        stream.synchronize()
        scratch.fill_uniform(1 << 18, rep)       # eager work between replays
        hml.synchronize()
        for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
            want = C.astype(np.int64) + rep * exact(A, B)
            np.testing.assert_array_equal(dC.to_numpy((M, N)), want.astype(np.float32),
                                          err_msg="replay %d" % rep)
    # new contents behind the same pointers are picked up by the next replay
    M, N, K, dA, dB, dC = probs[0]
    A2 = int_matrix(rng, (M, K))
    dA.upload(A2)
    before = dC.to_numpy((M, N)).astype(np.int64)
    plan.run(graph=True, stream=stream.ptr)
    stream.synchronize()
    np.testing.assert_array_equal(dC.to_numpy((M, N)),
                                  (before + exact(A2, host[0][1])).astype(np.float32))
    plan.destroy()


def test_grouped_matches_sequential_random(hml):
    """Random floats: the grouped kernel and one hmlSgemmEx launch per
    problem agree to FP32 rounding (same kernel body, no K split in the
    grouped kernel)."""
    rng = np.random.default_rng(12)
    probs = []
    for i in range(30):
        M = int(rng.integers(64, 5000)); K = int(rng.integers(64, 1500)); N = int(rng.integers(1, 65))
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
        probs.append((M, N, K, dA, dB, hml.DeviceBuffer(4 * M * N), hml.DeviceBuffer(4 * M * N), A, B))
    plan = hml.GroupedPlan([p[:6] for p in probs])
    plan.run()
    for M, N, K, dA, dB, dC, dS, A, B in probs:
        hml.sgemm_ex(M, N, K, dA, dB, dS)
    hml.synchronize()
    for M, N, K, dA, dB, dC, dS, A, B in probs:
        g, s = dC.to_numpy((M, N)), dS.to_numpy((M, N))
        R = A.astype(np.float64) @ B.astype(np.float64)
        bnd = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
        lim = 2 * (K + 2) * 2.0 ** -24 * bnd
        assert (np.abs(g - R) <= lim).all() and (np.abs(s - R) <= lim).all()
    plan.destroy()


def test_grouped_rejects_wide_problem(hml):
    d = hml.DeviceBuffer(1 << 16)
    with pytest.raises(hml.HmlError) as e:
        hml.GroupedPlan([(16, 64, 32, d, d, d)])
    assert e.value.status == 6


def test_grouped_512_suite_vs_fp64(hml):
    """The benchmark's batched workload at full size: 512 problems, M and K
    log-uniform in [64, 65536] and [64, 4096], N uniform in 1..64 (fixed
    seed), uniform[-1,1] entries filled on the device.  Run directly and from
    the captured graph, every problem is checked element by element by the
    FP64 kernel against 2 gamma_{K+2} |A||B| (the grouped tile runs its whole
    K in one FP32 accumulation), and both runs agree bit for bit."""
    from bench import grouped_suite_shapes
    shapes = grouped_suite_shapes(512, seed=0)
    assert max(m for m, n, k in shapes) > 30000 and max(k for m, n, k in shapes) > 3000
    probs = []
    for i, (M, N, K) in enumerate(shapes):
        dA, dB, dC = (hml.DeviceBuffer(4 * M * K), hml.DeviceBuffer(4 * K * N),
                      hml.DeviceBuffer(4 * M * N))
        dA.fill_uniform(M * K, 3 * i + 1)
        dB.fill_uniform(K * N, 3 * i + 2)
        probs.append((M, N, K, dA, dB, dC))
    plan = hml.GroupedPlan(probs, 1.0, 0.0)
    first = []
    for graph in (False, True):
        print("grouped suite, graph=%s" % graph, flush=True)
        for M, N, K, dA, dB, dC in probs:  # NaN everywhere: every output rewritten
            hml.check(hml.lib.hmlMemsetD8(dC.ptr, 0xFF, 4 * M * N))
        plan.run(graph=graph)
        hml.synchronize()
        worst = 0.0
        for j, (M, N, K, dA, dB, dC) in enumerate(probs):
            st = hml.check_f64(M, N, K, dA, dB, dC)
            assert st["nonfinite"] == 0, (j, M, N, K)
            assert st["max_rel_bound"] <= 2 * hml.gamma(K + 2), (j, M, N, K, st)
            worst = max(worst, st["max_rel_bound"] / hml.gamma(K + 2))
            if not graph:
                first.append(dC.to_numpy((M, N)) if M * N <= 1 << 20 else None)
            elif first[j] is not None:
                np.testing.assert_array_equal(dC.to_numpy((M, N)), first[j])
        print("worst max_err/bound in units of gamma_K+2: %.3f" % worst)
    plan.destroy()


def test_grouped_wide_and_short_problems(hml):
    """Problems the suite's distribution leaves out: N past 64 (several
    column passes of the 4-column class), and K below one 64-k stage down to
    1, mixed with ordinary ones; every kind of every class, exact on
    integer-valued data, with beta = 1 (C read back)."""
    rng = np.random.default_rng(21)
    shapes = [(300, 65, 1), (500, 96, 3), (2000, 130, 17), (777, 200, 64), (640, 70, 65),
              (1500, 128, 300), (90, 1, 2), (400, 33, 5), (1000, 16, 63), (256, 24, 129)]
    probs, host = [], []
    for M, N, K in shapes:
        A = int_matrix(rng, (M, K))
        B = int_matrix(rng, (K, N))
        C = int_matrix(rng, (M, N))
        dA, dB, dC = (hml.DeviceBuffer.from_numpy(x) for x in (A, B, C))
        probs.append((M, N, K, dA, dB, dC))
        host.append((A, B, C))
    plan = hml.GroupedPlan(probs, 1.0, 1.0)
    for kind in hml.GroupedPlan.KINDS:
        print("kind %s" % kind, flush=True)
        for c in range(hml.GroupedPlan.CLASSES):
            plan.set_kind(c, kind)
        for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
            dC.upload(C)
        plan.run(graph=False)
        hml.synchronize()
        for (M, N, K, dA, dB, dC), (A, B, C) in zip(probs, host):
            want = (exact(A, B) + C.astype(np.int64)).astype(np.float32)
            np.testing.assert_array_equal(dC.to_numpy((M, N)), want,
                                          err_msg="kind %s M=%d N=%d K=%d" % (kind, M, N, K))
    plan.destroy()
