"""GPU parity tests of the k-means assignment (hmlKmeansAssign) against the
oracle's restatement of the reference assign kernels
(include/hml_kmeans_kernel_template.h:18-88, 90-316, 323-367) and CPU path
(src/hml_kmeans_cpu.cu:55-95).  Indices are compared bit-exactly, including
near-ties (rows halfway between two centroids), duplicate centroids, NaN and
overflowing rows, dims that are not multiples of 4, centroid counts that
need several LDS chunks, and both distance arithmetics."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


KERNELS = [None, "screened", "direct"]


def kernels(D):
    """Kernel choices to check for D dims (the direct one takes <= 32)."""
    return KERNELS if D <= 32 else KERNELS[:2]


def run(hml, rows, ctrds, separate=False, offset=0, kernel=None):
    n, D = rows.shape
    K = ctrds.shape[0]
    flat = np.concatenate([np.full(offset, 7, np.float32), rows.ravel()]).astype(np.float32)
    dr = hml.DeviceBuffer.from_numpy(flat)
    dc = hml.DeviceBuffer.from_numpy(np.ascontiguousarray(ctrds, np.float32))
    da = hml.DeviceBuffer(4 * max(n, 1))
    hml.kmeans_assign(dr.ptr + 4 * offset, dc, da, n, D, K, separate_rounding=separate,
                      kernel=kernel)
    hml.synchronize()
    return da.to_numpy((n,), np.uint32)


def clustered(rng, n, D, K, spread=0.5):
    ctrds = rng.standard_normal((K, D)).astype(np.float32)
    rows = ctrds[rng.integers(0, K, n)] + spread * rng.standard_normal((n, D)).astype(np.float32)
    return rows.astype(np.float32), ctrds


def near_ties(rng, n, D, K, eps=1e-5):
    ctrds = rng.standard_normal((K, D)).astype(np.float32)
    a, b = ctrds[rng.integers(0, K, n)], ctrds[rng.integers(0, K, n)]
    rows = (a + b) / 2 + eps * rng.standard_normal((n, D))
    return rows.astype(np.float32), ctrds


@pytest.mark.parametrize("D", [1, 3, 4, 7, 16, 17, 31, 32, 48, 50, 63, 64])
@pytest.mark.parametrize("K", [1, 5, 16, 33, 64, 100])
def test_matches_oracle(hml, D, K):
    rng = np.random.default_rng(D * 1000 + K)
    n = 3001
    print("case D=%d K=%d" % (D, K), flush=True)
    for make in (clustered, near_ties):
        rows, ctrds = make(rng, n, D, K)
        for sep in (False, True):
            want = oracle.kmeans_assign(rows, ctrds, fused=not sep)
            for kern in kernels(D):
                got = run(hml, rows, ctrds, sep, kernel=kern)
                np.testing.assert_array_equal(got, want, err_msg="%s sep=%s kernel=%s" % (
                    make.__name__, sep, kern))


@pytest.mark.parametrize("D,K", [(64, 129), (64, 300), (16, 1000), (8, 4097), (2, 70), (64, 1000), (3, 9000)])
def test_many_centroids(hml, D, K):
    """More centroids than one LDS chunk holds (several chunks per row step:
    above 256 x 64 dims for the direct kernel, 128 x 64 for the screened
    one) and several screening passes."""
    rng = np.random.default_rng(K)
    print("case D=%d K=%d" % (D, K), flush=True)
    for make in (clustered, near_ties):
        rows, ctrds = make(rng, 2000, D, K)
        want = oracle.kmeans_assign(rows, ctrds)
        for kern in kernels(D):
            got = run(hml, rows, ctrds, kernel=kern)
            np.testing.assert_array_equal(got, want, err_msg="%s kernel=%s" % (make.__name__, kern))


def test_fused_vs_separate_flip(hml):
    """The constructed input on which the two arithmetics disagree: the
    kernel follows the requested one."""
    from test_oracle import fused_flip_case
    rows, ctrds = fused_flip_case()
    rows = np.repeat(rows, 37, 0)
    for kern in KERNELS:
        assert (run(hml, rows, ctrds, separate=False, kernel=kern) == 1).all()
        assert (run(hml, rows, ctrds, separate=True, kernel=kern) == 0).all()


def test_edge_rows(hml):
    """Duplicate centroids (the first wins), rows equal to centroids,
    NaN / overflowing rows (numClusts), zero rows, an odd row offset."""
    rng = np.random.default_rng(3)
    D, K = 21, 40
    ctrds = rng.standard_normal((K, D)).astype(np.float32)
    ctrds[30] = ctrds[4]
    ctrds[31] = ctrds[4]
    rows = ctrds[rng.integers(0, K, 500)].copy()
    rows[7] = np.nan
    rows[8] = 3e38
    rows[9, 3] = np.inf
    rows[10] = 0
    want = oracle.kmeans_assign(rows, ctrds)
    assert want[7] == K and want[8] == K and want[9] == K
    assert not np.isin(want, [30, 31]).any()
    for off in (0, 1):
        for kern in KERNELS:
            np.testing.assert_array_equal(run(hml, rows, ctrds, offset=off, kernel=kern), want)


def test_sizes(hml):
    """Row counts around the 16-row tile and the 4-tile workgroup step, one
    cluster, and a large problem checked through a sample of rows."""
    rng = np.random.default_rng(5)
    for n in (1, 15, 16, 17, 63, 64, 65, 255, 256, 257, 1000):
        rows, ctrds = clustered(rng, n, 10, 7)
        for kern in KERNELS:
            np.testing.assert_array_equal(run(hml, rows, ctrds, kernel=kern),
                                          oracle.kmeans_assign(rows, ctrds))
    rows, ctrds = clustered(rng, 100, 5, 1)
    assert (run(hml, rows, ctrds) == 0).all()
    n, D, K = 1 << 22, 32, 64
    rows, ctrds = clustered(rng, n, D, K)
    got = run(hml, rows, ctrds)
    idx = rng.choice(n, 20000, replace=False)
    np.testing.assert_array_equal(got[idx], oracle.kmeans_assign(rows[idx], ctrds))
    # every index in range, and each cluster's rows are nearer to it than
    # to a random other centroid (size-independent property)
    assert got.max() < K
    other = rng.integers(0, K, n)
    d_own = ((rows - ctrds[got]) ** 2).sum(1)
    d_oth = ((rows - ctrds[other]) ** 2).sum(1)
    assert (d_own <= d_oth * (1 + 1e-5) + 1e-30).all()


def test_direct_kernel_dim_limit(hml):
    """Forcing the direct kernel above 32 dims is rejected, not rerouted."""
    rows = np.zeros((10, 33), np.float32)
    ctrds = np.zeros((3, 33), np.float32)
    with pytest.raises(hml.HmlError):
        run(hml, rows, ctrds, kernel="direct")


def test_centroid_scales(hml):
    """Centroids of very different norms (the screen's margin is bounded
    with each chunk's largest norm), and an infinite and a NaN centroid
    (never chosen; the infinite norm widens the screen to every centroid),
    over several screening passes and LDS chunks."""
    rng = np.random.default_rng(11)
    for D, K in ((16, 200), (64, 300), (5, 40)):
        print("case D=%d K=%d" % (D, K), flush=True)
        for make in (clustered, near_ties):
            rows, ctrds = make(rng, 2000, D, K)
            ctrds = ctrds.copy()
            ctrds[3] *= 1e4
            ctrds[K // 2] *= 1e-3
            want = oracle.kmeans_assign(rows, ctrds)
            for kern in kernels(D):
                np.testing.assert_array_equal(run(hml, rows, ctrds, kernel=kern), want,
                                              err_msg="%s kernel=%s" % (make.__name__, kern))
            ctrds[7, 0] = np.inf
            ctrds[9] = np.nan
            want = oracle.kmeans_assign(rows, ctrds)
            assert not np.isin(want, [7, 9]).any()
            for kern in kernels(D):
                np.testing.assert_array_equal(run(hml, rows, ctrds, kernel=kern), want,
                                              err_msg="inf/nan %s kernel=%s" % (make.__name__, kern))
