"""Host-side launch planning of the skinny kernels, checked through
hmlSgemmWorkspaceSize (no GPU needed).

DOT form: K is cut into chunks of a multiple of the 64-long pipeline stage;
the wpb waves of a workgroup take consecutive chunks and sum them in LDS, so
partial sums go to the workspace only when the chunks span kg > 1
workgroups: kg * M * N floats (rounded up to 256 bytes)."""
import ctypes
import math

import pytest

import hiperml_amd as hml


def ws(M, N, K, v, ta=False):
    v = hml.SgemmVariant(*v)
    return hml.lib.hmlSgemmWorkspaceSize(0, 1 if ta else 0, 0, M, N, K, ctypes.byref(v))


def dot_partials(M, N, K, ksplit, wpb):
    kchunk = math.ceil(math.ceil(K / ksplit) / 64) * 64
    nchunks = math.ceil(K / kchunk)
    kw = max(1, min(wpb, nchunks))
    kg = math.ceil(nchunks / kw)
    return 0 if kg == 1 else (kg * M * N * 4 + 255) // 256 * 256


@pytest.mark.parametrize("M,N,K,ks,wpb", [
    (8192, 16, 8192, 12, 4),   # 12 chunks of 704, 3 workgroups
    (8192, 16, 8192, 4, 8),    # 4 chunks in one workgroup: no partials
    (8192, 16, 8192, 8, 4),    # 8 chunks, 2 workgroups
    (2048, 48, 2048, 64, 8),   # 32 chunks of 64 (the stage), 4 workgroups
    (1000, 7, 130, 3, 2),      # 3 chunks of 64 -> 3 chunks, 2 workgroups
    (100, 16, 63, 5, 4),       # K below one stage: one chunk
])
def test_dot_workspace(M, N, K, ks, wpb):
    assert ws(M, N, K, (1, 1, 1, ks, wpb)) == dot_partials(M, N, K, ks, wpb)


def test_axpy_workspace_packs_b():
    # AXPY form (transposed A): packed B (16-k steps x column tiles x 64
    # lanes x 4 floats) plus partials for every split
    M, N, K = 4096, 16, 1000
    nsteps = math.ceil(K / 16)
    packed = (nsteps * 1 * 64 * 4 * 4 + 255) // 256 * 256
    kchunk = math.ceil(math.ceil(K / 4) / 16) * 16
    parts = math.ceil(K / kchunk)
    assert ws(M, N, K, (2, 1, 1, 4, 4), ta=True) == packed + (parts * M * N * 4 + 255) // 256 * 256


def test_dot_variant_limits():
    shape = hml.GemmShape(8192, 64, 4096, 0, 0, 0, 1)
    cands = (hml.SgemmVariant * 512)()
    n = hml.lib.hmlSgemmCandidates(ctypes.byref(shape), cands, 512)
    for i in range(n):
        v = cands[i]
        if v.family != 1:
            continue
        assert v.ct <= 4 and v.rt * v.ct <= 8 and v.wpb in (1, 2, 4, 8)
