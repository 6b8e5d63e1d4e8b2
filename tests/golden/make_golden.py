"""Writes tests/golden/oracle_vectors.npz: small reference-input cases (the
reference apps' rand()/RAND_MAX stream, srand(1)) and the restated reference
CPU path's outputs.  Regression fixtures for the oracle and small GPU parity
cases.  Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def main():
    M, N, K = 37, 19, 53
    A, B, C0 = oracle.hml_gemm_inputs(M, N, K)
    C = oracle.sgemm_cpu(C0.copy(), A, B, 1.0, 0.0)
    Mv, Nv = 41, 29
    Av, x, y0 = oracle.hml_gemv_inputs(Mv, Nv)
    y = oracle.sgemv_cpu(y0.copy(), Av, x, 1.0, 0.0)
    np.savez_compressed(
        os.path.join(HERE, "oracle_vectors.npz"),
        gemm_M=M, gemm_N=N, gemm_K=K, gemm_A=A, gemm_B=B, gemm_C0=C0,
        gemm_C_cpu=C, gemm_csum=oracle.csum(C),
        gemv_M=Mv, gemv_N=Nv, gemv_A=Av, gemv_x=x, gemv_y0=y0, gemv_y_cpu=y,
        rand_first8=oracle.hml_random(8))


if __name__ == "__main__":
    main()
