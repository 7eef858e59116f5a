"""Kernel-level GPU tests: the MFMA contractions against a host GEMM (pins the operand and
accumulator lane layouts of v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32 and the bf16 variant)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 4, 16), (8, 3, 70), (37, 21, 130), (168, 16, 300), (128, 128, 200), (512, 36, 65)]
TOL = {0: 1e-14, 1: 2e-6, 2: 1e-2}   # relative max error per precision (FP64, FP32, BF16 operands)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_contraction(which, prec):
    from radiant_hip import _lib
    L = _lib.lib()
    rng = np.random.default_rng(3 + which + 10 * prec)
    for (M, K, N) in SHAPES:
        A = np.ascontiguousarray(rng.standard_normal((M, K)))
        B = np.ascontiguousarray(rng.standard_normal((K, N)))
        C = np.zeros((M, N))
        _lib.check(L.rt_test_contraction(which, prec, M, K, N, _lib.ptr(A), _lib.ptr(B), _lib.ptr(C)))
        ref = A @ B
        err = np.abs(C - ref).max() / np.abs(ref).max()
        print(which, prec, (M, K, N), err, flush=True)
        assert err < TOL[prec], (M, K, N, err)
