"""GPU complex GEMM (MFMA, real 2x2-block expansion) against numpy, all shapes the path
uses: skinny A^H B with split-K, tall A B, degenerate sizes.  FP64 <= 1e-13 relative;
mixed (complex64 products, fp32 accumulation) <= 2e-6 relative."""
import numpy as np
import pytest

from q3_amd import FP64, MIXED
from q3_amd.planewave import DeviceArray, zgemm

pytestmark = pytest.mark.gpu

SHAPES = [  # (transa, m, n, k)
    ("N", 1, 1, 1), ("C", 1, 1, 1), ("N", 70, 33, 5), ("C", 70, 33, 5),
    ("C", 256, 40, 5000),      # projector overlap: split-K
    ("N", 5000, 40, 256),      # projector update
    ("C", 320, 160, 3001),     # Gram block
    ("N", 3001, 160, 320),     # rotation
    ("C", 17, 3, 100000),      # very deep split-K
    ("N", 1100, 100, 64),      # row blocks not a multiple of 8 (the XCD grouping pads them)
    ("C", 100, 70, 20000),     # split-K over a multiple of 8 chunks on their XCDs
]


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


@pytest.mark.parametrize("transa,m,n,k", SHAPES)
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_zgemm_against_numpy(gpu_context, transa, m, n, k, precision):
    A = _rand((m, k) if transa == "N" else (k, m), m + 3 * k)
    B = _rand((k, n), n + 7)
    C = _rand((m, n), 11)
    alpha, beta = -1.0, 1.0
    want = alpha * ((A if transa == "N" else A.conj().T) @ B) + beta * C
    dA, dB, dC = (DeviceArray.from_host(gpu_context, x) for x in (A, B, C))
    zgemm(gpu_context, transa, m, n, k, alpha, dA, A.shape[0], dB, k, beta, dC, m, precision)
    got = dC.download((m, n))
    tol = 1e-13 if precision == FP64 else 2e-6
    scale = np.abs(A).max() * np.abs(B).max() * np.sqrt(k) + np.abs(C).max()
    assert np.abs(got - want).max() <= tol * scale


def test_zgemm_beta_zero_ignores_garbage(gpu_context):
    A = _rand((64, 8), 1)
    B = _rand((64, 4), 2)
    C = np.full((8, 4), np.nan, np.complex128, order="F")
    dA, dB, dC = (DeviceArray.from_host(gpu_context, x) for x in (A, B, C))
    zgemm(gpu_context, "C", 8, 4, 64, 1.0, dA, 64, dB, 64, 0.0, dC, 8)
    got = dC.download((8, 4))
    assert np.allclose(got, A.conj().T @ B, rtol=0, atol=1e-12)


def test_split_k_products_in_sequence_repeat_bit_exactly(gpu_context):
    # Deep-k products whose split-K partials are summed by the last workgroup of each tile:
    # tile counters are left at zero by every product, so products of different tile counts
    # run back to back on one stream give the same bits every time, and the right values.
    shapes = [(96, 96, 1500), (320, 160, 3001), (17, 3, 100000), (96, 96, 1500), (48, 48, 2000)]
    firsts = {}
    for rep in range(3):
        for i, (m, n, k) in enumerate(shapes):
            A = _rand((k, m), 100 + i)
            B = _rand((k, n), 200 + i)
            dA, dB = DeviceArray.from_host(gpu_context, A), DeviceArray.from_host(gpu_context, B)
            dC = DeviceArray.from_host(gpu_context, np.zeros((m, n), np.complex128, order="F"))
            zgemm(gpu_context, "C", m, n, k, 1.0, dA, k, dB, k, 0.0, dC, m)
            got = dC.download((m, n))
            if i not in firsts:
                want = A.conj().T @ B
                assert np.abs(got - want).max() <= 1e-13 * np.abs(A).max() * np.abs(B).max() * np.sqrt(k)
                firsts[i] = got
            else:
                np.testing.assert_array_equal(got, firsts[i])
