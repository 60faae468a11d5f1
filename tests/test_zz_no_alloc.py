"""The allocation-free bf16 product (poe2d_gemm_bf16_noalloc, a fixed device scratch in place of
per-call allocations) against the allocating entry point it mirrors (poe2d_gemm_bf16): bit for bit,
on consecutive calls that reuse the scratch, the later ones with smaller operands (stale contents of
a larger call must not leak into a smaller one), with and without split-K partials; and the size
limit refused before any device work."""
import numpy as np
import pytest

from poe2d_amd import _abi
from poe2d_amd.mlp import gemm_bf16, gemm_bf16_noalloc

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.mark.parametrize("shapes", [
    [(256, 512, 4096, 8), (130, 129, 64, 2)],      # split-K, then a smaller split-K product
    [(512, 640, 1000, 1), (96, 33, 40, 1), (1, 5, 7, 1)],  # no split, smaller and smaller (K padded)
    [(200, 300, 2048, 4), (200, 300, 100, 1)],    # a split-K call, then a plain one of the same M, N
])
def test_noalloc_gemm_equals_allocating_gemm(shapes):
    rng = np.random.default_rng(len(shapes))
    for m, n, k, ksplit in shapes:
        a = rng.normal(size=(m, k)).astype(np.float32)
        b = rng.normal(size=(n, k)).astype(np.float32)
        ref = gemm_bf16(a, b, ksplit)
        got = gemm_bf16_noalloc(a, b, ksplit)
        say(f"{m}x{n}x{k} split {ksplit}: bitwise equal {np.array_equal(got, ref)}")
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), (m, n, k, ksplit)


def test_noalloc_gemm_refuses_operands_beyond_its_scratch():
    # 4 * (2048 * 2048) * 2 bytes of FP32 operands alone exceed the 32 MiB scratch
    a = np.ones((2048, 2048), np.float32)
    with pytest.raises(_abi.Poe2dError) as ei:
        gemm_bf16_noalloc(a, a, 1)
    assert ei.value.status == _abi.POE2D_E_INVALID
    assert "scratch" in str(ei.value)
    # a product that fits still works after the refusal
    x = np.arange(64 * 32, dtype=np.float32).reshape(64, 32) / 100
    assert np.array_equal(gemm_bf16_noalloc(x, x, 1), gemm_bf16(x, x, 1))
