"""The stream-copy reference kernel copies every byte exactly, for sizes that end inside
a block's chunk, on a chunk boundary, and below one block."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n16", [1, 7, 16384, 16384 + 5, 3 * 16384, 5 * 16384 - 1, 1 << 20])
def test_stream_copy_exact(gpu, n16):
    g = torch.Generator(device="cuda").manual_seed(n16)
    src = torch.randint(-(2**62), 2**62, (2 * n16 + 2,), dtype=torch.int64, device="cuda", generator=g)
    dst = torch.full_like(src, -1)
    gpu.stream_copy(dst[: 2 * n16], src[: 2 * n16])
    torch.cuda.synchronize()
    assert torch.equal(dst[: 2 * n16], src[: 2 * n16])
    # Nothing past the end is written.
    assert torch.equal(dst[2 * n16:], torch.full((2,), -1, dtype=torch.int64, device="cuda"))
