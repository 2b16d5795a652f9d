"""The stream reference kernels (copy, non-temporal copy, read, fill) touch every byte exactly once, for
sizes that end inside a block's chunk, on a chunk boundary, and below one block.  The copies move 4 KiB
(256 16-byte elements) per block, with the blocks renumbered per XCD: block counts that are and are not
multiples of the 8 XCDs are included."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nontemporal", [False, True])
@pytest.mark.parametrize("n16", [1, 7, 255, 256, 257, 8 * 256, 8 * 256 + 3, 9 * 256 - 1, 61 * 256 + 17, 16384,
                                 16384 + 5, 3 * 16384, 5 * 16384 - 1, 1 << 20])
def test_stream_copy_exact(gpu, n16, nontemporal):
    g = torch.Generator(device="cuda").manual_seed(n16)
    src = torch.randint(-(2**62), 2**62, (2 * n16 + 2,), dtype=torch.int64, device="cuda", generator=g)
    dst = torch.full_like(src, -1)
    gpu.stream_copy(dst[: 2 * n16], src[: 2 * n16], nontemporal=nontemporal)
    torch.cuda.synchronize()
    assert torch.equal(dst[: 2 * n16], src[: 2 * n16])
    # Nothing past the end is written.
    assert torch.equal(dst[2 * n16:], torch.full((2,), -1, dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("n16", [1, 7, 16384, 16384 + 5, 3 * 16384, 5 * 16384 - 1, 1 << 20])
def test_stream_read_folds_every_word(gpu, n16):
    """The read stream's 256 XOR slots fold to the XOR of every 64-bit word of the input
    (computed on the host), so every byte is read once."""
    import numpy as np

    g = torch.Generator(device="cuda").manual_seed(1000 + n16)
    src = torch.randint(-(2**62), 2**62, (2 * n16,), dtype=torch.int64, device="cuda", generator=g)
    slots = torch.zeros(256, dtype=torch.int64, device="cuda")
    gpu.stream_read(src, slots)
    torch.cuda.synchronize()
    got = np.bitwise_xor.reduce(slots.cpu().numpy().view(np.uint64))
    want = np.bitwise_xor.reduce(src.cpu().numpy().view(np.uint64))
    assert got == want


@pytest.mark.parametrize("n16", [1, 7, 16384 + 5, 5 * 16384 - 1])
def test_stream_fill_exact(gpu, n16):
    dst = torch.full((2 * n16 + 2,), -1, dtype=torch.int64, device="cuda")
    pattern = 0x3FF0_0000_0000_0001
    gpu.stream_fill(dst[: 2 * n16], pattern)
    torch.cuda.synchronize()
    assert torch.equal(dst[: 2 * n16], torch.full((2 * n16,), pattern, dtype=torch.int64, device="cuda"))
    assert torch.equal(dst[2 * n16:], torch.full((2,), -1, dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("nb", [(1, 1), (5, 70), (37, 125), (997, 613), (4000, 2000)])
def test_stream_pattern2d_copies_the_interior(gpu, nb):
    """The step-pattern copy moves every interior cell of both arrays and writes no ghost
    cell: grids of one cell, ragged windows and strips, and the 4000x2000 comparison grid."""
    shape = (nb[0] + 2, nb[1] + 2, 2)
    g = torch.Generator(device="cuda").manual_seed(nb[0] * 7919 + nb[1])
    a = [torch.rand(shape, dtype=torch.float64, device="cuda", generator=g) for _ in range(2)]
    b = [torch.full(shape, -7.0, dtype=torch.float64, device="cuda") for _ in range(2)]
    gpu.stream_pattern2d(a, b)
    torch.cuda.synchronize()
    for src, dst in zip(a, b):
        assert torch.equal(dst[1:-1, 1:-1], src[1:-1, 1:-1])
        ghost = torch.ones(shape[:2], dtype=torch.bool, device="cuda")
        ghost[1:-1, 1:-1] = False
        assert bool((dst[ghost] == -7.0).all())
