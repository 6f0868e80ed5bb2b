"""BLCO construction on the GPU (blco_device_encode) against the host
encoder, which the oracle's golden vectors pin: the same layout, BLCO words,
values and blocks, bit for bit, on 64-bit, 65-bit and 119-bit keys, with and
without a block-size cap, and the host encoder's errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = {
    # name: (dims, nnz, zipf, max_block_size)
    "uniform3": ((2000, 1500, 1000), 60000, None, 0),
    "zipf3": ((12092, 9184, 28818), 150000, (1.1, 1.1, 0.0), 0),
    "bits65": ((4800000, 1800000, 1800000), 80000, (1.2, 1.2, 1.2), 0),
    "bits119": ((10000000, 9000000, 8000000, 7000000, 5000000), 50000, None, 0),
    "capped": ((3000, 2000, 1000), 50000, None, 7000),
    "capped65": ((4800000, 1800000, 1800000), 60000, None, 5000),
    "order2": ((5000, 3000), 20000, None, 0),
    "tiny": ((3, 2, 2), 5, None, 2),
    "ones": ((1, 1, 1), 1, None, 0),
}


def same(B, a, b):
    assert a.nmodes == b.nmodes and np.array_equal(a.dims, b.dims) and a.nnz == b.nnz
    for f in ("masks", "pos", "alto_lo", "alto_hi", "block_start", "block_nnz", "block_coords"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert a.nblocks == b.nblocks and a.max_block_nnz == b.max_block_nnz
    assert np.array_equal(a.coords, b.coords)
    assert np.array_equal(a.values.view(np.uint64), b.values.view(np.uint64))


@pytest.mark.parametrize("name", list(SHAPES))
def test_device_encoder_matches_host(B, gpu, name):
    dims, nnz, zipf, mbs = SHAPES[name]
    t = B.Coo.generate(dims, nnz, zipf=zipf, gen_rank=3, value_kind=1, seed=41, topup=False)
    print("device encode", name, dims, t.nnz, mbs, flush=True)
    host = B.encode(t, max_block_size=mbs)
    dev = B.encode_device(t, max_block_size=mbs)
    same(B, dev, host)
    if name in ("bits65", "bits119", "capped65"):
        assert host.nblocks > 1


def test_device_encoder_keeps_input_order_of_equal_keys(B, gpu):
    """Duplicate coordinates (not merged) keep their input order, as in the
    host encoder's stable sort."""
    dims = (40, 30, 20)
    rng = np.random.default_rng(3)
    cidx = rng.integers(0, 5, size=(3, 3000), dtype=np.uint64)
    vals = rng.random(3000)
    same(B, B.encode_device(dims, cidx, vals), B.encode(dims, cidx, vals))


def test_device_encoder_errors(B, gpu):
    dims = (10, 10)
    cidx = np.array([[1, 2], [3, 10]], np.uint64)  # 10 is outside mode 1
    with pytest.raises(B.BlcoError):
        B.encode_device(dims, cidx, np.ones(2))
    empty = B.encode_device(dims, np.zeros((2, 0), np.uint64), np.zeros(0))
    assert empty.nnz == 0 and empty.nblocks == 0
    with pytest.raises(B.BlcoError):
        B.encode_device((1 << 40, 1 << 40, 1 << 40, 1 << 40), np.zeros((4, 1), np.uint64), np.ones(1))
