"""Host side of the product library (CPU only): the BLCO encoder must be
bit-exact with the oracle's restatement of gen_blcotensor_host, the
generator must follow its spec, file formats must round-trip, and the C ABI
must export every declared symbol."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, random_coo


def header_symbols():
    syms = []
    for h in ("blco_host.h", "blco_hip.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        syms += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(blco_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(syms))


def test_library_exports_every_declared_symbol(B):
    lib = B.lib()
    syms = header_symbols()
    assert len(syms) >= 39
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (blco_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        getattr(lib, s)


@pytest.mark.parametrize("dims,nnz,mbs", [((2000, 1500, 1000), 20000, 0),
                                          ((12092, 9184, 28818), 20000, 3000),
                                          ((40, 4000, 5, 600), 8000, 0),
                                          ((9, 8, 7, 6, 5), 3000, 500),
                                          ((1 << 22, 1 << 22, 1 << 21), 10000, 0),
                                          ((4800000, 1800000, 1800000), 10000, 777),
                                          ((1 << 30, 1 << 30, 1 << 30, 1 << 25), 5000, 0)])
def test_encoder_bit_exact_vs_oracle(B, O, dims, nnz, mbs):
    rng = np.random.default_rng(41)
    c, v = random_coo(rng, dims, nnz)
    h = B.encode(dims, c, v, max_block_size=mbs)
    o = O.blco_encode(dims, c, v, mbs, 128)
    assert (h.coords == o["coords"]).all()
    assert (h.values == o["vals"]).all()
    assert (h.masks == o["masks"]).all() and (h.pos == o["pos"]).all()
    assert (h.block_start == o["block_start"]).all()
    assert (h.block_nnz == o["block_nnz"]).all()
    assert (h.block_coords == o["block_coords"]).all()
    lo, hi, _, _ = O.alto_setup(dims, 128)
    assert (h.alto_lo == lo).all() and (h.alto_hi == hi).all()
    # decode round trip against the input (through the oracle's sort order)
    _, _, _, perm = O.alto_encode(dims, c, v, 128)
    assert (h.decode() == c[:, perm]).all()


def test_encoder_rejects_bad_input(B):
    with pytest.raises(B.BlcoError):
        B.encode((10, 10), np.array([[1, 10], [0, 0]], np.uint64), np.ones(2))
    with pytest.raises(B.BlcoError):
        B.encode((10, 0), np.zeros((2, 1), np.uint64), np.ones(1))
    with pytest.raises(B.BlcoError):
        B.encode((1 << 40,) * 4, np.zeros((4, 1), np.uint64), np.ones(1))  # 160 bits


def test_encoder_mode_of_length_one(B, O):
    dims = (1, 50, 7)
    rng = np.random.default_rng(42)
    c, v = random_coo(rng, dims, 200)
    h = B.encode(dims, c, v)
    assert (h.decode()[0] == 0).all()
    o = O.blco_encode(dims, c, v, 0, 128)
    assert (h.coords == o["coords"]).all()


def test_encoder_empty(B):
    h = B.encode((5, 6, 7), np.zeros((3, 0), np.uint64), np.zeros(0))
    assert h.nnz == 0 and h.nblocks == 0


def test_generator_deterministic_and_unique(B):
    a = B.Coo.generate((200, 150, 100), 50000, gen_rank=10, seed=1001)
    b = B.Coo.generate((200, 150, 100), 50000, gen_rank=10, seed=1001)
    assert (a.cidx == b.cidx).all() and (a.vals == b.vals).all()
    c = a.cidx
    key = (c[0] * 150 + c[1]) * 100 + c[2]
    assert len(np.unique(key)) == a.nnz
    assert a.nnz < 50000  # duplicates merged without top-up
    assert (a.vals >= 1).all() and (a.vals == np.round(a.vals)).all()
    # rate mean: sum over 10 components of products of 3 Exp(1) -> E = 10
    assert 8.5 < a.vals.mean() < 11.5


def test_generator_zipf_topup(B):
    t = B.Coo.generate((500, 400, 300), 30000, zipf=(1.1, 1.1, 0.0), gen_rank=20, value_kind=1,
                       seed=1002, topup=True)
    assert t.nnz == 30000
    c = t.cidx
    key = (c[0] * 400 + c[1]) * 300 + c[2]
    assert len(np.unique(key)) == t.nnz
    assert (t.vals >= np.log(2) - 1e-12).all()  # log1p of a count >= 1
    # Zipf: row 0 is the most frequent in modes 0 and 1, mode 2 is flat
    h0 = np.bincount(c[0].astype(int), minlength=500)
    h2 = np.bincount(c[2].astype(int), minlength=300)
    assert h0.argmax() == 0 and h0[0] > 5 * h0[100:].mean()
    assert h2.max() < 3 * h2.mean()


def test_init_factor_counter_based(B):
    U = B.init_factor(7, 1, 100, 16)
    V = B.init_factor(7, 1, 100, 16)
    W = B.init_factor(7, 2, 100, 16)
    assert (U == V).all() and not (U == W).all()
    assert (U >= 0).all() and (U < 1).all() and abs(U.mean() - 0.5) < 0.05


def test_coo_file_round_trip(B, tmp_path):
    t = B.Coo.generate((30, 20, 10), 500, gen_rank=3, value_kind=1, seed=5)
    p = str(tmp_path / "t.bin")
    t.save(p, 1)
    u = B.Coo.load(p, 1)
    assert (u.dims == t.dims).all() and (u.cidx == t.cidx).all() and (u.vals == t.vals).all()
    q = str(tmp_path / "t.tns")
    t.save(q, 0)
    w = B.Coo.load(q, 0)
    assert (w.cidx == t.cidx).all()
    assert np.allclose(w.vals, t.vals, rtol=1e-5)  # %g keeps 6 digits
    # .bin layout: u64 nmodes, dims, nnz, cidx[mode][nnz], vals
    raw = np.fromfile(p, dtype=np.uint64)
    assert raw[0] == 3 and (raw[1:4] == t.dims).all() and raw[4] == t.nnz


def test_no_gpu_here_is_reported_not_crashed(B):
    if B.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = B.encode((5, 6, 7), np.array([[1], [2], [3]], np.uint64), np.ones(1))
    with pytest.raises(B.BlcoError):
        B.DeviceTensor(h)


def test_generator_matches_recorded_hashes(B):
    """The generator is a pure function of its spec (counter-based draws; the
    first unique draws in draw order are kept, whatever the top-up round
    sizes): its output for three recorded specs, one near saturation, is
    unchanged."""
    import hashlib
    import json
    rec = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "generator_hashes.json")))
    for c in rec["cases"]:
        t = B.Coo.generate(tuple(c["dims"]), c["nnz"], zipf=tuple(c["zipf"]), gen_rank=c["gen_rank"],
                           value_kind=c["value_kind"], seed=c["seed"], topup=c["topup"])
        assert t.nnz == c["out_nnz"] == c["nnz"]
        assert hashlib.sha1(t.cidx.tobytes()).hexdigest() == c["cidx_sha1"]
        assert hashlib.sha1(t.vals.tobytes()).hexdigest() == c["vals_sha1"]
