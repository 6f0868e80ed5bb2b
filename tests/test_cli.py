"""The cpd64 command line (the reference's src/main.cpp interface): usage,
the error path without a GPU, and on the GPU a CP-ALS run whose fit equals
the oracle's from the same seed and the -c MTTKRP check."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPD64 = os.path.join(ROOT, "blocked-linearized-coordinate_amd", "bin", "cpd64")


def run(*args, timeout=120):
    return subprocess.run([CPD64, *args], capture_output=True, text=True, timeout=timeout)


def test_usage():
    r = run("-h")
    assert r.returncode == 0
    assert "--kernel-id" in r.stderr and "--stream-data" in r.stderr and "--precision" in r.stderr


def test_without_gpu_fails_cleanly(B, tmp_path):
    if B.device_count() > 0:
        pytest.skip("a GPU is visible")
    t = B.Coo.generate((30, 20, 10), 500, seed=1)
    path = str(tmp_path / "t.bin")
    t.save(path, 1)
    r = run("-b", path, "-r", "4", "-m", "2")
    assert r.returncode != 0
    assert "GPU" in r.stderr


def test_device_budget_flag_parsing():
    r = run("--device-budget", "12Q", "-d", "10,10,10")
    assert r.returncode != 0 and "Invalid --device-budget" in r.stderr
    assert "--device-budget" in run("-h").stderr


@pytest.mark.gpu
def test_device_budget_flag_matches_resident(B, gpu, tmp_path):
    """--device-budget 256K holds 8192 of the 20000 nonzeros beside two
    4096-nonzero slots (16 bytes each); the run's fits equal the resident
    run's."""
    t = B.Coo.generate((300, 200, 100), 20000, gen_rank=5, value_kind=1, seed=9)
    path = str(tmp_path / "t.bin")
    t.save(path, 1)
    common = ("-b", path, "-r", "8", "-m", "4", "-x", "5", "-e", "0", "-k", "10")
    r0 = run(*common)
    r1 = run(*common, "--device-budget", "256K", "--max-block-size", "4096")
    assert r0.returncode == 0 and r1.returncode == 0, (r0.stderr, r1.stderr)
    f0 = float(re.search(r"Final fit: ([0-9.eE+-]+)", r0.stdout).group(1))
    f1 = float(re.search(r"Final fit: ([0-9.eE+-]+)", r1.stdout).group(1))
    assert abs(f0 - f1) <= 1e-9, (f0, f1)


@pytest.mark.gpu
def test_cpals_fit_matches_oracle(B, O, gpu, tmp_path):
    dims, R, iters, seed = (300, 200, 100), 8, 6, 42
    t = B.Coo.generate(dims, 20000, gen_rank=5, value_kind=1, seed=9)
    path = str(tmp_path / "t.bin")
    t.save(path, 1)
    r = run("-b", path, "-r", str(R), "-m", str(iters), "-x", str(seed), "-e", "0", "-k", "10")
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("--> Iter")]
    assert len(lines) == iters
    final = float(re.search(r"Final fit: ([0-9.eE+-]+)", r.stdout).group(1))
    at = O.Alto(dims, t.cidx, t.vals, nprtn=4)
    _, _, fits, _, _ = at.cpd(B.init_factors(seed, dims, R), iters, 0.0)
    assert abs(final - fits[-1]) <= 1e-8, (final, fits[-1])
    per_iter = [float(re.search(r"f = ([0-9.eE+-]+)", l).group(1)) for l in lines]
    assert np.allclose(per_iter, fits, atol=1e-6)  # printed with 6 decimals


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["1", "2", "3", "10", "11", "12", "13", "14", "20"])
def test_check_mode_passes(B, gpu, tmp_path, kernel):
    t = B.Coo.generate((400, 300, 50), 30000, gen_rank=5, seed=3)
    path = str(tmp_path / "t.bin")
    t.save(path, 1)
    r = run("-b", path, "-r", "16", "-c", "-t", "-1", "-k", kernel)
    assert r.returncode == 0, r.stderr
    assert r.stderr.count("Results of mttkrp_alto_dev is correct") == 3, r.stderr


API_CHECK = os.path.join(ROOT, "blocked-linearized-coordinate_amd", "bin", "api_check")


@pytest.mark.gpu
def test_reference_signature_api_on_gpu(gpu):
    """bin/api_check, compiled against the reference-signature C++ headers
    (blco.hpp, alto_dev.hpp, cpd_gpu.hpp, kruskal_model.hpp), calls
    mttkrp_alto_dev_onemode<LIT>, ata_gpu, kruskal_fit_gpu,
    columnwise_normscal, pseudoinverse_gpu and the vector helpers on the GPU
    and checks each on the host (long-double MTTKRP and Grams, the four
    Moore-Penrose conditions, the fit formula)."""
    r = subprocess.run([API_CHECK], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    lines = [l for l in r.stdout.splitlines() if l.startswith("CHECK ")]
    names = {l.split()[1] for l in lines}
    for want in ("device_blocks", "mttkrp_mode0", "mttkrp_mode1", "mttkrp_mode2", "ata_gpu_mode0",
                 "kruskal_fit_gpu", "columnwise_normscal", "pinv_general_AXA", "pinv_rank12_XAX",
                 "pinv_hadamard_of_grams_AX_sym", "fill_hadamard_vecsum"):
        assert want in names, want
    assert all(l.endswith("PASS") for l in lines), r.stdout
    assert r.returncode == 0 and "API CHECK OK" in r.stdout


def _read_model_file(path):
    """A file of ExportKruskalModel (src/kruskal_model.cpp:11-42): 'matrix',
    2, 'rows cols', then the entries column by column; or 'vector', 1, n,
    then the entries."""
    with open(path) as f:
        lines = f.read().split("\n")
    kind = lines[0]
    if kind == "matrix":
        assert lines[1] == "2"
        rows, cols = (int(x) for x in lines[2].split())
        vals = np.array([float(x) for x in lines[3:3 + rows * cols]])
        assert all("." in x and len(x.split(".")[1]) == 20 for x in lines[3:3 + rows * cols])
        return vals.reshape(cols, rows).T
    assert kind == "vector" and lines[1] == "1"
    n = int(lines[2])
    return np.array([float(x) for x in lines[3:3 + n]])


@pytest.mark.gpu
def test_model_export_matches_oracle(B, O, gpu, tmp_path):
    """-o writes the reference's model files (one per mode, column-major,
    %.20lf, plus lambda); their contents are the CP-ALS result, equal to the
    oracle's factors and lambda from the same seed."""
    dims, R, iters, seed = (120, 90, 60), 6, 4, 5
    t = B.Coo.generate(dims, 8000, gen_rank=4, value_kind=1, seed=11)
    path = str(tmp_path / "t.bin")
    t.save(path, 1)
    prefix = str(tmp_path / "model")
    r = run("-b", path, "-r", str(R), "-m", str(iters), "-x", str(seed), "-e", "0", "-k", "10", "-o", prefix)
    assert r.returncode == 0, r.stderr
    assert "--> Kruskal model saved to %s" % prefix in r.stdout
    U = [_read_model_file("%s.%d.out" % (prefix, n)) for n in range(3)]
    lam = _read_model_file("%s.lambda.out" % prefix)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=4)
    Uo, lo, _, _, _ = at.cpd(B.init_factors(seed, dims, R), iters, 0.0)
    for n in range(3):
        assert U[n].shape == (dims[n], R)
        assert np.linalg.norm(U[n] - Uo[n]) <= 1e-8 * np.linalg.norm(Uo[n])
    assert lam.shape == (R,)
    assert np.allclose(lam, lo, rtol=1e-8)


@pytest.mark.gpu
def test_mttkrp_bench_driver_output(B, gpu, tmp_path):
    """-p runs the MTTKRP benchmark driver (src/main.cpp:377,
    src/alto_dev.cu:1856-1952): one 'Start MTTKRP' line and one 'Total time
    for MTTKRP: T (time per iter: T/iters ms)' line per mode, in the
    reference's format."""
    t = B.Coo.generate((300, 200, 100), 20000, gen_rank=5, seed=4)
    path = str(tmp_path / "t.bin")
    t.save(path, 1)
    iters = 3
    r = run("-b", path, "-r", "16", "-p", "-m", str(iters), "-t", "-1", "-k", "10")
    assert r.returncode == 0, r.stderr
    starts = re.findall(r"--> Start MTTKRP mode (\d+) \((\d+) iters\)", r.stdout)
    assert starts == [(str(m), str(iters)) for m in range(3)]
    totals = re.findall(r"Total time for MTTKRP: ([0-9.]+) \(time per iter: ([0-9.]+) ms\)", r.stdout)
    assert len(totals) == 3
    for tot, per in totals:
        tot, per = float(tot), float(per)
        assert tot > 0 and abs(per - tot / iters) <= 1.5e-3
