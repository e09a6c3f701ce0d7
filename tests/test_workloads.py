"""The synthetic request streams: the 64-bit Mersenne Twister against the C++
standard's known answer and against the compiler's std::mt19937_64, the
committed manifests against a fresh generation, and the draws' ranges."""
import math
import os
import shutil
import subprocess

import pytest

import workloads as W


def test_mt19937_64_known_answer():
    # [rand.predef]: the 10000th consecutive invocation of a default-constructed
    # mt19937_64 produces 9981545732273789042
    g = W.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_mt19937_64_matches_the_standard_library(tmp_path):
    src = tmp_path / "mt.cpp"
    src.write_text("#include <cstdio>\n#include <random>\nint main(){for(unsigned long long s: {0ull, 1ull}){"
                   "std::mt19937_64 g(s); for(int i=0;i<700;++i) std::printf(\"%llu\\n\", (unsigned long long)g());}}\n")
    exe = tmp_path / "mt"
    subprocess.run(["g++", "-O1", str(src), "-o", str(exe)], check=True)
    expect = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    ours = []
    for seed in (0, 1):
        g = W.MT19937_64(seed)
        ours += [g() for _ in range(700)]
    assert ours == expect


def test_committed_manifests_are_the_generated_ones():
    assert W.load(W.MIXED_BATCH_FILE) == W.mixed_batch(0, 64)
    assert W.load(W.SERVING_FILE) == W.serving_stream(1, 20.0, 600.0)


def test_draws_stay_in_their_ranges():
    batch = W.load(W.MIXED_BATCH_FILE)
    stream = W.load(W.SERVING_FILE)
    for r in batch + stream:
        assert 0 <= r["m"] <= 6 and 8 <= r["t"] <= 18
        assert (r["p"] == 0) if r["m"] == 0 else (2 <= r["p"] <= 6)
        assert -5.0 <= r["alpha"] <= 12.0
        assert 2e5 <= r["reynolds"] <= 5e6 * (1 + 1e-12)
        assert r["level"] in W.LEVELS
    assert {r["level"] for r in batch} == set(W.LEVELS)
    arrivals = [r["arrival"] for r in stream]
    assert all(a < b for a, b in zip(arrivals, arrivals[1:])) and 0.0 < arrivals[0] and arrivals[-1] <= 600.0
    # Poisson count: mean 12,000, standard deviation ~110
    assert abs(len(stream) - 12000) < 5 * math.sqrt(12000)


def test_levels_give_the_stated_cgrid_sizes(G):
    for level, cells in (("coarse", (256, 64)), ("medium", (512, 128))):
        mesh = G.cgrid(0, 0, 12, options=W.mesh_options(G, level, 1e6))
        assert (mesh.ni, mesh.nj) == cells
    assert W.first_layer(1e6) == 2.5e-5
