"""The oracle's k-resolved random g factors (gfac_normal_dist;
src_v4.5.7/electron.cpp:7-18, comm/myarray.cpp:299-321, comm/Random.cpp:24-45).

Pinned three ways: the engine against the C++ standard's known answer for
mt19937_64 ([rand.predef]: the 10000th output of a default-seeded engine is
9981545732273789042); the normal draws against the host C++ library's own
std::normal_distribution over std::mt19937_64 (a program compiled here with
g++, as the reference's Random.cpp is); and random_normal_array's shift and
scale against the raw draws: with no cap, (g - mean) / sigma is the draws
standardised with their own mean and population sigma, and with a cap the
draws with |sigma z| > cap are the ones skipped."""
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle


def test_mt19937_64_known_answer():
    out = oracle.mt19937_64(5489, 10000)
    assert int(out[-1]) == 9981545732273789042


_PROG = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char **argv) {
    std::mt19937_64 g(std::strtoull(argv[1], nullptr, 10));
    std::normal_distribution<double> z;
    long n = std::atol(argv[2]);
    for (long i = 0; i < n; i++) std::printf("%a\n", z(g));
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("seed,n", [(7, 9), (4096, 5000), (12345, 2001)])
def test_normal_draws_match_the_standard_library(tmp_path, seed, n):
    src = tmp_path / "draw.cpp"
    src.write_text(_PROG)
    exe = tmp_path / "draw"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe), str(src)], check=True)
    r = subprocess.run([str(exe), str(seed), str(n)], capture_output=True, text=True, check=True)
    lib = np.array([float.fromhex(x) for x in r.stdout.split()])
    ours = oracle.normal_draws(seed, n)
    assert np.array_equal(lib, ours)


@pytest.mark.parametrize("n,mean,sigma", [(2, 2.0, 0.1), (97, 2.0, 0.3), (4096, 1.7, 0.05)])
def test_gfac_is_the_standardised_draws_of_a_generator_seeded_with_nk(n, mean, sigma):
    g = oracle.gfac_normal(n, mean, sigma, 0.0)
    z = oracle.normal_draws(n, n)
    zs = (z - z.mean()) / np.sqrt(((z - z.mean()) ** 2).mean())
    assert np.abs((g - mean) / sigma - zs).max() < 1e-12
    assert abs(g.mean() - mean) < 1e-13 * mean
    assert abs(np.sqrt(((g - mean) ** 2).mean()) - sigma) < 1e-13


def test_gfac_cap_skips_the_draws_beyond_it():
    n, mean, sigma, cap = 500, 2.0, 0.4, 0.5
    g = oracle.gfac_normal(n, mean, sigma, cap)
    z = oracle.normal_draws(n, 4 * n)
    kept = z[np.abs(sigma * z) <= cap][:n]
    assert len(kept) == n and (np.abs(sigma * z) > cap).any()
    d = sigma * kept
    zs = (d - d.mean()) / np.sqrt(((d - d.mean()) ** 2).mean())
    assert np.abs((g - mean) / sigma - zs).max() < 1e-12


def test_gfac_special_cases():
    assert np.array_equal(oracle.gfac_normal(1, 2.1, 0.3, 0.0), [2.1])
    assert np.array_equal(oracle.gfac_normal(50, 2.0, 0.0, 0.0), np.full(50, 2.0))
    assert oracle.gfac_normal(0, 2.0, 0.3, 0.0).shape == (0,)
