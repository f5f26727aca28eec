"""The oracle's draw semantics against libstdc++ itself (CPU).

The reference draws CP-ARLS-LEV rows through std::discrete_distribution and
STS-CP uniforms through std::uniform_real_distribution, both on
generate_canonical<double, 53> of two mt19937 words (cp_arls_lev.hpp:106,145;
sts_cp.hpp:198; random_util.hpp:17).  A small C++ program
(tests/libstdcxx_replay.cpp, compiled here with the system g++) replays
chosen 32-bit words into the real library; the oracle maps the same words
(the Philox words it draws from) to the same doubles and indices, bit for
bit.  This pins the lower_bound / last-entry-is-one / fewer-than-two-weights
semantics that a restatement alone leaves open."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import normalized_factors

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path_factory.mktemp("replay") / "replay")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "libstdcxx_replay.cpp"), "-o", exe], check=True)

    def run(weights, words):
        text = f"{len(weights)} " + " ".join(repr(float(x)) for x in weights)
        text += f" {len(words)} " + " ".join(f"{int(a)} {int(b)}" for a, b in words)
        out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
        rows = [r.split() for r in out if r.strip()]
        return (np.array([int(r[0]) for r in rows]), np.array([float.fromhex(r[1]) for r in rows]),
                np.array([int(r[2]) for r in rows]))
    return run


def philox_words(orc, seed, stream, first, count):
    out = []
    for i in range(first, first + count):
        w = orc.philox4x32([i & 0xFFFFFFFF, i >> 32, stream & 0xFFFFFFFF, stream >> 32],
                           [seed & 0xFFFFFFFF, seed >> 32])
        out.append((int(w[0]), int(w[1])))
    return out


def test_canonical_doubles_match_libstdcxx(orc, replay):
    seed, stream = 0x9E3779B97F4A7C15, orc.stream_id(7, 0, 2, 2)
    words = philox_words(orc, seed, stream, 0, 2000)
    # extremes: all-zero words, and all-ones words (rounds to 1.0, clamped below it)
    words += [(0, 0), (0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0), (0, 0x80000000)]
    _, u_lib, used = replay([1.0, 1.0], words)
    assert np.all(used == 2)
    u_orc = orc.uniform(seed, stream, 0, 2000)
    assert np.array_equal(u_orc, u_lib[:2000])
    assert u_lib[-3] == np.nextafter(1.0, 0.0)
    assert u_lib[-4] == 0.0


@pytest.mark.parametrize("case", ["leverage", "zeros", "two", "one"])
def test_arls_indices_match_discrete_distribution(orc, replay, case):
    rng = np.random.default_rng(31)
    if case == "leverage":
        U = normalized_factors(rng, [400, 3], 6)[0]
        weights = orc.arls_scores(U)
    elif case == "zeros":
        weights = rng.random(50)
        weights[[0, 7, 8, 49]] = 0.0
    elif case == "two":
        weights = np.array([0.25, 0.75])
    else:
        weights = np.array([3.0])
    seed, stream = 0xC0FFEE, orc.stream_id(2, 1, 0, 0)
    J = 3000
    words = philox_words(orc, seed, stream, 0, J)
    idx_lib, u_lib, used = replay(weights, words)
    # one weight: index 0 and no word consumed; otherwise two words per draw
    assert np.all(used == (0 if len(weights) < 2 else 2))
    u = orc.uniform(seed, stream, 0, J)
    other = np.full(J, 0.5)
    draws = np.stack([u, other])
    s, _, _ = orc.arls_draw([weights, np.ones(2)], 1, J, draws)
    assert np.array_equal(s[:, 0].astype(np.int64), idx_lib)
    if case == "zeros":
        assert not np.isin(idx_lib, [0, 7, 8, 49]).any()
