"""Binary dataset files (src/Utilities.cpp:142-214 format) and the command line's
parameter defaults (src/Utilities.cpp:221-496). No device needed."""
import numpy as np
import pytest

from ceos_amd.cli import parse
from ceos_amd.io import load_bin, rows_in, save_bin


def test_bin_round_trip_and_shape(tmp_path):
    X = np.random.default_rng(1).standard_normal((37, 13)).astype(np.float32)
    p = str(tmp_path / "x.bin")
    save_bin(p, X)
    assert rows_in(p, 13) == 37
    Y = load_bin(p, 20, 13)
    assert Y.shape == (20, 13)  # rows x features (the reference allocates it transposed)
    assert np.array_equal(np.asarray(Y), X[:20])
    assert np.array_equal(np.asarray(load_bin(p, 37, 13)), X)
    with pytest.raises(ValueError):
        load_bin(p, 38, 13)


def test_cli_defaults_follow_reference():
    a = parse("--X x --Q q --n_points 100 --n_features 300 --n_queries 5".split())
    assert a.n_proj == 512 and a.n_repeats == 1 and a.top_m == 1000 and a.n_threads == -1
    assert a.random_seed == -1 and not a.centering
    assert a.n_neighbors == 1 and a.n_probed_vectors == 10 and a.n_probed_points == 10 and a.n_cand == 1
    b = parse("--X x --Q q --n_points 100 --n_features 100 --n_queries 5 --n_neighbors 7".split())
    assert b.n_proj == 256 and b.n_cand == 7
