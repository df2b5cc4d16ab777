"""Index files: a saved and reloaded index returns what the saved one returned (static
coCEOs and CEOs-Hash; streamCEOs including random deletes and later updates)."""
import numpy as np
import pytest

from tests._parity import bits

pytestmark = pytest.mark.gpu


def _rand(n, d, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, d)) * rng.lognormal(0, 0.5, (n, 1))).astype(np.float32)


def test_save_load_static(CEOs, tmp_path):
    X = _rand(12000, 56, 91)
    Q = _rand(50, 56, 92)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(256, 3, 90, -1, 7)
    g.build_coCEOs_Est(X)
    g.n_probed_vectors, g.n_probed_points, g.n_cand = 20, 70, 400
    a = g.search_coCEOs_Est(Q, 10)
    h = g.search_CEOs_Hash(Q, 10)
    path = str(tmp_path / "idx.ceos")
    g.save(path)
    r = CEOs.CEOs(1, 56)
    r.load(path)
    assert (r.n_probed_vectors, r.n_probed_points, r.n_cand) == (20, 70, 400)
    assert np.array_equal(r.get_sign_bits(), g.get_sign_bits())
    for x, y in zip(r.get_buckets(), g.get_buckets()):
        assert np.array_equal(x, y)
    b = r.search_coCEOs_Est(Q, 10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(bits(a[1]), bits(b[1]))
    hb = r.search_CEOs_Hash(Q, 10)
    assert np.array_equal(h[0], hb[0])
    with pytest.raises(ValueError):
        CEOs.CEOs(1, 57).load(path)  # wrong n_features
    with pytest.raises(ValueError):
        CEOs.streamCEOs(56).load(path)  # a CEOs file


def test_save_load_stream(CEOs, oracle, tmp_path):
    d, D, R, m = 40, 128, 2, 45
    X = _rand(9000, d, 93)
    Q = _rand(30, d, 94)
    g = CEOs.streamCEOs(d)
    g.setIndexParam(D, R, m, -1, 11)
    g.build(X[:4000])
    g.update(X[4000:5000], 1500)  # the ring now starts mid-buffer
    g.delete_ids(np.array([3, 10, 200, 2000], np.int32))
    g.n_probed_vectors, g.n_probed_points, g.n_cand = 12, m, 120
    a = g.estimate_search(Q, 10)
    path = str(tmp_path / "stream.ceos")
    g.save(path)
    r = CEOs.streamCEOs(d)
    r.load(path)
    assert r.size() == g.size()
    b = r.estimate_search(Q, 10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(bits(a[1]), bits(b[1]))
    # both go on identically
    for obj in (g, r):
        obj.update(X[5000:6500], 700)
    a = g.estimate_search(Q, 10)
    b = r.estimate_search(Q, 10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(bits(a[1]), bits(b[1]))


def test_cli_end_to_end(CEOs, tmp_path):
    """python -m ceos_amd.cli on .bin files returns what the module returns."""
    from ceos_amd.cli import parse, run
    from ceos_amd.io import save_bin

    X = _rand(3000, 20, 95)
    Q = _rand(12, 20, 96)
    px, pq = str(tmp_path / "x.bin"), str(tmp_path / "q.bin")
    save_bin(px, X)
    save_bin(pq, Q)
    a = parse(f"--X {px} --Q {pq} --n_points 3000 --n_features 20 --n_queries 12 --n_proj 64 --n_repeats 2 "
              f"--top_m 100 --random_seed 3 --n_neighbors 5 --n_probed_vectors 8 --n_probed_points 50 "
              f"--n_cand 60 --method coceos,hash".split())
    out = run(a)
    g = CEOs.CEOs(3000, 20)
    g.setIndexParam(64, 2, 100, -1, 3)
    g.build_coCEOs_Est(X)
    g.n_probed_vectors, g.n_probed_points, g.n_cand = 8, 50, 60
    assert np.array_equal(out["ids_search_coCEOs_Est"], g.search_coCEOs_Est(Q, 5)[0])
    assert np.array_equal(out["ids_search_CEOs_Hash"], g.search_CEOs_Hash(Q, 5)[0])
