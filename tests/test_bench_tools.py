"""CPU checks of bench.py's offline tools: the f_c bracket comparison of two
saved depinning searches (--rk-compare) and the 128-byte line count of the
evaluation walk under a table layout (--count-lines), against direct
enumerations on small inputs."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_rk_compare_counts_brackets(tmp_path):
    lo = np.array([0.50, 0.51, 0.52, 0.53])
    hi = lo + 1e-3
    # line 0 identical, line 1 overlapping, line 2 disjoint above, line 3 narrower inside
    lo2 = np.array([0.50, 0.5105, 0.5215, 0.5301])
    hi2 = np.array([0.501, 0.5115, 0.5225, 0.5302])
    a, b = tmp_path / "a.npz", tmp_path / "b.npz"
    np.savez(a, lo=lo, hi=hi, h_lo=np.zeros((4, 2)))
    np.savez(b, lo=lo2, hi=hi2, h_lo=np.zeros((4, 2)))
    out = bench.rk_compare(str(a), str(b))
    assert out["ref"]["lines"] == 4 and out["other"]["lines"] == 4
    assert out["other_vs_ref"]["identical_brackets"] == 1
    assert out["other_vs_ref"]["brackets_overlap"] == 3
    fa, fb = 0.5 * (lo + hi), 0.5 * (lo2 + hi2)
    assert np.isclose(out["other_vs_ref"]["f_c_rel_err_max"], np.max(np.abs(fb - fa) / fa))
    width = (hi2 - lo2) / hi2
    assert np.isclose(out["other"]["bracket_rel_width_max"], width.max())
    assert out["ref"]["lines_at_1e-10"] == 0


def test_count_lines_matches_enumeration():
    rng = np.random.default_rng(4)
    n = nx = 256
    ny = 64
    h = np.mod(ny / 2 + np.cumsum(rng.standard_normal(n)) * 3.0, ny)
    ix, iy = bench.edge_cells(h, n, nx, ny)
    for tile in (None, (16, 2), (8, 4)):
        seen = set()
        for k in range(n):
            for a in (-1, 0, 1, 2):
                for b in (-1, 0, 1, 2):
                    x, y = (ix[k] + a) % nx, (iy[k] + b) % ny
                    if tile is None:
                        seen.add((y, x // 32))
                    else:
                        seen.add((y // tile[1], x // tile[0]))
        assert bench.count_lines(h, n, nx, ny, 4, tile, False) == len(seen), tile
    # midpoint table: one column per edge
    seen = {((iy[k] + b) % ny, ix[k] // 32) for k in range(n) for b in (-1, 0, 1, 2)}
    assert bench.count_lines(h, n, nx, ny, 4, None, True) == len(seen)
