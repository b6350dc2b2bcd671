"""The lane-spread reduction of the ACA's norm-estimate dot products
(wave_sum_spread in 3d_bem_amd/csrc/bem_hip.hip) restated over 64 emulated
lanes: each of its additions pairs the same two partials as the pairwise
tree of the device reduction order (partial p plus partial p + s, s = 32,
16, ..., 1, the order oracle/bem_oracle.cpp det_dot_conj follows), so the
totals are the same bits for 2, 4, 8 and 16 values per lane; lane
(64 / NV) v holds the total of value v.  The device code itself is pinned by
the GPU tests' bit-identical ACA factors and ranks against the oracle."""
import numpy as np
import pytest


def tree(x):
    """x: (64,) partials -> the pairwise tree total"""
    p = x.copy()
    s = 32
    while s:
        p[:s] = p[:s] + p[s:2 * s]
        s >>= 1
    return p[0]


def spread(x):
    """x: (64, NV) -> (64,) lane results of the spread reduction"""
    cur = x.copy()
    nv = x.shape[1]
    lanes = np.arange(64)
    s = 32
    while s:
        partner = lanes ^ s
        if nv > 1:
            h = nv // 2
            hi = (lanes & s) != 0
            keep = np.where(hi[:, None], cur[:, h:nv], cur[:, :h])
            send = np.where(hi[:, None], cur[:, :h], cur[:, h:nv])
            cur = keep + send[partner]
            nv = h
        else:
            cur = cur + cur[partner]
        s >>= 1
    return cur[:, 0]


@pytest.mark.parametrize("nv", [2, 4, 8, 16])
def test_spread_equals_tree(nv):
    rng = np.random.default_rng(nv)
    for _ in range(400):
        x = rng.uniform(-1, 1, (64, nv)) * 2.0 ** rng.integers(-30, 30, (64, nv))
        got = spread(x)
        stride = 64 // nv
        for v in range(nv):
            want = tree(x[:, v])
            assert got[stride * v].tobytes() == np.float64(want).tobytes(), (nv, v)


def test_tree_differs_from_sequential_sum():
    """the check has teeth: a plain left-to-right sum differs from the tree
    on such data, so an order slip would show"""
    rng = np.random.default_rng(1)
    diff = 0
    for _ in range(200):
        x = rng.uniform(-1, 1, 64) * 2.0 ** rng.integers(-30, 30, 64)
        seq = 0.0
        for v in x:
            seq += v
        diff += seq != tree(x)
    assert diff > 0
