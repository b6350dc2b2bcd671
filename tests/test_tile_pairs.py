"""The dense near-field kernels' walk over a tile (TilePairs in
3d_bem_amd/csrc/bem_hip.hip): thread t of a 256-thread workgroup starts at
pair t of the tile's column-major (row, column) order and steps by 256
pairs with integer additions only.  Restated here with the same operations,
it must visit every pair of every tile shape (1..64 rows x 1..64 columns)
exactly once, never leave the tile, and keep consecutive threads on
consecutive rows of one column except where a column ends."""
import numpy as np

KWG = 256
KTILE = 64


def walk(tid, rows, cols):
    """the pairs thread tid visits, as the device loop forms them"""
    c = tid // rows
    r = tid - c * rows
    dc = KWG // rows
    dr = KWG - dc * rows
    out = []
    p = tid
    while p < rows * cols:
        out.append((r, c))
        p += KWG
        r += dr
        c += dc
        if r >= rows:
            r -= rows
            c += 1
    return out


def test_every_pair_once_for_every_tile_shape():
    for rows in range(1, KTILE + 1):
        for cols in range(1, KTILE + 1):
            seen = np.zeros((rows, cols), np.int32)
            for tid in range(KWG):
                for r, c in walk(tid, rows, cols):
                    assert 0 <= r < rows and 0 <= c < cols, (rows, cols, tid, r, c)
                    seen[r, c] += 1
            assert np.all(seen == 1), (rows, cols)


def test_pair_index_matches_column_major_order():
    """thread tid's k-th pair is pair tid + 256 k of the column-major order"""
    rng = np.random.default_rng(0)
    for _ in range(200):
        rows, cols = (int(x) for x in rng.integers(1, KTILE + 1, 2))
        for tid in range(0, KWG, 7):
            for k, (r, c) in enumerate(walk(tid, rows, cols)):
                p = tid + KWG * k
                assert (r, c) == (p % rows, p // rows)


def test_lane_slots_busy_on_the_sphere_tiles():
    """the busy fraction the pair order gives a 64-lane wavefront on tiles
    of the usual near-field sizes, against one row per lane"""
    shapes = [(50, 50), (63, 67 - 64), (3, 64), (57, 43), (64, 64)]
    for rows, cols in shapes:
        pairs = rows * cols
        pair_rounds = -(-pairs // 64)             # wavefront-rounds in pair order
        row_rounds = sum(len(range(cg, cols, 4)) for cg in range(4))   # one row per lane
        assert pair_rounds <= row_rounds
        assert pairs / (64 * pair_rounds) > 0.75 or pairs < 64
