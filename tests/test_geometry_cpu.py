"""The block decomposition of a domain box (freelb_amd.geometry): FreeLB's
division rules restated (AABB::divide's remainder placement, DivideBlock3D's
most cube-like factorisation, DivideBlock2D's closest aspect ratio), every
domain cell in exactly one block, and cut / paste round trips, periodic axes
included."""
import numpy as np
import pytest

from freelb_amd.geometry import BlockGeometry, block_counts, divide_axis


def test_divide_axis_puts_the_remainder_last():
    # AABB::divide (geometry/basic_geometry.hh:192-250): parts of n // N cells,
    # the last n % N parts one longer
    assert divide_axis(10, 3) == [(0, 3), (3, 6), (6, 10)]
    assert divide_axis(11, 3) == [(0, 3), (3, 7), (7, 11)]
    assert divide_axis(9, 3) == [(0, 3), (3, 6), (6, 9)]
    with pytest.raises(ValueError):
        divide_axis(2, 3)


def test_block_counts():
    assert block_counts((100, 100, 100), 8) == (2, 2, 2)
    assert block_counts((100, 100, 100), 1) == (1, 1, 1)
    assert block_counts((64, 32, 16), 8) == (4, 2, 1)
    assert block_counts((40, 20, 1), 8) == (4, 2, 1)
    assert block_counts((30, 30, 1), 4) == (2, 2, 1)
    assert block_counts((64, 64, 64), 27) == (3, 3, 3)
    assert block_counts((96, 48, 48), 2) == (2, 1, 1)
    assert block_counts((32, 32, 128), 4) == (1, 1, 4)
    # DivideBlock3D settles on 5 x 3 x 1 here and splits the remaining block
    # in its "rest" branches, which are not reproduced
    with pytest.raises(ValueError):
        block_counts((48, 24, 12), 16)
    with pytest.raises(ValueError):
        block_counts((40, 20, 1), 7)


@pytest.mark.parametrize("dims,counts,periodic", [((23, 17, 13), (2, 3, 2), (False, False, False)),
                                                  ((40, 9, 7), (3, 1, 1), (True, False, False)),
                                                  ((21, 14, 1), (3, 2, 1), (False, False, False)),
                                                  ((12, 10, 9), (2, 2, 3), (True, True, True))])
def test_every_cell_in_one_block_and_round_trip(dims, counts, periodic):
    g = BlockGeometry(dims, counts=counts, periodic=periodic)
    assert len(g) == int(np.prod(counts))
    N = int(np.prod(g.padded))
    a = np.arange(N, dtype=np.float64)
    cover = np.zeros(N, dtype=int)
    parts = []
    for k in range(len(g)):
        p = g.cut(a, k)
        parts.append(p)
        bd = g.block_dims(k)
        inner = np.zeros(bd[::-1], dtype=bool)
        if g.d == 3:
            inner[1:-1, 1:-1, 1:-1] = True
        else:
            inner[0, 1:-1, 1:-1] = True
        cover[p[0][inner.reshape(-1)].astype(int)] += 1
    domain = np.zeros(g.padded[::-1], dtype=bool)
    sl = [slice(None) if periodic[a] else slice(1, -1) for a in range(3)][::-1]
    if g.d == 2:
        sl[0] = slice(None)
    domain[tuple(sl)] = True
    assert (cover[domain.reshape(-1)] == 1).all() and (cover[~domain.reshape(-1)] == 0).all()
    back = g.paste(parts, dtype=np.float64)
    assert np.array_equal(back[0][domain.reshape(-1)], a[domain.reshape(-1)])
    # a block's ghost layer holds its neighbours' cells (wrapped on periodic axes)
    for k in range(len(g)):
        lo, hi = g.boxes[k]
        for ax in range(3):
            if g.d == 2 and ax == 2:
                continue
            idx = np.arange(lo[ax], hi[ax])
            if periodic[ax]:
                idx = idx % g.padded[ax]
            assert idx.min() >= 0 and idx.max() < g.padded[ax]
