"""The survivor regions of the pair culls (k_self_cull, k_ccd_cull).

Block b of a cull writes the surviving pairs of its share to its own region
of ccd_region(cap) entries, with no bounds check in the kernel. This replays
the index arithmetic of pair_blocks and ccd_region (vh_api.hip) and of
pair_range (vh_contact.hpp) on the CPU. For every capacity and pair count the
shares cover each pair exactly once, and no block visits more pairs than its
region holds. A region one round short is caught."""
import numpy as np
import pytest

_MAX_BLOCKS, _X, _BLOCK = 4096, 8, 256


def pair_blocks(cap):
    b = max(1, min((cap + _BLOCK - 1) // _BLOCK, _MAX_BLOCKS))
    return (b + 7) // 8 * 8


def ccd_region(cap):
    span = pair_blocks(cap) // _X * _BLOCK
    per = (cap + 7) // 8
    return max(1, (per + span - 1) // span) * _BLOCK


def block_rounds(np_, grid):
    """Per block: the pair positions of its rounds (pair_range; each round
    covers up to 256 consecutive positions)."""
    nbx, per = grid // _X, (np_ + _X - 1) // _X
    out = []
    for blk in range(grid):
        x, q = blk % _X, blk // _X
        b = min(np_, x * per)
        e = min(np_, b + per)
        starts = list(range(b + q * _BLOCK, e, nbx * _BLOCK))
        out.append([(s, min(s + _BLOCK, e)) for s in starts])
    return out


_CAPS = [1, 255, 256, 257, 2048, 2049, 99_999, 1 << 20, 8_388_608 + 5, 16_777_216, 24_000_017]


@pytest.mark.parametrize("cap", _CAPS)
def test_regions_hold_every_block_share(cap):
    grid, reg = pair_blocks(cap), ccd_region(cap)
    assert grid % _X == 0 and grid <= _MAX_BLOCKS
    for np_ in sorted({0, 1, cap // 3, cap - 1, cap}):
        rounds = block_rounds(np_, grid)
        # a round appends at most 256 survivors: the bound the kernel relies on
        assert max(len(r) for r in rounds) * _BLOCK <= reg, (cap, np_)
        if np_ <= 300_000:  # exact cover, checked where it is cheap
            seen = np.zeros(np_, dtype=np.int32)
            for r in rounds:
                for s, e in r:
                    seen[s:e] += 1
            assert (seen == 1).all(), (cap, np_)


def test_a_short_region_is_caught():
    cap = 16_777_216
    rounds = block_rounds(cap, pair_blocks(cap))
    short = ccd_region(cap) - _BLOCK
    assert max(len(r) for r in rounds) * _BLOCK > short
