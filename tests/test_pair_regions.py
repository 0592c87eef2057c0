"""The survivor regions of the pair culls (k_self_cull, k_ccd_cull), and the
broad phase's single walk (k_bvh_walk1, k_bvh_compact, k_bvh_more).

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


# ---- the broad phase's single walk (k_bvh_walk1, k_bvh_compact, k_bvh_more) ----

_STAGE, _KEEP, _MORE = 4096, 16, 48  # kPairStage, kLeafStage, the host's wmore_leaf


def _single_walk_list(cnt, c2=_MORE, more_size=None):
    """Replays the index arithmetic of the three kernels for per-leaf pair
    counts cnt (pair c of leaf k encoded as k * 4096 + c) and returns the
    pair list they leave, asserting every index in bounds. more_size: the
    entries of the second part (the host allocates t * c2)."""
    t = len(cnt)
    nch = (t + _BLOCK - 1) // _BLOCK
    per = (nch + _X - 1) // _X
    grid = (nch + 7) // 8 * 8
    kept = np.full(nch * _STAGE, -1, np.int64)
    more = np.full(t * c2 if more_size is None else more_size, -1, np.int64)
    walk_more = False
    for b in range(grid):  # k_bvh_walk1, leaf_chunk
        ch = (b % _X) * per + b // _X
        if ch >= nch:
            continue
        stage = np.full(_STAGE, -2, np.int64)
        for tid in range(_BLOCK):
            k = ch * _BLOCK + tid
            if k >= t:
                continue
            for c in range(cnt[k]):
                if c < _KEEP:
                    assert tid * _KEEP + c < _STAGE
                    stage[tid * _KEEP + c] = k * 4096 + c
                elif c - _KEEP < c2:
                    assert k * c2 + c - _KEEP < more.size
                    more[k * c2 + c - _KEEP] = k * 4096 + c
            walk_more |= cnt[k] > _KEEP + c2
        assert (ch + 1) * _STAGE <= kept.size
        kept[ch * _STAGE:(ch + 1) * _STAGE] = stage
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    pairs = np.full(int(np.sum(cnt)), -3, np.int64)
    for b in range(nch):  # k_bvh_compact
        k0, kn = b * _BLOCK, min(_BLOCK, t - b * _BLOCK)
        so = list(off[k0:k0 + kn]) + [off[k0 + kn - 1] + cnt[k0 + kn - 1]]
        for o in range(so[0], so[kn]):
            lo, hi = 0, kn - 1
            while lo < hi:
                mid = (lo + hi + 1) >> 1
                if so[mid] <= o:
                    lo = mid
                else:
                    hi = mid - 1
            if so[lo + 1] - so[lo] > _KEEP + c2:
                continue
            j, k = o - so[lo], k0 + lo
            assert j < cnt[k]
            pairs[o] = kept[k * _KEEP + j] if j < _KEEP else more[k * c2 + j - _KEEP]
    if walk_more:  # k_bvh_more: the leaves with more pairs, walked again
        for k in np.nonzero(np.asarray(cnt) > _KEEP + c2)[0]:
            pairs[off[k]:off[k] + cnt[k]] = k * 4096 + np.arange(cnt[k])
    return pairs, walk_more


@pytest.mark.parametrize("t,mean", [(1, 3), (255, 25), (257, 25), (3000, 25), (2048, 60), (700, 0)])
def test_single_walk_keeps_the_pair_list(t, mean):
    rng = np.random.default_rng(t + mean)
    cnt = rng.poisson(mean, size=t) if mean else np.zeros(t, np.int64)
    cnt[::97] += 70  # a few leaves with more pairs than the walk keeps
    pairs, _ = _single_walk_list(cnt)
    want = np.concatenate([k * 4096 + np.arange(c) for k, c in enumerate(cnt)])
    assert np.array_equal(pairs, want)


def test_single_walk_check_catches_a_short_overflow_run():
    """A second part one entry short of t * c2 (a planted mistake): the
    replay's bounds check fails at the last leaf that fills its run."""
    cnt = np.full(300, _KEEP + _MORE, np.int64)
    _single_walk_list(cnt)
    with pytest.raises(AssertionError):
        _single_walk_list(cnt, more_size=300 * _MORE - 1)
