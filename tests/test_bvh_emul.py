"""CPU restatement of the GPU broad phase's tree logic (vh_contact.hpp:
k_karras, escape_of, k_bvh_pairs): Karras' internal nodes from sorted
30-bit keys with index tie-breaks, and the stackless pre-order walk through
escapes. On random key sets (many duplicates included) every walk ends and
the pairs found equal the brute-force set of overlapping intervals."""
import numpy as np

END = 0x7FFFFFFF


def _clz32(v):
    return 32 - int(v).bit_length() if v else 32


def _delta(code, t, i, j):
    if j < 0 or j >= t:
        return -1
    a, b = int(code[i]), int(code[j])
    if a != b:
        return _clz32(a ^ b)
    return 32 + _clz32(i ^ j)


def _karras(code):
    t = len(code)
    left, rlo, rhi = [0] * (t - 1), [0] * (t - 1), [0] * (t - 1)
    for i in range(t - 1):
        d = 1 if _delta(code, t, i, i + 1) - _delta(code, t, i, i - 1) >= 0 else -1
        dmin = _delta(code, t, i, i - d)
        lmax = 2
        while _delta(code, t, i, i + lmax * d) > dmin:
            lmax <<= 1
        l, s = 0, lmax >> 1
        while s >= 1:
            if _delta(code, t, i, i + (l + s) * d) > dmin:
                l += s
            s >>= 1
        j = i + l * d
        dnode = _delta(code, t, i, j)
        sp, s = 0, (l + 1) >> 1
        while True:
            if _delta(code, t, i, i + (sp + s) * d) > dnode:
                sp += s
            if s == 1:
                break
            s = (s + 1) >> 1
        g = i + sp * d + (d if d < 0 else 0)
        lo, hi = min(i, j), max(i, j)
        left[i] = ~g if lo == g else g
        rlo[i], rhi[i] = lo, hi
    return left, rlo, rhi


def _escape(hi, t, rlo):
    if hi >= t - 1:
        return END
    return hi + 1 if (hi + 1 <= t - 2 and rlo[hi + 1] == hi + 1) else ~(hi + 1)


def test_rope_walk_equals_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(200):
        t = int(rng.integers(2, 100))
        code = np.sort(rng.integers(0, int(rng.choice([4, 64, 1 << 30])), size=t))
        left, rlo, rhi = _karras(code)
        c, w = rng.random(t), rng.random(t) * 0.2
        blo, bhi = c - w, c + w
        nlo = [blo[rlo[i]:rhi[i] + 1].min() for i in range(t - 1)]
        nhi = [bhi[rlo[i]:rhi[i] + 1].max() for i in range(t - 1)]
        def walk(k, start):
            out, nd, steps = [], start, 0
            while nd != END:
                steps += 1
                assert steps < 4 * t
                if nd >= 0:
                    enter = rhi[nd] > k and blo[k] <= nhi[nd] and bhi[k] >= nlo[nd]
                    nd = left[nd] if enter else _escape(rhi[nd], t, rlo)
                else:
                    j = ~nd
                    if j > k and blo[k] <= bhi[j] and bhi[k] >= blo[j]:
                        out.append((k, j))
                    nd = _escape(j, t, rlo)
            return out, steps

        got = set()
        for k in range(t):
            # The kernel starts at leaf k's escape: the nodes after leaf k in
            # pre-order are the right subtrees of its ancestors, exactly the
            # ranges past k, so the walk from the root finds the same pairs
            # in the same order after descending to k.
            from_root, s_root = walk(k, 0)
            from_leaf, s_leaf = walk(k, _escape(k, t, rlo))
            assert from_leaf == from_root and s_leaf <= s_root
            got.update(from_leaf)
        want = {(a, b) for a in range(t) for b in range(a + 1, t) if blo[a] <= bhi[b] and bhi[a] >= blo[b]}
        assert got == want
