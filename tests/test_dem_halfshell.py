"""The bucket walks of the 27-cell touching probe (dem.hip, k_bp_probe_touch_sorted) list the
same touching pairs as the reference's probe (bpcd.py:114-152 with all 27 cells: particle i
lists j > i from each of its distinct probed buckets), each once.  A numpy restatement of the
two walks, checked by brute force on small random packings, together with the two facts they
rest on: from a table of 64 entries on the 27 probed hashes are distinct, and the probed
relation is symmetric for every table size.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O

OFFSETS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]  # (dx, dy, dz), probe order
# the half shell: d > 0 with z first, then y, then x
FORWARD = [d for d in OFFSETS if (d[2], d[1], d[0]) > (0, 0, 0)]


def _spread10(v):
    x = np.asarray(v, dtype=np.int64) & 0x3FF
    x = (x | x << 16) & 0x30000FF
    x = (x | x << 8) & 0x300F00F
    x = (x | x << 4) & 0x30C30C3
    x = (x | x << 2) & 0x9249249
    return x


def hash3(c, T):
    c = np.asarray(c, dtype=np.int64)
    return (_spread10(c[..., 0]) | _spread10(c[..., 1]) << 1 | _spread10(c[..., 2]) << 2) & (T - 1)


def test_hash_is_the_reference_morton_code():
    rng = np.random.default_rng(0)
    c = rng.integers(-50, 2000, (200, 3))
    T = 1 << 30
    assert all(int(hash3(v, T)) == O.morton3d32(*v) % T for v in c)


@pytest.mark.parametrize("T", [32, 64, 128, 256, 1 << 20])
def test_probed_hashes_distinct_from_64_entries(T):
    rng = np.random.default_rng(1)
    c = rng.integers(-3, 1500, (500, 3))
    h = hash3(c[:, None, :] + np.array(OFFSETS)[None], T)
    distinct = np.array([len(set(r.tolist())) == 27 for r in h])
    if T >= 64:
        assert distinct.all()
    else:
        assert not distinct.any()  # the z coordinate keeps one bit: offsets -1 and +1 alias


@pytest.mark.parametrize("T", [8, 32, 256, 1 << 20])
def test_probed_relation_is_symmetric(T):
    rng = np.random.default_rng(2)
    a = rng.integers(0, 40, (3000, 3))
    b = a + rng.integers(-3, 4, (3000, 3)) * rng.integers(1, 9, (3000, 1))
    ha = hash3(a[:, None, :] + np.array(OFFSETS)[None], T)
    hb = hash3(b[:, None, :] + np.array(OFFSETS)[None], T)
    b_in_a = (ha == hash3(b, T)[:, None]).any(1)
    a_in_b = (hb == hash3(a, T)[:, None]).any(1)
    assert np.array_equal(b_in_a, a_in_b) and b_in_a.any()
    # with 8 entries (every coordinate modulo 2) the 27 cells cover every bucket
    assert b_in_a.all() if T == 8 else not b_in_a.all()


def _touching(x, r, i, j):
    a, b = min(i, j), max(i, j)  # the lower index first, as the kernels evaluate it
    d = x[b] - x[a]
    x2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    s = r[a] + r[b]
    if x2 > s * s * (1.0 + 1e-9):
        return False
    return np.sqrt(x2) - r[a] - r[b] < 0


def _packing(n, seed):
    rng = np.random.default_rng(seed)
    r = 0.5 * np.minimum(np.exp(np.log(0.5) + 0.2 * rng.standard_normal(n)), 0.999)
    x = rng.uniform(0.0, 6.0, (n, 3))
    return x, r, float(2.0 * r.max())


def _buckets(x, cell, T, seed):
    """Cells, hashes, the bucket order (particle at each position), each particle's position
    and each bucket's first position; the order within a bucket is arbitrary on the GPU, so it
    is shuffled here."""
    c = np.floor(x / cell).astype(np.int64)
    h = hash3(c, T)
    perm = np.random.default_rng(seed).permutation(len(x))
    order = perm[np.argsort(h[perm], kind="stable")]
    pos = np.empty(len(x), dtype=np.int64)
    pos[order] = np.arange(len(x))
    start = {}
    for t, p in enumerate(order):
        start.setdefault(int(h[p]), t)
    return c, h, order, pos, start


def _reference_pairs(x, r, c, T):
    """Particle i lists j > i from each distinct probed bucket (the reference's rule)."""
    h = hash3(c, T)
    out = set()
    for i in range(len(x)):
        probed = set(hash3(c[i] + np.array(OFFSETS), T).tolist())
        for j in range(i + 1, len(x)):
            if int(h[j]) in probed and _touching(x, r, i, j):
                out.add((i, j))
    return out


def _walk(x, r, cell, T, seed, half):
    """The pairs the probe lists, in listing order.  half: the 13 forward cells in full and the
    particle's own bucket after its own position; otherwise the 27 cells' distinct buckets, each
    after the particle's own position."""
    c, h, order, pos, start = _buckets(x, cell, T, seed)
    cnt = np.bincount(h, minlength=T)
    listed = []
    for i in range(len(x)):
        t0 = pos[i]
        if half:
            walk = [(int(hash3(c[i] + np.array(d), T)), False) for d in FORWARD]
            walk.append((int(h[i]), True))
        else:
            walk = [(hk, True) for hk in dict.fromkeys(hash3(c[i] + np.array(OFFSETS), T).tolist())]
        for hk, after in walk:
            if hk not in start:
                continue
            b0, e1 = start[hk], start[hk] + int(cnt[hk])
            for t in range(max(b0, t0 + 1) if after else b0, e1):
                j = int(order[t])
                if _touching(x, r, i, j):
                    listed.append((min(i, j), max(i, j)))
    return listed, c


@pytest.mark.parametrize("T,half", [(1 << 20, True), (256, True), (64, True), (32, False), (8, False),
                                    (1 << 20, False)])
@pytest.mark.parametrize("seed", [3, 4])
def test_walk_lists_the_reference_pairs_once(T, half, seed):
    x, r, cell = _packing(700, seed)
    listed, c = _walk(x, r, cell, T, seed, half)
    ref = _reference_pairs(x, r, c, T)
    assert len(listed) == len(set(listed)), "a pair was listed twice"
    assert set(listed) == ref and len(ref) > 100


def test_half_shell_on_an_aliasing_table_lists_pairs_twice():
    """The half shell needs distinct probed hashes: on a table of 8 entries (every coordinate
    modulo 2) it lists some pairs twice, which is why such tables take the other walk."""
    x, r, cell = _packing(700, 5)
    listed, _ = _walk(x, r, cell, 8, 5, True)
    assert len(listed) > len(set(listed))
