"""CPU emulation of the Coulomb kernel's shuffle ranking (the bucketed rank of the grouped
pass, group_perm in picfoam_amd/csrc/k_coulomb.hip) against the oracle's order, a sort by
(random key, cell-list position).

A sort key is the parcel's 53-bit shuffle key with its cell-local index in the 11 bits
below.  The ranking must give a permutation equal to the oracle's for any ids, duplicates
included, up to the 512 parcels of the 64-lane pass and the 2048 of the large-cell pass.
The earlier layout (56-bit key, index in the low byte) is checked to fail that, which is
the fault this layout removes: two parcels with the same id at indices k and k + 256 got
equal sort keys, hence the same rank, and one rank slot stayed unset."""
import numpy as np
import pytest

KEY_BITS = 53
INDEX_BITS = 11


def sort_keys(shuffle_keys, index_bits=INDEX_BITS):
    k = np.arange(len(shuffle_keys), dtype=np.uint64)
    return (shuffle_keys.astype(np.uint64) << np.uint64(index_bits)) | k


def bucket_ranks(keys, L):
    """The grouped pass: 4 L bins by the top log2(4 L) bits, a key's rank = its bin's start
    plus the number of smaller keys in its bin."""
    nb = 4 * L
    sh = np.uint64(64 - int(np.log2(nb)))
    b = (keys >> sh).astype(np.int64)
    counts = np.bincount(b, minlength=nb)
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ranks = np.empty(len(keys), np.int64)
    for i, kk in enumerate(keys):
        same = keys[b == b[i]]
        ranks[i] = start[b[i]] + int(np.sum(same < kk))
    return ranks


def oracle_order(shuffle_keys):
    """The oracle's keyed_perm: sort by (key, position)."""
    return np.lexsort((np.arange(len(shuffle_keys)), shuffle_keys))


def draw_keys(rng, ids, bits=KEY_BITS):
    # one random key per distinct id (the key is a function of the id for a stage)
    uniq = np.unique(ids)
    table = {int(u): int(rng.integers(0, 1 << bits, dtype=np.uint64)) for u in uniq}
    return np.array([table[int(i)] for i in ids], np.uint64)


@pytest.mark.parametrize("n,L", [(130, 16), (256, 32), (300, 64), (512, 64)])
@pytest.mark.parametrize("ids", ["distinct", "pairs", "all-equal", "odd-key-all-equal"])
def test_bucket_rank_equals_oracle_order(n, L, ids):
    rng = np.random.default_rng(n * 7 + L)
    if ids == "distinct":
        idv = rng.permutation(n).astype(np.uint32)
    elif ids == "pairs":
        idv = (np.arange(n) // 2).astype(np.uint32)
    else:
        idv = np.zeros(n, np.uint32)
    sk = draw_keys(rng, idv)
    if ids == "odd-key-all-equal":
        sk[:] = sk[0] | np.uint64(1)
    keys = sort_keys(sk)
    assert len(np.unique(keys)) == n
    ranks = bucket_ranks(keys, L)
    assert sorted(ranks.tolist()) == list(range(n))
    val = np.empty(n, np.int64)
    val[ranks] = np.arange(n)
    assert np.array_equal(val, oracle_order(sk))


def test_large_cell_keys_are_distinct():
    """The large-cell pass sorts up to 2048 keys of a species (bitonic, not stable): the
    index bits keep every sort key distinct, so any correct sort gives the oracle's order.
    A list shorter than its power of two is padded with all-ones keys; a real key stays
    below them because its index is below the padded length."""
    for n in (2048, 2047, 1500):
        sk = np.full(n, (1 << KEY_BITS) - 1, np.uint64)
        keys = sort_keys(sk)
        assert len(np.unique(keys)) == n
        np2 = 1 << int(np.ceil(np.log2(n)))
        if n < np2:
            assert keys.max() < np.uint64((1 << 64) - 1)
        assert np.array_equal(np.argsort(keys, kind="stable"), oracle_order(sk))


def test_old_layout_collides_beyond_256():
    """The earlier layout (56-bit keys, index OR-ed into the low byte) gives equal sort keys
    for two parcels of one id at indices k and k + 256 when the key is odd, so two parcels
    share a rank: the ranking is then no permutation."""
    n = 300
    sk56 = np.full(n, (1 << 55) | 1, np.uint64)
    k = np.arange(n, dtype=np.uint64)
    old = (sk56 << np.uint64(8)) | k
    assert len(np.unique(old)) < n
    ranks = bucket_ranks(old, 64)
    assert sorted(ranks.tolist()) != list(range(n))
