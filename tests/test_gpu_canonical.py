"""GPU: the device canonical pair order (curtree_pairs_canonical_sort), the
(query id, box id) order the reference's two host stable sorts give
(src/rtree_test_frontier.cu:102-105), against numpy lexsort.

Pairs in query order (every query entry point's output) are sorted segment
by segment: the tests cover segments of every length class (1, 2..64,
65..1024, 1025..4096 and longer), signed and extreme box ids, digits every id
of a segment shares, pairs whose query ids are not in order (the full sort),
and a result of more than 2^30 pairs (checked by properties)."""
import ctypes

import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu


def _dev(ct, a):
    return ct._dev_from_host(np.ascontiguousarray(a, dtype=np.int32))


def _sort(ct, fid, tid):
    fp, tp = _dev(ct, fid), _dev(ct, tid)
    try:
        ct._check(ct.lib.curtree_pairs_canonical_sort(ctypes.c_void_p(fp), ctypes.c_void_p(tp),
                                                      int(fid.shape[0]), None))
        return (ct._host_from_dev(fp, fid.shape[0], np.int32),
                ct._host_from_dev(tp, fid.shape[0], np.int32))
    finally:
        ct.lib.curtree_free(ctypes.c_void_p(fp))
        ct.lib.curtree_free(ctypes.c_void_p(tp))


def _segments(lengths, rng, id_lo=-2**31, id_hi=2**31 - 1):
    """Pairs in query order: query q holds lengths[q] pairs with random ids."""
    tid = np.repeat(np.arange(len(lengths), dtype=np.int32), lengths)
    fid = rng.integers(id_lo, id_hi, tid.shape[0], dtype=np.int64, endpoint=True).astype(np.int32)
    return fid, tid


@pytest.mark.parametrize("lengths", [
    [1, 2, 3, 5, 17, 31, 32, 33, 63, 64],                  # the bitonic wave class
    [65, 100, 511, 777, 1000, 1023, 1024],                 # the LSD wave class
    [1025, 2000, 3333, 4095, 4096],                        # one block each
    [4097, 9000, 65_536, 300_001],                         # the long class
    [1, 64, 65, 1024, 1025, 4096, 4097, 1, 1, 50_000, 7],  # mixed
])
def test_segment_classes_match_lexsort(lengths):
    import curtree_amd as ct

    rng = np.random.default_rng(sum(lengths))
    fid, tid = _segments(lengths, rng)
    f, t = _sort(ct, fid, tid)
    cf, ctt = orc.canonical(fid, tid)
    assert np.array_equal(t, ctt) and np.array_equal(f, cf)


def test_shared_digits_extremes_and_duplicates():
    # ids sharing their top bytes (passes skipped), duplicate ids, INT_MIN/MAX
    import curtree_amd as ct

    rng = np.random.default_rng(7)
    lengths = [5000, 70_000, 900, 40]
    fid, tid = _segments(lengths, rng, 0, 255)  # only the lowest digit varies
    fid[:3] = [2**31 - 1, -2**31, 0]
    fid[5000:5010] = 2**31 - 1
    fid[5010:5020] = -2**31
    f, t = _sort(ct, fid, tid)
    cf, ctt = orc.canonical(fid, tid)
    assert np.array_equal(t, ctt) and np.array_equal(f, cf)


def test_query_ids_out_of_order_take_the_full_sort():
    # grouped by query but not in query order (as the self-join's box ids)
    import curtree_amd as ct

    rng = np.random.default_rng(8)
    lengths = rng.integers(1, 3000, 400)
    fid, tid = _segments(lengths, rng)
    perm = rng.permutation(400).astype(np.int32)
    tid = perm[tid]
    f, t = _sort(ct, fid, tid)
    cf, ctt = orc.canonical(fid, tid)
    assert np.array_equal(t, ctt) and np.array_equal(f, cf)


def test_more_than_2_30_pairs_properties():
    # 2^30 + 4099 pairs in query order, segment lengths cycling through every
    # class; each segment's ids are a descending run (so every segment is
    # reversed by the sort): after it, the query ids are untouched and every
    # segment holds its ascending run (first id, length: a closed form)
    import curtree_amd as ct

    n = (1 << 30) + 4099
    pattern = np.array([1, 3, 64, 700, 1024, 3000, 4096, 20_000], np.int64)
    reps = n // int(pattern.sum()) + 1
    lengths = np.tile(pattern, reps)
    ends = np.cumsum(lengths)
    k = int(np.searchsorted(ends, n))
    lengths = lengths[: k + 1].copy()
    lengths[-1] -= int(ends[k] - n)
    lengths = lengths[lengths > 0]
    assert int(lengths.sum()) == n
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    print(f"[case] {n} pairs in {lengths.shape[0]} segments", flush=True)
    tid = np.repeat(np.arange(lengths.shape[0], dtype=np.int32), lengths)
    # segment s holds ids base_s + len_s - 1 - j, j = 0..len_s-1 (descending)
    base = (np.arange(lengths.shape[0], dtype=np.int64) * 7919) % 1_000_003 - 500_000
    pos = np.arange(n, dtype=np.int64) - np.repeat(starts, lengths)
    fid = (np.repeat(base + lengths - 1, lengths) - pos).astype(np.int32)
    del pos
    f, t = _sort(ct, fid, tid)
    del fid
    assert np.array_equal(t, tid)
    pos = np.arange(n, dtype=np.int64) - np.repeat(starts, lengths)
    assert np.array_equal(f, (np.repeat(base, lengths) + pos).astype(np.int32))


def test_unaligned_arrays_and_chunk_edges():
    # pairs that start one element into their buffers (no wide loads), with
    # segment lengths around the head scan's 4096-position chunks (heads on
    # a chunk's first and last position, segments spanning many chunks, a run
    # of one-pair segments)
    import curtree_amd as ct

    rng = np.random.default_rng(9)
    lengths = ([511, 1, 513, 512, 2, 1023, 1025, 4095, 4097, 16383, 16385, 3, 70_000]
               + [138, 4096, 1, 4094, 1, 8192, 12_287] + [1] * 5000 + [2, 40_000])
    fid, tid = _segments(lengths, rng)
    n = fid.shape[0]
    fp = _dev(ct, np.concatenate([[0], fid]))
    tp = _dev(ct, np.concatenate([[0], tid]))
    try:
        ct._check(ct.lib.curtree_pairs_canonical_sort(ctypes.c_void_p(fp + 4),
                                                      ctypes.c_void_p(tp + 4), n, None))
        f = ct._host_from_dev(fp + 4, n, np.int32)
        t = ct._host_from_dev(tp + 4, n, np.int32)
    finally:
        ct.lib.curtree_free(ctypes.c_void_p(fp))
        ct.lib.curtree_free(ctypes.c_void_p(tp))
    cf, ctt = orc.canonical(fid, tid)
    assert np.array_equal(t, ctt) and np.array_equal(f, cf)


def test_long_segments_with_crowded_ids():
    # long segments are bucketed by the top bits of their key range; ids
    # crowded into a small part of the range leave buckets longer than a
    # block sorts (they take the LSD passes from the work array), all-equal
    # ids one bucket, and a few far ids many tiny buckets
    import curtree_amd as ct

    rng = np.random.default_rng(11)
    lengths = [300_000, 70_000, 40_000, 17_000, 5]
    fid, tid = _segments(lengths, rng, 0, 999)
    s1 = lengths[0]
    fid[:5] = [2**31 - 1, -2**31, 123_456_789, -7, 2**30]  # wide range, most ids crowded
    fid[s1:s1 + lengths[1]] = 42                            # one id repeated
    s3 = s1 + lengths[1] + lengths[2]
    fid[s3:s3 + lengths[3]] = rng.permutation(lengths[3]).astype(np.int32) * 4099  # spread evenly
    f, t = _sort(ct, fid, tid)
    cf, ctt = orc.canonical(fid, tid)
    assert np.array_equal(t, ctt) and np.array_equal(f, cf)
