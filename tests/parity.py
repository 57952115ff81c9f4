"""Bit-exact comparison of a built GPU graph against the oracle's build."""
import numpy as np


def assert_build_equal(gg, og):
    V, P, S = gg.V, gg.P, gg.S
    assert gg.edges_num == og.edges_num
    assert np.array_equal(gg.out_degree(), og.out_degree())
    assert np.array_equal(gg.in_degree(), og.in_degree())
    pa = og.page_amount()
    pre = np.concatenate([[0], np.cumsum(pa, dtype=np.uint64)]).astype(np.uint64)
    assert np.array_equal(gg.page_prefix(), pre)
    assert np.array_equal(gg.partition_offset(), og.partition_offset())
    for p in range(P):
        assert np.array_equal(gg.local_partition_offset(p), og.local_partition_offset(p)), p
    for p in range(P):
        for s in range(S):
            g_list, g_ci = gg.csc(p, s)
            o_list, o_ci = og.csc(p, s)
            assert g_list.shape == o_list.shape, ("csc", p, s)
            assert np.array_equal(g_list.view(np.uint8), o_list.view(np.uint8)), ("csc list", p, s)
            assert np.array_equal(g_ci.view(np.uint8), o_ci.view(np.uint8)), ("csc index", p, s)
            g_list, g_idx, g_bm = gg.csr(p, s)
            o_list, o_idx, o_bm = og.csr(p, s)
            assert np.array_equal(g_list.view(np.uint8), o_list.view(np.uint8)), ("csr list", p, s)
            assert np.array_equal(g_bm, o_bm), ("csr bitmap", p, s)
            defined = o_idx != np.iinfo(np.uint64).max
            assert np.array_equal(g_idx[defined], o_idx[defined]), ("csr index", p, s)
            # every GPU entry is defined: a proper row pointer over the list
            n = g_list.shape[0]
            assert g_idx[0] == 0 or n == 0 or g_idx[0] <= n
            assert np.all(np.diff(g_idx.astype(np.int64)) >= 0) and g_idx[-1] == n
            if n:
                counts = np.diff(g_idx.astype(np.int64))
                assert np.array_equal(np.repeat(np.arange(V), counts), g_list["src"].astype(np.int64))
