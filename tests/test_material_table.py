"""The material-table builder (fwe_gen_material_table, host code, no GPU)
against numpy: the number of distinct tuples, their order of first
appearance, the per-element index, bitwise comparison (-0.0 and 0.0 and
different NaN payloads are different materials) and the entry limit."""
import ctypes as C

import numpy as np
import pytest

from elanfmt_amd import fwe as F


def _table(cols, max_entries=256):
    cols = [np.ascontiguousarray(c, np.float32) for c in cols]
    n, m = cols[0].size, len(cols)
    ptrs = (F._FP * m)(*[F.fptr(c) for c in cols])
    idx = np.zeros(max(n, 1), np.uint8)
    tab = np.zeros(max_entries * m, np.float32)
    nt = F.lib().fwe_gen_material_table(n, m, ptrs, max_entries, idx.ctypes.data_as(C.POINTER(C.c_uint8)), F.fptr(tab))
    return nt, idx[:n], tab.reshape(max_entries, m)


def _numpy_table(cols):
    t = np.ascontiguousarray(np.stack([np.asarray(c, np.float32) for c in cols], -1))
    v = t.view(np.dtype((np.void, 4 * t.shape[1]))).ravel()
    _, first, inv = np.unique(v, return_index=True, return_inverse=True)
    order = np.argsort(first)              # numbering by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return order.size, rank[inv.ravel()], t[np.sort(first)]


@pytest.mark.parametrize("nmat,n,m", [(1, 1000, 12), (41, 50000, 12), (256, 70000, 5), (7, 3, 1)])
def test_table_matches_numpy(nmat, n, m):
    rng = np.random.default_rng(nmat)
    mats = rng.normal(size=(nmat, m)).astype(np.float32)
    # runs of one material, as in a layered medium, then shuffled blocks
    which = np.repeat(rng.integers(0, nmat, n // 5 + 1), 5)[:n]
    which[: min(n, nmat)] = np.arange(min(n, nmat))  # every material appears
    cols = [mats[which, q] for q in range(m)]
    nt, idx, tab = _table(cols)
    en, eidx, etab = _numpy_table(cols)
    assert nt == en == min(nmat, n)
    assert np.array_equal(idx, eidx)
    assert np.array_equal(tab[:nt].view(np.uint32), etab.view(np.uint32))
    # decoding gives back every array bit for bit
    for q in range(m):
        assert np.array_equal(tab[idx, q].view(np.uint32), cols[q].view(np.uint32))


def test_bitwise_distinct_values():
    nan2 = np.array([0x7FC00001], np.uint32).view(np.float32)[0]
    a = np.array([0.0, -0.0, 0.0, np.nan, nan2, np.nan, 1.0], np.float32)
    nt, idx, tab = _table([a])
    assert nt == 5
    assert list(idx) == [0, 1, 0, 2, 3, 2, 4]


def test_limit_and_arguments():
    a = np.arange(300, dtype=np.float32)
    assert _table([a])[0] == -1                      # 300 distinct > 256
    assert _table([a[:256]])[0] == 256
    assert _table([a[:10]], max_entries=9)[0] == -1
    assert _table([a[:10]], max_entries=10)[0] == 10
    lib = F.lib()
    ptrs = (F._FP * 1)(F.fptr(a))
    idx = np.zeros(300, np.uint8)
    tab = np.zeros(300 * 17, np.float32)
    u8 = idx.ctypes.data_as(C.POINTER(C.c_uint8))
    assert lib.fwe_gen_material_table(10, 0, ptrs, 256, u8, F.fptr(tab)) == -1
    assert lib.fwe_gen_material_table(10, 17, ptrs, 256, u8, F.fptr(tab)) == -1
    assert lib.fwe_gen_material_table(10, 1, ptrs, 257, u8, F.fptr(tab)) == -1
    assert lib.fwe_gen_material_table(0, 1, ptrs, 256, u8, F.fptr(tab)) == 0
