"""ctypes loader for the CPU restatement in oracle/cpu_ref.cpp.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker and the timed
host-CPU baseline; only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import it.  It restates core/graph.hpp (preprocessing and
engine semantics) and toolkits/*.cpp of the reference; see cpu_ref.cpp for
the file:line citations and the repairs it adopts.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libegoracle.so")
_lib = None

CIDX_DTYPE = np.dtype([("index", "<u8"), ("vertex", "<u4")], align=False)  # 12 B packed
REC8_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4")])
REC12_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4"), ("w", "<f4")])


def build_library() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing at {LIB_PATH}; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.orc_optimal_offset.argtypes = [vp, u32, u32, vp]
        L.orc_build.argtypes = [vp, u64, i32, u32, i32, i32, u64, u32, i32, i32, C.POINTER(vp)]
        L.orc_free.argtypes = [vp]
        L.orc_num_edges.argtypes = [vp]
        L.orc_num_edges.restype = u64
        L.orc_build_seconds.argtypes = [vp]
        L.orc_build_seconds.restype = C.c_double
        for name in ("orc_out_degree", "orc_in_degree", "orc_partition_offset"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.POINTER(u32)
        L.orc_local_partition_offset.argtypes = [vp, i32]
        L.orc_local_partition_offset.restype = C.POINTER(u32)
        L.orc_page_amount.argtypes = [vp, C.POINTER(u32)]
        L.orc_page_amount.restype = C.POINTER(u64)
        L.orc_csc.argtypes = [vp, i32, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(u64)]
        L.orc_csc.restype = u64
        L.orc_csr.argtypes = [vp, i32, i32, C.POINTER(vp), C.POINTER(C.POINTER(u64)), C.POINTER(C.POINTER(u64))]
        L.orc_csr.restype = u64
        L.orc_pagerank.argtypes = [vp, i32, vp, vp]
        L.orc_bfs.argtypes = [vp, u32, i32, vp, vp, vp]
        L.orc_cc.argtypes = [vp, vp, vp]
        L.orc_sssp.argtypes = [vp, u32, vp, vp]
        L.orc_csr_batch.argtypes = [vp, i32, i32, vp, vp, vp, vp, i32]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def optimal_offset(a: np.ndarray, k: int) -> np.ndarray:
    """core/graph.hpp:597-626 restated: CPBT boundaries of n weights into k groups."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros(k + 1, dtype=np.uint32)
    _check(lib().orc_optimal_offset(a.ctypes.data, a.size, k, out.ctypes.data), "optimal_offset")
    return out


def csr_batch(edges: np.ndarray, vertices, edge_counts, threads: int = 0):
    """prepare_push_mode at P = S = 1 (core/graph.hpp:1248-1355) of every graph
    of a batch stored back to back: (CSR records at the input offsets, row
    pointers of vertices[i] + 1 entries each, back to back)."""
    edges = np.ascontiguousarray(edges)
    nv = np.ascontiguousarray(vertices, dtype=np.uint32)
    ne = np.ascontiguousarray(edge_counts, dtype=np.uint64)
    out = np.empty_like(edges)
    rowptr = np.empty(int(nv.astype(np.uint64).sum()) + nv.size, dtype=np.uint64)
    _check(lib().orc_csr_batch(edges.ctypes.data, edges.dtype.itemsize, nv.size, nv.ctypes.data, ne.ctypes.data,
                               out.ctypes.data, rowptr.ctypes.data, int(threads)), "csr_batch")
    return out, rowptr


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy().view(dtype) if dtype is not None else None


class OracleGraph:
    """The restated Graph<EdgeData> after load_directed / load_undirected_from_directed.

    edges: numpy structured array (REC8_DTYPE or REC12_DTYPE) in file order.
    """

    def __init__(self, edges: np.ndarray, V: int, P: int = 1, S: int = 1, eta: int = 0, m: int = 1,
                 symmetrize: bool = False, threads: int = 0):
        edges = np.ascontiguousarray(edges)
        if edges.dtype.itemsize not in (8, 12):
            raise ValueError("edge records must be 8 (u32,u32) or 12 (u32,u32,f32) bytes")
        self.rec = edges.dtype.itemsize
        self.V, self.P, self.S = int(V), int(P), int(S)
        h = C.c_void_p()
        _check(lib().orc_build(edges.ctypes.data, edges.shape[0], self.rec, self.V, self.P, self.S,
                               int(eta), int(m), int(bool(symmetrize)), int(threads), C.byref(h)), "build")
        self._h = h
        self._free_fn = lib().orc_free  # usable at interpreter exit, when module globals are gone

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._free_fn(h)
            self._h = None

    @property
    def edges_num(self) -> int:
        return int(lib().orc_num_edges(self._h))

    @property
    def build_seconds(self) -> float:
        return float(lib().orc_build_seconds(self._h))

    def out_degree(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_out_degree(self._h), shape=(self.V,)).copy()

    def in_degree(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_in_degree(self._h), shape=(self.V,)).copy()

    def partition_offset(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_partition_offset(self._h), shape=(self.P + 1,)).copy()

    def local_partition_offset(self, p: int) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_local_partition_offset(self._h, p), shape=(self.S + 1,)).copy()

    def page_amount(self) -> np.ndarray:
        n = C.c_uint32()
        ptr = lib().orc_page_amount(self._h, C.byref(n))
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint64)

    def _rec_dtype(self):
        return REC8_DTYPE if self.rec == 8 else REC12_DTYPE

    def csc(self, p: int, s: int = 0):
        """(incoming_adj_list records, compressed_incoming_adj_index) of sub-partition (p, s)."""
        lst, cidx, ncidx = C.c_void_p(), C.c_void_p(), C.c_uint64()
        n = int(lib().orc_csc(self._h, p, s, C.byref(lst), C.byref(cidx), C.byref(ncidx)))
        recs = np.frombuffer((C.c_char * (n * self.rec)).from_address(lst.value), dtype=self._rec_dtype()).copy() \
            if n else np.zeros(0, self._rec_dtype())
        ci = np.frombuffer((C.c_char * (ncidx.value * 12)).from_address(cidx.value), dtype=CIDX_DTYPE).copy()
        return recs, ci

    def csr(self, p: int, s: int = 0):
        """(outgoing_adj_list records, outgoing_adj_index u64[V+1] with ~0 = undefined, bitmap words)."""
        lst = C.c_void_p()
        idx = C.POINTER(C.c_uint64)()
        bm = C.POINTER(C.c_uint64)()
        n = int(lib().orc_csr(self._h, p, s, C.byref(lst), C.byref(idx), C.byref(bm)))
        recs = np.frombuffer((C.c_char * (n * self.rec)).from_address(lst.value), dtype=self._rec_dtype()).copy() \
            if n else np.zeros(0, self._rec_dtype())
        index = np.ctypeslib.as_array(idx, shape=(self.V + 1,)).copy()
        bitmap = np.ctypeslib.as_array(bm, shape=((self.V >> 6) + 1,)).copy()
        return recs, index, bitmap

    def pagerank(self, iterations: int):
        out = np.zeros(self.V, dtype=np.float64)
        deltas = np.zeros(iterations + 1, dtype=np.float64)
        _check(lib().orc_pagerank(self._h, iterations, out.ctypes.data, deltas.ctypes.data), "pagerank")
        return out, deltas

    def bfs(self, root: int, policy: int = 0):
        parent = np.zeros(self.V, dtype=np.uint32)
        depth = np.zeros(self.V, dtype=np.uint32)
        levels = np.zeros(1, dtype=np.uint32)
        _check(lib().orc_bfs(self._h, root, policy, parent.ctypes.data, depth.ctypes.data, levels.ctypes.data),
               "bfs")
        return parent, depth, int(levels[0])

    def cc(self):
        label = np.zeros(self.V, dtype=np.uint32)
        it = np.zeros(1, dtype=np.uint32)
        _check(lib().orc_cc(self._h, label.ctypes.data, it.ctypes.data), "cc")
        return label, int(it[0])

    def sssp(self, root: int):
        dist = np.zeros(self.V, dtype=np.float32)
        it = np.zeros(1, dtype=np.uint32)
        _check(lib().orc_sssp(self._h, root, dist.ctypes.data, it.ctypes.data), "sssp")
        return dist, int(it[0])
