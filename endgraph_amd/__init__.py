"""Python host mirror of the MI355X-native EndGraph library (libendgraph.so).

This is plumbing over the C ABI in include/endgraph.h (ctypes); all work runs
in the hand-written gfx950 kernels.  There is no CPU fallback: importing the
native library fails loudly when it has not been built.

The Graph class mirrors the reference's Graph<EdgeData> (core/graph.hpp:100):
construct with the partition/socket counts, call load_directed /
load_undirected_from_directed with a device edge list, read back the built
arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libendgraph.so")
_lib = None

REC8 = np.dtype([("src", "<u4"), ("dst", "<u4")])
REC12 = np.dtype([("src", "<u4"), ("dst", "<u4"), ("w", "<f4")])
CIDX = np.dtype([("index", "<u8"), ("vertex", "<u4")], align=False)

EG_OK = 0


class EndGraphError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"endgraph status {status}: {msg}")
        self.status = status


def build_library(jobs: int = 8) -> str:
    """Compile every HIP source for gfx950 into endgraph_amd/libendgraph.so."""
    subprocess.run(["make", "-s", "-C", _HERE, f"-j{jobs}"], check=True)
    return LIB_PATH


class EngineTune(C.Structure):
    """eg_engine_tune: per-graph sizes and auto-policy thresholds of the traversals."""
    _fields_ = [("bfs_dense_div", C.c_uint32), ("sssp_dense_div", C.c_uint32), ("bfs_multi_dense_div", C.c_uint32),
                ("sssp_multi_dense_div", C.c_uint32), ("hub_degree", C.c_uint32), ("hub_chunk", C.c_uint32),
                ("multi_item_edges", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class _Info(C.Structure):
    _fields_ = [("vertices_num", C.c_uint32), ("partitions", C.c_int), ("sockets", C.c_int),
                ("record_bytes", C.c_int), ("edges_num", C.c_uint64), ("pages_num", C.c_uint32),
                ("symmetrized", C.c_int)]


# True when the library bound to the HIP runtime torch loaded (torch was
# imported first).  A process holds one HIP runtime: loaded first, this
# library binds to ROCm's runtime and torch, imported later, then loads its
# own copy and finds no GPU; imported first, torch's runtime serves both.
lib_shares_torch_runtime = False


def lib():
    """The loaded native library (raises if it is missing)."""
    global _lib, lib_shares_torch_runtime
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"native library missing at {LIB_PATH}: run `make -C endgraph_amd` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    lib_shares_torch_runtime = "torch" in sys.modules
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32, f32, f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float, C.c_double
    pp = C.POINTER(C.c_void_p)
    sigs = {
        "eg_last_error": ([], C.c_char_p),
        "eg_version": ([], i32),
        "eg_device_alloc": ([pp, u64], i32),
        "eg_device_free": ([vp], i32),
        "eg_memcpy_h2d": ([vp, vp, u64, vp], i32),
        "eg_memcpy_d2h": ([vp, vp, u64, vp], i32),
        "eg_memcpy_d2d": ([vp, vp, u64, vp], i32),
        "eg_stream_synchronize": ([vp], i32),
        "eg_gen_rmat": ([vp, u64, u32, f64, f64, f64, u64, i32, f32, f32, vp], i32),
        "eg_gen_powerlaw": ([vp, u64, u32, f64, f64, f64, f64, u64, i32, vp], i32),
        "eg_graph_create": ([u32, i32, i32, u64, u32, i32, pp], i32),
        "eg_graph_destroy": ([vp], i32),
        "eg_graph_build": ([vp, vp, u64, i32, vp], i32),
        "eg_graph_build_file": ([vp, C.c_char_p, vp, u64, i32, C.POINTER(u64), vp], i32),
        "eg_graph_get_info": ([vp, C.POINTER(_Info)], i32),
        "eg_graph_partition_offset": ([vp, vp], i32),
        "eg_graph_local_partition_offset": ([vp, i32, vp], i32),
        "eg_graph_degree": ([vp, pp, pp], i32),
        "eg_graph_page_prefix": ([vp, pp], i32),
        "eg_graph_csc": ([vp, i32, i32, pp, C.POINTER(u64), pp, C.POINTER(u64)], i32),
        "eg_graph_csr": ([vp, i32, i32, pp, C.POINTER(u64), pp, pp], i32),
        "eg_gen_er_batch": ([vp, i32, vp, vp, u64, i32, vp], i32),
        "eg_edge_file_records": ([C.c_char_p, i32, C.POINTER(u64)], i32),
        "eg_load_edge_file": ([C.c_char_p, i32, vp, u64, C.POINTER(u64), vp], i32),
        "eg_dump_vertex_array": ([vp, u64, u32, C.c_char_p, vp], i32),
        "eg_restore_vertex_array": ([vp, u64, u32, C.c_char_p, vp], i32),
        "eg_batch_csr_build": ([vp, i32, i32, vp, vp, pp, vp], i32),
        "eg_batch_csr_rebuild": ([vp, vp, vp], i32),
        "eg_batch_csr_get": ([vp, i32, pp, C.POINTER(u64), pp, C.POINTER(u32)], i32),
        "eg_batch_destroy": ([vp], i32),
        "eg_batch_csr_build_noalloc": ([vp, i32, i32, vp, vp, vp, vp, vp], i32),
        "eg_device_fill": ([vp, u64, vp, i32, vp], i32),
        "eg_device_iota": ([vp, u32, vp], i32),
        "eg_bitmap_or": ([vp, vp, u32, vp], i32),
        "eg_bitmap_count": ([vp, u32, C.POINTER(u64), vp], i32),
        "eg_bitmap_degree_sum": ([vp, vp, C.POINTER(u64), vp], i32),
        "eg_set_sort_variant": ([i32, i32, i32, i32], i32),
        "eg_get_sort_variant": ([i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)], i32),
        "eg_autotune_sort": ([vp, vp, u64, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(f64), vp],
                             i32),
        "eg_profile_enable": ([i32], i32),
        "eg_profile_reset": ([], i32),
        "eg_profile_read": ([C.c_char_p, C.POINTER(f64), C.POINTER(u64), C.POINTER(u64)], i32),
        "eg_engine_prepare": ([vp, C.POINTER(f64), vp], i32),
        "eg_bitmap_clear": ([vp, u32, vp], i32),
        "eg_bitmap_fill": ([vp, u32, vp], i32),
        "eg_bitmap_set_bit": ([vp, u32, vp], i32),
        "eg_pr_init": ([vp, vp, vp, vp], i32),
        "eg_pr_pull": ([vp, i32, vp, vp, vp], i32),
        "eg_pr_apply": ([vp, vp, vp, vp, i32, vp, vp], i32),
        "eg_pagerank": ([vp, i32, i32, vp, vp, vp], i32),
        "eg_pagerank_view_prepare": ([vp, C.POINTER(f64), vp], i32),
        "eg_pagerank_view_prepare_layout": ([vp, i32, C.POINTER(f64), vp], i32),
        "eg_pagerank_view_release": ([vp], i32),
        "eg_traversal_view_prepare": ([vp, C.POINTER(f64), vp], i32),
        "eg_bfs_level": ([vp, i32, u32, vp, vp, vp, vp, vp, u64, C.POINTER(u64), C.POINTER(u64),
                          C.POINTER(i32), vp], i32),
        "eg_bfs": ([vp, u32, i32, vp, vp, C.POINTER(u32), C.POINTER(u64), vp], i32),
        "eg_cc_iteration": ([vp, i32, vp, vp, vp, u64, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32), vp], i32),
        "eg_cc": ([vp, i32, vp, C.POINTER(u32), vp], i32),
        "eg_sssp_iteration": ([vp, i32, vp, vp, vp, u64, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32), vp], i32),
        "eg_sssp": ([vp, u32, i32, vp, C.POINTER(u32), vp], i32),
        "eg_ppr_slots": ([vp, vp, vp, vp, vp], i32),
        "eg_ppr_forward": ([vp, vp, vp, f64, i32, vp, vp], i32),
        "eg_ppr_adjoint": ([vp, vp, vp, vp, f64, i32, vp, vp, vp], i32),
        "eg_bfs_multi": ([vp, vp, i32, i32, vp, vp, vp, vp, C.POINTER(u32), vp], i32),
        "eg_sssp_multi": ([vp, vp, i32, i32, vp, C.POINTER(u32), vp], i32),
        "eg_get_engine_tune": ([vp, C.POINTER(EngineTune)], i32),
        "eg_set_engine_tune": ([vp, C.POINTER(EngineTune)], i32),
        "eg_autotune_engine": ([vp, i32, u64, C.POINTER(EngineTune), C.POINTER(f64), vp], i32),
        "eg_prep_data": ([vp, C.POINTER(u64)], i32),
        "eg_mode_sparse": ([vp, i32, i32, u64, C.POINTER(i32)], i32),
        "eg_dense_messages": ([vp, i32, vp, vp, vp, C.POINTER(u64), vp], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status: int):
    if status != EG_OK:
        raise EndGraphError(status, lib().eg_last_error().decode())


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(int(stream.cuda_stream))  # torch.cuda.Stream


class DeviceBuffer:
    """HBM allocation owned by Python (eg_device_alloc / eg_device_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().eg_device_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value
        # kept on the object: module globals may be gone when it is collected
        # at interpreter exit
        self._free_fn, self._vp = lib().eg_device_free, C.c_void_p

    def free(self):
        if getattr(self, "ptr", None):
            self._free_fn(self._vp(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()

    def upload(self, arr: np.ndarray, stream=None):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        check(lib().eg_memcpy_h2d(C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p), arr.nbytes, _stream(stream)))

    def download(self, dtype, count: int, offset_bytes: int = 0, stream=None) -> np.ndarray:
        return to_host(self.ptr + offset_bytes, dtype, count, stream)

    @classmethod
    def from_array(cls, arr: np.ndarray):
        b = cls(max(arr.nbytes, 1))
        b.upload(arr)
        return b


def to_host(d_ptr: int, dtype, count: int, stream=None) -> np.ndarray:
    """Copy `count` elements of `dtype` from device address d_ptr."""
    dtype = np.dtype(dtype)
    out = np.empty(count, dtype=dtype)
    if count:
        check(lib().eg_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(d_ptr), out.nbytes, _stream(stream)))
    return out


POLICY_REFERENCE, POLICY_AUTO, POLICY_PUSH, POLICY_PULL = 0, 1, 2, 3
PROGRAM_PAGERANK, PROGRAM_BFS, PROGRAM_CC, PROGRAM_SSSP = 0, 1, 2, 3
PR_FP64, PR_FP32C = 0, 1
# layout flags OR-ed into the precision of Graph.pagerank (defaults: FP64
# plain, FP32C degree-ordered)
PR_ORDERED, PR_PLAIN, PR_SORTED = 0x10, 0x20, 0x40


def set_sort_variant(rec_bytes: int, scheme: int, threads: int, digit_bits: int = 8):
    check(lib().eg_set_sort_variant(rec_bytes, scheme, threads, digit_bits))


def get_sort_variant(rec_bytes: int = 8):
    """(scheme, threads, digit_bits) of the radix sort used for this record size."""
    sc, nt, b = C.c_int(), C.c_int(), C.c_int()
    check(lib().eg_get_sort_variant(rec_bytes, C.byref(sc), C.byref(nt), C.byref(b)))
    return sc.value, nt.value, b.value


def profile_enable(on: bool = True):
    check(lib().eg_profile_enable(int(bool(on))))


def profile_reset():
    check(lib().eg_profile_reset())


def profile_read(kernel: str):
    """(total_ms, launches, algorithmic bytes) of a timed kernel class."""
    ms, n, b = C.c_double(), C.c_uint64(), C.c_uint64()
    check(lib().eg_profile_read(kernel.encode(), C.byref(ms), C.byref(n), C.byref(b)))
    return ms.value, n.value, b.value


def load_edge_file(path: str, rec_bytes: int = 8, stream=None):
    """Read a binary edge-list file into a new DeviceBuffer; returns (buffer, records)."""
    n = C.c_uint64()
    check(lib().eg_edge_file_records(str(path).encode(), rec_bytes, C.byref(n)))
    buf = DeviceBuffer(max(n.value * rec_bytes, 1))
    check(lib().eg_load_edge_file(str(path).encode(), rec_bytes, C.c_void_p(buf.ptr), n.value, C.byref(n),
                                  _stream(stream)))
    return buf, n.value


def dump_vertex_array(d_array: int, elem_bytes: int, V: int, path: str, stream=None):
    check(lib().eg_dump_vertex_array(C.c_void_p(d_array), elem_bytes, V, str(path).encode(), _stream(stream)))


def restore_vertex_array(d_array: int, elem_bytes: int, V: int, path: str, stream=None):
    check(lib().eg_restore_vertex_array(C.c_void_p(d_array), elem_bytes, V, str(path).encode(), _stream(stream)))


def memcpy_d2d(d_dst: int, d_src: int, nbytes: int, stream=None):
    check(lib().eg_memcpy_d2d(C.c_void_p(d_dst), C.c_void_p(d_src), nbytes, _stream(stream)))


def synchronize(stream=None):
    check(lib().eg_stream_synchronize(_stream(stream)))


def device_fill(d_ptr: int, n: int, value, dtype, stream=None):
    """fill_vertex_array (core/graph.hpp:289-294): n copies of value."""
    v = np.array([value], dtype=dtype)
    check(lib().eg_device_fill(C.c_void_p(d_ptr), n, v.ctypes.data_as(C.c_void_p), v.itemsize, _stream(stream)))


def device_iota(d_ptr: int, n: int, stream=None):
    """d[i] = i for i < n (cc.cpp:35-41's label)."""
    check(lib().eg_device_iota(C.c_void_p(d_ptr), n, _stream(stream)))


# BitmapType (core/bitmap.hpp:23-68) words in HBM: (V >> 6) + 1 u64
def bitmap_clear(bm: DeviceBuffer, V: int, stream=None):
    check(lib().eg_bitmap_clear(C.c_void_p(bm.ptr), V, _stream(stream)))


def bitmap_fill(bm: DeviceBuffer, V: int, stream=None):
    check(lib().eg_bitmap_fill(C.c_void_p(bm.ptr), V, _stream(stream)))


def bitmap_set_bit(bm: DeviceBuffer, v: int, stream=None):
    check(lib().eg_bitmap_set_bit(C.c_void_p(bm.ptr), v, _stream(stream)))


def bitmap_count(bm: DeviceBuffer, V: int, stream=None) -> int:
    c = C.c_uint64()
    check(lib().eg_bitmap_count(C.c_void_p(bm.ptr), V, C.byref(c), _stream(stream)))
    return c.value


def bitmap_ids(words: np.ndarray, V: int) -> np.ndarray:
    """Set bits of host bitmap words as ascending vertex ids (below V)."""
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:V]
    return np.flatnonzero(bits)


# ------------------------------------------------------------ generators --
def gen_rmat(d_edges: int, E: int, scale: int, a=0.57, b=0.19, c=0.19, seed=1, rec_bytes=8, w_lo=0.0, w_hi=1.0,
             stream=None):
    check(lib().eg_gen_rmat(C.c_void_p(d_edges), E, scale, a, b, c, seed, rec_bytes, w_lo, w_hi, _stream(stream)))


def gen_powerlaw(d_edges: int, E: int, V: int, alpha_out=1.9, alpha_in=2.1, max_out=2_997_469.0, max_in=770_155.0,
                 seed=3, rec_bytes=8, stream=None):
    check(lib().eg_gen_powerlaw(C.c_void_p(d_edges), E, V, alpha_out, alpha_in, max_out, max_in, seed, rec_bytes,
                                _stream(stream)))


def gen_er_batch(d_edges: int, vertices, edges, seed=4, rec_bytes=8, stream=None):
    v = np.ascontiguousarray(vertices, dtype=np.uint32)
    e = np.ascontiguousarray(edges, dtype=np.uint64)
    check(lib().eg_gen_er_batch(C.c_void_p(d_edges), int(v.size), v.ctypes.data_as(C.c_void_p),
                                e.ctypes.data_as(C.c_void_p), seed, rec_bytes, _stream(stream)))


def batch_csr_build_noalloc(d_edges: int, vertices, edges, d_out_list: int, d_out_rowptr: int, rec_bytes: int = 8,
                            stream=None):
    """The allocation-free batched CSR build: results into the caller's
    buffers (graph g's row pointers start at entry sum(vertices[:g]) + g)."""
    v = np.ascontiguousarray(vertices, dtype=np.uint32)
    e = np.ascontiguousarray(edges, dtype=np.uint64)
    check(lib().eg_batch_csr_build_noalloc(C.c_void_p(d_edges), rec_bytes, int(v.size), v.ctypes.data_as(C.c_void_p),
                                           e.ctypes.data_as(C.c_void_p), C.c_void_p(d_out_list),
                                           C.c_void_p(d_out_rowptr), _stream(stream)))


class BatchCSR:
    """Many independent graphs built together (eg_batch_csr_build); graph g's
    CSR equals prepare_push_mode's at P = S = 1."""

    def __init__(self, d_edges: int, vertices, edges, rec_bytes: int = 8, stream=None):
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        e = np.ascontiguousarray(edges, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().eg_batch_csr_build(C.c_void_p(d_edges), rec_bytes, int(v.size), v.ctypes.data_as(C.c_void_p),
                                       e.ctypes.data_as(C.c_void_p), C.byref(h), _stream(stream)))
        self._h, self.n, self.rec = h, int(v.size), rec_bytes
        self.rec_dtype = REC8 if rec_bytes == 8 else REC12
        self._destroy_fn = lib().eg_batch_destroy  # usable at interpreter exit

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._destroy_fn(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def rebuild(self, d_edges: int, stream=None):
        """Build again from new records of the same sizes (no allocation)."""
        check(lib().eg_batch_csr_rebuild(self._h, C.c_void_p(d_edges), _stream(stream)))

    def get(self, g: int):
        lst, rp = C.c_void_p(), C.c_void_p()
        n, nv = C.c_uint64(), C.c_uint32()
        check(lib().eg_batch_csr_get(self._h, g, C.byref(lst), C.byref(n), C.byref(rp), C.byref(nv)))
        return to_host(lst.value, self.rec_dtype, n.value), to_host(rp.value, np.uint64, nv.value + 1)


# ----------------------------------------------------------------- graph --
class Graph:
    """Graph<EdgeData> (core/graph.hpp:100) with P partitions and S sockets on one GPU."""

    def __init__(self, vertices_num: int, partitions: int = 1, sockets: int = 1, eta: int = 0, m: int = 1,
                 record_bytes: int = 8):
        h = C.c_void_p()
        check(lib().eg_graph_create(vertices_num, partitions, sockets, eta, m, record_bytes, C.byref(h)))
        self._h = h
        self._destroy_fn = lib().eg_graph_destroy  # usable at interpreter exit
        self.V, self.P, self.S, self.rec = int(vertices_num), int(partitions), int(sockets), int(record_bytes)
        self.rec_dtype = REC8 if self.rec == 8 else REC12

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._destroy_fn(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    # load_directed / load_undirected_from_directed (core/graph.hpp:1358, 629)
    def build(self, d_edges: int, E0: int, symmetrize: bool = False, stream=None):
        check(lib().eg_graph_build(self._h, C.c_void_p(d_edges), E0, int(bool(symmetrize)), _stream(stream)))

    def build_file(self, path: str, d_edges: int, capacity_records: int, symmetrize: bool = False, stream=None) -> int:
        """load_directed / load_undirected_from_directed from a binary edge file:
        the records land in d_edges rank slice by rank slice while the earlier
        slices' first sorts run; returns the record count."""
        n = C.c_uint64()
        check(lib().eg_graph_build_file(self._h, str(path).encode(), C.c_void_p(d_edges), capacity_records,
                                        int(bool(symmetrize)), C.byref(n), _stream(stream)))
        return n.value

    def engine_tune(self) -> dict:
        t = EngineTune()
        check(lib().eg_get_engine_tune(self._h, C.byref(t)))
        return t.as_dict()

    def set_engine_tune(self, **kw):
        t = EngineTune()
        check(lib().eg_get_engine_tune(self._h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().eg_set_engine_tune(self._h, C.byref(t)))

    def autotune_engine(self, n_roots: int = 16, seed: int = 1, stream=None):
        """Time the traversals' thresholds and sizes on this graph; keep the fastest. Returns (tune, best ms)."""
        t = EngineTune()
        ms = (C.c_double * 5)()
        check(lib().eg_autotune_engine(self._h, n_roots, seed, C.byref(t), ms, _stream(stream)))
        return t.as_dict(), list(ms)

    def autotune_sort(self, d_edges: int, E0: int, symmetrize: bool = False, stream=None):
        """Time every sort variant on this input; keep the fastest (graph left built)."""
        sc, nt, b, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().eg_autotune_sort(self._h, C.c_void_p(d_edges), E0, int(bool(symmetrize)), C.byref(sc),
                                     C.byref(nt), C.byref(b), C.byref(ms), _stream(stream)))
        return sc.value, nt.value, b.value, ms.value

    def load_directed(self, d_edges: int, E0: int, stream=None):
        self.build(d_edges, E0, False, stream)

    def load_undirected_from_directed(self, d_edges: int, E0: int, stream=None):
        self.build(d_edges, E0, True, stream)

    def info(self) -> dict:
        i = _Info()
        check(lib().eg_graph_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _Info._fields_}

    @property
    def edges_num(self) -> int:
        return int(self.info()["edges_num"])

    def partition_offset(self) -> np.ndarray:
        out = np.zeros(self.P + 1, np.uint32)
        check(lib().eg_graph_partition_offset(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def local_partition_offset(self, p: int) -> np.ndarray:
        out = np.zeros(self.S + 1, np.uint32)
        check(lib().eg_graph_local_partition_offset(self._h, p, out.ctypes.data_as(C.c_void_p)))
        return out

    def degree_ptrs(self):
        o, i = C.c_void_p(), C.c_void_p()
        check(lib().eg_graph_degree(self._h, C.byref(o), C.byref(i)))
        return o.value, i.value

    def out_degree(self) -> np.ndarray:
        return to_host(self.degree_ptrs()[0], np.uint32, self.V)

    def in_degree(self) -> np.ndarray:
        return to_host(self.degree_ptrs()[1], np.uint32, self.V)

    def page_prefix(self) -> np.ndarray:
        p = C.c_void_p()
        check(lib().eg_graph_page_prefix(self._h, C.byref(p)))
        pages = (self.V + 4095) // 4096
        return to_host(p.value, np.uint64, pages + 1)

    def csc_ptrs(self, p: int, s: int = 0):
        lst, ci = C.c_void_p(), C.c_void_p()
        n, nci = C.c_uint64(), C.c_uint64()
        check(lib().eg_graph_csc(self._h, p, s, C.byref(lst), C.byref(n), C.byref(ci), C.byref(nci)))
        return lst.value, n.value, ci.value, nci.value

    def csr_ptrs(self, p: int, s: int = 0):
        lst, idx, bm = C.c_void_p(), C.c_void_p(), C.c_void_p()
        n = C.c_uint64()
        check(lib().eg_graph_csr(self._h, p, s, C.byref(lst), C.byref(n), C.byref(idx), C.byref(bm)))
        return lst.value, n.value, idx.value, bm.value

    def csc(self, p: int, s: int = 0):
        """(incoming_adj_list records, compressed_incoming_adj_index) of (p, s), copied to the host."""
        lst, n, ci, nci = self.csc_ptrs(p, s)
        return to_host(lst, self.rec_dtype, n), to_host(ci, CIDX, nci)

    def csr(self, p: int, s: int = 0):
        """(outgoing_adj_list records, outgoing_adj_index u64[V+1], outgoing_adj_bitmap words) of (p, s)."""
        lst, n, idx, bm = self.csr_ptrs(p, s)
        return (to_host(lst, self.rec_dtype, n), to_host(idx, np.uint64, self.V + 1),
                to_host(bm, np.uint64, (self.V >> 6) + 1))

    # ----------------------------------------------------------- engine --
    def prepare_engine(self, stream=None) -> float:
        ms = C.c_double()
        check(lib().eg_engine_prepare(self._h, C.byref(ms), _stream(stream)))
        return ms.value

    def prepare_pagerank_view(self, layout: int = PR_ORDERED, stream=None) -> float:
        """Derive the degree-ordered PageRank view now (PR_ORDERED: the graph's
        default form; PR_SORTED: the sorted-row form); returns its wall ms."""
        ms = C.c_double(0)
        check(lib().eg_pagerank_view_prepare_layout(self._h, layout, C.byref(ms), _stream(stream)))
        return ms.value

    def release_pagerank_view(self):
        """Free the degree-ordered view; the next ordered run derives it again."""
        check(lib().eg_pagerank_view_release(self._h))

    def prepare_traversal_view(self, stream=None) -> float:
        """Derive the traversal view (in-edges by destination, sources ascending) now; returns its wall ms."""
        ms = C.c_double(0)
        check(lib().eg_traversal_view_prepare(self._h, C.byref(ms), _stream(stream)))
        return ms.value

    def pagerank(self, iterations: int, precision: int = PR_FP64, out: DeviceBuffer = None, stream=None):
        """toolkits/pagerank.cpp compute(): final ranks (FP64) and delta(i) per iteration."""
        own = out is None
        if own:
            out = DeviceBuffer(self.V * 8)
        deltas = np.zeros(iterations + 1, np.float64)
        check(lib().eg_pagerank(self._h, iterations, precision, C.c_void_p(out.ptr),
                                deltas.ctypes.data_as(C.c_void_p), _stream(stream)))
        return (out.download(np.float64, self.V) if own else out), deltas

    def bfs(self, root: int, policy: int = POLICY_REFERENCE, parent: DeviceBuffer = None,
            depth: DeviceBuffer = None, download: bool = True, stream=None):
        parent = parent or DeviceBuffer(self.V * 4)
        depth = depth or DeviceBuffer(self.V * 4)
        levels, edges = C.c_uint32(), C.c_uint64()
        check(lib().eg_bfs(self._h, root, policy, C.c_void_p(parent.ptr), C.c_void_p(depth.ptr), C.byref(levels),
                           C.byref(edges), _stream(stream)))
        if not download:
            return levels.value, edges.value
        return parent.download(np.uint32, self.V), depth.download(np.uint32, self.V), levels.value, edges.value

    def cc(self, policy: int = POLICY_REFERENCE, label: DeviceBuffer = None, download: bool = True, stream=None):
        label = label or DeviceBuffer(self.V * 4)
        it = C.c_uint32()
        check(lib().eg_cc(self._h, policy, C.c_void_p(label.ptr), C.byref(it), _stream(stream)))
        return (label.download(np.uint32, self.V) if download else None), it.value

    # ------ one process_edges / process_vertices step at a time: the path
    # the C++ facade's Graph<EdgeData>::process_edges takes -------------------
    def prep_data(self) -> int:
        """Graph::prep_data (core/graph.hpp:129): bytes the reference's ranks would exchange while preprocessing."""
        n = C.c_uint64()
        check(lib().eg_prep_data(self._h, C.byref(n)))
        return n.value

    def mode_sparse(self, policy: int, program: int, active_edges: int) -> bool:
        out = C.c_int()
        check(lib().eg_mode_sparse(self._h, policy, program, active_edges, C.byref(out)))
        return bool(out.value)

    def dense_messages(self, program: int, active: DeviceBuffer = None, visited: DeviceBuffer = None,
                       values: DeviceBuffer = None, stream=None) -> int:
        """Messages a dense step of `program` sends between partitions (exe_data, core/graph.hpp:1736)."""
        n = C.c_uint64()
        ptr = lambda b: C.c_void_p(b.ptr) if b is not None else None
        check(lib().eg_dense_messages(self._h, program, ptr(active), ptr(visited), ptr(values), C.byref(n),
                                      _stream(stream)))
        return n.value

    def bitmap_degree_sum(self, bm: DeviceBuffer, stream=None) -> int:
        """Sum of out-degrees over the set bits (the mode switch input, core/graph.hpp:1484-1490)."""
        s = C.c_uint64()
        check(lib().eg_bitmap_degree_sum(self._h, C.c_void_p(bm.ptr), C.byref(s), _stream(stream)))
        return s.value

    def bfs_level(self, policy, level, frontier, nxt, visited, parent, depth, frontier_edges, stream=None):
        """One BFS level (bfs.cpp:41-86): returns (new vertices, their out-edges, ran sparse)."""
        nv, ne, sp = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib().eg_bfs_level(self._h, policy, level, C.c_void_p(frontier.ptr), C.c_void_p(nxt.ptr),
                                 C.c_void_p(visited.ptr), C.c_void_p(parent.ptr),
                                 C.c_void_p(depth.ptr if depth is not None else None), frontier_edges, C.byref(nv),
                                 C.byref(ne), C.byref(sp), _stream(stream)))
        return nv.value, ne.value, bool(sp.value)

    def cc_iteration(self, policy, active, nxt, label, active_edges, stream=None):
        """One CC step (cc.cpp:45-80): returns (new active vertices, their out-edges, ran sparse)."""
        nv, ne, sp = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib().eg_cc_iteration(self._h, policy, C.c_void_p(active.ptr), C.c_void_p(nxt.ptr),
                                    C.c_void_p(label.ptr), active_edges, C.byref(nv), C.byref(ne), C.byref(sp),
                                    _stream(stream)))
        return nv.value, ne.value, bool(sp.value)

    def sssp_iteration(self, policy, active, nxt, dist, active_edges, stream=None):
        """One SSSP step (sssp.cpp:38-83): returns (new active vertices, their out-edges, ran sparse)."""
        nv, ne, sp = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib().eg_sssp_iteration(self._h, policy, C.c_void_p(active.ptr), C.c_void_p(nxt.ptr),
                                      C.c_void_p(dist.ptr), active_edges, C.byref(nv), C.byref(ne), C.byref(sp),
                                      _stream(stream)))
        return nv.value, ne.value, bool(sp.value)

    def pr_init(self, x: DeviceBuffer, x32: DeviceBuffer = None, stream=None):
        check(lib().eg_pr_init(self._h, C.c_void_p(x.ptr), C.c_void_p(x32.ptr if x32 else None), _stream(stream)))

    def pr_pull(self, precision: int, x: DeviceBuffer, nxt: DeviceBuffer, stream=None):
        check(lib().eg_pr_pull(self._h, precision, C.c_void_p(x.ptr), C.c_void_p(nxt.ptr), _stream(stream)))

    def pr_apply(self, x: DeviceBuffer, nxt: DeviceBuffer, nxt32: DeviceBuffer, last: bool, delta: DeviceBuffer,
                 stream=None):
        check(lib().eg_pr_apply(self._h, C.c_void_p(x.ptr), C.c_void_p(nxt.ptr),
                                C.c_void_p(nxt32.ptr if nxt32 else None), int(bool(last)), C.c_void_p(delta.ptr),
                                _stream(stream)))

    # ------------------------------------------------ personalised PageRank --
    def ppr_slots(self, stream=None):
        """(src, dst, record weight) of every CSC slot, the order of PPR edge weights (device buffers)."""
        E = self.edges_num
        src, dst, w = DeviceBuffer(max(E, 1) * 4), DeviceBuffer(max(E, 1) * 4), DeviceBuffer(max(E, 1) * 4)
        check(lib().eg_ppr_slots(self._h, C.c_void_p(src.ptr), C.c_void_p(dst.ptr), C.c_void_p(w.ptr),
                                 _stream(stream)))
        return src, dst, w

    def ppr_forward(self, d_w: int, d_p: int, damping: float, iterations: int, d_x: int, stream=None):
        check(lib().eg_ppr_forward(self._h, C.c_void_p(d_w), C.c_void_p(d_p), damping, iterations, C.c_void_p(d_x),
                                   _stream(stream)))

    def ppr_adjoint(self, d_w: int, d_x: int, d_c: int, damping: float, iterations: int, d_grad_p: int,
                    d_grad_w: int, stream=None):
        check(lib().eg_ppr_adjoint(self._h, C.c_void_p(d_w), C.c_void_p(d_x), C.c_void_p(d_c), damping, iterations,
                                   C.c_void_p(d_grad_p or None), C.c_void_p(d_grad_w or None), _stream(stream)))

    def bfs_multi(self, roots, policy: int = POLICY_REFERENCE, parent: DeviceBuffer = None,
                  depth: DeviceBuffer = None, download: bool = True, stream=None):
        """BFS from up to 64 roots in one launch; parent/depth are (V, 64) vertex-major, column r = roots[r]."""
        r = np.ascontiguousarray(roots, dtype=np.uint32)
        parent = parent or DeviceBuffer(self.V * 64 * 4)
        depth = depth or DeviceBuffer(self.V * 64 * 4)
        levels = np.zeros(r.size, np.uint32)
        edges = np.zeros(r.size, np.uint64)
        steps = C.c_uint32()
        check(lib().eg_bfs_multi(self._h, r.ctypes.data_as(C.c_void_p), r.size, policy, C.c_void_p(parent.ptr),
                                 C.c_void_p(depth.ptr), levels.ctypes.data_as(C.c_void_p),
                                 edges.ctypes.data_as(C.c_void_p), C.byref(steps), _stream(stream)))
        if not download:
            return levels, edges, steps.value
        return (parent.download(np.uint32, self.V * 64).reshape(self.V, 64),
                depth.download(np.uint32, self.V * 64).reshape(self.V, 64), levels, edges, steps.value)

    def sssp_multi(self, roots, policy: int = POLICY_REFERENCE, dist: DeviceBuffer = None, download: bool = True,
                   stream=None):
        """SSSP from up to 64 roots in one launch; dist is (V, 64) vertex-major, column r = roots[r]."""
        r = np.ascontiguousarray(roots, dtype=np.uint32)
        dist = dist or DeviceBuffer(self.V * 64 * 4)
        steps = C.c_uint32()
        check(lib().eg_sssp_multi(self._h, r.ctypes.data_as(C.c_void_p), r.size, policy, C.c_void_p(dist.ptr),
                                  C.byref(steps), _stream(stream)))
        return (dist.download(np.float32, self.V * 64).reshape(self.V, 64) if download else None), steps.value

    def sssp(self, root: int, policy: int = POLICY_REFERENCE, dist: DeviceBuffer = None, download: bool = True,
             stream=None):
        dist = dist or DeviceBuffer(self.V * 4)
        it = C.c_uint32()
        check(lib().eg_sssp(self._h, root, policy, C.c_void_p(dist.ptr), C.byref(it), _stream(stream)))
        return (dist.download(np.float32, self.V) if download else None), it.value
