"""Device engine: one GPU as one PE of the vertex-block layout.

``Context`` wraps a dmm_ctx (include/dmm_gapbs.h).  Its methods mirror the
reference's hot-path entry points:

  DOBFS(source, alpha=15, beta=18)      src/bfs.cc:142-188
  step_* (TDStep, BUStep, QueueToBitmap, BitmapToQueue, InitParent)
                                        src/bfs.cc:57-140
  PageRankPull(max_iters, epsilon)      src/pr.cc:41-70
  bfs_batch64(sources)                  64 BFS lanes in one pass
"""
import ctypes as C

import numpy as np

from . import _lib as L


def device_count():
    n = C.c_int32()
    L.check(L.load().dmm_device_count(C.byref(n)))
    return n.value


class Context:
    def __init__(self, device=0):
        self._lib = L.load()
        h = C.c_void_p()
        L.check(self._lib.dmm_ctx_create(device, C.byref(h)))
        self._h = h
        self.n = 0
        self.vbegin = 0
        self.vend = 0

    # ---- lifetime -----------------------------------------------------------
    def close(self):
        if self._h:
            self._lib.dmm_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return L.check(rc, self._h)

    @property
    def nloc(self):
        return self.vend - self.vbegin

    # ---- graph --------------------------------------------------------------
    def build_graph(self, pairs, num_nodes=-1, symmetrize=True):
        """Builder::MakeGraphFromEL + SquishGraph (src/builder.h:290-312) on this
        context's GPU; returns the HostGraph (not uploaded)."""
        from .graph import HostGraph
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        h = C.c_void_p()
        self._check(self._lib.dmm_build_graph_device(self._h, len(pairs), L.ptr(pairs, C.c_int32), num_nodes,
                                                     1 if symmetrize else 0, C.byref(h)))
        return HostGraph(h)

    def bc_forward(self, source, read=True):
        """Brandes' forward phase (src/bc.cc:55-96) from source on the uploaded
        whole graph (natural layout): (depth, path_counts, succ words, levels);
        with read=False the arrays stay in HBM (bc_read) and come back None."""
        levels = C.c_int32()
        if not read:
            self._check(self._lib.dmm_bc_forward(self._h, int(source), None, None, None, C.byref(levels)))
            return None, None, None, levels.value
        depth, pc, succ = self.bc_read()
        self._check(self._lib.dmm_bc_forward(self._h, int(source), L.ptr(depth, C.c_int32), L.ptr(pc, C.c_uint64),
                                             L.ptr(succ, C.c_uint64), C.byref(levels)))
        return depth, pc, succ, levels.value

    def bc_read(self, fetch=False):
        """Buffers for a Brandes forward result; with fetch=True filled with
        the last result kept in HBM."""
        depth = np.empty(self.nloc, np.int32)
        pc = np.empty(self.nloc, np.uint64)
        succ = np.empty((self.m_out + 63) // 64, np.uint64)
        if fetch:
            self._check(self._lib.dmm_bc_read(self._h, L.ptr(depth, C.c_int32), L.ptr(pc, C.c_uint64),
                                              L.ptr(succ, C.c_uint64)))
        return depth, pc, succ

    def upload(self, g=None, *, num_nodes=None, out_offsets=None, out_neighs=None,
               in_offsets=None, in_neighs=None, vbegin=0, vend=None, num_edges_directed=-1):
        """Uploads a HostGraph (whole graph) or explicit block arrays."""
        if g is not None:
            num_nodes = g.num_nodes
            out_offsets, out_neighs = g.out_offsets, g.out_neighs
            if g.directed:
                in_offsets, in_neighs = g.in_offsets, g.in_neighs
            num_edges_directed = g.num_edges_directed
        if vend is None:
            vend = num_nodes
        oo = L.i64(out_offsets)
        on = L.i32(out_neighs)
        io = L.i64(in_offsets) if in_offsets is not None else None
        inn = L.i32(in_neighs) if in_neighs is not None else None
        self._check(self._lib.dmm_graph_upload(
            self._h, num_nodes, vbegin, vend, L.ptr(oo, C.c_int64), L.ptr(on, C.c_int32),
            L.ptr(io, C.c_int64), L.ptr(inn, C.c_int32), num_edges_directed))
        self.n, self.vbegin, self.vend = num_nodes, vbegin, vend
        self.words = (num_nodes + 63) // 64
        self.m_out = int(oo[-1])

    def set_layout(self, layout):
        """L.DMM_LAYOUT_DEGREE (default) or L.DMM_LAYOUT_NATURAL; applies to
        the next upload.  The step API needs the natural layout."""
        self._check(self._lib.dmm_ctx_set_layout(self._h, layout))

    def generate_edgelist(self, scale, degree=16, uniform=False):
        """Generator::GenerateEL on this context's GPU: int32[num_pairs, 2]."""
        pairs = np.empty(((1 << scale) * degree, 2), np.int32)
        self._check(self._lib.dmm_generate_edgelist_device(self._h, scale, degree, 1 if uniform else 0,
                                                           L.ptr(pairs, C.c_int32)))
        return pairs

    def make_graph(self, scale, degree=16, uniform=False):
        """Generate and build the symmetrized graph on this context's GPU."""
        from .graph import HostGraph
        h = C.c_void_p()
        self._check(self._lib.dmm_make_graph_device(self._h, scale, degree, 1 if uniform else 0, C.byref(h)))
        return HostGraph(h)

    def synchronize(self):
        """Waits for all work queued on the context's stream."""
        self._check(self._lib.dmm_ctx_synchronize(self._h))

    def device_bytes(self):
        b = C.c_int64()
        self._check(self._lib.dmm_ctx_device_bytes(self._h, C.byref(b)))
        return b.value

    # ---- DOBFS -----------------------------------------------------------------
    def DOBFS(self, source, alpha=15, beta=18, trace=False, want_depth=True, read=True,
              with_stats=True):
        """Returns (parent, depth, stats).  parent/depth are None if read=False
        (the result stays in HBM); stats is None if with_stats=False (the call
        then does nothing beyond the BFS itself)."""
        parent = np.empty(self.nloc, np.int32) if read else None
        depth = np.empty(self.nloc, np.int32) if (read and want_depth) else None
        st = L.BfsStats() if with_stats else None
        self._check(self._lib.dmm_bfs(self._h, int(source), alpha, beta, L.ptr(parent, C.c_int32),
                                      L.ptr(depth, C.c_int32), 1 if trace else 0,
                                      C.byref(st) if st is not None else None))
        return parent, depth, st

    def bfs_read(self, parent=True, depth=True):
        p = np.empty(self.nloc, np.int32) if parent else None
        d = np.empty(self.nloc, np.int32) if depth else None
        self._check(self._lib.dmm_bfs_read(self._h, L.ptr(p, C.c_int32), L.ptr(d, C.c_int32)))
        return p, d

    # ---- single steps (host-driven mirror of the reference loop) --------------
    def step_init(self, source):
        self._check(self._lib.dmm_step_init(self._h, int(source)))

    def step_td(self):
        s, q = C.c_int64(), C.c_int64()
        self._check(self._lib.dmm_step_td(self._h, C.byref(s), C.byref(q)))
        return s.value, q.value

    def step_queue_to_bitmap(self):
        a = C.c_int64()
        self._check(self._lib.dmm_step_queue_to_bitmap(self._h, C.byref(a)))
        return a.value

    def step_bu(self):
        a = C.c_int64()
        self._check(self._lib.dmm_step_bu(self._h, C.byref(a)))
        return a.value

    def step_bitmap_to_queue(self):
        q = C.c_int64()
        self._check(self._lib.dmm_step_bitmap_to_queue(self._h, C.byref(q)))
        return q.value

    def step_finish(self):
        self._check(self._lib.dmm_step_finish(self._h))

    def get_parent(self):
        p = np.empty(self.nloc, np.int32)
        self._check(self._lib.dmm_state_get_parent(self._h, L.ptr(p, C.c_int32)))
        return p

    def set_parent(self, p):
        p = L.i32(p)
        self._check(self._lib.dmm_state_set_parent(self._h, L.ptr(p, C.c_int32)))

    def get_queue(self):
        size = C.c_int64()
        cap = self.nloc + 64
        q = np.empty(cap, np.int32)
        self._check(self._lib.dmm_state_get_queue(self._h, L.ptr(q, C.c_int32), cap, C.byref(size)))
        return q[:size.value].copy()

    def set_queue(self, q):
        q = L.i32(q)
        self._check(self._lib.dmm_state_set_queue(self._h, L.ptr(q, C.c_int32), len(q)))

    def get_bitmap(self, which=0):
        w = np.empty(self.words, np.uint64)
        self._check(self._lib.dmm_state_get_bitmap(self._h, which, L.ptr(w, C.c_uint64)))
        return w

    def set_bitmap(self, which, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        self._check(self._lib.dmm_state_set_bitmap(self._h, which, L.ptr(w, C.c_uint64)))

    def set_level(self, level):
        self._check(self._lib.dmm_state_set_level(self._h, level))

    # ---- batched BFS ---------------------------------------------------------------
    def bfs_batch64(self, sources, want_depth=True, td_div=None, graph_levels=None, depth_in_hbm=False,
                    with_stats=True):
        """Returns (depth[nsrc, nloc] uint8 with 255 = unreached, stats).

        depth_in_hbm with want_depth False leaves the depths in HBM
        (batch_read_depth reads them); with_stats False returns stats None."""
        src = L.i32(sources)
        ns = len(src)
        depth = np.empty((ns, self.nloc), np.uint8) if want_depth else None
        params = None
        if td_div is not None or graph_levels is not None or depth_in_hbm:
            t = self.get_tuning()
            params = L.BatchParams(td_div if td_div is not None else t.batch.td_div,
                                   t.batch.grid_per_cu,
                                   graph_levels if graph_levels is not None else t.batch.graph_levels,
                                   L.DMM_BATCH_DEPTH_HBM if depth_in_hbm else 0)
        st = L.BatchStats() if with_stats else None
        self._check(self._lib.dmm_bfs_batch64(self._h, L.ptr(src, C.c_int32), ns,
                                              L.ptr(depth, C.c_uint8),
                                              C.byref(params) if params is not None else None,
                                              C.byref(st) if st is not None else None))
        return depth, st

    def batch_read_depth(self, nsrc):
        """Depths (uint8[nsrc, nloc]) of the last bfs_batch64 run that wrote them."""
        d = np.empty((nsrc, self.nloc), np.uint8)
        self._check(self._lib.dmm_batch_read_depth(self._h, L.ptr(d, C.c_uint8)))
        return d

    def batch_read_level(self, level):
        """Level words (uint64[nloc], caller's IDs) of the last bfs_batch64 run."""
        w = np.empty(self.nloc, np.uint64)
        self._check(self._lib.dmm_batch_read_level(self._h, int(level), L.ptr(w, C.c_uint64)))
        return w

    def batch_block_level(self, front, seen):
        """One bottom-up batched level over the uploaded PE block.

        front: uint64[num_nodes] all-gathered frontier words; seen:
        uint64[nloc] lanes reached at or before the current level.  Returns
        (next uint64[nloc], BlockLevelStats)."""
        f = np.ascontiguousarray(front, dtype=np.uint64)
        sn = np.ascontiguousarray(seen, dtype=np.uint64)
        if f.shape[0] != self.n or sn.shape[0] != self.nloc:
            raise ValueError("front must have num_nodes words and seen one word per block vertex")
        nxt = np.empty(self.nloc, np.uint64)
        st = L.BlockLevelStats()
        self._check(self._lib.dmm_batch_block_level(self._h, L.ptr(f, C.c_uint64), L.ptr(sn, C.c_uint64),
                                                    L.ptr(nxt, C.c_uint64), C.byref(st)))
        return nxt, st

    # ---- PageRank ----------------------------------------------------------------------
    def PageRankPull(self, max_iters=20, epsilon=0.0, mode=L.DMM_PR_FP32_SERIAL, read=True):
        """Returns (scores, stats); scores float32 (FP32 modes) or float64."""
        dt = np.float64 if mode == L.DMM_PR_FP64 else np.float32
        scores = np.empty(self.nloc, dt) if read else None
        st = L.PrStats()
        self._check(self._lib.dmm_pr_pull(self._h, max_iters, float(epsilon), mode,
                                          scores.ctypes.data_as(C.c_void_p) if read else None,
                                          C.byref(st)))
        return scores, st

    def pr_block_iteration(self, global_contrib, scores, mode=L.DMM_PR_FP32_SERIAL):
        gc = np.ascontiguousarray(global_contrib, dtype=np.float32)
        sc = np.ascontiguousarray(scores, dtype=np.float32).copy()
        err, ms = C.c_double(), C.c_double()
        self._check(self._lib.dmm_pr_block_iteration(self._h, L.ptr(gc, C.c_float), L.ptr(sc, C.c_float),
                                                     mode, C.byref(err), C.byref(ms)))
        return sc, err.value, ms.value

    # ---- tuning --------------------------------------------------------------------------
    def get_tuning(self):
        t = L.Tuning()
        self._check(self._lib.dmm_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, t):
        self._check(self._lib.dmm_set_tuning(self._h, C.byref(t)))

    def autotune(self, sources, budget=1):
        src = L.i32(sources)
        t = L.Tuning()
        self._check(self._lib.dmm_autotune(self._h, L.ptr(src, C.c_int32), len(src), budget, C.byref(t)))
        return t


def stats_dict(st):
    """Plain-dict view of a BfsStats (trace arrays cut to num_steps)."""
    n = st.num_steps
    names = {L.DMM_STEP_TD: "td", L.DMM_STEP_BU: "bu", L.DMM_STEP_QB: "e", L.DMM_STEP_BQ: "c"}
    return {
        "steps": [(names.get(st.step_kind[i], "?"), st.step_in[i], st.step_out[i], st.step_edges[i],
                   st.step_scout[i], st.step_ms[i], st.step_kernel_ms[i]) for i in range(n)],
        "alg_bytes": st.alg_bytes, "bu_alg_bytes": st.bu_alg_bytes, "reached": st.reached,
        "reached_edges": st.reached_edges, "total_ms": st.total_ms, "bu_ms": st.bu_ms,
        "host_ms": st.host_ms,
    }
