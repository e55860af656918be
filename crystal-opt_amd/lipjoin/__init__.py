"""lipjoin -- host API of the MI355X star-join engine (ctypes over liblipjoin.so).

Mirrors the reference's host drivers for this path:
  * Context.hash_join  <- TimeKeeper hashJoin(...)  src/ops/join.cu:117-148,
                                                      src/ops/bloom_join.cu:121-156
  * Context.star_join  <- float runQuery(...)       src/ssb-jo/q33/0-lip.cu:205-302
  * Context.build_dim  <- build_kernel / build_hashtable_{s,c,d}
Device columns are passed as torch tensors (int32, on the context's device) or
as raw device addresses.  There is no CPU fallback: every call runs the HIP
kernels of liblipjoin.so, and a missing library raises LibraryMissing.
"""
import ctypes as C
import weakref

from . import _abi
from ._abi import (  # noqa: F401  (re-exported constants)
    BLOOM_BLOCKED, BLOOM_MOD, BLOOM_MULT, BLOOM_NONE, BLOOM_STANDARD, CMP_EQ, CMP_GT, CMP_GTE,
    CMP_LT, CMP_LTE, COMB_AND, COMB_OR, COMB_SET, GEN_DATEKEY, GEN_DATEKEY_SEQ, GEN_SEQUENCE,
    GEN_UNIFORM, GEN_YEAR_OF_DAY, HT_PHT1, HT_PHT2, HT_TEXTBOOK, JOIN_SUM_ONLY, METHOD_LIP,
    METHOD_LIP_ADAPTIVE, METHOD_NAIVE, MAX_DIMS, TUNE_SEMI_JOIN, LibraryMissing,
)

METHODS = {"naive": METHOD_NAIVE, "lip": METHOD_LIP, "lip-adaptive": METHOD_LIP_ADAPTIVE}


class LJError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_abi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    return _abi.load()


def _ptr(x):
    """Device address of a torch tensor, or an int address, or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def round_up_pow2(n):
    """roundUpToPowerOfTwo of src/ssb/ssb_utils.h:85-96 (0 -> 1)."""
    n = int(n)
    if n == 0:
        return 1
    n -= 1
    for s in (1, 2, 4, 8, 16):
        n |= n >> s
    return (n + 1) & 0xFFFFFFFF


def pred(*steps):
    """Predicate program: steps of (comb, cmp, value); () selects every row."""
    p = _abi.Pred()
    if len(steps) > _abi.MAX_PRED_STEPS:
        raise ValueError("at most 4 predicate steps")
    p.n_steps = len(steps)
    for i, (comb, cmp, val) in enumerate(steps):
        p.steps[i].comb, p.steps[i].cmp, p.steps[i].value = comb, cmp, val
    return p


def dim_desc(keys, n_rows, ht_len, attr=None, payload=None, predicate=(), keys_min=0,
             bloom=None, ht_kind=HT_PHT2):
    """Dimension descriptor.  bloom: None, ("mod", bits) or ("mult", bits, k, layout);
    ht_kind: HT_PHT2 (the reference's default), HT_PHT1 (keys only) or HT_TEXTBOOK
    (open addressing with key compare; ht_len a power of two)."""
    d = _abi.DimDesc()
    d.keys, d.attr, d.payload = _ptr(keys), _ptr(attr), _ptr(payload)
    d.n_rows = int(n_rows)
    d.pred = predicate if isinstance(predicate, _abi.Pred) else pred(*predicate)
    d.ht_kind = int(ht_kind)
    d.keys_min = int(keys_min)
    d.ht_len = int(ht_len)
    d.bloom_hash, d.bloom_layout, d.bloom_k, d.bloom_bits = BLOOM_NONE, 0, 0, 0
    if bloom is not None:
        kind = bloom[0]
        if kind == "mod":
            d.bloom_hash, d.bloom_k, d.bloom_bits = BLOOM_MOD, 1, int(bloom[1])
        elif kind == "mult":
            d.bloom_hash, d.bloom_bits, d.bloom_k = BLOOM_MULT, int(bloom[1]), int(bloom[2])
            d.bloom_layout = bloom[3] if len(bloom) > 3 else BLOOM_STANDARD
        else:
            raise ValueError(f"unknown Bloom kind {kind!r}")
    return d


def stats_dict(s):
    return {
        "rows_in": list(s.rows_in), "lines_in": list(s.lines_in),
        "rows_bloom_pass": s.rows_bloom_pass, "sample_rows": s.sample_rows,
        "sample_lines": s.sample_lines, "order_used": list(s.order_used), "n_oob": s.n_oob,
    }


class Dim:
    """A device-resident dimension (PHT_2 + optional Bloom filter)."""

    def __init__(self, ctx, handle):
        self.ctx, self.handle = ctx, handle

    def info(self):
        i = _abi.DimInfo()
        self.ctx._check(lib().lj_dim_get_info(self.ctx.h, self.handle, C.byref(i)))
        return {k: getattr(i, k) for k, _ in _abi.DimInfo._fields_}

    def export(self):
        """(PHT words int32[2*ht_len], Bloom words uint32[...]) as numpy arrays."""
        import numpy as np
        info = self.info()
        ht = np.zeros(info["ht_words"], dtype=np.int32)
        bf = np.zeros(max(info["bloom_words"], 0), dtype=np.uint32)
        self.ctx._check(lib().lj_dim_export(self.ctx.h, self.handle, ht.ctypes.data, ht.size,
                                            bf.ctypes.data if bf.size else None, bf.size))
        return ht, bf

    def release(self):
        if self.handle:
            lib().lj_dim_release(self.ctx.h, self.handle)
            self.handle = None


class Context:
    """One lj_ctx: a device, a stream, a caching allocator and an HBM budget."""

    def __init__(self, device=0, hbm_budget=0, stream=None):
        L = lib()
        h = C.c_void_p()
        st = _abi.OK
        if stream is None:
            try:
                import torch
                if torch.cuda.is_available():
                    stream = torch.cuda.current_stream(device).cuda_stream
            except ImportError:
                stream = None
        self._batches = weakref.WeakSet()
        st = L.lj_create(int(device), int(hbm_budget), stream, C.byref(h))
        if st != _abi.OK:
            raise LJError(st, L.lj_last_error(None).decode())
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            for b in list(getattr(self, "_batches", ())):
                b.close()   # a batch holds context memory: it goes first
            lib().lj_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, allow=()):
        if st != _abi.OK and st not in allow:
            raise LJError(st, lib().lj_last_error(self.h).decode())
        return st

    def sync(self):
        self._check(lib().lj_sync(self.h))

    def bytes_in_use(self):
        return lib().lj_bytes_in_use(self.h)

    # -- dimensions -------------------------------------------------------------
    def build_dim(self, desc):
        h = C.c_void_p()
        self._check(lib().lj_dim_build(self.h, C.byref(desc), C.byref(h)))
        return Dim(self, h)

    # -- single join (hashJoin) -------------------------------------------------------
    def hash_join(self, fact_fk, fact_measure, n_fact, dim, method="lip", allow_domain=False, sum_only=False):
        """lj_hash_join; sum_only: SUM(measure) and the counts only
        (LJ_JOIN_SUM_ONLY: sum_product is 0)."""
        j = _abi.JoinDesc()
        j.fact_fk, j.fact_measure, j.n_fact = _ptr(fact_fk), _ptr(fact_measure), int(n_fact)
        j.dim = dim
        j.method = METHODS[method] if isinstance(method, str) else method
        j.flags = JOIN_SUM_ONLY if sum_only else 0
        r = _abi.JoinResult()
        st = self._check(lib().lj_hash_join(self.h, C.byref(j), C.byref(r)),
                         allow=(_abi.EDOMAIN,) if allow_domain else ())
        out = {k: getattr(r, k) for k, _ in _abi.JoinResult._fields_ if k != "stats"}
        out["stats"] = stats_dict(r.stats)
        out["status"] = st
        return out

    # -- star join (runQuery) ---------------------------------------------------------
    def star_join_raw(self, desc, groups_cap=4096):
        """lj_star_join into a reused result structure, without building the
        Python result (the C-ABI call alone); star_join_result() reads it."""
        if getattr(self, "_gbuf_cap", 0) < groups_cap:
            self._gbuf = (_abi.Group * groups_cap)()
            self._gbuf_cap = groups_cap
        if getattr(self, "_raw_res", None) is None:
            self._raw_res = _abi.StarResult()
        r = self._raw_res
        r.groups, r.groups_cap = C.cast(self._gbuf, C.POINTER(_abi.Group)), groups_cap
        self._raw_nd = desc.n_dims
        return self._check(lib().lj_star_join(self.h, C.byref(desc), C.byref(r)))

    def star_join_result(self):
        r = self._raw_res
        return _star_result_dict(r, self._gbuf, self._raw_nd, r.groups_cap)

    def star_join_noalloc(self, desc, groups_cap=4096, allow_domain=False):
        """lj_star_join_noalloc: the same result as star_join, over resident
        dimensions, with the library allocating nothing."""
        return self.star_join(desc, groups_cap, allow_domain, _entry="lj_star_join_noalloc")

    def star_join(self, desc, groups_cap=4096, allow_domain=False, _entry="lj_star_join"):
        # the host result buffer is kept and reused across calls
        if getattr(self, "_gbuf_cap", 0) < groups_cap:
            self._gbuf = (_abi.Group * groups_cap)()
            self._gbuf_cap = groups_cap
        arr = self._gbuf
        r = _abi.StarResult()
        r.groups, r.groups_cap = C.cast(arr, C.POINTER(_abi.Group)), groups_cap
        st = self._check(getattr(lib(), _entry)(self.h, C.byref(desc), C.byref(r)),
                         allow=(_abi.EDOMAIN,) if allow_domain else ())
        n = min(r.n_groups, groups_cap)
        groups = [(arr[i].g, tuple(arr[i].payload[: desc.n_dims]), arr[i].sum, arr[i].count,
                   arr[i].sum_i32) for i in range(n)]
        return {
            "groups": groups, "n_groups": r.n_groups, "n_rows": r.n_rows, "sum_total": r.sum_total,
            "time_query_ms": r.time_query_ms, "time_build_ms": r.time_build_ms,
            "time_probe_ms": r.time_probe_ms, "stats": stats_dict(r.stats), "status": st,
        }

    # -- batches (many queries, one build launch + one probe launch per kernel kind) --------
    def batch(self, descs, use_graph=False, groups_cap=64, probe_wg=0, launch_per_kind=False):
        return Batch(self, descs, use_graph, groups_cap, probe_wg, launch_per_kind)

    # -- serving (request loop over resident dimensions) ----------------------------------
    def serve(self, hbm_budget=0, max_batch=64, max_inflight=4, groups_cap=64, streams=1, prewarm_groups=0):
        """lj_serve_start; prewarm_groups > 0 allocates the request buffers for
        queries of up to that many groups at start."""
        return Server(self, hbm_budget, max_batch, max_inflight, groups_cap, streams, prewarm_groups)

    # -- Bloom autotuner ------------------------------------------------------------------
    def tune_bloom(self, n_selected, n_dim_rows, fact_pass, semi_join=False):
        """lj_tune_bloom; semi_join: the join needs no payload (one group, no
        product), so an exact modulo filter costs no hash-table gathers."""
        ch = _abi.BloomChoice()
        flags = TUNE_SEMI_JOIN if semi_join else 0
        self._check(lib().lj_tune_bloom(self.h, int(n_selected), int(n_dim_rows), float(fact_pass), flags,
                                        C.byref(ch)))
        return {k: getattr(ch, k) for k, _ in _abi.BloomChoice._fields_}

    def tune_calibrate(self):
        self._check(lib().lj_tune_calibrate(self.h))

    def tune_model(self):
        """lj_tune_get_model: the tuner's cost constants (a dict)."""
        m = _abi.TuneModel()
        self._check(lib().lj_tune_get_model(self.h, C.byref(m)))
        return {k: getattr(m, k) for k, _ in _abi.TuneModel._fields_ if k != "reserved"}

    def set_tune_model(self, model):
        """lj_tune_set_model from a dict like tune_model()'s."""
        m = _abi.TuneModel(**{k: model[k] for k in ("scan_l2", "scan_big", "probe_extra", "line", "gather",
                                                     "scan_mod", "scan_blocked", "gather_queue", "gather_lds",
                                                     "gather_queue_lds")})
        self._check(lib().lj_tune_set_model(self.h, C.byref(m)))

    # -- synthetic columns -------------------------------------------------------------
    def gen_device(self, out, kind, seed, stream, lo, hi, first=0, n=None):
        n = out.numel() if n is None else n
        self._check(lib().lj_gen_device(self.h, kind, seed, stream, lo, hi, first, n, _ptr(out)))


def _star_result_dict(r, arr, n_dims, cap):
    n = min(r.n_groups, cap)
    groups = [(arr[i].g, tuple(arr[i].payload[:n_dims]), arr[i].sum, arr[i].count, arr[i].sum_i32)
              for i in range(n)]
    return {
        "groups": groups, "n_groups": r.n_groups, "n_rows": r.n_rows, "sum_total": r.sum_total,
        "time_query_ms": r.time_query_ms, "time_build_ms": r.time_build_ms,
        "time_probe_ms": r.time_probe_ms, "stats": stats_dict(r.stats), "status": r.status,
    }


class Batch:
    """lj_batch: the queries run as one build launch and one probe launch
    (launch_per_kind: one probe launch per stage-0 kernel kind); with use_graph
    the launches are recorded once as a HIP graph and replayed."""

    def __init__(self, ctx, descs, use_graph=False, groups_cap=64, probe_wg=0, launch_per_kind=False):
        self.ctx = ctx
        self.n = len(descs)
        self._descs = (_abi.StarDesc * self.n)(*descs)
        self._n_dims = [d.n_dims for d in descs]
        o = _abi.BatchOpts()
        o.use_graph, o.probe_wg, o.groups_cap = int(bool(use_graph)), int(probe_wg), int(groups_cap)
        o.launch_per_kind = int(bool(launch_per_kind))
        self.cap = int(groups_cap)
        h = C.c_void_p()
        ctx._check(lib().lj_batch_create(ctx.h, self._descs, self.n, C.byref(o), C.byref(h)))
        self.h = h
        ctx._batches.add(self)
        self._res = (_abi.StarResult * self.n)()
        self._arrs = [(_abi.Group * self.cap)() for _ in range(self.n)]
        for i in range(self.n):
            self._res[i].groups = C.cast(self._arrs[i], C.POINTER(_abi.Group))
            self._res[i].groups_cap = self.cap

    def run(self, allow_domain=False, allow_errors=False):
        """Results of every query (each dict carries its own "status").  A
        failed query raises unless allow_errors (allow_domain: EDOMAIN only)."""
        st = lib().lj_batch_run(self.h, self._res)
        allow = (_abi.EDOMAIN,) if allow_domain else ()
        if not allow_errors or st in (_abi.EINVAL, _abi.EHIP):
            self.ctx._check(st, allow=allow)
        return [_star_result_dict(self._res[i], self._arrs[i], self._n_dims[i], self.cap) for i in range(self.n)]

    def close(self):
        if self.h:
            lib().lj_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """lj_server: submit star joins over resident dimensions from any thread,
    poll for results.  The context must not be used otherwise until stop()."""

    def __init__(self, ctx, hbm_budget=0, max_batch=64, max_inflight=4, groups_cap=64, streams=1,
                 prewarm_groups=0):
        self.ctx = ctx
        cfg = _abi.ServeCfg()
        cfg.hbm_budget, cfg.max_batch, cfg.max_inflight, cfg.groups_cap = (int(hbm_budget), int(max_batch),
                                                                          int(max_inflight), int(groups_cap))
        cfg.streams = int(streams)
        cfg.prewarm_groups = int(prewarm_groups)
        self.cap = int(groups_cap)
        h = C.c_void_p()
        ctx._check(lib().lj_serve_start(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._n_dims = {}
        ctx._batches.add(self)

    def submit(self, desc):
        i = C.c_uint64()
        st = lib().lj_serve_submit(self.h, C.byref(desc), C.byref(i))
        if st != _abi.OK:
            raise LJError(st, "lj_serve_submit: " + lib().lj_last_error(None).decode())
        self._n_dims[i.value] = desc.n_dims
        return i.value

    def poll(self, rid, timeout_ms=-1):
        """The result dict, or None on timeout (the request stays queued)."""
        arr = (_abi.Group * self.cap)()
        r = _abi.StarResult()
        r.groups, r.groups_cap = C.cast(arr, C.POINTER(_abi.Group)), self.cap
        st = lib().lj_serve_poll(self.h, int(rid), C.byref(r), int(timeout_ms))
        if st == _abi.ETIMEOUT:
            return None
        nd = self._n_dims.pop(rid, 1)
        out = _star_result_dict(r, arr, nd, self.cap)
        out["status"] = st
        return out

    def stats(self):
        s = _abi.ServeStats()
        lib().lj_serve_get_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in _abi.ServeStats._fields_ if k != "reserved"}

    def close(self):
        if self.h:
            lib().lj_serve_stop(self.h)
            self.h = None

    stop = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_host(kind, seed, stream, lo, hi, n, first=0):
    """numpy int32 column from the counter-based generator (same bits as gen_device)."""
    import numpy as np
    out = np.empty(int(n), dtype=np.int32)
    st = lib().lj_gen_host(kind, seed, stream, lo, hi, first, int(n), out.ctypes.data if n else None)
    if st != _abi.OK:
        raise LJError(st, "lj_gen_host")
    return out


def gen_reference_host(n_dim, n_fact, want=("dim_keys", "dim_vals", "fact_keys", "fact_vals")):
    """The reference's single-join inputs (generator.h:37-136 after srand(0)):
    a dict of the numpy int32 columns named in `want`."""
    import numpy as np
    names = ("dim_keys", "dim_vals", "fact_keys", "fact_vals")
    sizes = (n_dim, n_dim, n_fact, n_fact)
    out = {k: np.empty(int(n), dtype=np.int32) for k, n in zip(names, sizes) if k in want}
    ptr = [out[k].ctypes.data if k in out and out[k].size else None for k in names]
    st = lib().lj_gen_reference_host(int(n_dim), int(n_fact), *ptr)
    if st != _abi.OK:
        raise LJError(st, lib().lj_last_error(None).decode())
    return out


def shuffle_host(arr, seed, stream):
    """In-place Fisher-Yates shuffle of an int32 numpy array."""
    st = lib().lj_gen_shuffle_host(seed, stream, arr.size, arr.ctypes.data if arr.size else None)
    if st != _abi.OK:
        raise LJError(st, "lj_gen_shuffle_host")
    return arr


def star_desc(fact_fk, fact_measure, n_fact, dims, order, method="lip", bloom_order=None,
              group_mult=(0, 0, 0, 0), group_offset=0, n_groups=1, resident=None):
    q = _abi.StarDesc()
    nd = len(dims) if dims else len(resident)
    q.n_dims = nd
    for d in range(nd):
        q.fact_fk[d] = _ptr(fact_fk[d])
        if dims and dims[d] is not None:
            q.dims[d] = dims[d]
        if resident and resident[d] is not None:
            q.resident[d] = resident[d].handle
        q.group_mult[d] = int(group_mult[d])
    q.fact_measure, q.n_fact = _ptr(fact_measure), int(n_fact)
    q.method = METHODS[method] if isinstance(method, str) else method
    for i in range(MAX_DIMS):
        q.order[i] = order[i] if i < nd else i
        q.bloom_order[i] = (bloom_order[i] if bloom_order is not None and i < nd else -1)
    q.group_offset, q.n_groups = int(group_offset), int(n_groups)
    return q
