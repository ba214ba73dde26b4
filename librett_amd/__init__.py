"""Python binding of the gfx950 transpose library (``librett.so``).

The library itself is C++/HIP; this module binds its C-ABI twin
(``include/librett_c.h``) with ctypes so tests and benchmarks can drive it.
Names and result codes follow the reference API (``src/librett.h:40-121``):
``librettPlan``, ``librettPlanMeasure``, ``librettExecute``,
``librettDestroy``, ``librettInitialize``, ``librettFinalize``.

There is no fallback: if ``librett.so`` is missing, importing the binding
raises, so nothing can silently run a host path instead of the HIP kernels.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# LIBRETT_SO selects another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("LIBRETT_SO") or os.path.join(_HERE, "librett.so")

LIBRETT_SUCCESS = 0
LIBRETT_INVALID_PLAN = 1
LIBRETT_INVALID_PARAMETER = 2
LIBRETT_INVALID_DEVICE = 3
LIBRETT_INTERNAL_ERROR = 4
LIBRETT_UNDEFINED_ERROR = 5

TRIVIAL, PACKED, PACKED_SPLIT, TILED, TILED_COPY = 1, 2, 3, 4, 5
METHOD_NAMES = {1: "Trivial", 2: "Packed", 3: "PackedSplit", 4: "Tiled", 5: "TiledCopy"}

_LIB = None


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int), ("elem", ctypes.c_int),
        ("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("vec_x", ctypes.c_int), ("vec_y", ctypes.c_int),
        ("word", ctypes.c_int), ("copy_tx", ctypes.c_int),
        ("nreg", ctypes.c_int), ("num_split", ctypes.c_int), ("split_rank", ctypes.c_int),
        ("size_mm", ctypes.c_int), ("size_mk", ctypes.c_int), ("vol_mmk", ctypes.c_int), ("vol_mbar", ctypes.c_int),
        ("grid_x", ctypes.c_int), ("grid_y", ctypes.c_int), ("block", ctypes.c_int), ("lds_bytes", ctypes.c_int),
        ("active_blocks", ctypes.c_int),
        ("predicted_s", ctypes.c_double), ("measured_s", ctypes.c_double),
        ("read_lines", ctypes.c_double), ("write_lines", ctypes.c_double), ("partial_lines", ctypes.c_double),
        ("bytes", ctypes.c_longlong),
        ("feat", ctypes.c_double * 8),
        ("desc", ctypes.c_char * 192),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["desc"] = self.desc.decode()
        d["feat"] = list(self.feat)
        d["method_name"] = METHOD_NAMES.get(self.method, "?")
        return d


class DeviceDesc(ctypes.Structure):
    _fields_ = [
        ("cu_count", ctypes.c_int), ("lds_per_block", ctypes.c_int), ("lds_per_cu", ctypes.c_int),
        ("max_threads_per_block", ctypes.c_int), ("max_waves_per_cu", ctypes.c_int),
        ("hbm_bytes_per_s", ctypes.c_double),
    ]


def mi355x_desc():
    """Published MI355X figures for host-only planning."""
    return DeviceDesc(256, 160 * 1024, 160 * 1024, 1024, 32, 8.0e12)


EXPORTED = [
    "librett_c_initialize", "librett_c_finalize", "librett_c_plan", "librett_c_plan_measure",
    "librett_c_destroy", "librett_c_execute", "librett_c_plan_info_get", "librett_c_candidates",
    "librett_c_plan_candidate", "librett_c_host_candidates", "librett_c_host_replay",
    "librett_c_reduce_ranks", "librett_c_device_copy", "librett_c_set_model_constants", "librett_c_version",
    "librett_c_batch_plan", "librett_c_batch_set_buffers", "librett_c_batch_execute", "librett_c_batch_destroy",
    "librett_c_batch_info_get", "librett_c_batch_host_replay", "librett_c_batch_host_candidate",
    "librett_c_server_create", "librett_c_server_alloc", "librett_c_server_free", "librett_c_server_submit",
    "librett_c_server_wait", "librett_c_server_drain", "librett_c_server_stats_get", "librett_c_server_reset_stats",
    "librett_c_server_destroy", "librett_c_server_run_poisson",
    "librett_c_mlp_sizes", "librett_c_measure_candidates", "librett_c_selector_features", "librett_c_selector_choose",
    "librett_c_selector_set", "librett_c_selector_load", "librett_c_selector_generation", "librett_c_mlp_forward",
    "librett_c_mlp_grad", "librett_c_mlp_adam", "librett_c_mlp_train", "librett_c_mlp_host_forward",
    "librett_c_mlp_host_grad", "librett_c_mlp_host_adam",
    "librett_c_check_fill", "librett_c_check_transpose", "librett_c_host_check_fill", "librett_c_host_check_transpose",
    "librett_c_check_transpose_noalloc",
    "librett_c_set_allocator", "librett_c_allocator_live",
    "librett_c_plan_counters", "librett_c_candidate_counters", "librett_c_host_candidate_counters",
    "librett_c_host_access_trace", "librett_c_selector_margin",
]
LIBRETT_BUSY = 6


class Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_ulonglong) for n in (
        "ld_req", "st_req", "ld_tran", "st_tran", "ld_bytes", "st_bytes", "read_lines", "write_lines",
        "partial_lines", "items")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong),
        ("bytes", ctypes.c_longlong), ("batches", ctypes.c_longlong), ("mean_batch", ctypes.c_double),
        ("admission_stalls", ctypes.c_longlong), ("stall_seconds", ctypes.c_double),
        ("arena_peak_bytes", ctypes.c_longlong), ("arena_in_use_bytes", ctypes.c_longlong),
        ("p50_us", ctypes.c_double), ("p90_us", ctypes.c_double), ("p99_us", ctypes.c_double),
        ("max_us", ctypes.c_double), ("mean_us", ctypes.c_double), ("wall_seconds", ctypes.c_double),
        ("gbs", ctypes.c_double), ("plan_misses", ctypes.c_longlong), ("plan_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class PoissonConfig(ctypes.Structure):
    _fields_ = [
        ("qps", ctypes.c_double), ("requests", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("catalog", ctypes.c_int),
        ("min_bytes", ctypes.c_double), ("max_bytes", ctypes.c_double), ("rank_lo", ctypes.c_int),
        ("rank_hi", ctypes.c_int), ("elems", ctypes.c_int * 3), ("nelems", ctypes.c_int), ("fresh", ctypes.c_int),
        ("inputs_in_arena", ctypes.c_int), ("clients", ctypes.c_int),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int), ("descriptors", ctypes.c_int), ("total_units", ctypes.c_int), ("grid", ctypes.c_int),
        ("tuned", ctypes.c_int), ("method_count", ctypes.c_int * 6), ("bytes", ctypes.c_longlong),
        ("plan_seconds", ctypes.c_double), ("tune_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["method_count"] = {METHOD_NAMES.get(i, str(i)): self.method_count[i] for i in range(6) if self.method_count[i]}
        return d


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)
_ALLOCATOR = None  # keeps the installed callbacks alive


def set_allocator(alloc=None, free=None):
    """Installs Python callables alloc(nbytes) -> device pointer (int, 0 on
    failure) and free(ptr) as the library's device allocator; no arguments
    restore hipMalloc / hipFree."""
    global _ALLOCATOR
    if alloc is None:
        r = lib().librett_c_set_allocator(ALLOC_FN(), FREE_FN(), None)
        _ALLOCATOR = None
        return r
    fa = ALLOC_FN(lambda n, ctx: alloc(n) or None)
    ff = FREE_FN(lambda p, ctx: free(p))
    r = lib().librett_c_set_allocator(fa, ff, None)
    if r == LIBRETT_SUCCESS:
        _ALLOCATOR = (fa, ff)
    return r


def allocator_live():
    return lib().librett_c_allocator_live()


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librett.so is not built (run __graft_entry__.build() or make -C librett_amd)")
        from . import hip
        hip.runtime()  # the HIP runtime the library links, loaded first and globally
        L = ctypes.CDLL(LIB_PATH)
        ip = ctypes.POINTER(ctypes.c_int)
        up = ctypes.POINTER(ctypes.c_uint)
        vp = ctypes.c_void_p
        L.librett_c_initialize.argtypes = []
        L.librett_c_finalize.argtypes = []
        L.librett_c_plan.argtypes = [up, ctypes.c_int, ip, ip, ctypes.c_size_t, vp]
        L.librett_c_plan_measure.argtypes = [up, ctypes.c_int, ip, ip, ctypes.c_size_t, vp, vp, vp]
        L.librett_c_destroy.argtypes = [ctypes.c_uint]
        L.librett_c_execute.argtypes = [ctypes.c_uint, vp, vp]
        L.librett_c_plan_info_get.argtypes = [ctypes.c_uint, ctypes.POINTER(PlanInfo)]
        L.librett_c_candidates.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.POINTER(PlanInfo),
                                           ctypes.c_int, ip]
        L.librett_c_plan_candidate.argtypes = [up, ctypes.c_int, ip, ip, ctypes.c_size_t, vp, ctypes.c_int]
        L.librett_c_host_candidates.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.POINTER(DeviceDesc),
                                                ctypes.POINTER(PlanInfo), ctypes.c_int, ip]
        L.librett_c_host_replay.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.POINTER(DeviceDesc),
                                            ctypes.c_int, vp, vp]
        L.librett_c_reduce_ranks.argtypes = [ctypes.c_int, ip, ip, ip, ip, ip]
        L.librett_c_device_copy.argtypes = [vp, vp, ctypes.c_size_t, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.librett_c_set_model_constants.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        pp = ctypes.POINTER(ctypes.c_void_p)
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.librett_c_batch_plan.argtypes = [up, ctypes.c_int, ip, ip, ip, szp, vp, ctypes.c_int]
        L.librett_c_batch_set_buffers.argtypes = [ctypes.c_uint, pp, pp]
        L.librett_c_batch_execute.argtypes = [ctypes.c_uint, ctypes.c_int]
        L.librett_c_batch_destroy.argtypes = [ctypes.c_uint]
        L.librett_c_batch_info_get.argtypes = [ctypes.c_uint, ctypes.POINTER(BatchInfo)]
        L.librett_c_batch_host_replay.argtypes = [ctypes.c_int, ip, ip, ip, szp, ctypes.POINTER(DeviceDesc), pp, pp,
                                                  ctypes.c_int]
        L.librett_c_batch_host_candidate.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.POINTER(DeviceDesc),
                                                     ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, vp,
                                                     vp, ip]
        ullp = ctypes.POINTER(ctypes.c_ulonglong)
        L.librett_c_server_create.argtypes = [up, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
        L.librett_c_server_alloc.argtypes = [ctypes.c_uint, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.librett_c_server_free.argtypes = [ctypes.c_uint, vp]
        L.librett_c_server_submit.argtypes = [ctypes.c_uint, ctypes.c_int, ip, ip, ctypes.c_size_t, vp, vp, ullp]
        L.librett_c_server_wait.argtypes = [ctypes.c_uint, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]
        L.librett_c_server_drain.argtypes = [ctypes.c_uint]
        L.librett_c_server_stats_get.argtypes = [ctypes.c_uint, ctypes.POINTER(ServerStats)]
        L.librett_c_server_reset_stats.argtypes = [ctypes.c_uint]
        L.librett_c_server_destroy.argtypes = [ctypes.c_uint]
        L.librett_c_server_run_poisson.argtypes = [ctypes.c_uint, ctypes.POINTER(PoissonConfig), vp, ctypes.c_size_t]
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        f_ = ctypes.c_float
        L.librett_c_mlp_sizes.argtypes = [ip, ip, ip]
        L.librett_c_measure_candidates.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, vp, vp, vp, dp, ctypes.c_int,
                                                   ip]
        L.librett_c_selector_features.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.POINTER(DeviceDesc),
                                                  fp, ctypes.c_int, ip, ip]
        L.librett_c_selector_choose.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ip]
        L.librett_c_selector_set.argtypes = [fp, ctypes.c_int, fp, fp, ctypes.c_int]
        L.librett_c_selector_load.argtypes = [ctypes.c_char_p]
        L.librett_c_selector_generation.argtypes = []
        L.librett_c_mlp_forward.argtypes = [vp, vp, ctypes.c_int, vp, vp]
        L.librett_c_mlp_grad.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, f_, ctypes.c_int, vp, vp, vp]
        L.librett_c_mlp_adam.argtypes = [vp, ctypes.c_int, vp, vp, vp, f_, f_, f_, f_, f_, f_, vp]
        L.librett_c_mlp_train.argtypes = [vp, vp, vp, fp, ctypes.c_int, vp, ctypes.c_int, f_, ctypes.c_int,
                                          ctypes.c_uint, fp, vp]
        L.librett_c_mlp_host_forward.argtypes = [fp, fp, ctypes.c_int, fp]
        L.librett_c_mlp_host_grad.argtypes = [fp, fp, fp, fp, ip, ctypes.c_int, f_, fp, fp]
        L.librett_c_mlp_host_adam.argtypes = [fp, ctypes.c_int, fp, fp, fp, f_, f_, f_, f_, f_, f_]
        llp = ctypes.POINTER(ctypes.c_longlong)
        L.librett_c_check_fill.argtypes = [vp, ctypes.c_size_t, vp]
        L.librett_c_check_transpose.argtypes = [vp, ctypes.c_int, ip, ip, ctypes.c_size_t, vp, ullp, llp]
        L.librett_c_host_check_fill.argtypes = [vp, ctypes.c_size_t]
        L.librett_c_check_transpose_noalloc.argtypes = [vp, ctypes.c_int, ip, ip, ctypes.c_size_t, vp, vp]
        L.librett_c_host_check_transpose.argtypes = [vp, ctypes.c_int, ip, ip, ctypes.c_size_t, ullp, llp]
        L.librett_c_set_allocator.argtypes = [ALLOC_FN, FREE_FN, vp]
        L.librett_c_allocator_live.argtypes = []
        cp = ctypes.POINTER(Counters)
        L.librett_c_plan_counters.argtypes = [ctypes.c_uint, cp]
        L.librett_c_candidate_counters.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, vp, ctypes.c_int, cp]
        L.librett_c_host_candidate_counters.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t,
                                                        ctypes.POINTER(DeviceDesc), ctypes.c_int, cp]
        L.librett_c_host_access_trace.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.POINTER(DeviceDesc),
                                                  ctypes.c_int, ctypes.c_int, llp, ctypes.c_longlong, llp]
        L.librett_c_selector_margin.argtypes = []
        L.librett_c_version.restype = ctypes.c_char_p
        for name in EXPORTED:
            f = getattr(L, name)
            if name != "librett_c_version" and f.restype is not ctypes.c_char_p:
                f.restype = ctypes.c_int
        L.librett_c_set_model_constants.restype = None
        for name in ("librett_c_mlp_sizes", "librett_c_mlp_host_forward", "librett_c_mlp_host_adam",
                     "librett_c_host_check_fill"):
            getattr(L, name).restype = None
        L.librett_c_selector_generation.restype = ctypes.c_ulonglong
        L.librett_c_selector_margin.restype = ctypes.c_double
        L.librett_c_allocator_live.restype = ctypes.c_ulonglong
        L.librett_c_initialize.restype = None
        L.librett_c_finalize.restype = None
        _LIB = L
    return _LIB


def _iarr(v):
    return (ctypes.c_int * max(1, len(v)))(*[int(x) for x in v])


# ---------------------------------------------------------------------------
# Reference-named entry points
# ---------------------------------------------------------------------------

def librettInitialize():
    lib().librett_c_initialize()


def librettFinalize():
    lib().librett_c_finalize()


def librettPlan(rank, dim, permutation, sizeofType, stream):
    """Returns (result, handle)."""
    h = ctypes.c_uint(0)
    r = lib().librett_c_plan(ctypes.byref(h), int(rank), _iarr(dim), _iarr(permutation), int(sizeofType),
                             ctypes.c_void_p(stream))
    return r, h.value


def librettPlanMeasure(rank, dim, permutation, sizeofType, stream, idata, odata):
    """Returns (result, handle)."""
    h = ctypes.c_uint(0)
    r = lib().librett_c_plan_measure(ctypes.byref(h), int(rank), _iarr(dim), _iarr(permutation),
                                     int(sizeofType), ctypes.c_void_p(stream), ctypes.c_void_p(idata),
                                     ctypes.c_void_p(odata))
    return r, h.value


def librettExecute(handle, idata, odata):
    return lib().librett_c_execute(ctypes.c_uint(handle), ctypes.c_void_p(idata), ctypes.c_void_p(odata))


def librettDestroy(handle):
    return lib().librett_c_destroy(ctypes.c_uint(handle))


# ---------------------------------------------------------------------------
# Extensions
# ---------------------------------------------------------------------------

def plan_candidate(rank, dim, permutation, sizeofType, stream, index):
    h = ctypes.c_uint(0)
    r = lib().librett_c_plan_candidate(ctypes.byref(h), int(rank), _iarr(dim), _iarr(permutation),
                                       int(sizeofType), ctypes.c_void_p(stream), int(index))
    return r, h.value


def plan_info(handle):
    info = PlanInfo()
    r = lib().librett_c_plan_info_get(ctypes.c_uint(handle), ctypes.byref(info))
    if r != LIBRETT_SUCCESS:
        raise ValueError("plan_info failed: %d" % r)
    return info.as_dict()


def candidates(dim, permutation, sizeofType, max_count=512):
    infos = (PlanInfo * max_count)()
    n = ctypes.c_int(0)
    r = lib().librett_c_candidates(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType), infos, max_count,
                                   ctypes.byref(n))
    if r != LIBRETT_SUCCESS:
        raise ValueError("candidates failed: %d" % r)
    return [infos[i].as_dict() for i in range(min(n.value, max_count))]


def host_candidates(dim, permutation, sizeofType, device=None, max_count=512):
    dev = device or mi355x_desc()
    infos = (PlanInfo * max_count)()
    n = ctypes.c_int(0)
    r = lib().librett_c_host_candidates(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                        ctypes.byref(dev), infos, max_count, ctypes.byref(n))
    if r != LIBRETT_SUCCESS:
        raise ValueError("host_candidates failed: %d" % r)
    return [infos[i].as_dict() for i in range(min(n.value, max_count))]


def host_replay(dim, permutation, sizeofType, index, inp, device=None):
    """Runs candidate `index` on the host; `inp` is a numpy array."""
    import numpy as np
    dev = device or mi355x_desc()
    inp = np.ascontiguousarray(inp)
    out = np.zeros_like(inp)
    r = lib().librett_c_host_replay(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType), ctypes.byref(dev),
                                    int(index), ctypes.c_void_p(inp.ctypes.data), ctypes.c_void_p(out.ctypes.data))
    if r != LIBRETT_SUCCESS:
        raise ValueError("host_replay failed: %d" % r)
    return out


def reduce_ranks(dim, permutation):
    rr = ctypes.c_int(0)
    rd = (ctypes.c_int * 64)()
    rp = (ctypes.c_int * 64)()
    r = lib().librett_c_reduce_ranks(len(dim), _iarr(dim), _iarr(permutation), ctypes.byref(rr), rd, rp)
    if r != LIBRETT_SUCCESS:
        raise ValueError("reduce_ranks failed: %d" % r)
    return [rd[i] for i in range(rr.value)], [rp[i] for i in range(rr.value)]


def device_copy(dst, src, nbytes, stream, words_per_thread=4, nontemporal=0, grid=0):
    """Hand-written D2D copy kernel (bandwidth ceiling); asynchronous."""
    return lib().librett_c_device_copy(ctypes.c_void_p(dst), ctypes.c_void_p(src), int(nbytes),
                                       ctypes.c_void_p(stream), int(words_per_thread), int(nontemporal), int(grid))


def set_model_constants(values):
    """Overrides the analytic model's constants (see librett_c.h)."""
    arr = (ctypes.c_double * len(values))(*values)
    lib().librett_c_set_model_constants(arr, len(values))


def _flatten(problems):
    ranks = [len(d) for d, _, _ in problems]
    dims = [x for d, _, _ in problems for x in d]
    perms = [x for _, p, _ in problems for x in p]
    elems = [e for _, _, e in problems]
    return ((ctypes.c_int * len(ranks))(*ranks), _iarr(dims), _iarr(perms), (ctypes.c_size_t * len(elems))(*elems))


def _ptrs(v):
    return (ctypes.c_void_p * len(v))(*[ctypes.c_void_p(int(x)) for x in v])


class Batch:
    """Many transposes in one persistent launch (extension of the API).

    ``problems`` is a list of (dims, perm, elem). ``tune`` = 0 uses the cost
    model, k > 0 times the model's k best batch plans of each problem."""

    def __init__(self, problems, stream, tune=0):
        self.problems = list(problems)
        h = ctypes.c_uint(0)
        r = lib().librett_c_batch_plan(ctypes.byref(h), len(self.problems), *_flatten(self.problems),
                                       ctypes.c_void_p(stream), int(tune))
        if r != LIBRETT_SUCCESS:
            raise ValueError("batch plan failed: %d" % r)
        self.handle = h.value

    def set_buffers(self, inputs, outputs):
        r = lib().librett_c_batch_set_buffers(self.handle, _ptrs(inputs), _ptrs(outputs))
        if r != LIBRETT_SUCCESS:
            raise ValueError("batch set_buffers failed: %d" % r)

    def execute(self, use_graph=True):
        return lib().librett_c_batch_execute(self.handle, 1 if use_graph else 0)

    def info(self):
        info = BatchInfo()
        r = lib().librett_c_batch_info_get(self.handle, ctypes.byref(info))
        if r != LIBRETT_SUCCESS:
            raise ValueError("batch info failed: %d" % r)
        return info.as_dict()

    def destroy(self):
        if self.handle:
            lib().librett_c_batch_destroy(self.handle)
            self.handle = 0


class Server:
    """Continuous-batching serving engine (extension of the API)."""

    def __init__(self, arena_bytes, slots=8, max_batch=64):
        h = ctypes.c_uint(0)
        r = lib().librett_c_server_create(ctypes.byref(h), int(arena_bytes), int(slots), int(max_batch))
        if r != LIBRETT_SUCCESS:
            raise ValueError("server create failed: %d" % r)
        self.handle = h.value

    def alloc(self, nbytes):
        """Device pointer from the arena, or None when it is full."""
        p = ctypes.c_void_p()
        r = lib().librett_c_server_alloc(self.handle, int(nbytes), ctypes.byref(p))
        if r == LIBRETT_BUSY:
            return None
        if r != LIBRETT_SUCCESS:
            raise ValueError("server alloc failed: %d" % r)
        return p.value

    def free(self, ptr):
        return lib().librett_c_server_free(self.handle, ctypes.c_void_p(ptr))

    def submit(self, dim, perm, elem, idata, odata):
        t = ctypes.c_ulonglong(0)
        r = lib().librett_c_server_submit(self.handle, len(dim), _iarr(dim), _iarr(perm), int(elem),
                                          ctypes.c_void_p(idata), ctypes.c_void_p(odata), ctypes.byref(t))
        if r != LIBRETT_SUCCESS:
            raise ValueError("server submit failed: %d" % r)
        return t.value

    def wait(self, ticket):
        lat = ctypes.c_double(0)
        r = lib().librett_c_server_wait(self.handle, ctypes.c_ulonglong(ticket), ctypes.byref(lat))
        if r != LIBRETT_SUCCESS:
            raise ValueError("server wait failed: %d" % r)
        return lat.value

    def drain(self):
        lib().librett_c_server_drain(self.handle)

    def stats(self):
        st = ServerStats()
        lib().librett_c_server_stats_get(self.handle, ctypes.byref(st))
        return st.as_dict()

    def reset_stats(self):
        lib().librett_c_server_reset_stats(self.handle)

    def run_poisson(self, pool_ptr, pool_bytes, qps, requests, seed=20261202, catalog=1024, min_bytes=4096,
                    max_bytes=64 << 20, rank_lo=2, rank_hi=6, elems=(4, 8, 16), fresh=False, inputs_in_arena=False,
                    clients=1):
        cfg = PoissonConfig(float(qps), int(requests), int(seed), int(catalog), float(min_bytes), float(max_bytes),
                            int(rank_lo), int(rank_hi), (ctypes.c_int * 3)(*(list(elems) + [4, 4, 4])[:3]),
                            len(elems), int(bool(fresh)), int(bool(inputs_in_arena)), int(clients))
        r = lib().librett_c_server_run_poisson(self.handle, ctypes.byref(cfg), ctypes.c_void_p(pool_ptr),
                                               int(pool_bytes))
        if r != LIBRETT_SUCCESS:
            raise ValueError("run_poisson failed: %d" % r)
        return self.stats()

    def destroy(self):
        if self.handle:
            lib().librett_c_server_destroy(self.handle)
            self.handle = 0


def batch_host_replay(problems, inputs, grid=0, device=None):
    """Host replay of the batched kernel; `inputs` are numpy arrays."""
    import numpy as np
    dev = device or mi355x_desc()
    ins = [np.ascontiguousarray(x) for x in inputs]
    outs = [np.zeros_like(x) for x in ins]
    r = lib().librett_c_batch_host_replay(len(problems), *_flatten(problems), ctypes.byref(dev),
                                          _ptrs([x.ctypes.data for x in ins]), _ptrs([y.ctypes.data for y in outs]),
                                          int(grid))
    if r != LIBRETT_SUCCESS:
        raise ValueError("batch host replay failed: %d" % r)
    return outs


def batch_host_candidates(dim, permutation, sizeofType, device=None):
    """Number of batch-mode candidates of one problem."""
    n = ctypes.c_int(0)
    r = lib().librett_c_batch_host_candidate(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                             ctypes.byref(device or mi355x_desc()), 0, 0, 0, 0, None, None,
                                             ctypes.byref(n))
    if r != LIBRETT_SUCCESS:
        raise ValueError("batch candidates failed: %d" % r)
    return n.value


def batch_host_replay_candidate(dim, permutation, sizeofType, x, index, fallback=False, unit_bytes=0, grid=0,
                                device=None):
    """Host replay of batch-mode candidate `index`; returns (status, output)."""
    import numpy as np
    x = np.ascontiguousarray(x)
    y = np.zeros_like(x)
    n = ctypes.c_int(0)
    r = lib().librett_c_batch_host_candidate(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                             ctypes.byref(device or mi355x_desc()), int(index), int(bool(fallback)),
                                             int(unit_bytes), int(grid), x.ctypes.data, y.ctypes.data, ctypes.byref(n))
    return r, y


# ---------------------------------------------------------------------------
# Learned plan selector (extension)
# ---------------------------------------------------------------------------

def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_sizes():
    """(features, hidden width, parameter count) of the selector's MLP."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().librett_c_mlp_sizes(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def measure_candidates(dim, permutation, sizeofType, stream, idata, odata, max_count=1024):
    """Times of every candidate (librettPlanMeasure's loop), in seconds."""
    import numpy as np
    t = np.full(max_count, -1.0)
    n = ctypes.c_int(0)
    r = lib().librett_c_measure_candidates(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                           ctypes.c_void_p(stream), ctypes.c_void_p(idata), ctypes.c_void_p(odata),
                                           t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), max_count,
                                           ctypes.byref(n))
    if r != LIBRETT_SUCCESS:
        raise ValueError("measure_candidates failed: %d" % r)
    return t[:min(n.value, max_count)].copy()


def selector_features(dim, permutation, sizeofType, device=None, max_count=1024):
    """(features [n, 50] float32, analytic model's choice). device None: the current GPU."""
    import numpy as np
    nf = mlp_sizes()[0]
    f = np.zeros((max_count, nf), dtype=np.float32)
    n, mc = ctypes.c_int(0), ctypes.c_int(-1)
    r = lib().librett_c_selector_features(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                          ctypes.byref(device) if device is not None else None, _fptr(f), max_count,
                                          ctypes.byref(n), ctypes.byref(mc))
    if r != LIBRETT_SUCCESS:
        raise ValueError("selector_features failed: %d" % r)
    return f[:min(n.value, max_count)].copy(), mc.value


def selector_margin():
    """Margin (log2 of predicted time) by which the selector must beat the
    analytic model's choice for librettPlan to take its pick."""
    return lib().librett_c_selector_margin()


def selector_choose(dim, permutation, sizeofType):
    i = ctypes.c_int(-1)
    r = lib().librett_c_selector_choose(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType), ctypes.byref(i))
    if r != LIBRETT_SUCCESS:
        raise ValueError("selector_choose failed: %d" % r)
    return i.value


def selector_set(params, mean, std):
    import numpy as np
    if params is None:
        return lib().librett_c_selector_set(None, 0, None, None, 0)
    p = np.ascontiguousarray(params, dtype=np.float32)
    m = np.ascontiguousarray(mean, dtype=np.float32)
    s = np.ascontiguousarray(std, dtype=np.float32)
    return lib().librett_c_selector_set(_fptr(p), p.size, _fptr(m), _fptr(s), m.size)


def selector_load(path):
    return lib().librett_c_selector_load(os.fsencode(path))


def selector_generation():
    return lib().librett_c_selector_generation()


def mlp_host_forward(params, x):
    import numpy as np
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(xx.shape[0], dtype=np.float32)
    lib().librett_c_mlp_host_forward(_fptr(p), _fptr(xx), xx.shape[0], _fptr(out))
    return out


def mlp_host_grad(params, x, y, idx, inv_m, w=None):
    """Host replay of one training workgroup: (gradient, weighted squared-error sum)."""
    import numpy as np
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    yy = np.ascontiguousarray(y, dtype=np.float32)
    ww = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
    ii = np.ascontiguousarray(idx, dtype=np.int32)
    g = np.zeros(p.size, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    r = lib().librett_c_mlp_host_grad(_fptr(p), _fptr(xx), _fptr(yy), None if ww is None else _fptr(ww),
                                      ii.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ii.size, float(inv_m), _fptr(g),
                                      _fptr(loss))
    if r != LIBRETT_SUCCESS:
        raise ValueError("mlp_host_grad failed: %d" % r)
    return g, float(loss[0])


def mlp_host_adam(partials, params, m, v, lr, beta1, beta2, c1, c2, eps):
    """Host replay of the Adam kernel; updates params, m, v (float32 arrays) in place."""
    import numpy as np
    g = np.ascontiguousarray(partials, dtype=np.float32)
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    lib().librett_c_mlp_host_adam(_fptr(g), g.shape[0], _fptr(params), _fptr(m), _fptr(v), lr, beta1, beta2, c1, c2,
                                  eps)


def _ok(r, what):
    if r != LIBRETT_SUCCESS:
        raise RuntimeError("%s failed: %d" % (what, r))


def mlp_forward(d_params, d_x, n, d_out, stream):
    _ok(lib().librett_c_mlp_forward(d_params, d_x, n, d_out, ctypes.c_void_p(stream)), "mlp_forward")


def mlp_grad(d_params, d_x, d_y, d_idx, base, end, inv_m, nblocks, d_partial, d_loss, stream, d_w=None):
    _ok(lib().librett_c_mlp_grad(d_params, d_x, d_y, d_w, d_idx, base, end, inv_m, nblocks, d_partial, d_loss,
                                 ctypes.c_void_p(stream)), "mlp_grad")


def mlp_adam(d_partial, nparts, d_params, d_m, d_v, lr, beta1, beta2, c1, c2, eps, stream):
    _ok(lib().librett_c_mlp_adam(d_partial, nparts, d_params, d_m, d_v, lr, beta1, beta2, c1, c2, eps,
                                 ctypes.c_void_p(stream)), "mlp_adam")


def mlp_train(d_x, d_y, n, d_params, epochs, lr, minibatch, seed, stream, d_w=None, w_host=None):
    """Trains on the GPU; returns the per-epoch (weighted) mean squared error.
    d_w / w_host: per-sample weights on the device and the same on the host
    (numpy float32), both or neither."""
    import numpy as np
    loss = np.zeros(max(1, epochs), dtype=np.float32)
    wh = None if w_host is None else np.ascontiguousarray(w_host, dtype=np.float32)
    _ok(lib().librett_c_mlp_train(d_x, d_y, d_w, None if wh is None else _fptr(wh), n, d_params, epochs, lr, minibatch,
                                  seed, _fptr(loss), ctypes.c_void_p(stream)), "mlp_train")
    return loss[:epochs]


# ---------------------------------------------------------------------------
# Index-pattern checker (extension; counterpart of the reference's TensorTester)
# ---------------------------------------------------------------------------

def check_fill(ptr, nbytes, stream):
    """Device buffer: 32-bit word k := k (asynchronous on `stream`)."""
    _ok(lib().librett_c_check_fill(ctypes.c_void_p(ptr), int(nbytes), ctypes.c_void_p(stream)), "check_fill")


def check_transpose(optr, dim, permutation, sizeofType, stream):
    """(wrong elements, first wrong output position or -1) of a device output."""
    n, f = ctypes.c_ulonglong(0), ctypes.c_longlong(-1)
    r = lib().librett_c_check_transpose(ctypes.c_void_p(optr), len(dim), _iarr(dim), _iarr(permutation),
                                        int(sizeofType), ctypes.c_void_p(stream), ctypes.byref(n), ctypes.byref(f))
    if r != LIBRETT_SUCCESS:
        raise ValueError("check_transpose failed: %d" % r)
    return n.value, f.value


def check_transpose_noalloc(optr, dim, permutation, sizeofType, stream, result_ptr):
    """Enqueues the allocation-free check; the device result (2 x uint64 at
    result_ptr) is ready when `stream` is."""
    return lib().librett_c_check_transpose_noalloc(ctypes.c_void_p(optr), len(dim), _iarr(dim),
                                                   _iarr(permutation), int(sizeofType), ctypes.c_void_p(stream),
                                                   ctypes.c_void_p(result_ptr))


def host_check_fill(arr):
    """numpy array (any dtype, size a multiple of 4 bytes): word k := k."""
    lib().librett_c_host_check_fill(ctypes.c_void_p(arr.ctypes.data), arr.nbytes)


def host_check_transpose(arr, dim, permutation, sizeofType):
    n, f = ctypes.c_ulonglong(0), ctypes.c_longlong(-1)
    r = lib().librett_c_host_check_transpose(ctypes.c_void_p(arr.ctypes.data), len(dim), _iarr(dim),
                                             _iarr(permutation), int(sizeofType), ctypes.byref(n), ctypes.byref(f))
    if r != LIBRETT_SUCCESS:
        raise ValueError("host_check_transpose failed: %d" % r)
    return n.value, f.value


# ---------------------------------------------------------------------------
# Access counters (extension; counterpart of the reference's
# librettGpuModelKernel)
# ---------------------------------------------------------------------------

def plan_counters(handle):
    """Exact access counts of a stored plan's kernel, on the device."""
    c = Counters()
    _ok(lib().librett_c_plan_counters(ctypes.c_uint(handle), ctypes.byref(c)), "plan_counters")
    return c.as_dict()


def candidate_counters(dim, permutation, sizeofType, index, stream=0):
    """Exact access counts of candidate `index` of candidates(), on the device."""
    c = Counters()
    _ok(lib().librett_c_candidate_counters(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                           ctypes.c_void_p(stream), int(index), ctypes.byref(c)), "candidate_counters")
    return c.as_dict()


def host_candidate_counters(dim, permutation, sizeofType, index, device=None):
    """The same counts from the host run of the counting code (host_candidates() indices)."""
    dev = device or mi355x_desc()
    c = Counters()
    _ok(lib().librett_c_host_candidate_counters(len(dim), _iarr(dim), _iarr(permutation), int(sizeofType),
                                                ctypes.byref(dev), int(index), ctypes.byref(c)),
        "host_candidate_counters")
    return c.as_dict()


def host_access_trace(dim, permutation, sizeofType, index, source, device=None):
    """(n, 6) int64 array of accesses (item, store, instruction, thread, byte offset, bytes).
    source 0: the counters' enumeration; 1: the kernel's host replay."""
    import numpy as np
    dev = device or mi355x_desc()
    n = ctypes.c_longlong(0)
    args = (len(dim), _iarr(dim), _iarr(permutation), int(sizeofType), ctypes.byref(dev), int(index), int(source))
    _ok(lib().librett_c_host_access_trace(*args, None, 0, ctypes.byref(n)), "host_access_trace")
    out = np.zeros((max(1, n.value), 6), dtype=np.int64)
    _ok(lib().librett_c_host_access_trace(*args, out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), n.value,
                                          ctypes.byref(n)), "host_access_trace")
    return out[:n.value]


def version():
    return lib().librett_c_version().decode()
