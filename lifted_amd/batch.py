"""Batched transforms of many 1-D signals of uneven length, wavelet and level
count, packed back to back in one tensor (see lfx_batch_* in the C ABI)."""
import ctypes

from . import _lib
from . import _wavelet_id, _storage_torch, LFX_FLAG_PLAIN


class LfxSignal(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_int64), ("length", ctypes.c_int64), ("wavelet", ctypes.c_uint32),
                ("levels", ctypes.c_int32)]


def _bind():
    L = _lib.lib()
    if not hasattr(L, "_batch_bound"):
        vp = ctypes.c_void_p
        L.lfx_batch_plan_create.argtypes = [ctypes.POINTER(LfxSignal), ctypes.c_int32, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                            ctypes.POINTER(vp)]
        L.lfx_batch_execute.argtypes = [vp, vp, vp, vp]
        L.lfx_batch_num_stages.argtypes = [vp]
        L.lfx_batch_execute_stage.argtypes = [vp, ctypes.c_int32, vp, vp, vp]
        L.lfx_batch_stage_info.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
        L.lfx_batch_plan_destroy.argtypes = [vp]
        L.lfx_batch_plan_destroy.restype = None
        L.lfx_set_tuning.argtypes = [ctypes.c_int64, ctypes.c_int32]
        L.lfx_batch_launch_info.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(LaunchInfo)]
        L.lfx_batch_execute_stage_timed.argtypes = [vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32,
                                                    ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L._batch_bound = True
    return L


class LaunchInfo(ctypes.Structure):
    """lfx_launch_info: one kernel launch of a batched stage."""
    _fields_ = [("kind", ctypes.c_int32), ("threads", ctypes.c_int32), ("nitems", ctypes.c_int64),
                ("alg_bytes", ctypes.c_double), ("lds_bytes", ctypes.c_int64)]


def set_tuning(tile=0, levels_per_pass=0):
    """Tile length / levels per tiled pass for plans created afterwards (0 = default)."""
    _lib.check(_bind().lfx_set_tuning(int(tile), int(levels_per_pass)))


class BatchPlan:
    """Plan for one transform type over a fixed list of packed signals.

    signals: iterable of (offset, length, wavelet, levels); levels 0 = max level.
    dtype: torch dtype of the packed arrays (float32, float64, bfloat16, float16).
    tuning: planning knobs from lifted_amd.tune.tune_batch (None: defaults).
    """

    def __init__(self, signals, transform, mode=0, dtype=None, plain=False, tuning=None, lds_only=False):
        """plain: one launch per level (no multi-level fusion; float32 plans
        still run each level as register tiles). lds_only: LDS kernels only.
        tuning: planning knobs (see lifted_amd.tune) used for this plan only."""
        import torch
        if tuning:
            from . import tune
            with tune.tuning(**tuning):
                self.__init__(signals, transform, mode, dtype, plain, None, lds_only)
            return
        L = _bind()
        sigs = list(signals)
        arr = (LfxSignal * max(len(sigs), 1))()
        for i, (off, n, w, lv) in enumerate(sigs):
            arr[i].offset, arr[i].length, arr[i].wavelet, arr[i].levels = int(off), int(n), _wavelet_id(w), int(lv)
        self.dtype = dtype or torch.float32
        storage = _storage_torch(self.dtype)
        h = ctypes.c_void_p()
        flags = (LFX_FLAG_PLAIN if plain else 0) | (_lib.LFX_FLAG_LDS_ONLY if lds_only else 0)
        _lib.check(L.lfx_batch_plan_create(arr, len(sigs), int(transform), int(mode), storage, flags,
                                           ctypes.byref(h)))
        self._h = h
        self.transform = int(transform)
        self.signals = sigs
        self.total = max((o + n for o, n, _, _ in sigs), default=0)
        self.num_stages = L.lfx_batch_num_stages(h)

    def _check_args(self, x, out):
        if x.dtype != self.dtype or out.dtype != self.dtype:
            raise ValueError("array dtype does not match the plan")
        if not (x.is_cuda and out.is_cuda and x.is_contiguous() and out.is_contiguous()):
            raise ValueError("packed arrays must be contiguous device tensors")
        if x.numel() < self.total or out.numel() < self.total:
            raise ValueError("packed arrays are shorter than the signal list")
        if x.data_ptr() == out.data_ptr():
            raise ValueError("input and output must not alias")

    def execute(self, x, out=None):
        import torch
        if out is None:
            out = torch.empty_like(x)
        self._check_args(x, out)
        st = torch.cuda.current_stream(x.device).cuda_stream
        _lib.check(_bind().lfx_batch_execute(self._h, x.data_ptr(), out.data_ptr(), st))
        return out

    def execute_stage(self, stage, x, out):
        import torch
        self._check_args(x, out)
        st = torch.cuda.current_stream(x.device).cuda_stream
        _lib.check(_bind().lfx_batch_execute_stage(self._h, int(stage), x.data_ptr(), out.data_ptr(), st))
        return out

    def stage_info(self, stage):
        n = ctypes.c_int64()
        b = ctypes.c_double()
        l = ctypes.c_int64()
        _lib.check(_bind().lfx_batch_stage_info(self._h, int(stage), ctypes.byref(n), ctypes.byref(b),
                                                ctypes.byref(l)))
        return {"items": n.value, "alg_bytes": b.value, "lds_bytes": l.value}

    def launch_info(self, stage):
        """Kernel launches of a stage (LDS launches first, then the register
        launch, then the row launch): dicts with kind ("lds" / "register" /
        "rows"), threads, items, alg_bytes, lds_bytes."""
        arr = (LaunchInfo * 16)()
        n = _lib.check(_bind().lfx_batch_launch_info(self._h, int(stage), 16, arr))
        kinds = {0: "lds", 1: "register", 2: "rows"}
        return [{"kind": kinds[arr[i].kind], "threads": arr[i].threads,
                 "items": arr[i].nitems, "alg_bytes": arr[i].alg_bytes, "lds_bytes": arr[i].lds_bytes}
                for i in range(n)]

    def execute_stage_timed(self, stage, x, out, events):
        """execute_stage, recording events[i] = (start, end) (torch.cuda.Event
        with timing, already recorded once so that they exist) around launch i
        on the stream it goes to. Nothing is synchronized."""
        import torch
        self._check_args(x, out)
        st = torch.cuda.current_stream(x.device).cuda_stream
        n = len(events)
        starts = (ctypes.c_void_p * max(n, 1))(*[e[0].cuda_event for e in events])
        ends = (ctypes.c_void_p * max(n, 1))(*[e[1].cuda_event for e in events])
        _lib.check(_bind().lfx_batch_execute_stage_timed(self._h, int(stage), x.data_ptr(), out.data_ptr(), st, n,
                                                         starts, ends))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _bind().lfx_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChainLaunchInfo(ctypes.Structure):
    """lfx_chain_launch_info: one kernel launch of a chain execution."""
    _fields_ = [("chunk", ctypes.c_int32), ("transform_index", ctypes.c_int32), ("stage", ctypes.c_int32),
                ("big", ctypes.c_int32), ("launch", LaunchInfo)]


def _bind_chain():
    L = _bind()
    if not hasattr(L, "_chain_bound"):
        vp = ctypes.c_void_p
        L.lfx_batch_chain_create.argtypes = [ctypes.POINTER(LfxSignal), ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                             ctypes.c_int32, ctypes.POINTER(vp)]
        L.lfx_batch_chain_execute.argtypes = [vp, vp, ctypes.POINTER(vp), vp]
        L.lfx_batch_chain_num_chunks.argtypes = [vp]
        L.lfx_batch_chain_launch_info.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ChainLaunchInfo)]
        L.lfx_batch_chain_execute_timed.argtypes = [vp, vp, ctypes.POINTER(vp), vp, ctypes.c_int32,
                                                    ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.lfx_batch_chain_destroy.argtypes = [vp]
        L.lfx_batch_chain_destroy.restype = None
        L._chain_bound = True
    return L


class BatchChain:
    """A sequence of batched transforms of one signal list (x -> outs[0] ->
    outs[1] ...), pipelined over `chunks` groups of signals: each plan's big
    stage (the finest levels of every signal) runs one chunk after another,
    and each chunk's short coarse-level stages run beside the next chunk's big
    stage on a second stream. Results equal running BatchPlan per transform.

    signals: (offset, length, wavelet, levels) tuples; transforms: Transform
    values in order; tuning: planning knobs (lifted_amd.tune) for its plans.
    """

    def __init__(self, signals, transforms, mode=0, dtype=None, chunks=2, plain=False, lds_only=False, tuning=None):
        import torch
        if tuning:
            from . import tune
            with tune.tuning(**tuning):
                self.__init__(signals, transforms, mode, dtype, chunks, plain, lds_only)
            return
        L = _bind_chain()
        sigs = list(signals)
        arr = (LfxSignal * max(len(sigs), 1))()
        for i, (off, n, w, lv) in enumerate(sigs):
            arr[i].offset, arr[i].length, arr[i].wavelet, arr[i].levels = int(off), int(n), _wavelet_id(w), int(lv)
        self.transforms = [int(t) for t in transforms]
        tf = (ctypes.c_uint32 * len(self.transforms))(*self.transforms)
        self.dtype = dtype or torch.float32
        storage = _storage_torch(self.dtype)
        flags = (LFX_FLAG_PLAIN if plain else 0) | (_lib.LFX_FLAG_LDS_ONLY if lds_only else 0)
        h = ctypes.c_void_p()
        _lib.check(L.lfx_batch_chain_create(arr, len(sigs), tf, len(self.transforms), int(mode), storage, flags,
                                            int(chunks), ctypes.byref(h)))
        self._h = h
        self.signals = sigs
        self.total = max((o + n for o, n, _, _ in sigs), default=0)
        self.num_chunks = L.lfx_batch_chain_num_chunks(h)

    def _outs(self, x, outs):
        import torch
        if outs is None:
            outs = [torch.empty_like(x) for _ in self.transforms]
        if len(outs) != len(self.transforms):
            raise ValueError("one output per transform is needed")
        for a in [x] + list(outs):
            if a.dtype != self.dtype or not a.is_cuda or not a.is_contiguous() or a.numel() < self.total:
                raise ValueError("packed arrays must be contiguous device tensors of the chain's dtype and length")
        return outs

    def execute(self, x, outs=None):
        """Runs every transform; returns the list of outputs (the last one is the result)."""
        import torch
        outs = self._outs(x, outs)
        ptrs = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        st = torch.cuda.current_stream(x.device).cuda_stream
        _lib.check(_bind_chain().lfx_batch_chain_execute(self._h, x.data_ptr(), ptrs, st))
        return outs

    def launch_info(self):
        """Launches of one execution in launch order: dicts with chunk,
        transform_index, stage, big, kind, threads, items, alg_bytes."""
        L = _bind_chain()
        n = _lib.check(L.lfx_batch_chain_launch_info(self._h, 0, None))
        arr = (ChainLaunchInfo * max(n, 1))()
        _lib.check(L.lfx_batch_chain_launch_info(self._h, n, arr))
        kinds = {0: "lds", 1: "register", 2: "rows"}
        return [{"chunk": a.chunk, "transform_index": a.transform_index, "stage": a.stage, "big": bool(a.big),
                 "kind": kinds[a.launch.kind], "threads": a.launch.threads, "items": a.launch.nitems,
                 "alg_bytes": a.launch.alg_bytes} for a in arr[:n]]

    def execute_timed(self, x, outs, events):
        """execute, recording events[i] = (start, end) (torch.cuda.Event with
        timing, recorded once beforehand so that they exist) around launch i
        (launch_info order) on the stream it goes to."""
        import torch
        outs = self._outs(x, outs)
        ptrs = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        n = len(events)
        starts = (ctypes.c_void_p * max(n, 1))(*[e[0].cuda_event for e in events])
        ends = (ctypes.c_void_p * max(n, 1))(*[e[1].cuda_event for e in events])
        st = torch.cuda.current_stream(x.device).cuda_stream
        _lib.check(_bind_chain().lfx_batch_chain_execute_timed(self._h, x.data_ptr(), ptrs, st, n, starts, ends))
        return outs

    def close(self):
        if getattr(self, "_h", None):
            _bind_chain().lfx_batch_chain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
