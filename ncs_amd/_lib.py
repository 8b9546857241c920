"""ctypes binding of libncs_hip.so (include/ncs_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""
import ctypes
import os
import threading

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libncs_hip.so")

NCS_OK = 0
# per-tile status codes (opencl/ncs.cl:66-71)
SUCCESS = 0
UNSTARTED = -1
REACHED_MAXITERS = -2
INCREASING_GRADIENT = -3
MINIMUM_STEP = -4
REACHED_MAXPRECISION = -5

GEOM_REFERENCE = 0
GEOM_BLOCKS = 1
LS_MORE_THUENTE = 0
LS_ARMIJO = 1
START_PLAIN = 0
START_TILES = 1
START_LEARNED = 2
START_LINEAR = 3
PRECISION_FP32 = 0
PRECISION_F16_SPLIT = 1
NET_PARAMS = 19393
LINEAR_START_PARAMS = 513
INPUT_PHOTONS_F32 = 0
INPUT_ADU_U16 = 1
PIPELINE_REPLAYS = 4  # NCS_PIPELINE_REPLAYS
MAX_GAIN_LEVELS = 4  # NCS_MAX_GAIN_LEVELS


class NCSHipError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """ncs_params (include/ncs_hip.h)."""
    _fields_ = [("max_iterations", ctypes.c_int),
                ("history", ctypes.c_int),
                ("epsilon", ctypes.c_float),
                ("linesearch", ctypes.c_int),
                ("max_linesearch", ctypes.c_int),
                ("ftol", ctypes.c_float),
                ("gtol", ctypes.c_float),
                ("xtol", ctypes.c_float),
                ("min_step", ctypes.c_float),
                ("max_step", ctypes.c_float),
                ("stepm", ctypes.c_float),
                ("start", ctypes.c_int),
                ("fallback", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int),
                ("precond", ctypes.c_int),
                ("waves_per_group", ctypes.c_int),
                ("precision", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Frame(ctypes.Structure):
    """ncs_frame (include/ncs_hip.h)."""
    _fields_ = [("data", ctypes.c_void_p),
                ("gamma", ctypes.c_void_p),
                ("out", ctypes.c_void_p),
                ("height", ctypes.c_int),
                ("width", ctypes.c_int),
                ("data_stride", ctypes.c_int),
                ("gamma_stride", ctypes.c_int),
                ("out_stride", ctypes.c_int),
                ("alpha", ctypes.c_float),
                ("geometry", ctypes.c_int)]


class TuneResult(ctypes.Structure):
    """ncs_tune_result (include/ncs_hip.h)."""
    _fields_ = [("history", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int),
                ("waves_per_group", ctypes.c_int),
                ("lds_waves_per_cu", ctypes.c_int),
                ("ms", ctypes.c_float),
                ("ms_default", ctypes.c_float),
                ("candidates", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PipelineFrame(ctypes.Structure):
    """ncs_pipeline_frame (include/ncs_hip.h)."""
    _fields_ = [("height", ctypes.c_int),
                ("width", ctypes.c_int),
                ("alpha", ctypes.c_float),
                ("geometry", ctypes.c_int)]


# Every symbol the header declares, with its ctypes signature.
_vp, _i, _f, _ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
_ip, _fp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
_P = ctypes.POINTER(Params)
_F = ctypes.POINTER(Frame)
SIGNATURES = {
    "ncs_hip_default_params": (None, [_P]),
    "ncs_hip_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "ncs_hip_destroy": (None, [_vp]),
    "ncs_hip_last_error": (ctypes.c_char_p, []),
    "ncs_hip_device_info": (_i, [_vp, ctypes.c_char_p, _i]),
    "ncs_hip_set_otf_mask": (_i, [_vp, _vp]),
    "ncs_hip_set_otf_mask_window": (_i, [_vp, _vp, _i]),
    "ncs_hip_tile_count": (_ll, [_i, _i, _i]),
    "ncs_hip_tile_count_window": (_ll, [_i, _i, _i, _i]),
    "ncs_hip_reduce_frames": (_i, [_vp, _F, _i, _P, _vp, _vp, _vp]),
    "ncs_hip_reduce_frames_device": (_i, [_vp, _F, _i, _P, _vp, _vp, _vp, _vp, _vp]),
    "ncs_hip_reduce_frames_device_noalloc": (_i, [_vp, _F, _i, _P, _vp, _vp, _vp, _vp, _vp]),
    "ncs_hip_solve_tiles": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _P]),
    "ncs_hip_solve_tiles_device": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _P,
                                        _vp]),
    "ncs_hip_eval_tiles": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _f]),
    "ncs_hip_net_set_weights": (_i, [_vp, _vp]),
    "ncs_hip_set_linear_start": (_i, [_vp, _vp]),
    "ncs_hip_set_linear_start_alpha": (_i, [_vp, _vp, ctypes.c_float]),
    "ncs_hip_net_predict_frames_device": (_i, [_vp, _F, _i, _vp, _vp]),
    "ncs_hip_net_forward": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "ncs_hip_net_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "ncsReduceNoise": (None, [_vp, _vp, _vp, _vp, ctypes.c_double, _i, _i, _i]),
    "ncsSRInitialize": (_vp, [_i]),
    "ncsSRNewRegion": (None, [_vp, _vp, _vp]),
    "ncsSRSetOTFMask": (None, [_vp, _vp]),
    "ncsSRSetU": (None, [_vp, _vp]),
    "ncsSRGetU": (None, [_vp, _vp]),
    "ncsSRSolve": (_i, [_vp, ctypes.c_double, _i]),
    "ncsSRCalcLogLikelihood": (ctypes.c_double, [_vp]),
    "ncsSRCalcLLGradient": (None, [_vp, _vp]),
    "ncsSRCalcNoiseContribution": (ctypes.c_double, [_vp]),
    "ncsSRCalcNCGradient": (None, [_vp, _vp]),
    "ncsSRCleanup": (None, [_vp]),
    "ncs_hip_autotune": (_i, [_vp, _F, _i, _P, _i, ctypes.POINTER(TuneResult)]),
    "ncs_hip_pipeline_create": (_i, [_vp, ctypes.POINTER(PipelineFrame), _i, _P, _i, _i,
                                     ctypes.POINTER(_vp)]),
    "ncs_hip_pipeline_destroy": (None, [_vp]),
    "ncs_hip_pipeline_set_gamma": (_i, [_vp, _i, _vp]),
    "ncs_hip_pipeline_set_camera": (_i, [_vp, _i, _vp, _vp, _vp]),
    "ncs_hip_pipeline_set_camera_levels": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp]),
    "ncs_hip_pipeline_launch": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ncs_hip_pipeline_launched": (_ll, [_vp]),
    "ncs_hip_pipeline_wait": (_i, [_vp]),
    "ncs_hip_pipeline_wait_replay": (_i, [_vp, _ll]),
    "ncs_hip_pipeline_replay_results": (_i, [_vp, _ll, _vp, _vp, _vp]),
    "ncs_hip_pipeline_tile_results": (_i, [_vp, _vp, _vp, _vp]),
    "ncs_hip_pipeline_tiles": (_ll, [_vp]),
}

_lib = None
_lock = threading.Lock()


def load():
    """The loaded library (raises NCSHipError if it is not built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NCSHipError(
                "libncs_hip.so is not built (expected at %s); run "
                "`python -m ncs_amd.build` or __graft_entry__.build()" % LIB_PATH)
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7); whichever loads first serves the whole process.
        # Loading torch's first keeps one runtime, the one that owns torch's
        # tensors and streams; the other order leaves torch without a device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.cdll.LoadLibrary(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error():
    return load().ncs_hip_last_error().decode(errors="replace")


def check(rc, what):
    if rc != NCS_OK:
        raise NCSHipError("%s failed (%d): %s" % (what, rc, last_error()))


def default_params(**overrides):
    p = Params()
    load().ncs_hip_default_params(ctypes.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise TypeError("unknown solver parameter " + repr(k))
        setattr(p, k, v)
    return p


def liblbfgs_params(**overrides):
    """The reference C library's solver semantics (liblbfgs 1.10 defaults:
    history 6, eps 1e-5, More-Thuente, scalar initial Hessian)."""
    p = default_params(history=6, precond=0)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def kernel_semantics_params(**overrides):
    """Parameters of the reference GPU kernel (opencl/ncs.cl:58-63): Armijo
    backtracking with c1 = 1e-4, step x0.5 down to 1e-6, history 8,
    epsilon 1e-4, at most 200 iterations."""
    p = default_params(linesearch=LS_ARMIJO, ftol=1e-4, stepm=0.5, min_step=1e-6,
                       history=8, epsilon=1e-4, max_iterations=200, max_linesearch=64,
                       precond=0)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def tile_count(height, width, geometry=GEOM_REFERENCE, window=16):
    """Tiles of a frame; window is the side of the reference geometry's
    windows (the mask's size), -1 when the frame is too small."""
    return int(load().ncs_hip_tile_count_window(int(height), int(width), int(geometry),
                                                int(window)))


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Context(object):
    """One libncs_hip context (device, stream, OTF tables, buffers)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.ncs_hip_create(int(device), ctypes.byref(h)), "ncs_hip_create")
        self.handle = h
        self.device = device
        self.mask_fft = None

    def close(self):
        if self.handle is not None:
            self.lib.ncs_hip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        buf = ctypes.create_string_buffer(256)
        check(self.lib.ncs_hip_device_info(self.handle, buf, 256), "ncs_hip_device_info")
        return buf.value.decode()

    def set_otf_mask(self, mask_fft_order):
        """FFT-order mask, 16 x 16, or r x r (r even, 4..16) for windows of
        side r with the plain start and the reference geometry."""
        m = numpy.ascontiguousarray(mask_fft_order, dtype=numpy.float32)
        r = m.shape[0] if m.ndim == 2 else 0
        if m.shape != (r, r) or r < 4 or r > 16 or r % 2:
            raise ValueError("the OTF mask must be r x r with r even, 4..16")
        check(self.lib.ncs_hip_set_otf_mask_window(self.handle, _ptr(m), r),
              "ncs_hip_set_otf_mask_window")
        self.mask_fft = m.copy()
        self.window = r

    def set_linear_start(self, coef, slope=None, alpha_ref=0.1):
        """Coefficients [A (16x16), B (16x16), c] of the learned linear start,
        fitted at alpha_ref, and optionally their slopes in ln(alpha) (same
        layout): a frame at alpha then uses coef + ln(alpha / alpha_ref) *
        slope.  Without slopes the coefficients are used at every alpha."""
        c = numpy.ascontiguousarray(coef, dtype=numpy.float32)
        if c.shape != (LINEAR_START_PARAMS,):
            raise ValueError("expected %d coefficients" % LINEAR_START_PARAMS)
        check(self.lib.ncs_hip_set_linear_start(self.handle, _ptr(c)),
              "ncs_hip_set_linear_start")
        if slope is not None:
            sl = numpy.ascontiguousarray(slope, dtype=numpy.float32)
            if sl.shape != (LINEAR_START_PARAMS,):
                raise ValueError("expected %d slopes" % LINEAR_START_PARAMS)
            check(self.lib.ncs_hip_set_linear_start_alpha(self.handle, _ptr(sl),
                                                          float(alpha_ref)),
                  "ncs_hip_set_linear_start_alpha")

    def reduce_frames(self, datas, gammas, outs, alphas, geometry, params):
        """Host-memory batch; returns per-tile (iterations, status, cost)."""
        n = len(datas)
        frames = (Frame * n)()
        total = 0
        keep = []
        for i in range(n):
            d = numpy.ascontiguousarray(datas[i], dtype=numpy.float32)
            g = numpy.ascontiguousarray(gammas[i], dtype=numpy.float32)
            o = outs[i]
            if d.ndim != 2 or g.shape != d.shape or o.shape != d.shape:
                raise ValueError("frame %d: data, gamma and out must be 2-D of one shape" % i)
            if o.dtype != numpy.float32 or not o.flags.c_contiguous:
                raise ValueError("out arrays must be C-contiguous float32")
            keep += [d, g]
            fr = frames[i]
            fr.data, fr.gamma, fr.out = d.ctypes.data, g.ctypes.data, o.ctypes.data
            fr.height, fr.width = d.shape
            fr.data_stride = fr.gamma_stride = fr.out_stride = 0
            fr.alpha = float(alphas[i])
            fr.geometry = int(geometry[i] if hasattr(geometry, "__len__") else geometry)
            nt = tile_count(fr.height, fr.width, fr.geometry, getattr(self, "window", 16))
            if nt < 0:
                raise ValueError("frame %d of shape %s is too small for this tiling"
                                 % (i, d.shape))
            total += nt
        iters = numpy.zeros(total, numpy.int32)
        status = numpy.zeros(total, numpy.int32)
        cost = numpy.zeros(total, numpy.float32)
        check(self.lib.ncs_hip_reduce_frames(self.handle, frames, n, ctypes.byref(params),
                                             _ptr(iters), _ptr(status), _ptr(cost)),
              "ncs_hip_reduce_frames")
        return iters, status, cost

    def solve_tiles(self, data, gamma, alpha, params):
        """Independent tiles [n, r, r] (host; r the mask's side, 16 unless an
        r x r mask is set); returns (u, iters, status, cost)."""
        d = numpy.ascontiguousarray(data, dtype=numpy.float32)
        g = numpy.ascontiguousarray(gamma, dtype=numpy.float32)
        r = getattr(self, "window", 16)
        if d.ndim != 3 or d.shape[1:] != (r, r) or g.shape != d.shape:
            raise ValueError("tiles must be [n, %d, %d] with matching gamma" % (r, r))
        n = d.shape[0]
        u = numpy.zeros_like(d)
        iters = numpy.zeros(n, numpy.int32)
        status = numpy.zeros(n, numpy.int32)
        cost = numpy.zeros(n, numpy.float32)
        check(self.lib.ncs_hip_solve_tiles(self.handle, _ptr(d), _ptr(g), _ptr(u), _ptr(iters),
                                           _ptr(status), _ptr(cost), n, float(alpha),
                                           ctypes.byref(params)),
              "ncs_hip_solve_tiles")
        return u, iters, status, cost

    def eval_tiles(self, data, gamma, u, alpha):
        """FP32 cost and gradient of tiles [n, 16, 16] (host)."""
        d = numpy.ascontiguousarray(data, dtype=numpy.float32)
        g = numpy.ascontiguousarray(gamma, dtype=numpy.float32)
        uu = numpy.ascontiguousarray(u, dtype=numpy.float32)
        n = d.shape[0]
        cost = numpy.zeros(n, numpy.float32)
        grad = numpy.zeros_like(uu)
        check(self.lib.ncs_hip_eval_tiles(self.handle, _ptr(d), _ptr(g), _ptr(uu), _ptr(cost),
                                          _ptr(grad), n, float(alpha)),
              "ncs_hip_eval_tiles")
        return cost, grad

    def solve_tiles_device(self, data_ptr, gamma_ptr, u0_ptr, out_ptr, iters_ptr,
                           status_ptr, cost_ptr, n, alpha, params, stream_ptr=None):
        """Device-pointer tiles (integers from torch's data_ptr()); async."""
        check(self.lib.ncs_hip_solve_tiles_device(
            self.handle, data_ptr, gamma_ptr, u0_ptr, out_ptr, iters_ptr, status_ptr,
            cost_ptr, int(n), float(alpha), ctypes.byref(params), stream_ptr),
            "ncs_hip_solve_tiles_device")

    def predict_frames_device(self, frames, u0_ptr, stream_ptr=None):
        """Learned starts of every window of device frames into u0 (device)."""
        check(self.lib.ncs_hip_net_predict_frames_device(self.handle, frame_array(frames),
                                                         len(frames), u0_ptr, stream_ptr),
              "ncs_hip_net_predict_frames_device")

    def autotune(self, frames, params=None, tune_history=False):
        """Time candidate schedules on device frames (dicts as for
        reduce_frames_device) and cache the fastest; returns the result dict."""
        res = TuneResult()
        p = params if params is not None else default_params()
        check(self.lib.ncs_hip_autotune(self.handle, frame_array(frames), len(frames),
                                        ctypes.byref(p), int(bool(tune_history)),
                                        ctypes.byref(res)),
              "ncs_hip_autotune")
        return res.as_dict()

    def reduce_frames_device(self, frames, params, u0_ptr=None, iters_ptr=None,
                             status_ptr=None, cost_ptr=None, stream_ptr=None):
        """frames: list of dicts(data, gamma, out, height, width, alpha, geometry,
        [data_stride, gamma_stride, out_stride]) with device pointers; async."""
        n = len(frames)
        arr = frame_array(frames)
        check(self.lib.ncs_hip_reduce_frames_device(self.handle, arr, n, ctypes.byref(params),
                                                    u0_ptr, iters_ptr, status_ptr, cost_ptr,
                                                    stream_ptr),
              "ncs_hip_reduce_frames_device")


def frame_array(frames):
    """ctypes array of ncs_frame from dicts of device pointers and sizes."""
    n = len(frames)
    arr = (Frame * n)()
    for i, f in enumerate(frames):
        a = arr[i]
        a.data, a.gamma, a.out = f["data"], f["gamma"], f.get("out")
        a.height, a.width = f["height"], f["width"]
        a.data_stride = f.get("data_stride", 0)
        a.gamma_stride = f.get("gamma_stride", 0)
        a.out_stride = f.get("out_stride", 0)
        a.alpha = float(f["alpha"])
        a.geometry = int(f.get("geometry", GEOM_REFERENCE))
    return arr


_contexts = {}


def context(device=0):
    """Process-wide context of a device (created on first use)."""
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _contexts[device] = ctx
    return ctx


class Pipeline(object):
    """Upload -> captured solve graph -> download for a fixed batch shape
    (ncs_hip_pipeline_*).  shapes: list of (height, width); alphas: one per
    frame; inputs and outputs are host numpy arrays (pinned memory makes the
    copies asynchronous).  Replays run in launch order, numbered from 0;
    up to PIPELINE_REPLAYS may be in flight, and the arrays of each are kept
    alive here until it is complete."""

    def __init__(self, ctx, shapes, alphas, params=None, geometry=GEOM_REFERENCE,
                 input_format=INPUT_PHOTONS_F32, keep_tile_results=False):
        self.ctx = ctx
        self.lib = ctx.lib
        n = len(shapes)
        fr = (PipelineFrame * n)()
        for i, (h, w) in enumerate(shapes):
            fr[i].height, fr[i].width = int(h), int(w)
            fr[i].alpha = float(alphas[i])
            fr[i].geometry = int(geometry[i] if hasattr(geometry, "__len__") else geometry)
        p = params if params is not None else default_params()
        h = ctypes.c_void_p()
        check(self.lib.ncs_hip_pipeline_create(ctx.handle, fr, n, ctypes.byref(p),
                                               int(input_format), int(bool(keep_tile_results)),
                                               ctypes.byref(h)),
              "ncs_hip_pipeline_create")
        self.handle = h
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.input_dtype = numpy.uint16 if input_format == INPUT_ADU_U16 else numpy.float32
        self.n_tiles = int(self.lib.ncs_hip_pipeline_tiles(h))
        self._keep = [None] * PIPELINE_REPLAYS
        self._cur = [None, None]

    def close(self):
        if self.handle is not None:
            self.lib.ncs_hip_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_frame(self, i, a, dtype):
        if a.shape != self.shapes[i] or a.dtype != dtype or not a.flags.c_contiguous:
            raise ValueError("frame %d: expected C-contiguous %s of shape %s"
                             % (i, numpy.dtype(dtype).name, self.shapes[i]))

    def set_gamma(self, i, gamma):
        g = numpy.ascontiguousarray(gamma, dtype=numpy.float32)
        self._check_frame(i, g, numpy.float32)
        check(self.lib.ncs_hip_pipeline_set_gamma(self.handle, int(i), _ptr(g)),
              "ncs_hip_pipeline_set_gamma")

    def set_camera(self, i, offset, gain, variance):
        maps = [numpy.ascontiguousarray(m, dtype=numpy.float32) for m in (offset, gain, variance)]
        for m in maps:
            self._check_frame(i, m, numpy.float32)
        check(self.lib.ncs_hip_pipeline_set_camera(self.handle, int(i), *[_ptr(m) for m in maps]),
              "ncs_hip_pipeline_set_camera")

    def set_camera_levels(self, i, points, offset, gains, variances):
        """Camera maps of frame i for several gain amplifiers: gains and
        variances [levels][H][W], `points` the levels - 1 ascending ADU
        division points (matlab/genmaps.m:29-62)."""
        off = numpy.ascontiguousarray(offset, dtype=numpy.float32)
        self._check_frame(i, off, numpy.float32)
        g = numpy.ascontiguousarray(gains, dtype=numpy.float32)
        v = numpy.ascontiguousarray(variances, dtype=numpy.float32)
        levels = g.shape[0] if g.ndim == 3 else 0
        if g.shape != (levels,) + self.shapes[i] or v.shape != g.shape:
            raise ValueError("gains and variances must be [levels][%d][%d]" % self.shapes[i])
        pts = numpy.ascontiguousarray(points, dtype=numpy.float64).reshape(-1)
        if pts.size != levels - 1:
            raise ValueError("%d levels take %d division points" % (levels, levels - 1))
        check(self.lib.ncs_hip_pipeline_set_camera_levels(
            self.handle, int(i), int(levels), pts.ctypes.data if pts.size else None, _ptr(off),
            _ptr(g), _ptr(v)), "ncs_hip_pipeline_set_camera_levels")

    def launch(self, inputs=None, outputs=None, stream_ptr=None):
        """Enqueue one replay and return its number.  None keeps the previous
        replay's arrays.  The host waits only for the oldest replay when
        PIPELINE_REPLAYS are in flight."""
        n = len(self.shapes)
        ins = outs = None
        if inputs is not None:
            for i, a in enumerate(inputs):
                self._check_frame(i, a, self.input_dtype)
            ins = (ctypes.c_void_p * n)(*[a.ctypes.data for a in inputs])
        if outputs is not None:
            for i, a in enumerate(outputs):
                self._check_frame(i, a, numpy.float32)
            outs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in outputs])
        seq = int(self.lib.ncs_hip_pipeline_launched(self.handle))
        rc = self.lib.ncs_hip_pipeline_launch(self.handle, ins, outs, stream_ptr)
        if int(self.lib.ncs_hip_pipeline_launched(self.handle)) > seq:
            # enqueued (even if part of it failed): its arrays stay alive
            # until the launch that reuses its slot has waited for it
            if inputs is not None:
                self._cur[0] = list(inputs)
            if outputs is not None:
                self._cur[1] = list(outputs)
            self._keep[seq % PIPELINE_REPLAYS] = tuple(self._cur)
        check(rc, "ncs_hip_pipeline_launch")
        return seq

    def launched(self):
        return int(self.lib.ncs_hip_pipeline_launched(self.handle))

    def wait(self, replay=None):
        """Wait for replay `replay` and every earlier one (None: all)."""
        if replay is None:
            check(self.lib.ncs_hip_pipeline_wait(self.handle), "ncs_hip_pipeline_wait")
        else:
            check(self.lib.ncs_hip_pipeline_wait_replay(self.handle, int(replay)),
                  "ncs_hip_pipeline_wait_replay")

    def tile_results(self, replay=None):
        """(iterations, status, cost) per tile of replay `replay` (one of the
        last PIPELINE_REPLAYS launched; None: the last); waits for it."""
        it = numpy.zeros(self.n_tiles, numpy.int32)
        st = numpy.zeros(self.n_tiles, numpy.int32)
        co = numpy.zeros(self.n_tiles, numpy.float32)
        if replay is None:
            check(self.lib.ncs_hip_pipeline_tile_results(self.handle, _ptr(it), _ptr(st),
                                                         _ptr(co)),
                  "ncs_hip_pipeline_tile_results")
        else:
            check(self.lib.ncs_hip_pipeline_replay_results(self.handle, int(replay), _ptr(it),
                                                           _ptr(st), _ptr(co)),
                  "ncs_hip_pipeline_replay_results")
        return it, st, co
