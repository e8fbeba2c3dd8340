"""Python binding of the engine's C ABI (include/les_gpu.h) via ctypes.

The engine is ``liblesgpu.so`` next to this file (built by
``make -C combustion-les-optimization_amd/csrc``).  There is no fallback: if
the library or a GPU is missing every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblesgpu.so")

EXPORTED = [
    "les_params_default", "les_params_validate", "les_mixture_synthetic", "les_mixture_load",
    "les_mixture_save", "les_gpu_create", "les_gpu_destroy", "les_gpu_last_error",
    "les_gpu_init_state", "les_gpu_upload_state", "les_gpu_download_state", "les_gpu_download_field",
    "les_gpu_step", "les_gpu_rhs", "les_gpu_probe_fetch", "les_gpu_set_params", "les_gpu_get_step",
    "les_gpu_kernel_timing",
    "les_gpu_launch_config",
]

_lib = None


def lib():
    """Load the engine library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("engine library missing: %s (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        abi.declare_common(L)
        P = C.POINTER
        vp = C.c_void_p
        L.les_gpu_create.argtypes = [P(abi.Params), P(abi.Mixture), P(vp)]
        L.les_gpu_destroy.argtypes = [vp]
        L.les_gpu_destroy.restype = None
        L.les_gpu_last_error.argtypes = [vp]
        L.les_gpu_last_error.restype = C.c_char_p
        L.les_gpu_init_state.argtypes = [vp]
        L.les_gpu_upload_state.argtypes = [vp, P(P(C.c_double)), C.c_int]
        L.les_gpu_download_state.argtypes = [vp, P(P(C.c_double)), C.c_int]
        L.les_gpu_download_field.argtypes = [vp, C.c_char_p, P(C.c_double)]
        L.les_gpu_step.argtypes = [vp, C.c_int, P(abi.StepReport)]
        L.les_gpu_rhs.argtypes = [vp, C.c_int, P(P(C.c_double)), C.c_int]
        L.les_gpu_rhs_noalloc.argtypes = [vp, C.c_int, P(P(C.c_double)), C.c_int]
        L.les_gpu_use_resume_attributes.argtypes = [vp]
        L.les_aux_step.argtypes = [vp, C.c_size_t, P(C.c_int)]
        L.les_gpu_probe_fetch.argtypes = [vp, P(abi.ProbeRow), C.c_int, P(C.c_int)]
        L.les_gpu_set_params.argtypes = [vp, P(abi.Params)]
        L.les_gpu_get_step.argtypes = [vp]
        L.les_gpu_kernel_timing.argtypes = [vp, P(C.c_double), P(C.c_int), P(C.c_double)]
        L.les_gpu_set_timing.argtypes = [vp, C.c_int]
        L.les_gpu_launch_config.argtypes = [vp, P(C.c_int), P(C.c_int)]
        L.les_gpu_stats_count.argtypes = [vp]
        L.les_gpu_point_fetch.argtypes = [vp, P(C.c_double), C.c_int, P(C.c_int), P(C.c_int)]
        L.les_batch_config_default.argtypes = [P(abi.BatchConfig)]
        L.les_batch_pool_bytes.argtypes = [P(abi.BatchConfig), P(C.c_size_t)]
        L.les_batch_create.argtypes = [P(abi.BatchConfig), P(abi.Params), P(vp)]
        L.les_batch_destroy.argtypes = [vp]
        L.les_batch_destroy.restype = None
        L.les_batch_last_error.argtypes = [vp]
        L.les_batch_last_error.restype = C.c_char_p
        L.les_batch_tuning.argtypes = [vp, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_double)]
        L.les_batch_groups.argtypes = [vp, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]
        L.les_batch_admit.argtypes = [vp, P(abi.Params), C.c_int, C.c_int, P(C.c_int)]
        L.les_batch_release.argtypes = [vp, C.c_int]
        L.les_batch_step.argtypes = [vp, C.c_int, P(C.c_double)]
        L.les_batch_slot_info.argtypes = [vp, C.c_int, P(abi.SlotInfo)]
        L.les_batch_poll.argtypes = [vp, P(C.c_int), C.c_int, P(C.c_int)]
        L.les_batch_upload.argtypes = [vp, C.c_int, P(P(C.c_double)), C.c_int]
        L.les_batch_download.argtypes = [vp, C.c_int, P(P(C.c_double)), C.c_int]
        L.les_batch_download_field.argtypes = [vp, C.c_int, C.c_char_p, P(C.c_double)]
        L.les_batch_report.argtypes = [vp, C.c_int, P(abi.StepReport)]
        L.les_batch_point_fetch.argtypes = [vp, C.c_int, P(C.c_double), C.c_int, P(C.c_int)]
        L.les_batch_probe_fetch.argtypes = [vp, C.c_int, P(abi.ProbeRow), C.c_int, P(C.c_int)]
        L.les_batch_xface_copy.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           P(C.c_double)]
        _lib = L
    return _lib


def _ptrs(arrays):
    arr = (C.POINTER(C.c_double) * len(arrays))()
    for i, a in enumerate(arrays):
        if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
            raise TypeError("fields must be C-contiguous float64")
        arr[i] = a.ctypes.data_as(C.POINTER(C.c_double))
    return arr


def params(preset=abi.LES_PRESET_REFERENCE, **over):
    p = abi.Params()
    if lib().les_params_default(C.byref(p), preset) != 0:
        raise ValueError("unknown preset %r" % preset)
    for k, v in over.items():
        if k == "probe_x":
            p.num_probe_x = len(v)
            for i, x in enumerate(v):
                p.probe_x[i] = x
        else:
            setattr(p, k, v)
    return p


def mixture():
    m = abi.Mixture()
    if lib().les_mixture_synthetic(C.byref(m)) != 0:
        raise RuntimeError("les_mixture_synthetic failed")
    return m


class LesError(RuntimeError):
    pass


class LesGpu:
    """One engine context on the current HIP device."""

    def __init__(self, p=None, mix=None):
        self.p = p if p is not None else params()
        self.mix = mix if mix is not None else mixture()
        h = C.c_void_p()
        rc = lib().les_gpu_create(C.byref(self.p), C.byref(self.mix), C.byref(h))
        if rc != 0:
            raise LesError("les_gpu_create failed (%d)" % rc)
        self.h = h
        self.shape = (self.p.nx, self.p.ny, self.p.nz)
        self.n = self.p.nx * self.p.ny * self.p.nz

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().les_gpu_last_error(self.h)
            raise LesError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "h", None):
            lib().les_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_state(self):
        self._check(lib().les_gpu_init_state(self.h), "les_gpu_init_state")

    def upload_state(self, fields):
        arrs = [np.ascontiguousarray(f, dtype=np.float64).reshape(-1) for f in fields]
        self._check(lib().les_gpu_upload_state(self.h, _ptrs(arrs), len(arrs)), "les_gpu_upload_state")

    def download_state(self, n=abi.LES_NFIELDS):
        out = [np.empty(self.n) for _ in range(n)]
        self._check(lib().les_gpu_download_state(self.h, _ptrs(out), n), "les_gpu_download_state")
        return [o.reshape(self.shape) for o in out]

    def field(self, name, count=None):
        out = np.empty(count if count is not None else self.n)
        self._check(lib().les_gpu_download_field(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double))),
                    "les_gpu_download_field(%s)" % name)
        return out.reshape(self.shape) if count is None else out

    def step(self, n=1):
        rep = abi.StepReport()
        self._check(lib().les_gpu_step(self.h, n, C.byref(rep)), "les_gpu_step")
        return rep

    def rhs(self, t_loop=0):
        out = [np.zeros(self.n) for _ in range(10)]
        self._check(lib().les_gpu_rhs(self.h, t_loop, _ptrs(out), 10), "les_gpu_rhs")
        return [o.reshape(self.shape) for o in out]

    def rhs_noalloc(self, t_loop=0, fill=0.0):
        """les_gpu_rhs_noalloc: the same RHS staged in the library's static
        device scratch (no device allocation); ``fill`` pre-sets the host
        outputs so a test can see that every value was written."""
        out = [np.full(self.n, fill) for _ in range(10)]
        self._check(lib().les_gpu_rhs_noalloc(self.h, t_loop, _ptrs(out), 10), "les_gpu_rhs_noalloc")
        return [o.reshape(self.shape) for o in out]

    def probe_rows(self, max_rows=1 << 16):
        rows = (abi.ProbeRow * max_rows)()
        n = C.c_int(0)
        self._check(lib().les_gpu_probe_fetch(self.h, rows, max_rows, C.byref(n)), "les_gpu_probe_fetch")
        return [rows[i] for i in range(n.value)]

    def set_params(self, p):
        self._check(lib().les_gpu_set_params(self.h, C.byref(p)), "les_gpu_set_params")
        self.p = p

    @property
    def step_index(self):
        return lib().les_gpu_get_step(self.h)

    def set_timing(self, on=True):
        self._check(lib().les_gpu_set_timing(self.h, 1 if on else 0), "les_gpu_set_timing")

    def launch_config(self):
        """Duct mode: the run's tile width and XCD mapping (les_gpu_launch_config);
        None in mode R."""
        tk, xm = C.c_int(), C.c_int()
        if lib().les_gpu_launch_config(self.h, C.byref(tk), C.byref(xm)) != 0:
            return None
        return {"tile_k": tk.value, "xcd_map": xm.value}

    def kernel_timing(self):
        ms, n, st = C.c_double(), C.c_int(), C.c_double()
        self._check(lib().les_gpu_kernel_timing(self.h, C.byref(ms), C.byref(n), C.byref(st)), "les_gpu_kernel_timing")
        return ms.value, n.value, st.value

    def point_series(self, max_steps=4096):
        """Fixed-point series recorded since the last fetch: array [step, point, (u, v, w, G)]."""
        out = np.zeros(max_steps * abi.LES_NPOINT_MAX * 4)
        ns, npnt = C.c_int(), C.c_int()
        self._check(lib().les_gpu_point_fetch(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), max_steps,
                                              C.byref(ns), C.byref(npnt)), "les_gpu_point_fetch")
        return out[:ns.value * npnt.value * 4].reshape(ns.value, npnt.value, 4)


def batch_config(**over):
    cfg = abi.BatchConfig()
    lib().les_batch_config_default(C.byref(cfg))
    for k, v in over.items():
        setattr(cfg, k, v)
    return cfg


def pool_bytes(cfg):
    b = C.c_size_t()
    if lib().les_batch_pool_bytes(C.byref(cfg), C.byref(b)) != 0:
        raise ValueError("invalid batch configuration")
    return b.value


class Batch:
    """A pool of duct-mode runs advanced together (les_batch_* entry points).

    Every ``step`` launch advances all resident, unfinished runs; runs are
    admitted and released between steps (continuous batching).  A run's
    results do not depend on what else is resident."""

    def __init__(self, cfg=None, probe=None):
        self.cfg = cfg if cfg is not None else batch_config()
        h = C.c_void_p()
        rc = lib().les_batch_create(C.byref(self.cfg), C.byref(probe) if probe is not None else None, C.byref(h))
        if rc != 0:
            raise LesError("les_batch_create failed (%d)" % rc)
        self.h = h
        self.shapes = {}

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().les_batch_last_error(self.h)
            raise LesError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "h", None):
            lib().les_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def tuning(self):
        tk, gb, xm, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._check(lib().les_batch_tuning(self.h, C.byref(tk), C.byref(gb), C.byref(xm), C.byref(ms)),
                    "les_batch_tuning")
        return {"tile_k": tk.value, "grid_blocks": gb.value, "xcd_map": xm.value, "tuned_ms": ms.value}

    def groups(self):
        """(first tile, tile count) of each group of runs the stage launches
        advance in turn (one group: the whole tile range)."""
        cap = 8
        n, base, cnt = C.c_int(), (C.c_int * cap)(), (C.c_int * cap)()
        self._check(lib().les_batch_groups(self.h, cap, C.byref(n), base, cnt), "les_batch_groups")
        return [(base[g], cnt[g]) for g in range(min(n.value, cap))]

    def admit(self, p, n_steps=0, init=True):
        """Slot of the admitted run, or None when it does not fit (admission refused)."""
        s = C.c_int(-1)
        rc = lib().les_batch_admit(self.h, C.byref(p), n_steps, 1 if init else 0, C.byref(s))
        if rc == abi.LES_ERR_ADMISSION:
            return None
        self._check(rc, "les_batch_admit")
        self.shapes[s.value] = (p.nx, p.ny, p.nz)
        return s.value

    def release(self, slot):
        self._check(lib().les_batch_release(self.h, slot), "les_batch_release")
        self.shapes.pop(slot, None)

    def step(self, n=1):
        ms = C.c_double()
        self._check(lib().les_batch_step(self.h, n, C.byref(ms)), "les_batch_step")
        return ms.value

    def info(self, slot):
        i = abi.SlotInfo()
        self._check(lib().les_batch_slot_info(self.h, slot, C.byref(i)), "les_batch_slot_info")
        return i

    def poll(self):
        cap = self.cfg.max_slots
        arr = (C.c_int * cap)()
        n = C.c_int()
        self._check(lib().les_batch_poll(self.h, arr, cap, C.byref(n)), "les_batch_poll")
        return [arr[i] for i in range(min(n.value, cap))]

    def upload(self, slot, fields):
        arrs = [np.ascontiguousarray(f, dtype=np.float64).reshape(-1) for f in fields]
        self._check(lib().les_batch_upload(self.h, slot, _ptrs(arrs), len(arrs)), "les_batch_upload")

    def download(self, slot, n=5):
        shape = self.shapes[slot]
        cells = shape[0] * shape[1] * shape[2]
        out = [np.empty(cells) for _ in range(n)]
        self._check(lib().les_batch_download(self.h, slot, _ptrs(out), n), "les_batch_download")
        return [o.reshape(shape) for o in out]

    def field(self, slot, name, count):
        out = np.empty(count)
        self._check(lib().les_batch_download_field(self.h, slot, name.encode(),
                                                   out.ctypes.data_as(C.POINTER(C.c_double))),
                    "les_batch_download_field(%s)" % name)
        return out

    def report(self, slot):
        rep = abi.StepReport()
        self._check(lib().les_batch_report(self.h, slot, C.byref(rep)), "les_batch_report")
        return rep

    def xface_copy(self, src, src_plane, dst, dst_plane, planes, reps=1):
        """Copy x-planes of the state from one run to another (the local part of
        an x-slab halo exchange); returns the mean ms of one copy."""
        ms = C.c_double()
        self._check(lib().les_batch_xface_copy(self.h, src, src_plane, dst, dst_plane, planes, reps, C.byref(ms)),
                    "les_batch_xface_copy")
        return ms.value

    def probe_rows(self, slot, max_rows=1 << 16):
        """Flame-region probe rows of one run recorded since the last fetch."""
        rows = (abi.ProbeRow * max_rows)()
        n = C.c_int(0)
        self._check(lib().les_batch_probe_fetch(self.h, slot, rows, max_rows, C.byref(n)), "les_batch_probe_fetch")
        return [rows[i] for i in range(n.value)]

    def point_series(self, slot, max_steps=256):
        out = np.zeros(max_steps * abi.LES_NPOINT_MAX * 4)
        ns = C.c_int()
        self._check(lib().les_batch_point_fetch(self.h, slot, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                max_steps, C.byref(ns)), "les_batch_point_fetch")
        return out[:ns.value * abi.LES_NPOINT_MAX * 4].reshape(ns.value, abi.LES_NPOINT_MAX, 4)
