"""ctypes binding of gpu_aln_pack.so, mirroring the reference's binding layer.

The struct mirrors and the pointer-array helper follow bin/sp_isac2_gpu.py:139-187
and bin/sp_isac_applications.py:96-157; the driver call sequences follow
bin/sp_isac2_gpu.py:959-1022 (multi-reference), bin/sp_isac_applications.py:
1445-1693 (pre-alignment) and :4576-4650 (reference-free).  Images are numpy
float32 arrays instead of EMAN2 EMData objects.

There is no fallback: if the library is missing or a call fails, an exception
is raised.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "gpu_aln_pack.so")
RING_LEN = 256

# Symbols declared in include/gpu_aln_pack.h (reference entry points first).
REFERENCE_SYMBOLS = (
    "print_gpu_info", "get_texture_limit", "gpu_clear", "reset_shifts",
    "multi_ref_align_init", "multi_ref_align", "pre_align_init", "pre_align_size_check",
    "pre_align_fetch", "pre_align_run", "ref_free_alignment_2D_init",
    "ref_free_alignment_2D_size_check", "ref_free_alignment_2D",
    "ref_free_alignment_2D_filter_references",
)
EXTRA_SYMBOLS = (
    "gpu_last_error", "gpu_last_error_string", "gpu_clear_error", "isac_mref_align_device",
    "isac_mref_align_mixed_device", "isac_mref_best_mixed_device", "isac_profile_enable",
    "isac_profile_read", "isac_mref_align_batched_device", "isac_batched_report",
    "isac_mref_proposal_device", "isac_assign_groups_device", "isac_assign_groups",
    "isac_update_references_device", "isac_mref_align_device_noalloc",
    "isac_multi_align_stability_device", "isac_pre_align_loop_device",
)

# Size of the static device array of isac_mref_align_device_noalloc (ISAC_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 64 << 20

BATCH_PER_SUBSET = 0   # ISAC_BATCH_PER_SUBSET
BATCH_GRAPH = 1        # ISAC_BATCH_GRAPH


class AlignConfig(ctypes.Structure):
    """Layout of cuda/gpu_aln_common.h:62-76."""
    _fields_ = [("sbj_num", ctypes.c_uint), ("ref_num", ctypes.c_uint),
                ("img_dim", ctypes.c_uint), ("ring_num", ctypes.c_uint),
                ("ring_len", ctypes.c_uint), ("shift_step", ctypes.c_float),
                ("shift_rng_x", ctypes.c_float), ("shift_rng_y", ctypes.c_float),
                ("aln_type", ctypes.c_char_p)]


class AlignParam(ctypes.Structure):
    """Layout of cuda/gpu_aln_common.h:78-85."""
    _fields_ = [("sbj_id", ctypes.c_int), ("ref_id", ctypes.c_int),
                ("shift_x", ctypes.c_float), ("shift_y", ctypes.c_float),
                ("angle", ctypes.c_float), ("mirror", ctypes.c_bool)]


class IsacMrefSubset(ctypes.Structure):
    """Layout of IsacMrefSubset in include/gpu_aln_pack.h (72 bytes)."""
    _fields_ = [("d_sbj", ctypes.c_void_p), ("d_ref", ctypes.c_void_p),
                ("d_acc_xy", ctypes.c_void_p), ("d_peak_list", ctypes.c_void_p),
                ("d_d", ctypes.c_void_p), ("n_sbj", ctypes.c_uint), ("n_ref", ctypes.c_uint),
                ("img_dim", ctypes.c_uint), ("ring_num", ctypes.c_uint),
                ("shift_step", ctypes.c_float), ("shift_rng_x", ctypes.c_float),
                ("shift_rng_y", ctypes.c_float), ("d_pitch", ctypes.c_uint)]


float_ptr = ctypes.POINTER(ctypes.c_float)
float_ptr_ptr = ctypes.POINTER(float_ptr)
aln_cfg_ptr = ctypes.POINTER(AlignConfig)
aln_param_ptr = ctypes.POINTER(AlignParam)

_lib = None


class GpuAlignError(RuntimeError):
    pass


def load(path: Optional[str] = None):
    """Load gpu_aln_pack.so and declare the argument types (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} not built; run `make -C {PKG_DIR}` (or __graft_entry__.build())")
    # PyTorch-ROCm wheels carry their own HIP runtime (soname libamdhip64.so.7, as
    # the library needs).  Loaded after torch, the library binds to that runtime;
    # loaded first, it binds /opt/rocm's and torch later loads a second runtime
    # into the process, and the library's first device call then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    c_uint, c_int, c_float, c_bool = ctypes.c_uint, ctypes.c_int, ctypes.c_float, ctypes.c_bool
    lib.print_gpu_info.argtypes = [c_uint]
    lib.print_gpu_info.restype = None
    lib.get_texture_limit.argtypes = [c_uint]
    lib.get_texture_limit.restype = c_int
    lib.gpu_clear.argtypes = []
    lib.gpu_clear.restype = None
    lib.reset_shifts.argtypes = [c_float, c_float]
    lib.reset_shifts.restype = None
    lib.multi_ref_align_init.argtypes = [aln_cfg_ptr, c_uint]
    lib.multi_ref_align_init.restype = aln_param_ptr
    lib.multi_ref_align.argtypes = [float_ptr_ptr, float_ptr_ptr, float_ptr, float_ptr, c_uint, c_float]
    lib.multi_ref_align.restype = None
    lib.pre_align_init.argtypes = [c_uint, aln_cfg_ptr, c_uint]
    lib.pre_align_init.restype = aln_param_ptr
    lib.pre_align_size_check.argtypes = [c_uint, aln_cfg_ptr, c_uint, c_float, c_bool]
    lib.pre_align_size_check.restype = c_int
    lib.pre_align_fetch.argtypes = [float_ptr_ptr, c_uint, ctypes.c_char_p]
    lib.pre_align_fetch.restype = None
    lib.pre_align_run.argtypes = [c_int, c_int]
    lib.pre_align_run.restype = None
    lib.ref_free_alignment_2D_init.argtypes = [aln_cfg_ptr, float_ptr_ptr, float_ptr_ptr,
                                               ctypes.POINTER(c_int), c_uint]
    lib.ref_free_alignment_2D_init.restype = aln_param_ptr
    lib.ref_free_alignment_2D_size_check.argtypes = [aln_cfg_ptr, c_uint, c_float, c_bool]
    lib.ref_free_alignment_2D_size_check.restype = c_int
    lib.ref_free_alignment_2D.argtypes = []
    lib.ref_free_alignment_2D.restype = None
    lib.ref_free_alignment_2D_filter_references.argtypes = [c_float, c_float]
    lib.ref_free_alignment_2D_filter_references.restype = None
    lib.gpu_last_error.argtypes = []
    lib.gpu_last_error.restype = c_int
    lib.gpu_last_error_string.argtypes = []
    lib.gpu_last_error_string.restype = ctypes.c_char_p
    lib.gpu_clear_error.argtypes = []
    lib.gpu_clear_error.restype = None
    vp = ctypes.c_void_p
    lib.isac_mref_align_device.argtypes = [vp, vp, vp, c_uint, c_uint, c_uint, c_uint, c_uint,
                                           c_float, c_float, c_float, vp, vp, c_uint, vp]
    lib.isac_mref_align_device.restype = c_int
    lib.isac_mref_align_device_noalloc.argtypes = lib.isac_mref_align_device.argtypes
    lib.isac_multi_align_stability_device.argtypes = [vp, c_uint, vp, c_uint, c_uint, c_float,
                                                      c_float, c_float, c_float, vp, vp, vp, vp,
                                                      vp, vp]
    lib.isac_multi_align_stability_device.restype = c_int
    lib.isac_pre_align_loop_device.argtypes = [vp, c_uint, c_uint, c_uint, vp, vp, c_uint, c_uint,
                                               c_int, c_float, c_float, vp, vp, vp]
    lib.isac_pre_align_loop_device.restype = c_int
    lib.isac_mref_align_device_noalloc.restype = c_int
    lib.isac_mref_align_mixed_device.argtypes = [vp, vp, vp, c_uint, c_uint, c_uint, c_uint, c_uint,
                                                 c_float, c_float, c_float, c_int, vp, vp, c_uint, vp]
    lib.isac_mref_align_mixed_device.restype = c_int
    lib.isac_mref_best_mixed_device.argtypes = [vp, vp, vp, c_uint, c_uint, c_uint, c_uint, c_uint,
                                                c_float, c_float, c_float, c_int, c_float,
                                                vp, vp, vp, vp, vp, vp, vp]
    lib.isac_mref_best_mixed_device.restype = c_int
    lib.isac_mref_proposal_device.argtypes = [vp, vp, vp, c_uint, c_uint, c_uint, c_uint, c_uint,
                                              c_float, c_float, c_float, vp, c_uint, c_uint,
                                              c_float, c_uint, c_float, c_int, c_int, vp, c_float,
                                              c_int, vp, vp, vp, vp, ctypes.POINTER(c_uint),
                                              ctypes.POINTER(c_uint), vp]
    lib.isac_mref_proposal_device.restype = c_int
    lib.isac_assign_groups_device.argtypes = [vp, c_uint, c_uint, c_uint, vp, vp]
    lib.isac_assign_groups_device.restype = c_int
    lib.isac_assign_groups.argtypes = [vp, c_uint, c_uint, vp]
    lib.isac_assign_groups.restype = c_int
    lib.isac_update_references_device.argtypes = [vp, c_uint, c_uint, vp, vp, c_uint, c_float,
                                                  c_float, vp, vp, vp, vp, vp]
    lib.isac_update_references_device.restype = c_int
    lib.isac_mref_align_batched_device.argtypes = [ctypes.POINTER(IsacMrefSubset), c_uint, c_int, vp]
    lib.isac_mref_align_batched_device.restype = c_int
    ip = ctypes.POINTER(c_int)
    lib.isac_batched_report.argtypes = [ip, ip, ip, ip, ip, ctypes.POINTER(c_float)]
    lib.isac_batched_report.restype = None
    lib.isac_profile_enable.argtypes = [c_int]
    lib.isac_profile_enable.restype = None
    lib.isac_profile_read.argtypes = [ctypes.POINTER(c_int), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(c_int), ctypes.POINTER(ctypes.c_double)]
    lib.isac_profile_read.restype = None
    if path is None:
        _lib = lib
    return lib


def check(lib=None):
    """Raise GpuAlignError if the library recorded an error (and clear it)."""
    lib = lib or load()
    code = lib.gpu_last_error()
    if code:
        msg = lib.gpu_last_error_string().decode(errors="replace")
        lib.gpu_clear_error()
        raise GpuAlignError(f"gpu_aln_pack error {code}: {msg}")


def get_c_ptr_array(images: Sequence[np.ndarray]):
    """C array of float* to C-contiguous float32 images (bin/sp_isac2_gpu.py:166-174)."""
    ptrs = []
    for img in images:
        assert img.flags["C_CONTIGUOUS"]
        assert img.dtype == np.float32
        ptrs.append(img.ctypes.data_as(float_ptr))
    return (float_ptr * len(ptrs))(*ptrs)


def make_config(sbj_num, ref_num, img_dim, ring_num, shift_step, shift_rng_x, shift_rng_y,
                aln_type: str, ring_len: int = RING_LEN) -> AlignConfig:
    cfg = AlignConfig(sbj_num, ref_num, img_dim, ring_num, ring_len, shift_step,
                      shift_rng_x, shift_rng_y, aln_type.encode("utf-8"))
    return cfg


# ------------------------------------------------------------ driver sequences

def multi_ref_align(sbj: np.ndarray, ref: np.ndarray, ring_num: int, shift_step: float,
                    shift_rng_x: float, shift_rng_y: float, acc_xy=None, device: int = 0,
                    d_pitch: Optional[int] = None, mem_request: float = 0.9):
    """Multi-reference call sequence of bin/sp_isac2_gpu.py:943-1006.

    Returns (peak_list [N, 4P] float32, d [N, d_pitch] float32) exactly as the
    library writes them (before the driver's post-rotation, see post_rotate).
    """
    lib = load()
    sbj = np.ascontiguousarray(sbj, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    pitch = P if d_pitch is None else d_pitch
    peak_list = np.zeros((N, 4 * P), dtype=np.float32)
    d = np.zeros((N, pitch), dtype=np.float32)
    cfg = make_config(P, N, D, ring_num, shift_step, shift_rng_x, shift_rng_y, "mref_aln")
    prm = lib.multi_ref_align_init(ctypes.byref(cfg), ctypes.c_uint(device))
    try:
        check(lib)
        if acc_xy is not None:
            acc = np.asarray(acc_xy, dtype=np.float64).reshape(P, 2)
            for i in range(P):
                prm[i].shift_x = float(acc[i, 0])
                prm[i].shift_y = float(acc[i, 1])
        else:
            for i in range(P):
                prm[i].shift_x = 0.0
                prm[i].shift_y = 0.0
        imgs_s = [sbj[i] for i in range(P)]
        imgs_r = [ref[i] for i in range(N)]
        lib.multi_ref_align(get_c_ptr_array(imgs_s), get_c_ptr_array(imgs_r),
                            peak_list.ctypes.data_as(float_ptr), d.ctypes.data_as(float_ptr),
                            ctypes.c_uint(pitch), ctypes.c_float(mem_request))
        check(lib)
    finally:
        lib.gpu_clear()
    return peak_list, d


def post_rotate(peak_list: np.ndarray) -> np.ndarray:
    """Driver conversion of shifts after the call (bin/sp_isac2_gpu.py:1016-1022)."""
    out = peak_list.astype(np.float32).copy()
    angle = out[:, 0::4]
    sx_neg = -out[:, 1::4]
    sy_neg = -out[:, 2::4]
    c_ang = np.cos(np.radians(angle))
    s_ang = -np.sin(np.radians(angle))
    out[:, 1::4] = sx_neg * c_ang - sy_neg * s_ang
    out[:, 2::4] = sx_neg * s_ang + sy_neg * c_ang
    return out


def params_to_numpy(prm, n: int) -> dict:
    return {"shift_x": np.array([prm[i].shift_x for i in range(n)], np.float32),
            "shift_y": np.array([prm[i].shift_y for i in range(n)], np.float32),
            "angle": np.array([prm[i].angle for i in range(n)], np.float32),
            "mirror": np.array([int(prm[i].mirror) for i in range(n)], np.int32),
            "ref_id": np.array([prm[i].ref_id for i in range(n)], np.int32)}


def pre_align(data: np.ndarray, ref: np.ndarray, ring_num: int, shift_rng: float,
              shift_step: float = 1.0, batch: Optional[int] = None, acc_xy=None,
              iterations: int = 1, device: int = 0):
    """Pre-alignment call sequence of bin/sp_isac_applications.py:1445-1693
    (one fixed reference; ``iterations`` repeated runs, each fetching the
    reference and every batch as the driver does).  Returns the AlignParam
    records as numpy arrays after the last run."""
    lib = load()
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, D = data.shape[0], data.shape[-1]
    cfg = make_config(n, 1, D, ring_num, shift_step, shift_rng, shift_rng, "pre_aln")
    if batch is None:
        limit = 0
        for split in [2 ** i for i in range(int(np.log2(n)) + 1)][::-1]:
            cfg.sbj_num = min(limit + split, n)
            if lib.pre_align_size_check(ctypes.c_uint(n), ctypes.byref(cfg), ctypes.c_uint(device),
                                        ctypes.c_float(0.9), ctypes.c_bool(False)) == 1:
                limit += split
        batch = max(limit, 1)
    cfg.sbj_num = batch
    prm = lib.pre_align_init(ctypes.c_uint(n), ctypes.byref(cfg), ctypes.c_uint(device))
    try:
        check(lib)
        if acc_xy is not None:
            acc = np.asarray(acc_xy, dtype=np.float64).reshape(n, 2)
            for i in range(n):
                prm[i].shift_x = float(acc[i, 0])
                prm[i].shift_y = float(acc[i, 1])
        n_batch = (n + batch - 1) // batch
        ref_img = np.ascontiguousarray(ref, dtype=np.float32).reshape(1, D, D)[0]
        if n_batch == 1:
            lib.pre_align_fetch(get_c_ptr_array([data[i] for i in range(n)]), ctypes.c_uint(n),
                                b"sbj_batch")
            check(lib)
        for _ in range(iterations):
            lib.pre_align_fetch(get_c_ptr_array([ref_img]), ctypes.c_uint(1), b"ref_batch")
            check(lib)
            for b in range(n_batch):
                s0, s1 = b * batch, min(n, (b + 1) * batch)
                if n_batch > 1:
                    lib.pre_align_fetch(get_c_ptr_array([data[i] for i in range(s0, s1)]),
                                        ctypes.c_uint(s1 - s0), b"sbj_batch")
                    check(lib)
                lib.pre_align_run(ctypes.c_int(s0), ctypes.c_int(s1))
                check(lib)
        return params_to_numpy(prm, n)
    finally:
        lib.gpu_clear()


def ref_free_align(sbj: np.ndarray, cid: np.ndarray, avg: np.ndarray, ring_num: int,
                   shift_rng: float, num_iter: int, cutoff_freq: float = 0.0,
                   falloff: float = 0.0, acc_xy=None, device: int = 0):
    """Reference-free call sequence of bin/sp_isac_applications.py:4576-4650.

    ``avg`` holds the initial class averages.  Returns the AlignParam records."""
    lib = load()
    sbj = np.ascontiguousarray(sbj, dtype=np.float32)
    P, D = sbj.shape[0], sbj.shape[-1]
    cid = np.asarray(cid, dtype=np.int32)
    avg = np.ascontiguousarray(avg, dtype=np.float32)
    cfg = make_config(P, int(cid[-1]) + 1, D, ring_num, 1.0, shift_rng, shift_rng, "multigroup_aln")
    cid_c = (ctypes.c_int * P)(*[int(c) for c in cid])
    prm = lib.ref_free_alignment_2D_init(ctypes.byref(cfg), get_c_ptr_array([sbj[i] for i in range(P)]),
                                         get_c_ptr_array([avg[k] for k in range(avg.shape[0])]),
                                         cid_c, ctypes.c_uint(device))
    try:
        check(lib)
        if acc_xy is not None:
            acc = np.asarray(acc_xy, dtype=np.float64).reshape(P, 2)
            for i in range(P):
                prm[i].shift_x = float(acc[i, 0])
                prm[i].shift_y = float(acc[i, 1])
        for i in range(num_iter):
            if cutoff_freq > 0.0:
                freq = 0.1 + (cutoff_freq - 0.1) * i / float(max(num_iter - 1, 1))
                lib.ref_free_alignment_2D_filter_references(ctypes.c_float(freq), ctypes.c_float(falloff))
                check(lib)
            lib.ref_free_alignment_2D()
            check(lib)
        return params_to_numpy(prm, P)
    finally:
        lib.gpu_clear()


def mref_align_device(sbj, ref, ring_num: int, shift_step: float, shift_rng_x: float,
                      shift_rng_y: float, acc_xy=None, stream=None, out=None):
    """Device-resident multi-reference alignment on torch tensors (cuda:0).

    sbj (P, D, D), ref (N, D, D) float32 device tensors; returns (peak_list
    [N, 4P], d [N, P]) device tensors.  ``stream`` is a torch.cuda.Stream."""
    import torch
    lib = load()
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    assert sbj.is_cuda and ref.is_cuda and sbj.dtype == torch.float32 and ref.dtype == torch.float32
    assert sbj.is_contiguous() and ref.is_contiguous()
    if out is None:
        peak = torch.empty((N, 4 * P), dtype=torch.float32, device=sbj.device)
        d = torch.empty((N, P), dtype=torch.float32, device=sbj.device)
    else:
        peak, d = out
    acc_ptr = None
    if acc_xy is not None:
        assert acc_xy.is_cuda and acc_xy.dtype == torch.float32 and acc_xy.is_contiguous()
        acc_ptr = acc_xy.data_ptr()
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_mref_align_device(sbj.data_ptr(), ref.data_ptr(), acc_ptr, P, N, D, ring_num,
                                    RING_LEN, shift_step, shift_rng_x, shift_rng_y,
                                    peak.data_ptr(), d.data_ptr(), P, st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_mref_align_device returned {rc}")
    return peak, d


def mref_align_device_noalloc(sbj, ref, ring_num: int, shift_step: float, shift_rng_x: float,
                              shift_rng_y: float, peak, d, acc_xy=None, stream=None):
    """mref_align_device without any device allocation inside the library: the
    caller passes the outputs peak [N, 4P] and d [N, P] (float32, cuda); the work
    runs in the library's static scratch array (NOALLOC_SCRATCH_BYTES per GPU).
    Calls that overlap on one GPU race on that array."""
    import torch
    lib = load()
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    assert sbj.is_cuda and ref.is_cuda and sbj.dtype == torch.float32 and ref.dtype == torch.float32
    assert sbj.is_contiguous() and ref.is_contiguous()
    assert tuple(peak.shape) == (N, 4 * P) and tuple(d.shape) == (N, P)
    assert peak.dtype == torch.float32 and d.dtype == torch.float32
    assert peak.is_contiguous() and d.is_contiguous()
    acc_ptr = None
    if acc_xy is not None:
        assert acc_xy.is_cuda and acc_xy.dtype == torch.float32 and acc_xy.is_contiguous()
        acc_ptr = acc_xy.data_ptr()
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_mref_align_device_noalloc(sbj.data_ptr(), ref.data_ptr(), acc_ptr, P, N, D,
                                            ring_num, RING_LEN, shift_step, shift_rng_x,
                                            shift_rng_y, peak.data_ptr(), d.data_ptr(), P, st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_mref_align_device_noalloc returned {rc}")
    return peak, d


def pre_align_loop_device(imgs, ring_num: int, shift_ranges, shift_steps, max_iter: int,
                          center: int = 0, cutoff: float = 0.0, falloff: float = 0.0,
                          acc_xy=None, stream=None):
    """The pre-alignment driver loop on the device (isac_pre_align_loop_device):
    imgs (P, D, D) float32 cuda tensor; one entry of shift_ranges / shift_steps per
    shift step, max_iter iterations each.  Returns (params dict with the
    AlignParam fields as numpy arrays, final average as a (D, D) cuda tensor)."""
    import torch
    lib = load()
    assert imgs.is_cuda and imgs.dtype == torch.float32 and imgs.is_contiguous()
    P, D = imgs.shape[0], imgs.shape[-1]
    recs = (AlignParam * P)()
    if acc_xy is not None:
        acc = np.asarray(acc_xy, np.float64).reshape(P, 2)
        for i in range(P):
            recs[i].shift_x, recs[i].shift_y = float(acc[i, 0]), float(acc[i, 1])
    raw = np.frombuffer(bytes(recs), dtype=np.uint8)
    d_prm = torch.from_numpy(raw.copy()).to(imgs.device)
    avg = torch.empty((D, D), dtype=torch.float32, device=imgs.device)
    rng_ = (ctypes.c_float * len(shift_ranges))(*[float(x) for x in shift_ranges])
    stp = (ctypes.c_float * len(shift_steps))(*[float(x) for x in shift_steps])
    assert len(shift_ranges) == len(shift_steps)
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_pre_align_loop_device(imgs.data_ptr(), P, D, ring_num, ctypes.cast(rng_, ctypes.c_void_p),
                                        ctypes.cast(stp, ctypes.c_void_p), len(shift_ranges),
                                        max_iter, center, cutoff, falloff, d_prm.data_ptr(),
                                        avg.data_ptr(), st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_pre_align_loop_device returned {rc}")
    out = (AlignParam * P).from_buffer_copy(d_prm.cpu().numpy().tobytes())
    prm = {"angle": np.array([r.angle for r in out]), "shift_x": np.array([r.shift_x for r in out]),
           "shift_y": np.array([r.shift_y for r in out]),
           "mirror": np.array([int(r.mirror) for r in out])}
    return prm, avg


def multi_align_stability_device(classes, mir_stab_thld: float, grp_err_thld: float,
                                 err_thld: float, d: float, stream=None):
    """Stability test of many classes in one device call.

    classes: one entry per class in the driver's layout (aln_param[k]: L flat
    lists [angle, sx, sy, mirror] * n_k, bin/sp_isac2_gpu.py:1381-1393); every
    class has the same L.  Returns per class (stable_set, mirror_consistent_rate,
    err, extra) like sp_pixel_error.multi_align_stability
    (bin/sp_isac2_gpu.py:1491): stable_set = [[pixel_error, member index,
    [angle, sx, sy, mirror]], ...] sorted by (pixel error, index); extra holds the
    per-member pixel errors and average alignments and the fitted transforms."""
    import torch
    lib = load()
    arrs = [np.asarray(c, np.float64).reshape(len(c), -1, 4) for c in classes]
    L = arrs[0].shape[0]
    assert all(a.shape[0] == L for a in arrs), "every class needs the same number of runs"
    counts = [a.shape[1] for a in arrs]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    prm = np.concatenate(arrs, axis=1) if arrs else np.zeros((L, 0, 4))
    n_total, n_class = int(off[-1]), len(arrs)
    dev = torch.device("cuda")
    d_prm = torch.from_numpy(np.ascontiguousarray(prm)).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    pix = torch.empty(max(n_total, 1), dtype=torch.float32, device=dev)
    avg = torch.empty((max(n_total, 1), 4), dtype=torch.float32, device=dev)
    stab = torch.empty(max(n_total, 1), dtype=torch.int32, device=dev)
    cls = torch.empty((max(n_class, 1), 2), dtype=torch.float32, device=dev)
    glob = torch.empty((max(n_class, 1), L, 4), dtype=torch.float32, device=dev)
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_multi_align_stability_device(d_prm.data_ptr(), n_total, d_off.data_ptr(), n_class,
                                               L, mir_stab_thld, grp_err_thld, err_thld, d,
                                               pix.data_ptr(), avg.data_ptr(), stab.data_ptr(),
                                               cls.data_ptr(), glob.data_ptr(), st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_multi_align_stability_device returned {rc}")
    pix, avg, stab = pix.cpu().numpy(), avg.cpu().numpy(), stab.cpu().numpy()
    cls, glob = cls.cpu().numpy(), glob.cpu().numpy()
    out = []
    for c in range(n_class):
        a, b = off[c], off[c + 1]
        st_set = [[float(pix[a + i]), int(i), [float(x) for x in avg[a + i, :3]] + [int(avg[a + i, 3])]]
                  for i in range(b - a) if stab[a + i]]
        st_set.sort(key=lambda e: (e[0], e[1]))
        extra = {"pix": pix[a:b].copy(), "avg": avg[a:b].copy(), "stable": stab[a:b].copy(),
                 "G": glob[c].copy(), "flips": glob[c, :, 3].astype(int)}
        out.append((st_set, float(cls[c, 0]), float(cls[c, 1]), extra))
    return out


def mref_align_mixed_device(sbj, ref, ring_num: int, shift_step: float, shift_rng_x: float,
                            shift_rng_y: float, precision: str = "fp16", acc_xy=None,
                            stream=None, out=None):
    """Mixed-precision (MFMA) multi-reference alignment on torch tensors.

    precision "fp16" or "bf16" selects the ring-contraction operand type; the
    psi transform runs on FP16 MFMA.  Returns approximate (peak_list, d)."""
    import torch
    lib = load()
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    assert sbj.is_cuda and ref.is_cuda and sbj.is_contiguous() and ref.is_contiguous()
    if out is None:
        peak = torch.empty((N, 4 * P), dtype=torch.float32, device=sbj.device)
        d = torch.empty((N, P), dtype=torch.float32, device=sbj.device)
    else:
        peak, d = out
    acc_ptr = acc_xy.data_ptr() if acc_xy is not None else None
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_mref_align_mixed_device(sbj.data_ptr(), ref.data_ptr(), acc_ptr, P, N, D, ring_num,
                                          RING_LEN, shift_step, shift_rng_x, shift_rng_y,
                                          1 if precision == "bf16" else 0, peak.data_ptr(),
                                          d.data_ptr(), P, st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_mref_align_mixed_device returned {rc}")
    return peak, d


def mref_best_mixed_device(sbj, ref, ring_num: int, shift_step: float, shift_rng_x: float,
                           shift_rng_y: float, precision: str = "fp16", acc_xy=None,
                           tie_rel: float = 1e-5, stream=None):
    """Mixed-precision search + exact best reference per particle (torch tensors).

    Returns a dict of device tensors: peak_list, d (approximate, every pair),
    best_ref, best_val, best_peak [P, 4], tie (per particle)."""
    import torch
    lib = load()
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    dev = sbj.device
    out = {"peak_list": torch.empty((N, 4 * P), dtype=torch.float32, device=dev),
           "d": torch.empty((N, P), dtype=torch.float32, device=dev),
           "best_ref": torch.empty(P, dtype=torch.int32, device=dev),
           "best_val": torch.empty(P, dtype=torch.float32, device=dev),
           "best_peak": torch.empty((P, 4), dtype=torch.float32, device=dev),
           "tie": torch.empty(P, dtype=torch.int32, device=dev)}
    acc_ptr = acc_xy.data_ptr() if acc_xy is not None else None
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_mref_best_mixed_device(
        sbj.data_ptr(), ref.data_ptr(), acc_ptr, P, N, D, ring_num, RING_LEN, shift_step,
        shift_rng_x, shift_rng_y, 1 if precision == "bf16" else 0, tie_rel,
        out["peak_list"].data_ptr(), out["d"].data_ptr(), out["best_ref"].data_ptr(),
        out["best_val"].data_ptr(), out["best_peak"].data_ptr(), out["tie"].data_ptr(), st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_mref_best_mixed_device returned {rc}")
    return out


def mref_proposal_device(sbj, ref, ring_num: int, shift_step: float, shift_rng_x: float,
                         shift_rng_y: float, prop, threshold: float, n_psi_bins: int = 72,
                         psi_half: float = 10.0, shift_half: int = 1, edge_fallback: bool = True,
                         level2: bool = False, conf=None, conf_min: float = 0.0, acc_xy=None,
                         stream=None, fallback_form: str = "fp32"):
    """Narrowed search around proposals prop [P, K, 4] int32 (ref, mirror, psi
    bin, shift index).  Particles below ``threshold`` (or, with
    ``edge_fallback``, whose maximum lies on the edge of the admitted region)
    get, with ``level2``, every shift and psi of their proposed references, and
    those still below the threshold the exhaustive search, as do the particles
    whose ``conf`` [P] (float32) is below ``conf_min``.  Returns a dict of
    device tensors best_ref, best_val, best_peak [P, 4], fallback, plus
    n_level2 and n_fallback (ints).  ``fallback_form`` "fp32" runs the exhaustive
    search in FP32; "fp16" / "bf16" run it on MFMA with each particle's 8 best
    references re-scored in FP32 (near-ties in FP64)."""
    import torch
    lib = load()
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    assert prop.dtype == torch.int32 and prop.is_contiguous() and prop.shape[0] == P
    form = {"fp32": 0, "fp16": 1, "bf16": 2}[fallback_form]
    dev = sbj.device
    out = {"best_ref": torch.empty(P, dtype=torch.int32, device=dev),
           "best_val": torch.empty(P, dtype=torch.float32, device=dev),
           "best_peak": torch.empty((P, 4), dtype=torch.float32, device=dev),
           "fallback": torch.empty(P, dtype=torch.int32, device=dev)}
    nl2, nfb = ctypes.c_uint(0), ctypes.c_uint(0)
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_mref_proposal_device(
        sbj.data_ptr(), ref.data_ptr(), acc_xy.data_ptr() if acc_xy is not None else None, P, N, D,
        ring_num, RING_LEN, shift_step, shift_rng_x, shift_rng_y, prop.data_ptr(), prop.shape[1],
        n_psi_bins, psi_half, shift_half, threshold, 1 if edge_fallback else 0,
        1 if level2 else 0, conf.data_ptr() if conf is not None else None, conf_min, form,
        out["best_ref"].data_ptr(), out["best_val"].data_ptr(),
        out["best_peak"].data_ptr(), out["fallback"].data_ptr(), ctypes.byref(nl2),
        ctypes.byref(nfb), st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_mref_proposal_device returned {rc}")
    out["n_level2"] = int(nl2.value)
    out["n_fallback"] = int(nfb.value)
    return out


def best_of_exhaustive(d, peak_list):
    """Best reference per particle of an exhaustive result (first maximum over
    references, like the fallback): (best_ref [P], best_val [P], best_peak [P, 4])."""
    import torch
    N, P = d.shape
    best = torch.argmax(d, dim=0)
    cols = torch.arange(P, device=d.device)
    return best.to(torch.int32), d[best, cols], peak_list.view(N, P, 4)[best, cols]


def assign_groups_device(d, n_img=None, stream=None):
    """Equal-size group assignment from d [N, pitch] (device float32): returns the
    n_img + 1 int32 layout (classes of n_img // N particles, leftovers, their
    class) as a device tensor."""
    import torch
    lib = load()
    N, pitch = d.shape
    P = pitch if n_img is None else n_img
    assert d.dtype == torch.float32 and d.is_cuda and d.is_contiguous()
    out = torch.empty(P + 1, dtype=torch.int32, device=d.device)
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_assign_groups_device(d.data_ptr(), N, P, pitch, out.data_ptr(), st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_assign_groups_device returned {rc}")
    return out


def assign_groups(d):
    """Host form (numpy float32 [N, P] -> numpy int32 [P + 1]), the call that
    replaces Util.assign_groups(address, numref, nima) in the driver."""
    lib = load()
    d = np.ascontiguousarray(d, dtype=np.float32)
    N, P = d.shape
    out = np.empty(P + 1, np.int32)
    rc = lib.isac_assign_groups(d.ctypes.data, N, P, out.ctypes.data)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_assign_groups returned {rc}")
    return out


def classes_from_assignment(groups, n_ref):
    """Class of each particle from assign_groups' n_img + 1 layout, as the driver
    builds ``belongsto`` (bin/sp_isac2_gpu.py:1122-1135): the first cap * n_ref
    entries fill the classes in blocks of cap = n_img // n_ref; only the first
    n_img % cap leftovers join the class in the last entry, and every particle
    the driver does not place keeps its default class 0 (``belongsto = [0]*nima``).
    That default applies whenever n_img - cap * n_ref >= cap, i.e. n_ref > cap."""
    g = np.asarray(groups)
    P = g.size - 1
    cap = P // n_ref
    if cap < 1:
        raise ValueError(f"{P} particles cannot fill {n_ref} classes (the driver divides by n_img // n_ref)")
    cls = np.zeros(P, np.int32)
    for i in range(cap * n_ref):
        cls[g[i]] = i // cap
    for i in range(P % cap):
        cls[g[cap * n_ref + i]] = g[-1]
    return cls


def update_references_device(imgs, peak_list, cls, n_ref, cutoff, falloff, stream=None):
    """Reference update after the assignment (isac_update_references_device):
    imgs [P, D, D], peak_list [N, 4P] (the library's layout, before the driver's
    shift rotation), cls [P] int32 class per particle.  Returns (refs [N, D, D],
    params [P, 4] = (angle, sx, sy, mirror) about the new centres, members [N],
    shift_avg [N, 2]) as device tensors."""
    import torch
    lib = load()
    P, D = imgs.shape[0], imgs.shape[-1]
    assert imgs.dtype == torch.float32 and imgs.is_cuda and imgs.is_contiguous()
    assert peak_list.shape == (n_ref, 4 * P) and peak_list.is_contiguous()
    cls = cls.to(device=imgs.device, dtype=torch.int32).contiguous()
    assert cls.numel() == P
    refs = torch.empty((n_ref, D, D), dtype=torch.float32, device=imgs.device)
    params = torch.empty((P, 4), dtype=torch.float32, device=imgs.device)
    members = torch.empty(n_ref, dtype=torch.int32, device=imgs.device)
    shift_avg = torch.empty((n_ref, 2), dtype=torch.float32, device=imgs.device)
    st = (stream or torch.cuda.current_stream()).cuda_stream
    rc = lib.isac_update_references_device(imgs.data_ptr(), P, D, peak_list.data_ptr(), cls.data_ptr(),
                                           n_ref, float(cutoff), float(falloff), refs.data_ptr(),
                                           params.data_ptr(), members.data_ptr(),
                                           shift_avg.data_ptr(), st)
    if rc:
        check(lib)
        raise GpuAlignError(f"isac_update_references_device returned {rc}")
    return refs, params, members, shift_avg


class BatchedMref:
    """Many multi-reference subsets through one call (torch tensors on cuda:0).

    ``subsets`` is a list of dicts with keys sbj (P, D, D), ref (N, D, D),
    ring_num, shift_rng (or shift_rng_x / shift_rng_y), shift_step (default 1)
    and optionally acc_xy (P, 2).  Output tensors (peak_list [N, 4P], d [N, P])
    are allocated once; the descriptor array is kept, so repeated ``run`` calls
    in graph mode replay the captured graph."""

    def __init__(self, subsets):
        import torch
        self.lib = load()
        self.subsets = subsets
        self.outputs = []
        arr = (IsacMrefSubset * len(subsets))()
        for i, s in enumerate(subsets):
            sbj, ref = s["sbj"], s["ref"]
            assert sbj.is_cuda and ref.is_cuda and sbj.is_contiguous() and ref.is_contiguous()
            assert sbj.dtype == torch.float32 and ref.dtype == torch.float32
            P, D, N = sbj.shape[0], sbj.shape[-1], ref.shape[0]
            peak = torch.empty((N, 4 * P), dtype=torch.float32, device=sbj.device)
            d = torch.empty((N, P), dtype=torch.float32, device=sbj.device)
            self.outputs.append((peak, d))
            acc = s.get("acc_xy")
            rx = s.get("shift_rng_x", s.get("shift_rng"))
            ry = s.get("shift_rng_y", s.get("shift_rng"))
            arr[i] = IsacMrefSubset(sbj.data_ptr(), ref.data_ptr(),
                                    acc.data_ptr() if acc is not None else None,
                                    peak.data_ptr(), d.data_ptr(), P, N, D, s["ring_num"],
                                    s.get("shift_step", 1.0), rx, ry, P)
        self.desc = arr

    def run(self, mode: int = BATCH_GRAPH, stream=None):
        import torch
        st = (stream or torch.cuda.current_stream()).cuda_stream
        rc = self.lib.isac_mref_align_batched_device(self.desc, len(self.subsets), mode, st)
        if rc:
            check(self.lib)
            raise GpuAlignError(f"isac_mref_align_batched_device returned {rc}")
        return self.outputs

    def report(self) -> dict:
        c = ctypes.c_int
        g, n, r, nw, b = c(), c(), c(), c(), c()
        ms = (ctypes.c_float * 3)()
        self.lib.isac_batched_report(ctypes.byref(g), ctypes.byref(n), ctypes.byref(r),
                                     ctypes.byref(nw), ctypes.byref(b), ms)
        return {"groups": g.value, "graph_nodes": n.value, "replayed": bool(r.value),
                "waves_per_workgroup": nw.value, "persistent_blocks": b.value,
                "tuning_ms": {"1": ms[0], "2": ms[1], "4": ms[2]}}
