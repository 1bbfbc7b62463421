"""ctypes binding of libabhsf.so (the C ABI declared in include/abhsf_c.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import os

from . import hip  # loads libamdhip64 first (same instance the library links)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libabhsf.so")

OK, ERR_INVALID, ERR_HIP, ERR_ALLOC, ERR_STATE, ERR_UNSUPPORTED = range(6)
POLICY_PREPROC, POLICY_MSMF = 0, 1
F64, F32 = 0, 1
VARIANT_AUTO, VARIANT_STRIP, VARIANT_BALANCED, VARIANT_PANEL, VARIANT_TILED = 0, 1, 2, 3, 4
VARIANT_NAMES = {0: "auto", 1: "strip", 2: "balanced", 3: "panel", 4: "tiled"}
GENERAL, SYMMETRIC, SKEW_SYMMETRIC = 0, 1, 2
DELTA_VARINT, DELTA_U32_ROW, DELTA_U32_BLOCK = 0, 1, 2
DELTA_CODECS = {0: "varint", 1: "u32_row", 2: "u32_block"}
SCHEMES = ("coo", "csr", "bitmap", "dense")

# every symbol include/abhsf_c.h declares
EXPORTED = (
    "abhsf_create", "abhsf_build_csr", "abhsf_build_coo", "abhsf_get_stats", "abhsf_spmv",
    "abhsf_spmm", "abhsf_export", "abhsf_import", "abhsf_block_hist", "abhsf_mlp_num_params", "abhsf_mlp_forward", "abhsf_mlp_predict",
    "abhsf_mlp_backward", "abhsf_mlp_train", "abhsf_mlp_train_noalloc", "abhsf_batch_create", "abhsf_batch_set_vectors",
    "abhsf_batch_spmv", "abhsf_batch_synchronize", "abhsf_batch_profile", "abhsf_batch_destroy", "abhsf_synchronize",
    "abhsf_destroy", "abhsf_last_error", "abhsf_version", "abhsf_set_symmetry",
    "abhsf_delta_create", "abhsf_delta_create_codec", "abhsf_delta_build_csr", "abhsf_delta_spmv", "abhsf_delta_get_stats",
    "abhsf_delta_export", "abhsf_delta_synchronize", "abhsf_delta_destroy",
)


class Options(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32), ("policy", C.c_int32), ("precision", C.c_int32),
        ("variant", C.c_int32), ("is_binary", C.c_int32), ("reserved", C.c_int32),
        ("stream", C.c_void_p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64),
        ("s", C.c_int32), ("k", C.c_int32),
        ("M", C.c_int64), ("N", C.c_int64),
        ("blocks", C.c_uint64 * 4), ("values", C.c_uint64 * 4),
        ("smf_bits_preproc", C.c_uint64), ("smf_bits_msmf", C.c_uint64),
        ("stored_bytes", C.c_uint64), ("index_bytes", C.c_uint64),
        ("build_ms", C.c_double),
        ("variant", C.c_int32), ("tile_zs1", C.c_int32), ("work_items", C.c_int64),
        ("max_strip_work", C.c_int64),
    ]

    def as_dict(self):
        return {
            "m": self.m, "n": self.n, "nnz": self.nnz, "s": self.s, "k": self.k,
            "M": self.M, "N": self.N,
            "blocks": dict(zip(SCHEMES, [int(v) for v in self.blocks])),
            "values": dict(zip(SCHEMES, [int(v) for v in self.values])),
            "smf_bits_preproc": int(self.smf_bits_preproc), "smf_bits_msmf": int(self.smf_bits_msmf),
            "stored_bytes": int(self.stored_bytes), "index_bytes": int(self.index_bytes),
            "build_ms": float(self.build_ms),
            "variant": VARIANT_NAMES.get(int(self.variant), int(self.variant)),
            "tile_zs1": bool(self.tile_zs1), "work_items": int(self.work_items),
            "max_strip_work": int(self.max_strip_work),
        }


class AbhsfError(RuntimeError):
    """Raised for a non-zero abhsf_status (the reference throws
    std::runtime_error("Error running X function!"), spmv/unified/spmv.cpp:109-136)."""

    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"Error running {fn} function! (status {status}: {msg})")
        self.status = status
        self.detail = msg


class DeltaStats(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("key_bits", C.c_int32), ("chunk", C.c_int32),
        ("words", C.c_uint64 * 4), ("stored_bytes", C.c_uint64), ("build_ms", C.c_double),
        ("codec", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {"m": self.m, "n": self.n, "nnz": self.nnz, "key_bits": self.key_bits, "chunk": self.chunk,
                "words": {w: int(self.words[i]) for i, w in enumerate((1, 2, 4, 8))},
                "stored_bytes": int(self.stored_bytes), "build_ms": float(self.build_ms),
                "codec": DELTA_CODECS.get(int(self.codec), int(self.codec))}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C abhsf_amd/csrc)")
        hip._load()
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, d, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_size_t
        L.abhsf_create.argtypes = [C.POINTER(Options), C.POINTER(vp)]
        L.abhsf_build_csr.argtypes = [vp, i64, i64, i64, vp, vp, vp, i32]
        L.abhsf_build_coo.argtypes = [vp, i64, i64, i64, vp, vp, vp]
        L.abhsf_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.abhsf_spmv.argtypes = [vp, d, vp, d, vp]
        L.abhsf_spmm.argtypes = [vp, i32, d, vp, i64, d, vp, i64]
        L.abhsf_export.argtypes = [vp, C.c_char_p, vp, sz, C.POINTER(sz)]
        L.abhsf_import.argtypes = [vp, i64, i64, i64, i32, vp, vp, vp]
        L.abhsf_block_hist.argtypes = [vp, vp, i64, i64, i64, i32, i32, i32, i64, vp, i64, vp, vp, vp]
        L.abhsf_mlp_num_params.argtypes = [i32]
        L.abhsf_mlp_forward.argtypes = [vp, i32, vp, i64, vp, vp, vp, vp]
        L.abhsf_mlp_predict.argtypes = [vp, i32, i32, vp, i64, vp, vp]
        L.abhsf_mlp_backward.argtypes = [vp, i32, vp, i64, vp, vp, vp, vp, vp, C.c_float, vp, vp, vp, vp, vp]
        L.abhsf_mlp_train.argtypes = [vp, i32, vp, vp, vp, i64, i64, C.c_float, vp, vp]
        L.abhsf_mlp_train_noalloc.argtypes = [vp, i32, vp, vp, vp, i64, i64, C.c_float, vp, vp]
        L.abhsf_batch_create.argtypes = [vp, i32, vp, C.POINTER(vp)]
        L.abhsf_batch_set_vectors.argtypes = [vp, vp, vp]
        L.abhsf_batch_spmv.argtypes = [vp, d, d, i32]
        L.abhsf_batch_synchronize.argtypes = [vp]
        L.abhsf_batch_profile.argtypes = [vp, d, d, vp]
        L.abhsf_set_symmetry.argtypes = [vp, i32]
        L.abhsf_delta_create.argtypes = [vp, C.POINTER(vp)]
        L.abhsf_delta_create_codec.argtypes = [vp, i32, C.POINTER(vp)]
        L.abhsf_delta_build_csr.argtypes = [vp, i64, i64, i64, vp, vp, vp, i32]
        L.abhsf_delta_spmv.argtypes = [vp, d, vp, d, vp]
        L.abhsf_delta_get_stats.argtypes = [vp, C.POINTER(DeltaStats)]
        L.abhsf_delta_export.argtypes = [vp, C.c_char_p, vp, sz, C.POINTER(sz)]
        L.abhsf_delta_synchronize.argtypes = [vp]
        L.abhsf_delta_destroy.argtypes = [vp]
        L.abhsf_batch_destroy.argtypes = [vp]
        L.abhsf_synchronize.argtypes = [vp]
        L.abhsf_destroy.argtypes = [vp]
        L.abhsf_last_error.restype = C.c_char_p
        L.abhsf_last_error.argtypes = []
        L.abhsf_version.restype = C.c_char_p
        L.abhsf_version.argtypes = []
        for name in EXPORTED:
            if name not in ("abhsf_last_error", "abhsf_version"):
                getattr(L, name).restype = C.c_int32
        L.abhsf_mlp_num_params.restype = C.c_int64
        _lib = L
    return _lib


def call(fn: str, *args) -> None:
    L = lib()
    rc = getattr(L, fn)(*args)
    if rc != OK:
        raise AbhsfError(fn, rc, L.abhsf_last_error().decode())
