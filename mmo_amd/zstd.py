"""zstd frames through the system's libzstd (ctypes).

The reference keeps its per-type grid files compressed as ``<file>.ba1.zst``, written
and read with the zstd command-line tool (Utls.zstd_compress_file /
zstd_uncompress_file, src/utls.ml:22-45); this module reads and writes the same
standard zstd frames with libzstd's one-shot API.  A missing libzstd raises an error
naming it when compression is first used.
"""
import ctypes as ct
import ctypes.util

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        name = ctypes.util.find_library("zstd") or "libzstd.so.1"
        try:
            lib = ct.CDLL(name)
        except OSError as e:
            raise RuntimeError("libzstd is needed for .zst grid files: %s" % e)
        lib.ZSTD_compressBound.restype = ct.c_size_t
        lib.ZSTD_compressBound.argtypes = [ct.c_size_t]
        lib.ZSTD_compress.restype = ct.c_size_t
        lib.ZSTD_compress.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t, ct.c_int]
        lib.ZSTD_decompress.restype = ct.c_size_t
        lib.ZSTD_decompress.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t]
        lib.ZSTD_getFrameContentSize.restype = ct.c_ulonglong
        lib.ZSTD_getFrameContentSize.argtypes = [ct.c_void_p, ct.c_size_t]
        lib.ZSTD_isError.restype = ct.c_uint
        lib.ZSTD_isError.argtypes = [ct.c_size_t]
        lib.ZSTD_getErrorName.restype = ct.c_char_p
        lib.ZSTD_getErrorName.argtypes = [ct.c_size_t]
        _LIB = lib
    return _LIB


# ZSTD_CONTENTSIZE_UNKNOWN / _ERROR
_UNKNOWN, _ERROR = 2 ** 64 - 1, 2 ** 64 - 2
MAGIC = b"\x28\xb5\x2f\xfd"   # a zstd frame starts with 0xFD2FB528, little endian


def _check(lib, n, what):
    if lib.ZSTD_isError(n):
        raise ValueError("zstd %s: %s" % (what, lib.ZSTD_getErrorName(n).decode()))
    return n


def compress(data, level=3):
    """One zstd frame (content size recorded) holding `data` (bytes-like)."""
    lib = _lib()
    src = bytes(data)
    cap = lib.ZSTD_compressBound(len(src))
    dst = ct.create_string_buffer(cap)
    n = _check(lib, lib.ZSTD_compress(dst, cap, src, len(src), int(level)), "compress")
    return dst.raw[:n]


def decompress(data, size=None):
    """The content of a single zstd frame; `size` (bytes) is needed only when the frame
    does not record its content size."""
    lib = _lib()
    src = bytes(data)
    if src[:4] != MAGIC:
        raise ValueError("not a zstd frame")
    n = lib.ZSTD_getFrameContentSize(src, len(src))
    if n == _ERROR:
        raise ValueError("zstd: unreadable frame header")
    if n == _UNKNOWN:
        if size is None:
            raise ValueError("zstd frame without content size: pass the expected size")
        n = size
    elif size is not None and n != size:
        raise ValueError("zstd frame holds %d bytes, %d expected" % (n, size))
    dst = ct.create_string_buffer(max(int(n), 1))
    got = _check(lib, lib.ZSTD_decompress(dst, int(n), src, len(src)), "decompress")
    if got != n:
        raise ValueError("zstd: %d bytes decompressed, %d expected" % (got, n))
    return dst.raw[:got]
