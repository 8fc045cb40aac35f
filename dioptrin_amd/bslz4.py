"""Bitshuffle + LZ4 chunks (raw HDF5 filter 32008 chunks, one per frame).

compress() writes one frame as a chunk; scan() gives the block descriptors
of a list of chunks (the reference's pre-scan, src/integration/cuda/
device.rs:2029-2180); DeviceChunks keeps chunks and their descriptors in
device memory and decodes them into a DeviceArray on the GPU. GpuPlan's
integrate_bslz4 / batch1d_bslz4 / batch2d_bslz4(_stack) decode and integrate
in one call (Cuda1dPlan::integrate_bslz4, src/integration/cuda.rs:550-564).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N
from .device import DeviceArray

DTYPES = {"uint16": (np.uint16, N.DTYPES["uint16"]), "uint32": (np.uint32, N.DTYPES["uint32"]),
          "int32": (np.int32, N.DTYPES["int32"])}


def _dtype(dtype) -> tuple:
    key = np.dtype(dtype).name if not isinstance(dtype, str) else dtype
    if key not in DTYPES:
        raise ValueError("Bitshuffle/LZ4 frames must be uint16, uint32 or int32")
    return DTYPES[key]


def compress(frame: np.ndarray, block_bytes: int = 8192) -> bytes:
    """One frame (uint16, uint32 or int32; any shape) as a Bitshuffle + LZ4
    chunk with blocks of block_bytes (0 = 8192)."""
    a = np.ascontiguousarray(frame)
    _, code = _dtype(a.dtype)
    cap = N.lib().dio_bslz4_bound(a.size, code, block_bytes)
    out = np.empty(cap, np.uint8)
    n = C.c_uint64()
    N.check(N.lib().dio_bslz4_compress(a.ctypes.data, a.size, code, block_bytes, out.ctypes.data,
                                       cap, C.byref(n)))
    return out[:n.value].tobytes()


def _chunk_arrays(chunks: Sequence) -> tuple:
    bufs = [np.frombuffer(memoryview(c).cast("B"), np.uint8) for c in chunks]
    ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
    sizes = (C.c_uint64 * len(bufs))(*[b.size for b in bufs])
    return bufs, ptrs, sizes


def scan(chunks: Sequence, pixels: int, dtype="uint32"):
    """(descriptors (n_blocks, 4) uint32: src_off, comp_len (bit 31: verbatim
    tail), dst_off, uncomp_bytes; largest compressed block; largest block)."""
    _, code = _dtype(dtype)
    bufs, ptrs, sizes = _chunk_arrays(chunks)
    nb, mc, mb = C.c_uint64(), C.c_uint32(), C.c_uint32()
    N.check(N.lib().dio_bslz4_scan(ptrs, sizes, len(bufs), pixels, code, None, 0, C.byref(nb),
                                   C.byref(mc), C.byref(mb)))
    desc = np.empty((max(nb.value, 1), 4), np.uint32)
    N.check(N.lib().dio_bslz4_scan(ptrs, sizes, len(bufs), pixels, code, desc.ctypes.data,
                                   nb.value, C.byref(nb), C.byref(mc), C.byref(mb)))
    return desc[:nb.value], mc.value, mb.value


class DeviceChunks:
    """Chunks of n frames of one shape resident in device memory with their
    block descriptors; decode() writes the frames into a DeviceArray."""

    def __init__(self, chunks: Sequence, shape, dtype="uint32", device: int = 0):
        self.np_dtype, self.code = _dtype(dtype)
        self.shape = tuple(int(s) for s in shape)
        self.n_frames = len(chunks)
        pixels = self.shape[0] * self.shape[1]
        desc, self.max_comp, self.max_block = scan(chunks, pixels, dtype)
        flat = np.frombuffer(b"".join(bytes(c) for c in chunks), np.uint8)
        self.comp_bytes = flat.size
        self.comp = DeviceArray.from_numpy(flat if flat.size else np.zeros(1, np.uint8), device)
        self.blocks = DeviceArray.from_numpy(desc if desc.size else np.zeros((1, 4), np.uint32),
                                             device)
        self.n_blocks = desc.shape[0]
        self.err = DeviceArray.from_numpy(np.zeros(1, np.uint32), device)
        self.device = device

    def decode(self, out: DeviceArray, stream=None) -> None:
        """Enqueues the decode into out ((n_frames, dim1, dim2) of the chunk
        dtype). check() reads the error flag (synchronising)."""
        if out.shape != (self.n_frames,) + self.shape or out.dtype != np.dtype(self.np_dtype):
            raise ValueError(f"output must be {(self.n_frames,) + self.shape} "
                             f"{np.dtype(self.np_dtype)}, got {out.shape} {out.dtype}")
        from .integrator import _Context
        N.check(N.lib().dio_bslz4_decode_device(
            _Context.get(self.device), self.comp.ptr, self.comp_bytes, self.blocks.ptr,
            self.n_blocks, self.max_comp, self.max_block, self.code, out.ptr,
            out.shape[0] * out.shape[1] * out.shape[2], self.err.ptr, stream))

    def check(self) -> None:
        if self.err.numpy()[0] != 0:
            raise ValueError("corrupt or truncated Bitshuffle/LZ4 compressed block")

    def reset_error(self) -> None:
        self.err.copy_from(np.zeros(1, np.uint32))


def integrate(plan, chunks: Sequence, dtype="uint32", errors=False, raw=False):
    """(I, S) of n chunks through dio_integrate_bslz4 (S None without errors)."""
    _, code = _dtype(dtype)
    if isinstance(chunks, (bytes, bytearray, memoryview)):
        raise TypeError("chunks must be a list of byte buffers, one per frame")
    bufs, ptrs, sizes = _chunk_arrays(list(chunks))
    n = len(bufs)
    I = np.empty((n, plan.nbins))
    S = np.empty((n, plan.nbins)) if errors else None
    flags = (N.WITH_ERRORS if errors else 0) | (N.RAW_SUMS if raw else 0)
    if n:
        N.check(N.lib().dio_integrate_bslz4(plan._h, ptrs, sizes, n, code, flags, I.ctypes.data,
                                            None if S is None else S.ctypes.data))
    return I, S
