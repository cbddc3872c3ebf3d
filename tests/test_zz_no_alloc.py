"""The allocation-free consumer path against the allocating one.

ct_hip_reconstruct_from_sinograms_dev_noalloc must give
ct_hip_reconstruct_from_sinograms_dev's output bit for bit (same kernels, same
z-blocks), on two calls in a row with the second smaller than the first (the
shared scratch keeps the first call's contents), without growing the
context's device memory, and refuse the shapes its scratch cannot hold.
"""
import ctypes

import numpy as np
import pytest

import qtctreconstruction_amd as ct

pytestmark = pytest.mark.gpu

_f32p = ctypes.POINTER(ctypes.c_float)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def _noalloc(ctx, d_s, nz, na, bins, ang, f, d_o):
    a = np.ascontiguousarray(ang, np.float32)
    return ctx.lib.ct_hip_reconstruct_from_sinograms_dev_noalloc(
        ctx.handle, d_s.ptr, nz, na, bins, a.ctypes.data_as(_f32p), int(f), d_o.ptr)


def test_noalloc_equals_allocating_entry(O):
    ctx = ct.Context(0)
    try:
        # 512^2 x 720: 7 z-blocks of filtered rows per scratch slab, so 96
        # slices take 4 slabs; then a smaller call with a ragged depth
        cases = [(512, 96, 720, ct.FilterType.SheppLogan, 31),
                 (256, 37, 360, ct.FilterType.Hann, 32)]
        for N, nz, na, f, seed in cases:
            print(f"[noalloc] N {N} nz {nz} na {na} filter {int(f)}", flush=True)
            d_v = ctx.phantom_dev(seed, 16, N, N, nz)
            d_s = ctx.alloc(4 * nz * na * N)
            ctx.radon_dev(d_v, N, N, nz, na, N, d_s)
            d_v.free()
            ang = O.serial_angles(na)
            d_ref = ctx.alloc(4 * nz * N * N)
            d_got = ctx.alloc(4 * nz * N * N)
            ctx.reconstruct_from_sinograms_dev(d_s, nz, na, N, ang, f, d_ref)
            used = ctx.bytes_in_use()
            rc = _noalloc(ctx, d_s, nz, na, N, ang, f, d_got)
            assert rc == ct.CT_OK, ctx.lib.ct_hip_last_error(ctx.handle)
            assert ctx.bytes_in_use() == used
            got = d_got.download((nz, N, N))
            assert _bits_equal(got, d_ref.download((nz, N, N)))
            z = nz // 2
            s = d_s.download((nz, na, N))[z]
            assert O.rel_l2(got[z], O.reconstruct_slice(s, ang, N, int(f))) <= 1e-5
            for b in (d_s, d_ref, d_got):
                b.free()
    finally:
        ctx.close()


def test_noalloc_refuses_what_its_scratch_cannot_hold():
    ctx = ct.Context(0)
    try:
        d_s = ctx.alloc(4 * 2048 * 1448)
        d_o = ctx.alloc(4 * 1448 * 1448)
        ang = np.zeros(2048, np.float32)
        # one z-block of filtered rows: 2048 x 2048 x 16 B = 64 MiB > the scratch's rows
        assert _noalloc(ctx, d_s, 1, 2048, 1448, ang, 0, d_o) == ct.CT_ERR_UNSUPPORTED
        assert b"MiB scratch" in ctx.lib.ct_hip_last_error(ctx.handle)
        ctx.set_filter_backend(ct.FilterBackend.GEMM_SPLIT_BF16)
        assert _noalloc(ctx, d_s, 1, 8, 64, ang[:8], 0, d_o) == ct.CT_ERR_UNSUPPORTED
        ctx.set_filter_backend(ct.FilterBackend.FFT_F32)
        big = np.zeros(8193, np.float32)
        assert _noalloc(ctx, d_s, 1, 8193, 16, big, 0, d_o) == ct.CT_ERR_UNSUPPORTED
    finally:
        ctx.close()
