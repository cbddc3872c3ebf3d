"""Precision form: the row filter as a Toeplitz GEMM on MFMA (BF16, split
BF16, FP8 E4M3 with per-row scale), checked against the FP64 oracle.

Tolerances (relative L2 vs the exact FP64 padded-Toeplitz filter):
  FFT FP32 1e-5, split BF16 1e-4, BF16 1e-2, FP8 E4M3 8e-2.
The FFT and split forms are also held to the 1e-4 reconstruction bar vs FP64.
"""
import numpy as np
import pytest

import qtctreconstruction_amd as ct

pytestmark = pytest.mark.gpu

FILTER_TOL = {ct.FilterBackend.FFT_F32: 1e-5, ct.FilterBackend.GEMM_SPLIT_BF16: 1e-4,
              ct.FilterBackend.GEMM_BF16: 1e-2, ct.FilterBackend.GEMM_FP8: 8e-2}


def _fp64_filter(O, sinos, ftype):
    nz, na, bins = sinos.shape
    sq, pad, nfft = O.geometry(bins)
    h = O.kernel64(nfft, ftype)
    i = np.arange(sq)[:, None]
    j = np.arange(bins)[None, :] + pad
    T = h[(i - j) % nfft]
    return sinos.astype(np.float64).reshape(-1, bins) @ T.T


@pytest.fixture()
def ctx():
    c = ct.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("fb", list(ct.FilterBackend))
@pytest.mark.parametrize("ftype,shape", [(ct.FilterType.Ramp, (5, 40, 128)),
                                         (ct.FilterType.Hann, (4, 33, 100)),
                                         (ct.FilterType.SheppLogan, (2, 16, 512)),
                                         (ct.FilterType.Ramp, (3, 20, 300)),
                                         (ct.FilterType.Hann, (2, 12, 600)),
                                         (ct.FilterType.SheppLogan, (1, 7, 800)),
                                         (ct.FilterType.Hann, (1, 5, 1000)),
                                         (ct.FilterType.Ramp, (1, 3, 1086))])
def test_filter_variants_vs_fp64(ctx, O, fb, ftype, shape):
    """Detector widths whose padded outputs take 1 to 6 column blocks of 256
    bins (GEMM launch blocks of 1, 2 or 3 of them; at 1000 and 1086 bins the
    split form's window table exceeds its LDS share and the B fragments come
    from T in L2), angle counts that leave a partial group of 16, and partial
    z-blocks."""
    rng = np.random.default_rng(int(fb) * 10 + int(ftype))
    vol = O.phantom(int(ftype) + 3, 10, shape[2], shape[2], shape[0])
    sinos = np.stack([O.compute_sinogram(vol[z], shape[1], shape[2])[0] for z in range(shape[0])])
    sinos += rng.normal(0, 0.01, sinos.shape).astype(np.float32)
    ctx.set_filter_backend(fb)
    got = ctx.filter_sinograms(sinos, ftype)
    ref = _fp64_filter(O, sinos, int(ftype)).reshape(got.shape)
    err = O.rel_l2(got, ref)
    print(f"{fb.name} {ftype.name} {shape}: filtered relL2 vs FP64 {err:.2e}, "
          f"max-abs {np.abs(got - ref).max():.2e}")
    assert err <= FILTER_TOL[fb]


@pytest.mark.parametrize("fb", [ct.FilterBackend.GEMM_SPLIT_BF16, ct.FilterBackend.FFT_F32])
def test_reconstruction_within_1e4_of_fp64(O, fb):
    be = ct.HipBackend(0)
    be.ctx.set_filter_backend(fb)
    n, na = 96, 120
    vol = O.phantom(99, 12, n, n, 6)
    out = be.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Hann, na))
    ref64 = O.reconstruct_volume64(vol, na, O.HANN)
    err = O.rel_l2(out, ref64)
    print(f"{fb.name}: reconstruction relL2 vs FP64 {err:.2e}")
    assert err <= 1e-4


def test_gemm_variants_reconstruct_close_to_serial(O):
    be = ct.HipBackend(0)
    n, na = 64, 90
    vol = O.phantom(5, 10, n, n, 4)
    ref, _ = O.reconstruct_volume(vol, na, O.RAMP)
    for fb, tol in ((ct.FilterBackend.GEMM_SPLIT_BF16, 1e-4), (ct.FilterBackend.GEMM_BF16, 1e-2),
                    (ct.FilterBackend.GEMM_FP8, 1e-1)):
        be.ctx.set_filter_backend(fb)
        out = be.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Ramp, na))
        assert O.rel_l2(out, ref) <= tol, fb


@pytest.mark.parametrize("fb", [ct.FilterBackend.GEMM_SPLIT_BF16, ct.FilterBackend.GEMM_FP8])
def test_host_slabs_with_gemm_filter_bit_identical(O, fb):
    """The host-buffer path at 512 x 512 x 72 runs slabs of 32, 32 and 8
    slices; with a GEMM filter (the FP8 form's per-row scales sized for the
    largest slab before the first runs) the result is the device path's, on a
    context with the same backend, bit for bit."""
    n, d, na = 512, 72, 90
    vol = O.phantom(17, 10, n, n, d)
    params = ct.ReconstructionParams(ct.FilterType.Ramp, na)
    ctx = ct.Context(0)
    ctx.set_filter_backend(fb)
    d_vol = ctx.alloc(vol.nbytes).upload(vol)
    d_out = ctx.alloc(vol.nbytes)
    ctx.reconstruct_volume_dev(d_vol, n, n, d, na, int(params.filter), d_out)
    ref = d_out.download(vol.shape)
    be = ct.HipBackend(0)
    be.ctx.set_filter_backend(fb)
    out = be.reconstructVolume(vol, params)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    be.ctx.close()
    ctx.close()
