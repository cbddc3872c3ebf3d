// ct_internal.h — shared declarations of the MI355X FBP backend (not part of
// the public C ABI). Kernel launchers live in the .hip files; capi.cpp owns the
// context, workspaces and orchestration.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cthip {

// Slices per z-block. Radon and backprojection work on 4 slices at once: the
// parallel-beam geometry is identical for every slice, so index arithmetic is
// done once per (voxel|ray, angle) and reused for the 4 slices, and the
// normalised volume / filtered sinogram are stored z-innermost as float4.
constexpr int kZB = 4;

// Largest FFT length of the LDS row filter (bins <= 1448).
constexpr int kMaxFFT = 4096;

// ---- phantom -------------------------------------------------------------
// Per-ellipsoid parameters prepared on the host: centre, world->body map
// (row-major 3x3), inverse squared semi-axes, density, and a conservative
// axis-aligned bounding box used only to skip ellipsoids per tile.
struct PhEll {
    float cx, cy, cz;
    float m[9];
    float ia[3];
    float rho;
    float lo[3], hi[3];
};

hipError_t launch_phantom(const PhEll* d_ell, int n, int nx, int ny, int nz, float* d_vol,
                          hipStream_t s);

// ---- normalisation + Radon ----------------------------------------------
// keys: 2*nz uint32 ordered-float keys {min, max} per slice.
hipError_t launch_minmax_init(unsigned* d_keys, int nz, hipStream_t s);
hipError_t launch_minmax(const float* d_vol, size_t slice_elems, int nz, unsigned* d_keys,
                         hipStream_t s);
// writes mm[2*z] = min, mm[2*z+1] = max as floats
hipError_t launch_minmax_decode(const unsigned* d_keys, int nz, float* d_mm, hipStream_t s);
// w = (raw == 0) ? 0 : (raw - min) * (1/span); z-blocked row layout [zb][h][w]
// and column layout [zb][w][h] (float4 per pixel = 4 slices).
hipError_t launch_normalise(const float* d_vol, int w, int h, int nz, const float* d_mm,
                            float4* d_row, float4* d_col, hipStream_t s);
hipError_t launch_radon(const float4* d_row, const float4* d_col, int w, int h, int nz, int na,
                        int bins, const float* d_cos, const float* d_sin, float* d_sino,
                        hipStream_t s);

// ---- row filter ----------------------------------------------------------
hipError_t launch_filter_design(int n, int type, float* d_H, hipStream_t s);
hipError_t launch_twiddles(int n, float2* d_tw, hipStream_t s);
// sino [nz][na][bins] -> filt [nzb][na][sq] float4 (4 slices interleaved)
hipError_t launch_row_filter_fft(const float* d_sino, int nz, int na, int bins, int sq, int pad,
                                 int nfft, const float* d_H, const float2* d_tw, float4* d_filt,
                                 hipStream_t s);

// ---- backprojection --------------------------------------------------------
// out[z][y][x] = inside ? acc*scale : 0, then (if d_mm) * eff_span + min.
hipError_t launch_backproject(const float4* d_filt, int nz, int na, int sq, int N,
                              const float* d_cos, const float* d_sin, float scale,
                              const float* d_mm, float* d_out, hipStream_t s);

}  // namespace cthip
