// ct_internal.h — shared declarations of the MI355X FBP backend (not part of
// the public C ABI). Kernel launchers live in the .hip files; capi.cpp owns the
// context, workspaces and orchestration.
#pragma once

#include <algorithm>

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cthip {

// Slices per z-block. Radon and backprojection work on 4 slices at once: the
// parallel-beam geometry is identical for every slice, so index arithmetic is
// done once per (voxel|ray, angle) and reused for the 4 slices, and the
// normalised volume / filtered sinogram are stored z-innermost as float4.
constexpr int kZB = 4;
// Slices per Radon z-block (the normalised copies are [zb][..][..][8]).
constexpr int kRZ = 8;
// zero pixels after each row of the Radon stage's normalised copies (the row
// copy and the column copy, which is the transpose's row copy)
constexpr int kPadX = 4;
// bytes of one z-block of the row copy of a w x h image: (w + kPadX) x h
// pixels of 32 bytes (8 slices); radon_zbytes(h, w) for the column copy
__host__ __device__ inline size_t radon_zbytes(int w, int h) {
    return (size_t)(w + kPadX) * h * 32;
}

// XCD-grouped work order. Workgroups are dispatched to the kXcds XCDs
// round-robin (workgroup u to XCD u % kXcds), each XCD with its own L2. A
// launch of nx * slots * rounds workgroups on nx XCDs (kXcds, or a share of
// them: see the paired Radon walks) covers nunit units (z-blocks) of nslot
// workgroups each: in round q, units q*P .. q*P + P - 1 (P = nx / gsz) are
// each walked by one group of gsz XCDs, slots = ceil(nslot / gsz) workgroups
// per XCD, so a unit's data stays in its group's L2s. The mapping is a
// bijection onto the (unit, slot) pairs whatever the dispatch order; only the
// speed depends on it.
constexpr int kXcds = 8;
__device__ __forceinline__ bool xcd_map(int u, int gsz, int slots, int nslot, int nunit,
                                        int& unit, int& slot, int nx = kXcds) {
    const int xcd = u % nx, r = u / nx;
    const int round = r / slots, rs = r - round * slots;
    unit = round * (nx / gsz) + xcd / gsz;
    slot = rs * gsz + xcd % gsz;
    return unit < nunit && slot < nslot;
}
// XCDs per unit for nunit units of nslot workgroups on nx XCDs. Each round
// takes ceil(nslot / g) workgroups per XCD and there are ceil(nunit / (nx /
// g)) rounds; among the g within 2% of the fewest workgroups per XCD over all
// rounds (so few units still fill every XCD), the one nearest `target` per
// XCD and round (300 measured best for the Radon walks of 256- and 512-lane
// blocks, about 1.5 times what an XCD holds of them at once), the smaller g
// on a tie.
inline int xcd_group(int nslot, int nunit, double target = 300.0, int nx = kXcds) {
    auto load = [&](int g) {
        const long long P = nx / g;
        return (double)((nunit + P - 1) / P) * (double)((nslot + g - 1) / g);
    };
    double best = 1e300;
    for (int g = 1; g <= nx; g *= 2) best = std::min(best, load(g));
    int pick = nx;
    double dist = 1e300;
    for (int g = 1; g <= nx; g *= 2) {
        if (load(g) > best * 1.02) continue;
        const double per = (double)((nslot + g - 1) / g);
        const double d = per > target ? per / target : target / per;
        if (d < dist) {
            dist = d;
            pick = g;
        }
    }
    return pick;
}

// Largest FFT length of the LDS row filter (bins <= 1448).
constexpr int kMaxFFT = 4096;
// Largest image edge and detector count of the Radon walk: the candidate
// margin kDelta (radon.hip) covers the float error of Serial's s, below
// ~6e-4 of a bin for |s| < 4096, and the row stepping error, ~1e-3 px for
// rows of up to 2048 px, stays below the walk's 0.01 px margin. The walk's
// 32-bit byte offsets into one z-block stay below 2^32 well within these.
constexpr int kRadonMaxEdge = 2048;
constexpr int kRadonMaxBins = 4096;
static_assert((unsigned long long)(kRadonMaxEdge + kPadX) * kRadonMaxEdge * 32 < (1ull << 32),
              "the Radon walk's 32-bit offsets must cover a whole z-block");

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
hipError_t launch_noisy_log(float* d_p, size_t n, float I0, float kappa, unsigned long long seed,
                            hipStream_t s);

// ---- normalisation + Radon ----------------------------------------------
// keys: 2*nz uint32 ordered-float keys {min, max} per slice.
hipError_t launch_minmax_init(unsigned* d_keys, int nz, hipStream_t s);
hipError_t launch_minmax(const float* d_vol, size_t slice_elems, int nz, unsigned* d_keys,
                         hipStream_t s);
// writes mm[2*z] = min, mm[2*z+1] = max as floats
hipError_t launch_minmax_decode(const unsigned* d_keys, int nz, float* d_mm, hipStream_t s);
// w = (raw == 0) ? 0 : (raw - min) * (1/span); z-blocked row copy
// [zb][h][w + kPadX] and column copy (the transpose) [zb][w][h + kPadX],
// kRZ = 8 slices per pixel (2 float4), kPadX zero pixels past each row.
// d_rflag [zb][h][2] and d_cflag [zb][w][2] (zeroed beforehand) get each
// z-block row's {max(x + 1), max(w - x)} over its nonzero normalised pixels,
// and each column's {max(y + 1), max(h - y)}: {0, 0} for one of zeros.
hipError_t launch_normalise(const float* d_vol, int w, int h, int nz, const float* d_mm,
                            float4* d_row, float4* d_col, int* d_rflag, int* d_cflag,
                            hipStream_t s);
// Radon of the n angles alist[0..n) over an image copy [zb][h][w + kPadX]
// walked by rows, |sin| <= |cos| for every angle: the LDS-staged row walk.
// Steeper angles run it on the column copy with (w, h) and (d_cos, d_sin)
// exchanged. d_side holds radon_side_floats(w, bins, nz, na) floats (the
// sums that cross bin blocks; none when every angle fits one block).
// d_flag [zb][h][2]: the copy's row extents (launch_normalise); the 16-slice
// walk skips the rows that are zero in both of its z-blocks, and the waves
// whose candidates lie outside a row's extent.
size_t radon_side_floats(int w, int bins, int nz, int na);
hipError_t launch_radon_rows(const float4* d_img, const int* d_flag, int w, int h, int nz,
                             int na, int bins, const float* d_cos, const float* d_sin,
                             const int* d_alist, int n, float* d_sino, float* d_side,
                             hipStream_t s);
// Whether both walks of the Radon stage take the 16-slice walk (rows and
// columns of up to 516 px); they then run as one launch.
bool radon_walks_paired(int w, int h);
// Both walks in one launch: the n_row angles alist[0..n_row) over the row copy
// d_row, the others over the column copy d_col (radon_walks_paired must hold).
hipError_t launch_radon_walks(const float4* d_row, const float4* d_col, const int* d_rflag,
                              const int* d_cflag, int w, int h, int nz, int na, int bins,
                              const float* d_cos, const float* d_sin, const int* d_alist,
                              int n_row, float* d_sino, float* d_side, hipStream_t s);

// ---- row filter ----------------------------------------------------------
struct FiltArgs {
    const float* sino;  // [nz][na][bins]
    int nz, na, bins, sq, pad, nfft, logn;
    const float* H;     // LDS form: nfft/2+1 taps; wave form: the short-transform filter
    const float2* tw;   // LDS form: nfft-1 twiddles, stage order (W_m^j at m/2-1+j);
                        // wave form: the wave twiddle table
    float4* filt;       // [nzb][na][sq] float4
    int wave_r;         // 0: LDS FFT of nfft points; 1..3: wave form of 512 wave_r points
};
struct FiltJob {
    FiltArgs a;
    int item_begin;  // first work item of this job; items are (z-block, angle)
    int pad_;
};
FiltArgs make_filt_args(const float* d_sino, int nz, int na, int bins, int sq, int pad, int nfft,
                        const float* d_H, const float2* d_tw, float4* d_filt);
size_t fft_lds_bytes(int nfft);
// One persistent launch per transform class: every job in its list has the
// class cls (nfft of the LDS form, or -R of the wave form).
int filter_persistent_grid(int cls, int device);
// hdr = {first job index in d_jobs, njobs, total items}; counter must be 0
// at launch. Everything batch-dependent is read from device memory.
hipError_t launch_row_filter_persistent(const FiltJob* d_jobs, const int* d_hdr, int* d_counter,
                                        int cls, int grid, hipStream_t s);
// FilterDesign::createFilter(n, type): n/2 + 1 taps, the reference's FP32
// arithmetic bit for bit (host; n a power of two >= 4)
void serial_filter_taps(int n, int type, float* H);
hipError_t launch_twiddles(int n, float2* d_tw, hipStream_t s);
// Wave form of the row filter (filter_wave.hip): rows of up to kWaveM bins
// whose padded transform has nfft >= 512 run as R sub-transforms of kWaveM
// points (a short circular convolution of 512 R >= 2D + 1 points, D the
// largest kernel offset the kept outputs meet). wave_filter_r gives R (0 when
// the LDS form takes the row); the tables are made on the host in FP64 from
// the reference's taps: tw holds kWaveTabEntries float2, hs kWaveHsFloats.
constexpr int kWaveM = 512;
constexpr int kWaveTabEntries = 1600;
constexpr int kWaveHsFloats = 64 * 8 * 3;
int wave_filter_r(int bins, int sq, int pad, int nfft);
// D: the largest kernel offset |i - pad - e| the kept outputs meet. Any
// circular length L >= 2D + 1 gives Serial's outputs with the kernel
// g[d mod L] = h[d] (|d| <= D); short_filter_taps gives g's real spectrum
// Hs[0 .. L/2] (FP64 from the reference's taps, rounded once), the taps of
// the LDS FFT form at length L.
int filter_kernel_reach(int bins, int sq, int pad);
void short_filter_taps(int bins, int sq, int pad, int nfft, const float* taps, int L, float* H);
void wave_filter_tables(int bins, int sq, int pad, int nfft, const float* taps, float2* tw,
                        float* hs);
hipError_t launch_row_filter_wave(const FiltArgs& p, hipStream_t s);
int filter_wave_persistent_grid(int r, int device);
hipError_t launch_row_filter_wave_persistent(const FiltJob* d_jobs, const int* d_hdr,
                                             int* d_counter, int r, int grid, hipStream_t s);
// sino [nz][na][bins] -> filt [nzb][na][sq] float4 (4 slices interleaved)
hipError_t launch_row_filter_fft(const float* d_sino, int nz, int na, int bins, int sq, int pad,
                                 int nfft, const float* d_H, const float2* d_tw, float4* d_filt,
                                 hipStream_t s);

// Toeplitz-GEMM row filter (MFMA), taken by parts. T is Np x Kp (Np = sq
// rounded up to 256, Kp = bins+1 rounded up to 32); d_h holds 2*nfft doubles
// (kernel, then its cumulative sum); variant 1 BF16, 2 split BF16, 3 FP8 E4M3.
hipError_t launch_toeplitz_design(int nfft, int type, int sq, int bins, int pad, int Np, int Kp,
                                  double* d_H, double* d_h, unsigned short* d_Thi,
                                  unsigned short* d_Tlo, float* d_T32, unsigned char* d_T8,
                                  float* d_tmax, hipStream_t s);
hipError_t launch_row_absmax(const float* d_sino, size_t rows, int cols, float* d_rmax,
                             hipStream_t s);
hipError_t launch_toeplitz_gemm(int variant, const float* d_sino, int nz, int na, int bins,
                                int sq, int Np, int Kp, const unsigned short* d_Thi,
                                const unsigned short* d_Tlo, const unsigned char* d_T8,
                                const float* d_rmax, const float* d_tmax, float4* d_filt,
                                hipStream_t s);

// ---- backprojection --------------------------------------------------------
struct BpArgs {
    const float4* filt;  // [nzb][na][sq] float4 (4 slices)
    float* out;          // [nz][N][N]
    const float* cos_t;
    const float* sin_t;
    const float* mm;  // optional {min, max} per slice -> denormalise
    int nz, na, sq, N;
    float scale;
    int check;  // 1: test i0 >= 0 && i0+1 < sq per update
};
struct BpJob {
    BpArgs a;
    int item_begin;  // first work item; items are (z-block group, tile) in z-major order
    int pad_;
};
struct BpConfig {
    int tile;    // voxel tile edge (16 or 32)
    int angles;  // angles per LDS stage
    int zb2;     // z-blocks (of 4 slices) per tile: 1 or 2
};
constexpr int kNumBpConfigs = 8;
extern const BpConfig kBpConfigs[kNumBpConfigs];
// out[z][y][x] = inside ? acc*scale : 0, then (if d_mm) * eff_span + min.
hipError_t launch_backproject(const float4* d_filt, int nz, int na, int sq, int N,
                              const float* d_cos, const float* d_sin, float scale,
                              const float* d_mm, float* d_out, hipStream_t s, int cfg = 0);
hipError_t launch_backproject_cfg(int cfg, const BpArgs& p, hipStream_t s);
// Every voxel inside the circle |(xx,yy)| <= R has |u - sq/2| <= R(1+1e-6),
// so when [sq/2 - R - 1, sq/2 + R + 1] lies in [0, sq-2] no bounds test is
// needed (always true for the reference geometry output_size == bins).
inline bool bp_needs_check(int N, int sq) {
    const double R = (double)(N / 2), dcb = (double)(sq / 2);
    return !((dcb - R - 1.0 >= 0.0) && (dcb + R + 1.0 <= (double)sq - 2.0));
}

int bp_persistent_grid(int cfg, int device);
// d_counter: kBpCounters ints, zero before the launch
constexpr int kBpCounters = 8;
hipError_t launch_bp_persistent(int cfg, const BpJob* d_jobs, const int* d_hdr, int* d_counter,
                                int grid, hipStream_t s);
hipError_t launch_zero_ints(int* d_p, int n, hipStream_t s);

// ---- point cloud ------------------------------------------------------------
unsigned pc_blocks(size_t n);
hipError_t launch_pc_count(const float* d_v, size_t n, float thr, unsigned* d_counts,
                           hipStream_t s);
hipError_t launch_pc_write(const float* d_v, size_t n, int w, int h, float thr,
                           const unsigned* d_offs, const float* d_xc, const float* d_yc,
                           const float* d_zc, float4* d_out, hipStream_t s);

// ---- reconstruction metrics ---------------------------------------------------
// Sums and extrema of diff = recon - orig over a volume: s2 = sum diff^2,
// s1 = sum |diff| (double), mx = max |diff|, lo/hi = min/max of orig (float).
struct MetricsSums {
    double s2, s1;
    float mx, lo, hi, pad;
};
// Most partial blocks of the metrics pass (the partial buffer's length).
constexpr int kMetricsBlocks = 1024;
int metrics_blocks(size_t n);
// d_orig and d_recon 16-byte aligned; d_part holds metrics_blocks(n) entries.
hipError_t launch_metrics(const float* d_orig, const float* d_recon, size_t n,
                          MetricsSums* d_part, MetricsSums* d_out, hipStream_t s);

}  // namespace cthip
