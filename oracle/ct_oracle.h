/*
 * ct_oracle.h — CPU oracle for the FBP hot path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product (libct_hip.so) never links, loads or calls it.
 *
 * Three families of functions:
 *   ct_o_*      FP32 restatement of the reference's SerialBackend path
 *               (QtCTReconstruction/SerialBackend.cpp, FFT.cpp, FilterDesign.cpp,
 *               Utils.cpp), same operation order, no FMA contraction.
 *   ct_o64_*    FP64 twin of the same method (the "FP64 oracle" every result
 *               reports its error against).
 *   ct_ph_*     the project's seeded ellipsoid phantom (no reference
 *               counterpart; the reference's generator is a different phantom).
 *   ct_omp_*    restatement of OpenMPBackend::reconstructVolume, used only as
 *               the timed CPU baseline (ct_ompbase.c, built separately).
 *
 * How each function is pinned (tests/test_oracle.py, tests/test_metrics_io.py):
 *   ct_o_forward_real / ct_o_inverse_real  bit-exact vs the reference's own
 *       FFT.cpp built under oracle/_ref; numpy FP64 rfft; the Nyquist closed form
 *   ct_o_create_filter     bit-exact vs the reference's FilterDesign.cpp build
 *       (42 vectors); ramp and window closed forms
 *   ct_o_compute_sinogram  0/90 degree closed forms (column and row sums with
 *       the last-bin rule); mass conservation; raw-zero skip; FP64 twin
 *   ct_o_reconstruct_slice exact padded-Toeplitz FP64 convolution (filter);
 *       disk reconstruction position and amplitude (y-flip, scale); linearity
 *   ct_o_reconstruct_volume composition of the above; denormalisation;
 *       the setSlice quirk
 *   ct_o_metrics           numpy FP64 means; constant-offset closed form;
 *       identical/constant volumes (PSNR +inf, max 1); NaN and n = 0 cases
 *   ct_ph_*                FP64 membership of one ellipsoid; volume closed form;
 *       additivity; determinism
 *   ct_o64_*               the closed forms above at tight tolerance
 *   ct_omp_*               the timed OpenMP baseline, pinned by its own
 *       conventions' closed forms: theta = 0 column sums with the (int)
 *       truncated bin, theta = 90 row reversal (-sin, no y-flip), negative-u
 *       truncation with mass kept, a disk reconstructing unflipped at its
 *       centre with amplitude ~1, linearity (tests/test_oracle.py)
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FilterType order of ReconstructionParams (QtCTReconstruction/CTTypes.h:151). */
enum { CT_O_RAMP = 0, CT_O_SHEPP_LOGAN = 1, CT_O_HAMMING = 2, CT_O_COSINE = 3,
       CT_O_HANN = 4, CT_O_BARTLETT = 5 };

/* ---- Utils.cpp restatements ---- */
float  ct_o_deg2rad(float deg);
size_t ct_o_next_pow2(size_t v);
void   ct_o_linspace(float start, float stop, size_t count, float* out);

/* ---- FFT.cpp restatement (radix-2, float) ---- */
/* in: n floats (n even); out: (n/2+1) complex as interleaved re,im */
int ct_o_forward_real(const float* in, size_t n, float* out_c);
/* in: (n2+1) complex interleaved; out: 2*n2 floats. Never reads in[n2]. */
int ct_o_inverse_real(const float* in_c, size_t n2, float* out);

/* ---- FilterDesign.cpp restatement: n/2+1 floats ---- */
int ct_o_create_filter(size_t n, int type, float* out);

/* ---- SerialBackend restatement ---- */
/* sino_out [na][bins]; minmax_out = {hu_min, hu_max} (as stored in Sinogram) */
int ct_o_compute_sinogram(const float* slice, size_t w, size_t h, size_t na,
                          size_t bins, float* sino_out, float* minmax_out);
/* square_bins / pad / fft size rules of reconstructSlice */
void ct_o_geometry(size_t bins, size_t* sq, size_t* pad_before, size_t* nfft);
/* filtered_out [na][sq] (optional, may be NULL); recon_out [N][N] (normalised,
 * i.e. scaled by pi/(2na) but not denormalised — Serial's reconstructSlice). */
int ct_o_reconstruct_slice(const float* sino, size_t na, size_t bins,
                           const float* angles_deg, size_t output_size, int filter,
                           float* filtered_out, float* recon_out);
/* Whole SerialBackend::reconstructVolume for a subset of slices.
 * vol (d,h,w); out (d,h,w) gets slices listed in z_list (others untouched).
 * sino_out optional: [len(z_list)][na][w]. Threads over slices (deterministic). */
int ct_o_reconstruct_volume(const float* vol, size_t w, size_t h, size_t d,
                            size_t na, int filter, const int* z_list, size_t nz_list,
                            float* out, float* sino_out, int threads);

/* ---- FP64 twin ---- */
int ct_o64_filter(size_t n, int type, double* out);   /* n/2+1, DC and Nyquist zeroed */
int ct_o64_kernel(size_t n, int type, double* h_out); /* h[k], k in [0,n): real-space kernel */
int ct_o64_compute_sinogram(const float* slice, size_t w, size_t h, size_t na,
                            size_t bins, double* sino_out);
int ct_o64_reconstruct_slice(const double* sino, size_t na, size_t bins,
                             const double* angles_deg, size_t output_size, int filter,
                             double* filtered_out, double* recon_out);
int ct_o64_reconstruct_volume(const float* vol, size_t w, size_t h, size_t d,
                              size_t na, int filter, const int* z_list, size_t nz_list,
                              double* out, int threads);

/* ---- CtReconstructionController.cpp:1199-1277 metrics ---- */
/* out_d = {mse, mae, psnr}; out_f = {max_abs (1 when < 1e-6), vmin, vmax} */
int ct_o_metrics(const float* orig, const float* recon, size_t n, double* out_d, float* out_f);

/* ---- project phantom ---- */
typedef struct {
    float cx, cy, cz;      /* centre in [-1,1]^3 coordinates */
    float ax, ay, az;      /* semi-axes */
    float phi, theta, psi; /* ZYZ Euler angles */
    float rho;             /* additive density */
} ct_o_ellipsoid;

int ct_ph_params(uint64_t seed, int n, ct_o_ellipsoid* out);
/* vol (nz,ny,nx) float; voxel centres linspace(-1,1,n) per axis */
int ct_ph_rasterize(const ct_o_ellipsoid* e, int n, int nx, int ny, int nz,
                    float* vol, int threads);

#ifdef __cplusplus
}
#endif
#endif
