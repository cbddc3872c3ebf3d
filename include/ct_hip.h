/*
 * ct_hip.h — C ABI of the MI355X FBP backend (libct_hip.so).
 *
 * Plain pointers and sizes only. Every entry point returns 0 on success or a
 * negative ct_status; no C++ exception crosses this boundary. Host buffers
 * are owned by the caller; device buffers made by ct_hip_alloc are owned by
 * the context and freed by ct_hip_free / ct_hip_destroy. A context is
 * internally synchronised (one recursive lock), so one context may be shared
 * by threads, as the reference's controller shares one backend between its
 * worker thread and an OpenMP loop (CtReconstructionController.cpp:1232-1241).
 *
 * Numerics follow the reference's SerialBackend (the method's oracle):
 * geometry, normalisation and index rules of SerialBackend.cpp:59-318.
 *
 * Reference interfaces replaced (paths relative to QtCTReconstruction/):
 *   IReconstructionBackend.h:10-55   the backend virtuals
 *   IVolumeGenerator.h:16-27         the phantom generator
 *   FilterDesign.h:9-18              filter design
 *   IFFTBackend.h:16-42              the row filter's FFT
 */
#ifndef CT_HIP_H
#define CT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ct_hip_ctx ct_hip_ctx;

typedef enum {
    CT_OK = 0,
    CT_ERR_INVALID = -1,     /* bad argument (sizes, null pointers, filter id) */
    CT_ERR_HIP = -2,         /* a HIP runtime call failed */
    CT_ERR_NOMEM = -3,       /* device allocation failed */
    CT_ERR_BUDGET = -4,      /* request exceeds the context's HBM budget */
    CT_ERR_UNSUPPORTED = -5, /* shape outside what the kernels support */
    CT_ERR_NODEVICE = -6,    /* no HIP device */
    CT_ERR_IO = -7           /* file missing, unreadable, unwritable or not in the format */
} ct_status;

/* Filter ids: the order of ReconstructionParams::FilterType (CTTypes.h:151).
 * Ids outside 0..5 select Shepp-Logan, as FilterDesign::createFilter's
 * default branch does (FilterDesign.cpp:25-27). */
enum { CT_FILTER_RAMP = 0, CT_FILTER_SHEPP_LOGAN = 1, CT_FILTER_HAMMING = 2,
       CT_FILTER_COSINE = 3, CT_FILTER_HANN = 4, CT_FILTER_BARTLETT = 5 };

/* ------------------------------------------------------------ context */

/* Replaces IReconstructionBackend::isAvailable (IReconstructionBackend.h:18):
 * number of visible HIP devices (0 when none). */
int ct_hip_device_count(void);

/* Creates a context on `device`. hbm_budget_bytes = 0 means "no limit
 * beyond the device's free memory". Device state (trig tables, filter banks,
 * workspaces) is created lazily and cached, like CUDABackend's mutable caches
 * (CUDABackend.h:46-127). */
int ct_hip_create(int device, size_t hbm_budget_bytes, ct_hip_ctx** out);
void ct_hip_destroy(ct_hip_ctx* ctx);
/* Message of the last failure on this context ("" if none). */
const char* ct_hip_last_error(const ct_hip_ctx* ctx);
/* Bytes of device memory the context holds now. */
size_t ct_hip_bytes_in_use(const ct_hip_ctx* ctx);
/* Frees cached workspaces (CUDABackend::clearWorkspace, CUDABackend.h:43). */
int ct_hip_clear_workspace(ct_hip_ctx* ctx);

/* ------------------------------------------------- device-resident buffers */
int ct_hip_alloc(ct_hip_ctx* ctx, size_t bytes, void** dptr);
int ct_hip_free(ct_hip_ctx* ctx, void* dptr);
int ct_hip_upload(ct_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int ct_hip_download(ct_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int ct_hip_synchronize(ct_hip_ctx* ctx);

/* --------------------------------------------------------------- phantom */

/* One ellipsoid of the project phantom: centre and semi-axes in the
 * [-1,1]^3 voxel-centre coordinates (linspace(-1,1,n) per axis, as
 * utils::linspace, Utils.cpp:12-26), ZYZ Euler angles, additive density. */
typedef struct {
    float cx, cy, cz;
    float ax, ay, az;
    float phi, theta, psi;
    float rho;
} ct_ellipsoid;

/* Seeded parameters: centres U[-0.6,0.6]^3, semi-axes U[0.05,0.4]^3, angles
 * U[0,pi)^3, density U[-0.3,1.0], from a counter-based generator so that
 * every (seed, index) pair is reproducible on any host. Host-only. */
int ct_phantom_params(uint64_t seed, int n, ct_ellipsoid* out);

/* Replaces IVolumeGenerator::generateBrainHU (IVolumeGenerator.h:26;
 * Generator3DGPU.cu:124-177): rasterises n ellipsoids into a (nz,ny,nx)
 * float volume. _dev writes device memory (d_out), the other host memory. */
int ct_hip_phantom(ct_hip_ctx* ctx, const ct_ellipsoid* e, int n, int nx, int ny, int nz,
                   float* host_out);
int ct_hip_phantom_dev(ct_hip_ctx* ctx, const ct_ellipsoid* e, int n, int nx, int ny, int nz,
                       float* d_out);

/* ------------------------------------------------------- backend virtuals */

/* Replaces IReconstructionBackend::computeSinogram (IReconstructionBackend.h:21;
 * SerialBackend.cpp:59-136). slice (h,w) row-major; sino_out [na][bins];
 * minmax_out = {original_min_hu, original_max_hu}. Angles are 180/na*a deg. */
int ct_hip_compute_sinogram(ct_hip_ctx* ctx, const float* slice, int w, int h, int num_angles,
                            int bins, float* sino_out, float* minmax_out);

/* Replaces IReconstructionBackend::reconstructSlice (IReconstructionBackend.h:24;
 * SerialBackend.cpp:142-255): pad to ceil(sqrt2*bins), filter, backproject
 * onto output_size^2, scale pi/(2na). Output is not denormalised, as Serial's. */
int ct_hip_reconstruct_slice(ct_hip_ctx* ctx, const float* sino, int num_angles, int bins,
                             const float* angles_deg, int output_size, int filter, float* out);

/* Replaces IReconstructionBackend::reconstructVolume (IReconstructionBackend.h:28-31;
 * SerialBackend.cpp:261-318): per z slice, sinogram with bins = w, FBP onto
 * w x w, then v*eff_span + min. vol and out are (d,h,w); when h != w the
 * output stays zero, as Buffer3D::setSlice drops mismatched slices
 * (CTTypes.h:112-116), and only the sinograms are computed (the w x w
 * reconstructions Serial computes and then drops are not; the timings then
 * hold the Radon stage and 0 for the rest). sino_out (optional) receives
 * [d][na][w]. */
int ct_hip_reconstruct_volume(ct_hip_ctx* ctx, const float* vol, int w, int h, int d,
                              int num_angles, int filter, float* out, float* sino_out);

/* Device-resident form of ct_hip_reconstruct_volume: all pointers are device
 * memory; d_sino_out may be NULL. Enqueued on the context's stream and
 * waited for before returning. */
int ct_hip_reconstruct_volume_dev(ct_hip_ctx* ctx, const float* d_vol, int w, int h, int d,
                                  int num_angles, int filter, float* d_out, float* d_sino_out);

/* Replaces IReconstructionBackend::reconstructVolumeFromSinograms
 * (IReconstructionBackend.h:36-41; default body BackendFactory.cpp:17-60):
 * sinos (nz,na,bins) -> out (nz,bins,bins), each slice as reconstructSlice.
 * Like the default body with a Serial backend, no denormalisation. */
int ct_hip_reconstruct_from_sinograms(ct_hip_ctx* ctx, const float* sinos, int nz, int na, int bins,
                                      const float* angles_deg, int filter, float* out);
int ct_hip_reconstruct_from_sinograms_dev(ct_hip_ctx* ctx, const float* d_sinos, int nz, int na,
                                          int bins, const float* angles_deg, int filter,
                                          float* d_out);

/* Allocation-free form of ct_hip_reconstruct_from_sinograms_dev: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments and, bit for bit, the same output; the
 * call allocates nothing. Its scratch (angle table, filter taps, twiddles,
 * filtered rows in slabs of 4-slice z-blocks) is one 64 MiB array per GPU,
 * held from library load to process end. Limits: the FFT filter backend
 * only, at most 8192 angles, and one z-block of filtered rows (na x
 * ceil(sqrt2*bins) x 16 bytes) must fit in 64 MiB less 104 KiB; beyond them
 * CT_ERR_UNSUPPORTED. The backprojection tiles are the tuned ones when the
 * (bins, na) class has been tuned on this context, else the context's
 * default (the same bits either way). Calls on one context are serialised;
 * calls on two contexts of one GPU that overlap race on the shared array.
 * Waits for completion. */
int ct_hip_reconstruct_from_sinograms_dev_noalloc(ct_hip_ctx* ctx, const float* d_sinos, int nz,
                                                  int na, int bins, const float* angles_deg,
                                                  int filter, float* d_out);

/* Replaces IReconstructionBackend::extractPointCloud (IReconstructionBackend.h:44;
 * SerialBackend.cpp:325-346): every voxel with value > threshold, in z,y,x
 * order, as 4 floats (x, y, z, hu). xc/yc/zc are the axis coordinates (host
 * arrays of w, h, d floats) or NULL for the indices. *count always receives
 * the number of points; points are written only when out_xyzh is not NULL
 * and capacity >= *count. The _dev form takes the volume in device memory. */
int ct_hip_extract_point_cloud(ct_hip_ctx* ctx, const float* vol, int w, int h, int d,
                               const float* xc, const float* yc, const float* zc, float threshold,
                               float* out_xyzh, size_t capacity, size_t* count);
int ct_hip_extract_point_cloud_dev(ct_hip_ctx* ctx, const float* d_vol, int w, int h, int d,
                                   const float* xc, const float* yc, const float* zc,
                                   float threshold, float* out_xyzh, size_t capacity,
                                   size_t* count);

/* Replaces FilterDesign::createFilter (FilterDesign.h:11; FilterDesign.cpp:13-111):
 * its n/2+1 taps bit for bit (the reference's FP32 FFT of the spatial ramp,
 * windowed in double; computed on the host). The row filter uses these taps
 * with the Nyquist bin zeroed, as FFT::inverseReal never reads it. */
int ct_hip_design_filter(ct_hip_ctx* ctx, int n, int filter, float* out);

/* Row-filter backend of the context, the analogue of the reference's
 * compile-time FFT backend choice (IFFTBackend.h:16-42, CMake FFT_BACKEND):
 *   CT_FILTER_BACKEND_FFT_F32        LDS FFT in FP32 (default)
 *   CT_FILTER_BACKEND_GEMM_BF16      Toeplitz GEMM on MFMA, BF16 inputs, FP32 accumulate
 *   CT_FILTER_BACKEND_GEMM_SPLIT_BF16 hi+lo BF16, three products, FP32 accumulate
 *   CT_FILTER_BACKEND_GEMM_FP8       FP8 E4M3 inputs with per-row scale
 * Backprojection always accumulates in FP32. */
enum { CT_FILTER_BACKEND_FFT_F32 = 0, CT_FILTER_BACKEND_GEMM_BF16 = 1,
       CT_FILTER_BACKEND_GEMM_SPLIT_BF16 = 2, CT_FILTER_BACKEND_GEMM_FP8 = 3 };
int ct_hip_set_filter_backend(ct_hip_ctx* ctx, int backend);

/* The row-filter step alone (SerialBackend.cpp:150-195): sinos (nz,na,bins) ->
 * out (nz,na,sq), sq = ceil(sqrt2*bins), with the context's filter backend. */
int ct_hip_filter_sinograms(ct_hip_ctx* ctx, const float* sinos, int nz, int na, int bins,
                            int filter, float* out);

/* ------------------------------------------------------------ batched form */

/* One volume of a batch: the input of reconstructVolumeFromSinograms
 * (IReconstructionBackend.h:36-41) in device memory, sinograms (nz, na, bins),
 * its angles in degrees (host array of na), a filter id, and a device output
 * of (nz, bins, bins). Each volume is reconstructed exactly as
 * ct_hip_reconstruct_from_sinograms_dev would reconstruct it alone. */
typedef struct {
    const float* d_sinos;
    float* d_out;
    const float* angles_deg;
    int nz, na, bins, filter;
} ct_fbp_job;

enum { CT_BATCH_GRAPH = 1,     /* capture the launches once in a hipGraph and replay it */
       CT_BATCH_AUTOTUNE = 2   /* pick the backprojection tiles with the autotuner */ };

/* Batched consumer path: the row filters of every volume run in persistent
 * launches (one per transform class) and the backprojections in one more,
 * all fed from work lists; with a Toeplitz-GEMM filter backend set on the
 * context (ct_hip_set_filter_backend) the filter is one MFMA launch per
 * volume instead. There is no reference counterpart (the reference loops per
 * slice, BackendFactory.cpp:17-60). Waits for completion. */
int ct_hip_fbp_batch_dev(ct_hip_ctx* ctx, const ct_fbp_job* jobs, int njobs, int flags);

/* Backprojection tile configurations (voxel tile edge, angles per LDS
 * stage, slices per thread: 4 or 8; slices may be NULL) and the autotuner
 * that picks one per (N, na) by timing them all on the device (cached in the
 * context and, when CT_HIP_TUNE_CACHE names a file, there). ms_out
 * (optional, n_ms entries) receives each configuration's time. Every
 * configuration gives bit-identical volumes. */
int ct_hip_bp_config_count(void);
int ct_hip_bp_config(int index, int* tile, int* angles, int* slices);
int ct_hip_autotune(ct_hip_ctx* ctx, int N, int na, int* config_out, double* ms_out, int n_ms);
/* Configuration used by the per-volume calls (default 0). */
int ct_hip_set_bp_config(ct_hip_ctx* ctx, int config);

/* ------------------------------------------------------------ serving form */

/* A request loop over device-resident sinograms with continuous batching
 * (no reference counterpart). The server owns a context and one device pool
 * of pool_bytes for outputs and filtered-sinogram workspaces; jobs are
 * admitted in the policy's order when both fit (the first that does not fit
 * stops admission), batched up to max_batch, and run in up to `slots`
 * batches in flight on separate streams. Each job's result equals its
 * standalone reconstructVolumeFromSinograms result bit for bit, whatever the
 * policy and the mix. */
typedef struct ct_hip_server ct_hip_server;
int ct_hip_server_create(int device, size_t pool_bytes, int max_batch, int slots,
                         ct_hip_server** out);
void ct_hip_server_destroy(ct_hip_server* srv);
/* CT_BATCH_* flags for the server's batches (default graph + autotune). */
int ct_hip_server_set_flags(ct_hip_server* srv, int flags);
/* The row filter backend of the server's batches (CT_FILTER_BACKEND_*;
 * default FFT_F32), for the batches formed after the call. With a
 * Toeplitz-GEMM backend each job's filter is one MFMA launch (split BF16
 * stays within 1e-4 of FP64) and its result equals a standalone
 * reconstructVolumeFromSinograms on a context with the same backend.
 * No counterpart in the reference: its backends pick their FFT at build
 * time (IFFTBackend). */
int ct_hip_server_set_filter_backend(ct_hip_server* srv, int backend);
/* Admission policy:
 *   CT_SERVE_FIFO        submission order;
 *   CT_SERVE_SIZE_AGED   (default) ascending estimated device time minus
 *                        age_weight x time waited (both in ms; default 0.25),
 *                        so small jobs pass large ones and a large job's
 *                        credit grows while it waits.
 * batch_cap_ms: a batch stops growing once its estimated device time would
 * pass this or a quarter of the queued work's, whichever is more (default
 * 5 ms; 0 = no cap; a batch always takes one job). The estimates use the
 * device rate measured over the server's completed batches. */
enum { CT_SERVE_FIFO = 0, CT_SERVE_SIZE_AGED = 1 };
int ct_hip_server_set_policy(ct_hip_server* srv, int policy, double batch_cap_ms,
                             double age_weight);
/* job->d_sinos must stay valid until the job is finished; job->d_out is
 * ignored (the output is placed in the pool). angles are copied. */
int ct_hip_server_submit(ct_hip_server* srv, const ct_fbp_job* job, uint64_t* ticket);
/* Waits up to timeout_ms for any finished job: CT_OK with its ticket, device
 * output and submit-to-completion latency; 1 on timeout; a negative code if
 * that job failed. */
int ct_hip_server_wait(ct_hip_server* srv, double timeout_ms, uint64_t* ticket,
                       const float** d_out, double* latency_ms);
/* Copies a finished job's output (bytes >= its size) to host memory through
 * the server's pinned staging; downloads of different tickets may run from
 * several threads. */
int ct_hip_server_download(ct_hip_server* srv, uint64_t ticket, float* host_out, size_t bytes);
/* Returns a finished job's output to the pool; refused (CT_ERR_INVALID)
 * while a download of the same ticket is in progress. */
int ct_hip_server_release(ct_hip_server* srv, uint64_t ticket);
/* [0] batches [1] completed [2] max jobs in flight [3] launches overlapping a
 * running batch [4] pool bytes in use [5] pool bytes [6] ms first submit ->
 * last completion */
int ct_hip_server_stats(ct_hip_server* srv, double* stats, int n);
const char* ct_hip_server_last_error(const ct_hip_server* srv);

/* ------------------------------------------------- reconstruction metrics */

/* The producer's quality figures of a reconstruction against its original
 * (CtReconstructionController.cpp:1199-1277), over every voxel:
 *   diff = recon - orig (float);
 *   mse = mean of double(diff)^2, mae = mean of |double(diff)|;
 *   max_abs = max |diff| (float), replaced by 1 when below 1e-6, as the
 *     controller's maxDifference;
 *   vmin / vmax = min / max of orig;
 *   psnr = 10 log10(range^2 / mse) with range = double(vmax - vmin), +inf
 *     when mse <= 0 or range <= 0.
 * The sums run on the device in double in a fixed order: the same inputs give
 * the same bits on every call; against the controller's OpenMP sums they
 * differ only by the order of the additions. */
typedef struct {
    double mse, mae, psnr;
    float max_abs, vmin, vmax;
    uint64_t count;
} ct_metrics;

/* orig and recon: n floats each, host memory. n = 0 gives count 0, mse and
 * mae 0, psnr +inf, max_abs 1, vmin FLT_MAX, vmax -FLT_MAX (the controller's
 * starting values). */
int ct_hip_volume_metrics(ct_hip_ctx* ctx, const float* orig, const float* recon, size_t n,
                          ct_metrics* out);
/* Device-resident form: d_orig and d_recon in device memory, 16-byte aligned. */
int ct_hip_volume_metrics_dev(ct_hip_ctx* ctx, const float* d_orig, const float* d_recon,
                              size_t n, ct_metrics* out);

/* ----------------------------------------------------------------- files */

/* Replaces FileIO::saveVolumeNPY (FileIO.h:11; FileIO.cpp:56-82): vol (d,h,w)
 * written as NPY 1.0 '<f4' of shape (d, h, w), byte for byte the reference's
 * file (header padded so that 10 + its length is a multiple of 16).
 * CT_ERR_INVALID for an empty volume, CT_ERR_IO when the file cannot be
 * written. Needs no context and no device. */
int ct_hip_save_volume_npy(const float* vol, size_t w, size_t h, size_t d, const char* path);

/* Replaces FileIO::loadVolumeNPY (FileIO.h:12; FileIO.cpp:84-145). dims3
 * receives {width, height, depth} of the Volume the reference builds: for
 * input_is_xyz_layout = 0 the file's (z, y, x) order is kept (width = last
 * axis); for 1 the file is (x, y, z) and is transposed into (z, y, x).
 * With vol == NULL only the header is read; otherwise vol (capacity floats,
 * at least w*h*d) receives the data. CT_ERR_IO for a missing file, a bad
 * magic, a dtype other than '<f4', a shape without three non-zero sizes or
 * short data. Versions 1.x and 2.x headers are read. */
int ct_hip_load_volume_npy(const char* path, int input_is_xyz_layout, float* vol,
                           size_t capacity, size_t* dims3);

/* Replaces FileIO::savePointCloudNPY / loadPointCloudNPY (FileIO.h:14-15;
 * FileIO.cpp:148-206): n points of 4 floats (x, y, z, hu) as an (n, 4) '<f4'
 * NPY. The loader sets *n, and fills xyzh when it is not NULL and
 * capacity >= *n points; a shape other than (n > 0, 4) is CT_ERR_IO. */
int ct_hip_save_point_cloud_npy(const float* xyzh, size_t n, const char* path);
int ct_hip_load_point_cloud_npy(const char* path, float* xyzh, size_t capacity, size_t* n);

/* Sinogram stack on disk: an NPY (nz, na, bins) and, beside it, the same path
 * with ".npy" replaced by ".meta.json" (version, depth, num_angles,
 * detector_bins, detector_spacing_mm, original_min_hu, original_max_hu,
 * angles_deg). */
typedef struct {
    size_t nz, na, bins;
    float detector_spacing_mm; /* from the sidecar, else 1 */
    float min_hu, max_hu;      /* from the sidecar, else the data's min/max */
    int meta_found;            /* the sidecar was read and parsed as a JSON object */
    int meta_minmax;           /* both min/max keys were present in it */
    int meta_angles;           /* its angle list had na entries */
} ct_sino_info;

/* Replaces CtReconstructionController::exportAllSinograms
 * (CtReconstructionController.cpp:1353-1482): computeSinogram of every slice
 * of vol (d,h,w) with num_angles angles and bins = w, on the device; writes
 * the NPY (d, num_angles, w) and the sidecar, whose angles are
 * float(180/na) * a in double and whose min/max are the extremes of the
 * slices' original min/max. */
int ct_hip_export_sinograms(ct_hip_ctx* ctx, const float* vol, int w, int h, int d,
                            int num_angles, const char* npy_path);

/* Replaces CtReconstructionController::loadSinograms
 * (CtReconstructionController.cpp:1494-1616). Two calls: with sinos == NULL
 * only the header and sidecar are read (info filled; min/max then come from
 * the sidecar or stay FLT_MAX / -FLT_MAX); with sinos (capacity >= nz*na*bins
 * floats) the data is read too, and min/max fall back to the data's. Each
 * sidecar key falls back on its own, as the reference's reader: spacing 1;
 * angles float(180/na) * a (float product) unless the sidecar lists exactly
 * na of them; min/max from the data unless both keys are present. A missing
 * or unparsable sidecar uses every default. angles_deg (optional) receives
 * na floats. Needs no context and no device. */
int ct_hip_load_sinograms(const char* npy_path, float* sinos, size_t capacity, float* angles_deg,
                          ct_sino_info* info);

/* ------------------------------------------------- synthetic acquisition */

/* Forward projection alone, device-resident: computeSinogram
 * (SerialBackend.cpp:59-136) of every slice, d_sino (d, na, bins);
 * minmax_out (optional host, 2*d) receives {min,max} per slice. */
int ct_hip_radon_dev(ct_hip_ctx* ctx, const float* d_vol, int w, int h, int d, int num_angles,
                     int bins, float* d_sino, float* minmax_out);
/* In place: counts ~ Poisson(I0 exp(-kappa p)) (normal approximation), then
 * p = -ln(max(counts,1)/I0)/kappa. Synthetic input for the serving form. */
int ct_hip_noisy_log_dev(ct_hip_ctx* ctx, float* d_sino, size_t n, float I0, float kappa,
                         uint64_t seed);

/* Shape limits of the kernels, the same checks every entry point makes before
 * any allocation or launch; needs no context and no device. No reference
 * counterpart (the reference's backends accept any shape):
 *   CT_STAGE_RADON  forward projection of a w x h slice into `bins` bins
 *                   (computeSinogram, reconstructVolume, the sinogram export):
 *                   w, h <= 2048 and bins <= 4096, the range over which the
 *                   walk's candidate margin is proven against Serial's float s;
 *   CT_STAGE_FBP    filter + backprojection of `bins`-bin rows with the given
 *                   CT_FILTER_BACKEND_*: the FFT backend needs nfft <= 4096
 *                   (bins <= 1448); w and h are ignored.
 * Returns CT_OK, CT_ERR_INVALID (a size <= 0, an unknown stage) or
 * CT_ERR_UNSUPPORTED. */
enum { CT_STAGE_RADON = 0, CT_STAGE_FBP = 1 };
int ct_hip_check_shape(int stage, int w, int h, int bins, int filter_backend);

/* Timings of the last volume call, milliseconds, measured with HIP events on
 * the context's stream: [0] normalise+Radon, [1] row filter,
 * [2] backprojection (incl. denormalisation), [3] total.
 * Replaces lastSinogramTimeMs / lastReconstructionTimeMs
 * (IReconstructionBackend.h:45-54). */
int ct_hip_last_timings(const ct_hip_ctx* ctx, double* ms4);

/* Breakdown of the Radon stage of the last volume call that ran one,
 * milliseconds, HIP events on the context's stream, summed over the call's
 * z-slabs: [0] the row walk over the image (angles with |sin| <= |cos|), [1]
 * the row walk over the transposed image (the other angles); *launches
 * (optional) gets the launch count of each walk (one per slab). When rows
 * and columns are both at most 516 px the two walks run as one launch: [0]
 * then holds that launch's time and [1] is zero. No counterpart in the
 * reference: per-kernel timing for profiling and the roofline report.
 * CT_ERR_INVALID before any Radon stage has run. */
int ct_hip_last_radon_timings(const ct_hip_ctx* ctx, double* ms2, int* launches);
/* Per-kernel milliseconds of the last call, summed over its slabs: [0] the
 * per-slice min/max pass, [1] the normalisation into the walks' copies, [2]
 * the walk over the image, [3] the walk over the transpose (both walks in
 * [2] and zero here when they run as one launch, as above), [4] the row
 * filter, [5] the backprojection; the first n (<= 6) are written. No
 * counterpart in the reference (profiling and the roofline report). */
int ct_hip_last_kernel_timings(const ct_hip_ctx* ctx, double* ms, int n);

#ifdef __cplusplus
}
#endif
#endif /* CT_HIP_H */
