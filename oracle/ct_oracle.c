/*
 * ct_oracle.c — CPU oracle for the FBP hot path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; never linked into or called by the product library.
 *
 * Build: gcc -O2 -ffp-contract=off (see oracle/Makefile). Float expressions
 * are written in the reference's operation order; with contraction off and
 * SSE float arithmetic every rounding matches the reference's source.
 *
 * Reference files restated (paths relative to QtCTReconstruction/):
 *   Utils.cpp:12-46, FFT.cpp:9-158, FilterDesign.cpp:13-111,
 *   SerialBackend.cpp:59-318, IReconstructionBackend.h / CTTypes.h layouts.
 */
#include "ct_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Utils.cpp                                                            */
/* ------------------------------------------------------------------ */

/* Utils.cpp:44-46: deg * pi_f / 180.0f, evaluated left to right. */
float ct_o_deg2rad(float deg) { return deg * 3.14159265358979323846f / 180.0f; }

/* Utils.cpp:28-42 (the _WIN64 branch, i.e. with the >>32 step). */
size_t ct_o_next_pow2(size_t v) {
    if (v <= 1) return 1;
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    v |= v >> 32;
    return v + 1;
}

/* Utils.cpp:12-26 */
void ct_o_linspace(float start, float stop, size_t count, float* out) {
    if (count == 0) return;
    if (count == 1) { out[0] = start; return; }
    const float step = (stop - start) / (float)(count - 1);
    for (size_t i = 0; i < count; ++i) out[i] = start + step * (float)i;
}

/* ------------------------------------------------------------------ */
/* FFT.cpp: radix-2 complex FFT and the packed real transforms          */
/* ------------------------------------------------------------------ */

typedef struct { float re, im; } cf;

/* std::complex<float> product: (ac - bd) + (ad + bc)i, each product rounded. */
static inline cf cmul(cf a, cf b) {
    cf r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline cf cadd(cf a, cf b) { cf r = { a.re + b.re, a.im + b.im }; return r; }
static inline cf csub(cf a, cf b) { cf r = { a.re - b.re, a.im - b.im }; return r; }
static inline cf cscale(cf a, float s) { cf r = { a.re * s, a.im * s }; return r; }
static inline cf cconj(cf a) { cf r = { a.re, -a.im }; return r; }

/* FFT.cpp:9-18 Plan::prepare */
static void plan_twiddles(cf* tw, size_t n, int inverse) {
    const float angle_base = 2.0f * 3.14159265359f / (float)n * (inverse ? 1.0f : -1.0f);
    for (size_t i = 0; i < n / 2; ++i) {
        const float angle = angle_base * (float)i;
        tw[i].re = cosf(angle);
        tw[i].im = sinf(angle);
    }
}

/* FFT.cpp:122-130 */
static void bit_reverse(cf* d, size_t n) {
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cf t = d[i]; d[i] = d[j]; d[j] = t; }
    }
}

/* FFT.cpp:132-158 */
static void fft_core(cf* d, size_t n, int inverse, const cf* tw) {
    if (n <= 1) return;
    bit_reverse(d, n);
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t j = 0; j < len / 2; ++j) {
                const cf v = cmul(d[i + j + len / 2], tw[j * step]);
                d[i + j + len / 2] = csub(d[i + j], v);
                d[i + j] = cadd(d[i + j], v);
            }
        }
    }
    if (inverse) {
        const float inv_n = 1.0f / (float)n;
        for (size_t i = 0; i < n; ++i) d[i] = cscale(d[i], inv_n);
    }
}

/* FFT.cpp:51-85 */
int ct_o_forward_real(const float* in, size_t n, float* out_c) {
    if (n == 0) return 0;
    if (n % 2 != 0) return -1;
    const size_t n2 = n / 2;
    cf* data = (cf*)malloc(sizeof(cf) * n2);
    cf* tw = (cf*)malloc(sizeof(cf) * (n2 / 2 + 1));
    cf* out = (cf*)out_c;
    for (size_t i = 0; i < n2; ++i) { data[i].re = in[2 * i]; data[i].im = in[2 * i + 1]; }
    plan_twiddles(tw, n2, 0);
    fft_core(data, n2, 0, tw);
    const float angle_step = -3.14159265359f / (float)n2;
    const cf mhalf_i = { 0.0f, -0.5f };
    for (size_t i = 0; i <= n2 / 2; ++i) {
        const size_t j = (n2 - i) % n2;
        const cf vi = data[i];
        const cf vj = cconj(data[j]);
        const cf even = cscale(cadd(vi, vj), 0.5f);
        const cf odd = cmul(csub(vi, vj), mhalf_i);
        const float angle = angle_step * (float)i;
        const cf t = { cosf(angle), sinf(angle) };
        out[i] = cadd(even, cmul(t, odd));
        if (i < n2 - i) out[n2 - i] = cconj(csub(even, cmul(t, odd)));
    }
    free(data);
    free(tw);
    return 0;
}

/* FFT.cpp:87-120. j=(n2-i)%n2 never reaches n2, so in[n2] is never read. */
int ct_o_inverse_real(const float* in_c, size_t n2, float* out) {
    if (n2 == 0) return 0;
    const cf* in = (const cf*)in_c;
    cf* data = (cf*)malloc(sizeof(cf) * n2);
    cf* tw = (cf*)malloc(sizeof(cf) * (n2 / 2 + 1));
    const float angle_step = 3.14159265359f / (float)n2;
    const cf one_i = { 0.0f, 1.0f };
    for (size_t i = 0; i < n2; ++i) {
        const size_t j = (n2 - i) % n2;
        const cf vi = in[i];
        const cf vj = cconj(in[j]);
        const cf even = cadd(vi, vj);
        const cf odd = cmul(csub(vi, vj), one_i);
        const float angle = angle_step * (float)i;
        const cf t = { cosf(angle), sinf(angle) };
        data[i] = cadd(even, cmul(t, odd));
    }
    plan_twiddles(tw, n2, 1);
    fft_core(data, n2, 1, tw);
    for (size_t i = 0; i < n2; ++i) {
        out[2 * i] = data[i].re * 0.5f;
        out[2 * i + 1] = data[i].im * 0.5f;
    }
    free(data);
    free(tw);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FilterDesign.cpp                                                     */
/* ------------------------------------------------------------------ */

static const double kPi = 3.14159265358979323846;

/* FilterDesign.cpp:31-50 */
static int filter_ramp(size_t n, float* filter) {
    float* f = (float*)calloc(n, sizeof(float));
    float* ff = (float*)malloc(sizeof(float) * 2 * (n / 2 + 1));
    f[0] = 0.25f;
    for (size_t i = 1; i < n; i += 2) {
        const size_t m = (i < n / 2) ? i : n - i;
        if (m == 0) continue;
        const double term = 1.0 / (kPi * (double)m);
        f[i] = (float)(-1.0 * term * term);
    }
    int rc = ct_o_forward_real(f, n, ff);
    for (size_t i = 0; i < n / 2 + 1; ++i) filter[i] = (float)(2.0 * (double)ff[2 * i]);
    filter[0] = 0.0f;
    free(f);
    free(ff);
    return rc;
}

/* FilterDesign.cpp:13-29 (dispatch; unknown types -> Shepp-Logan), 52-111 windows */
int ct_o_create_filter(size_t n, int type, float* out) {
    if (n < 2 || (n & (n - 1)) != 0) return -1;
    int rc = filter_ramp(n, out);
    if (rc) return rc;
    const size_t ns = n / 2 + 1;
    switch (type) {
    case CT_O_RAMP:
        break;
    case CT_O_HAMMING:
        for (size_t k = 0; k < ns; ++k) {
            const double freq = (double)k / (double)n;
            const double window = 0.54 + 0.46 * cos(kPi * freq / 0.5);
            out[k] = (float)((double)out[k] * window);
        }
        break;
    case CT_O_COSINE:
        for (size_t k = 0; k < ns; ++k) {
            const double freq = (double)k / (double)n;
            const double window = cos(kPi * freq);
            out[k] = (float)((double)out[k] * window);
        }
        break;
    case CT_O_HANN:
        for (size_t k = 0; k < ns; ++k) {
            const double freq = (double)k / (double)n;
            const double window = 0.5 + 0.5 * cos(kPi * freq / 0.5);
            out[k] = (float)((double)out[k] * window);
        }
        break;
    case CT_O_BARTLETT:
        for (size_t k = 0; k < ns; ++k) {
            const double freq = (double)k / (double)n;
            const double r = fabs(freq) / 0.5;
            const double window = 1.0 - (r < 1.0 ? r : 1.0);
            out[k] = (float)((double)out[k] * window);
        }
        break;
    case CT_O_SHEPP_LOGAN:
    default:
        for (size_t k = 1; k < ns; ++k) {
            const double freq = (double)k / (double)n;
            const double omega = kPi * freq;
            if (fabs(omega) > 1e-12) out[k] = (float)((double)out[k] * (sin(omega) / omega));
        }
        break;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* SerialBackend.cpp                                                    */
/* ------------------------------------------------------------------ */

/* SerialBackend.cpp:59-136 */
int ct_o_compute_sinogram(const float* slice, size_t w, size_t h, size_t na,
                          size_t bins, float* sino, float* minmax) {
    if (w == 0 || h == 0 || na == 0 || bins == 0) return -1;
    float hu_min = FLT_MAX, hu_max = -FLT_MAX;
    for (size_t i = 0; i < w * h; ++i) {
        const float v = slice[i];
        if (v < hu_min) hu_min = v;
        if (v > hu_max) hu_max = v;
    }
    float span = hu_max - hu_min;
    if (span < 1e-6f) span = 1.0f;
    const float inv_span = 1.0f / span;
    memset(sino, 0, sizeof(float) * na * bins);
    const float cx = (float)(w - 1) * 0.5f;
    const float cy = (float)(h - 1) * 0.5f;
    const float dc = (float)(bins / 2);
    const int db = (int)bins;
    for (size_t a = 0; a < na; ++a) {
        const float th = ct_o_deg2rad(180.0f / (float)na * (float)a);
        const float c = cosf(th), s = sinf(th);
        float* row = sino + a * bins;
        for (size_t y = 0; y < h; ++y) {
            const float* src = slice + y * w;
            const float yy = (float)y;
            for (size_t x = 0; x < w; ++x) {
                const float v_raw = src[x];
                if (v_raw == 0.0f) continue;
                const float v = (v_raw - hu_min) * inv_span;
                const float xx = (float)x;
                const float sd = (xx - cx) * c + (yy - cy) * s + dc;
                const int i0 = (int)floorf(sd);
                if (i0 >= 0 && i0 < db - 1) {
                    const float frac = sd - (float)i0;
                    row[i0] += v * (1.0f - frac);
                    row[i0 + 1] += v * frac;
                } else if (i0 == db - 1) {
                    row[i0] += v;
                }
            }
        }
    }
    if (minmax) { minmax[0] = hu_min; minmax[1] = hu_max; }
    return 0;
}

/* SerialBackend.cpp:150-161 */
void ct_o_geometry(size_t bins, size_t* sq, size_t* pad_before, size_t* nfft) {
    const size_t s = (size_t)ceil(sqrt(2.0) * (double)bins);
    if (sq) *sq = s;
    if (pad_before) *pad_before = s / 2 - bins / 2;
    if (nfft) {
        const size_t p = ct_o_next_pow2(2 * s);
        *nfft = p > 64 ? p : 64;
    }
}

/* SerialBackend.cpp:142-255 */
int ct_o_reconstruct_slice(const float* sino, size_t na, size_t bins,
                           const float* angles_deg, size_t N, int filter_type,
                           float* filtered_out, float* recon) {
    if (bins == 0 || na == 0 || N == 0) return -1;
    size_t sq, pad, nfft;
    ct_o_geometry(bins, &sq, &pad, &nfft);
    const size_t ns = nfft / 2 + 1;
    float* H = (float*)malloc(sizeof(float) * ns);
    if (ct_o_create_filter(nfft, filter_type, H)) { free(H); return -1; }
    float* pp = (float*)malloc(sizeof(float) * nfft);
    float* spec = (float*)malloc(sizeof(float) * 2 * ns);
    float* q = (float*)malloc(sizeof(float) * nfft);
    float* filt = filtered_out ? filtered_out : (float*)malloc(sizeof(float) * na * sq);
    for (size_t a = 0; a < na; ++a) {
        memset(pp, 0, sizeof(float) * nfft);
        for (size_t i = 0; i < bins; ++i) {
            const size_t dst = i + pad;
            if (dst < nfft) pp[dst] = sino[a * bins + i];
        }
        ct_o_forward_real(pp, nfft, spec);
        for (size_t k = 0; k < ns; ++k) { spec[2 * k] *= H[k]; spec[2 * k + 1] *= H[k]; }
        ct_o_inverse_real(spec, nfft / 2, q);
        memcpy(filt + a * sq, q, sizeof(float) * sq);
    }
    float* ct = (float*)malloc(sizeof(float) * na);
    float* st = (float*)malloc(sizeof(float) * na);
    for (size_t a = 0; a < na; ++a) {
        const float th = ct_o_deg2rad(angles_deg[a]);
        ct[a] = cosf(th);
        st[a] = sinf(th);
    }
    const float radius = (float)(N / 2);
    const float dcb = (float)(sq / 2);
    const int sqi = (int)sq;
    for (size_t y = 0; y < N; ++y) {
        const float yy = (float)y - radius;
        for (size_t x = 0; x < N; ++x) {
            const float xx = (float)x - radius;
            if (xx * xx + yy * yy > radius * radius) { recon[y * N + x] = 0.0f; continue; }
            float sum = 0.0f;
            for (size_t a = 0; a < na; ++a) {
                const float t = xx * ct[a] - yy * st[a];
                const float u = t + dcb;
                const int i0 = (int)floorf(u);
                const int i1 = i0 + 1;
                const float frac = u - (float)i0;
                if (i0 >= 0 && i1 < sqi) {
                    const float* fr = filt + a * sq;
                    sum += fr[i0] * (1.0f - frac) + fr[i1] * frac;
                }
            }
            recon[y * N + x] = sum;
        }
    }
    const float scale = 3.14159265358979323846f / (2.0f * (float)na);
    for (size_t i = 0; i < N * N; ++i) recon[i] *= scale;
    free(ct); free(st);
    if (!filtered_out) free(filt);
    free(H); free(pp); free(spec); free(q);
    return 0;
}

/* SerialBackend.cpp:261-318, restricted to the slices in z_list. Each slice
 * is computed exactly as the sequential loop does; slices are independent, so
 * running them on several threads changes nothing in the result. */
int ct_o_reconstruct_volume(const float* vol, size_t w, size_t h, size_t d,
                            size_t na, int filter, const int* z_list, size_t nz_list,
                            float* out, float* sino_out, int threads) {
    if (w == 0 || h == 0 || d == 0 || na == 0) return -1;
    float* angles = (float*)malloc(sizeof(float) * na);
    const float step = 180.0f / (float)na;
    for (size_t a = 0; a < na; ++a) angles[a] = step * (float)a;
    int rc = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(threads) reduction(| : rc)
#endif
    for (long k = 0; k < (long)nz_list; ++k) {
        const size_t z = (size_t)z_list[k];
        float* sino = (float*)malloc(sizeof(float) * na * w);
        float* recon = (float*)malloc(sizeof(float) * w * w);
        float mm[2];
        rc |= ct_o_compute_sinogram(vol + z * w * h, w, h, na, w, sino, mm);
        rc |= ct_o_reconstruct_slice(sino, na, w, angles, w, filter, NULL, recon);
        const float span = mm[1] - mm[0];
        const float hu_min = mm[0];
        const float eff_span = (span <= 0.0f) ? 1.0f : span;
        for (size_t i = 0; i < w * w; ++i) recon[i] = recon[i] * eff_span + hu_min;
        /* Buffer3D::setSlice (CTTypes.h:112-116) is a no-op unless h == w. */
        if (h == w) memcpy(out + z * w * h, recon, sizeof(float) * w * h);
        if (sino_out) memcpy(sino_out + (size_t)k * na * w, sino, sizeof(float) * na * w);
        free(sino);
        free(recon);
    }
    (void)threads;
    free(angles);
    return rc;
}

/* ------------------------------------------------------------------ */
/* FP64 twin: the same method evaluated in double.                      */
/*   - angles: theta = pi*a/na exactly (in double)                      */
/*   - filter: H = 2 Re DFT(f) by direct summation, windows as the      */
/*     reference, H[0] = H[n/2] = 0 (the effective Serial filter)       */
/*   - row filter: exact circular convolution with h = IDFT(H)          */
/* ------------------------------------------------------------------ */

static void ramp64(size_t n, double* H) {
    double* f = (double*)calloc(n, sizeof(double));
    f[0] = 0.25;
    for (size_t i = 1; i < n; i += 2) {
        const size_t m = (i < n / 2) ? i : n - i;
        const double t = 1.0 / (kPi * (double)m);
        f[i] = -t * t;
    }
    /* f is even (f[i] = f[n-i]), so DFT is real: F[k] = sum_j f[j] cos(2 pi jk/n).
     * Exact argument reduction: (j*k) mod n keeps the cosine argument small. */
    for (size_t k = 0; k <= n / 2; ++k) {
        double acc = 0.0;
        for (size_t j = 0; j < n; ++j) {
            if (f[j] == 0.0) continue;
            const size_t r = (j * k) % n;
            acc += f[j] * cos(2.0 * kPi * (double)r / (double)n);
        }
        H[k] = 2.0 * acc;
    }
    H[0] = 0.0;
    free(f);
}

int ct_o64_filter(size_t n, int type, double* H) {
    if (n < 2 || (n & (n - 1)) != 0) return -1;
    ramp64(n, H);
    const size_t ns = n / 2 + 1;
    for (size_t k = 0; k < ns; ++k) {
        const double freq = (double)k / (double)n;
        double w = 1.0;
        switch (type) {
        case CT_O_RAMP: w = 1.0; break;
        case CT_O_HAMMING: w = 0.54 + 0.46 * cos(kPi * freq / 0.5); break;
        case CT_O_COSINE: w = cos(kPi * freq); break;
        case CT_O_HANN: w = 0.5 + 0.5 * cos(kPi * freq / 0.5); break;
        case CT_O_BARTLETT: { const double r = fabs(freq) / 0.5; w = 1.0 - (r < 1.0 ? r : 1.0); } break;
        case CT_O_SHEPP_LOGAN:
        default: { const double om = kPi * freq; w = (k == 0) ? 1.0 : sin(om) / om; } break;
        }
        H[k] *= w;
    }
    H[0] = 0.0;
    H[n / 2] = 0.0; /* FFT.cpp:95-108 never reads the Nyquist bin */
    return 0;
}

int ct_o64_kernel(size_t n, int type, double* h) {
    const size_t ns = n / 2 + 1;
    double* H = (double*)malloc(sizeof(double) * ns);
    if (ct_o64_filter(n, type, H)) { free(H); return -1; }
    /* h[k] = (1/n) sum_{m=0}^{n-1} Hfull[m] e^{2 pi i mk/n}, Hfull even and real. */
    for (size_t k = 0; k < n; ++k) {
        double acc = H[0];
        for (size_t m = 1; m < n / 2; ++m) {
            const size_t r = (m * k) % n;
            acc += 2.0 * H[m] * cos(2.0 * kPi * (double)r / (double)n);
        }
        acc += H[n / 2] * ((k & 1) ? -1.0 : 1.0);
        h[k] = acc / (double)n;
    }
    free(H);
    return 0;
}

int ct_o64_compute_sinogram(const float* slice, size_t w, size_t h, size_t na,
                            size_t bins, double* sino) {
    if (w == 0 || h == 0 || na == 0 || bins == 0) return -1;
    float hu_min = FLT_MAX, hu_max = -FLT_MAX;
    for (size_t i = 0; i < w * h; ++i) {
        if (slice[i] < hu_min) hu_min = slice[i];
        if (slice[i] > hu_max) hu_max = slice[i];
    }
    double span = (double)hu_max - (double)hu_min;
    if (span < 1e-6) span = 1.0;
    memset(sino, 0, sizeof(double) * na * bins);
    const double cx = ((double)w - 1.0) * 0.5, cy = ((double)h - 1.0) * 0.5;
    const double dc = (double)(bins / 2);
    const long db = (long)bins;
    for (size_t a = 0; a < na; ++a) {
        const double th = kPi * (double)a / (double)na;
        const double c = cos(th), s = sin(th);
        double* row = sino + a * bins;
        for (size_t y = 0; y < h; ++y) {
            for (size_t x = 0; x < w; ++x) {
                const float v_raw = slice[y * w + x];
                if (v_raw == 0.0f) continue;
                const double v = ((double)v_raw - (double)hu_min) / span;
                const double sd = ((double)x - cx) * c + ((double)y - cy) * s + dc;
                const long i0 = (long)floor(sd);
                if (i0 >= 0 && i0 < db - 1) {
                    const double fr = sd - (double)i0;
                    row[i0] += v * (1.0 - fr);
                    row[i0 + 1] += v * fr;
                } else if (i0 == db - 1) {
                    row[i0] += v;
                }
            }
        }
    }
    return 0;
}

int ct_o64_reconstruct_slice(const double* sino, size_t na, size_t bins,
                             const double* angles_deg, size_t N, int filter_type,
                             double* filtered_out, double* recon) {
    if (bins == 0 || na == 0 || N == 0) return -1;
    size_t sq, pad, nfft;
    ct_o_geometry(bins, &sq, &pad, &nfft);
    double* hk = (double*)malloc(sizeof(double) * nfft);
    if (ct_o64_kernel(nfft, filter_type, hk)) { free(hk); return -1; }
    double* filt = filtered_out ? filtered_out : (double*)malloc(sizeof(double) * na * sq);
    /* q[i] = sum_j h[(i - j) mod n] p_pad[j]; p_pad nonzero on [pad, pad+bins). */
    for (size_t a = 0; a < na; ++a) {
        for (size_t i = 0; i < sq; ++i) {
            double acc = 0.0;
            for (size_t j = 0; j < bins; ++j) {
                const long d = (long)i - (long)(j + pad);
                const size_t k = (size_t)(d >= 0 ? d : (long)nfft + d);
                acc += hk[k] * sino[a * bins + j];
            }
            filt[a * sq + i] = acc;
        }
    }
    const double radius = (double)(N / 2);
    const double dcb = (double)(sq / 2);
    double* ct = (double*)malloc(sizeof(double) * na);
    double* st = (double*)malloc(sizeof(double) * na);
    for (size_t a = 0; a < na; ++a) {
        const double th = angles_deg[a] * kPi / 180.0;
        ct[a] = cos(th);
        st[a] = sin(th);
    }
    for (size_t y = 0; y < N; ++y) {
        const double yy = (double)y - radius;
        for (size_t x = 0; x < N; ++x) {
            const double xx = (double)x - radius;
            if (xx * xx + yy * yy > radius * radius) { recon[y * N + x] = 0.0; continue; }
            double sum = 0.0;
            for (size_t a = 0; a < na; ++a) {
                const double u = xx * ct[a] - yy * st[a] + dcb;
                const long i0 = (long)floor(u);
                const double fr = u - (double)i0;
                if (i0 >= 0 && i0 + 1 < (long)sq) {
                    const double* f = filt + a * sq;
                    sum += f[i0] * (1.0 - fr) + f[i0 + 1] * fr;
                }
            }
            recon[y * N + x] = sum * (kPi / (2.0 * (double)na));
        }
    }
    free(ct); free(st); free(hk);
    if (!filtered_out) free(filt);
    return 0;
}

int ct_o64_reconstruct_volume(const float* vol, size_t w, size_t h, size_t d,
                              size_t na, int filter, const int* z_list, size_t nz_list,
                              double* out, int threads) {
    if (w == 0 || h == 0 || d == 0 || na == 0) return -1;
    double* angles = (double*)malloc(sizeof(double) * na);
    for (size_t a = 0; a < na; ++a) angles[a] = 180.0 * (double)a / (double)na;
    int rc = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(threads) reduction(| : rc)
#endif
    for (long k = 0; k < (long)nz_list; ++k) {
        const size_t z = (size_t)z_list[k];
        const float* sl = vol + z * w * h;
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (size_t i = 0; i < w * h; ++i) { if (sl[i] < mn) mn = sl[i]; if (sl[i] > mx) mx = sl[i]; }
        double* sino = (double*)malloc(sizeof(double) * na * w);
        double* recon = (double*)malloc(sizeof(double) * w * w);
        rc |= ct_o64_compute_sinogram(sl, w, h, na, w, sino);
        rc |= ct_o64_reconstruct_slice(sino, na, w, angles, w, filter, NULL, recon);
        const double span = (double)mx - (double)mn;
        const double eff = span <= 0.0 ? 1.0 : span;
        if (h == w)
            for (size_t i = 0; i < w * w; ++i) out[z * w * h + i] = recon[i] * eff + (double)mn;
        free(sino);
        free(recon);
    }
    (void)threads;
    free(angles);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Project phantom: additive, rotated ellipsoids from a counter RNG.    */
/* (No reference counterpart: Generator3DCPU.cpp draws a different,     */
/* max-composited brain phantom.)                                       */
/* ------------------------------------------------------------------ */

static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* uniform in [lo, hi): 53-bit draw for (seed, ellipsoid e, parameter p) */
static float draw(uint64_t seed, int e, int p, double lo, double hi) {
    const uint64_t ctr = (uint64_t)e * 16u + (uint64_t)p;
    const uint64_t x = splitmix64(seed ^ splitmix64(ctr));
    const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    return (float)(lo + (hi - lo) * u);
}

int ct_ph_params(uint64_t seed, int n, ct_o_ellipsoid* out) {
    if (n < 0) return -1;
    for (int e = 0; e < n; ++e) {
        ct_o_ellipsoid* q = out + e;
        q->cx = draw(seed, e, 0, -0.6, 0.6);
        q->cy = draw(seed, e, 1, -0.6, 0.6);
        q->cz = draw(seed, e, 2, -0.6, 0.6);
        q->ax = draw(seed, e, 3, 0.05, 0.4);
        q->ay = draw(seed, e, 4, 0.05, 0.4);
        q->az = draw(seed, e, 5, 0.05, 0.4);
        q->phi = draw(seed, e, 6, 0.0, kPi);
        q->theta = draw(seed, e, 7, 0.0, kPi);
        q->psi = draw(seed, e, 8, 0.0, kPi);
        q->rho = draw(seed, e, 9, -0.3, 1.0);
    }
    return 0;
}

/* World->body map M = R^T with R = Rz(phi) Ry(theta) Rz(psi), rounded to
 * float, and inverse squared semi-axes rounded to float. */
typedef struct { float m[3][3]; float ia[3]; } ph_derived;

static void derive(const ct_o_ellipsoid* e, ph_derived* d) {
    const double a = e->phi, b = e->theta, g = e->psi;
    const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
    const double Rz1[3][3] = { { ca, -sa, 0 }, { sa, ca, 0 }, { 0, 0, 1 } };
    const double Ry[3][3] = { { cb, 0, sb }, { 0, 1, 0 }, { -sb, 0, cb } };
    const double Rz2[3][3] = { { cg, -sg, 0 }, { sg, cg, 0 }, { 0, 0, 1 } };
    double T[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Rz1[i][k] * Ry[k][j];
            T[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += T[i][k] * Rz2[k][j];
            R[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d->m[i][j] = (float)R[j][i];
    const double ax = e->ax, ay = e->ay, az = e->az;
    d->ia[0] = (float)(1.0 / (ax * ax));
    d->ia[1] = (float)(1.0 / (ay * ay));
    d->ia[2] = (float)(1.0 / (az * az));
}

int ct_ph_rasterize(const ct_o_ellipsoid* e, int n, int nx, int ny, int nz,
                    float* vol, int threads) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || n < 0) return -1;
    ph_derived* dv = (ph_derived*)malloc(sizeof(ph_derived) * (n > 0 ? n : 1));
    for (int k = 0; k < n; ++k) derive(e + k, dv + k);
    float* xs = (float*)malloc(sizeof(float) * nx);
    float* ys = (float*)malloc(sizeof(float) * ny);
    float* zs = (float*)malloc(sizeof(float) * nz);
    ct_o_linspace(-1.0f, 1.0f, nx, xs);
    ct_o_linspace(-1.0f, 1.0f, ny, ys);
    ct_o_linspace(-1.0f, 1.0f, nz, zs);
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int z = 0; z < nz; ++z) {
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) {
                float v = 0.0f;
                for (int k = 0; k < n; ++k) {
                    const ph_derived* d = dv + k;
                    const float dx = xs[x] - e[k].cx;
                    const float dy = ys[y] - e[k].cy;
                    const float dz = zs[z] - e[k].cz;
                    const float q0 = d->m[0][0] * dx + d->m[0][1] * dy + d->m[0][2] * dz;
                    const float q1 = d->m[1][0] * dx + d->m[1][1] * dy + d->m[1][2] * dz;
                    const float q2 = d->m[2][0] * dx + d->m[2][1] * dy + d->m[2][2] * dz;
                    const float f = q0 * q0 * d->ia[0] + q1 * q1 * d->ia[1] + q2 * q2 * d->ia[2];
                    if (f <= 1.0f) v += e[k].rho;
                }
                vol[((size_t)z * ny + y) * nx + x] = v;
            }
        }
    }
    (void)threads;
    free(dv); free(xs); free(ys); free(zs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* CtReconstructionController.cpp:1199-1277: the producer's metrics      */
/* ------------------------------------------------------------------ */

/* Sequential restatement of the controller's metrics loop and its finish:
 * d = recon - orig (float); mse and mae accumulate double(d)^2 and
 * |double(d)|; max |d| (float), min/max of orig (float), count. Then
 * max_abs < 1e-6 -> 1, means over the count, range = double(vmax - vmin),
 * psnr = +inf when mse <= 0 or range <= 0, else 10 log10(range^2 / mse).
 * out_d = {mse, mae, psnr}; out_f = {max_abs, vmin, vmax}. */
int ct_o_metrics(const float* orig, const float* recon, size_t n, double* out_d, float* out_f) {
    double mse = 0.0, mae = 0.0;
    float max_abs = 0.0f;
    float vmin = FLT_MAX, vmax = -FLT_MAX;
    for (size_t i = 0; i < n; ++i) {
        const float diff = recon[i] - orig[i];
        const float o = orig[i];
        mse += (double)diff * (double)diff;
        mae += fabs((double)diff);
        /* std::max(a, b) = (a < b) ? b : a; std::min(a, b) = (b < a) ? b : a */
        const float ad = fabsf(diff);
        max_abs = (max_abs < ad) ? ad : max_abs;
        vmin = (o < vmin) ? o : vmin;
        vmax = (vmax < o) ? o : vmax;
    }
    if (max_abs < 1e-6f) max_abs = 1.0f;
    mse = (n > 0) ? mse / (double)n : 0.0;
    mae = (n > 0) ? mae / (double)n : 0.0;
    const double range = (double)(vmax - vmin);
    const double psnr = (mse <= 0.0 || range <= 0.0) ? INFINITY : 10.0 * log10((range * range) / mse);
    out_d[0] = mse;
    out_d[1] = mae;
    out_d[2] = psnr;
    out_f[0] = max_abs;
    out_f[1] = vmin;
    out_f[2] = vmax;
    return 0;
}
