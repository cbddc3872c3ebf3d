/*
 * ct_ompbase.c — restatement of OpenMPBackend::reconstructVolume, used ONLY
 * as the timed CPU baseline (bench.py's cpu_baseline leg). TEST
 * INFRASTRUCTURE: never linked into or called by the product library.
 *
 * It follows the reference's algorithm and costs step by step, including
 * its geometry conventions, which differ from SerialBackend's (no y-flip,
 * (int) truncation, (n-1)/2 detector centre, no circular mask), so its output
 * is not a parity target.
 *
 * Reference files (relative to QtCTReconstruction/):
 *   OpenMPBackend.cpp:18-35, 61-190   computeSinogram, reconstructVolume
 *   RadonTransform.cpp:17-113          angle-batched scatter
 *   FFTBackendBuiltin.cpp:20-54        per-call FFT twiddles (FFT.cpp)
 *   ProjectionGeometry.cpp:7-56        parallel-beam 3x4 matrices
 *   OptimizedBackprojectionCPU.cpp:22-174  z-symmetric row backprojection
 */
#include "ct_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#if defined(__AVX512F__)
#define ANGLE_BATCH 16
#define BP_SIMD_WIDTH 16
#elif defined(__AVX2__) || defined(__AVX__)
#define ANGLE_BATCH 8
#define BP_SIMD_WIDTH 8
#else
#define ANGLE_BATCH 4
#define BP_SIMD_WIDTH 4
#endif

int ct_omp_reconstruct_volume(const float* vol, size_t w, size_t h, size_t d,
                              size_t na, int filter, float* out, int threads,
                              double* phase_ms);
int ct_omp_radon_forward(const float* slice, size_t w, size_t h, size_t na, size_t bins,
                         float* sino);

/* RadonTransform.cpp:25-113 with use_parallel=false (the call made from the
 * volume loop, OpenMPBackend.cpp:113). */
static void radon_forward(const float* slice, size_t w, size_t h, size_t na, size_t bins,
                          float min_hu, float max_hu, float* sino) {
    const float cx = (float)(w - 1) * 0.5f;
    const float cy = (float)(h - 1) * 0.5f;
    const float dc = (float)(bins / 2);
    float span = max_hu - min_hu;
    if (span < 1e-6f) span = 1.0f;
    const float inv_span = 1.0f / span;
    memset(sino, 0, sizeof(float) * na * bins);
    const int ina = (int)na, db = (int)bins;
    for (int b0 = 0; b0 < ina; b0 += ANGLE_BATCH) {
        const int bc = (ina - b0) < ANGLE_BATCH ? (ina - b0) : ANGLE_BATCH;
        float bcos[ANGLE_BATCH], bsin[ANGLE_BATCH], bub[ANGLE_BATCH], buy[ANGLE_BATCH];
        float* rows[ANGLE_BATCH];
        for (int b = 0; b < bc; ++b) {
            const float th = ct_o_deg2rad(180.0f / (float)ina * (float)(b0 + b));
            bcos[b] = cosf(th);
            bsin[b] = sinf(th);
            bub[b] = -cx * bcos[b] + cy * bsin[b] + dc;
            rows[b] = sino + (size_t)(b0 + b) * bins;
        }
        for (size_t y = 0; y < h; ++y) {
            const float yy = (float)y;
            const float* sp = slice + y * w;
            for (int b = 0; b < bc; ++b) buy[b] = -yy * bsin[b] + bub[b];
            for (size_t x = 0; x < w; ++x) {
                const float v_raw = sp[x];
                if (v_raw == 0.0f) continue;
                const float v = (v_raw - min_hu) * inv_span;
                const float xx = (float)x;
                for (int b = 0; b < bc; ++b) {
                    const float u = xx * bcos[b] + buy[b];
                    const int i0 = (int)u;
                    if (i0 >= 0 && i0 < db - 1) {
                        const float fr = u - (float)i0;
                        rows[b][i0] += v * (1.0f - fr);
                        rows[b][i0 + 1] += v * fr;
                    } else if (i0 == db - 1) {
                        rows[b][i0] += v;
                    }
                }
            }
        }
    }
}

/* The volume loop's Radon stage of one slice (min/max as OpenMPBackend.cpp:101-113
 * take them), exported so the tests can pin it by closed forms. */
int ct_omp_radon_forward(const float* slice, size_t w, size_t h, size_t na, size_t bins,
                         float* sino) {
    if (w == 0 || h == 0 || na == 0 || bins == 0) return -1;
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (size_t i = 0; i < w * h; ++i) { if (slice[i] < mn) mn = slice[i]; if (slice[i] > mx) mx = slice[i]; }
    radon_forward(slice, w, h, na, bins, mn, mx, sino);
    return 0;
}

typedef struct { const float* rp; const float* rn; float du, ubp, ubn, dudy; } bdata;

/* OptimizedBackprojectionCPU.cpp:113-149 process_row */
static inline void process_row(float* restrict row, const float* restrict pr, float u_base,
                               float du, int nx, int nw) {
    if (!pr) return;
    int xs0 = 0, xe0 = nx;
    if (du != 0.0f) {
        float xs = (0.0f - u_base) / du;
        float xe = ((float)(nw - 1) - u_base) / du;
        if (du < 0) { const float t = xs; xs = xe; xe = t; }
        const int a = (int)ceilf(xs), b = (int)floorf(xe);
        xs0 = a > 0 ? a : 0;
        xe0 = b < nx ? b : nx;
    }
    if (xs0 >= xe0) return;
    int x = xs0;
    for (; x <= xe0 - BP_SIMD_WIDTH; x += BP_SIMD_WIDTH) {
#pragma omp simd
        for (int k = 0; k < BP_SIMD_WIDTH; ++k) {
            const float uf = (float)(x + k) * du + u_base;
            const int u0 = (int)uf;
            row[x + k] += pr[u0] + (uf - (float)u0) * (pr[u0 + 1] - pr[u0]);
        }
    }
#pragma omp simd
    for (int xt = x; xt < xe0; ++xt) {
        const float uf = (float)xt * du + u_base;
        const int u0 = (int)uf;
        row[xt] += pr[u0] + (uf - (float)u0) * (pr[u0 + 1] - pr[u0]);
    }
}

int ct_omp_reconstruct_volume(const float* vol, size_t w, size_t h, size_t d,
                              size_t na, int filter, float* out, int threads,
                              double* phase_ms) {
    if (w == 0 || h == 0 || d == 0 || na == 0) return -1;
    if (threads <= 0) threads = omp_get_max_threads();
    const size_t bins = w;
    size_t sq, pad, nfft;
    ct_o_geometry(bins, &sq, &pad, &nfft);
    const size_t ns = nfft / 2 + 1;
    float* H = (float*)malloc(sizeof(float) * ns);
    ct_o_create_filter(nfft, filter, H);
    float* fall = (float*)calloc(sq * na * d, sizeof(float)); /* rows [a*d + z][sq] */
    float* min_hus = (float*)malloc(sizeof(float) * d);
    float* spans = (float*)malloc(sizeof(float) * d);
    const double t0 = omp_get_wtime();
    /* OpenMPBackend.cpp:101-141 */
#pragma omp parallel num_threads(threads)
    {
        float* pp = (float*)malloc(sizeof(float) * nfft);
        float* spec = (float*)malloc(sizeof(float) * 2 * ns);
        float* q = (float*)malloc(sizeof(float) * nfft);
        float* sino = (float*)malloc(sizeof(float) * na * bins);
#pragma omp for schedule(dynamic)
        for (long z = 0; z < (long)d; ++z) {
            const float* sl = vol + (size_t)z * w * h;
            float mn = FLT_MAX, mx = -FLT_MAX;
            for (size_t i = 0; i < w * h; ++i) { if (sl[i] < mn) mn = sl[i]; if (sl[i] > mx) mx = sl[i]; }
            radon_forward(sl, w, h, na, bins, mn, mx, sino);
            min_hus[z] = mn;
            const float s = mx - mn;
            spans[z] = (s <= 0.0f) ? 1.0f : s;
            for (size_t a = 0; a < na; ++a) {
                memset(pp, 0, sizeof(float) * nfft);
                for (size_t i = 0; i < bins; ++i) {
                    const size_t dst = i + pad;
                    if (dst < nfft) pp[dst] = sino[a * bins + i];
                }
                ct_o_forward_real(pp, nfft, spec);
                for (size_t k = 0; k < ns; ++k) { spec[2 * k] *= H[k]; spec[2 * k + 1] *= H[k]; }
                ct_o_inverse_real(spec, nfft / 2, q);
                memcpy(fall + (a * d + (size_t)z) * sq, q, sizeof(float) * sq);
            }
        }
        free(pp); free(spec); free(q); free(sino);
    }
    const double t1 = omp_get_wtime();

    /* ProjectionGeometry.cpp:7-56 + OptimizedBackprojectionCPU.cpp:22-174 */
    const int nx = (int)w, ny = (int)h, nz = (int)d, nw = (int)sq, nh = (int)d, np = (int)na;
    const float cx = (float)(nx - 1) * 0.5f, cy = (float)(ny - 1) * 0.5f, cz = (float)(nz - 1) * 0.5f;
    const float det_cx = (float)(nw - 1) * 0.5f, det_cy = (float)(nh - 1) * 0.5f;
    float* du_dx = (float*)malloc(sizeof(float) * np);
    float* du_dy = (float*)malloc(sizeof(float) * np);
    float* u_b0 = (float*)malloc(sizeof(float) * np);
    float* v_b = (float*)malloc(sizeof(float) * np);
    const float astep = 180.0f / (float)np;
    for (int p = 0; p < np; ++p) {
        const float rad = ct_o_deg2rad(astep * (float)p);
        const float ca = cosf(rad), sa = sinf(rad);
        du_dx[p] = ca * 1.0f;
        du_dy[p] = -sa * 1.0f;
        u_b0[p] = det_cx - (cx * ca - cy * sa) * 1.0f;
        v_b[p] = det_cy - cz * 1.0f;
    }
    const int nz_half = (nz + 1) / 2;
    const size_t ss = (size_t)nx * ny;
#pragma omp parallel for schedule(dynamic) num_threads(threads)
    for (int z = 0; z < nz_half; ++z) {
        memset(out + (size_t)z * ss, 0, sizeof(float) * ss);
        if (z != nz - 1 - z) memset(out + (size_t)(nz - 1 - z) * ss, 0, sizeof(float) * ss);
    }
#pragma omp parallel num_threads(threads)
    {
        bdata* bd = (bdata*)malloc(sizeof(bdata) * np);
#pragma omp for schedule(dynamic)
        for (int z = 0; z < nz_half; ++z) {
            const int zs = nz - 1 - z;
            const int has_sym = (z != zs);
            const float fz = (float)z, fzs = (float)zs;
            float* vp = out + (size_t)z * ss;
            float* vn = has_sym ? out + (size_t)zs * ss : NULL;
            for (int p = 0; p < np; ++p) {
                const int v0p = (int)floorf(1.0f * fz + v_b[p] + 0.5f);
                const int v0n = (int)floorf(1.0f * fzs + v_b[p] + 0.5f);
                bd[p].rp = (v0p >= 0 && v0p < nh) ? fall + ((size_t)p * nh + v0p) * nw : NULL;
                bd[p].rn = (vn && v0n >= 0 && v0n < nh) ? fall + ((size_t)p * nh + v0n) * nw : NULL;
                bd[p].du = du_dx[p];
                bd[p].dudy = du_dy[p];
                bd[p].ubp = fz * 0.0f + u_b0[p];
                bd[p].ubn = fzs * 0.0f + u_b0[p];
            }
            for (int p0 = 0; p0 < np; p0 += 16) {
                const int p1 = (p0 + 16) < np ? p0 + 16 : np;
                for (int y = 0; y < ny; ++y) {
                    const float fy = (float)y;
                    float* rp = vp + (size_t)y * nx;
                    float* rn = vn ? vn + (size_t)y * nx : NULL;
                    for (int p = p0; p < p1; ++p) {
                        const bdata* b = bd + p;
                        if (!b->rp && !b->rn) continue;
                        process_row(rp, b->rp, fy * b->dudy + b->ubp, b->du, nx, nw);
                        if (rn) process_row(rn, b->rn, fy * b->dudy + b->ubn, b->du, nx, nw);
                    }
                }
            }
        }
        free(bd);
    }
    const double t2 = omp_get_wtime();
    /* OpenMPBackend.cpp:169-182 */
    const float scale = 3.14159265358979323846f / (2.0f * (float)np);
#pragma omp parallel for num_threads(threads)
    for (int z = 0; z < nz; ++z) {
        const float cs = scale * spans[z];
        float* sp = out + (size_t)z * ss;
        for (size_t i = 0; i < ss; ++i) sp[i] = (sp[i] * cs) + min_hus[z];
    }
    const double t3 = omp_get_wtime();
    if (phase_ms) {
        phase_ms[0] = (t1 - t0) * 1e3;
        phase_ms[1] = (t2 - t1) * 1e3;
        phase_ms[2] = (t3 - t0) * 1e3;
    }
    free(du_dx); free(du_dy); free(u_b0); free(v_b);
    free(H); free(fall); free(min_hus); free(spans);
    return 0;
}
