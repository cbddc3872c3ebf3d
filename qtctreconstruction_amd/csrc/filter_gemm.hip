// filter_gemm.hip — the row filter as a Toeplitz GEMM on gfx950 MFMA.
//
// The padded FFT filter of SerialBackend.cpp:150-195 is exactly a linear map
// of each row: q[i] = sum_j h[(i - j - pad) mod nfft] p[j] for i < sq and
// j < bins, where h is the real-space kernel of the filter (DC and Nyquist
// removed, the effective Serial filter). Stacking every (slice, angle) row
// gives Q[M x sq] = P[M x bins] * T^T, a GEMM with T[n][k] = h[(n-k-pad) mod nfft].
//
// Accuracy: the kernel sums to zero (DC removed) and sinogram rows are
// smooth, so the terms h*p cancel heavily and their input rounding is
// amplified. The GEMM is therefore taken by parts: with dp[t] = p[t] - p[t-1]
// (t = 0..bins, p outside [0,bins) = 0, so sum dp = 0) and the cumulative
// kernel G[m] = sum_{u=-n/2}^{m} h[u], q[i] = sum_t G[i-t-pad] dp[t]
// exactly. G decays like 1/m and dp is small where p is smooth, which cuts
// the input-rounding error ~50x (BF16 1.2e-1 -> 2.4e-3 relative L2, split
// BF16 1.8e-4 -> 3.8e-6 on 512^2 x 720 sinograms).
//
// Variants (all accumulate in FP32):
//   1  BF16:        Q = bf16(P) * bf16(T)
//   2  split BF16:  Q = Ph*Th + Ph*Tl + Pl*Th, with X = Xh + Xl, Xh = bf16(X),
//                   Xl = bf16(X - Xh)  (three products; error ~ 2^-16)
//   3  FP8 E4M3:    per-row scale s_r = 448 / max|P_r|, one scale for T;
//                   Q = (fp8(P*s_r) * fp8(T*s_T)) / (s_r s_T)
//
// Work split: a block owns 16 angles x 4 slices of one z-block (4 MFMA row
// groups, one per slice) and 256 output bins; each of its 4 waves owns 64 bins
// (4 MFMA 16x16 tiles); a launch block holds up to 3 such column blocks. The
// A operand (the rows' differences, rounded to the variant's format) is
// converted once per block and K-chunk into LDS and read by every wave; the
// B fragments come from T in L2. MFMA shapes: v_mfma_f32_16x16x32_bf16 and
// v_mfma_f32_16x16x32_fp8_fp8. Per lane l (c = l%16, g = l/16): A holds row c,
// k = 8g..8g+7; B holds column c, k = 8g..8g+7; D holds rows 4g..4g+3 of
// column c. Because the 4 slice tiles of a (row, column) land in the same
// lane, the epilogue writes float4 = 4 slices per bin, the layout the
// backprojection stages from ([z/4][angle][sq][4]).
//
// The Toeplitz operand T is designed on the device in FP64 (H by direct DFT,
// h by inverse DFT), then rounded to the variant's format, and stays resident.
#include "ct_internal.h"

namespace cthip {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4v;

__device__ __forceinline__ unsigned short f2bf(float f) {
    // round to nearest even (inputs are finite)
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// ---- design: H in FP64 (the exact DFT of the reference's float spatial ramp
// taps, windowed), then h ----
__global__ void filter_design64_kernel(int n, int type, double* __restrict__ H) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nh = n / 2;
    if (k > nh) return;
    const double pi = 3.14159265358979323846;
    double acc = 0.25;
    for (int j = 1; j < n; j += 2) {
        const int m = (j < nh) ? j : n - j;
        const double t = 1.0 / (pi * (double)m);
        const double fj = (double)(float)(-1.0 * t * t);
        const long long r = ((long long)j * k) % n;
        acc += fj * cos(2.0 * pi * (double)r / (double)n);
    }
    double h = 2.0 * acc;
    const double freq = (double)k / (double)n;
    switch (type) {
        case 0: break;
        case 2: h *= 0.54 + 0.46 * cos(pi * freq / 0.5); break;
        case 3: h *= cos(pi * freq); break;
        case 4: h *= 0.5 + 0.5 * cos(pi * freq / 0.5); break;
        case 5: {
            const double r = fabs(freq) / 0.5;
            h *= 1.0 - (r < 1.0 ? r : 1.0);
        } break;
        case 1:
        default: {
            const double om = pi * freq;
            if (k > 0) h *= sin(om) / om;
        } break;
    }
    if (k == 0 || k == nh) h = 0.0;
    H[k] = h;
}

// G[m] = sum_{u=-n/2}^{m} h[u mod n], m in (-n/2, n/2], stored at G[m + n/2]
__global__ void cumulative_kernel64(int n, const double* __restrict__ h, double* __restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // m = i - n/2
    if (i >= n) return;
    double acc = 0.0;
    for (int u = 0; u <= i; ++u) acc += h[(u - n / 2 + n) % n];
    G[i] = acc;
}

// h[m] = (2/n) sum_{k=1}^{n/2-1} H[k] cos(2 pi k m / n)   (H[0] = H[n/2] = 0)
__global__ void kernel_from_filter64(int n, const double* __restrict__ H, double* __restrict__ h) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double pi = 3.14159265358979323846;
    double acc = 0.0;
    for (int k = 1; k < n / 2; ++k) {
        const long long r = ((long long)k * m) % n;
        acc += H[k] * cos(2.0 * pi * (double)r / (double)n);
    }
    h[m] = 2.0 * acc / (double)n;
}

// T[n][k] = G[n-k-pad] for n < sq, k <= bins (Np x Kp, k contiguous):
// hi/lo bf16 and fp32 copies
__global__ void toeplitz_build_kernel(const double* __restrict__ G, int nfft, int sq, int bins,
                                      int pad, int Np, int Kp, unsigned short* __restrict__ Thi,
                                      unsigned short* __restrict__ Tlo, float* __restrict__ T32) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)Np * Kp) return;
    const int nn = (int)(idx / Kp), kk = (int)(idx - (size_t)nn * Kp);
    float v = 0.0f;
    if (nn < sq && kk <= bins) v = (float)G[nn - kk - pad + nfft / 2];
    const unsigned short hi = f2bf(v);
    Thi[idx] = hi;
    Tlo[idx] = f2bf(v - bf2f(hi));
    T32[idx] = v;
}

// max |dp| per row, dp[t] = p[t] - p[t-1], t = 0..cols
__global__ void absmax_rows_kernel(const float* __restrict__ x, size_t rows, int cols,
                                   float* __restrict__ out) {
    const size_t r = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    float m = 0.0f;
    if (r < rows)
        for (int c = lane; c <= cols; c += 64) {
            const float a = (c < cols) ? x[r * cols + c] : 0.0f;
            const float b = (c > 0) ? x[r * cols + c - 1] : 0.0f;
            m = fmaxf(m, fabsf(a - b));
        }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (r < rows && lane == 0) out[r] = m;
}

// single-block max over n values (T scale)
__global__ void absmax_all_kernel(const float* __restrict__ x, size_t n, float* __restrict__ out) {
    float m = 0.0f;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) m = fmaxf(m, fabsf(x[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float s[16];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmaxf(m, s[i]);
        out[0] = m;
    }
}

// T quantised to FP8 E4M3 with scale 448/max|T|
__global__ void toeplitz_fp8_kernel(const float* __restrict__ T32, size_t n,
                                    const float* __restrict__ tmax, unsigned char* __restrict__ T8) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    if (i >= n) return;
    const float s = (tmax[0] > 0.0f) ? 448.0f / tmax[0] : 1.0f;
    const float a = T32[i] * s, b = (i + 1 < n) ? T32[i + 1] * s : 0.0f;
    const int p = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    T8[i] = (unsigned char)(p & 0xff);
    if (i + 1 < n) T8[i + 1] = (unsigned char)((p >> 8) & 0xff);
}

// The A operand of one K-chunk (32 dp values of each of the block's 4 slices
// x 16 angles) is converted once per block into LDS: wave w converts slice
// w, lane (c, g) the fragment row c, k = 8g..8g+7 that lane (c, g) of every
// wave reads back. Rows are kArow uint4 apart (4 fragments + 1 of padding:
// 80 bytes, so the 16 lanes of a fragment column land on distinct b128 slots,
// 5c + g mod 16). FP8 fragments use the first 8 bytes of a slot.
constexpr int kArow = 5;

// p[kb-1 .. kb+7] of one row (zeros outside [0, bins) and for rows past the
// volume)
__device__ __forceinline__ void load_a_row(const float* __restrict__ prow, bool live, bool vec,
                                           int bins, int kb, float (&pv)[9]) {
    if (live && vec && kb + 8 <= bins) {
        const float4 q0 = *reinterpret_cast<const float4*>(prow + kb);
        const float4 q1 = *reinterpret_cast<const float4*>(prow + kb + 4);
        pv[0] = kb > 0 ? prow[kb - 1] : 0.0f;
        pv[1] = q0.x;
        pv[2] = q0.y;
        pv[3] = q0.z;
        pv[4] = q0.w;
        pv[5] = q1.x;
        pv[6] = q1.y;
        pv[7] = q1.z;
        pv[8] = q1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int k = kb - 1 + j;
            pv[j] = (live && k >= 0 && k < bins) ? prow[k] : 0.0f;
        }
    }
}

// the fragment dp[kb .. kb+7] in the variant's format
template <int VARIANT>
__device__ __forceinline__ void convert_a(const float (&pv)[9], float rsc, uint4& hi, uint4& lo) {
    float x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = pv[j + 1] - pv[j];  // dp[kb+j]
    if (VARIANT == 3) {
        int l = 0, h = 0;
        l = __builtin_amdgcn_cvt_pk_fp8_f32(x[0] * rsc, x[1] * rsc, l, false);
        l = __builtin_amdgcn_cvt_pk_fp8_f32(x[2] * rsc, x[3] * rsc, l, true);
        h = __builtin_amdgcn_cvt_pk_fp8_f32(x[4] * rsc, x[5] * rsc, h, false);
        h = __builtin_amdgcn_cvt_pk_fp8_f32(x[6] * rsc, x[7] * rsc, h, true);
        hi = make_uint4((unsigned)l, (unsigned)h, 0u, 0u);
    } else {
        // round to nearest even (v_cvt_pk_bf16_f32); the low part is the
        // rounded remainder
        bf16x8 ah, al;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            ah[j] = (__bf16)x[j];
            if (VARIANT == 2) al[j] = (__bf16)(x[j] - (float)ah[j]);
        }
        hi = __builtin_bit_cast(uint4, ah);
        if (VARIANT == 2) lo = __builtin_bit_cast(uint4, al);
    }
}

// B fragments of the wave's 4 column tiles for the chunk at kb
template <int VARIANT>
__device__ __forceinline__ void load_b(const unsigned short* __restrict__ Thi,
                                       const unsigned short* __restrict__ Tlo,
                                       const unsigned char* __restrict__ T8, int n0, int c, int Kp,
                                       int kb, bf16x8 (&bh)[4], bf16x8 (&bl)[4], long (&b8)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const size_t off = (size_t)(n0 + 16 * t + c) * Kp + kb;
        if (VARIANT == 3) {
            b8[t] = *reinterpret_cast<const long*>(T8 + off);
        } else {
            bh[t] = *reinterpret_cast<const bf16x8*>(Thi + off);
            if (VARIANT == 2) bl[t] = *reinterpret_cast<const bf16x8*>(Tlo + off);
        }
    }
}

// grid: kXcds * ngrp * ceil(nmb / kXcds) blocks of 256 nbw threads; nmb row
// blocks (z-block, 16 angles), ngrp groups of nbw column blocks of 256 output
// bins (nbw divides Np / 256). Every wave owns 64 output bins; waves 0-3 also
// convert the A chunk that all of them read. A row block's column groups run
// one after the other on one XCD (blocks kXcds apart share an XCD), so its
// sinogram rows are read from HBM once.
//
// BLDS: the B fragments come from LDS instead of T in L2. T is Toeplitz
// within n < sq, k <= bins (T[n][k] depends on k - n alone); outside it the
// products do not count (dp[k] = 0 for k > bins, columns n >= sq are not
// written), so any finite value serves there. The block's columns
// [nb0, nb0 + NB) and every k give k - n = i - (nb0 + NB - 1), i in [0, Lq),
// Lq = Kp + NB. The block copies the diagonal Q[i] from T (at n = max(0,
// -J), k = max(0, J), the same rounded values; 0 where no (n, k) of T's
// Toeplitz range has this diagonal) and expands it into the window table
// W[i] = Q[i .. i+7], one 16-byte (FP8: 8-byte) slot per start: lane (c, g)
// of column tile t reads slot kb - (n - nb0) + NB - 1, so the 16 lanes of a
// fragment column read 16 consecutive slots (no bank conflicts).
template <int VARIANT, bool BLDS>
__global__ __launch_bounds__(768) void toeplitz_gemm_kernel(
    const float* __restrict__ sino, int nz, int na, int bins, int sq, int Kp,
    const unsigned short* __restrict__ Thi, const unsigned short* __restrict__ Tlo,
    const unsigned char* __restrict__ T8, const float* __restrict__ rmax,
    const float* __restrict__ tmax, float4* __restrict__ filt, int ngrp, int nmb, int vec) {
    extern __shared__ uint4 Wdyn[];
    const int nbw = blockDim.x / 256;
    const int xcd = blockIdx.x % kXcds, rr = blockIdx.x / kXcds;
    const int cg = rr % ngrp, mb = (rr / ngrp) * kXcds + xcd;
    if (mb >= nmb) return;  // whole block
    const int nab = (na + 15) / 16;
    const int zb = mb / nab, a0 = (mb % nab) * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int n0 = cg * nbw * 256 + wave * 64;  // first output bin of this wave
    const bool stager = wave < kZB;
    // staging role (waves 0-3): slice `wave`, angle a0 + c
    const int zs = zb * kZB + wave, as = a0 + c;
    const bool slive = stager && zs < nz && as < na;
    const float* prow = sino + (slive ? ((size_t)zs * na + as) * bins : 0);
    float rsc = 1.0f;
    if (VARIANT == 3 && slive) {
        const float m = rmax[(size_t)zs * na + as];
        rsc = (m > 0.0f) ? 448.0f / m : 1.0f;
    }
    __shared__ uint4 Ah[2][kZB][16][kArow];
    __shared__ uint4 Al[VARIANT == 2 ? 2 : 1][kZB][16][kArow];
    const int NB = nbw * 256, nb0 = cg * NB, Lq = Kp + NB;
    if (BLDS) {
        // the diagonal, staged in the A buffers (not yet in use): hi in Ah,
        // lo in Al, FP8 bytes in Ah
        unsigned short* qh = reinterpret_cast<unsigned short*>(&Ah[0][0][0][0]);
        unsigned short* ql = reinterpret_cast<unsigned short*>(&Al[0][0][0][0]);
        unsigned char* q8 = reinterpret_cast<unsigned char*>(&Ah[0][0][0][0]);
        for (int i = threadIdx.x; i < Lq + 8; i += blockDim.x) {
            const int J = i - (nb0 + NB - 1);
            const bool in = J > -sq && J <= bins;
            const size_t at = in ? (size_t)(J < 0 ? -J : 0) * Kp + (J > 0 ? J : 0) : 0;
            if (VARIANT == 3) {
                q8[i] = in ? T8[at] : (unsigned char)0;
            } else {
                qh[i] = in ? Thi[at] : (unsigned short)0;
                if (VARIANT == 2) ql[i] = in ? Tlo[at] : (unsigned short)0;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < Lq; i += blockDim.x) {
            if (VARIANT == 3) {
                unsigned w0 = 0, w1 = 0;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    w0 |= (unsigned)q8[i + e] << (8 * e);
                    w1 |= (unsigned)q8[i + 4 + e] << (8 * e);
                }
                reinterpret_cast<uint2*>(Wdyn)[i] = make_uint2(w0, w1);
            } else {
                unsigned wh[4], wl[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    wh[e] = (unsigned)qh[i + 2 * e] | ((unsigned)qh[i + 2 * e + 1] << 16);
                    if (VARIANT == 2)
                        wl[e] = (unsigned)ql[i + 2 * e] | ((unsigned)ql[i + 2 * e + 1] << 16);
                }
                Wdyn[i] = make_uint4(wh[0], wh[1], wh[2], wh[3]);
                if (VARIANT == 2) Wdyn[Lq + i] = make_uint4(wl[0], wl[1], wl[2], wl[3]);
            }
        }
        __syncthreads();  // the A buffers are free again
    }

    f4v acc[kZB][4];
#pragma unroll
    for (int s = 0; s < kZB; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[s][t] = (f4v){0.f, 0.f, 0.f, 0.f};

    const int nch = Kp / 32;
    // The stager waves fetch chunk kc + 1's rows from memory before chunk
    // kc's products and convert them into the other A buffer after them, so
    // the fetch's latency passes under the products; one barrier per chunk
    // (the buffer written in round kc was last read in round kc - 1).
    if (stager) {
        float pv[9];
        uint4 h, l;
        load_a_row(prow, slive, vec != 0, bins, 8 * g, pv);
        convert_a<VARIANT>(pv, rsc, h, l);
        Ah[0][wave][c][g] = h;
        if (VARIANT == 2) Al[0][wave][c][g] = l;
    }
    __syncthreads();
    for (int kc = 0; kc < nch; ++kc) {
        const int buf = kc & 1;
        const int kb = kc * 32 + 8 * g;
        float pn[9];
        const bool next = stager && kc + 1 < nch;
        if (next) load_a_row(prow, slive, vec != 0, bins, kb + 32, pn);
        bf16x8 bh[4], bl[4];
        long b8[4];
        if (BLDS) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int i0 = kb - (n0 - nb0 + 16 * t + c) + NB - 1;
                if (VARIANT == 3) {
                    const uint2 q = reinterpret_cast<const uint2*>(Wdyn)[i0];
                    b8[t] = (long)q.x | ((long)q.y << 32);
                } else {
                    bh[t] = __builtin_bit_cast(bf16x8, Wdyn[i0]);
                    if (VARIANT == 2) bl[t] = __builtin_bit_cast(bf16x8, Wdyn[Lq + i0]);
                }
            }
        } else {
            load_b<VARIANT>(Thi, Tlo, T8, n0, c, Kp, kb, bh, bl, b8);
        }
#pragma unroll
        for (int s = 0; s < kZB; ++s) {
            if (VARIANT == 3) {
                const uint4 q = Ah[buf][s][c][g];
                const long av = (long)q.x | ((long)q.y << 32);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[s][t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(av, b8[t], acc[s][t], 0, 0, 0);
            } else {
                const bf16x8 ah = __builtin_bit_cast(bf16x8, Ah[buf][s][c][g]);
                bf16x8 al;
                if (VARIANT == 2) al = __builtin_bit_cast(bf16x8, Al[buf][s][c][g]);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    acc[s][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], acc[s][t], 0, 0, 0);
                    if (VARIANT == 2) {
                        acc[s][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[s][t], 0, 0, 0);
                        acc[s][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[s][t], 0, 0, 0);
                    }
                }
            }
        }
        if (next) {
            uint4 h, l;
            convert_a<VARIANT>(pn, rsc, h, l);
            Ah[buf ^ 1][wave][c][g] = h;
            if (VARIANT == 2) Al[buf ^ 1][wave][c][g] = l;
        }
        __syncthreads();
    }

    // epilogue: lane holds rows 4g..4g+3 (angles) of column c of every tile
    float tsc = 1.0f;
    if (VARIANT == 3) tsc = (tmax[0] > 0.0f) ? tmax[0] / 448.0f : 1.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int nn = n0 + 16 * t + c;
        if (nn >= sq) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int aa = a0 + 4 * g + r;
            if (aa >= na) continue;
            float v[kZB];
#pragma unroll
            for (int s = 0; s < kZB; ++s) {
                float q = acc[s][t][r];
                if (VARIANT == 3) {
                    const int z = zb * kZB + s;
                    const float m = (z < nz) ? rmax[(size_t)z * na + aa] : 0.0f;
                    q = q * ((m > 0.0f ? m / 448.0f : 1.0f) * tsc);
                }
                v[s] = q;
            }
            filt[((size_t)zb * na + aa) * sq + nn] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

hipError_t launch_toeplitz_design(int nfft, int type, int sq, int bins, int pad, int Np, int Kp,
                                  double* d_H, double* d_h, unsigned short* d_Thi,
                                  unsigned short* d_Tlo, float* d_T32, unsigned char* d_T8,
                                  float* d_tmax, hipStream_t s) {
    hipLaunchKernelGGL(filter_design64_kernel, dim3((nfft / 2 + 1 + 127) / 128), dim3(128), 0, s,
                       nfft, type, d_H);
    hipLaunchKernelGGL(kernel_from_filter64, dim3((nfft + 127) / 128), dim3(128), 0, s, nfft, d_H,
                       d_h);
    double* d_G = d_h + nfft;  // d_h holds 2*nfft doubles: h, then G
    hipLaunchKernelGGL(cumulative_kernel64, dim3((nfft + 127) / 128), dim3(128), 0, s, nfft, d_h,
                       d_G);
    const size_t n = (size_t)Np * Kp;
    hipLaunchKernelGGL(toeplitz_build_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       d_G, nfft, sq, bins, pad, Np, Kp, d_Thi, d_Tlo, d_T32);
    hipLaunchKernelGGL(absmax_all_kernel, dim3(1), dim3(1024), 0, s, d_T32, n, d_tmax);
    hipLaunchKernelGGL(toeplitz_fp8_kernel, dim3((unsigned)((n / 2 + 256) / 256)), dim3(256), 0, s,
                       d_T32, n, d_tmax, d_T8);
    return hipGetLastError();
}

hipError_t launch_row_absmax(const float* d_sino, size_t rows, int cols, float* d_rmax,
                             hipStream_t s) {
    hipLaunchKernelGGL(absmax_rows_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s,
                       d_sino, rows, cols, d_rmax);
    return hipGetLastError();
}

hipError_t launch_toeplitz_gemm(int variant, const float* d_sino, int nz, int na, int bins,
                                int sq, int Np, int Kp, const unsigned short* d_Thi,
                                const unsigned short* d_Tlo, const unsigned char* d_T8,
                                const float* d_rmax, const float* d_tmax, float4* d_filt,
                                hipStream_t s) {
    const int nzb = (nz + kZB - 1) / kZB;
    const int nnb = Np / 256, nmb = nzb * ((na + 15) / 16);
    const int nbw = nnb % 3 == 0 ? 3 : (nnb % 2 == 0 ? 2 : 1);  // column blocks per block
    const int ngrp = nnb / nbw;
    const dim3 grid((unsigned)(kXcds * ngrp * ((nmb + kXcds - 1) / kXcds))), block(256 * nbw);
    const int vec = (bins % 4 == 0) && ((uintptr_t)d_sino % 16 == 0);
    // window table of the B fragments: Lq slots of 16 bytes per bf16 part (8
    // for FP8), beside the 20 KB of A buffers, within 64 KB of LDS
    const size_t Lq = (size_t)Kp + 256 * nbw;
    const size_t wbytes = Lq * (variant == 2 ? 32 : (variant == 1 ? 16 : 8));
    const bool blds = wbytes <= 44 * 1024;
    const size_t dyn = blds ? wbytes : 0;
#define CT_GEMM_LAUNCH(V)                                                                       \
    if (blds)                                                                                   \
        hipLaunchKernelGGL((toeplitz_gemm_kernel<V, true>), grid, block, dyn, s, d_sino, nz, na, \
                           bins, sq, Kp, d_Thi, d_Tlo, d_T8, d_rmax, d_tmax, d_filt, ngrp, nmb,  \
                           vec);                                                                \
    else                                                                                        \
        hipLaunchKernelGGL((toeplitz_gemm_kernel<V, false>), grid, block, 0, s, d_sino, nz, na,  \
                           bins, sq, Kp, d_Thi, d_Tlo, d_T8, d_rmax, d_tmax, d_filt, ngrp, nmb,  \
                           vec);
    switch (variant) {
        case 1:
            CT_GEMM_LAUNCH(1)
            break;
        case 2:
            CT_GEMM_LAUNCH(2)
            break;
        case 3:
            CT_GEMM_LAUNCH(3)
            break;
        default:
            return hipErrorInvalidValue;
    }
#undef CT_GEMM_LAUNCH
    return hipGetLastError();
}

}  // namespace cthip
