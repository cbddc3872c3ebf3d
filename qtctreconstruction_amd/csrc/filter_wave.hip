// filter_wave.hip — row filter for detector rows of up to 512 bins as short
// circular convolutions, one wavefront per pair of rows.
//
// Reference semantics (SerialBackend.cpp:150-195): row x of `bins` samples
// sits at offset pad in a zero row of nfft points, is convolved circularly
// with h = IDFT_nfft(H) (H real and even, Nyquist bin 0), and the first sq
// outputs are kept:
//     out[i] = sum_e x[e] h[(i - pad - e) mod nfft],  i < sq,  e < bins.
// Only the offsets d = i - pad - e in [-D, D], D = max(pad + bins - 1,
// sq - 1 - pad), ever meet, and 2D + 1 < nfft. So the same outputs come from
// a circular convolution of any length L >= 2D + 1 with the kernel
// g[d mod L] = h[d] for |d| <= D (0 elsewhere): g is real and even, its
// spectrum Hs = DFT_L(g) is real, and it is designed on the host in FP64 from
// the reference's FP32 taps. With L = 512 R (R = 1, 2, 3), the input (bins <=
// 512 samples, placed at 0) fills one sub-block, so the length-L transform
// splits into R transforms of 512 points:
//     Z[R m + r] = DFT_512(z[e] W_L^(e r))[m],
//     y[n' + 512 t] = sum_r W_R^(-t r) W_L^(-n' r) IDFT_512(Y[R . + r])[n'],
// and out[i] = y[(i - pad) mod L]. At 512 bins (nfft 2048) this is 3 x 512
// points per direction instead of 2048, 27 radix-2 stage-passes instead of
// 44 over the row pair.
//
// One wavefront owns two rows of one angle (row A in the real part, row B in
// the imaginary part: g is real, so they come back separated) and keeps all R
// sub-transforms in registers, 8 points per lane. A 512-point transform is
// three radix-8 passes (strides 64, 8, 1) with two transposes through the
// wavefront's own LDS slots between them; the forward transform's last pass,
// the product with Hs and the inverse transform's first pass run back to back
// in registers. No workgroup barrier is needed: a wavefront's LDS
// instructions execute in order, and a wavefront-scope fence keeps the
// compiler from moving them across a transpose. Output is written into the
// z-innermost float4 layout the backprojection reads: the two wavefronts of
// a z-block write the x,y and z,w halves of the same bins.
#include "ct_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

namespace cthip {

namespace {

typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const float __attribute__((address_space(1))) g_cfloat;
typedef const f2v __attribute__((address_space(1))) g_cfloat2;
typedef const f4v __attribute__((address_space(1))) g_cfloat4;

constexpr int kWS = 72;           // LDS slots per transpose row group (64 + 8)
constexpr int kWSlots = 8 * kWS;  // one wavefront's transpose area, float2 slots

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}

// a * W_8^j (W_8 = exp(-2 pi i / 8)), or a * W_8^-j for INV.
template <int J, bool INV>
__device__ __forceinline__ float2 mul_w8(float2 a) {
    constexpr float s = 0.70710678118654752f;
    if constexpr (J == 0) return a;
    else if constexpr (J == 2) return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
    else if constexpr (J == 1)
        return INV ? make_float2(s * (a.x - a.y), s * (a.x + a.y))
                   : make_float2(s * (a.x + a.y), s * (a.y - a.x));
    else
        return INV ? make_float2(-s * (a.x + a.y), s * (a.x - a.y))
                   : make_float2(s * (a.y - a.x), -s * (a.x + a.y));
}

// x[k] <- sum_j x[j] W_8^(+-jk), natural order in and out: three radix-2
// decimation-in-frequency stages leave X[k] at position brev3(k), and the
// final renaming is free (compile-time register indices).
template <bool INV>
__device__ __forceinline__ void dft8(float2 (&x)[8]) {
    float2 t[8], u[8];
    t[0] = cadd(x[0], x[4]);
    t[4] = csub(x[0], x[4]);
    t[1] = cadd(x[1], x[5]);
    t[5] = mul_w8<1, INV>(csub(x[1], x[5]));
    t[2] = cadd(x[2], x[6]);
    t[6] = mul_w8<2, INV>(csub(x[2], x[6]));
    t[3] = cadd(x[3], x[7]);
    t[7] = mul_w8<3, INV>(csub(x[3], x[7]));
#pragma unroll
    for (int h = 0; h < 8; h += 4) {
        u[h] = cadd(t[h], t[h + 2]);
        u[h + 2] = csub(t[h], t[h + 2]);
        u[h + 1] = cadd(t[h + 1], t[h + 3]);
        u[h + 3] = mul_w8<2, INV>(csub(t[h + 1], t[h + 3]));
    }
    x[0] = cadd(u[0], u[1]);
    x[4] = csub(u[0], u[1]);
    x[2] = cadd(u[2], u[3]);
    x[6] = csub(u[2], u[3]);
    x[1] = cadd(u[4], u[5]);
    x[5] = csub(u[4], u[5]);
    x[3] = cadd(u[6], u[7]);
    x[7] = csub(u[6], u[7]);
}

// Orders one wavefront's LDS accesses across a transpose (compiler only: the
// hardware runs a wavefront's LDS instructions in issue order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace

// Wave-form tables, laid out lane-fastest so that a wavefront's reads are
// consecutive (global) or bank-conflict free (LDS). Twiddles, float2 entries:
//   [((r - 1) 8 + j) 64 + l]   W_L^((l + 64 j) r), r = 1 .. R-1 (pre-twiddles,
//                              conjugated as post-twiddles)
//   [1024 + 64 k + l]          W_512^(l k)   (pass over stride 64)
//   [1536 + 8 k + q]           W_64^(q k)    (pass over stride 8)
// Filter, float4 entries: [(2 r + c) 64 + lane] holds Hs[R (m0 + 8 m1 +
// 64 j) + r] / L for j = 4c .. 4c + 3, lane = 8 m0 + m1: the spectrum points
// a lane holds after the forward passes.
constexpr int kWaveTwPre = 0, kWaveTw1 = 1024, kWaveTw2 = 1536;

__device__ __forceinline__ float2 to2(float2 v) { return v; }
__device__ __forceinline__ float2 to2(f2v v) { return make_float2(v.x, v.y); }

// Table reads through pointers of either address space (TP: float2 / f2v
// entries, HP: f4v entries), indexed as laid out above.
template <class TP, class HP>
struct WaveTabs {
    TP tw;
    HP hs;
    __device__ float2 pre(int r, int j, int lane) const { return to2(tw[((r - 1) * 8 + j) * 64 + lane]); }
    __device__ float2 w1(int k, int lane) const { return to2(tw[kWaveTw1 + 64 * k + lane]); }
    __device__ float2 w2(int k, int lane) const { return to2(tw[kWaveTw2 + 8 * k + (lane & 7)]); }
    __device__ f4v h(int q, int lane) const { return hs[q * 64 + lane]; }
};

// The row pair of a wavefront (rows z, z + 1 of angle a, z = 4 zb + 2 half):
// wave_fetch_rows loads raw[j] = (A, B)[min(lane + 64 j, bins - 1)] of rows
// min(z, nz - 1) and min(z + 1, nz - 1) (every lane loads an in-bounds
// address, so no load sits under a branch and the compiler counts them
// exactly); wave_mask_rows then zeroes what lies past bins or past nz,
// giving zr[j] = (A + i B)[lane + 64 j]. The two are split so a prefetch's
// values are not needed (waited for) until the item that uses them.
__device__ __forceinline__ void wave_fetch_rows(const FiltArgs& p, int a, int zb, int half,
                                                float2 (&raw)[8]) {
    const int lane = threadIdx.x & 63;
    const int z = zb * kZB + 2 * half;
    g_cfloat* ra = (g_cfloat*)p.sino + ((size_t)min(z, p.nz - 1) * p.na + a) * p.bins;
    g_cfloat* rb = (g_cfloat*)p.sino + ((size_t)min(z + 1, p.nz - 1) * p.na + a) * p.bins;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int ec = min(lane + 64 * j, p.bins - 1);
        raw[j] = make_float2(ra[ec], rb[ec]);
    }
}
__device__ __forceinline__ void wave_mask_rows(const FiltArgs& p, int zb, int half,
                                               const float2 (&raw)[8], float2 (&zr)[8]) {
    const int lane = threadIdx.x & 63;
    const int z = zb * kZB + 2 * half;
    const bool za = z < p.nz, zb1 = z + 1 < p.nz;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool in = lane + 64 * j < p.bins;
        zr[j] = make_float2(in && za ? raw[j].x : 0.0f, in && zb1 ? raw[j].y : 0.0f);
    }
}

struct NoPrefetch {
    __device__ void operator()() const {}
};

// One row pair (zr from wave_fetch_rows and wave_mask_rows) through the
// short-transform filter; `half` selects the float4 half written. Table
// reads sit right before their use (behind the transposes' fences, so they
// are not hoisted and held in registers). `prefetch` runs mid-item: the next
// row pair's loads then have half an item to land.
template <int R, class TP, class HP, class PF>
__device__ __forceinline__ void wave_filter_rows(const FiltArgs& p, int a, int zb, int half,
                                                 float2* __restrict__ X, const float2 (&zr)[8],
                                                 const WaveTabs<TP, HP>& T, const PF& prefetch) {
    const int lane = threadIdx.x & 63;
    float2 d[R][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) d[0][j] = zr[j];
    // sub-transform r starts from z[e] W_L^(e r)
#pragma unroll
    for (int r = 1; r < R; ++r)
#pragma unroll
        for (int j = 0; j < 8; ++j) d[r][j] = cmul(d[0][j], T.pre(r, j, lane));
    float2 w[8];
#pragma unroll
    for (int k = 1; k < 8; ++k) w[k] = T.w1(k, lane);
    const int hi = lane >> 3, lo = lane & 7;
    // Transposes, element (g, a, b) of a wavefront's 512 at LDS slot
    //   T1 (stride-64 groups <-> stride-8 groups): 72 g + 8 a + b
    //   T2 (stride-8 groups <-> stride-1 groups):  72 g + 9 a + b
    // In every access the 32 lanes of a half-wave hit 32 distinct 8-byte bank
    // pairs, and the register index enters as an immediate offset.
    float2* const t1w = X + lane;               // + 72 k: lane l, index k
    float2* const t1r = X + hi * kWS + lo;      // + 8 k: lane (g, b), index a = k
    float2* const t2w = X + hi * kWS + lo;      // + 9 k: lane (g, b), index a = k
    float2* const t2r = X + hi * kWS + 9 * lo;  // + k: lane (g, a), index b = k
    // forward pass over stride 64: lane l holds e = l + 64 j; out index m0
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dft8<false>(d[r]);
#pragma unroll
        for (int k = 1; k < 8; ++k) d[r][k] = cmul(d[r][k], w[k]);
    }
    // T1: lane l, index m0 -> lane (m0, e0), index e1 (l = e0 + 8 e1)
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) t1w[kWS * k] = d[r][k];
        wave_sync();
#pragma unroll
        for (int k = 0; k < 8; ++k) d[r][k] = t1r[8 * k];
        wave_sync();
    }
#pragma unroll
    for (int k = 1; k < 8; ++k) w[k] = T.w2(k, lane);
    // forward pass over stride 8: out index m1, twiddle W_64^(e0 m1)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dft8<false>(d[r]);
#pragma unroll
        for (int k = 1; k < 8; ++k) d[r][k] = cmul(d[r][k], w[k]);
    }
    // T2: lane (m0, e0), index m1 -> lane (m0, m1), index e0
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) t2w[9 * k] = d[r][k];
        wave_sync();
#pragma unroll
        for (int k = 0; k < 8; ++k) d[r][k] = t2r[k];
        wave_sync();
    }
    prefetch();
    // forward pass over stride 1 (index m2: Z_r[m0 + 8 m1 + 64 m2]), the
    // filter, and the inverse pass over m2 (out index n0)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dft8<false>(d[r]);
        const f4v h0 = T.h(2 * r, lane), h1 = T.h(2 * r + 1, lane);
        const float hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) d[r][k] = make_float2(d[r][k].x * hv[k], d[r][k].y * hv[k]);
        dft8<true>(d[r]);
    }
    // T2 back: lane (m0, m1), index n0 -> lane (m0, n0), index m1
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) t2r[k] = d[r][k];
        wave_sync();
#pragma unroll
        for (int k = 0; k < 8; ++k) d[r][k] = t2w[9 * k];
        wave_sync();
    }
    // inverse pass over m1: twiddle W_64^(-n0 m1), out index n1
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 1; k < 8; ++k) d[r][k] = cmulc(d[r][k], w[k]);
        dft8<true>(d[r]);
    }
    // T1 back: lane (m0, n0), index n1 -> lane n0 + 8 n1, index m0
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) t1r[8 * k] = d[r][k];
        wave_sync();
#pragma unroll
        for (int k = 0; k < 8; ++k) d[r][k] = t1w[kWS * k];
        wave_sync();
    }
    // inverse pass over m0: twiddle W_512^(-l m0), out index n2:
    // u_r[l + 64 n2]
#pragma unroll
    for (int k = 1; k < 8; ++k) w[k] = T.w1(k, lane);
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 1; k < 8; ++k) d[r][k] = cmulc(d[r][k], w[k]);
        dft8<true>(d[r]);
    }
    // post-twiddles W_L^(-n' r), the R-point inverse DFT, and the outputs
    // i = (n' + 512 t + pad) mod L that fall in [0, sq). The stores go through
    // a buffer resource spanning this item's output row: a lane whose i is
    // past sq stores at an offset past the resource's range, which the
    // hardware drops, so every store instruction is unconditional.
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(p.filt + ((size_t)zb * p.na + a) * p.sq), (short)0, p.sq * 16, 0x00020000);
    const int L = 512 * R;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int n = lane + 64 * j;
        float2 y[R];
        if constexpr (R == 1) {
            y[0] = d[0][j];
        } else {
            const float2 v1 = cmulc(d[1][j], T.pre(1, j, lane));
            if constexpr (R == 2) {
                y[0] = cadd(d[0][j], v1);
                y[1] = csub(d[0][j], v1);
            } else {
                const float2 v2 = cmulc(d[2][j], T.pre(2, j, lane));
                constexpr float c3 = 0.86602540378443865f;  // sin(2 pi / 3)
                const float2 s = cadd(v1, v2), df = csub(v1, v2);
                const float2 m = make_float2(fmaf(-0.5f, s.x, d[0][j].x), fmaf(-0.5f, s.y, d[0][j].y));
                y[0] = cadd(d[0][j], s);
                y[1] = make_float2(fmaf(-c3, df.y, m.x), fmaf(c3, df.x, m.y));
                y[2] = make_float2(fmaf(c3, df.y, m.x), fmaf(-c3, df.x, m.y));
            }
        }
#pragma unroll
        for (int t = 0; t < R; ++t) {
            int i = n + 512 * t + p.pad;
            if (i >= L) i -= L;
            const int off = i < p.sq ? i * 16 + half * 8 : -1;
            typedef unsigned u2v __attribute__((ext_vector_type(2)));
            const u2v bits = {__float_as_uint(y[t].x), __float_as_uint(y[t].y)};
            __builtin_amdgcn_raw_buffer_store_b64(bits, rs, off, 0, 0);
        }
    }
}

// Per-launch form: blocks of four wavefronts take row pairs ("halves" of
// the items (angle, z-block), h -> item h / 2, half h mod 2; a = item mod na)
// in a grid-stride loop. The tables are staged in LDS once per block, so the
// only vector-memory loads in the loop are the rows, and each wavefront
// loads its next row pair while it filters the current one (vector loads
// complete in issue order, so nothing later in the item waits on them).
template <int R>
__global__ __launch_bounds__(256) void row_filter_wave_kernel(FiltArgs p, int halves) {
    __shared__ float2 xs[4][kWSlots];
    __shared__ float2 stw[kWaveTabEntries];
    __shared__ f4v shs[64 * 2 * 3];
    g_cfloat2* gt = (g_cfloat2*)p.tw;
    g_cfloat4* gh = (g_cfloat4*)p.H;
    for (int i = threadIdx.x; i < kWaveTabEntries; i += 256) stw[i] = to2(gt[i]);
    for (int i = threadIdx.x; i < 64 * 2 * R; i += 256) shs[i] = gh[i];
    // wave-uniform (the compiler cannot tell from threadIdx alone): row and
    // output addresses stay scalar
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int stride = (int)gridDim.x * 4;
    int h = (int)blockIdx.x * 4 + w;
    float2 nxt[8];
    if (h < halves) wave_fetch_rows(p, (h >> 1) % p.na, (h >> 1) / p.na, h & 1, nxt);
    __syncthreads();
    const WaveTabs<const float2*, const f4v*> T{stw, shs};
    for (; h < halves; h += stride) {
        const int item = h >> 1, zb = item / p.na;
        float2 cur[8];
        wave_mask_rows(p, zb, h & 1, nxt, cur);
        // the last pass loads its own rows again, so the load is unconditional
        const int nh = h + stride < halves ? h + stride : h;
        auto pf = [&]() { wave_fetch_rows(p, (nh >> 1) % p.na, (nh >> 1) / p.na, nh & 1, nxt); };
        wave_filter_rows<R>(p, item - zb * p.na, zb, h & 1, xs[w], cur, T, pf);
    }
}

// Persistent form: items (job, angle, z-block) of many volumes whose rows
// all take the short transform of one length 512 R, from an atomic counter
// until the list is exhausted (every block then returns). Jobs differ in
// their filters, so the tables are read from global memory.
template <int R>
__global__ __launch_bounds__(128) void row_filter_wave_persistent_kernel(
    const FiltJob* __restrict__ jobs, const int* __restrict__ hdr, int* __restrict__ counter) {
    __shared__ float2 lds[2][kWSlots];
    __shared__ int item_s;
    const int half = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    jobs += hdr[0];
    const int njobs = hdr[1], total = hdr[2];
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if (item >= total) break;
        int j = 0;
        while (j + 1 < njobs && jobs[j + 1].item_begin <= item) ++j;
        const FiltJob& jb = jobs[j];
        const int local = item - jb.item_begin;
        const int zb = local / jb.a.na, a = local - zb * jb.a.na;
        float2 raw[8], zr[8];
        wave_fetch_rows(jb.a, a, zb, half, raw);
        wave_mask_rows(jb.a, zb, half, raw, zr);
        const WaveTabs<g_cfloat2*, g_cfloat4*> T{(g_cfloat2*)jb.a.tw, (g_cfloat4*)jb.a.H};
        wave_filter_rows<R>(jb.a, a, zb, half, lds[half], zr, T, NoPrefetch());
    }
}

// ---------------------------------------------------------------- host side

int filter_kernel_reach(int bins, int sq, int pad) {
    return std::max(pad + bins - 1, sq - 1 - pad);
}

int wave_filter_r(int bins, int sq, int pad, int nfft) {
    if (bins < 1 || bins > kWaveM || nfft < 512) return 0;
    const int D = filter_kernel_reach(bins, sq, pad);
    const int r = (2 * D + 1 + kWaveM - 1) / kWaveM;
    return r <= 3 ? r : 0;
}

namespace {
constexpr double kPi = 3.14159265358979323846;

// h[d] = IDFT_nfft(H)[d] for d = 0 .. D, in FP64, from the reference's FP32
// taps H (real and even; the Nyquist bin taps[nfft/2] is taken as 0, as
// Serial's inverse transform drops it).
std::vector<double> kernel_offsets(int nfft, const float* taps, int D) {
    std::vector<double> cn(nfft), h(D + 1);
    for (int i = 0; i < nfft; ++i) cn[i] = std::cos(2.0 * kPi * i / nfft);
    for (int d = 0; d <= D; ++d) {
        double s = (double)taps[0];
        for (int j = 1; j < nfft / 2; ++j) s += 2.0 * (double)taps[j] * cn[((long long)j * d) % nfft];
        h[d] = s / nfft;
    }
    return h;
}

// Hs[k] = sum_{|d| <= D} h[d] W_L^(k d) for the real, even kernel g of
// length L (cos table cl of L entries), in FP64.
double short_spectrum_at(const std::vector<double>& h, const std::vector<double>& cl, long long k,
                         int L) {
    double s = h[0];
    const int D = (int)h.size() - 1;
    for (int d = 1; d <= D; ++d) s += 2.0 * h[d] * cl[(k * d) % L];
    return s;
}

std::vector<double> cos_table(int L) {
    std::vector<double> cl(L);
    for (int i = 0; i < L; ++i) cl[i] = std::cos(2.0 * kPi * i / L);
    return cl;
}
}  // namespace

void short_filter_taps(int bins, int sq, int pad, int nfft, const float* taps, int L, float* H) {
    const int D = filter_kernel_reach(bins, sq, pad);
    const std::vector<double> h = kernel_offsets(nfft, taps, D);
    const std::vector<double> cl = cos_table(L);
    for (int k = 0; k <= L / 2; ++k) H[k] = (float)short_spectrum_at(h, cl, k, L);
}

void wave_filter_tables(int bins, int sq, int pad, int nfft, const float* taps, float2* tw,
                        float* hs) {
    const int R = wave_filter_r(bins, sq, pad, nfft);
    const int L = kWaveM * R;
    auto W = [&](long long k, long long n) {
        const double a = -2.0 * kPi * (double)(k % n) / (double)n;
        return make_float2((float)std::cos(a), (float)std::sin(a));
    };
    for (int i = 0; i < kWaveTabEntries; ++i) tw[i] = make_float2(0.0f, 0.0f);
    for (int l = 0; l < 64; ++l) {
        for (int r = 1; r < R; ++r)
            for (int j = 0; j < 8; ++j)
                tw[kWaveTwPre + ((r - 1) * 8 + j) * 64 + l] = W((long long)(l + 64 * j) * r, L);
        for (int k = 0; k < 8; ++k) tw[kWaveTw1 + 64 * k + l] = W((long long)l * k, 512);
    }
    for (int q = 0; q < 8; ++q)
        for (int k = 0; k < 8; ++k) tw[kWaveTw2 + 8 * k + q] = W((long long)q * k, 64);
    // the short kernel's spectrum at the points each lane holds, over L
    const std::vector<double> h = kernel_offsets(nfft, taps, filter_kernel_reach(bins, sq, pad));
    const std::vector<double> cl = cos_table(L);
    for (int lane = 0; lane < 64; ++lane) {
        const int m0 = lane >> 3, m1 = lane & 7;
        for (int r = 0; r < R; ++r)
            for (int j = 0; j < 8; ++j) {
                const long long k = (long long)R * (m0 + 8 * m1 + 64 * j) + r;
                hs[((2 * r + j / 4) * 64 + lane) * 4 + (j & 3)] =
                    (float)(short_spectrum_at(h, cl, k, L) / L);
            }
    }
}

// Blocks of the grid-stride launch: every block the device holds at once
// (per CU, from the occupancy of the instance), so each block takes several
// row pairs and its prefetch pays.
template <int R>
static int wave_grid_r(int device) {
    static std::atomic<int> cached[64] = {};
    if (device >= 0 && device < 64) {
        const int g = cached[device].load(std::memory_order_relaxed);
        if (g > 0) return g;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)row_filter_wave_kernel<R>,
                                                     256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    const int g = (per_cu < 1 ? 1 : per_cu) * cus;
    if (device >= 0 && device < 64) cached[device].store(g, std::memory_order_relaxed);
    return g;
}

hipError_t launch_row_filter_wave(const FiltArgs& p, hipStream_t s) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    // row pairs: two per (angle, z-block) item
    const long long halves_ll = 2LL * p.na * ((p.nz + kZB - 1) / kZB);
    if (halves_ll > 0x7fffffff) return hipErrorInvalidValue;
    const int halves = (int)halves_ll;
    int g = 0;
    switch (p.wave_r) {
        case 1: g = wave_grid_r<1>(device); break;
        case 2: g = wave_grid_r<2>(device); break;
        case 3: g = wave_grid_r<3>(device); break;
        default: return hipErrorInvalidValue;
    }
    if (g <= 0) return hipErrorInvalidValue;
    const dim3 grid(std::min(g, (halves + 3) / 4));
    switch (p.wave_r) {
        case 1: hipLaunchKernelGGL(row_filter_wave_kernel<1>, grid, dim3(256), 0, s, p, halves); break;
        case 2: hipLaunchKernelGGL(row_filter_wave_kernel<2>, grid, dim3(256), 0, s, p, halves); break;
        default: hipLaunchKernelGGL(row_filter_wave_kernel<3>, grid, dim3(256), 0, s, p, halves); break;
    }
    return hipGetLastError();
}

template <int R>
static int wave_persistent_grid_r(int device) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)row_filter_wave_persistent_kernel<R>, 128, 0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return (per_cu < 1 ? 1 : per_cu) * cus;
}

int filter_wave_persistent_grid(int r, int device) {
    switch (r) {
        case 1: return wave_persistent_grid_r<1>(device);
        case 2: return wave_persistent_grid_r<2>(device);
        case 3: return wave_persistent_grid_r<3>(device);
        default: return 0;
    }
}

hipError_t launch_row_filter_wave_persistent(const FiltJob* d_jobs, const int* d_hdr,
                                             int* d_counter, int r, int grid, hipStream_t s) {
    switch (r) {
        case 1: hipLaunchKernelGGL(row_filter_wave_persistent_kernel<1>, dim3(grid), dim3(128), 0, s, d_jobs, d_hdr, d_counter); break;
        case 2: hipLaunchKernelGGL(row_filter_wave_persistent_kernel<2>, dim3(grid), dim3(128), 0, s, d_jobs, d_hdr, d_counter); break;
        case 3: hipLaunchKernelGGL(row_filter_wave_persistent_kernel<3>, dim3(grid), dim3(128), 0, s, d_jobs, d_hdr, d_counter); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cthip
