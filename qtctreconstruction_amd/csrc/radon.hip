// radon.hip — per-slice normalisation and the ray-driven forward projection.
//
// The reference's oracle (SerialBackend::computeSinogram, SerialBackend.cpp:59-136)
// is a pixel-driven scatter: for every angle it walks the slice in raster order
// (y outer, x inner), skips pixels whose raw value is exactly 0, normalises
// v = (raw - min) * (1/span), and deposits v*(1-frac) / v*frac into bins
// i0 = floor(s), i0+1 (full v when i0 == bins-1, nothing when i0 < 0).
//
// Here every lane owns one unit interval [i, i+1) of s at one angle and
// gathers: it visits, row by row, exactly the pixels whose floor(s) is i,
// recomputing s with Serial's float expression, so every pixel's bin and
// fraction are Serial's. The lane sums the pixels' first deposits (for bin
// i) and their second ones (for bin i+1) apart, with fused multiply-adds,
// and bin i is the first sum of lane i plus the second of lane i-1: no
// atomics. A bin with a single deposit is Serial's rounded product bit for
// bit; sums of several differ from Serial's raster-order sums by rounding
// only. The candidate pixels come from a conservative interval in s
// (+-1/32 of a bin); the exact floor(s) test decides.
//
// Eight slices share one lane: the geometry is slice-independent, so the index
// work is done once and the eight values arrive as two float4 loads. Angles
// with |sin| <= |cos| walk the rows of the row-major copy; the others walk
// the rows of the transposed copy, so that adjacent lanes read adjacent
// pixels in both cases.
#include "ct_internal.h"

#include <algorithm>

namespace cthip {

// Order-preserving map of floats to unsigned keys (for -0 < +0 too; the
// extremes then differ from Serial's only in the sign of a zero, which every
// later use treats alike). Serial's scan (SerialBackend.cpp:69-76) starts
// from FLT_MAX / lowest and compares with < and >, so a NaN never becomes an
// extreme: kmin/kmax send NaN to the keys of the other end, which lose every
// comparison, and an all-NaN (or empty) slice keeps FLT_MAX / -FLT_MAX.
__host__ __device__ __forceinline__ unsigned f2key(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ unsigned kmin_of(float f) { return f != f ? 0xffffffffu : f2key(f); }
__device__ __forceinline__ unsigned kmax_of(float f) { return f != f ? 0u : f2key(f); }
constexpr unsigned kKeyFltMax = 0xff7fffffu;   // f2key(FLT_MAX)
constexpr unsigned kKeyFltLowest = 0x00800000u;  // f2key(-FLT_MAX)

__global__ void minmax_init_kernel(unsigned* keys, int nz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nz) {
        keys[2 * i] = kKeyFltMax;
        keys[2 * i + 1] = kKeyFltLowest;
    }
}

// grid (chunks, nz): block-reduce a chunk of one slice, then one atomic each.
// vec4: every slice starts 16-byte aligned and holds a multiple of 4 values,
// so the slice is read as float4.
__global__ __launch_bounds__(256) void minmax_kernel(const float* __restrict__ vol,
                                                     size_t slice_elems, int vec4,
                                                     unsigned* keys) {
    const int z = blockIdx.y;
    const float* p = vol + (size_t)z * slice_elems;
    unsigned kmin = 0xffffffffu, kmax = 0u;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t first = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (vec4) {
        const float4* p4 = reinterpret_cast<const float4*>(p);
        const size_t n4 = slice_elems / 4;
        size_t i = first;
        // four loads in flight per thread
        for (; i + 3 * stride < n4; i += 4 * stride) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = p4[i + u * stride];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                kmin = min(kmin, min(min(kmin_of(q[u].x), kmin_of(q[u].y)),
                                     min(kmin_of(q[u].z), kmin_of(q[u].w))));
                kmax = max(kmax, max(max(kmax_of(q[u].x), kmax_of(q[u].y)),
                                     max(kmax_of(q[u].z), kmax_of(q[u].w))));
            }
        }
        for (; i < n4; i += stride) {
            const float4 q = p4[i];
            kmin = min(kmin, min(min(kmin_of(q.x), kmin_of(q.y)), min(kmin_of(q.z), kmin_of(q.w))));
            kmax = max(kmax, max(max(kmax_of(q.x), kmax_of(q.y)), max(kmax_of(q.z), kmax_of(q.w))));
        }
    } else {
        for (size_t i = first; i < slice_elems; i += stride) {
            kmin = min(kmin, kmin_of(p[i]));
            kmax = max(kmax, kmax_of(p[i]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        kmin = min(kmin, (unsigned)__shfl_xor((int)kmin, off));
        kmax = max(kmax, (unsigned)__shfl_xor((int)kmax, off));
    }
    __shared__ unsigned smin[4], smax[4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smin[wv] = kmin;
        smax[wv] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; ++i) {
            kmin = min(kmin, smin[i]);
            kmax = max(kmax, smax[i]);
        }
        atomicMin(&keys[2 * z], kmin);
        atomicMax(&keys[2 * z + 1], kmax);
    }
}

__global__ void minmax_decode_kernel(const unsigned* keys, int nz, float* mm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * nz) mm[i] = key2f(keys[i]);
}

// One block normalises a tile of kNormTW x kNormTH pixels of one Radon
// z-block (kRZ = 8 slices; SerialBackend.cpp:78-80,104,106) and writes it to
// both copies. A pixel is 32 bytes (the 8 slices as 2 float4):
//   row copy     byte ((zb*h + y)*(w + kPadX) + x)*32
//   column copy  byte ((zb*w + x)*(h + kPadX) + y)*32
// The kPadX pixels past each row of either copy (x >= w in the row copy,
// y >= h in the column copy) are zeros, so a lane's fixed-size chunk of
// candidates may run past the edge of the image and read only zeros.
// Each thread reads 4 pixels of one row from each slice (float4 loads when
// vec4: w % 4 == 0 and a 16-byte aligned volume) and puts them in an LDS
// tile; both copies are then written from the tile with consecutive lanes
// on consecutive 16-byte units: 1 KB contiguous per wave store in the row
// copy, a column's 512 bytes per 32 lanes in the column copy. Tile rows are padded by one
// float4, so the column reads (16 rows apart) spread over the banks.
// rflag[zb][y] = {max(x + 1), max(w - x)} over the nonzero normalised pixels
// (x, y) of the z-block, and cflag[zb][x] = {max(y + 1), max(h - y)} over
// those of column x (both zero beforehand; {0, 0} for a row or column of
// zeros): the nonzero extent of each row and column. Zero pixels add nothing
// to any bin (SerialBackend.cpp:104 skips them): the 16-slice walk skips the
// rows of zeros, and the waves whose candidates all lie outside a row's
// extent.
constexpr int kNormTW = 64, kNormTH = 16;
constexpr int kNormRow = kNormTW * 2 + 1;  // float4 per padded tile row
__global__ __launch_bounds__(256) void normalise_kernel(const float* __restrict__ vol, int w,
                                                        int h, int nz,
                                                        const float* __restrict__ mm,
                                                        int vec4, float4* __restrict__ rowL,
                                                        float4* __restrict__ colL,
                                                        int* __restrict__ rflag,
                                                        int* __restrict__ cflag) {
    __shared__ float4 tile[kNormTH * kNormRow];
    __shared__ int rowe[kNormTH][2], cole[kNormTW][2];
    if (threadIdx.x < 2 * kNormTH) (&rowe[0][0])[threadIdx.x] = 0;
    if (threadIdx.x < 2 * kNormTW) (&cole[0][0])[threadIdx.x] = 0;
    __syncthreads();
    const int zb = blockIdx.z;
    const int bx = blockIdx.x * kNormTW, by = blockIdx.y * kNormTH;
    const int wp = w + kPadX, hp = h + kPadX;
    const int t = threadIdx.x;
    float mn[kRZ], inv[kRZ];
    bool live[kRZ];
#pragma unroll
    for (int k = 0; k < kRZ; ++k) {
        const int z = zb * kRZ + k;
        live[k] = z < nz;
        const float lo = live[k] ? mm[2 * z] : 0.0f;
        const float hi = live[k] ? mm[2 * z + 1] : 1.0f;
        float span = hi - lo;
        if (span < 1e-6f) span = 1.0f;
        mn[k] = lo;
        inv[k] = 1.0f / span;
    }
    const size_t ss = (size_t)w * h;
    {
        const int tx = t & 15, ty = t >> 4;
        const int x = bx + 4 * tx, y = by + ty;
        float v[kRZ][4];
#pragma unroll
        for (int k = 0; k < kRZ; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[k][j] = 0.0f;
        if (y < h && x < w) {
            const float* row = vol + (size_t)zb * kRZ * ss + (size_t)y * w + x;
            if (vec4) {  // x and every row start are 16-byte aligned
#pragma unroll
                for (int k = 0; k < kRZ; ++k)
                    if (live[k]) {
                        const float4 q = *reinterpret_cast<const float4*>(row + k * ss);
                        v[k][0] = q.x;
                        v[k][1] = q.y;
                        v[k][2] = q.z;
                        v[k][3] = q.w;
                    }
            } else {
#pragma unroll
                for (int k = 0; k < kRZ; ++k)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (live[k] && x + j < w) v[k][j] = row[k * ss + j];
            }
        }
#pragma unroll
        for (int k = 0; k < kRZ; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float raw = v[k][j];
                v[k][j] = (raw == 0.0f) ? 0.0f : (raw - mn[k]) * inv[k];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bool nz = false;
#pragma unroll
            for (int k = 0; k < kRZ; ++k) nz |= v[k][j] != 0.0f;  // NaN counts
            if (nz) {  // a nonzero pixel lies inside the image: x + j < w, y < h
                atomicMax(&rowe[ty][0], x + j + 1);
                atomicMax(&rowe[ty][1], w - (x + j));
                atomicMax(&cole[4 * tx + j][0], y + 1);
                atomicMax(&cole[4 * tx + j][1], h - y);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float4* d = tile + ty * kNormRow + (4 * tx + j) * 2;
            d[0] = make_float4(v[0][j], v[1][j], v[2][j], v[3][j]);
            d[1] = make_float4(v[4][j], v[5][j], v[6][j], v[7][j]);
        }
    }
    __syncthreads();
    // row copy: tile row r is 2*kNormTW float4 of one copy row (pixels w..
    // w+kPadX-1 are the zero pad; rows y >= h are not part of this copy)
    float4* __restrict__ r0 = rowL + (size_t)zb * h * wp * 2;
#pragma unroll
    for (int i = 0; i < kNormTH * kNormTW * 2 / 256; ++i) {
        const int g = i * 256 + t;
        const int r = g / (2 * kNormTW), f = g - r * (2 * kNormTW);
        const int y = by + r, x = bx + (f >> 1);
        if (y < h && x < wp) r0[((size_t)y * wp + x) * 2 + (f & 1)] = tile[r * kNormRow + f];
    }
    // column copy: column c is kNormTH pixels = 2*kNormTH float4 (rows
    // h..h+kPadX-1 are the zero pad; columns x >= w are not part of it)
    float4* __restrict__ c0 = colL + (size_t)zb * w * hp * 2;
#pragma unroll
    for (int i = 0; i < kNormTH * kNormTW * 2 / 256; ++i) {
        const int g = i * 256 + t;
        const int c = g / (2 * kNormTH), q = g - c * (2 * kNormTH);
        const int x = bx + c, y = by + (q >> 1);
        if (x < w && y < hp)
            c0[((size_t)x * hp + y) * 2 + (q & 1)] = tile[(q >> 1) * kNormRow + c * 2 + (q & 1)];
    }
    if (t < kNormTH) {
        const int y = by + t;
        if (y < h && rowe[t][0] > 0) {
            atomicMax(&rflag[((size_t)zb * h + y) * 2], rowe[t][0]);
            atomicMax(&rflag[((size_t)zb * h + y) * 2 + 1], rowe[t][1]);
        }
    } else if (t < kNormTH + kNormTW) {
        const int c = t - kNormTH, x = bx + c;
        if (x < w && cole[c][0] > 0) {
            atomicMax(&cflag[((size_t)zb * w + x) * 2], cole[c][0]);
            atomicMax(&cflag[((size_t)zb * w + x) * 2 + 1], cole[c][1]);
        }
    }
}

// s-interval margin: the float s of SerialBackend.cpp:110 is within ~3e-4 of
// the exact value for slices up to 2048 px, so a candidate interval built
// from [i-kDelta, i+1+kDelta] in exact arithmetic holds every pixel with
// floor(s) = i.
constexpr float kDelta = 1.0f / 32.0f;

// Lane geometry of one interval over an image walked by rows, for an angle
// with |sn| <= |c|. Lane i owns the pixels with floor(s) = i. Row y holds
// them in [xa(y), xa(y) + span], xa(y) = xa0 + y*dxa: the interval's
// x-range over the row, stepped from row to row (|dxa| <= 1 and
// |1/c| <= sqrt2, so the stepping error, ~1e-3 px, is far below the
// kDelta/|c| and 0.01 px margins), at most span + 0.02 <= 1.0625*sqrt2 +
// 0.02 < 2 pixels: two candidates per row.
struct RayGeom {
    float c, sn, cx, cy, dc;
    float xa0, dxa;  // interval start at row 0, less 0.01 px, and its step
    int w;
    float fi;        // the interval's bin i, as a float
    // weights of a pixel at u = s - i (the interval_weights rules)
    float ka, la, kb, lb;
};

// First candidate column of row yf, as a float in [0, w]: ceil of the
// stepped interval start clamped to [0, w] (ceil and the clamp commute for an
// integer w). Monotone in the bin: every operation is.
__device__ __forceinline__ float row_first(const RayGeom& g, float yf) {
    return ceilf(__builtin_amdgcn_fmed3f(fmaf(yf, g.dxa, g.xa0), 0.0f, (float)g.w));
}

// Serial (SerialBackend.cpp:110-120) adds, for every pixel with i0 = floor(s)
// and frac = s - i0, v*(1-frac) to bin i0 and v*frac to bin i0+1 when
// 0 <= i0 < bins-1, v to bin i0 when i0 = bins-1, nothing otherwise. Lane i
// takes the pixels with i0 = i: it sums their first deposits (A, for bin i)
// and their second ones (B, for bin i+1); bin i is then A_i + B_(i-1).
// With u = s - i (exact: Sterbenz where s is near i; where it is not, u is
// far from [0, 1) anyway), u in [0, 1) is exactly i0 = i, u is Serial's frac
// and fl(1 - u) its 1 - frac. A nonzero s is a multiple of 2^-49 (|cos|,
// |sin| are 0 or above 2^-25, pixel offsets multiples of 1/2), so is u, and:
//   A = clamp(min(fma(u, ka, la), fma(u, 2^64, 1)), 0, 1)
//   B = clamp(min(u,              fma(u, kb, lb)),  0, 1)
// - interior lanes (ka, la) = (-1, 1): fl(1 - u) for u in [0, 1), <= 0 for
//   u >= 1; (kb, lb) = (-2^64, 2^64): 2^64 (1 - u) >= 2^15 for u < 1 and
//   <= 0 for u >= 1; the second operand of A, u*2^64 + 1, is >= 1 for
//   u >= 0 and <= 1 - 2^15 for u < 0, where B's u is negative too;
// - the last lane, i = bins-1: (ka, la) = (-2^64, 2^64) gives A = 1 on
//   [0, 1) and 0 elsewhere; (kb, lb) = (0, 0) gives B = 0;
// - lanes past the last bin: (ka, la) = (kb, lb) = (0, -1), A = B = 0.
// Each weight is two or three VALU (fma, fma, min with clamp).
__device__ __forceinline__ void interval_weights(RayGeom& g, int i, int bins) {
    constexpr float kBig = 18446744073709551616.0f;  // 2^64
    g.fi = (float)i;
    g.ka = -1.0f;
    g.la = 1.0f;
    g.kb = -kBig;
    g.lb = kBig;
    if (i == bins - 1) {  // SerialBackend.cpp:117-118
        g.ka = -kBig;
        g.la = kBig;
        g.kb = 0.0f;
        g.lb = 0.0f;
    }
}

// The candidates x0 + k, k < U, of one row; xm0 = x0 - cx exactly (pixel
// offsets are multiples of 1/2 far below 2^24). The deposits are fused
// multiply-adds: every partial sum is within a rounding of Serial's terms,
// and a bin with one deposit gets Serial's rounded product exactly.
template <int U>
__device__ __forceinline__ void chunk_add(const RayGeom& g, float xm0, float sy,
                                          const float4 (&v)[U][2], float (&accA)[kRZ],
                                          float (&accB)[kRZ]) {
    constexpr float kBig = 18446744073709551616.0f;  // 2^64
#pragma unroll
    for (int k = 0; k < U; ++k) {
        // SerialBackend.cpp:110, same float operations and order (xm0 is
        // never -0, so xm0 + 0 is xm0)
        const float s = (k == 0 ? xm0 : xm0 + (float)k) * g.c + sy + g.dc;
        const float u = s - g.fi;
        const float wa = __builtin_amdgcn_fmed3f(fminf(fmaf(u, g.ka, g.la), fmaf(u, kBig, 1.0f)),
                                                 0.0f, 1.0f);
        const float wb = __builtin_amdgcn_fmed3f(fminf(u, fmaf(u, g.kb, g.lb)), 0.0f, 1.0f);
        // a later candidate that lies in no lane's interval adds +0 to every
        // sum: the wave skips it (most waves near 0 and 90 degrees)
        if (k > 0 && __builtin_amdgcn_ballot_w64(wa != 0.0f || wb != 0.0f) == 0) continue;
        const float p[kRZ] = {v[k][0].x, v[k][0].y, v[k][0].z, v[k][0].w,
                              v[k][1].x, v[k][1].y, v[k][1].z, v[k][1].w};
#pragma unroll
        for (int z = 0; z < kRZ; ++z) {
            accA[z] = fmaf(p[z], wa, accA[z]);
            accB[z] = fmaf(p[z], wb, accB[z]);
        }
    }
}

// Lane geometry of interval i at angle (c, sn), |sn| <= |c|, over a w x h
// image.
__device__ __forceinline__ RayGeom ray_geom(int i, int bins, int w, int h, float c, float sn) {
    RayGeom g;
    g.c = c;
    g.sn = sn;
    g.cx = (float)(w - 1) * 0.5f;
    g.cy = (float)(h - 1) * 0.5f;
    g.dc = (float)(bins / 2);
    g.w = w;
    interval_weights(g, i, bins);
    // x - cx = ((s - dc) - (y - cy) sn) / c over s - dc in [slo, shi]
    const float slo = (float)i - kDelta - g.dc;
    const float shi = (float)(i + 1) + kDelta - g.dc;
    const float inv_c = 1.0f / c;
    const float e0 = (slo + g.cy * sn) * inv_c, e1 = (shi + g.cy * sn) * inv_c;
    g.xa0 = (g.cx + fminf(e0, e1)) - 0.01f;
    g.dxa = -sn * inv_c;
    return g;
}

// A block of NT lanes takes the intervals [j NT, (j+1) NT) of one angle and
// z-block. Bin i needs lane i-1's B: within a block it is read from LDS; the
// block's last lane leaves its B in side[(z na + a) (bb-1) + j], and
// radon_seam_kernel adds it to bin (j+1) NT once the walk is done.
// Row walk over a w x h image copy [zb][h][w + kPadX] for angles with
// |sin| <= |cos|: every lane of a block (NT intervals of one angle and
// z-block) walks all rows top to bottom with one chunk of U = 2 candidates
// per row. The block stages RR rows at a time into LDS, each as a window of
// kRowW pixels starting at the extreme lane's first candidate. A lane's first
// candidate x0(y) is a composition of monotone float operations of its bin,
// so the window start is that of the extreme bin (the lowest one for
// cos > 0, the highest for cos < 0), and every lane's chunk lies within
// kRowW pixels of it: lanes 255 bins apart are at most 255/|cos| <= 360.7
// pixels apart, plus a chunk of 2. Rows outside a lane's candidate rows add
// +0, and so do rows of zeros: the block walks, in ascending order, only
// the rows that the z-block's row extents (launch_normalise) mark.
//
// The other angles walk the transposed image: the column copy
// [zb][w][h + kPadX] is the row copy of the transpose, and the kernel runs on
// it with (w, h) and (cos, sin) exchanged. Serial's s = (x-cx)*cos +
// (y-cy)*sin + dc is then formed as (y-cy)*sin + (x-cx)*cos + dc, the same
// float value (the first addition commutes exactly), so every pixel's bin and
// fraction are Serial's.
constexpr int kRadonRows = 2;  // rows staged per step

// Workgroups follow the XCD-grouped order (xcd_map): units are z-blocks,
// a z-block's slots its (bin block, angle) pairs, n_ang angles. A block has
// NT lanes (NT intervals). FULL: the image row fits the window (w + kPadX <=
// kRowW), every block stages whole rows and the window start is 0; NT = 512,
// 768 or 1024 then (rows up to 516, 772 or 1028 pixels), 8-16 waves sharing 2
// staged rows: 32 B of staging per lane and row, against 46 B for the
// windowed 256-lane form of wider rows, whose window must span the lanes'
// rays at any angle.
template <int RR, int U, int kRowW, int NT, bool FULL>
__global__ __launch_bounds__(NT) void radon_rows_kernel(
    const float4* __restrict__ rowL, const int* __restrict__ flag, int w, int h, int nz,
    int na, int bins,
    const float* __restrict__ cos_t, const float* __restrict__ sin_t,
    const int* __restrict__ alist, float* __restrict__ sino, float* __restrict__ side,
    int n_ang, int gsz) {
    // window: lanes NT-1 bins apart are at most (NT-1)/|cos| pixels apart,
    // plus a chunk of U; the window must also keep every chunk inside the
    // zero pad
    static_assert(U - 2 <= kPadX - 1, "chunk reaches past the zero pad");
    static_assert(FULL || (NT - 1) * 1.41421357 + U <= kRowW,
                  "window narrower than the lanes' span");
    static_assert(RR == 2 && NT % 128 == 0, "two staged rows, one per half of the block");
    static_assert(RR * kRowW >= 2 * NT, "the B exchange reuses the staging buffers");
    __shared__ float4 t0[RR][kRowW];  // slices 0-3 of the staged pixels
    __shared__ float4 t1[RR][kRowW];  // slices 4-7
    __shared__ int xw_s[2][RR];       // window starts: staged rows, next rows
    __shared__ unsigned short rows[kRadonMaxEdge];  // the rows to walk
    const int bb = (bins + NT - 1) / NT;
    const int nzb = (nz + kRZ - 1) / kRZ;
    const int slots = (bb * n_ang + gsz - 1) / gsz;
    int zb, slot;
    if (!xcd_map(blockIdx.x, gsz, slots, bb * n_ang, nzb, zb, slot)) return;  // whole block
    const int a = alist[slot / bb];
    const int jb = slot % bb;
    const int b0 = jb * NT;
    const int i = b0 + threadIdx.x;
    const bool live = i < bins;
    const float c = cos_t[a], sn = sin_t[a];
    RayGeom g = ray_geom(live ? i : bins - 1, bins, w, h, c, sn);
    if (!live) {  // walks along, adds nothing
        g.ka = g.kb = 0.0f;
        g.la = g.lb = -1.0f;
    }
    const int wp = w + kPadX;
    const char* __restrict__ slab =
        reinterpret_cast<const char*>(rowL) + (size_t)zb * radon_zbytes(w, h);
    // the extreme interval's geometry, held by the first RR lanes only: lane
    // r computes the window start of row ys + r of the next step
    const int ie = c > 0.0f ? b0 : min(b0 + NT - 1, bins - 1);
    // the rows with a nonzero pixel, 64 rows a ballot; every wave counts
    // them, wave 0 writes the list
    const int lane = threadIdx.x & 63;
    int nr = 0;
    {
        const int* f = flag + (size_t)zb * h * 2;
        for (int y0 = 0; y0 < h; y0 += 64) {
            const int y = y0 + lane;
            const bool mark = y < h && f[2 * y] > 0;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(mark);
            if (threadIdx.x < 64 && mark)
                rows[nr + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = (unsigned short)y;
            nr += __builtin_popcountll(m);
        }
    }
    __syncthreads();
    // listed row k (clamped to the last one: a row staged past the list is
    // never read)
    auto listed = [&](int k) { return nr > 0 ? (int)rows[min(k, nr - 1)] : 0; };
    RayGeom ge;
    if (threadIdx.x < RR) {
        if (FULL) {
            xw_s[0][threadIdx.x] = 0;
        } else {
            ge = ray_geom(ie, bins, w, h, c, sn);
            xw_s[0][threadIdx.x] = (int)row_first(ge, (float)listed(threadIdx.x));
        }
    }
    // Staging: the first half of the block stages the step's first row, the
    // second half its second; thread column j = (threadIdx.x % (NT/2)) +
    // (NT/2) k, k < NE. Columns past the window (FULL: past the row) load a
    // clamped in-bounds address and are not written.
    constexpr int HB = NT / 2;
    constexpr int NE = (kRowW + HB - 1) / HB;
    const int srow = __builtin_amdgcn_readfirstlane(threadIdx.x / HB);
    const int sj0 = threadIdx.x % HB;
    const int nstage = FULL ? wp : kRowW;

    float accA[kRZ], accB[kRZ];
#pragma unroll
    for (int k = 0; k < kRZ; ++k) accA[k] = accB[k] = 0.0f;
    int buf = 0;
    for (int ks = 0; ks < nr; ks += RR, buf ^= 1) {
        __syncthreads();  // the previous rows are consumed; xw_s[buf] is set
        if (!FULL && threadIdx.x < RR)  // window starts of the next step's rows
            xw_s[buf ^ 1][threadIdx.x] = (int)row_first(ge, (float)listed(ks + RR + threadIdx.x));
        // Lanes read window slots up to x0 + U - 1 <= w + kPadX - 1, inside
        // the zero pad, and rows below h only: slots past the pad or the last
        // row are never read, so their addresses are clamped instead of
        // tested. All loads are issued before the LDS writes.
        {
            // this half's row (clamped to the last row) and window start,
            // both wave-uniform
            const unsigned ro = (unsigned)listed(ks + srow) * (unsigned)wp * 32u;
            const int xr = FULL ? 0 : __builtin_amdgcn_readfirstlane(xw_s[buf][srow]);
            float4 q0[NE], q1[NE];
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                const int x = min(xr + sj0 + HB * k, w + kPadX - 1);
                const unsigned off = ro + ((unsigned)x << 5);
                q0[k] = *reinterpret_cast<const float4*>(slab + off);
                q1[k] = *reinterpret_cast<const float4*>(slab + off + 16);
            }
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                if (sj0 + HB * k < nstage) {
                    t0[srow][sj0 + HB * k] = q0[k];
                    t1[srow][sj0 + HB * k] = q1[k];
                }
            }
        }
        __syncthreads();
        for (int r = 0; r < RR && ks + r < nr; ++r) {
            const float yf = (float)listed(ks + r);
            const float sy = (yf - g.cy) * g.sn;  // SerialBackend.cpp:110
            const float x0f = row_first(g, yf);
            const int d = FULL ? (int)x0f : (int)x0f - xw_s[buf][r];
            float4 v[U][2];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                v[k][0] = t0[r][d + k];
                v[k][1] = t1[r][d + k];
            }
            chunk_add<U>(g, x0f - g.cx, sy, v, accA, accB);
        }
    }
    // bin i = A_i + B_(i-1): every lane leaves its B in LDS (the staging
    // buffers, consumed), lane t > 0 adds lane t-1's; the last lane's B goes
    // to the side buffer when a next bin block exists
    __syncthreads();
    float4* bx = &t0[0][0];
    bx[2 * threadIdx.x] = make_float4(accB[0], accB[1], accB[2], accB[3]);
    bx[2 * threadIdx.x + 1] = make_float4(accB[4], accB[5], accB[6], accB[7]);
    __syncthreads();
    if (live) {
        float pb[kRZ];
#pragma unroll
        for (int k = 0; k < kRZ; ++k) pb[k] = 0.0f;
        if (threadIdx.x > 0) {
            const float4 p0 = bx[2 * threadIdx.x - 2], p1 = bx[2 * threadIdx.x - 1];
            pb[0] = p0.x;
            pb[1] = p0.y;
            pb[2] = p0.z;
            pb[3] = p0.w;
            pb[4] = p1.x;
            pb[5] = p1.y;
            pb[6] = p1.z;
            pb[7] = p1.w;
        }
        const bool seam = threadIdx.x == NT - 1 && i + 1 < bins;
#pragma unroll
        for (int k = 0; k < kRZ; ++k) {
            const int z = zb * kRZ + k;
            if (z < nz) {
                sino[((size_t)z * na + a) * bins + i] = accA[k] + pb[k];
                if (seam) side[((size_t)z * na + a) * (bb - 1) + jb] = accB[k];
            }
        }
    }
}

// The row walk with 16 slices per lane, for rows of up to 516 pixels (512
// bins a block): a lane sums the same candidate pixels of two z-blocks of the
// row copy (slices 16 zp .. 16 zp + 15), so one candidate's geometry and
// weights serve 32 fused multiply-adds instead of 16. A block of 1024 threads
// walks two angles over the same staged rows (threads 0-511: the pair's first
// angle, 512-1023: its second), which halves the staging per lane and lets
// two blocks (8 waves per SIMD) share a CU. Rows are staged by DMA from
// global memory straight into LDS (global_load_lds_dwordx4: lane l of a wave
// writes slot base + l; 64 pixels per load, so rows are padded to kR16W = 576
// slots), one row ahead: row y is summed from one array while row y + 1
// lands in the other, so the staging latency hides behind a row of sums.
// Each candidate's 4 b128 LDS reads are issued before its sums; a wave whose
// intervals hold no second pixel in the row skips the second candidate
// (within ~10 degrees of the axes most waves do). Weights:
// u = s - i in [0, 1) exactly when floor(s) = i (one unsigned compare of u's
// bits); A = fl(1 - u), or 1 on the last bin (SerialBackend.cpp:117-118);
// B = u (the last bin's B goes nowhere and is not written).
// Zero pixels add +0 to every sum: the block first lists, in LDS and in
// ascending order, the rows of its two z-blocks that hold a nonzero pixel,
// with their joint nonzero extent [xlo, xhi] (launch_normalise), and stages
// and walks those only; in a row, a wave none of whose lanes' two candidates
// lies in the extent skips its sums. Every sum is the one the full walk
// forms.
constexpr int kR16W = 576;

// The walk of workgroup u of its grid on nx XCDs (the body of both kernels
// below).
__device__ __forceinline__ void rows16_walk(
    int u, const float4* __restrict__ rowL, const int* __restrict__ flag, int w, int h, int nz,
    int na, int bins,
    const float* __restrict__ cos_t, const float* __restrict__ sin_t,
    const int* __restrict__ alist, float* __restrict__ sino, float* __restrict__ side,
    int n_ang, int gsz, int nx) {
    constexpr int NT = 512;
    // staged rows, one row ahead: the even entries of the row list in tE,
    // the odd ones in tO (two arrays, so the LDS reads of one row need not
    // wait for the DMA into the other); [slice group (4 slices)][pixel]
    __shared__ float4 tE[4][kR16W];
    __shared__ float4 tO[4][kR16W];
    // the rows to walk: y | xlo << 11 | xhi << 21 (y < 2048, x < 1024)
    __shared__ unsigned rows[kRadonMaxEdge];
    const int bb = (bins + NT - 1) / NT;
    const int nzb = (nz + kRZ - 1) / kRZ;
    const int npair = (nzb + 1) / 2;
    const int nap = (n_ang + 1) / 2;  // angle pairs
    const int slots = (bb * nap + gsz - 1) / gsz;
    int zp, slot;
    if (!xcd_map(u, gsz, slots, bb * nap, npair, zp, slot, nx)) return;  // whole block
    const int half = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9);
    const int lt = threadIdx.x & (NT - 1);
    const int ai = 2 * (slot / bb) + half;  // this half's angle, in the walk's list
    const bool angle_live = ai < n_ang;
    const int a = alist[angle_live ? ai : n_ang - 1];
    const int jb = slot % bb;
    const int i = jb * NT + lt;
    const bool live = angle_live && i < bins;
    const RayGeom g = ray_geom(i < bins ? i : bins - 1, bins, w, h, cos_t[a], sin_t[a]);
    const float ka = (i == bins - 1) ? 0.0f : -1.0f;
    const int wp = w + kPadX;
    // A row is 4 slice groups x 9 loads of 64 pixels; group grp is half
    // (grp & 1) of z-block 2 zp + (grp >> 1), the last z-block standing in for
    // a missing second one (its slices are never written); slots past the
    // row's pad repeat its last pixel, and no lane reads them.
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const size_t zbytes = radon_zbytes(w, h);
    // wave v stages group v & 3, pixel chunks v >> 2 and (v >> 2) + 4 (pixels
    // 0..511), and waves 0-3 the tail chunk 8 (pixels 512..575, clamped) of
    // group v: every lane offset is fixed for the whole walk
    const int sgrp = wave & 3, sk = wave >> 2;
    const char* sbase = reinterpret_cast<const char*>(rowL) +
                        (size_t)min(2 * zp + (sgrp >> 1), nzb - 1) * zbytes + (sgrp & 1) * 16;
    const unsigned xo0 = (unsigned)min(sk * 64 + lane, wp - 1) << 5;
    const unsigned xo1 = (unsigned)min(sk * 64 + 256 + lane, wp - 1) << 5;
    const unsigned xo2 = (unsigned)min(512 + lane, wp - 1) << 5;
    const int nload = wave < 4 ? 3 : 2;
    auto stage = [&](float4 (&t)[4][kR16W], int y) {
        const char* src = sbase + (size_t)y * (size_t)wp * 32u;
        // a rolled loop: unrolled, the loads' LDS targets are taken as
        // possibly read by the row in progress, which then waits for them
#pragma unroll 1
        for (int m = 0; m < nload; ++m) {
            const unsigned xo = m == 0 ? xo0 : (m == 1 ? xo1 : xo2);
            const int slot = m == 2 ? 512 : sk * 64 + m * 256;
            __builtin_amdgcn_global_load_lds((const void*)(src + xo),
                                             (__attribute__((address_space(3))) void*)&t[sgrp][slot],
                                             16, 0, 0);
        }
    };
    // the rows with a nonzero pixel in either z-block, 64 rows a ballot;
    // every wave counts them, wave 0 writes the list
    int nr = 0;
    {
        const int* f0 = flag + (size_t)(2 * zp) * h * 2;
        const int* f1 = flag + (size_t)min(2 * zp + 1, nzb - 1) * h * 2;
        for (int y0 = 0; y0 < h; y0 += 64) {
            const int y = y0 + lane;
            const int hi1 = y < h ? max(f0[2 * y], f1[2 * y]) : 0;
            const int lo1 = y < h ? max(f0[2 * y + 1], f1[2 * y + 1]) : 0;
            const bool mark = hi1 > 0;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(mark);
            if (wave == 0 && mark)
                rows[nr + __builtin_popcountll(m & ((1ull << lane) - 1ull))] =
                    (unsigned)y | (unsigned)(w - lo1) << 11 | (unsigned)(hi1 - 1) << 21;
            nr += __builtin_popcountll(m);
        }
    }
    __syncthreads();
    float accA[16], accB[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) accA[k] = accB[k] = 0.0f;
    auto row = [&](const float4 (&t)[4][kR16W], unsigned e) {
        const int y = (int)(e & 2047u);
        const int xlo = (int)((e >> 11) & 1023u), xhi = (int)(e >> 21);
        const float yf = (float)y;
        const float sy = (yf - g.cy) * g.sn;  // SerialBackend.cpp:110
        const float x0f = row_first(g, yf);
        const int d = (int)x0f;
        const float xm0 = x0f - g.cx;
        // SerialBackend.cpp:110, the same float operations in order, for the
        // row's two candidates x0 and x0 + 1
        const float s0 = xm0 * g.c + sy + g.dc;
        const float s1 = (xm0 + 1.0f) * g.c + sy + g.dc;
        const float u0 = s0 - g.fi, u1 = s1 - g.fi;
        const bool in0 = __float_as_uint(u0) < 0x3f800000u;
        const bool in1 = __float_as_uint(u1) < 0x3f800000u;
        // A row holds a second pixel of an interval only where the interval
        // spans two pixels (|cos| < 1); a wave none of whose 64 intervals
        // does skips the second candidate's reads and sums, which would add
        // +0 to every sum.
        const bool two = __builtin_amdgcn_ballot_w64(in1) != 0;
        auto add = [&](int k, float u, bool in) {
            float4 q[4];
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) q[gq] = t[gq][d + k];
            const float wa = in ? fmaf(u, ka, 1.0f) : 0.0f;
            const float wb = in ? u : 0.0f;
            const float p[16] = {q[0].x, q[0].y, q[0].z, q[0].w, q[1].x, q[1].y, q[1].z, q[1].w,
                                 q[2].x, q[2].y, q[2].z, q[2].w, q[3].x, q[3].y, q[3].z, q[3].w};
#pragma unroll
            for (int z = 0; z < 16; ++z) {
                accA[z] = fmaf(p[z], wa, accA[z]);
                accB[z] = fmaf(p[z], wb, accB[z]);
            }
        };
        // a wave whose candidates d, d + 1 all lie outside [xlo, xhi] would add
        // +0 to every sum
        if (__builtin_amdgcn_ballot_w64(d <= xhi && d + 1 >= xlo) != 0) {
            add(0, u0, in0);
            if (two) add(1, u1, in1);
        }
    };
    // the list 64 entries at a time, entry k0 + l in lane l; an entry is read
    // out with v_readlane (a wave-uniform value, no LDS wait per row)
    int k0 = 0;
    int chunk = lane < nr ? (int)rows[lane] : 0;
    auto listed = [&](int k) {
        return k < nr ? (unsigned)__builtin_amdgcn_readlane(chunk, k - k0) : 0u;
    };
    unsigned y0 = listed(0), y1 = listed(1);
    if (nr > 0) stage(tE, (int)(y0 & 2047u));
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int k = 0; k < nr; k += 2) {
        // listed row k from tE while row k + 1 lands in tO (read last by row
        // k - 1, before the previous barrier)
        if (k + 2 >= k0 + 64) {  // k + 2 and k + 3 (k even) lie in the next 64
            k0 += 64;
            chunk = k0 + lane < nr ? (int)rows[k0 + lane] : 0;
        }
        const unsigned y2 = listed(k + 2), y3 = listed(k + 3);
        if (k + 1 < nr) stage(tO, (int)(y1 & 2047u));
        row(tE, y0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (k + 1 < nr) {
            if (k + 2 < nr) stage(tE, (int)(y2 & 2047u));
            row(tO, y1);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
        y0 = y2;
        y1 = y3;
    }
    // bin i = A_i + B_(i-1), as in radon_rows_kernel, per half (the staging
    // arrays are free: every wave passed the last barrier)
    float4* bx = (half == 0 ? &tE[0][0] : &tO[0][0]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        bx[4 * lt + q] = make_float4(accB[4 * q], accB[4 * q + 1], accB[4 * q + 2], accB[4 * q + 3]);
    __syncthreads();
    if (live) {
        float pb[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) pb[k] = 0.0f;
        if (lt > 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 v = bx[4 * (lt - 1) + q];
                pb[4 * q] = v.x;
                pb[4 * q + 1] = v.y;
                pb[4 * q + 2] = v.z;
                pb[4 * q + 3] = v.w;
            }
        }
        const bool seam = lt == NT - 1 && i + 1 < bins;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int z = zp * 16 + k;
            if (z < nz) {
                sino[((size_t)z * na + a) * bins + i] = accA[k] + pb[k];
                if (seam) side[((size_t)z * na + a) * (bb - 1) + jb] = accB[k];
            }
        }
    }
}

__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
void radon_rows16_kernel(
    const float4* __restrict__ rowL, const int* __restrict__ flag, int w, int h, int nz, int na,
    int bins, const float* __restrict__ cos_t, const float* __restrict__ sin_t,
    const int* __restrict__ alist, float* __restrict__ sino, float* __restrict__ side,
    int n_ang, int gsz) {
    rows16_walk(blockIdx.x, rowL, flag, w, h, nz, na, bins, cos_t, sin_t, alist, sino, side,
                n_ang, gsz, kXcds);
}

// Both walks of a Radon stage in one launch, each on half the XCDs: XCDs
// 0-3 walk the image (a), XCDs 4-7 the transposed image (b), each in its own
// XCD-grouped order over its four XCDs (workgroup g is on XCD g % 8; its
// index in its walk's grid is (g / 8) * 4 + g % 4). The walks write disjoint
// angles and take about the same time (as many angles each, to one). One
// launch leaves no idle tail between the walks, which over few slices (a
// round or two of workgroups) is much of the time: both walks over 32 slices
// of 512^2 at 720 angles, 1.18 ms as two launches. Each XCD still walks one
// pair of z-blocks at a time, so the L2s and the memory-side cache hold what
// they held for the separate launches.
struct Walk16 {
    const float4* rows;
    const int* flag;  // [zb][h] row flags of the copy
    int w, h;
    const float* cos_t;
    const float* sin_t;
    const int* alist;
    int n_ang, gsz;
    unsigned blocks;  // workgroups of the walk's grid
};
constexpr int kHalfXcds = kXcds / 2;

__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
void radon_rows16_pair_kernel(Walk16 a, Walk16 b, int nz, int na, int bins,
                              float* __restrict__ sino, float* __restrict__ side) {
    const unsigned g = blockIdx.x;
    const bool second = (g / kHalfXcds) & 1u;
    const Walk16& W = second ? b : a;
    const int u = (int)((g / kXcds) * kHalfXcds + g % kHalfXcds);
    if (u >= (int)W.blocks) return;  // whole block
    rows16_walk(u, W.rows, W.flag, W.w, W.h, nz, na, bins, W.cos_t, W.sin_t, W.alist, sino,
                side, W.n_ang, W.gsz, kHalfXcds);
}

// Adds each bin block's last B to the next block's first bin (grid: seams x
// angles of the walk, z).
__global__ void radon_seam_kernel(float* __restrict__ sino, const float* __restrict__ side,
                                  const int* __restrict__ alist, int n_ang, int na, int bins,
                                  int NT, int nseam) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseam * n_ang) return;
    const int j = t % nseam, a = alist[t / nseam];
    const size_t za = (size_t)blockIdx.y * na + a;
    sino[za * bins + (size_t)(j + 1) * NT] += side[za * nseam + j];
}

hipError_t launch_minmax_init(unsigned* d_keys, int nz, hipStream_t s) {
    hipLaunchKernelGGL(minmax_init_kernel, dim3((nz + 255) / 256), dim3(256), 0, s, d_keys, nz);
    return hipGetLastError();
}

hipError_t launch_minmax(const float* d_vol, size_t slice_elems, int nz, unsigned* d_keys,
                         hipStream_t s) {
    // 64 KB of a slice per block (16 float4 per thread)
    int chunks = (int)((slice_elems + 256 * 64 - 1) / (256 * 64));
    chunks = chunks < 1 ? 1 : (chunks > 64 ? 64 : chunks);
    const int vec4 = (slice_elems % 4 == 0) && ((uintptr_t)d_vol % 16 == 0);
    hipLaunchKernelGGL(minmax_kernel, dim3(chunks, nz), dim3(256), 0, s, d_vol, slice_elems, vec4,
                       d_keys);
    return hipGetLastError();
}

hipError_t launch_minmax_decode(const unsigned* d_keys, int nz, float* d_mm, hipStream_t s) {
    hipLaunchKernelGGL(minmax_decode_kernel, dim3((2 * nz + 255) / 256), dim3(256), 0, s, d_keys,
                       nz, d_mm);
    return hipGetLastError();
}

hipError_t launch_normalise(const float* d_vol, int w, int h, int nz, const float* d_mm,
                            float4* d_row, float4* d_col, int* d_rflag, int* d_cflag,
                            hipStream_t s) {
    const int nzb = (nz + kRZ - 1) / kRZ;
    dim3 grid((w + kPadX + kNormTW - 1) / kNormTW, (h + kPadX + kNormTH - 1) / kNormTH, nzb);
    const int vec4 = (w % 4 == 0) && ((uintptr_t)d_vol % 16 == 0);
    hipLaunchKernelGGL(normalise_kernel, grid, dim3(256), 0, s, d_vol, w, h, nz, d_mm, vec4,
                       d_row, d_col, d_rflag, d_cflag);
    return hipGetLastError();
}

// Block width of the walk over rows of w pixels: whole-row staging with 512,
// 768 or 1024 lanes while the row fits the window (NT + 8 pixels), else the
// windowed 256-lane form.
static int radon_lanes(int w) {
    return w + kPadX <= 520 ? 512 : (w + kPadX <= 776 ? 768 : (w + kPadX <= 1032 ? 1024 : 256));
}

size_t radon_side_floats(int w, int bins, int nz, int na) {
    const int NT = radon_lanes(w);
    return (size_t)nz * na * ((bins + NT - 1) / NT - 1);
}

// Grid of the 16-slice walk over n angles on nx XCDs: units are pairs of
// z-blocks, ~150 of its 1024-thread workgroups per XCD and round (40-600
// measured alike at 512^3); a multiple of nx workgroups (none when n = 0).
static Walk16 walk16(const float4* rows, const int* flag, int w, int h, int nz, int bins,
                     const float* cos_t, const float* sin_t, const int* alist, int n, int nx) {
    Walk16 W{rows, flag, w, h, cos_t, sin_t, alist, n, 1, 0u};
    if (n <= 0) return W;
    const int nzb = (nz + kRZ - 1) / kRZ, bb = (bins + 511) / 512;
    const int npair = (nzb + 1) / 2, nap = (n + 1) / 2;
    W.gsz = xcd_group(bb * nap, npair, 150.0, nx);
    const long long slots = (bb * nap + W.gsz - 1) / W.gsz, P = nx / W.gsz;
    W.blocks = (unsigned)(nx * slots * ((npair + P - 1) / P));
    return W;
}

static void launch_seams(float* d_sino, const float* d_side, const int* d_alist, int n, int nz,
                         int na, int bins, int NT, hipStream_t s) {
    const int nseam = (bins + NT - 1) / NT - 1;
    hipLaunchKernelGGL(radon_seam_kernel, dim3((nseam * n + 255) / 256, nz), dim3(256), 0, s,
                       d_sino, d_side, d_alist, n, na, bins, NT, nseam);
}

hipError_t launch_radon_rows(const float4* d_img, const int* d_flag, int w, int h, int nz,
                             int na, int bins,
                             const float* d_cos, const float* d_sin, const int* d_alist, int n,
                             float* d_sino, float* d_side, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int nzb = (nz + kRZ - 1) / kRZ;
    const int NT = radon_lanes(w);
    const int bb = (bins + NT - 1) / NT;
    // Rows up to 516 px: the 16-slice walk over angle pairs (units: pairs of
    // z-blocks; ~150 of its 1024-thread blocks per XCD and round: 40-600
    // measured alike at 512^3). Wider rows: the 8-slice walk, ~300 blocks per
    // XCD and round for windowed blocks, proportionally fewer for the wider
    // whole-row ones (measured: two XCDs per z-block at 768^3 and 1024^3).
    const int gsz = xcd_group(bb * n, nzb, NT > 512 ? 300.0 * 512.0 / NT : 300.0);
    const long long slots = (bb * n + gsz - 1) / gsz, P = kXcds / gsz;
    const dim3 grid((unsigned)(kXcds * slots * ((nzb + P - 1) / P)));
    if (NT == 512) {
        const Walk16 W = walk16(d_img, d_flag, w, h, nz, bins, d_cos, d_sin, d_alist, n, kXcds);
        hipLaunchKernelGGL(radon_rows16_kernel, dim3(W.blocks), dim3(1024), 0, s, d_img, d_flag,
                           w, h, nz, na, bins, d_cos, d_sin, d_alist, d_sino, d_side, n, W.gsz);
    } else if (NT == 768)
        hipLaunchKernelGGL((radon_rows_kernel<kRadonRows, 2, 776, 768, true>), grid, dim3(768), 0,
                           s, d_img, d_flag, w, h, nz, na, bins, d_cos, d_sin, d_alist, d_sino, d_side,
                           n, gsz);
    else if (NT == 1024)
        hipLaunchKernelGGL((radon_rows_kernel<kRadonRows, 2, 1032, 1024, true>), grid,
                           dim3(1024), 0, s, d_img, d_flag, w, h, nz, na, bins, d_cos, d_sin, d_alist,
                           d_sino, d_side, n, gsz);
    else
        hipLaunchKernelGGL((radon_rows_kernel<kRadonRows, 2, 368, 256, false>), grid, dim3(256),
                           0, s, d_img, d_flag, w, h, nz, na, bins, d_cos, d_sin, d_alist, d_sino, d_side,
                           n, gsz);
    if (bb > 1) launch_seams(d_sino, d_side, d_alist, n, nz, na, bins, NT, s);
    return hipGetLastError();
}

bool radon_walks_paired(int w, int h) { return radon_lanes(w) == 512 && radon_lanes(h) == 512; }

hipError_t launch_radon_walks(const float4* d_row, const float4* d_col, const int* d_rflag,
                              const int* d_cflag, int w, int h, int nz, int na, int bins,
                              const float* d_cos, const float* d_sin, const int* d_alist,
                              int n_row, float* d_sino, float* d_side, hipStream_t s) {
    const Walk16 a =
        walk16(d_row, d_rflag, w, h, nz, bins, d_cos, d_sin, d_alist, n_row, kHalfXcds);
    const Walk16 b = walk16(d_col, d_cflag, h, w, nz, bins, d_sin, d_cos, d_alist + n_row,
                            na - n_row, kHalfXcds);
    const unsigned half = std::max(a.blocks, b.blocks);  // multiples of kHalfXcds
    if (half == 0) return hipSuccess;
    hipLaunchKernelGGL(radon_rows16_pair_kernel, dim3(2 * half), dim3(1024), 0, s, a, b, nz, na,
                       bins, d_sino, d_side);
    const int bb = (bins + 511) / 512;
    if (bb > 1) {
        if (n_row > 0) launch_seams(d_sino, d_side, d_alist, n_row, nz, na, bins, 512, s);
        if (na - n_row > 0)
            launch_seams(d_sino, d_side, d_alist + n_row, na - n_row, nz, na, bins, 512, s);
    }
    return hipGetLastError();
}

}  // namespace cthip
