// bp_tile.h — one backprojection tile, shared by the per-volume kernel and
// the persistent multi-volume kernel.
//
// Reference semantics (SerialBackend.cpp:197-252): for voxel (x,y) of an
// N x N slice, xx = x - R, yy = y - R with R = float(N/2); voxels with
// xx^2 + yy^2 > R^2 are 0; otherwise for every angle
//   u = (xx*cos - yy*sin) + float(sq/2),  i0 = floor(u),  fr = u - i0,
//   sum += f[i0]*(1-fr) + f[i0+1]*fr   when i0 >= 0 and i0+1 < sq,
// and the sum is scaled by pi/(2 na). u, i0 and fr are computed with the same
// float operations (contraction off), so detector indexing is bit-identical
// to Serial's; the interpolation is one FMA f[i0] + fr*(f[i0+1]-f[i0]) and
// the angle sum runs in Serial's order.
//
// A tile is T x T voxels (T = 16 or 32) of ZB2 z-blocks (4 slices each,
// ZB2 = 1 or 2); each of its 4 waves covers one (T/2) x (T/2) quadrant, a
// lane (T/16)^2 voxels 8 apart x 4*ZB2 slices, so one voxel's detector
// position and fraction serve 4*ZB2 slices. Tiles and waves whose voxels all
// lie outside the reconstruction circle skip the angle sums (12.5% of the
// 32-voxel tiles of a 512 x 512 slice, 16% of its waves). Each slice's sum is the same sequence of operations whatever
// T, A and ZB2, so every configuration gives bit-identical volumes.
// Angles go in stages of A: per stage the block stages into LDS, for every
// angle, the detector window the tile can touch (F bins from
// floor(min corner u) - 1; u is monotone in x and y, so the corners bound it)
// as (f, f[i+1]-f[i]) float4 pairs: both interpolation operands for 4 slices
// in two 16-byte LDS reads. Each stage also stages its angles' (cos, sin),
// read in the angle loop as one broadcast LDS value.
#pragma once

#include "ct_internal.h"

namespace cthip {

// detector bins one angle of a T-tile can touch: ceil((T-1)*sqrt2) + 4
__host__ __device__ constexpr int bp_window(int T) { return T == 32 ? 48 : (T == 16 ? 26 : 0); }

template <int T, int A, int ZB2>
struct BpSmem {
    float4 f[ZB2][A][bp_window(T)];
    float4 d[ZB2][A][bp_window(T)];
    float2 cs[A];  // (cos, sin) of the stage's angles
    int lo[A];
};

// Global-memory views of the argument pointers. The persistent kernel reads
// its arguments from a work list in memory, where the compiler only sees
// generic pointers; naming the address space gives both kernels the same
// global (not flat) loads and stores.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const f4v __attribute__((address_space(1))) g_cfloat4;
typedef float __attribute__((address_space(1))) g_float;
typedef const float __attribute__((address_space(1))) g_cfloat;

template <int T, int A, bool CHECK, int ZB2>
__device__ __forceinline__ void bp_tile(const BpArgs& p, int bx, int by, int zp,
                                        BpSmem<T, A, ZB2>& sm) {
    constexpr int F = bp_window(T);
    constexpr int V = T / 16;  // voxels per thread per axis
    const int tid = threadIdx.x;
    // each wave covers one (T/2) x (T/2) quadrant of the tile, its lanes an
    // 8 x 8 grid there, a lane's voxels 8 apart
    const int w = tid >> 6, l = tid & 63;
    const int ox = (T / 2) * (w & 1) + (l & 7), oy = (T / 2) * (w >> 1) + (l >> 3);
    const float R = (float)(p.N / 2);
    const float dcb = (float)(p.sq / 2);
    float xx[V], yy[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        xx[i] = (float)(bx + ox + 8 * i) - R;
        yy[i] = (float)(by + oy + 8 * i) - R;
    }
    const float cxa = (float)bx - R, cxb = (float)(bx + T - 1) - R;
    const float cya = (float)by - R, cyb = (float)(by + T - 1) - R;
    float4 acc[ZB2][V][V];
#pragma unroll
    for (int q = 0; q < ZB2; ++q)
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
            for (int j = 0; j < V; ++j) acc[q][i][j] = make_float4(0.f, 0.f, 0.f, 0.f);
    // z-blocks zp*ZB2 + q; one past the last z-block stages zeros
    const int nzb = (p.nz + kZB - 1) / kZB;
    g_cfloat4* fz[ZB2];
    bool zok[ZB2];
#pragma unroll
    for (int q = 0; q < ZB2; ++q) {
        const int zb = zp * ZB2 + q;
        zok[q] = zb < nzb;
        fz[q] = (g_cfloat4*)p.filt + (size_t)min(zb, nzb - 1) * p.na * p.sq;
    }
    g_cfloat* __restrict__ cos_t = (g_cfloat*)p.cos_t;
    g_cfloat* __restrict__ sin_t = (g_cfloat*)p.sin_t;

    // Voxels outside the circle are written as 0 below whatever their sums,
    // so a tile (or a wave's share of it) with no voxel inside skips the
    // angle sums. The circle test x^2 + y^2 > R^2 is the epilogue's; it
    // rounds monotonically in |x| and |y|, so when the tile's voxel nearest
    // the centre fails it, every voxel of the tile does.
    const float nx = cxa > 0.0f ? cxa : (cxb < 0.0f ? -cxb : 0.0f);
    const float ny = cya > 0.0f ? cya : (cyb < 0.0f ? -cyb : 0.0f);
    const int na_run = (nx * nx + ny * ny > R * R) ? 0 : p.na;  // block-uniform
    bool mine = false;
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
        for (int j = 0; j < V; ++j)
            mine |= bx + ox + 8 * i < p.N && by + oy + 8 * j < p.N &&
                    !(xx[i] * xx[i] + yy[j] * yy[j] > R * R);
    const bool wave_on = __builtin_amdgcn_ballot_w64(mine) != 0;  // wave-uniform

    for (int a0 = 0; a0 < na_run; a0 += A) {
        const int nc = min(A, p.na - a0);
        if (tid < nc) {
            const float c = cos_t[a0 + tid], s = sin_t[a0 + tid];
            sm.cs[tid] = make_float2(c, s);
            const float u00 = (cxa * c - cya * s) + dcb;
            const float u01 = (cxb * c - cya * s) + dcb;
            const float u10 = (cxa * c - cyb * s) + dcb;
            const float u11 = (cxb * c - cyb * s) + dcb;
            sm.lo[tid] = (int)floorf(fminf(fminf(u00, u01), fminf(u10, u11))) - 1;
        }
        __syncthreads();
        for (int e = tid; e < nc * F; e += 256) {
            const int aa = e / F, j = e - aa * F;
            const int g = sm.lo[aa] + j;
            const bool in0 = g >= 0 && g < p.sq, in1 = g + 1 >= 0 && g + 1 < p.sq;
#pragma unroll
            for (int q = 0; q < ZB2; ++q) {
                g_cfloat4* row = fz[q] + (size_t)(a0 + aa) * p.sq;
                const f4v z4 = {0.f, 0.f, 0.f, 0.f};
                const f4v f0 = (zok[q] && in0) ? row[g] : z4;
                const f4v f1 = (zok[q] && in1) ? row[g + 1] : z4;
                sm.f[q][aa][j] = make_float4(f0.x, f0.y, f0.z, f0.w);
                sm.d[q][aa][j] = make_float4(f1.x - f0.x, f1.y - f0.y, f1.z - f0.z, f1.w - f0.w);
            }
        }
        __syncthreads();
        for (int aa = 0; aa < (wave_on ? nc : 0); ++aa) {
            const float2 cs = sm.cs[aa];
            const float c = cs.x, s = cs.y;
            // byte offset of this angle's window entry for bin 0: entry
            // i0 - lo of row aa of z-block q sits at i0*16 + boff + q*A*F*16
            const int boff = __builtin_amdgcn_readfirstlane((aa * F - sm.lo[aa]) * 16);
            float xc[V], ys[V];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                xc[i] = xx[i] * c;
                ys[i] = yy[i] * s;
            }
#pragma unroll
            for (int j = 0; j < V; ++j)
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const float t = xc[i] - ys[j];
                    const float u = t + dcb;
                    int i0;
                    float fr;
                    if (CHECK) {
                        const float fl = floorf(u);
                        i0 = (int)fl;
                        fr = u - fl;
                    } else {
                        // u >= dcb - R >= 1 inside the circle: truncation is
                        // floor and fract is u - floor(u), exactly; voxels
                        // outside the circle are zeroed below
                        i0 = (int)u;
                        fr = __builtin_amdgcn_fractf(u);
                    }
                    // i0 - lo is in [1, F - 2]: u is monotone in x and y, so
                    // floor(u) lies between the tile corners' floors, at most
                    // ceil((T-1)*sqrt2) apart (truncation adds at most 1)
                    const int bo = i0 * 16 + boff;
#pragma unroll
                    for (int q = 0; q < ZB2; ++q) {
                        const float4 f = *reinterpret_cast<const float4*>(
                            reinterpret_cast<const char*>(&sm.f[q][0][0]) + bo);
                        const float4 d = *reinterpret_cast<const float4*>(
                            reinterpret_cast<const char*>(&sm.d[q][0][0]) + bo);
                        if (!CHECK || (i0 >= 0 && i0 + 1 < p.sq)) {
                            acc[q][i][j].x += fmaf(fr, d.x, f.x);
                            acc[q][i][j].y += fmaf(fr, d.y, f.y);
                            acc[q][i][j].z += fmaf(fr, d.z, f.z);
                            acc[q][i][j].w += fmaf(fr, d.w, f.w);
                        }
                    }
                }
        }
        __syncthreads();
    }

    const size_t ss = (size_t)p.N * p.N;
#pragma unroll
    for (int q = 0; q < ZB2; ++q) {
        const int zb = zp * ZB2 + q;
        float lo_z[kZB], eff_z[kZB];
#pragma unroll
        for (int k = 0; k < kZB; ++k) {
            const int z = zb * kZB + k;
            lo_z[k] = 0.0f;
            eff_z[k] = 1.0f;
            if (p.mm && z < p.nz) {
                g_cfloat* mm = (g_cfloat*)p.mm;
                const float span = mm[2 * z + 1] - mm[2 * z];
                lo_z[k] = mm[2 * z];
                eff_z[k] = (span <= 0.0f) ? 1.0f : span;  // SerialBackend.cpp:298-304
            }
        }
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int x = bx + ox + 8 * i, y = by + oy + 8 * j;
                if (x >= p.N || y >= p.N) continue;
                const float xf = (float)x - R, yf = (float)y - R;
                const bool inside = !(xf * xf + yf * yf > R * R);
                const float v4[4] = {acc[q][i][j].x, acc[q][i][j].y, acc[q][i][j].z,
                                     acc[q][i][j].w};
#pragma unroll
                for (int k = 0; k < kZB; ++k) {
                    const int z = zb * kZB + k;
                    if (z >= p.nz) break;
                    float v = inside ? v4[k] * p.scale : 0.0f;
                    if (p.mm) v = v * eff_z[k] + lo_z[k];
                    ((g_float*)p.out)[(size_t)z * ss + (size_t)y * p.N + x] = v;
                }
            }
    }
}

}  // namespace cthip
