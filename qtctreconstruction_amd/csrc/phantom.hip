// phantom.hip — seeded additive-ellipsoid phantom on gfx950.
//
// Replaces the reference's per-slice generator (Generator3DGPU.cu:80-115,
// one launch plus a synchronous D2H per slice) with one launch over the whole
// volume, written straight into device memory. Voxel values are bit-equal to
// the host restatement: same float operations in the same order, contraction
// off (the library is built with -ffp-contract=off).
//
// Each block covers a 64x4 (x,y) tile of kPhZ = 4 z slices. Ellipsoids whose
// bounding box misses the tile are skipped; skipping is exact because a voxel
// outside the box cannot pass the membership test (the box carries a 1e-3
// margin, far above the float error of the form). The box tests run one
// ellipsoid per lane and a ballot gives the wave the ones to evaluate, so
// the test's loads are not serialised per ellipsoid.
#include "ct_internal.h"

namespace cthip {

__device__ __forceinline__ float lin_coord(int i, int n, float step) {
    // utils::linspace (Utils.cpp:12-26): start + step * i, start = -1
    return (n == 1) ? -1.0f : -1.0f + step * (float)i;
}

constexpr int kPhZ = 4;  // z slices per thread

__global__ __launch_bounds__(256) void phantom_kernel(const PhEll* __restrict__ e, int n, int nx,
                                                      int ny, int nz, float* __restrict__ vol) {
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
    const int x = x0 + (threadIdx.x & 63);
    const int y = y0 + (threadIdx.x >> 6);
    const int z0 = blockIdx.z * kPhZ;
    const int nzl = min(kPhZ, nz - z0);
    const float stx = 2.0f / (float)(nx - 1);
    const float sty = 2.0f / (float)(ny - 1);
    const float stz = 2.0f / (float)(nz - 1);
    const float xf = lin_coord(x, nx, stx);
    const float yf = lin_coord(y, ny, sty);
    float zf[kPhZ];
#pragma unroll
    for (int j = 0; j < kPhZ; ++j) zf[j] = lin_coord(min(z0 + j, nz - 1), nz, stz);
    // tile coordinate box (linspace is increasing)
    const float tx0 = lin_coord(x0, nx, stx);
    const float tx1 = lin_coord(min(x0 + 63, nx - 1), nx, stx);
    const float ty0 = lin_coord(y0, ny, sty);
    const float ty1 = lin_coord(min(y0 + 3, ny - 1), ny, sty);
    const float tz0 = zf[0], tz1 = zf[kPhZ - 1];
    float v[kPhZ];
#pragma unroll
    for (int j = 0; j < kPhZ; ++j) v[j] = 0.0f;
    // 64 ellipsoids at a time: lane j tests ellipsoid kb + j against the tile
    // box, and the wave then visits the ones that pass in ascending order
    // (the additions keep the host's order)
    for (int kb = 0; kb < n; kb += 64) {
        const int kj = kb + (int)(threadIdx.x & 63);
        bool pass = false;
        if (kj < n) {
            const PhEll& t = e[kj];
            pass = !(tx1 < t.lo[0] || tx0 > t.hi[0] || ty1 < t.lo[1] || ty0 > t.hi[1] ||
                     tz1 < t.lo[2] || tz0 > t.hi[2]);
        }
        unsigned long long m = __ballot(pass);
        while (m) {
            const int k = kb + __ffsll((long long)m) - 1;
            m &= m - 1;
            const PhEll& q = e[k];
            const float dx = xf - q.cx;
            const float dy = yf - q.cy;
            // (m0 dx + m1 dy) is the same float for every slice; the slice's
            // m2 dz is added last, as in the host's left-to-right sum
            const float a0 = q.m[0] * dx + q.m[1] * dy;
            const float a1 = q.m[3] * dx + q.m[4] * dy;
            const float a2 = q.m[6] * dx + q.m[7] * dy;
#pragma unroll
            for (int j = 0; j < kPhZ; ++j) {
                const float dz = zf[j] - q.cz;
                const float q0 = a0 + q.m[2] * dz;
                const float q1 = a1 + q.m[5] * dz;
                const float q2 = a2 + q.m[8] * dz;
                const float f = q0 * q0 * q.ia[0] + q1 * q1 * q.ia[1] + q2 * q2 * q.ia[2];
                if (f <= 1.0f) v[j] += q.rho;
            }
        }
    }
    if (x < nx && y < ny) {
#pragma unroll
        for (int j = 0; j < kPhZ; ++j)
            if (j < nzl) vol[((size_t)(z0 + j) * ny + y) * nx + x] = v[j];
    }
}

// Synthetic acquisition of a sinogram p (in place): counts ~ Poisson(I0 e^{-kappa p}),
// then p~ = -ln(max(counts,1)/I0)/kappa. The Poisson draw uses its normal
// approximation round(lam + sqrt(lam) z), z from Box-Muller on a counter-based
// hash of (seed, index); with kappa chosen so that kappa*p <= ~2 the mean is
// >= I0 e^-2 (1353 at I0 = 1e4), where the approximation is tight.
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void noisy_log_kernel(float* __restrict__ p, size_t n, float I0, float kappa,
                                 unsigned long long seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long h = mix64(seed ^ mix64((unsigned long long)i));
    const float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = (float)((h >> 16) & 0xffffff) * (1.0f / 16777216.0f);
    const float z = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    const float lam = I0 * expf(-kappa * p[i]);
    const float cnt = fmaxf(rintf(lam + sqrtf(lam) * z), 1.0f);
    p[i] = -logf(cnt / I0) / kappa;
}

hipError_t launch_noisy_log(float* d_p, size_t n, float I0, float kappa, unsigned long long seed,
                            hipStream_t s) {
    hipLaunchKernelGGL(noisy_log_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_p, n,
                       I0, kappa, seed);
    return hipGetLastError();
}

hipError_t launch_phantom(const PhEll* d_ell, int n, int nx, int ny, int nz, float* d_vol,
                          hipStream_t s) {
    dim3 grid((nx + 63) / 64, (ny + 3) / 4, (nz + kPhZ - 1) / kPhZ);
    hipLaunchKernelGGL(phantom_kernel, grid, dim3(256), 0, s, d_ell, n, nx, ny, nz, d_vol);
    return hipGetLastError();
}

}  // namespace cthip
