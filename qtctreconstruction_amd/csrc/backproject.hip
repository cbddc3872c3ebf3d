// backproject.hip — voxel-driven backprojection kernels (the tile routine and
// its numerics are in bp_tile.h).
//
//   bp_kernel<T,A,CHECK,Z>        one volume, grid (tiles_x, tiles_y, z-block
//                                 groups of Z)
//   bp_persistent_kernel<T,A,Z>   many volumes of uneven size in one launch:
//                                 work items (job, z-block group, tile) are
//                                 taken from 8 atomic cursors, one per eighth
//                                 of the list, until all are exhausted
//
// (T, A, Z) = (voxel tile edge, angles per LDS stage, z-blocks per tile) is a
// tuning choice; the instantiated set is kBpConfigs and the autotuner picks
// among them.
#include "bp_tile.h"

namespace cthip {

// Waves per SIMD that the tile's LDS lets a CU hold (a 256-thread block is
// one wave per SIMD; 160 KiB of LDS per CU, at most 8 waves per SIMD). The
// kernels ask the compiler for that occupancy, which bounds their VGPRs.
template <int T, int A, int Z>
constexpr int bp_waves_per_eu() {
    return (160 * 1024) / (int)sizeof(BpSmem<T, A, Z>) < 8
               ? (160 * 1024) / (int)sizeof(BpSmem<T, A, Z>)
               : 8;
}

template <int T, int A, bool CHECK, int Z>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<T, A, Z>())))
void bp_kernel(BpArgs p) {
    __shared__ BpSmem<T, A, Z> sm;
    // The dispatcher deals workgroups to the 8 XCDs in turn. Renumbered so
    // that XCD k takes the k-th eighth of the (z-group, tile) order, the
    // tiles of a z-group run on one XCD and read its filtered rows through
    // one L2 instead of eight.
    const int gxy = gridDim.x * gridDim.y, nb = gxy * gridDim.z;
    int w = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if ((nb & 7) == 0) w = (w & 7) * (nb >> 3) + (w >> 3);
    const int zp = w / gxy, t = w - zp * gxy;
    const int by = t / gridDim.x, bx = t - by * gridDim.x;
    bp_tile<T, A, CHECK, Z>(p, bx * T, by * T, zp, sm);
}

template <int T, int A, int Z>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<T, A, Z>())))
void bp_persistent_kernel(const BpJob* __restrict__ jobs, const int* __restrict__ hdr,
                          int* __restrict__ counter) {
    __shared__ BpSmem<T, A, Z> sm;
    __shared__ int item_s;
    const int njobs = hdr[0], total = hdr[1];
    // counter[0..7] are cursors into 8 contiguous eighths of the item list
    // (items run job by job and z-group by z-group, so an eighth is a run of
    // whole z-groups apart from its two ends). Workgroups are dealt to the 8
    // XCDs in turn: a workgroup
    // drains its XCD's eighth first, so a z-group's filtered rows are read
    // through one L2, then helps with the other eighths in turn. Every
    // workgroup leaves once all eight cursors have passed their ends.
    const int home = blockIdx.x & 7;
    for (int k = 0; k < 8; ++k) {
        const int part = (home + k) & 7;
        const int lo = (int)(((long long)total * part) >> 3);
        const int hi = (int)(((long long)total * (part + 1)) >> 3);
        for (;;) {
            if (threadIdx.x == 0) item_s = lo + atomicAdd(counter + part, 1);
            __syncthreads();
            const int item = item_s;
            __syncthreads();
            if (item >= hi) break;
            int j = 0;
            while (j + 1 < njobs && jobs[j + 1].item_begin <= item) ++j;
            const BpJob& jb = jobs[j];
            const int tx = (jb.a.N + T - 1) / T;
            const int per = tx * tx;
            int local = item - jb.item_begin;
            const int zp = local / per;  // z-block group
            local -= zp * per;
            const int by = local / tx, bx = local - by * tx;
            if (jb.a.check)
                bp_tile<T, A, true, Z>(jb.a, bx * T, by * T, zp, sm);
            else
                bp_tile<T, A, false, Z>(jb.a, bx * T, by * T, zp, sm);
        }
    }
}

__global__ void zero_ints_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

const BpConfig kBpConfigs[kNumBpConfigs] = {{32, 32, 1}, {32, 16, 1}, {32, 64, 1}, {16, 32, 1},
                                            {16, 64, 1}, {32, 8, 2},  {32, 16, 2}, {16, 16, 2}};

template <int T, int A, int Z>
static hipError_t launch_one(const BpArgs& p, hipStream_t s) {
    const int t = (p.N + T - 1) / T;
    dim3 grid(t, t, ((p.nz + kZB - 1) / kZB + Z - 1) / Z);
    if (p.check)
        hipLaunchKernelGGL((bp_kernel<T, A, true, Z>), grid, dim3(256), 0, s, p);
    else
        hipLaunchKernelGGL((bp_kernel<T, A, false, Z>), grid, dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_backproject_cfg(int cfg, const BpArgs& p, hipStream_t s) {
    switch (cfg) {
        case 0: return launch_one<32, 32, 1>(p, s);
        case 1: return launch_one<32, 16, 1>(p, s);
        case 2: return launch_one<32, 64, 1>(p, s);
        case 3: return launch_one<16, 32, 1>(p, s);
        case 4: return launch_one<16, 64, 1>(p, s);
        case 5: return launch_one<32, 8, 2>(p, s);
        case 6: return launch_one<32, 16, 2>(p, s);
        case 7: return launch_one<16, 16, 2>(p, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_backproject(const float4* d_filt, int nz, int na, int sq, int N,
                              const float* d_cos, const float* d_sin, float scale,
                              const float* d_mm, float* d_out, hipStream_t s, int cfg) {
    BpArgs p;
    p.filt = d_filt;
    p.out = d_out;
    p.cos_t = d_cos;
    p.sin_t = d_sin;
    p.mm = d_mm;
    p.nz = nz;
    p.na = na;
    p.sq = sq;
    p.N = N;
    p.scale = scale;
    p.check = bp_needs_check(N, sq) ? 1 : 0;
    return launch_backproject_cfg(cfg, p, s);
}

template <int T, int A, int Z>
static const void* persistent_fn() {
    return (const void*)bp_persistent_kernel<T, A, Z>;
}

static const void* bp_persistent_ptr(int cfg) {
    switch (cfg) {
        case 0: return persistent_fn<32, 32, 1>();
        case 1: return persistent_fn<32, 16, 1>();
        case 2: return persistent_fn<32, 64, 1>();
        case 3: return persistent_fn<16, 32, 1>();
        case 4: return persistent_fn<16, 64, 1>();
        case 5: return persistent_fn<32, 8, 2>();
        case 6: return persistent_fn<32, 16, 2>();
        case 7: return persistent_fn<16, 16, 2>();
        default: return nullptr;
    }
}

int bp_persistent_grid(int cfg, int device) {
    const void* fn = bp_persistent_ptr(cfg);
    if (!fn) return 0;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return (per_cu < 1 ? 1 : per_cu) * cus;
}

hipError_t launch_bp_persistent(int cfg, const BpJob* d_jobs, const int* d_hdr, int* d_counter,
                                int grid, hipStream_t s) {
    switch (cfg) {
        case 0: hipLaunchKernelGGL((bp_persistent_kernel<32, 32, 1>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 1: hipLaunchKernelGGL((bp_persistent_kernel<32, 16, 1>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 2: hipLaunchKernelGGL((bp_persistent_kernel<32, 64, 1>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 3: hipLaunchKernelGGL((bp_persistent_kernel<16, 32, 1>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 4: hipLaunchKernelGGL((bp_persistent_kernel<16, 64, 1>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 5: hipLaunchKernelGGL((bp_persistent_kernel<32, 8, 2>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 6: hipLaunchKernelGGL((bp_persistent_kernel<32, 16, 2>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        case 7: hipLaunchKernelGGL((bp_persistent_kernel<16, 16, 2>), dim3(grid), dim3(256), 0, s, d_jobs, d_hdr, d_counter); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero_ints(int* d_p, int n, hipStream_t s) {
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_p, n);
    return hipGetLastError();
}

}  // namespace cthip
