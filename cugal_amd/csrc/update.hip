// Frank-Wolfe step of the quasi-permutation, cugal/pred.py:77-86 + :133-141
// (fused form of cuda_kernels/update_quasi_permutation.cu:7-35):
//   plan = exp(-C/reg + u_i + v_j)          (LOG duals), or
//   plan = u_i exp(-C/reg) v_j              (STANDARD scalings,
//                                            cugal/sinkhorn.py:20-23),
//   gap = sum (plan - P) C,  P += a (plan - P)
// One read of C, one read and one write of P: 12 n^2 bytes.  The gap is
// accumulated in FP64 and reduced in a fixed order (deterministic).
#include <mutex>

#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kUpdThreads = 256;
constexpr int kUpdMaxBlocks = 2048;
// Largest n of the entry point without a workspace (the same limit as the
// others: 64-bit offsets, int row indices).
constexpr int64_t kUpdateStaticMaxN = int64_t(1) << 30;

template <bool STD>
__device__ __forceinline__ float plan_of(float c, float neg_rinv, float ui, float vj) {
    if constexpr (STD) return ui * exp2_fast(c * neg_rinv * kLog2e) * vj;
    return exp2_fast((fmaf(c, neg_rinv, ui) + vj) * kLog2e);
}

template <bool VEC, bool STD>
__global__ __launch_bounds__(kUpdThreads) void fw_update(float* __restrict__ P, int64_t ldp,
                                                        const float* __restrict__ C, int n,
                                                        const float* __restrict__ u, const float* __restrict__ v,
                                                        float alpha, float neg_rinv, int rows_per_blk,
                                                        double* __restrict__ part) {
    __shared__ double scratch[kUpdThreads / kWave];
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    double gap = 0.0;
    for (int i = r0; i < r1; ++i) {
        const int64_t base = static_cast<int64_t>(i) * n;
        float* __restrict__ prow = P + static_cast<int64_t>(i) * ldp;
        const float ui = u[i];
        float rowgap = 0.0f;
        if (VEC) {
            for (int j = threadIdx.x * 4; j < n; j += kUpdThreads * 4) {
                const float4 c = ld_stream4(C + base + j);
                const float4 vv = *reinterpret_cast<const float4*>(v + j);
                float4 p = *reinterpret_cast<const float4*>(prow + j);
                const float e0 = plan_of<STD>(c.x, neg_rinv, ui, vv.x) - p.x;
                const float e1 = plan_of<STD>(c.y, neg_rinv, ui, vv.y) - p.y;
                const float e2 = plan_of<STD>(c.z, neg_rinv, ui, vv.z) - p.z;
                const float e3 = plan_of<STD>(c.w, neg_rinv, ui, vv.w) - p.w;
                rowgap += (e0 * c.x + e1 * c.y) + (e2 * c.z + e3 * c.w);
                p.x = fmaf(alpha, e0, p.x);
                p.y = fmaf(alpha, e1, p.y);
                p.z = fmaf(alpha, e2, p.z);
                p.w = fmaf(alpha, e3, p.w);
                *reinterpret_cast<float4*>(prow + j) = p;
            }
        } else {
            for (int j = threadIdx.x; j < n; j += kUpdThreads) {
                const float c = C[base + j];
                const float p = prow[j];
                const float d = plan_of<STD>(c, neg_rinv, ui, v[j]) - p;
                rowgap += d * c;
                prow[j] = fmaf(alpha, d, p);
            }
        }
        gap += static_cast<double>(rowgap);
    }
    gap = block_sum_double(gap, scratch);
    if (threadIdx.x == 0) part[blockIdx.x] = gap;
}

// The same step in FP64 (the reference's default dtype): plan, gap and P
// all in double.
template <bool STD>
__global__ __launch_bounds__(kUpdThreads) void fw_update64(double* __restrict__ P, int64_t ldp,
                                                          const double* __restrict__ C, int n,
                                                          const double* __restrict__ u, const double* __restrict__ v,
                                                          double alpha, double nrinv, int rows_per_blk,
                                                          double* __restrict__ part) {
    __shared__ double scratch[kUpdThreads / kWave];
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    double gap = 0.0;
    for (int i = r0; i < r1; ++i) {
        const double* __restrict__ crow = C + static_cast<int64_t>(i) * n;
        double* __restrict__ prow = P + static_cast<int64_t>(i) * ldp;
        const double ui = u[i];
        for (int j = threadIdx.x; j < n; j += kUpdThreads) {
            const double c = __builtin_nontemporal_load(crow + j);
            const double p = prow[j];
            const double plan = STD ? ui * exp(c * nrinv) * v[j] : exp(fma(c, nrinv, ui) + v[j]);
            const double d = plan - p;
            gap += d * c;
            prow[j] = fma(alpha, d, p);
        }
    }
    gap = block_sum_double(gap, scratch);
    if (threadIdx.x == 0) part[blockIdx.x] = gap;
}

__global__ void sum_parts(const double* __restrict__ in, int n, double* __restrict__ out) {
    __shared__ double scratch[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += in[i];
    s = block_sum_double(s, scratch);
    if (threadIdx.x == 0) *out = s;
}

int update_blocks(int64_t n, int* rows_per_blk) {
    const int64_t rpb = std::max<int64_t>(1, div_up(n, kUpdMaxBlocks));
    *rows_per_blk = static_cast<int>(rpb);
    return static_cast<int>(div_up(n, rpb));
}

// Partial gaps of the entry point that allocates nothing: one double per
// block of fw_update (at most kUpdMaxBlocks for any n) plus the 256 bytes of
// slack the workspace size includes, a fixed 16.25 KiB that each device
// holds for the life of the process.  Host code never passes this name to a
// kernel; it passes the device address looked up per device.
constexpr int kScratchDoubles = kUpdMaxBlocks + 32;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the current device, looked up once per device.
int scratch_address(double** out) {
    int dev = 0;
    CG_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CG_EUNSUPPORTED, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !ptr)
            return fail(CG_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(rc));
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return CG_OK;
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_update_workspace(int64_t n) {
    if (n <= 0) return 0;
    int rpb = 0;
    return sizeof(double) * static_cast<size_t>(update_blocks(n, &rpb)) + 256;
}

namespace {

int run_update(bool standard, float* P, int64_t ldp, const float* C, int64_t n, const float* u, const float* v,
               float alpha,
               float regularization, double* gap_device, void* workspace, size_t workspace_bytes,
               hipStream_t s, const char* who) {
    if (!P || !C || !u || !v || !gap_device || n <= 0 || n > (int64_t(1) << 30) || ldp < n)
        return fail(CG_EINVAL, "%s: bad arguments", who);
    if (!(regularization > 0.0f)) return fail(CG_EINVAL, "regularization must be > 0");
    if (!workspace || workspace_bytes < cg_update_workspace(n)) return fail(CG_EWORKSPACE, "%s: workspace too small", who);
    int rpb = 0;
    const int blocks = update_blocks(n, &rpb);
    double* part = static_cast<double*>(workspace);
    const bool vec = (n % 4 == 0) && (ldp % 4 == 0) && aligned16(P) && aligned16(C) && aligned16(v);
    const float neg_rinv = -1.0f / regularization;
    const int ni = static_cast<int>(n);
    if (vec && standard)
        hipLaunchKernelGGL((fw_update<true, true>), dim3(blocks), dim3(kUpdThreads), 0, s, P, ldp, C, ni, u, v, alpha,
                           neg_rinv, rpb, part);
    else if (vec)
        hipLaunchKernelGGL((fw_update<true, false>), dim3(blocks), dim3(kUpdThreads), 0, s, P, ldp, C, ni, u, v, alpha,
                           neg_rinv, rpb, part);
    else if (standard)
        hipLaunchKernelGGL((fw_update<false, true>), dim3(blocks), dim3(kUpdThreads), 0, s, P, ldp, C, ni, u, v, alpha,
                           neg_rinv, rpb, part);
    else
        hipLaunchKernelGGL((fw_update<false, false>), dim3(blocks), dim3(kUpdThreads), 0, s, P, ldp, C, ni, u, v, alpha,
                           neg_rinv, rpb, part);
    if (int rc = check_launch("fw_update")) return rc;
    hipLaunchKernelGGL(sum_parts, dim3(1), dim3(256), 0, s, part, blocks, gap_device);
    return check_launch("sum_parts");
}

}  // namespace

int cg_update_quasi_permutation_log(float* P, int64_t ld_p, const float* C, int64_t n, const float* u,
                                    const float* v, float alpha, float regularization, double* gap_device,
                                    void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    return run_update(false, P, ld_p, C, n, u, v, alpha, regularization, gap_device, workspace, workspace_bytes,
                      reinterpret_cast<hipStream_t>(stream), "cg_update_quasi_permutation_log");
}

int cg_update_quasi_permutation_log_static(float* P, int64_t ld_p, const float* C, int64_t n, const float* u,
                                           const float* v, float alpha, float regularization, double* gap_device,
                                           cg_stream_t stream) {
    const char* who = "cg_update_quasi_permutation_log_static";
    if (n <= 0 || n > kUpdateStaticMaxN)
        return fail(CG_EINVAL, "%s: n = %lld outside 1..%lld", who, static_cast<long long>(n),
                    static_cast<long long>(kUpdateStaticMaxN));
    int rpb = 0;
    if (update_blocks(n, &rpb) > kUpdMaxBlocks) return fail(CG_EINVAL, "%s: too many partial sums", who);
    double* part = nullptr;
    if (int rc = scratch_address(&part)) return rc;
    return run_update(false, P, ld_p, C, n, u, v, alpha, regularization, gap_device, part,
                      sizeof(double) * kScratchDoubles, reinterpret_cast<hipStream_t>(stream), who);
}

namespace {

int run_update64(bool standard, double* P, int64_t ldp, const double* C, int64_t n, const double* u,
                 const double* v, double alpha, double regularization, double* gap_device, void* workspace,
                 size_t workspace_bytes, hipStream_t s, const char* who) {
    if (!P || !C || !u || !v || !gap_device || n <= 0 || n > (int64_t(1) << 30) || ldp < n)
        return fail(CG_EINVAL, "%s: bad arguments", who);
    if (!(regularization > 0.0)) return fail(CG_EINVAL, "regularization must be > 0");
    if (!workspace || workspace_bytes < cg_update_workspace(n)) return fail(CG_EWORKSPACE, "%s: workspace too small", who);
    int rpb = 0;
    const int blocks = update_blocks(n, &rpb);
    double* part = static_cast<double*>(workspace);
    const double nrinv = -1.0 / regularization;
    if (standard)
        hipLaunchKernelGGL(fw_update64<true>, dim3(blocks), dim3(kUpdThreads), 0, s, P, ldp, C, static_cast<int>(n),
                           u, v, alpha, nrinv, rpb, part);
    else
        hipLaunchKernelGGL(fw_update64<false>, dim3(blocks), dim3(kUpdThreads), 0, s, P, ldp, C, static_cast<int>(n),
                           u, v, alpha, nrinv, rpb, part);
    if (int rc = check_launch("fw_update64")) return rc;
    hipLaunchKernelGGL(sum_parts, dim3(1), dim3(256), 0, s, part, blocks, gap_device);
    return check_launch("sum_parts");
}

}  // namespace

int cg_update_quasi_permutation_f64(int standard, double* P, int64_t ld_p, const double* C, int64_t n,
                                    const double* u, const double* v, double alpha, double regularization,
                                    double* gap_device, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    return run_update64(standard != 0, P, ld_p, C, n, u, v, alpha, regularization, gap_device, workspace,
                        workspace_bytes, reinterpret_cast<hipStream_t>(stream), "cg_update_quasi_permutation_f64");
}

int cg_update_quasi_permutation_standard(float* P, int64_t ld_p, const float* C, int64_t n, const float* u,
                                         const float* v, float alpha, float regularization, double* gap_device,
                                         void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    return run_update(true, P, ld_p, C, n, u, v, alpha, regularization, gap_device, workspace, workspace_bytes,
                      reinterpret_cast<hipStream_t>(stream), "cg_update_quasi_permutation_standard");
}

}  // extern "C"
