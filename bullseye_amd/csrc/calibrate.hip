// calibrate.hip -- the device's FP64 atomic-add rate, the ceiling of every grid flush.
//
// Every gridder flushes its register sums with FP64 atomic adds to the device grid, and on this
// GPU those execute on the memory side whatever the address pattern (the per-XCD L2s are not
// coherent with each other).  The w-projection class kernels are bound by that rate, so the
// benchmark states their flush rate against this measurement: 2^27 adds of 1.0 at hashed
// addresses of a 16 MB buffer, timed after one warm-up launch.
#include "host_common.h"

namespace bam {

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// each thread adds 1.0 at `iters` hashed indexes of [0, mask]
__global__ void k_atomic_rate(double *g, uint32_t mask, int iters) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = 0; i < iters; ++i) unsafeAtomicAdd(g + (mix32(t * 131u + (uint32_t)i) & mask), 1.0);
}

}  // namespace

}  // namespace bam

using namespace bam;

extern "C" int bullseye_fp64_atomic_rate(double *ops_per_second) {
    if (!ops_per_second) return fail("ops_per_second is NULL");
    const uint32_t n = 1u << 21;                  // 2^21 doubles = 16 MB
    const int blocks = 8192, threads = 256, iters = 64;
    DevBuf buf;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    float ms = 0;
    std::vector<double> host(n);
    double sum = 0;
    if (buf.ensure((size_t)n * 8) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail("allocating the atomic-rate buffer, stream or events");
        goto done;
    }
    if (fill_f64(buf.as<double>(), n, 0.0, st) != hipSuccess) {
        rc = fail("clearing the atomic-rate buffer");
        goto done;
    }
    k_atomic_rate<<<blocks, threads, 0, st>>>(buf.as<double>(), n - 1, iters);   // warm-up
    (void)hipEventRecord(e0, st);
    k_atomic_rate<<<blocks, threads, 0, st>>>(buf.as<double>(), n - 1, iters);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
        hipMemcpy(host.data(), buf.p, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = fail("timing the atomic-rate kernel");
        goto done;
    }
    for (double v : host) sum += v;
    // every add landed: 2 launches x blocks x threads x iters
    if (sum != 2.0 * blocks * threads * iters || !(ms > 0)) {
        rc = fail("atomic-rate check failed (sum %.0f)", sum);
        goto done;
    }
    *ops_per_second = (double)blocks * threads * iters / (ms * 1e-3);
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    buf.release();
    return rc;
}
