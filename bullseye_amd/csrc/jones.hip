// jones.hip -- invert every 2x2 Jones matrix once before gridding.
//
// Restates invert_all/invert (cpu_gpu_common/jones_2x2.h:88-126), which the reference runs in
// place on the caller's array before each facet_4_cor_corrections call
// (cpu_algorithm/wrapper.cpp:356-360).  Here the inverses go to a library buffer in the compute
// path's precision, the caller's array is left unchanged, and a singular matrix sets bit 0 of
// the device error flag (the reference throws).
#include "device_common.h"

namespace bam {

template <typename T>
__global__ void k_invert_jones(const visibility_base_type *in, T *out, uint64_t n, int *err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const visibility_base_type *m = in + 8 * i;
    const T ar = m[0], ai = m[1], br = m[2], bi = m[3], cr = m[4], ci = m[5], dr = m[6], di = m[7];
    // det = a d - b c
    const T detr = (ar * dr - ai * di) - (br * cr - bi * ci);
    const T deti = (ar * di + ai * dr) - (br * ci + bi * cr);
    if (detr == 0 && deti == 0) atomicOr(err, 1);
    const T den = detr * detr + deti * deti;
    const T ir = detr / den, ii = -deti / den;  // 1 / det
    T *o = out + 8 * i;
    // [d -b; -c a] / det
    o[0] = dr * ir - di * ii;
    o[1] = dr * ii + di * ir;
    o[2] = -(br * ir - bi * ii);
    o[3] = -(br * ii + bi * ir);
    o[4] = -(cr * ir - ci * ii);
    o[5] = -(cr * ii + ci * ir);
    o[6] = ar * ir - ai * ii;
    o[7] = ar * ii + ai * ir;
}

hipError_t invert_jones(const visibility_base_type *in, void *out, uint64_t n, bool fp64,
                        int *err, hipStream_t st) {
    if (!n) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (fp64)
        k_invert_jones<double><<<blocks, 256, 0, st>>>(in, (double *)out, n, err);
    else
        k_invert_jones<float><<<blocks, 256, 0, st>>>(in, (float *)out, n, err);
    return hipGetLastError();
}

}  // namespace bam
