// Particle records of the V000 dump (src/cuda_SaveDump.cu:17-33): per
// particle the node index of its cell, the absolute position in cell units
// and the velocity, as one int32 and five float32 (24 bytes).  The absolute
// position is the cell's integer corner plus the in-cell offset, the sum the
// reference forms in Copy_GCPtoHCP (src/cuda_Particle.cu:70-93): one float
// addition on the FP32 path, the double sum rounded once to float on the FP64
// path.  One wave per cell writes its particles' records at their slots, so
// the file keeps the cell-major order the reference writes.
#include <hip/hip_runtime.h>

#include "pic_kernels.h"

namespace pic {

namespace {

template <typename R>
__global__ __launch_bounds__(256) void k_dump_pack(PView<R> p, int ncy, int ngy, int Csize, uint32_t *rec) {
    int lane;
    const int c = wave_cell(ncy, ngy, Csize, lane);
    if (c < 0) return;
    int k0, k1;
    cell_range(p, c, k0, k1);
    const int xi = c / ngy, yi = c % ngy;
    for (int k = k0 + lane; k < k1; k += 64) {
        float xa, ya;
        if constexpr (sizeof(R) == 4) {
            xa = (float)xi + p.x[k];
            ya = (float)yi + p.y[k];
        } else {
            xa = (float)((double)xi + p.x[k]);
            ya = (float)((double)yi + p.y[k]);
        }
        uint32_t *r = rec + (size_t)k * 6;
        r[0] = (uint32_t)c;
        r[1] = __float_as_uint(xa);
        r[2] = __float_as_uint(ya);
        r[3] = __float_as_uint((float)p.vx[k]);
        r[4] = __float_as_uint((float)p.vy[k]);
        r[5] = __float_as_uint((float)p.vz[k]);
    }
}

}  // namespace

template <typename R>
void launch_dump_pack(const PView<R> &p, int ncy, int ngy, int Csize, uint32_t *rec, hipStream_t s) {
    hipLaunchKernelGGL(k_dump_pack<R>, dim3((Csize + 3) / 4), dim3(256), 0, s, p, ncy, ngy, Csize, rec);
}

template void launch_dump_pack<float>(const PView<float> &, int, int, int, uint32_t *, hipStream_t);
template void launch_dump_pack<double>(const PView<double> &, int, int, int, uint32_t *, hipStream_t);

}  // namespace pic
