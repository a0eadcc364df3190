// Linear combinations of state vectors (stvec_t%assign / update).
#include <hip/hip_runtime.h>

namespace swm {

// out = s1*v1 + s2*v2 (assign) or out = out + s1*v1 + s2*v2 (update) over the mesh points of
// all three fields (grid_field_mod.f90:430-485 via stvec_swm_mod.f90:38-106)
template <typename T, bool UPDATE>
__global__ void axpby_kernel(T *out, T s1, const T *v1, T s2, const T *v2, int n, int pitch, long long fs) {
    const long long total = 3LL * 6 * n * (long long)n;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total; k += (long long)gridDim.x * blockDim.x) {
        const long long row = k / n;
        const int i = (int)(k - row * n);
        const long long id = row * pitch + i;
        if (UPDATE)
            out[id] = out[id] + s1 * v1[id] + s2 * v2[id];
        else
            out[id] = s1 * v1[id] + s2 * v2[id];
    }
}

}  // namespace swm
