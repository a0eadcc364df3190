// Output of Ah fields in the reference's on-disk formats (included into swm_capi.hip).
//
// Lat-lon writer, latlon_outputer_t / latlon_vec_outputer_t (outputer/latlon_outputer_mod.f90:
// 40-190) with the cubic regrid of create_latlon_outputer / create_latlon_vec_outputer
// (outputer_factory_mod.f90:45-190): every target point interpolates the field with a 4 x 4
// Lagrange stencil inside one panel's mesh points (latlon_interpolate, regridders/
// regrid_to_latlon_mod.f90:138-160: fout = fout + wx*wy*p over j1, then i1), vector components
// then rotate to lat-lon (transform_components_to_latlon_uv, :119-136).  The record is the
// REAL*4 rounding of the (Nlon, Nlat) FP64 result, written as direct-access record rec_num.
// Paneled writer, master_paneled_outputer_t (outputer/master_paneled_outputer_mod.f90:24-77): the
// REAL*4 rounding of the (N+1, N+1, 6) mesh points.  Record lengths count 4-byte words (the
// reference's Intel Fortran build), so record r starts at byte (r-1) * 4 * record words.
//
// The target tables (panel, stencil origin, weights, rotation) are built once on the host in
// FP64 (swmh_latlon_table) and kept in device memory; one launch regrids h or (u, v) and
// rounds to REAL*4, and the host writes the record.

#include <fcntl.h>
#include <unistd.h>

namespace swm {

struct LatlonTab {
    const int *panel, *x0, *y0;
    const double *wx, *wy, *rot;
    int npts, n, pitch;
};

template <typename T>
__device__ __forceinline__ double latlon_point(const LatlonTab &t, const T *f, int p, int x0, int y0,
                                               const double *wx, const double *wy) {
    double acc = 0.0;
#pragma unroll
    for (int j1 = 0; j1 < 4; ++j1)
#pragma unroll
        for (int i1 = 0; i1 < 4; ++i1) {
            const double v = (double)f[((long long)p * t.n + y0 + j1) * t.pitch + x0 + i1];
            acc = acc + wx[i1] * wy[j1] * v;
        }
    return acc;
}

// b == nullptr: scalar a -> oa; else covariant (a, b) = (u, v) -> lat-lon (oa, ob)
template <typename T>
__global__ void latlon_regrid(LatlonTab t, const T *__restrict__ a, const T *__restrict__ b, float *__restrict__ oa,
                              float *__restrict__ ob) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= t.npts) return;
    const int p = t.panel[q], x0 = t.x0[q], y0 = t.y0[q];
    double wx[4], wy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        wx[k] = t.wx[4 * q + k];
        wy[k] = t.wy[4 * q + k];
    }
    const double za = latlon_point(t, a, p, x0, y0, wx, wy);
    if (!b) {
        oa[q] = (float)za;
        return;
    }
    const double zb = latlon_point(t, b, p, x0, y0, wx, wy);
    const double *r = t.rot + 4 * q;   // u2u, u2v, v2u, v2v
    oa[q] = (float)(r[0] * za + r[2] * zb);
    ob[q] = (float)(r[1] * za + r[3] * zb);
}

// (i, j, panel) mesh points -> packed REAL*4, i fastest
template <typename T>
__global__ void paneled_pack(const T *__restrict__ f, int n, int pitch, float *__restrict__ out) {
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long total = 6LL * n * n;
    if (q >= total) return;
    const int i = (int)(q % n);
    const long long pj = q / n;
    out[q] = (float)f[pj * pitch + i];
}

}  // namespace swm

struct swm_outputer {
    swm_ctx *ctx = nullptr;
    int Nlat = 0, Nlon = 0, npts = 0;
    bool vec = false;
    int *panel = nullptr, *x0 = nullptr, *y0 = nullptr;
    double *wx = nullptr, *wy = nullptr, *rot = nullptr;
    float *out = nullptr;               // 2 x npts
    std::vector<float> host;
};

namespace {

void free_outputer(swm_outputer *o) {
    if (!o) return;
    void *ptrs[] = {o->panel, o->x0, o->y0, o->wx, o->wy, o->rot, o->out};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete o;
}

// write `count` REAL*4 values as direct-access record rec (1-based) of `count` words
int write_record(const char *file_name, int rec, const float *data, size_t count) {
    if (!file_name || !*file_name) return fail(SWM_EARG, "empty file name");
    if (rec < 1) return fail(SWM_EARG, "record numbers start at 1");
    const int fd = ::open(file_name, O_WRONLY | O_CREAT, 0644);
    if (fd < 0) return fail(SWM_EIO, std::string("cannot open ") + file_name);
    const size_t bytes = count * sizeof(float);
    const off_t off = (off_t)(rec - 1) * (off_t)bytes;
    size_t done = 0;
    while (done < bytes) {
        const ssize_t w = ::pwrite(fd, (const char *)data + done, bytes - done, off + (off_t)done);
        if (w <= 0) {
            ::close(fd);
            return fail(SWM_EIO, std::string("cannot write ") + file_name);
        }
        done += (size_t)w;
    }
    if (::close(fd) != 0) return fail(SWM_EIO, std::string("cannot close ") + file_name);
    return SWM_OK;
}

swm::LatlonTab latlon_tab(const swm_outputer *o) {
    const swm_ctx *c = o->ctx;
    return swm::LatlonTab{o->panel, o->x0, o->y0, o->wx, o->wy, o->rot, o->npts, c->n, c->pitch};
}

int enqueue_regrid(swm_outputer *o, const void *a, const void *b) {
    swm_ctx *c = o->ctx;
    const swm::LatlonTab t = latlon_tab(o);
    const dim3 blk(256), grd((unsigned)((o->npts + 255) / 256));
    float *ob = o->out + o->npts;
    if (c->prec == 64)
        hipLaunchKernelGGL(swm::latlon_regrid<double>, grd, blk, 0, c->stream, t, (const double *)a, (const double *)b,
                           o->out, b ? ob : nullptr);
    else
        hipLaunchKernelGGL(swm::latlon_regrid<float>, grd, blk, 0, c->stream, t, (const float *)a, (const float *)b,
                           o->out, b ? ob : nullptr);
    HIP_TRY(hipGetLastError());
    const size_t nb = (size_t)(b ? 2 : 1) * o->npts * sizeof(float);
    o->host.resize((size_t)2 * o->npts);
    HIP_TRY(hipMemcpyAsync(o->host.data(), o->out, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

int create_outputer(swm_ctx *c, int Nlat, int Nlon, const char *grid_type, bool vec, swm_outputer **out) {
    if (!c || !out || !grid_type) return fail(SWM_EARG, "NULL argument");
    *out = nullptr;
    if (std::string(grid_type) != "Ah")
        return fail(SWM_ENAME, std::string("the device path writes Ah fields only, not ") + grid_type);
    if (Nlat < 2 || Nlon < 1) return fail(SWM_EARG, "Nlat >= 2 and Nlon >= 1 required");
    if ((long long)Nlat * Nlon > (1LL << 30)) return fail(SWM_EARG, "lat-lon grid too large");
    if (c->n < 4) return fail(SWM_EARG, "the cubic stencil needs N+1 >= 4");
    swm_outputer *o = new swm_outputer();
    o->ctx = c;
    o->Nlat = Nlat;
    o->Nlon = Nlon;
    o->npts = Nlat * Nlon;
    o->vec = vec;
    const size_t np = (size_t)o->npts;
    std::vector<int> panel(np), x0(np), y0(np);
    std::vector<double> wx(4 * np), wy(4 * np), rot(vec ? 4 * np : 0);
    swmh_latlon_table(c->N, Nlat, Nlon, panel.data(), x0.data(), y0.data(), wx.data(), wy.data(),
                      vec ? rot.data() : nullptr);
    auto upload = [&](void **dst, const void *src, size_t bytes) -> int {
        HIP_TRY(hipMalloc(dst, bytes));
        HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return SWM_OK;
    };
    int rc;
    if ((rc = upload((void **)&o->panel, panel.data(), np * sizeof(int))) ||
        (rc = upload((void **)&o->x0, x0.data(), np * sizeof(int))) ||
        (rc = upload((void **)&o->y0, y0.data(), np * sizeof(int))) ||
        (rc = upload((void **)&o->wx, wx.data(), 4 * np * sizeof(double))) ||
        (rc = upload((void **)&o->wy, wy.data(), 4 * np * sizeof(double))) ||
        (vec && (rc = upload((void **)&o->rot, rot.data(), 4 * np * sizeof(double))))) {
        free_outputer(o);
        return rc;
    }
    if (hipMalloc((void **)&o->out, 2 * np * sizeof(float)) != hipSuccess) {
        free_outputer(o);
        return fail(SWM_EHIP, "hipMalloc of the output buffer failed");
    }
    *out = o;
    return SWM_OK;
}

}  // namespace

extern "C" int swm_latlon_outputer_create(swm_ctx *c, int Nlat, int Nlon, const char *scalar_grid_type,
                                          swm_outputer **out) {
    return create_outputer(c, Nlat, Nlon, scalar_grid_type, false, out);
}

extern "C" int swm_latlon_vec_outputer_create(swm_ctx *c, int Nlat, int Nlon, const char *staggering,
                                              const char *components_type, swm_outputer **out) {
    if (!components_type) return fail(SWM_EARG, "NULL argument");
    if (std::string(components_type) != "covariant")
        return fail(SWM_ENAME, std::string("the device path writes covariant components only, not ") +
                                   components_type);
    return create_outputer(c, Nlat, Nlon, staggering, true, out);
}

extern "C" int swm_outputer_destroy(swm_outputer *o) {
    free_outputer(o);
    return SWM_OK;
}

extern "C" int swm_latlon_write(swm_outputer *o, const void *f, const char *file_name, int rec_num) {
    if (!o || !f || !file_name) return fail(SWM_EARG, "NULL argument");
    int rc = enqueue_regrid(o, f, nullptr);
    if (rc) return rc;
    return write_record(file_name, rec_num, o->host.data(), (size_t)o->npts);
}

extern "C" int swm_latlon_vec_write(swm_outputer *o, const void *u, const void *v, const char *file_name_u,
                                    const char *file_name_v, int rec_num) {
    if (!o || !u || !v || !file_name_u || !file_name_v) return fail(SWM_EARG, "NULL argument");
    if (!o->vec) return fail(SWM_ESTATE, "not a vector outputer");
    int rc = enqueue_regrid(o, u, v);
    if (rc || (rc = write_record(file_name_u, rec_num, o->host.data(), (size_t)o->npts))) return rc;
    return write_record(file_name_v, rec_num, o->host.data() + o->npts, (size_t)o->npts);
}

extern "C" int swm_paneled_write(swm_ctx *c, const void *f, const char *file_name, int rec_num) {
    if (!c || !f || !file_name) return fail(SWM_EARG, "NULL argument");
    const long long total = 6LL * c->n * c->n;
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, (size_t)total * sizeof(float)));
    const dim3 blk(256), grd((unsigned)((total + 255) / 256));
    if (c->prec == 64)
        hipLaunchKernelGGL(swm::paneled_pack<double>, grd, blk, 0, c->stream, (const double *)f, c->n, c->pitch, d);
    else
        hipLaunchKernelGGL(swm::paneled_pack<float>, grd, blk, 0, c->stream, (const float *)f, c->n, c->pitch, d);
    std::vector<float> host((size_t)total);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), d, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SWM_EHIP, std::string("paneled output: ") + hipGetErrorString(e));
    return write_record(file_name, rec_num, host.data(), (size_t)total);
}
