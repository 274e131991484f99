// Field files and grid shrinking: the HDF5 field format of src/output.c:
// 63-108, src/output_mpi.c:52-141, src/input.c:182-270 and src/input_mpi.c:
// 73-128 (group Header with the BoxSize attribute, dataset Field/Field of
// doubles, N x N x N or padded N x N x (N+2)), and shrinkGrid(_dg) of
// src/shrink_grids.c:29-75.
//
// libhdf5 is loaded when a field file is first touched (MITOS_HDF5_LIB, then
// the default search path, then /opt/conda/lib), so the library itself has
// no link-time HDF5 dependency.  Device grids move through a pinned buffer a
// slab of x-planes at a time, written or read as hyperslabs of the dataset.
#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "../../include/mitos_hip.h"
#include "api_util.h"

using namespace mitos;

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t H5P_DEFAULT_ = 0;
constexpr unsigned H5F_ACC_RDONLY_ = 0x0000u;
constexpr unsigned H5F_ACC_TRUNC_ = 0x0002u;
constexpr int H5S_SELECT_SET_ = 0;

struct H5 {
    bool ok = false;
    herr_t (*open)();
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
    hid_t (*Gopen2)(hid_t, const char*, hid_t);
    herr_t (*Gclose)(hid_t);
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*Sselect_hyperslab)(hid_t, int, const hsize_t*, const hsize_t*, const hsize_t*,
                                const hsize_t*);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    herr_t (*Sclose)(hid_t);
    hid_t (*Acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    hid_t (*Aopen)(hid_t, const char*, hid_t);
    herr_t (*Awrite)(hid_t, hid_t, const void*);
    herr_t (*Aread)(hid_t, hid_t, void*);
    herr_t (*Aclose)(hid_t);
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    hid_t (*Dget_space)(hid_t);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*Dclose)(hid_t);
    hid_t* native_double;
};

template <typename F>
bool sym(void* lib, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

const H5* hdf5() {
    static std::mutex mu;
    static H5 h;
    static bool tried = false;
    std::lock_guard<std::mutex> lock(mu);
    if (tried) return h.ok ? &h : nullptr;
    tried = true;
    const char* env = getenv("MITOS_HDF5_LIB");
    const char* names[] = {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    void* lib = nullptr;
    for (const char* n : names)
        if (n && (lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!lib) return nullptr;
    bool ok = sym(lib, "H5open", h.open) && sym(lib, "H5Fcreate", h.Fcreate) &&
              sym(lib, "H5Fopen", h.Fopen) && sym(lib, "H5Fclose", h.Fclose) &&
              sym(lib, "H5Gcreate2", h.Gcreate2) && sym(lib, "H5Gopen2", h.Gopen2) &&
              sym(lib, "H5Gclose", h.Gclose) && sym(lib, "H5Screate_simple", h.Screate_simple) &&
              sym(lib, "H5Sselect_hyperslab", h.Sselect_hyperslab) &&
              sym(lib, "H5Sget_simple_extent_ndims", h.Sget_simple_extent_ndims) &&
              sym(lib, "H5Sget_simple_extent_dims", h.Sget_simple_extent_dims) &&
              sym(lib, "H5Sclose", h.Sclose) && sym(lib, "H5Acreate2", h.Acreate2) &&
              sym(lib, "H5Aopen", h.Aopen) && sym(lib, "H5Awrite", h.Awrite) &&
              sym(lib, "H5Aread", h.Aread) && sym(lib, "H5Aclose", h.Aclose) &&
              sym(lib, "H5Dcreate2", h.Dcreate2) && sym(lib, "H5Dopen2", h.Dopen2) &&
              sym(lib, "H5Dget_space", h.Dget_space) && sym(lib, "H5Dwrite", h.Dwrite) &&
              sym(lib, "H5Dread", h.Dread) && sym(lib, "H5Dclose", h.Dclose) &&
              sym(lib, "H5T_NATIVE_DOUBLE_g", h.native_double);
    if (!ok || h.open() < 0) return nullptr;
    h.ok = true;
    return &h;
}

// src/output.c:40-61: group Header, attribute BoxSize = {L, L, L}
int write_header(const H5* h, hid_t f, double boxlen) {
    const hid_t g = h->Gcreate2(f, "/Header", H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
    const hsize_t adims[1] = {3};
    const hid_t sp = h->Screate_simple(1, adims, nullptr);
    const hid_t a = h->Acreate2(g, "BoxSize", *h->native_double, sp, H5P_DEFAULT_, H5P_DEFAULT_);
    const double bs[3] = {boxlen, boxlen, boxlen};
    const herr_t e = (g < 0 || sp < 0 || a < 0) ? -1 : h->Awrite(a, *h->native_double, bs);
    if (a >= 0) h->Aclose(a);
    if (sp >= 0) h->Sclose(sp);
    if (g >= 0) h->Gclose(g);
    return e < 0 ? 1 : 0;
}

// the dataset Field/Field of dims (N, N, nz), created
hid_t create_field(const H5* h, hid_t f, int N, int nz) {
    const hid_t g = h->Gcreate2(f, "/Field", H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (g < 0) return -1;
    const hsize_t dims[3] = {(hsize_t)N, (hsize_t)N, (hsize_t)nz};
    const hid_t sp = h->Screate_simple(3, dims, nullptr);
    const hid_t d = sp < 0 ? -1
                           : h->Dcreate2(g, "Field", *h->native_double, sp, H5P_DEFAULT_,
                                         H5P_DEFAULT_, H5P_DEFAULT_);
    if (sp >= 0) h->Sclose(sp);
    h->Gclose(g);
    return d;
}

// write or read planes [x0, x0+nx) of the dataset d (dims N x N x nz),
// memory rows of nz values
herr_t slab_io(const H5* h, hid_t d, bool write, int N, int nz, long x0, long nx, double* buf) {
    const hid_t fs = h->Dget_space(d);
    if (fs < 0) return -1;
    const hsize_t start[3] = {(hsize_t)x0, 0, 0};
    const hsize_t count[3] = {(hsize_t)nx, (hsize_t)N, (hsize_t)nz};
    const hid_t ms = h->Screate_simple(3, count, nullptr);
    herr_t e = ms < 0 ? -1 : h->Sselect_hyperslab(fs, H5S_SELECT_SET_, start, nullptr, count,
                                                    nullptr);
    if (e >= 0)
        e = write ? h->Dwrite(d, *h->native_double, ms, fs, H5P_DEFAULT_, buf)
                  : h->Dread(d, *h->native_double, ms, fs, H5P_DEFAULT_, buf);
    if (ms >= 0) h->Sclose(ms);
    h->Sclose(fs);
    return e;
}

// the field dataset of an open file and its dims
hid_t open_field(const H5* h, hid_t f, hsize_t dims[3]) {
    const hid_t g = h->Gopen2(f, "Field", H5P_DEFAULT_);
    if (g < 0) return -1;
    const hid_t d = h->Dopen2(g, "Field", H5P_DEFAULT_);
    h->Gclose(g);
    if (d < 0) return -1;
    const hid_t sp = h->Dget_space(d);
    const int nd = sp < 0 ? -1 : h->Sget_simple_extent_ndims(sp);
    if (nd == 3) h->Sget_simple_extent_dims(sp, dims, nullptr);
    if (sp >= 0) h->Sclose(sp);
    if (nd != 3) {
        h->Dclose(d);
        return -2;
    }
    return d;
}

// planes per pinned-buffer pass: about 256 MiB
long slab_planes(int N, int nz) {
    const long plane = (long)N * nz * (long)sizeof(double);
    return std::max(1L, (256L << 20) / plane);
}

// src/shrink_grids.c:29-75: out(x, y, z) = sum over the K^3 input cells of
// the block, in the reference's (i, j, k) order, of factor * in
__global__ void k_shrink(double* out, const double* in, int M, int N, long rs, double factor) {
#pragma clang fp contract(off)
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)M * M * M) return;
    const int K = N / M;
    const int x = (int)(o / ((long)M * M)), y = (int)((o / M) % M), z = (int)(o % M);
    double s = 0;
    for (int a = 0; a < K; a++)
        for (int b = 0; b < K; b++) {
            const double* row = in + ((long)(x * K + a) * N + (y * K + b)) * rs + (long)z * K;
            for (int c = 0; c < K; c++) s += factor * row[c];
        }
    out[o] = s;
}

int shrink(double* out, const double* in, int M, int N, long rs) {
    if (!out || !in || M < 1 || N < 1) return MITOS_ERR_ARG;
    const int K = N / M;
    if (K * M != N) return 1;  // src/shrink_grids.c:31-34
    const long n = (long)M * M * M;
    hipLaunchKernelGGL(k_shrink, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, api_stream(),
                       out, in, M, N, rs, 1.0 / ((double)K * K * K));
    MITOS_TRY(launch_status());
    return api_sync();
}

}  // namespace

// src/output.c:63-108 (host array, N x N x N)
extern "C" int writeFieldFile(const double* box, int N, double boxlen, const char* fname) {
    if (!box || !fname || N < 1) return set_last_error(MITOS_ERR_ARG);
    const H5* h = hdf5();
    if (!h) return set_last_error(MITOS_ERR_UNSUPPORTED);
    const hid_t f = h->Fcreate(fname, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (f < 0) return set_last_error(1);
    int rc = write_header(h, f, boxlen);
    const hid_t d = rc ? -1 : create_field(h, f, N, N);
    if (d < 0 || slab_io(h, d, true, N, N, 0, N, const_cast<double*>(box)) < 0) rc = 1;
    if (d >= 0) h->Dclose(d);
    h->Fclose(f);
    return set_last_error(rc);
}

// src/input.c:182-270: allocates *box (N^3 doubles, malloc, host) and reads
// the unpadded part of an N x N x N or N x N x (N+2) dataset
extern "C" int readFieldFile(double** box, int* N, double* box_len, const char* fname) {
    if (!box || !N || !box_len || !fname) return set_last_error(MITOS_ERR_ARG);
    const H5* h = hdf5();
    if (!h) return set_last_error(MITOS_ERR_UNSUPPORTED);
    const hid_t f = h->Fopen(fname, H5F_ACC_RDONLY_, H5P_DEFAULT_);
    if (f < 0) return set_last_error(1);
    int rc = 0;
    const hid_t g = h->Gopen2(f, "Header", H5P_DEFAULT_);
    const hid_t a = g < 0 ? -1 : h->Aopen(g, "BoxSize", H5P_DEFAULT_);
    double bs[3] = {0, 0, 0};
    if (a < 0 || h->Aread(a, *h->native_double, bs) < 0) rc = 1;
    if (a >= 0) h->Aclose(a);
    if (g >= 0) h->Gclose(g);
    hsize_t dims[3] = {0, 0, 0};
    const hid_t d = rc ? -1 : open_field(h, f, dims);
    if (!rc && d < 0) rc = (d == -2) ? 2 : 1;
    const int n = (int)dims[0];
    if (!rc && (n != (int)dims[1] || (n != (int)dims[2] && n + 2 != (int)dims[2]))) rc = 2;
    if (!rc && (bs[0] != bs[1] || bs[1] != bs[2])) rc = 2;
    if (!rc) {
        double* b = (double*)malloc(sizeof(double) * (size_t)n * n * n);
        if (!b) rc = MITOS_ERR_ALLOC;
        if (!rc) {
            // the hyperslab of the first n values of each row (src/input.c:239-248)
            const hid_t fs = h->Dget_space(d);
            const hsize_t start[3] = {0, 0, 0};
            const hsize_t count[3] = {(hsize_t)n, (hsize_t)n, (hsize_t)n};
            const hid_t ms = h->Screate_simple(3, count, nullptr);
            herr_t e = h->Sselect_hyperslab(fs, H5S_SELECT_SET_, start, nullptr, count, nullptr);
            if (e >= 0) e = h->Dread(d, *h->native_double, ms, fs, H5P_DEFAULT_, b);
            h->Sclose(ms);
            h->Sclose(fs);
            if (e < 0) {
                free(b);
                rc = 1;
            } else {
                *box = b;
                *N = n;
                *box_len = bs[0];
            }
        }
    }
    if (d >= 0) h->Dclose(d);
    h->Fclose(f);
    return set_last_error(rc);
}

// src/output_mpi.c:77-141: the padded N x N x (N+2) device box of a grid in
// configuration space
extern "C" int writeFieldFile_dg(struct distributed_grid* dg, const char* fname) {
    if (!dg || !dg->box || !fname) return set_last_error(MITOS_ERR_ARG);
    if (dg->momentum_space == 1) return set_last_error(1);  // src/output_mpi.c:79-82
    const H5* h = hdf5();
    if (!h) return set_last_error(MITOS_ERR_UNSUPPORTED);
    const int N = dg->N, nz = N + 2;
    const long per = slab_planes(N, nz);
    double* buf = nullptr;
    if (hipHostMalloc((void**)&buf, sizeof(double) * per * N * nz, hipHostMallocDefault) !=
        hipSuccess)
        return set_last_error(MITOS_ERR_ALLOC);
    const hid_t f = h->Fcreate(fname, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    int rc = f < 0 ? 1 : write_header(h, f, dg->boxlen);
    const hid_t d = rc ? -1 : create_field(h, f, N, nz);
    if (d < 0) rc = 1;
    for (long x = 0; x < dg->NX && !rc; x += per) {
        const long nx = std::min(per, (long)dg->NX - x);
        if (hipMemcpyAsync(buf, dg->box + x * N * nz, sizeof(double) * nx * N * nz,
                           hipMemcpyDeviceToHost, api_stream()) != hipSuccess ||
            hipStreamSynchronize(api_stream()) != hipSuccess)
            rc = MITOS_ERR_DEVICE;
        else if (slab_io(h, d, true, N, nz, dg->X0 + x, nx, buf) < 0)
            rc = 1;
    }
    if (d >= 0) h->Dclose(d);
    if (f >= 0) h->Fclose(f);
    hipHostFree(buf);
    return set_last_error(rc);
}

// src/input_mpi.c:73-128: a padded or unpadded field file into the padded
// device box; the grid is then in configuration space
extern "C" int readFieldFile_dg(struct distributed_grid* dg, const char* fname) {
    if (!dg || !dg->box || !fname) return set_last_error(MITOS_ERR_ARG);
    const H5* h = hdf5();
    if (!h) return set_last_error(MITOS_ERR_UNSUPPORTED);
    const hid_t f = h->Fopen(fname, H5F_ACC_RDONLY_, H5P_DEFAULT_);
    if (f < 0) return set_last_error(1);
    const int N = dg->N;
    hsize_t dims[3] = {0, 0, 0};
    const hid_t d = open_field(h, f, dims);
    int rc = d < 0 ? 1 : 0;
    if (!rc && ((int)dims[0] != N || (int)dims[1] != N ||
                ((int)dims[2] != N && (int)dims[2] != N + 2)))
        rc = MITOS_ERR_DIMS;
    const int fz = (int)dims[2];  // values per file row
    const long per = slab_planes(N, N + 2);
    double* buf = nullptr;
    if (!rc && hipHostMalloc((void**)&buf, sizeof(double) * per * N * (N + 2),
                             hipHostMallocDefault) != hipSuccess)
        rc = MITOS_ERR_ALLOC;
    for (long x = 0; x < dg->NX && !rc; x += per) {
        const long nx = std::min(per, (long)dg->NX - x);
        if (slab_io(h, d, false, N, fz, dg->X0 + x, nx, buf) < 0) {
            rc = 1;
            break;
        }
        // rows of fz values into rows of N + 2 (an unpadded file leaves the
        // padding as it was, as the reference's hyperslab read does)
        double* dst = dg->box + x * N * (N + 2);
        hipError_t e;
        if (fz == N + 2)
            e = hipMemcpyAsync(dst, buf, sizeof(double) * nx * N * (N + 2),
                               hipMemcpyHostToDevice, api_stream());
        else
            e = hipMemcpy2DAsync(dst, sizeof(double) * (N + 2), buf, sizeof(double) * N,
                                 sizeof(double) * N, nx * N, hipMemcpyHostToDevice,
                                 api_stream());
        if (e != hipSuccess || hipStreamSynchronize(api_stream()) != hipSuccess)
            rc = MITOS_ERR_DEVICE;
    }
    if (buf) hipHostFree(buf);
    if (d >= 0) h->Dclose(d);
    h->Fclose(f);
    if (!rc) dg->momentum_space = 0;
    return set_last_error(rc);
}

// src/shrink_grids.c:29-52: out (device, M^3) from the padded box of a grid
extern "C" int shrinkGrid_dg(double* out, struct distributed_grid* in, int M, int N) {
    if (!in || !in->box) return set_last_error(MITOS_ERR_ARG);
    return set_last_error(shrink(out, in->box, M, N, N + 2));
}

// src/shrink_grids.c:55-75: device arrays, in unpadded N^3
extern "C" int shrinkGrid(double* out, const double* in, int M, int N) {
    return set_last_error(shrink(out, in, M, N, N));
}
