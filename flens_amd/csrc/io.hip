// io.hip -- FLENS's binary matrix / vector files for device data.
//
// Format of flens::save / flens::load (flens/io/fullstorage/save.tcc:49-82,
// load.tcc:52-80 for GeMatrix; flens/io/array/save.tcc:43-66,
// load.tcc:43-68 for DenseVector), with FLENS's default IndexType int:
//   GeMatrix:    int32 numRows, numCols, firstRow, firstCol, then the
//                elements row by row (A(firstRow, firstCol), A(firstRow,
//                firstCol+1), ...), native byte order;
//   DenseVector: int32 length, firstIndex, then the elements.
// A device matrix (column-major) is written by transposing it on the device
// into row order and streaming 64 MiB slices through a pinned host buffer;
// loading reverses the path.  Files written here load in FLENS and vice
// versa, so they serve as fixtures shared with reference programs.
#include <cstdio>
#include <cstring>
#include "flx_common.h"
#include "context.h"

using namespace flx;

namespace {

constexpr size_t IO_SLICE = 64u << 20;   // bytes per staged slice

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(path ? fopen(path, mode) : nullptr) {}
    ~File() { if (f) fclose(f); }
};

struct Pinned {
    void *p = nullptr;
    ~Pinned() { if (p) (void)hipHostFree(p); }
};

// device bytes -> file, through the pinned slice buffer
int stream_out(hipStream_t s, FILE *f, const void *dev, size_t bytes)
{
    if (!bytes) return FLX_SUCCESS;
    Pinned buf;
    const size_t slice = bytes < IO_SLICE ? bytes : IO_SLICE;
    FLX_TRY(hipHostMalloc(&buf.p, slice));
    for (size_t off = 0; off < bytes; off += slice) {
        const size_t len = bytes - off < slice ? bytes - off : slice;
        FLX_TRY(hipMemcpyAsync(buf.p, (const char *)dev + off, len, hipMemcpyDeviceToHost, s));
        FLX_TRY(hipStreamSynchronize(s));
        if (fwrite(buf.p, 1, len, f) != len) return FLX_INVALID_ARG;
    }
    return FLX_SUCCESS;
}

// file -> device bytes
int stream_in(hipStream_t s, FILE *f, void *dev, size_t bytes)
{
    if (!bytes) return FLX_SUCCESS;
    Pinned buf;
    const size_t slice = bytes < IO_SLICE ? bytes : IO_SLICE;
    FLX_TRY(hipHostMalloc(&buf.p, slice));
    for (size_t off = 0; off < bytes; off += slice) {
        const size_t len = bytes - off < slice ? bytes - off : slice;
        if (fread(buf.p, 1, len, f) != len) return FLX_INVALID_ARG;
        FLX_TRY(hipMemcpyAsync((char *)dev + off, buf.p, len, hipMemcpyHostToDevice, s));
        FLX_TRY(hipStreamSynchronize(s));
    }
    return FLX_SUCCESS;
}

} // namespace

extern "C" int flx_save_gematrix(flx_handle h, const char *path, int m, int n, const double *A,
                                 int lda, int first_row, int first_col)
{
    if (!h || !path || m < 0 || n < 0 || (m > 0 && n > 0 && lda < m)) return FLX_INVALID_ARG;
    File fo(path, "wb");
    if (!fo.f) return FLX_INVALID_ARG;
    const int hdr[4] = {m, n, first_row, first_col};
    if (fwrite(hdr, sizeof(int), 4, fo.f) != 4) return FLX_INVALID_ARG;
    if (m == 0 || n == 0) return FLX_SUCCESS;
    const size_t bytes = sizeof(double) * (size_t)m * n;
    FLX_TRY(ensure(h, h->W, bytes));
    // row order: W (n x m column-major, ld n) = A^T
    FLX_TRY(transpose_f64(h->stream, A, lda, (double *)h->W.p, n, m, n));
    return stream_out(h->stream, fo.f, h->W.p, bytes);
}

extern "C" int flx_load_gematrix_header(const char *path, int *m, int *n, int *first_row,
                                        int *first_col)
{
    if (!path || !m || !n) return FLX_INVALID_ARG;
    File fi(path, "rb");
    if (!fi.f) return FLX_INVALID_ARG;
    int hdr[4];
    if (fread(hdr, sizeof(int), 4, fi.f) != 4 || hdr[0] < 0 || hdr[1] < 0) return FLX_INVALID_ARG;
    *m = hdr[0];
    *n = hdr[1];
    if (first_row) *first_row = hdr[2];
    if (first_col) *first_col = hdr[3];
    return FLX_SUCCESS;
}

extern "C" int flx_load_gematrix(flx_handle h, const char *path, double *A, int lda)
{
    if (!h || !path) return FLX_INVALID_ARG;
    File fi(path, "rb");
    if (!fi.f) return FLX_INVALID_ARG;
    int hdr[4];
    if (fread(hdr, sizeof(int), 4, fi.f) != 4 || hdr[0] < 0 || hdr[1] < 0) return FLX_INVALID_ARG;
    const int m = hdr[0], n = hdr[1];
    if (m == 0 || n == 0) return FLX_SUCCESS;
    if (lda < m) return FLX_INVALID_ARG;
    const size_t bytes = sizeof(double) * (size_t)m * n;
    FLX_TRY(ensure(h, h->W, bytes));
    FLX_RET(stream_in(h->stream, fi.f, h->W.p, bytes));
    // W holds rows: (n x m column-major, ld n) -> A (m x n, lda)
    FLX_TRY(transpose_f64(h->stream, (const double *)h->W.p, n, A, lda, n, m));
    FLX_TRY(hipStreamSynchronize(h->stream));
    return FLX_SUCCESS;
}

// DenseVector files; elem = 8 (double) or 4 (int, e.g. pivots)
extern "C" int flx_save_densevector(flx_handle h, const char *path, int n, const void *x, int elem,
                                    int first_index)
{
    if (!h || !path || n < 0 || (elem != 8 && elem != 4)) return FLX_INVALID_ARG;
    File fo(path, "wb");
    if (!fo.f) return FLX_INVALID_ARG;
    const int hdr[2] = {n, first_index};
    if (fwrite(hdr, sizeof(int), 2, fo.f) != 2) return FLX_INVALID_ARG;
    return stream_out(h->stream, fo.f, x, (size_t)n * elem);
}

extern "C" int flx_load_densevector(flx_handle h, const char *path, void *x, int capacity,
                                    int elem, int *n, int *first_index)
{
    if (!h || !path || !n || (elem != 8 && elem != 4)) return FLX_INVALID_ARG;
    File fi(path, "rb");
    if (!fi.f) return FLX_INVALID_ARG;
    int hdr[2];
    if (fread(hdr, sizeof(int), 2, fi.f) != 2 || hdr[0] < 0) return FLX_INVALID_ARG;
    *n = hdr[0];
    if (first_index) *first_index = hdr[1];
    if (hdr[0] > capacity) return FLX_INVALID_ARG;
    return stream_in(h->stream, fi.f, x, (size_t)hdr[0] * elem);
}
