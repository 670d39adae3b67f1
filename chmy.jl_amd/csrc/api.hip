// Runtime part of the C ABI: errors, device selection, memory, descriptor checks.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include <algorithm>

#include "common.hpp"

namespace chmy {

static thread_local char g_msg[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    return code;
}

int hip_status(hipError_t e, const char* what) {
    set_error((int)e, "%s: %s", what, hipGetErrorString(e));
    return (int)e;
}

int check_grid_any(const chmy_grid_t* g) {
    if (!g) return set_error(CHMY_E_ARG, "grid is NULL");
    if (g->nd < 1 || g->nd > 3) return set_error(CHMY_E_ARG, "grid.nd = %d not in 1..3", g->nd);
    for (int d = 0; d < g->nd; ++d)
        if (g->n[d] < 1) return set_error(CHMY_E_ARG, "grid.n[%d] = %lld < 1", d, (long long)g->n[d]);
    for (int d = g->nd; d < 3; ++d)
        if (g->axis[d]) return set_error(CHMY_E_ARG, "grid.axis[%d] set beyond nd = %d", d, g->nd);
    return 0;
}

int check_grid(const chmy_grid_t* g) {
    CHMY_TRY(check_grid_any(g));
    if (nonuniform(g))
        return set_error(CHMY_E_ARG, "a grid with a FunctionAxis (non-uniform spacing) is accepted by the operators, "
                                     "bc! and set! only; this entry point needs uniform axes");
    return 0;
}

int check_box(const chmy_grid_t* g, const chmy_box_t* b, chmy_box_t* out) {
    chmy_box_t r;
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = 0;
        r.len[d] = d < g->nd ? g->n[d] + 2 : 1;
    }
    if (b) r = *b;
    for (int d = 0; d < g->nd; ++d) {
        if (r.len[d] < 0 || r.lo[d] < 0 || r.lo[d] + r.len[d] > g->n[d] + 2)
            return set_error(CHMY_E_BOX, "box dim %d [%lld, %lld) outside the launcher range [0, %lld]", d,
                             (long long)r.lo[d], (long long)(r.lo[d] + r.len[d]), (long long)(g->n[d] + 1));
    }
    *out = r;
    return 0;
}

void tensor_loc(int nd, int c, int* loc) {
    static const int L2[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 1, 0}};
    static const int L3[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    for (int d = 0; d < 3; ++d) loc[d] = nd == 2 ? L2[c][d] : L3[c][d];
}

int check_field(const chmy_grid_t* g, const chmy_field_t* f, const char* name, const int* loc, int dtype,
                bool allow_bf16) {
    if (!f || !f->data) return set_error(CHMY_E_ARG, "field %s is NULL", name);
    if (f->nd != g->nd) return set_error(CHMY_E_LOC, "field %s: nd %d != grid nd %d", name, f->nd, g->nd);
    if (f->dtype != CHMY_F64 && f->dtype != CHMY_F32 && f->dtype != CHMY_BF16)
        return set_error(CHMY_E_DTYPE, "field %s: unknown dtype %d", name, f->dtype);
    if (f->dtype == CHMY_BF16 && !allow_bf16)
        return set_error(CHMY_E_DTYPE, "field %s: bf16 storage is not accepted here", name);
    if (dtype >= 0 && f->dtype != dtype)
        return set_error(CHMY_E_DTYPE, "field %s: dtype %d, expected %d", name, f->dtype, dtype);
    for (int d = 0; d < g->nd; ++d) {
        if (f->loc[d] != 0 && f->loc[d] != 1) return set_error(CHMY_E_LOC, "field %s: bad loc", name);
        if (loc && loc[d] >= 0 && f->loc[d] != loc[d])
            return set_error(CHMY_E_LOC, "field %s: location along dim %d is %s, expected %s", name, d,
                             f->loc[d] ? "Vertex" : "Center", loc[d] ? "Vertex" : "Center");
        if (f->n[d] != g->n[d] + f->loc[d])
            return set_error(CHMY_E_LOC, "field %s: size %lld along dim %d does not match grid (%lld)", name,
                             (long long)f->n[d], d, (long long)(g->n[d] + f->loc[d]));
        if (f->halo[d] < 1) return set_error(CHMY_E_HALO, "field %s: halo %d < 1 along dim %d", name, f->halo[d], d);
    }
    return 0;
}

int check_center(const chmy_grid_t* g, const chmy_field_t* f, const char* name, int dtype, bool allow_bf16) {
    const int c[3] = {0, 0, 0};
    return check_field(g, f, name, c, dtype, allow_bf16);
}

int check_vector(const chmy_grid_t* g, const chmy_field_t* V, const char* name, int dtype, bool allow_bf16) {
    if (!V) return set_error(CHMY_E_ARG, "vector field %s is NULL", name);
    char nm[64];
    for (int D = 0; D < g->nd; ++D) {
        int loc[3] = {0, 0, 0};
        loc[D] = 1;
        std::snprintf(nm, sizeof nm, "%s[%d]", name, D);
        CHMY_TRY(check_field(g, &V[D], nm, loc, dtype, allow_bf16));
    }
    return 0;
}

int check_tensor(const chmy_grid_t* g, const chmy_field_t* tau, const char* name, int dtype, bool allow_bf16) {
    if (!tau) return set_error(CHMY_E_ARG, "tensor field %s is NULL", name);
    if (g->nd < 2) return set_error(CHMY_E_ARG, "tensor fields need nd >= 2");
    char nm[64];
    int nt = g->nd == 2 ? 3 : 6;
    for (int c = 0; c < nt; ++c) {
        int loc[3];
        tensor_loc(g->nd, c, loc);
        std::snprintf(nm, sizeof nm, "%s[%d]", name, c);
        CHMY_TRY(check_field(g, &tau[c], nm, loc, dtype, allow_bf16));
    }
    return 0;
}

}  // namespace chmy

using namespace chmy;

// Zero `bytes` bytes with vector stores (16 B per thread where aligned, a
// byte loop for the unaligned head and tail), grid-stride: a fill kernel, so a
// captured graph records a kernel node rather than a memset node.
__global__ void __launch_bounds__(256) zero_bytes_kernel(unsigned char* p, size_t bytes) {
    const size_t head = std::min<size_t>(bytes, (16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15);
    const size_t nv = (bytes - head) / 16;
    uint4* v = reinterpret_cast<uint4*>(p + head);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t t = t0; t < nv; t += stride) v[t] = make_uint4(0u, 0u, 0u, 0u);
    if (t0 == 0) {
        for (size_t b = 0; b < head; ++b) p[b] = 0;
        for (size_t b = head + nv * 16; b < bytes; ++b) p[b] = 0;
    }
}

extern "C" {

const char* chmy_last_error(void) { return g_msg; }

int chmy_device_count(int* count) {
    if (!count) return set_error(CHMY_E_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int chmy_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(CHMY_E_NODEV, "no HIP device visible");
    if (device < 0 || device >= n) return set_error(CHMY_E_ARG, "device %d not in [0, %d)", device, n);
    CHMY_HIP(hipSetDevice(device));
    return sia_scratch_ptr(nullptr);
}

int chmy_malloc_zeros(void** ptr, size_t bytes, void* stream) {
    if (!ptr) return set_error(CHMY_E_ARG, "ptr is NULL");
    CHMY_HIP(hipMalloc(ptr, bytes));
    CHMY_HIP(hipMemsetAsync(*ptr, 0, bytes, (hipStream_t)stream));
    return 0;
}

int chmy_stream_create(int priority, void** stream) {
    if (!stream) return set_error(CHMY_E_ARG, "stream is NULL");
    if (priority != 0) return set_error(CHMY_E_ARG, "only the default stream priority (0) is supported");
    hipStream_t s = nullptr;
    CHMY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return 0;
}

int chmy_stream_destroy(void* stream) {
    if (!stream) return set_error(CHMY_E_ARG, "stream is NULL");
    CHMY_HIP(hipStreamDestroy((hipStream_t)stream));
    return 0;
}

int chmy_memset_zero(void* ptr, size_t bytes, void* stream) {
    if (bytes == 0) return 0;
    if (!ptr) return set_error(CHMY_E_ARG, "ptr is NULL");
    const size_t nthreads = (bytes + 15) / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((nthreads + 255) / 256, 65535);
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       static_cast<unsigned char*>(ptr), bytes);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_free(void* ptr) {
    CHMY_HIP(hipFree(ptr));
    return 0;
}

int chmy_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    CHMY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    CHMY_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int chmy_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    CHMY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    CHMY_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int chmy_stream_sync(void* stream) {
    CHMY_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int chmy_coords(const chmy_grid_t* g, int dim, int loc, int64_t first, int64_t count, double* out) {
    CHMY_TRY(check_grid(g));
    if (dim < 0 || dim >= g->nd || !out || count < 0) return set_error(CHMY_E_ARG, "chmy_coords: bad arguments");
    const double o = g->origin[dim], h = g->spacing[dim];
    for (int64_t q = 0; q < count; ++q) {
        double i1 = (double)(first + q - 1);
        out[q] = loc == CHMY_VERTEX ? std::fma(i1, h, o) : std::fma(i1, h, std::fma(0.5, h, o));
    }
    return 0;
}

}  // extern "C"
