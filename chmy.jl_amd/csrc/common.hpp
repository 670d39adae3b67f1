// Shared device/host helpers of the gfx950 stencil library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/chmy_mi355x.h"

namespace chmy {

// ------------------------------------------------------------------ errors
int set_error(int code, const char* fmt, ...);
int hip_status(hipError_t e, const char* what);

#define CHMY_HIP(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_status(e_, #call); \
    } while (0)
#define CHMY_TRY(expr)               \
    do {                             \
        int rc_ = (expr);            \
        if (rc_ != 0) return rc_;    \
    } while (0)

// ------------------------------------------------------------------ bfloat16
// Storage type of bf16 fields (the forcing of FP32 correction sweeps):
// f32 -> bf16 rounds to nearest even, bf16 -> f32 is exact.
struct bf16s {
    uint16_t u;
};
__device__ __forceinline__ float bf16_to_f32(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);   // NaN stays a quiet NaN
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
// value conversions used by kernels that write in a chosen precision
template <class TO, class T>
__device__ __forceinline__ void put(TO& dst, T v) { dst = TO(v); }
template <class T>
__device__ __forceinline__ void put(bf16s& dst, T v) { dst.u = f32_to_bf16(float(v)); }
template <class T>
__device__ __forceinline__ double as_double(T v) { return (double)v; }
__device__ __forceinline__ double as_double(bf16s v) { return (double)bf16_to_f32(v.u); }

// ------------------------------------------------------------------ views
// Device view of a Field: p points at logical index (0,0,0), so logical I is
// p[i + j*sy + k*sz] (storage offset I + 2H - 1 per dim, field.jl:18).
template <class T>
struct DF {
    T* p;
    int64_t sy, sz;
    __device__ __forceinline__ T& operator()(int64_t i, int64_t j, int64_t k) const {
        return p[i + j * sy + k * sz];
    }
    __device__ __forceinline__ T& at(int64_t o) const { return p[o]; }
};

// Same, carrying locations and dimensionality for the generic operators.
template <class T>
struct DFL {
    T* p;
    int64_t st[3];
    int32_t loc[3];
    int32_t nd;
    __device__ __forceinline__ T& operator()(int64_t i, int64_t j, int64_t k) const {
        return p[i * st[0] + j * st[1] + k * st[2]];
    }
};

inline int64_t pad_of(const chmy_field_t& f, int d) { return f.n[d] + 4 * (int64_t)f.halo[d]; }

template <class T>
DF<T> view(const chmy_field_t& f) {
    DF<T> v;
    int64_t s1 = pad_of(f, 0);
    int64_t s2 = f.nd >= 2 ? s1 * pad_of(f, 1) : 0;
    v.sy = f.nd >= 2 ? s1 : 0;
    v.sz = f.nd >= 3 ? s2 : 0;
    int64_t off = (2 * (int64_t)f.halo[0] - 1);
    if (f.nd >= 2) off += (2 * (int64_t)f.halo[1] - 1) * v.sy;
    if (f.nd >= 3) off += (2 * (int64_t)f.halo[2] - 1) * v.sz;
    v.p = static_cast<T*>(f.data) + off;
    return v;
}

template <class T>
DFL<T> lview(const chmy_field_t& f) {
    DFL<T> v;
    DF<T> b = view<T>(f);
    v.p = b.p;
    v.st[0] = 1;
    v.st[1] = b.sy;
    v.st[2] = b.sz;
    for (int d = 0; d < 3; ++d) v.loc[d] = d < f.nd ? f.loc[d] : 0;
    v.nd = f.nd;
    return v;
}

// ------------------------------------------------------------------ validation
// All descriptors are checked on the host before any launch, so that every
// access a kernel makes (box ± 1 in each dim) lies inside the field storage.
// check_grid accepts uniform grids only; check_grid_any also grids with a
// FunctionAxis (the operators, bc! and set!).
int check_grid(const chmy_grid_t* g);
int check_grid_any(const chmy_grid_t* g);

// A FunctionAxis metric table on the device (chmy_grid_t.axis): entry (row, i)
// at p[row * rs + i + CHMY_AXIS_PAD]; p == nullptr on a uniform dim.
struct AxisT {
    const double* p;
    int64_t rs;
    __host__ __device__ __forceinline__ double at(int row, int64_t i) const { return p[row * rs + i + CHMY_AXIS_PAD]; }
};
struct Axes3 {
    AxisT a[3];
};
inline Axes3 axes_of(const chmy_grid_t* g) {
    Axes3 r;
    for (int d = 0; d < 3; ++d) {
        r.a[d].p = d < g->nd ? g->axis[d] : nullptr;
        r.a[d].rs = d < g->nd ? g->n[d] + 2 + 2 * CHMY_AXIS_PAD : 0;
    }
    return r;
}
inline bool nonuniform(const chmy_grid_t* g) {
    for (int d = 0; d < g->nd; ++d)
        if (g->axis[d]) return true;
    return false;
}
int check_box(const chmy_grid_t* g, const chmy_box_t* b, chmy_box_t* out);
// loc: expected location per dim (-1 = any); dtype: expected (-1 = any).
// bf16 storage is rejected unless allow_bf16 (only the forcing of FP32
// correction sweeps and the reductions read it).
int check_field(const chmy_grid_t* g, const chmy_field_t* f, const char* name, const int* loc, int dtype,
                bool allow_bf16 = false);
int check_center(const chmy_grid_t* g, const chmy_field_t* f, const char* name, int dtype, bool allow_bf16 = false);
int check_vector(const chmy_grid_t* g, const chmy_field_t* V, const char* name, int dtype, bool allow_bf16 = false);
int check_tensor(const chmy_grid_t* g, const chmy_field_t* tau, const char* name, int dtype, bool allow_bf16 = false);
void tensor_loc(int nd, int c, int* loc);
// Address on the current device of the SIA reverse sweep's file-scope device
// scratch (sia.hip), resolved once per device; p may be NULL.
int sia_scratch_ptr(void** p);

struct Box3 {
    int64_t lo[3];
    int64_t len[3];
};
inline Box3 box3(const chmy_box_t& b, int nd) {
    Box3 r;
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = d < nd ? b.lo[d] : 0;
        r.len[d] = d < nd ? b.len[d] : 1;
    }
    return r;
}

// A box with no points: the entry point launches nothing and returns 0 (a
// KernelAbstractions launch over an empty worksize does no work).
inline bool no_work(const Box3& b) { return b.len[0] <= 0 || b.len[1] <= 0 || b.len[2] <= 0; }

// Block of 256 threads as 64 (x, one wave) x 4 (y); grid covers the box.
inline dim3 grid_for(const Box3& b, dim3 blk) {
    return dim3((unsigned)((b.len[0] + blk.x - 1) / blk.x), (unsigned)((b.len[1] + blk.y - 1) / blk.y),
                (unsigned)((b.len[2] + blk.z - 1) / blk.z));
}

// Block id -> tile, so that the blocks an XCD receives (ids congruent mod 8,
// dispatched round-robin over the 8 XCDs) form one contiguous run of tiles and
// share halo rows in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int per = nb / 8;
    if (bid >= 8 * per) return bid;
    return (bid % 8) * per + bid / 8;
}

// Julia max/min against zero for floats (NaN propagates, max(-0.0, 0.0) = 0.0).
template <class T>
__device__ __forceinline__ T jmax0(T x) { return x != x ? x : (x > T(0) ? x : T(0)); }
template <class T>
__device__ __forceinline__ T jmin0(T x) { return x != x ? x : (x <= T(0) ? x : T(0)); }
template <class T>
__device__ __forceinline__ T jmax(T a, T b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}
template <class T>
__device__ __forceinline__ T visc_ve(T eta, T G, T dt) { return T(1) / (T(1) / eta + T(1) / (G * dt)); }

}  // namespace chmy
