// Residual-norm reductions on the GPU: the reference's host-side
// maximum(abs.(interior(X))) (examples/stokes_3d_inc_ve_T.jl:172-175) and a
// deterministic f64 sum of squares.  max|x| is exact and order-independent:
// blocks reduce in registers (wave64 shuffles) and LDS, then combine with one
// 64-bit atomic max per block on the bit pattern (non-negative doubles order
// like their unsigned bits; NaN's bits exceed +inf's, so NaN propagates:
// fmax drops NaN, so a NaN seen anywhere is carried as a flag instead).
#include <mutex>

#include "common.hpp"

namespace chmy {

constexpr int kRB = 256;           // threads per reduction block
constexpr int kSumBlocks = 1024;   // fixed partial count: deterministic sum order

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0) p[0] = 0ull;
}

// Work items are (x-row, 256-element segment) pairs of the box; a wave takes
// kRW consecutive items at a time and issues all their loads (kRW x 4 per
// lane) before the max, so that each wave keeps 16 loads in flight.  Item
// bases are formed once per item; no per-element division is needed.
constexpr int kRowChunk = 4;   // 64-wide slices of an item
constexpr int kSeg = 64 * kRowChunk;
constexpr int kRW = 4;         // items of a wave in flight

// item -> (segment, row j, plane k) in 32-bit arithmetic (the host checks
// that the items of a box number below 2^32): 64-bit divisions per item
// cost more than the item's loads; one segment per row needs none for it.
__device__ __forceinline__ void item_rc(int64_t it, uint32_t nseg, uint32_t ny, int64_t& sg, int64_t& j,
                                        int64_t& k) {
    const uint32_t u = (uint32_t)it;
    const uint32_t r = nseg == 1 ? u : u / nseg;
    const uint32_t kk = r / ny;
    sg = (int64_t)(u - r * nseg);
    j = (int64_t)(r - kk * ny);
    k = (int64_t)kk;
}

template <class T>
__global__ void __launch_bounds__(kRB) maxabs_kernel(Box3 b, DF<T> f, unsigned long long* out) {
    __shared__ double part[kRB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nx = b.len[0], ny = b.len[1];
    const int64_t nseg = (nx + kSeg - 1) / kSeg;
    const int64_t items = ny * b.len[2] * nseg;
    const int64_t nw = (int64_t)gridDim.x * (kRB / 64);
    double m = 0.0;
    bool isnan_ = false;
    for (int64_t i0 = ((int64_t)blockIdx.x * (kRB / 64) + w) * kRW; i0 < items; i0 += nw * kRW) {
        const T* base[kRW];
        int64_t len[kRW];
#pragma unroll
        for (int u = 0; u < kRW; ++u) {
            const int64_t it = i0 + u;
            int64_t sg, j, k;
            item_rc(it, (uint32_t)nseg, (uint32_t)ny, sg, j, k);
            const bool ok = it < items;
            base[u] = ok ? &f(b.lo[0] + sg * kSeg, b.lo[1] + j, b.lo[2] + k) : nullptr;
            len[u] = ok ? (nx - sg * kSeg < kSeg ? nx - sg * kSeg : kSeg) : 0;
        }
        double a[kRW][kRowChunk];
#pragma unroll
        for (int u = 0; u < kRW; ++u)
#pragma unroll
            for (int q = 0; q < kRowChunk; ++q) {
                const int x = 64 * q + lane;
                a[u][q] = x < len[u] ? fabs(as_double(base[u][x])) : 0.0;
            }
#pragma unroll
        for (int u = 0; u < kRW; ++u)
#pragma unroll
            for (int q = 0; q < kRowChunk; ++q) {
                isnan_ |= a[u][q] != a[u][q];
                m = fmax(m, a[u][q]);
            }
    }
    m = wave_max(m);
    const bool anynan = __any(isnan_);
    if (lane == 0) part[w] = anynan ? __longlong_as_double(0x7ff8000000000000ll) : m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = part[0];
        bool nan = v != v;
        for (int q = 1; q < kRB / 64; ++q) {
            nan |= part[q] != part[q];
            v = fmax(v, part[q]);
        }
        if (nan) v = __longlong_as_double(0x7ff8000000000000ll);
        atomicMax(out, (unsigned long long)__double_as_longlong(v));
    }
}

// Sum of squares over the same (row, segment) items, in a fixed order: the
// grid is always kSumBlocks blocks, so which thread adds which element, and
// the order of every addition, depend on the box only.
template <class T>
__global__ void __launch_bounds__(kRB) sumsq_kernel(Box3 b, DF<T> f, double* partial) {
    __shared__ double part[kRB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nx = b.len[0], ny = b.len[1];
    const int64_t nseg = (nx + kSeg - 1) / kSeg;
    const int64_t items = ny * b.len[2] * nseg;
    const int64_t nw = (int64_t)gridDim.x * (kRB / 64);
    double s = 0.0;
    for (int64_t i0 = ((int64_t)blockIdx.x * (kRB / 64) + w) * kRW; i0 < items; i0 += nw * kRW) {
        const T* base[kRW];
        int64_t len[kRW];
#pragma unroll
        for (int u = 0; u < kRW; ++u) {
            const int64_t it = i0 + u;
            int64_t sg, j, k;
            item_rc(it, (uint32_t)nseg, (uint32_t)ny, sg, j, k);
            const bool ok = it < items;
            base[u] = ok ? &f(b.lo[0] + sg * kSeg, b.lo[1] + j, b.lo[2] + k) : nullptr;
            len[u] = ok ? (nx - sg * kSeg < kSeg ? nx - sg * kSeg : kSeg) : 0;
        }
        double a[kRW][kRowChunk];
#pragma unroll
        for (int u = 0; u < kRW; ++u)
#pragma unroll
            for (int q = 0; q < kRowChunk; ++q) {
                const int x = 64 * q + lane;
                a[u][q] = x < len[u] ? as_double(base[u][x]) : 0.0;
            }
#pragma unroll
        for (int u = 0; u < kRW; ++u)
#pragma unroll
            for (int q = 0; q < kRowChunk; ++q) s += a[u][q] * a[u][q];
    }
    s = wave_sum(s);
    if (lane == 0) part[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = part[0];
        for (int q = 1; q < kRB / 64; ++q) v += part[q];
        partial[blockIdx.x] = v;
    }
}

__global__ void __launch_bounds__(kRB) sum_partials(const double* partial, int n, double* out) {
    __shared__ double part[kRB / 64];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += kRB) s += partial[q];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = part[0];
        for (int w = 1; w < kRB / 64; ++w) v += part[w];
        out[0] = v;
    }
}

// Library-held device scratch for reductions: one buffer per device, grown
// on demand, kept until process exit.
static std::mutex g_scratch_mu;
static void* g_scratch[64] = {nullptr};
static size_t g_scratch_sz[64] = {0};

static int scratch(size_t bytes, void** p) {
    int dev = 0;
    CHMY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_error(CHMY_E_ARG, "device index %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_sz[dev] < bytes) {
        if (g_scratch[dev]) CHMY_HIP(hipFree(g_scratch[dev]));
        g_scratch[dev] = nullptr;
        g_scratch_sz[dev] = 0;
        CHMY_HIP(hipMalloc(&g_scratch[dev], bytes));
        g_scratch_sz[dev] = bytes;
    }
    *p = g_scratch[dev];
    return 0;
}

static int reduce_box(const chmy_field_t* f, const chmy_box_t* box, Box3* out) {
    if (!f || !f->data) return set_error(CHMY_E_ARG, "field is NULL");
    if (f->nd < 1 || f->nd > 3) return set_error(CHMY_E_ARG, "field nd %d", f->nd);
    if (f->dtype != CHMY_F64 && f->dtype != CHMY_F32 && f->dtype != CHMY_BF16)
        return set_error(CHMY_E_DTYPE, "field dtype %d", f->dtype);
    chmy_box_t b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = 1; b.len[d] = d < f->nd ? f->n[d] : 1; }
    if (box) b = *box;
    for (int d = 0; d < f->nd; ++d) {
        const int64_t H = f->halo[d];
        if (b.len[d] < 0 || b.lo[d] < 1 - 2 * H || b.lo[d] + b.len[d] - 1 > f->n[d] + 2 * H)
            return set_error(CHMY_E_BOX, "reduction box outside the storage of the field along dim %d", d);
    }
    *out = box3(b, f->nd);
    const int64_t items = out->len[1] * out->len[2] * ((out->len[0] + kSeg - 1) / kSeg);
    if (items >= (int64_t(1) << 32)) return set_error(CHMY_E_SIZE, "reduction box of 2^32 rows or more");
    return 0;
}

// maxabs: (row, segment) items, two rounds of kRW items per wave at most
static unsigned row_blocks(const Box3& b) {
    const int64_t items = b.len[1] * b.len[2] * ((b.len[0] + kSeg - 1) / kSeg);
    int64_t nb = (items + (kRB / 64) * kRW * 2 - 1) / ((kRB / 64) * kRW * 2);
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    return (unsigned)nb;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_maxabs_dev(const chmy_field_t* f, const chmy_box_t* box, double* out_dev, void* stream) {
    Box3 b;
    CHMY_TRY(reduce_box(f, box, &b));
    if (!out_dev) return set_error(CHMY_E_ARG, "out_dev is NULL");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* o = reinterpret_cast<unsigned long long*>(out_dev);
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, o);
    if (f->dtype == CHMY_F64)
        hipLaunchKernelGGL(maxabs_kernel<double>, dim3(row_blocks(b)), dim3(kRB), 0, s, b, view<double>(*f), o);
    else if (f->dtype == CHMY_BF16)
        hipLaunchKernelGGL(maxabs_kernel<bf16s>, dim3(row_blocks(b)), dim3(kRB), 0, s, b, view<bf16s>(*f), o);
    else
        hipLaunchKernelGGL(maxabs_kernel<float>, dim3(row_blocks(b)), dim3(kRB), 0, s, b, view<float>(*f), o);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_maxabs(const chmy_field_t* f, const chmy_box_t* box, double* out_host, void* stream) {
    if (!out_host) return set_error(CHMY_E_ARG, "out_host is NULL");
    void* p = nullptr;
    CHMY_TRY(scratch(kSumBlocks * sizeof(double) + 64, &p));
    double* d = static_cast<double*>(p) + kSumBlocks;
    CHMY_TRY(chmy_maxabs_dev(f, box, d, stream));
    CHMY_HIP(hipMemcpyAsync(out_host, d, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream));
    CHMY_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int chmy_sumsq(const chmy_field_t* f, const chmy_box_t* box, double* out_host, void* stream) {
    Box3 b;
    CHMY_TRY(reduce_box(f, box, &b));
    if (!out_host) return set_error(CHMY_E_ARG, "out_host is NULL");
    void* p = nullptr;
    CHMY_TRY(scratch(kSumBlocks * sizeof(double) + 64, &p));
    double* partial = static_cast<double*>(p);
    double* res = partial + kSumBlocks;
    hipStream_t s = (hipStream_t)stream;
    if (f->dtype == CHMY_BF16)
        hipLaunchKernelGGL(sumsq_kernel<bf16s>, dim3(kSumBlocks), dim3(kRB), 0, s, b, view<bf16s>(*f), partial);
    else if (f->dtype == CHMY_F64)
        hipLaunchKernelGGL(sumsq_kernel<double>, dim3(kSumBlocks), dim3(kRB), 0, s, b, view<double>(*f), partial);
    else
        hipLaunchKernelGGL(sumsq_kernel<float>, dim3(kSumBlocks), dim3(kRB), 0, s, b, view<float>(*f), partial);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kRB), 0, s, partial, kSumBlocks, res);
    CHMY_HIP(hipGetLastError());
    CHMY_HIP(hipMemcpyAsync(out_host, res, sizeof(double), hipMemcpyDeviceToHost, s));
    CHMY_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
