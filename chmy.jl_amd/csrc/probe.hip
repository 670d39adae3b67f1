// Measurement aid: the HBM ceiling of an access pattern on its own.  The
// kernels below read nr FP64 arrays and write nw, point for point, in the
// fused 3D iteration's access patterns, with no work beyond one sum per point,
// so that the fused kernel's time can be set against the time its loads and
// stores take by themselves.  Arrays are Center storage of an n^3 grid
// ((n+4)^3 elements, logical -1 .. n+2); the points are the Launcher box
// [0, n+1]^3.
#include <cstring>
#include <type_traits>

#include "common.hpp"

namespace chmy {

constexpr int kProbeMaxR = 24, kProbeMaxW = 16;

struct ProbeArrays {
    const double* r[kProbeMaxR];
    double* w[kProbeMaxW];
    int nr, nw;
};

struct ProbeGeo {
    int64_t m;             // points per dim of the box (n + 2)
    int64_t px, pxy;       // storage strides (n + 4, (n + 4)^2)
    int ntx, nty, ntz, kchunk, xcd;
};

// NR and NW are compile-time counts, so that every load of a point is issued
// before any of them is waited for, as in the fused kernel.
template <int NR>
__device__ __forceinline__ double probe_point(const ProbeArrays& a, int64_t o) {
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) v[q] = a.r[q][o];
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NR; ++q) s += v[q];
    return s;
}

template <int NW>
__device__ __forceinline__ void probe_store(const ProbeArrays& a, int64_t o, double s) {
#pragma unroll
    for (int q = 0; q < NW; ++q) a.w[q][o] = s + q;
}

// storage offset of logical point (x, y, z), each in [-1, n+2] (storage 0 .. n+3)
__device__ __forceinline__ int64_t probe_off(const ProbeGeo& g, int64_t x, int64_t y, int64_t z) {
    return (z + 1) * g.pxy + (y + 1) * g.px + (x + 1);
}

// Linear order: one box point per lane, x fastest.
template <int NR, int NW>
__global__ void __launch_bounds__(256) probe_linear(ProbeArrays a, ProbeGeo g) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= g.m * g.m * g.m) return;
    const int64_t x = t % g.m, y = (t / g.m) % g.m, z = t / (g.m * g.m);
    const int64_t o = probe_off(g, x, y, z);
    probe_store<NW>(a, o, probe_point<NR>(a, o));
}

__device__ __forceinline__ int probe_tile(int b, int nb, int xcd) {
    if (!xcd) return b;
    const int per = nb / 8;
    return b >= 8 * per ? b : (b % 8) * per + b / 8;
}

// The fused kernel's tiles: 64 x BY lanes (rows of one wave), outputs on lanes
// 1..62 and rows 1..BY-1, a z-march over kchunk planes plus the one before;
// tiles x fastest, then y, then z, remapped per XCD when xcd != 0.
template <int BY, int NR, int NW>
__global__ void __launch_bounds__(64 * BY) probe_tiled(ProbeArrays a, ProbeGeo g) {
    const int nb = g.ntx * g.nty * g.ntz;
    const int t = probe_tile(blockIdx.x, nb, g.xcd);
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
    const int64_t x = (int64_t)tx * 62 + threadIdx.x - 1, y = (int64_t)ty * (BY - 1) + threadIdx.y - 1;
    const bool own = threadIdx.x >= 1 && threadIdx.x <= 62 && threadIdx.y >= 1 && x < g.m && y < g.m;
    const int64_t xc = min(max(x, (int64_t)-1), g.m), yc = min(max(y, (int64_t)-1), g.m);
    const int64_t z0 = (int64_t)tz * g.kchunk, ze = min(z0 + g.kchunk, g.m);
    for (int64_t z = z0 - 1; z < ze; ++z) {
        const int64_t o = probe_off(g, xc, yc, z);
        const double s = probe_point<NR>(a, o);
        if (own && z >= z0) probe_store<NW>(a, o, s);
    }
}

// The same tiles, one plane per block (no z-march).
template <int BY, int NR, int NW>
__global__ void __launch_bounds__(64 * BY) probe_plane(ProbeArrays a, ProbeGeo g) {
    const int nb = g.ntx * g.nty * (int)g.m;
    const int t = probe_tile(blockIdx.x, nb, g.xcd);
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty;
    const int64_t z = t / (g.ntx * g.nty);
    const int64_t x = (int64_t)tx * 62 + threadIdx.x - 1, y = (int64_t)ty * (BY - 1) + threadIdx.y - 1;
    const bool own = threadIdx.x >= 1 && threadIdx.x <= 62 && threadIdx.y >= 1 && x < g.m && y < g.m;
    const int64_t xc = min(max(x, (int64_t)-1), g.m), yc = min(max(y, (int64_t)-1), g.m);
    const int64_t o = probe_off(g, xc, yc, z);
    const double s = probe_point<NR>(a, o);
    if (own) probe_store<NW>(a, o, s);
}

// One plane per block as pattern 2, but each point also reads every input at
// plane k-1 and the first three inputs at plane k+1: the z neighbours a fused
// kernel without a z-march would re-read (2 NR + 3 reads per point).
template <int BY, int NR, int NW>
__global__ void __launch_bounds__(64 * BY) probe_plane_halo(ProbeArrays a, ProbeGeo g) {
    const int nb = g.ntx * g.nty * (int)g.m;
    const int t = probe_tile(blockIdx.x, nb, g.xcd);
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty;
    const int64_t z = t / (g.ntx * g.nty);
    const int64_t x = (int64_t)tx * 62 + threadIdx.x - 1, y = (int64_t)ty * (BY - 1) + threadIdx.y - 1;
    const bool own = threadIdx.x >= 1 && threadIdx.x <= 62 && threadIdx.y >= 1 && x < g.m && y < g.m;
    const int64_t xc = min(max(x, (int64_t)-1), g.m), yc = min(max(y, (int64_t)-1), g.m);
    const int64_t o = probe_off(g, xc, yc, z);
    double v[2 * NR + 3];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        v[q] = a.r[q][o];
        v[NR + q] = a.r[q][o - g.pxy];   // plane z-1 >= -1
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) v[2 * NR + q] = a.r[q % NR][o + g.pxy];   // plane z+1 <= n+2
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 2 * NR + 3; ++q) s += v[q];
    if (own) probe_store<NW>(a, o, s);
}

}  // namespace chmy

using namespace chmy;

extern "C" int chmy_probe_stream(int pattern, const double* const* in, int nr, double* const* out, int nw, int64_t n,
                                 int by, int kchunk, int xcd, void* stream) {
    if (pattern < 0 || pattern > 3) return set_error(CHMY_E_ARG, "probe pattern %d not in 0..3", pattern);
    const bool fused = nr == 18 && nw == 10, copy = nr == 1 && nw == 1;
    if (!in || !out || !(fused || copy))
        return set_error(CHMY_E_ARG, "probe: 18 inputs and 10 outputs (the fused iteration) or 1 and 1 (a copy)");
    if (n < 1 || n > 1000) return set_error(CHMY_E_SIZE, "probe: n = %lld not in 1..1000", (long long)n);
    if (by != 4 && by != 8 && by != 12 && by != 16) return set_error(CHMY_E_ARG, "probe: rows %d not 4/8/12/16", by);
    if (kchunk < 1) return set_error(CHMY_E_ARG, "probe: kchunk < 1");
    ProbeArrays a;
    std::memset(&a, 0, sizeof a);
    for (int q = 0; q < nr; ++q) {
        if (!in[q]) return set_error(CHMY_E_ARG, "probe: input %d is NULL", q);
        a.r[q] = in[q];
    }
    for (int q = 0; q < nw; ++q) {
        if (!out[q]) return set_error(CHMY_E_ARG, "probe: output %d is NULL", q);
        a.w[q] = out[q];
    }
    a.nr = nr;
    a.nw = nw;
    ProbeGeo g;
    g.m = n + 2;
    g.px = n + 4;
    g.pxy = (n + 4) * (n + 4);
    g.ntx = (int)((g.m + 61) / 62);
    g.nty = (int)((g.m + by - 2) / (by - 1));
    g.kchunk = kchunk;
    g.ntz = (int)((g.m + kchunk - 1) / kchunk);
    g.xcd = xcd != 0;
    hipStream_t s = (hipStream_t)stream;
    auto run = [&](auto nrc, auto nwc) {
        constexpr int NR = decltype(nrc)::value, NW = decltype(nwc)::value;
        if (pattern == 0) {
            const int64_t pts = g.m * g.m * g.m;
            hipLaunchKernelGGL((probe_linear<NR, NW>), dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, s, a, g);
            return;
        }
        const unsigned nb = (unsigned)g.ntx * g.nty * (pattern == 1 ? g.ntz : (int)g.m);
        const dim3 blk(64, by);
#define CHMY_PROBE(BY)                                                                                  \
    if (by == BY) {                                                                                     \
        if (pattern == 1) hipLaunchKernelGGL((probe_tiled<BY, NR, NW>), dim3(nb), blk, 0, s, a, g);     \
        else if (pattern == 2) hipLaunchKernelGGL((probe_plane<BY, NR, NW>), dim3(nb), blk, 0, s, a, g); \
        else hipLaunchKernelGGL((probe_plane_halo<BY, NR, NW>), dim3(nb), blk, 0, s, a, g);            \
    }
        CHMY_PROBE(4) CHMY_PROBE(8) CHMY_PROBE(12) CHMY_PROBE(16)
#undef CHMY_PROBE
    };
    if (fused) run(std::integral_constant<int, 18>{}, std::integral_constant<int, 10>{});
    else run(std::integral_constant<int, 1>{}, std::integral_constant<int, 1>{});
    CHMY_HIP(hipGetLastError());
    return 0;
}
