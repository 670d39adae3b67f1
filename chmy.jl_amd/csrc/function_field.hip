// Device evaluation of set!(f, grid, fun) and of FunctionField values
// (src/Fields/field.jl:119-142, src/Fields/function_field.jl:51-61).
//
// The reference compiles the user's function into its own GPU kernel.  Here
// the host turns the function into a point program: a straight-line list of
// operations in SSA form (operation q writes register q and reads only
// registers < q).  One kernel runs the program at every point of a box of the
// destination's logical indices and stores the last register, converted to
// the field's dtype.  Program, constants and source descriptors travel in the
// kernel argument, so every wave runs the same operation at the same time
// (the switch is uniform) and the only per-lane state is the register file.
//
// Continuous functions read the coordinates at the destination's location,
// coord(grid, loc, I...) with the fma forms of uniform_axis.jl:18-19;
// discrete ones read logical indices and other fields at constant index
// offsets (the reference's discrete FunctionFields index their parameter
// fields at ix, iy, iz, e.g. examples/sia.jl:38-53).  Arithmetic is IEEE f64
// operation by operation with no contraction (FMA is an operation of its own),
// so programs made of + - * / sqrt, comparisons and ifelse give the bits of
// the same expressions evaluated on the host; exp, log, sin, cos, tanh and
// pow are the device math library's and may differ from the host's by an ulp.
#include <cmath>

#include "common.hpp"

namespace chmy {

constexpr int kPMaxOps = 128;
constexpr int kPMaxConst = 64;
constexpr int kPMaxSrc = 8;

struct PSrc {
    const char* p;       // logical index (0,0,0)
    int64_t st[3];       // element strides
    int32_t dtype;
    int32_t esize;
};

struct PProg {
    int32_t n, nd;
    double o[3], h[3], ih[3];
    Axes3 ax;            // FunctionAxis tables (p == nullptr: uniform dim)
    chmy_pop_t op[kPMaxOps];
    double k[kPMaxConst];
    PSrc src[kPMaxSrc];
};

__device__ __forceinline__ double load_src(const PSrc& s, int64_t i, int64_t j, int64_t k) {
    const int64_t o = i * s.st[0] + j * s.st[1] + k * s.st[2];
    if (s.dtype == CHMY_F64) return reinterpret_cast<const double*>(s.p)[o];
    if (s.dtype == CHMY_F32) return (double)reinterpret_cast<const float*>(s.p)[o];
    return (double)bf16_to_f32(reinterpret_cast<const uint16_t*>(s.p)[o]);
}

__host__ __device__ __forceinline__ int off8(int32_t packed, int d) { return (int)(int8_t)((packed >> (8 * d)) & 0xff); }

// Julia's min/max on IEEE floats: NaN wins, -0.0 < 0.0
__device__ __forceinline__ double jl_min(double x, double y) {
    if (x != x || y != y) return x + y;
    return (y < x || (signbit(y) && !signbit(x))) ? y : x;
}
__device__ __forceinline__ double jl_max(double x, double y) {
    if (x != x || y != y) return x + y;
    return (y > x || (!signbit(y) && signbit(x))) ? y : x;
}

template <class T>
__global__ void __launch_bounds__(256) program_kernel(const PProg P, DF<T> dst, Box3 b) {
    const int64_t i = b.lo[0] + (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int64_t j = b.lo[1] + (int64_t)blockIdx.y * 4 + threadIdx.y;
    const int64_t k = b.lo[2] + blockIdx.z;
    if (i >= b.lo[0] + b.len[0] || j >= b.lo[1] + b.len[1]) return;
    const int64_t I[3] = {i, j, k};
    double r[kPMaxOps];
    for (int q = 0; q < P.n; ++q) {
        const chmy_pop_t o = P.op[q];
        double v = 0.0;
        switch (o.op) {
            case CHMY_P_CONST: v = P.k[o.a]; break;
            case CHMY_P_INDEX: v = (double)(I[o.a] + o.b); break;
            case CHMY_P_COORD: {
                if (P.ax.a[o.a].p) {
                    v = P.ax.a[o.a].at(o.b == CHMY_VERTEX ? CHMY_AX_VERTEX : CHMY_AX_CENTER, I[o.a] + o.c);
                    break;
                }
                const double i1 = (double)(I[o.a] + o.c - 1);
                const double h = P.h[o.a], og = P.o[o.a];
                v = o.b == CHMY_VERTEX ? fma(i1, h, og) : fma(i1, h, fma(0.5, h, og));
                break;
            }
            case CHMY_P_SPACING:
                v = P.ax.a[o.a].p ? P.ax.a[o.a].at(o.b == CHMY_VERTEX ? CHMY_AX_DV : CHMY_AX_DC, I[o.a] + o.c)
                                  : P.h[o.a];
                break;
            case CHMY_P_INV_SPACING:
                v = P.ax.a[o.a].p ? P.ax.a[o.a].at(o.b == CHMY_VERTEX ? CHMY_AX_IDV : CHMY_AX_IDC, I[o.a] + o.c)
                                  : P.ih[o.a];
                break;
            case CHMY_P_LOAD:
                v = load_src(P.src[o.a], i + off8(o.b, 0), j + off8(o.b, 1), k + off8(o.b, 2));
                break;
            case CHMY_P_NEG: v = -r[o.a]; break;
            case CHMY_P_ABS: v = fabs(r[o.a]); break;
            case CHMY_P_SQRT: v = sqrt(r[o.a]); break;
            case CHMY_P_EXP: v = exp(r[o.a]); break;
            case CHMY_P_LOG: v = log(r[o.a]); break;
            case CHMY_P_SIN: v = sin(r[o.a]); break;
            case CHMY_P_COS: v = cos(r[o.a]); break;
            case CHMY_P_TANH: v = tanh(r[o.a]); break;
            case CHMY_P_FLOOR: v = floor(r[o.a]); break;
            case CHMY_P_NOT: v = r[o.a] == 0.0 ? 1.0 : 0.0; break;
            case CHMY_P_ADD: v = r[o.a] + r[o.b]; break;
            case CHMY_P_SUB: v = r[o.a] - r[o.b]; break;
            case CHMY_P_MUL: v = r[o.a] * r[o.b]; break;
            case CHMY_P_DIV: v = r[o.a] / r[o.b]; break;
            case CHMY_P_POW: v = pow(r[o.a], r[o.b]); break;
            case CHMY_P_MIN: v = jl_min(r[o.a], r[o.b]); break;
            case CHMY_P_MAX: v = jl_max(r[o.a], r[o.b]); break;
            case CHMY_P_LT: v = r[o.a] < r[o.b] ? 1.0 : 0.0; break;
            case CHMY_P_LE: v = r[o.a] <= r[o.b] ? 1.0 : 0.0; break;
            case CHMY_P_GT: v = r[o.a] > r[o.b] ? 1.0 : 0.0; break;
            case CHMY_P_GE: v = r[o.a] >= r[o.b] ? 1.0 : 0.0; break;
            case CHMY_P_EQ: v = r[o.a] == r[o.b] ? 1.0 : 0.0; break;
            case CHMY_P_NE: v = r[o.a] != r[o.b] ? 1.0 : 0.0; break;
            case CHMY_P_AND: v = (r[o.a] != 0.0 && r[o.b] != 0.0) ? 1.0 : 0.0; break;
            case CHMY_P_OR: v = (r[o.a] != 0.0 || r[o.b] != 0.0) ? 1.0 : 0.0; break;
            case CHMY_P_POWI: {
                // x^n as repeated products, left to right (Julia's literal
                // powers x^2 = x*x, x^3 = x*x*x; x^-1 = inv(x), x^-2 = inv(x)^2)
                const double x = o.b < 0 ? 1.0 / r[o.a] : r[o.a];
                const int n = o.b < 0 ? -o.b : o.b;
                v = n == 0 ? 1.0 : x;
                for (int e = 1; e < n; ++e) v = v * x;
                break;
            }
            case CHMY_P_SELECT: v = r[o.a] != 0.0 ? r[o.b] : r[o.c]; break;
            case CHMY_P_FMA: v = fma(r[o.a], r[o.b], r[o.c]); break;
            default: break;
        }
        r[q] = v;
    }
    put(dst(i, j, k), r[P.n - 1]);
}

static bool unary(int op) { return op >= CHMY_P_NEG && op <= CHMY_P_NOT; }
static bool binary(int op) { return op >= CHMY_P_ADD && op <= CHMY_P_OR; }

static int check_program(const chmy_grid_t* g, const chmy_pop_t* prog, int nprog, int nconst, const chmy_field_t* srcs,
                         int nsrc, const Box3& b) {
    if (!prog || nprog < 1 || nprog > kPMaxOps)
        return set_error(CHMY_E_SIZE, "point program: %d operations, need 1..%d", nprog, kPMaxOps);
    for (int q = 0; q < nprog; ++q) {
        const chmy_pop_t& o = prog[q];
        auto reg = [&](int r) { return r >= 0 && r < q; };
        bool ok;
        switch (o.op) {
            case CHMY_P_CONST: ok = o.a >= 0 && o.a < nconst; break;
            case CHMY_P_INDEX: ok = o.a >= 0 && o.a < g->nd && o.b >= -64 && o.b <= 64; break;
            case CHMY_P_COORD:
            case CHMY_P_SPACING:
            case CHMY_P_INV_SPACING:
                ok = o.a >= 0 && o.a < g->nd && (o.b == CHMY_CENTER || o.b == CHMY_VERTEX) && o.c >= -64 && o.c <= 64;
                // a FunctionAxis table covers logical -PAD .. n+1+PAD
                if (ok && g->axis[o.a])
                    ok = b.lo[o.a] + o.c >= -CHMY_AXIS_PAD &&
                         b.lo[o.a] + b.len[o.a] - 1 + o.c <= g->n[o.a] + 1 + CHMY_AXIS_PAD;
                break;
            case CHMY_P_LOAD: {
                ok = o.a >= 0 && o.a < nsrc && o.b >= 0 && o.b < (1 << 24);
                if (!ok) break;
                const chmy_field_t& s = srcs[o.a];
                // every point of the box plus the offset stays in the source's storage
                for (int d = 0; d < 3 && ok; ++d) {
                    const int off = off8(o.b, d);
                    if (d >= g->nd) {
                        ok = off == 0;
                        break;
                    }
                    const int64_t H = 2 * (int64_t)s.halo[d];
                    ok = b.lo[d] + off >= 1 - H && b.lo[d] + b.len[d] - 1 + off <= s.n[d] + H;
                }
                if (!ok)
                    return set_error(CHMY_E_BOX, "point program op %d: load of source %d at offset (%d, %d, %d) leaves "
                                     "its storage over the box", q, o.a, off8(o.b, 0), off8(o.b, 1), off8(o.b, 2));
                break;
            }
            case CHMY_P_POWI: ok = reg(o.a) && o.b >= -64 && o.b <= 64; break;
            case CHMY_P_SELECT:
            case CHMY_P_FMA: ok = reg(o.a) && reg(o.b) && reg(o.c); break;
            default:
                if (unary(o.op)) ok = reg(o.a);
                else if (binary(o.op)) ok = reg(o.a) && reg(o.b);
                else ok = false;
        }
        if (!ok) return set_error(CHMY_E_ARG, "point program op %d (code %d, %d, %d, %d) is malformed", q, o.op, o.a, o.b, o.c);
    }
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_set_program(const chmy_grid_t* g, const chmy_field_t* dst, const chmy_box_t* box, const chmy_pop_t* prog,
                     int nprog, const double* consts, int nconst, const chmy_field_t* srcs, int nsrc, void* stream) {
    CHMY_TRY(check_grid_any(g));
    CHMY_TRY(check_field(g, dst, "dst", nullptr, -1, true));
    if (nconst < 0 || nconst > kPMaxConst || (nconst > 0 && !consts))
        return set_error(CHMY_E_SIZE, "point program: %d constants, at most %d", nconst, kPMaxConst);
    if (nsrc < 0 || nsrc > kPMaxSrc || (nsrc > 0 && !srcs))
        return set_error(CHMY_E_SIZE, "point program: %d source fields, at most %d", nsrc, kPMaxSrc);
    for (int s = 0; s < nsrc; ++s) CHMY_TRY(check_field(g, &srcs[s], "source", nullptr, -1, true));
    // the box: logical indices of dst, default its interior; within its storage
    chmy_box_t bb;
    for (int d = 0; d < 3; ++d) { bb.lo[d] = 1; bb.len[d] = d < g->nd ? dst->n[d] : 1; }
    if (box) bb = *box;
    for (int d = 0; d < g->nd; ++d) {
        const int64_t H = 2 * (int64_t)dst->halo[d];
        if (bb.len[d] < 0 || bb.lo[d] < 1 - H || bb.lo[d] + bb.len[d] - 1 > dst->n[d] + H)
            return set_error(CHMY_E_BOX, "point program: box outside the storage of dst along dim %d", d);
    }
    const Box3 b = box3(bb, g->nd);
    CHMY_TRY(check_program(g, prog, nprog, nconst, srcs, nsrc, b));
    if (b.len[0] == 0 || b.len[1] == 0 || b.len[2] == 0) return 0;
    PProg P;
    std::memset(&P, 0, sizeof P);
    P.n = nprog;
    P.nd = g->nd;
    for (int d = 0; d < 3; ++d) {
        P.o[d] = d < g->nd ? g->origin[d] : 0.0;
        P.h[d] = d < g->nd ? g->spacing[d] : 0.0;
        P.ih[d] = d < g->nd ? g->inv_spacing[d] : 0.0;
    }
    P.ax = axes_of(g);
    std::memcpy(P.op, prog, sizeof(chmy_pop_t) * nprog);
    if (nconst) std::memcpy(P.k, consts, sizeof(double) * nconst);
    for (int s = 0; s < nsrc; ++s) {
        const chmy_field_t& f = srcs[s];
        const int es = f.dtype == CHMY_F64 ? 8 : f.dtype == CHMY_F32 ? 4 : 2;
        // element strides and the element offset of logical (0,0,0): I + 2H - 1 per dim
        int64_t st[3] = {1, 0, 0}, off = 0;
        for (int d = 0; d < f.nd; ++d) {
            if (d > 0) st[d] = st[d - 1] * pad_of(f, d - 1);
            off += (2 * (int64_t)f.halo[d] - 1) * st[d];
        }
        P.src[s].p = static_cast<const char*>(f.data) + off * es;
        for (int d = 0; d < 3; ++d) P.src[s].st[d] = st[d];
        P.src[s].dtype = f.dtype;
        P.src[s].esize = es;
    }
    hipStream_t st = (hipStream_t)stream;
    const dim3 blk(64, 4, 1);
    const dim3 grd((unsigned)((b.len[0] + 63) / 64), (unsigned)((b.len[1] + 3) / 4), (unsigned)b.len[2]);
    if (dst->dtype == CHMY_F64)
        hipLaunchKernelGGL(program_kernel<double>, grd, blk, 0, st, P, view<double>(*dst), b);
    else if (dst->dtype == CHMY_F32)
        hipLaunchKernelGGL(program_kernel<float>, grd, blk, 0, st, P, view<float>(*dst), b);
    else
        hipLaunchKernelGGL(program_kernel<bf16s>, grd, blk, 0, st, P, view<bf16s>(*dst), b);
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
