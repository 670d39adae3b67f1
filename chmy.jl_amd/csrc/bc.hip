// bc! batches on the GPU (src/BoundaryConditions/batch.jl:20-29,159-184;
// first_order_boundary_condition.jl:45-84).  One launch per dim covers both
// sides: side-1 and side-2 writes and reads never alias for fields with at
// least one cell, and each thread only touches its own normal line, so the
// result equals the reference's side-1-then-side-2 launches.  Dims run N..1.
//
// A whole batch set also runs as ONE launch (bc_set_kernel) that writes each
// point's final value directly.  Dims run N..1, so the last write to a point
// p is made by the lowest dim D that writes it (ghost node 0 / m+1, or node
// 1 / m for Dirichlet on a vertex location).  That write read the inner node
// p' (p moved to 1 / m along D) as it stood after dims > D, which is again
// the write of the lowest dim > D writing p', or the untouched input value.
// Evaluating that chain (at most N steps, each the reference's own fma) per
// point gives the same bits as the N launches.  Each point is written by the
// thread of its lowest writing dim only, and the chain's final read is of a
// point that no dim writes, so no thread reads what another writes.
#include <algorithm>

#include "common.hpp"

#ifndef CHMY_BC_SET_LAUNCH
#define CHMY_BC_SET_LAUNCH 1   // 0: one launch per dim
#endif

namespace chmy {

constexpr int kMaxBatch = 12;

struct BCItem {
    void* p;          // logical (0,0,0) of the field
    int64_t st[3];    // strides per dim
    int64_t m;        // size of the field along D
    int32_t vertex;   // location along D
    int32_t kind;     // CHMY_DIRICHLET / CHMY_NEUMANN
    int32_t vkind;
    int32_t pad_;
    double value;
    void* vp;         // value field logical origin
    int64_t vst[2];   // value field strides over the tangential dims
};

struct BCLaunch {
    int32_t D;
    int32_t nt;       // number of tangential dims (0..2)
    int32_t td[2];    // tangential dims
    int64_t tn[2];    // tangential extents (n_t + 3)
    int32_t K[2];     // fields per side
    double h;         // spacing along D (Neumann, uniform axis)
    AxisT ax;         // FunctionAxis table along D, or p == nullptr
    BCItem it[2][kMaxBatch];
};

// Center Dirichlet coefficient t and Neumann spacing h of the ghost ig
// (inner neighbour in) along an axis; uniform: 2 and h_uniform.
__device__ __forceinline__ double dirichlet_t(const AxisT& ax, int64_t ig, int64_t in) {
    return ax.p ? (2.0 * ax.at(CHMY_AX_DC, in)) * ax.at(CHMY_AX_IDV, ig) : 2.0;
}
__device__ __forceinline__ double neumann_h(const AxisT& ax, int vertex, int64_t ig, double h) {
    return ax.p ? ax.at(vertex ? CHMY_AX_DC : CHMY_AX_DV, ig) : h;
}

template <class T>
__device__ __forceinline__ void apply_one(const BCItem& b, int side, int D, const int64_t* It, double h,
                                          const AxisT& ax) {
    T* p = static_cast<T*>(b.p);
    int64_t base = 0;
    for (int d = 0; d < 3; ++d)
        if (d != D) base += It[d] * b.st[d];
    T val = T(0);
    if (b.vkind == CHMY_VALUE_SCALAR) val = T(b.value);
    else if (b.vkind == CHMY_VALUE_FIELD) {
        int q = 0;
        int64_t vo = 0;
        for (int d = 0; d < 3; ++d)
            if (d != D && q < 2) { vo += It[d] * b.vst[q]; ++q; }
        val = static_cast<const T*>(b.vp)[vo];
    }
    const int64_t sD = b.st[D];
    if (b.kind == CHMY_DIRICHLET && b.vertex) {
        int64_t i = side == 1 ? 1 : b.m;
        p[base + i * sD] = val;
        return;
    }
    int64_t ig = side == 1 ? 0 : b.m + 1;
    int64_t in = side == 1 ? 1 : b.m;
    T bn = p[base + in * sD];
    if (b.kind == CHMY_DIRICHLET) {
        p[base + ig * sD] = fma(T(dirichlet_t(ax, ig, in)), val - bn, bn);
    } else {
        T qs = side == 1 ? -val : val;
        p[base + ig * sD] = fma(T(neumann_h(ax, b.vertex, ig, h)), qs, bn);
    }
}

template <class T>
__global__ void __launch_bounds__(256) bc_kernel(BCLaunch L) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = (L.nt > 0 ? L.tn[0] : 1) * (L.nt > 1 ? L.tn[1] : 1);
    if (t >= tot) return;
    int64_t It[3] = {0, 0, 0};
    if (L.nt > 0) It[L.td[0]] = t % L.tn[0];
    if (L.nt > 1) It[L.td[1]] = t / L.tn[0];
    for (int s = 0; s < 2; ++s)
        for (int k = 0; k < L.K[s]; ++k) apply_one<T>(L.it[s][k], s + 1, L.D, It, L.h, L.ax);
}

static int fill_item(const chmy_grid_t* g, int D, const chmy_field_t* f, const chmy_bc_t* bc, int dtype,
                     BCItem* it) {
    CHMY_TRY(check_field(g, f, "bc field", nullptr, dtype));
    if (bc->kind != CHMY_DIRICHLET && bc->kind != CHMY_NEUMANN) return set_error(CHMY_E_ARG, "bc kind %d", bc->kind);
    const int64_t es = dtype == CHMY_F64 ? 8 : 4;
    const int64_t s1 = pad_of(*f, 0);
    const int64_t sy = f->nd >= 2 ? s1 : 0;
    const int64_t sz = f->nd >= 3 ? s1 * pad_of(*f, 1) : 0;
    int64_t off = 2 * (int64_t)f->halo[0] - 1;
    if (f->nd >= 2) off += (2 * (int64_t)f->halo[1] - 1) * sy;
    if (f->nd >= 3) off += (2 * (int64_t)f->halo[2] - 1) * sz;
    it->p = static_cast<char*>(f->data) + off * es;
    it->st[0] = 1;
    it->st[1] = sy;
    it->st[2] = sz;
    it->m = f->n[D];
    it->vertex = f->loc[D];
    it->kind = bc->kind;
    it->vkind = bc->vkind;
    it->value = bc->value;
    it->vp = nullptr;
    it->vst[0] = it->vst[1] = 0;
    if (bc->vkind == CHMY_VALUE_FIELD) {
        const chmy_field_t* vf = bc->vfield;
        if (!vf || !vf->data) return set_error(CHMY_E_ARG, "bc value field is NULL");
        if (g->nd < 2) return set_error(CHMY_E_ARG, "field-valued bc needs nd >= 2");
        if (vf->nd != g->nd - 1 || vf->dtype != dtype) return set_error(CHMY_E_LOC, "bc value field: nd/dtype mismatch");
        int q = 0;
        int64_t st = 1, off = 0;
        for (int d = 0; d < g->nd; ++d) {
            if (d == D) continue;
            // tangential logical range 0..n_t+2 must be inside [1-2H, n_v+2H]
            if (vf->halo[q] < 1 || vf->n[q] + 2 * (int64_t)vf->halo[q] < g->n[d] + 2)
                return set_error(CHMY_E_HALO, "bc value field too small along tangential dim %d", d);
            it->vst[q] = st;
            off += (2 * (int64_t)vf->halo[q] - 1) * st;
            st *= vf->n[q] + 4 * (int64_t)vf->halo[q];
            ++q;
        }
        it->vp = static_cast<char*>(vf->data) + off * es;
    } else if (bc->vkind != CHMY_VALUE_ZERO && bc->vkind != CHMY_VALUE_SCALAR) {
        return set_error(CHMY_E_ARG, "bc value kind %d", bc->vkind);
    }
    return 0;
}

static int launch_dim(const chmy_grid_t* g, int D, const int* K, const chmy_field_t* const* f,
                      const chmy_bc_t* const* bc, hipStream_t s) {
    if (K[0] + K[1] == 0) return 0;
    if (K[0] > kMaxBatch || K[1] > kMaxBatch) return set_error(CHMY_E_SIZE, "more than %d fields in a bc batch", kMaxBatch);
    int dtype = K[0] ? f[0][0].dtype : f[1][0].dtype;
    BCLaunch L;
    std::memset(&L, 0, sizeof L);
    L.D = D;
    L.nt = 0;
    for (int d = 0; d < g->nd; ++d)
        if (d != D) { L.td[L.nt] = d; L.tn[L.nt] = g->n[d] + 3; ++L.nt; }
    L.h = g->spacing[D];
    L.ax = axes_of(g).a[D];
    for (int s = 0; s < 2; ++s) {
        L.K[s] = K[s];
        for (int k = 0; k < K[s]; ++k) CHMY_TRY(fill_item(g, D, &f[s][k], &bc[s][k], dtype, &L.it[s][k]));
    }
    int64_t tot = (L.nt > 0 ? L.tn[0] : 1) * (L.nt > 1 ? L.tn[1] : 1);
    dim3 blk(256), grd((unsigned)((tot + 255) / 256));
    if (dtype == CHMY_F64) hipLaunchKernelGGL(bc_kernel<double>, grd, blk, 0, s, L);
    else hipLaunchKernelGGL(bc_kernel<float>, grd, blk, 0, s, L);
    CHMY_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ single launch
constexpr int kSetFields = 8;   // distinct fields in a one-launch batch set

struct BCRule {       // the condition of one (dim, side) of a field
    int32_t kind;     // -1: none; else CHMY_DIRICHLET / CHMY_NEUMANN
    int32_t vkind;
    double value;
    const void* vp;   // value field logical origin
    int64_t vst[2];
};

struct BCSetField {
    void* p;          // logical (0,0,0)
    int64_t st[3];
    int64_t m[3];     // size per dim
    int32_t vertex[3];
    int32_t pad_;
    BCRule r[3][2];
};

struct BCSetLaunch {
    int32_t nd, nf;
    int64_t tn[3];    // n_d + 3
    double h[3];
    Axes3 ax;
    BCSetField f[kSetFields];
};

// node written by rule (j, side) of a field, or -1
__device__ __forceinline__ int64_t bc_node(const BCSetField& F, int j, int side) {
    const BCRule& r = F.r[j][side];
    if (r.kind < 0) return -1;
    const bool dv = r.kind == CHMY_DIRICHLET && F.vertex[j];
    return side == 0 ? (dv ? 1 : 0) : (dv ? F.m[j] : F.m[j] + 1);
}

// side (0/1) of the rule of dim j that writes node c, or -1
__device__ __forceinline__ int bc_writer(const BCSetField& F, int j, int64_t c) {
    if (c == bc_node(F, j, 0)) return 0;
    if (c == bc_node(F, j, 1)) return 1;
    return -1;
}

template <class T>
__device__ __forceinline__ T bc_value(const BCRule& r, int j, const int64_t* I) {
    if (r.vkind == CHMY_VALUE_SCALAR) return T(r.value);
    if (r.vkind == CHMY_VALUE_FIELD) {
        int q = 0;
        int64_t vo = 0;
        for (int d = 0; d < 3; ++d)
            if (d != j && q < 2) { vo += I[d] * r.vst[q]; ++q; }
        return static_cast<const T*>(r.vp)[vo];
    }
    return T(0);
}

// blockIdx.y = (field, dim, side); x threads over the dim's tangential plane
template <class T>
__global__ void __launch_bounds__(256) bc_set_kernel(BCSetLaunch L) {
    const int item = blockIdx.y;
    const int fi = item / (2 * L.nd), D = (item / 2) % L.nd, sd = item % 2;
    const BCSetField& F = L.f[fi];
    const int64_t pD = bc_node(F, D, sd);
    if (pD < 0) return;
    int64_t tot = 1;
    for (int d = 0; d < L.nd; ++d)
        if (d != D) tot *= L.tn[d];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tot) return;
    int64_t I[3] = {0, 0, 0};
    {
        int64_t u = t;
        for (int d = 0; d < L.nd; ++d)
            if (d != D) { I[d] = u % L.tn[d]; u /= L.tn[d]; }
    }
    I[D] = pD;
    for (int j = 0; j < D; ++j)
        if (bc_writer(F, j, I[j]) >= 0) return;      // a lower dim writes this point last
    const int64_t pofs = I[0] * F.st[0] + I[1] * F.st[1] + I[2] * F.st[2];
    // the chain: writes of dims D = j0 < j1 < ... applied innermost first
    int nop = 0;
    int oj[3], os[3];
    T ov[3];
    T v;
    int j = D, side = sd;
    int64_t c[3] = {I[0], I[1], I[2]};
    for (;;) {
        const BCRule& r = F.r[j][side];
        const T val = bc_value<T>(r, j, c);
        if (r.kind == CHMY_DIRICHLET && F.vertex[j]) {
            v = val;
            break;
        }
        oj[nop] = j;
        os[nop] = side;
        ov[nop] = val;
        ++nop;
        c[j] = side == 0 ? 1 : F.m[j];
        int nj = -1, ns = -1;
        for (int k = j + 1; k < L.nd; ++k) {
            const int w = bc_writer(F, k, c[k]);
            if (w >= 0) { nj = k; ns = w; break; }
        }
        if (nj < 0) {
            v = static_cast<const T*>(F.p)[c[0] * F.st[0] + c[1] * F.st[1] + c[2] * F.st[2]];
            break;
        }
        j = nj;
        side = ns;
    }
    for (int q = nop - 1; q >= 0; --q) {
        const int jq = oj[q];
        const BCRule& r = F.r[jq][os[q]];
        const int64_t ig = os[q] == 0 ? 0 : F.m[jq] + 1, in = os[q] == 0 ? 1 : F.m[jq];
        if (r.kind == CHMY_DIRICHLET) v = fma(T(dirichlet_t(L.ax.a[jq], ig, in)), ov[q] - v, v);
        else v = fma(T(neumann_h(L.ax.a[jq], F.vertex[jq], ig, L.h[jq])), os[q] == 0 ? -ov[q] : ov[q], v);
    }
    static_cast<T*>(F.p)[pofs] = v;
}

static bool same_field(const chmy_field_t& a, const chmy_field_t& b) {
    if (a.nd != b.nd || a.dtype != b.dtype) return false;
    for (int d = 0; d < 3; ++d)
        if (a.n[d] != b.n[d] || a.halo[d] != b.halo[d] || a.loc[d] != b.loc[d]) return false;
    return true;
}

// Packs the batch set for bc_set_kernel; returns 1 when it qualifies (at most
// kSetFields distinct fields, each field named once per (dim, side), at least
// two nodes along every dim it has a condition on, value fields distinct from
// the batch's fields), 0 when the per-dim launches must run, <0 on error.
static int pack_set(const chmy_grid_t* g, const int32_t* counts, const chmy_field_t* fields, const chmy_bc_t* bcs,
                    BCSetLaunch* L, int* dtype) {
    std::memset(L, 0, sizeof *L);
    L->nd = g->nd;
    L->ax = axes_of(g);
    for (int d = 0; d < 3; ++d) {
        L->tn[d] = d < g->nd ? g->n[d] + 3 : 1;
        L->h[d] = d < g->nd ? g->spacing[d] : 0.0;
    }
    const chmy_field_t* uniq[kSetFields];
    int q = 0;
    *dtype = -1;
    for (int D = 0; D < g->nd; ++D)
        for (int sd = 0; sd < 2; ++sd)
            for (int k = 0; k < counts[2 * D + sd]; ++k, ++q) {
                const chmy_field_t& f = fields[q];
                if (*dtype < 0) *dtype = f.dtype;
                if (f.dtype != CHMY_F64 && f.dtype != CHMY_F32) return 0;
                BCItem it;
                CHMY_TRY(fill_item(g, D, &f, &bcs[q], *dtype, &it));
                int fi = 0;
                while (fi < L->nf && uniq[fi]->data != f.data) ++fi;
                if (fi == L->nf) {
                    if (L->nf == kSetFields) return 0;
                    uniq[fi] = &f;
                    BCSetField& F = L->f[fi];
                    F.p = it.p;
                    for (int d = 0; d < 3; ++d) {
                        F.st[d] = it.st[d];
                        F.m[d] = d < f.nd ? f.n[d] : 1;
                        F.vertex[d] = d < f.nd ? f.loc[d] : 0;
                        F.r[d][0].kind = F.r[d][1].kind = -1;
                    }
                    ++L->nf;
                } else if (!same_field(*uniq[fi], f)) {
                    return 0;
                }
                BCSetField& F = L->f[fi];
                BCRule& r = F.r[D][sd];
                if (r.kind >= 0 || F.m[D] < 2) return 0;
                r.kind = it.kind;
                r.vkind = it.vkind;
                r.value = it.value;
                r.vp = it.vp;
                r.vst[0] = it.vst[0];
                r.vst[1] = it.vst[1];
            }
    for (int a = 0; a < q; ++a)
        if (bcs[a].vkind == CHMY_VALUE_FIELD)
            for (int fi = 0; fi < L->nf; ++fi)
                if (bcs[a].vfield->data == uniq[fi]->data) return 0;
    return L->nf > 0 ? 1 : 0;
}

static int launch_set(const BCSetLaunch& L, int dtype, hipStream_t s) {
    int64_t tot = 1;
    for (int d = 0; d < L.nd; ++d) {
        int64_t p = 1;
        for (int e = 0; e < L.nd; ++e)
            if (e != d) p *= L.tn[e];
        tot = std::max(tot, p);
    }
    dim3 blk(256), grd((unsigned)((tot + 255) / 256), (unsigned)(L.nf * 2 * L.nd));
    if (dtype == CHMY_F64) hipLaunchKernelGGL(bc_set_kernel<double>, grd, blk, 0, s, L);
    else hipLaunchKernelGGL(bc_set_kernel<float>, grd, blk, 0, s, L);
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_bc_side(const chmy_grid_t* g, int dim, int side, int K, const chmy_field_t* fields,
                 const chmy_bc_t* bcs, void* stream) {
    CHMY_TRY(check_grid_any(g));
    if (dim < 0 || dim >= g->nd || (side != 1 && side != 2) || K < 0 || (K > 0 && (!fields || !bcs)))
        return set_error(CHMY_E_ARG, "chmy_bc_side: bad arguments");
    int Ks[2] = {0, 0};
    const chmy_field_t* f[2] = {fields, fields};
    const chmy_bc_t* b[2] = {bcs, bcs};
    Ks[side - 1] = K;
    return launch_dim(g, dim, Ks, f, b, (hipStream_t)stream);
}

int chmy_bc_batchset(const chmy_grid_t* g, const int32_t* counts, const chmy_field_t* fields,
                     const chmy_bc_t* bcs, void* stream) {
    CHMY_TRY(check_grid_any(g));
    if (!counts) return set_error(CHMY_E_ARG, "counts is NULL");
    int offs[7] = {0};
    for (int q = 0; q < 2 * g->nd; ++q) {
        if (counts[q] < 0) return set_error(CHMY_E_ARG, "negative batch count");
        offs[q + 1] = offs[q] + counts[q];
    }
    if (offs[2 * g->nd] > 0 && (!fields || !bcs)) return set_error(CHMY_E_ARG, "fields/bcs NULL");
    if (offs[2 * g->nd] == 0) return 0;
    if (CHMY_BC_SET_LAUNCH) {
        BCSetLaunch L;
        int dtype;
        const int ok = pack_set(g, counts, fields, bcs, &L, &dtype);
        if (ok < 0) return ok;
        if (ok) return launch_set(L, dtype, (hipStream_t)stream);
    }
    for (int D = g->nd - 1; D >= 0; --D) {
        int Ks[2] = {counts[2 * D], counts[2 * D + 1]};
        const chmy_field_t* f[2] = {fields + offs[2 * D], fields + offs[2 * D + 1]};
        const chmy_bc_t* b[2] = {bcs + offs[2 * D], bcs + offs[2 * D + 1]};
        CHMY_TRY(launch_dim(g, D, Ks, f, b, (hipStream_t)stream));
    }
    return 0;
}

}  // extern "C"
