// bc! batches on the GPU (src/BoundaryConditions/batch.jl:20-29,159-184;
// first_order_boundary_condition.jl:45-84).  One launch per dim covers both
// sides: side-1 and side-2 writes and reads never alias for fields with at
// least one cell, and each thread only touches its own normal line, so the
// result equals the reference's side-1-then-side-2 launches.  Dims run N..1.
#include "common.hpp"

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
    double h;         // spacing along D (Neumann)
    BCItem it[2][kMaxBatch];
};

template <class T>
__device__ __forceinline__ void apply_one(const BCItem& b, int side, int D, const int64_t* It, double h) {
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
        p[base + ig * sD] = fma(T(2), val - bn, bn);
    } else {
        T qs = side == 1 ? -val : val;
        p[base + ig * sD] = fma(T(h), qs, bn);
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
        for (int k = 0; k < L.K[s]; ++k) apply_one<T>(L.it[s][k], s + 1, L.D, It, L.h);
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

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_bc_side(const chmy_grid_t* g, int dim, int side, int K, const chmy_field_t* fields,
                 const chmy_bc_t* bcs, void* stream) {
    CHMY_TRY(check_grid(g));
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
    CHMY_TRY(check_grid(g));
    if (!counts) return set_error(CHMY_E_ARG, "counts is NULL");
    int offs[7] = {0};
    for (int q = 0; q < 2 * g->nd; ++q) {
        if (counts[q] < 0) return set_error(CHMY_E_ARG, "negative batch count");
        offs[q + 1] = offs[q] + counts[q];
    }
    if (offs[2 * g->nd] > 0 && (!fields || !bcs)) return set_error(CHMY_E_ARG, "fields/bcs NULL");
    for (int D = g->nd - 1; D >= 0; --D) {
        int Ks[2] = {counts[2 * D], counts[2 * D + 1]};
        const chmy_field_t* f[2] = {fields + offs[2 * D], fields + offs[2 * D + 1]};
        const chmy_bc_t* b[2] = {bcs + offs[2 * D], bcs + offs[2 * D + 1]};
        CHMY_TRY(launch_dim(g, D, Ks, f, b, (hipStream_t)stream));
    }
    return 0;
}

}  // extern "C"
