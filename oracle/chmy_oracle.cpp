// CPU oracle: a plain C++ restatement of Chmy.jl's CPU semantics for the
// staggered-grid stencil path.  TEST INFRASTRUCTURE ONLY: it is the checker
// that the HIP library is compared against; nothing in the product links,
// loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may use it.
//
// Rounding rules (so that results agree with Chmy's CPU path to the last
// ulp wherever Chmy's own arithmetic is deterministic):
//   * compiled with -ffp-contract=off: no implicit FMA contraction;
//   * std::fma exactly where the reference writes `muladd`
//     (src/GridOperators/interpolation.jl:14,
//      src/BoundaryConditions/first_order_boundary_condition.jl:69,82,
//      src/Grids/uniform_axis.jl:18-19);
//   * association order of every expression as written in the reference.
//
// Parity pins (tests/test_oracle_*.py): the reference's own unit-test
// assertions (test/test_grid_operators.jl, test_interpolations.jl,
// test_boundary_conditions.jl, test_fields.jl, test_grids.jl), closed forms
// (exact derivatives of affine/quadratic fields, Neumann cosine eigenmodes of
// the diffusion step, exact linear pure-shear Stokes state), brute-force
// dense-matrix operators on tiny grids and invariants (mass conservation).
// Variable-viscosity Stokes and the SIA power law are build-defined
// extensions; their pins are listed in the test files.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <algorithm>
#include <omp.h>

extern "C" {

// Layouts match include/chmy_mi355x.h so one Python descriptor serves both.
typedef struct {
    void*   data;
    int32_t nd;
    int32_t dtype;    // 0 = f64, 1 = f32
    int64_t n[3];     // logical size per dim (size(grid, loc))
    int32_t halo[3];
    int32_t loc[3];   // 0 = Center, 1 = Vertex
} orc_field;

typedef struct {
    int32_t nd;
    int32_t reserved;
    int64_t n[3];     // cell counts
    double  origin[3];
    double  extent[3];
    double  spacing[3];
    double  inv_spacing[3];
    // FunctionAxis (src/Grids/function_axis.jl:1-9): host array of vertex(i)
    // for i = -kPad-1 .. n+2+kPad at axis[d][i + kPad + 1], or NULL (uniform)
    const double* axis[3];
} orc_grid;

typedef struct { int64_t lo[3]; int64_t len[3]; } orc_box;

typedef struct {
    int32_t kind;          // 0 = Dirichlet, 1 = Neumann
    int32_t vkind;         // 0 = nothing (zero), 1 = scalar, 2 = field
    double  value;
    const orc_field* vfield;
} orc_bc;

}  // extern "C"

namespace {

int g_err = 0;
char g_msg[256];

// Planted mistakes (test infrastructure for the pins themselves): a test may
// switch on one plausible error at a time (orc_set_plant) to show that the
// check pinning that oracle function fails on it.  0, the default, is the
// restatement; with 0 every site computes exactly what it computed before
// the switch existed (one well-predicted branch on a global).
enum Plant {
    PL_NONE = 0,
    PL_VERTEX_DIFF = 1,      // ∂ of a Vertex field with knots (I-e, I)
    PL_LERP_KNOTS = 2,       // Center->Vertex lerp knots (I, I+e)
    PL_CENTER_DIRICHLET = 3, // Center Dirichlet ghost with t = 1
    PL_NEUMANN_SIGN = 4,     // Neumann flux not negated on side 1
    PL_BC_ORDER = 5,         // batch dims 1..N instead of N..1 (oracle.py)
    PL_DIFF_FLUX = 6,        // compute_q! without the minus sign
    PL_STRESS_DEV = 7,       // diagonal r_tau without the -divV/3 term
    PL_EDGE_SWAP = 8,        // eta of tau_xz and tau_yz lerped to each other's edge
    PL_RV_SWAP = 9,          // r_Vx with dz(tau_yz) for dz(tau_xz); 2D dx(tau_xy) for dy(tau_xy)
    PL_VEL_ML_SIDE = 10,     // velocity coefficient from max(eta_ml[I], eta_ml[I+e])
    PL_MAXLOC_2N = 11,       // maxloc over the 2^N cells I-e..I instead of 3^N
    PL_THERMAL_UPWIND = 12,  // upwind knots swapped (left for V<0, right for V>0)
    PL_SIA_EXP = 13,         // D with A H^(n+1) instead of A H^(n+2)
    PL_SIA_GRAD = 14,        // D without |∇S|^(n-1)
    PL_SIA_M = 15,           // M = max(mb_a (S - mb_b), mb_c)
    PL_RESID_RP = 16,        // R_P = +divg(V)
    PL_PR_UPDATE = 17,       // Pr += divV eta_ve dtau_Pr
    PL_GRAD_EDGE = 18,       // gradient edge terms weighted 1/2 instead of 1/4
    PL_PT_PARAMS = 19,       // theta with (r + 1/3) (oracle.py pt_params)
    PL_VMAG = 20,            // vmag without the square root
    PL_DIVG = 21,            // divg = dx Vx - dy Vy
    PL_THERMAL_SIGN = 22,    // update_thermal! with T_old + dt divg(qT)
    PL_MASK_AT_OUT = 23,     // masked left/right read the mask at the output index, not the knot
    PL_HLERP_ARITH = 24,     // hlerp with the arithmetic rule muladd(t, b - a, a)
    PL_DELTA_SIGN = 25,      // δ = left - right
    PL_NU_DIFF_LOC = 26,     // FunctionAxis ∂ scaled by iΔ at the output location instead of the source's
    PL_NU_WEIGHT = 27,       // FunctionAxis Center->Vertex lerp weight from Δc(i) instead of Δc(i-1)
    PL_NU_DIRICHLET = 28,    // FunctionAxis Center Dirichlet t = 2 Δc(I_f) iΔv(I_f2) (indices swapped)
    PL_NU_NEUMANN = 29,      // FunctionAxis Neumann h = Δ(loc) instead of Δ(flip(loc))
};
constexpr long kPad = 4;   // FunctionAxis arrays reach kPad indices past the halo
int g_plant = 0;

void fail(const char* m) {
    if (!g_err) { g_err = 1; std::snprintf(g_msg, sizeof g_msg, "%s", m); }
}

struct Ix {
    long v[3];
};
inline Ix shift(Ix I, int d, long s) { I.v[d] += s; return I; }

// Field{T,N,L,H}: logical index I -> data[I .+ 2H] (1-based, x fastest),
// data size dims .+ 4H (src/Fields/field.jl:6-18,56-62).
template <class T>
struct FV {
    T* p;
    int nd;
    long n[3];
    int H[3];
    int loc[3];
    long st[3];
    explicit FV(const orc_field* f) {
        p = static_cast<T*>(f->data);
        nd = f->nd;
        long s = 1;
        for (int d = 0; d < 3; ++d) {
            n[d] = d < nd ? f->n[d] : 1;
            H[d] = d < nd ? f->halo[d] : 0;
            loc[d] = d < nd ? f->loc[d] : 0;
            st[d] = d < nd ? s : 0;
            if (d < nd) s *= n[d] + 4L * H[d];
        }
    }
    long off(const Ix& I) const {
        long o = 0;
        for (int d = 0; d < nd; ++d) {
            long i = I.v[d];
            if (i < 1 - 2L * H[d] || i > n[d] + 2L * H[d]) {
                fail("oracle: logical index outside padded storage");
                return 0;
            }
            o += (i + 2L * H[d] - 1) * st[d];
        }
        return o;
    }
    T& operator()(const Ix& I) const { return p[off(I)]; }
};

template <class T>
struct G {
    int nd;
    long n[3];
    T sp[3], isp[3];
    const double* xv[3];
    explicit G(const orc_grid* g) {
        nd = g->nd;
        for (int d = 0; d < 3; ++d) {
            n[d] = d < nd ? g->n[d] : 1;
            sp[d] = T(g->spacing[d]);
            isp[d] = T(g->inv_spacing[d]);
            xv[d] = d < nd ? g->axis[d] : nullptr;
        }
    }
    // FunctionAxis metrics from the vertex function's values
    // (src/Grids/abstract_axis.jl:18,26-29): center(i) = 0.5 (vertex(i) + vertex(i+1)),
    // spacing(Center, i) = vertex(i+1) - vertex(i), spacing(Vertex, i) =
    // center(i) - center(i-1), inv_spacing = inv(spacing).
    double vert(int d, long i) const {
        if (i < -kPad - 1 || i > n[d] + 2 + kPad) fail("oracle: FunctionAxis index outside its array");
        return xv[d][i + kPad + 1];
    }
    double cent(int d, long i) const { return 0.5 * (vert(d, i) + vert(d, i + 1)); }
    double dC(int d, long i) const { return vert(d, i + 1) - vert(d, i); }
    double dV(int d, long i) const { return cent(d, i) - cent(d, i - 1); }
    // iΔ(grid, loc, D, I) along D at index i (uniform: the grid's inv_spacing)
    T ispat(int D, int loc, long i) const {
        if (!xv[D]) return isp[D];
        return T(1.0 / (loc == 0 ? dC(D, i) : dV(D, i)));
    }
    // itp_weight(ax, Center, Vertex, i) = 0.5 Δ(Center, i-1) iΔ(Vertex, i)
    // (interpolation.jl:29); 0.5 on a uniform axis (:32) and Vertex -> Center (:28)
    T wcv(int D, long i) const {
        if (!xv[D]) return T(0.5);
        return T((0.5 * dC(D, g_plant == PL_NU_WEIGHT ? i : i - 1)) * (1.0 / dV(D, i)));
    }
};

// il/ir for an operator applied to a field at its own location
// (src/GridOperators/GridOperators.jl:23-36): Center -> (I-e, I), Vertex -> (I, I+e).
template <class T> inline T leftv(const FV<T>& f, int D, const Ix& I) {
    return f.loc[D] == 0 ? f(shift(I, D, -1)) : f(I);
}
template <class T> inline T rightv(const FV<T>& f, int D, const Ix& I) {
    return f.loc[D] == 0 ? f(I) : f(shift(I, D, +1));
}
// ∂ = δ * iΔ (src/GridOperators/partial_derivatives.jl:2-5)
// iΔ(grid, loc, dim, I) with loc the source field's location, at the output
// index I (partial_derivatives.jl:5, field_operators.jl:20-24)
template <class T> inline T isp_of(const FV<T>& f, const G<T>& g, int D, const Ix& I) {
    const int loc = g.xv[D] && g_plant == PL_NU_DIFF_LOC ? 1 - f.loc[D] : f.loc[D];
    return g.ispat(D, loc, I.v[D]);
}
template <class T> inline T dd(const FV<T>& f, const G<T>& g, int D, const Ix& I) {
    if (g_plant == PL_VERTEX_DIFF && f.loc[D] == 1) return (f(I) - f(shift(I, D, -1))) * g.isp[D];
    return (rightv(f, D, I) - leftv(f, D, I)) * isp_of(f, g, D, I);
}

// lerp (src/GridOperators/interpolation.jl:14-94): for every dim where the
// location differs, knots (I-e, I) for Center->Vertex and (I, I+e) for
// Vertex->Center, rule fma(0.5, b-a, a), last differing dim outermost.
// With a grid g, the weight of a Center->Vertex dim is g's (G::wcv).
template <class T>
T itp_rec(const FV<T>& f, const int* dims, int level, const int* to, const Ix& I, const G<T>* g) {
    if (level == 0) return f(I);
    int D = dims[level - 1];
    Ix Il = I, Ir = I;
    if (f.loc[D] == 0 && to[D] == 1 && g_plant != PL_LERP_KNOTS) { Il = shift(I, D, -1); }
    else { Ir = shift(I, D, +1); }
    T a = itp_rec(f, dims, level - 1, to, Il, g);
    T b = itp_rec(f, dims, level - 1, to, Ir, g);
    const T t = g && f.loc[D] == 0 && to[D] == 1 ? g->wcv(D, I.v[D]) : T(0.5);
    return std::fma(t, b - a, a);
}
template <class T> T lerp(const FV<T>& f, const int* to, const Ix& I, const G<T>* g = nullptr) {
    int dims[3], m = 0;
    for (int d = 0; d < f.nd; ++d)
        if (f.loc[d] != to[d]) dims[m++] = d;
    if (m == 0) return f(I);
    return itp_rec(f, dims, m, to, I, g);
}
template <class T> T hlerp_rec(const FV<T>& f, const int* dims, int level, const int* to, const Ix& I,
                                const G<T>* g) {
    if (level == 0) return f(I);
    int D = dims[level - 1];
    Ix Il = I, Ir = I;
    if (f.loc[D] == 0 && to[D] == 1) { Il = shift(I, D, -1); }
    else { Ir = shift(I, D, +1); }
    T a = hlerp_rec(f, dims, level - 1, to, Il, g);
    T b = hlerp_rec(f, dims, level - 1, to, Ir, g);
    const T t = g && f.loc[D] == 0 && to[D] == 1 ? g->wcv(D, I.v[D]) : T(0.5);
    // itp_rule(::HarmonicLinear, t, a, b) = inv(muladd(t, inv(b) - inv(a), inv(a)))
    if (g_plant == PL_HLERP_ARITH) return std::fma(t, b - a, a);
    return T(1) / std::fma(t, T(1) / b - T(1) / a, T(1) / a);
}
template <class T> T hlerp(const FV<T>& f, const int* to, const Ix& I, const G<T>* g = nullptr) {
    int dims[3], m = 0;
    for (int d = 0; d < f.nd; ++d)
        if (f.loc[d] != to[d]) dims[m++] = d;
    if (m == 0) return f(I);
    return hlerp_rec(f, dims, m, to, I, g);
}

// divg: left-associated Σ ∂_D V_D (src/GridOperators/field_operators.jl:50-59)
template <class T> T divg(const FV<T>* V, const G<T>& g, const Ix& I) {
    T s = dd(V[0], g, 0, I);
    for (int d = 1; d < g.nd; ++d) s = g_plant == PL_DIVG && d == 1 ? s - dd(V[d], g, d, I) : s + dd(V[d], g, d, I);
    return s;
}
// vmag = sqrt(Σ lerp(V_D, Center)^2) (field_operators.jl:116-125)
template <class T> T vmag(const FV<T>* V, const G<T>& g, const Ix& I) {
    const int c[3] = {0, 0, 0};
    T a = lerp(V[0], c, I, &g);
    T s = a * a;
    for (int d = 1; d < g.nd; ++d) { T b = lerp(V[d], c, I, &g); s = s + b * b; }
    return g_plant == PL_VMAG ? s : std::sqrt(s);
}
// ∂² (partial_derivatives.jl:7-12): from = loc = location(f).
template <class T> T d2(const FV<T>& f, const G<T>& g, int D, const Ix& I) {
    // floc[D] = flip(loc[D]); Ir = ir(floc, loc), Il = il(floc, loc)
    Ix Ir = I, Il = I;
    if (f.loc[D] == 0) { Ir = shift(I, D, +1); }   // ir(Vertex, Center) = I+e, il = I
    else { Il = shift(I, D, -1); }                 // ir(Center, Vertex) = I,   il = I-e
    return (dd(f, g, D, Ir) - dd(f, g, D, Il)) * isp_of(f, g, D, I);
}
// ∂k∂ (partial_derivatives.jl:14-21)
template <class T> T dkd(const FV<T>& f, const FV<T>& k, const G<T>& g, int D, const Ix& I) {
    Ix Ir = I, Il = I;
    if (f.loc[D] == 0) { Ir = shift(I, D, +1); } else { Il = shift(I, D, -1); }
    int floc[3] = {f.loc[0], f.loc[1], f.loc[2]};
    floc[D] = 1 - floc[D];
    return (lerp(k, floc, Ir, &g) * dd(f, g, D, Ir) - lerp(k, floc, Il, &g) * dd(f, g, D, Il)) * isp_of(f, g, D, I);
}

// Julia max/min against 0.0 for floats (NaN propagates, max(-0.0,0.0)=0.0).
template <class T> inline T jmax0(T x) { return x != x ? x : (x > T(0) ? x : T(0)); }
template <class T> inline T jmin0(T x) { return x != x ? x : (x <= T(0) ? x : T(0)); }
template <class T> inline T jmax(T a, T b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}

// Launcher iteration space: I_d in [lo, lo+len) (src/KernelLaunch.jl:41,105-126;
// the default box is lo = 0, len = n+2, i.e. Offset(-1) over n .+ 2).
template <class F> void for_box(const orc_box* b, int nd, F fn) {
    long lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < nd ? b->lo[d] : 1;
        hi[d] = d < nd ? b->lo[d] + b->len[d] : 2;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (long k = lo[2]; k < hi[2]; ++k)
        for (long j = lo[1]; j < hi[1]; ++j)
            for (long i = lo[0]; i < hi[0]; ++i) {
                Ix I{{i, j, k}};
                fn(I);
            }
}

template <class T> T visc_ve(T eta, T G_, T dt) { return T(1) / (T(1) / eta + T(1) / (G_ * dt)); }

// ---------------------------------------------------------------- BCs
// bc_kernel! + first-order Dirichlet/Neumann
// (src/BoundaryConditions/batch.jl:159-184, first_order_boundary_condition.jl:34-84)
template <class T>
void bc_one(const G<T>& g, int D, int side, const FV<T>& f, const orc_bc& bc, Ix It) {
    long m = f.n[D];
    auto value = [&](const Ix& If) -> T {
        if (bc.vkind == 0) return T(0);
        if (bc.vkind == 1) return T(bc.value);
        FV<T> vf(bc.vfield);
        Ix J{{1, 1, 1}};
        int q = 0;
        for (int d = 0; d < g.nd; ++d) if (d != D) J.v[q++] = If.v[d];
        return vf(J);
    };
    Ix If = It;
    if (bc.kind == 0 && f.loc[D] == 1) {         // Vertex Dirichlet: boundary node
        If.v[D] = side == 1 ? 1 : m;
        f(If) = value(If);
        return;
    }
    If.v[D] = side == 1 ? 0 : m + 1;              // ghost
    Ix If2 = If; If2.v[D] = side == 1 ? 1 : m;    // neighbour
    const long ig = If.v[D], in = If2.v[D];
    if (bc.kind == 0) {                           // Center Dirichlet, t = 2 on uniform grids
        T a = value(If);
        T b = f(If2);
        T t = g_plant == PL_CENTER_DIRICHLET ? T(1) : T(2);
        if (g.xv[D])                              // t = 2 Δ(Center, I_f2) iΔ(Vertex, I_f) (:64-66)
            t = g_plant == PL_NU_DIRICHLET ? T((2.0 * g.dC(D, ig)) * (1.0 / g.dV(D, in)))
                                           : T((2.0 * g.dC(D, in)) * (1.0 / g.dV(D, ig)));
        f(If) = std::fma(t, a - b, b);
    } else {                                      // Neumann, any location: h = Δ(grid, flip(loc), dim, I_f)
        T q = value(If);
        T qs = side == 1 && g_plant != PL_NEUMANN_SIGN ? -q : q;
        T h = g.sp[D];
        if (g.xv[D]) {
            const int fl = g_plant == PL_NU_NEUMANN ? f.loc[D] : 1 - f.loc[D];
            h = T(fl == 0 ? g.dC(D, ig) : g.dV(D, ig));
        }
        f(If) = std::fma(h, qs, f(If2));
    }
}

template <class T>
void bc_side(const orc_grid* gg, int D, int side, int K, const orc_field* fs, const orc_bc* bcs) {
    G<T> g(gg);
    // tangential range 0..n_t+2 (worksize = remove_dim(dim, size(grid, Vertex()) .+ 2), I = J .- 1)
    long hi[3] = {1, 1, 1};
    int td[3], nt = 0;
    for (int d = 0; d < gg->nd; ++d) if (d != D) { td[nt] = d; hi[nt] = gg->n[d] + 3; ++nt; }
    for (long b = 0; b < hi[1]; ++b)
        for (long a = 0; a < hi[0]; ++a) {
            Ix It{{1, 1, 1}};
            if (nt > 0) It.v[td[0]] = a;
            if (nt > 1) It.v[td[1]] = b;
            for (int k = 0; k < K; ++k) {
                FV<T> f(&fs[k]);
                bc_one(g, D, side, f, bcs[k], It);
            }
        }
}

// ---------------------------------------------------------------- kernels
template <class T> void views(const orc_field* f, int n, FV<T>* out) {
    for (int i = 0; i < n; ++i) new (&out[i]) FV<T>(&f[i]);
}

}  // namespace

#define DISPATCH(dt, call) ((dt) == 0 ? call<double> : call<float>)

namespace {

// examples/diffusion_2d.jl:8-13
template <class T> void compute_q_t(const orc_grid* gg, const orc_box* b, const orc_field* q, const orc_field* C, double chi) {
    G<T> g(gg);
    FV<T> c(C);
    alignas(FV<T>) unsigned char buf[3 * sizeof(FV<T>)];
    FV<T>* Q = reinterpret_cast<FV<T>*>(buf);
    views(q, g.nd, Q);
    T nchi = g_plant == PL_DIFF_FLUX ? T(chi) : -T(chi);
    for_box(b, g.nd, [&](const Ix& I) {
        for (int d = 0; d < g.nd; ++d) Q[d](I) = nchi * dd(c, g, d, I);
    });
}
// examples/diffusion_2d.jl:15-19
template <class T> void update_C_t(const orc_grid* gg, const orc_box* b, const orc_field* C, const orc_field* q, double dt) {
    G<T> g(gg);
    FV<T> c(C);
    alignas(FV<T>) unsigned char buf[3 * sizeof(FV<T>)];
    FV<T>* Q = reinterpret_cast<FV<T>*>(buf);
    views(q, g.nd, Q);
    T Dt = T(dt);
    for_box(b, g.nd, [&](const Ix& I) { c(I) = c(I) - Dt * divg(Q, g, I); });
}

}  // namespace

// ---------------------------------------------------------------- Stokes
namespace {

struct Visc {
    // Scalar form (the reference, examples/stokes_3d_inc_ve_T.jl:23-57): eta, eta_ve given.
    // Field form (build extension): eta at cell centres on the full padded
    // storage; eta at a tau location = lerp(eta) to that location; eta_ve =
    // 1/(1/eta + 1/(G dt)) per point; velocity uses eta_ml (max over the
    // 3^N cell neighbourhood) as max(eta_ml[I-e_D], eta_ml[I]).
    const orc_field* eta;
    const orc_field* eta_ml;
    double eta_s, eta_ve_s, G, dt;
};

template <class T> void stress_t(const orc_grid* gg, const orc_box* b, const orc_field* tau, const orc_field* Pr_,
                                 const orc_field* dV_, const orc_field* V_, const orc_field* tau_old, Visc vs,
                                 double dtau_Pr, double dtau_r) {
    G<T> g(gg);
    const int nd = g.nd;
    const int nt = nd == 2 ? 3 : 6;
    alignas(FV<T>) unsigned char b1[6 * sizeof(FV<T>)], b2[6 * sizeof(FV<T>)], b3[3 * sizeof(FV<T>)];
    FV<T>* tau_v = reinterpret_cast<FV<T>*>(b1);
    FV<T>* tauo_v = reinterpret_cast<FV<T>*>(b2);
    FV<T>* V = reinterpret_cast<FV<T>*>(b3);
    views(tau, nt, tau_v);
    views(tau_old, nt, tauo_v);
    views(V_, nd, V);
    FV<T> Pr(Pr_), dV(dV_);
    const bool fieldvisc = vs.eta != nullptr;
    FV<T> eta(fieldvisc ? vs.eta : Pr_);
    const T Gdt = T(vs.G) * T(vs.dt);
    const T Gs = T(vs.G), dts = T(vs.dt);
    const T dPr = T(dtau_Pr), dr = T(dtau_r);
    const T eta_s = T(vs.eta_s), eve_s = T(vs.eta_ve_s);
    // tau component locations: 2D (xx,yy,xy), 3D (xx,yy,zz,xy,xz,yz)
    static const int L2[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 1, 0}};
    static const int L3[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    static const int L3swap[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
    const auto& L3u = g_plant == PL_EDGE_SWAP ? L3swap : L3;
    const T dev3 = g_plant == PL_STRESS_DEV ? T(0) : T(1);
    const T prs = g_plant == PL_PR_UPDATE ? T(-1) : T(1);
    for_box(b, nd, [&](const Ix& I) {
        T e[6];
        if (nd == 2) {
            e[0] = dd(V[0], g, 0, I);
            e[1] = dd(V[1], g, 1, I);
            e[2] = T(0.5) * (dd(V[0], g, 1, I) + dd(V[1], g, 0, I));
        } else {
            e[0] = dd(V[0], g, 0, I);
            e[1] = dd(V[1], g, 1, I);
            e[2] = dd(V[2], g, 2, I);
            e[3] = T(0.5) * (dd(V[0], g, 1, I) + dd(V[1], g, 0, I));
            e[4] = T(0.5) * (dd(V[0], g, 2, I) + dd(V[2], g, 0, I));
            e[5] = T(0.5) * (dd(V[1], g, 2, I) + dd(V[2], g, 1, I));
        }
        dV(I) = divg(V, g, I);
        T eta_c = eta_s, eve_c = eve_s;
        if (fieldvisc) { eta_c = eta(I); eve_c = visc_ve(eta_c, Gs, dts); }
        if (prs > T(0)) Pr(I) = Pr(I) - (dV(I) * eve_c) * dPr;
        else Pr(I) = Pr(I) + (dV(I) * eve_c) * dPr;
        const int ndiag = nd;
        for (int c = 0; c < nt; ++c) {
            T et = eta_c, ev = eve_c;
            if (fieldvisc && c >= ndiag) {
                et = lerp(eta, nd == 2 ? L2[c] : L3u[c], I);
                ev = visc_ve(et, Gs, dts);
            }
            T t = tau_v[c](I), to = tauo_v[c](I);
            T r;
            if (c < ndiag) r = ((-(t - to)) / Gdt - t / et) + T(2) * (e[c] - dev3 * (dV(I) / T(3)));
            else r = ((-(t - to)) / Gdt - t / et) + T(2) * e[c];
            tau_v[c](I) = t + (r * ev) * dr;
        }
    });
}

// fV (build extension, optional): a forcing added to r_V after rho g, the
// form in which defect correction and adjoint solves drive the iteration.
template <class T> void velocity_t(const orc_grid* gg, const orc_box* b, const orc_field* V_, const orc_field* rV_,
                                   const orc_field* Pr_, const orc_field* tau, const orc_field* rhog_, Visc vs,
                                   double nudtau, const orc_field* fV_ = nullptr) {
    G<T> g(gg);
    const int nd = g.nd;
    const int nt = nd == 2 ? 3 : 6;
    alignas(FV<T>) unsigned char b1[6 * sizeof(FV<T>)], b2[3 * sizeof(FV<T>)], b3[3 * sizeof(FV<T>)];
    FV<T>* t = reinterpret_cast<FV<T>*>(b1);
    FV<T>* V = reinterpret_cast<FV<T>*>(b2);
    FV<T>* rV = reinterpret_cast<FV<T>*>(b3);
    views(tau, nt, t);
    views(V_, nd, V);
    views(rV_, nd, rV);
    FV<T> Pr(Pr_);
    const bool hasrg = rhog_ != nullptr;
    FV<T> rg(hasrg ? rhog_ : Pr_);
    const bool hasf = fV_ != nullptr;
    alignas(FV<T>) unsigned char b4[3 * sizeof(FV<T>)];
    FV<T>* fV = reinterpret_cast<FV<T>*>(b4);
    views(hasf ? fV_ : V_, nd, fV);
    const bool fieldvisc = vs.eta_ml != nullptr;
    FV<T> eml(fieldvisc ? vs.eta_ml : Pr_);
    const T nu = T(nudtau), eve_s = T(vs.eta_ve_s), Gs = T(vs.G), dts = T(vs.dt);
    const bool rvswap = g_plant == PL_RV_SWAP;
    const long mlside = g_plant == PL_VEL_ML_SIDE ? 1 : 0;
    for_box(b, nd, [&](const Ix& I) {
        T r[3];
        if (nd == 2) {
            // examples/stokes_2d_inc_ve_T.jl:39-40
            r[0] = (-dd(Pr, g, 0, I) + dd(t[0], g, 0, I)) + dd(t[2], g, rvswap ? 0 : 1, I);
            r[1] = ((-dd(Pr, g, 1, I) + dd(t[1], g, 1, I)) + dd(t[2], g, 0, I));
            if (hasrg) r[1] = r[1] - rg(I);
        } else {
            // examples/stokes_3d_inc_ve_T.jl:51-53
            r[0] = ((-dd(Pr, g, 0, I) + dd(t[0], g, 0, I)) + dd(t[3], g, 1, I)) + dd(t[rvswap ? 5 : 4], g, 2, I);
            r[1] = ((-dd(Pr, g, 1, I) + dd(t[1], g, 1, I)) + dd(t[3], g, 0, I)) + dd(t[5], g, 2, I);
            r[2] = ((-dd(Pr, g, 2, I) + dd(t[2], g, 2, I)) + dd(t[4], g, 0, I)) + dd(t[5], g, 1, I);
            if (hasrg) r[2] = r[2] - rg(I);
        }
        if (hasf)
            for (int d = 0; d < nd; ++d) r[d] = r[d] + fV[d](I);
        for (int d = 0; d < nd; ++d) {
            rV[d](I) = r[d];
            T ev = eve_s;
            if (fieldvisc) ev = visc_ve(jmax(eml(shift(I, d, mlside - 1)), eml(shift(I, d, mlside))), Gs, dts);
            V[d](I) = V[d](I) + (rV[d](I) * nu) / ev;
        }
    });
}

// Residuals of a PT state (defect correction), f64: r_tau, the bracket of
// update_stress! (stokes_3d_inc_ve_T.jl:34-39); -divg(V); and r_V of
// update_velocity! (stokes_3d_inc_ve_T.jl:51-53).
template <class T> void residual_t(const orc_grid* gg, const orc_box* b, const orc_field* V_, const orc_field* tau,
                                   const orc_field* Pr_, const orc_field* tauo, const orc_field* eta_, const orc_field* rg_,
                                   double eta_s, double G_, double dt_, const orc_field* Rt, const orc_field* RV,
                                   const orc_field* RP) {
    G<T> g(gg);
    const int nd = g.nd, nt = nd == 2 ? 3 : 6;
    alignas(FV<T>) unsigned char b1[6 * sizeof(FV<T>)], b2[6 * sizeof(FV<T>)], b3[3 * sizeof(FV<T>)],
        b4[6 * sizeof(FV<T>)], b5[3 * sizeof(FV<T>)];
    FV<T>* t = reinterpret_cast<FV<T>*>(b1);
    FV<T>* to = reinterpret_cast<FV<T>*>(b2);
    FV<T>* V = reinterpret_cast<FV<T>*>(b3);
    FV<T>* rt = reinterpret_cast<FV<T>*>(b4);
    FV<T>* rv = reinterpret_cast<FV<T>*>(b5);
    views(tau, nt, t);
    views(tauo, nt, to);
    views(V_, nd, V);
    views(Rt, nt, rt);
    views(RV, nd, rv);
    FV<T> Pr(Pr_), rp(RP);
    const bool fe = eta_ != nullptr, hr = rg_ != nullptr;
    FV<T> eta(fe ? eta_ : Pr_), rg(hr ? rg_ : Pr_);
    const T Gdt = T(G_) * T(dt_);
    static const int L2[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 1, 0}};
    static const int L3[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    for_box(b, nd, [&](const Ix& I) {
        T e[6];
        if (nd == 2) {
            e[0] = dd(V[0], g, 0, I); e[1] = dd(V[1], g, 1, I);
            e[2] = T(0.5) * (dd(V[0], g, 1, I) + dd(V[1], g, 0, I));
        } else {
            e[0] = dd(V[0], g, 0, I); e[1] = dd(V[1], g, 1, I); e[2] = dd(V[2], g, 2, I);
            e[3] = T(0.5) * (dd(V[0], g, 1, I) + dd(V[1], g, 0, I));
            e[4] = T(0.5) * (dd(V[0], g, 2, I) + dd(V[2], g, 0, I));
            e[5] = T(0.5) * (dd(V[1], g, 2, I) + dd(V[2], g, 1, I));
        }
        const T dv = divg(V, g, I);
        for (int c = 0; c < nt; ++c) {
            T et = T(eta_s);
            if (fe) et = c < nd ? eta(I) : lerp(eta, nd == 2 ? L2[c] : L3[c], I);
            const T tt = t[c](I), tto = to[c](I);
            rt[c](I) = c < nd ? ((-(tt - tto)) / Gdt - tt / et) + T(2) * (e[c] - dv / T(3))
                              : ((-(tt - tto)) / Gdt - tt / et) + T(2) * e[c];
        }
        rp(I) = g_plant == PL_RESID_RP ? dv : -dv;
        T r[3];
        if (nd == 2) {
            r[0] = (-dd(Pr, g, 0, I) + dd(t[0], g, 0, I)) + dd(t[2], g, 1, I);
            r[1] = ((-dd(Pr, g, 1, I) + dd(t[1], g, 1, I)) + dd(t[2], g, 0, I));
            if (hr) r[1] = r[1] - rg(I);
        } else {
            r[0] = ((-dd(Pr, g, 0, I) + dd(t[0], g, 0, I)) + dd(t[3], g, 1, I)) + dd(t[4], g, 2, I);
            r[1] = ((-dd(Pr, g, 1, I) + dd(t[1], g, 1, I)) + dd(t[3], g, 0, I)) + dd(t[5], g, 2, I);
            r[2] = ((-dd(Pr, g, 2, I) + dd(t[2], g, 2, I)) + dd(t[4], g, 0, I)) + dd(t[5], g, 1, I);
            if (hr) r[2] = r[2] - rg(I);
        }
        for (int d = 0; d < nd; ++d) rv[d](I) = r[d];
    });
}

// Gradient of a functional J of the converged PT Stokes state with respect
// to the Center viscosity field (build extension).  With the steady stress
// equations R_tau = -(tau - tau_old)/(G dt) - tau/eta_loc + 2 dev(eps) = 0,
// the adjoint multiplier of a stored stress component is -tau_a/2 (diagonal)
// or -tau_a (edge), tau_a the stresses of the adjoint Stokes solve (forcing
// -dJ/dV, homogeneous BCs, tau_old = 0), and dR_tau/deta_loc = tau/eta_loc^2:
//   g(c) = [c interior] sum_d tau_a_dd tau_dd / (2 eta_c eta_c)
//        + 0.25 sum over the edges e of each off-diagonal component that
//          lerp eta from c (vertex index c or c+1 in its vertex dims) and lie
//          in the active range (vertex 1..n+1, centre 1..n) of
//          tau_a_e tau_e / (eta_e eta_e),  eta_e = lerp(eta) to e;
// logscale != 0 returns dJ/dlog(eta) = eta_c g(c).  Evaluated over centres
// 0..n+1; fixed summation order (diagonal d ascending; components xy, xz, yz;
// edges with the first vertex dimension fastest).
template <class T> void stokes_grad_t(const orc_grid* gg, const orc_field* tau, const orc_field* taua,
                                      const orc_field* eta_, const orc_field* grad_, int logscale) {
    G<T> g(gg);
    const int nd = g.nd, nt = nd == 2 ? 3 : 6;
    alignas(FV<T>) unsigned char b1[6 * sizeof(FV<T>)], b2[6 * sizeof(FV<T>)];
    FV<T>* t = reinterpret_cast<FV<T>*>(b1);
    FV<T>* ta = reinterpret_cast<FV<T>*>(b2);
    views(tau, nt, t);
    views(taua, nt, ta);
    FV<T> eta(eta_), gr(grad_);
    static const int L2[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 1, 0}};
    static const int L3[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    orc_box b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.len[d] = d < nd ? g.n[d] + 2 : 1; }
    for_box(&b, nd, [&](const Ix& I) {
        const T ec = eta(I);
        bool interior = true;
        for (int d = 0; d < nd; ++d) interior = interior && I.v[d] >= 1 && I.v[d] <= g.n[d];
        T gd = T(0);
        if (interior)
            for (int d = 0; d < nd; ++d) gd = gd + (ta[d](I) * t[d](I)) / (T(2) * ec * ec);
        T ge = T(0);
        for (int c = nd; c < nt; ++c) {
            const int* L = nd == 2 ? L2[c] : L3[c];
            int vd[2], nv = 0;
            for (int d = 0; d < nd; ++d) if (L[d] == 1) vd[nv++] = d;
            for (int bb = 0; bb < 2; ++bb)
                for (int aa = 0; aa < 2; ++aa) {
                    Ix E = I;
                    E.v[vd[0]] += aa;
                    E.v[vd[1]] += bb;
                    bool act = true;
                    for (int d = 0; d < nd; ++d) {
                        const long hi = L[d] == 1 ? g.n[d] + 1 : g.n[d];
                        act = act && E.v[d] >= 1 && E.v[d] <= hi;
                    }
                    if (!act) continue;
                    const T et = lerp(eta, L, E);
                    ge = ge + (ta[c](E) * t[c](E)) / (et * et);
                }
        }
        T r = gd + (g_plant == PL_GRAD_EDGE ? T(0.5) : T(0.25)) * ge;
        if (logscale) r = r * ec;
        gr(I) = r;
    });
}

// examples/stokes_3d_inc_ve_T.jl:11-21 (2D: stokes_2d_inc_ve_T.jl:11-18)
template <class T> void update_old_t(const orc_box* b, int nd, const orc_field* Tf, const orc_field* tau,
                                     const orc_field* Told, const orc_field* tauo) {
    const int nt = nd == 2 ? 3 : 6;
    alignas(FV<T>) unsigned char b1[6 * sizeof(FV<T>)], b2[6 * sizeof(FV<T>)];
    FV<T>* t = reinterpret_cast<FV<T>*>(b1);
    FV<T>* to = reinterpret_cast<FV<T>*>(b2);
    views(tau, nt, t);
    views(tauo, nt, to);
    const bool hasT = Tf != nullptr;
    FV<T> TT(hasT ? Tf : tau), TO(hasT ? Told : tauo);
    for_box(b, nd, [&](const Ix& I) {
        if (hasT) TO(I) = TT(I);
        for (int c = 0; c < nt; ++c) to[c](I) = t[c](I);
    });
}

// examples/stokes_3d_inc_ve_T.jl:59-71
template <class T> void thermal_flux_t(const orc_grid* gg, const orc_box* b, const orc_field* qT, const orc_field* Tf,
                                       const orc_field* V_, double lam) {
    G<T> g(gg);
    alignas(FV<T>) unsigned char b1[3 * sizeof(FV<T>)], b2[3 * sizeof(FV<T>)];
    FV<T>* q = reinterpret_cast<FV<T>*>(b1);
    FV<T>* V = reinterpret_cast<FV<T>*>(b2);
    views(qT, g.nd, q);
    views(V_, g.nd, V);
    FV<T> TT(Tf);
    const T nl = -T(lam);
    const bool swp = g_plant == PL_THERMAL_UPWIND;
    for_box(b, g.nd, [&](const Ix& I) {
        for (int d = 0; d < g.nd; ++d) {
            T v = V[d](I);
            const T lt = swp ? rightv(TT, d, I) : leftv(TT, d, I), rt = swp ? leftv(TT, d, I) : rightv(TT, d, I);
            q[d](I) = (nl * dd(TT, g, d, I) + jmax0(v) * lt) + jmin0(v) * rt;
        }
    });
}
// examples/stokes_3d_inc_ve_T.jl:73-77
template <class T> void thermal_t(const orc_grid* gg, const orc_box* b, const orc_field* Tf, const orc_field* Told,
                                  const orc_field* qT, double dt) {
    G<T> g(gg);
    alignas(FV<T>) unsigned char b1[3 * sizeof(FV<T>)];
    FV<T>* q = reinterpret_cast<FV<T>*>(b1);
    views(qT, g.nd, q);
    FV<T> TT(Tf), TO(Told);
    const T Dt = g_plant == PL_THERMAL_SIGN ? -T(dt) : T(dt);
    for_box(b, g.nd, [&](const Ix& I) { TT(I) = TO(I) - Dt * divg(q, g, I); });
}

template <class T> void maxloc_t(const orc_box* b, const orc_field* dst, const orc_field* src) {
    FV<T> d(dst), s(src);
    const int nd = d.nd;
    const long up = g_plant == PL_MAXLOC_2N ? 0 : 1;
    for_box(b, nd, [&](const Ix& I) {
        T m = s(I);
        long r[3];
        for (int q = 0; q < 3; ++q) r[q] = q < nd ? 1 : 0;
        for (long c = -r[2]; c <= r[2] * up; ++c)
            for (long bb = -r[1]; bb <= r[1] * up; ++bb)
                for (long a = -r[0]; a <= r[0] * up; ++a) {
                    Ix J{{I.v[0] + a, I.v[1] + bb, I.v[2] + c}};
                    m = jmax(m, s(J));
                }
        d(I) = m;
    });
}

// ---------------------------------------------------------------- reductions
template <class T> double maxabs_t(const orc_field* f, const orc_box* b) {
    FV<T> v(f);
    double m = 0.0;
    long lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < v.nd ? b->lo[d] : 1;
        hi[d] = d < v.nd ? b->lo[d] + b->len[d] : 2;
    }
    bool nan = false;
    for (long k = lo[2]; k < hi[2]; ++k)
        for (long j = lo[1]; j < hi[1]; ++j)
            for (long i = lo[0]; i < hi[0]; ++i) {
                double a = std::fabs(double(v(Ix{{i, j, k}})));
                if (a != a) nan = true;
                if (a > m) m = a;
            }
    return nan ? NAN : m;
}
template <class T> double sumsq_t(const orc_field* f, const orc_box* b) {
    FV<T> v(f);
    double s = 0.0;
    long lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < v.nd ? b->lo[d] : 1;
        hi[d] = d < v.nd ? b->lo[d] + b->len[d] : 2;
    }
    for (long k = lo[2]; k < hi[2]; ++k)
        for (long j = lo[1]; j < hi[1]; ++j)
            for (long i = lo[0]; i < hi[0]; ++i) {
                double a = double(v(Ix{{i, j, k}}));
                s += a * a;
            }
    return s;
}

// ---------------------------------------------------------------- generic operators (test kernels)
// op: 0 ∂_D, 1 lerp to dst location, 2 divg, 3 vmag, 4 ∂²_D, 5 ∂k∂_D, 6 lapl, 7 divg_grad,
//     8 left_D, 9 right_D, 10 hlerp to dst location, 11 δ_D
template <class T> void op_t(int op, const orc_grid* gg, const orc_box* b, const orc_field* dst, int nsrc,
                             const orc_field* src, const orc_field* k, int D) {
    G<T> g(gg);
    FV<T> d(dst);
    alignas(FV<T>) unsigned char b1[3 * sizeof(FV<T>)];
    FV<T>* s = reinterpret_cast<FV<T>*>(b1);
    views(src, nsrc, s);
    const bool hk = k != nullptr;
    FV<T> kk(hk ? k : src);
    for_box(b, g.nd, [&](const Ix& I) {
        T v = 0;
        switch (op) {
            case 0: v = dd(s[0], g, D, I); break;
            case 1: v = lerp(s[0], d.loc, I, &g); break;
            case 2: v = divg(s, g, I); break;
            case 3: v = vmag(s, g, I); break;
            case 4: v = d2(s[0], g, D, I); break;
            case 5: v = dkd(s[0], kk, g, D, I); break;
            case 6: { v = d2(s[0], g, 0, I); for (int q = 1; q < g.nd; ++q) v = v + d2(s[0], g, q, I); } break;
            case 7: { v = dkd(s[0], kk, g, 0, I); for (int q = 1; q < g.nd; ++q) v = v + dkd(s[0], kk, g, q, I); } break;
            case 8: v = leftv(s[0], D, I); break;
            case 9: v = rightv(s[0], D, I); break;
            case 10: v = hlerp(s[0], d.loc, I, &g); break;
            case 11:   // δ (partial_derivatives.jl:2)
                v = g_plant == PL_DELTA_SIGN ? leftv(s[0], D, I) - rightv(s[0], D, I) : rightv(s[0], D, I) - leftv(s[0], D, I);
                break;
        }
        d(I) = v;
    });
}

// ---------------------------------------------------------------- masked operators
// src/GridOperators/masked_operators.jl:11-39: left/right read the field at
// the knot and multiply by the mask at the field's own location at that knot,
// left(f, ω)[I] = f[Il]·at(ω, loc(f), Il); δ, ∂, ∂², ∂k∂ are built on them
// as the unmasked forms are (∂k∂'s coefficient lerp is not masked), and
// divg/lapl/divg_grad are the left-associated sums of
// masked_field_operators.jl:56-127.  `w` is at(ω, loc(f)): the FieldMask
// component at the source's location (field_mask.jl:1-117).
template <class T> inline T leftm(const FV<T>& f, const FV<T>& w, int D, const Ix& I) {
    const Ix J = f.loc[D] == 0 ? shift(I, D, -1) : I;
    return f(J) * (g_plant == PL_MASK_AT_OUT ? w(I) : w(J));
}
template <class T> inline T rightm(const FV<T>& f, const FV<T>& w, int D, const Ix& I) {
    const Ix J = f.loc[D] == 0 ? I : shift(I, D, +1);
    return f(J) * (g_plant == PL_MASK_AT_OUT ? w(I) : w(J));
}
template <class T> inline T deltam(const FV<T>& f, const FV<T>& w, int D, const Ix& I) {
    return rightm(f, w, D, I) - leftm(f, w, D, I);
}
template <class T> inline T ddm(const FV<T>& f, const FV<T>& w, const G<T>& g, int D, const Ix& I) {
    return deltam(f, w, D, I) * isp_of(f, g, D, I);
}
template <class T> T d2m(const FV<T>& f, const FV<T>& w, const G<T>& g, int D, const Ix& I) {
    Ix Ir = I, Il = I;
    if (f.loc[D] == 0) { Ir = shift(I, D, +1); } else { Il = shift(I, D, -1); }
    return (ddm(f, w, g, D, Ir) - ddm(f, w, g, D, Il)) * isp_of(f, g, D, I);
}
template <class T> T dkdm(const FV<T>& f, const FV<T>& w, const FV<T>& k, const G<T>& g, int D, const Ix& I) {
    Ix Ir = I, Il = I;
    if (f.loc[D] == 0) { Ir = shift(I, D, +1); } else { Il = shift(I, D, -1); }
    int floc[3] = {f.loc[0], f.loc[1], f.loc[2]};
    floc[D] = 1 - floc[D];
    return (lerp(k, floc, Ir, &g) * ddm(f, w, g, D, Ir) - lerp(k, floc, Il, &g) * ddm(f, w, g, D, Il)) *
           isp_of(f, g, D, I);
}

// op as op_t (0 ∂, 2 divg, 4 ∂², 5 ∂k∂, 6 lapl, 7 divg_grad, 8 left, 9 right, 11 δ);
// masks[q] is the mask component at src[q]'s location.
template <class T> void op_masked_t(int op, const orc_grid* gg, const orc_box* b, const orc_field* dst, int nsrc,
                                    const orc_field* src, const orc_field* masks, const orc_field* k, int D) {
    G<T> g(gg);
    FV<T> d(dst);
    alignas(FV<T>) unsigned char b1[3 * sizeof(FV<T>)], b2[3 * sizeof(FV<T>)];
    FV<T>* s = reinterpret_cast<FV<T>*>(b1);
    FV<T>* w = reinterpret_cast<FV<T>*>(b2);
    views(src, nsrc, s);
    views(masks, nsrc, w);
    FV<T> kk(k ? k : src);
    for_box(b, g.nd, [&](const Ix& I) {
        T v = 0;
        switch (op) {
            case 0: v = ddm(s[0], w[0], g, D, I); break;
            case 2: { v = ddm(s[0], w[0], g, 0, I); for (int q = 1; q < g.nd; ++q) v = v + ddm(s[q], w[q], g, q, I); } break;
            case 4: v = d2m(s[0], w[0], g, D, I); break;
            case 5: v = dkdm(s[0], w[0], kk, g, D, I); break;
            case 6: { v = d2m(s[0], w[0], g, 0, I); for (int q = 1; q < g.nd; ++q) v = v + d2m(s[0], w[0], g, q, I); } break;
            case 7: { v = dkdm(s[0], w[0], kk, g, 0, I); for (int q = 1; q < g.nd; ++q) v = v + dkdm(s[0], w[0], kk, g, q, I); } break;
            case 8: v = leftm(s[0], w[0], D, I); break;
            case 9: v = rightm(s[0], w[0], D, I); break;
            case 11: v = deltam(s[0], w[0], D, I); break;
            default: fail("oracle: operator without a masked form");
        }
        d(I) = v;
    });
}

// ---------------------------------------------------------------- SIA (examples/sia.jl)
// Lazy FunctionFields B (given array), S = B + H, ∇S = ∂S, D, M are evaluated
// at every access, as a FunctionField does (src/Fields/function_field.jl:53-65).
// Build form: D = (A[I]·H^(n+2) + As·H^n)·vmag(∇S)^(n-1) with per-cell rate
// factor A (A ≡ 1 gives the reference's D, sia.jl:46-49); powers are plain
// repeated products; M = min(mb_a·(S - mb_b), mb_c) (sia.jl:51-54 has
// mb_a = 0.5, mb_b = 0.3, mb_c = 0.25).
struct SiaP { double As; int npow; double mb_a, mb_b, mb_c; };

template <class T> inline T ipow(T x, int n) { T r = T(1); for (int i = 0; i < n; ++i) r = r * x; return r; }

template <class T> struct Sia {
    FV<T> H, B, A;
    const G<T>& g;
    SiaP p;
    bool hasA;
    T S(Ix I) const { return B(I) + H(I); }
    T dSx(Ix I) const { return (S(I) - S(shift(I, 0, -1))) * g.isp[0]; }   // (Vertex, Center)
    T dSy(Ix I) const { return (S(I) - S(shift(I, 1, -1))) * g.isp[1]; }   // (Center, Vertex)
    T vmagS(Ix I) const {
        T a = std::fma(T(0.5), dSx(shift(I, 0, 1)) - dSx(I), dSx(I));
        T b = std::fma(T(0.5), dSy(shift(I, 1, 1)) - dSy(I), dSy(I));
        return std::sqrt(a * a + b * b);
    }
    T D(Ix I) const {
        T h = H(I);
        T a = hasA ? A(I) : T(1);
        const int e2 = g_plant == PL_SIA_EXP ? p.npow + 1 : p.npow + 2;
        const T grad = g_plant == PL_SIA_GRAD ? T(1) : ipow(vmagS(I), p.npow - 1);
        return (a * ipow(h, e2) + T(p.As) * ipow(h, p.npow)) * grad;
    }
    T M(Ix I) const {
        T m = T(p.mb_a) * (S(I) - T(p.mb_b));
        T c = T(p.mb_c);
        if (g_plant == PL_SIA_M) return m > c ? m : c;
        return m < c ? m : c;
    }
};

template <class T> void sia_q_t(const orc_grid* gg, const orc_field* q, const orc_field* H, const orc_field* B,
                               const orc_field* A, SiaP p) {
    G<T> g(gg);
    Sia<T> s{FV<T>(H), FV<T>(B), FV<T>(A ? A : H), g, p, A != nullptr};
    FV<T> qx(&q[0]), qy(&q[1]);
    orc_box b{{1, 1, 1}, {g.n[0], g.n[1], 1}};
    for_box(&b, 2, [&](const Ix& I) {
        // q.x = -lerp(D, (Vertex,Center)) * ∂x(S): knots D[I-e_x], D[I]
        T Dl = s.D(shift(I, 0, -1)), Dr = s.D(I);
        qx(I) = -std::fma(T(0.5), Dr - Dl, Dl) * s.dSx(I);
        T El = s.D(shift(I, 1, -1)), Er = s.D(I);
        qy(I) = -std::fma(T(0.5), Er - El, El) * s.dSy(I);
    });
}
template <class T> void sia_H_t(const orc_grid* gg, const orc_field* H, const orc_field* Hout, const orc_field* q,
                               const orc_field* B, SiaP p, double dt) {
    G<T> g(gg);
    Sia<T> s{FV<T>(H), FV<T>(B), FV<T>(H), g, p, false};
    FV<T> Ho(Hout);
    alignas(FV<T>) unsigned char b1[2 * sizeof(FV<T>)];
    FV<T>* Q = reinterpret_cast<FV<T>*>(b1);
    views(q, 2, Q);
    orc_box b{{1, 1, 1}, {g.n[0], g.n[1], 1}};
    const T Dt = T(dt);
    for_box(&b, 2, [&](const Ix& I) {
        T v = s.H(I) + Dt * (-divg(Q, g, I) + s.M(I));
        Ho(I) = jmax0(v);
    });
}
template <class T> double sia_Dmax_t(const orc_grid* gg, const orc_field* H, const orc_field* B, const orc_field* A, SiaP p) {
    G<T> g(gg);
    Sia<T> s{FV<T>(H), FV<T>(B), FV<T>(A ? A : H), g, p, A != nullptr};
    double m = -INFINITY;
    for (long j = 1; j <= g.n[1]; ++j)
        for (long i = 1; i <= g.n[0]; ++i) m = std::max(m, double(s.D(Ix{{i, j, 1}})));
    return m;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

// OpenMP threads of the host loops: n > 0 sets n, 0 restores the count the
// process started with (the CPU baseline's 1-thread figure)
int orc_set_threads(int n) {
    static const int initial = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : initial);
    return 0;
}

int orc_last_error(char* buf, int len) {
    int e = g_err;
    if (buf && len > 0) { std::snprintf(buf, len, "%s", e ? g_msg : ""); }
    g_err = 0;
    return e;
}

// Planted mistakes (see enum Plant): returns the previous setting.
int orc_set_plant(int k) { int old = g_plant; g_plant = k; return old; }
int orc_get_plant(void) { return g_plant; }

// UniformAxis coordinates (src/Grids/uniform_axis.jl:18-19).
double orc_vertex(double origin, double spacing, long i) { return std::fma(double(i - 1), spacing, origin); }
double orc_center(double origin, double spacing, long i) {
    return std::fma(double(i - 1), spacing, std::fma(0.5, spacing, origin));
}
void orc_coords(double origin, double spacing, int loc, long first, long count, double* out) {
    for (long q = 0; q < count; ++q)
        out[q] = loc == 0 ? orc_center(origin, spacing, first + q) : orc_vertex(origin, spacing, first + q);
}
// FunctionAxis coordinates (function_axis.jl:9, abstract_axis.jl:18,20-21): xv holds
// vertex(i) for i = -kPad-1 .. at xv[i + kPad + 1]
void orc_axis_coords(const double* xv, int loc, long first, long count, double* out) {
    for (long q = 0; q < count; ++q) {
        const long i = first + q + kPad + 1;
        out[q] = loc == 0 ? 0.5 * (xv[i] + xv[i + 1]) : xv[i];
    }
}
float orc_vertex_f32(float origin, float spacing, long i) { return std::fma(float(i - 1), spacing, origin); }
float orc_center_f32(float origin, float spacing, long i) {
    return std::fma(float(i - 1), spacing, std::fma(0.5f, spacing, origin));
}

int orc_bc_side(const orc_grid* g, int dtype, int D, int side, int K, const orc_field* f, const orc_bc* bcs) {
    if (dtype == 0) bc_side<double>(g, D, side, K, f, bcs); else bc_side<float>(g, D, side, K, f, bcs);
    return g_err;
}
// bc!(arch, grid, batch) ordering: dims N..1, side 1 then side 2 (batch.jl:20-29).
// ks[D*2+s] is the number of fields in that side's batch; fields/bcs are
// concatenated in that same (D, side) order.
int orc_bc_set(const orc_grid* g, int dtype, const int* ks, const orc_field* f, const orc_bc* bcs) {
    int offs[7] = {0};
    for (int q = 0; q < 2 * g->nd; ++q) offs[q + 1] = offs[q] + ks[q];
    for (int D = g->nd - 1; D >= 0; --D)
        for (int s = 0; s < 2; ++s) {
            int q = 2 * D + s;
            if (ks[q] > 0) orc_bc_side(g, dtype, D, s + 1, ks[q], f + offs[q], bcs + offs[q]);
        }
    return g_err;
}

int orc_compute_q(const orc_grid* g, const orc_box* b, const orc_field* q, const orc_field* C, double chi) {
    if (C->dtype == 0) compute_q_t<double>(g, b, q, C, chi); else compute_q_t<float>(g, b, q, C, chi);
    return g_err;
}
int orc_update_C(const orc_grid* g, const orc_box* b, const orc_field* C, const orc_field* q, double dt) {
    if (C->dtype == 0) update_C_t<double>(g, b, C, q, dt); else update_C_t<float>(g, b, C, q, dt);
    return g_err;
}
int orc_update_old(const orc_box* b, int nd, const orc_field* T, const orc_field* tau, const orc_field* Told,
                   const orc_field* tauo) {
    if (tau->dtype == 0) update_old_t<double>(b, nd, T, tau, Told, tauo); else update_old_t<float>(b, nd, T, tau, Told, tauo);
    return g_err;
}
int orc_update_stress(const orc_grid* g, const orc_box* b, const orc_field* tau, const orc_field* Pr,
                      const orc_field* dV, const orc_field* V, const orc_field* tauo, const orc_field* eta,
                      double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r) {
    Visc vs{eta, nullptr, eta_s, eta_ve_s, G, dt};
    if (Pr->dtype == 0) stress_t<double>(g, b, tau, Pr, dV, V, tauo, vs, dtau_Pr, dtau_r);
    else stress_t<float>(g, b, tau, Pr, dV, V, tauo, vs, dtau_Pr, dtau_r);
    return g_err;
}
int orc_update_velocity(const orc_grid* g, const orc_box* b, const orc_field* V, const orc_field* rV,
                        const orc_field* Pr, const orc_field* tau, const orc_field* rhog, const orc_field* eta_ml,
                        double eta_ve_s, double G, double dt, double nudtau, const orc_field* fV) {
    Visc vs{nullptr, eta_ml, 0.0, eta_ve_s, G, dt};
    if (Pr->dtype == 0) velocity_t<double>(g, b, V, rV, Pr, tau, rhog, vs, nudtau, fV);
    else velocity_t<float>(g, b, V, rV, Pr, tau, rhog, vs, nudtau, fV);
    return g_err;
}
int orc_stokes_residual(const orc_grid* g, const orc_box* b, const orc_field* V, const orc_field* tau,
                        const orc_field* Pr, const orc_field* tauo, const orc_field* eta, const orc_field* rg,
                        double eta_s, double G, double dt, const orc_field* Rt, const orc_field* RV,
                        const orc_field* RP) {
    residual_t<double>(g, b, V, tau, Pr, tauo, eta, rg, eta_s, G, dt, Rt, RV, RP);
    return g_err;
}
int orc_stokes_grad_eta(const orc_grid* g, const orc_field* tau, const orc_field* taua, const orc_field* eta,
                        const orc_field* grad, int logscale) {
    stokes_grad_t<double>(g, tau, taua, eta, grad, logscale);
    return g_err;
}

int orc_thermal_flux(const orc_grid* g, const orc_box* b, const orc_field* qT, const orc_field* T,
                     const orc_field* V, double lam) {
    if (T->dtype == 0) thermal_flux_t<double>(g, b, qT, T, V, lam); else thermal_flux_t<float>(g, b, qT, T, V, lam);
    return g_err;
}
int orc_thermal(const orc_grid* g, const orc_box* b, const orc_field* T, const orc_field* Told,
                const orc_field* qT, double dt) {
    if (T->dtype == 0) thermal_t<double>(g, b, T, Told, qT, dt); else thermal_t<float>(g, b, T, Told, qT, dt);
    return g_err;
}
int orc_maxloc(const orc_box* b, const orc_field* dst, const orc_field* src) {
    if (src->dtype == 0) maxloc_t<double>(b, dst, src); else maxloc_t<float>(b, dst, src);
    return g_err;
}
double orc_maxabs(const orc_field* f, const orc_box* b) {
    return f->dtype == 0 ? maxabs_t<double>(f, b) : maxabs_t<float>(f, b);
}
double orc_sumsq(const orc_field* f, const orc_box* b) {
    return f->dtype == 0 ? sumsq_t<double>(f, b) : sumsq_t<float>(f, b);
}
int orc_op(int op, const orc_grid* g, const orc_box* b, const orc_field* dst, int nsrc, const orc_field* src,
           const orc_field* k, int D) {
    if (dst->dtype == 0) op_t<double>(op, g, b, dst, nsrc, src, k, D); else op_t<float>(op, g, b, dst, nsrc, src, k, D);
    return g_err;
}
int orc_op_masked(int op, const orc_grid* g, const orc_box* b, const orc_field* dst, int nsrc,
                  const orc_field* src, const orc_field* masks, const orc_field* k, int D) {
    if (dst->dtype == 0) op_masked_t<double>(op, g, b, dst, nsrc, src, masks, k, D);
    else op_masked_t<float>(op, g, b, dst, nsrc, src, masks, k, D);
    return g_err;
}
int orc_sia_flux(const orc_grid* g, const orc_field* q, const orc_field* H, const orc_field* B, const orc_field* A,
                 double As, int npow, double mb_a, double mb_b, double mb_c) {
    SiaP p{As, npow, mb_a, mb_b, mb_c};
    sia_q_t<double>(g, q, H, B, A, p);
    return g_err;
}
int orc_sia_update(const orc_grid* g, const orc_field* H, const orc_field* Hout, const orc_field* q,
                   const orc_field* B, double As, int npow, double mb_a, double mb_b, double mb_c, double dt) {
    SiaP p{As, npow, mb_a, mb_b, mb_c};
    sia_H_t<double>(g, H, Hout, q, B, p, dt);
    return g_err;
}
double orc_sia_Dmax(const orc_grid* g, const orc_field* H, const orc_field* B, const orc_field* A,
                    double As, int npow, double mb_a, double mb_b, double mb_c) {
    SiaP p{As, npow, mb_a, mb_b, mb_c};
    return sia_Dmax_t<double>(g, H, B, A, p);
}

}  // extern "C"
