// Per-point Launcher kernels: generic GridOperators, diffusion, update_old!,
// thermal flux/update and the viscosity max-neighbourhood.  Each thread owns
// one logical index I of the box; blocks are one wave64 wide along x (the
// contiguous dim) so loads and stores coalesce.  These are bandwidth-bound
// streaming kernels; the fused Stokes iteration lives in stokes.hip.
#include <cstring>

#include "common.hpp"

namespace chmy {

constexpr int kBX = 64, kBY = 4;

struct Ix3 {
    int64_t i, j, k;
};

__device__ __forceinline__ Ix3 box_index(const Box3& b, bool& ok) {
    int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t c = blockIdx.y * (int64_t)blockDim.y + threadIdx.y;
    int64_t e = blockIdx.z * (int64_t)blockDim.z + threadIdx.z;
    ok = a < b.len[0] && c < b.len[1] && e < b.len[2];
    return Ix3{b.lo[0] + a, b.lo[1] + c, b.lo[2] + e};
}

template <class T>
__device__ __forceinline__ T& g_at(const DFL<T>& f, Ix3 I) { return f(I.i, I.j, I.k); }
__device__ __forceinline__ Ix3 sh(Ix3 I, int d, int64_t s) {
    if (d == 0) I.i += s; else if (d == 1) I.j += s; else I.k += s;
    return I;
}
// il/ir of an operator on a field at its own location (GridOperators.jl:23-36)
template <class T>
__device__ __forceinline__ T leftv(const DFL<T>& f, int D, Ix3 I) { return f.loc[D] == 0 ? g_at(f, sh(I, D, -1)) : g_at(f, I); }
template <class T>
__device__ __forceinline__ T rightv(const DFL<T>& f, int D, Ix3 I) { return f.loc[D] == 0 ? g_at(f, I) : g_at(f, sh(I, D, 1)); }
template <class T>
__device__ __forceinline__ T ddv(const DFL<T>& f, T isp, int D, Ix3 I) { return (rightv(f, D, I) - leftv(f, D, I)) * isp; }

// lerp (interpolation.jl:14-94): knots over the differing dims, reduced with
// fma(0.5, b-a, a) from the first differing dim (innermost) to the last.
template <class T>
__device__ T lerpv(const DFL<T>& f, const int32_t* to, Ix3 I) {
    int dims[3], m = 0;
    for (int d = 0; d < f.nd; ++d)
        if (f.loc[d] != to[d]) dims[m++] = d;
    T v[8];
    for (int q = 0; q < (1 << m); ++q) {
        Ix3 J = I;
        for (int b = 0; b < m; ++b) {
            int D = dims[b];
            bool right = (q >> b) & 1;
            // Center->Vertex knots (I-e, I); Vertex->Center knots (I, I+e)
            int64_t s = f.loc[D] == 0 ? (right ? 0 : -1) : (right ? 1 : 0);
            J = sh(J, D, s);
        }
        v[q] = g_at(f, J);
    }
    for (int b = 0; b < m; ++b) {
        int half = 1 << (m - b - 1);
        for (int q = 0; q < half; ++q) {
            T a = v[2 * q], c = v[2 * q + 1];
            v[q] = fma(T(0.5), c - a, a);
        }
    }
    return v[0];
}

struct OpArgs {
    int op, nd, D, nsrc;
};

template <class T>
__global__ void __launch_bounds__(256) op_kernel(Box3 b, DFL<T> dst, DFL<T> s0, DFL<T> s1, DFL<T> s2, OpArgs a,
                                                 T isp0, T isp1, T isp2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    const DFL<T>* s[3] = {&s0, &s1, &s2};
    const T isp[3] = {isp0, isp1, isp2};
    T v = T(0);
    switch (a.op) {
        case CHMY_OP_PARTIAL: v = ddv(s0, isp[a.D], a.D, I); break;
        case CHMY_OP_LERP: v = lerpv(s0, dst.loc, I); break;
        case CHMY_OP_DIVG: {
            v = ddv(*s[0], isp[0], 0, I);
            for (int d = 1; d < a.nd; ++d) v = v + ddv(*s[d], isp[d], d, I);
        } break;
        case CHMY_OP_VMAG: {
            const int32_t c[3] = {0, 0, 0};
            T x = lerpv(*s[0], c, I);
            T acc = x * x;
            for (int d = 1; d < a.nd; ++d) { T y = lerpv(*s[d], c, I); acc = acc + y * y; }
            v = sqrt(acc);
        } break;
        case CHMY_OP_LEFT: v = leftv(s0, a.D, I); break;
        case CHMY_OP_RIGHT: v = rightv(s0, a.D, I); break;
    }
    g_at(dst, I) = v;
}

// ------------------------------------------------------------------ masked and second-order operators
// Masked forms (masked_operators.jl:11-39): left/right multiply the knot
// value by the mask at the field's own location at that knot,
// left(f, ω)[I] = f[Il]·ω[Il]; with M = false the plain forms, bitwise those
// of leftv/rightv/ddv.  ∂² and ∂k∂ (partial_derivatives.jl:7-21) take
// differences of ∂ at the two knots of the flipped location; ∂k∂ weights
// them with lerp(k) at those knots (the coefficient is never masked).
//
// On a FunctionAxis (NU, non-NULL table along the dim) the metrics are the
// reference's: ∂ scales δ by iΔ(grid, loc(f), dim, I) (partial_derivatives.jl:5,
// field_operators.jl:20-24: the source field's location, at the output
// index), ∂² and ∂k∂ their difference by iΔ(loc(f), I) (:7-21), and the
// Center -> Vertex lerp weight at vertex i is 0.5 Δc(i-1) iΔv(i)
// (interpolation.jl:29); Vertex -> Center stays 0.5 (:28).
__device__ __forceinline__ int64_t ix_of(Ix3 I, int D) { return D == 0 ? I.i : D == 1 ? I.j : I.k; }

template <class T, bool NU>
__device__ __forceinline__ T isp_at(const Axes3& ax, const T* isp, int D, int loc, Ix3 I) {
    if (!NU || !ax.a[D].p) return isp[D];
    return T(ax.a[D].at(loc == 0 ? CHMY_AX_IDC : CHMY_AX_IDV, ix_of(I, D)));
}

template <class T, bool M>
__device__ __forceinline__ T knot(const DFL<T>& f, const DFL<T>& w, Ix3 J) {
    const T v = g_at(f, J);
    return M ? v * g_at(w, J) : v;
}
template <class T, bool M>
__device__ __forceinline__ T leftk(const DFL<T>& f, const DFL<T>& w, int D, Ix3 I) {
    return knot<T, M>(f, w, f.loc[D] == 0 ? sh(I, D, -1) : I);
}
template <class T, bool M>
__device__ __forceinline__ T rightk(const DFL<T>& f, const DFL<T>& w, int D, Ix3 I) {
    return knot<T, M>(f, w, f.loc[D] == 0 ? I : sh(I, D, 1));
}
template <class T, bool M, bool NU>
__device__ __forceinline__ T ddk(const DFL<T>& f, const DFL<T>& w, const Axes3& ax, const T* isp, int D, Ix3 I) {
    return (rightk<T, M>(f, w, D, I) - leftk<T, M>(f, w, D, I)) * isp_at<T, NU>(ax, isp, D, f.loc[D], I);
}
// knots of the flipped location around I: Center f -> (I, I+e), Vertex f -> (I-e, I)
__device__ __forceinline__ void flip_knots(int loc, int D, Ix3 I, Ix3& Il, Ix3& Ir) {
    Il = loc == 0 ? I : sh(I, D, -1);
    Ir = loc == 0 ? sh(I, D, 1) : I;
}
// lerp and hlerp with the weights of the axis (HARM: the harmonic rule
// inv(muladd(t, inv(b) - inv(a), inv(a))), interpolation.jl:15,101); knots
// over the differing dims, first differing dim innermost
template <class T, bool NU, bool HARM>
__device__ T itpw(const DFL<T>& f, const int32_t* to, const Axes3& ax, Ix3 I) {
    int dims[3], m = 0;
    for (int d = 0; d < f.nd; ++d)
        if (f.loc[d] != to[d]) dims[m++] = d;
    T v[8];
    for (int q = 0; q < (1 << m); ++q) {
        Ix3 J = I;
        for (int b = 0; b < m; ++b) {
            int D = dims[b];
            bool right = (q >> b) & 1;
            int64_t s = f.loc[D] == 0 ? (right ? 0 : -1) : (right ? 1 : 0);
            J = sh(J, D, s);
        }
        v[q] = g_at(f, J);
    }
    for (int b = 0; b < m; ++b) {
        const int D = dims[b];
        const T t = (NU && ax.a[D].p && f.loc[D] == 0) ? T(ax.a[D].at(CHMY_AX_WCV, ix_of(I, D))) : T(0.5);
        int half = 1 << (m - b - 1);
        for (int q = 0; q < half; ++q) {
            T a = v[2 * q], c = v[2 * q + 1];
            v[q] = HARM ? T(1) / fma(t, T(1) / c - T(1) / a, T(1) / a) : fma(t, c - a, a);
        }
    }
    return v[0];
}
template <class T, bool M, bool NU>
__device__ T d2k(const DFL<T>& f, const DFL<T>& w, const Axes3& ax, const T* isp, int D, Ix3 I) {
    Ix3 Il, Ir;
    flip_knots(f.loc[D], D, I, Il, Ir);
    return (ddk<T, M, NU>(f, w, ax, isp, D, Ir) - ddk<T, M, NU>(f, w, ax, isp, D, Il)) *
           isp_at<T, NU>(ax, isp, D, f.loc[D], I);
}
template <class T, bool M, bool NU>
__device__ T dkdk(const DFL<T>& f, const DFL<T>& w, const DFL<T>& k, const Axes3& ax, const T* isp, int D, Ix3 I) {
    Ix3 Il, Ir;
    flip_knots(f.loc[D], D, I, Il, Ir);
    int32_t floc[3] = {f.loc[0], f.loc[1], f.loc[2]};
    floc[D] = 1 - floc[D];
    return (itpw<T, NU, false>(k, floc, ax, Ir) * ddk<T, M, NU>(f, w, ax, isp, D, Ir) -
            itpw<T, NU, false>(k, floc, ax, Il) * ddk<T, M, NU>(f, w, ax, isp, D, Il)) *
           isp_at<T, NU>(ax, isp, D, f.loc[D], I);
}

template <class T, bool M, bool NU>
__global__ void __launch_bounds__(256) op_ext_kernel(Box3 b, DFL<T> dst, DFL<T> s0, DFL<T> s1, DFL<T> s2, DFL<T> w0,
                                                     DFL<T> w1, DFL<T> w2, DFL<T> k, OpArgs a, T isp0, T isp1,
                                                     T isp2, Axes3 ax) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    const DFL<T>* s[3] = {&s0, &s1, &s2};
    const DFL<T>* w[3] = {&w0, &w1, &w2};
    const T isp[3] = {isp0, isp1, isp2};
    T v = T(0);
    switch (a.op) {
        case CHMY_OP_PARTIAL: v = ddk<T, M, NU>(s0, w0, ax, isp, a.D, I); break;
        case CHMY_OP_DELTA: v = rightk<T, M>(s0, w0, a.D, I) - leftk<T, M>(s0, w0, a.D, I); break;
        case CHMY_OP_LEFT: v = leftk<T, M>(s0, w0, a.D, I); break;
        case CHMY_OP_RIGHT: v = rightk<T, M>(s0, w0, a.D, I); break;
        case CHMY_OP_DIVG: {
            v = ddk<T, M, NU>(*s[0], *w[0], ax, isp, 0, I);
            for (int d = 1; d < a.nd; ++d) v = v + ddk<T, M, NU>(*s[d], *w[d], ax, isp, d, I);
        } break;
        case CHMY_OP_D2: v = d2k<T, M, NU>(s0, w0, ax, isp, a.D, I); break;
        case CHMY_OP_DKD: v = dkdk<T, M, NU>(s0, w0, k, ax, isp, a.D, I); break;
        case CHMY_OP_LAPL: {
            v = d2k<T, M, NU>(s0, w0, ax, isp, 0, I);
            for (int d = 1; d < a.nd; ++d) v = v + d2k<T, M, NU>(s0, w0, ax, isp, d, I);
        } break;
        case CHMY_OP_DIVG_GRAD: {
            v = dkdk<T, M, NU>(s0, w0, k, ax, isp, 0, I);
            for (int d = 1; d < a.nd; ++d) v = v + dkdk<T, M, NU>(s0, w0, k, ax, isp, d, I);
        } break;
        case CHMY_OP_LERP: v = itpw<T, NU, false>(s0, dst.loc, ax, I); break;
        case CHMY_OP_HLERP: v = itpw<T, NU, true>(s0, dst.loc, ax, I); break;
        case CHMY_OP_VMAG: {
            const int32_t c[3] = {0, 0, 0};
            T x = itpw<T, NU, false>(*s[0], c, ax, I);
            T acc = x * x;
            for (int d = 1; d < a.nd; ++d) { T y = itpw<T, NU, false>(*s[d], c, ax, I); acc = acc + y * y; }
            v = sqrt(acc);
        } break;
    }
    g_at(dst, I) = v;
}

// ------------------------------------------------------------------ diffusion
template <class T>
__global__ void __launch_bounds__(256) diffusion_flux_kernel(Box3 b, DFL<T> C, DFL<T> q0, DFL<T> q1, DFL<T> q2,
                                                             int nd, T nchi, T i0, T i1, T i2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    // q.x[I] = -χ * ∂x(C)[I] (diffusion_2d.jl:11-12): Center C -> (C[I] - C[I-e]) iΔ
    g_at(q0, I) = nchi * ((g_at(C, I) - g_at(C, sh(I, 0, -1))) * i0);
    if (nd > 1) g_at(q1, I) = nchi * ((g_at(C, I) - g_at(C, sh(I, 1, -1))) * i1);
    if (nd > 2) g_at(q2, I) = nchi * ((g_at(C, I) - g_at(C, sh(I, 2, -1))) * i2);
}

template <class T>
__global__ void __launch_bounds__(256) diffusion_update_kernel(Box3 b, DFL<T> C, DFL<T> q0, DFL<T> q1, DFL<T> q2,
                                                               int nd, T dt, T i0, T i1, T i2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    // C[I] -= Δt * divg(q)[I] (diffusion_2d.jl:18); Vertex q_D -> (q[I+e] - q[I]) iΔ
    T dv = (g_at(q0, sh(I, 0, 1)) - g_at(q0, I)) * i0;
    if (nd > 1) dv = dv + (g_at(q1, sh(I, 1, 1)) - g_at(q1, I)) * i1;
    if (nd > 2) dv = dv + (g_at(q2, sh(I, 2, 1)) - g_at(q2, I)) * i2;
    g_at(C, I) = g_at(C, I) - dt * dv;
}

// compute_q! and update_C! in one launch (build extension): q at I from C at
// I, I-e (stored, as compute_q! stores it), q at I+e recomputed the same way
// when I+e is inside the box, else read from q's storage (what update_C! would
// read there); C_out[I] = C[I] - Δt divg(q)[I].  C and C_out are distinct
// buffers, so no thread reads a C value another has advanced.
// The step's new C at I (box b): q_D at I (stored when store_q, as
// compute_q! stores it), q_D at I+e recomputed when I+e is in the box, else
// read from q's storage; C[I] - Δt divg(q)[I].
template <class T>
__device__ __forceinline__ T diffusion_value(const Box3& b, const DFL<T>& C, const DFL<T>& q0, const DFL<T>& q1,
                                             const DFL<T>& q2, int nd, T nchi, T dt, const T* isp, Ix3 I,
                                             bool store_q) {
    const T c = g_at(C, I);
    const int64_t hi[3] = {b.lo[0] + b.len[0], b.lo[1] + b.len[1], b.lo[2] + b.len[2]};
    const int64_t at[3] = {I.i, I.j, I.k};
    T dv = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= nd) break;
        const DFL<T>& qd = d == 0 ? q0 : (d == 1 ? q1 : q2);
        const T ql = nchi * ((c - g_at(C, sh(I, d, -1))) * isp[d]);
        if (store_q) g_at(qd, I) = ql;
        const Ix3 J = sh(I, d, 1);
        const T qr = at[d] + 1 < hi[d] ? nchi * ((g_at(C, J) - c) * isp[d]) : g_at(qd, J);
        const T t = (qr - ql) * isp[d];
        dv = d == 0 ? t : dv + t;
    }
    return c - dt * dv;
}

template <class T>
__global__ void __launch_bounds__(256) diffusion_step_kernel(Box3 b, DFL<T> C, DFL<T> Co, DFL<T> q0, DFL<T> q1,
                                                             DFL<T> q2, int nd, T nchi, T dt, T i0, T i1, T i2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    const T isp[3] = {i0, i1, i2};
    g_at(Co, I) = diffusion_value(b, C, q0, q1, q2, nd, nchi, dt, isp, I, true);
}

// One scalar rule per (dim, side) of a Center field's batch.
struct SideRules {
    int32_t kind[3][2];   // CHMY_DIRICHLET / CHMY_NEUMANN
    double val[3][2];
    double h[3];          // spacings (the Neumann step)
};

// The step and the C batch in one launch, when every (dim, side) carries one
// scalar rule: one thread per point of [0, n+2]^N, which holds everything the
// step (box [0, n+1]^N) and the batch (ghost planes 0 and n+1, tangential
// 0..n+2) write.  A point that no dim writes keeps the step's value (or is
// left alone outside the box).  A point that some dims write takes the
// batch's final value, the chain bc_set_kernel evaluates: its inner point Q
// (every ghost coordinate moved inward) as the step leaves it (recomputed
// here when Q is in the box, else read as stored: no thread writes it), then
// each writing dim's rule from the highest to the lowest, the batch's own fma
// steps.  q is stored on the box as the step stores it.
template <class T>
__global__ void __launch_bounds__(256) diffusion_step_bc_kernel(Box3 box, Box3 reg, DFL<T> C, DFL<T> Co, DFL<T> q0,
                                                                DFL<T> q1, DFL<T> q2, int nd, T nchi, T dt, T i0,
                                                                T i1, T i2, SideRules r) {
    bool ok;
    const Ix3 P = box_index(reg, ok);
    if (!ok) return;
    const T isp[3] = {i0, i1, i2};
    const int64_t p[3] = {P.i, P.j, P.k};
    int64_t m[3], qc[3];
    int gs[3];
    bool in_box = true, ghost = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        m[d] = box.len[d] - 2;   // cells: the box is [0, n+1]
        qc[d] = p[d];
        gs[d] = -1;
        if (d >= nd) continue;
        in_box = in_box && p[d] <= m[d] + 1;
        if (p[d] == 0 || p[d] == m[d] + 1) {
            gs[d] = p[d] == 0 ? 0 : 1;
            qc[d] = p[d] == 0 ? 1 : m[d];
            ghost = true;
        }
    }
    T v = T(0);
    if (in_box) v = diffusion_value(box, C, q0, q1, q2, nd, nchi, dt, isp, P, true);
    if (!ghost) {
        if (in_box) g_at(Co, P) = v;
        return;
    }
    const Ix3 Q{qc[0], qc[1], qc[2]};
    bool q_in = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (d < nd) q_in = q_in && qc[d] <= m[d] + 1;
    T w = q_in ? diffusion_value(box, C, q0, q1, q2, nd, nchi, dt, isp, Q, false) : g_at(Co, Q);
#pragma unroll
    for (int d = 2; d >= 0; --d) {
        if (d >= nd || gs[d] < 0) continue;
        const T val = T(r.val[d][gs[d]]);
        if (r.kind[d][gs[d]] == CHMY_DIRICHLET) w = fma(T(2), val - w, w);
        else w = fma(T(r.h[d]), gs[d] == 0 ? -val : val, w);
    }
    g_at(Co, P) = w;
}

// ------------------------------------------------------------------ update_old!
template <class T>
struct Copy7 {
    DFL<T> src[7], dst[7];
    int n;
};

template <class T>
__global__ void __launch_bounds__(256) copy_kernel(Box3 b, Copy7<T> c) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    for (int q = 0; q < c.n; ++q) g_at(c.dst[q], I) = g_at(c.src[q], I);
}

// ------------------------------------------------------------------ thermal
template <class T>
__global__ void __launch_bounds__(256) thermal_flux_kernel(Box3 b, DFL<T> Tf, DFL<T> V0, DFL<T> V1, DFL<T> V2,
                                                           DFL<T> q0, DFL<T> q1, DFL<T> q2, int nd, T nlam, T i0,
                                                           T i1, T i2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    const T isp[3] = {i0, i1, i2};
    const T tc = g_at(Tf, I);
#pragma unroll
    for (int d = 0; d < 3; ++d) {   // constant indices once unrolled: no stack arrays
        if (d >= nd) break;
        const DFL<T>& Vd = d == 0 ? V0 : (d == 1 ? V1 : V2);
        const DFL<T>& qd = d == 0 ? q0 : (d == 1 ? q1 : q2);
        // qT_D = ((-λ ∂_D T) + max(V_D,0) left_D T) + min(V_D,0) right_D T
        // (stokes_3d_inc_ve_T.jl:59-71); T Center: left = T[I-e], right = T[I]
        const T tl = g_at(Tf, sh(I, d, -1));
        const T v = g_at(Vd, I);
        g_at(qd, I) = (nlam * ((tc - tl) * isp[d]) + jmax0(v) * tl) + jmin0(v) * tc;
    }
}

// update_thermal_flux! and update_thermal! in one launch (build extension):
// qT at I from T at I, I-e and V at I (stored, as update_thermal_flux! stores
// it), qT at I+e recomputed the same way when I+e is inside the box, else
// read from qT's storage; T_out[I] = T_old[I] - Δt divg(qT)[I].  T and T_out
// are distinct buffers.
template <class T>
__global__ void __launch_bounds__(256) thermal_step_kernel(Box3 b, DFL<T> Tf, DFL<T> Tout, DFL<T> To, DFL<T> V0,
                                                           DFL<T> V1, DFL<T> V2, DFL<T> q0, DFL<T> q1, DFL<T> q2,
                                                           int nd, T nlam, T dt, T i0, T i1, T i2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    const int64_t hi[3] = {b.lo[0] + b.len[0], b.lo[1] + b.len[1], b.lo[2] + b.len[2]};
    const int64_t at[3] = {I.i, I.j, I.k};
    const T isp[3] = {i0, i1, i2};
    const T tc = g_at(Tf, I);
    T dv = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= nd) break;
        const DFL<T>& Vd = d == 0 ? V0 : (d == 1 ? V1 : V2);
        const DFL<T>& qd = d == 0 ? q0 : (d == 1 ? q1 : q2);
        const T tl = g_at(Tf, sh(I, d, -1));
        const T v = g_at(Vd, I);
        const T ql = (nlam * ((tc - tl) * isp[d]) + jmax0(v) * tl) + jmin0(v) * tc;
        g_at(qd, I) = ql;
        const Ix3 J = sh(I, d, 1);
        T qr;
        if (at[d] + 1 < hi[d]) {
            const T tr = g_at(Tf, J), vr = g_at(Vd, J);
            qr = (nlam * ((tr - tc) * isp[d]) + jmax0(vr) * tc) + jmin0(vr) * tr;
        } else {
            qr = g_at(qd, J);
        }
        const T t = (qr - ql) * isp[d];
        dv = d == 0 ? t : dv + t;
    }
    g_at(Tout, I) = g_at(To, I) - dt * dv;
}

// The same step in 3D as a z-march: a block is 64 (x, one wave) x kThY (y)
// points and walks a chunk of at most 16 planes, carrying T at the plane and z's flux
// (qT_z at k+1 computed at plane k is qT_z at plane k+1, the same
// expression on the same operands), so T and V_z stream from HBM once
// instead of being fetched again by the blocks of the neighbouring planes.
// x and y neighbours are loads of the same or adjacent rows, found in L2:
// blocks are remapped so that each XCD walks a contiguous run of tiles.
// Per point: the same expressions, association and stored qT as
// thermal_step_kernel; fields of one staggered location share strides
// (the host checks it).  Lanes outside the box load at clamped indices and
// store nothing.
constexpr int kThY = 4;   // rows per block; with 16-plane chunks the fastest of 1-16 rows x 8-64 planes

template <class T>
__device__ __forceinline__ T th_flux(T tc, T tl, T v, T nlam, T isp) {
    return (nlam * ((tc - tl) * isp) + jmax0(v) * tl) + jmin0(v) * tc;
}

template <class T>
__global__ void __launch_bounds__(64 * kThY) thermal_step3_kernel(Box3 b, DF<T> Tf, DF<T> Tout, DF<T> To, DF<T> V0,
                                                                  DF<T> V1, DF<T> V2, DF<T> q0, DF<T> q1, DF<T> q2,
                                                                  T nlam, T dt, T i0, T i1, T i2, int nbx, int nby,
                                                                  int kchunk) {
    const int lane = threadIdx.x, r = threadIdx.y;
    const int lt = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = lt % nbx, by = (lt / nbx) % nby, bz = lt / (nbx * nby);
    const int64_t hi0 = b.lo[0] + b.len[0], hi1 = b.lo[1] + b.len[1], hi2 = b.lo[2] + b.len[2];
    const int64_t i = b.lo[0] + (int64_t)bx * 64 + lane, j = b.lo[1] + (int64_t)by * kThY + r;
    const int64_t kb = b.lo[2] + (int64_t)bz * kchunk;
    if (kb >= hi2) return;   // a last chunk past the box (whole block: no barrier follows)
    const int64_t ke = kb + kchunk < hi2 ? kb + kchunk : hi2;
    const bool act = i < hi0 && j < hi1;
    const int64_t ic = i < hi0 ? i : hi0 - 1, jc = j < hi1 ? j : hi1 - 1;
    const bool xr = ic + 1 < hi0, yr = jc + 1 < hi1;   // the right face's flux is computed, else read
    const int64_t sC = Tf.sz, sX = V0.sz, sY = V1.sz, sZ = V2.sz;
    const int64_t yC = Tf.sy, yY = V1.sy;
    int64_t oC = ic + jc * yC + kb * sC, oX = ic + jc * V0.sy + kb * sX;
    int64_t oY = ic + jc * yY + kb * sY, oZ = ic + jc * V2.sy + kb * sZ;
    T tc = Tf.p[oC];
    T qlz = th_flux(tc, Tf.p[oC - sC], V2.p[oZ], nlam, i2);
    for (int64_t k = kb; k < ke; ++k) {
        const T tp = Tf.p[oC + sC], vz1 = V2.p[oZ + sZ];
        const T txm = Tf.p[oC - 1], txp = Tf.p[oC + 1];
        const T tym = Tf.p[oC - yC], typ = Tf.p[oC + yC];
        const T vx = V0.p[oX], vx1 = V0.p[oX + 1];
        const T vy = V1.p[oY], vy1 = V1.p[oY + yY];
        const T to = To.p[oC];
        const T qlx = th_flux(tc, txm, vx, nlam, i0);
        const T qly = th_flux(tc, tym, vy, nlam, i1);
        const T qrx = xr ? th_flux(txp, tc, vx1, nlam, i0) : q0.p[oX + 1];
        const T qry = yr ? th_flux(typ, tc, vy1, nlam, i1) : q1.p[oY + yY];
        const T qrz = k + 1 < hi2 ? th_flux(tp, tc, vz1, nlam, i2) : q2.p[oZ + sZ];
        const T dv = ((qrx - qlx) * i0 + (qry - qly) * i1) + (qrz - qlz) * i2;
        if (act) {
            q0.p[oX] = qlx;
            q1.p[oY] = qly;
            q2.p[oZ] = qlz;
            Tout.p[oC] = to - dt * dv;
        }
        tc = tp;
        qlz = qrz;
        oC += sC;
        oX += sX;
        oY += sY;
        oZ += sZ;
    }
}

template <class T>
__global__ void __launch_bounds__(256) thermal_update_kernel(Box3 b, DFL<T> Tf, DFL<T> To, DFL<T> q0, DFL<T> q1,
                                                             DFL<T> q2, int nd, T dt, T i0, T i1, T i2) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    T dv = (g_at(q0, sh(I, 0, 1)) - g_at(q0, I)) * i0;
    if (nd > 1) dv = dv + (g_at(q1, sh(I, 1, 1)) - g_at(q1, I)) * i1;
    if (nd > 2) dv = dv + (g_at(q2, sh(I, 2, 1)) - g_at(q2, I)) * i2;
    g_at(Tf, I) = g_at(To, I) - dt * dv;
}

// ------------------------------------------------------------------ maxloc
template <class T>
__global__ void __launch_bounds__(256) maxloc_kernel(Box3 b, DFL<T> dst, DFL<T> src, int nd) {
    bool ok;
    Ix3 I = box_index(b, ok);
    if (!ok) return;
    T m = g_at(src, I);
    const int64_t rz = nd > 2 ? 1 : 0, ry = nd > 1 ? 1 : 0;
    for (int64_t c = -rz; c <= rz; ++c)
        for (int64_t bb = -ry; bb <= ry; ++bb)
            for (int64_t a = -1; a <= 1; ++a) m = jmax(m, g_at(src, Ix3{I.i + a, I.j + bb, I.k + c}));
    g_at(dst, I) = m;
}

// ------------------------------------------------------------------ host side
template <class T>
DFL<T> lv_or(const chmy_field_t* f, const DFL<T>& fallback) {
    return f ? lview<T>(*f) : fallback;
}

static dim3 blk2() { return dim3(kBX, kBY, 1); }

template <class T>
static int diffusion_launch(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* q,
                            const chmy_field_t* Cf, double scalar, bool flux, hipStream_t s) {
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    DFL<T> c = lview<T>(*Cf);
    DFL<T> q0 = lview<T>(q[0]);
    DFL<T> q1 = g->nd > 1 ? lview<T>(q[1]) : q0, q2 = g->nd > 2 ? lview<T>(q[2]) : q0;
    dim3 grd = grid_for(b, blk2());
    if (flux)
        hipLaunchKernelGGL(diffusion_flux_kernel<T>, grd, blk2(), 0, s, b, c, q0, q1, q2, g->nd, -T(scalar),
                           T(g->inv_spacing[0]), T(g->inv_spacing[1]), T(g->inv_spacing[2]));
    else
        hipLaunchKernelGGL(diffusion_update_kernel<T>, grd, blk2(), 0, s, b, c, q0, q1, q2, g->nd, T(scalar),
                           T(g->inv_spacing[0]), T(g->inv_spacing[1]), T(g->inv_spacing[2]));
    CHMY_HIP(hipGetLastError());
    return 0;
}

template <class T>
static int copy_launch(const chmy_grid_t* g, const chmy_box_t& bx, const chmy_field_t* Tf, const chmy_field_t* tau,
                       const chmy_field_t* To, const chmy_field_t* tauo, hipStream_t s) {
    Copy7<T> c;
    int nt = g->nd == 2 ? 3 : 6;
    c.n = 0;
    if (Tf) { c.src[c.n] = lview<T>(*Tf); c.dst[c.n] = lview<T>(*To); ++c.n; }
    for (int q = 0; q < nt; ++q) { c.src[c.n] = lview<T>(tau[q]); c.dst[c.n] = lview<T>(tauo[q]); ++c.n; }
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    hipLaunchKernelGGL(copy_kernel<T>, grid_for(b, blk2()), blk2(), 0, s, b, c);
    CHMY_HIP(hipGetLastError());
    return 0;
}

template <class T>
static int thermal_launch(const chmy_grid_t* g, const chmy_box_t& bx, const chmy_field_t* q, const chmy_field_t* Tf,
                          const chmy_field_t* other, double scalar, bool flux, hipStream_t s) {
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    DFL<T> t = lview<T>(*Tf), o = lview<T>(other[0]);
    DFL<T> q0 = lview<T>(q[0]);
    DFL<T> q1 = g->nd > 1 ? lview<T>(q[1]) : q0, q2 = g->nd > 2 ? lview<T>(q[2]) : q0;
    dim3 grd = grid_for(b, blk2());
    if (flux) {
        DFL<T> v1 = g->nd > 1 ? lview<T>(other[1]) : o, v2 = g->nd > 2 ? lview<T>(other[2]) : o;
        hipLaunchKernelGGL(thermal_flux_kernel<T>, grd, blk2(), 0, s, b, t, o, v1, v2, q0, q1, q2, g->nd, -T(scalar),
                           T(g->inv_spacing[0]), T(g->inv_spacing[1]), T(g->inv_spacing[2]));
    } else {
        hipLaunchKernelGGL(thermal_update_kernel<T>, grd, blk2(), 0, s, b, t, o, q0, q1, q2, g->nd, T(scalar),
                           T(g->inv_spacing[0]), T(g->inv_spacing[1]), T(g->inv_spacing[2]));
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_op(int op, const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* dst, int nsrc,
            const chmy_field_t* src, int dim, void* stream) {
    CHMY_TRY(check_grid_any(g));
    if (nonuniform(g)) return chmy_op_masked(op, g, box, dst, nsrc, src, nullptr, nullptr, dim, stream);
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_field(g, dst, "dst", nullptr, -1));
    const int dt = dst->dtype;
    const bool vec = op == CHMY_OP_DIVG || op == CHMY_OP_VMAG;
    if (vec && nsrc != g->nd) return set_error(CHMY_E_ARG, "divg/vmag need nd source fields");
    if (!vec && nsrc != 1) return set_error(CHMY_E_ARG, "operator needs one source field");
    if (!src) return set_error(CHMY_E_ARG, "src is NULL");
    if (op != CHMY_OP_PARTIAL && op != CHMY_OP_LERP && op != CHMY_OP_DIVG && op != CHMY_OP_VMAG &&
        op != CHMY_OP_LEFT && op != CHMY_OP_RIGHT)
        return set_error(CHMY_E_ARG, "unknown operator %d", op);
    if ((op == CHMY_OP_PARTIAL || op == CHMY_OP_LEFT || op == CHMY_OP_RIGHT) && (dim < 0 || dim >= g->nd))
        return set_error(CHMY_E_ARG, "dim %d out of range", dim);
    for (int q = 0; q < nsrc; ++q) {
        const int* want = nullptr;
        int loc[3];
        if (op == CHMY_OP_DIVG || op == CHMY_OP_VMAG) {
            for (int d = 0; d < 3; ++d) loc[d] = d == q ? 1 : 0;
            want = loc;
        }
        CHMY_TRY(check_field(g, &src[q], "src", want, dt));
    }
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    OpArgs a{op, g->nd, dim, nsrc};
    hipStream_t s = (hipStream_t)stream;
    dim3 grd = grid_for(b, blk2());
    if (dt == CHMY_F64) {
        DFL<double> d = lview<double>(*dst), s0 = lview<double>(src[0]);
        DFL<double> s1 = nsrc > 1 ? lview<double>(src[1]) : s0, s2 = nsrc > 2 ? lview<double>(src[2]) : s0;
        hipLaunchKernelGGL(op_kernel<double>, grd, blk2(), 0, s, b, d, s0, s1, s2, a, g->inv_spacing[0],
                           g->inv_spacing[1], g->inv_spacing[2]);
    } else {
        DFL<float> d = lview<float>(*dst), s0 = lview<float>(src[0]);
        DFL<float> s1 = nsrc > 1 ? lview<float>(src[1]) : s0, s2 = nsrc > 2 ? lview<float>(src[2]) : s0;
        hipLaunchKernelGGL(op_kernel<float>, grd, blk2(), 0, s, b, d, s0, s1, s2, a, (float)g->inv_spacing[0],
                           (float)g->inv_spacing[1], (float)g->inv_spacing[2]);
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_op_masked(int op, const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* dst, int nsrc,
                   const chmy_field_t* src, const chmy_field_t* masks, const chmy_field_t* k, int dim, void* stream) {
    CHMY_TRY(check_grid_any(g));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_field(g, dst, "dst", nullptr, -1));
    const int dt = dst->dtype;
    const bool vec = op == CHMY_OP_DIVG || op == CHMY_OP_VMAG;
    const bool per_dim = op == CHMY_OP_PARTIAL || op == CHMY_OP_DELTA || op == CHMY_OP_LEFT || op == CHMY_OP_RIGHT ||
                         op == CHMY_OP_D2 || op == CHMY_OP_DKD;
    const bool known = per_dim || vec || op == CHMY_OP_LERP || op == CHMY_OP_HLERP || op == CHMY_OP_LAPL ||
                       op == CHMY_OP_DIVG_GRAD;
    if (!known) return set_error(CHMY_E_ARG, "unknown operator %d", op);
    const bool maskable = op != CHMY_OP_LERP && op != CHMY_OP_HLERP && op != CHMY_OP_VMAG;
    const bool needs_k = op == CHMY_OP_DKD || op == CHMY_OP_DIVG_GRAD;
    if (vec && nsrc != g->nd) return set_error(CHMY_E_ARG, "divg/vmag need nd source fields");
    if (!vec && nsrc != 1) return set_error(CHMY_E_ARG, "operator needs one source field");
    if (!src) return set_error(CHMY_E_ARG, "src is NULL");
    if (per_dim && (dim < 0 || dim >= g->nd)) return set_error(CHMY_E_ARG, "dim %d out of range", dim);
    if (masks && !maskable) return set_error(CHMY_E_ARG, "operator %d has no masked form", op);
    if (needs_k != (k != nullptr)) return set_error(CHMY_E_ARG, needs_k ? "k is NULL" : "k is given to an operator without a coefficient");
    for (int q = 0; q < nsrc; ++q) {
        const int* want = nullptr;
        int loc[3];
        if (vec) {
            for (int d = 0; d < 3; ++d) loc[d] = d == q ? 1 : 0;
            want = loc;
        }
        CHMY_TRY(check_field(g, &src[q], "src", want, dt));
        // the mask component read with src[q] is the one at src[q]'s location
        if (masks) CHMY_TRY(check_field(g, &masks[q], "mask", src[q].loc, dt));
    }
    if (k) CHMY_TRY(check_field(g, k, "k", nullptr, dt));
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    OpArgs a{op, g->nd, dim, nsrc};
    hipStream_t s = (hipStream_t)stream;
    dim3 grd = grid_for(b, blk2());
    auto run = [&](auto tag) {
        using T = decltype(tag);
        DFL<T> d = lview<T>(*dst), s0 = lview<T>(src[0]);
        DFL<T> s1 = nsrc > 1 ? lview<T>(src[1]) : s0, s2 = nsrc > 2 ? lview<T>(src[2]) : s0;
        DFL<T> w0 = masks ? lview<T>(masks[0]) : s0;
        DFL<T> w1 = masks && nsrc > 1 ? lview<T>(masks[1]) : w0, w2 = masks && nsrc > 2 ? lview<T>(masks[2]) : w0;
        DFL<T> kk = k ? lview<T>(*k) : s0;
        const T i0 = (T)g->inv_spacing[0], i1 = (T)g->inv_spacing[1], i2 = (T)g->inv_spacing[2];
        const Axes3 ax = axes_of(g);
        if (nonuniform(g)) {
            if (masks)
                hipLaunchKernelGGL((op_ext_kernel<T, true, true>), grd, blk2(), 0, s, b, d, s0, s1, s2, w0, w1, w2, kk, a,
                                   i0, i1, i2, ax);
            else
                hipLaunchKernelGGL((op_ext_kernel<T, false, true>), grd, blk2(), 0, s, b, d, s0, s1, s2, w0, w1, w2, kk, a,
                                   i0, i1, i2, ax);
        } else if (masks) {
            hipLaunchKernelGGL((op_ext_kernel<T, true, false>), grd, blk2(), 0, s, b, d, s0, s1, s2, w0, w1, w2, kk, a,
                               i0, i1, i2, ax);
        } else {
            hipLaunchKernelGGL((op_ext_kernel<T, false, false>), grd, blk2(), 0, s, b, d, s0, s1, s2, w0, w1, w2, kk, a,
                               i0, i1, i2, ax);
        }
    };
    if (dt == CHMY_F64) run(double{});
    else run(float{});
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_diffusion_flux(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* q, const chmy_field_t* C,
                        double chi, void* stream) {
    CHMY_TRY(check_grid(g));
    CHMY_TRY(check_center(g, C, "C", -1));
    CHMY_TRY(check_vector(g, q, "q", C->dtype));
    return C->dtype == CHMY_F64 ? diffusion_launch<double>(g, box, q, C, chi, true, (hipStream_t)stream)
                                : diffusion_launch<float>(g, box, q, C, chi, true, (hipStream_t)stream);
}

int chmy_diffusion_update(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* C,
                          const chmy_field_t* q, double dt, void* stream) {
    CHMY_TRY(check_grid(g));
    CHMY_TRY(check_center(g, C, "C", -1));
    CHMY_TRY(check_vector(g, q, "q", C->dtype));
    return C->dtype == CHMY_F64 ? diffusion_launch<double>(g, box, q, C, dt, false, (hipStream_t)stream)
                                : diffusion_launch<float>(g, box, q, C, dt, false, (hipStream_t)stream);
}

int chmy_diffusion_step(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* C,
                        const chmy_field_t* C_out, const chmy_field_t* q, double chi, double dt, void* stream) {
    CHMY_TRY(check_grid(g));
    CHMY_TRY(check_center(g, C, "C", -1));
    CHMY_TRY(check_center(g, C_out, "C_out", C->dtype));
    CHMY_TRY(check_vector(g, q, "q", C->dtype));
    if (C->data == C_out->data) return set_error(CHMY_E_ARG, "C and C_out must be distinct buffers");
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    auto launch = [&](auto zero) -> int {
        using T = decltype(zero);
        DFL<T> c = lview<T>(*C), co = lview<T>(*C_out);
        DFL<T> q0 = lview<T>(q[0]);
        DFL<T> q1 = g->nd > 1 ? lview<T>(q[1]) : q0, q2 = g->nd > 2 ? lview<T>(q[2]) : q0;
        dim3 grd = grid_for(b, blk2());
        hipLaunchKernelGGL(diffusion_step_kernel<T>, grd, blk2(), 0, (hipStream_t)stream, b, c, co, q0, q1, q2,
                           g->nd, -T(chi), T(dt), T(g->inv_spacing[0]), T(g->inv_spacing[1]),
                           T(g->inv_spacing[2]));
        CHMY_HIP(hipGetLastError());
        return 0;
    };
    return C->dtype == CHMY_F64 ? launch(0.0) : launch(0.0f);
}

int chmy_diffusion_iterate(const chmy_grid_t* g, const chmy_field_t* C0, const chmy_field_t* C1,
                           const chmy_field_t* q, const int32_t* bc_counts, const chmy_bc_t* bcs, double chi,
                           double dt, int niter, int* parity_inout, void* stream) {
    CHMY_TRY(check_grid(g));
    if (!parity_inout || (*parity_inout != 0 && *parity_inout != 1)) return set_error(CHMY_E_ARG, "parity must be 0/1");
    if (niter < 0) return set_error(CHMY_E_ARG, "niter < 0");
    if (!bc_counts || !bcs) return set_error(CHMY_E_ARG, "C batch is NULL");
    int total = 0;
    for (int k = 0; k < 2 * g->nd; ++k) {
        if (bc_counts[k] < 0 || bc_counts[k] > 1) return set_error(CHMY_E_ARG, "one C entry per (dim, side) at most");
        total += bc_counts[k];
    }
    const chmy_field_t* Cs[2] = {C0, C1};
    chmy_field_t fset[2][6];
    for (int s = 0; s < 2; ++s)
        for (int k = 0; k < total; ++k) fset[s][k] = *Cs[s];
    chmy_box_t box;
    for (int d = 0; d < 3; ++d) {
        box.lo[d] = 0;
        box.len[d] = d < g->nd ? g->n[d] + 2 : 1;
    }
    int p = *parity_inout;
    // one scalar rule on every (dim, side): the batch rides in the step's launch
    bool fold = total == 2 * g->nd;
    SideRules rules;
    std::memset(&rules, 0, sizeof rules);
    for (int k = 0; fold && k < total; ++k) {
        const chmy_bc_t& c = bcs[k];
        fold = (c.kind == CHMY_DIRICHLET || c.kind == CHMY_NEUMANN) &&
               (c.vkind == CHMY_VALUE_ZERO || c.vkind == CHMY_VALUE_SCALAR);   // else the batch set validates
        rules.kind[k / 2][k % 2] = c.kind;
        rules.val[k / 2][k % 2] = c.vkind == CHMY_VALUE_SCALAR ? c.value : 0.0;
    }
    if (fold && niter > 0) {
        for (int d = 0; d < 3; ++d) rules.h[d] = g->spacing[d];
        CHMY_TRY(check_center(g, C0, "C0", -1));
        CHMY_TRY(check_center(g, C1, "C1", C0->dtype));
        CHMY_TRY(check_vector(g, q, "q", C0->dtype));
        if (C0->data == C1->data) return set_error(CHMY_E_ARG, "C0 and C1 must be distinct buffers");
        chmy_box_t bx;
        CHMY_TRY(check_box(g, &box, &bx));
        const Box3 b = box3(bx, g->nd);
        Box3 reg = b;
        for (int d = 0; d < g->nd; ++d) reg.len[d] = b.len[d] + 1;   // [0, n+2]
        auto run = [&](auto zero) -> int {
            using T = decltype(zero);
            DFL<T> c[2] = {lview<T>(*C0), lview<T>(*C1)};
            DFL<T> q0 = lview<T>(q[0]);
            DFL<T> q1 = g->nd > 1 ? lview<T>(q[1]) : q0, q2 = g->nd > 2 ? lview<T>(q[2]) : q0;
            const dim3 grd = grid_for(reg, blk2());
            for (int it = 0; it < niter; ++it) {
                hipLaunchKernelGGL(diffusion_step_bc_kernel<T>, grd, blk2(), 0, (hipStream_t)stream, b, reg, c[p],
                                   c[1 - p], q0, q1, q2, g->nd, -T(chi), T(dt), T(g->inv_spacing[0]),
                                   T(g->inv_spacing[1]), T(g->inv_spacing[2]), rules);
                p = 1 - p;
            }
            CHMY_HIP(hipGetLastError());
            return 0;
        };
        CHMY_TRY(C0->dtype == CHMY_F64 ? run(0.0) : run(0.0f));
        *parity_inout = p;
        return 0;
    }
    for (int it = 0; it < niter; ++it) {
        CHMY_TRY(chmy_diffusion_step(g, &box, Cs[p], Cs[1 - p], q, chi, dt, stream));
        CHMY_TRY(chmy_bc_batchset(g, bc_counts, fset[1 - p], bcs, stream));
        p = 1 - p;
    }
    *parity_inout = p;
    return 0;
}

int chmy_stokes_update_old(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* Tf,
                           const chmy_field_t* tau, const chmy_field_t* To, const chmy_field_t* tauo, void* stream) {
    CHMY_TRY(check_grid(g));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_tensor(g, tau, "tau", -1));
    const int dt = tau[0].dtype;
    CHMY_TRY(check_tensor(g, tauo, "tau_old", dt));
    if ((Tf == nullptr) != (To == nullptr)) return set_error(CHMY_E_ARG, "T and T_old must both be given or NULL");
    if (Tf) {
        CHMY_TRY(check_center(g, Tf, "T", dt));
        CHMY_TRY(check_center(g, To, "T_old", dt));
    }
    return dt == CHMY_F64 ? copy_launch<double>(g, bx, Tf, tau, To, tauo, (hipStream_t)stream)
                          : copy_launch<float>(g, bx, Tf, tau, To, tauo, (hipStream_t)stream);
}

int chmy_thermal_flux(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* qT, const chmy_field_t* Tf,
                      const chmy_field_t* V, double lam, void* stream) {
    CHMY_TRY(check_grid(g));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_center(g, Tf, "T", -1));
    CHMY_TRY(check_vector(g, qT, "qT", Tf->dtype));
    CHMY_TRY(check_vector(g, V, "V", Tf->dtype));
    return Tf->dtype == CHMY_F64 ? thermal_launch<double>(g, bx, qT, Tf, V, lam, true, (hipStream_t)stream)
                                 : thermal_launch<float>(g, bx, qT, Tf, V, lam, true, (hipStream_t)stream);
}

int chmy_thermal_update(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* Tf,
                        const chmy_field_t* To, const chmy_field_t* qT, double dt, void* stream) {
    CHMY_TRY(check_grid(g));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_center(g, Tf, "T", -1));
    CHMY_TRY(check_center(g, To, "T_old", Tf->dtype));
    CHMY_TRY(check_vector(g, qT, "qT", Tf->dtype));
    return Tf->dtype == CHMY_F64 ? thermal_launch<double>(g, bx, qT, Tf, To, dt, false, (hipStream_t)stream)
                                 : thermal_launch<float>(g, bx, qT, Tf, To, dt, false, (hipStream_t)stream);
}

int chmy_thermal_step(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* Tf, const chmy_field_t* T_out,
                      const chmy_field_t* To, const chmy_field_t* qT, const chmy_field_t* V, double lam, double dt,
                      void* stream) {
    CHMY_TRY(check_grid(g));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_center(g, Tf, "T", -1));
    CHMY_TRY(check_center(g, T_out, "T_out", Tf->dtype));
    CHMY_TRY(check_center(g, To, "T_old", Tf->dtype));
    CHMY_TRY(check_vector(g, qT, "qT", Tf->dtype));
    CHMY_TRY(check_vector(g, V, "V", Tf->dtype));
    if (Tf->data == T_out->data) return set_error(CHMY_E_ARG, "T and T_out must be distinct buffers");
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    auto launch = [&](auto zero) -> int {
        using T = decltype(zero);
        DFL<T> t = lview<T>(*Tf), to = lview<T>(*T_out), o = lview<T>(*To);
        DFL<T> v0 = lview<T>(V[0]), q0 = lview<T>(qT[0]);
        DFL<T> v1 = g->nd > 1 ? lview<T>(V[1]) : v0, v2 = g->nd > 2 ? lview<T>(V[2]) : v0;
        DFL<T> q1 = g->nd > 1 ? lview<T>(qT[1]) : q0, q2 = g->nd > 2 ? lview<T>(qT[2]) : q0;
        // 3D with one stride set per location (Center; each V_D with qT_D): the z-march
        bool march = g->nd == 3;
        if (march) {
            const DF<T> c = view<T>(*Tf), c1 = view<T>(*T_out), c2 = view<T>(*To);
            march = c1.sy == c.sy && c1.sz == c.sz && c2.sy == c.sy && c2.sz == c.sz;
            for (int d = 0; d < 3; ++d) {
                const DF<T> a = view<T>(V[d]), q = view<T>(qT[d]);
                march = march && a.sy == q.sy && a.sz == q.sz;
            }
        }
        if (march) {
            const int nbx = (int)((b.len[0] + 63) / 64), nby = (int)((b.len[1] + kThY - 1) / kThY);
            const int64_t nch = (b.len[2] + 15) / 16;
            const int kchunk = (int)((b.len[2] + nch - 1) / nch);
            const int64_t nb = (int64_t)nbx * nby * nch;
            if (nb > INT32_MAX) return set_error(CHMY_E_SIZE, "thermal step: box too large");
            hipLaunchKernelGGL(thermal_step3_kernel<T>, dim3((unsigned)nb), dim3(64, kThY, 1), 0, (hipStream_t)stream,
                               b, view<T>(*Tf), view<T>(*T_out), view<T>(*To), view<T>(V[0]), view<T>(V[1]),
                               view<T>(V[2]), view<T>(qT[0]), view<T>(qT[1]), view<T>(qT[2]), -T(lam), T(dt),
                               T(g->inv_spacing[0]), T(g->inv_spacing[1]), T(g->inv_spacing[2]), nbx, nby, kchunk);
            CHMY_HIP(hipGetLastError());
            return 0;
        }
        dim3 grd = grid_for(b, blk2());
        hipLaunchKernelGGL(thermal_step_kernel<T>, grd, blk2(), 0, (hipStream_t)stream, b, t, to, o, v0, v1, v2, q0,
                           q1, q2, g->nd, -T(lam), T(dt), T(g->inv_spacing[0]), T(g->inv_spacing[1]),
                           T(g->inv_spacing[2]));
        CHMY_HIP(hipGetLastError());
        return 0;
    };
    return Tf->dtype == CHMY_F64 ? launch(0.0) : launch(0.0f);
}

int chmy_maxloc(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* dst, const chmy_field_t* src,
                void* stream) {
    CHMY_TRY(check_grid(g));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_center(g, src, "src", -1));
    CHMY_TRY(check_center(g, dst, "dst", src->dtype));
    Box3 b = box3(bx, g->nd);
    if (no_work(b)) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (src->dtype == CHMY_F64)
        hipLaunchKernelGGL(maxloc_kernel<double>, grid_for(b, blk2()), blk2(), 0, s, b, lview<double>(*dst),
                           lview<double>(*src), g->nd);
    else
        hipLaunchKernelGGL(maxloc_kernel<float>, grid_for(b, blk2()), blk2(), 0, s, b, lview<float>(*dst),
                           lview<float>(*src), g->nd);
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
