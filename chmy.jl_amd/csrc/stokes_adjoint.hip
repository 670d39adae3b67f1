// Gradient of a functional of the converged PT Stokes state with respect to
// the Center viscosity field (build extension; the reference has no adjoint).
//
// The steady equations R_tau = -(tau - tau_old)/(G dt) - tau/eta_loc +
// 2 dev(eps(V)) = 0, R_V = -grad Pr + div tau = 0, R_P = -div V = 0 are
// linear in (V, Pr, tau); eliminating tau gives the symmetric viscous block
// K = -sum D^T (2 eta_ve) D and the pressure coupling -grad = div^T, so the
// adjoint system is the same Stokes problem with Pr -> -lambda_P.  The
// adjoint solve is therefore a forward PT solve (forcing -dJ/dV, homogeneous
// BCs, tau_old = 0) whose stresses tau_a give the multipliers of R_tau as
// -tau_a/2 (diagonal) and -tau_a (edges).  With dR_tau/deta_loc = tau/eta_loc^2:
//   g(c) = [c interior] sum_d tau_a_dd tau_dd / (2 eta_c eta_c)
//        + 0.25 sum_{active edges e lerping eta from c} tau_a_e tau_e / (eta_e eta_e)
// (edges: vertex index c or c+1 in their vertex dims, active range vertex
// 1..n+1 and centre 1..n; eta_e = lerp(eta) as in the stress update), and
// logscale multiplies by eta_c (d/dlog eta).  One gather pass over the
// centres 0..n+1 in a fixed summation order.
#include "common.hpp"

namespace chmy {

template <class T>
struct GradArgs {
    DF<T> t[6], ta[6], eta, g;
    int32_t n[3];
    int32_t logscale;
};

template <class T>
__device__ __forceinline__ T lerp_knots(const DF<T>& f, int64_t o, int64_t sa, int64_t sb) {
    const T a0 = f.at(o - sa - sb), a1 = f.at(o - sb), b0 = f.at(o - sa), b1 = f.at(o);
    const T A = fma(T(0.5), a1 - a0, a0);
    const T B = fma(T(0.5), b1 - b0, b0);
    return fma(T(0.5), B - A, A);
}

template <class T, int ND>
__global__ void __launch_bounds__(256) grad_eta_kernel(GradArgs<T> a) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i > a.n[0] + 1 || j > a.n[1] + 1) return;
    const int n[3] = {a.n[0], a.n[1], ND == 3 ? a.n[2] : 0};
    const int I[3] = {i, j, k};
    auto O = [&](const DF<T>& f, int x, int y, int z) -> int64_t { return x + y * f.sy + z * f.sz; };
    const T ec = a.eta.at(O(a.eta, i, j, k));
    bool interior = i >= 1 && i <= n[0] && j >= 1 && j <= n[1];
    if (ND == 3) interior = interior && k >= 1 && k <= n[2];
    T gd = T(0);
    if (interior) {
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const int64_t o = O(a.t[d], i, j, k);
            gd = gd + (a.ta[d].at(o) * a.t[d].at(o)) / (T(2) * ec * ec);
        }
    }
    // off-diagonal components: 2D xy; 3D xy, xz, yz (vertex dims va < vb)
    constexpr int NE = ND == 2 ? 1 : 3;
    const int VA[3] = {0, 0, 1}, VB[3] = {1, 2, 2};
    T ge = T(0);
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        const int c = ND + q;
        const int va = VA[q], vb = VB[q];
        const DF<T>& t = a.t[c];
        const DF<T>& ta = a.ta[c];
        const int64_t st[3] = {1, a.eta.sy, a.eta.sz};
        for (int bb = 0; bb < 2; ++bb)
            for (int aa = 0; aa < 2; ++aa) {
                int E[3] = {I[0], I[1], I[2]};
                E[va] += aa;
                E[vb] += bb;
                bool act = true;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const int hi = (d == va || d == vb) ? n[d] + 1 : n[d];
                    act = act && E[d] >= 1 && E[d] <= hi;
                }
                if (!act) continue;
                const T et = lerp_knots(a.eta, O(a.eta, E[0], E[1], E[2]), st[va], st[vb]);
                const int64_t o = O(t, E[0], E[1], E[2]);
                ge = ge + (ta.at(o) * t.at(o)) / (et * et);
            }
    }
    T r = gd + T(0.25) * ge;
    if (a.logscale) r = r * ec;
    a.g.at(O(a.g, i, j, k)) = r;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_stokes_grad_eta(const chmy_grid_t* g, const chmy_field_t* tau, const chmy_field_t* tau_a,
                         const chmy_field_t* eta, const chmy_field_t* grad, int logscale, void* stream) {
    CHMY_TRY(check_grid(g));
    if (g->nd < 2) return set_error(CHMY_E_ARG, "chmy_stokes_grad_eta needs a 2D or 3D grid");
    CHMY_TRY(check_tensor(g, tau, "tau", CHMY_F64));
    CHMY_TRY(check_tensor(g, tau_a, "tau_a", CHMY_F64));
    CHMY_TRY(check_center(g, eta, "eta", CHMY_F64));
    CHMY_TRY(check_center(g, grad, "grad", CHMY_F64));
    const int nt = g->nd == 2 ? 3 : 6;
    GradArgs<double> a;
    std::memset(&a, 0, sizeof a);
    for (int c = 0; c < nt; ++c) {
        a.t[c] = view<double>(tau[c]);
        a.ta[c] = view<double>(tau_a[c]);
    }
    a.eta = view<double>(*eta);
    a.g = view<double>(*grad);
    for (int d = 0; d < 3; ++d) a.n[d] = d < g->nd ? (int32_t)g->n[d] : 0;
    a.logscale = logscale != 0;
    dim3 blk(64, 4, 1);
    dim3 grd((unsigned)((g->n[0] + 2 + 63) / 64), (unsigned)((g->n[1] + 2 + 3) / 4),
             (unsigned)(g->nd == 3 ? g->n[2] + 2 : 1));
    hipStream_t s = (hipStream_t)stream;
    if (g->nd == 2) hipLaunchKernelGGL((grad_eta_kernel<double, 2>), grd, blk, 0, s, a);
    else hipLaunchKernelGGL((grad_eta_kernel<double, 3>), grd, blk, 0, s, a);
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
