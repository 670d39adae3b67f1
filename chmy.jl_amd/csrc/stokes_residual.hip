// Defect correction for the PT Stokes solve (build extension; the reference has
// no mixed-precision path).  An f64 master state x = (V, tau, Pr) is advanced
// by corrections d computed with f32 PT sweeps:
//   1. R(x) in f64: R_tau = -(tau - tau_old)/(G dt) - tau/eta + 2 dev(eps(V)),
//      R_P = -divg(V), R_V = -grad(Pr) + divg(tau) - rhog, each in the
//      reference's expression order (examples/stokes_3d_inc_ve_T.jl:23-57),
//      stored as the sweep's forcing in the sweep precision:
//      tau_old_c = G dt R_tau, f_P = R_P, f_V = R_V;
//   2. PT sweeps on d from 0 (chmy_stokes_iter with f_V/f_P, homogeneous BCs);
//   3. x += d (chmy_stokes_accumulate), then the real V batch on x.
// Because the PT iteration is stationary in (V, tau, Pr), folding d into x and
// restarting d at 0 continues the same iteration; the f64 residual sets the
// attainable accuracy.
#include "common.hpp"

namespace chmy {

constexpr int kRX = 64, kRY = 4;

template <class T, class TO>
struct ResArgs {
    DF<T> V[3], tau[6], told[6], Pr, eta, rg, fe[3];
    DF<TO> toc[6], fV[3], fP;
    DF<T> dV, rV[3];
    T ix, iy, iz, eta_s, eve_s_unused, Gs, dts, Gdt;
    int has_eta, has_rg, has_norm, has_fe;
};

template <class T>
__device__ __forceinline__ T lerp_e(const DF<T>& f, int64_t o, int64_t sa, int64_t sb) {
    T a0 = f.at(o - sa - sb), a1 = f.at(o - sb);
    T b0 = f.at(o - sa), b1 = f.at(o);
    T A = fma(T(0.5), a1 - a0, a0);
    T B = fma(T(0.5), b1 - b0, b0);
    return fma(T(0.5), B - A, A);
}

template <class T>
__device__ __forceinline__ T r_tau_diag(T t, T to, T e, T dV, T eta, T Gdt) {
    return ((-(t - to)) / Gdt - t / eta) + T(2) * (e - dV / T(3));
}
template <class T>
__device__ __forceinline__ T r_tau_off(T t, T to, T e, T eta, T Gdt) {
    return ((-(t - to)) / Gdt - t / eta) + T(2) * e;
}

template <class T, class TO, int ND>
__global__ void __launch_bounds__(256) residual_kernel(Box3 b, ResArgs<T, TO> a) {
    const int64_t i = b.lo[0] + blockIdx.x * (int64_t)kRX + threadIdx.x;
    const int64_t j = b.lo[1] + blockIdx.y * (int64_t)kRY + threadIdx.y;
    const int64_t k = b.lo[2] + blockIdx.z;
    if (i >= b.lo[0] + b.len[0] || j >= b.lo[1] + b.len[1]) return;
    auto O = [i, j, k](const auto& f) { return i + j * f.sy + k * f.sz; };
    const DF<T>& Vx = a.V[0];
    const DF<T>& Vy = a.V[1];
    const int64_t ox = O(Vx), oy = O(Vy), oc = O(a.Pr);
    T e[6], dv;
    if constexpr (ND == 2) {
        e[0] = (Vx.at(ox + 1) - Vx.at(ox)) * a.ix;
        e[1] = (Vy.at(oy + Vy.sy) - Vy.at(oy)) * a.iy;
        e[2] = T(0.5) * ((Vx.at(ox) - Vx.at(ox - Vx.sy)) * a.iy + (Vy.at(oy) - Vy.at(oy - 1)) * a.ix);
        dv = e[0] + e[1];
    } else {
        const DF<T>& Vz = a.V[2];
        const int64_t oz = O(Vz);
        e[0] = (Vx.at(ox + 1) - Vx.at(ox)) * a.ix;
        e[1] = (Vy.at(oy + Vy.sy) - Vy.at(oy)) * a.iy;
        e[2] = (Vz.at(oz + Vz.sz) - Vz.at(oz)) * a.iz;
        e[3] = T(0.5) * ((Vx.at(ox) - Vx.at(ox - Vx.sy)) * a.iy + (Vy.at(oy) - Vy.at(oy - 1)) * a.ix);
        e[4] = T(0.5) * ((Vx.at(ox) - Vx.at(ox - Vx.sz)) * a.iz + (Vz.at(oz) - Vz.at(oz - 1)) * a.ix);
        e[5] = T(0.5) * ((Vy.at(oy) - Vy.at(oy - Vy.sz)) * a.iz + (Vz.at(oz) - Vz.at(oz - Vz.sy)) * a.iy);
        dv = (e[0] + e[1]) + e[2];
    }
    const int nt = ND == 2 ? 3 : 6;
    const int64_t oe = O(a.eta);
    for (int c = 0; c < nt; ++c) {
        T eta = a.eta_s;
        const bool diag = c < ND;
        if (a.has_eta) {
            if (diag) eta = a.eta.at(oe);
            else if (ND == 2) eta = lerp_e(a.eta, oe, 1, a.eta.sy);
            else if (c == 3) eta = lerp_e(a.eta, oe, 1, a.eta.sy);
            else if (c == 4) eta = lerp_e(a.eta, oe, 1, a.eta.sz);
            else eta = lerp_e(a.eta, oe, a.eta.sy, a.eta.sz);
        }
        const T t = a.tau[c].at(O(a.tau[c])), to = a.told[c].at(O(a.told[c]));
        const T r = diag ? r_tau_diag(t, to, e[c], dv, eta, a.Gdt) : r_tau_off(t, to, e[c], eta, a.Gdt);
        put(a.toc[c].at(O(a.toc[c])), a.Gdt * r);
    }
    put(a.fP.at(O(a.fP)), -dv);
    // R_V (update_velocity!'s r_V, stokes_3d_inc_ve_T.jl:51-53 / 2D stokes_2d_inc_ve_T.jl:39-40)
    const T P = a.Pr.at(oc);
    T r[3];
    if constexpr (ND == 2) {
        const DF<T>&txx = a.tau[0], &tyy = a.tau[1], &txy = a.tau[2];
        const int64_t o0 = O(txx), o1 = O(tyy), o2 = O(txy);
        r[0] = (-((P - a.Pr.at(oc - 1)) * a.ix) + (txx.at(o0) - txx.at(o0 - 1)) * a.ix) +
               (txy.at(o2 + txy.sy) - txy.at(o2)) * a.iy;
        r[1] = (-((P - a.Pr.at(oc - a.Pr.sy)) * a.iy) + (tyy.at(o1) - tyy.at(o1 - tyy.sy)) * a.iy) +
               (txy.at(o2 + 1) - txy.at(o2)) * a.ix;
        if (a.has_rg) r[1] = r[1] - a.rg.at(O(a.rg));
    } else {
        const DF<T>&txx = a.tau[0], &tyy = a.tau[1], &tzz = a.tau[2];
        const DF<T>&txy = a.tau[3], &txz = a.tau[4], &tyz = a.tau[5];
        const int64_t o0 = O(txx), o1 = O(tyy), o2 = O(tzz), o3 = O(txy), o4 = O(txz), o5 = O(tyz);
        r[0] = ((-((P - a.Pr.at(oc - 1)) * a.ix) + (txx.at(o0) - txx.at(o0 - 1)) * a.ix) +
                (txy.at(o3 + txy.sy) - txy.at(o3)) * a.iy) + (txz.at(o4 + txz.sz) - txz.at(o4)) * a.iz;
        r[1] = ((-((P - a.Pr.at(oc - a.Pr.sy)) * a.iy) + (tyy.at(o1) - tyy.at(o1 - tyy.sy)) * a.iy) +
                (txy.at(o3 + 1) - txy.at(o3)) * a.ix) + (tyz.at(o5 + tyz.sz) - tyz.at(o5)) * a.iz;
        r[2] = ((-((P - a.Pr.at(oc - a.Pr.sz)) * a.iz) + (tzz.at(o2) - tzz.at(o2 - tzz.sz)) * a.iz) +
                (txz.at(o4 + 1) - txz.at(o4)) * a.ix) + (tyz.at(o5 + tyz.sy) - tyz.at(o5)) * a.iy;
        if (a.has_rg) r[2] = r[2] - a.rg.at(O(a.rg));
    }
    if (a.has_fe)
        for (int d = 0; d < ND; ++d) r[d] = r[d] + a.fe[d].at(O(a.fe[d]));
    for (int d = 0; d < ND; ++d) {
        put(a.fV[d].at(O(a.fV[d])), r[d]);
        if (a.has_norm) a.rV[d].at(O(a.rV[d])) = r[d];
    }
    if (a.has_norm) a.dV.at(O(a.dV)) = dv;
}

template <class T, class TO>
struct AccArgs {
    DF<T> x[10];
    DF<TO> d[10];
    int n;
};

template <class T, class TO>
__global__ void __launch_bounds__(256) accumulate_kernel(Box3 b, AccArgs<T, TO> a) {
    const int64_t i = b.lo[0] + blockIdx.x * (int64_t)kRX + threadIdx.x;
    const int64_t j = b.lo[1] + blockIdx.y * (int64_t)kRY + threadIdx.y;
    const int64_t k = b.lo[2] + blockIdx.z;
    if (i >= b.lo[0] + b.len[0] || j >= b.lo[1] + b.len[1]) return;
    for (int q = 0; q < a.n; ++q) {
        const int64_t ox = i + j * a.x[q].sy + k * a.x[q].sz;
        const int64_t od = i + j * a.d[q].sy + k * a.d[q].sz;
        a.x[q].at(ox) = a.x[q].at(ox) + T(a.d[q].at(od));
    }
}

template <class T, class TO>
static int residual_launch(const chmy_grid_t* g, const chmy_box_t& bx, const chmy_field_t* V,
                           const chmy_field_t* tau, const chmy_field_t* Pr, const chmy_field_t* told,
                           const chmy_field_t* eta, const chmy_field_t* rg, const chmy_field_t* fe,
                           double eta_s, double G, double dt, const chmy_field_t* toc, const chmy_field_t* fV, const chmy_field_t* fP,
                           const chmy_field_t* dV, const chmy_field_t* rV, hipStream_t s) {
    ResArgs<T, TO> a;
    std::memset(&a, 0, sizeof a);
    const int nd = g->nd, nt = nd == 2 ? 3 : 6;
    for (int d = 0; d < nd; ++d) {
        a.V[d] = view<T>(V[d]);
        a.fV[d] = view<TO>(fV[d]);
        a.rV[d] = rV ? view<T>(rV[d]) : a.V[d];
        a.fe[d] = fe ? view<T>(fe[d]) : a.V[d];
    }
    a.has_fe = fe != nullptr;
    for (int c = 0; c < nt; ++c) {
        a.tau[c] = view<T>(tau[c]);
        a.told[c] = view<T>(told[c]);
        a.toc[c] = view<TO>(toc[c]);
    }
    a.Pr = view<T>(*Pr);
    a.eta = eta ? view<T>(*eta) : a.Pr;
    a.rg = rg ? view<T>(*rg) : a.Pr;
    a.fP = view<TO>(*fP);
    a.dV = dV ? view<T>(*dV) : a.Pr;
    a.has_eta = eta != nullptr;
    a.has_rg = rg != nullptr;
    a.has_norm = dV != nullptr;
    a.ix = T(g->inv_spacing[0]);
    a.iy = T(g->inv_spacing[1]);
    a.iz = T(g->inv_spacing[2]);
    a.eta_s = T(eta_s);
    a.Gs = T(G);
    a.dts = T(dt);
    a.Gdt = T(G) * T(dt);
    Box3 b = box3(bx, nd);
    if (no_work(b)) return 0;
    dim3 blk(kRX, kRY, 1), grd((unsigned)((b.len[0] + kRX - 1) / kRX), (unsigned)((b.len[1] + kRY - 1) / kRY),
                               (unsigned)b.len[2]);
    if (nd == 2) hipLaunchKernelGGL((residual_kernel<T, TO, 2>), grd, blk, 0, s, b, a);
    else hipLaunchKernelGGL((residual_kernel<T, TO, 3>), grd, blk, 0, s, b, a);
    CHMY_HIP(hipGetLastError());
    return 0;
}

template <class T, class TO>
static int accumulate_launch(const chmy_grid_t* g, const chmy_box_t& bx, const chmy_field_t* V,
                             const chmy_field_t* tau, const chmy_field_t* Pr, const chmy_field_t* Vc,
                             const chmy_field_t* tauc, const chmy_field_t* Prc, hipStream_t s) {
    AccArgs<T, TO> a;
    std::memset(&a, 0, sizeof a);
    const int nd = g->nd, nt = nd == 2 ? 3 : 6;
    a.n = 0;
    for (int d = 0; d < nd; ++d) { a.x[a.n] = view<T>(V[d]); a.d[a.n] = view<TO>(Vc[d]); ++a.n; }
    for (int c = 0; c < nt; ++c) { a.x[a.n] = view<T>(tau[c]); a.d[a.n] = view<TO>(tauc[c]); ++a.n; }
    a.x[a.n] = view<T>(*Pr);
    a.d[a.n] = view<TO>(*Prc);
    ++a.n;
    Box3 b = box3(bx, nd);
    if (no_work(b)) return 0;
    dim3 blk(kRX, kRY, 1), grd((unsigned)((b.len[0] + kRX - 1) / kRX), (unsigned)((b.len[1] + kRY - 1) / kRY),
                               (unsigned)b.len[2]);
    hipLaunchKernelGGL((accumulate_kernel<T, TO>), grd, blk, 0, s, b, a);
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_stokes_residual(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau, const chmy_field_t* Pr,
                         const chmy_field_t* tau_old, const chmy_field_t* eta, const chmy_field_t* rhog,
                         const chmy_field_t* f_ext, double eta_s, double G, double dt,
                         const chmy_field_t* tau_old_c, const chmy_field_t* f_V,
                         const chmy_field_t* f_P, const chmy_field_t* divV, const chmy_field_t* r_V, void* stream) {
    CHMY_TRY(check_grid(g));
    if (g->nd < 2) return set_error(CHMY_E_ARG, "Stokes kernels need nd >= 2");
    CHMY_TRY(check_center(g, Pr, "Pr", CHMY_F64));
    CHMY_TRY(check_vector(g, V, "V", CHMY_F64));
    CHMY_TRY(check_tensor(g, tau, "tau", CHMY_F64));
    CHMY_TRY(check_tensor(g, tau_old, "tau_old", CHMY_F64));
    if (eta) CHMY_TRY(check_center(g, eta, "eta", CHMY_F64));
    if (rhog) {
        int loc[3] = {0, 0, 0};
        loc[g->nd - 1] = 1;
        CHMY_TRY(check_field(g, rhog, "rhog", loc, CHMY_F64));
    }
    if (f_ext) CHMY_TRY(check_vector(g, f_ext, "f_ext", CHMY_F64));
    CHMY_TRY(check_center(g, f_P, "f_P", -1, true));
    const int to = f_P->dtype;
    CHMY_TRY(check_vector(g, f_V, "f_V", to, true));
    CHMY_TRY(check_tensor(g, tau_old_c, "tau_old_c", to, true));
    if ((divV == nullptr) != (r_V == nullptr)) return set_error(CHMY_E_ARG, "divV and r_V go together");
    if (divV) {
        CHMY_TRY(check_center(g, divV, "divV", CHMY_F64));
        CHMY_TRY(check_vector(g, r_V, "r_V", CHMY_F64));
    }
    chmy_box_t bx;
    CHMY_TRY(check_box(g, nullptr, &bx));
    hipStream_t s = (hipStream_t)stream;
    if (to == CHMY_F32)
        return residual_launch<double, float>(g, bx, V, tau, Pr, tau_old, eta, rhog, f_ext, eta_s, G, dt, tau_old_c,
                                              f_V, f_P, divV, r_V, s);
    if (to == CHMY_BF16)
        return residual_launch<double, bf16s>(g, bx, V, tau, Pr, tau_old, eta, rhog, f_ext, eta_s, G, dt, tau_old_c,
                                              f_V, f_P, divV, r_V, s);
    return residual_launch<double, double>(g, bx, V, tau, Pr, tau_old, eta, rhog, f_ext, eta_s, G, dt, tau_old_c,
                                           f_V, f_P, divV, r_V, s);
}

int chmy_stokes_accumulate(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau,
                           const chmy_field_t* Pr, const chmy_field_t* V_c, const chmy_field_t* tau_c,
                           const chmy_field_t* Pr_c, void* stream) {
    CHMY_TRY(check_grid(g));
    if (g->nd < 2) return set_error(CHMY_E_ARG, "Stokes kernels need nd >= 2");
    CHMY_TRY(check_center(g, Pr, "Pr", CHMY_F64));
    CHMY_TRY(check_vector(g, V, "V", CHMY_F64));
    CHMY_TRY(check_tensor(g, tau, "tau", CHMY_F64));
    CHMY_TRY(check_center(g, Pr_c, "Pr_c", -1));
    const int to = Pr_c->dtype;
    CHMY_TRY(check_vector(g, V_c, "V_c", to));
    CHMY_TRY(check_tensor(g, tau_c, "tau_c", to));
    chmy_box_t bx;
    CHMY_TRY(check_box(g, nullptr, &bx));
    hipStream_t s = (hipStream_t)stream;
    return to == CHMY_F32 ? accumulate_launch<double, float>(g, bx, V, tau, Pr, V_c, tau_c, Pr_c, s)
                          : accumulate_launch<double, double>(g, bx, V, tau, Pr, V_c, tau_c, Pr_c, s);
}

}  // extern "C"
