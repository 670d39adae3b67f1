// Pseudo-transient viscoelastic Stokes kernels: update_stress! and
// update_velocity! (examples/stokes_3d_inc_ve_T.jl:23-57; 2D
// examples/stokes_2d_inc_ve_T.jl:20-43) as per-point Launcher kernels.
// Expressions keep the reference's association order; the library is built
// with -ffp-contract=off so f64 results equal the CPU semantics bit for bit.
#include "common.hpp"

namespace chmy {

constexpr int kSX = 64, kSY = 4;

template <class T>
struct StressArgs {
    DF<T> V[3], tau[6], tauo[6];
    DF<T> Pr, dV, eta;
    T ix, iy, iz;
    T eta_s, eve_s, Gs, dts, Gdt, dPr, dr;
};

template <class T>
struct VelArgs {
    DF<T> V[3], rV[3], tau[6];
    DF<T> Pr, rg, eml;
    T ix, iy, iz;
    T eve_s, Gs, dts, nu;
    int has_rg;
};

// lerp of a Center field to an edge with two Vertex dims a<b (interpolation.jl:14-94)
template <class T>
__device__ __forceinline__ T lerp_cc_vv(const DF<T>& f, int64_t o, int64_t sa, int64_t sb) {
    T a0 = f.at(o - sa - sb), a1 = f.at(o - sb);
    T b0 = f.at(o - sa), b1 = f.at(o);
    T A = fma(T(0.5), a1 - a0, a0);
    T B = fma(T(0.5), b1 - b0, b0);
    return fma(T(0.5), B - A, A);
}

template <class T>
__device__ __forceinline__ T tau_new_diag(T t, T to, T e, T dV, T eta, T eve, T Gdt, T dr) {
    T r = ((-(t - to)) / Gdt - t / eta) + T(2) * (e - dV / T(3));
    return t + (r * eve) * dr;
}
template <class T>
__device__ __forceinline__ T tau_new_off(T t, T to, T e, T eta, T eve, T Gdt, T dr) {
    T r = ((-(t - to)) / Gdt - t / eta) + T(2) * e;
    return t + (r * eve) * dr;
}

template <class T, int ND, bool FVISC>
__global__ void __launch_bounds__(256) stress_kernel(Box3 b, StressArgs<T> a) {
    const int64_t i = b.lo[0] + blockIdx.x * (int64_t)kSX + threadIdx.x;
    const int64_t j = b.lo[1] + blockIdx.y * (int64_t)kSY + threadIdx.y;
    const int64_t k = b.lo[2] + blockIdx.z;
    if (i >= b.lo[0] + b.len[0] || j >= b.lo[1] + b.len[1]) return;
    // every field shares the padded geometry along y/z except the staggered dim,
    // so offsets are computed per field
    auto O = [&](const DF<T>& f) { return i + j * f.sy + k * f.sz; };
    const DF<T>& Vx = a.V[0];
    const DF<T>& Vy = a.V[1];
    const int64_t ox = O(Vx), oy = O(Vy);
    if constexpr (ND == 2) {
        T exx = (Vx.at(ox + 1) - Vx.at(ox)) * a.ix;
        T eyy = (Vy.at(oy + Vy.sy) - Vy.at(oy)) * a.iy;
        T exy = T(0.5) * ((Vx.at(ox) - Vx.at(ox - Vx.sy)) * a.iy + (Vy.at(oy) - Vy.at(oy - 1)) * a.ix);
        T dv = exx + eyy;
        const int64_t oc = O(a.Pr);
        a.dV.at(O(a.dV)) = dv;
        T eta_c = a.eta_s, eve_c = a.eve_s, eta_e = a.eta_s, eve_e = a.eve_s;
        if constexpr (FVISC) {
            eta_c = a.eta.at(O(a.eta));
            eve_c = visc_ve(eta_c, a.Gs, a.dts);
            eta_e = lerp_cc_vv(a.eta, O(a.eta), 1, a.eta.sy);
            eve_e = visc_ve(eta_e, a.Gs, a.dts);
        }
        a.Pr.at(oc) = a.Pr.at(oc) - (dv * eve_c) * a.dPr;
        const int64_t o0 = O(a.tau[0]), o1 = O(a.tau[1]), o2 = O(a.tau[2]);
        a.tau[0].at(o0) = tau_new_diag(a.tau[0].at(o0), a.tauo[0].at(O(a.tauo[0])), exx, dv, eta_c, eve_c, a.Gdt, a.dr);
        a.tau[1].at(o1) = tau_new_diag(a.tau[1].at(o1), a.tauo[1].at(O(a.tauo[1])), eyy, dv, eta_c, eve_c, a.Gdt, a.dr);
        a.tau[2].at(o2) = tau_new_off(a.tau[2].at(o2), a.tauo[2].at(O(a.tauo[2])), exy, eta_e, eve_e, a.Gdt, a.dr);
    } else {
        const DF<T>& Vz = a.V[2];
        const int64_t oz = O(Vz);
        T exx = (Vx.at(ox + 1) - Vx.at(ox)) * a.ix;
        T eyy = (Vy.at(oy + Vy.sy) - Vy.at(oy)) * a.iy;
        T ezz = (Vz.at(oz + Vz.sz) - Vz.at(oz)) * a.iz;
        T exy = T(0.5) * ((Vx.at(ox) - Vx.at(ox - Vx.sy)) * a.iy + (Vy.at(oy) - Vy.at(oy - 1)) * a.ix);
        T exz = T(0.5) * ((Vx.at(ox) - Vx.at(ox - Vx.sz)) * a.iz + (Vz.at(oz) - Vz.at(oz - 1)) * a.ix);
        T eyz = T(0.5) * ((Vy.at(oy) - Vy.at(oy - Vy.sz)) * a.iz + (Vz.at(oz) - Vz.at(oz - Vz.sy)) * a.iy);
        T dv = (exx + eyy) + ezz;
        const int64_t oc = O(a.Pr);
        a.dV.at(O(a.dV)) = dv;
        T eta_c = a.eta_s, eve_c = a.eve_s;
        T exy_eta = a.eta_s, exz_eta = a.eta_s, eyz_eta = a.eta_s;
        T exy_eve = a.eve_s, exz_eve = a.eve_s, eyz_eve = a.eve_s;
        if constexpr (FVISC) {
            const int64_t oe = O(a.eta);
            eta_c = a.eta.at(oe);
            eve_c = visc_ve(eta_c, a.Gs, a.dts);
            exy_eta = lerp_cc_vv(a.eta, oe, 1, a.eta.sy);
            exz_eta = lerp_cc_vv(a.eta, oe, 1, a.eta.sz);
            eyz_eta = lerp_cc_vv(a.eta, oe, a.eta.sy, a.eta.sz);
            exy_eve = visc_ve(exy_eta, a.Gs, a.dts);
            exz_eve = visc_ve(exz_eta, a.Gs, a.dts);
            eyz_eve = visc_ve(eyz_eta, a.Gs, a.dts);
        }
        a.Pr.at(oc) = a.Pr.at(oc) - (dv * eve_c) * a.dPr;
        const T e[6] = {exx, eyy, ezz, exy, exz, eyz};
        const T et[6] = {eta_c, eta_c, eta_c, exy_eta, exz_eta, eyz_eta};
        const T ev[6] = {eve_c, eve_c, eve_c, exy_eve, exz_eve, eyz_eve};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t o = O(a.tau[c]);
            a.tau[c].at(o) = tau_new_diag(a.tau[c].at(o), a.tauo[c].at(O(a.tauo[c])), e[c], dv, et[c], ev[c], a.Gdt, a.dr);
        }
#pragma unroll
        for (int c = 3; c < 6; ++c) {
            const int64_t o = O(a.tau[c]);
            a.tau[c].at(o) = tau_new_off(a.tau[c].at(o), a.tauo[c].at(O(a.tauo[c])), e[c], et[c], ev[c], a.Gdt, a.dr);
        }
    }
}

template <class T, bool FVISC>
__device__ __forceinline__ T vel_eve(const VelArgs<T>& a, int64_t o, int64_t s) {
    if constexpr (FVISC) return visc_ve(jmax(a.eml.at(o - s), a.eml.at(o)), a.Gs, a.dts);
    else return a.eve_s;
}

template <class T, int ND, bool FVISC>
__global__ void __launch_bounds__(256) velocity_kernel(Box3 b, VelArgs<T> a) {
    const int64_t i = b.lo[0] + blockIdx.x * (int64_t)kSX + threadIdx.x;
    const int64_t j = b.lo[1] + blockIdx.y * (int64_t)kSY + threadIdx.y;
    const int64_t k = b.lo[2] + blockIdx.z;
    if (i >= b.lo[0] + b.len[0] || j >= b.lo[1] + b.len[1]) return;
    auto O = [&](const DF<T>& f) { return i + j * f.sy + k * f.sz; };
    const int64_t op = O(a.Pr);
    const T P = a.Pr.at(op);
    const int64_t oe = O(a.eml);
    if constexpr (ND == 2) {
        const DF<T>&txx = a.tau[0], &tyy = a.tau[1], &txy = a.tau[2];
        const int64_t o0 = O(txx), o1 = O(tyy), o2 = O(txy);
        T rx = (-((P - a.Pr.at(op - 1)) * a.ix) + (txx.at(o0) - txx.at(o0 - 1)) * a.ix) +
               (txy.at(o2 + txy.sy) - txy.at(o2)) * a.iy;
        T ry = (-((P - a.Pr.at(op - a.Pr.sy)) * a.iy) + (tyy.at(o1) - tyy.at(o1 - tyy.sy)) * a.iy) +
               (txy.at(o2 + 1) - txy.at(o2)) * a.ix;
        if (a.has_rg) ry = ry - a.rg.at(O(a.rg));
        const int64_t vx = O(a.V[0]), vy = O(a.V[1]);
        a.rV[0].at(O(a.rV[0])) = rx;
        a.rV[1].at(O(a.rV[1])) = ry;
        a.V[0].at(vx) = a.V[0].at(vx) + (rx * a.nu) / vel_eve<T, FVISC>(a, oe, 1);
        a.V[1].at(vy) = a.V[1].at(vy) + (ry * a.nu) / vel_eve<T, FVISC>(a, oe, a.eml.sy);
    } else {
        const DF<T>&txx = a.tau[0], &tyy = a.tau[1], &tzz = a.tau[2];
        const DF<T>&txy = a.tau[3], &txz = a.tau[4], &tyz = a.tau[5];
        const int64_t o0 = O(txx), o1 = O(tyy), o2 = O(tzz), o3 = O(txy), o4 = O(txz), o5 = O(tyz);
        T rx = ((-((P - a.Pr.at(op - 1)) * a.ix) + (txx.at(o0) - txx.at(o0 - 1)) * a.ix) +
                (txy.at(o3 + txy.sy) - txy.at(o3)) * a.iy) +
               (txz.at(o4 + txz.sz) - txz.at(o4)) * a.iz;
        T ry = ((-((P - a.Pr.at(op - a.Pr.sy)) * a.iy) + (tyy.at(o1) - tyy.at(o1 - tyy.sy)) * a.iy) +
                (txy.at(o3 + 1) - txy.at(o3)) * a.ix) +
               (tyz.at(o5 + tyz.sz) - tyz.at(o5)) * a.iz;
        T rz = ((-((P - a.Pr.at(op - a.Pr.sz)) * a.iz) + (tzz.at(o2) - tzz.at(o2 - tzz.sz)) * a.iz) +
                (txz.at(o4 + 1) - txz.at(o4)) * a.ix) +
               (tyz.at(o5 + tyz.sy) - tyz.at(o5)) * a.iy;
        if (a.has_rg) rz = rz - a.rg.at(O(a.rg));
        const int64_t vx = O(a.V[0]), vy = O(a.V[1]), vz = O(a.V[2]);
        a.rV[0].at(O(a.rV[0])) = rx;
        a.rV[1].at(O(a.rV[1])) = ry;
        a.rV[2].at(O(a.rV[2])) = rz;
        a.V[0].at(vx) = a.V[0].at(vx) + (rx * a.nu) / vel_eve<T, FVISC>(a, oe, 1);
        a.V[1].at(vy) = a.V[1].at(vy) + (ry * a.nu) / vel_eve<T, FVISC>(a, oe, a.eml.sy);
        a.V[2].at(vz) = a.V[2].at(vz) + (rz * a.nu) / vel_eve<T, FVISC>(a, oe, a.eml.sz);
    }
}

template <class T>
static int stress_launch(const chmy_grid_t* g, const chmy_box_t& bx, const chmy_field_t* tau, const chmy_field_t* Pr,
                         const chmy_field_t* dV, const chmy_field_t* V, const chmy_field_t* tauo,
                         const chmy_field_t* eta, double eta_s, double eve_s, double G, double dt, double dPr,
                         double dr, hipStream_t s) {
    StressArgs<T> a;
    std::memset(&a, 0, sizeof a);
    const int nd = g->nd, nt = nd == 2 ? 3 : 6;
    for (int d = 0; d < nd; ++d) a.V[d] = view<T>(V[d]);
    for (int c = 0; c < nt; ++c) { a.tau[c] = view<T>(tau[c]); a.tauo[c] = view<T>(tauo[c]); }
    a.Pr = view<T>(*Pr);
    a.dV = view<T>(*dV);
    a.eta = eta ? view<T>(*eta) : a.Pr;
    a.ix = T(g->inv_spacing[0]);
    a.iy = T(g->inv_spacing[1]);
    a.iz = T(g->inv_spacing[2]);
    a.eta_s = T(eta_s);
    a.eve_s = T(eve_s);
    a.Gs = T(G);
    a.dts = T(dt);
    a.Gdt = T(G) * T(dt);
    a.dPr = T(dPr);
    a.dr = T(dr);
    Box3 b = box3(bx, nd);
    if (no_work(b)) return 0;
    dim3 blk(kSX, kSY, 1), grd((unsigned)((b.len[0] + kSX - 1) / kSX), (unsigned)((b.len[1] + kSY - 1) / kSY),
                               (unsigned)b.len[2]);
    if (nd == 2) {
        if (eta) hipLaunchKernelGGL((stress_kernel<T, 2, true>), grd, blk, 0, s, b, a);
        else hipLaunchKernelGGL((stress_kernel<T, 2, false>), grd, blk, 0, s, b, a);
    } else {
        if (eta) hipLaunchKernelGGL((stress_kernel<T, 3, true>), grd, blk, 0, s, b, a);
        else hipLaunchKernelGGL((stress_kernel<T, 3, false>), grd, blk, 0, s, b, a);
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

template <class T>
static int velocity_launch(const chmy_grid_t* g, const chmy_box_t& bx, const chmy_field_t* V, const chmy_field_t* rV,
                           const chmy_field_t* Pr, const chmy_field_t* tau, const chmy_field_t* rg,
                           const chmy_field_t* eml, double eve_s, double G, double dt, double nu, hipStream_t s) {
    VelArgs<T> a;
    std::memset(&a, 0, sizeof a);
    const int nd = g->nd, nt = nd == 2 ? 3 : 6;
    for (int d = 0; d < nd; ++d) { a.V[d] = view<T>(V[d]); a.rV[d] = view<T>(rV[d]); }
    for (int c = 0; c < nt; ++c) a.tau[c] = view<T>(tau[c]);
    a.Pr = view<T>(*Pr);
    a.rg = rg ? view<T>(*rg) : a.Pr;
    a.has_rg = rg != nullptr;
    a.eml = eml ? view<T>(*eml) : a.Pr;
    a.ix = T(g->inv_spacing[0]);
    a.iy = T(g->inv_spacing[1]);
    a.iz = T(g->inv_spacing[2]);
    a.eve_s = T(eve_s);
    a.Gs = T(G);
    a.dts = T(dt);
    a.nu = T(nu);
    Box3 b = box3(bx, nd);
    if (no_work(b)) return 0;
    dim3 blk(kSX, kSY, 1), grd((unsigned)((b.len[0] + kSX - 1) / kSX), (unsigned)((b.len[1] + kSY - 1) / kSY),
                               (unsigned)b.len[2]);
    if (nd == 2) {
        if (eml) hipLaunchKernelGGL((velocity_kernel<T, 2, true>), grd, blk, 0, s, b, a);
        else hipLaunchKernelGGL((velocity_kernel<T, 2, false>), grd, blk, 0, s, b, a);
    } else {
        if (eml) hipLaunchKernelGGL((velocity_kernel<T, 3, true>), grd, blk, 0, s, b, a);
        else hipLaunchKernelGGL((velocity_kernel<T, 3, false>), grd, blk, 0, s, b, a);
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_stokes_stress(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* tau, const chmy_field_t* Pr,
                       const chmy_field_t* dV, const chmy_field_t* V, const chmy_field_t* tauo,
                       const chmy_field_t* eta, double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr,
                       double dtau_r, void* stream) {
    CHMY_TRY(check_grid(g));
    if (g->nd < 2) return set_error(CHMY_E_ARG, "Stokes kernels need nd >= 2");
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_center(g, Pr, "Pr", -1));
    const int dt_ = Pr->dtype;
    CHMY_TRY(check_center(g, dV, "divV", dt_));
    CHMY_TRY(check_vector(g, V, "V", dt_));
    CHMY_TRY(check_tensor(g, tau, "tau", dt_));
    CHMY_TRY(check_tensor(g, tauo, "tau_old", dt_));
    if (eta) CHMY_TRY(check_center(g, eta, "eta", dt_));
    return dt_ == CHMY_F64
               ? stress_launch<double>(g, bx, tau, Pr, dV, V, tauo, eta, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r,
                                       (hipStream_t)stream)
               : stress_launch<float>(g, bx, tau, Pr, dV, V, tauo, eta, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r,
                                      (hipStream_t)stream);
}

int chmy_stokes_velocity(const chmy_grid_t* g, const chmy_box_t* box, const chmy_field_t* V, const chmy_field_t* rV,
                         const chmy_field_t* Pr, const chmy_field_t* tau, const chmy_field_t* rhog,
                         const chmy_field_t* eta_ml, double eta_ve_s, double G, double dt, double nudtau,
                         void* stream) {
    CHMY_TRY(check_grid(g));
    if (g->nd < 2) return set_error(CHMY_E_ARG, "Stokes kernels need nd >= 2");
    chmy_box_t bx;
    CHMY_TRY(check_box(g, box, &bx));
    CHMY_TRY(check_center(g, Pr, "Pr", -1));
    const int dt_ = Pr->dtype;
    CHMY_TRY(check_vector(g, V, "V", dt_));
    CHMY_TRY(check_vector(g, rV, "r_V", dt_));
    CHMY_TRY(check_tensor(g, tau, "tau", dt_));
    if (rhog) {
        int loc[3] = {0, 0, 0};
        loc[g->nd - 1] = 1;
        CHMY_TRY(check_field(g, rhog, "rhog", loc, dt_));
    }
    if (eta_ml) CHMY_TRY(check_center(g, eta_ml, "eta_ml", dt_));
    return dt_ == CHMY_F64 ? velocity_launch<double>(g, bx, V, rV, Pr, tau, rhog, eta_ml, eta_ve_s, G, dt, nudtau,
                                                     (hipStream_t)stream)
                           : velocity_launch<float>(g, bx, V, rV, Pr, tau, rhog, eta_ml, eta_ve_s, G, dt, nudtau,
                                                    (hipStream_t)stream);
}

}  // extern "C"
