// One pseudo-transient Stokes iteration in a single pass over HBM:
// update_stress! followed by update_velocity! (examples/stokes_3d_inc_ve_T.jl:
// 222-223), fused and double-buffered.
//
// Layout of the work (3D): a block is 64 (x, one wave) x BY (y) threads; thread
// (tx, ty) owns the logical column (i, j) = (i0-1+tx, j0-1+ty) and marches z
// over a chunk [kb, ke).  The tile's outputs are tx in [1, 62], ty in
// [1, BY-2]; the one-cell ring around them recomputes the stresses the
// velocity stencil needs (Center values at I-e, edge values at I+e), so no
// stress is ever read back from HBM after being advanced.  At z-step k a
// thread computes the Center stresses and Pr at plane k, tau_xy at plane k
// and tau_xz, tau_yz at z-vertex plane k+1, then V at plane k.  x/y
// neighbours come from LDS planes; z neighbours are carried in registers.
//
// Inputs (V, tau, Pr) and outputs (V_out, tau_out, Pr_out) are distinct
// buffers, so a block never reads a value another block has advanced.  Points
// of the ring outside the Launcher box keep their stored (input) value, as in
// the two-kernel sequence where update_stress! does not touch them.  Every
// arithmetic expression is the one of the per-point kernels (stokes.hip), so
// the results are bitwise those of update_stress! + update_velocity!.
#include "common.hpp"

namespace chmy {

constexpr int kFX = 64;

template <class T>
struct FusedArgs {
    DF<T> V[3], tau[6], Pr;            // inputs
    DF<T> Vo[3], tauo_[6], Pro;        // outputs
    DF<T> told[6], eta, eml, rg;
    DF<T> dV, rV[3];
    int64_t n[3];                      // grid cells
    T ix, iy, iz;
    T eta_s, eve_s, Gs, dts, Gdt, dPr, dr, nu;
    int has_rg, write_res, kchunk;
};

template <class T>
__device__ __forceinline__ T f_diag(T t, T to, T e, T dV, T eta, T eve, T Gdt, T dr) {
    T r = ((-(t - to)) / Gdt - t / eta) + T(2) * (e - dV / T(3));
    return t + (r * eve) * dr;
}
template <class T>
__device__ __forceinline__ T f_off(T t, T to, T e, T eta, T eve, T Gdt, T dr) {
    T r = ((-(t - to)) / Gdt - t / eta) + T(2) * e;
    return t + (r * eve) * dr;
}
template <class T>
__device__ __forceinline__ T lerp4(T a0, T a1, T b0, T b1) {
    // knots (lo,lo),(hi,lo),(lo,hi),(hi,hi); first dim innermost (interpolation.jl:37-56)
    T A = fma(T(0.5), a1 - a0, a0);
    T B = fma(T(0.5), b1 - b0, b0);
    return fma(T(0.5), B - A, A);
}

template <class T, bool FVISC, int BY>
__global__ void __launch_bounds__(kFX * BY) fused3d_kernel(FusedArgs<T> a) {
    __shared__ T s_vx[BY][kFX], s_vy[BY][kFX], s_vz1[BY][kFX];
    __shared__ T s_eta[FVISC ? BY : 1][kFX], s_eta1[FVISC ? BY : 1][kFX], s_eml[FVISC ? BY : 1][kFX];
    __shared__ T s_P[BY][kFX], s_txx[BY][kFX], s_tyy[BY][kFX], s_txy[BY][kFX];
    __shared__ T s_txz[2][BY][kFX], s_tyz[2][BY][kFX];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int64_t i = (int64_t)blockIdx.x * (kFX - 2) - 1 + tx;
    const int64_t j = (int64_t)blockIdx.y * (BY - 2) - 1 + ty;
    const int64_t nb0 = a.n[0] + 1, nb1 = a.n[1] + 1, nb2 = a.n[2] + 1;   // box is [0, n+1]
    const int64_t kb = (int64_t)blockIdx.z * a.kchunk;
    const int64_t ke = min(kb + (int64_t)a.kchunk, nb2 + 1);
    // storage holds logical [-1, n+2] (Center) / [-1, n+3] (Vertex) per dim
    const bool valid = i <= a.n[0] + 2 && j <= a.n[1] + 2;
    const bool inxy = i >= 0 && i <= nb0 && j >= 0 && j <= nb1;
    const bool own = tx >= 1 && tx <= kFX - 2 && ty >= 1 && ty <= BY - 2 && inxy;
    const T ix = a.ix, iy = a.iy, iz = a.iz;

    auto O = [i, j](DF<T> f, int64_t k) { return i + j * f.sy + k * f.sz; };
    auto ld = [i, j, valid](DF<T> f, int64_t k) -> T { return valid ? f.p[i + j * f.sy + k * f.sz] : T(0); };

    // registers carried along z (own column)
    T vx_k = ld(a.V[0], kb - 1), vy_k = ld(a.V[1], kb - 1), vz_k = ld(a.V[2], kb - 1);
    T eta_k = FVISC ? ld(a.eta, kb - 1) : T(0);
    T eml_km1 = FVISC ? ld(a.eml, kb - 1) : T(0);
    T P_km1 = T(0), tzz_km1 = T(0), txz_k = T(0), tyz_k = T(0);
    int par = 0;

    for (int64_t k = kb - 1; k < ke; ++k) {
        const bool pro = k < kb;          // prologue: plane kb-1 feeds the first step
        // ---------------- A: loads of plane k+1, V/eta planes to LDS
        const T vx_k1 = ld(a.V[0], k + 1), vy_k1 = ld(a.V[1], k + 1), vz_k1 = ld(a.V[2], k + 1);
        T eta_k1 = T(0), eml_k = T(0);
        s_vx[ty][tx] = vx_k;
        s_vy[ty][tx] = vy_k;
        s_vz1[ty][tx] = vz_k1;
        if constexpr (FVISC) {
            eta_k1 = ld(a.eta, k + 1);
            eml_k = ld(a.eml, k);
            s_eta[ty][tx] = eta_k;
            s_eta1[ty][tx] = eta_k1;
            s_eml[ty][tx] = eml_k;
        }
        __syncthreads();

        // ---------------- B: stresses
        const bool in_k = inxy && k >= 0 && k <= nb2;
        const bool in_k1 = inxy && k + 1 <= nb2;
        const int tx1 = tx < kFX - 1 ? tx + 1 : tx, ty1 = ty < BY - 1 ? ty + 1 : ty;
        const int txm = tx > 0 ? tx - 1 : tx, tym = ty > 0 ? ty - 1 : ty;
        // Center plane k
        T exx = (s_vx[ty][tx1] - vx_k) * ix;
        T eyy = (s_vy[ty1][tx] - vy_k) * iy;
        T ezz = (vz_k1 - vz_k) * iz;
        T dv = (exx + eyy) + ezz;
        T P = ld(a.Pr, k), txx = ld(a.tau[0], k), tyy = ld(a.tau[1], k), tzz = ld(a.tau[2], k);
        T txy = ld(a.tau[3], k);
        T txz1 = ld(a.tau[4], k + 1), tyz1 = ld(a.tau[5], k + 1);
        T eta_c = a.eta_s, eve_c = a.eve_s;
        if constexpr (FVISC) {
            eta_c = eta_k;
            eve_c = visc_ve(eta_c, a.Gs, a.dts);
        }
        if (in_k) {
            P = P - (dv * eve_c) * a.dPr;
            txx = f_diag(txx, ld(a.told[0], k), exx, dv, eta_c, eve_c, a.Gdt, a.dr);
            tyy = f_diag(tyy, ld(a.told[1], k), eyy, dv, eta_c, eve_c, a.Gdt, a.dr);
            tzz = f_diag(tzz, ld(a.told[2], k), ezz, dv, eta_c, eve_c, a.Gdt, a.dr);
            T exy = T(0.5) * ((vx_k - s_vx[tym][tx]) * iy + (vy_k - s_vy[ty][txm]) * ix);
            T et = a.eta_s, ev = a.eve_s;
            if constexpr (FVISC) {
                et = lerp4(s_eta[tym][txm], s_eta[tym][tx], s_eta[ty][txm], eta_k);
                ev = visc_ve(et, a.Gs, a.dts);
            }
            txy = f_off(txy, ld(a.told[3], k), exy, et, ev, a.Gdt, a.dr);
        }
        if (in_k1) {
            T exz = T(0.5) * ((vx_k1 - vx_k) * iz + (vz_k1 - s_vz1[ty][txm]) * ix);
            T eyz = T(0.5) * ((vy_k1 - vy_k) * iz + (vz_k1 - s_vz1[tym][tx]) * iy);
            T et4 = a.eta_s, ev4 = a.eve_s, et5 = a.eta_s, ev5 = a.eve_s;
            if constexpr (FVISC) {
                et4 = lerp4(s_eta[ty][txm], eta_k, s_eta1[ty][txm], eta_k1);
                ev4 = visc_ve(et4, a.Gs, a.dts);
                et5 = lerp4(s_eta[tym][tx], eta_k, s_eta1[tym][tx], eta_k1);
                ev5 = visc_ve(et5, a.Gs, a.dts);
            }
            txz1 = f_off(txz1, ld(a.told[4], k + 1), exz, et4, ev4, a.Gdt, a.dr);
            tyz1 = f_off(tyz1, ld(a.told[5], k + 1), eyz, et5, ev5, a.Gdt, a.dr);
        }
        if (!pro && own && in_k) {
            const int64_t oc = O(a.Pro, k);
            a.Pro.at(oc) = P;
            a.tauo_[0].at(O(a.tauo_[0], k)) = txx;
            a.tauo_[1].at(O(a.tauo_[1], k)) = tyy;
            a.tauo_[2].at(O(a.tauo_[2], k)) = tzz;
            a.tauo_[3].at(O(a.tauo_[3], k)) = txy;
            a.tauo_[4].at(O(a.tauo_[4], k)) = txz_k;
            a.tauo_[5].at(O(a.tauo_[5], k)) = tyz_k;
            if (a.write_res) a.dV.at(O(a.dV, k)) = dv;
        }
        s_P[ty][tx] = P;
        s_txx[ty][tx] = txx;
        s_tyy[ty][tx] = tyy;
        s_txy[ty][tx] = txy;
        s_txz[par][ty][tx] = txz1;
        s_tyz[par][ty][tx] = tyz1;
        __syncthreads();

        // ---------------- C: velocity at plane k
        if (!pro) {
            const T rx = ((-((P - s_P[ty][txm]) * ix) + (txx - s_txx[ty][txm]) * ix) +
                          (s_txy[ty1][tx] - txy) * iy) + (txz1 - txz_k) * iz;
            const T ry = ((-((P - s_P[tym][tx]) * iy) + (tyy - s_tyy[tym][tx]) * iy) +
                          (s_txy[ty][tx1] - txy) * ix) + (tyz1 - tyz_k) * iz;
            T rz = ((-((P - P_km1) * iz) + (tzz - tzz_km1) * iz) + (s_txz[par ^ 1][ty][tx1] - txz_k) * ix) +
                   (s_tyz[par ^ 1][ty1][tx] - tyz_k) * iy;
            if (own && in_k) {
                if (a.has_rg) rz = rz - a.rg.at(O(a.rg, k));
                T ex = a.eve_s, ey = a.eve_s, ez = a.eve_s;
                if constexpr (FVISC) {
                    ex = visc_ve(jmax(s_eml[ty][txm], eml_k), a.Gs, a.dts);
                    ey = visc_ve(jmax(s_eml[tym][tx], eml_k), a.Gs, a.dts);
                    ez = visc_ve(jmax(eml_km1, eml_k), a.Gs, a.dts);
                }
                a.Vo[0].at(O(a.Vo[0], k)) = vx_k + (rx * a.nu) / ex;
                a.Vo[1].at(O(a.Vo[1], k)) = vy_k + (ry * a.nu) / ey;
                a.Vo[2].at(O(a.Vo[2], k)) = vz_k + (rz * a.nu) / ez;
                if (a.write_res) {
                    a.rV[0].at(O(a.rV[0], k)) = rx;
                    a.rV[1].at(O(a.rV[1], k)) = ry;
                    a.rV[2].at(O(a.rV[2], k)) = rz;
                }
            }
        }
        // ---------------- rotate the z pipeline
        P_km1 = P;
        tzz_km1 = tzz;
        txz_k = txz1;
        tyz_k = tyz1;
        vx_k = vx_k1;
        vy_k = vy_k1;
        vz_k = vz_k1;
        eta_k = eta_k1;
        eml_km1 = eml_k;
        par ^= 1;
        __syncthreads();
    }
}

// 2D: one tile per block, no z pipeline.
template <class T, bool FVISC, int BY>
__global__ void __launch_bounds__(kFX * BY) fused2d_kernel(FusedArgs<T> a) {
    __shared__ T s_vx[BY][kFX], s_vy[BY][kFX];
    __shared__ T s_eta[FVISC ? BY : 1][kFX], s_eml[FVISC ? BY : 1][kFX];
    __shared__ T s_P[BY][kFX], s_txx[BY][kFX], s_tyy[BY][kFX], s_txy[BY][kFX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int64_t i = (int64_t)blockIdx.x * (kFX - 2) - 1 + tx;
    const int64_t j = (int64_t)blockIdx.y * (BY - 2) - 1 + ty;
    const int64_t nb0 = a.n[0] + 1, nb1 = a.n[1] + 1;
    const bool valid = i <= a.n[0] + 2 && j <= a.n[1] + 2;
    const bool inxy = i >= 0 && i <= nb0 && j >= 0 && j <= nb1;
    const bool own = tx >= 1 && tx <= kFX - 2 && ty >= 1 && ty <= BY - 2 && inxy;
    const T ix = a.ix, iy = a.iy;
    auto O = [i, j](DF<T> f) { return i + j * f.sy; };
    auto ld = [i, j, valid](DF<T> f) -> T { return valid ? f.p[i + j * f.sy] : T(0); };

    const T vx = ld(a.V[0]), vy = ld(a.V[1]);
    T eta = T(0), eml = T(0);
    s_vx[ty][tx] = vx;
    s_vy[ty][tx] = vy;
    if constexpr (FVISC) {
        eta = ld(a.eta);
        eml = ld(a.eml);
        s_eta[ty][tx] = eta;
        s_eml[ty][tx] = eml;
    }
    __syncthreads();
    const int tx1 = tx < kFX - 1 ? tx + 1 : tx, ty1 = ty < BY - 1 ? ty + 1 : ty;
    const int txm = tx > 0 ? tx - 1 : tx, tym = ty > 0 ? ty - 1 : ty;
    T exx = (s_vx[ty][tx1] - vx) * ix;
    T eyy = (s_vy[ty1][tx] - vy) * iy;
    T dv = exx + eyy;
    T P = ld(a.Pr), txx = ld(a.tau[0]), tyy = ld(a.tau[1]), txy = ld(a.tau[2]);
    if (inxy) {
        T eta_c = a.eta_s, eve_c = a.eve_s, et = a.eta_s, ev = a.eve_s;
        if constexpr (FVISC) {
            eta_c = eta;
            eve_c = visc_ve(eta_c, a.Gs, a.dts);
            et = lerp4(s_eta[tym][txm], s_eta[tym][tx], s_eta[ty][txm], eta);
            ev = visc_ve(et, a.Gs, a.dts);
        }
        P = P - (dv * eve_c) * a.dPr;
        txx = f_diag(txx, ld(a.told[0]), exx, dv, eta_c, eve_c, a.Gdt, a.dr);
        tyy = f_diag(tyy, ld(a.told[1]), eyy, dv, eta_c, eve_c, a.Gdt, a.dr);
        T exy = T(0.5) * ((vx - s_vx[tym][tx]) * iy + (vy - s_vy[ty][txm]) * ix);
        txy = f_off(txy, ld(a.told[2]), exy, et, ev, a.Gdt, a.dr);
    }
    if (own) {
        a.Pro.at(O(a.Pro)) = P;
        a.tauo_[0].at(O(a.tauo_[0])) = txx;
        a.tauo_[1].at(O(a.tauo_[1])) = tyy;
        a.tauo_[2].at(O(a.tauo_[2])) = txy;
        if (a.write_res) a.dV.at(O(a.dV)) = dv;
    }
    s_P[ty][tx] = P;
    s_txx[ty][tx] = txx;
    s_tyy[ty][tx] = tyy;
    s_txy[ty][tx] = txy;
    __syncthreads();
    if (own) {
        const T rx = (-((P - s_P[ty][txm]) * ix) + (txx - s_txx[ty][txm]) * ix) + (s_txy[ty1][tx] - txy) * iy;
        T ry = (-((P - s_P[tym][tx]) * iy) + (tyy - s_tyy[tym][tx]) * iy) + (s_txy[ty][tx1] - txy) * ix;
        if (a.has_rg) ry = ry - a.rg.at(O(a.rg));
        T ex = a.eve_s, ey = a.eve_s;
        if constexpr (FVISC) {
            ex = visc_ve(jmax(s_eml[ty][txm], eml), a.Gs, a.dts);
            ey = visc_ve(jmax(s_eml[tym][tx], eml), a.Gs, a.dts);
        }
        a.Vo[0].at(O(a.Vo[0])) = vx + (rx * a.nu) / ex;
        a.Vo[1].at(O(a.Vo[1])) = vy + (ry * a.nu) / ey;
        if (a.write_res) {
            a.rV[0].at(O(a.rV[0])) = rx;
            a.rV[1].at(O(a.rV[1])) = ry;
        }
    }
}

constexpr int kBY3 = 8, kBY2 = 8, kZChunk = 32;

template <class T>
static int fused_launch(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau, const chmy_field_t* Pr,
                        const chmy_field_t* Vo, const chmy_field_t* tauo, const chmy_field_t* Pro,
                        const chmy_field_t* told, const chmy_field_t* eta, const chmy_field_t* eml,
                        const chmy_field_t* rg, const chmy_field_t* dV, const chmy_field_t* rV, int wres,
                        double eta_s, double eve_s, double G, double dt, double dPr, double dr, double nu,
                        hipStream_t s) {
    FusedArgs<T> a;
    std::memset(&a, 0, sizeof a);
    const int nd = g->nd, nt = nd == 2 ? 3 : 6;
    for (int d = 0; d < nd; ++d) {
        a.V[d] = view<T>(V[d]);
        a.Vo[d] = view<T>(Vo[d]);
        a.rV[d] = wres ? view<T>(rV[d]) : a.Vo[d];
    }
    for (int c = 0; c < nt; ++c) {
        a.tau[c] = view<T>(tau[c]);
        a.tauo_[c] = view<T>(tauo[c]);
        a.told[c] = view<T>(told[c]);
    }
    a.Pr = view<T>(*Pr);
    a.Pro = view<T>(*Pro);
    a.eta = eta ? view<T>(*eta) : a.Pr;
    a.eml = eml ? view<T>(*eml) : a.Pr;
    a.rg = rg ? view<T>(*rg) : a.Pr;
    a.dV = wres ? view<T>(*dV) : a.Pro;
    for (int d = 0; d < 3; ++d) a.n[d] = d < nd ? g->n[d] : 0;
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
    a.nu = T(nu);
    a.has_rg = rg != nullptr;
    a.write_res = wres;
    const bool fv = eta != nullptr;
    if (nd == 3) {
        a.kchunk = kZChunk;
        const int64_t nbx = g->n[0] + 2, nby = g->n[1] + 2, nbz = g->n[2] + 2;
        dim3 blk(kFX, kBY3, 1);
        dim3 grd((unsigned)((nbx + kFX - 3) / (kFX - 2)), (unsigned)((nby + kBY3 - 3) / (kBY3 - 2)),
                 (unsigned)((nbz + kZChunk - 1) / kZChunk));
        if (fv) hipLaunchKernelGGL((fused3d_kernel<T, true, kBY3>), grd, blk, 0, s, a);
        else hipLaunchKernelGGL((fused3d_kernel<T, false, kBY3>), grd, blk, 0, s, a);
    } else {
        const int64_t nbx = g->n[0] + 2, nby = g->n[1] + 2;
        dim3 blk(kFX, kBY2, 1);
        dim3 grd((unsigned)((nbx + kFX - 3) / (kFX - 2)), (unsigned)((nby + kBY2 - 3) / (kBY2 - 2)), 1);
        if (fv) hipLaunchKernelGGL((fused2d_kernel<T, true, kBY2>), grd, blk, 0, s, a);
        else hipLaunchKernelGGL((fused2d_kernel<T, false, kBY2>), grd, blk, 0, s, a);
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_stokes_iter(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau, const chmy_field_t* Pr,
                     const chmy_field_t* V_out, const chmy_field_t* tau_out, const chmy_field_t* Pr_out,
                     const chmy_field_t* tau_old, const chmy_field_t* eta, const chmy_field_t* eta_ml,
                     const chmy_field_t* rhog, const chmy_field_t* divV, const chmy_field_t* r_V, int write_residual,
                     double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                     double nudtau, void* stream) {
    CHMY_TRY(check_grid(g));
    if (g->nd < 2) return set_error(CHMY_E_ARG, "Stokes kernels need nd >= 2");
    CHMY_TRY(check_center(g, Pr, "Pr", -1));
    const int dt_ = Pr->dtype;
    CHMY_TRY(check_center(g, Pr_out, "Pr_out", dt_));
    CHMY_TRY(check_vector(g, V, "V", dt_));
    CHMY_TRY(check_vector(g, V_out, "V_out", dt_));
    CHMY_TRY(check_tensor(g, tau, "tau", dt_));
    CHMY_TRY(check_tensor(g, tau_out, "tau_out", dt_));
    CHMY_TRY(check_tensor(g, tau_old, "tau_old", dt_));
    if ((eta == nullptr) != (eta_ml == nullptr)) return set_error(CHMY_E_ARG, "eta and eta_ml go together");
    if (eta) {
        CHMY_TRY(check_center(g, eta, "eta", dt_));
        CHMY_TRY(check_center(g, eta_ml, "eta_ml", dt_));
    }
    if (rhog) {
        int loc[3] = {0, 0, 0};
        loc[g->nd - 1] = 1;
        CHMY_TRY(check_field(g, rhog, "rhog", loc, dt_));
    }
    if (write_residual) {
        CHMY_TRY(check_center(g, divV, "divV", dt_));
        CHMY_TRY(check_vector(g, r_V, "r_V", dt_));
    }
    for (int d = 0; d < g->nd; ++d)
        for (const chmy_field_t* f : {Pr, Pr_out, eta, eta_ml})
            if (f && f->halo[d] != 1) return set_error(CHMY_E_HALO, "fused iteration needs halo 1");
    if (V[0].data == V_out[0].data || Pr->data == Pr_out->data || tau[0].data == tau_out[0].data)
        return set_error(CHMY_E_ARG, "fused iteration needs distinct input and output buffers");
    return dt_ == CHMY_F64
               ? fused_launch<double>(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, divV, r_V,
                                      write_residual, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau,
                                      (hipStream_t)stream)
               : fused_launch<float>(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, divV, r_V,
                                     write_residual, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau,
                                     (hipStream_t)stream);
}

}  // extern "C"
