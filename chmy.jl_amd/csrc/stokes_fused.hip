// One pseudo-transient Stokes iteration in a single pass over HBM:
// update_stress! followed by update_velocity! (stokes_3d_inc_ve_T.jl:164-165),
// fused and double-buffered.
//
// Work layout (3D): a block is 64 (x, one wave) x BY (y) threads; thread
// (tx, ty) owns the logical column (i, j) = (i0-1+tx, j0-1+ty) and marches z
// over a chunk [kb, ke).  The tile's outputs are tx in [1, 62], ty in
// [1, BY-1]: lanes 0 and 63 and row 0 recompute the stresses that the
// velocity stencil needs (Center values at I-e, edge values at I+e), and the
// last wave also stages row j+1 of V and eta, so an advanced stress is never
// read back from HBM.  At z-step k a thread computes the Center stresses and
// Pr at plane k, tau_xy at row j+1, tau_xz, tau_yz at z-vertex plane k+1,
// then V at plane k.  x/y neighbours come from LDS planes; z neighbours are
// carried in registers.  Two barriers per z-step: LDS planes whose readers
// and next writers are not separated by a barrier are double-buffered by
// step parity.
//
// 2D: fused2d_kernel marches y with one wave per 64 columns and no LDS (x
// neighbours by lane shuffles).  Two-iteration passes (fused2d_tb_kernel,
// fused3d_tb_kernel) run the step twice per pass with the free-slip V batch
// applied in-kernel between the iterations; see their comments.
//
// Addressing: the 7 staggered locations (Center, Vx, Vy, Vz, xy, xz, yz edges)
// each have their own padded strides; a thread keeps one 32-bit byte offset per
// location (advanced by the z stride each step) and every array is a uniform
// base pointer, so loads are SGPR-base + VGPR-offset.  The host checks that
// every offset fits in 32 bits.
//
// Inputs (V, tau, Pr) and outputs (V_out, tau_out, Pr_out) are distinct
// buffers, so a block never reads a value another block has advanced.  Ring
// points outside the Launcher box keep their stored (input) value, as in the
// two-kernel sequence where update_stress! does not touch them.  With
// CHMY_ITER_EXACT every arithmetic expression is the one of the per-point
// kernels (stokes.hip), so the results are bitwise those of update_stress! +
// update_velocity!; the default fast form (Arith<T, false>) is algebraically
// equal and agrees to a few ulp per iteration.
#include <type_traits>

#include "common.hpp"

namespace chmy {

constexpr int kFX = 64;
enum { LC = 0, LVX = 1, LVY = 2, LVZ = 3, LXY = 4, LXZ = 5, LYZ = 6, NLOC = 7 };

// Scalars shared by both arithmetic forms.
template <class T>
struct Kc {
    T Gs, dts, Gdt, dPr, dr, nu, eta_s, eve_s;   // reference inputs
    T iGdt, omdr, third, a_s, b_s, cv_s;         // fast-form constants (host computed)
};

struct LocStride {
    uint32_t sy, sz;     // element strides
    uint32_t o0;         // element offset of logical (0, 0, 0)
};

template <class T>
struct Fused3 {
    const T *Vx, *Vy, *Vz, *Pr, *txx, *tyy, *tzz, *txy, *txz, *tyz;
    T *oVx, *oVy, *oVz, *oPr, *otxx, *otyy, *otzz, *otxy, *otxz, *otyz;
    const T *Oxx, *Oyy, *Ozz, *Oxy, *Oxz, *Oyz;    // tau_old
    const T *eta, *eml, *rg;
    const T *fVx, *fVy, *fVz, *fP;     // correction forcing (defect correction), optional
    T *dV, *rVx, *rVy, *rVz;
    LocStride L[NLOC];
    int32_t n[3];
    int32_t ntx, nty, nchunks, kchunk;
    int32_t has_rg, write_res, has_f;
    T ix, iy, iz;
    Kc<T> k;
};

#ifndef CHMY_FUSED_NT_STORE
#define CHMY_FUSED_NT_STORE 0
#endif

template <class T>
__device__ __forceinline__ T ldo(const T* p, uint32_t boff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + boff);
}
// Loads of the sweep's read-only forcing (tau_old, f_V, f_P): in the sweep
// precision, or (BF, FP32 sweeps only) from bf16 storage at half the byte
// offset of the FP32 layout.
template <class T, bool BF>
__device__ __forceinline__ T ldf(const T* p, uint32_t boff) {
    if constexpr (BF) {
        static_assert(sizeof(T) == 4, "bf16 forcing goes with FP32 sweeps");
        return T(bf16_to_f32(*reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(p) + (boff >> 1))));
    } else {
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + boff);
    }
}

template <class T>
__device__ __forceinline__ void sto(T* p, uint32_t boff, T v) {
#if CHMY_FUSED_NT_STORE
    // streaming store: the advanced state is not re-read in this launch
    __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(p) + boff));
#else
    *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + boff) = v;
#endif
}

// EXACT: the reference's expressions (stokes_3d_inc_ve_T.jl:23-57), bitwise.
// FAST: the same update with eta_ve*(1/(G dt) + 1/eta) = 1 used to cancel
// tau/eta: tau_new = tau(1 - dtau_r) + eta_ve dtau_r (tau_old/(G dt) + 2 eps'),
// eta_ve = eta/(1 + eta/(G dt)), Pr_new = Pr - divV eta_ve dtau_Pr and
// V_new = V + r nudtau (1/eta_V + 1/(G dt)): 7 divisions per point instead
// of 31, results within a few ulp per iteration of the exact form.
template <class T, bool EXACT>
struct Arith;

template <class T>
struct Arith<T, true> {
    struct Loc { T eta, eve; };
    __device__ static Loc field(T eta, const Kc<T>& k) { return Loc{eta, visc_ve(eta, k.Gs, k.dts)}; }
    __device__ static Loc scalar(const Kc<T>& k) { return Loc{k.eta_s, k.eve_s}; }
    __device__ static T diag(T t, T to, T e, T dV, Loc L, const Kc<T>& k) {
        T r = ((-(t - to)) / k.Gdt - t / L.eta) + T(2) * (e - dV / T(3));
        return t + (r * L.eve) * k.dr;
    }
    __device__ static T off(T t, T to, T e, Loc L, const Kc<T>& k) {
        T r = ((-(t - to)) / k.Gdt - t / L.eta) + T(2) * e;
        return t + (r * L.eve) * k.dr;
    }
    __device__ static T pr(T P, T dv, Loc L, const Kc<T>& k) { return P - (dv * L.eve) * k.dPr; }
    __device__ static T vcoef_field(T em, const Kc<T>& k) { return visc_ve(em, k.Gs, k.dts); }
    __device__ static T vcoef_scalar(const Kc<T>& k) { return k.eve_s; }
    __device__ static T vel(T v, T r, T c, const Kc<T>& k) { return v + (r * k.nu) / c; }
};

// Divisions of the fast form: IEEE in FP64; in FP32 (the defect-correction
// sweeps) a times the hardware reciprocal of b (v_rcp_f32, 1 ulp), for its 7
// divisions per point
__device__ __forceinline__ double fast_div(double a, double b) { return a / b; }
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <class T>
struct Arith<T, false> {
    struct Loc { T a, b; };   // a = eta_ve dtau_r, b = eta_ve dtau_Pr
    __device__ static Loc field(T eta, const Kc<T>& k) {
        const T eve = fast_div(eta, fma(eta, k.iGdt, T(1)));
        return Loc{eve * k.dr, eve * k.dPr};
    }
    __device__ static Loc scalar(const Kc<T>& k) { return Loc{k.a_s, k.b_s}; }
    __device__ static T diag(T t, T to, T e, T dV, Loc L, const Kc<T>& k) {
        return fma(t, k.omdr, L.a * fma(to, k.iGdt, T(2) * fma(-dV, k.third, e)));
    }
    __device__ static T off(T t, T to, T e, Loc L, const Kc<T>& k) {
        return fma(t, k.omdr, L.a * fma(to, k.iGdt, T(2) * e));
    }
    __device__ static T pr(T P, T dv, Loc L, const Kc<T>& k) { return fma(-dv, L.b, P); }
    __device__ static T vcoef_field(T em, const Kc<T>& k) { return k.nu * (fast_div(T(1), em) + k.iGdt); }
    __device__ static T vcoef_scalar(const Kc<T>& k) { return k.cv_s; }
    __device__ static T vel(T v, T r, T c, const Kc<T>& k) { return fma(r, c, v); }
};

template <class T>
__device__ __forceinline__ T lerp4(T a0, T a1, T b0, T b1) {
    // knots (lo,lo),(hi,lo),(lo,hi),(hi,hi); first dim innermost (interpolation.jl:37-56)
    T A = fma(T(0.5), a1 - a0, a0);
    T B = fma(T(0.5), b1 - b0, b0);
    return fma(T(0.5), B - A, A);
}

// Tuning knobs fixed at build time: rows per block (FP64 / FP32 sweeps) and
// the minimum waves per SIMD the register allocation must allow (0 = no bound).
#ifndef CHMY_FUSED_TW64
#define CHMY_FUSED_TW64 64
#endif
#ifndef CHMY_FUSED_TW32
#define CHMY_FUSED_TW32 64
#endif
#ifndef CHMY_FUSED_BY64
#define CHMY_FUSED_BY64 12
#endif
#ifndef CHMY_FUSED_BY32
#define CHMY_FUSED_BY32 8
#endif
#ifndef CHMY_FUSED_WPE
#define CHMY_FUSED_WPE 0
#endif

// OPT selects the optional inputs/outputs at compile time so that the common
// variants keep their pointers in SGPRs: 0 none, 1 forcing only (the
// correction sweeps), 2 rho g only (the reference examples' buoyancy), 7
// forcing, residual fields and rho g as flagged.
//
// Rows: thread row ty sits at y = j = j0 - 1 + ty (j0 = tile * (BY - 1)) and
// computes the Center stresses and tau_xz at row j, and the y-vertex edge
// stresses tau_xy, tau_yz at row j + 1.  Velocities need Center values at
// rows j-1, j and y-vertex edges at rows j, j+1, so rows ty = 1..BY-1 are
// outputs and only row 0 is a ring.  The last wave also stages row j + 1 of
// V and eta (LDS row BY).
template <class T, bool FVISC, bool EXACT, int TW, int BY, int OPT, bool BF>
__global__ void __launch_bounds__(TW * BY, CHMY_FUSED_WPE) fused3d_kernel(Fused3<T> a) {
    using AR = Arith<T, EXACT>;
    const Kc<T>& K = a.k;
    const bool has_f = OPT == 1 || (OPT == 7 && a.has_f);
    const bool write_res = OPT == 7 && a.write_res, has_rg = OPT == 2 || (OPT == 7 && a.has_rg);
    constexpr int NR = BY + 1;   // LDS rows of the V and eta planes
    __shared__ T s_vx[NR][TW], s_vy[2][NR][TW], s_vz1[NR][TW];
    __shared__ T s_eta[FVISC ? 2 : 1][FVISC ? NR : 1][TW], s_eml[FVISC ? 2 : 1][FVISC ? BY : 1][TW];
    __shared__ T s_P[BY][TW], s_txx[BY][TW], s_tyy[BY][TW], s_txy[BY][TW];
    __shared__ T s_txz[2][BY][TW], s_tyz[2][BY][TW];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int nb = a.ntx * a.nty * a.nchunks;
    const int lt = xcd_remap(blockIdx.x, nb);
    const int bx = lt % a.ntx, by = (lt / a.ntx) % a.nty, bz = lt / (a.ntx * a.nty);
    const int i = bx * (TW - 2) - 1 + tx;
    const int j = by * (BY - 1) - 1 + ty;
    const int nb0 = a.n[0] + 1, nb1 = a.n[1] + 1, nb2 = a.n[2] + 1;   // box is [0, n+1]
    const int kb = bz * a.kchunk;
    const int ke = min(kb + a.kchunk, nb2 + 1);
    // storage holds logical [-1, n+2] (Center) / [-1, n+3] (Vertex) per dim
    const bool valid = i <= a.n[0] + 2 && j <= a.n[1] + 2;
    const bool inx = i >= 0 && i <= nb0;
    const bool inc = inx && j >= 0 && j <= nb1;            // Center / tau_xz column in the box
    const bool ine = inx && j + 1 <= nb1;                  // y-vertex edge row j+1 in the box
    const bool ownx = tx >= 1 && tx <= TW - 2;
    const bool ownc = ownx && ty >= 1 && inc;              // outputs: Center, tau_xz, V at row j
    const bool owne = ownx && ty <= BY - 2 && ine;         // outputs: tau_xy, tau_yz at row j+1
    const bool last = ty == BY - 1;                        // stages row j+1 (wave-uniform)
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    constexpr uint32_t E = sizeof(T);

    // byte offsets at plane k = kb - 1 and the per-step z increments; lanes
    // outside the storage keep offset 0 plus whole z planes, which stays inside
    // every field (at most n+4 planes are walked), so all loads are
    // unconditional and such lanes' values reach no output
    // (row j+1 of a lane with offset 0 is still inside plane 0 + the walk).
    // The staging wave reads row j+1 only where it is inside the storage
    // (j + 1 <= n+2 for Center-in-y fields), else its own row (unused values).
    uint32_t o[NLOC], dz[NLOC], dy[NLOC], d8[NLOC];
    const bool row8 = last && j + 1 <= a.n[1] + 2;
#pragma unroll
    for (int c = 0; c < NLOC; ++c) {
        const LocStride s = a.L[c];
        o[c] = valid ? (s.o0 + (uint32_t)i + (uint32_t)j * s.sy + (uint32_t)(kb - 1) * s.sz) * E : 0u;
        dz[c] = s.sz * E;
        dy[c] = s.sy * E;
        d8[c] = row8 ? dy[c] : 0u;
    }
    auto ld = [](const T* p, uint32_t off) -> T { return ldo(p, off); };

    int p = 0;
    T vx_k = ld(a.Vx, o[LVX]), vy_k = ld(a.Vy, o[LVY]), vz_k = ld(a.Vz, o[LVZ]);
    T eta_k = T(0), eml_km1 = T(0);
    s_vy[0][ty][tx] = vy_k;
    if constexpr (FVISC) {
        eta_k = ld(a.eta, o[LC]);
        s_eta[0][ty][tx] = eta_k;
    }
    if (last) {   // row j+1 at plane kb - 1
        s_vy[0][BY][tx] = ld(a.Vy, o[LVY] + d8[LVY]);
        if constexpr (FVISC) s_eta[0][BY][tx] = ld(a.eta, o[LC] + d8[LC]);
    }
    T P_km1 = T(0), tzz_km1 = T(0), txz_k = T(0), tyz_k = T(0);

    for (int k = kb - 1; k < ke; ++k) {
        const bool pro = k < kb;
        // ---------------- A: all of this step's global loads first (plane k+1 V
        // and eta, plane k stresses and tau_old), then the V and eta planes to LDS
        const T vx_k1 = ld(a.Vx, o[LVX] + dz[LVX]);
        const T vy_k1 = ld(a.Vy, o[LVY] + dz[LVY]);
        const T vz_k1 = ld(a.Vz, o[LVZ] + dz[LVZ]);
        T eta_k1 = T(0), eml_k = T(0);
        if constexpr (FVISC) {
            eta_k1 = ld(a.eta, o[LC] + dz[LC]);
            eml_k = ld(a.eml, o[LC]);
        }
        T P = ld(a.Pr, o[LC]), txx = ld(a.txx, o[LC]), tyy = ld(a.tyy, o[LC]), tzz = ld(a.tzz, o[LC]);
        T txy1 = ld(a.txy, o[LXY] + dy[LXY]);
        T txz1 = ld(a.txz, o[LXZ] + dz[LXZ]), tyz1 = ld(a.tyz, o[LYZ] + dy[LYZ] + dz[LYZ]);
        const T oxx = ldf<T, BF>(a.Oxx, o[LC]), oyy = ldf<T, BF>(a.Oyy, o[LC]), ozz = ldf<T, BF>(a.Ozz, o[LC]);
        const T oxy = ldf<T, BF>(a.Oxy, o[LXY] + dy[LXY]);
        const T oxz = ldf<T, BF>(a.Oxz, o[LXZ] + dz[LXZ]), oyz = ldf<T, BF>(a.Oyz, o[LYZ] + dy[LYZ] + dz[LYZ]);
        T vx8 = T(0), vy8_1 = T(0), vz8_1 = T(0), eta8_1 = T(0);   // row j+1 (staging wave)
        if (last) {
            vx8 = ld(a.Vx, o[LVX] + d8[LVX]);
            vy8_1 = ld(a.Vy, o[LVY] + d8[LVY] + dz[LVY]);
            vz8_1 = ld(a.Vz, o[LVZ] + d8[LVZ] + dz[LVZ]);
            if constexpr (FVISC) eta8_1 = ld(a.eta, o[LC] + d8[LC] + dz[LC]);
        }
        const bool in_k = k >= 0;
        const bool in_k1 = k + 1 <= nb2;
        s_vx[ty][tx] = vx_k;
        s_vy[p ^ 1][ty][tx] = vy_k1;
        s_vz1[ty][tx] = vz_k1;
        if constexpr (FVISC) {
            s_eta[p ^ 1][ty][tx] = eta_k1;
            s_eml[p][ty][tx] = eml_k;
        }
        if (last) {
            s_vx[BY][tx] = vx8;
            s_vy[p ^ 1][BY][tx] = vy8_1;
            s_vz1[BY][tx] = vz8_1;
            if constexpr (FVISC) s_eta[p ^ 1][BY][tx] = eta8_1;
        }
        __syncthreads();

        // ---------------- B: stresses (computed on every lane, kept in the box)
        const int tx1 = tx < TW - 1 ? tx + 1 : tx, txm = tx > 0 ? tx - 1 : tx;
        const int tym = ty > 0 ? ty - 1 : ty;
        const T exx = (s_vx[ty][tx1] - vx_k) * ix;
        const T eyy = (s_vy[p][ty + 1][tx] - vy_k) * iy;
        const T ezz = (vz_k1 - vz_k) * iz;
        const T dv = (exx + eyy) + ezz;
        {
            typename AR::Loc Lc;
            if constexpr (FVISC) Lc = AR::field(eta_k, K);
            else Lc = AR::scalar(K);
            const T Pn = AR::pr(P, has_f ? dv - ldf<T, BF>(a.fP, o[LC]) : dv, Lc, K);
            const T txxn = AR::diag(txx, oxx, exx, dv, Lc, K);
            const T tyyn = AR::diag(tyy, oyy, eyy, dv, Lc, K);
            const T tzzn = AR::diag(tzz, ozz, ezz, dv, Lc, K);
            const bool u = inc && in_k;
            P = u ? Pn : P;
            txx = u ? txxn : txx;
            tyy = u ? tyyn : tyy;
            tzz = u ? tzzn : tzz;
        }
        {   // tau_xy at (i, j+1, k)
            const T exy = T(0.5) * ((s_vx[ty + 1][tx] - vx_k) * iy + (s_vy[p][ty + 1][tx] - s_vy[p][ty + 1][txm]) * ix);
            typename AR::Loc Le;
            if constexpr (FVISC)
                Le = AR::field(lerp4(s_eta[p][ty][txm], eta_k, s_eta[p][ty + 1][txm], s_eta[p][ty + 1][tx]), K);
            else Le = AR::scalar(K);
            const T n = AR::off(txy1, oxy, exy, Le, K);
            txy1 = ine && in_k ? n : txy1;
        }
        {   // tau_xz at (i, j, k+1) and tau_yz at (i, j+1, k+1)
            const T exz = T(0.5) * ((vx_k1 - vx_k) * iz + (vz_k1 - s_vz1[ty][txm]) * ix);
            const T eyz = T(0.5) * ((s_vy[p ^ 1][ty + 1][tx] - s_vy[p][ty + 1][tx]) * iz + (s_vz1[ty + 1][tx] - vz_k1) * iy);
            typename AR::Loc L4, L5;
            if constexpr (FVISC) {
                L4 = AR::field(lerp4(s_eta[p][ty][txm], eta_k, s_eta[p ^ 1][ty][txm], eta_k1), K);
                L5 = AR::field(lerp4(eta_k, s_eta[p][ty + 1][tx], eta_k1, s_eta[p ^ 1][ty + 1][tx]), K);
            } else {
                L4 = L5 = AR::scalar(K);
            }
            const T nxz = AR::off(txz1, oxz, exz, L4, K);
            const T nyz = AR::off(tyz1, oyz, eyz, L5, K);
            txz1 = inc && in_k1 ? nxz : txz1;
            tyz1 = ine && in_k1 ? nyz : tyz1;
        }
        if (!pro && in_k) {
            if (ownc) {
                sto(a.oPr, o[LC], P);
                sto(a.otxx, o[LC], txx);
                sto(a.otyy, o[LC], tyy);
                sto(a.otzz, o[LC], tzz);
                sto(a.otxz, o[LXZ], txz_k);
                if (write_res) sto(a.dV, o[LC], dv);
            }
            if (owne) {
                sto(a.otxy, o[LXY] + dy[LXY], txy1);
                sto(a.otyz, o[LYZ] + dy[LYZ], tyz_k);
            }
        }
        s_P[ty][tx] = P;
        s_txx[ty][tx] = txx;
        s_tyy[ty][tx] = tyy;
        s_txy[ty][tx] = txy1;
        s_txz[p][ty][tx] = txz1;
        s_tyz[p][ty][tx] = tyz1;
        __syncthreads();

        // ---------------- C: velocity at plane k, rows ty >= 1
        if (!pro && ownc && in_k) {
            const T txy_j = s_txy[tym][tx];   // tau_xy at (i, j, k), from row ty-1
            T rx = ((-((P - s_P[ty][txm]) * ix) + (txx - s_txx[ty][txm]) * ix) + (txy1 - txy_j) * iy) +
                   (txz1 - txz_k) * iz;
            T ry = ((-((P - s_P[tym][tx]) * iy) + (tyy - s_tyy[tym][tx]) * iy) + (s_txy[tym][tx1] - txy_j) * ix) +
                   (s_tyz[p][tym][tx] - s_tyz[p ^ 1][tym][tx]) * iz;
            T rz = ((-((P - P_km1) * iz) + (tzz - tzz_km1) * iz) + (s_txz[p ^ 1][ty][tx1] - txz_k) * ix) +
                   (tyz_k - s_tyz[p ^ 1][tym][tx]) * iy;
            if (has_rg) rz = rz - ldo(a.rg, o[LVZ]);
            if (has_f) {
                rx = rx + ldf<T, BF>(a.fVx, o[LVX]);
                ry = ry + ldf<T, BF>(a.fVy, o[LVY]);
                rz = rz + ldf<T, BF>(a.fVz, o[LVZ]);
            }
            T ex, ey, ez;
            if constexpr (FVISC) {
                ex = AR::vcoef_field(jmax(s_eml[p][ty][txm], eml_k), K);
                ey = AR::vcoef_field(jmax(s_eml[p][tym][tx], eml_k), K);
                ez = AR::vcoef_field(jmax(eml_km1, eml_k), K);
            } else {
                ex = ey = ez = AR::vcoef_scalar(K);
            }
            sto(a.oVx, o[LVX], AR::vel(vx_k, rx, ex, K));
            sto(a.oVy, o[LVY], AR::vel(vy_k, ry, ey, K));
            sto(a.oVz, o[LVZ], AR::vel(vz_k, rz, ez, K));
            if (write_res) {
                sto(a.rVx, o[LVX], rx);
                sto(a.rVy, o[LVY], ry);
                sto(a.rVz, o[LVZ], rz);
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
        p ^= 1;
#pragma unroll
        for (int c = 0; c < NLOC; ++c) o[c] += dz[c];
    }
}

// 2D: each wave owns an x-tile of 64 columns (62 outputs + a one-column ring)
// and marches y over a chunk of rows.  x neighbours come from wave shuffles,
// y neighbours from registers carried along the march, so there is no LDS,
// no barrier and no y ring.  At row j a lane computes the Center stresses and
// Pr at row j and tau_xy at vertex row j+1, then V at row j (the 2D form of
// the 3D z pipeline).  Addressing as in 3D: one 32-bit byte offset per
// staggered location (C, Vx, Vy, xy), advanced by one row per step; lanes
// right of the storage address the last stored column (their values reach no
// output), so every load is unconditional.
enum { L2C = 0, L2X = 1, L2Y = 2, L2XY = 3, NLOC2 = 4 };

template <class T>
struct Fused2 {
    const T *Vx, *Vy, *Pr, *txx, *tyy, *txy;
    T *oVx, *oVy, *oPr, *otxx, *otyy, *otxy;
    const T *Oxx, *Oyy, *Oxy;
    const T *eta, *eml, *rg;
    const T *fVx, *fVy, *fP;
    T *dV, *rVx, *rVy;
    LocStride L[NLOC2];
    int32_t n[2];
    int32_t ntx, nchunks, ychunk;
    int32_t has_rg, write_res, has_f;
    T ix, iy;
    Kc<T> k;
};

template <class T>
__device__ __forceinline__ T shup(T v) { return __shfl_up(v, 1, kFX); }     // value of lane l-1
template <class T>
__device__ __forceinline__ T shdn(T v) { return __shfl_down(v, 1, kFX); }   // value of lane l+1

#ifndef CHMY_FUSED_W2
#define CHMY_FUSED_W2 4
#endif
constexpr int kW2 = CHMY_FUSED_W2;   // waves (x-tiles) per block
#ifndef CHMY_FUSED_YCHUNK2
#define CHMY_FUSED_YCHUNK2 8
#endif
constexpr int kYChunk2 = CHMY_FUSED_YCHUNK2;   // rows per wave (the march is unrolled over them)

// Raw load of a forcing value (kept as bf16 bits until use, so the load can
// stay in flight across a row of the march) and its conversion.
template <class T, bool BF>
using Fraw = typename std::conditional<BF, uint16_t, T>::type;
template <class T, bool BF>
__device__ __forceinline__ Fraw<T, BF> ldr(const T* p, uint32_t boff) {
    if constexpr (BF) return *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(p) + (boff >> 1));
    else return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + boff);
}
template <class T, bool BF>
__device__ __forceinline__ T cvf(Fraw<T, BF> v) {
    if constexpr (BF) return T(bf16_to_f32(v));
    else return v;
}

// What one row j of the y-march reads: Center values at row j, the y-face,
// vertex and η values at row j+1, and Vx at row j+1 (Vx at row j is carried).
template <class T, bool BF>
struct Row2 {
    T vx1, vy1, P, txx, tyy, txy1, eta1, eml, rg;
    Fraw<T, BF> oxx, oyy, oxy, fP, fVx, fVy;
};

template <class T, bool FVISC, int OPT, bool BF>
__device__ __forceinline__ void load_row2(const Fused2<T>& a, const uint32_t* q, const uint32_t* dy, bool has_f,
                                          bool has_rg, bool outs, Row2<T, BF>& r) {
    r.vx1 = ldo(a.Vx, q[L2X] + dy[L2X]);
    r.vy1 = ldo(a.Vy, q[L2Y] + dy[L2Y]);
    r.P = ldo(a.Pr, q[L2C]);
    r.txx = ldo(a.txx, q[L2C]);
    r.tyy = ldo(a.tyy, q[L2C]);
    r.txy1 = ldo(a.txy, q[L2XY] + dy[L2XY]);
    r.oxx = ldr<T, BF>(a.Oxx, q[L2C]);
    r.oyy = ldr<T, BF>(a.Oyy, q[L2C]);
    r.oxy = ldr<T, BF>(a.Oxy, q[L2XY] + dy[L2XY]);
    if constexpr (FVISC) {
        r.eta1 = ldo(a.eta, q[L2C] + dy[L2C]);
        r.eml = ldo(a.eml, q[L2C]);
    }
    if (has_f) {
        r.fP = ldr<T, BF>(a.fP, q[L2C]);
        if (outs) {
            r.fVx = ldr<T, BF>(a.fVx, q[L2X]);
            r.fVy = ldr<T, BF>(a.fVy, q[L2Y]);
        }
    }
    if (has_rg && outs) r.rg = ldo(a.rg, q[L2Y]);
}

// One wave owns 64 x columns (62 outputs) and marches y over a chunk of
// rows; x neighbours come from shuffles, y neighbours from values carried in
// registers.  The loads of row j+1 are issued before row j is computed, so a
// wave always has a row of loads in flight.
template <class T, bool FVISC, bool EXACT, int OPT, bool BF>
__global__ void __launch_bounds__(kFX * kW2) fused2d_kernel(Fused2<T> a) {
    using AR = Arith<T, EXACT>;
    const Kc<T>& K = a.k;
    const bool has_f = OPT == 1 || (OPT == 7 && a.has_f);
    const bool write_res = OPT == 7 && a.write_res, has_rg = OPT == 2 || (OPT == 7 && a.has_rg);
    const int lane = threadIdx.x;
    const int xt = blockIdx.x * kW2 + threadIdx.y;
    if (xt >= a.ntx) return;                        // whole wave leaves together
    const int i = xt * (kFX - 2) - 1 + lane;
    const int nb0 = a.n[0] + 1, nb1 = a.n[1] + 1;
    const int jb = blockIdx.y * kYChunk2;
    const int je = min(jb + kYChunk2, nb1 + 1);
    const bool inx = i >= 0 && i <= nb0;
    const bool ownx = lane >= 1 && lane <= kFX - 2 && inx;
    const int ic = min(i, a.n[0] + 2);              // stored column
    const T ix = a.ix, iy = a.iy;
    constexpr uint32_t E = sizeof(T);
    uint32_t o[NLOC2], dy[NLOC2];
#pragma unroll
    for (int c = 0; c < NLOC2; ++c) {
        o[c] = (a.L[c].o0 + (uint32_t)ic + (uint32_t)(jb - 1) * a.L[c].sy) * E;   // row jb - 1
        dy[c] = a.L[c].sy * E;
    }

    T vx_j = ldo(a.Vx, o[L2X]), vy_j = ldo(a.Vy, o[L2Y]);
    T eta_j = FVISC ? ldo(a.eta, o[L2C]) : T(0);
    T P_jm1 = T(0), tyy_jm1 = T(0), txy_j = T(0), eml_jm1 = T(0);
    // Row j: issue the loads of row j+1 into `nxt`, then compute from `cur`.
    // The march is unrolled over a full chunk; rows j >= je (the last chunk)
    // are computed and dropped.  The offsets stop at row je - 1, so every load
    // stays inside the chunk's rows (cached reloads, never used).
    auto row = [&](int j, const Row2<T, BF>& cur, Row2<T, BF>& nxt) {
        const bool pro = j < jb;
        const bool live = j < je;
        const bool in_j = inx && j >= 0;
        const bool in_j1 = inx && j + 1 <= nb1;
        const bool more = j + 1 < je;
        {
            uint32_t q[NLOC2];
#pragma unroll
            for (int c = 0; c < NLOC2; ++c) q[c] = o[c] + (more ? dy[c] : 0u);
            load_row2<T, FVISC, OPT, BF>(a, q, dy, has_f, has_rg, true, nxt);
        }
        __builtin_amdgcn_sched_barrier(0);   // keep the next row's loads ahead of this row's work
        const T vx1 = cur.vx1, vy1 = cur.vy1;
        T P = cur.P, txx = cur.txx, tyy = cur.tyy, txy1 = cur.txy1;
        const T eta1 = FVISC ? cur.eta1 : T(0);
        const T eml = FVISC ? cur.eml : T(0);
        // Center stresses and Pr at row j
        const T exx = (shdn(vx_j) - vx_j) * ix;
        const T eyy = (vy1 - vy_j) * iy;
        const T dv = exx + eyy;
        {
            typename AR::Loc Lc;
            if constexpr (FVISC) Lc = AR::field(eta_j, K);
            else Lc = AR::scalar(K);
            const T oxx = cvf<T, BF>(cur.oxx), oyy = cvf<T, BF>(cur.oyy);
            const T Pn = AR::pr(P, has_f ? dv - cvf<T, BF>(cur.fP) : dv, Lc, K);
            const T txxn = AR::diag(txx, oxx, exx, dv, Lc, K);
            const T tyyn = AR::diag(tyy, oyy, eyy, dv, Lc, K);
            P = in_j ? Pn : P;
            txx = in_j ? txxn : txx;
            tyy = in_j ? tyyn : tyy;
        }
        // tau_xy at vertex row j+1 (needs Vx at rows j, j+1 and Vy at column i-1)
        const T vy1l = shup(vy1);
        const T etal = FVISC ? shup(eta_j) : T(0), eta1l = FVISC ? shup(eta1) : T(0);
        {
            const T exy = T(0.5) * ((vx1 - vx_j) * iy + (vy1 - vy1l) * ix);
            typename AR::Loc Le;
            if constexpr (FVISC) Le = AR::field(lerp4(etal, eta_j, eta1l, eta1), K);
            else Le = AR::scalar(K);
            const T n = AR::off(txy1, cvf<T, BF>(cur.oxy), exy, Le, K);
            txy1 = in_j1 ? n : txy1;
        }
        const T Pl = shup(P), txxl = shup(txx), txyr = shdn(txy_j), emll = FVISC ? shup(eml) : T(0);
        if (live && !pro && ownx && in_j) {
            sto(a.oPr, o[L2C], P);
            sto(a.otxx, o[L2C], txx);
            sto(a.otyy, o[L2C], tyy);
            sto(a.otxy, o[L2XY], txy_j);
            if (write_res) sto(a.dV, o[L2C], dv);
            T rx = (-((P - Pl) * ix) + (txx - txxl) * ix) + (txy1 - txy_j) * iy;
            T ry = (-((P - P_jm1) * iy) + (tyy - tyy_jm1) * iy) + (txyr - txy_j) * ix;
            if (has_rg) ry = ry - cur.rg;
            if (has_f) {
                rx = rx + cvf<T, BF>(cur.fVx);
                ry = ry + cvf<T, BF>(cur.fVy);
            }
            T ex, ey;
            if constexpr (FVISC) {
                ex = AR::vcoef_field(jmax(emll, eml), K);
                ey = AR::vcoef_field(jmax(eml_jm1, eml), K);
            } else {
                ex = ey = AR::vcoef_scalar(K);
            }
            sto(a.oVx, o[L2X], AR::vel(vx_j, rx, ex, K));
            sto(a.oVy, o[L2Y], AR::vel(vy_j, ry, ey, K));
            if (write_res) {
                sto(a.rVx, o[L2X], rx);
                sto(a.rVy, o[L2Y], ry);
            }
        }
        P_jm1 = P;
        tyy_jm1 = tyy;
        txy_j = txy1;
        vx_j = vx1;
        vy_j = vy1;
        eta_j = eta1;
        eml_jm1 = eml;
#pragma unroll
        for (int c = 0; c < NLOC2; ++c) o[c] += more ? dy[c] : 0u;
    };
    // Two register sets used in turn, fully unrolled: no row's values are
    // copied across a loop edge (a copy would wait for the loads in flight).
    Row2<T, BF> rr[2];
    load_row2<T, FVISC, OPT, BF>(a, o, dy, has_f, has_rg, false, rr[0]);   // row jb - 1 writes nothing
#pragma unroll
    for (int t = 0; t <= kYChunk2; ++t) row(jb - 1 + t, rr[t & 1], rr[(t + 1) & 1]);
}

// ------------------------------------------------------------------ 2D, two PT iterations per pass
// The V batch a 2D free-slip solve applies after every iteration
// (stokes_2d_inc_ve_T.jl's bc of update_velocity!): Vx Dirichlet on its x
// nodes 1 / n+1 and Neumann in y, Vy Neumann in x and Dirichlet on its y
// nodes 1 / n+1, scalar values [side 1, side 2]; h the spacings.
struct FreeSlip2 {
    double ax[2], qy[2], by[2], qx[2];
    double hx, hy;
};

#ifndef CHMY_FUSED_TB
#define CHMY_FUSED_TB 1   // 0: chmy_stokes_iterate runs single iterations only
#endif
#ifndef CHMY_FUSED_TB_CHUNK
#define CHMY_FUSED_TB_CHUNK 8
#endif
#ifndef CHMY_FUSED_TB_W
#define CHMY_FUSED_TB_W 4
#endif
constexpr int kTBW = CHMY_FUSED_TB_W;       // waves per block
// output rows per wave; the march is unrolled over them, and the FP64 exact
// form (31 divisions per point) takes shorter chunks to stay unrolled
template <class T, bool EXACT>
constexpr int kTBC = (EXACT && sizeof(T) == 8) ? 8 : CHMY_FUSED_TB_CHUNK;

// What row a of the first iteration reads (Center values at row a, the rest
// at row a + 1), as fused2d_kernel's row.
template <class T, bool BF>
struct In2 {
    T vx1, vy1, P, txx, tyy, txy1, eta1, eml, rg;
    Fraw<T, BF> oxx, oyy, oxy, fP, fVx, fVy;
};

template <class T, bool FVISC, bool BF>
__device__ __forceinline__ void load_in2(const Fused2<T>& a, const uint32_t* qc, const uint32_t* qn, bool has_f,
                                         bool has_rg, In2<T, BF>& r) {
    r.vx1 = ldo(a.Vx, qn[L2X]);
    r.vy1 = ldo(a.Vy, qn[L2Y]);
    r.P = ldo(a.Pr, qc[L2C]);
    r.txx = ldo(a.txx, qc[L2C]);
    r.tyy = ldo(a.tyy, qc[L2C]);
    r.txy1 = ldo(a.txy, qn[L2XY]);
    r.oxx = ldr<T, BF>(a.Oxx, qc[L2C]);
    r.oyy = ldr<T, BF>(a.Oyy, qc[L2C]);
    r.oxy = ldr<T, BF>(a.Oxy, qn[L2XY]);
    if constexpr (FVISC) {
        r.eta1 = ldo(a.eta, qn[L2C]);
        r.eml = ldo(a.eml, qc[L2C]);
    }
    if (has_f) {
        r.fP = ldr<T, BF>(a.fP, qc[L2C]);
        r.fVx = ldr<T, BF>(a.fVx, qc[L2X]);
        r.fVy = ldr<T, BF>(a.fVy, qc[L2Y]);
    }
    if (has_rg) r.rg = ldo(a.rg, qc[L2Y]);
}

// One iteration's state carried along the y-march, and its row step: the
// expressions of fused2d_kernel's row, with V at row j computed on every lane
// (the caller keeps or drops it).
template <class T>
struct Carry2 {
    T vx_j, vy_j, eta_j, P_jm1, tyy_jm1, txy_j, eml_jm1;
};

template <class T, bool FVISC, bool EXACT>
__device__ __forceinline__ void row2(const Kc<T>& K, T ix, T iy, bool in_j, bool in_j1, bool has_f, bool has_rg,
                                     Carry2<T>& c, T vx1, T vy1, T& P, T& txx, T& tyy, T& txy1, T oxx, T oyy,
                                     T oxy, T eta1, T eml, T fP, T fVx, T fVy, T rg, T& vxn, T& vyn,
                                     T& txy_prev) {
    using AR = Arith<T, EXACT>;
    const T exx = (shdn(c.vx_j) - c.vx_j) * ix;
    const T eyy = (vy1 - c.vy_j) * iy;
    const T dv = exx + eyy;
    {
        typename AR::Loc Lc;
        if constexpr (FVISC) Lc = AR::field(c.eta_j, K);
        else Lc = AR::scalar(K);
        const T Pn = AR::pr(P, has_f ? dv - fP : dv, Lc, K);
        const T txxn = AR::diag(txx, oxx, exx, dv, Lc, K);
        const T tyyn = AR::diag(tyy, oyy, eyy, dv, Lc, K);
        P = in_j ? Pn : P;
        txx = in_j ? txxn : txx;
        tyy = in_j ? tyyn : tyy;
    }
    const T vy1l = shup(vy1);
    const T etal = FVISC ? shup(c.eta_j) : T(0), eta1l = FVISC ? shup(eta1) : T(0);
    {
        const T exy = T(0.5) * ((vx1 - c.vx_j) * iy + (vy1 - vy1l) * ix);
        typename AR::Loc Le;
        if constexpr (FVISC) Le = AR::field(lerp4(etal, c.eta_j, eta1l, eta1), K);
        else Le = AR::scalar(K);
        const T n = AR::off(txy1, oxy, exy, Le, K);
        txy1 = in_j1 ? n : txy1;
    }
    const T Pl = shup(P), txxl = shup(txx), txyr = shdn(c.txy_j), emll = FVISC ? shup(eml) : T(0);
    T rx = (-((P - Pl) * ix) + (txx - txxl) * ix) + (txy1 - c.txy_j) * iy;
    T ry = (-((P - c.P_jm1) * iy) + (tyy - c.tyy_jm1) * iy) + (txyr - c.txy_j) * ix;
    if (has_rg) ry = ry - rg;
    if (has_f) {
        rx = rx + fVx;
        ry = ry + fVy;
    }
    T ex, ey;
    if constexpr (FVISC) {
        ex = AR::vcoef_field(jmax(emll, eml), K);
        ey = AR::vcoef_field(jmax(c.eml_jm1, eml), K);
    } else {
        ex = ey = AR::vcoef_scalar(K);
    }
    vxn = AR::vel(c.vx_j, rx, ex, K);
    vyn = AR::vel(c.vy_j, ry, ey, K);
    txy_prev = c.txy_j;
    c.P_jm1 = P;
    c.tyy_jm1 = tyy;
    c.txy_j = txy1;
    c.vx_j = vx1;
    c.vy_j = vy1;
    c.eta_j = eta1;
    c.eml_jm1 = eml;
}

// Two PT iterations in one pass: the state is read once, the first iteration
// (A) lives in registers, the free-slip V batch is applied to its velocities
// in-kernel, and the second iteration (B) writes the output set.  B's row r
// needs A's stresses at row r, A's tau_xy at row r+1 and A's batched V at rows
// r, r+1 (row 0 of Vx copies row 1), so B trails A by two rows; A covers rows
// jb-2 .. je+1 of a chunk, B rows jb-1 (its first row, nothing stored) .. je-1.
// In x, A is valid on lanes 1..62 and B on lanes 2..61 (60 outputs per wave).
// Rows and columns are clamped into the stored range [-1, n+2] for loads;
// values at clamped positions reach no output.  The batched velocities follow
// the batch's own order (dims y then x; the x condition writes last): see
// bc.hip.  Results are bitwise those of two fused2d_kernel launches with the
// batch between them.
template <class T, bool FVISC, bool EXACT, int OPT, bool BF, bool RG>
__global__ void __launch_bounds__(kFX * kTBW) fused2d_tb_kernel(Fused2<T> a, FreeSlip2 fs) {
    static_assert(OPT == 0 || OPT == 1 || OPT == 2, "two-iteration pass: plain, forcing-only (rho g by RG), "
                                                    "2: plain + first V");
    static_assert(OPT < 2 || sizeof(T) == 8, "the first iteration's V is written in FP64 only");
    // WA: A's V (the first iteration's, before its V batch) is stored into
    // a.rVx/rVy at the points and rows B stores its own V
    constexpr bool HF = OPT == 1, WA = OPT == 2;
    const Kc<T>& K = a.k;
    const int lane = threadIdx.x;
    const int xt = blockIdx.x * kTBW + threadIdx.y;
    if (xt >= a.ntx) return;                        // whole wave leaves together
    const int i = xt * (kFX - 4) - 2 + lane;
    const int n0 = a.n[0], n1 = a.n[1];
    const int nb0 = n0 + 1, nb1 = n1 + 1;
    constexpr int C = kTBC<T, EXACT>;
    const int jb = blockIdx.y * C;
    const int je = min(jb + C, nb1 + 1);
    const bool inx = i >= 0 && i <= nb0;
    const bool own = lane >= 2 && lane <= kFX - 3 && inx;
    const int ic = min(max(i, -1), n0 + 2);         // stored column
    const T ix = a.ix, iy = a.iy;
    constexpr uint32_t E = sizeof(T);
    uint32_t b[NLOC2], dy[NLOC2];
#pragma unroll
    for (int c = 0; c < NLOC2; ++c) {
        b[c] = (a.L[c].o0 + (uint32_t)ic - a.L[c].sy) * E;   // row -1
        dy[c] = a.L[c].sy * E;
    }
    auto rowoff = [&](int r, uint32_t* q) {
        const uint32_t k = (uint32_t)(min(max(r, -1), n1 + 2) + 1);
#pragma unroll
        for (int c = 0; c < NLOC2; ++c) q[c] = b[c] + k * dy[c];
    };
    // batched V (after the free-slip batch) of A at row q
    const T hx = T(fs.hx), hy = T(fs.hy);
    const bool i_in = i >= 0 && i <= n0 + 2;        // inside the tangential range 0..n+2 along x
    auto vx_bc = [&](int q, T rm, T r0, T rp) -> T {
        T v = r0;
        if (q == 0 && i_in) v = fma(hy, -T(fs.qy[0]), rp);
        else if (q == n1 + 1 && i_in) v = fma(hy, T(fs.qy[1]), rm);
        if ((i == 1 || i == n0 + 1) && q >= 0 && q <= n1 + 2) v = T(fs.ax[i == 1 ? 0 : 1]);
        return v;
    };
    auto vy_bc = [&](int q, T r0) -> T {
        T v = r0;
        if (i_in && q == 1) v = T(fs.by[0]);
        else if (i_in && q == n1 + 1) v = T(fs.by[1]);
        const T dn = shdn(v), up = shup(v);
        const bool q_in = q >= 0 && q <= n1 + 2;
        if (i == 0 && q_in) return fma(hx, -T(fs.qx[0]), dn);
        if (i == n0 + 1 && q_in) return fma(hx, T(fs.qx[1]), up);
        return v;
    };

    constexpr int S = C + 4;   // A steps (rows jb-2 .. jb+C+1)
    Carry2<T> A, B;
    In2<T, BF> in[2];
    {
        uint32_t qc[NLOC2], qn[NLOC2];
        rowoff(jb - 2, qc);
        rowoff(jb - 1, qn);
        A.vx_j = ldo(a.Vx, qc[L2X]);
        A.vy_j = ldo(a.Vy, qc[L2Y]);
        A.eta_j = FVISC ? ldo(a.eta, qc[L2C]) : T(0);
        A.P_jm1 = A.tyy_jm1 = A.txy_j = A.eml_jm1 = T(0);
        load_in2<T, FVISC, BF>(a, qc, qn, HF, RG, in[0]);
    }
    B.P_jm1 = B.tyy_jm1 = B.txy_j = B.eml_jm1 = T(0);
    // A's results by step, as B reads them two steps later
    T rP[S], rtxx[S], rtyy[S], rtxy1[S], rvx[S], rvy[S];
    T roxx[S], royy[S], roxy[S], reta1[S], reml[S], rfP[S], rfVx[S], rfVy[S], rrg[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int ra = jb - 2 + t;
        if (t + 1 < S) {   // the next row's loads go out before this row's work
            uint32_t qc[NLOC2], qn[NLOC2];
            rowoff(ra + 1, qc);
            rowoff(ra + 2, qn);
            load_in2<T, FVISC, BF>(a, qc, qn, HF, RG, in[(t + 1) & 1]);
        }
        __builtin_amdgcn_sched_barrier(0);
        const In2<T, BF>& cur = in[t & 1];
        // ---- A, row ra
        {
            const bool inA = inx && ra >= 0 && ra <= nb1;
            const bool inA1 = inx && ra + 1 >= 0 && ra + 1 <= nb1;
            T P = cur.P, txx = cur.txx, tyy = cur.tyy, txy1 = cur.txy1;
            const T eta1 = FVISC ? cur.eta1 : T(0), eml = FVISC ? cur.eml : T(0);
            const T fP = HF ? cvf<T, BF>(cur.fP) : T(0);
            const T fVx = HF ? cvf<T, BF>(cur.fVx) : T(0), fVy = HF ? cvf<T, BF>(cur.fVy) : T(0);
            const T oxx = cvf<T, BF>(cur.oxx), oyy = cvf<T, BF>(cur.oyy), oxy = cvf<T, BF>(cur.oxy);
            const T rg = RG ? cur.rg : T(0);
            const T vx0 = A.vx_j, vy0 = A.vy_j;
            T vxn, vyn, txy_prev;
            row2<T, FVISC, EXACT>(K, ix, iy, inA, inA1, HF, RG, A, cur.vx1, cur.vy1, P, txx, tyy, txy1, oxx, oyy,
                                  oxy, eta1, eml, fP, fVx, fVy, rg, vxn, vyn, txy_prev);
            const bool keep = inA && t > 0;          // the first step has no velocity
            if constexpr (WA) {
                if (ra >= jb && ra < je && own && inA) {
                    uint32_t q[NLOC2];
                    rowoff(ra, q);
                    sto(a.rVx, q[L2X], vxn);
                    sto(a.rVy, q[L2Y], vyn);
                }
            }
            rP[t] = P;
            rtxx[t] = txx;
            rtyy[t] = tyy;
            rtxy1[t] = txy1;
            rvx[t] = keep ? vxn : vx0;
            rvy[t] = keep ? vyn : vy0;
            roxx[t] = oxx;
            royy[t] = oyy;
            roxy[t] = oxy;
            reta1[t] = eta1;
            reml[t] = eml;
            rfP[t] = fP;
            rfVx[t] = fVx;
            rfVy[t] = fVy;
            rrg[t] = rg;
        }
        // ---- B, row r = ra - 2
        if (t >= 3) {
            const int r = ra - 2;
            if (t == 3) {
                B.vx_j = vx_bc(r, rvx[t - 3], rvx[t - 2], rvx[t - 1]);
                B.vy_j = vy_bc(r, rvy[t - 2]);
                B.eta_j = reta1[t - 3];
            }
            const T vx1 = vx_bc(r + 1, rvx[t - 2], rvx[t - 1], rvx[t]);
            const T vy1 = vy_bc(r + 1, rvy[t - 1]);
            const int u = t - 2;
            T P = rP[u], txx = rtxx[u], tyy = rtyy[u], txy1 = rtxy1[u];
            const bool inB = inx && r >= 0 && r <= nb1;
            const bool inB1 = inx && r + 1 >= 0 && r + 1 <= nb1;
            T vxn, vyn, txy_prev;
            row2<T, FVISC, EXACT>(K, ix, iy, inB, inB1, HF, RG, B, vx1, vy1, P, txx, tyy, txy1, roxx[u], royy[u],
                                  roxy[u], reta1[u], reml[u], rfP[u], rfVx[u], rfVy[u], rrg[u], vxn, vyn, txy_prev);
            if (t >= 4 && r < je && own && inB) {
                uint32_t q[NLOC2];
                rowoff(r, q);
                sto(a.oPr, q[L2C], P);
                sto(a.otxx, q[L2C], txx);
                sto(a.otyy, q[L2C], tyy);
                sto(a.otxy, q[L2XY], txy_prev);
                sto(a.oVx, q[L2X], vxn);
                sto(a.oVy, q[L2Y], vyn);
            }
        }
    }
}

// ------------------------------------------------------------------ 3D, two PT iterations per pass
// The V batch of a 3D free-slip solve (stokes_3d_inc_ve_T.jl:130-132): each
// component Dirichlet on its own dim (vertex nodes 1 / n+1), Neumann on the
// other two; v[component][dim][side] the scalar value, h the spacings.
struct FreeSlip3 {
    double v[3][3][2];
    double h[3];
};

#ifndef CHMY_FUSED_TB3_BY64
#define CHMY_FUSED_TB3_BY64 12
#endif
#ifndef CHMY_FUSED_TB3_BY32
#define CHMY_FUSED_TB3_BY32 16
#endif
// the shortest z chunk of the two-iteration pass (the host picks the length)
#ifndef CHMY_FUSED_TB3_ZCHUNK
#define CHMY_FUSED_TB3_ZCHUNK 2
#endif
template <class T>
constexpr int kTB3Y = std::is_same<T, double>::value ? CHMY_FUSED_TB3_BY64 : CHMY_FUSED_TB3_BY32;
constexpr int kTB3Z = CHMY_FUSED_TB3_ZCHUNK;
static_assert(kTB3Z >= 2, "B's first plane of a chunk above the bottom one must be >= 1");

// x neighbours inside a row's wave by DPP whole-wave shifts: xm1 gives lane l
// the value of lane l-1, xp1 that of lane l+1 (lanes 0 / 63 receive 0, and
// reach no output).
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ double xm1(double v) {
    return __hiloint2double(dpp_shr1(__double2hiint(v)), dpp_shr1(__double2loint(v)));
}
__device__ __forceinline__ double xp1(double v) {
    return __hiloint2double(dpp_shl1(__double2hiint(v)), dpp_shl1(__double2loint(v)));
}
__device__ __forceinline__ float xm1(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
__device__ __forceinline__ float xp1(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }

// Two PT iterations in one pass over a 64 x BY tile marching z.  Thread
// (tx, ty) owns column (i, j) = (i0 - 2 + tx, j0 - 2 + ty); the row index is
// wave-uniform.  Step s:
//   A (first iteration): fused3d_kernel's step at plane s (Center stresses
//     and tau_xy at s, tau_xz / tau_yz at s+1, V at s), nothing stored; its
//     results and the eta-derived coefficients it formed stay in registers
//     for one step.
//   batch: A's V at plane s gets the free-slip batch per point (the batch's
//     order: z from A's planes s-1 / s, y through an LDS exchange of the
//     neighbouring rows, x through lane shifts; the reference's fma steps);
//     blocks and planes no rule reaches skip it.
//   B (second iteration): the same step at plane r = s-1 from A's results of
//     step s-1 and the batched V at planes r and r+1, with A's coefficients
//     (the same points, so B forms no eta_ve or eta_ml term itself); it
//     stores planes kb .. ke-1.
// The bottom chunk's B starts at plane 0 one step later, since the batch's
// Neumann ghost at plane 0 needs A's V at plane 1: at that step B also forms
// its tau_xz / tau_yz at plane 0 from A's values kept from step -1 and the
// input V at plane -1, which no rule and no iteration changes.
// A is exact on lanes 1..62 and rows 1..BY-1; B's outputs are lanes 2..61
// and rows 2..BY-2 (Center values, V) and 1..BY-3 (edges at j+1).  Rows,
// columns and planes are clamped into the stored range for loads; clamped
// values reach no output.  The LDS planes of V and of the stresses are
// shared by A and B (their uses are separated by barriers).  Results are
// bitwise those of two fused3d_kernel launches with the batch between them.
template <class T, bool FVISC, bool EXACT, int OPT, bool BF>
__global__ void __launch_bounds__(kFX * kTB3Y<T>) fused3d_tb_kernel(Fused3<T> a, FreeSlip3 fs) {
    static_assert(OPT >= 0 && OPT <= 4, "two-iteration pass: plain, forcing, rho g; 3, 4: plain, rho g + first V");
    static_assert(OPT < 3 || sizeof(T) == 8, "the first iteration's V is written in FP64 only");
    constexpr int BY = kTB3Y<T>, NR = BY + 1, TW = kFX;
    // WA: A's V (the first iteration's, before its V batch) is stored into
    // a.rVx/rVy/rVz at the cells and planes B stores its own V
    constexpr bool HF = OPT == 1, RG = OPT == 2 || OPT == 4, WA = OPT >= 3;
    using AR = Arith<T, EXACT>;
    using Loc = typename AR::Loc;
    const Kc<T>& K = a.k;
    __shared__ T s_vx[NR][TW], s_vz1[NR][TW];         // Vx at the stressed plane, Vz at the next (A, then B)
    __shared__ T s_vy[2][NR][TW], b_vy[2][NR][TW];    // Vy at the plane and the next: A's, B's
    __shared__ T s_eta[FVISC ? 2 : 1][NR][TW];       // eta at s, s+1 (FVISC; else the batch exchange)
    __shared__ T s_eml[FVISC ? BY : 1][TW];
    __shared__ T s_P[BY][TW], s_tyy[BY][TW], s_txy[BY][TW];   // the plane's stresses (A, then B)
    __shared__ T s_tyz[2][BY][TW], b_tyz[2][BY][TW];         // tau_yz at the plane and the next
    // A's stresses of the previous step, one slot per thread (B's inputs);
    // tau_yz is read from A's s_tyz plane of that step instead
    __shared__ T q_st[6][BY][TW];
    // B's tau_old (and forcing, rho g) are the values A loaded one step
    // earlier; they pass through per-thread LDS slots by step parity instead
    // of being loaded again: in FP32 all of them (bf16 forcing as its 16
    // bits), in FP64 the two the remaining LDS holds (tau_old xz and yz; the
    // block then uses all 160 KB).  FP32 forcing stored in FP32 has no room.
    constexpr bool HO = sizeof(T) == 4 && (!HF || BF);
    constexpr bool HO64 = sizeof(T) == 8;
    constexpr int NO = HO ? 6 + (HF ? 4 : RG ? 1 : 0) : 2;
    using HS = typename std::conditional<BF, uint16_t, T>::type;
    __shared__ HS q_old[HO || HO64 ? 2 : 1][NO][HO || HO64 ? BY : 1][TW];
    auto hpack = [](T v) -> HS {
        if constexpr (BF) return (uint16_t)(__float_as_uint(v) >> 16);   // exact: v came from bf16
        else return v;
    };
    auto hunpack = [](HS h) -> T {
        if constexpr (BF) return bf16_to_f32(h);
        else return h;
    };

    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int nb = a.ntx * a.nty * a.nchunks;
    const int lt = xcd_remap(blockIdx.x, nb);
    const int bx = lt % a.ntx, by = (lt / a.ntx) % a.nty, bz = lt / (a.ntx * a.nty);
    const int i0 = bx * (TW - 4) - 2, j0 = by * (BY - 3) - 2;
    const int i = i0 + tx;
    const int j = j0 + ty;
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int nb0 = n0 + 1, nb1 = n1 + 1, nb2 = n2 + 1;
    const int kb = bz * a.kchunk;
    const int ke = min(kb + a.kchunk, nb2 + 1);
    const int ic = min(max(i, -1), n0 + 2), jc = min(max(j, -1), n1 + 2);
    const bool inx = i >= 0 && i <= nb0;
    const bool inc = inx && j >= 0 && j <= nb1;
    const bool ine = inx && j + 1 >= 0 && j + 1 <= nb1;
    const bool ownx = tx >= 2 && tx <= TW - 3;
    const bool ownc = ownx && ty >= 2 && ty <= BY - 2 && inc;
    const bool owne = ownx && ty >= 1 && ty <= BY - 3 && ine;
    const bool last = ty == BY - 1;
    const bool row8 = last && jc + 1 <= n1 + 2;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    constexpr uint32_t E = sizeof(T);
    // batch rules reach this block: y rows {0, 1, n1+1}, x columns {0, 1, n0+1}
    const bool yb = j0 <= 1 || (j0 <= nb1 && nb1 <= j0 + BY - 1);
    const bool xb = i0 <= 1 || (i0 <= nb0 && nb0 <= i0 + TW - 1);

    // byte offsets at plane kb - 2 (clamped to -1); loads of plane s+1 past
    // the storage (s = n2+2) take plane s instead.  FP64: a location's offset
    // is xo + uo[c], the column part xo per lane and shared by all seven
    // locations, the row and plane part uo[c] wave-uniform (SGPRs), so the
    // offsets hold one VGPR instead of seven (the unsigned sum wraps to the
    // non-negative offset when ic = -1): VGPR spills 9 -> 2 in the fast form.
    // uo[c] is recomputed from the plane index at each step (see the loop).
    // FP32 keeps uo[c] as the whole per-lane offset (the split cost its
    // correction sweep 1%).
    constexpr bool SPLIT = sizeof(T) == 8;
    const int s0 = max(kb - 2, -1);
    uint32_t uo[NLOC], dz[NLOC], dy[NLOC], d8[NLOC];
    const uint32_t xo = SPLIT ? (uint32_t)(ic * (int)E) : 0u;
    const auto o = [&](int c) -> uint32_t { return SPLIT ? xo + uo[c] : uo[c]; };
#pragma unroll
    for (int c = 0; c < NLOC; ++c) {
        const LocStride sl = a.L[c];
        uo[c] = (uint32_t)(((int64_t)sl.o0 + (SPLIT ? 0 : ic) + (int64_t)jc * sl.sy + (int64_t)s0 * sl.sz) * E);
        dz[c] = sl.sz * E;
        dy[c] = sl.sy * E;
        d8[c] = row8 ? dy[c] : 0u;
    }
    auto ld = [](const T* p, uint32_t off) -> T { return ldo(p, off); };
    const int tx1 = tx < TW - 1 ? tx + 1 : tx, txm = tx > 0 ? tx - 1 : tx;
    const int tym = ty > 0 ? ty - 1 : ty, typ = ty + 1;   // typ <= BY: row BY exists in every NR array

    // the batch of one plane q from A's raw V at q-1 (rm), q (r0), q+1 (rp);
    // the y rule's exchange goes through two LDS planes free at that point
    const T hx = T(fs.h[0]), hy = T(fs.h[1]), hz = T(fs.h[2]);
    const bool ij_in = i >= 0 && i <= n0 + 2 && j >= 0 && j <= n1 + 2;   // z rule's tangential range
    auto batched = [&](int q, const T* rm, const T* r0, const T* rp, T (*xa)[TW], T (*xc)[TW], T* out) {
        T az[3] = {r0[0], r0[1], r0[2]};
        if (q == 0 || q == 1 || q == n2 + 1) {   // z rule
            if (ij_in) {
                if (q == 1) az[2] = T(fs.v[2][2][0]);
                if (q == n2 + 1) az[2] = T(fs.v[2][2][1]);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if (q == 0) az[c] = fma(hz, -T(fs.v[c][2][0]), rp[c]);
                    if (q == n2 + 1) az[c] = fma(hz, T(fs.v[c][2][1]), rm[c]);
                }
            }
        }
        T ay[3] = {az[0], az[1], az[2]};
        if (yb) {   // y rule (block-uniform branch around the barrier)
            xa[ty][tx] = az[0];
            xc[ty][tx] = az[2];
            __syncthreads();
            const int tyx = ty + 1 < BY ? ty + 1 : ty;   // the last row is past every output
            if (i >= 0 && i <= n0 + 2 && q >= 0 && q <= n2 + 2) {
                if (j == 0) {
                    ay[0] = fma(hy, -T(fs.v[0][1][0]), xa[tyx][tx]);
                    ay[2] = fma(hy, -T(fs.v[2][1][0]), xc[tyx][tx]);
                } else if (j == n1 + 1) {
                    ay[0] = fma(hy, T(fs.v[0][1][1]), xa[tym][tx]);
                    ay[2] = fma(hy, T(fs.v[2][1][1]), xc[tym][tx]);
                }
                if (j == 1) ay[1] = T(fs.v[1][1][0]);
                else if (j == n1 + 1) ay[1] = T(fs.v[1][1][1]);
            }
        }
        out[0] = ay[0];
        out[1] = ay[1];
        out[2] = ay[2];
        if (xb) {   // x rule: Vx Dirichlet on its x nodes, Vy and Vz Neumann from the neighbouring lane
            const bool jq_in = j >= 0 && j <= n1 + 2 && q >= 0 && q <= n2 + 2;
            if (jq_in && i == 1) out[0] = T(fs.v[0][0][0]);
            else if (jq_in && i == n0 + 1) out[0] = T(fs.v[0][0][1]);
#pragma unroll
            for (int c = 1; c < 3; ++c) {
                const T dn = xp1(ay[c]), up = xm1(ay[c]);
                if (jq_in && i == 0) out[c] = fma(hx, -T(fs.v[c][0][0]), dn);
                else if (jq_in && i == n0 + 1) out[c] = fma(hx, T(fs.v[c][0][1]), up);
            }
        }
    };

    // ---- A: carried state at plane s0
    int pa = 0, pb = 0;
    T Avx = ld(a.Vx, o(LVX)), Avy = ld(a.Vy, o(LVY)), Avz = ld(a.Vz, o(LVZ));
    T Aeta = T(0), Aeml_m = T(0), AP_m = T(0), Atzz_m = T(0), Atxz = T(0), Atyz = T(0);
    s_vy[0][ty][tx] = Avy;
    if constexpr (FVISC) {
        Aeta = ld(a.eta, o(LC));
        s_eta[0][ty][tx] = Aeta;
    }
    if (last) {
        s_vy[0][BY][tx] = ld(a.Vy, o(LVY) + d8[LVY]);
        if constexpr (FVISC) s_eta[0][BY][tx] = ld(a.eta, o(LC) + d8[LC]);
    }
    // A's results of the previous step (plane s-1), for B (its stresses are in q_st)
    T qv[3] = {T(0), T(0), T(0)};   // A's raw V at s-1
    Loc qLc{}, qLe{}, qL4{}, qL5{};
    T qex = T(0), qey = T(0), qez = T(0);
    // B's carried state: batched V at r, Center values at r-1, tau_xz / tau_yz at r
    T Bv[3] = {T(0), T(0), T(0)};
    T BP_m = T(0), Btzz_m = T(0), Btxz = T(0), Btyz = T(0);
    const int sb = kb == 0 ? 1 : kb;   // B's first step (plane sb - 1)

    uint32_t ub[NLOC];   // uo at plane 0 (modulo 2^32)
#pragma unroll
    for (int c = 0; c < NLOC; ++c) ub[c] = uo[c] - (uint32_t)s0 * dz[c];
    for (int s = s0; s <= ke; ++s) {
        if constexpr (SPLIT) {
            // formed from the plane index through readfirstlane, so that the
            // compiler keeps one offset per location rather than one induction
            // variable per derived offset (o + dy, o - dz, ...): SGPR spill
            // reloads in the loop 546 -> 132, pass 1.257 -> 1.230 ms at 255^3
#pragma unroll
            for (int c = 0; c < NLOC; ++c) uo[c] = __builtin_amdgcn_readfirstlane(ub[c] + (uint32_t)s * dz[c]);
        }
        const bool top = s + 1 > n2 + 2;   // plane s+1 is past the storage
        uint32_t z1[NLOC];
#pragma unroll
        for (int c = 0; c < NLOC; ++c) z1[c] = top ? o(c) : o(c) + dz[c];
        // ================= A: plane s
        T va[3], vx1, vy1, vz1, P, txx, tyy, tzz, txy1, txz1, tyz1;
        Loc Lc, Le, L4, L5;
        T ex, ey, ez;
        {
            // the staging wave's row j+1 first: loads complete in order, and its LDS
            // writes (before the first barrier) then need not wait for the rest
            T vx8 = T(0), vy8 = T(0), vz8 = T(0), eta8 = T(0);
            if (last) {
                vx8 = ld(a.Vx, o(LVX) + d8[LVX]);
                vy8 = ld(a.Vy, z1[LVY] + d8[LVY]);
                vz8 = ld(a.Vz, z1[LVZ] + d8[LVZ]);
                if constexpr (FVISC) eta8 = ld(a.eta, z1[LC] + d8[LC]);
            }
            vx1 = ld(a.Vx, z1[LVX]);
            vy1 = ld(a.Vy, z1[LVY]);
            vz1 = ld(a.Vz, z1[LVZ]);
            T eta1 = T(0), eml = T(0);
            if constexpr (FVISC) {
                eta1 = ld(a.eta, z1[LC]);
                eml = ld(a.eml, o(LC));
            }
            P = ld(a.Pr, o(LC));
            txx = ld(a.txx, o(LC));
            tyy = ld(a.tyy, o(LC));
            tzz = ld(a.tzz, o(LC));
            txy1 = ld(a.txy, o(LXY) + dy[LXY]);
            txz1 = ld(a.txz, z1[LXZ]);
            tyz1 = ld(a.tyz, z1[LYZ] + dy[LYZ]);
            const T oxx = ldf<T, BF>(a.Oxx, o(LC)), oyy = ldf<T, BF>(a.Oyy, o(LC)), ozz = ldf<T, BF>(a.Ozz, o(LC));
            const T oxy = ldf<T, BF>(a.Oxy, o(LXY) + dy[LXY]);
            const T oxz = ldf<T, BF>(a.Oxz, z1[LXZ]), oyz = ldf<T, BF>(a.Oyz, z1[LYZ] + dy[LYZ]);
            T fP = T(0), fVx = T(0), fVy = T(0), fVz = T(0);
            if constexpr (HF) {
                fP = ldf<T, BF>(a.fP, o(LC));
                fVx = ldf<T, BF>(a.fVx, o(LVX));
                fVy = ldf<T, BF>(a.fVy, o(LVY));
                fVz = ldf<T, BF>(a.fVz, o(LVZ));
            }
            const T rgz = RG ? ld(a.rg, o(LVZ)) : T(0);
            const bool in_s = s >= 0 && s <= nb2;
            const bool in_s1 = s + 1 >= 0 && s + 1 <= nb2;
            s_vx[ty][tx] = Avx;
            s_vy[pa ^ 1][ty][tx] = vy1;
            s_vz1[ty][tx] = vz1;
            if constexpr (FVISC) {
                s_eta[pa ^ 1][ty][tx] = eta1;
                s_eml[ty][tx] = eml;
            }
            if (last) {
                s_vx[BY][tx] = vx8;
                s_vy[pa ^ 1][BY][tx] = vy8;
                s_vz1[BY][tx] = vz8;
                if constexpr (FVISC) s_eta[pa ^ 1][BY][tx] = eta8;
            }
            __syncthreads();
            const T exx = (xp1(Avx) - Avx) * ix;
            const T eyy = (s_vy[pa][typ][tx] - Avy) * iy;
            const T ezz = (vz1 - Avz) * iz;
            const T dv = (exx + eyy) + ezz;
            if constexpr (FVISC) Lc = AR::field(Aeta, K);
            else Lc = AR::scalar(K);
            {
                const T Pn = AR::pr(P, HF ? dv - fP : dv, Lc, K);
                const T txxn = AR::diag(txx, oxx, exx, dv, Lc, K);
                const T tyyn = AR::diag(tyy, oyy, eyy, dv, Lc, K);
                const T tzzn = AR::diag(tzz, ozz, ezz, dv, Lc, K);
                const bool u = inc && in_s;
                P = u ? Pn : P;
                txx = u ? txxn : txx;
                tyy = u ? tyyn : tyy;
                tzz = u ? tzzn : tzz;
            }
            {   // tau_xy at (i, j+1, s)
                const T exy = T(0.5) * ((s_vx[typ][tx] - Avx) * iy + (s_vy[pa][typ][tx] - s_vy[pa][typ][txm]) * ix);
                if constexpr (FVISC)
                    Le = AR::field(lerp4(s_eta[pa][ty][txm], Aeta, s_eta[pa][typ][txm], s_eta[pa][typ][tx]), K);
                else Le = AR::scalar(K);
                const T nxy = AR::off(txy1, oxy, exy, Le, K);
                txy1 = ine && in_s ? nxy : txy1;
            }
            {   // tau_xz at (i, j, s+1) and tau_yz at (i, j+1, s+1)
                const T exz = T(0.5) * ((vx1 - Avx) * iz + (vz1 - xm1(vz1)) * ix);
                const T eyz = T(0.5) * ((s_vy[pa ^ 1][typ][tx] - s_vy[pa][typ][tx]) * iz + (s_vz1[typ][tx] - vz1) * iy);
                T e4 = T(0), e5 = T(0);
                if constexpr (FVISC) {
                    e4 = lerp4(s_eta[pa][ty][txm], Aeta, s_eta[pa ^ 1][ty][txm], eta1);
                    e5 = lerp4(Aeta, s_eta[pa][typ][tx], eta1, s_eta[pa ^ 1][typ][tx]);
                    L4 = AR::field(e4, K);
                    L5 = AR::field(e5, K);
                } else {
                    L4 = L5 = AR::scalar(K);
                }
                const T nxz = AR::off(txz1, oxz, exz, L4, K);
                const T nyz = AR::off(tyz1, oyz, eyz, L5, K);
                txz1 = inc && in_s1 ? nxz : txz1;
                tyz1 = ine && in_s1 ? nyz : tyz1;
                if (kb == 0 && s == -1) {
                    // plane 0's tau_xz / tau_yz and eta at their points, for the bottom
                    // chunk's B at step 1, in planes B leaves untouched until then
                    b_tyz[0][ty][tx] = txz1;
                    b_tyz[1][ty][tx] = tyz1;
                    b_vy[0][ty][tx] = e4;
                    b_vy[1][ty][tx] = e5;
                }
            }
            s_P[ty][tx] = P;
            s_tyy[ty][tx] = tyy;
            s_txy[ty][tx] = txy1;
            s_tyz[pa][ty][tx] = tyz1;
            if constexpr (HO64) {
                q_old[pa][0][ty][tx] = oxz;
                q_old[pa][1][ty][tx] = oyz;
            }
            if constexpr (HO) {
                q_old[pa][0][ty][tx] = hpack(oxx);
                q_old[pa][1][ty][tx] = hpack(oyy);
                q_old[pa][2][ty][tx] = hpack(ozz);
                q_old[pa][3][ty][tx] = hpack(oxy);
                q_old[pa][4][ty][tx] = hpack(oxz);
                q_old[pa][5][ty][tx] = hpack(oyz);
                if constexpr (HF) q_old[pa][6][ty][tx] = hpack(fP);
            }
            const T txxl = xm1(txx), txzr = xp1(Atxz);
            __syncthreads();
            // A's V at plane s
            const T txy_j = s_txy[tym][tx];
            T rx = ((-((P - s_P[ty][txm]) * ix) + (txx - txxl) * ix) + (txy1 - txy_j) * iy) + (txz1 - Atxz) * iz;
            T ry = ((-((P - s_P[tym][tx]) * iy) + (tyy - s_tyy[tym][tx]) * iy) + (s_txy[tym][tx1] - txy_j) * ix) +
                   (s_tyz[pa][tym][tx] - s_tyz[pa ^ 1][tym][tx]) * iz;
            T rz = ((-((P - AP_m) * iz) + (tzz - Atzz_m) * iz) + (txzr - Atxz) * ix) +
                   (Atyz - s_tyz[pa ^ 1][tym][tx]) * iy;
            if constexpr (RG) rz = rz - rgz;
            if constexpr (HF) {
                rx = rx + fVx;
                ry = ry + fVy;
                rz = rz + fVz;
            }
            if constexpr (HO && HF) {
                q_old[pa][7][ty][tx] = hpack(fVx);
                q_old[pa][8][ty][tx] = hpack(fVy);
                q_old[pa][9][ty][tx] = hpack(fVz);
            }
            if constexpr (HO && RG) q_old[pa][6][ty][tx] = hpack(rgz);
            if constexpr (FVISC) {
                ex = AR::vcoef_field(jmax(xm1(eml), eml), K);
                ey = AR::vcoef_field(jmax(s_eml[tym][tx], eml), K);
                ez = AR::vcoef_field(jmax(Aeml_m, eml), K);
            } else {
                ex = ey = ez = AR::vcoef_scalar(K);
            }
            const bool keep = inc && in_s && s > s0;   // the first plane has no velocity
            const T vxa = keep ? AR::vel(Avx, rx, ex, K) : Avx;
            const T vya = keep ? AR::vel(Avy, ry, ey, K) : Avy;
            const T vza = keep ? AR::vel(Avz, rz, ez, K) : Avz;
            if constexpr (WA) {
                if (s >= kb && s < ke && ownc) {
                    sto(a.rVx, o(LVX), vxa);
                    sto(a.rVy, o(LVY), vya);
                    sto(a.rVz, o(LVZ), vza);
                }
            }
            // A's carried state for step s+1
            AP_m = P;
            Atzz_m = tzz;
            Atxz = txz1;
            Atyz = tyz1;
            Avx = vx1;
            Avy = vy1;
            Avz = vz1;
            if constexpr (FVISC) {
                Aeta = eta1;
                Aeml_m = eml;
            }
            va[0] = vxa;
            va[1] = vya;
            va[2] = vza;
        }
        // ================= B: plane r = s - 1
        const int r = s - 1;
        if (s >= sb) {
            // B's inputs are A's stresses of the previous step; this step's take their slots
            T bP = q_st[0][ty][tx], btxx = q_st[1][ty][tx], btyy = q_st[2][ty][tx], btzz = q_st[3][ty][tx];
            T btxy = q_st[4][ty][tx], btxz = q_st[5][ty][tx], btyz = s_tyz[pa ^ 1][ty][tx];
            q_st[0][ty][tx] = P;
            q_st[1][ty][tx] = txx;
            q_st[2][ty][tx] = tyy;
            q_st[3][ty][tx] = tzz;
            q_st[4][ty][tx] = txy1;
            q_st[5][ty][tx] = txz1;
            T V1[3];
            if (s == sb) {   // B's first step: V after the batch at plane r as well
                if (kb == 0) {
                    const T ztxz = b_tyz[0][ty][tx], ztyz = b_tyz[1][ty][tx];   // A's, plane 0
                    Loc zL4, zL5;
                    if constexpr (FVISC) {
                        zL4 = AR::field(b_vy[0][ty][tx], K);
                        zL5 = AR::field(b_vy[1][ty][tx], K);
                    } else {
                        zL4 = zL5 = AR::scalar(K);
                    }
                    // plane 0 from A's V at planes 0 (previous step) and 1 (this step)
                    batched(0, qv, qv, va, s_eta[FVISC ? pa : 0], s_vy[pa], Bv);
                    __syncthreads();
                    b_vy[pb][ty][tx] = Bv[1];
                    s_vz1[ty][tx] = Bv[2];   // Vz(0) for its row j+1 neighbour
                    __syncthreads();
                    // B's tau_xz at (i, j, 0) and tau_yz at (i, j+1, 0); V at plane -1 is the
                    // input, since no iteration and no rule changes that plane
                    const T vxm = ld(a.Vx, o(LVX) - 2 * dz[LVX]);
                    const T vym = ld(a.Vy, o(LVY) - 2 * dz[LVY] + dy[LVY]);
                    const T oxz = ldf<T, BF>(a.Oxz, o(LXZ) - dz[LXZ]);
                    const T oyz = ldf<T, BF>(a.Oyz, o(LYZ) - dz[LYZ] + dy[LYZ]);
                    BP_m = ld(a.Pr, o(LC) - 2 * dz[LC]);
                    Btzz_m = ld(a.tzz, o(LC) - 2 * dz[LC]);
                    const T exz = T(0.5) * ((Bv[0] - vxm) * iz + (Bv[2] - xm1(Bv[2])) * ix);
                    const T eyz = T(0.5) * ((b_vy[pb][typ][tx] - vym) * iz + (s_vz1[typ][tx] - Bv[2]) * iy);
                    Btxz = inc ? AR::off(ztxz, oxz, exz, zL4, K) : ztxz;
                    Btyz = ine ? AR::off(ztyz, oyz, eyz, zL5, K) : ztyz;
                    __syncthreads();
                    b_tyz[pb ^ 1][ty][tx] = Btyz;
                } else {
                    batched(r, qv, qv, qv, s_eta[FVISC ? pa : 0], s_vy[pa], Bv);
                    b_vy[pb][ty][tx] = Bv[1];
                    __syncthreads();   // the exchange planes are used again below
                }
            }
            batched(r + 1, qv, va, va, s_eta[FVISC ? pa : 0], s_vy[pa], V1);
            s_vx[ty][tx] = Bv[0];
            b_vy[pb ^ 1][ty][tx] = V1[1];
            s_vz1[ty][tx] = V1[2];
            // tau_old (and forcing) of plane r (Center, xy) and r+1 (xz, yz): o[] sits at
            // plane s; with HO the slots A filled one step earlier
            const uint32_t oC = o(LC) - dz[LC];
            T oxx, oyy, ozz, oxy, oxz, oyz, fP = T(0), fVx = T(0), fVy = T(0), fVz = T(0), rgz = T(0);
            if constexpr (HO) {
                const int qb = pa ^ 1;
                oxx = hunpack(q_old[qb][0][ty][tx]);
                oyy = hunpack(q_old[qb][1][ty][tx]);
                ozz = hunpack(q_old[qb][2][ty][tx]);
                oxy = hunpack(q_old[qb][3][ty][tx]);
                oxz = hunpack(q_old[qb][4][ty][tx]);
                oyz = hunpack(q_old[qb][5][ty][tx]);
                if constexpr (HF) {
                    fP = hunpack(q_old[qb][6][ty][tx]);
                    fVx = hunpack(q_old[qb][7][ty][tx]);
                    fVy = hunpack(q_old[qb][8][ty][tx]);
                    fVz = hunpack(q_old[qb][9][ty][tx]);
                }
                if constexpr (RG) rgz = hunpack(q_old[qb][6][ty][tx]);
            } else {
                oxx = ldf<T, BF>(a.Oxx, oC);
                oyy = ldf<T, BF>(a.Oyy, oC);
                ozz = ldf<T, BF>(a.Ozz, oC);
                oxy = ldf<T, BF>(a.Oxy, o(LXY) - dz[LXY] + dy[LXY]);
                if constexpr (HO64) {
                    oxz = q_old[pa ^ 1][0][ty][tx];
                    oyz = q_old[pa ^ 1][1][ty][tx];
                } else {
                    oxz = ldf<T, BF>(a.Oxz, o(LXZ));
                    oyz = ldf<T, BF>(a.Oyz, o(LYZ) + dy[LYZ]);
                }
                if constexpr (HF) {
                    fP = ldf<T, BF>(a.fP, oC);
                    fVx = ldf<T, BF>(a.fVx, o(LVX) - dz[LVX]);
                    fVy = ldf<T, BF>(a.fVy, o(LVY) - dz[LVY]);
                    fVz = ldf<T, BF>(a.fVz, o(LVZ) - dz[LVZ]);
                }
                if constexpr (RG) rgz = ld(a.rg, o(LVZ) - dz[LVZ]);
            }
            __syncthreads();
            const bool in_k = r >= 0 && r <= nb2;
            const bool in_k1 = r + 1 >= 0 && r + 1 <= nb2;
            const T exx = (xp1(Bv[0]) - Bv[0]) * ix;
            const T eyy = (b_vy[pb][typ][tx] - Bv[1]) * iy;
            const T ezz = (V1[2] - Bv[2]) * iz;
            const T dv = (exx + eyy) + ezz;
            {
                const T Pn = AR::pr(bP, HF ? dv - fP : dv, qLc, K);
                const T txxn = AR::diag(btxx, oxx, exx, dv, qLc, K);
                const T tyyn = AR::diag(btyy, oyy, eyy, dv, qLc, K);
                const T tzzn = AR::diag(btzz, ozz, ezz, dv, qLc, K);
                const bool u = inc && in_k;
                bP = u ? Pn : bP;
                btxx = u ? txxn : btxx;
                btyy = u ? tyyn : btyy;
                btzz = u ? tzzn : btzz;
            }
            {
                const T exy = T(0.5) * ((s_vx[typ][tx] - Bv[0]) * iy + (b_vy[pb][typ][tx] - b_vy[pb][typ][txm]) * ix);
                const T nxy = AR::off(btxy, oxy, exy, qLe, K);
                btxy = ine && in_k ? nxy : btxy;
            }
            {
                const T exz = T(0.5) * ((V1[0] - Bv[0]) * iz + (V1[2] - xm1(V1[2])) * ix);
                const T eyz = T(0.5) * ((b_vy[pb ^ 1][typ][tx] - b_vy[pb][typ][tx]) * iz + (s_vz1[typ][tx] - V1[2]) * iy);
                const T nxz = AR::off(btxz, oxz, exz, qL4, K);
                const T nyz = AR::off(btyz, oyz, eyz, qL5, K);
                btxz = inc && in_k1 ? nxz : btxz;
                btyz = ine && in_k1 ? nyz : btyz;
            }
            const bool st = r >= kb && r < ke;
            if (st) {
                if (ownc) {
                    sto(a.oPr, oC, bP);
                    sto(a.otxx, oC, btxx);
                    sto(a.otyy, oC, btyy);
                    sto(a.otzz, oC, btzz);
                    sto(a.otxz, o(LXZ) - dz[LXZ], Btxz);
                }
                if (owne) {
                    sto(a.otxy, o(LXY) - dz[LXY] + dy[LXY], btxy);
                    sto(a.otyz, o(LYZ) - dz[LYZ] + dy[LYZ], Btyz);
                }
            }
            const T txxl = xm1(btxx), txzr = xp1(Btxz);
            s_P[ty][tx] = bP;
            s_tyy[ty][tx] = btyy;
            s_txy[ty][tx] = btxy;
            b_tyz[pb][ty][tx] = btyz;
            __syncthreads();
            const T txy_j = s_txy[tym][tx];
            T rx = ((-((bP - s_P[ty][txm]) * ix) + (btxx - txxl) * ix) + (btxy - txy_j) * iy) + (btxz - Btxz) * iz;
            T ry = ((-((bP - s_P[tym][tx]) * iy) + (btyy - s_tyy[tym][tx]) * iy) + (s_txy[tym][tx1] - txy_j) * ix) +
                   (b_tyz[pb][tym][tx] - b_tyz[pb ^ 1][tym][tx]) * iz;
            T rz = ((-((bP - BP_m) * iz) + (btzz - Btzz_m) * iz) + (txzr - Btxz) * ix) +
                   (Btyz - b_tyz[pb ^ 1][tym][tx]) * iy;
            if constexpr (RG) rz = rz - rgz;
            if constexpr (HF) {
                rx = rx + fVx;
                ry = ry + fVy;
                rz = rz + fVz;
            }
            if (st && ownc) {
                sto(a.oVx, o(LVX) - dz[LVX], AR::vel(Bv[0], rx, qex, K));
                sto(a.oVy, o(LVY) - dz[LVY], AR::vel(Bv[1], ry, qey, K));
                sto(a.oVz, o(LVZ) - dz[LVZ], AR::vel(Bv[2], rz, qez, K));
            }
            BP_m = bP;
            Btzz_m = btzz;
            Btxz = btxz;
            Btyz = btyz;
            Bv[0] = V1[0];
            Bv[1] = V1[1];
            Bv[2] = V1[2];
            pb ^= 1;
        } else {
            __syncthreads();   // A's next step writes the eta_ml plane its velocity just read
            q_st[0][ty][tx] = P;   // A's results of plane s for B's next step
            q_st[1][ty][tx] = txx;
            q_st[2][ty][tx] = tyy;
            q_st[3][ty][tx] = tzz;
            q_st[4][ty][tx] = txy1;
            q_st[5][ty][tx] = txz1;
        }
        qv[0] = va[0];
        qv[1] = va[1];
        qv[2] = va[2];
        qLc = Lc;
        qLe = Le;
        qL4 = L4;
        qL5 = L5;
        qex = ex;
        qey = ey;
        qez = ez;
        pa ^= 1;
        if constexpr (!SPLIT)
#pragma unroll
            for (int c = 0; c < NLOC; ++c) uo[c] += dz[c];
    }
}

#ifndef CHMY_FUSED_ZCHUNK64
#define CHMY_FUSED_ZCHUNK64 16
#endif
#ifndef CHMY_FUSED_ZCHUNK32
#define CHMY_FUSED_ZCHUNK32 12
#endif
template <class T>
constexpr int kZChunk = std::is_same<T, double>::value ? CHMY_FUSED_ZCHUNK64 : CHMY_FUSED_ZCHUNK32;
template <class T>
constexpr int kBY3 = std::is_same<T, double>::value ? CHMY_FUSED_BY64 : CHMY_FUSED_BY32;
template <class T>
constexpr int kTW3 = std::is_same<T, double>::value ? CHMY_FUSED_TW64 : CHMY_FUSED_TW32;

// Planes per block of the two-iteration pass.  One block runs per CU at a
// time (LDS), so a launch of B blocks takes ceil(B / CUs) rounds, each as
// long as one block's march: planes + 3 first-iteration steps.  The chunk
// count minimising rounds x steps wins (255^3 on 256 CUs: 7 chunks of 37
// planes, 1015 blocks in 4 rounds; measured ms per iteration at chunk
// lengths 16 / 22 / 24 / 26 / 37 / 43 / 52: 0.784 / 0.739 / 0.775 / 0.734 /
// 0.694 / 0.760 / 0.705, the order this cost gives).
static int32_t tb3_chunk(int64_t tiles, int64_t planes) {
    static int ncu[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (ncu[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        ncu[dev] = n;
    }
    int64_t best = -1, bestk = planes;
    for (int64_t c = 1; c <= planes; ++c) {
        const int64_t k = (planes + c - 1) / c;
        if (k < kTB3Z) break;
        const int64_t cost = ((tiles * c + ncu[dev] - 1) / ncu[dev]) * (k + 3);
        if (best < 0 || cost < best) {
            best = cost;
            bestk = k;
        }
    }
    return (int32_t)bestk;
}

template <class T>
static Kc<T> make_consts(double eta_s, double eve_s, double G, double dt, double dPr, double dr, double nu) {
    Kc<T> k;
    k.Gs = T(G);
    k.dts = T(dt);
    k.Gdt = T(G) * T(dt);
    k.dPr = T(dPr);
    k.dr = T(dr);
    k.nu = T(nu);
    k.eta_s = T(eta_s);
    k.eve_s = T(eve_s);
    k.iGdt = T(1.0 / (G * dt));
    k.omdr = T(1.0 - dr);
    k.third = T(1.0 / 3.0);
    k.a_s = T(eve_s * dr);
    k.b_s = T(eve_s * dPr);
    k.cv_s = T(nu / eve_s);
    return k;
}



static LocStride loc_stride(const chmy_field_t& f) {
    const int64_t p0 = pad_of(f, 0), p1 = pad_of(f, 1);
    LocStride s;
    s.sy = (uint32_t)p0;
    s.sz = (uint32_t)(p0 * p1);
    int64_t o0 = (2 * f.halo[0] - 1) + (2 * f.halo[1] - 1) * p0;
    if (f.nd >= 3) o0 += (2 * f.halo[2] - 1) * p0 * p1;
    s.o0 = (uint32_t)o0;
    return s;
}

template <class T>
static int fused_launch(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau, const chmy_field_t* Pr,
                        const chmy_field_t* Vo, const chmy_field_t* tauo, const chmy_field_t* Pro,
                        const chmy_field_t* told, const chmy_field_t* eta, const chmy_field_t* eml,
                        const chmy_field_t* rg, const chmy_field_t* fV, const chmy_field_t* fP,
                        const chmy_field_t* dV, const chmy_field_t* rV, int wres, bool exact,
                        double eta_s, double eve_s, double G, double dt, double dPr, double dr, double nu,
                        hipStream_t s, const FreeSlip2* tb = nullptr, const FreeSlip3* tb3 = nullptr,
                        const chmy_field_t* vfirst = nullptr) {
    const bool fv = eta != nullptr;
    const bool bf = told[0].dtype == CHMY_BF16;   // validated: FP32 sweep with bf16 forcing
    // every byte offset a thread forms must fit in 32 bits
    {
        const int nd = g->nd;
        const chmy_field_t* all[] = {&V[0], &V[1], nd == 3 ? &V[2] : Pr, &tau[nd == 3 ? 3 : 2],
                                     &tau[nd == 3 ? 4 : 2], &tau[nd == 3 ? 5 : 2], Pr};
        for (const chmy_field_t* f : all) {
            int64_t el = 1;
            for (int d = 0; d < nd; ++d) el *= pad_of(*f, d);
            if (el * (int64_t)sizeof(T) >= (int64_t)UINT32_MAX)
                return set_error(CHMY_E_SIZE, "fused iteration: a field exceeds 4 GiB (32-bit offsets)");
        }
    }
    if (g->nd == 3) {
        Fused3<T> a;
        std::memset(&a, 0, sizeof a);
        auto P = [](const chmy_field_t& f) { return static_cast<T*>(f.data); };
        a.Vx = P(V[0]); a.Vy = P(V[1]); a.Vz = P(V[2]); a.Pr = P(*Pr);
        a.txx = P(tau[0]); a.tyy = P(tau[1]); a.tzz = P(tau[2]); a.txy = P(tau[3]); a.txz = P(tau[4]); a.tyz = P(tau[5]);
        a.oVx = P(Vo[0]); a.oVy = P(Vo[1]); a.oVz = P(Vo[2]); a.oPr = P(*Pro);
        a.otxx = P(tauo[0]); a.otyy = P(tauo[1]); a.otzz = P(tauo[2]);
        a.otxy = P(tauo[3]); a.otxz = P(tauo[4]); a.otyz = P(tauo[5]);
        a.Oxx = P(told[0]); a.Oyy = P(told[1]); a.Ozz = P(told[2]);
        a.Oxy = P(told[3]); a.Oxz = P(told[4]); a.Oyz = P(told[5]);
        a.eta = fv ? P(*eta) : a.Pr;
        a.eml = fv ? P(*eml) : a.Pr;
        a.rg = rg ? P(*rg) : a.Pr;
        a.dV = wres ? P(*dV) : a.oPr;
        a.rVx = wres ? P(rV[0]) : a.oVx;
        a.rVy = wres ? P(rV[1]) : a.oVy;
        a.rVz = wres ? P(rV[2]) : a.oVz;
        a.has_f = fV != nullptr;
        a.fVx = fV ? P(fV[0]) : a.Pr;
        a.fVy = fV ? P(fV[1]) : a.Pr;
        a.fVz = fV ? P(fV[2]) : a.Pr;
        a.fP = fV ? P(*fP) : a.Pr;
        a.L[LC] = loc_stride(*Pr);
        a.L[LVX] = loc_stride(V[0]);
        a.L[LVY] = loc_stride(V[1]);
        a.L[LVZ] = loc_stride(V[2]);
        a.L[LXY] = loc_stride(tau[3]);
        a.L[LXZ] = loc_stride(tau[4]);
        a.L[LYZ] = loc_stride(tau[5]);
        for (int d = 0; d < 3; ++d) a.n[d] = (int32_t)g->n[d];
        if (tb3) {   // two iterations, the free-slip batch between them
            a.has_rg = rg != nullptr;
            a.write_res = 0;
            a.ix = T(g->inv_spacing[0]);
            a.iy = T(g->inv_spacing[1]);
            a.iz = T(g->inv_spacing[2]);
            a.k = make_consts<T>(eta_s, eve_s, G, dt, dPr, dr, nu);
            a.ntx = (int32_t)((g->n[0] + 2 + kFX - 5) / (kFX - 4));
            a.nty = (int32_t)((g->n[1] + 2 + kTB3Y<T> - 4) / (kTB3Y<T> - 3));
            a.kchunk = tb3_chunk((int64_t)a.ntx * a.nty, g->n[2] + 2);
            a.nchunks = (int32_t)((g->n[2] + 2 + a.kchunk - 1) / a.kchunk);
            const unsigned nb = (unsigned)a.ntx * a.nty * a.nchunks;
            dim3 blk(kFX, kTB3Y<T>, 1), grd(nb, 1, 1);
            const bool hf = a.has_f, hr = a.has_rg;
            if (vfirst) {   // validated: FP64, no forcing
                if constexpr (std::is_same<T, double>::value) {
                    a.rVx = P(vfirst[0]);
                    a.rVy = P(vfirst[1]);
                    a.rVz = P(vfirst[2]);
#define CHMY_TB3F(FV, EX)                                                                              \
    do {                                                                                               \
        if (hr) hipLaunchKernelGGL((fused3d_tb_kernel<double, FV, EX, 4, false>), grd, blk, 0, s, a, *tb3); \
        else hipLaunchKernelGGL((fused3d_tb_kernel<double, FV, EX, 3, false>), grd, blk, 0, s, a, *tb3);    \
    } while (0)
                    if (fv && exact) CHMY_TB3F(true, true);
                    else if (fv) CHMY_TB3F(true, false);
                    else if (exact) CHMY_TB3F(false, true);
                    else CHMY_TB3F(false, false);
#undef CHMY_TB3F
                    CHMY_HIP(hipGetLastError());
                    return 0;
                } else {
                    return set_error(CHMY_E_DTYPE, "the first iteration's V is written in FP64 only");
                }
            }
#define CHMY_TB3(FV, EX)                                                                                    \
    do {                                                                                                    \
        constexpr bool F32 = std::is_same<T, float>::value;                                                 \
        if (hf && bf) hipLaunchKernelGGL((fused3d_tb_kernel<T, FV, EX, 1, F32>), grd, blk, 0, s, a, *tb3);   \
        else if (hf) hipLaunchKernelGGL((fused3d_tb_kernel<T, FV, EX, 1, false>), grd, blk, 0, s, a, *tb3);  \
        else if (hr) hipLaunchKernelGGL((fused3d_tb_kernel<T, FV, EX, 2, false>), grd, blk, 0, s, a, *tb3);  \
        else hipLaunchKernelGGL((fused3d_tb_kernel<T, FV, EX, 0, false>), grd, blk, 0, s, a, *tb3);          \
    } while (0)
            if (fv && exact) CHMY_TB3(true, true);
            else if (fv) CHMY_TB3(true, false);
            else if (exact) CHMY_TB3(false, true);
            else CHMY_TB3(false, false);
#undef CHMY_TB3
            CHMY_HIP(hipGetLastError());
            return 0;
        }
        a.kchunk = kZChunk<T>;
        a.ntx = (int32_t)((g->n[0] + 2 + kTW3<T> - 3) / (kTW3<T> - 2));
        a.nty = (int32_t)((g->n[1] + 2 + kBY3<T> - 2) / (kBY3<T> - 1));
        a.nchunks = (int32_t)((g->n[2] + 2 + kZChunk<T> - 1) / kZChunk<T>);
        a.has_rg = rg != nullptr;
        a.write_res = wres;
        a.ix = T(g->inv_spacing[0]);
        a.iy = T(g->inv_spacing[1]);
        a.iz = T(g->inv_spacing[2]);
        a.k = make_consts<T>(eta_s, eve_s, G, dt, dPr, dr, nu);
        const unsigned nb = (unsigned)a.ntx * a.nty * a.nchunks;
        dim3 blk(kTW3<T>, kBY3<T>, 1), grd(nb, 1, 1);
        const bool xtra = a.has_f || a.write_res || a.has_rg;
        const bool fonly = a.has_f && !a.write_res && !a.has_rg;
        const bool gonly = a.has_rg && !a.has_f && !a.write_res;
#define CHMY_F3(FV, EX)                                                                                   \
    do {                                                                                                  \
        constexpr bool F32 = std::is_same<T, float>::value;                                               \
        if (fonly && bf) hipLaunchKernelGGL((fused3d_kernel<T, FV, EX, kTW3<T>, kBY3<T>, 1, F32>), grd, blk, 0, s, a); \
        else if (fonly) hipLaunchKernelGGL((fused3d_kernel<T, FV, EX, kTW3<T>, kBY3<T>, 1, false>), grd, blk, 0, s, a); \
        else if (bf) hipLaunchKernelGGL((fused3d_kernel<T, FV, EX, kTW3<T>, kBY3<T>, 7, F32>), grd, blk, 0, s, a);    \
        else if (gonly) hipLaunchKernelGGL((fused3d_kernel<T, FV, EX, kTW3<T>, kBY3<T>, 2, false>), grd, blk, 0, s, a); \
        else if (xtra) hipLaunchKernelGGL((fused3d_kernel<T, FV, EX, kTW3<T>, kBY3<T>, 7, false>), grd, blk, 0, s, a); \
        else hipLaunchKernelGGL((fused3d_kernel<T, FV, EX, kTW3<T>, kBY3<T>, 0, false>), grd, blk, 0, s, a);           \
    } while (0)
        if (fv && exact) CHMY_F3(true, true);
        else if (fv) CHMY_F3(true, false);
        else if (exact) CHMY_F3(false, true);
        else CHMY_F3(false, false);
#undef CHMY_F3
    } else {
        Fused2<T> a;
        std::memset(&a, 0, sizeof a);
        auto P = [](const chmy_field_t& f) { return static_cast<T*>(f.data); };
        a.Vx = P(V[0]); a.Vy = P(V[1]); a.Pr = P(*Pr);
        a.txx = P(tau[0]); a.tyy = P(tau[1]); a.txy = P(tau[2]);
        a.oVx = P(Vo[0]); a.oVy = P(Vo[1]); a.oPr = P(*Pro);
        a.otxx = P(tauo[0]); a.otyy = P(tauo[1]); a.otxy = P(tauo[2]);
        a.Oxx = P(told[0]); a.Oyy = P(told[1]); a.Oxy = P(told[2]);
        a.eta = fv ? P(*eta) : a.Pr;
        a.eml = fv ? P(*eml) : a.Pr;
        a.rg = rg ? P(*rg) : a.Pr;
        a.dV = wres ? P(*dV) : a.oPr;
        a.rVx = wres ? P(rV[0]) : a.oVx;
        a.rVy = wres ? P(rV[1]) : a.oVy;
        a.has_f = fV != nullptr;
        a.fVx = fV ? P(fV[0]) : a.Pr;
        a.fVy = fV ? P(fV[1]) : a.Pr;
        a.fP = fV ? P(*fP) : a.Pr;
        a.L[L2C] = loc_stride(*Pr);
        a.L[L2X] = loc_stride(V[0]);
        a.L[L2Y] = loc_stride(V[1]);
        a.L[L2XY] = loc_stride(tau[2]);
        a.n[0] = (int32_t)g->n[0];
        a.n[1] = (int32_t)g->n[1];
        a.has_rg = rg != nullptr;
        a.write_res = wres;
        a.ix = T(g->inv_spacing[0]);
        a.iy = T(g->inv_spacing[1]);
        a.k = make_consts<T>(eta_s, eve_s, G, dt, dPr, dr, nu);
        const int64_t nbx = g->n[0] + 2, nby = g->n[1] + 2;
        if (tb) {   // two iterations, the free-slip batch between them
            a.ntx = (int32_t)((nbx + kFX - 5) / (kFX - 4));
            a.ychunk = exact ? kTBC<T, true> : kTBC<T, false>;
            a.nchunks = (int32_t)((nby + a.ychunk - 1) / a.ychunk);
            dim3 blk(kFX, kTBW, 1);
            dim3 grd((unsigned)((a.ntx + kTBW - 1) / kTBW), (unsigned)a.nchunks, 1);
            const bool hf = a.has_f;
            const bool hr = rg != nullptr;
            if (vfirst) {   // validated: FP64, no forcing
                if constexpr (std::is_same<T, double>::value) {
                    a.rVx = P(vfirst[0]);
                    a.rVy = P(vfirst[1]);
#define CHMY_TB2F(FV, EX)                                                                                          \
    do {                                                                                                           \
        if (hr) hipLaunchKernelGGL((fused2d_tb_kernel<double, FV, EX, 2, false, true>), grd, blk, 0, s, a, *tb);    \
        else hipLaunchKernelGGL((fused2d_tb_kernel<double, FV, EX, 2, false, false>), grd, blk, 0, s, a, *tb);      \
    } while (0)
                    if (fv && exact) CHMY_TB2F(true, true);
                    else if (fv) CHMY_TB2F(true, false);
                    else if (exact) CHMY_TB2F(false, true);
                    else CHMY_TB2F(false, false);
#undef CHMY_TB2F
                    CHMY_HIP(hipGetLastError());
                    return 0;
                } else {
                    return set_error(CHMY_E_DTYPE, "the first iteration's V is written in FP64 only");
                }
            }
#define CHMY_TB2R(FV, EX, RG)                                                                                 \
    do {                                                                                                      \
        constexpr bool F32 = std::is_same<T, float>::value;                                                   \
        if (hf && bf) hipLaunchKernelGGL((fused2d_tb_kernel<T, FV, EX, 1, F32, RG>), grd, blk, 0, s, a, *tb);  \
        else if (hf) hipLaunchKernelGGL((fused2d_tb_kernel<T, FV, EX, 1, false, RG>), grd, blk, 0, s, a, *tb); \
        else hipLaunchKernelGGL((fused2d_tb_kernel<T, FV, EX, 0, false, RG>), grd, blk, 0, s, a, *tb);         \
    } while (0)
#define CHMY_TB2(FV, EX)                          \
    do {                                          \
        if (hr) CHMY_TB2R(FV, EX, true);          \
        else CHMY_TB2R(FV, EX, false);            \
    } while (0)
            if (fv && exact) CHMY_TB2(true, true);
            else if (fv) CHMY_TB2(true, false);
            else if (exact) CHMY_TB2(false, true);
            else CHMY_TB2(false, false);
#undef CHMY_TB2
#undef CHMY_TB2R
            CHMY_HIP(hipGetLastError());
            return 0;
        }
        a.ntx = (int32_t)((nbx + kFX - 3) / (kFX - 2));
        a.ychunk = kYChunk2;
        a.nchunks = (int32_t)((nby + kYChunk2 - 1) / kYChunk2);
        dim3 blk(kFX, kW2, 1);
        dim3 grd((unsigned)((a.ntx + kW2 - 1) / kW2), (unsigned)a.nchunks, 1);
        const bool xtra = a.has_f || a.write_res || a.has_rg;
        const bool fonly = a.has_f && !a.write_res && !a.has_rg;
        const bool gonly = a.has_rg && !a.has_f && !a.write_res;
#define CHMY_F2(FV, EX)                                                                                   \
    do {                                                                                                  \
        constexpr bool F32 = std::is_same<T, float>::value;                                               \
        if (fonly && bf) hipLaunchKernelGGL((fused2d_kernel<T, FV, EX, 1, F32>), grd, blk, 0, s, a);       \
        else if (fonly) hipLaunchKernelGGL((fused2d_kernel<T, FV, EX, 1, false>), grd, blk, 0, s, a);      \
        else if (bf) hipLaunchKernelGGL((fused2d_kernel<T, FV, EX, 7, F32>), grd, blk, 0, s, a);          \
        else if (gonly) hipLaunchKernelGGL((fused2d_kernel<T, FV, EX, 2, false>), grd, blk, 0, s, a);      \
        else if (xtra) hipLaunchKernelGGL((fused2d_kernel<T, FV, EX, 7, false>), grd, blk, 0, s, a);       \
        else hipLaunchKernelGGL((fused2d_kernel<T, FV, EX, 0, false>), grd, blk, 0, s, a);                 \
    } while (0)
        if (fv && exact) CHMY_F2(true, true);
        else if (fv) CHMY_F2(true, false);
        else if (exact) CHMY_F2(false, true);
        else CHMY_F2(false, false);
#undef CHMY_F2
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // namespace chmy

using namespace chmy;

extern "C" {

}  // extern "C"

namespace chmy {

// chmy_stokes_iter's checks and launch; `tb` (2D, no residual fields) or
// `tb3` (3D, no residual fields, no rho g) runs two iterations with the
// free-slip batch between them.
static int stokes_iter_impl(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau,
                            const chmy_field_t* Pr, const chmy_field_t* V_out, const chmy_field_t* tau_out,
                            const chmy_field_t* Pr_out, const chmy_field_t* tau_old, const chmy_field_t* eta,
                            const chmy_field_t* eta_ml, const chmy_field_t* rhog, const chmy_field_t* f_V,
                            const chmy_field_t* f_P, const chmy_field_t* divV, const chmy_field_t* r_V, int flags,
                            double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                            double nudtau, void* stream, const FreeSlip2* tb, const FreeSlip3* tb3,
                            const chmy_field_t* V_first = nullptr) {
    CHMY_TRY(check_grid(g));
    if ((tb && g->nd != 2) || (tb3 && g->nd != 3) || (tb3 && rhog && f_V) || ((tb || tb3) && (flags & CHMY_ITER_RESIDUAL)))
        return set_error(CHMY_E_ARG, "two-iteration pass: free-slip batch of the grid's dims, no residual fields "
                                     "(and in 3D not rho g together with forcing)");
    if (g->nd < 2) return set_error(CHMY_E_ARG, "Stokes kernels need nd >= 2");
    if (flags & ~(CHMY_ITER_RESIDUAL | CHMY_ITER_EXACT)) return set_error(CHMY_E_ARG, "unknown flags %d", flags);
    const int write_residual = (flags & CHMY_ITER_RESIDUAL) != 0;
    const bool exact = (flags & CHMY_ITER_EXACT) != 0;
    CHMY_TRY(check_center(g, Pr, "Pr", -1));
    const int dt_ = Pr->dtype;
    CHMY_TRY(check_center(g, Pr_out, "Pr_out", dt_));
    CHMY_TRY(check_vector(g, V, "V", dt_));
    CHMY_TRY(check_vector(g, V_out, "V_out", dt_));
    CHMY_TRY(check_tensor(g, tau, "tau", dt_));
    CHMY_TRY(check_tensor(g, tau_out, "tau_out", dt_));
    // bf16 tau_old goes with bf16 f_V/f_P in an FP32 sweep (correction forcing)
    const bool bf = tau_old && tau_old[0].dtype == CHMY_BF16;
    if (bf && (dt_ != CHMY_F32 || !f_V || f_V[0].dtype != CHMY_BF16))
        return set_error(CHMY_E_DTYPE, "bf16 tau_old needs an FP32 sweep with bf16 f_V and f_P");
    CHMY_TRY(check_tensor(g, tau_old, "tau_old", bf ? CHMY_BF16 : dt_, bf));
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
    if ((f_V == nullptr) != (f_P == nullptr)) return set_error(CHMY_E_ARG, "f_V and f_P go together");
    if (f_V) {
        CHMY_TRY(check_vector(g, f_V, "f_V", bf ? CHMY_BF16 : dt_, bf));
        CHMY_TRY(check_center(g, f_P, "f_P", bf ? CHMY_BF16 : dt_, bf));
    }
    // one stride set per staggered location: every field must have halo 1
    const int nd = g->nd, nt = nd == 2 ? 3 : 6;
    auto halo1 = [nd](const chmy_field_t* f) {
        for (int d = 0; d < nd; ++d)
            if (f->halo[d] != 1) return false;
        return true;
    };
    bool ok = halo1(Pr) && halo1(Pr_out) && (!eta || (halo1(eta) && halo1(eta_ml))) && (!rhog || halo1(rhog)) &&
              (!write_residual || halo1(divV)) && (!f_P || halo1(f_P));
    for (int d = 0; d < nd; ++d) {
        ok = ok && halo1(&V[d]) && halo1(&V_out[d]);
        if (write_residual) ok = ok && halo1(&r_V[d]);
        if (f_V) ok = ok && halo1(&f_V[d]);
    }
    for (int c = 0; c < nt; ++c) ok = ok && halo1(&tau[c]) && halo1(&tau_out[c]) && halo1(&tau_old[c]);
    if (!ok) return set_error(CHMY_E_HALO, "fused iteration needs halo 1 on every field");
    for (int d = 0; d < nd; ++d)
        if (V[d].data == V_out[d].data) return set_error(CHMY_E_ARG, "V and V_out must be distinct buffers");
    for (int c = 0; c < nt; ++c)
        if (tau[c].data == tau_out[c].data) return set_error(CHMY_E_ARG, "tau and tau_out must be distinct buffers");
    if (Pr->data == Pr_out->data) return set_error(CHMY_E_ARG, "Pr and Pr_out must be distinct buffers");
    if (V_first) {   // the first iteration's V: FP64 passes without forcing
        if (!(tb || tb3) || dt_ != CHMY_F64 || f_V)
            return set_error(CHMY_E_ARG, "the first iteration's V is written by FP64 passes without forcing");
        CHMY_TRY(check_vector(g, V_first, "V_first", dt_));
        for (int d = 0; d < nd; ++d) {
            if (!halo1(&V_first[d])) return set_error(CHMY_E_HALO, "fused iteration needs halo 1 on every field");
            for (int e = 0; e < nd; ++e)
                if (V_first[d].data == V[e].data || V_first[d].data == V_out[e].data)
                    return set_error(CHMY_E_ARG, "V_first must be distinct from V and V_out");
        }
    }
    return dt_ == CHMY_F64
               ? fused_launch<double>(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, f_V, f_P, divV, r_V,
                                      write_residual, exact, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau,
                                      (hipStream_t)stream, tb, tb3, V_first)
               : fused_launch<float>(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, f_V, f_P, divV, r_V,
                                     write_residual, exact, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau,
                                     (hipStream_t)stream, tb, tb3);
}

// The batch names exactly the 2D free-slip conditions with scalar values:
// one entry per (component, dim, side); Vx Dirichlet in x and Neumann in y,
// Vy Neumann in x and Dirichlet in y; at least two cells per dim.
static bool free_slip_2d(const chmy_grid_t* g, const int32_t* counts, const int32_t* comp, const chmy_bc_t* bcs,
                         FreeSlip2* fs) {
    if (g->nd != 2 || g->n[0] < 2 || g->n[1] < 2) return false;
    const chmy_bc_t* rule[2][2][2] = {};   // [component][dim][side]
    int q = 0;
    for (int D = 0; D < 2; ++D)
        for (int sd = 0; sd < 2; ++sd)
            for (int k = 0; k < counts[2 * D + sd]; ++k, ++q) {
                const int c = comp[q];
                if (rule[c][D][sd]) return false;
                if (bcs[q].vkind != CHMY_VALUE_ZERO && bcs[q].vkind != CHMY_VALUE_SCALAR) return false;
                rule[c][D][sd] = &bcs[q];
            }
    auto val = [](const chmy_bc_t* b) { return b->vkind == CHMY_VALUE_SCALAR ? b->value : 0.0; };
    for (int c = 0; c < 2; ++c)
        for (int D = 0; D < 2; ++D)
            for (int sd = 0; sd < 2; ++sd) {
                const chmy_bc_t* b = rule[c][D][sd];
                const int want = c == D ? CHMY_DIRICHLET : CHMY_NEUMANN;
                if (!b || b->kind != want) return false;
            }
    for (int sd = 0; sd < 2; ++sd) {
        fs->ax[sd] = val(rule[0][0][sd]);
        fs->qy[sd] = val(rule[0][1][sd]);
        fs->qx[sd] = val(rule[1][0][sd]);
        fs->by[sd] = val(rule[1][1][sd]);
    }
    fs->hx = g->spacing[0];
    fs->hy = g->spacing[1];
    return true;
}

// The batch names exactly the 3D free-slip conditions with scalar values:
// one entry per (component, dim, side); component c Dirichlet on dim c and
// Neumann on the other dims; at least two cells per dim.
static bool free_slip_3d(const chmy_grid_t* g, const int32_t* counts, const int32_t* comp, const chmy_bc_t* bcs,
                         FreeSlip3* fs) {
    if (g->nd != 3 || g->n[0] < 2 || g->n[1] < 2 || g->n[2] < 2) return false;
    const chmy_bc_t* rule[3][3][2] = {};   // [component][dim][side]
    int q = 0;
    for (int D = 0; D < 3; ++D)
        for (int sd = 0; sd < 2; ++sd)
            for (int k = 0; k < counts[2 * D + sd]; ++k, ++q) {
                const int c = comp[q];
                if (rule[c][D][sd]) return false;
                if (bcs[q].vkind != CHMY_VALUE_ZERO && bcs[q].vkind != CHMY_VALUE_SCALAR) return false;
                rule[c][D][sd] = &bcs[q];
            }
    for (int c = 0; c < 3; ++c)
        for (int D = 0; D < 3; ++D)
            for (int sd = 0; sd < 2; ++sd) {
                const chmy_bc_t* b = rule[c][D][sd];
                if (!b || b->kind != (c == D ? CHMY_DIRICHLET : CHMY_NEUMANN)) return false;
                fs->v[c][D][sd] = b->vkind == CHMY_VALUE_SCALAR ? b->value : 0.0;
            }
    for (int D = 0; D < 3; ++D) fs->h[D] = g->spacing[D];
    return true;
}

}  // namespace chmy

extern "C" {

int chmy_stokes_iter(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau, const chmy_field_t* Pr,
                     const chmy_field_t* V_out, const chmy_field_t* tau_out, const chmy_field_t* Pr_out,
                     const chmy_field_t* tau_old, const chmy_field_t* eta, const chmy_field_t* eta_ml,
                     const chmy_field_t* rhog, const chmy_field_t* f_V, const chmy_field_t* f_P,
                     const chmy_field_t* divV, const chmy_field_t* r_V, int flags,
                     double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                     double nudtau, void* stream) {
    return stokes_iter_impl(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, f_V, f_P, divV, r_V,
                            flags, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau, stream, nullptr, nullptr);
}

// Several PT iterations in one call: each is chmy_stokes_iter from buffer
// set p into set 1-p followed by the V boundary batch on set 1-p (the
// Launcher's `bc=` of update_velocity!, examples/stokes_3d_inc_ve_T.jl:165),
// so the host pays one call per `niter` iterations instead of ~4 launches and
// descriptor packing per iteration.  The V batch is given once: bc_counts[2D+s]
// entries per (dim, side), each naming a V component (bc_comp) and its
// condition (bcs), in batch order; it is applied to whichever set was written.
// flags apply to every iteration, flags_last to the final one (e.g. to write
// the residual fields).  *parity_inout is the set holding the current state
// on entry and receives it on return.
int chmy_stokes_iterate(const chmy_grid_t* g, const chmy_field_t* V0, const chmy_field_t* tau0,
                        const chmy_field_t* Pr0, const chmy_field_t* V1, const chmy_field_t* tau1,
                        const chmy_field_t* Pr1, const chmy_field_t* tau_old, const chmy_field_t* eta,
                        const chmy_field_t* eta_ml, const chmy_field_t* rhog, const chmy_field_t* f_V,
                        const chmy_field_t* f_P, const chmy_field_t* divV, const chmy_field_t* r_V,
                        const int32_t* bc_counts, const int32_t* bc_comp, const chmy_bc_t* bcs, int niter,
                        int flags, int flags_last, double eta_s, double eta_ve_s, double G, double dt,
                        double dtau_Pr, double dtau_r, double nudtau, int* parity_inout, void* stream) {
    CHMY_TRY(check_grid(g));
    if (!parity_inout || (*parity_inout != 0 && *parity_inout != 1)) return set_error(CHMY_E_ARG, "parity must be 0/1");
    if (niter < 0) return set_error(CHMY_E_ARG, "niter < 0");
    if (!bc_counts || !bc_comp || !bcs) return set_error(CHMY_E_ARG, "V batch is NULL");
    const int nd = g->nd;
    int total = 0;
    for (int q = 0; q < 2 * nd; ++q) {
        if (bc_counts[q] < 0) return set_error(CHMY_E_ARG, "negative batch count");
        total += bc_counts[q];
    }
    if (total > 64) return set_error(CHMY_E_SIZE, "V batch larger than 64 entries");
    for (int q = 0; q < total; ++q)
        if (bc_comp[q] < 0 || bc_comp[q] >= nd) return set_error(CHMY_E_ARG, "bc_comp out of range");
    chmy_field_t fset[2][64];
    const chmy_field_t* Vs[2] = {V0, V1};
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < total; ++q) fset[p][q] = Vs[p][bc_comp[q]];
    const chmy_field_t* T[2] = {tau0, tau1};
    const chmy_field_t* P[2] = {Pr0, Pr1};
    // free-slip batches run iterations in pairs (one pass, the batch in-kernel
    // between them); the pair writes the other set, as one iteration does, so
    // the parity flips once per pair.  CHMY_ITER_SINGLE keeps single
    // iterations (CHMY_ITER_PAIRS, once the 3D opt-in, is accepted and ignored)
    const bool single = ((flags | flags_last) & CHMY_ITER_SINGLE) != 0;
    flags &= ~(CHMY_ITER_PAIRS | CHMY_ITER_SINGLE);
    flags_last &= ~(CHMY_ITER_PAIRS | CHMY_ITER_SINGLE);
    FreeSlip2 fs;
    FreeSlip3 fs3;
    const bool tb2 = CHMY_FUSED_TB && !single && nd == 2 && free_slip_2d(g, bc_counts, bc_comp, bcs, &fs);
    const bool tb3 = !single && !(rhog && f_V) && nd == 3 && free_slip_3d(g, bc_counts, bc_comp, bcs, &fs3);
    const bool tb = tb2 || tb3;
    int p = *parity_inout;
    for (int it = 0; it < niter;) {
        const int fl = it == niter - 1 ? flags_last : flags;
        const int fl2 = it + 1 == niter - 1 ? flags_last : flags;
        if (tb && it + 1 < niter && fl == fl2 && !(fl & CHMY_ITER_RESIDUAL)) {
            CHMY_TRY(stokes_iter_impl(g, Vs[p], T[p], P[p], Vs[1 - p], T[1 - p], P[1 - p], tau_old, eta, eta_ml, rhog,
                                      f_V, f_P, divV, r_V, fl, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau,
                                      stream, tb2 ? &fs : nullptr, tb3 ? &fs3 : nullptr));
            it += 2;
        } else {
            CHMY_TRY(chmy_stokes_iter(g, Vs[p], T[p], P[p], Vs[1 - p], T[1 - p], P[1 - p], tau_old, eta, eta_ml,
                                      rhog, f_V, f_P, divV, r_V, fl, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau,
                                      stream));
            it += 1;
        }
        CHMY_TRY(chmy_bc_batchset(g, bc_counts, fset[1 - p], bcs, stream));
        p = 1 - p;
    }
    *parity_inout = p;
    return 0;
}

// Two PT iterations in one pass with the free-slip V batch between them,
// evaluated in-kernel: (V, tau, Pr) -> (V_out, tau_out, Pr_out); the caller
// applies the batch to V_out afterwards (build extension).
int chmy_stokes_iter2(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau, const chmy_field_t* Pr,
                      const chmy_field_t* V_out, const chmy_field_t* tau_out, const chmy_field_t* Pr_out,
                      const chmy_field_t* tau_old, const chmy_field_t* eta, const chmy_field_t* eta_ml,
                      const chmy_field_t* rhog, const chmy_field_t* f_V, const chmy_field_t* f_P,
                      const int32_t* bc_counts, const int32_t* bc_comp, const chmy_bc_t* bcs, int flags,
                      double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                      double nudtau, void* stream) {
    CHMY_TRY(check_grid(g));
    if (!bc_counts || !bc_comp || !bcs) return set_error(CHMY_E_ARG, "V batch is NULL");
    if (flags & ~CHMY_ITER_EXACT) return set_error(CHMY_E_ARG, "chmy_stokes_iter2 takes CHMY_ITER_EXACT only");
    const int nd = g->nd;
    int total = 0;
    for (int q = 0; q < 2 * nd; ++q) {
        if (bc_counts[q] < 0) return set_error(CHMY_E_ARG, "negative batch count");
        total += bc_counts[q];
    }
    for (int q = 0; q < total; ++q)
        if (bc_comp[q] < 0 || bc_comp[q] >= nd) return set_error(CHMY_E_ARG, "bc_comp out of range");
    FreeSlip2 fs;
    FreeSlip3 fs3;
    const bool tb2 = nd == 2 && free_slip_2d(g, bc_counts, bc_comp, bcs, &fs);
    const bool tb3 = nd == 3 && free_slip_3d(g, bc_counts, bc_comp, bcs, &fs3);
    if (!tb2 && !tb3)
        return set_error(CHMY_E_ARG, "chmy_stokes_iter2 needs the free-slip V batch (each component Dirichlet on "
                                     "its own dim, Neumann on the others, scalar values, >= 2 cells per dim)");
    return stokes_iter_impl(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, f_V, f_P, nullptr,
                            nullptr, flags, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau, stream,
                            tb2 ? &fs : nullptr, tb3 ? &fs3 : nullptr);
}

// chmy_stokes_iter2 that also writes the first iteration's V (before its V
// batch) into V_first, at the points the pass writes V_out (build extension):
// in the physical time loop update_thermal_flux! reads V after every PT
// iteration.  2D and 3D, FP64, no forcing; the caller applies the V batch to
// V_first and to V_out.
int chmy_stokes_iter2_first(const chmy_grid_t* g, const chmy_field_t* V, const chmy_field_t* tau,
                            const chmy_field_t* Pr, const chmy_field_t* V_out, const chmy_field_t* tau_out,
                            const chmy_field_t* Pr_out, const chmy_field_t* V_first, const chmy_field_t* tau_old,
                            const chmy_field_t* eta, const chmy_field_t* eta_ml, const chmy_field_t* rhog,
                            const int32_t* bc_counts, const int32_t* bc_comp, const chmy_bc_t* bcs, int flags,
                            double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                            double nudtau, void* stream) {
    CHMY_TRY(check_grid(g));
    if (!V_first) return set_error(CHMY_E_ARG, "V_first is NULL");
    if (!bc_counts || !bc_comp || !bcs) return set_error(CHMY_E_ARG, "V batch is NULL");
    if (flags & ~CHMY_ITER_EXACT) return set_error(CHMY_E_ARG, "chmy_stokes_iter2_first takes CHMY_ITER_EXACT only");
    const int nd = g->nd;
    if (nd != 2 && nd != 3) return set_error(CHMY_E_ARG, "chmy_stokes_iter2_first is 2D or 3D");
    int total = 0;
    for (int q = 0; q < 2 * nd; ++q) {
        if (bc_counts[q] < 0) return set_error(CHMY_E_ARG, "negative batch count");
        total += bc_counts[q];
    }
    for (int q = 0; q < total; ++q)
        if (bc_comp[q] < 0 || bc_comp[q] >= nd) return set_error(CHMY_E_ARG, "bc_comp out of range");
    FreeSlip2 fs;
    FreeSlip3 fs3;
    const bool tb2 = nd == 2 && free_slip_2d(g, bc_counts, bc_comp, bcs, &fs);
    const bool tb3 = nd == 3 && free_slip_3d(g, bc_counts, bc_comp, bcs, &fs3);
    if (!tb2 && !tb3)
        return set_error(CHMY_E_ARG, "chmy_stokes_iter2_first needs the free-slip V batch (each component Dirichlet "
                                     "on its own dim, Neumann on the others, scalar values, >= 2 cells per dim)");
    return stokes_iter_impl(g, V, tau, Pr, V_out, tau_out, Pr_out, tau_old, eta, eta_ml, rhog, nullptr, nullptr,
                            nullptr, nullptr, flags, eta_s, eta_ve_s, G, dt, dtau_Pr, dtau_r, nudtau, stream,
                            tb2 ? &fs : nullptr, tb3 ? &fs3 : nullptr, V_first);
}

}  // extern "C"
