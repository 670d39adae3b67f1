// Shallow-ice (SIA) nonlinear diffusion: forward kernels of examples/sia.jl
// and a hand-written adjoint of one explicit step (build extension).
//
// Forward (sia.jl:7-16,30-54), on cells/vertices 1..n of a 2D grid:
//   S = B + H;  g_x = ∂x S (vertex);  a_x = lerp(g_x → centre);  vm = sqrt(a_x²+a_y²)
//   D = (A H^(n+2) + As H^n) vm^(n-1)          (A ≡ 1 reproduces sia.jl's D)
//   q_x = -lerp(D → x-vertex) g_x,  boundary vertices 0 (q Dirichlet on all sides)
//   M = min(mb_a (S - mb_b), mb_c);  H' = max(H + Δt(-divg q + M), 0)
// Powers are repeated products; lerp is fma(0.5, b - a, a) as in Chmy.  Lazy
// FunctionFields are evaluated at every access, so D is recomputed per use.
//
// Adjoint of one step H -> H' (reverse mode, exact derivative of the
// discrete map above; max(x,0) and min(..) take the branch they took):
//   A1 cells   mu = lam [x > 0]
//   A2 faces   qbar = Δt iΔ (mu_right - mu_left) on interior faces;
//              Dvbar = -g qbar,  gbar_direct = -lerp(D) qbar
//   A3 cells   Dbar = ½ Σ Dvbar of the 4 faces;  Abar → grad_logA += A Abar;
//              HDbar = Dbar ∂D/∂H;  abar = Dbar ∂D/∂a
//   A4 faces   gbar = gbar_direct + ½ (abar of the 2 cells sharing the face)
//   A5 cells   lam' = mu + HDbar + Sbar, Sbar = iΔ(gbar_left - gbar_right) (+ M branch);
//              halo cells fold into their neighbour (the Neumann copy H_0 = H_1).
// Each pass is a gather, so no float atomics and the result is deterministic.
#include <algorithm>
#include <mutex>

#include <vector>

#include "common.hpp"

namespace chmy {

struct SiaPar {
    double As;
    int npow;
    double mb_a, mb_b, mb_c, dt;
};

template <class T>
__device__ __forceinline__ T ipow_(T x, int n) {
    T r = T(1);
    for (int i = 0; i < n; ++i) r = r * x;
    return r;
}

template <class T>
struct SiaF {
    DF<T> H, B, A;
    int hasA;
    T ix, iy, As, mba, mbb, mbc;
    int n;
    __device__ T S(int64_t i, int64_t j) const { return B(i, j, 0) + H(i, j, 0); }
    __device__ T gx(int64_t i, int64_t j) const { return (S(i, j) - S(i - 1, j)) * ix; }
    __device__ T gy(int64_t i, int64_t j) const { return (S(i, j) - S(i, j - 1)) * iy; }
    __device__ T ax(int64_t i, int64_t j) const { T l = gx(i, j), r = gx(i + 1, j); return fma(T(0.5), r - l, l); }
    __device__ T ay(int64_t i, int64_t j) const { T l = gy(i, j), r = gy(i, j + 1); return fma(T(0.5), r - l, l); }
    __device__ T D(int64_t i, int64_t j) const {
        const T h = H(i, j, 0);
        const T a = hasA ? A(i, j, 0) : T(1);
        const T u = ax(i, j), w = ay(i, j);
        const T vm = sqrt(u * u + w * w);
        return (a * ipow_(h, n + 2) + As * ipow_(h, n)) * ipow_(vm, n - 1);
    }
    __device__ T M(int64_t i, int64_t j) const {
        const T m = mba * (S(i, j) - mbb);
        return m < mbc ? m : mbc;
    }
    // fluxes of the reference kernel (sia.jl:14-15) at vertex (i, j)
    __device__ T qx(int64_t i, int64_t j) const {
        const T Dl = D(i - 1, j), Dr = D(i, j);
        return -fma(T(0.5), Dr - Dl, Dl) * gx(i, j);
    }
    __device__ T qy(int64_t i, int64_t j) const {
        const T Dl = D(i, j - 1), Dr = D(i, j);
        return -fma(T(0.5), Dr - Dl, Dl) * gy(i, j);
    }
    // flux after the Dirichlet batch: zero on boundary vertices 1 and n+1
    __device__ T qxb(int64_t i, int64_t j, int64_t nx) const { return (i <= 1 || i >= nx + 1) ? T(0) : qx(i, j); }
    __device__ T qyb(int64_t i, int64_t j, int64_t ny) const { return (j <= 1 || j >= ny + 1) ? T(0) : qy(i, j); }
};

template <class T>
__device__ __forceinline__ void cell_index(int64_t& i, int64_t& j, int64_t nx, int64_t ny, bool& ok) {
    i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    j = 1 + blockIdx.y * (int64_t)blockDim.y + threadIdx.y;
    ok = i <= nx && j <= ny;
}

// update_q! (sia.jl:12-16) over vertices 1..n in both dims
template <class T>
__global__ void __launch_bounds__(256) sia_flux_kernel(SiaF<T> s, DF<T> qx, DF<T> qy, int64_t nx, int64_t ny) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    qx(i, j, 0) = s.qx(i, j);
    qy(i, j, 0) = s.qy(i, j);
}

// update_H! (sia.jl:7-10) over cells 1..n
template <class T>
__global__ void __launch_bounds__(256) sia_update_kernel(SiaF<T> s, DF<T> qx, DF<T> qy, DF<T> Ho, int64_t nx,
                                                         int64_t ny, T dt) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    const T dv = (qx(i + 1, j, 0) - qx(i, j, 0)) * s.ix + (qy(i, j + 1, 0) - qy(i, j, 0)) * s.iy;
    Ho(i, j, 0) = jmax0(s.H(i, j, 0) + dt * (-dv + s.M(i, j)));
}

// D at cells 1..n (the lazy FunctionField of sia.jl:46-49 materialised)
template <class T>
__global__ void __launch_bounds__(256) sia_d_kernel(SiaF<T> s, DF<T> D, int64_t nx, int64_t ny) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    D(i, j, 0) = s.D(i, j);
}

// update_q! + q Dirichlet batch + update_H! in one pass (no q storage)
template <class T>
__global__ void __launch_bounds__(256) sia_step_kernel(SiaF<T> s, DF<T> Ho, int64_t nx, int64_t ny, T dt) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    const T dv = (s.qxb(i + 1, j, nx) - s.qxb(i, j, nx)) * s.ix + (s.qyb(i, j + 1, ny) - s.qyb(i, j, ny)) * s.iy;
    Ho(i, j, 0) = jmax0(s.H(i, j, 0) + dt * (-dv + s.M(i, j)));
}

// ---------------------------------------------------------------- adjoint
template <class T>
struct SiaW {   // workspace fields (Center / x-vertex / y-vertex)
    DF<T> mu, Dvx, gdx, Dvy, gdy, ax, ay, HD;
};

template <class T>
__global__ void __launch_bounds__(256) sia_adj1(SiaF<T> s, DF<T> lam, SiaW<T> w, int64_t nx, int64_t ny, T dt) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    const T dv = (s.qxb(i + 1, j, nx) - s.qxb(i, j, nx)) * s.ix + (s.qyb(i, j + 1, ny) - s.qyb(i, j, ny)) * s.iy;
    const T x = s.H(i, j, 0) + dt * (-dv + s.M(i, j));
    w.mu(i, j, 0) = x > T(0) ? lam(i, j, 0) : T(0);
}

// faces: x-vertices (i in 1..nx+1, j in 1..ny) and y-vertices (i in 1..nx, j in 1..ny+1)
template <class T>
__global__ void __launch_bounds__(256) sia_adj2(SiaF<T> s, SiaW<T> w, int64_t nx, int64_t ny, T dt) {
    const int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j = 1 + blockIdx.y * (int64_t)blockDim.y + threadIdx.y;
    if (i <= nx + 1 && j <= ny) {
        T Dv = T(0), gd = T(0);
        if (i >= 2 && i <= nx) {
            const T qb = dt * s.ix * (w.mu(i, j, 0) - w.mu(i - 1, j, 0));
            const T Dl = s.D(i - 1, j), Dr = s.D(i, j);
            Dv = -s.gx(i, j) * qb;
            gd = -fma(T(0.5), Dr - Dl, Dl) * qb;
        }
        w.Dvx(i, j, 0) = Dv;
        w.gdx(i, j, 0) = gd;
    }
    if (i <= nx && j <= ny + 1) {
        T Dv = T(0), gd = T(0);
        if (j >= 2 && j <= ny) {
            const T qb = dt * s.iy * (w.mu(i, j, 0) - w.mu(i, j - 1, 0));
            const T Dl = s.D(i, j - 1), Dr = s.D(i, j);
            Dv = -s.gy(i, j) * qb;
            gd = -fma(T(0.5), Dr - Dl, Dl) * qb;
        }
        w.Dvy(i, j, 0) = Dv;
        w.gdy(i, j, 0) = gd;
    }
}

template <class T>
__global__ void __launch_bounds__(256) sia_adj3(SiaF<T> s, SiaW<T> w, DF<T> gA, int64_t nx, int64_t ny) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    const T Db = T(0.5) * (w.Dvx(i, j, 0) + w.Dvx(i + 1, j, 0)) + T(0.5) * (w.Dvy(i, j, 0) + w.Dvy(i, j + 1, 0));
    const int n = s.n;
    const T h = s.H(i, j, 0);
    const T a = s.hasA ? s.A(i, j, 0) : T(1);
    const T u = s.ax(i, j), v = s.ay(i, j);
    const T vm = sqrt(u * u + v * v);
    const T P = ipow_(vm, n - 1);
    const T C = a * ipow_(h, n + 2) + s.As * ipow_(h, n);
    if (s.hasA) gA(i, j, 0) = gA(i, j, 0) + a * (Db * ipow_(h, n + 2) * P);
    w.HD(i, j, 0) = Db * (T(n + 2) * a * ipow_(h, n + 1) + T(n) * s.As * ipow_(h, n - 1)) * P;
    // d(vm^(n-1))/du = coef * u with coef = (n-1) vm^(n-3) (n >= 3; n = 3 gives 2),
    // 1/vm for n = 2 (0 at vm = 0), 0 for n = 1
    T coef = T(0);
    if (n >= 3) coef = T(n - 1) * ipow_(vm, n - 3);
    else if (n == 2 && vm > T(0)) coef = T(1) / vm;
    const T Pb = Db * C;
    w.ax(i, j, 0) = Pb * coef * u;
    w.ay(i, j, 0) = Pb * coef * v;
}

template <class T>
__global__ void __launch_bounds__(256) sia_adj4(SiaW<T> w, int64_t nx, int64_t ny) {
    const int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j = 1 + blockIdx.y * (int64_t)blockDim.y + threadIdx.y;
    if (i <= nx + 1 && j <= ny) {
        T g = w.gdx(i, j, 0);
        if (i - 1 >= 1) g = g + T(0.5) * w.ax(i - 1, j, 0);
        if (i <= nx) g = g + T(0.5) * w.ax(i, j, 0);
        w.gdx(i, j, 0) = g;
    }
    if (i <= nx && j <= ny + 1) {
        T g = w.gdy(i, j, 0);
        if (j - 1 >= 1) g = g + T(0.5) * w.ay(i, j - 1, 0);
        if (j <= ny) g = g + T(0.5) * w.ay(i, j, 0);
        w.gdy(i, j, 0) = g;
    }
}

template <class T>
__global__ void __launch_bounds__(256) sia_adj5(SiaF<T> s, SiaW<T> w, DF<T> lo, int64_t nx, int64_t ny, T dt) {
    int64_t i, j;
    bool ok;
    cell_index<T>(i, j, nx, ny, ok);
    if (!ok) return;
    auto Sb = [&](int64_t a, int64_t b) -> T {
        // Sbar of cell (a, b): gx(v) = (S_v - S_{v-1}) iΔx, faces v in 1..nx+1 (zero outside)
        T r = T(0);
        if (a >= 1 && a <= nx + 1) r = r + s.ix * w.gdx(a, b, 0);
        if (a + 1 >= 1 && a + 1 <= nx + 1) r = r - s.ix * w.gdx(a + 1, b, 0);
        if (b >= 1 && b <= ny + 1) r = r + s.iy * w.gdy(a, b, 0);
        if (b + 1 >= 1 && b + 1 <= ny + 1) r = r - s.iy * w.gdy(a, b + 1, 0);
        return r;
    };
    const T mu = w.mu(i, j, 0);
    T l = mu + w.HD(i, j, 0) + Sb(i, j);
    const T m = s.mba * (s.S(i, j) - s.mbb);
    if (m < s.mbc) l = l + dt * mu * s.mba;
    // halo cells feed through the Neumann copy H_0 = H_1, H_{n+1} = H_n: the halo
    // S enters only the boundary-face gradients, Sbar_0 = -iΔx gbar_1 and
    // Sbar_{n+1} = iΔx gbar_{n+1}
    if (i == 1) l = l + (-s.ix * w.gdx(1, j, 0));
    if (i == nx) l = l + s.ix * w.gdx(nx + 1, j, 0);
    if (j == 1) l = l + (-s.iy * w.gdy(i, 1, 0));
    if (j == ny) l = l + s.iy * w.gdy(i, ny + 1, 0);
    lo(i, j, 0) = l;
}

// ---------------------------------------------------------------- tiled one-pass forms
// The forward step and the adjoint step as one launch each: a block owns a
// kSX x kSY tile of cells and stages the lazily evaluated fields it needs in
// LDS (D on the tile + 1 ring for the step; D on +3, mu on +2 and abar on +1
// for the adjoint), so D is evaluated once per cell instead of once per use.
// The arithmetic of every value is the one of the per-point kernels above,
// so the results are bitwise equal to them.  With blockIdx.z = 1 a launch of
// the adjoint also advances an independent forward step (the recompute of
// the previous checkpoint segment), which fills the GPU at small grids.
#ifndef CHMY_SIA_TX
#define CHMY_SIA_TX 32
#endif
#ifndef CHMY_SIA_TY
#define CHMY_SIA_TY 8
#endif
#ifndef CHMY_SIA_THREADS
#define CHMY_SIA_THREADS 256
#endif
constexpr int kSX = CHMY_SIA_TX, kSY = CHMY_SIA_TY, kST = CHMY_SIA_THREADS;

// S = B + H staged in LDS on the tile + R ring (zero outside cells 0..n+1,
// which no evaluation reads); the formulas are those of SiaF.
template <class T, int R>
struct SiaLds {
    static constexpr int W = kSX + 2 * R, HH = kSY + 2 * R;
    const T* Ss;
    int64_t i0, j0;
    const SiaF<T>* f;
    __device__ static void load(const SiaF<T>& s, T* Ss, int64_t i0, int64_t j0, int64_t nx, int64_t ny) {
        for (int q = threadIdx.x; q < W * HH; q += kST) {
            const int a = q % W, b = q / W;
            const int64_t i = i0 - R + a, j = j0 - R + b;
            Ss[q] = (i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 1) ? s.S(i, j) : T(0);
        }
    }
    __device__ T S(int64_t i, int64_t j) const { return Ss[(j - j0 + R) * W + (i - i0 + R)]; }
    __device__ T gx(int64_t i, int64_t j) const { return (S(i, j) - S(i - 1, j)) * f->ix; }
    __device__ T gy(int64_t i, int64_t j) const { return (S(i, j) - S(i, j - 1)) * f->iy; }
    __device__ T ax(int64_t i, int64_t j) const { T l = gx(i, j), r = gx(i + 1, j); return fma(T(0.5), r - l, l); }
    __device__ T ay(int64_t i, int64_t j) const { T l = gy(i, j), r = gy(i, j + 1); return fma(T(0.5), r - l, l); }
    __device__ T D(int64_t i, int64_t j) const {
        const T h = f->H(i, j, 0);
        const T a = f->hasA ? f->A(i, j, 0) : T(1);
        const T u = ax(i, j), w = ay(i, j);
        const T vm = sqrt(u * u + w * w);
        return (a * ipow_(h, f->n + 2) + f->As * ipow_(h, f->n)) * ipow_(vm, f->n - 1);
    }
    __device__ T M(int64_t i, int64_t j) const {
        const T m = f->mba * (S(i, j) - f->mbb);
        return m < f->mbc ? m : f->mbc;
    }
};

template <class T>
struct SiaStep {   // one forward step H -> Ho (+ the H Neumann batch when bc != 0)
    SiaF<T> s;
    DF<T> Ho;
    T dt;
    int bc;
    int wd;        // also store D(H) of the tile's cells into Dout (for a later adjoint step at H)
    DF<T> Dout;
};

template <class T>
__device__ __forceinline__ void sia_step_block(const SiaStep<T>& p, int64_t nx, int64_t ny, T* lds) {
    constexpr int W = kSX + 2, Hh = kSY + 2;
    using LS = SiaLds<T, 2>;
    T(*Ds)[W] = reinterpret_cast<T(*)[W]>(lds + LS::W * LS::HH);
    const SiaF<T>& s = p.s;
    const int tid = threadIdx.x;
    const int64_t i0 = 1 + (int64_t)blockIdx.x * kSX, j0 = 1 + (int64_t)blockIdx.y * kSY;
    LS::load(s, lds, i0, j0, nx, ny);
    const LS L{lds, i0, j0, &s};
    __syncthreads();
    for (int q = tid; q < W * Hh; q += kST) {
        const int a = q % W, b = q / W;
        const int64_t i = i0 - 1 + a, j = j0 - 1 + b;
        const bool in = i >= 1 && i <= nx && j >= 1 && j <= ny;
        const T d = in ? L.D(i, j) : T(0);
        Ds[b][a] = d;
        if (p.wd && in && a >= 1 && a <= kSX && b >= 1 && b <= kSY) p.Dout(i, j, 0) = d;
    }
    __syncthreads();
    for (int q = tid; q < kSX * kSY; q += kST) {
        const int a = q % kSX + 1, b = q / kSX + 1;   // LDS position of the cell
        const int64_t i = i0 + a - 1, j = j0 + b - 1;
        if (i > nx || j > ny) continue;
        auto qx = [&](int64_t f, int la) -> T {   // x-vertex f, right cell at LDS column la
            if (f <= 1 || f >= nx + 1) return T(0);
            const T Dl = Ds[b][la - 1], Dr = Ds[b][la];
            return -fma(T(0.5), Dr - Dl, Dl) * L.gx(f, j);
        };
        auto qy = [&](int64_t f, int lb) -> T {
            if (f <= 1 || f >= ny + 1) return T(0);
            const T Dl = Ds[lb - 1][a], Dr = Ds[lb][a];
            return -fma(T(0.5), Dr - Dl, Dl) * L.gy(i, f);
        };
        const T dv = (qx(i + 1, a + 1) - qx(i, a)) * s.ix + (qy(j + 1, b + 1) - qy(j, b)) * s.iy;
        const T h = jmax0(s.H(i, j, 0) + p.dt * (-dv + L.M(i, j)));
        const DF<T>& Ho = p.Ho;
        Ho(i, j, 0) = h;
        if (p.bc) {
            // Neumann(0) batch, dims y then x (batch.jl:20-29): ghosts copy the
            // adjacent cell (fma(Δ, -0, f) == f), corners the diagonal cell
            if (i == 1) Ho(0, j, 0) = h;
            if (i == nx) Ho(nx + 1, j, 0) = h;
            if (j == 1) {
                Ho(i, 0, 0) = h;
                if (i == 1) Ho(0, 0, 0) = h;
                if (i == nx) Ho(nx + 1, 0, 0) = h;
            }
            if (j == ny) {
                Ho(i, ny + 1, 0) = h;
                if (i == 1) Ho(0, ny + 1, 0) = h;
                if (i == nx) Ho(nx + 1, ny + 1, 0) = h;
            }
        }
    }
}

template <class T>
struct SiaAdj {    // one adjoint step at state s.H: lam -> lo, gA += dJ/dlog A
    SiaF<T> s;
    DF<T> lam, lo, gA;
    T dt;
    int rd;        // D(s.H) read from Din (stored by the forward step at s.H) instead of evaluated
    DF<T> Din;
};

template <class T>
__device__ __forceinline__ void sia_adj_block(const SiaAdj<T>& p, int64_t nx, int64_t ny, T* lds) {
    constexpr int W3 = kSX + 6, H3 = kSY + 6, W2 = kSX + 4, H2 = kSY + 4, W1 = kSX + 2, H1 = kSY + 2;
    using LS = SiaLds<T, 4>;
    T* base = lds + LS::W * LS::HH;
    T(*Ds)[W3] = reinterpret_cast<T(*)[W3]>(base);
    T(*Ms)[W2] = reinterpret_cast<T(*)[W2]>(base + W3 * H3);
    T(*Ax)[W1] = reinterpret_cast<T(*)[W1]>(base + W3 * H3 + W2 * H2);
    T(*Ay)[W1] = reinterpret_cast<T(*)[W1]>(base + W3 * H3 + W2 * H2 + W1 * H1);
    T(*HDs)[kSX] = reinterpret_cast<T(*)[kSX]>(base + W3 * H3 + W2 * H2 + 2 * W1 * H1);
    const SiaF<T>& s = p.s;
    const T dt = p.dt;
    const int tid = threadIdx.x;
    const int64_t i0 = 1 + (int64_t)blockIdx.x * kSX, j0 = 1 + (int64_t)blockIdx.y * kSY;
    LS::load(s, lds, i0, j0, nx, ny);
    const LS L{lds, i0, j0, &s};
    __syncthreads();
    auto inside = [nx, ny](int64_t i, int64_t j) { return i >= 1 && i <= nx && j >= 1 && j <= ny; };
    // D on the tile + 3 ring; cell (i, j) at Ds[j - j0 + 3][i - i0 + 3]; the same
    // bits whether evaluated here or stored by the forward step at this state
    for (int q = tid; q < W3 * H3; q += kST) {
        const int a = q % W3, b = q / W3;
        const int64_t i = i0 - 3 + a, j = j0 - 3 + b;
        Ds[b][a] = inside(i, j) ? (p.rd ? p.Din(i, j, 0) : L.D(i, j)) : T(0);
    }
    __syncthreads();
    // pass A1: mu on the tile + 2 ring; cell at Ms[j - j0 + 2][i - i0 + 2]
    for (int q = tid; q < W2 * H2; q += kST) {
        const int a = q % W2, b = q / W2;
        const int64_t i = i0 - 2 + a, j = j0 - 2 + b;
        T mu = T(0);
        if (inside(i, j)) {
            const int da = a + 1, db = b + 1;   // the cell in Ds
            auto qx = [&](int64_t f, int la) -> T {
                if (f <= 1 || f >= nx + 1) return T(0);
                const T Dl = Ds[db][la - 1], Dr = Ds[db][la];
                return -fma(T(0.5), Dr - Dl, Dl) * L.gx(f, j);
            };
            auto qy = [&](int64_t f, int lb) -> T {
                if (f <= 1 || f >= ny + 1) return T(0);
                const T Dl = Ds[lb - 1][da], Dr = Ds[lb][da];
                return -fma(T(0.5), Dr - Dl, Dl) * L.gy(i, f);
            };
            const T dv = (qx(i + 1, da + 1) - qx(i, da)) * s.ix + (qy(j + 1, db + 1) - qy(j, db)) * s.iy;
            const T x = s.H(i, j, 0) + dt * (-dv + L.M(i, j));
            mu = x > T(0) ? p.lam(i, j, 0) : T(0);
        }
        Ms[b][a] = mu;
    }
    __syncthreads();
    // passes A2+A3: Dbar and abar on the tile + 1 ring; HD and grad on the tile
    for (int q = tid; q < W1 * H1; q += kST) {
        const int a = q % W1, b = q / W1;
        const int64_t i = i0 - 1 + a, j = j0 - 1 + b;
        T abx = T(0), aby = T(0);
        if (inside(i, j)) {
            const int ma = a + 1, mb = b + 1;   // the cell in Ms
            auto Dvx = [&](int64_t f, int la) -> T {   // x-vertex f, right cell at Ms column la
                if (f < 2 || f > nx) return T(0);
                const T qb = dt * s.ix * (Ms[mb][la] - Ms[mb][la - 1]);
                return -L.gx(f, j) * qb;
            };
            auto Dvy = [&](int64_t f, int lb) -> T {
                if (f < 2 || f > ny) return T(0);
                const T qb = dt * s.iy * (Ms[lb][ma] - Ms[lb - 1][ma]);
                return -L.gy(i, f) * qb;
            };
            const T Db = T(0.5) * (Dvx(i, ma) + Dvx(i + 1, ma + 1)) + T(0.5) * (Dvy(j, mb) + Dvy(j + 1, mb + 1));
            const int n = s.n;
            const T h = s.H(i, j, 0);
            const T av = s.hasA ? s.A(i, j, 0) : T(1);
            const T u = L.ax(i, j), v = L.ay(i, j);
            const T vm = sqrt(u * u + v * v);
            const T P = ipow_(vm, n - 1);
            const T Cc = av * ipow_(h, n + 2) + s.As * ipow_(h, n);
            const bool tile = a >= 1 && a <= kSX && b >= 1 && b <= kSY;
            if (tile) {
                if (s.hasA) p.gA(i, j, 0) = p.gA(i, j, 0) + av * (Db * ipow_(h, n + 2) * P);
                HDs[b - 1][a - 1] = Db * (T(n + 2) * av * ipow_(h, n + 1) + T(n) * s.As * ipow_(h, n - 1)) * P;
            }
            T coef = T(0);
            if (n >= 3) coef = T(n - 1) * ipow_(vm, n - 3);
            else if (n == 2 && vm > T(0)) coef = T(1) / vm;
            const T Pb = Db * Cc;
            abx = Pb * coef * u;
            aby = Pb * coef * v;
        }
        Ax[b][a] = abx;
        Ay[b][a] = aby;
    }
    __syncthreads();
    // passes A4+A5: gbar on the 4 faces of each tile cell, then lam'
    for (int q = tid; q < kSX * kSY; q += kST) {
        const int t = q % kSX, u = q / kSX;
        const int64_t i = i0 + t, j = j0 + u;
        if (i > nx || j > ny) continue;
        const int a1 = t + 1, b1 = u + 1;   // cell in Ax/Ay
        const int a2 = t + 2, b2 = u + 2;   // cell in Ms
        const int a3 = t + 3, b3 = u + 3;   // cell in Ds
        auto gbx = [&](int64_t f, int o) -> T {   // x-vertex f = i + o, o in {0, 1}
            T g = T(0);
            if (f >= 2 && f <= nx) {
                const T qb = dt * s.ix * (Ms[b2][a2 + o] - Ms[b2][a2 + o - 1]);
                const T Dl = Ds[b3][a3 + o - 1], Dr = Ds[b3][a3 + o];
                g = -fma(T(0.5), Dr - Dl, Dl) * qb;
            }
            if (f - 1 >= 1) g = g + T(0.5) * Ax[b1][a1 + o - 1];
            if (f <= nx) g = g + T(0.5) * Ax[b1][a1 + o];
            return g;
        };
        auto gby = [&](int64_t f, int o) -> T {
            T g = T(0);
            if (f >= 2 && f <= ny) {
                const T qb = dt * s.iy * (Ms[b2 + o][a2] - Ms[b2 + o - 1][a2]);
                const T Dl = Ds[b3 + o - 1][a3], Dr = Ds[b3 + o][a3];
                g = -fma(T(0.5), Dr - Dl, Dl) * qb;
            }
            if (f - 1 >= 1) g = g + T(0.5) * Ay[b1 + o - 1][a1];
            if (f <= ny) g = g + T(0.5) * Ay[b1 + o][a1];
            return g;
        };
        const T gx0 = gbx(i, 0), gx1 = gbx(i + 1, 1), gy0 = gby(j, 0), gy1 = gby(j + 1, 1);
        T r = T(0);
        r = r + s.ix * gx0;
        r = r - s.ix * gx1;
        r = r + s.iy * gy0;
        r = r - s.iy * gy1;
        const T mu = Ms[b2][a2];
        T l = mu + HDs[u][t] + r;
        const T m = s.mba * (L.S(i, j) - s.mbb);
        if (m < s.mbc) l = l + dt * mu * s.mba;
        if (i == 1) l = l + (-s.ix * gx0);
        if (i == nx) l = l + s.ix * gx1;
        if (j == 1) l = l + (-s.iy * gy0);
        if (j == ny) l = l + s.iy * gy1;
        p.lo(i, j, 0) = l;
    }
}

constexpr int kAdjLds = (kSX + 8) * (kSY + 8) + (kSX + 6) * (kSY + 6) + (kSX + 4) * (kSY + 4) + 2 * (kSX + 2) * (kSY + 2) + kSX * kSY;

// NP > 0: the Glen exponent as a compile-time constant (it must equal the
// run-time one), so that every power's product loop unrolls; the products
// and their order are the same as with NP = 0.
template <class T, int NP>
__global__ void __launch_bounds__(kST) sia_step_tile(SiaStep<T> f, int64_t nx, int64_t ny) {
    __shared__ T lds[(kSX + 4) * (kSY + 4) + (kSX + 2) * (kSY + 2)];
    if constexpr (NP > 0) f.s.n = NP;
    sia_step_block(f, nx, ny, lds);
}

// blockIdx.z == 0: adjoint tile; blockIdx.z == 1: forward tile of an independent step
template <class T, int NP>
__global__ void __launch_bounds__(kST) sia_adj_tile(SiaAdj<T> a, SiaStep<T> f, int64_t nx, int64_t ny) {
    __shared__ T lds[kAdjLds];
    if constexpr (NP > 0) {
        a.s.n = NP;
        f.s.n = NP;
    }
    if (blockIdx.z == 0) sia_adj_block(a, nx, ny, lds);
    else sia_step_block(f, nx, ny, lds);
}

// n = 3 (Glen's law as sia.jl uses it) takes the unrolled kernels
static void launch_step_tile(dim3 grd, hipStream_t st, const SiaStep<double>& f, int64_t nx, int64_t ny) {
    if (f.s.n == 3) hipLaunchKernelGGL((sia_step_tile<double, 3>), grd, dim3(kST), 0, st, f, nx, ny);
    else hipLaunchKernelGGL((sia_step_tile<double, 0>), grd, dim3(kST), 0, st, f, nx, ny);
}

static void launch_adj_tile(dim3 grd, hipStream_t st, const SiaAdj<double>& a, const SiaStep<double>& f, int64_t nx,
                            int64_t ny) {
    if (a.s.n == 3) hipLaunchKernelGGL((sia_adj_tile<double, 3>), grd, dim3(kST), 0, st, a, f, nx, ny);
    else hipLaunchKernelGGL((sia_adj_tile<double, 0>), grd, dim3(kST), 0, st, a, f, nx, ny);
}

template <class T>
static SiaF<T> make_sia(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* B, const chmy_field_t* A,
                        const SiaPar& p) {
    SiaF<T> s;
    s.H = view<T>(*H);
    s.B = view<T>(*B);
    s.A = A ? view<T>(*A) : s.H;
    s.hasA = A != nullptr;
    s.ix = T(g->inv_spacing[0]);
    s.iy = T(g->inv_spacing[1]);
    s.As = T(p.As);
    s.mba = T(p.mb_a);
    s.mbb = T(p.mb_b);
    s.mbc = T(p.mb_c);
    s.n = p.npow;
    return s;
}

static int sia_check(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* B, const chmy_field_t* A,
                     int npow) {
    CHMY_TRY(check_grid(g));
    if (g->nd != 2) return set_error(CHMY_E_ARG, "SIA kernels are 2D");
    if (npow < 1 || npow > 12) return set_error(CHMY_E_ARG, "Glen exponent n=%d not in 1..12", npow);
    CHMY_TRY(check_center(g, H, "H", CHMY_F64));
    CHMY_TRY(check_center(g, B, "B", CHMY_F64));
    if (A) CHMY_TRY(check_center(g, A, "A", CHMY_F64));
    // D at cell 0 reads S at -1: halo 1 covers logical -1..n+2
    return 0;
}

static dim3 sia_tiles(int64_t nx, int64_t ny, int z) {
    return dim3((unsigned)((nx + kSX - 1) / kSX), (unsigned)((ny + kSY - 1) / kSY), (unsigned)z);
}

static SiaStep<double> sia_stepper(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* Ho,
                                   const chmy_field_t* B, const chmy_field_t* A, const SiaPar& p, double dt) {
    SiaStep<double> f;
    f.s = make_sia<double>(g, H, B, A, p);
    f.Ho = view<double>(*Ho);
    f.dt = dt;
    f.bc = 1;
    f.wd = 0;
    f.Dout = f.Ho;
    return f;
}

static SiaAdj<double> sia_adjointer(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* lam,
                                    const chmy_field_t* lo, const chmy_field_t* gA, const chmy_field_t* B,
                                    const chmy_field_t* A, const SiaPar& p, double dt) {
    SiaAdj<double> a;
    a.s = make_sia<double>(g, H, B, A, p);
    a.lam = view<double>(*lam);
    a.lo = view<double>(*lo);
    a.gA = gA ? view<double>(*gA) : a.lo;
    a.dt = dt;
    a.rd = 0;
    a.Din = a.lo;
    return a;
}

static int sia_check_states(const chmy_grid_t* g, const chmy_field_t* st, int n, const char* what) {
    if (!st) return set_error(CHMY_E_ARG, "%s is NULL", what);
    for (int q = 0; q < n; ++q) CHMY_TRY(check_center(g, &st[q], what, CHMY_F64));
    return 0;
}

// Validation shared by both reverse-sweep entry points.
static int sia_reverse_check(const chmy_grid_t* g, const chmy_field_t* ckpt, int nckpt, int stride, int nsteps,
                             const chmy_field_t* lam, const chmy_field_t* grad_logA, const chmy_field_t* B,
                             const chmy_field_t* A, int npow, const double* dts) {
    if (nsteps < 1 || stride < 1) return set_error(CHMY_E_ARG, "nsteps=%d stride=%d", nsteps, stride);
    const int nseg = (nsteps + stride - 1) / stride;
    if (nckpt != nseg) return set_error(CHMY_E_ARG, "nckpt=%d, need ceil(nsteps/stride)=%d", nckpt, nseg);
    if (!dts) return set_error(CHMY_E_ARG, "dts is NULL");
    CHMY_TRY(check_grid(g));
    CHMY_TRY(sia_check_states(g, ckpt, nckpt, "ckpt"));
    CHMY_TRY(sia_check(g, &ckpt[0], B, A, npow));
    CHMY_TRY(check_center(g, lam, "lam_a", CHMY_F64));
    if (A) CHMY_TRY(check_center(g, grad_logA, "grad_logA", CHMY_F64));
    return 0;
}

// The reverse sweep on the stream: seg(i) gives segment state buffer i
// (0 <= i < 2 (stride - 1)) by value; with dbuf (2 stride Center fields, or
// NULL) the recompute steps also store D of the state they start from, and
// the adjoint step at that state reads it instead of evaluating it (every
// state of a segment but its last).  Returns the slot that holds dJ/dH_0
// (0 = lam_a, 1 = lam_b).
template <class SegF>
static int sia_reverse_sweep(const chmy_grid_t* g, const chmy_field_t* ckpt, int stride, int nsteps, SegF seg,
                             const chmy_field_t* lam_a, const chmy_field_t* lam_b, const chmy_field_t* grad_logA,
                             const chmy_field_t* B, const chmy_field_t* A, const SiaPar& p, const double* dts,
                             hipStream_t st, const chmy_field_t* dbuf = nullptr) {
    const int nseg = (nsteps + stride - 1) / stride;
    const int64_t nx = g->n[0], ny = g->n[1];
    // state q of segment s: q = 0 its checkpoint, q >= 1 seg((s % 2)(stride - 1) + q - 1)
    auto state = [&](int sg, int q) -> chmy_field_t {
        return q == 0 ? ckpt[sg] : seg((sg % 2) * (stride - 1) + q - 1);
    };
    // D of state q of segment s: dbuf[(s % 2) stride + q]
    auto dfield = [&](int sg, int q) -> DF<double> { return view<double>(dbuf[(sg % 2) * stride + q]); };
    auto len = [&](int sg) { return std::min(stride, nsteps - sg * stride); };
    // recompute the last segment on its own
    for (int q = 1; q < len(nseg - 1); ++q) {
        const chmy_field_t s0 = state(nseg - 1, q - 1), s1 = state(nseg - 1, q);
        SiaStep<double> f = sia_stepper(g, &s0, &s1, B, A, p, dts[(nseg - 1) * stride + q - 1]);
        if (dbuf) {
            f.wd = 1;
            f.Dout = dfield(nseg - 1, q - 1);
        }
        launch_step_tile(sia_tiles(nx, ny, 1), st, f, nx, ny);
    }
    int slot = 0;   // lam_in of the next adjoint step: 0 = lam_a, 1 = lam_b
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int L = len(sg), Lp = sg > 0 ? len(sg - 1) : 0;
        for (int r = 0; r < std::max(L, Lp - 1); ++r) {
            // the recompute of segment sg - 1 (state r -> r + 1) rides along with
            // adjoint step q = L - 1 - r of segment sg
            const bool fw = sg > 0 && r + 1 < Lp;
            SiaStep<double> f{};
            if (fw) {
                const chmy_field_t s0 = state(sg - 1, r), s1 = state(sg - 1, r + 1);
                f = sia_stepper(g, &s0, &s1, B, A, p, dts[(sg - 1) * stride + r]);
                if (dbuf) {
                    f.wd = 1;
                    f.Dout = dfield(sg - 1, r);
                }
            }
            if (r >= L) {
                launch_step_tile(sia_tiles(nx, ny, 1), st, f, nx, ny);
                continue;
            }
            const int q = L - 1 - r;
            const chmy_field_t* li = slot ? lam_b : lam_a;
            const chmy_field_t* lo = slot ? lam_a : lam_b;
            const chmy_field_t sq = state(sg, q);
            SiaAdj<double> a = sia_adjointer(g, &sq, li, lo, A ? grad_logA : nullptr, B, A, p,
                                             dts[sg * stride + q]);
            if (dbuf && q + 1 < L) {   // stored by the recompute step q -> q + 1
                a.rd = 1;
                a.Din = dfield(sg, q);
            }
            launch_adj_tile(sia_tiles(nx, ny, fw ? 2 : 1), st, a, f, nx, ny);
            slot ^= 1;
        }
    }
    return slot;
}

// Device scratch of chmy_sia_reverse_noalloc: the segment states and the
// second adjoint buffer in one file-scope array, which each GPU holds until
// the process ends.  48 MiB covers 512 x 512 cells (516 x 516 Center storage,
// 2.13 MB a field) at 10 steps per checkpoint: 2 x 9 segment states + 1.
// In host code `scratch` is not a device address: its address on a device
// comes from hipGetSymbolAddress with that device current.
constexpr size_t kSiaScratchBytes = size_t(48) << 20;
static __device__ __align__(256) double scratch[kSiaScratchBytes / sizeof(double)];

static std::mutex g_scratch_mu;
static void* g_scratch_dev[64] = {nullptr};   // scratch's address on each device

__global__ void __launch_bounds__(256) sia_fill_zero(double* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0.0;
}

}  // namespace chmy

// Resolves scratch's address on the current device (chmy_init calls it; the
// no-allocation entry point calls it on its first use on a device).
int chmy::sia_scratch_ptr(void** p) {
    int dev = 0;
    CHMY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_error(CHMY_E_ARG, "device index %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q)
            return set_error(CHMY_E_NODEV, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = q;
    }
    if (p) *p = g_scratch_dev[dev];
    return 0;
}

namespace chmy {

// Device memory for the D fields of chmy_sia_reverse: one buffer per device,
// grown on demand and kept until the process ends (calls on one device that
// overlap race, as with any shared workspace).
static std::mutex g_dcache_mu;
static void* g_dcache[64] = {nullptr};
static size_t g_dcache_sz[64] = {0};

static int sia_dcache(size_t bytes, void** p) {
    int dev = 0;
    CHMY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_error(CHMY_E_ARG, "device index %d", dev);
    std::lock_guard<std::mutex> lk(g_dcache_mu);
    if (g_dcache_sz[dev] < bytes) {
        if (g_dcache[dev]) CHMY_HIP(hipFree(g_dcache[dev]));
        g_dcache[dev] = nullptr;
        g_dcache_sz[dev] = 0;
        CHMY_HIP(hipMalloc(&g_dcache[dev], bytes));
        g_dcache_sz[dev] = bytes;
    }
    *p = g_dcache[dev];
    return 0;
}

static dim3 sia_blk() { return dim3(64, 4, 1); }
static dim3 sia_grid(int64_t nx, int64_t ny) {
    return dim3((unsigned)((nx + 1 + 63) / 64), (unsigned)((ny + 1 + 3) / 4), 1);
}

}  // namespace chmy

using namespace chmy;

extern "C" {

int chmy_sia_flux(const chmy_grid_t* g, const chmy_field_t* q, const chmy_field_t* H, const chmy_field_t* B,
                  const chmy_field_t* A, double As, int npow, void* stream) {
    CHMY_TRY(sia_check(g, H, B, A, npow));
    CHMY_TRY(check_vector(g, q, "q", CHMY_F64));
    SiaPar p{As, npow, 0, 0, 0, 0};
    SiaF<double> s = make_sia<double>(g, H, B, A, p);
    hipLaunchKernelGGL(sia_flux_kernel<double>, sia_grid(g->n[0], g->n[1]), sia_blk(), 0, (hipStream_t)stream, s,
                       view<double>(q[0]), view<double>(q[1]), g->n[0], g->n[1]);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_update(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* H_out, const chmy_field_t* q,
                    const chmy_field_t* B, double mb_a, double mb_b, double mb_c, double dt, void* stream) {
    CHMY_TRY(sia_check(g, H, B, nullptr, 1));
    CHMY_TRY(check_center(g, H_out, "H_out", CHMY_F64));
    CHMY_TRY(check_vector(g, q, "q", CHMY_F64));
    SiaPar p{0, 1, mb_a, mb_b, mb_c, dt};
    SiaF<double> s = make_sia<double>(g, H, B, nullptr, p);
    hipLaunchKernelGGL(sia_update_kernel<double>, sia_grid(g->n[0], g->n[1]), sia_blk(), 0, (hipStream_t)stream, s,
                       view<double>(q[0]), view<double>(q[1]), view<double>(*H_out), g->n[0], g->n[1], dt);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_diffusivity(const chmy_grid_t* g, const chmy_field_t* D, const chmy_field_t* H, const chmy_field_t* B,
                         const chmy_field_t* A, double As, int npow, void* stream) {
    CHMY_TRY(sia_check(g, H, B, A, npow));
    CHMY_TRY(check_center(g, D, "D", CHMY_F64));
    SiaPar p{As, npow, 0, 0, 0, 0};
    SiaF<double> s = make_sia<double>(g, H, B, A, p);
    hipLaunchKernelGGL(sia_d_kernel<double>, sia_grid(g->n[0], g->n[1]), sia_blk(), 0, (hipStream_t)stream, s,
                       view<double>(*D), g->n[0], g->n[1]);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_step(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* H_out, const chmy_field_t* B,
                  const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c, double dt,
                  void* stream) {
    CHMY_TRY(sia_check(g, H, B, A, npow));
    CHMY_TRY(check_center(g, H_out, "H_out", CHMY_F64));
    if (H_out->data == H->data) return set_error(CHMY_E_ARG, "chmy_sia_step needs distinct H and H_out");
    SiaPar p{As, npow, mb_a, mb_b, mb_c, dt};
    SiaF<double> s = make_sia<double>(g, H, B, A, p);
    hipLaunchKernelGGL(sia_step_kernel<double>, sia_grid(g->n[0], g->n[1]), sia_blk(), 0, (hipStream_t)stream, s,
                       view<double>(*H_out), g->n[0], g->n[1], dt);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_step_adjoint(const chmy_grid_t* g, const chmy_field_t* H, const chmy_field_t* lam_in,
                          const chmy_field_t* lam_out, const chmy_field_t* grad_logA, const chmy_field_t* B,
                          const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c,
                          double dt, const chmy_field_t* work, void* stream) {
    CHMY_TRY(sia_check(g, H, B, A, npow));
    CHMY_TRY(check_center(g, lam_in, "lam_in", CHMY_F64));
    CHMY_TRY(check_center(g, lam_out, "lam_out", CHMY_F64));
    if (lam_in->data == lam_out->data) return set_error(CHMY_E_ARG, "lam_in and lam_out must differ");
    if (A) CHMY_TRY(check_center(g, grad_logA, "grad_logA", CHMY_F64));
    SiaPar p{As, npow, mb_a, mb_b, mb_c, dt};
    if (!work) {   // one tiled pass
        SiaAdj<double> a = sia_adjointer(g, H, lam_in, lam_out, A ? grad_logA : nullptr, B, A, p, dt);
        launch_adj_tile(sia_tiles(g->n[0], g->n[1], 1), (hipStream_t)stream, a, SiaStep<double>{}, g->n[0], g->n[1]);
        CHMY_HIP(hipGetLastError());
        return 0;
    }
    const int vx[3] = {1, 0, 0}, vy[3] = {0, 1, 0};
    CHMY_TRY(check_center(g, &work[0], "work.mu", CHMY_F64));
    CHMY_TRY(check_field(g, &work[1], "work.Dvx", vx, CHMY_F64));
    CHMY_TRY(check_field(g, &work[2], "work.gdx", vx, CHMY_F64));
    CHMY_TRY(check_field(g, &work[3], "work.Dvy", vy, CHMY_F64));
    CHMY_TRY(check_field(g, &work[4], "work.gdy", vy, CHMY_F64));
    for (int q = 5; q < 8; ++q) CHMY_TRY(check_center(g, &work[q], "work", CHMY_F64));
    SiaF<double> s = make_sia<double>(g, H, B, A, p);
    SiaW<double> w{view<double>(work[0]), view<double>(work[1]), view<double>(work[2]), view<double>(work[3]),
                   view<double>(work[4]), view<double>(work[5]), view<double>(work[6]), view<double>(work[7])};
    hipStream_t st = (hipStream_t)stream;
    const int64_t nx = g->n[0], ny = g->n[1];
    const DF<double> gA = A ? view<double>(*grad_logA) : view<double>(*lam_out);
    hipLaunchKernelGGL(sia_adj1<double>, sia_grid(nx, ny), sia_blk(), 0, st, s, view<double>(*lam_in), w, nx, ny, dt);
    hipLaunchKernelGGL(sia_adj2<double>, sia_grid(nx, ny), sia_blk(), 0, st, s, w, nx, ny, dt);
    hipLaunchKernelGGL(sia_adj3<double>, sia_grid(nx, ny), sia_blk(), 0, st, s, w, gA, nx, ny);
    hipLaunchKernelGGL(sia_adj4<double>, sia_grid(nx, ny), sia_blk(), 0, st, w, nx, ny);
    hipLaunchKernelGGL(sia_adj5<double>, sia_grid(nx, ny), sia_blk(), 0, st, s, w, view<double>(*lam_out), nx, ny, dt);
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_run(const chmy_grid_t* g, const chmy_field_t* states, int nstates, const chmy_field_t* B,
                 const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c,
                 const double* dts, void* stream) {
    if (nstates < 1) return set_error(CHMY_E_ARG, "nstates=%d < 1", nstates);
    CHMY_TRY(check_grid(g));
    CHMY_TRY(sia_check_states(g, states, nstates, "states"));
    CHMY_TRY(sia_check(g, &states[0], B, A, npow));
    if (nstates > 1 && !dts) return set_error(CHMY_E_ARG, "dts is NULL");
    for (int q = 0; q + 1 < nstates; ++q)
        if (states[q].data == states[q + 1].data)
            return set_error(CHMY_E_ARG, "states[%d] and states[%d] share storage", q, q + 1);
    const SiaPar p{As, npow, mb_a, mb_b, mb_c, 0.0};
    hipStream_t st = (hipStream_t)stream;
    for (int q = 0; q + 1 < nstates; ++q) {
        SiaStep<double> f = sia_stepper(g, &states[q], &states[q + 1], B, A, p, dts[q]);
        launch_step_tile(sia_tiles(g->n[0], g->n[1], 1), st, f, g->n[0], g->n[1]);
    }
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_reverse(const chmy_grid_t* g, const chmy_field_t* ckpt, int nckpt, int stride, int nsteps,
                     const chmy_field_t* seg, const chmy_field_t* lam_a, const chmy_field_t* lam_b,
                     const chmy_field_t* grad_logA, const chmy_field_t* B, const chmy_field_t* A, double As, int npow,
                     double mb_a, double mb_b, double mb_c, const double* dts, int* lam_slot, void* stream) {
    CHMY_TRY(sia_reverse_check(g, ckpt, nckpt, stride, nsteps, lam_a, grad_logA, B, A, npow, dts));
    if (!lam_slot) return set_error(CHMY_E_ARG, "lam_slot is NULL");
    if (stride > 1) CHMY_TRY(sia_check_states(g, seg, 2 * (stride - 1), "seg"));
    CHMY_TRY(check_center(g, lam_b, "lam_b", CHMY_F64));
    if (lam_a->data == lam_b->data) return set_error(CHMY_E_ARG, "lam_a and lam_b must differ");
    const SiaPar p{As, npow, mb_a, mb_b, mb_c, 0.0};
    // D of the segments' states (2 stride Center fields) in library-held device
    // memory, so that the adjoint steps read what the recompute steps evaluated
    std::vector<chmy_field_t> dbuf;
    if (stride > 1) {
        const size_t fbytes = (size_t)(pad_of(*lam_a, 0) * pad_of(*lam_a, 1)) * sizeof(double);
        const size_t fstride = (fbytes + 255) / 256 * 256;
        char* base = nullptr;
        CHMY_TRY(sia_dcache(2 * (size_t)stride * fstride, reinterpret_cast<void**>(&base)));
        for (int i = 0; i < 2 * stride; ++i) {
            chmy_field_t f = *lam_a;
            f.data = base + (size_t)i * fstride;
            dbuf.push_back(f);
        }
    }
    *lam_slot = sia_reverse_sweep(g, ckpt, stride, nsteps, [&](int i) { return seg[i]; }, lam_a, lam_b, grad_logA,
                                  B, A, p, dts, (hipStream_t)stream, dbuf.empty() ? nullptr : dbuf.data());
    CHMY_HIP(hipGetLastError());
    return 0;
}

int chmy_sia_reverse_noalloc(const chmy_grid_t* g, const chmy_field_t* ckpt, int nckpt, int stride, int nsteps,
                             const chmy_field_t* lam, const chmy_field_t* grad_logA, const chmy_field_t* B,
                             const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c,
                             const double* dts, void* stream) {
    CHMY_TRY(sia_reverse_check(g, ckpt, nckpt, stride, nsteps, lam, grad_logA, B, A, npow, dts));
    // 2 (stride - 1) segment states and the second adjoint buffer, one Center
    // storage each, every one starting on a 256 B boundary
    const int nbuf = 2 * (stride - 1) + 1;
    const size_t fbytes = (size_t)(pad_of(*lam, 0) * pad_of(*lam, 1)) * sizeof(double);
    const size_t fstride = (fbytes + 255) / 256 * 256;
    if ((size_t)nbuf > kSiaScratchBytes / fstride)
        return set_error(CHMY_E_SIZE,
                         "chmy_sia_reverse_noalloc: %d Center fields of %zu bytes (stride %d) exceed its %zu-byte "
                         "device scratch; use chmy_sia_reverse with caller-allocated seg and lam_b",
                         nbuf, fstride, stride, kSiaScratchBytes);
    char* base = nullptr;
    CHMY_TRY(sia_scratch_ptr(reinterpret_cast<void**>(&base)));
    hipStream_t st = (hipStream_t)stream;
    // the caller-allocated workspace of chmy_sia_reverse starts at zero; so
    // does this one, whatever the previous call on this device left in it
    const int64_t nz = (int64_t)(nbuf * fstride / sizeof(double));
    hipLaunchKernelGGL(sia_fill_zero, dim3((unsigned)std::min<int64_t>((nz + 255) / 256, 2048)), dim3(256), 0, st,
                       reinterpret_cast<double*>(base), nz);
    auto buf = [&](int i) {
        chmy_field_t f = *lam;
        f.data = base + (size_t)i * fstride;
        return f;
    };
    const chmy_field_t lam_b = buf(nbuf - 1);
    const SiaPar p{As, npow, mb_a, mb_b, mb_c, 0.0};
    const int slot = sia_reverse_sweep(g, ckpt, stride, nsteps, buf, lam, &lam_b, grad_logA, B, A, p, dts, st);
    if (slot == 1) CHMY_HIP(hipMemcpyAsync(lam->data, lam_b.data, fbytes, hipMemcpyDeviceToDevice, st));
    CHMY_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
