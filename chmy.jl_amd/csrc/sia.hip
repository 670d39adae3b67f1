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
    if (!work) return set_error(CHMY_E_ARG, "work is NULL");
    const int vx[3] = {1, 0, 0}, vy[3] = {0, 1, 0};
    CHMY_TRY(check_center(g, &work[0], "work.mu", CHMY_F64));
    CHMY_TRY(check_field(g, &work[1], "work.Dvx", vx, CHMY_F64));
    CHMY_TRY(check_field(g, &work[2], "work.gdx", vx, CHMY_F64));
    CHMY_TRY(check_field(g, &work[3], "work.Dvy", vy, CHMY_F64));
    CHMY_TRY(check_field(g, &work[4], "work.gdy", vy, CHMY_F64));
    for (int q = 5; q < 8; ++q) CHMY_TRY(check_center(g, &work[q], "work", CHMY_F64));
    SiaPar p{As, npow, mb_a, mb_b, mb_c, dt};
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

}  // extern "C"
