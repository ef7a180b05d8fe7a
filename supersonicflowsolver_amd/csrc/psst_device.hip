// psst_device.hip -- gfx950 kernels and the device context behind the C-ABI
// in include/psst_gpu.h.
//
// One iteration of the reference's explicit_solve (psstFiniteVolumeGrid.cpp:305-397):
//   k_grad_*      gradients of (u, p, T) once per iteration (second order)
//   per RK stage:
//     k_flux      one thread per face: reconstruction + Riemann flux, F*A
//     k_update    one thread per interior cell: residual R = ordered gather of
//                 the +-F*A of its faces, time step, RK update, primitive
//                 recovery, residual max-norm
//                 (k_update_ordered_*: the reference's in-place sweep for the
//                 neighbour-dependent time step, level-scheduled)
//     k_bc        one thread per ghost: boundary state
// Data layout: structure of arrays in HBM.  Geometry is FP64 always; the
// state type T is double (parity) or float (FP32 mode).  Prim buffers are
// ping-ponged so a stage reads its start state while writing the new one.
//
// Bitwise parity with the reference in FP64 rests on: the same association
// of every expression (each kernel function names the lines it restates),
// no FMA contraction (-ffp-contract=off), IEEE-correct division and square
// root, and sums taken in the reference's order (per-cell CSR lists; no
// floating-point atomics).  The residual max-norm is order-free and uses
// integer atomics on the bit patterns of non-negative doubles.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "psst_gpu.h"
#include "tiles.hpp"

namespace psst {
void log_phase(const char *what);
}

namespace {

constexpr int kBlock = 256;
constexpr int kResSlots = 64;  // residual max-norm is spread over this many slots per column

// ---------------------------------------------------------------------------
// Device pointer bundles
// ---------------------------------------------------------------------------
struct Geom {
    int N, G, NT, F;
    const double *__restrict__ vol;   // [NT]
    const double *__restrict__ cx;    // [3][NT]
    const double *__restrict__ fn;    // [3][F]
    const double *__restrict__ fa;    // [F]
    const double *__restrict__ beta;  // [F]
    const int *__restrict__ f0;       // [F]
    const int *__restrict__ f1;       // [F]
    const int *__restrict__ voff;     // [N+1]
    const int *__restrict__ vface;    // [nnz]
    const int *__restrict__ loff;     // [N+1]
    const int *__restrict__ lface;    // [nnz]
    const int *__restrict__ lnb;      // [nnz] neighbour across lface
    const double *__restrict__ lsqc;  // [3][nnz] least-squares weights w*theta_k per list entry
    const double *__restrict__ S;     // [3][N] projected areas (time step mode 2)
    const int *__restrict__ gface;    // [G]
    const int *__restrict__ gnb;      // [G]
    const int *__restrict__ gbc;      // [G]
    const int *__restrict__ rank;     // [NT] position in the reference's update loop (ghosts: >= N)
    const double *__restrict__ rlr;   // [3][F] x_0 - x_1 per face (neighbour 1 in neighbour 0's frame)
};

struct BCDev {
    int dyn, th;
    double u[3], p, T;
};

struct Phys {
    double gamma, rm, cp, cv, cfl, dt_uniform, venkat_eps2;
};

// Gradients are stored cell-major ("AoS"): the (dim+2) x dim components of a
// cell are contiguous, padded to an even count so that a cell's gradients
// are 16-byte aligned and load as a few wide accesses.
template <int DIM>
struct GradLayout {
    static constexpr int NC = (DIM + 2) * DIM;       // components per cell
    static constexpr int STRIDE = (NC + 1) & ~1;     // padded
};

template <typename T>
struct State {
    const T *__restrict__ q;  // [(dim+3)][NT]  stage-start primitives
    T *__restrict__ qn;       // [(dim+3)][NT]  new primitives
    T *__restrict__ W0;       // [(dim+2)][N]   conservative state at iteration start
    T *__restrict__ grad;     // [N][stride]: component (var i, axis k) at i*dim + k
    T *__restrict__ FA;       // [(dim+2)][F]
};

// Store one cell's gradients (cell-major layout), wide stores.
template <int DIM, typename T>
__device__ __forceinline__ void store_grads(T *__restrict__ gr, int cell, const T (&acc)[DIM + 2][DIM]) {
    constexpr int S = GradLayout<DIM>::STRIDE, NC = GradLayout<DIM>::NC;
    T flat[S];
#pragma unroll
    for (int n = 0; n < DIM + 2; ++n)
#pragma unroll
        for (int k = 0; k < DIM; ++k) flat[n * DIM + k] = acc[n][k];
#pragma unroll
    for (int i = NC; i < S; ++i) flat[i] = T(0);
    if (sizeof(T) == 8) {
        double2 *p = reinterpret_cast<double2 *>(gr + (size_t)cell * S);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) p[i] = make_double2((double)flat[2 * i], (double)flat[2 * i + 1]);
    } else {
        float2 *p = reinterpret_cast<float2 *>(gr + (size_t)cell * S);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) p[i] = make_float2((float)flat[2 * i], (float)flat[2 * i + 1]);
    }
}

// Gradients of one cell (cell-major layout) into registers, wide loads.
template <int DIM, typename T>
__device__ __forceinline__ void load_grads(const T *__restrict__ gr, int cell, T (&g)[GradLayout<DIM>::STRIDE]) {
    constexpr int S = GradLayout<DIM>::STRIDE;
    if (sizeof(T) == 8) {
        const double2 *p = reinterpret_cast<const double2 *>(gr + (size_t)cell * S);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) {
            double2 v = p[i];
            g[2 * i] = (T)v.x;
            g[2 * i + 1] = (T)v.y;
        }
    } else {
        const float2 *p = reinterpret_cast<const float2 *>(gr + (size_t)cell * S);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) {
            float2 v = p[i];
            g[2 * i] = (T)v.x;
            g[2 * i + 1] = (T)v.y;
        }
    }
}

// ---------------------------------------------------------------------------
// Small device helpers
// ---------------------------------------------------------------------------
// Division and square root of the state type: IEEE-correct for double (the
// parity path); the hardware's approximate forms for float (FP32 mode makes
// no bitwise claim, and they cost a fraction of the correctly rounded
// sequences).
template <typename T>
__device__ __forceinline__ T dsqrt(T x) {
    return sqrt(x);
}
template <>
__device__ __forceinline__ float dsqrt<float>(float x) {
    return __builtin_amdgcn_sqrtf(x);
}
template <typename T>
__device__ __forceinline__ T fdiv(T a, T b) {
    return a / b;
}
template <>
__device__ __forceinline__ float fdiv<float>(float a, float b) {
    return __fdividef(a, b);
}

// psstScalarLimiters.cpp:17-93; eps is absolute (physical units).
// Venkatakrishnan (limiter 4, no reference counterpart): the smooth ratio
// form psi = (a^2 + 2ab + eps2) / (a^2 + ab + 2b^2 + eps2), a = num, b = den,
// which without eps2 is (r^2 + 2r) / (r^2 + r + 2) in r = a/b; eps2 =
// (K h)^3 with h the mean cell size, as in Venkatakrishnan's cell limiter.
template <typename T, int LIM>
__device__ __forceinline__ T limiter(T num, T den, T eps2) {
    const T eps = T(1.0e-10);
    if ((num > T(0) && den < T(0)) || (num < T(0) && den > T(0))) return T(0);
    if (LIM == PSST_LIMITER_VENKATAKRISHNAN) {
        if (fabs(den) < eps) return T(1);
        const T nn = num * num, nd = num * den, dd = den * den;
        return fdiv(nn + T(2) * nd + eps2, nn + nd + T(2) * dd + eps2);
    } else if (LIM == PSST_LIMITER_VAN_ALBADA) {
        if (fabs(den) < eps) return T(1);
        T r = fdiv(num, den);
        return fdiv(r * r + r, r * r + T(1));
    } else if (LIM == PSST_LIMITER_MINMOD) {
        if (fabs(den) < eps) return T(1);
        T r = fdiv(num, den);
        T res = r;
        if (r > T(1)) res = T(1);
        return res;
    } else {  // Van Leer
        if (fabs(den) < eps) return T(2);
        T r = fdiv(num, den);
        return fdiv(T(2) * r, T(1) + r);
    }
}

// psstConvectiveFlow.cpp:13-24
template <typename T>
__device__ __forceinline__ T roe_avg(T lv, T rv, T ld, T rd) {
    T l = dsqrt(ld);
    T r = dsqrt(rd);
    T num = lv * l + rv * r;
    T den = (l + r);
    return fdiv(num, den);
}

// ---------------------------------------------------------------------------
// Gradients
// ---------------------------------------------------------------------------
// psstGreenGauss::calculate (psstGradCalculation.cpp:80-177), per interior
// cell: sum over the cell's faces in face-loop order of
// ((+-phi_f) * (A/V)) * n_k with phi_f = beta*phi_0 + (1-beta)*phi_1.
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_grad_gg(Geom g, State<T> s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    constexpr int NV = DIM + 2;
    T acc[NV][DIM];
#pragma unroll
    for (int n = 0; n < NV; ++n)
#pragma unroll
        for (int k = 0; k < DIM; ++k) acc[n][k] = T(0);
    const double Vc = g.vol[c];
    const int b = g.voff[c], e = g.voff[c + 1];
    for (int i = b; i < e; ++i) {
        const int f = g.vface[i];
        const int a0 = g.f0[f], a1 = g.f1[f];
        const T geom = T(g.fa[f] / Vc);
        const T beta = T(g.beta[f]);
        const T rbeta = T(1) - beta;
        T nrm[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) nrm[k] = T(g.fn[(size_t)k * g.F + f]);
        const bool first = (a0 == c);
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            T phi = beta * s.q[(size_t)n * g.NT + a0] + rbeta * s.q[(size_t)n * g.NT + a1];
            T sp = first ? -phi : phi;
#pragma unroll
            for (int k = 0; k < DIM; ++k) acc[n][k] += sp * geom * nrm[k];
        }
    }
    store_grads<DIM, T>(s.grad, c, acc);
}

// psstLeastSquare::calculate (psstGradCalculation.cpp:268-368), per interior
// cell over its face list; w*theta_k is precomputed on the host with the
// reference's operation order, so each term is (w*theta_k) * (phi_nb - phi_c).
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_grad_lsq(Geom g, State<T> s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    constexpr int NV = DIM + 2;
    T acc[NV][DIM];
    T qc[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        qc[n] = s.q[(size_t)n * g.NT + c];
#pragma unroll
        for (int k = 0; k < DIM; ++k) acc[n][k] = T(0);
    }
    const int b = g.loff[c], e = g.loff[c + 1];
    const int nnz = g.loff[g.N];
    for (int i = b; i < e; ++i) {
        const int nb = g.lnb[i];
        T w[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) w[k] = T(g.lsqc[(size_t)k * nnz + i]);
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            T d = (s.q[(size_t)n * g.NT + nb] - qc[n]);
#pragma unroll
            for (int k = 0; k < DIM; ++k) acc[n][k] += w[k] * d;
        }
    }
    store_grads<DIM, T>(s.grad, c, acc);
}

// ---------------------------------------------------------------------------
// Face fluxes
// ---------------------------------------------------------------------------
template <int DIM, typename T>
struct Side {
    T u[DIM], p, T_, rho;
};

// psstReconstructorDM::reconstruct_vals (psstValueReconstructor.cpp:70-134).
// G0/G1 return the gradient component (var i, axis k) of neighbour 0 / 1;
// RLR = x_0 - x_1.
template <int DIM, typename T, int LIM, typename G0, typename G1>
__device__ __forceinline__ void reconstruct_dm(Side<DIM, T> &r0, Side<DIM, T> &r1, const T *RLR, T rm, T eps2, G0 g0,
                                               G1 g1) {
    constexpr int NV = DIM + 2;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        T *p0 = (i < DIM) ? &r0.u[i] : (i == DIM ? &r0.p : &r0.T_);
        T *p1 = (i < DIM) ? &r1.u[i] : (i == DIM ? &r1.p : &r1.T_);
        T dl = T(0), dr = T(0);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            dl += g1(i, k) * RLR[k];
            dr += g0(i, k) * RLR[k];
        }
        const T v0 = *p0, v1 = *p1;
        T vlm = v1 - (T(2) * dl - (v0 - v1));
        T vlp = v0;
        T vrm = v1;
        T vrp = v0 + (T(2) * dr - (v0 - v1));
        T nl = vlp - v1;
        T dnl = v1 - vlm;
        T nr = vrm - v0;
        T dnr = v0 - vrp;
        *p0 = v0 + T(0.5) * limiter<T, LIM>(nr, dnr, eps2) * dnr;
        *p1 = v1 + T(0.5) * limiter<T, LIM>(nl, dnl, eps2) * dnl;
    }
    r0.rho = fdiv(r0.p, rm * r0.T_);
    r1.rho = fdiv(r1.p, rm * r1.T_);
}

template <int DIM, typename T>
__device__ __forceinline__ void load_side(const State<T> &s, int NT, int a, Side<DIM, T> &r) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) r.u[k] = s.q[(size_t)k * NT + a];
    r.p = s.q[(size_t)DIM * NT + a];
    r.T_ = s.q[(size_t)(DIM + 1) * NT + a];
    r.rho = s.q[(size_t)(DIM + 2) * NT + a];
}

// Boundary face: physical flux of the ghost state (quirk Q5),
// psstConvectiveHLL::calculate :139-201 (same in HLLC and AUSM).
template <int DIM, typename T>
__device__ __forceinline__ void ghost_flux(const Side<DIM, T> &gs, const T *nrm, T cp, T *flow) {
    T vn = T(0), vsq = T(0);
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        vn += gs.u[k] * nrm[k];
        vsq += gs.u[k] * gs.u[k];
    }
    T H = cp * gs.T_ + T(0.5) * vsq;
#pragma unroll
    for (int k = 0; k < DIM; ++k) flow[k] = gs.rho * gs.u[k] * vn + gs.p * nrm[k];
    flow[DIM] = gs.rho * vn;
    flow[DIM + 1] = gs.rho * vn * H;
}

// Physical flux F(q).n and conservative state of one side
// (psstConvectiveHLL::calculate :235-244).
template <int DIM, typename T>
__device__ __forceinline__ void side_flux(const Side<DIM, T> &r, T vn, T H, const T *nrm, T *fi) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) fi[k] = r.rho * r.u[k] * vn + r.p * nrm[k];
    fi[DIM] = r.rho * vn;
    fi[DIM + 1] = r.rho * vn * H;
}
template <int DIM, typename T>
__device__ __forceinline__ void side_state(const Side<DIM, T> &r, T E, T *wi) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) wi[k] = r.rho * r.u[k];
    wi[DIM] = r.rho;
    wi[DIM + 1] = r.rho * E;
}

// Interior face: Riemann flux of the two (reconstructed) states, r[0] on the
// side the normal points to.  HLL psstConvectiveFlow.cpp:215-286, HLLC
// :499-585, AUSM :782-851.  Each branch forms only the side fluxes and
// states it uses; every value it forms is computed exactly as the reference
// computes it.
template <int DIM, typename T, int FLUX>
__device__ __forceinline__ void riemann_flux(const Side<DIM, T> (&r)[2], const T *nrm, const Phys &ph, T *flow) {
    constexpr int WS = DIM + 2;
    const T cp = T(ph.cp), cv = T(ph.cv), rm = T(ph.rm), gam = T(ph.gamma);
    T vn[2], a[2], H[2], vsq2[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        vn[j] = T(0);
        T vsq = T(0);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            vn[j] += r[j].u[k] * nrm[k];
            vsq += r[j].u[k] * r[j].u[k];
        }
        vsq2[j] = vsq;
        a[j] = dsqrt(gam * rm * r[j].T_);
        H[j] = cp * r[j].T_ + T(0.5) * vsq;
    }
    if (FLUX == PSST_FLUX_HLL || FLUX == PSST_FLUX_HLLC) {
        T Hr = roe_avg(H[1], H[0], r[1].rho, r[0].rho);
        T vrn = T(0), vsq = T(0);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            T vr = roe_avg(r[1].u[k], r[0].u[k], r[1].rho, r[0].rho);
            vrn += vr * nrm[k];
            vsq += vr * vr;
        }
        T ar = dsqrt((gam - T(1)) * (Hr - T(0.5) * vsq));
        T s1 = vn[1] - a[1];
        T tmp = vrn - ar;
        if (s1 > tmp) s1 = tmp;
        T s0 = vn[0] + a[0];
        tmp = vrn + ar;
        if (s0 < tmp) s0 = tmp;
        if (s1 >= T(0)) {
            side_flux<DIM, T>(r[1], vn[1], H[1], nrm, flow);
            return;
        }
        if (FLUX == PSST_FLUX_HLL) {
            if (s1 <= T(0) && s0 >= T(0)) {
                T fi0[WS], fi1[WS], wi0[WS], wi1[WS];
                side_flux<DIM, T>(r[0], vn[0], H[0], nrm, fi0);
                side_flux<DIM, T>(r[1], vn[1], H[1], nrm, fi1);
                side_state<DIM, T>(r[0], cv * r[0].T_ + T(0.5) * vsq2[0], wi0);
                side_state<DIM, T>(r[1], cv * r[1].T_ + T(0.5) * vsq2[1], wi1);
#pragma unroll
                for (int j = 0; j < WS; ++j)
                    flow[j] = fdiv(s0 * fi1[j] - s1 * fi0[j] + s0 * s1 * (wi0[j] - wi1[j]), s0 - s1);
            } else {
                // s0 <= 0, or NaN speeds (the reference throws, quirk Q13; the
                // NaN propagates into the state and psst_run reports it)
                side_flux<DIM, T>(r[0], vn[0], H[0], nrm, flow);
            }
            return;
        }
        T sh = fdiv(r[0].p - r[1].p + r[1].rho * vn[1] * (s1 - vn[1]) - r[0].rho * vn[0] * (s0 - vn[0]),
                    r[1].rho * (s1 - vn[1]) - r[0].rho * (s0 - vn[0]));
        if ((s1 <= T(0) && sh >= T(0)) || (sh <= T(0) && s0 >= T(0))) {
            const int K = (s1 <= T(0) && sh >= T(0)) ? 1 : 0;
            const Side<DIM, T> &rk = r[K];
            const T sK = K ? s1 : s0;
            const T vK = vn[K];
            T fiK[WS], wiK[WS];
            side_flux<DIM, T>(rk, vK, H[K], nrm, fiK);
            side_state<DIM, T>(rk, cv * rk.T_ + T(0.5) * vsq2[K], wiK);
            const T tt = sK - sh;
            const T PK = rk.p + rk.rho * (sK - vK) * (sh - vK);
#pragma unroll
            for (int j = 0; j < DIM; ++j) flow[j] = fdiv(sh * (sK * wiK[j] - fiK[j]) + sK * PK * nrm[j], tt);
            flow[DIM] = fdiv(sh * (sK * wiK[DIM] - fiK[DIM]), tt);
            flow[DIM + 1] = fdiv(sh * (sK * wiK[DIM + 1] - fiK[DIM + 1]) + sK * PK * sh, tt);
        } else {
            side_flux<DIM, T>(r[0], vn[0], H[0], nrm, flow);
        }
    } else {
        T M[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) M[j] = fdiv(vn[j], a[j]);
        T Ms1, ps1, Ms0, ps0;
        if (M[1] >= T(1)) {
            Ms1 = M[1];
            ps1 = r[1].p;
        } else if (M[1] <= T(-1)) {
            Ms1 = T(0);
            ps1 = T(0);
        } else {
            Ms1 = T(0.25) * (M[1] + T(1)) * (M[1] + T(1));
            ps1 = T(0.25) * r[1].p * (M[1] + T(1)) * (M[1] + T(1)) * (T(2) - M[1]);
        }
        if (M[0] >= T(1)) {
            Ms0 = T(0);
            ps0 = T(0);
        } else if (M[0] <= T(-1)) {
            Ms0 = M[0];
            ps0 = r[0].p;
        } else {
            Ms0 = T(-0.25) * (M[0] - T(1)) * (M[0] - T(1));
            ps0 = T(0.25) * r[0].p * (M[0] - T(1)) * (M[0] - T(1)) * (T(2) + M[0]);
        }
        T Mf = Ms0 + Ms1;
        T pf = ps0 + ps1;
        const int K = (Mf > T(0)) ? 1 : 0;
        T tmp = Mf * a[K] * r[K].rho;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            T pflow = pf * nrm[j];
            flow[j] = tmp * r[K].u[j] + pflow;
        }
        flow[DIM] = tmp;
        flow[DIM + 1] = tmp * H[K];
    }
}

// One face: FA[k][f] = F_k * A, the flux from neighbour 1 into neighbour 0
// (R_0 -= FA, R_1 += FA).
template <int DIM, typename T, int FLUX, int ORDER, int LIM>
__global__ void __launch_bounds__(kBlock) k_flux(Geom g, State<T> s, Phys ph) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    constexpr int WS = DIM + 2;
    const int a0 = g.f0[f], a1 = g.f1[f];
    T nrm[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) nrm[k] = T(g.fn[(size_t)k * g.F + f]);
    const T A = T(g.fa[f]);
    T flow[WS];
    const bool g0 = a0 >= g.N, g1 = a1 >= g.N;
    if (g0 || g1) {
        Side<DIM, T> gs;
        load_side<DIM, T>(s, g.NT, g0 ? a0 : a1, gs);
        ghost_flux<DIM, T>(gs, nrm, T(ph.cp), flow);
    } else {
        Side<DIM, T> r[2];
        load_side<DIM, T>(s, g.NT, a0, r[0]);
        load_side<DIM, T>(s, g.NT, a1, r[1]);
        if (ORDER == 2) {
            T RLR[DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) RLR[k] = T(g.rlr[(size_t)k * g.F + f]);
            const T *gr = s.grad;
            reconstruct_dm<DIM, T, LIM>(
                r[0], r[1], RLR, T(ph.rm), T(ph.venkat_eps2),
                [&](int i, int k) { return gr[(size_t)a0 * GradLayout<DIM>::STRIDE + i * DIM + k]; },
                [&](int i, int k) { return gr[(size_t)a1 * GradLayout<DIM>::STRIDE + i * DIM + k]; });
        }
        riemann_flux<DIM, T, FLUX>(r, nrm, ph, flow);
    }
#pragma unroll
    for (int k = 0; k < WS; ++k) s.FA[(size_t)k * g.F + f] = flow[k] * A;
}

// ---------------------------------------------------------------------------
// Time step, update, primitives
// ---------------------------------------------------------------------------
// psstLocalAccTime1::calculate_at (psstLocalTimeCalculation.cpp:90-131).  The
// running normal-velocity average is not reset between faces (quirk Q3).
// Neighbour states are read through `nbq`, which decides between the new and
// the stage-start buffer.
template <int DIM, typename T, typename NbQ>
__device__ __forceinline__ T dt_local1(const Geom &g, int c, const T *uc, T Tc, const Phys &ph, NbQ nbq) {
    const T gr = T(ph.gamma) * T(ph.rm);
    T cc = dsqrt(gr * Tc);
    T den = T(0), nn = T(0);
    const int b = g.loff[c], e = g.loff[c + 1];
    for (int i = b; i < e; ++i) {
        const int f = g.lface[i];
        const int nb = g.lnb[i];
        T nv = T(0);
        T Tn, un[DIM];
        nbq(nb, un, Tn);
        T cn = (dsqrt(gr * Tn) + cc) * T(0.5);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            T t = T(g.fn[(size_t)k * g.F + f]);
            nv += uc[k] * t;
            nn += un[k] * t;
        }
        nn = (nn + nv) * T(0.5);
        den += (fabs(nn) + cn) * T(g.fa[f]);
    }
    return fdiv(T(ph.cfl) * T(g.vol[c]), den);
}

// psstLocalAccTime2::calculate_at (psstLocalTimeCalculation.cpp:175-213);
// S_k = sum_f A|n_k| is precomputed on the host in face-list order.
template <int DIM, typename T>
__device__ __forceinline__ T dt_local2(const Geom &g, int c, const T *uc, T Tc, const Phys &ph) {
    T cc = dsqrt(T(ph.gamma) * T(ph.rm) * Tc);
    T den = T(0);
#pragma unroll
    for (int j = 0; j < DIM; ++j) den += (fabs(uc[j]) + cc) * T(g.S[(size_t)j * g.N + c]) * T(0.5);
    return fdiv(T(ph.cfl) * T(g.vol[c]), den);
}

template <typename T>
__device__ __forceinline__ unsigned long long abs_bits(T x) {
    double d = fabs((double)x);
    return (unsigned long long)__double_as_longlong(d);
}

// Block-wide max of the |R_j| bit patterns, then one integer atomic per block
// and column into one of kResSlots slots.
template <int WS>
__device__ __forceinline__ void residual_max(unsigned long long (&m)[WS], unsigned long long *res_row) {
    __shared__ unsigned long long sm[16][WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        unsigned long long v = m[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(v, off, 64);
            v = v > o ? v : o;
        }
        m[j] = v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < WS; ++j) sm[wave][j] = m[j];
    __syncthreads();
    if (threadIdx.x < WS) {
        unsigned long long v = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) v = v > sm[w][threadIdx.x] ? v : sm[w][threadIdx.x];
        if (v) atomicMax(&res_row[threadIdx.x * kResSlots + (blockIdx.x % kResSlots)], v);
    }
}

// The body of the reference's update loop for one cell
// (psstFiniteVolumeGrid.cpp:317-347).
template <int DIM, typename T, int LC, typename NbQ>
__device__ __forceinline__ void update_cell(const Geom &g, const State<T> &s, const Phys &ph, int c, T alpha,
                                            bool first_stage, bool last_stage, unsigned long long (&rmax)[DIM + 2],
                                            NbQ nbq) {
    constexpr int WS = DIM + 2;
    // residual: ordered gather of the face contributions (face-loop order)
    T R[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) R[j] = T(0);
    const int b = g.voff[c], e = g.voff[c + 1];
    for (int i = b; i < e; ++i) {
        const int f = g.vface[i];
        const bool first = (g.f0[f] == c);
#pragma unroll
        for (int j = 0; j < WS; ++j) {
            T v = s.FA[(size_t)j * g.F + f];
            R[j] = first ? R[j] - v : R[j] + v;
        }
    }
    T uc[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) uc[k] = s.q[(size_t)k * g.NT + c];
    const T Tc = s.q[(size_t)(DIM + 1) * g.NT + c];
    T dt;
    if (LC == PSST_DT_UNIFORM) dt = T(ph.dt_uniform);
    else if (LC == PSST_DT_LOCAL1) dt = dt_local1<DIM, T>(g, c, uc, Tc, ph, nbq);
    else dt = dt_local2<DIM, T>(g, c, uc, Tc, ph);
    const T V = T(g.vol[c]);
    T w[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        T w0 = s.W0[(size_t)j * g.N + c];
        w[j] = w0 - fdiv(dt * alpha * R[j], V);
        unsigned long long ab = abs_bits(R[j]);
        rmax[j] = rmax[j] > ab ? rmax[j] : ab;
    }
    (void)first_stage;
    if (last_stage)
#pragma unroll
        for (int j = 0; j < WS; ++j) s.W0[(size_t)j * g.N + c] = w[j];
    const T rho = w[DIM];
    T vsq = T(0);
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        T u = fdiv(w[k], rho);
        s.qn[(size_t)k * g.NT + c] = u;
        vsq += u * u;
    }
    T Tn = fdiv(fdiv(w[DIM + 1], rho) - (vsq * T(0.5)), T(ph.cv));
    s.qn[(size_t)DIM * g.NT + c] = rho * T(ph.rm) * Tn;
    s.qn[(size_t)(DIM + 1) * g.NT + c] = Tn;
    s.qn[(size_t)(DIM + 2) * g.NT + c] = rho;
}

template <int DIM, typename T>
struct NbOld {
    const T *q;
    int NT;
    __device__ __forceinline__ void operator()(int nb, T *un, T &Tn) const {
#pragma unroll
        for (int k = 0; k < DIM; ++k) un[k] = q[(size_t)k * NT + nb];
        Tn = q[(size_t)(DIM + 1) * NT + nb];
    }
};

// Reference order: an interior neighbour earlier in the update loop has its
// new state already (psstFiniteVolumeGrid.cpp:317-347 is an in-place sweep).
template <int DIM, typename T>
struct NbOrdered {
    const T *q;
    const T *qn;
    const int *rank;
    int NT;
    int c;
    __device__ __forceinline__ void operator()(int nb, T *un, T &Tn) const {
        const T *src = (rank[nb] < rank[c]) ? qn : q;  // ghosts rank >= N: stage-start state
#pragma unroll
        for (int k = 0; k < DIM; ++k) un[k] = src[(size_t)k * NT + nb];
        Tn = src[(size_t)(DIM + 1) * NT + nb];
    }
};

template <int DIM, typename T, int LC>
__global__ void __launch_bounds__(kBlock) k_update(Geom g, State<T> s, Phys ph, T alpha, int first_stage,
                                                   int last_stage, unsigned long long *res_row) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long rmax[DIM + 2];
#pragma unroll
    for (int j = 0; j < DIM + 2; ++j) rmax[j] = 0;
    if (c < g.N) update_cell<DIM, T, LC>(g, s, ph, c, alpha, first_stage, last_stage, rmax, NbOld<DIM, T>{s.q, g.NT});
    residual_max<DIM + 2>(rmax, res_row);
}

// Reference-ordered update for time step mode 1: the cells of one level
// depend only on cells of lower levels.  One launch per level.
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_update_level(Geom g, State<T> s, Phys ph, T alpha, int first_stage,
                                                         int last_stage, unsigned long long *res_row,
                                                         const int *__restrict__ cells, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long rmax[DIM + 2];
#pragma unroll
    for (int j = 0; j < DIM + 2; ++j) rmax[j] = 0;
    if (i < count) {
        const int c = cells[i];
        update_cell<DIM, T, PSST_DT_LOCAL1>(g, s, ph, c, alpha, first_stage, last_stage, rmax,
                                            NbOrdered<DIM, T>{s.q, s.qn, g.rank, g.NT, c});
    }
    residual_max<DIM + 2>(rmax, res_row);
}

// Same, all levels in one workgroup with a barrier between levels (small
// meshes, where launch latency would dominate).
template <int DIM, typename T>
__global__ void __launch_bounds__(1024) k_update_levels_block(Geom g, State<T> s, Phys ph, T alpha, int first_stage,
                                                              int last_stage, unsigned long long *res_row,
                                                              const int *__restrict__ cells,
                                                              const int *__restrict__ level_off, int n_levels) {
    unsigned long long rmax[DIM + 2];
#pragma unroll
    for (int j = 0; j < DIM + 2; ++j) rmax[j] = 0;
    for (int l = 0; l < n_levels; ++l) {
        const int b = level_off[l], e = level_off[l + 1];
        for (int i = b + (int)threadIdx.x; i < e; i += (int)blockDim.x) {
            const int c = cells[i];
            update_cell<DIM, T, PSST_DT_LOCAL1>(g, s, ph, c, alpha, first_stage, last_stage, rmax,
                                                NbOrdered<DIM, T>{s.q, s.qn, g.rank, g.NT, c});
        }
        __syncthreads();
    }
    // block-wide max over 16 waves
    __shared__ unsigned long long sm[16][DIM + 2];
#pragma unroll
    for (int j = 0; j < DIM + 2; ++j) {
        unsigned long long v = rmax[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(v, off, 64);
            v = v > o ? v : o;
        }
        rmax[j] = v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < DIM + 2; ++j) sm[wave][j] = rmax[j];
    __syncthreads();
    if (threadIdx.x < DIM + 2) {
        unsigned long long v = 0;
        for (int w = 0; w < (int)blockDim.x / 64; ++w) v = v > sm[w][threadIdx.x] ? v : sm[w][threadIdx.x];
        if (v) atomicMax(&res_row[threadIdx.x * kResSlots], v);
    }
}

// Tables of the ordered sweep on large meshes, by position in the level
// order (so a level's loads are contiguous): the cell, its volume, and its
// face list as fixed-width entries (neighbour device id, with bit 30 set
// when the neighbour comes earlier in the reference's update loop; -1 past
// the end), face normals and areas.  R holds each cell's residual.
struct OrderedTables {
    const int *__restrict__ cell;   // [N]
    const double *__restrict__ vol; // [N]
    const int *__restrict__ nb;     // [W][N]
    const double *__restrict__ fn;  // [DIM][W][N]
    const double *__restrict__ fa;  // [W][N]
    void *R;                        // [DIM+2][N], state precision
    int W;                          // at most kOrderedW
};
constexpr int kOrderedW = 8;
constexpr int kEarlier = 1 << 30;

// Pre-pass of the ordered sweep: every cell's residual (the ordered gather
// of update_cell) at its level-order position, and the residual max-norm;
// neither depends on the sweep.
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_ordered_residual(Geom g, State<T> s, OrderedTables ot,
                                                             unsigned long long *res_row) {
    constexpr int WS = DIM + 2;
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    unsigned long long rmax[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) rmax[j] = 0;
    if (p < g.N) {
        const int c = ot.cell[p];
        T R[WS];
#pragma unroll
        for (int j = 0; j < WS; ++j) R[j] = T(0);
        for (int i = g.voff[c]; i < g.voff[c + 1]; ++i) {
            const int f = g.vface[i];
            const bool first = (g.f0[f] == c);
#pragma unroll
            for (int j = 0; j < WS; ++j) {
                T v = s.FA[(size_t)j * g.F + f];
                R[j] = first ? R[j] - v : R[j] + v;
            }
        }
        T *Rp = static_cast<T *>(ot.R);
#pragma unroll
        for (int j = 0; j < WS; ++j) {
            Rp[(size_t)j * g.N + p] = R[j];
            rmax[j] = abs_bits(R[j]);
        }
    }
    residual_max<WS>(rmax, res_row);
}

// One cell of the ordered sweep from the tables: update_cell with
// dt_local1 (psstLocalAccTime1::calculate_at, psstLocalTimeCalculation.cpp:
// 90-131) in the same operation order.  Everything but the earlier
// neighbours' new state is loaded straight from the position
// (ordered_fetch), so it can be loaded before the previous level is done;
// ordered_apply loads the neighbour states and updates the cell.
template <int DIM, typename T>
struct OrderedIn {
    int c;
    T R[DIM + 2], w0[DIM + 2], uc[DIM], Tc, V;
    int nbw[kOrderedW];
    double fnw[kOrderedW][DIM], faw[kOrderedW];
};

template <int DIM, typename T>
__device__ __forceinline__ void ordered_fetch(const Geom &g, const State<T> &s, const OrderedTables &ot, int p,
                                              OrderedIn<DIM, T> &o) {
    constexpr int WS = DIM + 2;
    o.c = ot.cell[p];
    const T *Rp = static_cast<const T *>(ot.R);
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        o.R[j] = Rp[(size_t)j * g.N + p];
        o.w0[j] = s.W0[(size_t)j * g.N + o.c];
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k) o.uc[k] = s.q[(size_t)k * g.NT + o.c];
    o.Tc = s.q[(size_t)(DIM + 1) * g.NT + o.c];
    o.V = T(ot.vol[p]);
#pragma unroll
    for (int w = 0; w < kOrderedW; ++w) {
        const bool in = w < ot.W;
        o.nbw[w] = in ? ot.nb[(size_t)w * g.N + p] : -1;
        o.faw[w] = in ? ot.fa[(size_t)w * g.N + p] : 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) o.fnw[w][k] = in ? ot.fn[((size_t)k * ot.W + w) * g.N + p] : 0.0;
    }
}

template <int DIM, typename T>
__device__ __forceinline__ void ordered_apply(const Geom &g, const State<T> &s, const Phys &ph, T alpha,
                                              int last_stage, const OrderedIn<DIM, T> &o) {
    constexpr int WS = DIM + 2;
    const int c = o.c;
    // neighbour states: the new one for neighbours earlier in the update
    // loop, the stage-start one otherwise (ghosts included)
    T un[kOrderedW][DIM], Tn[kOrderedW];
#pragma unroll
    for (int w = 0; w < kOrderedW; ++w) {
        const int e = o.nbw[w] < 0 ? 0 : o.nbw[w];
        const int nb = e & (kEarlier - 1);
        const T *src = (e & kEarlier) ? s.qn : s.q;
#pragma unroll
        for (int k = 0; k < DIM; ++k) un[w][k] = o.nbw[w] < 0 ? T(0) : src[(size_t)k * g.NT + nb];
        Tn[w] = o.nbw[w] < 0 ? T(0) : src[(size_t)(DIM + 1) * g.NT + nb];
    }
    const T gr = T(ph.gamma) * T(ph.rm);
    T cc = dsqrt(gr * o.Tc);
    T den = T(0), nn = T(0);
#pragma unroll
    for (int w = 0; w < kOrderedW; ++w) {
        if (o.nbw[w] < 0) break;
        T nv = T(0);
        T cn = (dsqrt(gr * Tn[w]) + cc) * T(0.5);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            T t = T(o.fnw[w][k]);
            nv += o.uc[k] * t;
            nn += un[w][k] * t;
        }
        nn = (nn + nv) * T(0.5);
        den += (fabs(nn) + cn) * T(o.faw[w]);
    }
    const T V = o.V;
    const T dt = fdiv(T(ph.cfl) * V, den);
    T wn[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) wn[j] = o.w0[j] - fdiv(dt * alpha * o.R[j], V);
    if (last_stage)
#pragma unroll
        for (int j = 0; j < WS; ++j) s.W0[(size_t)j * g.N + c] = wn[j];
    const T rho = wn[DIM];
    T vsq = T(0);
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        T u = fdiv(wn[k], rho);
        s.qn[(size_t)k * g.NT + c] = u;
        vsq += u * u;
    }
    T Tnew = fdiv(fdiv(wn[DIM + 1], rho) - (vsq * T(0.5)), T(ph.cv));
    s.qn[(size_t)DIM * g.NT + c] = rho * T(ph.rm) * Tnew;
    s.qn[(size_t)(DIM + 1) * g.NT + c] = Tnew;
    s.qn[(size_t)(DIM + 2) * g.NT + c] = rho;
}

template <int DIM, typename T>
__device__ __forceinline__ void ordered_update(const Geom &g, const State<T> &s, const Phys &ph, T alpha,
                                               int last_stage, const OrderedTables &ot, int p) {
    OrderedIn<DIM, T> o;
    ordered_fetch<DIM, T>(g, s, ot, p, o);
    ordered_apply<DIM, T>(g, s, ph, alpha, last_stage, o);
}

// The ordered sweep as one cooperative launch over several workgroups: the
// positions of a level are spread over the grid and a grid-wide barrier
// separates the levels.  Every workgroup runs the same number of barriers;
// the cooperative launch guarantees that all of them are resident.
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_update_levels_grid(Geom g, State<T> s, Phys ph, T alpha, int last_stage,
                                                               OrderedTables ot, const int *__restrict__ level_off,
                                                               int n_levels) {
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x), nthr = (int)(gridDim.x * blockDim.x);
    // each thread's first position of the next level is fetched right after
    // its cell of this level is updated (nothing it reads is written by this
    // level), so its loads overlap the stores the grid barrier waits for
    OrderedIn<DIM, T> o;
    bool have = n_levels > 0 && level_off[0] + tid < level_off[1];
    if (have) ordered_fetch<DIM, T>(g, s, ot, level_off[0] + tid, o);
    for (int l = 0; l < n_levels; ++l) {
        const int b = level_off[l], e = level_off[l + 1];
        if (have) ordered_apply<DIM, T>(g, s, ph, alpha, last_stage, o);
        for (int p = b + tid + nthr; p < e; p += nthr) ordered_update<DIM, T>(g, s, ph, alpha, last_stage, ot, p);
        have = false;
        if (l + 1 < n_levels) {
            const int pn = level_off[l + 1] + tid;
            if (pn < level_off[l + 2]) {
                ordered_fetch<DIM, T>(g, s, ot, pn, o);
                have = true;
            }
        }
        grid.sync();
    }
}

// The ordered sweep in one workgroup, a workgroup barrier between levels
// (small meshes or narrow levels).  With the residual gathered by the
// parallel pre-pass, each level is only the time step and update chain.
template <int DIM, typename T>
__global__ void __launch_bounds__(1024) k_update_levels_wg(Geom g, State<T> s, Phys ph, T alpha, int last_stage,
                                                           OrderedTables ot, const int *__restrict__ level_off,
                                                           int n_levels) {
    for (int l = 0; l < n_levels; ++l) {
        const int b = level_off[l], e = level_off[l + 1];
        for (int p = b + (int)threadIdx.x; p < e; p += (int)blockDim.x)
            ordered_update<DIM, T>(g, s, ph, alpha, last_stage, ot, p);
        __syncthreads();
    }
}

// Same for levels of at most 256 cells, with each thread's position of the
// next level fetched as soon as its cell of this level is updated: a
// workgroup barrier waits for LDS operations only, so those loads stay in
// flight across it.  Nothing they read is written by the level before
// (tables, residuals, the stage-start state, the cell's own snapshot).
template <int DIM, typename T>
__global__ void __launch_bounds__(256) k_update_levels_wg_prefetch(Geom g, State<T> s, Phys ph, T alpha,
                                                                   int last_stage, OrderedTables ot,
                                                                   const int *__restrict__ level_off, int n_levels) {
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    OrderedIn<DIM, T> o;
    bool have = n_levels > 0 && level_off[0] + tid < level_off[1];
    if (have) ordered_fetch<DIM, T>(g, s, ot, level_off[0] + tid, o);
    for (int l = 0; l < n_levels; ++l) {
        const int b = level_off[l], e = level_off[l + 1];
        if (have) ordered_apply<DIM, T>(g, s, ph, alpha, last_stage, o);
        for (int p = b + tid + nthr; p < e; p += nthr) ordered_update<DIM, T>(g, s, ph, alpha, last_stage, ot, p);
        have = false;
        if (l + 1 < n_levels) {
            const int pn = level_off[l + 1] + tid;
            if (pn < level_off[l + 2]) {
                ordered_fetch<DIM, T>(g, s, ot, pn, o);
                have = true;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Boundary conditions (psstBoundaryCondition.cpp; ghost density
// psstFiniteVolumeGrid.cpp:364).  Writes the ghost part of `qn` from the
// interior part of `qn`.
// ---------------------------------------------------------------------------
template <int DIM, typename T>
__device__ __forceinline__ void bc_apply(const Geom &g, T *__restrict__ qn, const BCDev *__restrict__ bcs,
                                         const Phys &ph, int i) {
    const int gc = g.N + i;
    const int c = g.gnb[i];
    const BCDev bc = bcs[g.gbc[i]];
    const size_t NT = (size_t)g.NT;
    T u[DIM], p, Tt;
#pragma unroll
    for (int k = 0; k < DIM; ++k) u[k] = qn[k * NT + gc];
    p = qn[DIM * NT + gc];
    Tt = qn[(DIM + 1) * NT + gc];
    switch (bc.dyn) {
    case PSST_BC_SS_OUTLET:  // :72-80
#pragma unroll
        for (int k = 0; k < DIM; ++k) u[k] = qn[k * NT + c];
        p = qn[DIM * NT + c];
        break;
    case PSST_BC_SLIP_WALL: {  // :108-127
        const int f = g.gface[i];
        T vn = T(0);
        T nrm[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            nrm[k] = T(g.fn[(size_t)k * g.F + f]);
            vn += qn[k * NT + c] * nrm[k];
        }
#pragma unroll
        for (int k = 0; k < DIM; ++k) u[k] = qn[k * NT + c] - vn * nrm[k];
        p = qn[DIM * NT + c];
        break;
    }
    case PSST_BC_OUTLET:  // :212-219
#pragma unroll
        for (int k = 0; k < DIM; ++k) u[k] = qn[k * NT + c];
        break;
    default:  // supersonic inlet: fixed state
        break;
    }
    if (bc.th == PSST_BCT_OUTLET || bc.th == PSST_BCT_WALL_Q) Tt = qn[(DIM + 1) * NT + c];  // :285-288, :354-357
#pragma unroll
    for (int k = 0; k < DIM; ++k) qn[k * NT + gc] = u[k];
    qn[DIM * NT + gc] = p;
    qn[(DIM + 1) * NT + gc] = Tt;
    qn[(DIM + 2) * NT + gc] = fdiv(p, Tt * T(ph.rm));
}

template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_bc(Geom g, T *__restrict__ qn, const BCDev *__restrict__ bcs, Phys ph) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.G) return;
    bc_apply<DIM, T>(g, qn, bcs, ph, i);
}

// Conservative variables from primitives for the interior cells
// (psstFiniteVolumeGrid.cpp:279-300).
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_set_w(Geom g, const T *__restrict__ q, T *__restrict__ W0, Phys ph) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t NT = (size_t)g.NT;
    T rho = q[(DIM + 2) * NT + c];
    T vsq = T(0);
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        T t = q[j * NT + c];
        W0[(size_t)j * g.N + c] = rho * t;
        vsq += t * t;
    }
    T E = T(ph.cv) * q[(DIM + 1) * NT + c] + T(0.5) * vsq;
    W0[(size_t)DIM * g.N + c] = rho;
    W0[(size_t)(DIM + 1) * g.N + c] = rho * E;
}

__global__ void k_canary(unsigned long long *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0x5053535443414e59ull;  // "PSSTCANY"
}

// Plain streaming copy of known traffic (counter calibration).
__global__ void __launch_bounds__(kBlock) k_copy(const double4 *__restrict__ a, double4 *__restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Gather-only pass (no update): R for diagnostics.
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_gather_residual(Geom g, const T *__restrict__ FA, double *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    constexpr int WS = DIM + 2;
    T R[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) R[j] = T(0);
    for (int i = g.voff[c]; i < g.voff[c + 1]; ++i) {
        const int f = g.vface[i];
        const bool first = (g.f0[f] == c);
#pragma unroll
        for (int j = 0; j < WS; ++j) {
            T v = FA[(size_t)j * g.F + f];
            R[j] = first ? R[j] - v : R[j] + v;
        }
    }
#pragma unroll
    for (int j = 0; j < WS; ++j) out[(size_t)j * g.N + c] = (double)R[j];
}

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ a, B *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (B)a[i];
}

// ---------------------------------------------------------------------------
// Fused, tiled kernels.  A workgroup owns a tile: a contiguous range of
// (Morton-ordered) interior cells.  It stages the tile's cells and their
// face neighbours ("slots": own cells first, then the halo) in LDS, computes
// each face of the tile once, and gathers per cell in face-loop order.
// ---------------------------------------------------------------------------
struct TileDev {
    const int *__restrict__ cell_begin;  // [T+1]
    const int *__restrict__ halo_off;    // [T+1]
    const int *__restrict__ halo;        // device cell ids
    const int *__restrict__ face_off;    // [T+1]
    const unsigned *__restrict__ ends;   // [TF] slot0 | slot1 << 15 | kind << 30
    const double *__restrict__ fn;       // [DIM][TF]
    const double *__restrict__ fa;       // [TF]
    const double *__restrict__ rlr;      // [DIM][TF] x_0 - x_1
    const double *__restrict__ beta;     // [TF]
    const int *__restrict__ ghost_off;   // [T+1]
    const int *__restrict__ ghost;       // ghost device ids
    const short *__restrict__ visit;     // [Wv][N] local face | 0x4000 if neighbour 0, face-loop order
    const short *__restrict__ list;      // [Wl][N] local face, face-list order
    const double *__restrict__ lsqc;     // [DIM][Wl][N] least-squares weights
    int TF, Wv, Wl, MS, MF;              // total tile faces, ELL widths, LDS slot/face capacity
    int n_tiles;
};

// Tile of this workgroup.  Workgroups are dealt round-robin to the 8 XCDs
// (each with its own L2); giving XCD x the contiguous tile range
// [x*per, (x+1)*per) keeps neighbouring tiles -- which share halo cells -- in
// the same L2.  The grid has 8*per workgroups; those past the last tile exit.
__device__ __forceinline__ int tile_of_block(int n_tiles) {
    const int per = (n_tiles + 7) >> 3;
    return (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
}

// Work list of a batched launch.  A batch holds many independent instances
// in one union mesh; its tiles never straddle two instances.  The tiles of
// the instances still iterating form the list [0, *count).  The grid covers
// every tile of the batch (a captured graph keeps its launch shape); each of
// the 8 XCDs takes a contiguous eighth of the active list (neighbouring tiles
// share halo cells, so they are best served by the same L2), one tile per
// workgroup, and the workgroups beyond the active list exit at once.
// Measured on the 4.9M-cell wedge as one instance: a persistent grid that
// strides through the list was 30% slower, and one that popped tiles from
// per-XCD atomic queues 3.5x slower (the pops serialise on the counters).
// list == nullptr: unbatched, one tile per workgroup.
struct BatchDev {
    const int *__restrict__ list;       // active tiles
    const int *__restrict__ count;      // number of active tiles
    const int *__restrict__ tile_inst;  // [n_tiles] instance of each tile
    unsigned long long *res;            // [n_inst][dim+2][kResSlots] residual slots of the iteration
};

// Runs the per-tile body on this workgroup's tile: tile_of_block of an
// unbatched launch, or item (b & 7) * per + (b >> 3) of the active list of a
// batched one (workgroup b runs on XCD b & 7).
template <int WS, typename Body>
__device__ __forceinline__ void for_each_tile(const TileDev &td, const BatchDev &bd, unsigned long long *res_row,
                                              Body body) {
    if (!bd.list) {
        const int t = tile_of_block(td.n_tiles);
        if (t < td.n_tiles) body(t, res_row);
        return;
    }
    const int cnt = *bd.count;
    const int per = (cnt + 7) >> 3, j = (int)(blockIdx.x >> 3);
    if (j >= per) return;
    const int item = (int)(blockIdx.x & 7u) * per + j;
    if (item >= cnt) return;
    const int t = bd.list[item];
    body(t, bd.res + (size_t)bd.tile_inst[t] * WS * kResSlots);
}

__device__ __forceinline__ int end0(unsigned e) { return (int)(e & 0x7fffu); }
__device__ __forceinline__ int end1(unsigned e) { return (int)((e >> 15) & 0x7fffu); }
__device__ __forceinline__ int endkind(unsigned e) { return (int)(e >> 30); }

// Stage the tile's slots: q (DIM+3 rows) and the slot -> cell map.
// With second order an interior cell's density is never read (face states
// take rho from p and T), so only ghost slots load it.
template <int DIM, typename T, bool ALL_RHO>
__device__ __forceinline__ void stage_slots(const Geom &g, const TileDev &td, const State<T> &s, int cb, int nc, int hb,
                                            int ns, T *sq, int *scell) {
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        const int cell = i < nc ? cb + i : td.halo[hb + i - nc];
        scell[i] = cell;
#pragma unroll
        for (int v = 0; v < DIM + 2; ++v) sq[v * td.MS + i] = s.q[(size_t)v * g.NT + cell];
        if (ALL_RHO || cell >= g.N) sq[(DIM + 2) * td.MS + i] = s.q[(size_t)(DIM + 2) * g.NT + cell];
    }
}


// One RK stage, fused: face fluxes -> per-cell ordered gather -> time step ->
// update -> primitives -> ghost states of the tile's boundary faces.
template <int DIM, typename T, int FLUX, int ORDER, int LIM, int LC>
__device__ __forceinline__ void stage_tile(const Geom &g, const TileDev &td, const State<T> &s, const Phys &ph, T alpha,
                                           int last_stage, unsigned long long *res_row, const BCDev *__restrict__ bcs,
                                           int t, unsigned char *smem) {
    constexpr int NQ = DIM + 3, NV = DIM + 2, WS = DIM + 2;
    T *sq = reinterpret_cast<T *>(smem);
    T *sfa = sq + NQ * td.MS;
    int *scell = reinterpret_cast<int *>(sfa + NV * td.MF);
    const int cb = td.cell_begin[t], nc = td.cell_begin[t + 1] - cb;
    const int hb = td.halo_off[t], ns = nc + (td.halo_off[t + 1] - hb);
    const int fb = td.face_off[t], nf = td.face_off[t + 1] - fb;
    stage_slots<DIM, T, ORDER == 1>(g, td, s, cb, nc, hb, ns, sq, scell);
    __syncthreads();
    // ---- faces of the tile ----
    for (int j = threadIdx.x; j < nf; j += blockDim.x) {
        const unsigned e = td.ends[fb + j];
        const int l0 = end0(e), l1 = end1(e), kind = endkind(e);
        T nrm[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) nrm[k] = T(td.fn[(size_t)k * td.TF + fb + j]);
        T flow[WS];
        if (kind != 0) {
            const int gsl = kind == 1 ? l0 : l1;
            Side<DIM, T> gs;
#pragma unroll
            for (int k = 0; k < DIM; ++k) gs.u[k] = sq[k * td.MS + gsl];
            gs.p = sq[DIM * td.MS + gsl];
            gs.T_ = sq[(DIM + 1) * td.MS + gsl];
            gs.rho = sq[(DIM + 2) * td.MS + gsl];
            ghost_flux<DIM, T>(gs, nrm, T(ph.cp), flow);
        } else {
            Side<DIM, T> r[2];
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                r[0].u[k] = sq[k * td.MS + l0];
                r[1].u[k] = sq[k * td.MS + l1];
            }
            r[0].p = sq[DIM * td.MS + l0];
            r[1].p = sq[DIM * td.MS + l1];
            r[0].T_ = sq[(DIM + 1) * td.MS + l0];
            r[1].T_ = sq[(DIM + 1) * td.MS + l1];
            r[0].rho = sq[(DIM + 2) * td.MS + l0];
            r[1].rho = sq[(DIM + 2) * td.MS + l1];
            if (ORDER == 2) {
                T RLR[DIM];
#pragma unroll
                for (int k = 0; k < DIM; ++k) RLR[k] = T(td.rlr[(size_t)k * td.TF + fb + j]);
                T g0[GradLayout<DIM>::STRIDE], g1[GradLayout<DIM>::STRIDE];
                load_grads<DIM, T>(s.grad, scell[l0], g0);
                load_grads<DIM, T>(s.grad, scell[l1], g1);
                reconstruct_dm<DIM, T, LIM>(
                    r[0], r[1], RLR, T(ph.rm), T(ph.venkat_eps2), [&](int i, int k) { return g0[i * DIM + k]; },
                    [&](int i, int k) { return g1[i * DIM + k]; });
            }
            riemann_flux<DIM, T, FLUX>(r, nrm, ph, flow);
        }
        const T A = T(td.fa[fb + j]);
#pragma unroll
        for (int k = 0; k < WS; ++k) sfa[k * td.MF + j] = flow[k] * A;
    }
    __syncthreads();
    // ---- cells of the tile ----
    unsigned long long rmax[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) rmax[j] = 0;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const int c = cb + i;
        T R[WS];
#pragma unroll
        for (int j = 0; j < WS; ++j) R[j] = T(0);
        // The cell's independent loads (iteration-start state, volume, face
        // sums, visit list) are issued together before the gather: measured
        // 3% (FP64) and 6% (FP32) faster than loading each where it is used.
        T w0v[WS];
#pragma unroll
        for (int j = 0; j < WS; ++j) w0v[j] = __builtin_nontemporal_load(&s.W0[(size_t)j * g.N + c]);
        const T Vh = T(g.vol[c]);
        T Sh[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) Sh[k] = LC == PSST_DT_LOCAL2 ? T(g.S[(size_t)k * g.N + c]) : T(0);
        auto gather = [&](int e) {
            const int lf = e & 0x3fff;
            const bool first = (e & 0x4000) != 0;
#pragma unroll
            for (int j = 0; j < WS; ++j) {
                const T v = sfa[j * td.MF + lf];
                R[j] = first ? R[j] - v : R[j] + v;
            }
        };
        if (td.Wv <= 8) {
            // the visit list is padded with -1 to Wv entries: all loads in bounds
            int ev[8];
#pragma unroll
            for (int w = 0; w < 8; ++w) ev[w] = w < td.Wv ? (int)td.visit[(size_t)w * g.N + c] : -1;
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                if (ev[w] < 0) break;
                gather(ev[w]);
            }
        } else {
            for (int w = 0; w < td.Wv; ++w) {
                const int e = td.visit[(size_t)w * g.N + c];
                if (e < 0) break;
                gather(e);
            }
        }
        T uc[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) uc[k] = sq[k * td.MS + i];
        const T Tc = sq[(DIM + 1) * td.MS + i];
        T dt;
        if (LC == PSST_DT_UNIFORM) {
            dt = T(ph.dt_uniform);
        } else if (LC == PSST_DT_LOCAL2) {
            // dt_local2 with the hoisted S_k and V
            T cc = dsqrt(T(ph.gamma) * T(ph.rm) * Tc);
            T den = T(0);
#pragma unroll
            for (int j = 0; j < DIM; ++j) den += (fabs(uc[j]) + cc) * Sh[j] * T(0.5);
            dt = fdiv(T(ph.cfl) * Vh, den);
        } else {
            // time step mode 1 with every neighbour at the stage-start state
            // (Jacobi; psstLocalTimeCalculation.cpp:90-131 otherwise)
            const T gr = T(ph.gamma) * T(ph.rm);
            T cc = dsqrt(gr * Tc);
            T den = T(0), nn = T(0);
            for (int w = 0; w < td.Wl; ++w) {
                const int lf = td.list[(size_t)w * g.N + c];
                if (lf < 0) break;
                const unsigned e = td.ends[fb + lf];
                const int nbs = end0(e) == i ? end1(e) : end0(e);
                T nv = T(0);
                T cn = (dsqrt(gr * sq[(DIM + 1) * td.MS + nbs]) + cc) * T(0.5);
#pragma unroll
                for (int k = 0; k < DIM; ++k) {
                    T tn = T(td.fn[(size_t)k * td.TF + fb + lf]);
                    nv += uc[k] * tn;
                    nn += sq[k * td.MS + nbs] * tn;
                }
                nn = (nn + nv) * T(0.5);
                den += (fabs(nn) + cn) * T(td.fa[fb + lf]);
            }
            dt = fdiv(T(ph.cfl) * T(g.vol[c]), den);
        }
        const T V = Vh;
        T w[WS];
#pragma unroll
        for (int j = 0; j < WS; ++j) {
            T w0 = w0v[j];
            w[j] = w0 - fdiv(dt * alpha * R[j], V);
            unsigned long long ab = abs_bits(R[j]);
            rmax[j] = rmax[j] > ab ? rmax[j] : ab;
        }
        // the new state and the iteration-start snapshot are written once
        // and not read again in this launch (the ghost pass reads only the
        // tile's own interior neighbours): non-temporal stores and loads,
        // measured 0.6% (FP64) and 1.2% (FP32) faster per iteration on the
        // 4.9M-cell wedge; the same hint on the face geometry loads was slower
        if (last_stage)
#pragma unroll
            for (int j = 0; j < WS; ++j) __builtin_nontemporal_store(w[j], &s.W0[(size_t)j * g.N + c]);
        const T rho = w[DIM];
        T vsq = T(0);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            T u = fdiv(w[k], rho);
            __builtin_nontemporal_store(u, &s.qn[(size_t)k * g.NT + c]);
            vsq += u * u;
        }
        T Tn = fdiv(fdiv(w[DIM + 1], rho) - (vsq * T(0.5)), T(ph.cv));
        __builtin_nontemporal_store(rho * T(ph.rm) * Tn, &s.qn[(size_t)DIM * g.NT + c]);
        __builtin_nontemporal_store(Tn, &s.qn[(size_t)(DIM + 1) * g.NT + c]);
        __builtin_nontemporal_store(rho, &s.qn[(size_t)(DIM + 2) * g.NT + c]);
    }
    // ---- ghosts of the tile's boundary faces (psstFiniteVolumeGrid.cpp:351-366) ----
    __syncthreads();
    const int gb = td.ghost_off[t], ng = td.ghost_off[t + 1] - gb;
    for (int i = threadIdx.x; i < ng; i += blockDim.x) bc_apply<DIM, T>(g, s.qn, bcs, ph, td.ghost[gb + i] - g.N);
    residual_max<WS>(rmax, res_row);
}

// Occupancy of the stage kernel: in 2D FP64 the compiler is asked to keep
// five waves per SIMD (96 VGPRs), which with 224-cell tiles (about 30 KB of
// LDS) gives five workgroups per CU; measured 15% faster than four waves
// on the 3840x1280 wedge.  3D and FP32 keep the compiler's choice.
template <int DIM, typename T>
struct StageWaves {
    static constexpr int value = (DIM == 2 && sizeof(T) == 8) ? 5 : 1;
};

template <int DIM, typename T, int FLUX, int ORDER, int LIM, int LC, int NTHR>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(StageWaves<DIM, T>::value))) k_stage(Geom g, TileDev td, State<T> s, Phys ph, T alpha, int last_stage,
                                                  unsigned long long *res_row, const BCDev *__restrict__ bcs,
                                                  BatchDev bd) {
    extern __shared__ __align__(16) unsigned char smem[];
    for_each_tile<DIM + 2>(td, bd, res_row, [&](int t, unsigned long long *row) {
        stage_tile<DIM, T, FLUX, ORDER, LIM, LC>(g, td, s, ph, alpha, last_stage, row, bcs, t, smem);
    });
}

// Green-Gauss gradients on a tile (psstGreenGauss::calculate,
// psstGradCalculation.cpp:80-177): q of own and halo cells in LDS, each
// cell's faces in face-loop order.
template <int DIM, typename T>
__device__ __forceinline__ void grad_gg_tile_body(const Geom &g, const TileDev &td, const State<T> &s, int t,
                                            unsigned char *smem) {
    constexpr int NV = DIM + 2;
    T *sq = reinterpret_cast<T *>(smem);
    const int cb = td.cell_begin[t], nc = td.cell_begin[t + 1] - cb;
    const int hb = td.halo_off[t], ns = nc + (td.halo_off[t + 1] - hb);
    const int fb = td.face_off[t];
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        const int cell = i < nc ? cb + i : td.halo[hb + i - nc];
#pragma unroll
        for (int v = 0; v < NV; ++v) sq[v * td.MS + i] = s.q[(size_t)v * g.NT + cell];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const int c = cb + i;
        T acc[NV][DIM];
#pragma unroll
        for (int n = 0; n < NV; ++n)
#pragma unroll
            for (int k = 0; k < DIM; ++k) acc[n][k] = T(0);
        const double Vc = g.vol[c];
        // one face's term, from its ends, area, beta and normal
        auto term = [&](unsigned e, double fa, double be, const double *fn) {
            const int l0 = end0(e), l1 = end1(e);
            const T geom = T(fa / Vc);
            const T beta = T(be);
            const T rbeta = T(1) - beta;
            const bool first = (l0 == i);
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                T phi = beta * sq[n * td.MS + l0] + rbeta * sq[n * td.MS + l1];
                T sp = first ? -phi : phi;
#pragma unroll
                for (int k = 0; k < DIM; ++k) acc[n][k] += sp * geom * T(fn[k]);
            }
        };
        auto unrolled = [&](auto width) {
            // every load of the cell's faces is issued before the first
            // term: the visit list is padded with -1 to Wv entries, and a
            // padding entry reads the tile's face 0 (in bounds) and is
            // skipped; padding only follows the cell's real faces.  On the
            // 4.9M-cell wedge: 0.278 vs 0.288 ms (FP64), 0.176 vs 0.197 ms
            // (FP32) against loading each face's data where it is used
            constexpr int U = decltype(width)::value;
            int ev[U];
#pragma unroll
            for (int w = 0; w < U; ++w) ev[w] = w < td.Wv ? (int)td.visit[(size_t)w * g.N + c] : -1;
            unsigned ee[U];
            double fa[U], be[U], fn[U][DIM];
#pragma unroll
            for (int w = 0; w < U; ++w) {
                const int lf = ev[w] < 0 ? 0 : (ev[w] & 0x3fff);
                ee[w] = td.ends[fb + lf];
                fa[w] = td.fa[fb + lf];
                be[w] = td.beta[fb + lf];
#pragma unroll
                for (int k = 0; k < DIM; ++k) fn[w][k] = td.fn[(size_t)k * td.TF + fb + lf];
            }
#pragma unroll
            for (int w = 0; w < U; ++w) {
                if (ev[w] < 0) break;
                term(ee[w], fa[w], be[w], fn[w]);
            }
        };
        if (td.Wv <= 4) {
            unrolled(std::integral_constant<int, 4>{});
        } else if (td.Wv <= 6) {
            unrolled(std::integral_constant<int, 6>{});
        } else {
            for (int w = 0; w < td.Wv; ++w) {
                const int ve = td.visit[(size_t)w * g.N + c];
                if (ve < 0) break;
                const int lf = ve & 0x3fff;
                double fn[DIM];
#pragma unroll
                for (int k = 0; k < DIM; ++k) fn[k] = td.fn[(size_t)k * td.TF + fb + lf];
                term(td.ends[fb + lf], td.fa[fb + lf], td.beta[fb + lf], fn);
            }
        }
        store_grads<DIM, T>(s.grad, c, acc);
    }
}

template <int DIM, typename T, int NTHR>
__global__ void __launch_bounds__(NTHR) k_grad_gg_tile(Geom g, TileDev td, State<T> s, BatchDev bd) {
    extern __shared__ __align__(16) unsigned char smem[];
    for_each_tile<DIM + 2>(td, bd, nullptr, [&](int t, unsigned long long *) {
        grad_gg_tile_body<DIM, T>(g, td, s, t, smem);
    });
}

// Least-squares gradients on a tile (psstLeastSquare::calculate,
// psstGradCalculation.cpp:268-368), face-list order.
template <int DIM, typename T>
__device__ __forceinline__ void grad_lsq_tile_body(const Geom &g, const TileDev &td, const State<T> &s, int t,
                                            unsigned char *smem) {
    constexpr int NV = DIM + 2;
    T *sq = reinterpret_cast<T *>(smem);
    const int cb = td.cell_begin[t], nc = td.cell_begin[t + 1] - cb;
    const int hb = td.halo_off[t], ns = nc + (td.halo_off[t + 1] - hb);
    const int fb = td.face_off[t];
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        const int cell = i < nc ? cb + i : td.halo[hb + i - nc];
#pragma unroll
        for (int v = 0; v < NV; ++v) sq[v * td.MS + i] = s.q[(size_t)v * g.NT + cell];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const int c = cb + i;
        T acc[NV][DIM];
#pragma unroll
        for (int n = 0; n < NV; ++n)
#pragma unroll
            for (int k = 0; k < DIM; ++k) acc[n][k] = T(0);
        // one neighbour's term, from the face's ends and the weights w*theta_k
        auto term = [&](unsigned e, const double *wk) {
            const int nbs = end0(e) == i ? end1(e) : end0(e);
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                T d = (sq[n * td.MS + nbs] - sq[n * td.MS + i]);
#pragma unroll
                for (int k = 0; k < DIM; ++k) acc[n][k] += T(wk[k]) * d;
            }
        };
        auto unrolled = [&](auto width) {
            // every load of the cell's neighbours is issued before the first
            // term: the face list and the weights are padded to Wl entries
            // (list entry -1), and a padding entry reads the tile's face 0
            // (in bounds) and is skipped; padding only follows real entries
            constexpr int U = decltype(width)::value;
            int lv[U];
#pragma unroll
            for (int w = 0; w < U; ++w) lv[w] = w < td.Wl ? (int)td.list[(size_t)w * g.N + c] : -1;
            unsigned ee[U];
            double wk[U][DIM];
#pragma unroll
            for (int w = 0; w < U; ++w) {
                ee[w] = td.ends[fb + (lv[w] < 0 ? 0 : lv[w])];
#pragma unroll
                for (int k = 0; k < DIM; ++k)
                    wk[w][k] = w < td.Wl ? td.lsqc[((size_t)k * td.Wl + w) * g.N + c] : 0.0;
            }
#pragma unroll
            for (int w = 0; w < U; ++w) {
                if (lv[w] < 0) break;
                term(ee[w], wk[w]);
            }
        };
        // FP64 only: in FP32 on the 4.9M-cell wedge the hoisted loads made
        // this kernel 20% slower (0.167 vs 0.139 ms); FP64 gained 4-6%
        constexpr bool hoist = std::is_same<T, double>::value;
        if (hoist && td.Wl <= 4) {
            unrolled(std::integral_constant<int, 4>{});
        } else if (hoist && td.Wl <= 6) {
            unrolled(std::integral_constant<int, 6>{});
        } else {
            for (int w = 0; w < td.Wl; ++w) {
                const int lf = td.list[(size_t)w * g.N + c];
                if (lf < 0) break;
                double wk[DIM];
#pragma unroll
                for (int k = 0; k < DIM; ++k) wk[k] = td.lsqc[((size_t)k * td.Wl + w) * g.N + c];
                term(td.ends[fb + lf], wk);
            }
        }
        store_grads<DIM, T>(s.grad, c, acc);
    }
}

template <int DIM, typename T, int NTHR>
__global__ void __launch_bounds__(NTHR) k_grad_lsq_tile(Geom g, TileDev td, State<T> s, BatchDev bd) {
    extern __shared__ __align__(16) unsigned char smem[];
    for_each_tile<DIM + 2>(td, bd, nullptr, [&](int t, unsigned long long *) {
        grad_lsq_tile_body<DIM, T>(g, td, s, t, smem);
    });
}


// ---------------------------------------------------------------------------
// Batch bookkeeping, one workgroup, run after every iteration of a batch:
// per-instance residual max-norms from the atomic slots (which it clears),
// the residual history row, the drop against the instance's reference row,
// and the freeze of every instance that reached the tolerance (or went
// non-finite).  When the active set changed it rebuilds the list of active
// tiles, in instance order.
// ---------------------------------------------------------------------------
struct BatchCheck {
    int n, H;
    unsigned long long *res;  // [n][WS][kResSlots]
    double *res0;             // [n][WS]; negative: take the next row
    double *hist;             // [H][n][WS]
    double *drop;             // [n]
    int *active, *conv, *frozen_buf, *iter;
    const int *tile_begin;    // [n+1]
    int *list, *count;
};

constexpr int kCheckThreads = 1024;

__device__ void rebuild_list(const BatchCheck &b) {
    __shared__ int s_off[kCheckThreads];
    __shared__ int s_tb[kCheckThreads], s_nt[kCheckThreads];
    int carry = 0;
    for (int base = 0; base < b.n; base += kCheckThreads) {
        const int i = base + (int)threadIdx.x;
        int nt = 0, tb = 0;
        if (i < b.n) {
            tb = b.tile_begin[i];
            nt = b.active[i] ? b.tile_begin[i + 1] - tb : 0;
        }
        s_off[threadIdx.x] = nt;
        __syncthreads();
        for (int off = 1; off < kCheckThreads; off <<= 1) {  // inclusive scan
            const int x = (int)threadIdx.x >= off ? s_off[threadIdx.x - off] : 0;
            __syncthreads();
            s_off[threadIdx.x] += x;
            __syncthreads();
        }
        const int incl = s_off[threadIdx.x], total = s_off[kCheckThreads - 1];
        __syncthreads();
        s_off[threadIdx.x] = carry + incl - nt;
        s_tb[threadIdx.x] = tb;
        s_nt[threadIdx.x] = nt;
        __syncthreads();
        const int m = min(kCheckThreads, b.n - base);
        for (int ii = 0; ii < m; ++ii) {
            const int n_t = s_nt[ii], o = s_off[ii], t0 = s_tb[ii];
            for (int k = threadIdx.x; k < n_t; k += kCheckThreads) b.list[o + k] = t0 + k;
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *b.count = carry;
}

template <int WS>
__global__ void __launch_bounds__(kCheckThreads) k_batch_check(BatchCheck b, double tol, int cur_after,
                                                               int rebuild_only) {
    __shared__ int s_changed;
    if (threadIdx.x == 0) s_changed = rebuild_only;
    __syncthreads();
    const int it = *b.iter;
    if (!rebuild_only) {
        double *hrow = b.hist + (size_t)(it % b.H) * b.n * WS;
        for (int i = threadIdx.x; i < b.n; i += kCheckThreads) {
            if (!b.active[i]) {
#pragma unroll
                for (int j = 0; j < WS; ++j) hrow[(size_t)i * WS + j] = 0.0;
                continue;
            }
            double r[WS];
            bool bad = false;
#pragma unroll
            for (int j = 0; j < WS; ++j) {
                unsigned long long *sl = b.res + ((size_t)i * WS + j) * kResSlots, mx = 0;
                for (int k = 0; k < kResSlots; ++k) {
                    const unsigned long long v = sl[k];
                    mx = mx > v ? mx : v;
                    sl[k] = 0;
                }
                r[j] = __longlong_as_double((long long)mx);
                if (!(r[j] <= 1.0e300)) bad = true;
                hrow[(size_t)i * WS + j] = r[j];
            }
            if (b.res0[(size_t)i * WS] < 0.0)
#pragma unroll
                for (int j = 0; j < WS; ++j) b.res0[(size_t)i * WS + j] = r[j];
            double d = 0.0;
#pragma unroll
            for (int j = 0; j < WS; ++j) {
                const double r0 = b.res0[(size_t)i * WS + j];
                const double q = r0 > 0.0 ? r[j] / r0 : 0.0;
                d = d > q ? d : q;
            }
            b.drop[i] = d;
            if (bad || (tol > 0.0 && d <= tol)) {
                b.active[i] = 0;
                b.conv[i] = bad ? -2 : it;
                b.frozen_buf[i] = cur_after;
                s_changed = 1;
            }
        }
    }
    __syncthreads();
    if (!rebuild_only && threadIdx.x == 0) *b.iter = it + 1;
    if (s_changed) rebuild_list(b);
}
}  // namespace

// ===========================================================================
// Context
// ===========================================================================
// Batch of independent instances sharing one union mesh (psst_batch_*).
struct BatchState {
    int n = 0, ws = 0, H = 256;
    int32_t N_tot = 0;
    std::vector<int32_t> cell_begin, ghost_begin, tile_begin;  // [n+1] union ranges (device numbering)
    std::vector<int32_t> c2d;                                  // union canonical -> device
    std::vector<char> have_state, enabled;
    std::vector<int32_t> buf_of;  // state buffer holding each instance's latest state
    int *tile_inst = nullptr, *tile_begin_dev = nullptr, *list = nullptr, *count = nullptr;
    unsigned long long *res = nullptr;
    double *res0 = nullptr, *hist = nullptr, *drop = nullptr;
    int *active = nullptr, *conv = nullptr, *frozen_buf = nullptr, *iter = nullptr;
    // pinned host copies read back at every chunk (active flags then the
    // active-tile count; the residual ring), so the copies go straight to
    // host memory
    int *h_flags = nullptr;
    double *h_ring = nullptr;
    int64_t iter_host = 0;
    bool w_dirty = false;
    bool use_graph = true;
    // Tile kernels of a run launch one workgroup per tile of the batch's
    // active-list capacity: all tiles at first, then the smallest power of
    // two (at least 64) that holds the active list when a chunk starts; a
    // captured graph per capacity and state-buffer parity.
    int grid_tiles = 0;  // 0: every tile of the batch
    std::map<int, std::array<hipGraphExec_t, 2>> exec;
    int graph_precision = 0;
    double graph_tol = -1.0;
};

struct psst_ctx {
    int dim = 2, N = 0, G = 0, F = 0, NT = 0, nnz_visit = 0, nnz_list = 0;
    psst_settings s{};
    std::vector<psst_bc> bc_host;
    Phys ph{};
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    Geom geom{};
    BCDev *bc = nullptr;
    // state (double or float)
    void *q[2] = {nullptr, nullptr};
    void *W0 = nullptr, *grad = nullptr, *FA = nullptr;
    int cur = 0;
    bool have_state = false;
    unsigned long long *res = nullptr;  // [cap][dim+2][kResSlots]
    int res_cap = 0;
    // reference-ordered time step schedule
    int *level_cells = nullptr, *level_off_dev = nullptr;
    std::vector<int> level_off;
    int max_level = 0;   // cells in the widest level
    int coop_blocks = -1;  // workgroups of the cooperative sweep (-1: not sized yet, 0: unavailable)
    OrderedTables ot{};    // tables of the cooperative sweep (ot.cell == nullptr: not built)
    // residual normalisation for the tolerance test
    std::vector<double> res0;
    char *snap = nullptr;  // state at the start of a chunk of a run with a tolerance
    // device numbering: d2c[d] = canonical index of device cell d (NT entries)
    std::vector<int32_t> d2c;
    // fused tiled path
    bool fused = false;
    TileDev td{};
    int n_tiles = 0, tile_threads = 256, tile_cells = 0;
    std::vector<int32_t> tile_cell_begin;  // [n_tiles+1] first device cell of each tile
    int max_slots = 0, max_faces = 0;
    size_t shm_stage = 0, shm_grad = 0;
    // batched launches: per-launch work list (list == nullptr when unbatched)
    BatchDev bd{};
    int persist_blocks = 0;             // grid of a persistent batched launch
    std::unique_ptr<struct BatchState> batch;
    // timing
    bool timing = false;
    struct KStat {
        int64_t count = 0;
        double total_ms = 0.0;
    };
    std::map<std::string, KStat> stats;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::string> kernel_names;
    std::string err;
};

namespace {

thread_local std::string g_create_error;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename X>
X *dev_alloc(psst_ctx *c, size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    hip_check(hipMalloc(&p, n * sizeof(X)), "hipMalloc");
    c->allocs.push_back(p);
    return static_cast<X *>(p);
}

template <typename X>
X *dev_upload(psst_ctx *c, const X *h, size_t n) {
    X *d = dev_alloc<X>(c, n);
    if (n) hip_check(hipMemcpy(d, h, n * sizeof(X), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}

hipEvent_t get_event(psst_ctx *c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

void flush_timing(psst_ctx *c) {
    if (c->pending.empty()) return;
    hip_check(hipEventSynchronize(c->pending.back().second.second), "hipEventSynchronize");
    for (auto &p : c->pending) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.second.first, p.second.second), "hipEventElapsedTime");
        auto &st = c->stats[p.first];
        st.count++;
        st.total_ms += ms;
        c->event_pool.push_back(p.second.first);
        c->event_pool.push_back(p.second.second);
    }
    c->pending.clear();
}

template <typename K, typename... Args>
void launch(psst_ctx *c, const char *name, K kernel, dim3 grid, dim3 block, Args... args) {
    if (grid.x == 0) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        if (c->pending.size() > 4096) flush_timing(c);
        e0 = get_event(c);
        e1 = get_event(c);
        hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
    }
    hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, args...);
    hip_check(hipGetLastError(), name);
    if (c->timing) {
        hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
        c->pending.push_back({name, {e0, e1}});
    }
}

// A cooperative launch (all workgroups resident, grid-wide barriers allowed).
template <typename K, typename... Args>
void launch_coop(psst_ctx *c, const char *name, K kernel, dim3 grid, dim3 block, Args... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        if (c->pending.size() > 4096) flush_timing(c);
        e0 = get_event(c);
        e1 = get_event(c);
        hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
    }
    void *argv[] = {static_cast<void *>(&args)...};
    hip_check(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kernel), grid, block, argv, 0, c->stream),
              name);
    if (c->timing) {
        hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
        c->pending.push_back({name, {e0, e1}});
    }
}

inline unsigned blocks(long n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

template <typename T>
State<T> state(psst_ctx *c) {
    State<T> st;
    st.q = static_cast<const T *>(c->q[c->cur]);
    st.qn = static_cast<T *>(c->q[c->cur ^ 1]);
    st.W0 = static_cast<T *>(c->W0);
    st.grad = static_cast<T *>(c->grad);
    st.FA = static_cast<T *>(c->FA);
    return st;
}

// ---- dispatch of the templated kernels ----
template <typename K>
void set_shm(K kernel, size_t bytes) {
    static std::map<const void *, size_t> done;  // largest size already allowed per kernel
    if (bytes <= 64 * 1024) return;
    auto it = done.find((const void *)kernel);
    if (it != done.end() && it->second >= bytes) return;
    hip_check(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
              "hipFuncSetAttribute");
    done[(const void *)kernel] = bytes;
}

template <typename K, typename... Args>
void launch_shm(psst_ctx *c, const char *name, K kernel, dim3 grid, dim3 block, size_t shm, Args... args) {
    if (grid.x == 0) return;
    set_shm(kernel, shm);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        if (c->pending.size() > 4096) flush_timing(c);
        e0 = get_event(c);
        e1 = get_event(c);
        hip_check(hipEventRecord(e0, c->stream), "hipEventRecord");
    }
    hipLaunchKernelGGL(kernel, grid, block, shm, c->stream, args...);
    hip_check(hipGetLastError(), name);
    if (c->timing) {
        hip_check(hipEventRecord(e1, c->stream), "hipEventRecord");
        c->pending.push_back({name, {e0, e1}});
    }
}

// Grid of a tile kernel: one workgroup per tile of the context (8-aligned
// for the XCD mapping), batched or not.
template <typename K>
dim3 tile_grid(psst_ctx *c, K, int, size_t) {
    const int tiles = (c->batch && c->batch->grid_tiles > 0) ? c->batch->grid_tiles : c->n_tiles;
    return dim3(8 * ((tiles + 7) / 8));
}

// Launch capacity of a batched run for `count` active tiles.
int batch_grid_tiles(int count, int n_tiles) {
    int cap = 64;
    while (cap < count && cap < n_tiles) cap <<= 1;
    return std::min(cap, n_tiles);
}


size_t stage_shm(const psst_ctx *c, size_t sb) {
    const int NQ = c->dim + 3, NV = c->dim + 2;
    return ((size_t)NQ * c->max_slots + (size_t)NV * c->max_faces) * sb + (size_t)c->max_slots * sizeof(int);
}
size_t grad_shm(const psst_ctx *c, size_t sb) { return (size_t)(c->dim + 2) * c->max_slots * sb; }

template <int DIM, typename T, int NTHR>
void run_grad_tile(psst_ctx *c) {
    State<T> st = state<T>(c);
    const size_t shm = grad_shm(c, sizeof(T));
    const BatchDev bd = c->bd;  // work list (list == nullptr: unbatched)
    if (c->s.grad == PSST_GRAD_GREEN_GAUSS) {
        auto k = k_grad_gg_tile<DIM, T, NTHR>;
        launch_shm(c, "grad_gg_tile", k, tile_grid(c, k, NTHR, shm), dim3(NTHR), shm, c->geom, c->td, st, bd);
    } else {
        auto k = k_grad_lsq_tile<DIM, T, NTHR>;
        launch_shm(c, "grad_lsq_tile", k, tile_grid(c, k, NTHR, shm), dim3(NTHR), shm, c->geom, c->td, st, bd);
    }
}

template <int DIM, typename T>
void run_grad(psst_ctx *c) {
    if (c->fused) {
        if (c->tile_threads == 512) run_grad_tile<DIM, T, 512>(c);
        else run_grad_tile<DIM, T, 256>(c);
        return;
    }
    State<T> st = state<T>(c);
    if (c->s.grad == PSST_GRAD_GREEN_GAUSS)
        launch(c, "grad_gg", k_grad_gg<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st);
    else
        launch(c, "grad_lsq", k_grad_lsq<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st);
}

template <int DIM, typename T, int FLUX, int ORDER, int LIM, int LC, int NTHR>
void run_stage_k(psst_ctx *c, int k, unsigned long long *res_row) {
    State<T> st = state<T>(c);
    const T alpha = T(c->s.rk_coeff[k]);
    const int last = (k == c->s.rk_levels - 1);
    const size_t shm = stage_shm(c, sizeof(T));
    auto kern = k_stage<DIM, T, FLUX, ORDER, LIM, LC, NTHR>;
    launch_shm(c, "stage", kern, tile_grid(c, kern, NTHR, shm), dim3(NTHR), shm, c->geom, c->td, st, c->ph, alpha,
               last, res_row, (const BCDev *)c->bc, c->bd);
}

template <int DIM, typename T, int FLUX, int ORDER, int LIM, int NTHR>
void run_stage_lc_t(psst_ctx *c, int k, unsigned long long *res_row) {
    switch (c->s.time_mode) {
    case PSST_DT_UNIFORM:
        return run_stage_k<DIM, T, FLUX, ORDER, LIM, PSST_DT_UNIFORM, NTHR>(c, k, res_row);
    case PSST_DT_LOCAL1:
        return run_stage_k<DIM, T, FLUX, ORDER, LIM, PSST_DT_LOCAL1, NTHR>(c, k, res_row);
    default:
        return run_stage_k<DIM, T, FLUX, ORDER, LIM, PSST_DT_LOCAL2, NTHR>(c, k, res_row);
    }
}

template <int DIM, typename T, int FLUX, int ORDER, int LIM>
void run_stage_lc(psst_ctx *c, int k, unsigned long long *res_row) {
    if (c->tile_threads == 512) run_stage_lc_t<DIM, T, FLUX, ORDER, LIM, 512>(c, k, res_row);
    else run_stage_lc_t<DIM, T, FLUX, ORDER, LIM, 256>(c, k, res_row);
}

template <int DIM, typename T, int FLUX, int ORDER>
void run_stage_lim(psst_ctx *c, int k, unsigned long long *res_row) {
    if (ORDER == 1) return run_stage_lc<DIM, T, FLUX, 1, PSST_LIMITER_VAN_LEER>(c, k, res_row);
    switch (c->s.limiter) {
    case PSST_LIMITER_VAN_ALBADA:
        return run_stage_lc<DIM, T, FLUX, 2, PSST_LIMITER_VAN_ALBADA>(c, k, res_row);
    case PSST_LIMITER_MINMOD:
        return run_stage_lc<DIM, T, FLUX, 2, PSST_LIMITER_MINMOD>(c, k, res_row);
    case PSST_LIMITER_VENKATAKRISHNAN:
        return run_stage_lc<DIM, T, FLUX, 2, PSST_LIMITER_VENKATAKRISHNAN>(c, k, res_row);
    default:
        return run_stage_lc<DIM, T, FLUX, 2, PSST_LIMITER_VAN_LEER>(c, k, res_row);
    }
}

template <int DIM, typename T>
void run_stage(psst_ctx *c, int k, unsigned long long *res_row) {
    const bool o2 = c->s.order == 2;
    switch (c->s.flux) {
    case PSST_FLUX_AUSM:
        return o2 ? run_stage_lim<DIM, T, PSST_FLUX_AUSM, 2>(c, k, res_row)
                  : run_stage_lim<DIM, T, PSST_FLUX_AUSM, 1>(c, k, res_row);
    case PSST_FLUX_HLL:
        return o2 ? run_stage_lim<DIM, T, PSST_FLUX_HLL, 2>(c, k, res_row)
                  : run_stage_lim<DIM, T, PSST_FLUX_HLL, 1>(c, k, res_row);
    default:
        return o2 ? run_stage_lim<DIM, T, PSST_FLUX_HLLC, 2>(c, k, res_row)
                  : run_stage_lim<DIM, T, PSST_FLUX_HLLC, 1>(c, k, res_row);
    }
}

template <int DIM, typename T, int FLUX, int ORDER>
void run_flux_lim(psst_ctx *c) {
    State<T> st = state<T>(c);
    dim3 gr(blocks(c->F)), bl(kBlock);
    if (ORDER == 1) {
        launch(c, "flux", k_flux<DIM, T, FLUX, 1, PSST_LIMITER_VAN_LEER>, gr, bl, c->geom, st, c->ph);
        return;
    }
    switch (c->s.limiter) {
    case PSST_LIMITER_VAN_ALBADA:
        launch(c, "flux", k_flux<DIM, T, FLUX, 2, PSST_LIMITER_VAN_ALBADA>, gr, bl, c->geom, st, c->ph);
        break;
    case PSST_LIMITER_MINMOD:
        launch(c, "flux", k_flux<DIM, T, FLUX, 2, PSST_LIMITER_MINMOD>, gr, bl, c->geom, st, c->ph);
        break;
    case PSST_LIMITER_VENKATAKRISHNAN:
        launch(c, "flux", k_flux<DIM, T, FLUX, 2, PSST_LIMITER_VENKATAKRISHNAN>, gr, bl, c->geom, st, c->ph);
        break;
    default:
        launch(c, "flux", k_flux<DIM, T, FLUX, 2, PSST_LIMITER_VAN_LEER>, gr, bl, c->geom, st, c->ph);
        break;
    }
}

template <int DIM, typename T>
void run_flux(psst_ctx *c) {
    const bool o2 = c->s.order == 2;
    switch (c->s.flux) {
    case PSST_FLUX_AUSM:
        o2 ? run_flux_lim<DIM, T, PSST_FLUX_AUSM, 2>(c) : run_flux_lim<DIM, T, PSST_FLUX_AUSM, 1>(c);
        break;
    case PSST_FLUX_HLL:
        o2 ? run_flux_lim<DIM, T, PSST_FLUX_HLL, 2>(c) : run_flux_lim<DIM, T, PSST_FLUX_HLL, 1>(c);
        break;
    default:
        o2 ? run_flux_lim<DIM, T, PSST_FLUX_HLLC, 2>(c) : run_flux_lim<DIM, T, PSST_FLUX_HLLC, 1>(c);
        break;
    }
}

// Workgroups of the cooperative ordered sweep: enough to hold the widest
// level at one cell per thread, within what the device keeps resident at once;
// 0 where cooperative launches are unavailable.
template <int DIM, typename T>
int coop_sweep_blocks(psst_ctx *c) {
    if (c->coop_blocks >= 0) return c->coop_blocks;
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev), "hipDeviceGetAttribute");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_update_levels_grid<DIM, T>, kBlock, 0),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    const int want = (c->max_level + kBlock - 1) / kBlock;
    c->coop_blocks = coop ? std::max(0, std::min(want, cus * per_cu)) : 0;
    return c->coop_blocks;
}

template <int DIM, typename T>
void run_update(psst_ctx *c, int k, unsigned long long *res_row) {
    State<T> st = state<T>(c);
    const T alpha = T(c->s.rk_coeff[k]);
    const int first = (k == 0), last = (k == c->s.rk_levels - 1);
    if (c->s.time_mode == PSST_DT_LOCAL1 && c->s.dt_ordering == 0) {
        const int nl = (int)c->level_off.size() - 1;
        // One workgroup sweeping every level (a barrier per level) suits small
        // meshes; on large ones it runs on one CU: 938 ms per iteration on the
        // 4.9M-cell wedge (about 5,000 levels of up to 1,280 cells), against
        // 325 ms with one multi-workgroup launch per level.
        const bool narrow = c->max_level <= 256 || c->N <= 65536;
        if (!c->batch && c->ot.cell && (narrow || coop_sweep_blocks<DIM, T>(c) > 0)) {
            // residuals in parallel, then the dependent chain of time steps
            // and updates level by level: in one workgroup when the levels are
            // narrow (the reference's own 19,200-cell case: 1.8 ms per
            // iteration before, with the residual gather inside the chain),
            // else in one cooperative launch with a grid barrier per level
            launch(c, "ordered_residual", k_ordered_residual<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st,
                   c->ot, res_row);
            if (narrow && c->max_level <= 256)
                launch(c, "update_levels_wg", k_update_levels_wg_prefetch<DIM, T>, dim3(1), dim3(256), c->geom, st,
                       c->ph, alpha, last, c->ot, (const int *)c->level_off_dev, nl);
            else if (narrow)
                launch(c, "update_levels_wg", k_update_levels_wg<DIM, T>, dim3(1), dim3(1024), c->geom, st, c->ph,
                       alpha, last, c->ot, (const int *)c->level_off_dev, nl);
            else
                launch_coop(c, "update_levels_grid", k_update_levels_grid<DIM, T>, dim3(c->coop_blocks),
                            dim3(kBlock), c->geom, st, c->ph, alpha, last, c->ot, (const int *)c->level_off_dev, nl);
        } else if (narrow) {
            launch(c, "update_levels_block", k_update_levels_block<DIM, T>, dim3(1), dim3(1024), c->geom, st, c->ph,
                   alpha, first, last, res_row, (const int *)c->level_cells, (const int *)c->level_off_dev, nl);
        } else {
            for (int l = 0; l < nl; ++l) {
                int b = c->level_off[l], n = c->level_off[l + 1] - b;
                launch(c, "update_level", k_update_level<DIM, T>, dim3(blocks(n)), dim3(kBlock), c->geom, st, c->ph,
                       alpha, first, last, res_row, (const int *)(c->level_cells + b), n);
            }
        }
    } else if (c->s.time_mode == PSST_DT_LOCAL1) {
        launch(c, "update", k_update<DIM, T, PSST_DT_LOCAL1>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st, c->ph,
               alpha, first, last, res_row);
    } else if (c->s.time_mode == PSST_DT_LOCAL2) {
        launch(c, "update", k_update<DIM, T, PSST_DT_LOCAL2>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st, c->ph,
               alpha, first, last, res_row);
    } else {
        launch(c, "update", k_update<DIM, T, PSST_DT_UNIFORM>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st, c->ph,
               alpha, first, last, res_row);
    }
}

template <int DIM, typename T>
void run_bc(psst_ctx *c) {
    T *qn = static_cast<T *>(c->q[c->cur ^ 1]);
    launch(c, "bc", k_bc<DIM, T>, dim3(blocks(c->G)), dim3(kBlock), c->geom, qn, (const BCDev *)c->bc, c->ph);
}

template <int DIM, typename T>
void run_iteration(psst_ctx *c, unsigned long long *res_row) {
    if (c->s.order == 2) run_grad<DIM, T>(c);
    if (c->fused) {
        for (int k = 0; k < c->s.rk_levels; ++k) {
            run_stage<DIM, T>(c, k, res_row);
            c->cur ^= 1;
        }
        return;
    }
    for (int k = 0; k < c->s.rk_levels; ++k) {
        run_flux<DIM, T>(c);
        run_update<DIM, T>(c, k, res_row);
        run_bc<DIM, T>(c);
        c->cur ^= 1;
    }
}

template <int DIM, typename T>
void run_set_w(psst_ctx *c) {
    launch(c, "set_w", k_set_w<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom, (const T *)c->q[c->cur],
           (T *)c->W0, c->ph);
}

template <typename T>
void dispatch_iteration(psst_ctx *c, unsigned long long *res_row) {
    if (c->dim == 2) run_iteration<2, T>(c, res_row);
    else run_iteration<3, T>(c, res_row);
}

void ensure_res(psst_ctx *c, int n_iter) {
    if (n_iter <= c->res_cap) return;
    int cap = std::max(n_iter, 64);
    c->res = dev_alloc<unsigned long long>(c, (size_t)cap * (c->dim + 2) * kResSlots);
    c->res_cap = cap;
}

size_t state_bytes(const psst_ctx *c) { return c->s.precision == 32 ? sizeof(float) : sizeof(double); }

// canonical <-> device numbering of [rows][n] arrays (n = NT or N)
void to_device_order(const psst_ctx *c, const double *src, double *dst, int rows, int n) {
    for (int r = 0; r < rows; ++r)
        for (int d = 0; d < n; ++d) dst[(size_t)r * n + d] = src[(size_t)r * n + c->d2c[(size_t)d]];
}
void to_canonical_order(const psst_ctx *c, const double *src, double *dst, int rows, int n) {
    for (int r = 0; r < rows; ++r)
        for (int d = 0; d < n; ++d) dst[(size_t)r * n + c->d2c[(size_t)d]] = src[(size_t)r * n + d];
}

// device buffer (double or float) -> host double, device order
void download_raw(psst_ctx *c, const void *dev, double *out, size_t cnt) {
    if (c->s.precision == 64) {
        hip_check(hipMemcpyAsync(out, dev, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "D2H");
    } else {
        std::vector<float> f(cnt);
        hip_check(hipMemcpyAsync(f.data(), dev, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "D2H");
        for (size_t i = 0; i < cnt; ++i) out[i] = f[i];
    }
}

// device buffer (double or float) [rows][n] -> host double, canonical order
void download_canonical(psst_ctx *c, const void *dev, double *out, int rows, int n) {
    std::vector<double> tmp((size_t)rows * n);
    download_raw(c, dev, tmp.data(), tmp.size());
    to_canonical_order(c, tmp.data(), out, rows, n);
}

psst_status set_error(psst_ctx *c, const std::exception &e, psst_status st) {
    if (c) c->err = e.what();
    else g_create_error = e.what();
    return st;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *psst_last_error(const psst_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

// Context creation; seg (may be empty) lists the first interior cell of each
// independent instance of a batch.
static psst_status create_impl(const psst_mesh_desc *m, const psst_settings *s, const std::vector<int32_t> &seg,
                               psst_ctx **out) {
    if (!m || !s || !out) {
        g_create_error = "null argument";
        return PSST_BAD_ARG;
    }
    *out = nullptr;
    auto c = std::make_unique<psst_ctx>();
    try {
        if (m->dim != 2 && m->dim != 3) throw std::invalid_argument("dim must be 2 or 3");
        if (m->n_cells <= 0 || m->n_ghosts < 0 || m->n_faces <= 0) throw std::invalid_argument("empty mesh");
        if (s->rk_levels < 1 || s->rk_levels > 8) throw std::invalid_argument("rk_levels must be 1..8");
        if (s->flux < 1 || s->flux > 3) throw std::invalid_argument("flux must be 1 (AUSM), 2 (HLL) or 3 (HLLC)");
        if (s->order != 1 && s->order != 2) throw std::invalid_argument("order must be 1 or 2");
        if (s->order == 2 && (s->limiter < 1 || s->limiter > 4))
            throw std::invalid_argument("limiter must be 1..4 for second order");
        if (s->order == 2 && s->limiter == PSST_LIMITER_VENKATAKRISHNAN && !(s->venkat_k > 0.0))
            throw std::invalid_argument("venkat_k must be positive");
        if (s->time_mode < 0 || s->time_mode > 2) throw std::invalid_argument("time_mode must be 0, 1 or 2");
        if (s->precision != 64 && s->precision != 32) throw std::invalid_argument("precision must be 64 or 32");
        if (s->order == 2 && s->grad == PSST_GRAD_LEAST_SQUARES && (!m->lsq_r || !m->lsq_beta))
            throw std::invalid_argument("least squares needs lsq_r and lsq_beta");
        if (s->n_bc <= 0 || !s->bc) throw std::invalid_argument("boundary table missing");
        c->dim = m->dim;
        c->N = m->n_cells;
        c->G = m->n_ghosts;
        c->F = m->n_faces;
        c->NT = c->N + c->G;
        c->s = *s;
        c->bc_host.assign(s->bc, s->bc + s->n_bc);
        c->s.bc = nullptr;
        // Venkatakrishnan's eps^2 = (K h)^3, h = mean interior cell size
        // (sum of volumes in canonical order, then the dim-th root)
        double vsum = 0.0;
        for (int i = 0; i < m->n_cells; ++i) vsum += m->cell_volume[i];
        const double hbar = std::pow(vsum / m->n_cells, 1.0 / m->dim), kh = s->venkat_k * hbar;
        c->ph = Phys{s->gamma, s->rm, s->cp, s->cv, s->cfl, s->uniform_dt, kh * kh * kh};
        const int dim = c->dim, N = c->N, G = c->G, F = c->F, NT = c->NT;
        // ---- validation of the canonical topology (every index in range) ----
        const int32_t *f0 = m->face_cells, *f1 = m->face_cells + F;
        for (int f = 0; f < F; ++f) {
            if (f0[f] < 0 || f0[f] >= NT || f1[f] < 0 || f1[f] >= NT || f0[f] == f1[f])
                throw std::invalid_argument("face_cells out of range");
            if (f0[f] >= N && f1[f] >= N) throw std::invalid_argument("face between two ghosts");
        }
        if (m->visit_offsets[0] != 0 || m->list_offsets[0] != 0) throw std::invalid_argument("CSR offsets must start at 0");
        for (int i = 0; i < N; ++i) {
            if (m->visit_offsets[i + 1] < m->visit_offsets[i] || m->list_offsets[i + 1] < m->list_offsets[i])
                throw std::invalid_argument("CSR offsets not monotone");
        }
        c->nnz_visit = m->visit_offsets[N];
        c->nnz_list = m->list_offsets[N];
        std::vector<int> lnb_c((size_t)c->nnz_list);
        for (int i = 0; i < N; ++i) {
            for (int k = m->visit_offsets[i]; k < m->visit_offsets[i + 1]; ++k) {
                int f = m->visit_faces[k];
                if (f < 0 || f >= F || (f0[f] != i && f1[f] != i)) throw std::invalid_argument("visit list corrupt");
            }
            for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k) {
                int f = m->list_faces[k];
                if (f < 0 || f >= F || (f0[f] != i && f1[f] != i)) throw std::invalid_argument("face list corrupt");
                lnb_c[(size_t)k] = (f0[f] == i) ? f1[f] : f0[f];
            }
        }
        for (int i = 0; i < G; ++i) {
            int f = m->ghost_face[i];
            if (f < 0 || f >= F || (f0[f] != N + i && f1[f] != N + i)) throw std::invalid_argument("ghost_face corrupt");
            if (m->ghost_bc[i] < 0 || m->ghost_bc[i] >= s->n_bc) throw std::invalid_argument("ghost_bc out of range");
            const psst_bc &b = s->bc[m->ghost_bc[i]];
            int dk = b.dynamic_kind, tk = b.thermal_kind;
            if (dk != PSST_BC_SS_INLET && dk != PSST_BC_SS_OUTLET && dk != PSST_BC_SLIP_WALL && dk != PSST_BC_OUTLET)
                throw std::invalid_argument("unsupported dynamic boundary kind");
            if (tk != PSST_BCT_WALL_T && tk != PSST_BCT_WALL_Q && tk != PSST_BCT_INLET && tk != PSST_BCT_OUTLET)
                throw std::invalid_argument("unsupported thermal boundary kind");
        }
        // ---- device numbering: interior cells along a Morton curve, ghosts unchanged ----
        // typical cell spacing per axis: mean |dx_k| over interior faces whose
        // centre-to-centre vector is dominated by axis k
        double axis_h[3] = {0, 0, 0};
        {
            double sum[3] = {0, 0, 0};
            int64_t cnt[3] = {0, 0, 0};
            for (int f = 0; f < F; ++f) {
                if (f0[f] >= N || f1[f] >= N) continue;
                int kk = 0;
                double best = -1.0;
                for (int k = 0; k < dim; ++k) {
                    double d = m->cell_centre[(size_t)k * NT + f0[f]] - m->cell_centre[(size_t)k * NT + f1[f]];
                    if (m->face_shift) d -= m->face_shift[(size_t)k * F + f];
                    if (std::abs(d) > best) {
                        best = std::abs(d);
                        kk = k;
                    }
                }
                sum[kk] += best;
                cnt[kk]++;
            }
            for (int k = 0; k < dim; ++k) axis_h[k] = cnt[k] ? sum[k] / (double)cnt[k] : 0.0;
        }
        psst::log_phase("context: mesh validated");
        std::vector<int32_t> d2c_int = psst::morton_order(dim, N, NT, m->cell_centre, seg, axis_h);
        psst::log_phase("context: cells renumbered");
        c->d2c.resize((size_t)NT);
        std::vector<int32_t> c2d((size_t)NT);
        for (int d = 0; d < N; ++d) c->d2c[(size_t)d] = d2c_int[(size_t)d];
        for (int d = N; d < NT; ++d) c->d2c[(size_t)d] = d;
        for (int d = 0; d < NT; ++d) c2d[(size_t)c->d2c[(size_t)d]] = d;
        const std::vector<int32_t> &d2c = c->d2c;
        std::vector<int32_t> f0d((size_t)F), f1d((size_t)F);
        for (int f = 0; f < F; ++f) {
            f0d[(size_t)f] = c2d[(size_t)f0[f]];
            f1d[(size_t)f] = c2d[(size_t)f1[f]];
        }
        std::vector<int32_t> voff((size_t)N + 1), vface((size_t)c->nnz_visit), loff((size_t)N + 1),
            lface((size_t)c->nnz_list), lnb((size_t)c->nnz_list), lsrc((size_t)c->nnz_list);
        voff[0] = loff[0] = 0;
        for (int d = 0; d < N; ++d) {
            const int cc = d2c[(size_t)d];
            int o = voff[(size_t)d];
            for (int k = m->visit_offsets[cc]; k < m->visit_offsets[cc + 1]; ++k) vface[(size_t)o++] = m->visit_faces[k];
            voff[(size_t)d + 1] = o;
            o = loff[(size_t)d];
            for (int k = m->list_offsets[cc]; k < m->list_offsets[cc + 1]; ++k) {
                lface[(size_t)o] = m->list_faces[k];
                lnb[(size_t)o] = c2d[(size_t)lnb_c[(size_t)k]];
                lsrc[(size_t)o] = k;
                ++o;
            }
            loff[(size_t)d + 1] = o;
        }
        std::vector<int> gnb((size_t)std::max(G, 1));
        for (int i = 0; i < G; ++i) {
            int f = m->ghost_face[i];
            gnb[(size_t)i] = (f0d[(size_t)f] == N + i) ? f1d[(size_t)f] : f0d[(size_t)f];
        }
        // neighbour 1 of face f in neighbour 0's frame (periodic faces)
        auto sh = [&](int k, int f) { return m->face_shift ? m->face_shift[(size_t)k * F + f] : 0.0; };
        auto x1 = [&](int k, int f) {
            const double x = m->cell_centre[(size_t)k * NT + f1[f]], t = sh(k, f);
            return t != 0.0 ? x + t : x;
        };
        std::vector<double> rlr3(3 * (size_t)F, 0.0);
        for (int f = 0; f < F; ++f)
            for (int k = 0; k < dim; ++k) rlr3[(size_t)k * F + f] = m->cell_centre[(size_t)k * NT + f0[f]] - x1(k, f);
        std::vector<double> vol_d((size_t)NT), cx_d(3 * (size_t)NT, 0.0), fn3(3 * (size_t)F, 0.0);
        for (int d = 0; d < NT; ++d) {
            vol_d[(size_t)d] = m->cell_volume[d2c[(size_t)d]];
            for (int k = 0; k < dim; ++k) cx_d[(size_t)k * NT + d] = m->cell_centre[(size_t)k * NT + d2c[(size_t)d]];
        }
        for (int k = 0; k < dim; ++k) std::memcpy(&fn3[(size_t)k * F], m->face_normal + (size_t)k * F, sizeof(double) * F);
        // projected areas for time step mode 2, face-list order (psstLocalTimeCalculation.cpp:192-202)
        std::vector<double> S(3 * (size_t)N, 0.0);
        for (int d = 0; d < N; ++d) {
            double sk[3] = {0, 0, 0};
            for (int k = loff[(size_t)d]; k < loff[(size_t)d + 1]; ++k) {
                int f = lface[(size_t)k];
                for (int x = 0; x < dim; ++x) sk[x] += m->face_area[f] * std::abs(m->face_normal[(size_t)x * F + f]);
            }
            for (int x = 0; x < 3; ++x) S[(size_t)x * N + d] = sk[x];
        }
        // least-squares weights w*theta_k per list entry (psstGradCalculation.cpp:309-356;
        // 3D: textbook QR weights), computed in canonical terms
        std::vector<double> lsqc(3 * (size_t)std::max(c->nnz_list, 1), 0.0);
        if (s->grad == PSST_GRAD_LEAST_SQUARES && m->lsq_r) {
            for (int d = 0; d < N; ++d) {
                const int i = d2c[(size_t)d];
                const double ra = m->lsq_r[(size_t)0 * N + i], rb = m->lsq_r[(size_t)1 * N + i],
                             rd = m->lsq_r[(size_t)2 * N + i], re = m->lsq_r[(size_t)4 * N + i],
                             rg = m->lsq_r[(size_t)5 * N + i], beta = m->lsq_beta[i];
                for (int k = loff[(size_t)d]; k < loff[(size_t)d + 1]; ++k) {
                    const int nb = lnb_c[(size_t)lsrc[(size_t)k]];
                    const int fl = lface[(size_t)k];
                    double diff[3] = {0, 0, 0}, w = 0.0, alpha[3] = {0, 0, 0}, theta[3] = {0, 0, 0};
                    for (int x = 0; x < dim; ++x) {
                        double xn = m->cell_centre[(size_t)x * NT + nb];
                        const double t = sh(x, fl);
                        if (t != 0.0) xn = (f0[fl] == i) ? xn + t : xn - t;
                        diff[x] = xn - m->cell_centre[(size_t)x * NT + i];
                        w += diff[x] * diff[x];
                    }
                    double ws = 1.0 / w;
                    if (dim == 2) {
                        alpha[0] = diff[0] / (ra * ra);
                        alpha[1] = (diff[1] - rb / ra * diff[0]) / (rd * rd);
                        theta[0] = alpha[0] - rb / ra * alpha[1];
                        theta[1] = alpha[1];
                    } else {
                        alpha[0] = diff[0] / (ra * ra);
                        alpha[1] = (diff[1] - rb / ra * diff[0]) / (rd * rd);
                        alpha[2] = (diff[2] - re / rd * diff[1] + beta * diff[0]) / (rg * rg);
                        theta[0] = alpha[0] - rb / ra * alpha[1] + beta * alpha[2];
                        theta[1] = alpha[1] - re / rd * alpha[2];
                        theta[2] = alpha[2];
                    }
                    for (int x = 0; x < 3; ++x) lsqc[(size_t)x * c->nnz_list + k] = ws * theta[x];
                }
            }
        }
        hip_check(hipStreamCreate(&c->stream), "hipStreamCreate");
        {
            // a kernel of this library must demonstrably run: a HIP runtime
            // other than the one it was built for (e.g. one bundled with
            // another framework and loaded first under the same soname) can
            // leave launches without effect and without an error
            unsigned long long *probe = dev_alloc<unsigned long long>(c.get(), 1);
            unsigned long long zero = 0, got = 0;
            hip_check(hipMemcpy(probe, &zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy");
            hipLaunchKernelGGL(k_canary, dim3(1), dim3(64), 0, c->stream, probe);
            hip_check(hipGetLastError(), "canary launch");
            hip_check(hipStreamSynchronize(c->stream), "canary");
            hip_check(hipMemcpy(&got, probe, sizeof(got), hipMemcpyDeviceToHost), "hipMemcpy");
            if (got != 0x5053535443414e59ull) {
                int rv = 0;
                (void)hipRuntimeGetVersion(&rv);
                throw HipError("gfx950 kernels of libpsst did not run (HIP runtime version " + std::to_string(rv) +
                               " in this process): load libpsst before any library that bundles its own HIP "
                               "runtime, or use the system ROCm runtime");
            }
        }
        // reference-ordered schedule for time step mode 1: level of a cell =
        // 1 + max level of its interior face-list neighbours earlier in the
        // update loop (canonical index order); stored as device ids
        if (s->time_mode == PSST_DT_LOCAL1 && s->dt_ordering == 0) {
            std::vector<int> lev((size_t)N, 0);
            int maxl = 0;
            for (int i = 0; i < N; ++i) {
                int l = 0;
                for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k) {
                    int nb = lnb_c[(size_t)k];
                    if (nb < i) l = std::max(l, lev[(size_t)nb] + 1);
                }
                lev[(size_t)i] = l;
                maxl = std::max(maxl, l);
            }
            c->level_off.assign((size_t)maxl + 2, 0);
            for (int i = 0; i < N; ++i) c->level_off[(size_t)lev[(size_t)i] + 1]++;
            for (int l = 0; l <= maxl; ++l) c->level_off[(size_t)l + 1] += c->level_off[(size_t)l];
            std::vector<int> cells((size_t)N), fill(c->level_off.begin(), c->level_off.end() - 1);
            int widest = 0;
            for (int i = 0; i < N; ++i) cells[(size_t)fill[(size_t)lev[(size_t)i]]++] = c2d[(size_t)i];
            for (int l = 0; l <= maxl; ++l) widest = std::max(widest, c->level_off[(size_t)l + 1] - c->level_off[(size_t)l]);
            c->max_level = widest;
            c->level_cells = dev_upload(c.get(), cells.data(), cells.size());
            c->level_off_dev = dev_upload(c.get(), c->level_off.data(), c->level_off.size());
            // tables of the ordered sweep (single contexts; a batch keeps the
            // per-level launches, which it captures into graphs)
            int W = 0;
            for (int i = 0; i < N; ++i) W = std::max(W, m->list_offsets[i + 1] - m->list_offsets[i]);
            if (seg.empty() && W <= kOrderedW) {
                std::vector<int> onb((size_t)W * N, -1);
                std::vector<double> ofn((size_t)dim * W * N, 0.0), ofa((size_t)W * N, 0.0), ovol((size_t)N);
                for (int p = 0; p < N; ++p) {
                    const int i = d2c[(size_t)cells[(size_t)p]];
                    ovol[(size_t)p] = m->cell_volume[i];
                    int w = 0;
                    for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k, ++w) {
                        const int f = m->list_faces[k], nbc = lnb_c[(size_t)k];
                        onb[(size_t)w * N + p] = c2d[(size_t)nbc] | (nbc < i ? kEarlier : 0);
                        ofa[(size_t)w * N + p] = m->face_area[f];
                        for (int x = 0; x < dim; ++x)
                            ofn[((size_t)x * W + w) * N + p] = m->face_normal[(size_t)x * F + f];
                    }
                }
                c->ot.cell = c->level_cells;
                c->ot.vol = dev_upload(c.get(), ovol.data(), ovol.size());
                c->ot.nb = dev_upload(c.get(), onb.data(), onb.size());
                c->ot.fn = dev_upload(c.get(), ofn.data(), ofn.size());
                c->ot.fa = dev_upload(c.get(), ofa.data(), ofa.size());
                c->ot.R = dev_alloc<double>(c.get(), (size_t)(dim + 2) * N);
                c->ot.W = W;
            }
        }
        // ---- device geometry ----
        // Device face order: faces sorted by their lower device cell id, so
        // that the face-indexed kernels (unfused fluxes, Green-Gauss and
        // least-squares sweeps, the reference-ordered time step) read
        // neighbouring cells in neighbouring lanes.  Every order-dependent
        // sum walks a per-cell face list, which keeps the reference's order
        // whatever the face numbering.  The tile plan works on the canonical
        // host arrays and is not affected.
        std::vector<int32_t> forder((size_t)F), fpos((size_t)F);
        for (int f = 0; f < F; ++f) forder[(size_t)f] = f;
        std::stable_sort(forder.begin(), forder.end(), [&](int32_t a, int32_t b) {
            return std::min(f0d[(size_t)a], f1d[(size_t)a]) < std::min(f0d[(size_t)b], f1d[(size_t)b]);
        });
        for (int p = 0; p < F; ++p) fpos[(size_t)forder[(size_t)p]] = p;
        // [rows][F] arrays in device face order, and face ids as device positions
        auto upload_by_face = [&](const auto *src, size_t rows) {
            using V = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
            std::vector<V> out(rows * (size_t)F);
            for (size_t r = 0; r < rows; ++r)
                for (int p = 0; p < F; ++p) out[r * F + p] = src[r * F + (size_t)forder[(size_t)p]];
            return dev_upload(c.get(), out.data(), out.size());
        };
        auto upload_positions = [&](const int32_t *ids, size_t n) {
            std::vector<int32_t> out(n);
            for (size_t i = 0; i < n; ++i) out[i] = fpos[(size_t)ids[i]];
            return dev_upload(c.get(), out.data(), out.size());
        };
        Geom &g = c->geom;
        g.N = N;
        g.G = G;
        g.NT = NT;
        g.F = F;
        g.vol = dev_upload(c.get(), vol_d.data(), vol_d.size());
        g.cx = dev_upload(c.get(), cx_d.data(), cx_d.size());
        g.fn = upload_by_face(fn3.data(), 3);
        g.fa = upload_by_face(m->face_area, 1);
        g.beta = upload_by_face(m->face_beta, 1);
        g.f0 = upload_by_face(f0d.data(), 1);
        g.f1 = upload_by_face(f1d.data(), 1);
        g.rlr = upload_by_face(rlr3.data(), 3);
        g.vface = upload_positions(vface.data(), vface.size());
        g.lface = upload_positions(lface.data(), lface.size());
        g.gface = upload_positions(m->ghost_face, (size_t)G);
        g.voff = dev_upload(c.get(), voff.data(), voff.size());
        g.loff = dev_upload(c.get(), loff.data(), loff.size());
        g.lnb = dev_upload(c.get(), lnb.data(), lnb.size());
        g.lsqc = dev_upload(c.get(), lsqc.data(), lsqc.size());
        g.S = dev_upload(c.get(), S.data(), S.size());
        g.gnb = dev_upload(c.get(), gnb.data(), (size_t)G);
        g.gbc = dev_upload(c.get(), m->ghost_bc, (size_t)G);
        {
            std::vector<int32_t> rank(d2c.begin(), d2c.end());
            g.rank = dev_upload(c.get(), rank.data(), rank.size());
        }
        std::vector<BCDev> bcd((size_t)s->n_bc);
        for (int i = 0; i < s->n_bc; ++i) {
            bcd[(size_t)i].dyn = s->bc[i].dynamic_kind;
            bcd[(size_t)i].th = s->bc[i].thermal_kind;
            for (int d = 0; d < 3; ++d) bcd[(size_t)i].u[d] = s->bc[i].velocity[d];
            bcd[(size_t)i].p = s->bc[i].pressure;
            bcd[(size_t)i].T = s->bc[i].temperature;
        }
        c->bc = dev_upload(c.get(), bcd.data(), bcd.size());
        const size_t sb = state_bytes(c.get());
        // ---- tile plan of the fused kernels ----
        const bool ordered_dt1 = (s->time_mode == PSST_DT_LOCAL1 && s->dt_ordering == 0);
        if (s->fused && !ordered_dt1) {
            // Tile size: the caller's, or the first of a list whose stage
            // kernel LDS lets a CU (160 KB) hold 16 or more waves.  The lists
            // start at what was fastest on the 3840x1280 wedge (FP64 224 cells
            // / 256 threads, FP32 768 / 512); a 3D hexahedral tile carries
            // about as many halo cells as own cells, and there 128-cell FP64
            // tiles (four workgroups per CU) ran 31% faster than 224-cell ones
            // (two).
            if (s->tile_threads != 0 && s->tile_threads != 256 && s->tile_threads != 512)
                throw std::invalid_argument("tile_threads must be 256 or 512");
            std::vector<int> cands;
            if (s->tile_cells > 0) {
                if (s->tile_cells < 32 || s->tile_cells > 4096) throw std::invalid_argument("tile_cells must be 32..4096");
                cands = {s->tile_cells};
            } else if (s->precision == 64) {
                cands = {224, 192, 160, 128, 96};
            } else {
                cands = {768, 512, 384, 256, 192, 128};
            }
            const double cost_slot = (double)(dim + 3) * sb + sizeof(int), cost_face = (double)(dim + 2) * sb;
            psst::TilePlan tp;
            size_t shm_stage = 0, shm_grad = 0;
            for (size_t ci = 0; ci < cands.size(); ++ci) {
                tp = psst::plan_tiles(N, G, F, cands[ci], f0d.data(), f1d.data(), voff.data(), vface.data(),
                                      loff.data(), lface.data(), seg, cost_slot, cost_face);
                c->max_slots = tp.max_slots;
                c->max_faces = tp.max_faces;
                shm_stage = stage_shm(c.get(), sb);
                shm_grad = grad_shm(c.get(), sb);
                psst::log_phase(("context: tile plan of " + std::to_string(cands[ci]) + " cells, stage LDS " +
                                 std::to_string(shm_stage) + " B").c_str());
                const bool last = ci + 1 == cands.size();
                const int thr = s->tile_threads ? s->tile_threads : (cands[ci] >= 512 ? 512 : 256);
                const size_t waves = (160 * 1024 / std::max<size_t>(shm_stage, 1)) * (size_t)(thr / 64);
                if (waves >= 16 || (last && shm_stage <= 160 * 1024)) break;
            }
            const int tile = tp.tile;
            c->tile_threads = s->tile_threads ? s->tile_threads : (tile >= 512 ? 512 : 256);
            c->tile_cells = tile;
            c->tile_cell_begin = tp.cell_begin;
            if (shm_stage <= 160 * 1024) {
                const int TF = (int)tp.face_id.size();
                std::vector<double> tfn((size_t)dim * TF), tfa((size_t)TF), trlr((size_t)dim * TF), tbeta((size_t)TF);
                for (int j = 0; j < TF; ++j) {
                    const int f = tp.face_id[(size_t)j];
                    for (int k = 0; k < dim; ++k) {
                        tfn[(size_t)k * TF + j] = m->face_normal[(size_t)k * F + f];
                        trlr[(size_t)k * TF + j] = rlr3[(size_t)k * F + f];
                    }
                    tfa[(size_t)j] = m->face_area[f];
                    tbeta[(size_t)j] = m->face_beta[f];
                }
                std::vector<double> tlsq((size_t)dim * std::max(tp.W_list, 1) * N, 0.0);
                if (s->grad == PSST_GRAD_LEAST_SQUARES)
                    for (int d = 0; d < N; ++d)
                        for (int k = loff[(size_t)d], w = 0; k < loff[(size_t)d + 1]; ++k, ++w)
                            for (int x = 0; x < dim; ++x)
                                tlsq[((size_t)x * tp.W_list + w) * N + d] = lsqc[(size_t)x * c->nnz_list + k];
                TileDev &td = c->td;
                td.cell_begin = dev_upload(c.get(), tp.cell_begin.data(), tp.cell_begin.size());
                td.halo_off = dev_upload(c.get(), tp.halo_off.data(), tp.halo_off.size());
                td.halo = dev_upload(c.get(), tp.halo.data(), tp.halo.size());
                td.face_off = dev_upload(c.get(), tp.face_off.data(), tp.face_off.size());
                td.ends = dev_upload(c.get(), tp.face_ends.data(), tp.face_ends.size());
                td.fn = dev_upload(c.get(), tfn.data(), tfn.size());
                td.fa = dev_upload(c.get(), tfa.data(), tfa.size());
                td.rlr = dev_upload(c.get(), trlr.data(), trlr.size());
                td.beta = dev_upload(c.get(), tbeta.data(), tbeta.size());
                td.ghost_off = dev_upload(c.get(), tp.ghost_off.data(), tp.ghost_off.size());
                td.ghost = dev_upload(c.get(), tp.ghost.data(), tp.ghost.size());
                td.visit = reinterpret_cast<const short *>(dev_upload(c.get(), tp.visit_ell.data(), tp.visit_ell.size()));
                td.list = reinterpret_cast<const short *>(dev_upload(c.get(), tp.list_ell.data(), tp.list_ell.size()));
                td.lsqc = dev_upload(c.get(), tlsq.data(), tlsq.size());
                td.TF = TF;
                td.Wv = tp.W_visit;
                td.Wl = tp.W_list;
                td.MS = tp.max_slots;
                td.MF = tp.max_faces;
                td.n_tiles = tp.n_tiles;
                c->n_tiles = tp.n_tiles;
                c->shm_stage = shm_stage;
                c->shm_grad = shm_grad;
                c->fused = true;
            }
        }
        psst::log_phase("context: geometry and tiles uploaded");
        // ---- state: sized for FP64 whatever the precision, so that
        // psst_set_precision can switch in place ----
        const size_t sb8 = sizeof(double);
        for (int i = 0; i < 2; ++i) {
            hip_check(hipMalloc(&c->q[i], sb8 * (size_t)(dim + 3) * NT), "hipMalloc");
            c->allocs.push_back(c->q[i]);
        }
        hip_check(hipMalloc(&c->W0, sb8 * (size_t)(dim + 2) * N), "hipMalloc");
        c->allocs.push_back(c->W0);
        hip_check(hipMalloc(&c->grad, sb8 * (size_t)(dim == 2 ? GradLayout<2>::STRIDE : GradLayout<3>::STRIDE) * N),
                  "hipMalloc");
        c->allocs.push_back(c->grad);
        hip_check(hipMalloc(&c->FA, sb8 * (size_t)(dim + 2) * F), "hipMalloc");
        c->allocs.push_back(c->FA);
        ensure_res(c.get(), 64);
        c->kernel_names = {"grad_gg",      "grad_lsq",
                           "flux",         "update",
                           "update_level", "update_levels_block",
                           "update_levels_grid", "bc",
                           "stage",        "grad_gg_tile",
                           "grad_lsq_tile", "batch_check",
                           "ordered_residual", "update_levels_wg"};
    } catch (const std::exception &e) {
        g_create_error = e.what();
        for (void *p : c->allocs) (void)hipFree(p);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        return dynamic_cast<const HipError *>(&e) ? PSST_HIP_ERROR : PSST_BAD_ARG;
    }
    *out = c.release();
    return PSST_OK;
}

extern "C" {

psst_status psst_ctx_create(const psst_mesh_desc *m, const psst_settings *s, psst_ctx **out) {
    return create_impl(m, s, {}, out);
}

psst_status psst_set_state(psst_ctx *c, const double *prim) {
    if (!c || !prim) return PSST_BAD_ARG;
    try {
        const size_t n = (size_t)(c->dim + 3) * c->NT;
        std::vector<double> dev_prim(n);
        to_device_order(c, prim, dev_prim.data(), c->dim + 3, c->NT);
        if (c->s.precision == 64) {
            hip_check(hipMemcpy(c->q[0], dev_prim.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(c->q[1], dev_prim.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        } else {
            double *tmp = nullptr;
            hip_check(hipMalloc(&tmp, n * sizeof(double)), "hipMalloc");
            hip_check(hipMemcpy(tmp, dev_prim.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)n)), dim3(kBlock), 0, c->stream, tmp,
                               (float *)c->q[0], n);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)n)), dim3(kBlock), 0, c->stream, tmp,
                               (float *)c->q[1], n);
            hip_check(hipStreamSynchronize(c->stream), "convert");
            hip_check(hipFree(tmp), "hipFree");
        }
        c->cur = 0;
        if (c->dim == 2) {
            c->s.precision == 64 ? run_set_w<2, double>(c) : run_set_w<2, float>(c);
        } else {
            c->s.precision == 64 ? run_set_w<3, double>(c) : run_set_w<3, float>(c);
        }
        hip_check(hipStreamSynchronize(c->stream), "set_w");
        c->have_state = true;
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_set_residual_reference(psst_ctx *c, const double *res0) {
    if (!c || !res0) return PSST_BAD_ARG;
    c->res0.assign(res0, res0 + c->dim + 2);
    return PSST_OK;
}

psst_status psst_run(psst_ctx *c, int32_t n_iter, double tol, double *res_hist, int32_t *iters_done) {
    if (!c) return PSST_BAD_ARG;
    if (!c->have_state) {
        c->err = "psst_run: no state (call psst_set_state first)";
        return PSST_BAD_ARG;
    }
    if (iters_done) *iters_done = 0;
    if (n_iter <= 0) return PSST_OK;
    const int ws = c->dim + 2;
    try {
        // Residual rows come back once per chunk.  With a tolerance, the state
        // at the start of each chunk is kept; if the tolerance is met inside a
        // chunk, the state is restored and exactly the iterations up to that
        // one are replayed (the solver is deterministic, so the replay is
        // bitwise the same), and the run ends on that iteration's state.
        const int chunk = tol > 0.0 ? std::min<int32_t>(n_iter, 64) : n_iter;
        ensure_res(c, chunk);
        std::vector<unsigned long long> slots((size_t)chunk * ws * kResSlots);
        const size_t sb = state_bytes(c);
        const size_t nq = (size_t)(c->dim + 3) * c->NT * sb, nw = (size_t)(c->dim + 2) * c->N * sb;
        if (tol > 0.0 && !c->snap)  // sized for FP64, whatever the current precision
            c->snap = dev_alloc<char>(c, ((size_t)(c->dim + 3) * c->NT + (size_t)(c->dim + 2) * c->N) * sizeof(double));
        auto run_iters = [&](int n) {
            const size_t nres = (size_t)n * ws * kResSlots;
            launch(c, "fill", k_fill_u64, dim3(blocks((long)nres)), dim3(kBlock), c->res, nres, 0ull);
            for (int it = 0; it < n; ++it) {
                unsigned long long *row = c->res + (size_t)it * ws * kResSlots;
                c->s.precision == 64 ? dispatch_iteration<double>(c, row) : dispatch_iteration<float>(c, row);
            }
        };
        int done = 0;
        bool stop = false;
        bool bad = false;
        while (done < n_iter && !stop) {
            const int n = std::min(chunk, n_iter - done);
            const size_t nres = (size_t)n * ws * kResSlots;
            const int cur0 = c->cur;
            if (tol > 0.0) {
                hip_check(hipMemcpyAsync(c->snap, c->q[cur0], nq, hipMemcpyDeviceToDevice, c->stream), "snapshot");
                hip_check(hipMemcpyAsync(c->snap + nq, c->W0, nw, hipMemcpyDeviceToDevice, c->stream), "snapshot");
            }
            run_iters(n);
            hip_check(hipMemcpyAsync(slots.data(), c->res, nres * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                     c->stream),
                      "hipMemcpy D2H");
            hip_check(hipStreamSynchronize(c->stream), "psst_run");
            if (c->timing) flush_timing(c);
            int stop_at = -1;
            for (int it = 0; it < n; ++it) {
                double row[5];
                for (int j = 0; j < ws; ++j) {
                    unsigned long long mx = 0;
                    for (int sl = 0; sl < kResSlots; ++sl)
                        mx = std::max(mx, slots[((size_t)it * ws + j) * kResSlots + sl]);
                    double v;
                    std::memcpy(&v, &mx, sizeof(v));
                    row[j] = v;
                    if (!(v <= 1.0e300)) bad = true;  // NaN or inf
                }
                if (res_hist)
                    for (int j = 0; j < ws; ++j) res_hist[(size_t)(done + it) * ws + j] = row[j];
                if (c->res0.empty()) c->res0.assign(row, row + ws);
                if (tol > 0.0) {
                    double drop = 0.0;
                    for (int j = 0; j < ws; ++j) drop = std::max(drop, c->res0[(size_t)j] > 0 ? row[j] / c->res0[(size_t)j] : 0.0);
                    if (drop <= tol) {
                        stop_at = it;
                        break;
                    }
                }
            }
            if (stop_at >= 0) {
                stop = true;
                if (stop_at < n - 1) {
                    // restore the chunk's start and replay up to the stopping iteration
                    c->cur = cur0;
                    hip_check(hipMemcpyAsync(c->q[cur0], c->snap, nq, hipMemcpyDeviceToDevice, c->stream), "restore");
                    hip_check(hipMemcpyAsync(c->W0, c->snap + nq, nw, hipMemcpyDeviceToDevice, c->stream), "restore");
                    run_iters(stop_at + 1);
                    hip_check(hipStreamSynchronize(c->stream), "psst_run replay");
                    if (c->timing) flush_timing(c);
                }
                done += stop_at + 1;
                if (iters_done) *iters_done = done;
                break;
            }
            done += n;
            if (iters_done) *iters_done = done;
        }
        if (bad) {
            c->err = "non-finite residual: the state contains NaN/inf (a negative temperature or density upstream)";
            return PSST_NAN_DETECTED;
        }
        if (tol > 0.0 && !stop) return PSST_NO_CONVERGENCE;
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_get_state(psst_ctx *c, double *prim) {
    if (!c || !prim) return PSST_BAD_ARG;
    try {
        download_canonical(c, c->q[c->cur], prim, c->dim + 3, c->NT);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_get_conservative(psst_ctx *c, double *w) {
    if (!c || !w) return PSST_BAD_ARG;
    try {
        download_canonical(c, c->W0, w, c->dim + 2, c->N);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_debug_gradients(psst_ctx *c, double *grads) {
    if (!c || !grads) return PSST_BAD_ARG;
    try {
        if (c->dim == 2) {
            c->s.precision == 64 ? run_grad<2, double>(c) : run_grad<2, float>(c);
        } else {
            c->s.precision == 64 ? run_grad<3, double>(c) : run_grad<3, float>(c);
        }
        {
            const int nc = (c->dim + 2) * c->dim;
            const int stride = c->dim == 2 ? GradLayout<2>::STRIDE : GradLayout<3>::STRIDE;
            std::vector<double> aos((size_t)stride * c->N), soa((size_t)nc * c->N);
            download_raw(c, c->grad, aos.data(), aos.size());
            for (int d = 0; d < c->N; ++d)
                for (int v = 0; v < nc; ++v) soa[(size_t)v * c->N + d] = aos[(size_t)d * stride + v];
            to_canonical_order(c, soa.data(), grads, nc, c->N);
        }
        if (c->timing) flush_timing(c);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_debug_residual(psst_ctx *c, double *residual) {
    if (!c || !residual) return PSST_BAD_ARG;
    try {
        double *out = nullptr;
        const size_t n = (size_t)(c->dim + 2) * c->N;
        hip_check(hipMalloc(&out, n * sizeof(double)), "hipMalloc");
        if (c->dim == 2) {
            if (c->s.precision == 64) {
                run_flux<2, double>(c);
                hipLaunchKernelGGL((k_gather_residual<2, double>), dim3(blocks(c->N)), dim3(kBlock), 0, c->stream,
                                   c->geom, (const double *)c->FA, out);
            } else {
                run_flux<2, float>(c);
                hipLaunchKernelGGL((k_gather_residual<2, float>), dim3(blocks(c->N)), dim3(kBlock), 0, c->stream,
                                   c->geom, (const float *)c->FA, out);
            }
        } else {
            if (c->s.precision == 64) {
                run_flux<3, double>(c);
                hipLaunchKernelGGL((k_gather_residual<3, double>), dim3(blocks(c->N)), dim3(kBlock), 0, c->stream,
                                   c->geom, (const double *)c->FA, out);
            } else {
                run_flux<3, float>(c);
                hipLaunchKernelGGL((k_gather_residual<3, float>), dim3(blocks(c->N)), dim3(kBlock), 0, c->stream,
                                   c->geom, (const float *)c->FA, out);
            }
        }
        hip_check(hipGetLastError(), "gather");
        std::vector<double> tmp(n);
        hip_check(hipMemcpyAsync(tmp.data(), out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "D2H");
        hip_check(hipFree(out), "hipFree");
        to_canonical_order(c, tmp.data(), residual, c->dim + 2, c->N);
        if (c->timing) flush_timing(c);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_kernel_timing(psst_ctx *c, int32_t enable) {
    if (!c) return PSST_BAD_ARG;
    try {
        if (!enable) flush_timing(c);
        c->timing = enable != 0;
        if (enable) c->stats.clear();
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_kernel_stats(psst_ctx *c, const char *name, int64_t *count, double *mean_ms) {
    if (!c || !name) return PSST_BAD_ARG;
    try {
        flush_timing(c);
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
    auto it = c->stats.find(name);
    if (count) *count = it == c->stats.end() ? 0 : it->second.count;
    if (mean_ms) *mean_ms = (it == c->stats.end() || it->second.count == 0) ? 0.0 : it->second.total_ms / it->second.count;
    return PSST_OK;
}

psst_status psst_kernel_names(psst_ctx *c, char *buf, int32_t len) {
    if (!c || !buf || len <= 0) return PSST_BAD_ARG;
    std::string s;
    for (auto &n : c->kernel_names) {
        if (!s.empty()) s += ';';
        s += n;
    }
    if ((int32_t)s.size() + 1 > len) return PSST_BAD_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return PSST_OK;
}

psst_status psst_copy_calibration(int64_t bytes, int32_t reps, double *mean_ms) {
    if (bytes <= 0 || reps <= 0) return PSST_BAD_ARG;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    psst_status rc = PSST_OK;
    try {
        const size_t n = (size_t)bytes / sizeof(double4);
        hip_check(hipMalloc(&a, n * sizeof(double4)), "hipMalloc");
        hip_check(hipMalloc(&b, n * sizeof(double4)), "hipMalloc");
        hip_check(hipMemset(a, 0, n * sizeof(double4)), "hipMemset");
        hip_check(hipStreamCreate(&st), "hipStreamCreate");
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        const unsigned grid = 256 * 8 * 4;
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, (const double4 *)a, (double4 *)b, n);
        hip_check(hipEventRecord(e0, st), "hipEventRecord");
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, (const double4 *)a, (double4 *)b, n);
        hip_check(hipGetLastError(), "k_copy");
        hip_check(hipEventRecord(e1, st), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (mean_ms) *mean_ms = ms / reps;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        rc = PSST_HIP_ERROR;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

void psst_ctx_destroy(psst_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->batch) {
        for (auto &kv : c->batch->exec)
            for (hipGraphExec_t &e : kv.second)
                if (e) (void)hipGraphExecDestroy(e);
        if (c->batch->h_flags) (void)hipHostFree(c->batch->h_flags);
        if (c->batch->h_ring) (void)hipHostFree(c->batch->h_ring);
    }
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

// ===========================================================================
// Precision switch, batches of instances, autotuner
// ===========================================================================
namespace {

void destroy_graphs(psst_ctx *c) {
    if (!c->batch) return;
    for (auto &kv : c->batch->exec)
        for (hipGraphExec_t &e : kv.second)
            if (e) (void)hipGraphExecDestroy(e);
    c->batch->exec.clear();
}

BatchCheck batch_check_args(psst_ctx *c) {
    BatchState &b = *c->batch;
    BatchCheck k;
    k.n = b.n;
    k.H = b.H;
    k.res = b.res;
    k.res0 = b.res0;
    k.hist = b.hist;
    k.drop = b.drop;
    k.active = b.active;
    k.conv = b.conv;
    k.frozen_buf = b.frozen_buf;
    k.iter = b.iter;
    k.tile_begin = b.tile_begin_dev;
    k.list = b.list;
    k.count = b.count;
    return k;
}

void launch_check(psst_ctx *c, double tol, int cur_after, int rebuild_only) {
    const BatchCheck k = batch_check_args(c);
    if (c->dim == 2)
        launch(c, "batch_check", k_batch_check<4>, dim3(1), dim3(kCheckThreads), k, tol, cur_after, rebuild_only);
    else
        launch(c, "batch_check", k_batch_check<5>, dim3(1), dim3(kCheckThreads), k, tol, cur_after, rebuild_only);
}

// One batch iteration: gradients, the RK stages over the active tiles, and
// the per-instance bookkeeping.
void batch_iteration(psst_ctx *c, double tol) {
    c->s.precision == 64 ? dispatch_iteration<double>(c, nullptr) : dispatch_iteration<float>(c, nullptr);
    launch_check(c, tol, c->cur, 0);
}

// The same iteration as a captured graph (one per parity of the state
// buffers at its start: with an odd stage count the buffers alternate).
void batch_graph_iteration(psst_ctx *c, double tol) {
    BatchState &b = *c->batch;
    if (b.graph_precision != c->s.precision || b.graph_tol != tol) {
        destroy_graphs(c);
        b.graph_precision = c->s.precision;
        b.graph_tol = tol;
    }
    const int parity = c->cur;
    hipGraphExec_t &ex = b.exec[b.grid_tiles][parity];
    if (!ex) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
            batch_iteration(c, tol);
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            c->cur = parity;
            throw;
        }
        hip_check(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) ex = nullptr;
        hip_check(e, "hipGraphInstantiate");
        c->cur = parity;
    }
    hip_check(hipGraphLaunch(ex, c->stream), "hipGraphLaunch");
    c->cur = parity ^ (c->s.rk_levels & 1);
}

// Brings every instance's latest state into the current state buffer.
void batch_gather_states(psst_ctx *c) {
    BatchState &b = *c->batch;
    const size_t sb = state_bytes(c);
    for (int i = 0; i < b.n; ++i) {
        const int from = b.buf_of[(size_t)i];
        if (from == c->cur) continue;
        const char *src = static_cast<const char *>(c->q[from]);
        char *dst = static_cast<char *>(c->q[c->cur]);
        for (int v = 0; v < c->dim + 3; ++v) {
            const size_t ci = (size_t)v * c->NT + b.cell_begin[(size_t)i];
            const size_t nc = (size_t)(b.cell_begin[(size_t)i + 1] - b.cell_begin[(size_t)i]);
            hip_check(hipMemcpyAsync(dst + ci * sb, src + ci * sb, nc * sb, hipMemcpyDeviceToDevice, c->stream), "D2D");
            const size_t gi = (size_t)v * c->NT + b.N_tot + b.ghost_begin[(size_t)i];
            const size_t ng = (size_t)(b.ghost_begin[(size_t)i + 1] - b.ghost_begin[(size_t)i]);
            if (ng)
                hip_check(hipMemcpyAsync(dst + gi * sb, src + gi * sb, ng * sb, hipMemcpyDeviceToDevice, c->stream),
                          "D2D");
        }
        b.buf_of[(size_t)i] = c->cur;
    }
}

void upload_values(psst_ctx *c, void *dev_base, size_t offset, const std::vector<double> &vals) {
    if (vals.empty()) return;
    if (c->s.precision == 64) {
        hip_check(hipMemcpy(static_cast<double *>(dev_base) + offset, vals.data(), vals.size() * sizeof(double),
                            hipMemcpyHostToDevice),
                  "hipMemcpy");
    } else {
        std::vector<float> f(vals.begin(), vals.end());
        hip_check(hipMemcpy(static_cast<float *>(dev_base) + offset, f.data(), f.size() * sizeof(float),
                            hipMemcpyHostToDevice),
                  "hipMemcpy");
    }
}

std::vector<double> download_values(psst_ctx *c, const void *dev_base, size_t offset, size_t n) {
    std::vector<double> out(n);
    if (!n) return out;
    if (c->s.precision == 64) {
        hip_check(hipMemcpy(out.data(), static_cast<const double *>(dev_base) + offset, n * sizeof(double),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy");
    } else {
        std::vector<float> f(n);
        hip_check(hipMemcpy(f.data(), static_cast<const float *>(dev_base) + offset, n * sizeof(float),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy");
        for (size_t k = 0; k < n; ++k) out[k] = f[k];
    }
    return out;
}

bool same_scheme(const psst_settings &a, const psst_settings &b) {
    bool eq = a.flux == b.flux && a.order == b.order && a.grad == b.grad && a.limiter == b.limiter &&
              a.time_mode == b.time_mode && a.rk_levels == b.rk_levels && a.cfl == b.cfl && a.gamma == b.gamma &&
              a.rm == b.rm && a.cp == b.cp && a.cv == b.cv && a.precision == b.precision &&
              a.dt_ordering == b.dt_ordering && a.venkat_k == b.venkat_k && a.tile_cells == b.tile_cells &&
              a.tile_threads == b.tile_threads;
    for (int k = 0; k < a.rk_levels && k < 8; ++k) eq = eq && a.rk_coeff[k] == b.rk_coeff[k];
    if (a.time_mode == PSST_DT_UNIFORM) eq = eq && a.uniform_dt == b.uniform_dt;
    return eq;
}

}  // namespace

struct psst_batch {
    psst_ctx *ctx = nullptr;
    std::string err;
    ~psst_batch() {
        if (ctx) psst_ctx_destroy(ctx);
    }
};

extern "C" {

psst_status psst_set_precision(psst_ctx *c, int32_t precision) {
    if (!c) return PSST_BAD_ARG;
    if (precision != 32 && precision != 64) {
        c->err = "precision must be 32 or 64";
        return PSST_BAD_ARG;
    }
    if (precision == c->s.precision) return PSST_OK;
    try {
        const size_t nsb = precision == 64 ? sizeof(double) : sizeof(float);
        if (c->fused && stage_shm(c, nsb) > 160 * 1024) {
            c->err = "the tile plan does not fit in LDS at this precision";
            return PSST_UNSUPPORTED;
        }
        const size_t nq = (size_t)(c->dim + 3) * c->NT, nw = (size_t)(c->dim + 2) * c->N;
        void *tmp = nullptr;
        hip_check(hipMalloc(&tmp, nq * sizeof(double)), "hipMalloc");
        try {
            void *bufs[3] = {c->q[0], c->q[1], c->W0};
            const size_t cnt[3] = {nq, nq, nw};
            for (int k = 0; k < 3; ++k) {
                if (precision == 64) {
                    hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks((long)cnt[k])), dim3(kBlock), 0,
                                       c->stream, (const float *)bufs[k], (double *)tmp, cnt[k]);
                    hip_check(hipGetLastError(), "convert");
                    hip_check(hipMemcpyAsync(bufs[k], tmp, cnt[k] * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                              "D2D");
                } else {
                    hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)cnt[k])), dim3(kBlock), 0,
                                       c->stream, (const double *)bufs[k], (float *)tmp, cnt[k]);
                    hip_check(hipGetLastError(), "convert");
                    hip_check(hipMemcpyAsync(bufs[k], tmp, cnt[k] * sizeof(float), hipMemcpyDeviceToDevice, c->stream),
                              "D2D");
                }
            }
            hip_check(hipStreamSynchronize(c->stream), "convert");
        } catch (...) {
            (void)hipFree(tmp);
            throw;
        }
        hip_check(hipFree(tmp), "hipFree");
        c->s.precision = precision;
        c->shm_stage = c->fused ? stage_shm(c, nsb) : 0;
        c->shm_grad = c->fused ? grad_shm(c, nsb) : 0;
        destroy_graphs(c);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

const char *psst_batch_last_error(const psst_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

psst_status psst_batch_create(int32_t n, const psst_mesh_desc *meshes, const psst_settings *settings,
                              psst_batch **out) {
    if (n <= 0 || !meshes || !settings || !out) {
        g_create_error = "psst_batch_create: null argument or empty batch";
        return PSST_BAD_ARG;
    }
    *out = nullptr;
    try {
        const int dim = meshes[0].dim;
        const psst_settings &s0 = settings[0];
        if (s0.time_mode == PSST_DT_LOCAL1 && s0.dt_ordering == 0)
            throw std::invalid_argument("batches need an order-free time step (time_mode 0 or 2, or dt_ordering 1)");
        int64_t N = 0, G = 0, F = 0, nv = 0, nl = 0, nbc = 0;
        bool lsq = s0.order == 2 && s0.grad == PSST_GRAD_LEAST_SQUARES;
        bool any_shift = false;
        for (int i = 0; i < n; ++i) any_shift = any_shift || meshes[i].face_shift;
        for (int i = 0; i < n; ++i) {
            const psst_mesh_desc &m = meshes[i];
            if (m.dim != dim) throw std::invalid_argument("instances of a batch must share the dimension");
            if (!same_scheme(settings[i], s0)) throw std::invalid_argument("instances of a batch must share the scheme");
            if (m.n_cells <= 0 || m.n_faces <= 0 || m.n_ghosts < 0) throw std::invalid_argument("empty instance mesh");
            if (settings[i].n_bc <= 0 || !settings[i].bc) throw std::invalid_argument("boundary table missing");
            N += m.n_cells;
            G += m.n_ghosts;
            F += m.n_faces;
            nv += m.visit_offsets[m.n_cells];
            nl += m.list_offsets[m.n_cells];
            nbc += settings[i].n_bc;
        }
        const int64_t NT = N + G;
        if (NT >= INT32_MAX / 2 || F >= INT32_MAX / 2 || nv >= INT32_MAX / 2 || nl >= INT32_MAX / 2)
            throw std::invalid_argument("batch too large for 32-bit indices");
        std::vector<double> vol((size_t)NT), cen((size_t)dim * NT), nrm((size_t)dim * F), area((size_t)F),
            beta((size_t)F), lsq_r(lsq ? 6 * (size_t)N : 0), lsq_b(lsq ? (size_t)N : 0),
            fshift(any_shift ? 3 * (size_t)F : 0, 0.0);
        std::vector<int32_t> fc(2 * (size_t)F), voff((size_t)N + 1), vf((size_t)nv), loff((size_t)N + 1), lf((size_t)nl),
            gface((size_t)G), gbc((size_t)G), seg((size_t)n);
        std::vector<psst_bc> bcs;
        bcs.reserve((size_t)nbc);
        auto bs = std::make_unique<BatchState>();
        bs->n = n;
        bs->ws = dim + 2;
        bs->N_tot = (int32_t)N;
        bs->cell_begin.assign(1, 0);
        bs->ghost_begin.assign(1, 0);
        int64_t cb = 0, gb = 0, fb = 0, vb = 0, lb = 0, bb = 0;
        voff[0] = loff[0] = 0;
        for (int i = 0; i < n; ++i) {
            const psst_mesh_desc &m = meshes[i];
            const int64_t Ni = m.n_cells, Gi = m.n_ghosts, Fi = m.n_faces, NTi = Ni + Gi;
            auto uc = [&](int32_t x) -> int32_t {  // instance cell -> union canonical cell
                return (int32_t)(x < Ni ? cb + x : N + gb + (x - Ni));
            };
            seg[(size_t)i] = (int32_t)cb;
            for (int64_t x = 0; x < NTi; ++x) {
                const int32_t u = uc((int32_t)x);
                vol[(size_t)u] = m.cell_volume[x];
                for (int k = 0; k < dim; ++k) cen[(size_t)k * NT + u] = m.cell_centre[(size_t)k * NTi + x];
            }
            for (int64_t f = 0; f < Fi; ++f) {
                for (int k = 0; k < dim; ++k) nrm[(size_t)k * F + fb + f] = m.face_normal[(size_t)k * Fi + f];
                if (m.face_shift)
                    for (int k = 0; k < 3; ++k) fshift[(size_t)k * F + fb + f] = m.face_shift[(size_t)k * Fi + f];
                area[(size_t)(fb + f)] = m.face_area[f];
                beta[(size_t)(fb + f)] = m.face_beta[f];
                const int32_t a = m.face_cells[f], b = m.face_cells[Fi + f];
                if (a < 0 || a >= NTi || b < 0 || b >= NTi) throw std::invalid_argument("face_cells out of range");
                fc[(size_t)(fb + f)] = uc(a);
                fc[(size_t)(F + fb + f)] = uc(b);
            }
            for (int64_t x = 0; x < Ni; ++x) {
                for (int32_t k = m.visit_offsets[x]; k < m.visit_offsets[x + 1]; ++k)
                    vf[(size_t)(vb + k)] = (int32_t)(fb + m.visit_faces[k]);
                for (int32_t k = m.list_offsets[x]; k < m.list_offsets[x + 1]; ++k)
                    lf[(size_t)(lb + k)] = (int32_t)(fb + m.list_faces[k]);
                voff[(size_t)(cb + x + 1)] = (int32_t)(vb + m.visit_offsets[x + 1]);
                loff[(size_t)(cb + x + 1)] = (int32_t)(lb + m.list_offsets[x + 1]);
                if (lsq) {
                    if (!m.lsq_r || !m.lsq_beta) throw std::invalid_argument("least squares needs lsq_r and lsq_beta");
                    for (int r = 0; r < 6; ++r) lsq_r[(size_t)r * N + cb + x] = m.lsq_r[(size_t)r * Ni + x];
                    lsq_b[(size_t)(cb + x)] = m.lsq_beta[x];
                }
            }
            for (int64_t g = 0; g < Gi; ++g) {
                gface[(size_t)(gb + g)] = (int32_t)(fb + m.ghost_face[g]);
                if (m.ghost_bc[g] < 0 || m.ghost_bc[g] >= settings[i].n_bc) throw std::invalid_argument("ghost_bc out of range");
                gbc[(size_t)(gb + g)] = (int32_t)(bb + m.ghost_bc[g]);
            }
            for (int k = 0; k < settings[i].n_bc; ++k) bcs.push_back(settings[i].bc[k]);
            cb += Ni;
            gb += Gi;
            fb += Fi;
            vb += m.visit_offsets[Ni];
            lb += m.list_offsets[Ni];
            bb += settings[i].n_bc;
            bs->cell_begin.push_back((int32_t)cb);
            bs->ghost_begin.push_back((int32_t)gb);
        }
        double vmin = meshes[0].min_volume;
        for (int i = 1; i < n; ++i) vmin = std::min(vmin, meshes[i].min_volume);
        psst_mesh_desc um{};
        um.dim = dim;
        um.n_cells = (int32_t)N;
        um.n_ghosts = (int32_t)G;
        um.n_faces = (int32_t)F;
        um.cell_volume = vol.data();
        um.cell_centre = cen.data();
        um.face_normal = nrm.data();
        um.face_area = area.data();
        um.face_cells = fc.data();
        um.face_beta = beta.data();
        um.visit_offsets = voff.data();
        um.visit_faces = vf.data();
        um.list_offsets = loff.data();
        um.list_faces = lf.data();
        um.ghost_face = gface.data();
        um.ghost_bc = gbc.data();
        um.lsq_r = lsq ? lsq_r.data() : nullptr;
        um.lsq_beta = lsq ? lsq_b.data() : nullptr;
        um.min_volume = vmin;
        um.face_shift = any_shift ? fshift.data() : nullptr;
        psst_settings us = s0;
        us.fused = 1;
        us.bc = bcs.data();
        us.n_bc = (int32_t)bcs.size();
        psst_ctx *c = nullptr;
        psst_status st = create_impl(&um, &us, seg, &c);
        if (st != PSST_OK) return st;
        auto bt = std::make_unique<psst_batch>();
        bt->ctx = c;
        if (!c->fused) {
            psst_ctx_destroy(c);
            bt->ctx = nullptr;
            throw std::invalid_argument("the batch's tiles do not fit in LDS (use smaller tile_cells)");
        }
        // tiles never straddle instances: the tiles of instance i are those
        // whose first cell lies in [cell_begin[i], cell_begin[i+1])
        bs->tile_begin.assign(1, 0);
        std::vector<int32_t> tile_inst((size_t)c->n_tiles);
        for (int t = 0, i = 0; t < c->n_tiles; ++t) {
            const int32_t cb0 = c->tile_cell_begin[(size_t)t];
            while (cb0 >= bs->cell_begin[(size_t)i + 1]) {
                bs->tile_begin.push_back(t);
                ++i;
            }
            if (c->tile_cell_begin[(size_t)t + 1] > bs->cell_begin[(size_t)i + 1])
                throw std::logic_error("a tile straddles two instances");
            tile_inst[(size_t)t] = i;
        }
        while ((int)bs->tile_begin.size() < n + 1) bs->tile_begin.push_back(c->n_tiles);
        bs->c2d.resize((size_t)NT);
        for (int64_t d = 0; d < NT; ++d) bs->c2d[(size_t)c->d2c[(size_t)d]] = (int32_t)d;
        bs->have_state.assign((size_t)n, 0);
        bs->enabled.assign((size_t)n, 1);
        bs->buf_of.assign((size_t)n, 0);
        const int ws = dim + 2;
        psst_ctx *cx = c;
        bs->tile_inst = dev_upload(cx, tile_inst.data(), tile_inst.size());
        bs->tile_begin_dev = dev_upload(cx, bs->tile_begin.data(), bs->tile_begin.size());
        bs->list = dev_alloc<int>(cx, (size_t)c->n_tiles);
        std::vector<int> zeros_i((size_t)std::max(n, 8 * 9), 0), minus1((size_t)n, -1);
        bs->count = dev_upload(cx, zeros_i.data(), 1);
        bs->iter = dev_upload(cx, zeros_i.data(), 1);
        bs->active = dev_upload(cx, zeros_i.data(), (size_t)n);
        bs->frozen_buf = dev_upload(cx, zeros_i.data(), (size_t)n);
        bs->conv = dev_upload(cx, minus1.data(), (size_t)n);
        std::vector<unsigned long long> zres((size_t)n * ws * kResSlots, 0ull);
        bs->res = dev_upload(cx, zres.data(), zres.size());
        std::vector<double> mres((size_t)n * ws, -1.0), zd((size_t)n, 0.0);
        bs->res0 = dev_upload(cx, mres.data(), mres.size());
        bs->drop = dev_upload(cx, zd.data(), zd.size());
        bs->hist = dev_alloc<double>(cx, (size_t)bs->H * n * ws);
        c->batch = std::move(bs);
        // after the move, so that the context's destroy frees them on any
        // later failure
        BatchState &b = *c->batch;
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&b.h_flags), sizeof(int) * ((size_t)n + 1)),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&b.h_ring), sizeof(double) * (size_t)b.H * n * ws),
                  "hipHostMalloc");
        *out = bt.release();
        return PSST_OK;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        return dynamic_cast<const HipError *>(&e) ? PSST_HIP_ERROR : PSST_BAD_ARG;
    }
}

psst_ctx *psst_batch_context(psst_batch *b) { return b ? b->ctx : nullptr; }

psst_status psst_batch_set_state(psst_batch *bt, int32_t inst, const double *prim) {
    if (!bt || !prim) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    try {
        const int32_t cb = b.cell_begin[(size_t)inst], Ni = b.cell_begin[(size_t)inst + 1] - cb;
        const int32_t gb = b.ghost_begin[(size_t)inst], Gi = b.ghost_begin[(size_t)inst + 1] - gb;
        const size_t NTi = (size_t)Ni + Gi;
        std::vector<double> vi((size_t)Ni), vg((size_t)Gi);
        for (int v = 0; v < c->dim + 3; ++v) {
            for (int32_t x = 0; x < Ni; ++x) vi[(size_t)(b.c2d[(size_t)(cb + x)] - cb)] = prim[(size_t)v * NTi + x];
            for (int32_t g = 0; g < Gi; ++g) vg[(size_t)g] = prim[(size_t)v * NTi + Ni + g];
            for (int k = 0; k < 2; ++k) {
                upload_values(c, c->q[k], (size_t)v * c->NT + cb, vi);
                upload_values(c, c->q[k], (size_t)v * c->NT + b.N_tot + gb, vg);
            }
        }
        const int m1 = -1;
        hip_check(hipMemcpy(b.conv + inst, &m1, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
        b.have_state[(size_t)inst] = 1;
        b.buf_of[(size_t)inst] = c->cur;
        b.w_dirty = true;
        c->have_state = true;
        return PSST_OK;
    } catch (const std::exception &e) {
        bt->err = e.what();
        return PSST_HIP_ERROR;
    }
}

psst_status psst_batch_get_state(psst_batch *bt, int32_t inst, double *prim) {
    if (!bt || !prim) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    try {
        const int32_t cb = b.cell_begin[(size_t)inst], Ni = b.cell_begin[(size_t)inst + 1] - cb;
        const int32_t gb = b.ghost_begin[(size_t)inst], Gi = b.ghost_begin[(size_t)inst + 1] - gb;
        const size_t NTi = (size_t)Ni + Gi;
        const void *src = c->q[b.buf_of[(size_t)inst]];
        for (int v = 0; v < c->dim + 3; ++v) {
            std::vector<double> vi = download_values(c, src, (size_t)v * c->NT + cb, (size_t)Ni);
            std::vector<double> vg = download_values(c, src, (size_t)v * c->NT + b.N_tot + gb, (size_t)Gi);
            for (int32_t x = 0; x < Ni; ++x) prim[(size_t)v * NTi + x] = vi[(size_t)(b.c2d[(size_t)(cb + x)] - cb)];
            for (int32_t g = 0; g < Gi; ++g) prim[(size_t)v * NTi + Ni + g] = vg[(size_t)g];
        }
        return PSST_OK;
    } catch (const std::exception &e) {
        bt->err = e.what();
        return PSST_HIP_ERROR;
    }
}

psst_status psst_batch_set_residual_reference(psst_batch *bt, int32_t inst, const double *res0) {
    if (!bt || !res0) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    if (hipMemcpy(b.res0 + (size_t)inst * b.ws, res0, sizeof(double) * b.ws, hipMemcpyHostToDevice) != hipSuccess) {
        bt->err = "hipMemcpy failed";
        return PSST_HIP_ERROR;
    }
    return PSST_OK;
}

psst_status psst_batch_set_enabled(psst_batch *bt, int32_t inst, int32_t enabled) {
    if (!bt) return PSST_BAD_ARG;
    BatchState &b = *bt->ctx->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    b.enabled[(size_t)inst] = enabled != 0;
    return PSST_OK;
}

psst_status psst_batch_set_graph(psst_batch *bt, int32_t enable) {
    if (!bt) return PSST_BAD_ARG;
    bt->ctx->batch->use_graph = enable != 0;
    return PSST_OK;
}

psst_status psst_batch_run(psst_batch *bt, int32_t max_iter, double tol, int32_t *iters, int32_t *status,
                           double *drop_out, double *res_hist) {
    if (!bt) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    for (int i = 0; i < b.n; ++i)
        if (!b.have_state[(size_t)i]) {
            bt->err = "psst_batch_run: instance " + std::to_string(i) + " has no state";
            return PSST_BAD_ARG;
        }
    if (max_iter <= 0) return PSST_OK;
    const int n = b.n, ws = b.ws;
    try {
        batch_gather_states(c);
        if (b.w_dirty) {
            if (c->dim == 2) c->s.precision == 64 ? run_set_w<2, double>(c) : run_set_w<2, float>(c);
            else c->s.precision == 64 ? run_set_w<3, double>(c) : run_set_w<3, float>(c);
            b.w_dirty = false;
        }
        std::vector<int> conv((size_t)n), active((size_t)n);
        hip_check(hipMemcpy(conv.data(), b.conv, sizeof(int) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        for (int i = 0; i < n; ++i) {
            active[(size_t)i] = conv[(size_t)i] != -2 && b.enabled[(size_t)i];
            if (conv[(size_t)i] != -2) conv[(size_t)i] = -1;
        }
        hip_check(hipMemcpy(b.active, active.data(), sizeof(int) * n, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(b.conv, conv.data(), sizeof(int) * n, hipMemcpyHostToDevice), "hipMemcpy");
        launch_check(c, tol, c->cur, 1);
        hip_check(hipMemcpyAsync(b.h_flags + n, b.count, sizeof(int), hipMemcpyDeviceToHost, c->stream),
                  "hipMemcpy");
        hip_check(hipStreamSynchronize(c->stream), "psst_batch_run");
        int count = b.h_flags[n];
        // the active list only shrinks during a run, so a capacity chosen at
        // a chunk's start holds it for the whole chunk
        b.grid_tiles = batch_grid_tiles(count, c->n_tiles);
        c->bd = BatchDev{b.list, b.count, b.tile_inst, b.res};
        const int64_t it0 = b.iter_host;
        int done = 0;
        bool any = true, first = true;
        while (done < max_iter && any) {
            const int m = std::min(b.H, max_iter - done);
            for (int k = 0; k < m; ++k) {
                if (first || !b.use_graph || c->timing) batch_iteration(c, tol);
                else batch_graph_iteration(c, tol);
                first = false;
            }
            if (res_hist)
                hip_check(hipMemcpyAsync(b.h_ring, b.hist, sizeof(double) * (size_t)b.H * n * ws,
                                         hipMemcpyDeviceToHost, c->stream),
                          "hipMemcpy");
            hip_check(hipMemcpyAsync(b.h_flags, b.active, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpy");
            hip_check(hipMemcpyAsync(b.h_flags + n, b.count, sizeof(int), hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpy");
            hip_check(hipStreamSynchronize(c->stream), "psst_batch_run");
            std::copy(b.h_flags, b.h_flags + n, active.begin());
            count = b.h_flags[n];
            b.grid_tiles = batch_grid_tiles(count, c->n_tiles);
            if (c->timing) flush_timing(c);
            if (res_hist)
                for (int k = 0; k < m; ++k) {
                    const size_t row = (size_t)((b.iter_host + k) % b.H);
                    std::memcpy(res_hist + (size_t)(done + k) * n * ws, b.h_ring + row * n * ws,
                                sizeof(double) * n * ws);
                }
            b.iter_host += m;
            done += m;
            any = false;
            for (int i = 0; i < n; ++i) any = any || active[(size_t)i];
        }
        c->bd = BatchDev{};
        b.grid_tiles = 0;
        std::vector<int> frozen((size_t)n);
        std::vector<double> drop((size_t)n);
        hip_check(hipMemcpy(conv.data(), b.conv, sizeof(int) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(frozen.data(), b.frozen_buf, sizeof(int) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(drop.data(), b.drop, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        psst_status worst = PSST_OK;
        for (int i = 0; i < n; ++i) {
            const int cv = conv[(size_t)i];
            b.buf_of[(size_t)i] = active[(size_t)i] ? c->cur : frozen[(size_t)i];
            psst_status si = PSST_OK;
            if (cv == -2) si = PSST_NAN_DETECTED;
            else if (cv < 0 && tol > 0.0) si = PSST_NO_CONVERGENCE;
            if (!b.enabled[(size_t)i] && cv != -2) si = PSST_OK;
            if (iters) iters[i] = cv >= 0 ? (int32_t)(cv - it0 + 1) : (b.enabled[(size_t)i] ? done : 0);
            if (status) status[i] = si;
            if (drop_out) drop_out[i] = drop[(size_t)i];
            if (si == PSST_NAN_DETECTED) worst = PSST_NAN_DETECTED;
            else if (si == PSST_NO_CONVERGENCE && worst == PSST_OK) worst = PSST_NO_CONVERGENCE;
        }
        if (worst == PSST_NAN_DETECTED) bt->err = "non-finite residual in at least one instance";
        else if (worst == PSST_NO_CONVERGENCE) bt->err = "at least one instance did not reach the tolerance";
        return worst;
    } catch (const std::exception &e) {
        c->bd = BatchDev{};
        b.grid_tiles = 0;
        bt->err = e.what();
        return PSST_HIP_ERROR;
    }
}

psst_status psst_batch_set_precision(psst_batch *bt, int32_t precision) {
    if (!bt) return PSST_BAD_ARG;
    psst_status st = psst_set_precision(bt->ctx, precision);
    if (st != PSST_OK) bt->err = bt->ctx->err;
    return st;
}

psst_status psst_batch_info(const psst_batch *bt, int64_t *info) {
    if (!bt || !info) return PSST_BAD_ARG;
    const psst_ctx *c = bt->ctx;
    info[0] = c->batch->n;
    info[1] = c->N;
    info[2] = c->G;
    info[3] = c->F;
    info[4] = c->n_tiles;
    info[5] = c->tile_cells;
    info[6] = c->tile_threads;
    info[7] = c->batch->iter_host;
    return PSST_OK;
}

void psst_batch_destroy(psst_batch *bt) { delete bt; }

psst_status psst_autotune(const psst_mesh_desc *m, const psst_settings *s, const double *prim, int32_t trial_iters,
                          int32_t max_cand, int32_t *cand, double *ms_per_iter, int32_t *n_cand, int32_t *best) {
    if (!m || !s || !prim || trial_iters <= 0 || !best) {
        g_create_error = "psst_autotune: bad argument";
        return PSST_BAD_ARG;
    }
    // tile shapes: cells per tile x threads per workgroup (the kernels are
    // built for 256 and 512 threads)
    static const int shapes[][2] = {{128, 256}, {160, 256}, {192, 256}, {224, 256}, {256, 256}, {256, 512},
                                    {384, 256}, {384, 512}, {512, 256}, {512, 512}, {768, 512}, {1024, 512}};
    double best_ms = 1e300;
    int k_out = 0;
    best[0] = best[1] = 0;
    for (const auto &sh : shapes) {
        psst_settings t = *s;
        t.fused = 1;
        t.tile_cells = sh[0];
        t.tile_threads = sh[1];
        psst_ctx *c = nullptr;
        double ms = -1.0;
        if (create_impl(m, &t, {}, &c) == PSST_OK) {
            if (c->fused && psst_set_state(c, prim) == PSST_OK) {
                int32_t done = 0;
                if (psst_run(c, 2, 0.0, nullptr, &done) == PSST_OK) {
                    const auto t0 = std::chrono::steady_clock::now();
                    if (psst_run(c, trial_iters, 0.0, nullptr, &done) == PSST_OK)
                        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                             trial_iters;
                }
            }
            psst_ctx_destroy(c);
        }
        if (cand && k_out < max_cand) {
            cand[2 * k_out] = sh[0];
            cand[2 * k_out + 1] = sh[1];
        }
        if (ms_per_iter && k_out < max_cand) ms_per_iter[k_out] = ms;
        ++k_out;
        if (ms > 0.0 && ms < best_ms) {
            best_ms = ms;
            best[0] = sh[0];
            best[1] = sh[1];
        }
    }
    if (n_cand) *n_cand = std::min(k_out, std::max(max_cand, 0));
    if (best[0] == 0) {
        g_create_error = "psst_autotune: no tile shape ran";
        return PSST_UNSUPPORTED;
    }
    return PSST_OK;
}

}  // extern "C"
