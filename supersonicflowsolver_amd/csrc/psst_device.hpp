// psst_device.hpp -- gfx950 kernels, the device context and the launch
// templates of the solver behind the C-ABI in include/psst_gpu.h.  Included
// by psst_device.hip (the C-ABI) and by psst_inst.hip, which is compiled once
// per (dimension, state type) pair to instantiate the launch templates, so
// the four kernel families build in parallel.
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
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
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

namespace psst_dev {

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
    const int *stop;                  // run control: >= 0 once the run has stopped; nullptr: never stops
};

// Kernels of an iteration return at once after the run stopped (tolerance
// met or a non-finite residual row), so the iterations a run has already
// queued behind the stopping one leave the state untouched.
__device__ __forceinline__ bool run_stopped(const Geom &g) { return g.stop != nullptr && *g.stop >= 0; }

// Run control of psst_run in device memory: the residual max-norm slots of
// the iteration in flight, the control words, the normalising row and the
// residual rows of the current chunk.  ctl and res0 are contiguous, so one
// copy brings both back.
constexpr int kRunRows = 1024;  // rows of one chunk
struct RunScratch {
    unsigned long long slots[5 * 64];  // [dim+2][kResSlots]
    int ctl[4];                        // [0] row the run stopped at (-1: running), [1] rows written,
                                       // [2] non-finite row seen, [3] a wait of the chunked sweep gave up
    double res0[5];                    // normalising row; res0[0] < 0: taken from the next row
    double rows[kRunRows * 5];         // [row][dim+2]
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
    if (run_stopped(g)) return;
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
    if (run_stopped(g)) return;
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
    if (run_stopped(g)) return;
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
    if (run_stopped(g)) return;
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
    if (run_stopped(g)) return;
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
    if (run_stopped(g)) return;
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
// face list as fixed-width entries (neighbour device id, with kEarlier set
// when the neighbour comes earlier in the reference's update loop and
// kRemote when it is also in another chunk; -1 past the end), face normals
// and areas.  R holds each cell's residual.
struct OrderedTables {
    const int *__restrict__ cell;   // [N] device cell, kProducer set when another chunk reads it
    const double *__restrict__ vol; // [N]
    const int *__restrict__ nb;     // [W][N]
    const double *__restrict__ fn;  // [DIM][W][N]
    const double *__restrict__ fa;  // [W][N]
    void *R;                        // [DIM+2][N], state precision
    const int *__restrict__ pos;    // [N] position of each device cell
    int W;                          // 4 or kOrderedW: the kernels' table width (unused rows: -1, 0)
};

constexpr int kOrderedW = 8;
constexpr int kEarlier = 1 << 30;   // nb: the neighbour comes earlier in the update loop
constexpr int kRemote = 1 << 29;    // nb: ... and belongs to another chunk of the chunked sweep
constexpr int kCellMask = kRemote - 1;
constexpr int kProducer = 1 << 30;  // cell: another chunk reads its new state

// Pre-pass of the ordered sweep: every cell's residual (the ordered gather
// of update_cell) at its position in the sweep's tables, and the residual
// max-norm; neither depends on the sweep.  It also clears the progress words
// of the chunked sweep that follows it.
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_ordered_residual(Geom g, State<T> s, OrderedTables ot,
                                                             unsigned long long *res_row, int *progress,
                                                             int n_progress) {
    if (run_stopped(g)) return;
    constexpr int WS = DIM + 2;
    // one thread per device cell (neighbouring cells share faces and cache
    // lines), the residual stored at the cell's position in the tables
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c < n_progress) progress[c] = 0;  // chunked sweep: no chunk has published in this stage yet
    unsigned long long rmax[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) rmax[j] = 0;
    if (c < g.N) {
        const int p = ot.pos[c];
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
// 90-131) in the same operation order.  A cell is loaded in three steps, so
// that a sweep can run them for different levels at once: ordered_head (the
// device cell and its neighbours, by position), ordered_fetch (everything
// else that no cell of the sweep writes: the residual, the snapshot, the
// cell's own and its later neighbours' stage-start state, the face data) and
// ordered_apply (the earlier neighbours' new state, then the update).  NW is
// the table width the kernel is built for (at least the tables' W).
template <int NW>
struct OrderedHead {
    int c;
    int nbw[NW];
};

template <int DIM, typename T, int NW>
struct OrderedIn {
    int c;
    bool producer;
    T R[DIM + 2], w0[DIM + 2], uc[DIM], Tc, V;
    int nbw[NW];
    double fnw[NW][DIM], faw[NW];
    T un[NW][DIM], Tn[NW];  // stage-start state of the neighbours that are not earlier
};

template <int NW>
__device__ __forceinline__ void ordered_head(const OrderedTables &ot, int N, int p, OrderedHead<NW> &h) {
    h.c = ot.cell[p] & kCellMask;
#pragma unroll
    for (int w = 0; w < NW; ++w) h.nbw[w] = ot.nb[(size_t)w * N + p];
}

template <int DIM, typename T, int NW>
__device__ __forceinline__ void ordered_fetch(const Geom &g, const State<T> &s, const OrderedTables &ot, int p,
                                              const OrderedHead<NW> &h, OrderedIn<DIM, T, NW> &o) {
    constexpr int WS = DIM + 2;
    // the cell and its neighbours are loaded again (cache hits) rather than
    // copied from the head: the head's registers are reloaded for the next
    // level while these are still in use, and a copy would have to wait for
    // that reload
    const int craw = ot.cell[p];
    o.c = craw & kCellMask;
    o.producer = (craw & kProducer) != 0;

    const T *Rp = static_cast<const T *>(ot.R);
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        o.R[j] = Rp[(size_t)j * g.N + p];
        o.w0[j] = s.W0[(size_t)j * g.N + h.c];
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k) o.uc[k] = s.q[(size_t)k * g.NT + h.c];
    o.Tc = s.q[(size_t)(DIM + 1) * g.NT + h.c];
    o.V = T(ot.vol[p]);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        o.nbw[w] = ot.nb[(size_t)w * g.N + p];
        o.faw[w] = ot.fa[(size_t)w * g.N + p];
#pragma unroll
        for (int k = 0; k < DIM; ++k) o.fnw[w][k] = ot.fn[((size_t)k * NW + w) * g.N + p];
        // a later neighbour (ghosts included) keeps its stage-start state
        // through the sweep; padding (-1) and earlier neighbours read cell 0
        const int nb = (h.nbw[w] < 0 || (h.nbw[w] & kEarlier)) ? 0 : h.nbw[w];
#pragma unroll
        for (int k = 0; k < DIM; ++k) o.un[w][k] = s.q[(size_t)k * g.NT + nb];
        o.Tn[w] = s.q[(size_t)(DIM + 1) * g.NT + nb];
    }
}

// Loads and stores of the new state that workgroups of the chunked sweep on
// other XCDs read: agent-scope (sc1) accesses, which go past the CU's L1 and
// the XCD's L2, so no cache write-back or invalidation is needed between
// chunks.  A chunk's own workgroup reads its cells with plain loads (a CU
// sees its own stores in order).
template <bool COH, typename T>
__device__ __forceinline__ T qn_load(const T *p) {
    if (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
template <bool COH, typename T>
__device__ __forceinline__ void qn_store(T *p, T v) {
    if (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// The earlier neighbours' new state of one cell, as loaded by
// ordered_load_new: plain loads, or agent-scope loads for a neighbour in
// another chunk.
template <int DIM, typename T, int NW>
struct OrderedNew {
    T qn[NW][DIM + 1];  // (u, T) per neighbour slot
};

// Issues the loads of the new state of the cell's earlier neighbours
// (ordered_finish uses them).  Each kind of load is masked to the lanes that
// need it and writes registers of its own, and the choice is made after:
// loads of different kinds into the same registers on diverged lanes would
// make the compiler wait for every load in flight between them.  A wave
// with no lane needing a kind skips its loads.
template <int DIM, typename T, int NW, bool COH>
__device__ __forceinline__ void ordered_load_new(const Geom &g, const State<T> &s, const OrderedIn<DIM, T, NW> &o,
                                                 OrderedNew<DIM, T, NW> &nw) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int e = o.nbw[w];
        const bool early = e >= 0 && (e & kEarlier);
        const bool remote = COH && early && (e & kRemote);
        const int nb = e & kCellMask;
        T gl[DIM + 1], rn[DIM + 1];
        if (early && !remote)
#pragma unroll
            for (int k = 0; k <= DIM; ++k) gl[k] = s.qn[(size_t)(k < DIM ? k : DIM + 1) * g.NT + nb];
        if (remote)
#pragma unroll
            for (int k = 0; k <= DIM; ++k) rn[k] = qn_load<true>(s.qn + (size_t)(k < DIM ? k : DIM + 1) * g.NT + nb);
#pragma unroll
        for (int k = 0; k <= DIM; ++k) nw.qn[w][k] = remote ? rn[k] : gl[k];
    }
}

// The update of one cell from its loaded inputs: neighbour states (the new
// one for neighbours earlier in the update loop, the stage-start one loaded
// ahead otherwise), the time step, the conservative and primitive update.
template <int DIM, typename T, int NW, bool COH>
__device__ __forceinline__ void ordered_finish(const Geom &g, const State<T> &s, const Phys &ph, T alpha,
                                               int last_stage, const OrderedIn<DIM, T, NW> &o,
                                               const OrderedNew<DIM, T, NW> &nw) {
    constexpr int WS = DIM + 2;
    const int c = o.c;
    T un[NW][DIM], Tn[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const bool early = o.nbw[w] >= 0 && (o.nbw[w] & kEarlier);
#pragma unroll
        for (int k = 0; k < DIM; ++k) un[w][k] = early ? nw.qn[w][k] : o.un[w][k];
        Tn[w] = early ? nw.qn[w][DIM] : o.Tn[w];
    }
    const T gr = T(ph.gamma) * T(ph.rm);
    T cc = dsqrt(gr * o.Tc);
    T den = T(0), nn = T(0);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
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
    T vsq = T(0), qv[DIM + 3];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        T u = fdiv(wn[k], rho);
        qv[k] = u;
        vsq += u * u;
    }
    T Tnew = fdiv(fdiv(wn[DIM + 1], rho) - (vsq * T(0.5)), T(ph.cv));
    qv[DIM] = rho * T(ph.rm) * Tnew;
    qv[DIM + 1] = Tnew;
    qv[DIM + 2] = rho;
    // a cell another chunk reads goes to memory with agent-scope stores;
    // then every cell is stored plainly, which leaves the line in the XCD's
    // L2 for the chunk's own next levels
    if (COH && o.producer)
#pragma unroll
        for (int k = 0; k < DIM + 3; ++k) qn_store<true>(s.qn + (size_t)k * g.NT + c, qv[k]);
#pragma unroll
    for (int k = 0; k < DIM + 3; ++k) s.qn[(size_t)k * g.NT + c] = qv[k];
}

template <int DIM, typename T, int NW, bool COH>
__device__ __forceinline__ void ordered_apply(const Geom &g, const State<T> &s, const Phys &ph, T alpha,
                                              int last_stage, const OrderedIn<DIM, T, NW> &o) {
    OrderedNew<DIM, T, NW> nw;
    ordered_load_new<DIM, T, NW, COH>(g, s, o, nw);
    ordered_finish<DIM, T, NW, COH>(g, s, ph, alpha, last_stage, o, nw);
}

// All three steps of one position at once (cells beyond a workgroup's first
// pass over a level, and the unpipelined sweep).
template <int DIM, typename T, int NW, bool COH>
__device__ __forceinline__ void ordered_update(const Geom &g, const State<T> &s, const Phys &ph, T alpha,
                                               int last_stage, const OrderedTables &ot, int p) {
    OrderedHead<NW> h;
    ordered_head<NW>(ot, g.N, p, h);
    OrderedIn<DIM, T, NW> o;
    ordered_fetch<DIM, T, NW>(g, s, ot, p, h, o);
    ordered_apply<DIM, T, NW, COH>(g, s, ph, alpha, last_stage, o);
}

// The three-step pipeline of a workgroup walking a sequence of levels, with
// one thread per position on the first pass over each level: while the
// cells of level e are updated, the thread's position of level e + 1 is
// fetched (its head was loaded one level earlier) and the head of its
// position of level e + 2 is loaded.  lev_pos(e) is the first position of
// level e of the sequence (e = n_lev: one past the last).  Every thread
// loads at every step, a thread without a position of that level at the
// level's last position, so the number of loads in flight never depends on
// the data and the compiler's waits before the next uses are exact.
template <int DIM, typename T, int NW>
struct SweepPipe {
    OrderedIn<DIM, T, NW> o;  // level e
    OrderedHead<NW> h;        // level e + 1
    bool have = false, have_h = false;
    int ph = 0;               // position of the head

    template <typename LevPos>
    __device__ __forceinline__ static int pos(int e, int n_lev, int tid, bool &valid, LevPos lev_pos) {
        const int ec = e < n_lev ? e : n_lev - 1;
        const int b = lev_pos(ec), en = lev_pos(ec + 1);
        valid = e < n_lev && b + tid < en;
        return b + tid < en ? b + tid : en - 1;
    }
    template <typename LevPos>
    __device__ __forceinline__ void start(const Geom &g, const State<T> &s, const OrderedTables &ot, int n_lev,
                                          int tid, LevPos lev_pos) {
        const int p = pos(0, n_lev, tid, have, lev_pos);
        OrderedHead<NW> h0;
        ordered_head<NW>(ot, g.N, p, h0);
        ordered_fetch<DIM, T, NW>(g, s, ot, p, h0, o);
        ph = pos(1, n_lev, tid, have_h, lev_pos);
        ordered_head<NW>(ot, g.N, ph, h);
    }
    // after level e's updates: fetch level e + 1, load the head of e + 2
    template <typename LevPos>
    __device__ __forceinline__ void advance(const Geom &g, const State<T> &s, const OrderedTables &ot, int e,
                                            int n_lev, int tid, LevPos lev_pos) {
        have = have_h;
        ordered_fetch<DIM, T, NW>(g, s, ot, ph, h, o);
        ph = pos(e + 2, n_lev, tid, have_h, lev_pos);
        ordered_head<NW>(ot, g.N, ph, h);
    }
};

// Tables of the chunked ordered sweep (host side: ordered_sched.cpp).  One
// workgroup per chunk of cells walks the chunk's entries (one entry = the
// chunk's cells of one level, at positions [ent_pos[e], ent_pos[e+1]) of the
// OrderedTables).  Before an entry it waits until every chunk
// wait_chunk[i], i in [ent_wait[e], ent_wait[e+1]), has finished level
// wait_level[i]; after an entry whose ent_level carries kChunkPublish it
// publishes progress = level + 1.  A chunk has at most kChunkMaxEntries
// entries and kChunkMaxWaits waits (the host checks).
struct ChunkTables {
    const int *__restrict__ ch_ent;      // [n_chunks+1]
    const int *__restrict__ ent_pos;     // [E+1]
    const int *__restrict__ ent_level;   // [E] level | kChunkPublish
    const int *__restrict__ ent_wait;    // [E+1]
    const int2 *__restrict__ waits;      // [waits] (chunk, level)
    int *progress;                       // [n_chunks]: 1 + the last level published, 0 at the start of a stage
    int *timed_out;                      // set to 1 when a wait gives up; then every wait gives up
    int n_chunks;
    int threads;                         // workgroup size: 128 when no level of a chunk is wider, else 256
};
constexpr int kChunkPublish = 1 << 30;
constexpr int kChunkMaxEntries = 1024;
constexpr int kChunkMaxWaits = 1024;
// Polls of one wait before it gives up (a second or two): a wait ends once
// the chunk it waits for publishes, which the schedule guarantees; the bound
// only keeps a broken schedule from holding the device.  The run then fails.
constexpr int kChunkSpinLimit = 1 << 21;

__device__ __forceinline__ int progress_load(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spins until *flag >= need; gives up (and makes every other wait give up)
// after kChunkSpinLimit polls.
__device__ __forceinline__ void progress_wait(const int *flag, int need, int *timed_out) {
    for (int spins = 0; progress_load(flag) < need; ++spins) {
        if ((spins & 1023) == 1023 && progress_load(timed_out)) break;
        if (spins > kChunkSpinLimit) {
            __hip_atomic_store(timed_out, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// The ordered sweep as one cooperative launch of one workgroup per chunk
// (all resident, so every wait ends).  A chunk's levels are separated by a
// workgroup barrier (the CU's own accesses reach the L2 in order).  Across
// chunks the new state goes through agent-scope loads and stores: before a
// publishing barrier every wave waits until its stores are complete
// (s_waitcnt vmcnt(0)), so the progress word thread 0 then stores is seen
// only after the cells it covers.
// The chunk's entry tables are staged in LDS (vector loads of them would
// make every wait on them also wait for the prefetches in flight).  The
// cell loads are pipelined (SweepPipe), and the first wave reads the
// progress words an entry ahead, before that entry's prefetches: when the
// waited-for chunks are ahead, the check finds the words already read and
// the level's chain is only the earlier neighbours' new state and the update.
template <int DIM, typename T, int NW>
__global__ void __launch_bounds__(kBlock) k_update_chunks(Geom g, State<T> s, Phys ph, T alpha, int last_stage,
                                                          OrderedTables ot, ChunkTables ct) {
    __shared__ int s_pos[kChunkMaxEntries + 1], s_lev[kChunkMaxEntries], s_wait[kChunkMaxEntries + 1];
    __shared__ int2 s_waits[kChunkMaxWaits];
    if (run_stopped(g)) return;
    const int w = (int)blockIdx.x, tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    const int e0 = ct.ch_ent[w], n_ent = ct.ch_ent[w + 1] - e0;
    const int w0 = ct.ent_wait[e0], n_waits = ct.ent_wait[e0 + n_ent] - w0;
    for (int i = tid; i <= n_ent; i += nthr) {
        s_pos[i] = ct.ent_pos[e0 + i];
        s_wait[i] = ct.ent_wait[e0 + i] - w0;
        if (i < n_ent) s_lev[i] = ct.ent_level[e0 + i];
    }
    for (int i = tid; i < n_waits; i += nthr) s_waits[i] = ct.waits[w0 + i];
    __syncthreads();
    auto lev_pos = [](int e) { return s_pos[e]; };
    // first wave: lane i reads the progress word of wait i of the next entry
    // (lanes without one read the chunk's own word and need nothing)
    const bool poller = tid < 64;
    int poll_v = 0, poll_need = 0;
    auto poll_ahead = [&](int e) {
        if (!poller) return;
        const int wb = e < n_ent ? s_wait[e] : 0, nw = e < n_ent ? s_wait[e + 1] - wb : 0;
        const int2 wt = tid < nw ? s_waits[wb + tid] : make_int2(w, -1);
        poll_need = wt.y + 1;
        poll_v = progress_load(ct.progress + wt.x);
    };
    poll_ahead(0);
    SweepPipe<DIM, T, NW> pipe;
    pipe.start(g, s, ot, n_ent, tid, lev_pos);
    for (int e = 0; e < n_ent; ++e) {
        const int wb = s_wait[e], nw = s_wait[e + 1] - wb;
        if (nw > 0) {
            if (poller && poll_v < poll_need) progress_wait(ct.progress + s_waits[wb + tid].x, poll_need, ct.timed_out);
            for (int i = 64 + tid; i < nw; i += nthr)  // more waits than one wave (rare)
                progress_wait(ct.progress + s_waits[wb + i].x, s_waits[wb + i].y + 1, ct.timed_out);
            __syncthreads();
        }
        const int b = s_pos[e], en = s_pos[e + 1];
        const int lv = s_lev[e];
        if (pipe.have) ordered_apply<DIM, T, NW, true>(g, s, ph, alpha, last_stage, pipe.o);
        for (int p = b + tid + nthr; p < en; p += nthr)
            ordered_update<DIM, T, NW, true>(g, s, ph, alpha, last_stage, ot, p);
        if (lv & kChunkPublish) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        poll_ahead(e + 1);
        pipe.advance(g, s, ot, e, n_ent, tid, lev_pos);
        __syncthreads();
        if ((lv & kChunkPublish) && tid == 0)
            __hip_atomic_store(ct.progress + w, (lv & (kChunkPublish - 1)) + 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The ordered sweep in one workgroup, a workgroup barrier between levels
// (small meshes or narrow levels).  With the residual gathered by the
// parallel pre-pass, each level is only the time step and update chain.
template <int DIM, typename T, int NW>
__global__ void __launch_bounds__(1024) k_update_levels_wg(Geom g, State<T> s, Phys ph, T alpha, int last_stage,
                                                           OrderedTables ot, const int *__restrict__ level_off,
                                                           int n_levels) {
    if (run_stopped(g)) return;
    for (int l = 0; l < n_levels; ++l) {
        const int b = level_off[l], e = level_off[l + 1];
        for (int p = b + (int)threadIdx.x; p < e; p += (int)blockDim.x)
            ordered_update<DIM, T, NW, false>(g, s, ph, alpha, last_stage, ot, p);
        __syncthreads();
    }
}

// Same for levels of at most 256 cells, pipelined (SweepPipe): a workgroup
// barrier waits for LDS operations only, so the loads of the next two levels
// stay in flight across it.  Nothing they read is written during the sweep
// (tables, residuals, the stage-start state, the cell's own snapshot).
template <int DIM, typename T, int NW>
__global__ void __launch_bounds__(256) k_update_levels_wg_prefetch(Geom g, State<T> s, Phys ph, T alpha,
                                                                   int last_stage, OrderedTables ot,
                                                                   const int *__restrict__ level_off, int n_levels) {
    if (run_stopped(g)) return;
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    auto lev_pos = [level_off](int e) { return level_off[e]; };
    SweepPipe<DIM, T, NW> pipe;
    pipe.start(g, s, ot, n_levels, tid, lev_pos);
    for (int l = 0; l < n_levels; ++l) {
        const int b = level_off[l], e = level_off[l + 1];
        if (pipe.have) ordered_apply<DIM, T, NW, false>(g, s, ph, alpha, last_stage, pipe.o);
        for (int p = b + tid + nthr; p < e; p += nthr)
            ordered_update<DIM, T, NW, false>(g, s, ph, alpha, last_stage, ot, p);
        pipe.advance(g, s, ot, l, n_levels, tid, lev_pos);
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
    if (run_stopped(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.G) return;
    bc_apply<DIM, T>(g, qn, bcs, ph, i);
}

// Conservative variables from primitives for the interior cells [c_begin, c_end)
// (psstFiniteVolumeGrid.cpp:279-300).
template <int DIM, typename T>
__global__ void __launch_bounds__(kBlock) k_set_w(Geom g, const T *__restrict__ q, T *__restrict__ W0, Phys ph,
                                                  int c_begin, int c_end) {
    const int c = c_begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= c_end) return;
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
    const double4 *__restrict__ gg;      // [TF] 2D: area, beta, n_x, n_y (the Green-Gauss kernel's face record)
    const double4 *__restrict__ sf;      // [TF] 2D: n_x, n_y, (x_0 - x_1)_x, (x_0 - x_1)_y (the stage kernel's)
    const double2 *__restrict__ sf3;     // [3 TF] 3D: (n_x, n_y), (n_z, (x_0 - x_1)_x), ((x_0 - x_1)_y, (x_0 - x_1)_z)
    const int *__restrict__ ghost_off;   // [T+1]
    const int *__restrict__ ghost;       // ghost device ids
    const short *__restrict__ visit;     // [Wv][N] local face | 0x4000 if neighbour 0, face-loop order
    const short *__restrict__ list;      // [Wl][N] local face, face-list order
    const short *__restrict__ lslot;     // [Wl][N] local slot of the neighbour across that face
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


// The face part of a stage on one tile: the tile's slots staged in LDS (sq
// [DIM+3][MS], scell [MS]), then every tile face's F * A into sfa
// [DIM+2][MF], a workgroup barrier after each.  The tile's own cells are
// slots [0, nc), its halo cells [nc, ns); its faces are tile faces
// [fb, fb + nf).
template <int DIM, typename T, int FLUX, int ORDER, int LIM>
__device__ __forceinline__ void stage_face_phase(const Geom &g, const TileDev &td, const State<T> &s, const Phys &ph,
                                                 int cb, int nc, int hb, int ns, int fb, int nf, T *sq, T *sfa,
                                                 int *scell) {
    constexpr int WS = DIM + 2;
    stage_slots<DIM, T, ORDER == 1>(g, td, s, cb, nc, hb, ns, sq, scell);
    __syncthreads();
    // ---- faces of the tile ----
    for (int j = threadIdx.x; j < nf; j += blockDim.x) {
        const unsigned e = td.ends[fb + j];
        const int l0 = end0(e), l1 = end1(e), kind = endkind(e);
        T nrm[DIM];
        // the face's normal and x_0 - x_1 as one record, in wide loads: on the
        // 4.9M-cell wedge 0.365 vs 0.373 ms (FP64; FP32 equal), on the
        // 12.6M-hexahedra wedge 1.98 vs 2.05 ms (FP64); 3D FP32 keeps the
        // separate arrays (1.41 vs 1.33 ms with the record)
        double sfr[2 * DIM];
        if constexpr (DIM == 2) {
            const double4 r = td.sf[fb + j];
            sfr[0] = r.x;
            sfr[1] = r.y;
            sfr[2] = r.z;
            sfr[3] = r.w;
        } else if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double2 r = td.sf3[(size_t)3 * (fb + j) + q];
                sfr[2 * q] = r.x;
                sfr[2 * q + 1] = r.y;
            }
        } else {
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                sfr[k] = td.fn[(size_t)k * td.TF + fb + j];
                sfr[DIM + k] = td.rlr[(size_t)k * td.TF + fb + j];
            }
        }
#pragma unroll
        for (int k = 0; k < DIM; ++k) nrm[k] = T(sfr[k]);
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
                for (int k = 0; k < DIM; ++k) RLR[k] = T(sfr[DIM + k]);
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
    stage_face_phase<DIM, T, FLUX, ORDER, LIM>(g, td, s, ph, cb, nc, hb, ns, fb, nf, sq, sfa, scell);
    // ---- cells of the tile ----
    unsigned long long rmax[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) rmax[j] = 0;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const int c = cb + i;
        // the residual is summed, and the update formed, in FP64 whatever the
        // state type: in FP32 mode the face fluxes and the state are FP32,
        // their sum and w0 - dt alpha R / V are not rounded to it before the
        // new state is (in FP64 mode this is the reference's own arithmetic)
        double R[WS];
#pragma unroll
        for (int j = 0; j < WS; ++j) R[j] = 0.0;
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
            w[j] = T(double(w0v[j]) - double(dt) * double(alpha) * R[j] / double(V));
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
    if (run_stopped(g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    for_each_tile<DIM + 2>(td, bd, res_row, [&](int t, unsigned long long *row) {
        stage_tile<DIM, T, FLUX, ORDER, LIM, LC>(g, td, s, ph, alpha, last_stage, row, bcs, t, smem);
    });
}

// The residual pass of the reference-ordered sweep (time step mode 1,
// psstFiniteVolumeGrid.cpp:311-332) on tiles: the stage's face part, then
// each own cell's gather in its visit order, summed in the state type as
// k_ordered_residual sums it, stored at the cell's position in the sweep's
// tables, and the residual max-norm.  One launch replaces k_flux and
// k_ordered_residual (no face flux array written and read back).  Workgroup
// 0 also clears the chunked sweep's progress words.
template <int DIM, typename T, int FLUX, int ORDER, int LIM, int NTHR>
__global__ void __launch_bounds__(NTHR) k_stage_residual(Geom g, TileDev td, State<T> s, Phys ph, OrderedTables ot,
                                                         unsigned long long *res_row, int *progress, int n_progress) {
    if (run_stopped(g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    if (blockIdx.x == 0)
        for (int i = (int)threadIdx.x; i < n_progress; i += (int)blockDim.x) progress[i] = 0;
    const int t = tile_of_block(td.n_tiles);
    if (t >= td.n_tiles) return;
    constexpr int NQ = DIM + 3, NV = DIM + 2, WS = DIM + 2;
    T *sq = reinterpret_cast<T *>(smem);
    T *sfa = sq + NQ * td.MS;
    int *scell = reinterpret_cast<int *>(sfa + NV * td.MF);
    const int cb = td.cell_begin[t], nc = td.cell_begin[t + 1] - cb;
    const int hb = td.halo_off[t], ns = nc + (td.halo_off[t + 1] - hb);
    const int fb = td.face_off[t], nf = td.face_off[t + 1] - fb;
    stage_face_phase<DIM, T, FLUX, ORDER, LIM>(g, td, s, ph, cb, nc, hb, ns, fb, nf, sq, sfa, scell);
    unsigned long long rmax[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) rmax[j] = 0;
    T *Rp = static_cast<T *>(ot.R);
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const int c = cb + i;
        const int p = ot.pos[c];
        T R[WS];
#pragma unroll
        for (int j = 0; j < WS; ++j) R[j] = T(0);
        for (int w = 0; w < td.Wv; ++w) {
            const int e = td.visit[(size_t)w * g.N + c];
            if (e < 0) break;
            const int lf = e & 0x3fff;
            const bool first = (e & 0x4000) != 0;
#pragma unroll
            for (int j = 0; j < WS; ++j) {
                const T v = sfa[j * td.MF + lf];
                R[j] = first ? R[j] - v : R[j] + v;
            }
        }
#pragma unroll
        for (int j = 0; j < WS; ++j) {
            Rp[(size_t)j * g.N + p] = R[j];
            const unsigned long long ab = abs_bits(R[j]);
            rmax[j] = rmax[j] > ab ? rmax[j] : ab;
        }
    }
    residual_max<WS>(rmax, res_row);
}

// Tiles of at most one cell and two slots per thread: the thread's slot
// cells (own, or the halo table's) are found, then their q is loaded and
// written to LDS.  `between` runs after the slot cells are known and before
// their q loads: the Green-Gauss kernel issues its cell's face loads there,
// so that those and the slot loads are in flight together before the
// workgroup's barrier (two dependent round trips to memory instead of four).
template <int NV, typename T, typename Between>
__device__ __forceinline__ void stage_two_slots(const Geom &g, const TileDev &td, const State<T> &s, int cb, int nc,
                                                int hb, int ns, T *sq, Between between) {
    const int ia = threadIdx.x, ib = threadIdx.x + blockDim.x;
    // past the tile's slots, a thread reads the tile's first cell (in bounds)
    const int ca = ia < nc ? cb + ia : (ia < ns ? td.halo[hb + ia - nc] : cb);
    const int cc = ib < nc ? cb + ib : (ib < ns ? td.halo[hb + ib - nc] : cb);
    between();
    T qa[NV], qb[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        qa[v] = s.q[(size_t)v * g.NT + ca];
        qb[v] = s.q[(size_t)v * g.NT + cc];
    }
    if (ia < ns)
#pragma unroll
        for (int v = 0; v < NV; ++v) sq[v * td.MS + ia] = qa[v];
    if (ib < ns)
#pragma unroll
        for (int v = 0; v < NV; ++v) sq[v * td.MS + ib] = qb[v];
}

// One face's Green-Gauss term for the cell in slot i (psstGradCalculation.cpp:
// 80-177): the face value beta q_0 + (1 - beta) q_1, signed outward, times
// n A / V.
template <int DIM, typename T>
__device__ __forceinline__ void gg_term(const TileDev &td, const T *sq, int i, double Vc, unsigned e, double fa,
                                        double be, const double *fn, T (&acc)[DIM + 2][DIM]) {
    constexpr int NV = DIM + 2;
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
}

// Green-Gauss on a tile of at most one cell per thread and at most U faces
// per cell: the cell's visit list, volume and face data are loaded with the
// slots, before the barrier.  On the 4.9M-cell wedge (FP64): 0.244 vs
// 0.265 ms against loading them after the barrier, bitwise unchanged.
template <int DIM, typename T, int U>
__device__ __forceinline__ void grad_gg_tile_prefetched(const Geom &g, const TileDev &td, const State<T> &s, T *sq,
                                                        int cb, int nc, int hb, int ns, int fb) {
    constexpr int NV = DIM + 2;
    const int i = threadIdx.x;
    const bool own = i < nc;
    const int c = cb + (own ? i : 0);
    // the visit list is padded with -1 to Wv <= U entries; a padding entry
    // reads the tile's face 0 (in bounds) and is skipped
    int ev[U];
#pragma unroll
    for (int w = 0; w < U; ++w) ev[w] = w < td.Wv ? (int)td.visit[(size_t)w * g.N + c] : -1;
    const double Vc = g.vol[c];
    unsigned ee[U];
    double fa[U], be[U], fn[U][DIM];
    stage_two_slots<NV, T>(g, td, s, cb, nc, hb, ns, sq, [&] {
#pragma unroll
        for (int w = 0; w < U; ++w) {
            const int lf = ev[w] < 0 ? 0 : (ev[w] & 0x3fff);
            ee[w] = td.ends[fb + lf];
            if constexpr (DIM == 2) {
                // one 32-byte record per face: two wide loads instead of four
                const double4 r = td.gg[fb + lf];
                fa[w] = r.x;
                be[w] = r.y;
                fn[w][0] = r.z;
                fn[w][DIM - 1] = r.w;
            } else {
                fa[w] = td.fa[fb + lf];
                be[w] = td.beta[fb + lf];
#pragma unroll
                for (int k = 0; k < DIM; ++k) fn[w][k] = td.fn[(size_t)k * td.TF + fb + lf];
            }
        }
    });
    __syncthreads();
    if (!own) return;
    T acc[NV][DIM];
#pragma unroll
    for (int n = 0; n < NV; ++n)
#pragma unroll
        for (int k = 0; k < DIM; ++k) acc[n][k] = T(0);
#pragma unroll
    for (int w = 0; w < U; ++w) {
        if (ev[w] < 0) break;
        gg_term<DIM, T>(td, sq, i, Vc, ee[w], fa[w], be[w], fn[w], acc);
    }
    store_grads<DIM, T>(s.grad, c, acc);
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
    // 2D tiles of quadrilaterals (or triangles) only: in 3D the face data of
    // a hexahedron held across the barrier would cost occupancy (the same
    // form of the least-squares kernel on the 12.6M-hexahedra wedge, 147
    // VGPRs against 100, took 1.40 ms against 1.15)
    if (DIM == 2 && nc <= (int)blockDim.x && ns <= 2 * (int)blockDim.x && td.Wv <= 4) {
        grad_gg_tile_prefetched<DIM, T, 4>(g, td, s, sq, cb, nc, hb, ns, fb);
        return;
    }
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
        auto term = [&](unsigned e, double fa, double be, const double *fn) {
            gg_term<DIM, T>(td, sq, i, Vc, e, fa, be, fn, acc);
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
                if constexpr (DIM == 2) {
                    const double4 r = td.gg[fb + lf];
                    fa[w] = r.x;
                    be[w] = r.y;
                    fn[w][0] = r.z;
                    fn[w][DIM - 1] = r.w;
                } else {
                    fa[w] = td.fa[fb + lf];
                    be[w] = td.beta[fb + lf];
#pragma unroll
                    for (int k = 0; k < DIM; ++k) fn[w][k] = td.fn[(size_t)k * td.TF + fb + lf];
                }
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
    if (run_stopped(g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    for_each_tile<DIM + 2>(td, bd, nullptr, [&](int t, unsigned long long *) {
        grad_gg_tile_body<DIM, T>(g, td, s, t, smem);
    });
}

// One neighbour's least-squares term for the cell in slot i
// (psstGradCalculation.cpp:268-368): the weights w theta_k times the
// difference of the neighbour's q and the cell's.
template <int DIM, typename T>
__device__ __forceinline__ void lsq_term(const TileDev &td, const T *sq, int i, int nbs, const double *wk,
                                         T (&acc)[DIM + 2][DIM]) {
    constexpr int NV = DIM + 2;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        T d = (sq[n * td.MS + nbs] - sq[n * td.MS + i]);
#pragma unroll
        for (int k = 0; k < DIM; ++k) acc[n][k] += T(wk[k]) * d;
    }
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
        auto term = [&](int nbs, const double *wk) { lsq_term<DIM, T>(td, sq, i, nbs, wk, acc); };
        auto unrolled = [&](auto width) {
            // every load of the cell's neighbours is issued before the first
            // term: the neighbour slots (a table per cell, so no load depends
            // on another) and the weights are padded to Wl entries (slot -1),
            // and padding only follows real entries
            constexpr int U = decltype(width)::value;
            int nv[U];
#pragma unroll
            for (int w = 0; w < U; ++w) nv[w] = w < td.Wl ? (int)td.lslot[(size_t)w * g.N + c] : -1;
            double wk[U][DIM];
#pragma unroll
            for (int w = 0; w < U; ++w)
#pragma unroll
                for (int k = 0; k < DIM; ++k)
                    wk[w][k] = w < td.Wl ? td.lsqc[((size_t)k * td.Wl + w) * g.N + c] : 0.0;
#pragma unroll
            for (int w = 0; w < U; ++w) {
                if (nv[w] < 0) break;
                term(nv[w], wk[w]);
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
                const int nbs = td.lslot[(size_t)w * g.N + c];
                if (nbs < 0) break;
                double wk[DIM];
#pragma unroll
                for (int k = 0; k < DIM; ++k) wk[k] = td.lsqc[((size_t)k * td.Wl + w) * g.N + c];
                term(nbs, wk);
            }
        }
        store_grads<DIM, T>(s.grad, c, acc);
    }
}

template <int DIM, typename T, int NTHR>
__global__ void __launch_bounds__(NTHR) k_grad_lsq_tile(Geom g, TileDev td, State<T> s, BatchDev bd) {
    if (run_stopped(g)) return;
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
// Residual drop max_j res_j / res0_j of a residual row against its
// normalising row.  A column whose normalising value is round-off against
// the row's largest (res0_j <= 1e-12 max_k res0_k, e.g. the z momentum of a
// z-symmetric 3D flow, zero up to rounding) carries no scale and is left
// out, as is a column with res0_j = 0.
constexpr double kDropColumnFloor = 1.0e-12;
template <int WS>
__device__ __forceinline__ double residual_drop(const double (&r)[WS], const double *res0) {
    double mx = 0.0;
#pragma unroll
    for (int j = 0; j < WS; ++j) mx = mx > res0[j] ? mx : res0[j];
    const double floor = kDropColumnFloor * mx;
    double d = 0.0;
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        const double r0 = res0[j];
        const double q = (r0 > 0.0 && r0 > floor) ? r[j] / r0 : 0.0;
        d = d > q ? d : q;
    }
    return d;
}

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
            const double d = residual_drop<WS>(r, b.res0 + (size_t)i * WS);
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
}  // namespace psst_dev


// ===========================================================================
// Context
// ===========================================================================
namespace psst_dev {
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
    int *h_flags = nullptr;   // [3n+1]: active flags [n], active-tile count, conv [n], frozen_buf [n]
    double *h_ring = nullptr; // [H][n][dim+2] residual ring
    double *h_drop = nullptr; // [n]
    int64_t iter_host = 0;
    std::vector<char> w_dirty;  // instances whose conservative state must be formed from their primitives
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

}  // namespace psst_dev

struct psst_ctx {
    int dim = 2, N = 0, G = 0, F = 0, NT = 0, nnz_visit = 0, nnz_list = 0;
    int device = 0;  // HIP device the context was created on
    psst_settings s{};
    std::vector<psst_bc> bc_host;
    psst_dev::Phys ph{};
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    psst_dev::Geom geom{};
    psst_dev::BCDev *bc = nullptr;
    // state (double or float)
    void *q[2] = {nullptr, nullptr};
    void *W0 = nullptr, *grad = nullptr, *FA = nullptr;
    int cur = 0;
    bool have_state = false;
    psst_dev::RunScratch *run = nullptr;  // run control of psst_run
    // reference-ordered time step schedule
    int *level_cells = nullptr, *level_off_dev = nullptr;
    std::vector<int> level_off;
    int max_level = 0;   // cells in the widest level
    psst_dev::OrderedTables ot{};    // tables of the ordered sweeps (ot.cell == nullptr: not built)
    psst_dev::ChunkTables ct{};      // chunked sweep (ct.n_chunks == 0: not built; ot then in level order)
    int chunk_ok[2] = {-1, -1};      // the chunks fit one cooperative launch, FP32 / FP64 (-1: not checked)
    // residual normalisation for the tolerance test
    double res0[5] = {0, 0, 0, 0, 0};
    bool have_res0 = false;
    // device numbering: d2c[d] = canonical index of device cell d (NT entries)
    std::vector<int32_t> d2c;
    // tiled kernels: the tile plan is uploaded (td); fused: the stage runs as
    // one tiled kernel (every time step mode but the reference-ordered one,
    // which uses the tiles for its gradients and its residual pass)
    bool tiled = false;
    bool fused = false;
    psst_dev::TileDev td{};
    int n_tiles = 0, tile_threads = 256, tile_cells = 0;
    std::vector<int32_t> tile_cell_begin;  // [n_tiles+1] first device cell of each tile
    int max_slots = 0, max_faces = 0;
    size_t shm_stage = 0, shm_grad = 0;
    // batched launches: per-launch work list (list == nullptr when unbatched)
    psst_dev::BatchDev bd{};
    int persist_blocks = 0;             // grid of a persistent batched launch
    std::unique_ptr<struct psst_dev::BatchState> batch;
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

namespace psst_dev {
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what) {
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

inline hipEvent_t get_event(psst_ctx *c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

inline void flush_timing(psst_ctx *c) {
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
    // largest size already allowed per kernel, in a fixed table (no heap:
    // runs of psst_run_noalloc launch through here); once the table is full
    // the attribute is simply set again
    struct Entry {
        const void *k;
        size_t bytes;
    };
    static Entry done[256];
    static int used = 0;
    if (bytes <= 64 * 1024) return;
    for (int i = 0; i < used; ++i)
        if (done[i].k == (const void *)kernel && done[i].bytes >= bytes) return;
    hip_check(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
              "hipFuncSetAttribute");
    for (int i = 0; i < used; ++i)
        if (done[i].k == (const void *)kernel) {
            done[i].bytes = bytes;
            return;
        }
    if (used < 256) done[used++] = Entry{(const void *)kernel, bytes};
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
inline int batch_grid_tiles(int count, int n_tiles) {
    int cap = 64;
    while (cap < count && cap < n_tiles) cap <<= 1;
    return std::min(cap, n_tiles);
}


inline size_t stage_shm(const psst_ctx *c, size_t sb) {
    const int NQ = c->dim + 3, NV = c->dim + 2;
    return ((size_t)NQ * c->max_slots + (size_t)NV * c->max_faces) * sb + (size_t)c->max_slots * sizeof(int);
}
inline size_t grad_shm(const psst_ctx *c, size_t sb) { return (size_t)(c->dim + 2) * c->max_slots * sb; }

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
    if (c->tiled) {
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

// Workgroups of `threads` threads that one cooperative launch of the chunked
// ordered sweep holds on the current device for tables of width W (0: no
// cooperative launches).
template <int DIM, typename T>
int chunk_sweep_capacity(int W, int threads) {
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev), "hipDeviceGetAttribute");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    if (W <= 4)
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_update_chunks<DIM, T, 4>, threads, 0),
                  "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    else
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_update_chunks<DIM, T, kOrderedW>, threads, 0),
                  "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    return coop ? cus * per_cu : 0;
}

// Whether one cooperative launch holds a workgroup per chunk of the chunked
// ordered sweep (every chunk must be resident for its waits to end).
template <int DIM, typename T>
bool chunk_sweep_fits(psst_ctx *c) {
    int &ok = c->chunk_ok[sizeof(T) == 8 ? 1 : 0];
    if (ok < 0)
        ok = (c->ct.n_chunks > 0 && c->ct.n_chunks <= chunk_sweep_capacity<DIM, T>(c->ot.W, c->ct.threads)) ? 1 : 0;
    return ok == 1;
}

// The reference-ordered update of one stage from the tables (time step mode
// 1, dt_ordering 0): the residuals in parallel, then the dependent chain of
// time steps and updates.  Large meshes: one workgroup per chunk walking its
// levels, each waiting only for the chunks that own earlier neighbours of
// its next level.  Small meshes: one workgroup walking every level, with a
// workgroup barrier per level (the reference's own 19,200-cell case).
// Called where ordered_tables_usable holds; `residual_done`: the residuals
// and the progress words are already written (k_stage_residual).
template <int DIM, typename T, int NW>
void run_ordered_tables(psst_ctx *c, const State<T> &st, T alpha, int last, unsigned long long *res_row,
                        bool residual_done) {
    const int nl = (int)c->level_off.size() - 1;
    if (c->ct.n_chunks > 0) {
        ChunkTables ct = c->ct;  // a wait that gives up is reported in the run's control word 3
        if (c->geom.stop) ct.timed_out = const_cast<int *>(c->geom.stop) + 3;
        if (!residual_done)
            launch(c, "ordered_residual", k_ordered_residual<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st,
                   c->ot, res_row, ct.progress, ct.n_chunks);
        launch_coop(c, "update_chunks", k_update_chunks<DIM, T, NW>, dim3(ct.n_chunks), dim3(ct.threads), c->geom,
                    st, c->ph, alpha, last, c->ot, ct);
        return;
    }
    if (!residual_done)
        launch(c, "ordered_residual", k_ordered_residual<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom, st,
               c->ot, res_row, (int *)nullptr, 0);
    if (c->max_level <= 256)
        launch(c, "update_levels_wg", k_update_levels_wg_prefetch<DIM, T, NW>, dim3(1), dim3(256), c->geom, st, c->ph,
               alpha, last, c->ot, (const int *)c->level_off_dev, nl);
    else
        launch(c, "update_levels_wg", k_update_levels_wg<DIM, T, NW>, dim3(1), dim3(1024), c->geom, st, c->ph, alpha,
               last, c->ot, (const int *)c->level_off_dev, nl);
}

// Whether the reference-ordered update (time step mode 1, dt_ordering 0)
// runs on the ordered sweep's tables: unbatched, tables built, and the
// chunks fit one cooperative launch, or (no chunks) the mesh is small or
// its levels narrow enough for one workgroup.  One workgroup sweeping every
// level on a large mesh runs on one CU: 938 ms per iteration on the
// 4.9M-cell wedge (about 5,000 levels of up to 1,280 cells), against 325 ms
// with one multi-workgroup launch per level, the fallback then.
template <int DIM, typename T>
bool ordered_tables_usable(psst_ctx *c) {
    if (c->s.time_mode != PSST_DT_LOCAL1 || c->s.dt_ordering != 0 || c->batch || !c->ot.cell) return false;
    if (c->ct.n_chunks > 0) return chunk_sweep_fits<DIM, T>(c);
    return c->max_level <= 256 || c->N <= 65536;
}

template <int DIM, typename T, int FLUX, int ORDER, int LIM>
void run_stage_residual_k(psst_ctx *c, unsigned long long *res_row) {
    State<T> st = state<T>(c);
    const size_t shm = stage_shm(c, sizeof(T));
    const int np = c->ct.n_chunks;
    if (c->tile_threads == 512) {
        auto kern = k_stage_residual<DIM, T, FLUX, ORDER, LIM, 512>;
        launch_shm(c, "stage_residual", kern, tile_grid(c, kern, 512, shm), dim3(512), shm, c->geom, c->td, st, c->ph,
                   c->ot, res_row, c->ct.progress, np);
    } else {
        auto kern = k_stage_residual<DIM, T, FLUX, ORDER, LIM, 256>;
        launch_shm(c, "stage_residual", kern, tile_grid(c, kern, 256, shm), dim3(256), shm, c->geom, c->td, st, c->ph,
                   c->ot, res_row, c->ct.progress, np);
    }
}

template <int DIM, typename T, int FLUX, int ORDER>
void run_stage_residual_lim(psst_ctx *c, unsigned long long *res_row) {
    if (ORDER == 1) return run_stage_residual_k<DIM, T, FLUX, 1, PSST_LIMITER_VAN_LEER>(c, res_row);
    switch (c->s.limiter) {
    case PSST_LIMITER_VAN_ALBADA:
        return run_stage_residual_k<DIM, T, FLUX, 2, PSST_LIMITER_VAN_ALBADA>(c, res_row);
    case PSST_LIMITER_MINMOD:
        return run_stage_residual_k<DIM, T, FLUX, 2, PSST_LIMITER_MINMOD>(c, res_row);
    case PSST_LIMITER_VENKATAKRISHNAN:
        return run_stage_residual_k<DIM, T, FLUX, 2, PSST_LIMITER_VENKATAKRISHNAN>(c, res_row);
    default:
        return run_stage_residual_k<DIM, T, FLUX, 2, PSST_LIMITER_VAN_LEER>(c, res_row);
    }
}

// The residual pass of the ordered sweep on the tiles (k_stage_residual).
template <int DIM, typename T>
void run_stage_residual(psst_ctx *c, unsigned long long *res_row) {
    const bool o2 = c->s.order == 2;
    switch (c->s.flux) {
    case PSST_FLUX_AUSM:
        return o2 ? run_stage_residual_lim<DIM, T, PSST_FLUX_AUSM, 2>(c, res_row)
                  : run_stage_residual_lim<DIM, T, PSST_FLUX_AUSM, 1>(c, res_row);
    case PSST_FLUX_HLL:
        return o2 ? run_stage_residual_lim<DIM, T, PSST_FLUX_HLL, 2>(c, res_row)
                  : run_stage_residual_lim<DIM, T, PSST_FLUX_HLL, 1>(c, res_row);
    default:
        return o2 ? run_stage_residual_lim<DIM, T, PSST_FLUX_HLLC, 2>(c, res_row)
                  : run_stage_residual_lim<DIM, T, PSST_FLUX_HLLC, 1>(c, res_row);
    }
}

// `residual_done`: the ordered tables' residuals are written
// (run_stage_residual ran instead of run_flux).
template <int DIM, typename T>
void run_update(psst_ctx *c, int k, unsigned long long *res_row, bool residual_done) {
    State<T> st = state<T>(c);
    const T alpha = T(c->s.rk_coeff[k]);
    const int first = (k == 0), last = (k == c->s.rk_levels - 1);
    if (c->s.time_mode == PSST_DT_LOCAL1 && c->s.dt_ordering == 0) {
        const int nl = (int)c->level_off.size() - 1;
        if (ordered_tables_usable<DIM, T>(c)) {
            if (c->ot.W <= 4) run_ordered_tables<DIM, T, 4>(c, st, alpha, last, res_row, residual_done);
            else run_ordered_tables<DIM, T, kOrderedW>(c, st, alpha, last, res_row, residual_done);
            return;
        }
        if (c->max_level <= 256 || c->N <= 65536) {
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
    // the reference-ordered sweep takes its residuals from the tiled pass
    // where the tiles exist: on the 4.9M-cell wedge (FP64) one launch of
    // 0.35 ms per stage where k_flux and k_ordered_residual took 0.30 + 0.26
    const bool tiled_residual = c->tiled && ordered_tables_usable<DIM, T>(c);
    for (int k = 0; k < c->s.rk_levels; ++k) {
        if (tiled_residual) run_stage_residual<DIM, T>(c, res_row);
        else run_flux<DIM, T>(c);
        run_update<DIM, T>(c, k, res_row, tiled_residual);
        run_bc<DIM, T>(c);
        c->cur ^= 1;
    }
}

template <int DIM, typename T>
void run_set_w(psst_ctx *c, int c_begin, int c_end) {
    launch(c, "set_w", k_set_w<DIM, T>, dim3(blocks(c_end - c_begin)), dim3(kBlock), c->geom, (const T *)c->q[c->cur],
           (T *)c->W0, c->ph, c_begin, c_end);
}

// One convective pass with the gradients last computed, gathered per cell
// into out [(dim+2)][N] (device order; psst_debug_residual).
template <int DIM, typename T>
void run_debug_residual(psst_ctx *c, double *out) {
    run_flux<DIM, T>(c);
    launch(c, "gather_residual", k_gather_residual<DIM, T>, dim3(blocks(c->N)), dim3(kBlock), c->geom,
           (const T *)c->FA, out);
}

}  // namespace psst_dev

// The launch templates that the C-ABI calls, per (dimension, state type):
// explicitly instantiated in psst_inst.hip (EXT empty), declared extern in
// psst_device.hip (EXT extern).
#define PSST_DIM_T_ENTRY_POINTS(EXT, DIM, T)                                                  \
    EXT template void psst_dev::run_iteration<DIM, T>(psst_ctx *, unsigned long long *);     \
    EXT template void psst_dev::run_set_w<DIM, T>(psst_ctx *, int, int);                      \
    EXT template void psst_dev::run_grad<DIM, T>(psst_ctx *);                                 \
    EXT template void psst_dev::run_debug_residual<DIM, T>(psst_ctx *, double *);         \
    EXT template int psst_dev::chunk_sweep_capacity<DIM, T>(int, int);
