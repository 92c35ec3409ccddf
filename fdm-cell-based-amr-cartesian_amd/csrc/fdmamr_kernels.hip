// fdmamr_kernels.hip -- gfx950 kernels of the per-step hot path.
//
// One RK stage is three launches on the context's stream:
//   1. k_slots_wls / k_slots_restrict: level-jump stencil values (WLS-WENO,
//      XuanGua.F90:179-356, and restriction means, XuanGua.F90:359-398) from the stage input,
//      written behind the leaf states so the flux kernel gathers all eight stencil entries
//      from one array;
//   2. k_stage: one thread per interior leaf, gathers the 5-point x and y columns, computes
//      the WCNS3 flux derivative in both directions (nnd_center_RK.F90:66-353), the
//      forward-Euler substep (nnd_center_RK.F90:15-44) and the primitive RK blend
//      (solution.F90:386-404), writes the stage output; stage 3 also reduces the residual;
//   3. k_bc_fill: ghost leaves from their partner cells (solution.F90:109-168,314-356).
//
// Layout in HBM: state as AoS double4 (rho,u,v,p) per leaf -- one 32-byte gather per stencil
// entry -- plus a separate double array for the passive temperature T. Leaves are in the
// reference's cell_NS order (Morton inside each base cell) followed by ghost and corner
// leaves, then the slot values. Stencil ids are one 32-byte int32 record per leaf ([N][8]).
//
// FP64 kernels are compiled with -ffp-contract=off and follow the reference's evaluation
// order, so they reproduce the CPU restatement bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "fdmamr_device.hpp"

namespace fdm {

// ---------------------------------------------------------------- constants
#define WCNS_D1 0x1.745d18p-4                     // REAL(4) 1./11.  nnd_center_RK.F90:331
#define WCNS_D2 0x1.d1745ep-1                     // REAL(4) 10./11.
#define WCNS_EPS 0x1.16c2p-133                    // REAL(4) 1.e-40 (subnormal) nnd_center_RK.F90:319

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------- flux math
template <typename S>
struct Pt { S U[4]; S F[4]; };

// compute_E / compute_F (nnd_center_RK.F90:171-185,293-307) for one point
template <typename S, int DIR>
__device__ __forceinline__ void cons_flux(double4 q, S gm1, Pt<S>& o) {
    S r = (S)q.x, u = (S)q.y, v = (S)q.z, p = (S)q.w;
    o.U[0] = r;
    o.U[1] = r * u;
    o.U[2] = r * v;
    o.U[3] = p / gm1 + (S)0.5 * r * (u * u + v * v);
    if (DIR == 0) {
        o.F[0] = r * u;
        o.F[1] = r * u * u + p;
        o.F[2] = r * u * v;
        o.F[3] = u * (p / gm1 + (S)0.5 * r * (u * u + v * v) + p);
    } else {
        o.F[0] = r * v;
        o.F[1] = r * u * v;
        o.F[2] = r * v * v + p;
        o.F[3] = v * (p / gm1 + (S)0.5 * r * (u * u + v * v) + p);
    }
}

// REcalc / RFcalc (nnd_center_RK.F90:114-169,236-291). R and Ri are row-major here:
// R[r][c] = R_i(r+1, c+1).
template <typename S, int DIR>
__device__ __forceinline__ void eigen(const S* U, S gama, S R[4][4], S Ri[4][4]) {
    S r = U[0], u = U[1] / U[0], v = U[2] / U[0];
    S p = (gama - (S)1.0) * r * (U[3] / r - (S)0.5 * (u * u + v * v));
    S a = sqrt(gama * p / r);
    S q2 = (S)0.5 * (u * u + v * v);
    S g1 = gama - (S)1.0, mg = (S)1.0 - gama;
    S vn = DIR == 0 ? u : v;  // normal velocity
    R[0][0] = 1; R[0][1] = 1; R[0][2] = 1; R[0][3] = 0;
    R[3][0] = a * a / g1 + q2 - vn * a;
    R[3][1] = q2;
    R[3][2] = a * a / g1 + q2 + vn * a;
    if (DIR == 0) {
        R[1][0] = u - a; R[1][1] = u; R[1][2] = u + a; R[1][3] = 0;
        R[2][0] = v; R[2][1] = v; R[2][2] = v; R[2][3] = -1;
        R[3][3] = -v;
        Ri[0][0] = (g1 * q2 + a * u) / ((S)2.0 * a * a);
        Ri[1][0] = (a * a - g1 * q2) / (a * a);
        Ri[2][0] = (g1 * q2 - a * u) / ((S)2.0 * a * a);
        Ri[3][0] = v;
        Ri[0][1] = (mg * u - a) / ((S)2.0 * a * a);
        Ri[1][1] = (g1 * u) / (a * a);
        Ri[2][1] = (mg * u + a) / ((S)2.0 * a * a);
        Ri[3][1] = 0;
        Ri[0][2] = (mg * v) / ((S)2.0 * a * a);
        Ri[1][2] = (g1 * v) / (a * a);
        Ri[2][2] = (mg * v) / ((S)2.0 * a * a);
        Ri[3][2] = -1;
    } else {
        R[1][0] = u; R[1][1] = u; R[1][2] = u; R[1][3] = 1;
        R[2][0] = v - a; R[2][1] = v; R[2][2] = v + a; R[2][3] = 0;
        R[3][3] = u;
        Ri[0][0] = (g1 * q2 + a * v) / ((S)2.0 * a * a);
        Ri[1][0] = (a * a - g1 * q2) / (a * a);
        Ri[2][0] = (g1 * q2 - a * v) / ((S)2.0 * a * a);
        Ri[3][0] = -u;
        Ri[0][1] = (mg * u) / ((S)2.0 * a * a);
        Ri[1][1] = (g1 * u) / (a * a);
        Ri[2][1] = (mg * u) / ((S)2.0 * a * a);
        Ri[3][1] = 1;
        Ri[0][2] = (mg * v - a) / ((S)2.0 * a * a);
        Ri[1][2] = (g1 * v) / (a * a);
        Ri[2][2] = (mg * v + a) / ((S)2.0 * a * a);
        Ri[3][2] = 0;
    }
    Ri[0][3] = g1 / ((S)2.0 * a * a);
    Ri[1][3] = mg / (a * a);
    Ri[2][3] = g1 / ((S)2.0 * a * a);
    Ri[3][3] = 0;
}

// WCNS_Z3 nonlinear weights for one component (nnd_center_RK.F90:326-346)
template <typename S>
__device__ __forceinline__ S wcns_component(S u1, S u2, S u3) {
    S IS1 = (-u1 + u2) * (-u1 + u2);
    S IS2 = (-u2 + u3) * (-u2 + u3);
    S t3 = u1 - (S)2. * u2 + u3;
    S IS3 = t3 * t3;
    S p1 = u2;
    S p2 = ((-u1 + (S)6.0 * u2 + (S)3.0 * u3) / (S)8.0 - (S)WCNS_D1 * p1) / (S)WCNS_D2;
    IS1 = IS1 < IS2 ? IS1 : IS2;
    IS2 = IS3 + u1 * u1 + u2 * u2 + u3 * u3 - u1 * u2 - u2 * u3 - u1 * u3;
    S tq = (S)10. * fabs(IS2 - (S)3. * IS1);
    S tau3 = tq * tq;
    S r1, r2;
    if constexpr (sizeof(S) == 8) {
        r1 = tau3 / (WCNS_EPS + IS1);
        r2 = tau3 / (WCNS_EPS + IS2);
    } else {
        // FP32: the REAL(4) epsilon is a subnormal; use a normal floor and clamp the ratios
        // so that large jumps cannot overflow to inf/inf
        r1 = fminf(tau3 / (1.0e-30f + IS1), 1.0e30f);
        r2 = fminf(tau3 / (1.0e-30f + IS2), 1.0e30f);
    }
    S a1 = (S)WCNS_D1 * ((S)1.0 + r1);
    S a2 = (S)WCNS_D2 * ((S)1.0 + r2);
    S w1 = a1 / (a1 + a2), w2 = a2 / (a1 + a2);
    return w1 * p1 + w2 * p2;
}

template <typename S>
__device__ __forceinline__ void wcns(const S* u1, const S* u2, const S* u3, const S R[4][4], S* hp) {
    S uL[4];
#pragma unroll
    for (int i = 0; i < 4; i++) uL[i] = wcns_component<S>(u1[i], u2[i], u3[i]);
#pragma unroll
    for (int r = 0; r < 4; r++) hp[r] = R[r][0] * uL[0] + R[r][1] * uL[1] + R[r][2] * uL[2] + R[r][3] * uL[3];
}

// One face of LXL_X_wcns3 / LXL_Y_wcns3 (nnd_center_RK.F90:86-106): the face between P[1] and
// P[2] of four consecutive points; eigenvectors at the arithmetic mean of the two conservative
// states, characteristic U and F, four WCNS_Z3 reconstructions, Lax-Friedrichs flux.
template <typename S, int DIR>
__device__ __forceinline__ void face_flux(const Pt<S>* P, S lam, S gama, S* flux) {
    S avg[4];
#pragma unroll
    for (int k = 0; k < 4; k++) avg[k] = (P[1].U[k] + P[2].U[k]) / (S)2.0;
    S R[4][4], Ri[4][4];
    eigen<S, DIR>(avg, gama, R, Ri);
    S w[4][4], g[4][4];  // characteristic variables at the four points
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            S s = (S)0.0, t = (S)0.0;
#pragma unroll
            for (int c = 0; c < 4; c++) { s = s + Ri[r][c] * P[j].U[c]; t = t + Ri[r][c] * P[j].F[c]; }
            w[j][r] = s; g[j][r] = t;
        }
    S wL[4], wR[4], gL[4], gR[4];
    wcns<S>(w[0], w[1], w[2], R, wL);
    wcns<S>(w[3], w[2], w[1], R, wR);
    wcns<S>(g[0], g[1], g[2], R, gL);
    wcns<S>(g[3], g[2], g[1], R, gR);
#pragma unroll
    for (int k = 0; k < 4; k++) flux[k] = (S)0.5 * (gL[k] + gR[k] - lam * (wR[k] - wL[k]));
}

// the derivative from the two face fluxes and the point fluxes of i-1, i+1
// (nnd_center_RK.F90:108-111 and :229-232, whose operator orders differ)
template <typename S, int DIR>
__device__ __forceinline__ void assemble(const S* flux0, const S* flux1, const Pt<S>& Pm, const Pt<S>& Pp,
                                         double d, double* dEF) {
#pragma unroll
    for (int n = 0; n < 4; n++) {
        double df = (double)(flux1[n] - flux0[n]);
        double de = (double)(Pp.F[n] - Pm.F[n]);
        if (DIR == 0) dEF[n] = 4.0 / (3.0 * d) * df - 1.0 / (6.0 * d) * de;
        else dEF[n] = 4.0 / 3.0 / d * df - 1.0 / 6.0 / d * de;
    }
}

// LXL_X_wcns3 / LXL_Y_wcns3 (nnd_center_RK.F90:66-112,187-234): q[0..4] = points i-2..i+2
template <typename S, int DIR>
__device__ __forceinline__ void flux_derivative(const double4* q, double d, S lam, S gama, double* dEF) {
    Pt<S> P[5];
    const S gm1 = gama - (S)1.0;
#pragma unroll
    for (int j = 0; j < 5; j++) cons_flux<S, DIR>(q[j], gm1, P[j]);
    S flux[2][4];
#pragma unroll
    for (int f = 0; f < 2; f++) face_flux<S, DIR>(P + f, lam, gama, flux[f]);
    assemble<S, DIR>(flux[0], flux[1], P[1], P[3], d, dEF);
}

// ---------------------------------------------------------------- stage kernel
// Qn/Tn alias Qout/Tout in stage 3 (in-place final stage), so they carry no __restrict__.
template <typename S>
__global__ __launch_bounds__(256) void k_stage(DevTables t, const double4* __restrict__ Qin,
                                               const double4* Qn, const double* Tn,
                                               double4* Qout, double* Tout,
                                               const DevScalars* __restrict__ sc, double alpha,
                                               double beta, int reduce, double* __restrict__ partials,
                                               double* __restrict__ lampart, int* posflag) {
    const int lb = xcd_block();
    const int i = lb * blockDim.x + threadIdx.x;
    double rmax[4] = {0, 0, 0, 0}, rsum[4] = {0, 0, 0, 0};
    double lx = 0.0, ly = 0.0;
    if (i < t.N) {
        const double gama = t.gamma, cv = t.cv;
        const double dt = sc->dt;
        int s[8];
        load_stencil(t, i, s);
        const double4 qc = Qin[i];
        const int lv = t.lvl[i];
        const double dy = t.h / (double)(1 << lv);
        const double dx = t.xf * dy;
        double dE[4], dF[4];
        {
            double4 col[5] = {Qin[s[5]], Qin[s[1]], qc, Qin[s[0]], Qin[s[4]]};
            flux_derivative<S, 0>(col, dx, (S)sc->lam_x, (S)gama, dE);
        }
        {
            double4 col[5] = {Qin[s[7]], Qin[s[3]], qc, Qin[s[2]], Qin[s[6]]};
            flux_derivative<S, 1>(col, dy, (S)sc->lam_y, (S)gama, dF);
        }
        // LXL (nnd_center_RK.F90:22-42)
        double U[4], Ut[4], fu[4];
        U[0] = qc.x;
        U[1] = qc.x * qc.y;
        U[2] = qc.x * qc.z;
        U[3] = qc.w / (gama - 1.0) + 0.5 * qc.x * (qc.y * qc.y + qc.z * qc.z);
#pragma unroll
        for (int k = 0; k < 4; k++) fu[k] = -dE[k] - dF[k];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Ut[k] = U[k] + fu[k] * dt;
            fu[k] = fu[k] * dx * dy;
            rmax[k] = fabs(fu[k]);
            rsum[k] = fu[k] * fu[k];
        }
        double r = Ut[0];
        double u = Ut[1] / Ut[0];
        double v = Ut[2] / Ut[0];
        double T = (Ut[3] - 0.5 * Ut[0] * (u * u + v * v)) / (cv * Ut[0]);
        double p = (gama - 1.0) * r * (Ut[3] / r - 0.5 * (u * u + v * v));
        // updateRK (solution.F90:398-400): U = alpha*Un + beta*Ut on (rho,u,v,T,p)
        const double4 qn = Qn[i];
        double4 o;
        o.x = alpha * qn.x + beta * r;
        o.y = alpha * qn.y + beta * u;
        o.z = alpha * qn.z + beta * v;
        o.w = alpha * qn.w + beta * p;
        Qout[i] = o;
        Tout[i] = alpha * Tn[i] + beta * T;
        // FP64 positivity of the stage result: rho <= 0 or p <= 0 sets the flag (a NaN does not: it
        // reaches the residual and the NaN stop)
        if (posflag && (o.x <= 0.0 || o.w <= 0.0)) atomicOr(posflag, 1);
        if (lampart) {
            // the next step's wave speeds on this final state (solution.F90:186-193)
            const double a = sqrt(t.gamma * o.w / o.x);
            lx = fabs(o.y) + a;
            ly = fabs(o.z) + a;
        }
    }
    if (reduce) block_partials(rmax, rsum, partials, lb, lampart, lx, ly);
}

// ---------------------------------------------------------------- face-shared FP64 stage
// The same stage with each face computed once where two leaves read it from the same points.
// k_faces computes, one face per thread, every leaf's x+1/2 and y+1/2 face (points i-1, i,
// i+1, i+2) into fxp / fyp, and the x-1/2 / y-1/2 faces that differ from the lower neighbour's
// upper face (Tables::face_lo) into fxm / fym; k_stage_shared then only reads faces: its lower
// face from the neighbour's upper one where the tables mark them identical (the neighbour's
// four stencil ids are the leaf's), else its own. Every face is evaluated with k_stage's
// arithmetic, so the results are the same bit for bit.
template <int DIR, bool LOWER>
__device__ __forceinline__ void face_item(const DevTables& t, const double4* __restrict__ Qin, int i, double lam,
                                          double4* __restrict__ out) {
    const int32_t* rec = t.stencil + (size_t)i * 8;
    // upper face: points i-1, i, i+1, i+2; lower face: i-2, i-1, i, i+1 (stencil rows
    // +1x 0, -1x 1, +1y 2, -1y 3, +2x 4, -2x 5, +2y 6, -2y 7)
    const int a = LOWER ? rec[DIR ? 7 : 5] : rec[DIR ? 3 : 1];
    const int b = LOWER ? rec[DIR ? 3 : 1] : i;
    const int c = LOWER ? i : rec[DIR ? 2 : 0];
    const int d = LOWER ? rec[DIR ? 2 : 0] : rec[DIR ? 6 : 4];
    const double gama = t.gamma, gm1 = gama - 1.0;
    Pt<double> P[4];
    cons_flux<double, DIR>(Qin[a], gm1, P[0]);
    cons_flux<double, DIR>(Qin[b], gm1, P[1]);
    cons_flux<double, DIR>(Qin[c], gm1, P[2]);
    cons_flux<double, DIR>(Qin[d], gm1, P[3]);
    double fl[4];
    face_flux<double, DIR>(P, lam, gama, fl);
    out[i] = make_double4(fl[0], fl[1], fl[2], fl[3]);
}

// Blocks: x+1/2 faces of leaves [0, N), y+1/2, then the listed x-1/2 and y-1/2 faces; each
// segment starts on a block boundary (nb0 .. nb3 blocks), so a block runs one face kind.
__global__ __launch_bounds__(256) void k_faces(DevTables t, const double4* __restrict__ Qin,
                                               const DevScalars* __restrict__ sc, double4* __restrict__ fxp,
                                               double4* __restrict__ fyp, double4* __restrict__ fxm,
                                               double4* __restrict__ fym, int nb0, int nb2) {
    const int b = blockIdx.x;
    const double lx = sc->lam_x, ly = sc->lam_y;
    if (b < 2 * nb0) {
        const int seg = b < nb0 ? 0 : 1;
        const int lb = xcd_block_n(b - seg * nb0, nb0);
        const int i = lb * 256 + (int)threadIdx.x;
        if (i >= t.N) return;
        if (seg == 0) face_item<0, false>(t, Qin, i, lx, fxp);
        else face_item<1, false>(t, Qin, i, ly, fyp);
    } else {
        const int k = (b - 2 * nb0) * 256 + (int)threadIdx.x;
        if (k < t.n_lo_x) face_item<0, true>(t, Qin, t.face_lo[k], lx, fxm);
        else if (k >= nb2 * 256 && k - nb2 * 256 < t.n_lo_y)
            face_item<1, true>(t, Qin, t.face_lo[t.n_lo_x + (k - nb2 * 256)], ly, fym);
    }
}

template <int DIR>
__device__ __forceinline__ void shared_derivative(const DevTables& t, const double4* __restrict__ Qin,
                                                  const int* s, int i, bool shared,
                                                  const double4* __restrict__ fp, const double4* __restrict__ fm,
                                                  double d, double* dEF) {
    const double gm1 = t.gamma - 1.0;
    Pt<double> Pm, Pp;  // points i-1, i+1 (their fluxes enter the derivative directly)
    cons_flux<double, DIR>(Qin[DIR ? s[3] : s[1]], gm1, Pm);
    cons_flux<double, DIR>(Qin[DIR ? s[2] : s[0]], gm1, Pp);
    const double4 q0 = shared ? fp[DIR ? s[3] : s[1]] : fm[i];
    const double4 q1 = fp[i];
    const double f0[4] = {q0.x, q0.y, q0.z, q0.w}, f1[4] = {q1.x, q1.y, q1.z, q1.w};
    assemble<double, DIR>(f0, f1, Pm, Pp, d, dEF);
}

__global__ __launch_bounds__(256) void k_stage_shared(DevTables t, const double4* __restrict__ Qin,
                                                      const double4* Qn, const double* Tn, double4* Qout,
                                                      double* Tout, const DevScalars* __restrict__ sc,
                                                      double alpha, double beta, int reduce,
                                                      double* __restrict__ partials, double* __restrict__ lampart,
                                                      int* posflag, const double4* __restrict__ fxp,
                                                      const double4* __restrict__ fyp,
                                                      const double4* __restrict__ fxm,
                                                      const double4* __restrict__ fym) {
    const int lb = xcd_block();
    const int i = lb * blockDim.x + threadIdx.x;
    double rmax[4] = {0, 0, 0, 0}, rsum[4] = {0, 0, 0, 0};
    double lx = 0.0, ly = 0.0;
    if (i < t.N) {
        const double gama = t.gamma, cv = t.cv;
        const double dt = sc->dt;
        int s[8];
        load_stencil(t, i, s);
        const double4 qc = Qin[i];
        const int lv = t.lvl[i];
        const double dy = t.h / (double)(1 << lv);
        const double dx = t.xf * dy;
        const uint8_t share = t.face_share[i];
        double dE[4], dF[4];
        shared_derivative<0>(t, Qin, s, i, share & 1, fxp, fxm, dx, dE);
        shared_derivative<1>(t, Qin, s, i, share & 2, fyp, fym, dy, dF);
        // LXL (nnd_center_RK.F90:22-42), as k_stage
        double U[4], Ut[4], fu[4];
        U[0] = qc.x;
        U[1] = qc.x * qc.y;
        U[2] = qc.x * qc.z;
        U[3] = qc.w / (gama - 1.0) + 0.5 * qc.x * (qc.y * qc.y + qc.z * qc.z);
#pragma unroll
        for (int k = 0; k < 4; k++) fu[k] = -dE[k] - dF[k];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Ut[k] = U[k] + fu[k] * dt;
            fu[k] = fu[k] * dx * dy;
            rmax[k] = fabs(fu[k]);
            rsum[k] = fu[k] * fu[k];
        }
        double r = Ut[0];
        double u = Ut[1] / Ut[0];
        double v = Ut[2] / Ut[0];
        double T = (Ut[3] - 0.5 * Ut[0] * (u * u + v * v)) / (cv * Ut[0]);
        double p = (gama - 1.0) * r * (Ut[3] / r - 0.5 * (u * u + v * v));
        const double4 qn = Qn[i];
        double4 o;
        o.x = alpha * qn.x + beta * r;
        o.y = alpha * qn.y + beta * u;
        o.z = alpha * qn.z + beta * v;
        o.w = alpha * qn.w + beta * p;
        Qout[i] = o;
        Tout[i] = alpha * Tn[i] + beta * T;
        if (posflag && (o.x <= 0.0 || o.w <= 0.0)) atomicOr(posflag, 1);
        if (lampart) {
            const double a = sqrt(t.gamma * o.w / o.x);
            lx = fabs(o.y) + a;
            ly = fabs(o.z) + a;
        }
    }
    if (reduce) block_partials(rmax, rsum, partials, lb, lampart, lx, ly);
}

// ---------------------------------------------------------------- WLS-WENO slots
__device__ __forceinline__ double isq_poly(double k1, double k2, double k3, const double* b) {
    // XuanGua.F90:342-355; b = dxb, dyb, x2, y2, x3, y3
    double term1 = k1 * k1 * b[0] * b[1];
    double term2 = 2 * k1 * k2 * b[2] / 2 * b[1] + 2 * k1 * k3 * b[0] * b[3] / 2;
    double term3 = k2 * k2 * b[4] / 3 * b[1] + 2 * k2 * k3 * b[2] / 2 * b[3] / 2 + k3 * k3 * b[0] * b[5] / 3;
    return term1 + term2 + term3;
}

__global__ __launch_bounds__(128) void k_slots_wls(DevTables t, double4* __restrict__ Q) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.S_w) return;
    // the id row as three 16-byte loads (size word, m1, ids 0..8, pattern)
    static_assert(WLSI_STRIDE == 12 && WLSI_M == 0 && WLSI_M1 == 1 && WLSI_IDS == 2 && WLSI_PAT == 11,
                  "id row layout");
    const int4* __restrict__ row = reinterpret_cast<const int4*>(t.wls_i + (size_t)s * WLSI_STRIDE);
    const int4 ra = row[0], rb = row[1], rc = row[2];
    const int ids[9] = {ra.z, ra.w, rb.x, rb.y, rb.z, rb.w, rc.x, rc.y, rc.z};
    const double* __restrict__ d = t.wls_d + (size_t)rc.w * WLS_STRIDE;
    // ghosts are filled before this launch, so the ghost mask above the size byte is not used
    const int m = ra.x & WLSI_MMASK, m1 = ra.y;
    const double xx = d[WLS_XX], yy = d[WLS_YY];
    const double* box = d + WLS_BOX;
    const double* __restrict__ qd = reinterpret_cast<const double*>(Q);
    const double lw0 = 0.98, lw1 = 0.01, lw2 = 0.01;
    // the inverse-distance weights, loaded without a branch (the record always holds nine)
    double wgt[9];
#pragma unroll
    for (int i = 0; i < 9; i++) wgt[i] = d[WLS_W + i];
    double out[4];
    // one component at a time (not unrolled): the register footprint of one fit
#pragma nounroll
    for (int k = 0; k < 4; k++) {
        // the nine point values issued together (points past m re-read point 0, every WLS
        // slot has one, and are then dropped)
        double q[9], wphi[9];
#pragma unroll
        for (int i = 0; i < 9; i++) q[i] = qd[(size_t)(i < m ? ids[i] : ids[0]) * 4 + k];
#pragma unroll
        for (int i = 0; i < 9; i++) wphi[i] = i < m ? wgt[i] * q[i] : 0.0;
        double b6[6] = {0, 0, 0, 0, 0, 0}, b3a[3] = {0, 0, 0}, b3b[3] = {0, 0, 0};
        for (int i = 0; i < 9; i++) {
            if (i >= m) break;
            const double px = d[WLS_PX + i], py = d[WLS_PY + i];
            const double P[6] = {1.0, px, py, px * py, px * px, py * py};
            for (int r = 0; r < 6; r++) b6[r] = b6[r] + P[r] * wphi[i];
            if (i < m1) for (int r = 0; r < 3; r++) b3a[r] = b3a[r] + P[r] * wphi[i];
            else for (int r = 0; r < 3; r++) b3b[r] = b3b[r] + P[r] * wphi[i];
        }
        double al[6], as1[3], as2[3];
        for (int r = 0; r < 6; r++) {
            double a = 0.0;
            for (int c = 0; c < 6; c++) a = a + d[WLS_INV6 + c * 6 + r] * b6[c];
            al[r] = a;
        }
        for (int r = 0; r < 3; r++) {
            double a = 0.0, b = 0.0;
            for (int c = 0; c < 3; c++) { a = a + d[WLS_INV3A + c * 3 + r] * b3a[c]; b = b + d[WLS_INV3B + c * 3 + r] * b3b[c]; }
            as1[r] = a; as2[r] = b;
        }
        double vl = al[0] + xx * al[1] + yy * al[2] + xx * yy * al[3] + xx * xx * al[4] + yy * yy * al[5];
        double sl = 0.0;                            // compute_smoothness order 2, XuanGua.F90:325-334
        sl = sl + isq_poly(al[1], 2 * al[3], al[4], box);
        sl = sl + isq_poly(al[2], al[4], 2 * al[5], box);
        sl = sl + (2 * al[3]) * (2 * al[3]) * box[0] * box[1];
        sl = sl + al[4] * al[4] * box[0] * box[1];
        sl = sl + (2 * al[5]) * (2 * al[5]) * box[0] * box[1];
        double vs1 = as1[0] + xx * as1[1] + yy * as1[2];
        double vs2 = as2[0] + xx * as2[1] + yy * as2[2];
        double be1 = (as1[1] * as1[1] + as1[2] * as1[2]) * box[0] * box[1];
        double be2 = (as2[1] * as2[1] + as2[2] * as2[2]) * box[0] * box[1];
        double tau = (fabs(sl - be1) + fabs(sl - be2)) / 2.0;
        tau = tau * tau;
        double a0 = lw0 * (1.0 + tau / (1e-6 + sl));
        double a1 = lw1 * (1.0 + tau / (1e-6 + be1));
        double a2 = lw2 * (1.0 + tau / (1e-6 + be2));
        double sa = a0 + a1 + a2;
        double o0 = a0 / sa, o1 = a1 / sa, o2 = a2 / sa;
        const double o = o0 * (1.0 / lw0 * vl - (lw1 / lw0 * vs1 + lw2 / lw0 * vs2)) + (o1 * vs1 + o2 * vs2);
        out[0] = k == 0 ? o : out[0];
        out[1] = k == 1 ? o : out[1];
        out[2] = k == 2 ? o : out[2];
        out[3] = k == 3 ? o : out[3];
    }
    Q[t.L + s] = make_double4(out[0], out[1], out[2], out[3]);
}

// ---------------------------------------------------------------- ghost fill + restriction slots
// items [0, G): ghost leaves from their partner (BC_cyc, solution.F90:109-168) -- a leaf copy,
// a restriction mean, or the fixed inflow state (new kind); items [G, G + S_r): restriction
// slot values. Optional FP32 shadow for the fast path.
__device__ __forceinline__ void fill_item(const DevTables& t, int item, double4* __restrict__ Q,
                                          double* __restrict__ T, const double4* __restrict__ Qi,
                                          const double* __restrict__ Ti, float4* __restrict__ Qf) {
    if (item < t.G) {
        const int r = t.bc_ref[item];
        double4 q;
        double tv = 0.0;
        if (r == REF_INFLOW) {
            q = make_double4(t.inflow[0], t.inflow[1], t.inflow[2], t.inflow[3]);
            tv = (t.inflow[3] * t.gamma * t.Ma * t.Ma) / t.inflow[0];
        } else if (r >= 0) {
            q = Q[r];
            if (T) tv = T[r];
        } else {
            q = Qi[-(r + 1)];
            if (T) tv = Ti[-(r + 1)];
        }
        Q[t.N + item] = q;
        if (T) T[t.N + item] = tv;
        if (Qf) Qf[t.N + item] = make_float4((float)q.x, (float)q.y, (float)q.z, (float)q.w);
    } else if (item < t.G + t.S_r) {
        const int s = item - t.G;
        const double4 q = Qi[t.rst_ref[s]];
        Q[t.L + t.S_w + s] = q;
        if (Qf) Qf[t.L + t.S_w + s] = make_float4((float)q.x, (float)q.y, (float)q.z, (float)q.w);
    }
}

__global__ __launch_bounds__(256) void k_fill(DevTables t, double4* __restrict__ Q, double* __restrict__ T,
                                              const double4* __restrict__ Qi, const double* __restrict__ Ti,
                                              float4* __restrict__ Qf) {
    fill_item(t, blockIdx.x * blockDim.x + threadIdx.x, Q, T, Qi, Ti, Qf);
}

// ---------------------------------------------------------------- restriction means
// interpA (XuanGua.F90:359-398) for the internal nodes the ghosts and restriction slots need,
// one tree level per launch, deepest first, keeping the reference's ((s1+s2)+s3)+s4 order.
// T may be null (the fast path does not carry it).
__global__ __launch_bounds__(256) void k_restrict_level(DevTables t, const double4* __restrict__ Q,
                                                        const double* __restrict__ T,
                                                        double4* __restrict__ Qi, double* __restrict__ Ti,
                                                        int begin, int end) {
    const int n = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= end) return;
    double4 acc = make_double4(0, 0, 0, 0);
    double ta = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = t.int_child[4 * (size_t)n + k];
        double4 v;
        double tv = 0.0;
        if (r >= 0) { v = Q[r]; if (T) tv = T[r]; }
        else { v = Qi[-(r + 1)]; if (T) tv = Ti[-(r + 1)]; }
        if (k == 0) { acc = v; ta = tv; }
        else { acc.x = acc.x + v.x; acc.y = acc.y + v.y; acc.z = acc.z + v.z; acc.w = acc.w + v.w; ta = ta + tv; }
    }
    Qi[n] = make_double4(0.25 * acc.x, 0.25 * acc.y, 0.25 * acc.z, 0.25 * acc.w);
    if (T) Ti[n] = 0.25 * ta;
}

// All levels in one launch of one block (few internal nodes): levels separated by barriers;
// with `fill` set the same block then fills the ghosts and restriction slots.
struct LevelGroups { int n; int start[MAX_TREE_DEPTH + 1]; };

__global__ __launch_bounds__(1024) void k_restrict_all(DevTables t, double4* __restrict__ Q,
                                                       double* __restrict__ T,
                                                       double4* __restrict__ Qi, double* __restrict__ Ti,
                                                       LevelGroups lg, int fill, float4* __restrict__ Qf) {
    for (int g = 0; g < lg.n; g++) {
        for (int n = lg.start[g] + threadIdx.x; n < lg.start[g + 1]; n += blockDim.x) {
            double4 acc = make_double4(0, 0, 0, 0);
            double ta = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int r = t.int_child[4 * (size_t)n + k];
                double4 v;
                double tv = 0.0;
                if (r >= 0) { v = Q[r]; if (T) tv = T[r]; }
                else { v = Qi[-(r + 1)]; if (T) tv = Ti[-(r + 1)]; }
                if (k == 0) { acc = v; ta = tv; }
                else { acc.x = acc.x + v.x; acc.y = acc.y + v.y; acc.z = acc.z + v.z; acc.w = acc.w + v.w; ta = ta + tv; }
            }
            Qi[n] = make_double4(0.25 * acc.x, 0.25 * acc.y, 0.25 * acc.z, 0.25 * acc.w);
            if (T) Ti[n] = 0.25 * ta;
        }
        __syncthreads();  // the next (shallower) level reads this one
    }
    if (fill)
        for (int item = threadIdx.x; item < t.G + t.S_r; item += blockDim.x) fill_item(t, item, Q, T, Qi, Ti, Qf);
}


// ---------------------------------------------------------------- lambda / scalars
__global__ __launch_bounds__(256) void k_wavespeed(DevTables t, const double4* __restrict__ Q,
                                                   double* __restrict__ partials) {
    // solution.F90:186-193: max |u|+c, |v|+c over interior leaves
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double lx = 0.0, ly = 0.0;
    if (i < t.N) {
        const double4 q = Q[i];
        const double a = sqrt(t.gamma * q.w / q.x);
        lx = fabs(q.y) + a;
        ly = fabs(q.z) + a;
    }
    lx = wave_max(lx);
    ly = wave_max(ly);
    __shared__ double red[4][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[wv][0] = lx; red[wv][1] = ly; }
    __syncthreads();
    if (threadIdx.x < 2) {
        double acc = red[0][threadIdx.x];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) acc = fmax(acc, red[w][threadIdx.x]);
        partials[(size_t)blockIdx.x * 2 + threadIdx.x] = acc;
    }
}

// one block finishes a step's reductions, over STEP_PARTS first-level entries
constexpr int STEP_THREADS = 1024;

// lambda running maximum and dt from CFL (solution.F90:185-201), from the reduced wave speeds
__device__ void step_begin_scalar(double lam_x, double lam_y, DevScalars* sc) {
    if (sc->status != 0) return;
    // lambda_x = max(lambda_x, ...) keeps the running maximum (solution.F90:189-190)
    if (lam_x > sc->lam_x) sc->lam_x = lam_x;
    if (lam_y > sc->lam_y) sc->lam_y = lam_y;
    double dt = sc->dt_fixed;
    if (sc->localdt) {
        const double p2 = (double)(1 << sc->l_max);
        dt = sc->cfl / (sc->lam_x / (sc->xf * sc->h / p2) + sc->lam_y / (sc->h / p2));
    }
    if (sc->tt + dt > sc->t_end) dt = sc->t_end - sc->tt;
    sc->dt = dt;
}

// Block-wide reduction (STEP_THREADS threads) of per-block partials: NV values per entry, the
// first NMAX reduced by maximum (non-negative values), the rest by sum. Every thread's loads
// are issued four entries at a time; the result is valid in thread 0.
template <int NV, int NMAX>
__device__ void reduce_entries(const double* __restrict__ part, int n, double* out) {
    __shared__ double red[STEP_THREADS / 64][NV];
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (int b0 = threadIdx.x; b0 < n; b0 += 4 * STEP_THREADS) {
        double v[4][NV];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int b = b0 + u * STEP_THREADS;
#pragma unroll
            for (int k = 0; k < NV; k++) v[u][k] = b < n ? part[(size_t)NV * b + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int k = 0; k < NV; k++) acc[k] = k < NMAX ? fmax(acc[k], v[u][k]) : acc[k] + v[u][k];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double r = k < NMAX ? wave_reduce63<true>(acc[k]) : wave_reduce63<false>(acc[k]);
        if (lane == 63) red[wv][k] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double r = red[0][k];
            for (int w = 1; w < STEP_THREADS / 64; w++) r = k < NMAX ? fmax(r, red[w][k]) : r + red[w][k];
            out[k] = r;
        }
    }
}

__global__ __launch_bounds__(STEP_THREADS) void k_step_begin(const double* __restrict__ partials, int nblk,
                                                             DevScalars* sc) {
    double lam[2];
    reduce_entries<2, 2>(partials, nblk, lam);
    if (threadIdx.x == 0) step_begin_scalar(lam[0], lam[1], sc);
}

// First level of the step's reduction: STEP_PARTS blocks each fold a share of the stages'
// per-block partials (residual max and sum of squares, and the final stage's wave speeds) into
// one entry [rmax 0..3, lam_x, lam_y, rsum 0..3]. Only blocks that held leaves are read:
// with meta (the fast path's list split) the full-path blocks [0, 4 n_troubled / 256) and the
// linear blocks [lin0, lin0 + n_cleared / 256); without it, [0, nblk).
__global__ __launch_bounds__(256) void k_reduce_parts(const double* __restrict__ partials,
                                                      const double* __restrict__ lampart, int nblk,
                                                      const int* __restrict__ meta, int n_leaf, int lin0,
                                                      double* __restrict__ out) {
    int n0 = nblk, n1 = 0;
    if (meta) {
        // the full blocks of the troubled leaves, then every linear block
        n0 = (4 * (n_leaf - meta[0]) + 255) >> 8;
        n1 = (n_leaf + 255) >> 8;
    }
    const int total = n0 + n1;
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int b = e < n0 ? e : lin0 + (e - n0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] = fmax(acc[k], partials[(size_t)b * 8 + k]);
            acc[6 + k] = acc[6 + k] + partials[(size_t)b * 8 + 4 + k];
        }
        if (lampart) {
            acc[4] = fmax(acc[4], lampart[(size_t)b * 2]);
            acc[5] = fmax(acc[5], lampart[(size_t)b * 2 + 1]);
        }
    }
    __shared__ double red[4][10];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        const double r = k < 6 ? wave_reduce63<true>(acc[k]) : wave_reduce63<false>(acc[k]);
        if (lane == 63) red[wv][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        const int k = threadIdx.x;
        double r = red[0][k];
        for (int w = 1; w < 4; w++) r = k < 6 ? fmax(r, red[w][k]) : r + red[w][k];
        out[blockIdx.x * 10 + k] = r;
    }
}

// residual finalize; tt advance; NaN status (solution.F90:205-214); with prepare_next, the
// next step's begin (its wave speeds came from the final stage)
__global__ __launch_bounds__(STEP_THREADS) void k_step_end(const double* __restrict__ parts, int n_leaf,
                                                           int advance_time, int prepare_next, DevScalars* sc) {
    double v[10];
    reduce_entries<10, 6>(parts, STEP_PARTS, v);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; k++) { sc->res_max[k] = v[k]; sc->res_sum[k] = v[6 + k]; }
        if (advance_time && sc->status == 0) {
            sc->tt = sc->tt + sc->dt;
            sc->steps_done = sc->steps_done + 1;
            for (int k = 0; k < 4; k++) {
                double r = sqrt(sc->res_sum[k] / (1.0 * n_leaf));
                sc->res_rms[k] = r;
                if (r != r) sc->status = -3;
            }
            // positivity lost in one of the step's stages (it precedes the NaN it usually causes)
            if (sc->pos_bad != 0) sc->status = -4;
        }
        if (prepare_next) step_begin_scalar(v[4], v[5], sc);
    }
}

// ---------------------------------------------------------------- sensors
// rot_div raw parts (solution.F90:558-580): |vx-uy| and |ux|+|vy| from the +-1 stencil values;
// the host multiplies by dy**1.5.
__global__ __launch_bounds__(256) void k_sensors(DevTables t, const double4* __restrict__ Q,
                                                 double* __restrict__ rot, double* __restrict__ div) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.N) return;
    const int lv = t.lvl[i];
    const double dy = t.h / (double)(1 << lv);
    const double dx = t.xf * dy;
    const int32_t* rec = t.stencil + (size_t)i * 8;
    const double4 ip = Q[rec[0]], im = Q[rec[1]];
    const double4 jp = Q[rec[2]], jm = Q[rec[3]];
    const double vx = (ip.z - im.z) / (2.0 * dx), uy = (jp.y - jm.y) / (2.0 * dy);
    rot[i] = fabs(vx - uy);
    const double ux = (ip.x - im.x) / (2.0 * dx), vy = (jp.x - jm.x) / (2.0 * dy);
    div[i] = fabs(ux) + fabs(vy);
}

// positivity of rho and p (FP64), counts bad leaves
__global__ __launch_bounds__(256) void k_positivity(DevTables t, const double4* __restrict__ Q,
                                                    int* __restrict__ nbad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.N) return;
    const double4 q = Q[i];
    if (!(q.x > 0.0) || !(q.w > 0.0)) atomicAdd(nbad, 1);
}

// gather the eight stencil values of one leaf (diagnostics)
__global__ void k_stencil_values(DevTables t, const double4* __restrict__ Q, int i, double* out) {
    const int k = threadIdx.x;
    if (k < 8 && i < t.N) {
        const double4 q = Q[t.stencil[(size_t)i * 8 + k]];
        out[4 * k + 0] = q.x; out[4 * k + 1] = q.y; out[4 * k + 2] = q.z; out[4 * k + 3] = q.w;
    }
}

// ---------------------------------------------------------------- no-allocation wave speeds
// Scratch of fdmamr_max_wavespeed_noalloc: 2 doubles of wave-speed partials per 256-leaf block,
// then the two reduced maxima. It is a file-scope device array, so that entry point allocates
// nothing; every GPU the library runs on holds its own copy until the process ends, and calls
// that overlap on one GPU share it (they race). Host code uses only the address
// hipGetSymbolAddress returned for the current device, never the array's name.
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS + 8];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_DEVICES] = {};

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = static_cast<double*>(ptr);
    return e;
}

static double* noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

__global__ __launch_bounds__(STEP_THREADS) void k_reduce_lam(const double* __restrict__ partials, int nblk,
                                                             double* __restrict__ out) {
    double lam[2];
    reduce_entries<2, 2>(partials, nblk, lam);
    if (threadIdx.x == 0) { out[0] = lam[0]; out[1] = lam[1]; }
}

hipError_t launch_wavespeed_noalloc(const DevTables& t, const double4* Q, int device, double* out2,
                                    hipStream_t st) {
    double* sp = noalloc_scratch(device);
    if (!sp) return hipErrorNotInitialized;
    const int nb = (t.N + 255) / 256;
    if (nb <= 0 || nb > NOALLOC_BLOCKS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_wavespeed, dim3(nb), dim3(256), 0, st, t, Q, sp);
    hipLaunchKernelGGL(k_reduce_lam, dim3(1), dim3(STEP_THREADS), 0, st, sp, nb, sp + 2 * NOALLOC_BLOCKS);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(out2, sp + 2 * NOALLOC_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
}

// ---------------------------------------------------------------- launchers
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_slots(const DevTables& t, double4* Q, const double* T, hipStream_t st) {
    (void)T;
    if (t.S_w > 0) hipLaunchKernelGGL(k_slots_wls, dim3(nblk(t.S_w, 128)), dim3(128), 0, st, t, Q);
}

void launch_restrict_fill(const DevTables& t, const std::vector<int32_t>& level_start, double4* Q,
                          double* T, float4* Qf, double4* Qi, double* Ti, hipStream_t st) {
    const int ngroups = (int)level_start.size() - 1;
    const int n_int = ngroups > 0 ? level_start.back() : 0;
    const int items = t.G + t.S_r;
    if (n_int > 0 && n_int <= 65536 && ngroups <= MAX_TREE_DEPTH) {
        // few internal nodes: one launch, one block (which also fills when the fill is small:
        // one block is slower than a separate wide launch beyond a few thousand items)
        LevelGroups lg;
        lg.n = ngroups;
        for (int g = 0; g <= ngroups; g++) lg.start[g] = level_start[g];
        const int fill = items <= 2048 ? 1 : 0;
        hipLaunchKernelGGL(k_restrict_all, dim3(1), dim3(1024), 0, st, t, Q, T, Qi, Ti, lg, fill, Qf);
        if (fill) return;
    } else {
        for (int g = 0; g < ngroups; g++) {
            const int b = level_start[g], e = level_start[g + 1];
            if (e > b)
                hipLaunchKernelGGL(k_restrict_level, dim3(nblk(e - b, 256)), dim3(256), 0, st, t, Q, T, Qi, Ti, b, e);
        }
    }
    if (items > 0) hipLaunchKernelGGL(k_fill, dim3(nblk(items, 256)), dim3(256), 0, st, t, Q, T, Qi, Ti, Qf);
}

void launch_stage(const DevTables& t, const double4* Qin, const double4* Qn, const double* Tn,
                  double4* Qout, double* Tout, const DevScalars* sc, double alpha, double beta,
                  int reduce, double* partials, double* lampart, int* posflag, hipStream_t st) {
    if (t.N <= 0) return;
    hipLaunchKernelGGL(k_stage<double>, dim3(nblk(t.N, 256)), dim3(256), 0, st, t, Qin, Qn, Tn, Qout, Tout,
                       sc, alpha, beta, reduce, partials, lampart, posflag);
}

void launch_stage_shared(const DevTables& t, const double4* Qin, const double4* Qn, const double* Tn,
                         double4* Qout, double* Tout, const DevScalars* sc, double alpha, double beta, int reduce,
                         double* partials, double* lampart, int* posflag, double4* fxp, double4* fyp, double4* fxm,
                         double4* fym, hipStream_t st) {
    if (t.N <= 0) return;
    const int nb0 = nblk(t.N, 256), nb2 = nblk(t.n_lo_x, 256), nb3 = nblk(t.n_lo_y, 256);
    hipLaunchKernelGGL(k_faces, dim3(2 * nb0 + nb2 + nb3), dim3(256), 0, st, t, Qin, sc, fxp, fyp, fxm, fym, nb0,
                       nb2);
    hipLaunchKernelGGL(k_stage_shared, dim3(nb0), dim3(256), 0, st, t, Qin, Qn, Tn, Qout, Tout, sc, alpha, beta,
                       reduce, partials, lampart, posflag, fxp, fyp, fxm, fym);
}

void launch_wavespeed(const DevTables& t, const double4* Q, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_wavespeed, dim3(nblk(t.N, 256)), dim3(256), 0, st, t, Q, partials);
}

void launch_step_begin(const double* partials, int nb, DevScalars* sc, hipStream_t st) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(STEP_THREADS), 0, st, partials, nb, sc);
}

void launch_step_end(const double* partials, int nb, const int* meta, int lin0, int n_leaf, int advance_time,
                     DevScalars* sc, const double* lampart, double* parts, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_parts, dim3(STEP_PARTS), dim3(256), 0, st, partials, lampart, nb, meta, n_leaf,
                       lin0, parts);
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(STEP_THREADS), 0, st, parts, n_leaf, advance_time,
                       lampart != nullptr ? 1 : 0, sc);
}
void launch_sensors(const DevTables& t, const double4* Q, double* rot, double* div, hipStream_t st) {
    hipLaunchKernelGGL(k_sensors, dim3(nblk(t.N, 256)), dim3(256), 0, st, t, Q, rot, div);
}

void launch_positivity(const DevTables& t, const double4* Q, int* nbad, hipStream_t st) {
    hipLaunchKernelGGL(k_positivity, dim3(nblk(t.N, 256)), dim3(256), 0, st, t, Q, nbad);
}

void launch_stencil_values(const DevTables& t, const double4* Q, int i, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_stencil_values, dim3(1), dim3(64), 0, st, t, Q, i, out);
}

}  // namespace fdm
