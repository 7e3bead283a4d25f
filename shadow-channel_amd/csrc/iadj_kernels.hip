// Incremental adjoint stage: one Runge-Kutta stage k of the second-order
// (incremental) adjoint over every mode, increadjproject0
// (channel/increadjproject0.c:14-357) for (kx, kz) = (0, 0) and
// increadjproject (channel/increadjproject.c:10-645) for the others, as the
// driver loop calls them (shan/increadjoint_test7.c:673-685):
//   - right-hand side: at k = 0 the explicit operator on IAC plus the tangent
//     coupling re dt (a0 + b0) V MIC[count] (V the viscous part); the
//     Galerkin projection of the stage's products (IU: H slots, S_x, S_z as
//     the adjoint's); the forcing; the wall-lift terms of Uxb, Uzb, their
//     values Uxbt, Uzbt at the previous stage and the adjoint walls AUxb,
//     AUzb;
//   - the implicit solve (stride-2 band LU without pivoting, per parity);
//   - the Gauss-point values IAU with the wall lifts and the wall values
//     IAUxb, IAUzb;
//   - for k < 2 the next stage's explicit part from the products LIU and
//     MIC[count - 1].
// Layouts as the reference's arrays (complex {re, im}, kx fastest):
//   IAC [Nz][2][dimR][Mx], IU, LIU, IAU [Nz][5][qpts][Mx],
//   walls [6][Nz][Mx] (Uxb, Uzb, Uxbt, Uzbt, AUxb, AUzb), IAUb [2][Nz][Mx],
//   forcing [2 stages][2 sides: alpha, beta][Nz][Mx][dimR].
//
// The Galerkin projections of IU and LIU are the adjoint's (the same
// operator on other products) and run on the adjoint path's FP64 matrix-core
// kernels in one launch; so does the y-transform to the Gauss points, on the
// solved coefficients with the omega_y coefficients beyond dimQ cleared (the
// reference's sums stop there), the wall lifts added after it by a streaming
// pass.  What is particular to this stage (right-hand sides, lifts, tangent
// coupling, band solves, wall values, next stage's explicit part) is one
// kernel over the modes: IA_XG neighbouring kx columns of one kz row per
// block, one band row per thread; the implicit operators are the adjoint's,
// whose LU factors the setup stores, so each parity chain (one per thread)
// only substitutes.
#include "sc_internal.h"

#include <algorithm>

namespace sc {

namespace {

#ifndef SC_IA_XG
#define SC_IA_XG 4
#endif
constexpr int IA_XG = SC_IA_XG;   // kx columns per block
constexpr int IA_THREADS = 256;   // rows: alpha on threads 0..127, beta on 128..255

// adjoint RK3/CN coefficients (channel/increadjproject.c:34-37)
__constant__ double kA[3] = {1. / 3., -1. / 2, 1. / 3.};
__constant__ double kB[3] = {2. / 3, 1. / 6, 0.};
__constant__ double kC[3] = {1., 1. / 3., 1. / 2.};
__constant__ double kD[3] = {0., -2. / 3, -1. / 6};
// Galerkin projections of the wall lifts (channel/increadjproject.c:40-46):
// LG[i] = 2 (R_i, (1-y)/2), LX[i] = (Q_i, lv), LR[i] = 2 (Q_i, lv''),
// lv = (1-y)^2 (1+y)/4; zero beyond these entries
__constant__ double kLG[4] = {4. / 3., -4. / 15., -4. / 15., 4. / 35.};
__constant__ double kLR[8] = {-16. / 15., 16. / 35., 32. / 105., -32. / 105.,
                              -16. / 315., 16. / 231., 0., 0.};
__constant__ double kLX[8] = {8. / 35., -8. / 315., -8. / 105., 8. / 385.,
                              8. / 385., -8. / 1001., -8. / 3003., 8. / 6435.};

struct IadjArgs {
    const cplx *IU, *LIU, *W, *F, *mic0, *mic1;
    // projections of IU (Fa, Fb) and LIU (FLa, FLb): [Nz][dimQ|dimR][Mx]
    const cplx *Fa, *Fb, *FLa, *FLb;
    cplx *IAC, *IAU, *IAUb;
    cplx *Ctr;             // solved coefficients for the y-transform
    const double *Kx, *Kz, *Qs, *Qps, *Qpps, *Rs, *Rps, *Rp0;
    const double *tRw, *tRpw;   // [q][dimR]: Rw, Rpw transposed
    // LU factors of the stage's implicit operators (the adjoint's): the solve
    // images of the setup (alpha La, beta Lb: [block][parity][chain row][W][sbt]
    // over columns z Mx + x, rows z <= Nz/2 only when mirror) and the mean's [dimR][5]
    const double *La, *Lb, *Lm;
    int sbt, mirror;
    const double *lift;    // [4][q]: lv, lv', lv'', lg
    const double *mlift;   // [dimR]: sum_j Rw[i][j] (1 - y_j)/2
    int Nz, Mx, q, dQ, dR, k;
    double re, dt;
};

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return make_double2(a.x * s, a.y * s); }

// Forward and back substitution on the rows of parity p of a stride-2 band
// system, in place on vec (row i at vec[i]), with the LU factors of the
// setup (no pivoting; per matrix row W = 2 DL + 1 entries: j < DL the
// multiplier L(i, i - 2(DL - j)), DL the reciprocal of U(i, i), j > DL
// U(i, i + 2(j - DL)), zero outside the matrix).  Chain row r (matrix row
// p + 2r) has its entry j at f0[r rs + j js].  The last DL values stay in
// registers, so one row waits only on the row before it.
template <int DL>
__device__ void chain_subst(cplx *vec, const double *f0, size_t rs, size_t js, int N, int p)
{
    const int nr = (N - p + 1) / 2;
    cplx y[DL];
#pragma unroll
    for (int m = 0; m < DL; ++m) y[m] = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int r = 0; r < nr; ++r) {
        const double *f = f0 + r * rs;
        cplx b = vec[p + 2 * r];
#pragma unroll
        for (int m = 1; m <= DL; ++m) {
            const double l = f[(DL - m) * js];
            b.x -= l * y[m - 1].x;
            b.y -= l * y[m - 1].y;
        }
#pragma unroll
        for (int m = DL - 1; m > 0; --m) y[m] = y[m - 1];
        y[0] = b;
        vec[p + 2 * r] = b;
    }
#pragma unroll
    for (int m = 0; m < DL; ++m) y[m] = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int r = nr - 1; r >= 0; --r) {
        const double *f = f0 + r * rs;
        cplx v = vec[p + 2 * r];
#pragma unroll
        for (int t = 1; t <= DL; ++t) {
            const double u = f[(DL + t) * js];
            v.x -= u * y[t - 1].x;
            v.y -= u * y[t - 1].y;
        }
        v = cscale(v, f[DL * js]);
#pragma unroll
        for (int m = DL - 1; m > 0; --m) y[m] = y[m - 1];
        y[0] = v;
        vec[p + 2 * r] = v;
    }
}

// One block row of the band operators: the Gram rows of row i in registers
// (alpha: Qs, Qps, Qpps; beta: Rs, Rps, entries (i, i + 2o), zero outside
// the matrix), combined per mode.
struct GramRow {
    double s[9], p[9], pp[9];
};

__device__ __forceinline__ void load_gram(const IadjArgs &a, bool isb, int i, int N, GramRow &g)
{
    const int dl = isb ? 2 : 4, w = 2 * dl + 1;
#pragma unroll
    for (int o = 0; o < 9; ++o) {
        g.s[o] = g.p[o] = g.pp[o] = 0.0;
        if (o < w && i + 2 * (o - dl) >= 0 && i + 2 * (o - dl) < N) {
            if (isb) {
                g.s[o] = a.Rs[i * 5 + o];
                g.p[o] = a.Rps[i * 5 + o];
            } else {
                g.s[o] = a.Qs[i * 9 + o];
                g.p[o] = a.Qps[i * 9 + o];
                g.pp[o] = a.Qpps[i * 9 + o];
            }
        }
    }
}

// entry o of row i: operator (alpha: -(k2 Qs + Qps) + cv V, beta: Rs + cv (Rps + k2 Rs))
// and the viscous part V (alpha: k2^2 Qs + 2 k2 Qps + Qpps, beta: -(Rps + k2 Rs))
__device__ __forceinline__ double g_op(const GramRow &g, bool isb, int o, double k2, double cv)
{
    return isb ? g.s[o] + cv * (g.p[o] + k2 * g.s[o])
               : -(k2 * g.s[o] + g.p[o]) + cv * (k2 * k2 * g.s[o] + 2. * k2 * g.p[o] + g.pp[o]);
}
__device__ __forceinline__ double g_visc(const GramRow &g, bool isb, int o, double k2)
{
    return isb ? -(g.p[o] + k2 * g.s[o]) : k2 * k2 * g.s[o] + 2. * k2 * g.p[o] + g.pp[o];
}

// sum_o op(i, o) v[i + 2o] and sum_o V(i, o) m[i + 2o] (v, m rows in LDS)
__device__ __forceinline__ void band_pair(const GramRow &g, bool isb, int i, int N, double k2,
                                          double cv, const cplx *v, const cplx *m, cplx &e, cplx &mm)
{
    const int dl = isb ? 2 : 4;
    e = mm = make_double2(0.0, 0.0);
#pragma unroll
    for (int o = 0; o < 9; ++o) {
        if (o > 2 * dl) break;
        const int col = i + 2 * (o - dl);
        if (col < 0 || col >= N) continue;
        const double ex = g_op(g, isb, o, k2, cv), vs = g_visc(g, isb, o, k2);
        const cplx cv2 = v[col], mv = m[col];
        e.x += ex * cv2.x; e.y += ex * cv2.y;
        mm.x += vs * mv.x; mm.y += vs * mv.y;
    }
}

// the (0, 0) mode (channel/increadjproject0.c:14-357), one block; wave-wide
// sums over the Gauss points for its projections
__device__ void iadj_mean(const IadjArgs &a, cplx *sh)
{
    const int q = a.q, dR = a.dR, Mx = a.Mx, Nz = a.Nz, k = a.k, tid = threadIdx.x;
    const int lane = tid & 63;
    const size_t plane = (size_t) q * Mx, wp = (size_t) Nz * Mx;
    const double dt = a.dt, re = a.re;
    cplx *sal = sh, *sbe = sh + dR, *smic = sh + 2 * dR;          // smic [2][dR]
    double *sp = (double *) (sh + 4 * dR);                        // [4][q]: Hx, Sz, Hz, Sx (real)
    double *sf = sp + 4 * q;                                      // [2][dR]
    const cplx ub = a.W[0], wb = a.W[wp], ubt = a.W[2 * wp], wbt = a.W[3 * wp];
    auto stage = [&](const cplx *U, const cplx *mic) {
        for (int j = tid; j < q; j += IA_THREADS) {
            sp[j] = U[plane + (size_t) j * Mx].x;
            sp[q + j] = U[4 * plane + (size_t) j * Mx].x;
            sp[2 * q + j] = U[(size_t) j * Mx].x;
            sp[3 * q + j] = U[3 * plane + (size_t) j * Mx].x;
        }
        for (int r = tid; r < dR; r += IA_THREADS) {
            smic[r] = mic[(size_t) r * Mx];
            smic[dR + r] = mic[((size_t) dR + r) * Mx];
        }
    };
    // f[r] = sum_j Rw Hx + Rpw Sz (alpha rows, threads < 128), Rw Hz - Rpw Sx
    // (beta rows, threads >= 128): real parts; Rw, Rpw transposed ([q][dimR])
    auto project = [&]() {
        const bool b2 = tid >= 128;
        const int r2 = b2 ? tid - 128 : tid;
        if (r2 >= dR) return;
        const double *p0 = sp + (b2 ? 2 * q : 0), *p1 = sp + (b2 ? 3 * q : q);
        const double sg = b2 ? -1.0 : 1.0;
        // batches of 8 Gauss points: 16 loads in flight before the sums
        double f = 0.0;
        for (int j0 = 0; j0 < q; j0 += 8) {
            double w0[8], w1[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = min(j0 + u, q - 1);
                w0[u] = a.tRw[(size_t) j * dR + r2];
                w1[u] = a.tRpw[(size_t) j * dR + r2];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < q) f += w0[u] * p0[j0 + u] + sg * w1[u] * p1[j0 + u];
        }
        sf[b2 ? dR + r2 : r2] = f;
    };
    for (int r = tid; r < dR; r += IA_THREADS) {
        sal[r] = a.IAC[(size_t) r * Mx];
        sbe[r] = a.IAC[((size_t) dR + r) * Mx];
    }
    stage(a.IU, a.mic0);
    __syncthreads();
    project();
    __syncthreads();
    const bool isb = tid >= 128;
    const int r = isb ? tid - 128 : tid;
    GramRow g;
    cplx c = make_double2(0.0, 0.0);
    if (r < dR) {
        load_gram(a, true, r, dR, g);
        const cplx *v = isb ? sbe : sal;
        c = v[r];
        if (k == 0) {
            // Rs - re a0 dt Rps on the coefficients, - re dt (a0 + b0) Rps MIC[count]
            cplx e = make_double2(0.0, 0.0), m = make_double2(0.0, 0.0);
            for (int o = 0; o < 5; ++o) {
                const int col = r + 2 * (o - 2);
                if (col < 0 || col >= dR) continue;
                const double ex = g.s[o] - re * kA[0] * dt * g.p[o];
                const cplx vv = v[col], mv = smic[(isb ? dR : 0) + col];
                e.x += ex * vv.x; e.y += ex * vv.y;
                m.x += g.p[o] * mv.x; m.y += g.p[o] * mv.y;
            }
            c = cadd(e, cscale(m, -(dt * re * (kA[0] + kB[0]))));
        }
        double f = sf[isb ? dR + r : r];
        // only the real part of the stage's forcing reaches the right-hand side
        // (channel/increadjproject0.c:130-141)
        if (a.F) f += a.F[((size_t) (isb ? 1 : 0) * Nz * Mx) * dR + r].x;
        c.x += dt * kC[k] * f;
        const cplx dw = isb ? make_double2(wbt.x - wb.x, wbt.y - wb.y)
                            : make_double2(ubt.x - ub.x, ubt.y - ub.y);
        c.x += a.mlift[r] * dw.x;
        c.y += a.mlift[r] * dw.y;
    }
    __syncthreads();
    if (r < dR) (isb ? sbe : sal)[r] = c;
    __syncthreads();
    if (tid < 4) chain_subst<2>(tid < 2 ? sal : sbe, a.Lm + (tid & 1) * 5, 10, 1, dR, tid & 1);
    __syncthreads();
    // the mean's Gauss-point values sum all dimR coefficients
    for (int i = tid; i < dR; i += IA_THREADS) {
        a.Ctr[(size_t) i * Mx] = sal[i];
        a.Ctr[((size_t) dR + i) * Mx] = sbe[i];
    }
    if (tid < 128) {
        // IAUxb, IAUzb: Rp0 . alpha - Uxb/2, Rp0 . beta - Uzb/2 (wave 0, wave 1)
        const cplx *v = tid < 64 ? sal : sbe;
        double xr = 0.0, xi = 0.0;
        for (int i = lane; i < dR; i += 64) {
            xr += a.Rp0[i] * v[i].x;
            xi += a.Rp0[i] * v[i].y;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            xr += __shfl_xor(xr, off);
            xi += __shfl_xor(xi, off);
        }
        const cplx wv = tid < 64 ? ub : wb;
        if (lane == 0) a.IAUb[tid < 64 ? 0 : wp] = make_double2(xr - wv.x / 2., xi - wv.y / 2.);
    }
    if (k == 2) {
        for (int i = tid; i < dR; i += IA_THREADS) {
            a.IAC[(size_t) i * Mx] = sal[i];
            a.IAC[((size_t) dR + i) * Mx] = sbe[i];
        }
        return;
    }
    __syncthreads();
    stage(a.LIU, a.mic1);
    __syncthreads();
    project();
    __syncthreads();
    if (r < dR) {
        const cplx *v = isb ? sbe : sal;
        cplx e = make_double2(0.0, 0.0), m = make_double2(0.0, 0.0);
        for (int o = 0; o < 5; ++o) {
            const int col = r + 2 * (o - 2);
            if (col < 0 || col >= dR) continue;
            const double ex = g.s[o] - re * kA[k + 1] * dt * g.p[o];
            const cplx vv = v[col], mv = smic[(isb ? dR : 0) + col];
            e.x += ex * vv.x; e.y += ex * vv.y;
            m.x += g.p[o] * mv.x; m.y += g.p[o] * mv.y;
        }
        cplx gf = make_double2(sf[isb ? dR + r : r], 0.0);
        if (a.F) gf = cadd(gf, a.F[((size_t) (2 + (isb ? 1 : 0)) * Nz * Mx) * dR + r]);
        const cplx c2 = cadd(cadd(e, cscale(gf, dt * kD[k + 1])),
                             cscale(m, -(dt * re * (kA[k + 1] + kB[k + 1]))));
        a.IAC[((size_t) (isb ? dR : 0) + r) * Mx] = c2;
    }
}

// Right-hand side, solve, wall values and next explicit part of the modes
// (kx, kz) != (0, 0), kz row != Nz/2: IA_XG kx columns of one kz row per
// block; the first block of the grid serves the mean mode (it runs longest).
__global__ __launch_bounds__(IA_THREADS) void k_iadj_modes(IadjArgs a)
{
    extern __shared__ double2 sh[];
    const int dQ = a.dQ, dR = a.dR, Mx = a.Mx, Nz = a.Nz, k = a.k;
    const int nxg = (Mx + IA_XG - 1) / IA_XG;
    if (blockIdx.x == 0) {
        iadj_mean(a, sh);
        return;
    }
    const int z = (blockIdx.x - 1) / nxg, x0 = ((blockIdx.x - 1) % nxg) * IA_XG;
    const int nx = min(IA_XG, Mx - x0);
    const int tid = threadIdx.x;
    const size_t wp = (size_t) Nz * Mx;
    if (z == Nz / 2) {
        // cleared coefficients (zero Gauss-point values) and wall values
        for (int e = tid; e < 2 * dR * nx; e += IA_THREADS) {
            const int t = e % nx, r = e / nx;
            a.Ctr[((size_t) z * 2 * dR + r) * Mx + x0 + t] = make_double2(0.0, 0.0);
        }
        if (tid < 2 * nx) a.IAUb[(tid / nx) * wp + (size_t) z * Mx + x0 + tid % nx] = make_double2(0.0, 0.0);
        return;
    }
    cplx *sal = sh;                          // [t][dR]
    cplx *sbe = sal + IA_XG * dR;            // [t][dR]
    cplx *smic = sbe + IA_XG * dR;           // [t][2][dR]: MIC[count], then MIC[count - 1]
    const double kz = a.Kz[z];
    auto valid = [&](int t) { return t < nx && !(z == 0 && x0 + t == 0); };
    auto stage_mic = [&](const cplx *mic) {
        for (int e = tid; e < 2 * dR * IA_XG; e += IA_THREADS) {
            const int t = e % IA_XG, r = e / IA_XG;        // r over [2][dR]
            if (t < nx) smic[t * 2 * dR + r] = mic[((size_t) z * 2 * dR + r) * Mx + x0 + t];
        }
    };
    for (int e = tid; e < dR * IA_XG; e += IA_THREADS) {
        const int i = e / IA_XG, t = e % IA_XG;
        if (t < nx) {
            const size_t o = ((size_t) (z * 2) * dR + i) * Mx + x0 + t;
            sal[t * dR + i] = a.IAC[o];
            sbe[t * dR + i] = a.IAC[o + (size_t) dR * Mx];
        }
    }
    if (k == 0) stage_mic(a.mic0);
    __syncthreads();
    // right-hand sides: thread rows i (alpha: tid < 128, beta: tid >= 128)
    const bool isb = tid >= 128;
    const int i = isb ? tid - 128 : tid;
    const int N = isb ? dR : dQ;
    const double dt = a.dt, re = a.re;
    GramRow g;
    cplx acc[IA_XG];
    if (i < N) {
        load_gram(a, isb, i, N, g);
#pragma unroll
        for (int t = 0; t < IA_XG; ++t) {
            acc[t] = make_double2(0.0, 0.0);
            if (!valid(t)) continue;
            const int x = x0 + t;
            const double kx = a.Kx[x], k2 = kx * kx + kz * kz;
            const cplx *v = (isb ? sbe : sal) + t * dR;
            cplx c = v[i];
            if (k == 0) {
                cplx e, m;
                band_pair(g, isb, i, N, k2, isb ? -re * kA[0] * dt : re * kA[0] * dt, v,
                          smic + t * 2 * dR + (isb ? dR : 0), e, m);
                c = cadd(e, cscale(m, re * dt * (kA[0] + kB[0])));
            }
            cplx gf = isb ? a.Fb[((size_t) z * dR + i) * Mx + x] : a.Fa[((size_t) z * dQ + i) * Mx + x];
            if (a.F) gf = cadd(gf, a.F[((((size_t) 0 * 2 + (isb ? 1 : 0)) * Nz + z) * Mx + x) * dR + i]);
            c = cadd(c, cscale(gf, dt * kC[k]));
            const size_t wo = (size_t) z * Mx + x;
            if (!isb && i < 8) {
                const cplx wb = a.W[wp + wo], wbt = a.W[3 * wp + wo], awb = a.W[5 * wp + wo];
                const double mm = kLR[i] / 2. - k2 * kLX[i], vv = -k2 * kLR[i] + k2 * k2 * kLX[i];
                c.x += mm * (wbt.x - wb.x) + dt * re * (kA[k] * wbt.x + kB[k] * wb.x) * vv +
                       dt * re * (kA[k] + kB[k]) * awb.x * vv;
                c.y += mm * (wbt.y - wb.y) + dt * re * (kA[k] * wbt.y + kB[k] * wb.y) * vv +
                       dt * re * (kA[k] + kB[k]) * awb.y * vv;
            }
            if (isb && i < 4) {
                const cplx ub = a.W[wo], ubt = a.W[2 * wp + wo], aub = a.W[4 * wp + wo];
                const double h = kLG[i] / 2.;
                c.x += h * (ubt.x - ub.x) + dt * re * (kA[k] * ubt.x + kB[k] * ub.x) * (-k2 * h) +
                       dt * re * (kA[k] + kB[k]) * aub.x * (-k2 * h);
                c.y += h * (ubt.y - ub.y) + dt * re * (kA[k] * ubt.y + kB[k] * ub.y) * (-k2 * h) +
                       dt * re * (kA[k] + kB[k]) * aub.y * (-k2 * h);
            }
            acc[t] = c;
        }
    }
    __syncthreads();
    if (i < N) {
#pragma unroll
        for (int t = 0; t < IA_XG; ++t)
            if (valid(t)) (isb ? sbe : sal)[t * dR + i] = acc[t];
    }
    if (k < 2) stage_mic(a.mic1);     // the rows of MIC[count] are no longer read
    __syncthreads();
    // implicit solves with the setup's factors: one parity chain per thread,
    // neighbouring modes on neighbouring lanes (neighbouring image columns)
    if (tid < 4 * IA_XG) {
        const int kind = tid / IA_XG, t = tid % IA_XG;
        if (valid(t)) {
            const int zi = (a.mirror && z > Nz / 2) ? Nz - z : z;
            const size_t col = (size_t) zi * Mx + x0 + t, blk = col / a.sbt, cc = col % a.sbt;
            const int par = kind & 1, sbt = a.sbt;
            if (kind < 2) {
                const size_t rmax = (dQ + 1) / 2;
                chain_subst<4>(sal + t * dR, a.La + ((blk * 2 + par) * rmax * 9) * sbt + cc,
                               (size_t) 9 * sbt, sbt, dQ, par);
            } else {
                const size_t rmax = (dR + 1) / 2;
                chain_subst<2>(sbe + t * dR, a.Lb + ((blk * 2 + par) * rmax * 5) * sbt + cc,
                               (size_t) 5 * sbt, sbt, dR, par);
            }
        }
    }
    __syncthreads();
    // solved coefficients for the y-transform: alpha rows beyond dimQ and
    // omega_y rows beyond dimQ cleared (channel/increadjproject.c:371-395 sum
    // the omega_y coefficients below dimQ)
    for (int e = tid; e < dR * IA_XG; e += IA_THREADS) {
        const int r = e / IA_XG, t = e % IA_XG;
        if (!valid(t)) continue;
        const size_t o = ((size_t) (z * 2) * dR + r) * Mx + x0 + t;
        a.Ctr[o] = r < dQ ? sal[t * dR + r] : make_double2(0.0, 0.0);
        a.Ctr[o + (size_t) dR * Mx] = r < dQ ? sbe[t * dR + r] : make_double2(0.0, 0.0);
    }
    // wall values IAUxb, IAUzb (channel/increadjproject.c:446-485): wave t
    // of the block sums mode t's coefficients
    for (int t = tid >> 6; t < IA_XG; t += IA_THREADS / 64) {
        if (!valid(t)) continue;
        const int lane = tid & 63, x = x0 + t;
        double br = 0.0, bi = 0.0, ar = 0.0, ai = 0.0;
        for (int r = lane; r < dR; r += 64) {
            const cplx be = sbe[t * dR + r];
            br += a.Rp0[r] * be.x;
            bi += a.Rp0[r] * be.y;
            if (r < dQ) {
                const double sg = (r & 1) ? -8.0 : 8.0;
                const cplx al = sal[t * dR + r];
                ar += sg * al.x;
                ai += sg * al.y;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            br += __shfl_xor(br, off);
            bi += __shfl_xor(bi, off);
            ar += __shfl_xor(ar, off);
            ai += __shfl_xor(ai, off);
        }
        if (lane == 0) {
            const double kx = a.Kx[x], k2 = kx * kx + kz * kz;
            const cplx ub = a.W[(size_t) z * Mx + x], wb = a.W[wp + (size_t) z * Mx + x];
            cplx xb, zb;
            xb.x = bi * kz / k2 - ub.y / 2. * kz / k2 - ai * kx / k2 + wb.y * 2. * kx / k2;
            xb.y = -br * kz / k2 + ub.x / 2. * kz / k2 + 2. * ar * kx / k2 - wb.x * 2. * kx / k2;
            zb.x = -bi * kx / k2 + ub.y / 2. * kx / k2 - ai * kz / k2 + wb.y * 2. * kz / k2;
            zb.y = br * kx / k2 - ub.x / 2. * kx / k2 + 2. * ar * kz / k2 - wb.x * 2. * kz / k2;
            a.IAUb[(size_t) z * Mx + x] = xb;
            a.IAUb[wp + (size_t) z * Mx + x] = zb;
        }
    }
    if (k == 2) {
        for (int e = tid; e < dR * IA_XG; e += IA_THREADS) {
            const int r = e / IA_XG, t = e % IA_XG;
            if (!valid(t)) continue;
            const size_t o = ((size_t) (z * 2) * dR + r) * Mx + x0 + t;
            if (r < dQ) a.IAC[o] = sal[t * dR + r];
            a.IAC[o + (size_t) dR * Mx] = sbe[t * dR + r];
        }
        return;
    }
    // the next stage's explicit part: projections of LIU, forcing k+1,
    // MIC[count-1]
    if (i < N) {
#pragma unroll
        for (int t = 0; t < IA_XG; ++t) {
            if (!valid(t)) continue;
            const int x = x0 + t;
            const double kx = a.Kx[x], k2 = kx * kx + kz * kz;
            cplx e, m;
            band_pair(g, isb, i, N, k2, isb ? -re * kA[k + 1] * dt : re * kA[k + 1] * dt,
                      (isb ? sbe : sal) + t * dR, smic + t * 2 * dR + (isb ? dR : 0), e, m);
            cplx gf = isb ? a.FLb[((size_t) z * dR + i) * Mx + x] : a.FLa[((size_t) z * dQ + i) * Mx + x];
            if (a.F) gf = cadd(gf, a.F[((((size_t) 1 * 2 + (isb ? 1 : 0)) * Nz + z) * Mx + x) * dR + i]);
            const cplx c = cadd(cadd(e, cscale(gf, dt * kD[k + 1])),
                                cscale(m, re * dt * (kA[k + 1] + kB[k + 1])));
            a.IAC[((size_t) (z * 2 + (isb ? 1 : 0)) * dR + i) * Mx + x] = c;
        }
    }
}

// The wall lifts added to the Gauss-point values of the y-transform
// (channel/increadjproject.c:315-420, channel/increadjproject0.c:220-290):
// modes: v += Uzb lv, -dv/dy -= Uzb lv', d2v/dy2 += Uzb lv'', omega_y += Uxb lg,
// d omega_y/dy -= Uxb/2, carried through u = -i(kx f + kz g)/k2,
// w = i(kx g - kz f)/k2, du/dy = i(kx q - kz p)/k2, dw/dy = i(kz q + kx p)/k2;
// the mean: u += Uxb lg, w += Uzb lg, du/dy -= Uxb/2, dw/dy -= Uzb/2.
// Row Nz/2 is zero.  One thread per (z, j, x).
__global__ __launch_bounds__(256) void k_iadj_lift(IadjArgs a)
{
    const int q = a.q, Mx = a.Mx, Nz = a.Nz;
    const size_t n = (size_t) Nz * q * Mx;
    const size_t e = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int x = (int) (e % Mx), j = (int) ((e / Mx) % q), z = (int) (e / ((size_t) q * Mx));
    const size_t plane = (size_t) q * Mx, wp = (size_t) Nz * Mx;
    cplx *o = a.IAU + (size_t) z * 5 * plane + (size_t) j * Mx + x;
    if (z == Nz / 2) {
        for (int s = 0; s < 5; ++s) o[s * plane] = make_double2(0.0, 0.0);
        return;
    }
    const cplx ub = a.W[(size_t) z * Mx + x], wb = a.W[wp + (size_t) z * Mx + x];
    const double lv = a.lift[j], lv1 = a.lift[q + j], lv2 = a.lift[2 * q + j];
    const double lg = a.lift[3 * q + j];
    if (z == 0 && x == 0) {
        o[0] = cadd(o[0], cscale(ub, lg));
        o[2 * plane] = cadd(o[2 * plane], cscale(wb, lg));
        o[3 * plane] = cadd(o[3 * plane], cscale(ub, -0.5));
        o[4 * plane] = cadd(o[4 * plane], cscale(wb, -0.5));
        return;
    }
    const double kx = a.Kx[x], kz = a.Kz[z], k2 = kx * kx + kz * kz;
    const cplx f = cscale(wb, -lv1), g = cscale(ub, lg), qq = cscale(wb, lv2), p = cscale(ub, -0.5);
    const cplx u = make_double2((kx * f.y + kz * g.y) / k2, -(kx * f.x + kz * g.x) / k2);
    const cplx w = make_double2((-kx * g.y + kz * f.y) / k2, (kx * g.x - kz * f.x) / k2);
    const cplx du = make_double2((-kx * qq.y + kz * p.y) / k2, (kx * qq.x - kz * p.x) / k2);
    const cplx dw = make_double2(-(kx * p.y + kz * qq.y) / k2, (kx * p.x + kz * qq.x) / k2);
    o[0] = cadd(o[0], u);
    o[plane] = cadd(o[plane], cscale(wb, lv));
    o[2 * plane] = cadd(o[2 * plane], w);
    o[3 * plane] = cadd(o[3 * plane], du);
    o[4 * plane] = cadd(o[4 * plane], dw);
}

// a += b, row Nz/2 of a cleared (one thread per complex value)
__global__ __launch_bounds__(256) void k_gauss_sum(cplx *a, const cplx *b, size_t n, size_t per_z,
                                                   int zhalf)
{
    const size_t e = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    a[e] = (e / per_z == (size_t) zhalf) ? make_double2(0.0, 0.0) : cadd(a[e], b[e]);
}

}  // namespace

void launch_gauss_sum(Context &c, cplx *a, const cplx *b)
{
    const size_t n = c.usize, per_z = (size_t) 5 * c.d.qpts * c.d.Mx;
    hipLaunchKernelGGL(k_gauss_sum, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, c.stream, a, b,
                       n, per_z, c.d.Nz / 2);
}

// Constants of the incremental adjoint on the device (the lift values at the
// Gauss points, the mean mode's lift projection, Rw and Rpw transposed) and
// its staging buffer, built at the first call of a context.
int iadj_prepare(Context &c)
{
    if (c.dIadj) return 0;
    const Dims &d = c.d;
    const int q = d.qpts, nQ = d.dimQ, nR = d.dimR;
    std::vector<double> h;
    for (int j = 0; j < q; ++j) {
        const double y = c.y[j];
        h.push_back((1 - y) * (1 - y) * (1 + y) * 0.25);
    }
    for (int j = 0; j < q; ++j) h.push_back((3 * c.y[j] * c.y[j] - 2 * c.y[j] - 1) * 0.25);
    for (int j = 0; j < q; ++j) h.push_back((3 * c.y[j] - 1) * 0.5);
    for (int j = 0; j < q; ++j) h.push_back((1 - c.y[j]) * 0.5);
    for (int i = 0; i < nR; ++i) {
        double s = 0.0;
        for (int j = 0; j < q; ++j) s += c.Rw[(size_t) i * q + j] * (1 - c.y[j]) * 0.5;
        h.push_back(s);
    }
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < nR; ++i) h.push_back(c.Rw[(size_t) i * q + j]);
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < nR; ++i) h.push_back(c.Rpw[(size_t) i * q + j]);
    const size_t np = (size_t) d.Nz * d.Mx;
    const size_t buf = sizeof(cplx) * (3 * c.csize + 3 * c.usize + 6 * np + 4 * np * nR + 2 * np +
                                       2 * np * (nQ + nR));
    if (hipMalloc(&c.dIadj, sizeof(double) * h.size()) != hipSuccess ||
        hipMemcpy(c.dIadj, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c.iadjbuf, buf) != hipSuccess) {
        last_error() = "incremental adjoint: device allocation failed";
        if (c.dIadj) (void) hipFree(c.dIadj);
        c.dIadj = nullptr;
        if (c.iadjbuf) (void) hipFree(c.iadjbuf);
        c.iadjbuf = nullptr;
        return -1;
    }
    return 0;
}

int launch_iadj(Context &c, int k, int count, cplx *IAC, const cplx *IU, const cplx *LIU,
                const cplx *W, const cplx *F, cplx *IAU, cplx *IAUb)
{
    const Dims &d = c.d;
    if (iadj_prepare(c)) return -1;
    const int q = d.qpts, nQ = d.dimQ, nR = d.dimR;
    const size_t np = (size_t) d.Nz * d.Mx;
    // work arrays after the staging of iadjbuf (sc_increadjproject's layout)
    cplx *work = c.iadjbuf + c.csize + 3 * c.usize + 6 * np + 4 * np * nR + 2 * np;
    cplx *Ctr = work, *Fa = Ctr + c.csize, *Fb = Fa + np * nQ, *FLa = Fb + np * nR, *FLb = FLa + np * nQ;
    IadjArgs a;
    a.IU = IU; a.LIU = LIU; a.W = W; a.F = F;
    a.mic0 = c.MIC + c.csize * (size_t) count;
    a.mic1 = count > 0 ? c.MIC + c.csize * (size_t) (count - 1) : nullptr;
    a.Fa = Fa; a.Fb = Fb; a.FLa = FLa; a.FLb = FLb;
    a.IAC = IAC; a.IAU = IAU; a.IAUb = IAUb; a.Ctr = Ctr;
    a.Kx = c.dKx; a.Kz = c.dKz;
    a.Qs = c.dQs; a.Qps = c.dQps; a.Qpps = c.dQpps; a.Rs = c.dRs; a.Rps = c.dRps;
    a.Rp0 = c.dRwall;                     // R_j'(-1) (first row of the wall table)
    a.lift = c.dIadj;
    a.mlift = a.lift + 4 * (size_t) q;
    a.tRw = a.mlift + nR;
    a.tRpw = a.tRw + (size_t) q * nR;
    a.Nz = d.Nz; a.Mx = d.Mx; a.q = q; a.dQ = nQ; a.dR = nR; a.k = k;
    a.re = d.re; a.dt = d.dt;
    a.La = c.dLa[1][k]; a.Lb = c.dLb[1][k]; a.Lm = c.dLm[1][k];
    a.sbt = c.sbt; a.mirror = c.factor_mirror ? 1 : 0;
    const size_t lds_modes = sizeof(cplx) * 4 * IA_XG * (size_t) nR;
    const size_t lds_mean = sizeof(cplx) * 4 * (size_t) nR + sizeof(double) * (4 * (size_t) q + 2 * (size_t) nR);
    const size_t lds = std::max(lds_modes, lds_mean);
    (void) hipFuncSetAttribute((const void *) k_iadj_modes, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int) lds);
    const int nxg = (d.Mx + IA_XG - 1) / IA_XG;
    const size_t nu = (size_t) d.Nz * q * d.Mx;
    // algorithmic bytes of the stage: per mode IAC read and written, the
    // products of this stage (and the next) read, MIC[count] (k = 0) and
    // MIC[count - 1] (k < 2), the forcing, IAU written, the wall planes
    const double modes = (double) (d.Nz - 1) * d.Mx;
    double per = 16.0 * (2.0 * nR + nQ + nR) + 16.0 * 5.0 * q * (k < 2 ? 2 : 1) + 16.0 * 5.0 * q + 16.0 * 8.0;
    if (k == 0) per += 16.0 * 2.0 * nR;
    if (k < 2) per += 16.0 * 2.0 * nR;
    if (F) per += 16.0 * (k < 2 ? 4.0 : 2.0) * nR;
    const double bytes = modes * per;
    // the adjoint path's projections and y-transform in FP64 whatever the
    // context's precision
    const bool mixed = c.mixed;
    c.mixed = false;
    auto go = [=, &c]() {
        if (k < 2) launch_force2(c, 1, IU, Fa, Fb, LIU, FLa, FLb);
        else launch_force(c, IU, 1, Fa, Fb);
        hipLaunchKernelGGL(k_iadj_modes, dim3((unsigned) (d.Nz * nxg + 1)), dim3(IA_THREADS), lds,
                           c.stream, a);
        launch_rebuild(c, Ctr, IAU);
        hipLaunchKernelGGL(k_iadj_lift, dim3((unsigned) ((nu + 255) / 256)), dim3(256), 0, c.stream, a);
    };
    prof_begin(c, c.stream, 8, bytes);
    go();
    prof_end(c, c.stream, 8);
    c.mixed = mixed;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace sc
