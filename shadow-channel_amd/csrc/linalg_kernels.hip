// Legendre-Galerkin kernels: y-transforms, banded Helmholtz solves with the
// (kx,kz) = (0,0) mode, and small vector kernels.
//
// y-transforms (channel/initAlphaBeta2.c:96-240 rebuild, channel/project.c:
// 95-115,176-191 and channel/adjproject.c:108-134,213-234 projections) are
// real dense products (qpts x dim or dim x qpts) applied to the complex
// columns of every kz row; they run on the FP64 matrix cores, the complex
// columns split into real and imaginary parts.
//
// Band solves (channel/mvOps.c:29-144, used by channel/project.c,
// increproject.c, adjproject.c): the matrices of the implicit stage are
// time-invariant, so their LU factors (no pivoting, same elimination order as
// the reference) are computed once on the host.  Even and odd rows decouple
// (bands sit on every other diagonal), so one thread solves one parity of
// one (kz, kx, field) system.  Explicit band products are formed on the fly
// from the Gram tables.
#include "sc_internal.h"

namespace sc {

// ------------------------------------------------------------ MFMA helpers
// v_mfma_f64_16x16x4f64: D[16][16] += A[16][4] B[4][16].  Lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; item r of its result holds D[4r + l/16][l%16].
typedef double mf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mf4 mfma_f64(double a, double b, mf4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// The 16 real columns of an n-tile are eight complex columns; real column n
// is complex column 2(n%4) + n/8, part (n/4)%2.  Item r of lane l then holds
// complex column 2(l/16) + r/2, part r%2: both parts of two neighbouring
// columns sit in one lane.
__device__ __forceinline__ int ncol_x(int n) { return 2 * (n & 3) + (n >> 3); }
__device__ __forceinline__ int ncol_part(int n) { return (n >> 2) & 1; }

__device__ __forceinline__ mf4 mf4_zero()
{
    mf4 z = {0.0, 0.0, 0.0, 0.0};
    return z;
}

// ------------------------------------------------------------ rebuild
struct RebuildArgs {
    const double *frag, *Kx, *Kz;
    int Nz, Mx, qpts, dimR, dimQ, nks;
    const int *abort_flag;
};

// U[z][0..4][y][x] from C[z][0..1][j][x]; grid (nyt, Nz), 256 threads.
// D[n][y] = sum_j C[j][n] B_m[y][j] for the bases m = Q, Qp, Qpp (alpha
// columns) and R, Rp (beta columns); wave w takes the n-tiles w, w+4, ... of
// y-tile blockIdx.x.  The basis fragments of a chunk of KC_REBUILD k-steps are
// staged in LDS (double-buffered) and shared by the four waves; the
// coefficient fragments are prefetched one chunk ahead in registers.
// Epilogue: u, w, du/dy, dw/dy from v', v'', omega_y, omega_y'
// (channel/initAlphaBeta2.c:96-240); the mean column (kx, kz) = (0, 0) takes
// u = R alpha, du/dy = Rp alpha, w = R beta, dw/dy = Rp beta, v = 0
// (channel/initAlphaBeta2.c:22-68) from two more products in its wave.
__global__ __launch_bounds__(256) void k_rebuild(RebuildArgs a, const cplx *C, cplx *U)
{
    constexpr int KC = KC_REBUILD, CH = KC * 5 * 64, PT = CH / 256;
    __shared__ double bs[2][CH];
    if (*a.abort_flag) return;
    const int z = blockIdx.y;
    if (z == a.Nz / 2) return;
    const int yt = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Mx = a.Mx, q = a.qpts, nR = a.dimR, nQ = a.dimQ;
    const int nnt = (Mx + 7) / 8, nch = a.nks / KC;
    const int kq = lane >> 4, nl = lane & 15;
    const double *frag = a.frag + (size_t) yt * a.nks * 5 * 64;
    const double *Cd = (const double *) (C + (size_t) z * 2 * nR * Mx);
    const size_t ss = (size_t) q * Mx;
    cplx *Uz = U + (size_t) z * 5 * ss;
    const double kz = a.Kz[z];
    for (int nt0 = 0; nt0 < nnt; nt0 += 4) {
        const int nt = nt0 + wave;
        const bool active = nt < nnt;
        const int xcol = nt * 8 + ncol_x(nl), part = ncol_part(nl);
        const bool xok = active && xcol < Mx;
        const int limA = (z == 0 && xcol == 0) ? nR : nQ;
        const bool meanw = z == 0 && nt == 0;
        mf4 acc[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) acc[m] = mf4_zero();
        double ca[KC], cb[KC], st[PT];
        auto load_chunk = [&](int ch) {
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                const int j = (ch * KC + s) * 4 + kq;
                ca[s] = (xok && j < limA) ? Cd[((size_t) j * Mx + xcol) * 2 + part] : 0.0;
                cb[s] = (xok && j < nR) ? Cd[((size_t) (nR + j) * Mx + xcol) * 2 + part] : 0.0;
            }
            const double *src = frag + (size_t) ch * CH;
#pragma unroll
            for (int i = 0; i < PT; ++i) st[i] = src[threadIdx.x + 256 * i];
        };
        load_chunk(0);
#pragma unroll
        for (int i = 0; i < PT; ++i) bs[0][threadIdx.x + 256 * i] = st[i];
        __syncthreads();
        for (int ch = 0; ch < nch; ++ch) {
            double xa[KC], xb[KC];
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                xa[s] = ca[s];
                xb[s] = cb[s];
            }
            if (ch + 1 < nch) load_chunk(ch + 1);
            const double *b = bs[ch & 1];
            if (active) {
#pragma unroll
                for (int s = 0; s < KC; ++s) {
                    const double b0 = b[(s * 5 + 0) * 64 + lane], b1 = b[(s * 5 + 1) * 64 + lane];
                    const double b2 = b[(s * 5 + 2) * 64 + lane], b3 = b[(s * 5 + 3) * 64 + lane];
                    const double b4 = b[(s * 5 + 4) * 64 + lane];
                    acc[0] = mfma_f64(xa[s], b0, acc[0]);
                    acc[1] = mfma_f64(xa[s], b1, acc[1]);
                    acc[2] = mfma_f64(xa[s], b2, acc[2]);
                    acc[3] = mfma_f64(xb[s], b3, acc[3]);
                    acc[4] = mfma_f64(xb[s], b4, acc[4]);
                    if (meanw) {
                        acc[5] = mfma_f64(xa[s], b3, acc[5]);
                        acc[6] = mfma_f64(xa[s], b4, acc[6]);
                    }
                }
            }
            if (ch + 1 < nch) {
#pragma unroll
                for (int i = 0; i < PT; ++i) bs[(ch + 1) & 1][threadIdx.x + 256 * i] = st[i];
            }
            __syncthreads();
        }
        const int y = yt * 16 + nl;
        if (!active || y >= q) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int x = nt * 8 + 2 * kq + h;
            if (x >= Mx) continue;
            const size_t o = (size_t) y * Mx + x;
            if (z == 0 && x == 0) {
                Uz[o + 0 * ss] = make_double2(acc[5][2 * h], acc[5][2 * h + 1]);
                Uz[o + 1 * ss] = make_double2(0.0, 0.0);
                Uz[o + 2 * ss] = make_double2(acc[3][2 * h], acc[3][2 * h + 1]);
                Uz[o + 3 * ss] = make_double2(acc[6][2 * h], acc[6][2 * h + 1]);
                Uz[o + 4 * ss] = make_double2(acc[4][2 * h], acc[4][2 * h + 1]);
                continue;
            }
            const double kx = a.Kx[x], k2 = kz * kz + kx * kx;
            const double2 v = make_double2(acc[0][2 * h], acc[0][2 * h + 1]);
            const double2 f = make_double2(-acc[1][2 * h], -acc[1][2 * h + 1]);
            const double2 q2 = make_double2(acc[2][2 * h], acc[2][2 * h + 1]);
            const double2 g = make_double2(acc[3][2 * h], acc[3][2 * h + 1]);
            const double2 gp = make_double2(acc[4][2 * h], acc[4][2 * h + 1]);
            Uz[o + 1 * ss] = v;
            Uz[o + 0 * ss] = make_double2((kx * f.y + kz * g.y) / k2, -(kx * f.x + kz * g.x) / k2);
            Uz[o + 2 * ss] = make_double2((-kx * g.y + kz * f.y) / k2, (kx * g.x - kz * f.x) / k2);
            Uz[o + 3 * ss] = make_double2((-kx * q2.y + kz * gp.y) / k2, (kx * q2.x - kz * gp.x) / k2);
            Uz[o + 4 * ss] = make_double2(-(kx * gp.y + kz * q2.y) / k2, (kx * gp.x + kz * q2.x) / k2);
        }
    }
}

void launch_rebuild(Context &c, const cplx *Csrc, cplx *Udst)
{
    const Dims &d = c.d;
    RebuildArgs a{c.dFragR, c.dKx, c.dKz, d.Nz, d.Mx, d.qpts, d.dimR, d.dimQ, c.nksR, c.abort_flag};
    dim3 grid(c.nyt, d.Nz);
    double bytes = 16.0 * (double) (d.Nz - 1) * d.Mx * (2.0 * d.dimR + 5.0 * d.qpts);
    prof_begin(c, c.stream, 4, bytes);
    hipLaunchKernelGGL(k_rebuild, grid, dim3(256), 0, c.stream, a, Csrc, Udst);
    prof_end(c, c.stream, 4);
}

// ------------------------------------------------------------ projection
struct ForceArgs {
    const double *frag, *Kx, *Kz;
    int Nz, Mx, qpts, dimR, dimQ, nks, adj;
    const int *abort_flag;
};

// Fa[z][i][x] (i < dimQ), Fb[z][i][x] (i < dimR); grid (nit, Nz), 256 threads.
// D[n][i] = sum_y A_m[y][n] W_m[i][y] (channel/project.c:95-115,176-191,
// channel/adjproject.c:108-134,213-234) with the pointwise combinations
//   fa = Qpw P - k2 Qw Hy (+ Qppw Z),     P  = i kx Hx + i kz Hz
//   fb = Rw P2 (+ Rpw Z2),                P2 = i kz Hx - i kx Hz
//   (adjoint) Z = -i kz Sx + i kx Sz,     Z2 = i kz Sz + i kx Sx
// summed into two accumulators.  Each lane forms its A entries from the parts
// of Hx, Hz, Hy, Sx, Sz it needs.  Mean column (channel/project0.c:110-146,
// channel/adjproject0.c:98-140): Fb(0,0) = (Rw Re Hx + Rpw Re Sz,
// Rw Re Hz - Rpw Re Sx).  Wave w takes the n-tiles w, w+4, ... of i-tile
// blockIdx.x; the weighted-basis fragments of a chunk of KC_FORCE k-steps are
// staged in LDS (double-buffered), the field values prefetched a chunk ahead.
__global__ __launch_bounds__(256) void k_force(ForceArgs a, const cplx *U, cplx *Fa, cplx *Fb)
{
    constexpr int KC = KC_FORCE, CH = KC * 5 * 64, PT = CH / 256;
    static_assert(CH % 256 == 0, "chunk must split over the block");
    __shared__ double bs[2][CH];
    if (*a.abort_flag) return;
    const int z = blockIdx.y;
    if (z == a.Nz / 2) return;
    const int it = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Mx = a.Mx, q = a.qpts, nR = a.dimR, nQ = a.dimQ, adj = a.adj;
    const int nnt = (Mx + 7) / 8, nch = a.nks / KC;
    const int kq = lane >> 4, nl = lane & 15;
    const double *frag = a.frag + (size_t) it * a.nks * 5 * 64;
    const size_t ss = (size_t) q * Mx;
    const double *Ud = (const double *) (U + (size_t) z * 5 * ss);
    const double kz = a.Kz[z];
    for (int nt0 = 0; nt0 < nnt; nt0 += 4) {
        const int nt = nt0 + wave;
        const bool active = nt < nnt;
        const int xcol = nt * 8 + ncol_x(nl), part = ncol_part(nl);
        const bool xok = active && xcol < Mx;
        const bool mean = z == 0 && xcol == 0;
        const double kx = xok ? a.Kx[xcol] : 0.0, k2 = kz * kz + kx * kx;
        const double sg = part ? 1.0 : -1.0;
        const int other = mean ? 0 : 1 - part;
        mf4 accA = mf4_zero(), accB = mf4_zero();
        double rhx[KC], rhz[KC], rhy[KC], rsx[KC], rsz[KC], st[PT];
        auto load_chunk = [&](int ch) {
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                const int y = (ch * KC + s) * 4 + kq;
                const bool ok = xok && y < q;
                const size_t e = ((size_t) y * Mx + xcol) * 2;
                rhz[s] = ok ? Ud[e + other] : 0.0;
                rhx[s] = ok ? Ud[e + 2 * ss + other] : 0.0;
                rhy[s] = ok ? Ud[e + 4 * ss + part] : 0.0;
                rsx[s] = (ok && adj) ? Ud[e + 6 * ss + other] : 0.0;
                rsz[s] = (ok && adj) ? Ud[e + 8 * ss + other] : 0.0;
            }
            const double *src = frag + (size_t) ch * CH;
#pragma unroll
            for (int i = 0; i < PT; ++i) st[i] = src[threadIdx.x + 256 * i];
        };
        load_chunk(0);
#pragma unroll
        for (int i = 0; i < PT; ++i) bs[0][threadIdx.x + 256 * i] = st[i];
        __syncthreads();
        for (int ch = 0; ch < nch; ++ch) {
            double vP[KC], vY[KC], vZ[KC], vP2[KC], vZ2[KC];
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                if (mean) {
                    vP[s] = 0.0;
                    vY[s] = 0.0;
                    vZ[s] = 0.0;
                    vP2[s] = part ? rhz[s] : rhx[s];
                    vZ2[s] = part ? -rsx[s] : rsz[s];
                } else {
                    vP[s] = sg * (kx * rhx[s] + kz * rhz[s]);
                    vY[s] = -k2 * rhy[s];
                    vP2[s] = sg * (kz * rhx[s] - kx * rhz[s]);
                    vZ[s] = -sg * (kz * rsx[s] - kx * rsz[s]);
                    vZ2[s] = sg * (kz * rsz[s] + kx * rsx[s]);
                }
            }
            if (ch + 1 < nch) load_chunk(ch + 1);
            const double *b = bs[ch & 1];
            if (active) {
#pragma unroll
                for (int s = 0; s < KC; ++s) {
                    accA = mfma_f64(vP[s], b[(s * 5 + 0) * 64 + lane], accA);
                    accB = mfma_f64(vP2[s], b[(s * 5 + 3) * 64 + lane], accB);
                    accA = mfma_f64(vY[s], b[(s * 5 + 1) * 64 + lane], accA);
                    if (adj) {
                        accA = mfma_f64(vZ[s], b[(s * 5 + 2) * 64 + lane], accA);
                        accB = mfma_f64(vZ2[s], b[(s * 5 + 4) * 64 + lane], accB);
                    }
                }
            }
            if (ch + 1 < nch) {
#pragma unroll
                for (int i = 0; i < PT; ++i) bs[(ch + 1) & 1][threadIdx.x + 256 * i] = st[i];
            }
            __syncthreads();
        }
        const int i = it * 16 + nl;
        if (!active || i >= nR) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int x = nt * 8 + 2 * kq + h;
            if (x >= Mx) continue;
            if (i < nQ && !(z == 0 && x == 0))
                Fa[((size_t) z * nQ + i) * Mx + x] = make_double2(accA[2 * h], accA[2 * h + 1]);
            Fb[((size_t) z * nR + i) * Mx + x] = make_double2(accB[2 * h], accB[2 * h + 1]);
        }
    }
}

void launch_force(Context &c, const cplx *Usrc, int adj, cplx *Fa, cplx *Fb)
{
    const Dims &d = c.d;
    ForceArgs a{c.dFragF, c.dKx, c.dKz, d.Nz, d.Mx, d.qpts, d.dimR, d.dimQ, c.nksF, adj,
                c.abort_flag};
    dim3 grid(c.nit, d.Nz);
    double bytes = 16.0 * (double) (d.Nz - 1) * d.Mx *
                   ((adj ? 5.0 : 3.0) * d.qpts + d.dimQ + d.dimR);
    prof_begin(c, c.stream, 5, bytes);
    hipLaunchKernelGGL(k_force, grid, dim3(256), 0, c.stream, a, Usrc, Fa, Fb);
    prof_end(c, c.stream, 5);
}

void launch_force(Context &c, const cplx *Usrc, int adj) { launch_force(c, Usrc, adj, c.Fa, c.Fb); }

// ------------------------------------------------------------ band solves
// Stage solve of every (kz, kx) column (channel/project.c:22-151,
// channel/increproject.c:13-194, channel/adjproject.c:9-226): the right-hand
// side E C + dt c_k F (E applied only at the first stage, where C still holds
// the state), the LU solve with the stage's implicit operator, then
// C <- E' X + dt d_{k+1} F2, the next stage's carried right-hand side.  The
// even and odd rows of the stride-2 bands decouple, so one chain is one
// (column, field, parity).
//
// The LU factors are formed once at init with the pivots inverted, so a chain
// is a pure multiply-add recurrence.  A block stages the factors and vectors
// of SBT columns of one (field, parity) in LDS with all its threads, runs the
// SBT chains with one thread each (the last DL solved values in registers),
// and writes back with all its threads.  The loops over rows stay rolled: the
// kernel fits the instruction cache at every grid size.
//
// The mean mode (kz = kx = 0; channel/project0.c:22-271,
// channel/increproject0.c:14-312, channel/adjproject0.c:8-310) runs in one
// extra block: real-part-only forcing fa = Re(Hx)-projection and
// fb = Re(Hz)-projection, read as the pair Fb(0,0) = (fa, fb); the flux
// correction alpha += (target - sum_i rowsum_i Re alpha_i) add / 2.
struct SolveArgs {
    const double *Qs, *Qps, *Qpps, *Rs, *Rps, *Kx, *Kz;
    const double *La, *Lb;             // factors of this stage, [z][row][2DL+1][x]
    const double *Lm, *add, *rowsum;   // mean mode: factors [row][5], flux vector, row sums
    const cplx *Fa, *Fb;               // projection for this stage's RHS
    const cplx *Fa2, *Fb2;             // projection for the next-stage RHS
    int Nz, Mx, dimR, dimQ, k, has_next;
    double dtc, dtd;                   // dt c_k, dt d_{k+1}
    double ce0[2], cn[2];              // explicit-operator coefficients per field (stage 0, next)
    double target;                     // mean-mode flux target
    const int *abort_flag;
};

// explicit band operator row entry A(i, i+2m)
__device__ __forceinline__ double op_entry(const SolveArgs &a, int var, int i, int m, double k2, double coef)
{
    if (var == 0) {
        int e = i * 9 + 4 + m;
        double s = k2 * k2;
        return -(k2 * a.Qs[e] + a.Qps[e]) + coef * (s * a.Qs[e] + 2. * k2 * a.Qps[e] + a.Qpps[e]);
    }
    int e = i * 5 + 2 + m;
    return a.Rs[e] + coef * (a.Rps[e] + k2 * a.Rs[e]);
}

// Forward and back substitution of one chain held in LDS.  Factors of row r:
// f[r*fr + j*fj], j = DL - m for L(r, r-m), j = DL for 1/U(r, r),
// j = DL + u for U(r, r+u); entries that fall outside the matrix are zero.
// v: right-hand side in, solution out, row r at v[r*vr].  The next row's
// operands are fetched while the current row is eliminated.
template <int DL>
__device__ void chain_subst(const double *f, int fr, int fj, double2 *v, int vr, int nr)
{
    if (nr <= 0) return;
    double2 w[DL];
#pragma unroll
    for (int m = 0; m < DL; ++m) w[m] = make_double2(0.0, 0.0);
    double cf[DL], nf[DL];
    double2 nv = v[0];
#pragma unroll
    for (int m = 1; m <= DL; ++m) nf[m - 1] = f[(DL - m) * fj];
    for (int r = 0; r < nr; ++r) {
        double2 x = nv;
#pragma unroll
        for (int m = 0; m < DL; ++m) cf[m] = nf[m];
        if (r + 1 < nr) {
            nv = v[(r + 1) * vr];
#pragma unroll
            for (int m = 1; m <= DL; ++m) nf[m - 1] = f[(r + 1) * fr + (DL - m) * fj];
        }
#pragma unroll
        for (int m = DL; m >= 1; --m) {
            x.x -= cf[m - 1] * w[m - 1].x;
            x.y -= cf[m - 1] * w[m - 1].y;
        }
#pragma unroll
        for (int m = DL - 1; m > 0; --m) w[m] = w[m - 1];
        w[0] = x;
        v[r * vr] = x;
    }
#pragma unroll
    for (int m = 0; m < DL; ++m) w[m] = make_double2(0.0, 0.0);
    double crd, nrd;
    {
        const int r = nr - 1;
        nv = v[r * vr];
        nrd = f[r * fr + DL * fj];
#pragma unroll
        for (int u = 1; u <= DL; ++u) nf[u - 1] = f[r * fr + (DL + u) * fj];
    }
    for (int r = nr - 1; r >= 0; --r) {
        double2 x = nv;
        crd = nrd;
#pragma unroll
        for (int u = 0; u < DL; ++u) cf[u] = nf[u];
        if (r > 0) {
            nv = v[(r - 1) * vr];
            nrd = f[(r - 1) * fr + DL * fj];
#pragma unroll
            for (int u = 1; u <= DL; ++u) nf[u - 1] = f[(r - 1) * fr + (DL + u) * fj];
        }
#pragma unroll
        for (int u = 1; u <= DL; ++u) {
            x.x -= cf[u - 1] * w[u - 1].x;
            x.y -= cf[u - 1] * w[u - 1].y;
        }
        x.x *= crd;
        x.y *= crd;
#pragma unroll
        for (int u = DL - 1; u > 0; --u) w[u] = w[u - 1];
        w[0] = x;
        v[r * vr] = x;
    }
}

struct ColumnOf {
    int z, x;
    bool valid, chain;
};

__device__ __forceinline__ ColumnOf column_of(const SolveArgs &a, int col)
{
    ColumnOf c;
    c.valid = col < a.Nz * a.Mx;
    c.z = c.valid ? col / a.Mx : 0;
    c.x = c.valid ? col - c.z * a.Mx : 0;
    c.chain = c.valid && c.z != a.Nz / 2 && !(c.z == 0 && c.x == 0);
    return c;
}

// one (field, parity) of SBT consecutive columns; lds: [nr][W][SBT] factors,
// then [nr][SBT] right-hand side, then [nr][SBT] stage-0 copy of C
template <int DL, int SBT>
__device__ void solve_block(const SolveArgs &a, int var, int par, int col0, cplx *C, cplx *Cs,
                            double *lds)
{
    constexpr int W = 2 * DL + 1;
    const int Mx = a.Mx, nR = a.dimR, nQ = a.dimQ;
    const int N = var == 0 ? nQ : nR;
    const int nr = (N - par + 1) / 2;
    const int T = blockDim.x, tid = threadIdx.x;
    double *lf = lds;
    double2 *vec = (double2 *) (lds + (size_t) nr * W * SBT);
    double2 *raw = vec + nr * SBT;
    const double *Lg = var == 0 ? a.La : a.Lb;
    const cplx *F = var == 0 ? a.Fa : a.Fb;
    const cplx *F2 = var == 0 ? a.Fa2 : a.Fb2;
    const double ce0 = a.ce0[var], cn = a.cn[var];

    // stage the factors
    const int nfac = nr * W * SBT;
#pragma unroll 4
    for (int e = tid; e < nfac; e += T) {
        const int cc = e % SBT, q = e / SBT, r = q / W, j = q - r * W;
        const ColumnOf co = column_of(a, col0 + cc);
        lf[e] = co.chain ? Lg[(((size_t) co.z * N + par + 2 * r) * W + j) * Mx + co.x] : 0.0;
    }
    // stage the carried right-hand side (plus the projection after stage 0)
    const int nvec = nr * SBT;
#pragma unroll 2
    for (int e = tid; e < nvec; e += T) {
        const int cc = e % SBT, r = e / SBT, i = par + 2 * r;
        const ColumnOf co = column_of(a, col0 + cc);
        if (!co.chain) continue;
        double2 cv = C[(((size_t) co.z * 2 + var) * nR + i) * Mx + co.x];
        if (a.k == 0) {
            raw[e] = cv;
        } else {
            double2 f = F[((size_t) co.z * N + i) * Mx + co.x];
            cv.x += a.dtc * f.x;
            cv.y += a.dtc * f.y;
            vec[e] = cv;
        }
    }
    __syncthreads();
    if (a.k == 0) {
        for (int e = tid; e < nvec; e += T) {
            const int cc = e % SBT, r = e / SBT, i = par + 2 * r;
            const ColumnOf co = column_of(a, col0 + cc);
            if (!co.chain) continue;
            const double kx = a.Kx[co.x], kz = a.Kz[co.z], k2 = kz * kz + kx * kx;
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int m = -DL; m <= DL; ++m)
                if (r + m >= 0 && r + m < nr) {
                    double ev = op_entry(a, var, i, m, k2, ce0);
                    double2 y = raw[(r + m) * SBT + cc];
                    acc.x += ev * y.x;
                    acc.y += ev * y.y;
                }
            double2 f = F[((size_t) co.z * N + i) * Mx + co.x];
            acc.x += a.dtc * f.x;
            acc.y += a.dtc * f.y;
            vec[e] = acc;
        }
        __syncthreads();
    }
    if (tid < SBT && column_of(a, col0 + tid).chain)
        chain_subst<DL>(lf + tid, W * SBT, SBT, vec + tid, SBT, nr);
    __syncthreads();
    // write the solution and the next stage's right-hand side
    for (int e = tid; e < nvec; e += T) {
        const int cc = e % SBT, r = e / SBT, i = par + 2 * r;
        const ColumnOf co = column_of(a, col0 + cc);
        if (!co.chain) continue;
        const size_t gi = (((size_t) co.z * 2 + var) * nR + i) * Mx + co.x;
        const double2 sol = vec[e];
        Cs[gi] = sol;
        if (a.has_next) {
            const double kx = a.Kx[co.x], kz = a.Kz[co.z], k2 = kz * kz + kx * kx;
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int m = -DL; m <= DL; ++m)
                if (r + m >= 0 && r + m < nr) {
                    double ev = op_entry(a, var, i, m, k2, cn);
                    double2 y = vec[(r + m) * SBT + cc];
                    acc.x += ev * y.x;
                    acc.y += ev * y.y;
                }
            double2 f = F2[((size_t) co.z * N + i) * Mx + co.x];
            acc.x += a.dtd * f.x;
            acc.y += a.dtd * f.y;
            C[gi] = acc;
        } else {
            C[gi] = sol;
        }
    }
    // rows of v beyond dimQ are carried unchanged; the Nyquist row is zeroed
    const int nrest = (nR - par + 1) / 2;
    for (int e = tid; e < nrest * SBT; e += T) {
        const int cc = e % SBT, r = e / SBT, i = par + 2 * r;
        const ColumnOf co = column_of(a, col0 + cc);
        if (!co.valid || (co.z == 0 && co.x == 0)) continue;
        const size_t gi = (((size_t) co.z * 2 + var) * nR + i) * Mx + co.x;
        if (co.z == a.Nz / 2)
            C[gi] = make_double2(0.0, 0.0);
        else if (var == 0 && i >= nQ)
            Cs[gi] = C[gi];
    }
}

// explicit mean operator Rs + coef Rps applied to v (row i)
__device__ double2 mean_mult(const SolveArgs &a, const double2 *v, int i, double coef)
{
    const int N = a.dimR;
    double2 t = make_double2(0.0, 0.0);
    int m0 = -min(2, i / 2), m1 = min(2, (N - i + 1) / 2 - 1);
    for (int m = m0; m <= m1; ++m) {
        double c = op_entry(a, 1, i, m, 0.0, coef);
        t.x += c * v[i + 2 * m].x;
        t.y += c * v[i + 2 * m].y;
    }
    return t;
}

// the mean mode: alpha (streamwise mean) and beta (spanwise mean) columns
__device__ void solve_mean(const SolveArgs &a, cplx *C, cplx *Cs, double *lds)
{
    const int nR = a.dimR, Mx = a.Mx, T = blockDim.x, tid = threadIdx.x;
    double2 *al = (double2 *) lds, *be = al + nR, *tmp = be + nR;
    double *lm = (double *) (tmp + nR), *red = lm + 5 * nR;
    cplx *Ca = C, *Cb = C + (size_t) nR * Mx;
    for (int i = tid; i < nR; i += T) {
        al[i] = Ca[(size_t) i * Mx];
        be[i] = Cb[(size_t) i * Mx];
    }
    for (int e = tid; e < 5 * nR; e += T) lm[e] = a.Lm[e];
    __syncthreads();
    if (a.k == 0) {
        for (int i = tid; i < nR; i += T) tmp[i] = mean_mult(a, al, i, a.ce0[1]);
        __syncthreads();
        for (int i = tid; i < nR; i += T) al[i] = tmp[i];
        __syncthreads();
        for (int i = tid; i < nR; i += T) tmp[i] = mean_mult(a, be, i, a.ce0[1]);
        __syncthreads();
        for (int i = tid; i < nR; i += T) be[i] = tmp[i];
        __syncthreads();
    }
    const cplx *Fm = a.Fb, *Fm2 = a.Fb2;   // column (0, 0): (fa, fb)
    for (int i = tid; i < nR; i += T) {
        double2 f = Fm[(size_t) i * Mx];
        al[i].x += a.dtc * f.x;
        be[i].x += a.dtc * f.y;
    }
    __syncthreads();
    const int n0 = (nR + 1) / 2, n1 = nR / 2;
    if (tid == 0) chain_subst<2>(lm, 10, 1, al, 2, n0);
    if (tid == 64) chain_subst<2>(lm + 5, 10, 1, al + 1, 2, n1);
    if (tid == 128) chain_subst<2>(lm, 10, 1, be, 2, n0);
    if (tid == 192) chain_subst<2>(lm + 5, 10, 1, be + 1, 2, n1);
    __syncthreads();
    double p = 0.0;
    for (int i = tid; i < nR; i += T) p += a.rowsum[i] * al[i].x;
    red[tid] = p;
    __syncthreads();
    for (int o = T / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double flux_t = a.target - red[0];
    for (int i = tid; i < nR; i += T) {
        al[i].x = al[i].x + flux_t * a.add[i] * 0.5;
        al[i].y = al[i].y + flux_t * 0.0 * 0.5;
        Cs[(size_t) i * Mx] = al[i];
        Cs[(size_t) (nR + i) * Mx] = be[i];
    }
    __syncthreads();
    for (int i = tid; i < nR; i += T) {
        if (a.has_next) {
            double2 f = Fm2[(size_t) i * Mx];
            double2 ta = mean_mult(a, al, i, a.cn[1]), tb = mean_mult(a, be, i, a.cn[1]);
            Ca[(size_t) i * Mx] = make_double2(ta.x + a.dtd * f.x, ta.y + a.dtd * 0.0);
            Cb[(size_t) i * Mx] = make_double2(tb.x + a.dtd * f.y, tb.y + a.dtd * 0.0);
        } else {
            Ca[(size_t) i * Mx] = al[i];
            Cb[(size_t) i * Mx] = be[i];
        }
    }
}

// grid (ceil(Nz*Mx/SBT) + 1, 4): blockIdx.y = field*2 + parity; the last
// x-block with blockIdx.y == 0 is the mean mode.  Block of 256 threads.
template <int SBT>
__global__ __launch_bounds__(256) void k_solve(SolveArgs a, cplx *C, cplx *Cs)
{
    extern __shared__ double lds_solve[];
    if (*a.abort_flag) return;
    if (blockIdx.x == gridDim.x - 1) {
        if (blockIdx.y == 0) solve_mean(a, C, Cs, lds_solve);
        return;
    }
    const int var = blockIdx.y >> 1, par = blockIdx.y & 1;
    const int col0 = blockIdx.x * SBT;
    if (var == 0)
        solve_block<4, SBT>(a, 0, par, col0, C, Cs, lds_solve);
    else
        solve_block<2, SBT>(a, 1, par, col0, C, Cs, lds_solve);
}

// coefficient tables for the stages
static const double FA_[3] = {1. / 3., -1. / 2, 1. / 3.};
static const double FC_[3] = {1. / 2, 1. / 3., 1.};
static const double FD_[3] = {0., -1. / 6, -2. / 3};
static const double BA_[3] = {1. / 3., -1. / 2, 1. / 3.};
static const double BC_[3] = {1., 1. / 3., 1. / 2.};
static const double BD_[3] = {0., -2. / 3, -1. / 6};

size_t solve_lds_bytes(int dimR, int sbt)
{
    const size_t nr = (size_t) (dimR + 1) / 2;
    const size_t blk = nr * sbt * (9 * 8 + 16 + 16);
    const size_t mean = (size_t) dimR * (3 * 16 + 5 * 8) + 256 * 8;
    return blk > mean ? blk : mean;
}

int solve_columns_per_block(int dimR)
{
    for (int sbt = 8; sbt > 1; sbt >>= 1)
        if (solve_lds_bytes(dimR, sbt) <= 48 * 1024) return sbt;
    return 1;
}

void launch_solve(Context &c, cplx *Ccarry, cplx *Csol, int k, int adj, const cplx *Fa2,
                  const cplx *Fb2, double target)
{
    const Dims &d = c.d;
    const double *A = adj ? BA_ : FA_, *Cc = adj ? BC_ : FC_, *D = adj ? BD_ : FD_;
    SolveArgs a;
    a.Qs = c.dQs; a.Qps = c.dQps; a.Qpps = c.dQpps; a.Rs = c.dRs; a.Rps = c.dRps;
    a.Kx = c.dKx; a.Kz = c.dKz;
    a.La = c.dLa[adj][k];
    a.Lb = c.dLb[adj][k];
    a.Lm = c.dLm[adj][k];
    a.add = c.dAdd[adj][k];
    a.rowsum = c.dRowsum;
    a.Fa = c.Fa; a.Fb = c.Fb;
    a.Fa2 = Fa2; a.Fb2 = Fb2;
    a.Nz = d.Nz; a.Mx = d.Mx; a.dimR = d.dimR; a.dimQ = d.dimQ; a.k = k;
    a.has_next = k != 2;
    a.dtc = d.dt * Cc[k];
    a.dtd = k != 2 ? d.dt * D[k + 1] : 0.0;
    // v: -(K2 Qs + Qps) + re a dt (K2^2 Qs + 2 K2 Qps + Qpps);
    // omega_y: Rs - re a dt (Rps + K2 Rs)  (channel/project.c:55-71,140-151)
    a.ce0[0] = d.re * A[0] * d.dt;
    a.ce0[1] = -(d.re * A[0] * d.dt);
    a.cn[0] = k != 2 ? d.re * A[k + 1] * d.dt : 0.0;
    a.cn[1] = k != 2 ? -(d.re * A[k + 1] * d.dt) : 0.0;
    a.target = target;
    a.abort_flag = c.abort_flag;
    // RHS in, projections in, solution and next right-hand side out, factors in
    double bytes = (double) (d.Nz - 1) * d.Mx *
                   (16.0 * (2.0 * d.dimR) * 3.0 + 16.0 * (d.dimQ + d.dimR) * (k != 2 ? 2.0 : 1.0) +
                    8.0 * (9.0 * d.dimQ + 5.0 * d.dimR));
    prof_begin(c, c.stream, 6, bytes);
    const int sbt = solve_columns_per_block(d.dimR);
    const int ncol = d.Nz * d.Mx;
    const size_t lds = solve_lds_bytes(d.dimR, sbt);
    dim3 grid((ncol + sbt - 1) / sbt + 1, 4);
    if (sbt == 8)
        hipLaunchKernelGGL(k_solve<8>, grid, dim3(256), lds, c.stream, a, Ccarry, Csol);
    else if (sbt == 4)
        hipLaunchKernelGGL(k_solve<4>, grid, dim3(256), lds, c.stream, a, Ccarry, Csol);
    else if (sbt == 2)
        hipLaunchKernelGGL(k_solve<2>, grid, dim3(256), lds, c.stream, a, Ccarry, Csol);
    else
        hipLaunchKernelGGL(k_solve<1>, grid, dim3(256), lds, c.stream, a, Ccarry, Csol);
    prof_end(c, c.stream, 6);
}

// ------------------------------------------------------------ vectors
__global__ void k_axpy(cplx *y, const cplx *x, double a, size_t n, const int *abort_flag)
{
    if (*abort_flag) return;
    for (size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x; e < n;
         e += (size_t) gridDim.x * blockDim.x) {
        y[e].x -= a * x[e].x;
        y[e].y -= a * x[e].y;
    }
}

__global__ void k_diff_scale(cplx *o, const cplx *p, const cplx *m, double s, size_t n)
{
    for (size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x; e < n;
         e += (size_t) gridDim.x * blockDim.x) {
        o[e].x = s * (p[e].x - m[e].x);
        o[e].y = s * (p[e].y - m[e].y);
    }
}

__global__ void k_ddt_reduce(const double *part, int nparts, double *scal)
{
    __shared__ double r0[256], r1[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    r0[threadIdx.x] = a;
    r1[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            r0[threadIdx.x] += r0[threadIdx.x + o];
            r1[threadIdx.x] += r1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[0] = r0[0];
        scal[1] = r1[0];
        scal[2] = r0[0] / r1[0];
    }
}

__global__ void k_ddt_apply(cplx *v, const cplx *dC, const double *scal, size_t n)
{
    double r = scal[2];
    for (size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x; e < n;
         e += (size_t) gridDim.x * blockDim.x) {
        v[e].x -= r * dC[e].x;
        v[e].y -= r * dC[e].y;
    }
}

// first plane whose CFL number reaches the limit sets abort_flag = 1 + plane
__global__ void k_cfl_check(const double *cflmax, int qpts, double dt, int *abort_flag)
{
    if (threadIdx.x != 0 || *abort_flag) return;
    for (int y = 0; y < qpts; ++y) {
        if (cflmax[y] * dt >= 2.8) {
            *abort_flag = 1 + y;
            return;
        }
    }
}

static int nblocks(size_t n) { return (int) std::min<size_t>((n + 255) / 256, 8192); }

void launch_axpy_c(Context &c, cplx *y, const cplx *x, double a, size_t n)
{
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), dim3(256), 0, c.stream, y, x, a, n, c.abort_flag);
}

void launch_diff_scale(Context &c, cplx *out, const cplx *p, const cplx *m, double s, size_t n)
{
    hipLaunchKernelGGL(k_diff_scale, dim3(nblocks(n)), dim3(256), 0, c.stream, out, p, m, s, n);
}

void launch_ddt_reduce(Context &c, int nparts)
{
    hipLaunchKernelGGL(k_ddt_reduce, dim3(1), dim3(256), 0, c.stream, c.dotpart, nparts, c.scal);
}

void launch_ddt_apply(Context &c, cplx *v, const cplx *dC, size_t n)
{
    hipLaunchKernelGGL(k_ddt_apply, dim3(nblocks(n)), dim3(256), 0, c.stream, v, dC, c.scal, n);
}

void launch_cfl_check(Context &c)
{
    hipLaunchKernelGGL(k_cfl_check, dim3(1), dim3(64), 0, c.stream, c.cflmax, c.d.qpts, c.d.dt,
                       c.abort_flag);
}

}  // namespace sc
