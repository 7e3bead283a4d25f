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
#include <type_traits>

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

// The FP32 form for the mixed-precision rebuild, v_mfma_f32_16x16x4f32: the
// same operand layout (A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k),
// but result row i sits in lane group i / 4, item i % 4, where the FP64 form
// puts it in lane group i % 4, item i / 4 (both measured on gfx950).  The
// FP32 instance therefore feeds A row i with the data of row
// 4 (i % 4) + i / 4 (mfma_arow), which leaves every result where the FP64
// instance, and the shared store code, have it.
typedef float ff4 __attribute__((ext_vector_type(4)));
template <class T> __device__ __forceinline__ int mfma_arow(int i)
{
    return sizeof(T) == sizeof(double) ? i : ((i & 3) << 2) | (i >> 2);
}

__device__ __forceinline__ mf4 mfma_t(double a, double b, mf4 c) { return mfma_f64(a, b, c); }
__device__ __forceinline__ ff4 mfma_t(float a, float b, ff4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// accumulator type of the y-transform products in arithmetic type T
template <class T> struct YAcc;
template <> struct YAcc<double> { typedef mf4 type; };
template <> struct YAcc<float> { typedef ff4 type; };
template <class T> __device__ __forceinline__ typename YAcc<T>::type yacc_zero()
{
    typename YAcc<T>::type z = {0, 0, 0, 0};
    return z;
}

// The 16 real columns of an n-tile are eight complex columns; real column n
// is complex column 2(n%4) + n/8, part (n/4)%2.  Item r of lane l then holds
// complex column 2(l/16) + r/2, part r%2: both parts of two neighbouring
// columns sit in one lane.
__device__ __forceinline__ int ncol_x(int n) { return 2 * (n & 3) + (n >> 3); }
__device__ __forceinline__ int ncol_part(int n) { return (n >> 2) & 1; }


// most blocks per XCD of the parity rebuild, split over its half-domain y-tiles
#ifndef SC_RB_BLOCKS
#define SC_RB_BLOCKS 64
#endif

// wave index within the block, in a scalar register: the compiler then
// treats everything derived from it as uniform across the wave
__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane((int) (threadIdx.x >> 6));
}

__device__ __forceinline__ mf4 mf4_zero()
{
    mf4 z = {0.0, 0.0, 0.0, 0.0};
    return z;
}

// Wavenumbers formed in registers, bit-identical to the host tables
// (kx = x 2pi/Lx; kz = z 2pi/Lz, z - Nz above Nz/2): no global load, and so
// no wait on the in-order memory counter behind prefetched operands.
__device__ __forceinline__ double wave_kx(int x, double ax) { return x * ax; }
__device__ __forceinline__ double wave_kz(int z, int Nz, double az)
{
    return (z <= Nz / 2 ? z : z - Nz) * az;
}

// ------------------------------------------------------------ rebuild
struct RebuildArgs {
    const double *frag, *Kx, *Kz;
    int Nz, Mx, qpts, dimR, dimQ, nks, nyt, nseg;
    double ax, az;                      // 2 pi / Lx, 2 pi / Lz
    const int *abort_flag;
    unsigned *queue;                    // [8][nyt][2] task and exit counters, zero between launches
    int bpq;                            // blocks per (XCD, y-tile) queue
    int nprob;                          // 2: a second rebuild C1 -> U1 in the same launch
    const cplx *C1;
    cplx *U1;
    // members: problem p = m nprob + sub (sub 1: C1 -> U1) reads C (or C1)
    // + m ms_c and writes U (or U1) + m ms_u; k_rebuild: grid y = member
    int nm;
    size_t ms_c, ms_u;
};

// the coefficient and Gauss-point arrays of problem p (member p / nprob)
__device__ __forceinline__ const cplx *rb_src(const RebuildArgs &a, const cplx *C, int p)
{
    const int m = p / a.nprob;
    return (p - m * a.nprob ? a.C1 : C) + m * a.ms_c;
}
__device__ __forceinline__ cplx *rb_dst(const RebuildArgs &a, cplx *U, int p)
{
    const int m = p / a.nprob;
    return (p - m * a.nprob ? a.U1 : U) + m * a.ms_u;
}

// U[z][0..4][y][x] from C[z][0..1][j][x]; 1-D grid over the (kz row, y-tile)
// tiles, y-tile fastest, XCD-aware (xcd_tile); 256 threads.
// D[n][y] = sum_j C[j][n] B_m[y][j] for the bases m = Q, Qp, Qpp (alpha
// columns) and R, Rp (beta columns); wave w takes the n-tiles w, w+4, ... of
// the block's y-tile.  The basis fragments of a chunk of KC_REBUILD k-steps are
// staged in LDS (double-buffered) and shared by the four waves; the
// coefficient fragments are prefetched one chunk ahead in registers.
// Epilogue: u, w, du/dy, dw/dy from v', v'', omega_y, omega_y'
// (channel/initAlphaBeta2.c:96-240).
// The Nyquist row kz = Nz/2 carries no data, so its blocks rebuild the mean
// column (kx, kz) = (0, 0) instead: u = R alpha, du/dy = Rp alpha,
// w = R beta, dw/dy = Rp beta, v = 0 (channel/initAlphaBeta2.c:22-68), with
// the n rows 0..3 of one tile holding Re/Im alpha, Re/Im beta.
__global__ __launch_bounds__(256, SC_YT_WAVES) void k_rebuild(RebuildArgs a, const cplx *C, cplx *U)
{
    constexpr int KC = KC_REBUILD, CH = KC * 5 * 64, PT = CH / 256;
    static_assert(CH % 256 == 0, "chunk must split over the block");
    __shared__ double bs[2][CH];
    if (*a.abort_flag) return;
    const int tile = xcd_tile(a.nyt * a.Nz);
    if (tile < 0) return;
    C += blockIdx.y * a.ms_c;
    U += blockIdx.y * a.ms_u;
    const int zrow = tile / a.nyt, yt = tile - zrow * a.nyt;
    const bool meanb = zrow == a.Nz / 2;
    const int z = meanb ? 0 : zrow;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int Mx = a.Mx, q = a.qpts, nR = a.dimR, nQ = a.dimQ;
    const int nnt = meanb ? 1 : (Mx + 7) / 8, nch = a.nks / KC;
    const int kq = lane >> 4, nl = lane & 15;
    const double *frag = a.frag + (size_t) yt * a.nks * 5 * 64;
    const double *Cd = (const double *) (C + (size_t) z * 2 * nR * Mx);
    const size_t ss = (size_t) q * Mx;
    cplx *Uz = U + (size_t) z * 5 * ss;
    const double kz = a.Kz[z];
    for (int nt0 = 0; nt0 < nnt; nt0 += 4) {
        const int nt = nt0 + wave;
        const bool active = nt < nnt;
        int xcol, part;
        bool okA, okB;
        size_t offA, offB;
        if (meanb) {
            // n = 0, 1: Re, Im alpha(0,0); n = 2, 3: Re, Im beta(0,0)
            xcol = 0;
            part = nl & 1;
            okA = active && nl < 4;
            okB = false;
            offA = (size_t) (nl >> 1) * nR * Mx * 2 + part;
            offB = 0;
        } else {
            xcol = nt * 8 + ncol_x(nl);
            part = ncol_part(nl);
            okA = okB = active && xcol < Mx;
            offA = (size_t) xcol * 2 + part;
            offB = (size_t) nR * Mx * 2 + offA;
        }
        const int limA = meanb ? nR : nQ;
        mf4 acc[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) acc[m] = mf4_zero();
        // two register sets of coefficient fragments, used alternately
        double ca0[KC], cb0[KC], ca1[KC], cb1[KC], st[PT];
        auto load_coef = [&](int ch, double *ca, double *cb) {
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                const int j = (ch * KC + s) * 4 + kq;
                ca[s] = (okA && j < limA) ? Cd[offA + (size_t) j * Mx * 2] : 0.0;
                cb[s] = (okB && j < nR) ? Cd[offB + (size_t) j * Mx * 2] : 0.0;
            }
        };
        auto load_basis = [&](int ch) {
            const double *src = frag + (size_t) ch * CH;
#pragma unroll
            for (int i = 0; i < PT; ++i) st[i] = src[threadIdx.x + 256 * i];
        };
        auto put_basis = [&](int buf) {
#pragma unroll
            for (int i = 0; i < PT; ++i) bs[buf][threadIdx.x + 256 * i] = st[i];
        };
        auto compute = [&](const double *xa, const double *xb, const double *b) {
            if (active && meanb) {
#pragma unroll
                for (int s = 0; s < KC; ++s) {
                    acc[3] = mfma_f64(xa[s], b[(s * 5 + 3) * 64 + lane], acc[3]);
                    acc[4] = mfma_f64(xa[s], b[(s * 5 + 4) * 64 + lane], acc[4]);
                }
            } else if (active) {
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
                }
            }
        };
        load_coef(0, ca0, cb0);
        load_basis(0);
        put_basis(0);
        __syncthreads();
        for (int ch = 0; ch < nch; ch += 2) {
            const bool n1 = ch + 1 < nch;
            if (n1) {
                load_coef(ch + 1, ca1, cb1);
                load_basis(ch + 1);
            }
            compute(ca0, cb0, bs[0]);
            if (n1) put_basis(1);
            __syncthreads();
            if (!n1) break;
            const bool n2 = ch + 2 < nch;
            if (n2) {
                load_coef(ch + 2, ca0, cb0);
                load_basis(ch + 2);
            }
            compute(ca1, cb1, bs[1]);
            if (n2) put_basis(0);
            __syncthreads();
        }
        const int y = yt * 16 + nl;
        if (!active || y >= q) continue;
        if (meanb) {
            // item 0 of lane group kq holds n = kq: u (kq < 2) or w (kq >= 2)
            double *Ur = (double *) (Uz + (size_t) y * Mx) + (kq & 1);
            const int su = kq < 2 ? 0 : 2;
            Ur[(size_t) su * ss * 2] = acc[3][0];
            Ur[(size_t) (su == 0 ? 3 : 4) * ss * 2] = acc[4][0];
            Ur[(size_t) 1 * ss * 2] = 0.0;
            continue;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int x = nt * 8 + 2 * kq + h;
            if (x >= Mx || (z == 0 && x == 0)) continue;
            const size_t o = (size_t) y * Mx + x;
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

// Operands of one (kz row, n-tile) task of the rebuild: lane offsets into C
// of its alpha and beta fragments.  Every lane loads (the loads carry no
// predicate, so nothing waits on them before use): columns past Mx read the
// last column and are never stored, rows past the basis meet zero basis
// fragments, and in the mean row only the n rows 0..3 are stored.
struct RebuildTask {
    size_t offA, offB;
    const double *cd;                   // coefficient array of the task's rebuild
};

__device__ __forceinline__ RebuildTask rebuild_task(const RebuildArgs &a, int zrow, int nt, int nl)
{
    RebuildTask t;
    const int Mx = a.Mx, nR = a.dimR;
    if (zrow == a.Nz / 2) {
        // n = 0, 1: Re, Im alpha(0,0); n = 2, 3: Re, Im beta(0,0)
        t.offA = (size_t) ((nl >> 1) & 1) * nR * Mx * 2 + (nl & 1);
        t.offB = t.offA;
    } else {
        const int xcol = min(nt * 8 + ncol_x(nl), Mx - 1);
        t.offA = (size_t) zrow * 2 * nR * Mx * 2 + (size_t) xcol * 2 + ncol_part(nl);
        t.offB = t.offA + (size_t) nR * Mx * 2;
    }
    return t;
}

// Gauss-point values of one task from its five accumulators (D[n][y] layout)
// VEL: the velocity slots u, v, w only (acc[2], acc[4] unused).  The slots
// are stored as O: double2, or float2 for the FP32 instance (the mixed-
// precision steps, whose z-inverse transform reads FP32 slots at the same
// element index; the member base address stays that of the FP64 layout).
template <bool VEL = false, class ACC = mf4, class O = double2>
__device__ __forceinline__ void rebuild_store(const RebuildArgs &a, int zrow, int nt, int y, int kq,
                                              const ACC *acc, cplx *U)
{
    typedef decltype(O().x) S;
    const int Mx = a.Mx, q = a.qpts;
    if (y >= q) return;
    const size_t ss = (size_t) q * Mx;
    O *Uo = (O *) U;
    if (zrow == a.Nz / 2) {
        // mean column: item 0 of lane group kq holds n = kq, Re/Im of
        // alpha (kq < 2: u, du/dy) or beta (kq >= 2: w, dw/dy)
        S *Ur = (S *) (Uo + (size_t) y * Mx) + (kq & 1);
        const int su = kq < 2 ? 0 : 2;
        Ur[(size_t) su * ss * 2] = (S) acc[3][0];
        if (!VEL) Ur[(size_t) (su == 0 ? 3 : 4) * ss * 2] = (S) acc[4][0];
        Ur[(size_t) 1 * ss * 2] = (S) 0;
        return;
    }
    O *Uz = Uo + (size_t) zrow * 5 * ss;
    auto put = [&](size_t i, double re, double im) { Uz[i] = O{(S) re, (S) im}; };
    const double kz = wave_kz(zrow, a.Nz, a.az);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int x = nt * 8 + 2 * kq + h;
        if (x >= Mx || (zrow == 0 && x == 0)) continue;
        const size_t o = (size_t) y * Mx + x;
        // one reciprocal per column instead of eight divisions
        const double kx = wave_kx(x, a.ax), r2 = 1.0 / (kz * kz + kx * kx);
        const double2 v = make_double2(acc[0][2 * h], acc[0][2 * h + 1]);
        const double2 f = make_double2(-acc[1][2 * h], -acc[1][2 * h + 1]);
        const double2 g = make_double2(acc[3][2 * h], acc[3][2 * h + 1]);
        put(o + 1 * ss, v.x, v.y);
        put(o + 0 * ss, (kx * f.y + kz * g.y) * r2, -(kx * f.x + kz * g.x) * r2);
        put(o + 2 * ss, (-kx * g.y + kz * f.y) * r2, (kx * g.x - kz * f.x) * r2);
        if (VEL) continue;
        const double2 q2 = make_double2(acc[2][2 * h], acc[2][2 * h + 1]);
        const double2 gp = make_double2(acc[4][2 * h], acc[4][2 * h + 1]);
        put(o + 3 * ss, (-kx * q2.y + kz * gp.y) * r2, (kx * q2.x - kz * gp.x) * r2);
        put(o + 4 * ss, -(kx * gp.y + kz * q2.y) * r2, (kx * gp.x + kz * q2.x) * r2);
    }
}

// Rebuild for dimR <= 4*NKS with dynamic task queues: a block owns one
// y-tile and stages its whole basis (NKS k-steps) in LDS once; its waves then
// take (kz row, n-tile) tasks (load the 2 NKS coefficient fragments, 5 NKS
// products, store) from the queue of their (XCD, y-tile) until it is empty,
// so the products spread evenly over the SIMDs.  Blocks are dispatched
// round-robin over the 8 XCDs (block b on XCD b % 8), and the queue of XCD x
// holds the kz rows [x Nz/8, (x+1) Nz/8): the coefficient rows of an XCD's
// tasks stay in its L2.  A queue is a task counter and an exit counter; the
// last wave of the queue to exit clears both, so they are zero between
// launches.  Two blocks share a CU.  Grid: 8 * nyt * bpq blocks.
template <int NKS>
__global__ __launch_bounds__(256, 2) void k_rebuild_rows(RebuildArgs a, const cplx *C, cplx *U)
{
    __shared__ double bs[NKS * 5 * 64];
    if (*a.abort_flag) return;
    const int xcd = blockIdx.x & 7, ib = blockIdx.x >> 3;
    const int yt = ib % a.nyt;
    const int z0 = xcd * a.Nz / 8, z1 = (xcd + 1) * a.Nz / 8;
    unsigned *qc = a.queue + ((size_t) xcd * a.nyt + yt) * 2;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int kq = lane >> 4, nl = lane & 15;
    const int nnt = (a.Mx + 7) / 8, nR = a.dimR;
    const double *frag = a.frag + (size_t) yt * NKS * 5 * 64;
    {
        // all loads in flight before the first LDS store
        constexpr int PT = NKS * 5 * 64 / 256;
        static_assert(NKS * 5 * 64 % 256 == 0, "basis must split over the block");
        double st[PT];
#pragma unroll
        for (int i = 0; i < PT; ++i) st[i] = frag[threadIdx.x + 256 * i];
#pragma unroll
        for (int i = 0; i < PT; ++i) bs[threadIdx.x + 256 * i] = st[i];
    }
    // task t of the queue: rows z0..z1-1 in order, nnt n-tiles each (one in
    // the mean row); tasks p ntask1 .. belong to problem p (rb_src)
    const int zm = a.Nz / 2;
    const bool has_mean = zm >= z0 && zm < z1;
    const int ntask1 = (z1 - z0) * nnt - (has_mean ? nnt - 1 : 0);
    const int ntask = ntask1 * a.nprob * a.nm;
    auto task_of = [&](int t, int &zrow, int &nt) {
        t -= (t / ntask1) * ntask1;
        const int tm = has_mean ? (zm - z0) * nnt : ntask1;  // first task of the mean row
        if (t < tm) {
            zrow = z0 + t / nnt;
            nt = t % nnt;
        } else if (t == tm) {
            zrow = zm;
            nt = 0;
        } else {
            const int u = t - tm - 1;
            zrow = zm + 1 + u / nnt;
            nt = u % nnt;
        }
    };
    auto grab = [&]() {
        unsigned t0 = 0;
        if (lane == 0) t0 = atomicAdd(&qc[0], 1u);
        return __builtin_amdgcn_readfirstlane((int) t0);
    };
    // The coefficient fragments of a task move in two halves, one register
    // buffer each: the loads of the second half are issued before the
    // products of the first, those of the next task's first half before the
    // products of the second.
    constexpr int CK = NKS / 2, NCH = 2;
    static_assert(CK % 4 == 0, "a half must be whole groups of four k-steps");
    double ca[NCH][CK], cb[NCH][CK];
    // row j = 4 s + kq of k-step s: a lane base (row kq) plus a uniform step;
    // only the last k-step can pass row dimR - 1 (4 (NKS-1) < dimR), so only
    // it is clamped
    const size_t rstep = (size_t) 4 * a.Mx * 2;
    auto load = [&](const RebuildTask &t, int ch) {
        const double *Cd = t.cd;
        const double *pa = Cd + t.offA + (size_t) kq * a.Mx * 2;
        const double *pb = Cd + t.offB + (size_t) kq * a.Mx * 2;
#pragma unroll
        for (int s = 0; s < CK; ++s) {
            const int ks = ch * CK + s;
            if (ks == NKS - 1) {
                const size_t j = min(ks * 4 + kq, nR - 1);
                ca[ch][s] = Cd[t.offA + j * a.Mx * 2];
                cb[ch][s] = Cd[t.offB + j * a.Mx * 2];
            } else {
                ca[ch][s] = pa[ks * rstep];
                cb[ch][s] = pb[ks * rstep];
            }
        }
    };
    // products of chunk ch (basis fragments LDS -> registers two k-steps at a
    // time, one group ahead; the scheduling barriers keep the compiler from
    // sinking those loads to their uses)
    auto chunk = [&](bool mean, int ch, mf4 *acc) {
        const double *xa = ca[ch], *xb = cb[ch];
        if (mean) {
#pragma unroll
            for (int t = 0; t < CK; ++t) {
                const int s = ch * CK + t;
                acc[3] = mfma_f64(xa[t], bs[(s * 5 + 3) * 64 + lane], acc[3]);
                acc[4] = mfma_f64(xa[t], bs[(s * 5 + 4) * 64 + lane], acc[4]);
            }
            return;
        }
        double p0[10], p1[10];
        auto ldb = [&](int s0, double *p) {
#pragma unroll
            for (int e = 0; e < 10; ++e) p[e] = bs[(s0 * 5 + e) * 64 + lane];
        };
        auto mm = [&](int t0, const double *p) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = t0 + u;
                acc[0] = mfma_f64(xa[t], p[u * 5 + 0], acc[0]);
                acc[1] = mfma_f64(xa[t], p[u * 5 + 1], acc[1]);
                acc[2] = mfma_f64(xa[t], p[u * 5 + 2], acc[2]);
                acc[3] = mfma_f64(xb[t], p[u * 5 + 3], acc[3]);
                acc[4] = mfma_f64(xb[t], p[u * 5 + 4], acc[4]);
            }
        };
        const int sb = ch * CK;
        ldb(sb, p0);
#pragma unroll
        for (int t0 = 0; t0 < CK; t0 += 4) {
            if (t0 + 2 < CK) ldb(sb + t0 + 2, p1);
            __builtin_amdgcn_sched_barrier(0);
            mm(t0, p0);
            __builtin_amdgcn_sched_barrier(0);
            if (t0 + 2 >= CK) break;
            if (t0 + 4 < CK) ldb(sb + t0 + 4, p0);
            __builtin_amdgcn_sched_barrier(0);
            mm(t0 + 2, p1);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    __syncthreads();
    (void) wave;
    // operands of task t: offsets and the coefficient array of its rebuild
    auto task_ops = [&](int t, int zr, int nt) {
        RebuildTask r = rebuild_task(a, zr, nt, nl);
        r.cd = (const double *) rb_src(a, C, t / ntask1);
        return r;
    };
    int t = grab();
    if (t >= ntask) {
        t = ntask;
    } else {
        int zr, nt;
        task_of(t, zr, nt);
        load(task_ops(t, zr, nt), 0);
    }
    while (t < ntask) {
        int zr, nt;
        task_of(t, zr, nt);
        const RebuildTask cur = task_ops(t, zr, nt);
        const bool mean = zr == zm;
        mf4 acc[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) acc[m] = mf4_zero();
        int tn = ntask;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (ch + 1 < NCH) {
                load(cur, ch + 1);
            } else {
                tn = grab();
                if (tn < ntask) {
                    int zn, ntn;
                    task_of(tn, zn, ntn);
                    load(task_ops(tn, zn, ntn), 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            chunk(mean, ch, acc);
            __builtin_amdgcn_sched_barrier(0);
        }
        rebuild_store(a, zr, nt, yt * 16 + nl, kq, acc, rb_dst(a, U, t / ntask1));
        t = tn;
    }
    // every wave of the queue's blocks exits here once; the last clears it
    if (lane == 0) {
        const unsigned done = atomicAdd(&qc[1], 1u);
        if (done == (unsigned) (a.bpq * 4 - 1)) {
            atomicExch(&qc[0], 0u);
            atomicExch(&qc[1], 0u);
        }
    }
}

// Parity form of the queued rebuild.  The Gauss points are mirror images
// (y[q-1-i] = -y[i], exactly) and every basis column has a parity, exactly:
// Q_j, Qpp_j and R_j that of j, Qp_j and Rp_j the opposite.  With E and O
// the sums over the even and the odd j at a point y_i of the half domain
// i < (q+1)/2,
//     U(y_i) = E + O,    U(y_{q-1-i}) = s (E - O),
// s = +1 for Q, Qpp, R and -1 for Qp, Rp.  A task covers 16 half-domain
// points and their mirrors with the 5 NKS products that the direct form
// spends on 16 points, so the products halve.  For odd q the middle point is
// y = 0, where the columns of the other parity vanish exactly and E + O is
// stored once.  k-step s < NKS/2 holds the even rows j = 8s + 2(l/16), the
// others the odd rows; the coefficients move in four quarters (even rows,
// then odd rows, two halves each) through two register buffers, the next
// quarter's loads issued before the current quarter's products.  Blocks and
// the mean task as in k_rebuild_rows (a.nyt is the number of half-domain
// y-tiles, a.frag the parity-ordered basis), but the tasks are assigned
// statically: wave w of the bpq blocks of an (XCD, y-tile) takes the tasks
// w, w + 4 bpq, ..., and issues its first coefficient loads before it
// stages the basis.
template <int NKS, bool VEL, class T>
__global__ __launch_bounds__(256, 2) void k_rebuild_par(RebuildArgs a, const cplx *C, cplx *U)
{
    using ACC = typename YAcc<T>::type;
    __shared__ T bs[NKS * 5 * 64];
    if (*a.abort_flag) return;
    const int xcd = blockIdx.x & 7, ib = blockIdx.x >> 3;
    const int yt = ib % a.nyt;
    const int z0 = xcd * a.Nz / 8, z1 = (xcd + 1) * a.Nz / 8;
    const int lane = threadIdx.x & 63;
    const int kq = lane >> 4, nl = lane & 15;
    const int nnt = (a.Mx + 7) / 8, nR = a.dimR, M2 = a.Mx * 2;
    const int tstep = 4 * a.bpq;
    const int zm = a.Nz / 2;
    const bool has_mean = zm >= z0 && zm < z1;
    const int ntask1 = (z1 - z0) * nnt - (has_mean ? nnt - 1 : 0);
    const int ntask = ntask1 * a.nprob * a.nm;
    auto task_of = [&](int t, int &zrow, int &nt) {
        t -= (t / ntask1) * ntask1;
        const int tm = has_mean ? (zm - z0) * nnt : ntask1;
        if (t < tm) {
            zrow = z0 + t / nnt;
            nt = t % nnt;
        } else if (t == tm) {
            zrow = zm;
            nt = 0;
        } else {
            const int u = t - tm - 1;
            zrow = zm + 1 + u / nnt;
            nt = u % nnt;
        }
    };
    constexpr int NKH = NKS / 2, CK = NKS / 4;
    static_assert(CK % 2 == 0, "a quarter must be whole pairs of k-steps");
    T ca[2][CK], cb[2][CK];
    // quarter ch: parity ch / 2, k-steps (ch % 2) CK .. + CK - 1 of that
    // parity, as buffer loads over the whole coefficient array: a lane
    // offset plus a uniform k-step offset.  Rows past dimR - 1 (whose basis
    // fragments are zero) read the next rows of the array, or zero past its
    // end, so no row index is clamped.
    const unsigned cbytes = (unsigned) ((size_t) a.Nz * 2 * nR * M2 * sizeof(double));
    auto load = [&](const RebuildTask &t, int ch, int buf) {
        __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *) t.cd, 0, cbytes, 0x00020000);
        const int par = ch >> 1, s0 = (ch & 1) * CK;
        const int row = (2 * kq + par) * M2;
        const int va = (int) (unsigned) ((t.offA + row) * 8), vb = (int) (unsigned) ((t.offB + row) * 8);
#pragma unroll
        for (int s = 0; s < CK; ++s) {
            const int so = (s0 + s) * 8 * M2 * 8;
            ca[buf][s] = (T) __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, va, so, 0));
            cb[buf][s] = (T) __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, vb, so, 0));
        }
    };
    auto chunk = [&](bool mean, int ch, int buf, ACC *acc) {
        const T *xa = ca[buf], *xb = cb[buf];
        const int sb = (ch >> 1) * NKH + (ch & 1) * CK;
        if (mean) {
#pragma unroll
            for (int t = 0; t < CK; ++t) {
                acc[3] = mfma_t(xa[t], bs[((sb + t) * 5 + 3) * 64 + lane], acc[3]);
                if (!VEL) acc[4] = mfma_t(xa[t], bs[((sb + t) * 5 + 4) * 64 + lane], acc[4]);
            }
            return;
        }
        T p0[10], p1[10];
        auto ldb = [&](int s0, T *p) {
#pragma unroll
            for (int e = 0; e < 10; ++e)
                if (!VEL || (e % 5 != 2 && e % 5 != 4)) p[e] = bs[(s0 * 5 + e) * 64 + lane];
        };
        auto mm = [&](int t0, const T *p) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = t0 + u;
                acc[0] = mfma_t(xa[t], p[u * 5 + 0], acc[0]);
                acc[1] = mfma_t(xa[t], p[u * 5 + 1], acc[1]);
                if (!VEL) acc[2] = mfma_t(xa[t], p[u * 5 + 2], acc[2]);
                acc[3] = mfma_t(xb[t], p[u * 5 + 3], acc[3]);
                if (!VEL) acc[4] = mfma_t(xb[t], p[u * 5 + 4], acc[4]);
            }
        };
        ldb(sb, p0);
#pragma unroll
        for (int t0 = 0; t0 < CK; t0 += 4) {
            if (t0 + 2 < CK) ldb(sb + t0 + 2, p1);
            __builtin_amdgcn_sched_barrier(0);
            mm(t0, p0);
            __builtin_amdgcn_sched_barrier(0);
            if (t0 + 2 >= CK) break;
            if (t0 + 4 < CK) ldb(sb + t0 + 4, p0);
            __builtin_amdgcn_sched_barrier(0);
            mm(t0 + 2, p1);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    auto task_ops = [&](int t, int zr, int nt) {
        RebuildTask r = rebuild_task(a, zr, nt, mfma_arow<T>(nl));
        r.cd = (const double *) rb_src(a, C, t / ntask1);
        return r;
    };
    const int hq = (a.qpts + 1) / 2;
    const int y = yt * 16 + nl, ym = a.qpts - 1 - y;
    int t = (ib / a.nyt) * 4 + wave_id();
    if (t < ntask) {
        int zr, nt;
        task_of(t, zr, nt);
        load(task_ops(t, zr, nt), 0, 0);
    }
    {
        const double *frag = a.frag + (size_t) yt * NKS * 5 * 64;
        constexpr int PT = NKS * 5 * 64 / 256;
        static_assert(NKS * 5 * 64 % 256 == 0, "basis must split over the block");
        double st[PT];
#pragma unroll
        for (int i = 0; i < PT; ++i) st[i] = frag[threadIdx.x + 256 * i];   // converted when stored
#pragma unroll
        for (int i = 0; i < PT; ++i) bs[threadIdx.x + 256 * i] = (T) st[i];
    }
    __syncthreads();
    while (t < ntask) {
        int zr, nt;
        task_of(t, zr, nt);
        const RebuildTask cur = task_ops(t, zr, nt);
        const bool mean = zr == zm;
        ACC ae[5], ao[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            ae[m] = yacc_zero<T>();
            ao[m] = yacc_zero<T>();
        }
        int tn = ntask;
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
            if (ch < 3) {
                load(cur, ch + 1, (ch + 1) & 1);
            } else {
                tn = t + tstep;
                if (tn < ntask) {
                    int zn, ntn;
                    task_of(tn, zn, ntn);
                    load(task_ops(tn, zn, ntn), 0, 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            chunk(mean, ch, ch & 1, ch < 2 ? ae : ao);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (y < hq) {
            cplx *Ud = rb_dst(a, U, t / ntask1);
            // values at y (E + O), then at the mirror point (s (E - O))
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const ACC e = ae[m], o = ao[m];
                ae[m] = e + o;
                ao[m] = (m == 1 || m == 4) ? o - e : e - o;
            }
            // FP32 products: FP32 slots
            typedef typename std::conditional<sizeof(T) == sizeof(float), float2, double2>::type O;
            rebuild_store<VEL, ACC, O>(a, zr, nt, y, kq, ae, Ud);
            if (ym != y) rebuild_store<VEL, ACC, O>(a, zr, nt, ym, kq, ao, Ud);
        }
        t = tn;
    }
}

// one or two rebuilds (C1 null: one); the queued kernel runs both in one
// launch (they share the staged basis), the chunked kernel one launch each
void launch_rebuild2(Context &c, const cplx *C0, cplx *U0, const cplx *C1, cplx *U1, bool vel_only)
{
    const Dims &d = c.d;
    // segments of kz rows for the chunked kernel
    const int nseg = std::max(1, std::min(d.Nz, 512 / c.nyt));
    // queued form: blocks per (XCD, y-tile), about two blocks per CU
    const int bpq = std::max(1, 64 / c.nyt);
    int nprob = C1 ? 2 : 1;
    const int nm = c.nrun;
    const size_t ms_c = mstride(c, C0), ms_u = mstride(c, U0);
    if (C1 && (mstride(c, C1) != ms_c || mstride(c, U1) != ms_u)) {
        // the two problems' arrays have other member strides: two launches
        launch_rebuild2(c, C0, U0, nullptr, nullptr, vel_only);
        launch_rebuild2(c, C1, U1, nullptr, nullptr, vel_only);
        return;
    }
    RebuildArgs a{c.dFragR, c.dKx, c.dKz, d.Nz, d.Mx, d.qpts, d.dimR, d.dimQ, c.nksR, c.nyt,
                  nseg, 2. * M_PI / d.Lx, 2. * M_PI / d.Lz, c.abort_flag, c.rb_queue, bpq,
                  nprob, C1, U1, nm, ms_c, ms_u};
    const double slot_bytes = slots32(c) ? 8.0 : 16.0;
    double bytes = nm * nprob * (double) (d.Nz - 1) * d.Mx *
                   (16.0 * 2.0 * d.dimR + slot_bytes * (vel_only ? 3.0 : 5.0) * d.qpts);
    const hipStream_t st = c.stream;
    const int nks = c.nksR;
    const dim3 grows(8 * c.nyt * bpq), gchunk(xcd_grid(c.nyt * d.Nz), nm);
    const int nksp = c.nksP;
    const bool par = c.ypar && (nksp == 8 || nksp == 16 || nksp == 24);
    RebuildArgs ap = a;
    ap.frag = c.dFragP;
    ap.nyt = c.nht;
    ap.nks = nksp;
    {
        // enough waves for one task each (of one rebuild of all members), at
        // most about two blocks per CU
        const int ntmax = nm * ((d.Nz + 7) / 8 * ((d.Mx + 7) / 8));
        const int rbb = tuning().rebuild_blocks > 0 ? tuning().rebuild_blocks : SC_RB_BLOCKS;
        ap.bpq = std::max(1, std::min((ntmax + 3) / 4, rbb / c.nht));
    }
    const dim3 gpar(8 * c.nht * ap.bpq);
    // mixed precision: the parity rebuild's products on the FP32 matrix cores
    // and FP32 slots, except for outputs such as spec2phys
    const bool f32 = slots32(c);
    typedef void (*RebuildKernel)(RebuildArgs, const cplx *, cplx *);
#define SC_RB_PAR(N, V) (f32 ? (RebuildKernel) k_rebuild_par<N, V, float> : (RebuildKernel) k_rebuild_par<N, V, double>)
    auto go = [=]() {
        if (par && vel_only) {
            if (nksp == 8)
                hipLaunchKernelGGL(SC_RB_PAR(8, true), gpar, dim3(256), 0, st, ap, C0, U0);
            else if (nksp == 16)
                hipLaunchKernelGGL(SC_RB_PAR(16, true), gpar, dim3(256), 0, st, ap, C0, U0);
            else
                hipLaunchKernelGGL(SC_RB_PAR(24, true), gpar, dim3(256), 0, st, ap, C0, U0);
        } else if (par) {
            if (nksp == 8)
                hipLaunchKernelGGL(SC_RB_PAR(8, false), gpar, dim3(256), 0, st, ap, C0, U0);
            else if (nksp == 16)
                hipLaunchKernelGGL(SC_RB_PAR(16, false), gpar, dim3(256), 0, st, ap, C0, U0);
            else
                hipLaunchKernelGGL(SC_RB_PAR(24, false), gpar, dim3(256), 0, st, ap, C0, U0);
        } else if (nks == 8)
            hipLaunchKernelGGL(k_rebuild_rows<8>, grows, dim3(256), 0, st, a, C0, U0);
        else if (nks == 16)
            hipLaunchKernelGGL(k_rebuild_rows<16>, grows, dim3(256), 0, st, a, C0, U0);
        else if (nks == 24)
            hipLaunchKernelGGL(k_rebuild_rows<24>, grows, dim3(256), 0, st, a, C0, U0);
        else
            for (int p = 0; p < nprob; ++p)
                hipLaunchKernelGGL(k_rebuild, gchunk, dim3(256), 0, st, a, p ? C1 : C0, p ? U1 : U0);
    };
    prof_begin(c, st, 4, bytes);
    go();
    prof_end(c, st, 4, go);
}
#undef SC_RB_PAR

void launch_rebuild(Context &c, const cplx *Csrc, cplx *Udst, bool vel_only)
{
    launch_rebuild2(c, Csrc, Udst, nullptr, nullptr, vel_only);
}

// ------------------------------------------------------------ projection
struct ForceArgs {
    const double *frag, *Kx, *Kz;
    int Nz, Mx, qpts, dimR, dimQ, nks, nit;
    double ax, az;                      // 2 pi / Lx, 2 pi / Lz
    const int *abort_flag;
    // members: problem p = m nprob + sub of the staged-basis kernels reads
    // U (sub 0: U0, 1: U1) + m ms_u and writes Fa, Fb + m ms_fa, m ms_fb;
    // k_force: grid y = member
    int nm;
    size_t ms_u, ms_fa, ms_fb;
};

// Fa[z][i][x] (i < dimQ), Fb[z][i][x] (i < dimR); 1-D grid over the (kz row,
// i-tile) tiles, i-tile fastest, XCD-aware (xcd_tile); 256 threads.
// D[n][i] = sum_y A_m[y][n] W_m[i][y] (channel/project.c:95-115,176-191,
// channel/adjproject.c:108-134,213-234) with the pointwise combinations
//   fa = Qpw P - k2 Qw Hy (+ Qppw Z),     P  = i kx Hx + i kz Hz
//   fb = Rw P2 (+ Rpw Z2),                P2 = i kz Hx - i kx Hz
//   (adjoint) Z = -i kz Sx + i kx Sz,     Z2 = i kz Sz + i kx Sx
// summed into two accumulators.  Each lane forms its A entries from the parts
// of Hx, Hz, Hy, Sx, Sz it needs.  Mean column (channel/project0.c:110-146,
// channel/adjproject0.c:98-140): Fb(0,0) = (Rw Re Hx + Rpw Re Sz,
// Rw Re Hz - Rpw Re Sx).  Wave w takes the n-tiles w, w+4, ... of the
// block's i-tile; the weighted-basis fragments of a chunk of KC_FORCE k-steps are
// staged in LDS (double-buffered), the field values prefetched a chunk ahead.
template <bool ADJ>
__global__ __launch_bounds__(256, SC_YT_WAVES) void k_force(ForceArgs a, const cplx *U, cplx *Fa, cplx *Fb)
{
    constexpr int KC = KC_FORCE, CH = KC * 5 * 64, PT = CH / 256;
    static_assert(CH % 256 == 0, "chunk must split over the block");
    __shared__ double bs[2][CH];
    if (*a.abort_flag) return;
    const int tile = xcd_tile(a.nit * a.Nz);
    if (tile < 0) return;
    U += blockIdx.y * a.ms_u;
    Fa += blockIdx.y * a.ms_fa;
    Fb += blockIdx.y * a.ms_fb;
    const int z = tile / a.nit, it = tile - z * a.nit;
    if (z == a.Nz / 2) return;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int Mx = a.Mx, q = a.qpts, nR = a.dimR, nQ = a.dimQ;
    constexpr bool adj = ADJ;
    const int nnt = (Mx + 7) / 8, nch = a.nks / KC;
    const int kq = lane >> 4, nl = lane & 15;
    const double *frag = a.frag + (size_t) it * a.nks * 5 * 64;
    const size_t ss = (size_t) q * Mx;
    const double *Ud = (const double *) (U + (size_t) z * 5 * ss);
    const double kz = wave_kz(z, a.Nz, a.az);
    for (int nt0 = 0; nt0 < nnt; nt0 += 4) {
        const int nt = nt0 + wave;
        const bool active = nt < nnt;
        // columns past Mx read the last column and are never stored
        const int xcol = min(nt * 8 + ncol_x(nl), Mx - 1), part = ncol_part(nl);
        const bool mean = z == 0 && xcol == 0;
        const double kx = wave_kx(xcol, a.ax), k2 = kz * kz + kx * kx;
        const double sg = part ? 1.0 : -1.0;
        const int other = mean ? 0 : 1 - part;
        mf4 accA = mf4_zero(), accB = mf4_zero();
        // two register sets of field values, used alternately:
        // r[0] Hx, r[1] Hz, r[2] Hy, r[3] Sx, r[4] Sz
        double r0[5][KC], r1[5][KC], st[PT];
        // unpredicated loads: Gauss points past qpts read the last point and
        // meet zero basis fragments
        auto load_fields = [&](int ch, double (*r)[KC]) {
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                const int y = min((ch * KC + s) * 4 + kq, q - 1);
                const size_t e = ((size_t) y * Mx + xcol) * 2;
                r[1][s] = Ud[e + other];
                r[0][s] = Ud[e + 2 * ss + other];
                r[2][s] = Ud[e + 4 * ss + part];
                if (adj) {
                    r[3][s] = Ud[e + 6 * ss + other];
                    r[4][s] = Ud[e + 8 * ss + other];
                }
            }
        };
        auto load_basis = [&](int ch) {
            const double *src = frag + (size_t) ch * CH;
#pragma unroll
            for (int i = 0; i < PT; ++i) st[i] = src[threadIdx.x + 256 * i];
        };
        auto put_basis = [&](int buf) {
#pragma unroll
            for (int i = 0; i < PT; ++i) bs[buf][threadIdx.x + 256 * i] = st[i];
        };
        auto compute = [&](double (*r)[KC], const double *b) {
            if (!active) return;
#pragma unroll
            for (int s = 0; s < KC; ++s) {
                const double hx = r[0][s], hz = r[1][s], hy = r[2][s];
                double vP, vY, vP2, vZ = 0.0, vZ2 = 0.0;
                if (mean) {
                    vP = 0.0;
                    vY = 0.0;
                    vP2 = part ? hz : hx;
                    if (adj) vZ2 = part ? -r[3][s] : r[4][s];
                } else {
                    vP = sg * (kx * hx + kz * hz);
                    vY = -k2 * hy;
                    vP2 = sg * (kz * hx - kx * hz);
                    if (adj) {
                        vZ = -sg * (kz * r[3][s] - kx * r[4][s]);
                        vZ2 = sg * (kz * r[4][s] + kx * r[3][s]);
                    }
                }
                accA = mfma_f64(vP, b[(s * 5 + 0) * 64 + lane], accA);
                accB = mfma_f64(vP2, b[(s * 5 + 3) * 64 + lane], accB);
                accA = mfma_f64(vY, b[(s * 5 + 1) * 64 + lane], accA);
                if (adj) {
                    accA = mfma_f64(vZ, b[(s * 5 + 2) * 64 + lane], accA);
                    accB = mfma_f64(vZ2, b[(s * 5 + 4) * 64 + lane], accB);
                }
            }
        };
        load_fields(0, r0);
        load_basis(0);
        put_basis(0);
        __syncthreads();
        for (int ch = 0; ch < nch; ch += 2) {
            const bool n1 = ch + 1 < nch;
            if (n1) {
                load_fields(ch + 1, r1);
                load_basis(ch + 1);
            }
            compute(r0, bs[0]);
            if (n1) put_basis(1);
            __syncthreads();
            if (!n1) break;
            const bool n2 = ch + 2 < nch;
            if (n2) {
                load_fields(ch + 2, r0);
                load_basis(ch + 2);
            }
            compute(r1, bs[1]);
            if (n2) put_basis(0);
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

// Projection with the i-tile's whole weighted basis in LDS (NKS k-steps;
// the three bases Qpw, Qw, Rw of the primal, all five for the adjoint), for
// NKS in {16, 28, 40}.  A block of 8 waves owns one i-tile and a contiguous
// range of its (kz row, n-tile) tasks; wave w takes tasks w, w+8, ... of the
// range.  No barrier after the staging: each wave streams its field values
// through two register buffers of CK k-steps (the loads of chunk c+1, or of
// the next task's chunk 0, issued before the products of chunk c).  The
// products enter the accumulators in the same order as in k_force.
// Grid: nit * bpi blocks, block b -> i-tile b / bpi.  With nprob = 2 the
// task range covers two projections sharing the staged basis (fields U0 ->
// Fa0, Fb0, then U1 -> Fa1, Fb1).
constexpr int FR_THREADS = 512;

template <bool ADJ, int NKS>
struct ForceRowsShape {
    static constexpr int NB = ADJ ? 5 : 3;                   // bases staged per k-step
    static constexpr int NF = ADJ ? 5 : 3;                   // field values per k-step
    // chunk: the largest CK <= CKMAX with an even number of chunks
    static constexpr int CKMAX = ADJ ? 4 : 5;
    static constexpr int pick(int ck) { return (NKS % ck == 0 && (NKS / ck) % 2 == 0) ? ck : pick(ck - 1); }
    static constexpr int CK = pick(CKMAX), NCH = NKS / CK;
};

template <bool ADJ, int NKS>
__global__ __launch_bounds__(FR_THREADS) void k_force_rows(ForceArgs a, int bpi, int nprob,
                                                           const cplx *U0, cplx *Fa0, cplx *Fb0,
                                                           const cplx *U1, cplx *Fa1, cplx *Fb1)
{
    using SH = ForceRowsShape<ADJ, NKS>;
    constexpr int NB = SH::NB, NF = SH::NF, CK = SH::CK, NCH = SH::NCH;
    static_assert(NCH % 2 == 0 && CK >= 1, "chunks alternate between two buffers");
    extern __shared__ double bsf[];
    if (*a.abort_flag) return;
    const int it = blockIdx.x / bpi, lb = blockIdx.x - it * bpi;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int kq = lane >> 4, nl = lane & 15;
    const int Mx = a.Mx, q = a.qpts, nR = a.dimR, nQ = a.dimQ;
    const int nnt = (Mx + 7) / 8;
    const int ntask = (a.Nz - 1) * nnt;                       // every row but the Nyquist row
    const int nall = ntask * nprob * a.nm;
    const int t0 = lb * nall / bpi, t1 = (lb + 1) * nall / bpi;
    {
        // bases of the i-tile: fragF [nit][nksF][5][64] -> [NKS][NB][64]
        const double *frag = a.frag + (size_t) it * a.nks * 5 * 64;
        constexpr int TOT = NKS * NB * 64, PT = (TOT + FR_THREADS - 1) / FR_THREADS;
        double st[PT];
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const int e = min((int) threadIdx.x + FR_THREADS * i, TOT - 1);
            const int s = e / (NB * 64), r = e - s * NB * 64, m = r >> 6, l = r & 63;
            const int mm = ADJ ? m : (m == 2 ? 3 : m);       // primal: Qpw, Qw, Rw
            st[i] = frag[((size_t) s * 5 + mm) * 64 + l];
        }
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const int e = threadIdx.x + FR_THREADS * i;
            if (e < TOT) bsf[e] = st[i];
        }
    }
    __syncthreads();
    // staged bases per k-step: Qpw, Qw, Qppw, Rw, Rpw (adjoint); Qpw, Qw, Rw (primal)
    struct Task {
        int z, nt, xcol, part, p, m;
        size_t off;
    };
    auto task_of = [&](int t) {
        Task k;
        const int pr = t / ntask;                             // problem: member, projection
        k.m = pr / nprob;
        k.p = pr - k.m * nprob;
        t -= pr * ntask;
        const int zr = t / nnt;
        k.z = zr < a.Nz / 2 ? zr : zr + 1;
        k.nt = t - zr * nnt;
        k.xcol = min(k.nt * 8 + ncol_x(nl), Mx - 1);
        k.part = ncol_part(nl);
        k.off = (size_t) k.z * 5 * q * Mx * 2 + (size_t) k.xcol * 2;
        return k;
    };
    const size_t ss2 = (size_t) q * Mx * 2;                   // slot stride in doubles
    double r[2][NF][CK];
    auto load = [&](const Task &k, int ch, int buf) {
        const double *Ud = (const double *) ((k.p ? U1 : U0) + k.m * a.ms_u);
        const bool mean = k.z == 0 && k.xcol == 0;
        const int other = mean ? 0 : 1 - k.part;
#pragma unroll
        for (int s = 0; s < CK; ++s) {
            const int y = min((ch * CK + s) * 4 + kq, q - 1);
            const size_t e = k.off + (size_t) y * Mx * 2;
            r[buf][1][s] = Ud[e + other];
            r[buf][0][s] = Ud[e + ss2 + other];
            r[buf][2][s] = Ud[e + 2 * ss2 + k.part];
            if (ADJ) {
                r[buf][3 % NF][s] = Ud[e + 3 * ss2 + other];
                r[buf][4 % NF][s] = Ud[e + 4 * ss2 + other];
            }
        }
    };
    int t = t0 + wave;
    if (t < t1) load(task_of(t), 0, 0);
    while (t < t1) {
        const Task k = task_of(t);
        const bool mean = k.z == 0 && k.xcol == 0;
        const double kz = wave_kz(k.z, a.Nz, a.az);
        const double kx = wave_kx(k.xcol, a.ax), k2 = kz * kz + kx * kx;
        const double sg = k.part ? 1.0 : -1.0;
        mf4 accA = mf4_zero(), accB = mf4_zero();
        const int tn = t + 8;
        // products of chunk ch from buffer buf
        auto chunk = [&](int ch, int buf) {
            const double *bc = bsf + (size_t) ch * CK * NB * 64 + lane;
#pragma unroll
            for (int u = 0; u < CK; ++u) {
                const double hx = r[buf][0][u], hz = r[buf][1][u], hy = r[buf][2][u];
                double vP, vY, vP2, vZ = 0.0, vZ2 = 0.0;
                if (mean) {
                    vP = 0.0;
                    vY = 0.0;
                    vP2 = k.part ? hz : hx;
                    if (ADJ) vZ2 = k.part ? -r[buf][3 % NF][u] : r[buf][4 % NF][u];
                } else {
                    vP = sg * (kx * hx + kz * hz);
                    vY = -k2 * hy;
                    vP2 = sg * (kz * hx - kx * hz);
                    if (ADJ) {
                        vZ = -sg * (kz * r[buf][3 % NF][u] - kx * r[buf][4 % NF][u]);
                        vZ2 = sg * (kz * r[buf][4 % NF][u] + kx * r[buf][3 % NF][u]);
                    }
                }
                const double *b = bc + u * NB * 64;
                accA = mfma_f64(vP, b[(ADJ ? 0 : 0) * 64], accA);
                accB = mfma_f64(vP2, b[(ADJ ? 3 : 2) * 64], accB);
                accA = mfma_f64(vY, b[1 * 64], accA);
                if (ADJ) {
                    accA = mfma_f64(vZ, b[2 * 64], accA);
                    accB = mfma_f64(vZ2, b[4 * 64], accB);
                }
            }
        };
        // two chunks per iteration, one per buffer; the loads of the chunk
        // after next (or of the next task's first chunk) go into the buffer
        // just used
#pragma unroll 1
        for (int ch = 0; ch < NCH; ch += 2) {
            load(k, ch + 1, 1);
            __builtin_amdgcn_sched_barrier(0);
            chunk(ch, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (ch + 2 < NCH)
                load(k, ch + 2, 0);
            else if (tn < t1)
                load(task_of(tn), 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            chunk(ch + 1, 1);
            __builtin_amdgcn_sched_barrier(0);
        }
        const int i = it * 16 + nl;
        if (i < nR) {
            cplx *Fa = (k.p ? Fa1 : Fa0) + k.m * a.ms_fa, *Fb = (k.p ? Fb1 : Fb0) + k.m * a.ms_fb;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int x = k.nt * 8 + 2 * kq + h;
                if (x >= Mx) continue;
                if (i < nQ && !(k.z == 0 && x == 0))
                    Fa[((size_t) k.z * nQ + i) * Mx + x] = make_double2(accA[2 * h], accA[2 * h + 1]);
                Fb[((size_t) k.z * nR + i) * Mx + x] = make_double2(accB[2 * h], accB[2 * h + 1]);
            }
        }
        t = tn;
    }
}

// Parity form of the staged-basis projection.  The weighted basis functions
// have the parity of the Gauss-point mirror (Qw, Qppw, Rw of row i that of
// i, Qpw and Rpw the opposite), so over a mirrored pair of points
//     W(y) v(y) + W(-y) v(-y) = W(y) (v(y) + v(-y))  or  W(y) (v(y) - v(-y)),
// and a projection onto 16 even and 16 odd rows i takes the half domain's
// points once each: the products of one k-step feed both row tiles.  Each
// lane forms the pointwise combinations (P, Hy, P2, Z, Z2 as in k_force) at
// y and at its mirror, then their sum S and difference D:
//     even rows:  Qpw D_P + Qw S_Y + Qppw S_Z,   Rw S_P2 + Rpw D_Z2
//     odd rows:   Qpw S_P + Qw D_Y + Qppw D_Z,   Rw D_P2 + Rpw S_Z2
// The middle point of an odd qpts (y = 0) is its own mirror: its staged
// weights are halved, so that S = 2 v(0) gives the weight once, and the
// functions that D multiplies vanish there.  A block owns one pair of row
// tiles (even rows 2(16p + n), odd rows 2(16p + n) + 1), stages their
// weighted bases [NKS][2][NB][64] and runs tasks as k_force_rows.  Field
// values are read with buffer loads over the whole field array: the points
// past the half domain, whose staged weights are zero, read values in or
// past the array (zero past its end) and need no clamp.
template <bool ADJ, int NKS>
struct ForceParShape {
    static constexpr int NB = ADJ ? 5 : 3;
    static constexpr int NF = ADJ ? 5 : 3;
    // KS waves share a task's k-steps (partial sums added through LDS)
    // where the exchange area fits beside the staged bases
    static constexpr int XCH1 = 4 * 4 * 64;                  // doubles of one wave's partials
    static constexpr int KS = (NKS % 2 == 0 &&
                               (NKS * 2 * NB * 64 + (FR_THREADS / 128) * XCH1) * 8 <= 160 * 1024) ? 2 : 1;
    static constexpr int KW = NKS / KS;                      // k-steps per wave
    static constexpr int TPR = FR_THREADS / 64 / KS;         // tasks per round of the block
    static constexpr int XCH = KS > 1 ? TPR * (KS - 1) * XCH1 : 0;
    static constexpr int CKMAX = ADJ ? 2 : 5;
    static constexpr int pick(int ck) { return KW % ck == 0 ? ck : pick(ck - 1); }
    static constexpr int CK = pick(CKMAX), NCH = KW / CK;
    static constexpr size_t lds_bytes() { return sizeof(double) * ((size_t) NKS * 2 * NB * 64 + XCH); }
};
// P32: the product slots are FP32 complex (mixed precision; same element
// index), converted to FP64 on load
template <bool ADJ, int NKS, bool P32>
__global__ __launch_bounds__(FR_THREADS) void k_force_par(ForceArgs a, int bpi, int nprob,
                                                          const cplx *U0, cplx *Fa0, cplx *Fb0,
                                                          const cplx *U1, cplx *Fa1, cplx *Fb1)
{
    using SH = ForceParShape<ADJ, NKS>;
    constexpr int NB = SH::NB, NF = SH::NF, CK = SH::CK, NCH = SH::NCH, KS = SH::KS;
    constexpr unsigned ESZ = P32 ? 8 : 16;                   // bytes of a slot value
    constexpr int KW = SH::KW, TPR = SH::TPR, XCH1 = SH::XCH1;
    extern __shared__ double bsf[];
    if (*a.abort_flag) return;
    // XCD-aware: block b runs task range lb = (b / 8 / npair) 8 + b % 8 of
    // row-tile pair tp = (b / 8) % npair, so the pairs' blocks of one task
    // range (which read the same fields) share an XCD and its L2; bpi is a
    // multiple of 8
    const int npair = a.nit;
    const int rest = blockIdx.x >> 3, tp = rest % npair;
    const int lb = (rest / npair) * 8 + (blockIdx.x & 7);
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int kq = lane >> 4, nl = lane & 15;
    const int Mx = a.Mx, q = a.qpts, nR = a.dimR, nQ = a.dimQ;
    const int nnt = (Mx + 7) / 8;
    const int ntask = (a.Nz - 1) * nnt;
    const int nall = ntask * nprob * a.nm;
    const int t0 = lb * nall / bpi, t1 = (lb + 1) * nall / bpi;
    {
        // fragQ [npair][NKS][2][5][64] -> [NKS][2][NB][64]
        const double *frag = a.frag + (size_t) tp * NKS * 2 * 5 * 64;
        constexpr int TOT = NKS * 2 * NB * 64, PT = (TOT + FR_THREADS - 1) / FR_THREADS;
        double st[PT];
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const int e = min((int) threadIdx.x + FR_THREADS * i, TOT - 1);
            const int sp = e / (NB * 64), r = e - sp * NB * 64, m = r >> 6, l = r & 63;
            const int mm = ADJ ? m : (m == 2 ? 3 : m);       // primal: Qpw, Qw, Rw
            st[i] = frag[((size_t) sp * 5 + mm) * 64 + l];
        }
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const int e = threadIdx.x + FR_THREADS * i;
            if (e < TOT) bsf[e] = st[i];
        }
    }
    __syncthreads();
    struct Task {
        int z, nt, xcol, part, p, m;
        unsigned off;                                         // bytes: row z, column xcol
    };
    auto task_of = [&](int t) {
        Task k;
        const int pr = t / ntask;                             // problem: member, projection
        k.m = pr / nprob;
        k.p = pr - k.m * nprob;
        t -= pr * ntask;
        const int zr = t / nnt;
        k.z = zr < a.Nz / 2 ? zr : zr + 1;
        k.nt = t - zr * nnt;
        k.xcol = min(k.nt * 8 + ncol_x(nl), Mx - 1);
        k.part = ncol_part(nl);
        k.off = (unsigned) (((size_t) k.z * 5 * q * Mx + k.xcol) * ESZ);
        return k;
    };
    const unsigned ubytes = (unsigned) ((size_t) a.Nz * 5 * q * Mx * ESZ);
    const unsigned rowb = (unsigned) Mx * ESZ, slotb = (unsigned) q * rowb;
    // r[buf][0: y, 1: mirror][field][k-step]
    double r[2][2][NF][CK];
    auto load = [&](const Task &k, int ch, int buf) {
        __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *) ((k.p ? U1 : U0) + k.m * a.ms_u), 0, ubytes, 0x00020000);
        const bool mean = k.z == 0 && k.xcol == 0;
        const unsigned other = k.off + (mean ? 0 : 1 - k.part) * (ESZ / 2),
                       own = k.off + k.part * (ESZ / 2);
        // lane offsets for the point and the part, uniform slot offsets
        auto ld = [&](unsigned v, unsigned slot) {
            if constexpr (P32)
                return (double) __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int) v, (int) (slot * slotb), 0));
            else
                return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int) v, (int) (slot * slotb), 0));
        };
#pragma unroll
        for (int s = 0; s < CK; ++s) {
            const int y = (ch * CK + s) * 4 + kq, ym = q - 1 - y;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const unsigned e = (unsigned) (h ? ym : y) * rowb;
                r[buf][h][1][s] = ld(e + other, 0);
                r[buf][h][0][s] = ld(e + other, 1);
                r[buf][h][2][s] = ld(e + own, 2);
                if (ADJ) {
                    r[buf][h][3 % NF][s] = ld(e + other, 3);
                    r[buf][h][4 % NF][s] = ld(e + other, 4);
                }
            }
        }
    };
    // Rounds of TPR tasks: task t + wave / KS, k-steps [kb, kb + KW) of it
    // with kb = (wave % KS) KW.  The partial sums of the waves kb > 0 go
    // through LDS to the wave kb = 0, which adds them in order and stores.
    const int part = wave % KS, slot = wave / KS, kb = part * KW;
    double *xch = bsf + (size_t) NKS * 2 * NB * 64;
    for (int t = t0; t < t1; t += TPR) {
        const int tk = t + slot;
        const bool act = tk < t1;
        mf4 aAe = mf4_zero(), aBe = mf4_zero(), aAo = mf4_zero(), aBo = mf4_zero();
        const Task k = task_of(act ? tk : t0);
        if (act) {
            const bool mean = k.z == 0 && k.xcol == 0;
            const double kz = wave_kz(k.z, a.Nz, a.az);
            const double kx = wave_kx(k.xcol, a.ax), k2 = kz * kz + kx * kx;
            const double sg = k.part ? 1.0 : -1.0;
            auto chunk = [&](int ch, int buf) {
                const double *bc = bsf + (size_t) (kb + ch * CK) * 2 * NB * 64 + lane;
#pragma unroll
                for (int u = 0; u < CK; ++u) {
                    double vP[2], vY[2], vP2[2], vZ[2] = {0.0, 0.0}, vZ2[2] = {0.0, 0.0};
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const double hx = r[buf][h][0][u], hz = r[buf][h][1][u], hy = r[buf][h][2][u];
                        if (mean) {
                            vP[h] = 0.0;
                            vY[h] = 0.0;
                            vP2[h] = k.part ? hz : hx;
                            if (ADJ) vZ2[h] = k.part ? -r[buf][h][3 % NF][u] : r[buf][h][4 % NF][u];
                        } else {
                            vP[h] = sg * (kx * hx + kz * hz);
                            vY[h] = -k2 * hy;
                            vP2[h] = sg * (kz * hx - kx * hz);
                            if (ADJ) {
                                vZ[h] = -sg * (kz * r[buf][h][3 % NF][u] - kx * r[buf][h][4 % NF][u]);
                                vZ2[h] = sg * (kz * r[buf][h][4 % NF][u] + kx * r[buf][h][3 % NF][u]);
                            }
                        }
                    }
                    const double sP = vP[0] + vP[1], dP = vP[0] - vP[1];
                    const double sY = vY[0] + vY[1], dY = vY[0] - vY[1];
                    const double sP2 = vP2[0] + vP2[1], dP2 = vP2[0] - vP2[1];
                    const double *be = bc + u * 2 * NB * 64, *bo = be + NB * 64;
                    aAe = mfma_f64(dP, be[0], aAe);
                    aAo = mfma_f64(sP, bo[0], aAo);
                    aBe = mfma_f64(sP2, be[(ADJ ? 3 : 2) * 64], aBe);
                    aBo = mfma_f64(dP2, bo[(ADJ ? 3 : 2) * 64], aBo);
                    aAe = mfma_f64(sY, be[64], aAe);
                    aAo = mfma_f64(dY, bo[64], aAo);
                    if (ADJ) {
                        const double sZ = vZ[0] + vZ[1], dZ = vZ[0] - vZ[1];
                        const double sZ2 = vZ2[0] + vZ2[1], dZ2 = vZ2[0] - vZ2[1];
                        aAe = mfma_f64(sZ, be[2 * 64], aAe);
                        aAo = mfma_f64(dZ, bo[2 * 64], aAo);
                        aBe = mfma_f64(dZ2, be[4 * 64], aBe);
                        aBo = mfma_f64(sZ2, bo[4 * 64], aBo);
                    }
                }
            };
            // chunks of CK k-steps through two register buffers, the loads of
            // chunk ch + 1 issued before the products of chunk ch
            load(k, kb / CK, 0);
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                if (ch + 1 < NCH) load(k, kb / CK + ch + 1, (ch + 1) & 1);
                __builtin_amdgcn_sched_barrier(0);
                chunk(ch, ch & 1);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (KS > 1) {
            double *xw = xch + (size_t) (slot * (KS - 1) + part - 1) * XCH1 + lane;
            if (act && part > 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    xw[(0 * 4 + e) * 64] = aAe[e];
                    xw[(1 * 4 + e) * 64] = aBe[e];
                    xw[(2 * 4 + e) * 64] = aAo[e];
                    xw[(3 * 4 + e) * 64] = aBo[e];
                }
            }
            __syncthreads();
            if (act && part == 0) {
#pragma unroll
                for (int p2 = 1; p2 < KS; ++p2) {
                    const double *xr = xch + (size_t) (slot * (KS - 1) + p2 - 1) * XCH1 + lane;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        aAe[e] += xr[(0 * 4 + e) * 64];
                        aBe[e] += xr[(1 * 4 + e) * 64];
                        aAo[e] += xr[(2 * 4 + e) * 64];
                        aBo[e] += xr[(3 * 4 + e) * 64];
                    }
                }
            }
        }
        if (act && part == 0) {
            cplx *Fa = (k.p ? Fa1 : Fa0) + k.m * a.ms_fa, *Fb = (k.p ? Fb1 : Fb0) + k.m * a.ms_fb;
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                const int i = 2 * (tp * 16 + nl) + par;
                if (i >= nR) continue;
                const mf4 &A = par ? aAo : aAe, &B = par ? aBo : aBe;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int x = k.nt * 8 + 2 * kq + h;
                    if (x >= Mx) continue;
                    if (i < nQ && !(k.z == 0 && x == 0))
                        Fa[((size_t) k.z * nQ + i) * Mx + x] = make_double2(A[2 * h], A[2 * h + 1]);
                    Fb[((size_t) k.z * nR + i) * Mx + x] = make_double2(B[2 * h], B[2 * h + 1]);
                }
            }
        }
        // the exchange area is reused by the next round
        if (KS > 1) __syncthreads();
    }
}

template <bool ADJ, int NKS, bool P32>
static void launch_force_par(const ForceArgs &a, int npair, int ntask, int nprob, const cplx *U0,
                             cplx *Fa0, cplx *Fb0, const cplx *U1, cplx *Fa1, cplx *Fb1,
                             hipStream_t st)
{
    using SH = ForceParShape<ADJ, NKS>;
    const size_t lds = SH::lds_bytes();
    static bool attr = false;
    if (!attr) {
        (void) hipFuncSetAttribute((const void *) k_force_par<ADJ, NKS, P32>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        attr = true;
    }
    // blocks per pair: a multiple of 8 (XCD-aware mapping), one block per CU
    const int cap = tuning().force_blocks > 0 ? tuning().force_blocks : 256 / npair / 8 * 8;
    const int bpi = std::max(8, std::min((ntask * nprob * a.nm + 7) / 8 * 8, cap));
    hipLaunchKernelGGL((k_force_par<ADJ, NKS, P32>), dim3(npair * bpi), dim3(FR_THREADS), lds, st, a, bpi,
                       nprob, U0, Fa0, Fb0, U1, Fa1, Fb1);
}

template <bool ADJ, int NKS>
static void launch_force_rows(const ForceArgs &a, int nit, int ntask, int nprob, const cplx *U0,
                              cplx *Fa0, cplx *Fb0, const cplx *U1, cplx *Fa1, cplx *Fb1,
                              hipStream_t st)
{
    using SH = ForceRowsShape<ADJ, NKS>;
    const size_t lds = sizeof(double) * NKS * SH::NB * 64;
    static bool attr = false;
    if (!attr) {
        (void) hipFuncSetAttribute((const void *) k_force_rows<ADJ, NKS>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        attr = true;
    }
    // blocks per i-tile: one 8-wave block per CU in total (the registers of
    // the wave's two field buffers allow no second block), never more than
    // the tasks
    const int bpi = std::max(1, std::min(ntask * nprob * a.nm, 256 / nit));
    hipLaunchKernelGGL((k_force_rows<ADJ, NKS>), dim3(nit * bpi), dim3(FR_THREADS), lds, st, a, bpi,
                       nprob, U0, Fa0, Fb0, U1, Fa1, Fb1);
}

// one or two projections (U1 null: one); the staged-basis kernel runs both in
// one launch, the chunked kernel one launch each
void launch_force2(Context &c, int adj, const cplx *U0, cplx *Fa0, cplx *Fb0, const cplx *U1,
                   cplx *Fa1, cplx *Fb1)
{
    const Dims &d = c.d;
    const int nm = c.nrun;
    const size_t ms_u = mstride(c, U0), ms_fa = mstride(c, Fa0), ms_fb = mstride(c, Fb0);
    if (U1 && (mstride(c, U1) != ms_u || mstride(c, Fa1) != ms_fa || mstride(c, Fb1) != ms_fb)) {
        // the two problems' arrays have other member strides: two launches
        launch_force2(c, adj, U0, Fa0, Fb0, nullptr, nullptr, nullptr);
        launch_force2(c, adj, U1, Fa1, Fb1, nullptr, nullptr, nullptr);
        return;
    }
    ForceArgs a{c.dFragF, c.dKx, c.dKz, d.Nz, d.Mx, d.qpts, d.dimR, d.dimQ, c.nksF, c.nit,
                2. * M_PI / d.Lx, 2. * M_PI / d.Lz, c.abort_flag, nm, ms_u, ms_fa, ms_fb};
    dim3 grid(xcd_grid(c.nit * d.Nz), nm);
    const int nprob = U1 ? 2 : 1;
    // mixed precision: FP32 product slots (the z-forward transforms wrote them)
    const bool p32 = prods32(c);
    double bytes = nm * nprob * (double) (d.Nz - 1) * d.Mx *
                   ((p32 ? 8.0 : 16.0) * (adj ? 5.0 : 3.0) * d.qpts + 16.0 * (d.dimQ + d.dimR));
    const hipStream_t st = c.stream;
    const int nks = c.nksF, nit = c.nit, ntask = (d.Nz - 1) * ((d.Mx + 7) / 8);
    const int nksq = c.nksQ, npair = c.npair;
    const bool par = c.ypar && (nksq == 8 || nksq == 16 || nksq == 20 || nksq == 28);
    ForceArgs ap = a;
    ap.frag = c.dFragQ;
    ap.nks = nksq;
    ap.nit = npair;
    auto go = [=]() {
        if (par) {
#define SC_FORCE_PAR(A, N)                                                                   \
    (p32 ? launch_force_par<A, N, true>(ap, npair, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st) \
         : launch_force_par<A, N, false>(ap, npair, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st))
            if (adj) {
                if (nksq == 8) SC_FORCE_PAR(true, 8);
                else if (nksq == 16) SC_FORCE_PAR(true, 16);
                else if (nksq == 20) SC_FORCE_PAR(true, 20);
                else SC_FORCE_PAR(true, 28);
            } else {
                if (nksq == 8) SC_FORCE_PAR(false, 8);
                else if (nksq == 16) SC_FORCE_PAR(false, 16);
                else if (nksq == 20) SC_FORCE_PAR(false, 20);
                else SC_FORCE_PAR(false, 28);
            }
#undef SC_FORCE_PAR
            return;
        }
        // staged-basis kernel for the k-step counts it is built for
        if (nks == 40 || nks == 28 || nks == 16) {
            if (adj) {
                if (nks == 40)
                    launch_force_rows<true, 40>(a, nit, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st);
                else if (nks == 28)
                    launch_force_rows<true, 28>(a, nit, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st);
                else
                    launch_force_rows<true, 16>(a, nit, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st);
            } else {
                if (nks == 40)
                    launch_force_rows<false, 40>(a, nit, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st);
                else if (nks == 28)
                    launch_force_rows<false, 28>(a, nit, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st);
                else
                    launch_force_rows<false, 16>(a, nit, ntask, nprob, U0, Fa0, Fb0, U1, Fa1, Fb1, st);
            }
            return;
        }
        for (int p = 0; p < nprob; ++p) {
            const cplx *U = p ? U1 : U0;
            cplx *Fa = p ? Fa1 : Fa0, *Fb = p ? Fb1 : Fb0;
            if (adj)
                hipLaunchKernelGGL(k_force<true>, grid, dim3(256), 0, st, a, U, Fa, Fb);
            else
                hipLaunchKernelGGL(k_force<false>, grid, dim3(256), 0, st, a, U, Fa, Fb);
        }
    };
    prof_begin(c, st, 5, bytes);
    go();
    prof_end(c, st, 5, go);
}

void launch_force(Context &c, const cplx *Usrc, int adj, cplx *Fa, cplx *Fb)
{
    launch_force2(c, adj, Usrc, Fa, Fb, nullptr, nullptr, nullptr);
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
    int nblk;                          // column blocks of SBT columns
    int nyq_mean;                      // the mean mode runs in the Nyquist blocks' first (1) or in an extra block (0)
    int mirror;                        // factor images of rows z > Nz/2 are those of row Nz - z
    double dtc, dtd;                   // dt c_k, dt d_{k+1}
    double ce0[2], cn[2];              // explicit-operator coefficients per field (stage 0, next)
    double target;                     // mean-mode flux target
    double target_alt;                 // ... of the members m with bit m of alt_mask set
    unsigned alt_mask;
    const int *abort_flag;
    int *err;                          // device error word (SC_KERR_* bits)
    // member strides (grid z = member) of C, Cs, Fa, Fb, Fa2, Fb2, Src
    size_t ms_c, ms_cs, ms_fa, ms_fb, ms_fa2, ms_fb2, ms_src;
    // optional right-hand-side source of the stage (C layout), added after
    // the explicit part and the forcing, before the solve (all modes)
    const cplx *Src;
    // adjoint only: the mean column as the exact transpose of the tangent's
    // flux correction: at stages 0 and 1 the right-hand side's real alpha
    // loses w1 (add_mass . rhs) / 2 before the solve, and no correction
    // follows the solve (the host loop applies P at each step's start)
    int exact_mean;
    const double *add_mass;            // Rs^{-1} w1
    int nm;                            // members of the launch (grid z: groups of them)
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
// v: right-hand side in, solution out, row r at v[r*vr].
//
// Forward sweep, row s: x_s = v_s - L(s,DL) x_{s-DL} - ... - L(s,1) x_{s-1},
// the terms applied farthest first (the reference's order).  It runs skewed:
// the iteration that produces x_{r-1} applies it to the rows r .. r+DL-1 at
// once (term m = k+1 of row r+k), so the DL multiply-adds of an iteration are
// independent and only one of them waits for the previous iteration.  Every
// row still sees its terms in the same order, so the result is the same as
// the row-by-row sweep.  P[k] holds row r+k with its terms DL .. k+2 applied.
//
// Operands stream from LDS through two register buffers of two rows each,
// used in turn by a body of four rows: the loads for rows r+4, r+5 go into
// the buffer rows r, r+1 have just used, issued before the arithmetic of rows
// r+2, r+3 (a scheduling barrier keeps them there), with no register copies
// that would wait for them.  Row indices are clamped into the chain; values
// of rows past the end are never stored.
template <int DL>
struct FwdOps {
    double c[2][DL];   // c[i][k] = L(r+i+k, k+1)
    double2 vn[2];     // right-hand side of row r+i+DL-1
};

template <int DL>
__device__ __forceinline__ void fwd_load(FwdOps<DL> &o, const double *f, int fr, int fj,
                                         const double2 *v, int vr, int r, int nr)
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int k = 0; k < DL; ++k) o.c[i][k] = f[min(r + i + k, nr - 1) * fr + (DL - 1 - k) * fj];
        o.vn[i] = v[min(r + i + DL - 1, nr - 1) * vr];
    }
}

template <int DL>
__device__ __forceinline__ void fwd_step(const double *c, double2 vn, double2 *P, double2 &xm1,
                                         double2 *v, int vr, int r)
{
    double2 x = P[0];
    x.x -= c[0] * xm1.x;
    x.y -= c[0] * xm1.y;
#pragma unroll
    for (int k = 1; k < DL - 1; ++k) {
        P[k - 1].x = P[k].x - c[k] * xm1.x;
        P[k - 1].y = P[k].y - c[k] * xm1.y;
    }
    P[DL - 2].x = vn.x - c[DL - 1] * xm1.x;
    P[DL - 2].y = vn.y - c[DL - 1] * xm1.y;
    v[r * vr] = x;
    xm1 = x;
}

template <int DL>
__device__ __forceinline__ void chain_forward(const double *f, int fr, int fj, double2 *v, int vr,
                                              int nr)
{
    double2 P[DL - 1];
#pragma unroll
    for (int k = 0; k < DL - 1; ++k) P[k] = v[min(k, nr - 1) * vr];
    double2 xm1 = make_double2(0.0, 0.0);
    FwdOps<DL> X, Y;
    fwd_load<DL>(X, f, fr, fj, v, vr, 0, nr);
    fwd_load<DL>(Y, f, fr, fj, v, vr, 2, nr);
    int r = 0;
    for (; r + 4 <= nr; r += 4) {
        fwd_step<DL>(X.c[0], X.vn[0], P, xm1, v, vr, r);
        fwd_step<DL>(X.c[1], X.vn[1], P, xm1, v, vr, r + 1);
        fwd_load<DL>(X, f, fr, fj, v, vr, r + 4, nr);
        __builtin_amdgcn_sched_barrier(0);
        fwd_step<DL>(Y.c[0], Y.vn[0], P, xm1, v, vr, r + 2);
        fwd_step<DL>(Y.c[1], Y.vn[1], P, xm1, v, vr, r + 3);
        fwd_load<DL>(Y, f, fr, fj, v, vr, r + 6, nr);
        __builtin_amdgcn_sched_barrier(0);
    }
    // fewer than four rows left: X holds rows r, r+1, Y rows r+2, r+3
    if (r < nr) fwd_step<DL>(X.c[0], X.vn[0], P, xm1, v, vr, r);
    if (r + 1 < nr) fwd_step<DL>(X.c[1], X.vn[1], P, xm1, v, vr, r + 1);
    if (r + 2 < nr) fwd_step<DL>(Y.c[0], Y.vn[0], P, xm1, v, vr, r + 2);
}

// Backward sweep, row r: x_r = (v_r - U(r,1) x_{r+1} - ... - U(r,DL) x_{r+DL})
// / U(r,r), nearest term first (the reference's order).  Same two-buffer
// stream of operands as the forward sweep, rows descending.
template <int DL>
struct BwdOps {
    double u[2][DL + 1];   // U(r-i, r-i+1..r-i+DL), then 1/U(r-i, r-i)
    double2 v[2];          // right-hand side of row r-i
};

template <int DL>
__device__ __forceinline__ void bwd_load(BwdOps<DL> &o, const double *f, int fr, int fj,
                                         const double2 *v, int vr, int r)
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rr = max(r - i, 0);
#pragma unroll
        for (int k = 1; k <= DL; ++k) o.u[i][k - 1] = f[rr * fr + (DL + k) * fj];
        o.u[i][DL] = f[rr * fr + DL * fj];
        o.v[i] = v[rr * vr];
    }
}

template <int DL>
__device__ __forceinline__ void bwd_step(const double *u, double2 x, double2 *w, double2 *v,
                                         int vr, int r)
{
#pragma unroll
    for (int k = 1; k <= DL; ++k) {
        x.x -= u[k - 1] * w[k - 1].x;
        x.y -= u[k - 1] * w[k - 1].y;
    }
    x.x *= u[DL];
    x.y *= u[DL];
#pragma unroll
    for (int m = DL - 1; m > 0; --m) w[m] = w[m - 1];
    w[0] = x;
    v[r * vr] = x;
}

template <int DL>
__device__ __forceinline__ void chain_backward(const double *f, int fr, int fj, double2 *v, int vr,
                                               int nr)
{
    double2 w[DL];
#pragma unroll
    for (int m = 0; m < DL; ++m) w[m] = make_double2(0.0, 0.0);
    BwdOps<DL> X, Y;
    bwd_load<DL>(X, f, fr, fj, v, vr, nr - 1);
    bwd_load<DL>(Y, f, fr, fj, v, vr, nr - 3);
    int r = nr - 1;
    // rows below r - 3 still hold forward values until their own step
    for (; r >= 3; r -= 4) {
        bwd_step<DL>(X.u[0], X.v[0], w, v, vr, r);
        bwd_step<DL>(X.u[1], X.v[1], w, v, vr, r - 1);
        bwd_load<DL>(X, f, fr, fj, v, vr, r - 4);
        __builtin_amdgcn_sched_barrier(0);
        bwd_step<DL>(Y.u[0], Y.v[0], w, v, vr, r - 2);
        bwd_step<DL>(Y.u[1], Y.v[1], w, v, vr, r - 3);
        bwd_load<DL>(Y, f, fr, fj, v, vr, r - 6);
        __builtin_amdgcn_sched_barrier(0);
    }
    // fewer than four rows left: X holds rows r, r-1, Y rows r-2, r-3
    if (r >= 0) bwd_step<DL>(X.u[0], X.v[0], w, v, vr, r);
    if (r >= 1) bwd_step<DL>(X.u[1], X.v[1], w, v, vr, r - 1);
    if (r >= 2) bwd_step<DL>(Y.u[0], Y.v[0], w, v, vr, r - 2);
}

// inlined at every call site: the pointers then stay LDS pointers (a call
// would pass them as generic addresses, and every access would go through
// the flat path)
template <int DL>
__device__ __forceinline__ void chain_subst(const double *f, int fr, int fj, double2 *v, int vr,
                                            int nr)
{
    if (nr <= 0) return;
    chain_forward<DL>(f, fr, fj, v, vr, nr);
    chain_backward<DL>(f, fr, fj, v, vr, nr);
}

// The same sweeps with NB operand buffers of two rows (prefetch distance
// 2 NB - 2 rows): the chain waits less on LDS latency.
template <int DL, int NB>
__device__ __forceinline__ void chain_forward_n(const double *f, int fr, int fj, double2 *v, int vr,
                                                int nr)
{
    double2 P[DL - 1];
#pragma unroll
    for (int k = 0; k < DL - 1; ++k) P[k] = v[min(k, nr - 1) * vr];
    double2 xm1 = make_double2(0.0, 0.0);
    FwdOps<DL> B[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) fwd_load<DL>(B[b], f, fr, fj, v, vr, 2 * b, nr);
    int r = 0;
    for (; r + 2 * NB <= nr; r += 2 * NB) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            fwd_step<DL>(B[b].c[0], B[b].vn[0], P, xm1, v, vr, r + 2 * b);
            fwd_step<DL>(B[b].c[1], B[b].vn[1], P, xm1, v, vr, r + 2 * b + 1);
            fwd_load<DL>(B[b], f, fr, fj, v, vr, r + 2 * b + 2 * NB, nr);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // fewer than 2 NB rows left: B[b] holds rows r + 2b, r + 2b + 1
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (r + 2 * b < nr) fwd_step<DL>(B[b].c[0], B[b].vn[0], P, xm1, v, vr, r + 2 * b);
        if (r + 2 * b + 1 < nr) fwd_step<DL>(B[b].c[1], B[b].vn[1], P, xm1, v, vr, r + 2 * b + 1);
    }
}

// Real-valued forms of the sweeps: the real and imaginary parts of a
// complex right-hand side are independent chains with the same factors, so
// two lanes share a column (one part each) and every step issues half the
// multiply-adds of the complex form.  v: one part of the vector (row r at
// v[r*vr]); operands stream through NB buffers of two rows.
template <int DL>
struct FwdOpsR {
    double c[2][DL];
    double vn[2];
};

template <int DL>
__device__ __forceinline__ void fwd_load_r(FwdOpsR<DL> &o, const double *f, int fr, int fj,
                                           const double *v, int vr, int r, int nr)
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int k = 0; k < DL; ++k) o.c[i][k] = f[min(r + i + k, nr - 1) * fr + (DL - 1 - k) * fj];
        o.vn[i] = v[min(r + i + DL - 1, nr - 1) * vr];
    }
}

template <int DL>
__device__ __forceinline__ void fwd_step_r(const double *c, double vn, double *P, double &xm1,
                                           double *v, int vr, int r)
{
    const double x = P[0] - c[0] * xm1;
#pragma unroll
    for (int k = 1; k < DL - 1; ++k) P[k - 1] = P[k] - c[k] * xm1;
    P[DL - 2] = vn - c[DL - 1] * xm1;
    v[r * vr] = x;
    xm1 = x;
}

template <int DL, int NB>
__device__ __forceinline__ void chain_forward_r(const double *f, int fr, int fj, double *v, int vr,
                                                int nr)
{
    double P[DL - 1];
#pragma unroll
    for (int k = 0; k < DL - 1; ++k) P[k] = v[min(k, nr - 1) * vr];
    double xm1 = 0.0;
    FwdOpsR<DL> B[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) fwd_load_r<DL>(B[b], f, fr, fj, v, vr, 2 * b, nr);
    int r = 0;
    for (; r + 2 * NB <= nr; r += 2 * NB) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            fwd_step_r<DL>(B[b].c[0], B[b].vn[0], P, xm1, v, vr, r + 2 * b);
            fwd_step_r<DL>(B[b].c[1], B[b].vn[1], P, xm1, v, vr, r + 2 * b + 1);
            fwd_load_r<DL>(B[b], f, fr, fj, v, vr, r + 2 * b + 2 * NB, nr);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (r + 2 * b < nr) fwd_step_r<DL>(B[b].c[0], B[b].vn[0], P, xm1, v, vr, r + 2 * b);
        if (r + 2 * b + 1 < nr) fwd_step_r<DL>(B[b].c[1], B[b].vn[1], P, xm1, v, vr, r + 2 * b + 1);
    }
}

template <int DL>
struct BwdOpsR {
    double u[2][DL + 1];
    double v[2];
};

template <int DL>
__device__ __forceinline__ void bwd_load_r(BwdOpsR<DL> &o, const double *f, int fr, int fj,
                                           const double *v, int vr, int r)
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rr = max(r - i, 0);
#pragma unroll
        for (int k = 1; k <= DL; ++k) o.u[i][k - 1] = f[rr * fr + (DL + k) * fj];
        o.u[i][DL] = f[rr * fr + DL * fj];
        o.v[i] = v[rr * vr];
    }
}

template <int DL>
__device__ __forceinline__ void bwd_step_r(const double *u, double x, double *w, double *v, int vr,
                                           int r)
{
#pragma unroll
    for (int k = 1; k <= DL; ++k) x -= u[k - 1] * w[k - 1];
    x *= u[DL];
#pragma unroll
    for (int m = DL - 1; m > 0; --m) w[m] = w[m - 1];
    w[0] = x;
    v[r * vr] = x;
}

// Rows of the factor image whose U parts arrive while the chain runs: chunk
// c holds the rows [nr - (c+1) cr, nr - c cr) (last rows first); ready[c]
// counts the waves that have stored their share of it.
struct StagedRows {
    const int *ready;
    int cr, nchunk, nwriters;
    int *err;          // device error word (SC_KERR_* bits)
};

// Polls of one chunk's counter before the wait gives up: each poll sleeps
// 64 clocks, so the limit is ~30 ms, against waits of a few microseconds.
constexpr int SOLVE_SPIN_LIMIT = 1 << 20;

// wait until the U parts of rows >= row are in LDS (chunks in order).  A
// counter that never reaches the writer count (a staging wave that did not
// signal) sets SC_KERR_SOLVE_WAIT in the error word and ends the wait, so
// the kernel still exits; the host reports the error after the run.
__device__ __forceinline__ void wait_rows(const StagedRows &sr, int row, int &low, int &next)
{
    while (row < low && next < sr.nchunk) {
        int spins = 0;
        while (__hip_atomic_load(&sr.ready[next], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) <
               sr.nwriters) {
            if (++spins > SOLVE_SPIN_LIMIT) {
                __hip_atomic_fetch_or(sr.err, SC_KERR_SOLVE_WAIT, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
                next = sr.nchunk;
                low = 0;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        ++next;
        low = max(low - sr.cr, 0);
    }
}

// real-valued backward sweep with NB operand buffers, the operand stream
// waiting for staged rows
template <int DL, int NB>
__device__ __forceinline__ void chain_backward_staged(const double *f, int fr, int fj, double *v,
                                                      int vr, int nr, const StagedRows &sr)
{
    int low = nr, next = 0;
    double w[DL];
#pragma unroll
    for (int m = 0; m < DL; ++m) w[m] = 0.0;
    BwdOpsR<DL> B[NB];
    wait_rows(sr, max(nr - 2 * NB, 0), low, next);
#pragma unroll
    for (int b = 0; b < NB; ++b) bwd_load_r<DL>(B[b], f, fr, fj, v, vr, nr - 1 - 2 * b);
    int r = nr - 1;
    // B[b] holds rows r - 2b, r - 2b - 1
    for (; r >= 2 * NB - 1; r -= 2 * NB) {
        wait_rows(sr, max(r - 4 * NB + 1, 0), low, next);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            bwd_step_r<DL>(B[b].u[0], B[b].v[0], w, v, vr, r - 2 * b);
            bwd_step_r<DL>(B[b].u[1], B[b].v[1], w, v, vr, r - 2 * b - 1);
            bwd_load_r<DL>(B[b], f, fr, fj, v, vr, r - 2 * b - 2 * NB);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (r - 2 * b >= 0) bwd_step_r<DL>(B[b].u[0], B[b].v[0], w, v, vr, r - 2 * b);
        if (r - 2 * b - 1 >= 0) bwd_step_r<DL>(B[b].u[1], B[b].v[1], w, v, vr, r - 2 * b - 1);
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

// rows of one thread in a block of SOLVE_THREADS: r = tid / SBT + t * (SOLVE_THREADS / SBT),
// t < VR (a chain has at most VR * SOLVE_THREADS / SBT rows)
constexpr int SOLVE_THREADS = 256;
constexpr int SOLVE_VR_MAX = 4;
// factor values in flight per thread and batch (16-byte loads)
constexpr int SOLVE_FBAT = 8;

// One (field, parity) of SBT consecutive columns, for the members m0 ..
// m0 + nmb - 1 (nmb <= MB), which all solve with the same operator.  The
// factors of the block are one contiguous image [nr][W][SBT] (laid out so at
// init) copied to LDS once with 16-byte loads and shared by the members;
// every thread owns the column tid % SBT, so its rows are tid / SBT + t *
// (SOLVE_THREADS / SBT).  lds: [nr][W][SBT] factors, then [MB][nr][SBT]
// right-hand sides, then [MB][nr][SBT] stage-0 copies of C.  The chain wave
// runs 2 SBT lanes per member (real and imaginary part of each column): with
// several members one staged image feeds up to 64 chains at once.
template <int DL, int SBT, int VR, int MB>
__device__ void solve_block(const SolveArgs &a, int var, int par, int blk, int m0, int nmb,
                            cplx *C, cplx *Cs, double *lds)
{
    constexpr int W = 2 * DL + 1;
    constexpr int RSTEP = SOLVE_THREADS / SBT;
    static_assert(2 * SBT * MB <= 64, "one chain lane per member, part and column");
    const int Mx = a.Mx, nR = a.dimR, nQ = a.dimQ;
    const int N = var == 0 ? nQ : nR;
    const int nr = (N - par + 1) / 2, nrmax = (N + 1) / 2;
    const int tid = threadIdx.x;
    double *lf = lds;
    double2 *vec = (double2 *) (lds + (size_t) nr * W * SBT);
    double2 *raw = vec + MB * nr * SBT;
    const double ce0 = a.ce0[var], cn = a.cn[var];
    // the members' arrays (member m0 + mb; uniform across the block)
    auto Cm = [&](int mb) { return C + (size_t) (m0 + mb) * a.ms_c; };
    auto Csm = [&](int mb) { return Cs + (size_t) (m0 + mb) * a.ms_cs; };
    auto Fm = [&](int mb) {
        return var == 0 ? a.Fa + (size_t) (m0 + mb) * a.ms_fa : a.Fb + (size_t) (m0 + mb) * a.ms_fb;
    };
    auto F2m = [&](int mb) {
        return var == 0 ? a.Fa2 + (size_t) (m0 + mb) * a.ms_fa2 : a.Fb2 + (size_t) (m0 + mb) * a.ms_fb2;
    };
    auto Sm = [&](int mb) { return a.Src + (size_t) (m0 + mb) * a.ms_src; };

    constexpr int NCHU = 4;                 // chunks of rows whose U parts stream in
    __shared__ int chain_wave[1];
    __shared__ int ready[NCHU];
    if (tid == 0) chain_wave[0] = 0;
    if (tid < NCHU) ready[tid] = 0;
    __syncthreads();
    {
        // HW_ID: SIMD in bits 5:4, workgroup slot in bits 19:16
        const unsigned hw = __builtin_amdgcn_s_getreg(4 | (0 << 6) | (31 << 11));
        if ((tid & 63) == 0 && ((hw >> 4) & 3) == ((hw >> 16) & 3)) chain_wave[0] = tid >> 6;
    }
    const int cc = tid % SBT, rt = tid / SBT;
    const ColumnOf co = column_of(a, blk * SBT + cc);
    const double kx = a.Kx[co.x], kz = a.Kz[co.z], k2 = kz * kz + kx * kx;
    const size_t cbase = (((size_t) co.z * 2 + var) * nR + par) * Mx + co.x;   // C row par
    const size_t fbase = ((size_t) co.z * N + par) * Mx + co.x;                 // F row par
    const size_t rs = (size_t) 2 * Mx;                                          // next chain row

    // right-hand side rows: member 0's loads first (before the factor copy);
    // the other members' after it, one member at a time
    double2 cv[VR], fv[VR], sv[VR];
    auto load_rows = [&](int mb) {
#pragma unroll
        for (int t = 0; t < VR; ++t) {
            const int r = min(rt + t * RSTEP, nr - 1);
            cv[t] = Cm(mb)[cbase + r * rs];
            fv[t] = Fm(mb)[fbase + r * rs];
            sv[t] = a.Src ? Sm(mb)[cbase + r * rs] : make_double2(0.0, 0.0);
        }
    };
    auto store_rows = [&](int mb) {
#pragma unroll
        for (int t = 0; t < VR; ++t) {
            const int r = rt + t * RSTEP;
            if (r >= nr || !co.chain) continue;
            if (a.k == 0) {
                raw[(mb * nr + r) * SBT + cc] = cv[t];
            } else {
                double2 v = cv[t];
                v.x += a.dtc * fv[t].x;
                v.y += a.dtc * fv[t].y;
                v.x += sv[t].x;
                v.y += sv[t].y;
                vec[(mb * nr + r) * SBT + cc] = v;
            }
        }
    };
    load_rows(0);
    // The factor image ([row][W][SBT], 16-byte loads): the L parts (columns
    // j < DL) of every row before the barrier, for the forward sweep; the U
    // parts (j >= DL) after it, last rows first in NCHU chunks, each wave
    // signalling a chunk once its share is in LDS, so that the backward
    // sweep starts while the upper rows still stream in.
    constexpr int RPR = W * SBT / 2, LPR = DL * SBT / 2, UPR = RPR - LPR;
    int fb = blk;
    if (a.mirror) {
        // factor image of the mirror row Nz - z for z > Nz/2
        const int z = blk * SBT / Mx;
        if (z > a.Nz / 2) fb = blk - (2 * z - a.Nz) * (Mx / SBT);
    }
    const double2 *src = (const double2 *) (var == 0 ? a.La : a.Lb) +
                         ((size_t) fb * 2 + par) * nrmax * RPR;
    double2 *dst = (double2 *) lf;
    {
        const int nl = nr * LPR;
        for (int b0 = 0; b0 < nl; b0 += SOLVE_FBAT * SOLVE_THREADS) {
            double2 t[SOLVE_FBAT];
            int o[SOLVE_FBAT];
#pragma unroll
            for (int i = 0; i < SOLVE_FBAT; ++i) {
                const int e = min(b0 + tid + i * SOLVE_THREADS, nl - 1), r = e / LPR;
                o[i] = r * RPR + (e - r * LPR);
                t[i] = src[o[i]];
            }
#pragma unroll
            for (int i = 0; i < SOLVE_FBAT; ++i)
                if (b0 + tid + i * SOLVE_THREADS < nl) dst[o[i]] = t[i];
        }
    }
    store_rows(0);
    // member 0's forcing rows kept for the stage-0 operator (with one member
    // they stay where they are)
    constexpr int KV = MB > 1 ? VR : 1;
    double2 fv0[KV], sv0[KV];
    if constexpr (MB > 1) {
#pragma unroll
        for (int t = 0; t < VR; ++t) {
            fv0[t] = fv[t];
            sv0[t] = sv[t];
        }
    }
    for (int mb = 1; mb < nmb; ++mb) {
        load_rows(mb);
        store_rows(mb);
    }
    __syncthreads();
    if (a.k == 0) {
        // E C + dt c_0 F with the stage-0 explicit operator
        if constexpr (MB <= 2) {
            // each band entry formed once for the block's members; their
            // forcing rows are still in registers (member 0's in fv0/sv0
            // when MB > 1, the last member's in fv/sv)
#pragma unroll
            for (int t = 0; t < VR; ++t) {
                const int r = rt + t * RSTEP, i = par + 2 * r;
                if (r >= nr || !co.chain) continue;
                double2 acc[MB];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) acc[mb] = make_double2(0.0, 0.0);
#pragma unroll
                for (int m = -DL; m <= DL; ++m)
                    if (r + m >= 0 && r + m < nr) {
                        const double ev = op_entry(a, var, i, m, k2, ce0);
#pragma unroll
                        for (int mb = 0; mb < MB; ++mb)
                            if (mb < nmb) {
                                const double2 y = raw[(mb * nr + r + m) * SBT + cc];
                                acc[mb].x += ev * y.x;
                                acc[mb].y += ev * y.y;
                            }
                    }
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    if (mb >= nmb) continue;
                    double2 f = fv[t], sf = sv[t];
                    if constexpr (MB > 1) {
                        if (mb == 0) {
                            f = fv0[t];
                            sf = sv0[t];
                        }
                    }
                    acc[mb].x += a.dtc * f.x;
                    acc[mb].y += a.dtc * f.y;
                    acc[mb].x += sf.x;
                    acc[mb].y += sf.y;
                    vec[(mb * nr + r) * SBT + cc] = acc[mb];
                }
            }
        } else {
            // member 0's forcing rows still in registers, the others'
            // loaded again (one member's at a time)
            for (int mb = 0; mb < nmb; ++mb) {
                if (mb > 0) load_rows(mb);
#pragma unroll
                for (int t = 0; t < VR; ++t) {
                    const int r = rt + t * RSTEP, i = par + 2 * r;
                    if (r >= nr || !co.chain) continue;
                    double2 f = fv[t], sf = sv[t];
                    if (mb == 0) {
                        f = fv0[t];
                        sf = sv0[t];
                    }
                    const double2 *rm = raw + mb * nr * SBT;
                    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                    for (int m = -DL; m <= DL; ++m)
                        if (r + m >= 0 && r + m < nr) {
                            double ev = op_entry(a, var, i, m, k2, ce0);
                            double2 y = rm[(r + m) * SBT + cc];
                            acc.x += ev * y.x;
                            acc.y += ev * y.y;
                        }
                    acc.x += a.dtc * f.x;
                    acc.y += a.dtc * f.y;
                    acc.x += sf.x;
                    acc.y += sf.y;
                    vec[(mb * nr + r) * SBT + cc] = acc;
                }
            }
        }
        __syncthreads();
    }
    // The chains run in one wave of the block: the wave whose SIMD equals the
    // block's workgroup slot on the CU (mod 4), so that the blocks resident on
    // one CU run their chains on different SIMDs (wave 0 when no wave matches).
    // The other waves stage the U parts meanwhile, chunk by chunk, and signal.
    const int cl = tid & 63, cw = chain_wave[0];
    const int cr = (nr + NCHU - 1) / NCHU;
    // Every wave but the chain wave stages: each signals every chunk exactly
    // once (both staging paths below), so a chunk is complete at NWR signals.
    constexpr int NWR = SOLVE_THREADS / 64 - 1;          // staging waves
    static_assert(NWR >= 1 && SOLVE_THREADS % 64 == 0, "one chain wave, the rest stage");
    const StagedRows sr{ready, cr, NCHU, NWR, a.err};
    if ((tid >> 6) == cw) {
        // lane cl < 2 SBT nmb: member cl / (2 SBT), part (cl / SBT) % 2
        // (real, imaginary) of column cl % SBT
        if (cl < 2 * SBT * nmb && co.chain && nr > 0) {
            const int mb = cl / (2 * SBT), col = cl % SBT, part = (cl / SBT) & 1;
            double *vd = (double *) (vec + mb * nr * SBT) + 2 * col + part;
            chain_forward_r<DL, 3>(lf + col, W * SBT, SBT, vd, 2 * SBT, nr);
            chain_backward_staged<DL, 3>(lf + col, W * SBT, SBT, vd, 2 * SBT, nr, sr);
        }
    } else {
        // element e of the U parts counts from the last row up (row
        // nr - 1 - e / UPR); chunk c holds e in [c cr UPR, (c+1) cr UPR).
        // The staging waves' threads wt take e = c cr UPR + wt + j NT of
        // every chunk, all loads issued before the first store (chains of
        // up to 48 rows; longer ones copy directly).
        constexpr int NT = NWR * 64;
        constexpr int CRMAX = (48 + NCHU - 1) / NCHU;
        static_assert(CRMAX * UPR <= 2 * NT, "two elements per thread and chunk");
        const int wt = tid - ((tid >> 6) > cw ? 64 : 0);
        const int nu = nr * UPR, ce = cr * UPR;
        const bool regs = cr <= CRMAX;
        // a chunk's two elements per thread in named registers (arrays of
        // them stay in scratch memory next to the signalling atomic); the
        // loads of chunk c+1 are issued before the stores of chunk c
        struct UC {
            double2 v0, v1;
            int o0, o1;
            bool s0, s1;
        };
        auto ld = [&](int c) {
            UC u;
            const int e0 = c * ce, e1 = min(e0 + ce, nu);
            const int ea = e0 + wt, eb = ea + NT;
            u.s0 = ea < e1;
            u.s1 = eb < e1;
            const int ca = max(min(ea, e1 - 1), 0), cb = max(min(eb, e1 - 1), 0);
            const int ra = ca / UPR, rb = cb / UPR;
            // (no rows at all: read the image's first element, store nothing)
            u.o0 = max((nr - 1 - ra) * RPR + LPR + (ca - ra * UPR), 0);
            u.o1 = max((nr - 1 - rb) * RPR + LPR + (cb - rb * UPR), 0);
            u.v0 = src[u.o0];
            u.v1 = src[u.o1];
            return u;
        };
        auto st = [&](const UC &u, int c) {
            if (u.s0) dst[u.o0] = u.v0;
            if (u.s1) dst[u.o1] = u.v1;
            if (cl == 0) __hip_atomic_fetch_add(&ready[c], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        if (regs) {
            static_assert(NCHU == 4, "four chunks below");
            UC a0 = ld(0), a1 = ld(1);
            st(a0, 0);
            a0 = ld(2);
            st(a1, 1);
            a1 = ld(3);
            st(a0, 2);
            st(a1, 3);
        } else {
            for (int c = 0; c < NCHU; ++c) {
                const int e0 = c * ce, e1 = min(e0 + ce, nu);
                for (int e = e0 + wt; e < e1; e += NT) {
                    const int rb = e / UPR, o = (nr - 1 - rb) * RPR + LPR + (e - rb * UPR);
                    dst[o] = src[o];
                }
                if (cl == 0) __hip_atomic_fetch_add(&ready[c], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    __syncthreads();
    // write the solutions and the next stage's right-hand sides
    if constexpr (MB <= 2) {
        // each band entry of the next stage's operator formed once for the
        // block's members
#pragma unroll
        for (int t = 0; t < VR; ++t) {
            const int r = rt + t * RSTEP, i = par + 2 * r;
            if (r >= nr || !co.chain) continue;
            const size_t gi = cbase + r * rs;
            double2 f2[MB], acc[MB];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                acc[mb] = make_double2(0.0, 0.0);
                if (mb < nmb && a.has_next) f2[mb] = F2m(mb)[fbase + r * rs];
            }
            if (a.has_next) {
#pragma unroll
                for (int m = -DL; m <= DL; ++m)
                    if (r + m >= 0 && r + m < nr) {
                        const double ev = op_entry(a, var, i, m, k2, cn);
#pragma unroll
                        for (int mb = 0; mb < MB; ++mb)
                            if (mb < nmb) {
                                const double2 y = vec[(mb * nr + r + m) * SBT + cc];
                                acc[mb].x += ev * y.x;
                                acc[mb].y += ev * y.y;
                            }
                    }
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                if (mb >= nmb) continue;
                const double2 sol = vec[(mb * nr + r) * SBT + cc];
                Csm(mb)[gi] = sol;
                if (a.has_next) {
                    acc[mb].x += a.dtd * f2[mb].x;
                    acc[mb].y += a.dtd * f2[mb].y;
                    Cm(mb)[gi] = acc[mb];
                } else {
                    Cm(mb)[gi] = sol;
                }
            }
        }
    } else {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int t = 0; t < VR; ++t) {
                const int r = rt + t * RSTEP, i = par + 2 * r;
                if (mb >= nmb || r >= nr || !co.chain) continue;
                const double2 *vm = vec + mb * nr * SBT;
                const size_t gi = cbase + r * rs;
                const double2 sol = vm[r * SBT + cc];
                Csm(mb)[gi] = sol;
                if (a.has_next) {
                    const double2 f2 = F2m(mb)[fbase + r * rs];
                    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                    for (int m = -DL; m <= DL; ++m)
                        if (r + m >= 0 && r + m < nr) {
                            double ev = op_entry(a, var, i, m, k2, cn);
                            double2 y = vm[(r + m) * SBT + cc];
                            acc.x += ev * y.x;
                            acc.y += ev * y.y;
                        }
                    acc.x += a.dtd * f2.x;
                    acc.y += a.dtd * f2.y;
                    Cm(mb)[gi] = acc;
                } else {
                    Cm(mb)[gi] = sol;
                }
            }
    }
    // rows of v beyond dimQ are carried unchanged; the Nyquist row is zeroed
    const int nrest = (nR - par + 1) / 2;
    if (!co.valid || (co.z == 0 && co.x == 0)) return;
    for (int mb = 0; mb < nmb; ++mb)
        for (int r = rt; r < nrest; r += RSTEP) {
            const int i = par + 2 * r;
            const size_t gi = cbase + r * rs;
            if (co.z == a.Nz / 2)
                Cm(mb)[gi] = make_double2(0.0, 0.0);
            else if (var == 0 && i >= nQ)
                Csm(mb)[gi] = Cm(mb)[gi];
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

// the mean mode: alpha (streamwise mean) and beta (spanwise mean) columns.
// Every global load (state, both forcings, factors) is issued up front; the
// two explicit products of a pass run together.  Rows per thread:
// i = tid + MEAN_ROWS_STEP t, t < MEAN_ROWS (dimR <= 512).
constexpr int MEAN_ROWS = 2;
__device__ void solve_mean(const SolveArgs &a, cplx *C, cplx *Cs, double *lds)
{
    const int nR = a.dimR, Mx = a.Mx, T = SOLVE_THREADS, tid = threadIdx.x;
    double2 *al = (double2 *) lds, *be = al + nR, *ta = be + nR, *tb = ta + nR;
    double *lm = (double *) (tb + nR), *red = lm + 5 * nR;
    cplx *Ca = C, *Cb = C + (size_t) nR * Mx;
    const cplx *Fm = a.Fb, *Fm2 = a.Fb2;   // column (0, 0): (fa, fb)
    double2 fr[MEAN_ROWS], f2r[MEAN_ROWS];
#pragma unroll
    for (int t = 0; t < MEAN_ROWS; ++t) {
        const int i = tid + T * t;
        if (i < nR) {
            al[i] = Ca[(size_t) i * Mx];
            be[i] = Cb[(size_t) i * Mx];
            fr[t] = Fm[(size_t) i * Mx];
            f2r[t] = a.has_next ? Fm2[(size_t) i * Mx] : make_double2(0.0, 0.0);
        }
    }
    for (int e = tid; e < 5 * nR; e += T) lm[e] = a.Lm[e];
    __syncthreads();
    if (a.k == 0) {
#pragma unroll
        for (int t = 0; t < MEAN_ROWS; ++t) {
            const int i = tid + T * t;
            if (i < nR) {
                ta[i] = mean_mult(a, al, i, a.ce0[1]);
                tb[i] = mean_mult(a, be, i, a.ce0[1]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < MEAN_ROWS; ++t) {
            const int i = tid + T * t;
            if (i < nR) {
                al[i] = ta[i];
                be[i] = tb[i];
            }
        }
    }
#pragma unroll
    for (int t = 0; t < MEAN_ROWS; ++t) {
        const int i = tid + T * t;
        if (i < nR) {
            al[i].x += a.dtc * fr[t].x;
            be[i].x += a.dtc * fr[t].y;
            if (a.Src) {
                const double2 sa = a.Src[(size_t) i * Mx], sb = a.Src[(size_t) (nR + i) * Mx];
                al[i].x += sa.x;
                al[i].y += sa.y;
                be[i].x += sb.x;
                be[i].y += sb.y;
            }
        }
    }
    __syncthreads();
    if (a.exact_mean && a.k < 2) {
        // the transpose of the flux correction on the right-hand side
        double q = 0.0;
        for (int i = tid; i < nR; i += T) q += a.add_mass[i] * al[i].x;
        red[tid] = q;
        __syncthreads();
        for (int o = T / 2; o > 0; o >>= 1) {
            if (tid < o) red[tid] += red[tid + o];
            __syncthreads();
        }
        const double h = 0.5 * red[0];
        __syncthreads();
        for (int i = tid; i < nR; i += T) al[i].x -= h * a.rowsum[i];
        __syncthreads();
    }
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
        if (!a.exact_mean) {
            al[i].x = al[i].x + flux_t * a.add[i] * 0.5;
            al[i].y = al[i].y + flux_t * 0.0 * 0.5;
        }
        Cs[(size_t) i * Mx] = al[i];
        Cs[(size_t) (nR + i) * Mx] = be[i];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MEAN_ROWS; ++t) {
        const int i = tid + T * t;
        if (i >= nR) continue;
        if (a.has_next) {
            const double2 f = f2r[t];
            double2 xa = mean_mult(a, al, i, a.cn[1]), xb = mean_mult(a, be, i, a.cn[1]);
            Ca[(size_t) i * Mx] = make_double2(xa.x + a.dtd * f.x, xa.y + a.dtd * 0.0);
            Cb[(size_t) i * Mx] = make_double2(xb.x + a.dtd * f.y, xb.y + a.dtd * 0.0);
        } else {
            Ca[(size_t) i * Mx] = al[i];
            Cb[(size_t) i * Mx] = be[i];
        }
    }
}

// The Nyquist row kz = Nz/2 carries no data: a block whose columns all lie
// in it only clears its C rows.
template <int SBT>
__device__ void nyquist_clear(const SolveArgs &a, int var, int par, int blk, cplx *C)
{
    const int nR = a.dimR, Mx = a.Mx, nrest = (nR - par + 1) / 2;
    const int x0 = blk * SBT - (a.Nz / 2) * Mx;
    const size_t base = (((size_t) (a.Nz / 2) * 2 + var) * nR + par) * Mx + x0;
    for (int e = threadIdx.x; e < nrest * SBT; e += SOLVE_THREADS)
        C[base + (size_t) (e / SBT) * 2 * Mx + e % SBT] = make_double2(0.0, 0.0);
}

// grid (nblk, 4, member groups), or (nblk + 1, 4, member groups) when the
// mean mode needs a block of its own: blockIdx.y = field*2 + parity,
// blockIdx.z = group of MB members (the last group may hold fewer: a.nm
// members in all).  With Mx % SBT == 0 the blocks of the Nyquist row hold no
// chains, and the first of them with blockIdx.y == 0 runs the mean mode (so
// the grid is exactly the column blocks); otherwise the extra x-block nblk
// with blockIdx.y == 0 does.  The mean mode runs member after member.
template <int SBT, int VR, int MB>
__global__ __launch_bounds__(SOLVE_THREADS, MB == 1 ? 4 : MB == 2 ? 3 : 2) void k_solve(SolveArgs a, cplx *C, cplx *Cs)
{
    extern __shared__ double lds_solve[];
    if (*a.abort_flag) return;
    const int m0 = blockIdx.z * MB, nmb = min(MB, a.nm - m0);
    const int var = blockIdx.y >> 1, par = blockIdx.y & 1, blk = blockIdx.x;
    auto mean = [&]() {
        for (int mb = 0; mb < nmb; ++mb) {
            const int m = m0 + mb;
            SolveArgs b = a;
            if ((a.alt_mask >> m) & 1u) b.target = a.target_alt;
            b.Fa += m * a.ms_fa;
            b.Fb += m * a.ms_fb;
            b.Fa2 += m * a.ms_fa2;
            b.Fb2 += m * a.ms_fb2;
            if (b.Src) b.Src += m * a.ms_src;
            if (mb) __syncthreads();                 // the previous member's LDS reads done
            solve_mean(b, C + m * a.ms_c, Cs + m * a.ms_cs, lds_solve);
        }
    };
    if (blk == a.nblk) {
        if (blockIdx.y == 0) mean();
        return;
    }
    if (a.nyq_mean && blk * SBT / a.Mx == a.Nz / 2) {
        if (blockIdx.y == 0 && blk * SBT == (a.Nz / 2) * a.Mx) mean();
        for (int mb = 0; mb < nmb; ++mb)
            nyquist_clear<SBT>(a, var, par, blk, C + (m0 + mb) * a.ms_c);
        return;
    }
    if (var == 0)
        solve_block<4, SBT, VR, MB>(a, 0, par, blk, m0, nmb, C, Cs, lds_solve);
    else
        solve_block<2, SBT, VR, MB>(a, 1, par, blk, m0, nmb, C, Cs, lds_solve);
}

// coefficient tables for the stages
static const double FA_[3] = {1. / 3., -1. / 2, 1. / 3.};
static const double FC_[3] = {1. / 2, 1. / 3., 1.};
static const double FD_[3] = {0., -1. / 6, -2. / 3};
static const double BA_[3] = {1. / 3., -1. / 2, 1. / 3.};
static const double BC_[3] = {1., 1. / 3., 1. / 2.};
static const double BD_[3] = {0., -2. / 3, -1. / 6};

size_t solve_lds_bytes(int dimR, int sbt, int mb)
{
    const size_t nr = (size_t) (dimR + 1) / 2;
    const size_t blk = nr * sbt * (9 * 8 + 32 * (size_t) mb);
    const size_t mean = (size_t) dimR * (4 * 16 + 5 * 8) + SOLVE_THREADS * 8;
    return blk > mean ? blk : mean;
}

// Columns per solve block: 8, 4 or 2 (an even count keeps every block's
// factor image a whole number of 16-byte pairs), within 64 KB of LDS and with
// a row slot for every chain row; 0 when none fits.
int solve_columns_per_block(int dimR)
{
    const int nr = (dimR + 1) / 2;
    for (int sbt = 8; sbt >= 2; sbt >>= 1)
        if (solve_lds_bytes(dimR, sbt, 1) <= 64 * 1024 && nr <= SOLVE_VR_MAX * (SOLVE_THREADS / sbt))
            return sbt;
    return 0;
}

// Members per solve block (1, 2 or 4): the tuning's choice when set, else 4;
// fewer when the run has fewer members or their vectors and the factor image
// would not fit twice in the LDS of a CU.
int solve_members_per_block(const Context &c)
{
    int mb = tuning().solve_mb > 0 ? tuning().solve_mb : 4;
    // several members only with two row slots per thread (the four-slot
    // form would spill its registers)
    const bool two = (c.d.dimR + 1) / 2 <= 2 * (SOLVE_THREADS / c.sbt);
    if (!two) mb = 1;
    while (mb > 1 && (mb > c.nrun || 2 * solve_lds_bytes(c.d.dimR, c.sbt, mb) > 160 * 1024))
        mb /= 2;
    return mb;
}

void launch_solve(Context &c, cplx *Ccarry, cplx *Csol, int k, int adj, const cplx *Fa2,
                  const cplx *Fb2, double target, double target_alt, unsigned alt_mask,
                  const cplx *Src, int exact_mean)
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
    a.target_alt = target_alt;
    a.alt_mask = alt_mask;
    a.Src = Src;
    a.ms_src = mstride(c, Src);
    a.exact_mean = adj && exact_mean;
    a.add_mass = c.dAdd[0][k];
    a.abort_flag = c.abort_flag;
    a.err = c.abort_flag + 2;
    a.ms_c = mstride(c, Ccarry);
    a.ms_cs = mstride(c, Csol);
    a.ms_fa = mstride(c, a.Fa);
    a.ms_fb = mstride(c, a.Fb);
    a.ms_fa2 = mstride(c, Fa2);
    a.ms_fb2 = mstride(c, Fb2);
    // RHS in, projections in (the next stage's only when they are other
    // arrays: the primal's are the same), solution and next right-hand side
    // out, and the factor images in (each stored image once: mirror rows
    // share one)
    const bool f2 = k != 2 && Fa2 != c.Fa;
    const double fcols = (double) (c.factor_mirror ? d.Nz / 2 + 1 : d.Nz) * d.Mx;
    double bytes = c.nrun * (double) (d.Nz - 1) * d.Mx *
                       (16.0 * (2.0 * d.dimR) * 3.0 + 16.0 * (d.dimQ + d.dimR) * (f2 ? 2.0 : 1.0)) +
                   fcols * 8.0 * (9.0 * d.dimQ + 5.0 * d.dimR);
    const int sbt = c.sbt;
    const int ncol = d.Nz * d.Mx;
    // members per block: the staged factor image feeds the chains of up to
    // four members (2 sbt lanes each), within 160 KB of LDS per CU
    const int mb = solve_members_per_block(c);
    const size_t lds = solve_lds_bytes(d.dimR, sbt, mb);
    a.nm = c.nrun;
    a.nblk = (ncol + sbt - 1) / sbt;
    a.mirror = c.factor_mirror;
    a.nyq_mean = d.Mx % sbt == 0;
    const dim3 grid(a.nblk + (a.nyq_mean ? 0 : 1), 4, (c.nrun + mb - 1) / mb);
    const hipStream_t st = c.stream;
    // row slots per thread: the chains have at most (dimR + 1) / 2 rows
    const bool two = (d.dimR + 1) / 2 <= 2 * (SOLVE_THREADS / sbt);
    typedef void (*SolveKernel)(SolveArgs, cplx *, cplx *);
    SolveKernel kern;
    if (mb == 4)
        kern = sbt == 8 ? k_solve<8, 2, 4> : sbt == 4 ? k_solve<4, 2, 4> : k_solve<2, 2, 4>;
    else if (mb == 2)
        kern = sbt == 8 ? k_solve<8, 2, 2> : sbt == 4 ? k_solve<4, 2, 2> : k_solve<2, 2, 2>;
    else
        kern = sbt == 8 ? (two ? k_solve<8, 2, 1> : k_solve<8, 4, 1>)
             : sbt == 4 ? (two ? k_solve<4, 2, 1> : k_solve<4, 4, 1>)
                        : (two ? k_solve<2, 2, 1> : k_solve<2, 4, 1>);
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int) lds);
    auto go = [=]() {
        hipLaunchKernelGGL(kern, grid, dim3(SOLVE_THREADS), lds, st, a, Ccarry, Csol);
    };
    prof_begin(c, st, 6, bytes);
    go();
    prof_end(c, st, 6, go);
}

// ------------------------------------------------------------ vectors
// Inner products of ddt_project (channel/channel.c:704-780) straight from the
// coefficients.  The reference forms <f, g> = sum_y W_y sum_{x,z} f g over the
// velocities of f and g rebuilt on the dealiased Px x Pz x qpts grid.  Those
// sums are exact integrals: Parseval in x and z (the fields are band-limited
// below the padded grid's Nyquist bins), and in y the Gauss rule with qpts
// points integrates the products of the rebuilt fields exactly (degree at
// most 2 dimQ + 6 < 2 qpts).  With u = (i kx v' - i kz w_y)/k2 and
// w = (i kz v' + i kx w_y)/k2, u conj(u~) + w conj(w~) = (v' conj(v~') +
// w_y conj(w~_y))/k2, so per column (kz, kx) != (0, 0)
//     <f, g>_col = c Re[ a_f^H (Qs + Qps/k2) a_g + b_f^H (Rs/k2) b_g ]
// with the Gauss-rule Gram bands Qs = sum W Q Q, Qps = sum W Q' Q',
// Rs = sum W R R of the setup, alpha rows a < dimQ, beta rows b < dimR,
// c = 2 for kx > 0 (the conjugate column) and c = 1 at kx = 0, where the
// coefficients are replaced by their Hermitian part (c(kz) + conj c(-kz))/2,
// the only part the complex-to-real x transform sees.  The mean column holds
// the u and w profiles in the R basis: Re a_f^T Rs Re a_g + Re b_f^T Rs Re b_g
// (the Hermitian part of column (0,0) is its real part).  Row Nz/2 is not
// transformed and counts nothing.  The sums differ from the grid sums only
// in rounding.  num = <dudt, V>, den = <dudt, dudt> (times Px Pz, the grid's
// Parseval factor) with dudt = s (P - M) formed on load.
//
// Block: kz row z and ncol columns; dudt and V of the block's columns in
// LDS ([var][row][column]), then one (var, row, column) item per thread and
// pass; one partial pair per block.
// Sums of a and b over the 256 threads of a block, in the same order in every
// block: lane exchanges within each wave (distances 32 .. 1; both partners
// add the same two values, so every lane holds the same bits), then the four
// wave sums in wave order.  Every thread gets both sums.  sh: 8 doubles.
__device__ __forceinline__ double2 block_sum2_256(double a, double b, double *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = a;
        sh[4 + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    return make_double2((sh[0] + sh[1]) + (sh[2] + sh[3]), (sh[4] + sh[5]) + (sh[6] + sh[7]));
}

constexpr int DDT_IT = 4;        // items per thread of k_ddt_gram's load phase
constexpr int DDT_BT = 12;       // Gram band values per thread loaded in one batch

__global__ __launch_bounds__(256) void k_ddt_gram(const cplx *P, const cplx *M, const cplx *V,
                                                  double s, const double *Qs, const double *Qps,
                                                  const double *Rs, const double *Kx,
                                                  const double *Kz, int Nz, int Mx, int dimR,
                                                  int dimQ, int ncol, double scale, double *part,
                                                  const int *abort_flag, size_t ms_p, size_t ms_m,
                                                  size_t ms_v, const cplx *P0, const cplx *M0)
{
    extern __shared__ double2 sh_ddt[];
    __shared__ double rs[8];
    // member blockIdx.y: its partial pairs follow those of the members before;
    // P0, M0 (when given) replace member 0's pair of states
    P += blockIdx.y * ms_p;
    M += blockIdx.y * ms_m;
    if (blockIdx.y == 0 && P0) {
        P = P0;
        M = M0;
    }
    V += blockIdx.y * ms_v;
    part += (size_t) blockIdx.y * 2 * gridDim.x;
    const int nxb = (Mx + ncol - 1) / ncol;
    const int z = blockIdx.x / nxb, x0 = (blockIdx.x - z * nxb) * ncol;
    const bool live = !*abort_flag && z != Nz / 2;
    const int nit = 2 * dimR * ncol;       // at most 256 DDT_IT (host)
    double2 *sd = sh_ddt, *sv = sh_ddt + nit;
    double2 *smd = sv + nit, *smv = smd + 2 * dimR;   // kx = 0 column of row -kz
    // the Gram bands: Qs, Qps [dimQ][9], Rs [dimR][5]
    double *bq = (double *) (smv + 2 * dimR), *bqp = bq + 9 * dimQ, *br = bqp + 9 * dimQ;
    const bool mirror = live && x0 == 0;
    if (live) {
        // every load of the thread in flight before the first store: the
        // block's items, the mirror column, the Gram bands
        double2 rp[DDT_IT], rm[DDT_IT], rv[DDT_IT], qp, qm, qv;
#pragma unroll
        for (int t = 0; t < DDT_IT; ++t) {
            const int e = threadIdx.x + t * 256;
            const int col = e % ncol, j = (e / ncol) % dimR, var = e / (ncol * dimR);
            const int x = min(x0 + col, Mx - 1);
            const size_t o = (((size_t) z * 2 + min(var, 1)) * dimR + j) * Mx + x;
            if (e < nit) {
                rp[t] = P[o];
                rm[t] = M[o];
                rv[t] = V[o];
            }
        }
        const bool mload = mirror && (int) threadIdx.x < 2 * dimR;
        if (mload) {
            const int zm = z == 0 ? 0 : Nz - z;
            const int j = threadIdx.x % dimR, var = threadIdx.x / dimR;
            const size_t om = (((size_t) zm * 2 + var) * dimR + j) * Mx;
            qp = P[om];
            qm = M[om];
            qv = V[om];
        }
        const int nbq = 9 * dimQ, nbt = 18 * dimQ + 5 * dimR;
        auto band = [&](int i) {
            return i < nbq ? Qs[i] : i < 2 * nbq ? Qps[min(i - nbq, nbq - 1)]
                                                 : Rs[min(max(i - 2 * nbq, 0), 5 * dimR - 1)];
        };
        double tb[DDT_BT];
#pragma unroll
        for (int u = 0; u < DDT_BT; ++u) {
            const int i = threadIdx.x + u * 256;
            if (i < nbt) tb[u] = band(i);
        }
#pragma unroll
        for (int u = 0; u < DDT_BT; ++u) {
            const int i = threadIdx.x + u * 256;
            if (i < nbt) bq[i] = tb[u];
        }
        for (int i = threadIdx.x + DDT_BT * 256; i < nbt; i += 256) bq[i] = band(i);
#pragma unroll
        for (int t = 0; t < DDT_IT; ++t) {
            const int e = threadIdx.x + t * 256;
            if (e >= nit) continue;
            const int col = e % ncol;
            const bool in = x0 + col < Mx;
            sd[e] = in ? make_double2(s * (rp[t].x - rm[t].x), s * (rp[t].y - rm[t].y))
                       : make_double2(0.0, 0.0);
            sv[e] = in ? rv[t] : make_double2(0.0, 0.0);
        }
        if (mload) {
            smd[threadIdx.x] = make_double2(s * (qp.x - qm.x), s * (qp.y - qm.y));
            smv[threadIdx.x] = qv;
        }
        if (mirror) {
            // rows beyond the first 256 (var, j) pairs (dimR > 128)
            const int zm = z == 0 ? 0 : Nz - z;
            for (int q = threadIdx.x + 256; q < 2 * dimR; q += 256) {
                const int j = q % dimR, var = q / dimR;
                const size_t om = (((size_t) zm * 2 + var) * dimR + j) * Mx;
                const cplx pm = P[om], mm = M[om];
                smd[q] = make_double2(s * (pm.x - mm.x), s * (pm.y - mm.y));
                smv[q] = V[om];
            }
        }
    }
    __syncthreads();
    if (mirror) {
        // Hermitian part (c(kz) + conj c(-kz)) / 2 of the kx = 0 column
        for (int q = threadIdx.x; q < 2 * dimR; q += 256) {
            const int e = q * ncol;                // (var, j) = q, column 0
            const double2 d = sd[e], v = sv[e], dm = smd[q], vm = smv[q];
            sd[e] = make_double2(0.5 * (d.x + dm.x), 0.5 * (d.y - dm.y));
            sv[e] = make_double2(0.5 * (v.x + vm.x), 0.5 * (v.y - vm.y));
        }
    }
    __syncthreads();
    double num = 0.0, den = 0.0;
    if (live) {
        const double kz = Kz[z];
        for (int e = threadIdx.x; e < nit; e += 256) {
            const int col = e % ncol, j = (e / ncol) % dimR, var = e / (ncol * dimR);
            const int x = x0 + col;
            if (x >= Mx) continue;
            const bool mean = z == 0 && x == 0;
            const bool qrow = var == 0 && !mean;             // v in the Q basis
            if (qrow && j >= dimQ) continue;
            const double kx = Kx[x], k2 = kx * kx + kz * kz;
            const double2 *dc = sd + var * dimR * ncol + col;
            const double2 *vc = sv + var * dimR * ncol + col;
            double gdx = 0.0, gdy = 0.0, gvx = 0.0, gvy = 0.0;   // (G d)_j, (G v)_j
            if (qrow) {
                const double ik2 = 1.0 / k2;
#pragma unroll
                for (int m = -4; m <= 4; ++m) {
                    const int i = j + 2 * m;
                    if (i < 0 || i >= dimQ) continue;
                    const double g = bq[j * 9 + 4 + m] + bqp[j * 9 + 4 + m] * ik2;
                    const double2 di = dc[i * ncol], vi = vc[i * ncol];
                    gdx += g * di.x;
                    gdy += g * di.y;
                    gvx += g * vi.x;
                    gvy += g * vi.y;
                }
            } else {
                const double f = mean ? 1.0 : 1.0 / k2;
#pragma unroll
                for (int m = -2; m <= 2; ++m) {
                    const int i = j + 2 * m;
                    if (i < 0 || i >= dimR) continue;
                    const double g = br[j * 5 + 2 + m] * f;
                    const double2 di = dc[i * ncol], vi = vc[i * ncol];
                    gdx += g * di.x;
                    gdy += g * di.y;
                    gvx += g * vi.x;
                    gvy += g * vi.y;
                }
            }
            const double2 dj = dc[j * ncol];
            const double w = x > 0 ? 2.0 : 1.0;
            num += w * (dj.x * gvx + dj.y * gvy);
            den += w * (dj.x * gdx + dj.y * gdy);
        }
    }
    const double2 t = block_sum2_256(num * scale, den * scale, rs);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = t.x;
        part[2 * blockIdx.x + 1] = t.y;
    }
}

// vout = vin - r dC, dC = s (P - M), with r = num / den from the nparts
// partial pairs of k_ddt_gram, summed by every block in the same order
// (strided over 256 threads, then block_sum2_256); block 0 leaves num, den, r in
// scal.  With acc, also acc -= a vout, twice when `twice` (the adjoint's
// trapezoid forcing at the end of one step and the start of the next,
// applied back to back).
// grid (blocks, members): member blockIdx.y sums its own nparts partial
// pairs and writes its scalars to scal + 8 blockIdx.y
__global__ __launch_bounds__(256) void k_ddt_apply(cplx *vout, const cplx *vin, const cplx *P,
                                                   const cplx *M, double s, const double *part,
                                                   int nparts, double *scal, size_t n, cplx *acc,
                                                   double a, int twice, size_t ms_vo, size_t ms_vi,
                                                   size_t ms_p, size_t ms_m, size_t ms_acc,
                                                   const cplx *P0, const cplx *M0)
{
    __shared__ double rs[8];
    {
        const int m = blockIdx.y;
        vout += m * ms_vo;
        vin += m * ms_vi;
        P += m * ms_p;
        M += m * ms_m;
        if (m == 0 && P0) {
            P = P0;
            M = M0;
        }
        if (acc) acc += m * ms_acc;
        part += (size_t) m * 2 * nparts;
        scal += 8 * m;
    }
    double sa = 0.0, sb = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        sa += part[2 * i];
        sb += part[2 * i + 1];
    }
    const double2 t = block_sum2_256(sa, sb, rs);
    const double r = t.x / t.y;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[0] = t.x;
        scal[1] = t.y;
        scal[2] = r;
    }
    for (size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x; e < n;
         e += (size_t) gridDim.x * blockDim.x) {
        const cplx p = P[e], m = M[e];
        const double dx = s * (p.x - m.x), dy = s * (p.y - m.y);
        cplx v = vin[e];
        v.x -= r * dx;
        v.y -= r * dy;
        vout[e] = v;
        if (acc) {
            cplx w = acc[e];
            w.x -= a * v.x;
            w.y -= a * v.y;
            if (twice) {
                w.x -= a * v.x;
                w.y -= a * v.y;
            }
            acc[e] = w;
        }
    }
}

// Reynolds-number sensitivity forcing of a stored state S = MC[count]
// (channel/tangent_forcing.c:10-75, channel/tangent_forcing0.c:10-44), band
// products in the order of smMult / smMult0 (channel/mvOps.c:198-237):
//   (kz,kx) != (0,0): fa = -re^2 (K2^2 Qs + 2 K2 Qps + Qpps) alpha   [Nz][dimQ][Mx]
//                     fb =  re^2 (Rps + K2 Rs) beta                   [Nz][dimR][Mx]
//   mean column:      f0 = re^2 Rps alpha(0,0), re^2 Rps beta(0,0)    [2][dimR]
// Column (0,0) of fa and fb is zero.  One thread per output element.
__global__ void k_tangent_forcing(const cplx *S, const double *Qs, const double *Qps,
                                  const double *Qpps, const double *Rs, const double *Rps,
                                  const double *Kx, const double *Kz, double re, int Nz, int Mx,
                                  int nQ, int nR, cplx *fa, cplx *fb, cplx *f0)
{
    const size_t na = (size_t) Nz * nQ * Mx, nb = (size_t) Nz * nR * Mx;
    const size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
    if (e >= na + nb + 2 * (size_t) nR) return;
    if (e >= na + nb) {
        const int v = (int) ((e - na - nb) / nR), i = (int) (e - na - nb) - v * nR;
        const cplx *y = S + (size_t) v * nR * Mx;                  // column (0,0) of alpha or beta
        double sr = 0.0, si = 0.0;
        for (int m = -min(2, i / 2); m <= min(2, (nR - i + 1) / 2 - 1); ++m) {
            const double a = re * re * Rps[i * 5 + 2 + m];
            sr += a * y[(size_t) (i + 2 * m) * Mx].x;
            si += a * y[(size_t) (i + 2 * m) * Mx].y;
        }
        f0[e - na - nb] = make_double2(sr, si);
        return;
    }
    const bool alpha = e < na;
    const size_t r = alpha ? e : e - na;
    const int N = alpha ? nQ : nR;
    const int x = (int) (r % Mx), i = (int) ((r / Mx) % N), z = (int) (r / ((size_t) Mx * N));
    cplx *out = alpha ? fa + r : fb + r;
    if (z == 0 && x == 0) {
        *out = make_double2(0.0, 0.0);
        return;
    }
    const double k2 = Kz[z] * Kz[z] + Kx[x] * Kx[x], s = k2 * k2;
    const cplx *y = S + ((size_t) (z * 2 + (alpha ? 0 : 1)) * nR) * Mx + x;
    double sr = 0.0, si = 0.0;
    if (alpha) {
        for (int m = -min(4, i / 2); m <= min(4, (nQ - i + 1) / 2 - 1); ++m) {
            const int b = i * 9 + 4 + m;
            const double a = -re * re * (s * Qs[b] + 2. * k2 * Qps[b] + Qpps[b]);
            sr += a * y[(size_t) (i + 2 * m) * Mx].x;
            si += a * y[(size_t) (i + 2 * m) * Mx].y;
        }
    } else {
        for (int m = -min(2, i / 2); m <= min(2, (nR - i + 1) / 2 - 1); ++m) {
            const int b = i * 5 + 2 + m;
            const double a = re * re * (Rps[b] + k2 * Rs[b]);
            sr += a * y[(size_t) (i + 2 * m) * Mx].x;
            si += a * y[(size_t) (i + 2 * m) * Mx].y;
        }
    }
    *out = make_double2(sr, si);
}

void launch_tangent_forcing(Context &c, const cplx *S, cplx *fa, cplx *fb, cplx *f0)
{
    const Dims &d = c.d;
    const size_t n = (size_t) d.Nz * (d.dimQ + d.dimR) * d.Mx + 2 * (size_t) d.dimR;
    hipLaunchKernelGGL(k_tangent_forcing, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, c.stream,
                       S, c.dQs, c.dQps, c.dQpps, c.dRs, c.dRps, c.dKx, c.dKz, d.re, d.Nz, d.Mx,
                       d.dimQ, d.dimR, fa, fb, f0);
}

// LDS bytes of a k_ddt_gram block of nc columns
static size_t ddt_gram_lds(const Dims &d, int nc)
{
    return sizeof(double2) * 4 * (size_t) d.dimR * (nc + 1) +
           sizeof(double) * (18 * (size_t) d.dimQ + 5 * (size_t) d.dimR);
}

// columns per k_ddt_gram block: 4, fewer while the block's items exceed
// 256 DDT_IT or its LDS 64 KB (grids beyond dimR = 128 with one column ask
// for more)
static int ddt_gram_cols(const Dims &d)
{
    int ncol = 4;
    while (ncol > 1 && (2 * d.dimR * ncol > 256 * DDT_IT || ddt_gram_lds(d, ncol) > 65536))
        ncol /= 2;
    return ncol;
}

int ddt_gram_blocks(const Dims &d)
{
    const int ncol = ddt_gram_cols(d);
    return d.Nz * ((d.Mx + ncol - 1) / ncol);
}

int launch_ddt_gram(Context &c, const cplx *P, const cplx *M, double s, const cplx *V)
{
    return launch_ddt_gram(c, P, M, s, V, c.dotpart, c.ndotpart, c.abort_flag);
}

int launch_ddt_gram(Context &c, const cplx *P, const cplx *M, double s, const cplx *V,
                    double *part, size_t part_cap, const int *stop, const cplx *P0,
                    const cplx *M0)
{
    const Dims &d = c.d;
    const int ncol = ddt_gram_cols(d), nb = ddt_gram_blocks(d);
    if (2 * (size_t) nb * c.nrun > part_cap) {
        last_error() = "ddt_project: partial-sum buffer too small";
        return 0;
    }
    const size_t lds = ddt_gram_lds(d, ncol);
    if (lds > 65536)
        (void) hipFuncSetAttribute((const void *) k_ddt_gram,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    hipLaunchKernelGGL(k_ddt_gram, dim3(nb, c.nrun), dim3(256), lds, c.stream, P, M, V, s, c.dQs,
                       c.dQps, c.dRs, c.dKx, c.dKz, d.Nz, d.Mx, d.dimR, d.dimQ, ncol,
                       (double) d.Px * (double) d.Pz, part, stop, mstride(c, P), mstride(c, M),
                       mstride(c, V), P0, M0);
    return nb;
}

void launch_ddt_apply(Context &c, cplx *vout, const cplx *vin, const cplx *P, const cplx *M,
                      double s, size_t n, cplx *acc, double a, int nparts, int twice)
{
    launch_ddt_apply(c, vout, vin, P, M, s, n, acc, a, nparts, twice, c.dotpart, c.scal);
}

void launch_ddt_apply(Context &c, cplx *vout, const cplx *vin, const cplx *P, const cplx *M,
                      double s, size_t n, cplx *acc, double a, int nparts, int twice,
                      const double *part, double *scal, const cplx *P0, const cplx *M0)
{
    const int nm = c.nrun;
    hipLaunchKernelGGL(k_ddt_apply, dim3(std::max(128, 1024 / nm), nm), dim3(256), 0, c.stream,
                       vout, vin, P, M, s, part, nparts, scal, n, acc, a, twice, mstride(c, vout),
                       mstride(c, vin), mstride(c, P), mstride(c, M), acc ? mstride(c, acc) : 0,
                       P0, M0);
}

__global__ __launch_bounds__(256) void k_zero(cplx *p, size_t n)
{
    for (size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x; e < n;
         e += (size_t) gridDim.x * blockDim.x)
        p[e] = make_double2(0.0, 0.0);
}

void launch_zero(Context &c, cplx *p, size_t n)
{
    const size_t nb = std::min<size_t>(1024, (n + 255) / 256);
    if (nb) hipLaunchKernelGGL(k_zero, dim3((unsigned) nb), dim3(256), 0, c.stream, p, n);
}

}  // namespace sc
