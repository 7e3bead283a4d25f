// Dealiased x/z transforms of the nonlinear terms, one launch per direction.
//
// The reference does, per Gauss plane y and field, an unnormalised backward
// c2c along z over the Nx/2 stored columns of a 3/2-padded array, then a c2r
// along x, the pointwise products, and the reverse with a 1/(Px Pz) scale and
// truncation (channel/pass1.c:107-114, 312-371; channel/pass2.c:128-221,
// 417-472; channel/spec2phys.c:67-76).  Here the three phases are separate
// batched kernels over all planes at once:
//   zinv : (y, field, kx block) -> length-Pz inverse transforms in LDS, with
//          the spectral curl fused into the load and the zero padding implied
//   xmix : (y, block of z rows) -> pairs of real rows packed into one complex
//          length-Px inverse transform, products, CFL maximum, forward
//          transforms of packed product pairs, unpacking of the kx < Nx/2 bins
//   zfwd : (y, field, kx block) -> forward length-Pz transforms, scale,
//          truncation to the stored kz rows
// The FFTs are Stockham passes run in place in LDS with radices up to 16 (the
// composite ones done as small FFTs in registers, so a length of 96 takes two
// passes): every thread holds its butterflies' inputs in registers across a
// barrier.
#include "sc_internal.h"

// The file is compiled twice: as is (sc::fp64, FP64 transforms) and with
// SC_FFT_F32 defined (sc::fp32: transforms, products and the intermediate
// arrays W1, W2 and the cached fields in FP32).  The state arrays read at
// the start (U, FP64) and the results written at the end (the Gauss-point
// slots, FP64) are FP64 in both; values are converted on load and store.
#ifdef SC_FFT_F32
#define SC_FFT_NS fp32
#else
#define SC_FFT_NS fp64
#endif

namespace sc {
namespace SC_FFT_NS {

#ifdef SC_FFT_F32
using real = float;
using real2 = float2;
#else
using real = double;
using real2 = double2;
#endif

__host__ __device__ __forceinline__ real2 mk2(real a, real b)
{
    real2 v;
    v.x = a;
    v.y = b;
    return v;
}
__device__ __forceinline__ real2 to_real2(const double2 &v) { return mk2((real) v.x, (real) v.y); }
__device__ __forceinline__ double2 to_cplx(const real2 &v) { return make_double2(v.x, v.y); }
// member stride of an array viewed as real2 values
static inline size_t rstride(const Context &c, const void *p) { return mstride_bytes(c, p) / sizeof(real2); }

constexpr int BLK = FFT_BLOCK;   // threads per block of the transform kernels
// complex values staged in LDS by one x-kernel block, data and twiddles
constexpr int XMIX_TILE = FFT_TILE;

struct PlanDev {
    int n, nf;
    int f[MAX_FACTORS];
    // per pass: ceil(2^32 / Ns), Ns = product of the earlier radices, so that
    // x / Ns = (x * nsm) >> 32 exactly for x <= 8192, Ns <= 4096
    unsigned long long nsm[MAX_FACTORS];
    const double2 *tw;         // the plan's FP64 twiddles (converted when staged)
};

static PlanDev to_dev(const FftPlan &p)
{
    PlanDev d;
    d.n = p.n;
    d.nf = p.nf;
    unsigned long long ns = 1;
    for (int i = 0; i < MAX_FACTORS; ++i) {
        d.f[i] = i < p.nf ? p.f[i] : 0;
        d.nsm[i] = ((1ull << 32) + ns - 1) / ns;
        if (i < p.nf) ns *= (unsigned long long) p.f[i];
    }
    d.tw = p.tw;
    return d;
}

__device__ __forceinline__ int magic_div(int x, unsigned long long m)
{
    return (int) (((unsigned long long) (unsigned) x * m) >> 32);
}

// (index / L, index % L) of the elements threadIdx.x, threadIdx.x + BLK, ...
// of a row-major walk with rows of length L, without divisions in the walk
struct Walk {
    int q, r, dq, dr, L;
    __device__ __forceinline__ Walk(int L_) : L(L_)
    {
        // one division per walk
        q = (int) threadIdx.x / L;
        r = (int) threadIdx.x - q * L;
        dq = BLK / L;
        dr = BLK - dq * L;
    }
    __device__ __forceinline__ void next()
    {
        q += dq;
        r += dr;
        if (r >= L) {
            r -= L;
            ++q;
        }
    }
};

__device__ __forceinline__ real2 cmul(real2 a, real2 b)
{
    return mk2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place DFTs in registers: v[q] <- sum_m v[m] exp(sign 2 pi i m q / R).
template <int R> __device__ __forceinline__ void dft(real2 *v, real sign);

template <> __device__ __forceinline__ void dft<2>(real2 *v, real sign)
{
    real2 a = v[0], b = v[1];
    v[0] = mk2(a.x + b.x, a.y + b.y);
    v[1] = mk2(a.x - b.x, a.y - b.y);
}

template <> __device__ __forceinline__ void dft<4>(real2 *v, real sign)
{
    real2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
    real2 s02 = mk2(a0.x + a2.x, a0.y + a2.y);
    real2 d02 = mk2(a0.x - a2.x, a0.y - a2.y);
    real2 s13 = mk2(a1.x + a3.x, a1.y + a3.y);
    real2 d13 = mk2(a1.x - a3.x, a1.y - a3.y);
    // w d13 with w = sign * i
    real2 wd = mk2(-sign * d13.y, sign * d13.x);
    v[0] = mk2(s02.x + s13.x, s02.y + s13.y);
    v[1] = mk2(d02.x + wd.x, d02.y + wd.y);
    v[2] = mk2(s02.x - s13.x, s02.y - s13.y);
    v[3] = mk2(d02.x - wd.x, d02.y - wd.y);
}

template <> __device__ __forceinline__ void dft<3>(real2 *v, real sign)
{
    const real h = 0.86602540378443864676;   // sqrt(3)/2
    real2 a0 = v[0], a1 = v[1], a2 = v[2];
    real2 t1 = mk2(a1.x + a2.x, a1.y + a2.y);
    real2 t2 = mk2(a0.x - (real) 0.5 * t1.x, a0.y - (real) 0.5 * t1.y);
    real2 t3 = mk2(sign * h * (a1.x - a2.x), sign * h * (a1.y - a2.y));
    v[0] = mk2(a0.x + t1.x, a0.y + t1.y);
    v[1] = mk2(t2.x - t3.y, t2.y + t3.x);
    v[2] = mk2(t2.x + t3.y, t2.y - t3.x);
}

template <> __device__ __forceinline__ void dft<5>(real2 *v, real sign)
{
    const real c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const real s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    real2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3], a4 = v[4];
    real2 b1 = mk2(a1.x + a4.x, a1.y + a4.y);
    real2 b2 = mk2(a2.x + a3.x, a2.y + a3.y);
    real2 d1 = mk2(a1.x - a4.x, a1.y - a4.y);
    real2 d2 = mk2(a2.x - a3.x, a2.y - a3.y);
    real2 t1 = mk2(a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y);
    real2 t2 = mk2(a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y);
    real2 u1 = mk2(sign * (s1 * d1.x + s2 * d2.x), sign * (s1 * d1.y + s2 * d2.y));
    real2 u2 = mk2(sign * (s2 * d1.x - s1 * d2.x), sign * (s2 * d1.y - s1 * d2.y));
    v[0] = mk2(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
    v[1] = mk2(t1.x - u1.y, t1.y + u1.x);
    v[4] = mk2(t1.x + u1.y, t1.y - u1.x);
    v[2] = mk2(t2.x - u2.y, t2.y + u2.x);
    v[3] = mk2(t2.x + u2.y, t2.y - u2.x);
}

// cos and sin of 2 pi j / R, j < R (correctly rounded; exact where the value
// is 0, +-1/2 or +-1)
template <int R> struct RootTab;
template <> struct RootTab<6> {
    static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    static constexpr double s[6] = {0.0, 0.8660254037844386, 0.8660254037844386, 0.0, -0.8660254037844386, -0.8660254037844386};
};
template <> struct RootTab<8> {
    static constexpr double c[8] = {1.0, 0.7071067811865476, 0.0, -0.7071067811865476, -1.0, -0.7071067811865476, 0.0, 0.7071067811865476};
    static constexpr double s[8] = {0.0, 0.7071067811865476, 1.0, 0.7071067811865476, 0.0, -0.7071067811865476, -1.0, -0.7071067811865476};
};
template <> struct RootTab<9> {
    static constexpr double c[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084, -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
    static constexpr double s[9] = {0.0, 0.6427876096865394, 0.984807753012208, 0.8660254037844386, 0.3420201433256687, -0.3420201433256687, -0.8660254037844386, -0.984807753012208, -0.6427876096865394};
};
template <> struct RootTab<10> {
    static constexpr double c[10] = {1.0, 0.8090169943749475, 0.30901699437494745, -0.30901699437494745, -0.8090169943749475, -1.0, -0.8090169943749475, -0.30901699437494745, 0.30901699437494745, 0.8090169943749475};
    static constexpr double s[10] = {0.0, 0.5877852522924731, 0.9510565162951535, 0.9510565162951535, 0.5877852522924731, 0.0, -0.5877852522924731, -0.9510565162951535, -0.9510565162951535, -0.5877852522924731};
};
template <> struct RootTab<12> {
    static constexpr double c[12] = {1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5, 0.0, 0.5, 0.8660254037844386};
    static constexpr double s[12] = {0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5};
};
template <> struct RootTab<15> {
    static constexpr double c[15] = {1.0, 0.9135454576426009, 0.6691306063588582, 0.30901699437494745, -0.10452846326765347, -0.5, -0.8090169943749475, -0.9781476007338057, -0.9781476007338057, -0.8090169943749475, -0.5, -0.10452846326765347, 0.30901699437494745, 0.6691306063588582, 0.9135454576426009};
    static constexpr double s[15] = {0.0, 0.4067366430758002, 0.7431448254773942, 0.9510565162951535, 0.9945218953682733, 0.8660254037844386, 0.5877852522924731, 0.20791169081775934, -0.20791169081775934, -0.5877852522924731, -0.8660254037844386, -0.9945218953682733, -0.9510565162951535, -0.7431448254773942, -0.4067366430758002};
};
template <> struct RootTab<16> {
    static constexpr double c[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898, 0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867};
    static constexpr double s[16] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867, 1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898};
};

// Composite R = R1 R2 in registers: with m = R2 m1 + m2 and q = k1 + R1 k2,
// DFT_R = (R1 DFTs of length R2) . twiddles exp(sign 2 pi i m2 k1 / R) .
// (R2 DFTs of length R1).  The twiddles are compile-time constants.
template <int R1, int R2>
__device__ __forceinline__ void dft_comp(real2 *v, real sign)
{
    constexpr int R = R1 * R2;
    real2 a[R];
#pragma unroll
    for (int m2 = 0; m2 < R2; ++m2) {
        real2 t[R1];
#pragma unroll
        for (int m1 = 0; m1 < R1; ++m1) t[m1] = v[R2 * m1 + m2];
        dft<R1>(t, sign);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            const int e = (m2 * k1) % R;
            if (e == 0) {
                a[m2 * R1 + k1] = t[k1];
            } else {
                const real2 w = mk2((real) RootTab<R>::c[e], sign * (real) RootTab<R>::s[e]);
                a[m2 * R1 + k1] = cmul(t[k1], w);
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        real2 t[R2];
#pragma unroll
        for (int m2 = 0; m2 < R2; ++m2) t[m2] = a[m2 * R1 + k1];
        dft<R2>(t, sign);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
    }
}

template <> __device__ __forceinline__ void dft<6>(real2 *v, real s) { dft_comp<3, 2>(v, s); }
template <> __device__ __forceinline__ void dft<8>(real2 *v, real s) { dft_comp<4, 2>(v, s); }
template <> __device__ __forceinline__ void dft<9>(real2 *v, real s) { dft_comp<3, 3>(v, s); }
template <> __device__ __forceinline__ void dft<10>(real2 *v, real s) { dft_comp<5, 2>(v, s); }
template <> __device__ __forceinline__ void dft<12>(real2 *v, real s) { dft_comp<4, 3>(v, s); }
template <> __device__ __forceinline__ void dft<15>(real2 *v, real s) { dft_comp<5, 3>(v, s); }
template <> __device__ __forceinline__ void dft<16>(real2 *v, real s) { dft_comp<4, 4>(v, s); }

// One Stockham pass of radix R over nseq sequences of length n, in place in
// LDS: element e of sequence s at data[s*sstride + e*estride].  Lanes run
// over the sequences fastest, so with estride = nseq (sequences interleaved)
// every access of a pass is a contiguous run, and with sstride odd and
// estride = 1 neighbouring lanes fall in different banks.  Each thread keeps
// its butterflies' inputs in registers across the barrier; a block of BLK
// threads transforms at most FFT_TILE complex values, so a thread owns at most
// ceil(FFT_TILE / (BLK R)) butterflies.
template <int R>
__device__ __forceinline__ void fft_pass(real2 *data, int nseq, int sstride, int estride, int n,
                                         int Ns, unsigned long long nsm, const real2 *tw,
                                         real sign)
{
    constexpr int NB = (FFT_TILE / BLK + R - 1) / R;
    const int nb = n / R;
    const int twstep = magic_div(nb, nsm);      // n / (Ns R)
    real2 v[NB][R];
    int off[NB];                                // output position of the butterfly
    Walk w(nseq);                               // w.q = butterfly index j, w.r = sequence s
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int j = w.q, s = w.r;
        off[i] = -1;
        if (j < nb) {
            const real2 *src = data + s * sstride;
            const int jq = magic_div(j, nsm), k = j - jq * Ns;
            off[i] = s * sstride + (jq * Ns * R + k) * estride;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                real2 a = src[(j + q * nb) * estride];
                if (q > 0) {
                    real2 t = tw[q * k * twstep];
                    t.y *= sign;
                    a = cmul(a, t);
                }
                v[i][q] = a;
            }
            dft<R>(v[i], sign);
        }
        w.next();
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if (off[i] >= 0) {
            real2 *dst = data + off[i];
#pragma unroll
            for (int q = 0; q < R; ++q) dst[q * Ns * estride] = v[i][q];
        }
    }
    __syncthreads();
}

// tw: the plan's twiddles exp(2 pi i m / n), m < n, staged in LDS
__device__ __forceinline__ void lds_fft(real2 *data, int nseq, int sstride, int estride,
                                        const PlanDev &p, const real2 *tw, real sign)
{
    int Ns = 1;
    for (int pass = 0; pass < p.nf; ++pass) {
        const int r = p.f[pass];
        const unsigned long long m = p.nsm[pass];
        switch (r) {
        case 16: fft_pass<16>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        case 12: fft_pass<12>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        case 8: fft_pass<8>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        case 6: fft_pass<6>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        case 5: fft_pass<5>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        case 4: fft_pass<4>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        case 3: fft_pass<3>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign); break;
        default: fft_pass<2>(data, nseq, sstride, estride, p.n, Ns, m, tw, sign);
        }
        Ns *= r;
    }
}

// copy the plan's twiddle table into LDS (visible after the caller's barrier)
__device__ __forceinline__ void stage_twiddles(const PlanDev &p, real2 *twl)
{
    for (int m = threadIdx.x; m < p.n; m += BLK) twl[m] = to_real2(p.tw[m]);
}

// ---------------------------------------------------------------- fixed lengths
// Transform lengths with a compile-time plan (k_zinv<PZ>, k_xmix<MODE, PX>,
// k_zfwd<PZ> with a nonzero length): radices, strides, the tile shape and all
// index divisions fold into constants.  The same rule as the host planner:
// fewest passes over the radices below, then the smallest per-thread tile.
struct PBest {
    int passes, worst, first;
};

constexpr int FIX_RADICES[] = {16, 12, 8, 6, 5, 4, 3, 2};

constexpr int fix_tile(int r) { return (FFT_TILE / BLK + r - 1) / r * r; }

constexpr PBest plan_best(int m)
{
    if (m == 1) return PBest{0, 0, 0};
    PBest b{1 << 20, 0, 0};
    for (int r : FIX_RADICES) {
        if (m % r) continue;
        const PBest sub = plan_best(m / r);
        if (sub.passes >= (1 << 20)) continue;
        const int np = sub.passes + 1, nw = sub.worst > fix_tile(r) ? sub.worst : fix_tile(r);
        if (np < b.passes || (np == b.passes && nw < b.worst)) b = PBest{np, nw, r};
    }
    return b;
}

// radix of pass p of the length-n plan
constexpr int plan_radix(int n, int p)
{
    int m = n;
    for (int i = 0; i < p; ++i) m /= plan_best(m).first;
    return plan_best(m).first;
}

// one pass over NSEQ sequences, element e of sequence s at s*SSTR + e*ESTR
template <int R, int N, int NS, int NSEQ, int SSTR, int ESTR>
__device__ __forceinline__ void fft_pass_c(real2 *data, const real2 *tw, real sign)
{
    constexpr int NBS = N / R;                       // butterflies per sequence
    constexpr int TOTAL = NSEQ * NBS;
    constexpr int NBT = (TOTAL + BLK - 1) / BLK;     // butterflies per thread
    constexpr int TWS = N / (NS * R);
    real2 v[NBT][R];
    int off[NBT];
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
        const int b = threadIdx.x + i * BLK;
        const int j = b / NSEQ, s = b - j * NSEQ;
        off[i] = -1;
        if (b < TOTAL) {
            const real2 *src = data + s * SSTR;
            const int jq = j / NS, k = j - jq * NS;
            off[i] = s * SSTR + (jq * NS * R + k) * ESTR;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                real2 a = src[(j + q * NBS) * ESTR];
                if (q > 0 && NS > 1) {
                    real2 t = tw[q * k * TWS];
                    t.y *= sign;
                    a = cmul(a, t);
                }
                v[i][q] = a;
            }
            dft<R>(v[i], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
        if (off[i] >= 0) {
            real2 *dst = data + off[i];
#pragma unroll
            for (int q = 0; q < R; ++q) dst[q * NS * ESTR] = v[i][q];
        }
    }
    __syncthreads();
}

template <int N, int NSEQ, int SSTR, int ESTR, int P = 0, int NS = 1>
__device__ __forceinline__ void lds_fft_c(real2 *data, const real2 *tw, real sign)
{
    if constexpr (P < plan_best(N).passes) {
        constexpr int R = plan_radix(N, P);
        fft_pass_c<R, N, NS, NSEQ, SSTR, ESTR>(data, tw, sign);
        lds_fft_c<N, NSEQ, SSTR, ESTR, P + 1, NS * R>(data, tw, sign);
    }
}

// columns per block of the fixed-length z kernels (the launch needs Mx % ZXB == 0)
constexpr int ZXB = 16;
static_assert(ZXB * 96 + 96 <= FFT_TILE, "z tile of the length-96 kernels");

// rows per block of the fixed-length x kernel: the largest power of 2 with
// the tile (rows * pairs * (PX + 1) + PX twiddles) within FFT_TILE
#ifndef SC_XMIX_ZBMAX
#define SC_XMIX_ZBMAX 64
#endif
constexpr int xmix_rows(int px, int nseq)
{
    int zb = 1;
    while (2 * zb <= SC_XMIX_ZBMAX && 2 * zb * nseq * (px + 1) + px <= FFT_TILE) zb *= 2;
    return zb;
}

struct ZinvArgs {
    FieldSpec fs;
    const double *Kx, *Kz;
    int Nz, Mx, qpts, Pz, XB;
    double ax, az;                 // 2 pi / Lx, 2 pi / Lz
    const int *abort_flag;
    size_t ms_w1;                  // member stride of W1
};

constexpr int MAXE = FFT_TILE / BLK;   // LDS elements per thread of a full tile
constexpr int SBAT = 8;                // stage-in loads in flight per thread and batch
static_assert(MAXE % SBAT == 0, "stage-in batches must tile the per-thread elements");

// A Gauss-point slot value at element index i of src: FP64 complex, or
// FP32 complex (S32) where the mixed-precision rebuild stored the slots.
template <bool S32>
__device__ __forceinline__ double2 ld_slot(const cplx *src, size_t i)
{
    if constexpr (S32) {
        const float2 v = ((const float2 *) src)[i];
        return make_double2(v.x, v.y);
    } else {
        return src[i];
    }
}

// Stage a (Pz x XB) tile of one field into LDS in its memory layout
// [zp][x], zero rows for the padding, the spectral curl formed on the fly.
// A batch's loads are all issued before its first LDS store: indices are
// clamped into the array and the padding masked arithmetically, so no load
// sits behind a branch.
template <int KIND, int PZ, bool S32>
__device__ __forceinline__ void zinv_stage(const ZinvArgs &a, const cplx *src, int slot, int y,
                                           int x0, int XB, real2 *lds)
{
    // PZ > 0: the tile is ZXB x PZ and every thread stages NE values in one batch
    constexpr int NE = PZ > 0 ? (ZXB * PZ + BLK - 1) / BLK : MAXE;
    constexpr int SB = PZ > 0 ? NE : SBAT;
    const int Pz = PZ > 0 ? PZ : a.Pz, Nz = PZ > 0 ? 2 * PZ / 3 : a.Nz, Mx = a.Mx;
    const int total = XB * Pz;
    const size_t zs = (size_t) 5 * a.qpts * Mx;        // stride of kz in U
    const size_t ss = (size_t) a.qpts * Mx;            // stride of slot in U
    Walk wk(XB);                                       // (zp, xl)
#pragma unroll 1
    for (int i0 = 0; i0 < NE; i0 += SB) {
        if ((int) threadIdx.x + i0 * BLK >= total) break;
        double2 p[SB], q[SB];
        bool ok[SB];
        int kzi[SB], xi[SB];
#pragma unroll
        for (int i = 0; i < SB; ++i) {
            const int e = threadIdx.x + (i0 + i) * BLK;
            const int zp = min(wk.q, Pz - 1), xl = wk.r;
            wk.next();
            const int kz = zp < Nz / 2 ? zp : (zp > Nz ? zp - Nz / 2 : -1);
            ok[i] = e < total && kz >= 0;
            kzi[i] = max(kz, 0);
            xi[i] = x0 + xl;
            const size_t b = kzi[i] * zs + (size_t) y * Mx + xi[i];
            if (KIND == FK_SLOT) {
                p[i] = ld_slot<S32>(src, b + slot * ss);
            } else if (KIND == FK_OMEGA_Z) {
                p[i] = ld_slot<S32>(src, b + 3 * ss);
                q[i] = ld_slot<S32>(src, b + 1 * ss);
            } else if (KIND == FK_OMEGA_X) {
                p[i] = ld_slot<S32>(src, b + 4 * ss);
                q[i] = ld_slot<S32>(src, b + 1 * ss);
            } else {
                p[i] = ld_slot<S32>(src, b);
                q[i] = ld_slot<S32>(src, b + 2 * ss);
            }
        }
#pragma unroll
        for (int i = 0; i < SB; ++i) {
            const int e = threadIdx.x + (i0 + i) * BLK;
            if (e >= total) break;
            const double kx = xi[i] * a.ax;
            const int kz = kzi[i];
            const double kzv = (kz <= Nz / 2 ? kz : kz - Nz) * a.az;
            // formed in FP64 from the FP64 state, then stored as real
            double2 val;
            if (KIND == FK_SLOT) {
                val = p[i];
            } else if (KIND == FK_OMEGA_Z) {            // -du/dy + i kx v
                const double2 du = p[i], v = q[i];
                val = make_double2(-(du.x + kx * v.y), -du.y + kx * v.x);
            } else if (KIND == FK_OMEGA_X) {            // dw/dy - i kz v
                const double2 dw = p[i], v = q[i];
                val = make_double2(dw.x + kzv * v.y, dw.y - kzv * v.x);
            } else {                                    // i kz u - i kx w
                const double2 u = p[i], w = q[i];
                val = make_double2(kx * w.y - kzv * u.y, kzv * u.x - kx * w.x);
            }
            lds[e] = ok[i] ? to_real2(val) : mk2(0.0, 0.0);
        }
    }
}

// Fixed lengths: the stage-in is the first pass of the transform.  Thread
// b < ZXB NB1 takes butterfly j = b / ZXB of column xl = b % ZXB (the pass's
// own thread mapping), forms its R1 inputs (rows zp = j + q NB1) straight
// from U, transforms them in registers and writes the pass's output; groups
// q whose rows all lie in the zero band (the Nyquist row and the padding)
// are not loaded.
template <int KIND, int PZ, bool S32>
__device__ __forceinline__ void zinv_first_pass(const ZinvArgs &a, const cplx *src, int slot, int y,
                                                int x0, real2 *lds)
{
    constexpr int Nz = 2 * PZ / 3, R1 = plan_radix(PZ, 0), NB1 = PZ / R1, NBUT = ZXB * NB1;
    static_assert(NBUT <= BLK, "one first-pass butterfly per thread");
    const int b = threadIdx.x;
    if (b >= NBUT) return;
    const int j = b / ZXB, xl = b - j * ZXB, Mx = a.Mx;
    const size_t zs = (size_t) 5 * a.qpts * Mx, ss = (size_t) a.qpts * Mx;
    const size_t col = (size_t) y * Mx + x0 + xl;
    const double kx = (x0 + xl) * a.ax;
    real2 v[R1];
#pragma unroll
    for (int q = 0; q < R1; ++q) {
        if (q * NB1 >= Nz / 2 && q * NB1 + NB1 - 1 <= Nz) {
            v[q] = mk2(0.0, 0.0);
            continue;
        }
        const int zp = j + q * NB1;
        const int kz = zp < Nz / 2 ? zp : (zp > Nz ? zp - Nz / 2 : -1);
        const int kzi = max(kz, 0);
        const size_t bp = col + kzi * zs;
        const double kzv = (kzi <= Nz / 2 ? kzi : kzi - Nz) * a.az;
        double2 val;
        if (KIND == FK_SLOT) {
            val = ld_slot<S32>(src, bp + slot * ss);
        } else if (KIND == FK_OMEGA_Z) {                // -du/dy + i kx v
            const double2 du = ld_slot<S32>(src, bp + 3 * ss), vv = ld_slot<S32>(src, bp + 1 * ss);
            val = make_double2(-(du.x + kx * vv.y), -du.y + kx * vv.x);
        } else if (KIND == FK_OMEGA_X) {                // dw/dy - i kz v
            const double2 dw = ld_slot<S32>(src, bp + 4 * ss), vv = ld_slot<S32>(src, bp + 1 * ss);
            val = make_double2(dw.x + kzv * vv.y, dw.y - kzv * vv.x);
        } else {                                        // i kz u - i kx w
            const double2 u = ld_slot<S32>(src, bp), w = ld_slot<S32>(src, bp + 2 * ss);
            val = make_double2(kx * w.y - kzv * u.y, kzv * u.x - kx * w.x);
        }
        v[q] = kz >= 0 ? to_real2(val) : mk2(0.0, 0.0);
    }
    dft<R1>(v, +1.0);
    real2 *dst = lds + xl + j * R1 * ZXB;
#pragma unroll
    for (int q = 0; q < R1; ++q) dst[q * ZXB] = v[q];
}

// grid (ceil(Mx/XB), nf * members, qpts): blockIdx.y = member nf + field;
// PZ > 0: Pz == PZ, XB == ZXB and Mx % ZXB == 0
template <int PZ, bool S32>
__global__ __launch_bounds__(BLK) void k_zinv(ZinvArgs a, PlanDev pz, real2 *W1)
{
    extern __shared__ real2 lds[];
    if (*a.abort_flag) return;
    const int XB0 = PZ > 0 ? ZXB : a.XB;
    const int nf = a.fs.nf, m = blockIdx.y / nf, f = blockIdx.y - m * nf;
    const int x0 = blockIdx.x * XB0, y = blockIdx.z;
    const int XB = PZ > 0 ? ZXB : min(a.XB, a.Mx - x0);
    const int Pz = PZ > 0 ? PZ : a.Pz, Mx = a.Mx;
    const cplx *src = a.fs.src[f] + m * a.fs.ms[f];
    const int kind = a.fs.kind[f], slot = a.fs.slot[f];
    real2 *twl = lds + XB0 * Pz;
    stage_twiddles(pz, twl);
    // columns interleaved in LDS exactly as in memory: no transpose
    if constexpr (PZ > 0) {
        if (kind == FK_SLOT) zinv_first_pass<FK_SLOT, PZ, S32>(a, src, slot, y, x0, lds);
        else if (kind == FK_OMEGA_Z) zinv_first_pass<FK_OMEGA_Z, PZ, S32>(a, src, slot, y, x0, lds);
        else if (kind == FK_OMEGA_X) zinv_first_pass<FK_OMEGA_X, PZ, S32>(a, src, slot, y, x0, lds);
        else zinv_first_pass<FK_OMEGA_Y, PZ, S32>(a, src, slot, y, x0, lds);
        __syncthreads();
        lds_fft_c<PZ, ZXB, 1, ZXB, 1, plan_radix(PZ, 0)>(lds, twl, +1.0);
    } else {
        if (kind == FK_SLOT) zinv_stage<FK_SLOT, PZ, S32>(a, src, slot, y, x0, XB, lds);
        else if (kind == FK_OMEGA_Z) zinv_stage<FK_OMEGA_Z, PZ, S32>(a, src, slot, y, x0, XB, lds);
        else if (kind == FK_OMEGA_X) zinv_stage<FK_OMEGA_X, PZ, S32>(a, src, slot, y, x0, XB, lds);
        else zinv_stage<FK_OMEGA_Y, PZ, S32>(a, src, slot, y, x0, XB, lds);
        __syncthreads();
        lds_fft(lds, XB, 1, XB, pz, twl, +1.0);
    }
    real2 *dst = W1 + m * a.ms_w1 + (((size_t) y * nf + f) * Pz) * Mx;
    Walk wk(XB);
    for (int e = threadIdx.x; e < XB * Pz; e += BLK, wk.next())
        dst[(size_t) wk.q * Mx + x0 + wk.r] = lds[e];
}

struct XmixArgs {
    int mode, nf, nout, Mx, Px, Pz, qpts, ZB, check_cfl;
    int nf1;                   // fields 0..nf1-1 from the first source, the rest from W1b
    const real2 *W1b;
    int nf2;                   // XM_ADJ2: fields nf1..nf2-1 from W1b, nf2..nf-1 from W1c
    const real2 *W1c;
    real2 *W2b;                 // XM_ADJ2: the second product set
    double cfl1, cfl3;
    const double *cfl2, *W;
    double *cflmax;        // [qpts], as ordered bit patterns
    double *flow;          // XM_PHYS output
    const int *abort_flag;
    // member strides of W1, W1b, W1c, W2, W2b (grid z = member)
    size_t ms_in, ms_inb, ms_inc, ms_out, ms_outb;
};

__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v)
{
    atomicMax((unsigned long long *) addr, (unsigned long long) __double_as_longlong(v));
}

template <int MODE> struct XmixShape;
template <> struct XmixShape<XM_PRIMAL> { static constexpr int NF = 6, NOUT = 3; };
template <> struct XmixShape<XM_TANGENT> { static constexpr int NF = 12, NOUT = 3; };
template <> struct XmixShape<XM_ADJOINT> { static constexpr int NF = 9, NOUT = 6; };
template <> struct XmixShape<XM_PHYS> { static constexpr int NF = 3, NOUT = 0; };
template <> struct XmixShape<XM_ADJ2> { static constexpr int NF = 15, NOUT = 12; };

// grid (Pz/ZB, qpts, members); LDS: ZB * nseq sequences of Px complex at a stride of
// Px + 1 (odd, so lanes on neighbouring sequences use different banks),
// nseq = ceil(nf/2).  PX > 0: Px == PX and ZB == xmix_rows(PX, nseq).
template <int MODE, int PX>
__global__ __launch_bounds__(BLK) void k_xmix(XmixArgs a, PlanDev px, const real2 *W1, real2 *W2)
{
    extern __shared__ real2 lds[];
    if (*a.abort_flag) return;
    constexpr int nf = XmixShape<MODE>::NF;
    constexpr int NOUT = XmixShape<MODE>::NOUT;
    constexpr int nseq = (nf + 1) / 2;
    constexpr int nso = (NOUT + 1) / 2;
    static_assert(MODE != XM_ADJ2 || PX > 0, "the merged adjoint mode has fixed-length kernels only");
    constexpr int ZBC = PX > 0 ? xmix_rows(PX, nseq) : 1;
    const int zb = blockIdx.x, y = blockIdx.y;
    if (const int m = blockIdx.z) {
        W1 += m * a.ms_in;
        W2 += m * a.ms_out;
        a.W1b += m * a.ms_inb;
        a.W1c += m * a.ms_inc;
        a.W2b += m * a.ms_outb;
    }
    const int Px = PX > 0 ? PX : a.Px, Mx = PX > 0 ? PX / 3 : a.Mx, Pz = a.Pz;
    const int ZB = PX > 0 ? ZBC : a.ZB;
    const int SP = Px + 1;
    const int z0 = zb * ZB;
    real2 *twl = lds + ZB * nseq * SP;
    stage_twiddles(px, twl);
    // load packed pairs with Hermitian extension (c2r uses Re only at kx=0):
    // bin n < Mx takes X[n] + i Y[n], bin n > Px - Mx the conjugates at
    // Px - n, the bins between are zero.  Loads unpredicated (index clamped
    // into the row), all issued before the LDS stores.
    if constexpr (PX > 0) {
        // fixed length: the load stage is the first pass of the inverse
        // transform.  Thread b < NSQ NB1 takes butterfly j = b / NSQ of packed
        // row t = b % NSQ = r nseq + s (the pass's own thread mapping), loads
        // its R1 bins n = j + q NB1 straight from the fields (32-bit offsets
        // from two block bases: fields below nf1 in W1, the others in W1b;
        // pairs never straddle), applies the Hermitian extension and the
        // radix-R1 transform in registers and writes the pass's output.
        // Groups q whose bins all lie in the zero band are not loaded.
        constexpr int MX = PX / 3, R1 = plan_radix(PX, 0), NB1 = PX / R1;
        constexpr int NSQ = ZBC * nseq, NBUT = NSQ * NB1;
        static_assert(NBUT <= BLK, "one first-pass butterfly per thread");
        const int nf1 = a.nf1, PM = Pz * Mx;
        const int b = threadIdx.x;
        if (b < NBUT) {
            const int j = b / NSQ, t = b - j * NSQ, r = t / nseq, s2 = 2 * (t - r * nseq);
            const int nf2 = MODE == XM_ADJ2 ? a.nf2 : nf;
            const int src = s2 < nf1 ? 0 : s2 < nf2 ? 1 : 2;
            const int f0 = src == 0 ? 0 : src == 1 ? nf1 : nf2;
            const int fn = src == 0 ? nf1 : src == 1 ? nf2 - nf1 : nf - nf2;
            const real2 *B = src == 0 ? W1 : src == 1 ? a.W1b : a.W1c;
            const real2 *X = B + ((size_t) y * fn * Pz + z0) * Mx + ((s2 - f0) * Pz + r) * Mx;
            const bool hasy = s2 + 1 < nf;
            const int yo = hasy ? PM : 0;
            real2 v[R1];
#pragma unroll
            for (int q = 0; q < R1; ++q) {
                if (q * NB1 >= MX && q * NB1 + NB1 - 1 <= PX - MX) {
                    v[q] = mk2(0.0, 0.0);             // zero band for every j
                    continue;
                }
                const int n = j + q * NB1;
                const int k = min(n < MX ? n : PX - n, MX - 1);
                const real2 xv = X[k], yl = X[yo + k];
                const real2 yv = hasy ? yl : mk2(0.0, 0.0);
                real2 val = mk2(0.0, 0.0);
                if (n == 0)
                    val = mk2(xv.x, yv.x);
                else if (n < MX)
                    val = mk2(xv.x - yv.y, xv.y + yv.x);
                else if (n > PX - MX)
                    val = mk2(xv.x + yv.y, -xv.y + yv.x);
                v[q] = val;
            }
            dft<R1>(v, +1.0);
            real2 *dst = lds + t * (PX + 1) + j * R1;
#pragma unroll
            for (int q = 0; q < R1; ++q) dst[q] = v[q];
        }
    } else {
        const int tot = ZB * nseq * Px;
        constexpr int MAXL = (XMIX_TILE + BLK - 1) / BLK;
        constexpr int LB = 4;                              // loads in flight per batch
        Walk wl(Px);                                       // (packed row t, bin n)
#pragma unroll 1
        for (int i0 = 0; i0 < MAXL; i0 += LB) {
            if ((int) threadIdx.x + i0 * BLK >= tot) break;
            real2 lx[LB], ly[LB];
            int tn[LB];
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const int t = min(wl.q, ZB * nseq - 1), n = wl.r;
                tn[i] = wl.q * Px + n;
                wl.next();
                const int s = t % nseq, r = t / nseq;
                const int k = min(n < Mx ? n : Px - n, Mx - 1);
                const int f0 = 2 * s, nf1 = a.nf1;
                const real2 *X = f0 < nf1 ? W1 + (((size_t) y * nf1 + f0) * Pz + z0 + r) * Mx
                                         : a.W1b + (((size_t) y * (nf - nf1) + f0 - nf1) * Pz + z0 + r) * Mx;
                const int sy = min(2 * s + 1, nf - 1) - 2 * s;     // 1, or 0 for a lone field
                lx[i] = X[k];
                ly[i] = X[(size_t) sy * Pz * Mx + k];
            }
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const int e = threadIdx.x + (i0 + i) * BLK;
                if (e >= tot) break;
                const int t = tn[i] / Px, n = tn[i] - t * Px;
                const int s = t % nseq, r = t / nseq;
                const bool hasy = 2 * s + 1 < nf;
                real2 xv = lx[i], yv = hasy ? ly[i] : mk2(0.0, 0.0);
                real2 val;
                if (n == 0)
                    val = mk2(xv.x, yv.x);
                else if (n < Mx)
                    val = mk2(xv.x - yv.y, xv.y + yv.x);
                else if (n > Px - Mx)
                    val = mk2(xv.x + yv.y, -xv.y + yv.x);
                else
                    val = mk2(0.0, 0.0);
                lds[(r * nseq + s) * SP + n] = val;
            }
        }
    }
    __syncthreads();
    if constexpr (PX > 0)
        lds_fft_c<PX, ZBC * nseq, PX + 1, 1, 1, plan_radix(PX, 0)>(lds, twl, +1.0);
    else
        lds_fft(lds, ZB * nseq, SP, 1, px, twl, +1.0);
    // pointwise stage: read all fields of a point, compute, then (after a
    // barrier) write the packed outputs compactly: row r, pair t at sequence
    // r*nso + t
    constexpr int MAXP = PX > 0 ? (ZBC * PX + BLK - 1) / BLK          // points per thread
                                : (XMIX_TILE / nseq + BLK - 1) / BLK;
    constexpr int NO = NOUT > 0 ? NOUT : 1;
    real outv[MAXP][NO];
    double cmax = 0.0;
    const int npt = ZB * Px;
    Walk wp(Px);                                       // (z row r, point n)
#pragma unroll
    for (int i = 0; i < MAXP; ++i) {
        int e = threadIdx.x + i * BLK;
        if (e >= npt) break;
        const int n = wp.r, r = wp.q;
        wp.next();
        const real2 *row = lds + (size_t) r * nseq * SP + n;
        real f[16];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s < nseq) {
                real2 v = row[s * SP];
                f[2 * s] = v.x;
                f[2 * s + 1] = v.y;
            } else {
                f[2 * s] = 0.0;
                f[2 * s + 1] = 0.0;
            }
        }
        real u = f[0], v = f[1], w = f[2], oz = f[3], ox = f[4], oy = f[5];
        if (a.check_cfl) {
            double t = fabs(u) * a.cfl1 + fabs(v) * a.cfl2[y] + fabs(w) * a.cfl3;
            cmax = fmax(cmax, t);
        }
        if constexpr (MODE == XM_PRIMAL) {
            outv[i][1] = v * oz - w * oy;           // Hx
            outv[i][2] = w * ox - u * oz;           // Hy
            outv[i][0] = u * oy - v * ox;           // Hz
        } else if constexpr (MODE == XM_TANGENT) {
            real Iu = f[6], Iv = f[7], Iw = f[8], Ioz = f[9], Iox = f[10], Ioy = f[11];
            outv[i][1] = Iv * oz - Iw * oy + v * Ioz - w * Ioy;
            outv[i][2] = Iw * ox - Iu * oz + w * Iox - u * Ioz;
            outv[i][0] = Iu * oy - Iv * ox + u * Ioy - v * Iox;
        } else if constexpr (MODE == XM_ADJOINT) {
            real lx = f[6], ly = f[7], lz = f[8];
            outv[i][0] = v * lz - w * ly;           // Sx
            outv[i][1] = w * lx - u * lz;           // Sy
            outv[i][2] = u * ly - v * lx;           // Sz
            outv[i][3] = ly * oz - lz * oy;         // Tx
            outv[i][4] = lz * ox - lx * oz;         // Ty
            outv[i][5] = lx * oy - ly * ox;         // Tz
        } else if constexpr (MODE == XM_ADJ2) {
            // the adjoint products of lambda with two states: (u, omega) in
            // f[0..5] and f[6..11], lambda in f[12..14]
            real lx = f[12], ly = f[13], lz = f[14];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const real pu = f[6 * p], pv = f[6 * p + 1], pw = f[6 * p + 2];
                const real qz = f[6 * p + 3], qx = f[6 * p + 4], qy = f[6 * p + 5];
                outv[i][6 * p + 0] = pv * lz - pw * ly;     // Sx
                outv[i][6 * p + 1] = pw * lx - pu * lz;     // Sy
                outv[i][6 * p + 2] = pu * ly - pv * lx;     // Sz
                outv[i][6 * p + 3] = ly * qz - lz * qy;     // Tx
                outv[i][6 * p + 4] = lz * qx - lx * qz;     // Ty
                outv[i][6 * p + 5] = lx * qy - ly * qx;     // Tz
            }
        } else if constexpr (MODE == XM_PHYS) {
            int zp = z0 + r;
            size_t plane = (size_t) Px * a.qpts * Pz;
            size_t o = ((size_t) n * a.qpts + y) * Pz + zp;
            a.flow[o] = u;
            a.flow[plane + o] = v;
            a.flow[2 * plane + o] = w;
        }
    }
    if (a.check_cfl) {
        // block maximum: within each wave by lane exchanges, then the waves'
        // maxima through LDS (the maximum is exact in any order)
        __shared__ double cw[BLK / 64];
        double m = cmax;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
        if ((threadIdx.x & 63) == 0) cw[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = cw[0];
#pragma unroll
            for (int w = 1; w < BLK / 64; ++w) b = fmax(b, cw[w]);
            atomic_max_nonneg(&a.cflmax[y], b);
        }
    }
    if constexpr (MODE == XM_PHYS) return;
    if constexpr (NOUT > 0) {
    __syncthreads();
    wp = Walk(Px);
#pragma unroll
    for (int i = 0; i < MAXP; ++i) {
        int e = threadIdx.x + i * BLK;
        if (e >= npt) break;
        const int n = wp.r, r = wp.q;
        wp.next();
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            if (t < nso) {
                real re = outv[i][2 * t < NO ? 2 * t : 0];
                real im = (2 * t + 1 < NOUT) ? outv[i][2 * t + 1 < NO ? 2 * t + 1 : 0] : (real) 0;
                lds[(size_t) (r * nso + t) * SP + n] = mk2(re, im);
            }
        }
    }
    __syncthreads();
    if constexpr (PX > 0)
        lds_fft_c<PX, ZBC * nso, PX + 1, 1>(lds, twl, -1.0);
    else
        lds_fft(lds, ZB * nso, SP, 1, px, twl, -1.0);
    // unpack two real transforms from each packed pair, bins kx < Mx
    const int tot2 = ZB * nso * Mx;
    Walk wu(Mx);                                       // (packed row, bin k)
    for (int e = threadIdx.x; e < tot2; e += BLK, wu.next()) {
        const int k = wu.r, t = wu.q % nso, r = wu.q / nso;
        const real2 *Z = lds + (size_t) (r * nso + t) * SP;
        real2 zk = Z[k], zm = Z[k == 0 ? 0 : Px - k];
        real2 X = mk2((real) 0.5 * (zk.x + zm.x), (real) 0.5 * (zk.y - zm.y));
        real2 Yv = mk2((real) 0.5 * (zk.y + zm.y), (real) -0.5 * (zk.x - zm.x));
        int zp = z0 + r;
        int g0 = 2 * t, g1 = 2 * t + 1;
        if constexpr (MODE == XM_ADJ2) {
            // pairs t < 3 to W2, the others to W2b, six fields each
            real2 *D = t < 3 ? W2 : a.W2b;
            const int h0 = t < 3 ? g0 : g0 - 6;
            D[(((size_t) y * 6 + h0) * Pz + zp) * Mx + k] = X;
            D[(((size_t) y * 6 + h0 + 1) * Pz + zp) * Mx + k] = Yv;
        } else {
            W2[(((size_t) y * NOUT + g0) * Pz + zp) * Mx + k] = X;
            if (g1 < NOUT) W2[(((size_t) y * NOUT + g1) * Pz + zp) * Mx + k] = Yv;
        }
    }
    }
}

struct ZfwdArgs {
    int nout, Nz, Mx, Pz, qpts, XB, dst_kind;
    double norm;
    const int *abort_flag;
    // cfl_slot >= 0: the first block runs the CFL check of the products
    // kernel's maxima first (in place of a separate launch)
    int cfl_slot;
    double *cflmax;
    double dt;
    int *abort_rw;
    size_t ms_in, ms_out;      // member strides of W2 and dst
    int p32;                   // ZF_USLOTS stored as FP32 complex (same element index)
};

// a product slot value at element index i of U: FP64 or FP32 complex
template <class V>
__device__ __forceinline__ void st_slot(cplx *U, size_t i, bool p32, V re, V im)
{
    if (p32)
        ((float2 *) U)[i] = make_float2((float) re, (float) im);
    else
        U[i] = make_double2((double) re, (double) im);
}

// grid (ceil(Mx/XB), nout * members, qpts): blockIdx.y = member nout + field;
// PZ > 0: Pz == PZ, XB == ZXB and Mx % ZXB == 0
template <int PZ>
__global__ __launch_bounds__(BLK) void k_zfwd(ZfwdArgs a, PlanDev pz, const real2 *W2, cplx *dst)
{
    extern __shared__ real2 lds[];
    if (a.cfl_slot >= 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0)
        cfl_check_block(a.cflmax, a.qpts, a.dt, a.abort_rw, a.cfl_slot);
    if (*a.abort_flag) return;
    const int XB0 = PZ > 0 ? ZXB : a.XB;
    const int m = blockIdx.y / a.nout, g = blockIdx.y - m * a.nout;
    const int x0 = blockIdx.x * XB0, y = blockIdx.z;
    const int XB = PZ > 0 ? ZXB : min(a.XB, a.Mx - x0);
    const int Nz = PZ > 0 ? 2 * PZ / 3 : a.Nz, Pz = PZ > 0 ? PZ : a.Pz, Mx = a.Mx;
    const real2 *src = W2 + m * a.ms_in + (((size_t) y * a.nout + g) * Pz) * Mx;
    dst += m * a.ms_out;
    real2 *twl = lds + XB0 * Pz;
    stage_twiddles(pz, twl);
    {
        // a batch's loads all issued before its LDS stores
        constexpr int NE = PZ > 0 ? (ZXB * PZ + BLK - 1) / BLK : MAXE;
        constexpr int SB = PZ > 0 ? NE : SBAT;
        const int total = XB * Pz;
        Walk wk(XB);
#pragma unroll 1
        for (int i0 = 0; i0 < NE; i0 += SB) {
            if ((int) threadIdx.x + i0 * BLK >= total) break;
            real2 v[SB];
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                const int zp = min(wk.q, Pz - 1), xl = wk.r;
                wk.next();
                v[i] = src[(size_t) zp * Mx + x0 + xl];
            }
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                const int e = threadIdx.x + (i0 + i) * BLK;
                if (e < total) lds[e] = v[i];
            }
        }
    }
    __syncthreads();
    if constexpr (PZ > 0)
        lds_fft_c<PZ, ZXB, 1, ZXB>(lds, twl, -1.0);
    else
        lds_fft(lds, XB, 1, XB, pz, twl, -1.0);
    Walk wt(XB);
    for (int e = threadIdx.x; e < XB * (Nz - 1); e += BLK, wt.next()) {
        const int r = wt.q, xl = wt.r;
        int kz = r < Nz / 2 ? r : r + 1;             // skip the Nyquist row
        int zp = kz < Nz / 2 ? kz : kz + Nz / 2;
        real2 v = lds[zp * XB + xl];
        v.x *= (real) a.norm;
        v.y *= (real) a.norm;
        if (a.dst_kind == ZF_USLOTS)
            st_slot(dst, (((size_t) kz * 5 + g) * a.qpts + y) * Mx + x0 + xl, a.p32, v.x, v.y);
        else
            dst[(((size_t) g * Nz + kz) * a.qpts + y) * Mx + x0 + xl] = to_cplx(v);
    }
}

// Forward z transforms of the six adjoint products with the assembly fused
// (channel/pass2.c:417-566): a block transforms all six fields of ZA
// columns, interleaved in LDS as 6 ZA sequences ([zp][field][x], no
// transpose), and forms the five state slots at the stored kz rows:
// Hz = -Tz - i kx Sy, Hx = -Tx + i kz Sy, Hy = -Ty - i kz Sx + i kx Sz, then
// S_x, S_z (each field scaled first, as the separate transform does).
// grid (Mx / ZA, qpts); PZ > 0: Pz == PZ and Mx % ZA == 0.
constexpr int ZA = 4;
struct ZadjArgs {
    int Nz, Mx, Pz, qpts, XA;
    double norm;
    const double *Kx, *Kz;
    const int *abort_flag;
    // k_zfwd_adj2 with nprob = 2: a second problem (W2b -> Ub) in the same grid
    int nprob;
    const real2 *W2b;
    cplx *Ub;
    // members: k_zfwd_adj grid z = member; k_zfwd_adj2 tiles of member m
    // follow those of member m - 1 (strides of W2 / W2b and U / Ub)
    int nm;
    size_t ms_in, ms_out;
    int p32;                   // the five slots stored as FP32 complex
};

template <int PZ>
__global__ __launch_bounds__(BLK) void k_zfwd_adj(ZadjArgs a, PlanDev pz, const real2 *W2, cplx *U)
{
    extern __shared__ real2 lds[];
    if (*a.abort_flag) return;
    const int XA = PZ > 0 ? ZA : a.XA;
    const int x0 = blockIdx.x * XA, y = blockIdx.y;
    W2 += blockIdx.z * a.ms_in;
    U += blockIdx.z * a.ms_out;
    const int XB = PZ > 0 ? ZA : min(a.XA, a.Mx - x0);   // columns of this block
    const int Nz = PZ > 0 ? 2 * PZ / 3 : a.Nz, Pz = PZ > 0 ? PZ : a.Pz, Mx = a.Mx;
    const int NS = 6 * XA;                               // interleaved sequences
    real2 *twl = lds + NS * Pz;
    stage_twiddles(pz, twl);
    {
        // element (zp, field f, column xl) at (zp * 6 + f) * XA + xl
        const int total = 6 * Pz * XA;
        const real2 *src = W2 + (size_t) y * 6 * Pz * Mx;
        Walk wk(XA);                                     // (zp * 6 + f, xl)
        constexpr int NE = PZ > 0 ? (6 * ZA * PZ + BLK - 1) / BLK : MAXE;
        constexpr int SB = PZ > 0 ? NE : SBAT;
#pragma unroll 1
        for (int i0 = 0; i0 < NE; i0 += SB) {
            if ((int) threadIdx.x + i0 * BLK >= total) break;
            real2 v[SB];
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                const int q = min(wk.q, 6 * Pz - 1), xl = min(wk.r, XB - 1);
                wk.next();
                const int zp = q / 6, f = q - zp * 6;
                v[i] = src[((size_t) f * Pz + zp) * Mx + x0 + xl];
            }
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                const int e = threadIdx.x + (i0 + i) * BLK;
                if (e < total) lds[e] = v[i];
            }
        }
    }
    __syncthreads();
    if constexpr (PZ > 0)
        lds_fft_c<PZ, 6 * ZA, 1, 6 * ZA>(lds, twl, -1.0);
    else
        lds_fft(lds, NS, 1, NS, pz, twl, -1.0);
    const size_t ss = (size_t) a.qpts * Mx;
    Walk wt(XA);                                         // (stored kz row r, xl)
    for (int e = threadIdx.x; e < XA * (Nz - 1); e += BLK, wt.next()) {
        const int r = wt.q, xl = wt.r;
        if (xl >= XB) continue;
        const int kz = r < Nz / 2 ? r : r + 1;           // skip the Nyquist row
        const int zp = kz < Nz / 2 ? kz : kz + Nz / 2;
        real2 S[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            S[f] = lds[(zp * 6 + f) * XA + xl];
            S[f].x *= (real) a.norm;
            S[f].y *= (real) a.norm;
        }
        const int x = x0 + xl;
        const double kx = a.Kx[x], kzv = a.Kz[kz];
        const real2 Sx = S[0], Sy = S[1], Sz = S[2], Tx = S[3], Ty = S[4], Tz = S[5];
        const size_t b = ((size_t) kz * 5 * a.qpts + y) * Mx + x;
        st_slot(U, b + 0 * ss, a.p32, -Tz.x + kx * Sy.y, -Tz.y - kx * Sy.x);
        st_slot(U, b + 1 * ss, a.p32, -Tx.x - kzv * Sy.y, -Tx.y + kzv * Sy.x);
        st_slot(U, b + 2 * ss, a.p32, -Ty.x + kzv * Sx.y - kx * Sz.y, -Ty.y - kzv * Sx.x + kx * Sz.x);
        st_slot(U, b + 3 * ss, a.p32, Sx.x, Sx.y);
        st_slot(U, b + 4 * ss, a.p32, Sz.x, Sz.y);
    }
}

// Fixed-length form of k_zfwd_adj in two phases over one 3-field tile:
// S = (Sx, Sy, Sz) of the ZA columns is transformed first and each thread
// keeps the values of its (stored kz row, column); then T = (Tx, Ty, Tz) is
// transformed in the same LDS and the thread assembles its five slots.  The
// tile is half that of k_zfwd_adj, so five blocks share a CU and the
// (Mx/ZA) x qpts grid runs in one round; the other blocks of the CU cover a
// block's loads.  1-D grid over the (y, column group) tiles, XCD-aware: the
// column groups of a y share an XCD, so the two halves of a 128-byte line
// (ZA = 4 columns are 64 bytes) meet in one L2.
template <int PZ>
__global__ __launch_bounds__(BLK) void k_zfwd_adj2(ZadjArgs a, PlanDev pz, const real2 *W2, cplx *U)
{
    static_assert(PZ > 0, "fixed lengths only");
    constexpr int Pz = PZ, Nz = 2 * PZ / 3, NS = 3 * ZA, TOT = NS * Pz;
    constexpr int NE = (TOT + BLK - 1) / BLK;
    static_assert(ZA * (Nz - 1) <= BLK, "one output column-row per thread");
    extern __shared__ real2 lds[];
    if (*a.abort_flag) return;
    const int Mx = a.Mx, nxg = Mx / ZA, n1 = nxg * a.qpts;
    int tile = xcd_tile(a.nm * a.nprob * n1);
    if (tile < 0) return;
    const int prob = tile / n1, m = prob / a.nprob;
    tile -= prob * n1;
    if (prob - m * a.nprob) {
        // the second problem's tiles
        W2 = a.W2b;
        U = a.Ub;
    }
    W2 += m * a.ms_in;
    U += m * a.ms_out;
    const int y = tile / nxg, x0 = (tile - y * nxg) * ZA;
    real2 *twl = lds + TOT;
    stage_twiddles(pz, twl);
    // element (zp, f, xl) of a phase at (zp * 3 + f) * ZA + xl; phase h
    // holds fields 3h .. 3h + 2 of W2[y]
    const real2 *src = W2 + (size_t) y * 6 * Pz * Mx;
    const int r = threadIdx.x / ZA, xl = threadIdx.x - r * ZA;
    const bool own = r < Nz - 1;
    const int kz = r < Nz / 2 ? r : r + 1;               // skip the Nyquist row
    const int zp = kz < Nz / 2 ? kz : kz + Nz / 2;
    // phase h: fields 3h .. 3h + 2 into LDS, transform, this thread's values
    auto phase = [&](int h, real2 *P) {
        real2 v[NE];
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = min((int) threadIdx.x + i * BLK, TOT - 1);
            const int q = e / ZA, c = e - q * ZA, zq = q / 3, f = q - zq * 3;
            v[i] = src[((size_t) (3 * h + f) * Pz + zq) * Mx + x0 + c];
        }
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = threadIdx.x + i * BLK;
            if (e < TOT) lds[e] = v[i];
        }
        __syncthreads();
        lds_fft_c<PZ, NS, 1, NS>(lds, twl, -1.0);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            P[f] = own ? lds[(zp * 3 + f) * ZA + xl] : mk2(0.0, 0.0);
            P[f].x *= (real) a.norm;
            P[f].y *= (real) a.norm;
        }
    };
    real2 S[3], T[3];
    phase(0, S);
    __syncthreads();                                     // S read before T is stored
    phase(1, T);
    if (!own) return;
    const int x = x0 + xl;
    const double kx = a.Kx[x], kzv = a.Kz[kz];
    const real2 Sx = S[0], Sy = S[1], Sz = S[2], Tx = T[0], Ty = T[1], Tz = T[2];
    const size_t ss = (size_t) a.qpts * Mx;
    const size_t b = ((size_t) kz * 5 * a.qpts + y) * Mx + x;
    st_slot(U, b + 0 * ss, a.p32, -Tz.x + kx * Sy.y, -Tz.y - kx * Sy.x);
    st_slot(U, b + 1 * ss, a.p32, -Tx.x - kzv * Sy.y, -Tx.y + kzv * Sy.x);
    st_slot(U, b + 2 * ss, a.p32, -Ty.x + kzv * Sx.y - kx * Sz.y, -Ty.y - kzv * Sx.x + kx * Sz.x);
    st_slot(U, b + 3 * ss, a.p32, Sx.x, Sx.y);
    st_slot(U, b + 4 * ss, a.p32, Sz.x, Sz.y);
}

// adjoint assembly (channel/pass2.c:474-566):
// Hz = -Tz - i kx Sy, Hx = -Tx + i kz Sy, Hy = -Ty - i kz Sx + i kx Sz
// grid (blocks, members): SP and U of member blockIdx.y at the given strides
__global__ void k_assemble_adj(const cplx *SP, cplx *U, const double *Kx, const double *Kz,
                               int Nz, int qpts, int Mx, const int *abort_flag, size_t ms_sp,
                               size_t ms_u, int p32)
{
    if (*abort_flag) return;
    SP += blockIdx.y * ms_sp;
    U += blockIdx.y * ms_u;
    size_t n = (size_t) Nz * qpts * Mx;
    for (size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x; e < n;
         e += (size_t) gridDim.x * blockDim.x) {
        int x = e % Mx;
        int y = (e / Mx) % qpts;
        int z = e / ((size_t) Mx * qpts);
        if (z == Nz / 2) continue;
        double kx = Kx[x], kz = Kz[z];
        double2 Sx = SP[e], Sy = SP[n + e], Sz = SP[2 * n + e];
        double2 Tx = SP[3 * n + e], Ty = SP[4 * n + e], Tz = SP[5 * n + e];
        size_t b = ((size_t) z * 5 * qpts + y) * Mx + x, ss = (size_t) qpts * Mx;
        st_slot(U, b + 0 * ss, p32, -Tz.x + kx * Sy.y, -Tz.y - kx * Sy.x);
        st_slot(U, b + 1 * ss, p32, -Tx.x - kz * Sy.y, -Tx.y + kz * Sy.x);
        st_slot(U, b + 2 * ss, p32, -Ty.x + kz * Sx.y - kx * Sz.y, -Ty.y - kz * Sx.x + kx * Sz.x);
        st_slot(U, b + 3 * ss, p32, Sx.x, Sx.y);
        st_slot(U, b + 4 * ss, p32, Sz.x, Sz.y);
    }
}

// ---------------------------------------------------------------- launches
int xmix_zblock(Context &c, int nf);

// fixed-length instances of the z kernels: Pz of 48 or 96, Nx/2 a multiple of ZXB
typedef void (*ZinvKernel)(ZinvArgs, PlanDev, real2 *);
typedef void (*ZfwdKernel)(ZfwdArgs, PlanDev, const real2 *, cplx *);

static int z_fixed(const Dims &d)
{
    if (d.Mx % ZXB) return 0;
    return d.Pz == 96 ? 96 : d.Pz == 48 ? 48 : 0;
}

static int pick_block(int Mx, int P, int nper)
{
    // columns per block so that the LDS tile plus the P twiddles stay within
    // FFT_TILE complex
    int xb = Mx;
    while (xb > 1 && (size_t) xb * P * nper + P > FFT_TILE) xb = (xb + 1) / 2;
    return xb;
}

void launch_zinv(Context &c, const FieldSpec &fs, cplx *W1)
{
    const Dims &d = c.d;
    ZinvArgs a;
    a.fs = fs;
    a.Kx = c.dKx;
    a.Kz = c.dKz;
    a.Nz = d.Nz;
    a.Mx = d.Mx;
    a.qpts = d.qpts;
    a.Pz = d.Pz;
    const int fixed = z_fixed(d);
    a.XB = fixed ? ZXB : pick_block(d.Mx, d.Pz, 1);
    a.ax = 2. * M_PI / d.Lx;
    a.az = 2. * M_PI / d.Lz;
    a.abort_flag = c.abort_flag;
    for (int f = 0; f < fs.nf; ++f) a.fs.ms[f] = mstride(c, fs.src[f]);
    a.ms_w1 = rstride(c, W1);
    const int nm = c.nrun;
    dim3 grid((d.Mx + a.XB - 1) / a.XB, fs.nf * nm, d.qpts);
    size_t lds = sizeof(real2) * (a.XB * d.Pz + d.Pz);
    // the FP32 build reads FP32 slots where the rebuild wrote them
#ifdef SC_FFT_F32
    const bool s32 = slots32(c);
#else
    const bool s32 = false;
#endif
    const double plane = (double) d.qpts * d.Mx;
    int slots = 0;
    for (int f = 0; f < fs.nf; ++f) {
        bool seen = false;
        for (int g = 0; g < f; ++g) seen |= fs.src[g] == fs.src[f];
        if (!seen) slots += fs.nf >= 6 ? 5 : 3;
    }
    if (fs.nf == 9) slots = 8;            // adjoint: 5 state slots + 3 adjoint
    if (fs.nf == 6 && fs.kind[3] == FK_SLOT) slots = 6;   // two velocity triples
    double bytes = nm * plane * ((s32 ? 8.0 : 16.0) * slots * (d.Nz - 1) +
                                 (double) sizeof(real2) * fs.nf * d.Pz);
    const hipStream_t st = c.stream;
    const PlanDev pd = to_dev(c.pz);
    const ZinvKernel kern =
        s32 ? (fixed == 96 ? k_zinv<96, true> : fixed == 48 ? k_zinv<48, true> : k_zinv<0, true>)
            : (fixed == 96 ? k_zinv<96, false> : fixed == 48 ? k_zinv<48, false> : k_zinv<0, false>);
    real2 *W1r = (real2 *) W1;
    auto go = [=]() { hipLaunchKernelGGL(kern, grid, dim3(BLK), lds, st, a, pd, W1r); };
    prof_begin(c, st, 1, bytes);
    go();
    prof_end(c, st, 1, go);
}

typedef void (*XmixKernel)(XmixArgs, PlanDev, const real2 *, real2 *);

template <int MODE>
static XmixKernel xmix_kernel_mode(int Px, int nseq, int zb)
{
    // the fixed-length instance needs the row block it was compiled for
    if (Px == 96 && zb == xmix_rows(96, nseq)) return k_xmix<MODE, 96>;
    if (Px == 48 && zb == xmix_rows(48, nseq)) return k_xmix<MODE, 48>;
    return k_xmix<MODE, 0>;
}

static XmixKernel xmix_kernel(int mode, int Px, int nseq, int zb)
{
    switch (mode) {
    case XM_PRIMAL: return xmix_kernel_mode<XM_PRIMAL>(Px, nseq, zb);
    case XM_TANGENT: return xmix_kernel_mode<XM_TANGENT>(Px, nseq, zb);
    case XM_ADJOINT: return xmix_kernel_mode<XM_ADJOINT>(Px, nseq, zb);
    default: return xmix_kernel_mode<XM_PHYS>(Px, nseq, zb);
    }
}

void launch_xmix(Context &c, int mode, const cplx *W1, cplx *W2, int check_cfl,
                 double *flow, const cplx *W1b, int nf1)
{
    const Dims &d = c.d;
    XmixArgs a;
    a.mode = mode;
    a.nf = mode == XM_PRIMAL ? 6 : mode == XM_TANGENT ? 12 : mode == XM_ADJOINT ? 9 : 3;
    a.nout = mode == XM_ADJOINT ? 6 : (mode == XM_PRIMAL || mode == XM_TANGENT) ? 3 : 0;
    a.nf1 = W1b ? nf1 : a.nf;
    a.W1b = (const real2 *) (W1b ? W1b : W1);
    a.nf2 = a.nf;
    a.W1c = nullptr;
    a.W2b = nullptr;
    a.Mx = d.Mx;
    a.Px = d.Px;
    a.Pz = d.Pz;
    a.qpts = d.qpts;
    int nseq = (a.nf + 1) / 2;
    int zb = SC_FFT_NS::xmix_zblock(c, a.nf);
    a.ZB = zb;
    a.check_cfl = check_cfl;
    a.cfl1 = d.cfl1;
    a.cfl3 = d.cfl3;
    a.cfl2 = c.dcfl2;
    a.W = c.dW;
    a.cflmax = c.cflmax;
    a.flow = flow;
    a.abort_flag = c.abort_flag;
    a.ms_in = rstride(c, W1);
    a.ms_inb = rstride(c, a.W1b);
    a.ms_inc = 0;
    a.ms_out = rstride(c, W2);
    a.ms_outb = 0;
    // the physical output exists for one member
    const int nm = mode == XM_PHYS ? 1 : c.nrun;
    dim3 grid(d.Pz / zb, d.qpts, nm);
    size_t lds = sizeof(real2) * (zb * nseq * (d.Px + 1) + d.Px);
    PlanDev pd = to_dev(c.px);
    const double rows = (double) sizeof(real2) * d.qpts * d.Pz * d.Mx;
    double bytes = nm * rows * (a.nf + a.nout);
    if (mode == XM_PHYS) bytes += 8.0 * 3 * d.Px * d.qpts * d.Pz;
    const hipStream_t st = c.stream;
    const XmixKernel kern = xmix_kernel(mode, d.Px, nseq, zb);
    const real2 *W1r = (const real2 *) W1;
    real2 *W2r = (real2 *) W2;
    auto go = [=]() { hipLaunchKernelGGL(kern, grid, dim3(BLK), lds, st, a, pd, W1r, W2r); };
    // the CFL maximum accumulates across launches, so a replay skips it
    XmixArgs ar = a;
    ar.check_cfl = 0;
    auto again = [=]() { hipLaunchKernelGGL(kern, grid, dim3(BLK), lds, st, ar, pd, W1r, W2r); };
    prof_begin(c, st, 2, bytes);
    go();
    // replays only the modes whose launches are side-effect free
    if (mode == XM_ADJOINT || mode == XM_PRIMAL || mode == XM_TANGENT)
        prof_end(c, st, 2, again);
    else
        prof_end(c, st, 2);
}

bool xmix_adj2_available(Context &c)
{
    const int zb = SC_FFT_NS::xmix_zblock(c, 15);
    return (c.d.Px == 96 && zb == xmix_rows(96, 8)) || (c.d.Px == 48 && zb == xmix_rows(48, 8));
}

void launch_xmix_adj2(Context &c, const cplx *Wa, const cplx *Wb, const cplx *Wl, cplx *W2,
                      cplx *W2b)
{
    const Dims &d = c.d;
    XmixArgs a;
    a.mode = XM_ADJ2;
    a.nf = 15;
    a.nout = 12;
    a.nf1 = 6;
    a.W1b = (const real2 *) Wb;
    a.nf2 = 12;
    a.W1c = (const real2 *) Wl;
    a.W2b = (real2 *) W2b;
    a.Mx = d.Mx;
    a.Px = d.Px;
    a.Pz = d.Pz;
    a.qpts = d.qpts;
    const int zb = SC_FFT_NS::xmix_zblock(c, a.nf);
    a.ZB = zb;
    a.check_cfl = 0;
    a.cfl1 = d.cfl1;
    a.cfl3 = d.cfl3;
    a.cfl2 = c.dcfl2;
    a.W = c.dW;
    a.cflmax = c.cflmax;
    a.flow = nullptr;
    a.abort_flag = c.abort_flag;
    a.ms_in = rstride(c, Wa);
    a.ms_inb = rstride(c, Wb);
    a.ms_inc = rstride(c, Wl);
    a.ms_out = rstride(c, W2);
    a.ms_outb = rstride(c, W2b);
    const dim3 grid(d.Pz / zb, d.qpts, c.nrun);
    const size_t lds = sizeof(real2) * (zb * 8 * (d.Px + 1) + d.Px);
    const PlanDev pd = to_dev(c.px);
    const XmixKernel kern = d.Px == 96 ? k_xmix<XM_ADJ2, 96> : k_xmix<XM_ADJ2, 48>;
    const double bytes = c.nrun * (double) sizeof(real2) * d.qpts * d.Pz * d.Mx * (a.nf + a.nout);
    const hipStream_t st = c.stream;
    const real2 *War = (const real2 *) Wa;
    real2 *W2r = (real2 *) W2;
    auto go = [=]() { hipLaunchKernelGGL(kern, grid, dim3(BLK), lds, st, a, pd, War, W2r); };
    prof_begin(c, st, 7, bytes);
    go();
    prof_end(c, st, 7, go);
}

int xmix_zblock(Context &c, int nf)
{
    int nseq = (nf + 1) / 2, zb = 1;
    while (zb * 2 <= c.d.Pz && zb * 2 <= SC_XMIX_ZBMAX && c.d.Pz % (zb * 2) == 0 &&
           (size_t) zb * 2 * nseq * (c.d.Px + 1) + c.d.Px <= (size_t) XMIX_TILE)
        zb *= 2;
    return zb;
}

void launch_zfwd(Context &c, const cplx *W2, int nout, cplx *dst, int dst_kind, int cfl_slot)
{
    const Dims &d = c.d;
    ZfwdArgs a;
    a.cfl_slot = cfl_slot;
    a.cflmax = c.cflmax;
    a.dt = d.dt;
    a.abort_rw = c.abort_flag;
    a.nout = nout;
    a.Nz = d.Nz;
    a.Mx = d.Mx;
    a.Pz = d.Pz;
    a.qpts = d.qpts;
    const int fixed = z_fixed(d);
    a.XB = fixed ? ZXB : pick_block(d.Mx, d.Pz, 1);
    a.dst_kind = dst_kind;
    a.norm = 1.0 / ((3. * d.Nx / 2.) * (3. * d.Nz / 2.));
    a.abort_flag = c.abort_flag;
    a.ms_in = rstride(c, W2);
    a.ms_out = mstride(c, dst);
    a.p32 = dst_kind == ZF_USLOTS && prods32(c);
    dim3 grid((d.Mx + a.XB - 1) / a.XB, nout * c.nrun, d.qpts);
    size_t lds = sizeof(real2) * (a.XB * d.Pz + d.Pz);
    double bytes = c.nrun * (double) d.qpts * d.Mx * nout *
                   (sizeof(real2) * d.Pz + (a.p32 ? 8.0 : 16.0) * (d.Nz - 1));
    const hipStream_t st = c.stream;
    const PlanDev pd = to_dev(c.pz);
    const ZfwdKernel kern = fixed == 96 ? k_zfwd<96> : fixed == 48 ? k_zfwd<48> : k_zfwd<0>;
    const real2 *W2r = (const real2 *) W2;
    auto go = [=]() { hipLaunchKernelGGL(kern, grid, dim3(BLK), lds, st, a, pd, W2r, dst); };
    prof_begin(c, st, 3, bytes);
    go();
    prof_end(c, st, 3, go);
}

// the fused transform + assembly; false when its tile does not fit (the
// caller then transforms into scratch and assembles separately)
bool launch_zfwd_adj(Context &c, const cplx *W2, cplx *Udst)
{
    return SC_FFT_NS::launch_zfwd_adj2(c, W2, Udst, nullptr, nullptr);
}

bool launch_zfwd_adj2(Context &c, const cplx *W2, cplx *Udst, const cplx *W2b, cplx *Udstb)
{
    const Dims &d = c.d;
    const int fixed = z_fixed(d);
    // the pair in one grid only for the fixed-length two-phase kernel
    if (W2b && !fixed) {
        if (!SC_FFT_NS::launch_zfwd_adj2(c, W2, Udst, nullptr, nullptr)) return false;
        return SC_FFT_NS::launch_zfwd_adj2(c, W2b, Udstb, nullptr, nullptr);
    }
    ZadjArgs a;
    a.nprob = W2b ? 2 : 1;
    a.W2b = (const real2 *) W2b;
    a.Ub = Udstb;
    a.Nz = d.Nz;
    a.Mx = d.Mx;
    a.Pz = d.Pz;
    a.qpts = d.qpts;
    // generic tile: 6 XA Pz values and the twiddles within FFT_TILE
    a.XA = fixed ? ZA : std::min(d.Mx, (FFT_TILE - d.Pz) / (6 * d.Pz));
    if (a.XA < 1) return false;
    a.norm = 1.0 / ((3. * d.Nx / 2.) * (3. * d.Nz / 2.));
    a.Kx = c.dKx;
    a.Kz = c.dKz;
    a.abort_flag = c.abort_flag;
    a.nm = c.nrun;
    a.ms_in = rstride(c, W2);
    a.ms_out = mstride(c, Udst);
    a.p32 = prods32(c);
    if (W2b && (rstride(c, W2b) != a.ms_in || mstride(c, Udstb) != a.ms_out)) return false;
    const dim3 grid = fixed ? dim3(xcd_grid(a.nm * a.nprob * (d.Mx / ZA) * d.qpts))
                            : dim3((d.Mx + a.XA - 1) / a.XA, d.qpts, a.nm);
    // fixed lengths: the two-phase kernel on a 3-field tile
    size_t lds = sizeof(real2) * ((fixed ? 3 : 6) * (size_t) a.XA * d.Pz + d.Pz);
    double bytes = a.nm * a.nprob * (double) d.qpts * d.Mx *
                   (sizeof(real2) * 6.0 * d.Pz + (a.p32 ? 8.0 : 16.0) * 5.0 * (d.Nz - 1));
    const hipStream_t st = c.stream;
    const PlanDev pd = to_dev(c.pz);
    typedef void (*ZadjKernel)(ZadjArgs, PlanDev, const real2 *, cplx *);
    const ZadjKernel kern = fixed == 96 ? k_zfwd_adj2<96> : fixed == 48 ? k_zfwd_adj2<48> : k_zfwd_adj<0>;
    const real2 *W2r = (const real2 *) W2;
    auto go = [=]() { hipLaunchKernelGGL(kern, grid, dim3(BLK), lds, st, a, pd, W2r, Udst); };
    prof_begin(c, st, 3, bytes);
    go();
    prof_end(c, st, 3, go);
    return true;
}

void launch_assemble_adj(Context &c, const cplx *SP, cplx *Udst)
{
    const Dims &d = c.d;
    size_t n = (size_t) d.Nz * d.qpts * d.Mx;
    int blocks = (int) std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_assemble_adj, dim3(blocks, c.nrun), dim3(256), 0, c.stream, SP, Udst,
                       c.dKx, c.dKz, d.Nz, d.qpts, d.Mx, c.abort_flag, mstride(c, SP),
                       mstride(c, Udst), (int) prods32(c));
}

}  // namespace SC_FFT_NS
}  // namespace sc
