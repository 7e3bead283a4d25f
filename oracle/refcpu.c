/*
 * refcpu.c -- CPU restatement of the shadow-channel DNS path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  Nothing in the product links it.
 *
 * Each routine names the reference file:line it restates.  The restatement
 * keeps the reference's semantics, including its quirks (the right-hand side
 * carried between Runge-Kutta stages, real-part-only mean-mode forcing,
 * mass-only flux correction, untouched alpha rows dimQ..dimR-1, Nyquist
 * zeroing, the stage-k storage index of MC/MIC, the LU lag of the adjoint and
 * the forward difference of ddt_project at step 0).
 *
 * Pinning (tests/test_oracle_*.py): Gauss nodes/weights against numpy's
 * leggauss; the bases against numpy Legendre polynomials; the band Gram
 * matrices against dense quadrature; the padded transforms against numpy.fft;
 * the nonlinear products against numpy on the physical grid; the band solver
 * against dense solves; the linear stage map against dense Galerkin matrices
 * built independently; the tangent against central finite differences of the
 * primal and its linearity; the adjoint against the tangent by duality; the
 * laminar Orr-Sommerfeld growth rate against Orszag (1971).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "refcpu.h"

typedef orc_cplx cx;

#define CFL_LIMIT 2.8
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* slots of U (channel/minChnl.h:46-56) */
enum { S_U = 0, S_V = 1, S_W = 2, S_DU = 3, S_DW = 4 };
/* after the nonlinear pass: Hz -> slot 0, Hx -> slot 1, Hy -> slot 2 */
enum { S_HZ = 0, S_HX = 1, S_HY = 2 };

/* ------------------------------------------------------------------ */
/* FFT: generic mixed-radix decimation in time, exp(sign*2*pi*i*j*k/n)  */
/* ------------------------------------------------------------------ */
typedef struct {
    int n, nf, f[40];
    cx *twp, *twm;      /* exp(+2 pi i j/n), exp(-2 pi i j/n) */
    cx *buf;
} fftplan;

static void plan_make(fftplan *p, int n)
{
    int m = n, d = 2;
    p->n = n;
    p->nf = 0;
    while (m > 1) {
        while (m % d == 0) { p->f[p->nf++] = d; m /= d; }
        ++d;
    }
    p->twp = (cx *) malloc(sizeof(cx) * n);
    p->twm = (cx *) malloc(sizeof(cx) * n);
    p->buf = (cx *) malloc(sizeof(cx) * n);
    for (int j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * (double) j / (double) n;
        p->twp[j].re = cos(a);
        p->twp[j].im = sin(a);
        p->twm[j].re = cos(a);
        p->twm[j].im = -sin(a);
    }
}

static void plan_free(fftplan *p)
{
    free(p->twp); free(p->twm); free(p->buf);
    memset(p, 0, sizeof(*p));
}

/* Stockham autosort pass of radix r: x (length n) -> y, Ns = product of the
   radices already applied; twiddle exp(sign 2 pi i q k / (Ns r)) */
static void fft_pass(const fftplan *p, int r, int Ns, const cx *x, cx *y, const cx *tw)
{
    int n = p->n, nb = n / r, step = n / (Ns * r);
    cx v[64];
    cx w_r[64];
    for (int q = 0; q < r; ++q) w_r[q] = tw[(q * (n / r)) % n];   /* exp(sign 2 pi i q / r) */
    for (int j = 0; j < nb; ++j) {
        int k = j % Ns;
        for (int q = 0; q < r; ++q) {
            cx a = x[j + q * nb];
            if (q > 0 && k > 0) {
                cx w = tw[q * k * step];
                cx t = { a.re * w.re - a.im * w.im, a.re * w.im + a.im * w.re };
                a = t;
            }
            v[q] = a;
        }
        int base = (j / Ns) * Ns * r + k;
        if (r == 2) {
            y[base].re = v[0].re + v[1].re; y[base].im = v[0].im + v[1].im;
            y[base + Ns].re = v[0].re - v[1].re; y[base + Ns].im = v[0].im - v[1].im;
            continue;
        }
        for (int s = 0; s < r; ++s) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < r; ++q) {
                cx w = w_r[(q * s) % r];
                sr += v[q].re * w.re - v[q].im * w.im;
                si += v[q].re * w.im + v[q].im * w.re;
            }
            y[base + s * Ns].re = sr;
            y[base + s * Ns].im = si;
        }
    }
}

/* in-place transform of x (contiguous), sign +1 (backward) or -1 (forward) */
static void fft_c(fftplan *p, cx *x, int sign)
{
    const cx *tw = sign > 0 ? p->twp : p->twm;
    cx *a = x, *b = p->buf;
    int Ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        fft_pass(p, p->f[f], Ns, a, b, tw);
        Ns *= p->f[f];
        cx *t = a; a = b; b = t;
    }
    if (a != x) memcpy(x, a, sizeof(cx) * p->n);
}

/* ------------------------------------------------------------------ */
/* global solver state (the reference keeps one simulation per process) */
/* ------------------------------------------------------------------ */
static struct {
    int ok;
    int Nx, Ny, Nz, Mx, qpts, dimR, dimQ, nsteps, Px, Pz;
    double Lx, Lz, dt, re, flux, cfl1, cfl3;
    double *y, *w, *cfl2, *Kx, *Kz, *K2;
    double *Q, *Qp, *Qpp, *R, *Rp;          /* qpts x dim   */
    double *Qw, *Qpw, *Qppw, *Rw, *Rpw;     /* dim x qpts   */
    double *Qs, *Qps, *Qpps;                /* dimQ x 9     */
    double *Rs, *Rps;                       /* dimR x 5     */
    double *Rp0, *Rpp0;
    double *addm;                           /* Rs^{-1} Rw 1 */
    cx *C, *U, *IC, *IU, *AC, *AU, *LU;
    cx *MC, *MIC;
    cx *Uxb, *Uzb;
    cx *Fa, *Fb, *TM;                       /* dimR x Mx    */
    double *band;                           /* dimR x 9 x Mx */
    /* pass work */
    cx *spec;                               /* 12 x Nz x Mx */
    double *phys;                           /* 12 x Pz x Px */
    cx *zcol, *xrow;                        /* Pz x Mx, Px  */
    fftplan pz, px;
    int faithful, status;
} G;

static int rand_pos = -1;
static unsigned rand_ring[31];

/* glibc random() TYPE_3 additive feedback generator, as seeded by
   srand(seed): r0 = seed, r_i = 16807 r_{i-1} mod (2^31-1) for i < 31,
   r_i = r_{i-31} for i = 31..33, r_i = r_{i-3} + r_{i-31} (mod 2^32)
   afterwards; the first 310 outputs are discarded and rand() returns
   r_i >> 1.  rand_ring holds r_{i-31} .. r_{i-1}. */
void orc_srand(unsigned seed)
{
    unsigned r[34];
    long long v = seed == 0 ? 1 : seed;
    r[0] = (unsigned) v;
    for (int i = 1; i < 31; ++i) {
        v = (16807LL * v) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        r[i] = (unsigned) v;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 31; ++i) rand_ring[i] = r[i + 3];
    rand_pos = 0;
    for (int i = 0; i < 310; ++i) (void) orc_rand();
}

int orc_rand(void)
{
    if (rand_pos < 0) orc_srand(1);
    unsigned v = rand_ring[0] + rand_ring[28];
    memmove(rand_ring, rand_ring + 1, sizeof(unsigned) * 30);
    rand_ring[30] = v;
    return (int) (v >> 1);
}

/* index helpers */
static inline long ci(int z, int v, int i, int x)
{ return (((long) z * 2 + v) * G.dimR + i) * G.Mx + x; }
static inline long ui(int z, int s, int j, int x)
{ return (((long) z * 5 + s) * G.qpts + j) * G.Mx + x; }
static inline long csize(void) { return (long) G.Nz * 2 * G.dimR * G.Mx; }
static inline long usize(void) { return (long) G.Nz * 5 * G.qpts * G.Mx; }

/* ------------------------------------------------------------------ */
/* setup: channel/LegendreSetup.c:25-527, channel/RsQs.c:55-261,       */
/*        channel/waveNumsAndCFL.c:22-96                                */
/* ------------------------------------------------------------------ */
static double legendre_p(int n, double x)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) return 1.0;
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

/* Gauss-Legendre nodes in ascending order, mirror-symmetric
   (channel/LegendreSetup.c:566-643; Newton iteration instead of the
   bracketing secant, same roots) and weights (:732-745). */
static void gauss_nodes(int n, double *x, double *w)
{
    int half = n / 2;
    for (int m = 0; m < half; ++m) {
        double t = -cos(M_PI * (m + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p = legendre_p(n, t), pm = legendre_p(n - 1, t);
            double dp = n * (pm - t * p) / (1.0 - t * t);
            double dx = p / dp;
            t -= dx;
            if (fabs(dx) <= 1e-16 * fabs(t) || dx == 0.0) break;
        }
        x[m] = t;
        x[n - 1 - m] = -t;
    }
    if (n % 2 == 1) x[half] = 0.0;
    for (int k = 0; k < (n + 1) / 2; ++k) {
        double v1 = n * (legendre_p(n - 1, x[k]) - x[k] * legendre_p(n, x[k]));
        w[k] = 2.0 * (1.0 - x[k] * x[k]) / (v1 * v1);
        w[n - 1 - k] = w[k];
    }
}

/* banded Gram matrix D (n x (2dl+1)), D[i][dl+m] = sum_k A[i][k] B[k][i+2m],
   computed for |m| < nd (diagonals beyond are structurally zero) and
   mirrored (channel/LegendreSetup.c:698-720). A is dim x qpts (weighted),
   B is qpts x dim. */
static void gram_band(double *D, int n, int dl, int nd, const double *A,
                      const double *B, int dim)
{
    int W = 2 * dl + 1;
    memset(D, 0, sizeof(double) * n * W);
    for (int m = 0; m < nd; ++m) {
        for (int i = 0; i + 2 * m < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < G.qpts; ++k)
                s += A[(long) i * G.qpts + k] * B[(long) k * dim + i + 2 * m];
            D[i * W + dl + m] = s;
            if (m > 0) D[(i + 2 * m) * W + dl - m] = s;
        }
    }
}

/* ------------------------------------------------------------------ */
/* stride-2 band LU without pivoting (channel/mvOps.c:29-144)           */
/* ------------------------------------------------------------------ */
static void band_factor(double *A, int N, int dl)
{
    int W = 2 * dl + 1;
    for (int i = 2; i < N; ++i) {
        int mmax = MIN(dl, i / 2);
        for (int m = mmax; m >= 1; --m) {
            double v = A[i * W + dl - m];
            for (int k = 1; m + k <= mmax; ++k)
                v -= A[i * W + dl - (m + k)] * A[(i - 2 * (m + k)) * W + dl + k];
            A[i * W + dl - m] = v / A[(i - 2 * m) * W + dl];
        }
        int umax = MIN(dl, (N - i + 1) / 2);
        for (int u = 0; u < umax; ++u) {
            double v = A[i * W + dl + u];
            int kmax = MIN(i / 2, MIN(dl, dl - u));
            for (int k = 1; k <= kmax; ++k)
                v -= A[i * W + dl - k] * A[(i - 2 * k) * W + dl + u + k];
            A[i * W + dl + u] = v;
        }
    }
}

/* forward/back substitution with factors from band_factor; b has stride bs */
static void band_subst(const double *A, cx *b, long bs, int N, int dl)
{
    int W = 2 * dl + 1;
    for (int i = 2; i < N; ++i) {
        int mmax = MIN(dl, i / 2);
        for (int m = mmax; m >= 1; --m) {
            double l = A[i * W + dl - m];
            b[i * bs].re -= l * b[(i - 2 * m) * bs].re;
            b[i * bs].im -= l * b[(i - 2 * m) * bs].im;
        }
    }
    for (int i = N - 1; i >= 0; --i) {
        int jmax = MIN(dl, (N - i + 1) / 2 - 1);
        for (int j = 1; j <= jmax; ++j) {
            double u = A[i * W + dl + j];
            b[i * bs].re -= u * b[(i + 2 * j) * bs].re;
            b[i * bs].im -= u * b[(i + 2 * j) * bs].im;
        }
        b[i * bs].re /= A[i * W + dl];
        b[i * bs].im /= A[i * W + dl];
    }
}

void orc_band_solve(double *A, orc_cplx *b, int N, int dl)
{
    band_factor(A, N, dl);
    band_subst(A, b, 1, N, dl);
}

/* out[i] = sum_m A[i][dl+m] y[i+2m], rows i < N (channel/mvOps.c:198-237) */
static void band_mult(const double *A, const cx *y, long ys, cx *out, long os,
                      int N, int dl)
{
    int W = 2 * dl + 1;
    for (int i = 0; i < N; ++i) {
        double sr = 0.0, si = 0.0;
        int m0 = -MIN(dl, i / 2);
        int m1 = MIN(dl, (N - i + 1) / 2 - 1);
        for (int m = m0; m <= m1; ++m) {
            double a = A[i * W + dl + m];
            sr += a * y[(i + 2 * m) * ys].re;
            si += a * y[(i + 2 * m) * ys].im;
        }
        out[i * os].re = sr;
        out[i * os].im = si;
    }
}

/* ------------------------------------------------------------------ */
/* operators (channel/project.c:55-71,88-93,140-151,163-172)            */
/* ------------------------------------------------------------------ */
/* v operator: -(K2 Qs + Qps) + coef (K2^2 Qs + 2 K2 Qps + Qpps) */
static void op_alpha(double *M, double K2, double coef)
{
    double s = K2 * K2;
    for (int i = 0; i < G.dimQ; ++i)
        for (int j = 0; j < 9; ++j) {
            int e = i * 9 + j;
            M[e] = -(K2 * G.Qs[e] + G.Qps[e]) +
                   coef * (s * G.Qs[e] + 2. * K2 * G.Qps[e] + G.Qpps[e]);
        }
}

/* omega_y operator: Rs + coef (Rps + K2 Rs) */
static void op_beta(double *M, double K2, double coef)
{
    for (int i = 0; i < G.dimR; ++i)
        for (int j = 0; j < 5; ++j) {
            int e = i * 5 + j;
            M[e] = G.Rs[e] + coef * (G.Rps[e] + K2 * G.Rs[e]);
        }
}

/* mean-mode operator: Rs + coef Rps */
static void op_mean(double *M, double coef)
{
    for (int i = 0; i < G.dimR; ++i)
        for (int j = 0; j < 5; ++j) {
            int e = i * 5 + j;
            M[e] = G.Rs[e] + coef * G.Rps[e];
        }
}

/* RK3/CN coefficients: forward (channel/project.c:44-47),
   adjoint (channel/adjproject.c:34-37) */
static const double FA[3] = { 1. / 3., -1. / 2, 1. / 3. };
static const double FB[3] = { 1. / 6, 2. / 3, 0. };
static const double FC[3] = { 1. / 2, 1. / 3., 1., };
static const double FD[3] = { 0., -1. / 6, -2. / 3 };
static const double BA[3] = { 1. / 3., -1. / 2, 1. / 3. };
static const double BB[3] = { 2. / 3, 1. / 6, 0. };
static const double BC[3] = { 1., 1. / 3., 1. / 2., };
static const double BD[3] = { 0., -2. / 3, -1. / 6 };

/* ------------------------------------------------------------------ */
/* padded 2-D transforms (channel/pass1.c:52-114, 312-371)              */
/* ------------------------------------------------------------------ */
/* spectral plane S[kz][kx] (row stride ss, only rows != Nz/2) -> phys[zp][n] */
static void inv2d(const cx *S, long ss, double *phys)
{
    int Nz = G.Nz, Mx = G.Mx, Pz = G.Pz, Px = G.Px;
    cx *col = G.zcol;      /* Pz x Mx, column-major per kx: col[x*Pz + zp] */
    for (int x = 0; x < Mx; ++x) {
        cx *c = col + (long) x * Pz;
        memset(c, 0, sizeof(cx) * Pz);
        for (int z = 0; z < Nz / 2; ++z) c[z] = S[z * ss + x];
        for (int z = Nz / 2 + 1; z < Nz; ++z) c[z + Nz / 2] = S[z * ss + x];
        fft_c(&G.pz, c, +1);
    }
    cx *row = G.xrow;
    for (int zp = 0; zp < Pz; ++zp) {
        memset(row, 0, sizeof(cx) * Px);
        row[0].re = col[zp].re;        /* c2r ignores Im at kx = 0 */
        for (int k = 1; k < Mx; ++k) {
            cx v = col[(long) k * Pz + zp];
            row[k] = v;
            row[Px - k].re = v.re;
            row[Px - k].im = -v.im;
        }
        fft_c(&G.px, row, +1);
        for (int n = 0; n < Px; ++n) phys[(long) zp * Px + n] = row[n].re;
    }
}

/* phys[zp][n] -> spectral plane S[kz][kx], scaled by 1/(Px Pz), rows != Nz/2 */
static void fwd2d(const double *phys, cx *S, long ss)
{
    int Nz = G.Nz, Mx = G.Mx, Pz = G.Pz, Px = G.Px;
    cx *col = G.zcol;
    cx *row = G.xrow;
    double norm = 1.0 / ((3. * G.Nx / 2.) * (3. * G.Nz / 2.));
    for (int zp = 0; zp < Pz; ++zp) {
        for (int n = 0; n < Px; ++n) {
            row[n].re = phys[(long) zp * Px + n];
            row[n].im = 0.0;
        }
        fft_c(&G.px, row, -1);
        for (int k = 0; k < Mx; ++k) col[(long) k * Pz + zp] = row[k];
    }
    for (int x = 0; x < Mx; ++x) {
        cx *c = col + (long) x * Pz;
        fft_c(&G.pz, c, -1);
        for (int z = 0; z < Nz / 2; ++z) {
            S[z * ss + x].re = norm * c[z].re;
            S[z * ss + x].im = norm * c[z].im;
        }
        for (int z = Nz / 2 + 1; z < Nz; ++z) {
            S[z * ss + x].re = norm * c[z + Nz / 2].re;
            S[z * ss + x].im = norm * c[z + Nz / 2].im;
        }
    }
}

void orc_inv2d(const orc_cplx *spec, double *phys) { inv2d(spec, G.Mx, phys); }

void orc_fwd2d(const double *phys, orc_cplx *spec)
{
    fwd2d(phys, spec, G.Mx);
    memset(spec + (long) (G.Nz / 2) * G.Mx, 0, sizeof(cx) * G.Mx);
}

/* load u, v, w, omega_z, omega_x, omega_y of plane y into spec[0..5]
   (channel/pass1.c:52-104: curl in spectral space) */
static void load_curl(const cx *Ua, int y, cx *sp)
{
    int Nz = G.Nz, Mx = G.Mx;
    long fs = (long) Nz * Mx;
    for (int z = 0; z < Nz; ++z) {
        if (z == Nz / 2) continue;
        double kz = G.Kz[z];
        for (int x = 0; x < Mx; ++x) {
            double kx = G.Kx[x];
            cx u = Ua[ui(z, S_U, y, x)], v = Ua[ui(z, S_V, y, x)];
            cx w = Ua[ui(z, S_W, y, x)], du = Ua[ui(z, S_DU, y, x)];
            cx dw = Ua[ui(z, S_DW, y, x)];
            long o = (long) z * Mx + x;
            sp[0 * fs + o] = u;
            sp[1 * fs + o] = v;
            sp[2 * fs + o] = w;
            sp[3 * fs + o].re = -(du.re + kx * v.im);      /* omega_z */
            sp[3 * fs + o].im = -du.im + kx * v.re;
            sp[4 * fs + o].re = dw.re + kz * v.im;         /* omega_x */
            sp[4 * fs + o].im = dw.im - kz * v.re;
            sp[5 * fs + o].re = kx * w.im - kz * u.im;     /* omega_y */
            sp[5 * fs + o].im = kz * u.re - kx * w.re;
        }
    }
}

/* ------------------------------------------------------------------ */
/* pass1: H = u x omega (+ tangent IH) (channel/pass1.c:13-376)         */
/* ------------------------------------------------------------------ */
static int pass1(cx *Ua, cx *IUa, int check_cfl)
{
    int Nz = G.Nz, Mx = G.Mx, Pz = G.Pz, Px = G.Px;
    long fs = (long) Nz * Mx, ps = (long) Pz * Px;
    for (int y = 0; y < G.qpts; ++y) {
        load_curl(Ua, y, G.spec);
        for (int f = 0; f < 6; ++f) inv2d(G.spec + f * fs, Mx, G.phys + f * ps);
        double *u = G.phys, *v = u + ps, *w = v + ps;
        double *oz = w + ps, *ox = oz + ps, *oy = ox + ps;
        if (check_cfl) {
            double cfl = 0.0;
            for (long p = 0; p < ps; ++p) {
                double t = fabs(u[p]) * G.cfl1 + fabs(v[p]) * G.cfl2[y] +
                           fabs(w[p]) * G.cfl3;
                if (t > cfl) cfl = t;
            }
            cfl *= G.dt;
            if (cfl >= CFL_LIMIT) {
                printf("Pass1:  CFL[%d]:%e, Max:%e dt:%e.  CFL condition violated.\n",
                       y, cfl, CFL_LIMIT, G.dt);
                return -1;
            }
        }
        if (IUa) {
            double *Iu = G.phys + 6 * ps, *Iv = Iu + ps, *Iw = Iv + ps;
            double *Ioz = Iw + ps, *Iox = Ioz + ps, *Ioy = Iox + ps;
            load_curl(IUa, y, G.spec + 6 * fs);
            for (int f = 0; f < 6; ++f)
                inv2d(G.spec + (6 + f) * fs, Mx, G.phys + (6 + f) * ps);
            for (long p = 0; p < ps; ++p) {
                double hx = Iv[p] * oz[p] - Iw[p] * oy[p] + v[p] * Ioz[p] - w[p] * Ioy[p];
                double hy = Iw[p] * ox[p] - Iu[p] * oz[p] + w[p] * Iox[p] - u[p] * Ioz[p];
                double hz = Iu[p] * oy[p] - Iv[p] * ox[p] + u[p] * Ioy[p] - v[p] * Iox[p];
                Iu[p] = hz;
                Iv[p] = hx;
                Iw[p] = hy;
            }
        }
        for (long p = 0; p < ps; ++p) {
            double hx = v[p] * oz[p] - w[p] * oy[p];
            double hy = w[p] * ox[p] - u[p] * oz[p];
            double hz = u[p] * oy[p] - v[p] * ox[p];
            u[p] = hz;
            v[p] = hx;
            w[p] = hy;
        }
        long zs = 5L * G.qpts * Mx;
        for (int s = 0; s < 3; ++s)
            fwd2d(G.phys + s * ps, Ua + ui(0, s, y, 0), zs);
        if (IUa) {
            for (int s = 0; s < 3; ++s)
                fwd2d(G.phys + (6 + s) * ps, IUa + ui(0, s, y, 0), zs);
        }
    }
    return 0;
}

/* scratch U used when only the cost of the tangent side is wanted */
static cx *g_scratchU;

int orc_pass1_state(orc_cplx *Ua, int check_cfl)
{
    return pass1(Ua, NULL, check_cfl);
}

/* ------------------------------------------------------------------ */
/* pass2: adjoint convective operator (channel/pass2.c:23-570)          */
/* ------------------------------------------------------------------ */
static void pass2(cx *Ua, const cx *AUa, int incremental_cost)
{
    int Nz = G.Nz, Mx = G.Mx, Pz = G.Pz, Px = G.Px;
    long fs = (long) Nz * Mx, ps = (long) Pz * Px;
    cx *out = G.spec + 9 * fs;     /* 6 transformed fields in spec[9..14]? */
    (void) out;
    for (int y = 0; y < G.qpts; ++y) {
        load_curl(Ua, y, G.spec);
        for (int s = 0; s < 3; ++s) {
            for (int z = 0; z < Nz; ++z) {
                if (z == Nz / 2) continue;
                for (int x = 0; x < Mx; ++x)
                    G.spec[(6 + s) * fs + (long) z * Mx + x] = AUa[ui(z, s, y, x)];
            }
        }
        for (int f = 0; f < 9; ++f) inv2d(G.spec + f * fs, Mx, G.phys + f * ps);
        if (incremental_cost) {
            /* the reference also transforms the incremental side (IU, IAU)
               and discards the result here; repeat the work for cost only */
            for (int f = 0; f < 9; ++f) inv2d(G.spec + f * fs, Mx, G.phys + 9 * ps);
        }
        double *u = G.phys, *v = u + ps, *w = v + ps;
        double *oz = w + ps, *ox = oz + ps, *oy = ox + ps;
        double *lx = oy + ps, *ly = lx + ps, *lz = ly + ps;
        for (long p = 0; p < ps; ++p) {
            double sx = v[p] * lz[p] - w[p] * ly[p];
            double sy = w[p] * lx[p] - u[p] * lz[p];
            double sz = u[p] * ly[p] - v[p] * lx[p];
            double tx = ly[p] * oz[p] - lz[p] * oy[p];
            double ty = lz[p] * ox[p] - lx[p] * oz[p];
            double tz = lx[p] * oy[p] - ly[p] * ox[p];
            u[p] = sx; v[p] = sy; w[p] = sz;
            oz[p] = tx; ox[p] = ty; oy[p] = tz;
        }
        /* transformed S_x,S_y,S_z,T_x,T_y,T_z into spec[0..5] */
        for (int f = 0; f < 6; ++f) fwd2d(G.phys + f * ps, G.spec + f * fs, Mx);
        if (incremental_cost)
            for (int f = 0; f < 6; ++f) fwd2d(G.phys + f * ps, G.spec + 9 * fs, Mx);
        const cx *Sx = G.spec, *Sy = Sx + fs, *Sz = Sy + fs;
        const cx *Tx = Sz + fs, *Ty = Tx + fs, *Tz = Ty + fs;
        for (int z = 0; z < Nz; ++z) {
            if (z == Nz / 2) continue;
            double kz = G.Kz[z];
            for (int x = 0; x < Mx; ++x) {
                double kx = G.Kx[x];
                long o = (long) z * Mx + x;
                cx *hx = Ua + ui(z, S_HX, y, x), *hy = Ua + ui(z, S_HY, y, x);
                cx *hz = Ua + ui(z, S_HZ, y, x);
                Ua[ui(z, S_DU, y, x)] = Sx[o];
                Ua[ui(z, S_DW, y, x)] = Sz[o];
                hx->re = -Tx[o].re - kz * Sy[o].im;
                hx->im = -Tx[o].im + kz * Sy[o].re;
                hy->re = -Ty[o].re + kz * Sx[o].im - kx * Sz[o].im;
                hy->im = -Ty[o].im - kz * Sx[o].re + kx * Sz[o].re;
                hz->re = -Tz[o].re + kx * Sy[o].im;
                hz->im = -Tz[o].im - kx * Sy[o].re;
            }
        }
    }
}

void orc_pass2_state(orc_cplx *Ua, const orc_cplx *AUa) { pass2(Ua, AUa, 0); }

/* The incremental side of the adjoint convective operator
   (channel/pass2.c:120-566): per plane the state (u, omega) of U, the
   tangent (Iu, I omega) of IU, the adjoint velocity lambda of AU and the
   incremental adjoint velocity I lambda of IAU at the Gauss points;
   IS = u x I lambda + Iu x lambda, IT = lambda x I omega + I lambda x omega
   (channel/pass2.c:227-330); IH = -IT - the x,z curl of IS with IS_x, IS_z
   left in the du/dy, dw/dy slots (channel/pass2.c:474-566), into IH. */
int orc_pass2_incremental(const orc_cplx *U, const orc_cplx *IU, const orc_cplx *AU,
                          const orc_cplx *IAU, orc_cplx *IH)
{
    if (!G.ok || !U || !IU || !AU || !IAU || !IH) return -1;
    int Nz = G.Nz, Mx = G.Mx;
    long fs = (long) Nz * Mx, ps = (long) G.Pz * G.Px;
    cx *sp = calloc((size_t) 18 * fs, sizeof(cx));
    double *ph = calloc((size_t) 18 * ps, sizeof(double));
    if (!sp || !ph) { free(sp); free(ph); return -1; }
    const cx *Ua = (const cx *) U, *IUa = (const cx *) IU;
    const cx *AUa = (const cx *) AU, *IAUa = (const cx *) IAU;
    cx *Ha = (cx *) IH;
    for (int y = 0; y < G.qpts; ++y) {
        load_curl(Ua, y, sp);
        load_curl(IUa, y, sp + 6 * fs);
        for (int s = 0; s < 3; ++s)
            for (int z = 0; z < Nz; ++z) {
                if (z == Nz / 2) continue;
                for (int x = 0; x < Mx; ++x) {
                    sp[(12 + s) * fs + (long) z * Mx + x] = AUa[ui(z, s, y, x)];
                    sp[(15 + s) * fs + (long) z * Mx + x] = IAUa[ui(z, s, y, x)];
                }
            }
        for (int f = 0; f < 18; ++f) inv2d(sp + f * fs, Mx, ph + f * ps);
        const double *u = ph, *v = u + ps, *w = v + ps, *oz = w + ps, *ox = oz + ps, *oy = ox + ps;
        const double *iu = oy + ps, *iv = iu + ps, *iw = iv + ps;
        const double *ioz = iw + ps, *iox = ioz + ps, *ioy = iox + ps;
        const double *lx = ioy + ps, *ly = lx + ps, *lz = ly + ps;
        const double *ilx = lz + ps, *ily = ilx + ps, *ilz = ily + ps;
        double tmp[6];
        for (long p = 0; p < ps; ++p) {
            tmp[0] = v[p] * ilz[p] - w[p] * ily[p] + iv[p] * lz[p] - iw[p] * ly[p];
            tmp[1] = w[p] * ilx[p] - u[p] * ilz[p] + iw[p] * lx[p] - iu[p] * lz[p];
            tmp[2] = u[p] * ily[p] - v[p] * ilx[p] + iu[p] * ly[p] - iv[p] * lx[p];
            tmp[3] = ly[p] * ioz[p] - lz[p] * ioy[p] + ily[p] * oz[p] - ilz[p] * oy[p];
            tmp[4] = lz[p] * iox[p] - lx[p] * ioz[p] + ilz[p] * ox[p] - ilx[p] * oz[p];
            tmp[5] = lx[p] * ioy[p] - ly[p] * iox[p] + ilx[p] * oy[p] - ily[p] * ox[p];
            /* planes 0..5 (the state's) are read only at p: overwrite in place */
            for (int f = 0; f < 6; ++f) ph[f * ps + p] = tmp[f];
        }
        for (int f = 0; f < 6; ++f) fwd2d(ph + f * ps, sp + f * fs, Mx);
        const cx *Sx = sp, *Sy = Sx + fs, *Sz = Sy + fs;
        const cx *Tx = Sz + fs, *Ty = Tx + fs, *Tz = Ty + fs;
        for (int z = 0; z < Nz; ++z) {
            if (z == Nz / 2) {
                for (int s = 0; s < 5; ++s)
                    for (int x = 0; x < Mx; ++x) Ha[ui(z, s, y, x)].re = Ha[ui(z, s, y, x)].im = 0.0;
                continue;
            }
            double kz = G.Kz[z];
            for (int x = 0; x < Mx; ++x) {
                double kx = G.Kx[x];
                long o = (long) z * Mx + x;
                cx *hx = Ha + ui(z, S_HX, y, x), *hy = Ha + ui(z, S_HY, y, x);
                cx *hz = Ha + ui(z, S_HZ, y, x);
                Ha[ui(z, S_DU, y, x)] = Sx[o];
                Ha[ui(z, S_DW, y, x)] = Sz[o];
                hx->re = -Tx[o].re - kz * Sy[o].im;
                hx->im = -Tx[o].im + kz * Sy[o].re;
                hy->re = -Ty[o].re + kz * Sx[o].im - kx * Sz[o].im;
                hy->im = -Ty[o].im - kz * Sx[o].re + kx * Sz[o].re;
                hz->re = -Tz[o].re + kx * Sy[o].im;
                hz->im = -Tz[o].im - kx * Sy[o].re;
            }
        }
    }
    free(sp);
    free(ph);
    return 0;
}

/* ------------------------------------------------------------------ */
/* coefficients -> Gauss points (channel/initAlphaBeta2.c:12-240)       */
/* ------------------------------------------------------------------ */
/* all modes x >= x0 of row z: Gauss-point values from the coefficients
   (channel/initAlphaBeta2.c:96-208), sums over j in ascending order */
static void rebuild_row(const cx *Ca, cx *Ua, int z, int x0)
{
    int q = G.qpts, nQ = G.dimQ, nR = G.dimR, Mx = G.Mx;
    double kz = G.Kz[z];
    cx acc[5][512];
    for (int i = 0; i < q; ++i) {
        for (int m = 0; m < 5; ++m)
            for (int x = x0; x < Mx; ++x) acc[m][x].re = acc[m][x].im = 0.0;
        for (int j = 0; j < nQ; ++j) {
            double cq = G.Q[i * nQ + j], cp = G.Qp[i * nQ + j], c2 = G.Qpp[i * nQ + j];
            const cx *a = Ca + ci(z, 0, j, 0);
            for (int x = x0; x < Mx; ++x) {
                acc[0][x].re += cq * a[x].re; acc[0][x].im += cq * a[x].im;
                acc[1][x].re -= cp * a[x].re; acc[1][x].im -= cp * a[x].im;
                acc[2][x].re += c2 * a[x].re; acc[2][x].im += c2 * a[x].im;
            }
        }
        for (int j = 0; j < nR; ++j) {
            double cr = G.R[i * nR + j], cp = G.Rp[i * nR + j];
            const cx *b = Ca + ci(z, 1, j, 0);
            for (int x = x0; x < Mx; ++x) {
                acc[3][x].re += cr * b[x].re; acc[3][x].im += cr * b[x].im;
                acc[4][x].re += cp * b[x].re; acc[4][x].im += cp * b[x].im;
            }
        }
        for (int x = x0; x < Mx; ++x) {
            double kx = G.Kx[x], k2 = G.K2[z * Mx + x];
            double vr = acc[0][x].re, vi = acc[0][x].im, fr = acc[1][x].re, fi = acc[1][x].im;
            double qr = acc[2][x].re, qi = acc[2][x].im, gr = acc[3][x].re, gi = acc[3][x].im;
            double pr = acc[4][x].re, pi = acc[4][x].im;
            cx *o = Ua;
            o[ui(z, S_V, i, x)].re = vr;
            o[ui(z, S_V, i, x)].im = vi;
            o[ui(z, S_U, i, x)].re = (kx * fi + kz * gi) / k2;
            o[ui(z, S_U, i, x)].im = -(kx * fr + kz * gr) / k2;
            o[ui(z, S_W, i, x)].re = (-kx * gi + kz * fi) / k2;
            o[ui(z, S_W, i, x)].im = (kx * gr - kz * fr) / k2;
            o[ui(z, S_DU, i, x)].re = (-kx * qi + kz * pi) / k2;
            o[ui(z, S_DU, i, x)].im = (kx * qr - kz * pr) / k2;
            o[ui(z, S_DW, i, x)].re = -(kx * pi + kz * qi) / k2;
            o[ui(z, S_DW, i, x)].im = (kx * pr + kz * qr) / k2;
        }
    }
}

static void rebuild_mean(const cx *Ca, cx *Ua)
{
    int q = G.qpts, nR = G.dimR;
    for (int i = 0; i < q; ++i) {
        double ur = 0, uim = 0, dr = 0, di = 0, wr = 0, wi = 0, er = 0, ei = 0;
        for (int j = 0; j < nR; ++j) {
            cx a = Ca[ci(0, 0, j, 0)], b = Ca[ci(0, 1, j, 0)];
            double r = G.R[i * nR + j], rp = G.Rp[i * nR + j];
            ur += r * a.re; uim += r * a.im;
            dr += rp * a.re; di += rp * a.im;
            wr += r * b.re; wi += r * b.im;
            er += rp * b.re; ei += rp * b.im;
        }
        Ua[ui(0, S_U, i, 0)].re = ur;  Ua[ui(0, S_U, i, 0)].im = uim;
        Ua[ui(0, S_DU, i, 0)].re = dr; Ua[ui(0, S_DU, i, 0)].im = di;
        Ua[ui(0, S_V, i, 0)].re = 0;   Ua[ui(0, S_V, i, 0)].im = 0;
        Ua[ui(0, S_W, i, 0)].re = wr;  Ua[ui(0, S_W, i, 0)].im = wi;
        Ua[ui(0, S_DW, i, 0)].re = er; Ua[ui(0, S_DW, i, 0)].im = ei;
    }
}

/* wall derivatives of the state (channel/project.c:336-349) */
static void wall_accumulate(const cx *Ca, int z, int x)
{
    double kx = G.Kx[x], kz = G.Kz[z], k2 = G.K2[z * G.Mx + x];
    cx *ub = G.Uxb + z * G.Mx + x, *wb = G.Uzb + z * G.Mx + x;
    for (int j = 0; j < G.dimR; ++j) {
        cx b = Ca[ci(z, 1, j, x)];
        ub->re += G.Rp0[j] * b.im * kz / k2;
        ub->im += -G.Rp0[j] * b.re * kz / k2;
        wb->re += -G.Rp0[j] * b.im * kx / k2;
        wb->im += G.Rp0[j] * b.re * kx / k2;
    }
}

static void wall_mean(const cx *Ca)
{
    memset(G.Uxb, 0, sizeof(cx) * G.Nz * G.Mx);
    memset(G.Uzb, 0, sizeof(cx) * G.Nz * G.Mx);
    for (int j = 0; j < G.dimR; ++j) {
        cx a = Ca[ci(0, 0, j, 0)], b = Ca[ci(0, 1, j, 0)];
        G.Uxb[0].re += G.Rp0[j] * a.re;
        G.Uxb[0].im += G.Rp0[j] * a.im;
        G.Uzb[0].re += G.Rp0[j] * b.re;
        G.Uzb[0].im += G.Rp0[j] * b.im;
    }
}

/* initAlphaBeta2 / incre_initAlphaBeta2 / adj_initAlphaBeta2 */
static void coef_to_gauss(const cx *Ca, cx *Ua, int wall)
{
    rebuild_mean(Ca, Ua);
    if (wall) wall_mean(Ca);
    for (int z = 0; z < G.Nz; ++z) {
        rebuild_row(Ca, Ua, z, z == 0 ? 1 : 0);
        if (wall)
            for (int x = (z == 0 ? 1 : 0); x < G.Mx; ++x) wall_accumulate(Ca, z, x);
    }
}

void orc_coef_to_gauss(const orc_cplx *Ca, orc_cplx *Ua) { coef_to_gauss(Ca, Ua, 0); }

/* Wall plane of a coefficient array (channel/initAlphaBeta2.c:70-81,
   210-223): Uxb = Rp0.alpha, Uzb = Rp0.beta at (0,0); elsewhere
   Uxb = -i kz/k2 Rp0.beta, Uzb = i kx/k2 Rp0.beta (every row, Nz/2 too). */
void orc_wall_plane(const orc_cplx *Ca, orc_cplx *Uxb, orc_cplx *Uzb)
{
    wall_mean(Ca);
    for (int z = 0; z < G.Nz; ++z)
        for (int x = (z == 0 ? 1 : 0); x < G.Mx; ++x) wall_accumulate(Ca, z, x);
    memcpy(Uxb, G.Uxb, sizeof(cx) * G.Nz * G.Mx);
    memcpy(Uzb, G.Uzb, sizeof(cx) * G.Nz * G.Mx);
}

/* increBoundary (channel/increBoundary.c:20-162), in place on two wall
   planes [Nz][Nx/2]: both planes and the fixed plane w.n (a single mode,
   Re = 1 at kz row 1, kx 1: channel/increBoundary.c:66) go through the
   padded inverse transform, the products Uxb w.n and Uzb w.n through the
   forward transform (scaled by 1/(Px Pz), row Nz/2 left zero), then every
   (z, x) != (0, 0) is rewritten as
       Uxb <- i kz Uxb - i kx Uzb,   Uzb <- -i kx Uxb - i kz Uzb
   (channel/increBoundary.c:137-158). */
int orc_incre_boundary(orc_cplx *Uxb, orc_cplx *Uzb)
{
    const int Nz = G.Nz, Mx = G.Mx;
    const long np = (long) G.Pz * G.Px;
    if (Nz < 4 || Mx < 2) return -1;
    cx *wn = calloc((size_t) Nz * Mx, sizeof(cx));
    if (!wn) return -1;
    wn[1 * Mx + 1].re = 1.0;
    double *px = G.phys, *pz = G.phys + np, *pw = G.phys + 2 * np;
    inv2d(Uxb, Mx, px);
    inv2d(Uzb, Mx, pz);
    inv2d(wn, Mx, pw);
    free(wn);
    for (long i = 0; i < np; ++i) {
        px[i] = px[i] * pw[i];
        pz[i] = pz[i] * pw[i];
    }
    fwd2d(px, Uxb, Mx);
    fwd2d(pz, Uzb, Mx);
    memset(Uxb + (long) (Nz / 2) * Mx, 0, sizeof(cx) * Mx);
    memset(Uzb + (long) (Nz / 2) * Mx, 0, sizeof(cx) * Mx);
    for (int z = 0; z < Nz; ++z)
        for (int x = 0; x < Mx; ++x) {
            if (z * z + x * x == 0) continue;
            const double kx = G.Kx[x], kz = G.Kz[z];
            cx *a = Uxb + (long) z * Mx + x, *b = Uzb + (long) z * Mx + x;
            const double t1 = -kz * a->im + kx * b->im;
            const double t2 = kz * a->re - kx * b->re;
            const double t3 = kx * a->im + kz * b->im;
            b->im = -kx * a->re - kz * b->re;
            b->re = t3;
            a->re = t1;
            a->im = t2;
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Galerkin projection of the nonlinear term (channel/project.c:95-115, */
/* 176-191; channel/adjproject.c:108-134, 213-234)                      */
/* ------------------------------------------------------------------ */
static void forcing_modes(const cx *Ua, int z, int x0, int adj)
{
    int q = G.qpts, nQ = G.dimQ, nR = G.dimR, Mx = G.Mx;
    double kz = G.Kz[z];
    for (int i = 0; i < nQ; ++i)
        for (int x = x0; x < Mx; ++x) G.Fa[i * Mx + x].re = G.Fa[i * Mx + x].im = 0.0;
    for (int i = 0; i < nR; ++i)
        for (int x = x0; x < Mx; ++x) G.Fb[i * Mx + x].re = G.Fb[i * Mx + x].im = 0.0;
    for (int i = 0; i < nQ; ++i) {
        cx *F = G.Fa + i * Mx;
        for (int j = 0; j < q; ++j) {
            double a = G.Qpw[i * q + j], b = G.Qw[i * q + j];
            double c = adj ? G.Qppw[i * q + j] : 0.0;
            const cx *hx = Ua + ui(z, S_HX, j, 0), *hy = Ua + ui(z, S_HY, j, 0);
            const cx *hz = Ua + ui(z, S_HZ, j, 0);
            const cx *sx = Ua + ui(z, S_DU, j, 0), *sz = Ua + ui(z, S_DW, j, 0);
            for (int x = x0; x < Mx; ++x) {
                double kx = G.Kx[x], k2 = G.K2[z * Mx + x];
                double tr = a * (-kx * hx[x].im - kz * hz[x].im) - k2 * (b * hy[x].re);
                double ti = a * (kx * hx[x].re + kz * hz[x].re) - k2 * (b * hy[x].im);
                if (adj) {
                    tr += c * (kz * sx[x].im - kx * sz[x].im);
                    ti += c * (-kz * sx[x].re + kx * sz[x].re);
                }
                F[x].re += tr;
                F[x].im += ti;
            }
        }
    }
    for (int i = 0; i < nR; ++i) {
        cx *F = G.Fb + i * Mx;
        for (int j = 0; j < q; ++j) {
            double a = G.Rw[i * q + j];
            double c = adj ? G.Rpw[i * q + j] : 0.0;
            const cx *hx = Ua + ui(z, S_HX, j, 0), *hz = Ua + ui(z, S_HZ, j, 0);
            const cx *sx = Ua + ui(z, S_DU, j, 0), *sz = Ua + ui(z, S_DW, j, 0);
            for (int x = x0; x < Mx; ++x) {
                double kx = G.Kx[x];
                double tr = a * (-kz * hx[x].im + kx * hz[x].im);
                double ti = a * (kz * hx[x].re - kx * hz[x].re);
                if (adj) {
                    tr += c * (-kz * sz[x].im - kx * sx[x].im);
                    ti += c * (kz * sz[x].re + kx * sx[x].re);
                }
                F[x].re += tr;
                F[x].im += ti;
            }
        }
    }
}

/* Optional right-hand-side source of the current stage (C layout), added
   after the explicit part and the forcing, before the solve (the
   Reynolds-number source of orc_tangent_src); NULL: none. */
static const cx *g_src = NULL;

/* ------------------------------------------------------------------ */
/* one Runge-Kutta stage for the modes of row z, x >= x0                */
/* project (channel/project.c:24-405), increproject                     */
/* (channel/increproject.c:17-500), adjproject (channel/adjproject.c)   */
/*   Ca:   coefficient array carried between stages (RHS between stages)*/
/*   Ua:   nonlinear term in (H slots; S_x,S_z in DU/DW for adjoint)     */
/*   LUa:  adjoint only: nonlinear term for the next-stage explicit part*/
/*   store: MC/MIC slot or NULL                                          */
/* ------------------------------------------------------------------ */
static void stage_modes(cx *Ca, const cx *Ua, const cx *LUa, cx *Uout,
                        cx *store, int k, int z, int x0, int adj, int wall)
{
    int Mx = G.Mx, nQ = G.dimQ, nR = G.dimR;
    const double *a = adj ? BA : FA, *b = adj ? BB : FB;
    const double *c = adj ? BC : FC, *d = adj ? BD : FD;
    double dt = G.dt, re = G.re;
    double *M = G.band;
    /* alpha */
    if (k == 0) {
        for (int x = x0; x < Mx; ++x) {
            op_alpha(M, G.K2[z * Mx + x], re * a[0] * dt);
            band_mult(M, Ca + ci(z, 0, 0, x), Mx, G.TM + x, Mx, nQ, 4);
        }
        for (int i = 0; i < nQ; ++i)
            for (int x = x0; x < Mx; ++x) Ca[ci(z, 0, i, x)] = G.TM[i * Mx + x];
    }
    forcing_modes(Ua, z, x0, adj);
    for (int i = 0; i < nQ; ++i)
        for (int x = x0; x < Mx; ++x) {
            Ca[ci(z, 0, i, x)].re += dt * c[k] * G.Fa[i * Mx + x].re;
            Ca[ci(z, 0, i, x)].im += dt * c[k] * G.Fa[i * Mx + x].im;
            if (g_src) {
                Ca[ci(z, 0, i, x)].re += g_src[ci(z, 0, i, x)].re;
                Ca[ci(z, 0, i, x)].im += g_src[ci(z, 0, i, x)].im;
            }
        }
    for (int x = x0; x < Mx; ++x) {
        op_alpha(M, G.K2[z * Mx + x], -(re * b[k] * dt));
        band_factor(M, nQ, 4);
        band_subst(M, Ca + ci(z, 0, 0, x), Mx, nQ, 4);
    }
    /* beta */
    if (k == 0) {
        for (int x = x0; x < Mx; ++x) {
            op_beta(M, G.K2[z * Mx + x], -(re * a[0] * dt));
            band_mult(M, Ca + ci(z, 1, 0, x), Mx, G.TM + x, Mx, nR, 2);
        }
        for (int i = 0; i < nR; ++i)
            for (int x = x0; x < Mx; ++x) Ca[ci(z, 1, i, x)] = G.TM[i * Mx + x];
    }
    for (int i = 0; i < nR; ++i)
        for (int x = x0; x < Mx; ++x) {
            Ca[ci(z, 1, i, x)].re += dt * c[k] * G.Fb[i * Mx + x].re;
            Ca[ci(z, 1, i, x)].im += dt * c[k] * G.Fb[i * Mx + x].im;
            if (g_src) {
                Ca[ci(z, 1, i, x)].re += g_src[ci(z, 1, i, x)].re;
                Ca[ci(z, 1, i, x)].im += g_src[ci(z, 1, i, x)].im;
            }
        }
    for (int x = x0; x < Mx; ++x) {
        op_beta(M, G.K2[z * Mx + x], re * b[k] * dt);
        band_factor(M, nR, 2);
        band_subst(M, Ca + ci(z, 1, 0, x), Mx, nR, 2);
    }
    if (store) {
        for (int v = 0; v < 2; ++v)
            for (int i = 0; i < nR; ++i)
                for (int x = x0; x < Mx; ++x) store[ci(z, v, i, x)] = Ca[ci(z, v, i, x)];
    }
    /* velocities at Gauss points from the solved coefficients */
    rebuild_row(Ca, Uout, z, x0);
    if (wall)
        for (int x = x0; x < Mx; ++x) wall_accumulate(Ca, z, x);
    /* explicit part of the next stage */
    if (k != 2) {
        if (adj) forcing_modes(LUa, z, x0, 1);
        for (int x = x0; x < Mx; ++x) {
            op_alpha(M, G.K2[z * Mx + x], re * a[k + 1] * dt);
            band_mult(M, Ca + ci(z, 0, 0, x), Mx, G.TM + x, Mx, nQ, 4);
        }
        for (int i = 0; i < nQ; ++i)
            for (int x = x0; x < Mx; ++x) {
                Ca[ci(z, 0, i, x)].re = G.TM[i * Mx + x].re + dt * d[k + 1] * G.Fa[i * Mx + x].re;
                Ca[ci(z, 0, i, x)].im = G.TM[i * Mx + x].im + dt * d[k + 1] * G.Fa[i * Mx + x].im;
            }
        for (int x = x0; x < Mx; ++x) {
            op_beta(M, G.K2[z * Mx + x], -(re * a[k + 1] * dt));
            band_mult(M, Ca + ci(z, 1, 0, x), Mx, G.TM + x, Mx, nR, 2);
        }
        for (int i = 0; i < nR; ++i)
            for (int x = x0; x < Mx; ++x) {
                Ca[ci(z, 1, i, x)].re = G.TM[i * Mx + x].re + dt * d[k + 1] * G.Fb[i * Mx + x].re;
                Ca[ci(z, 1, i, x)].im = G.TM[i * Mx + x].im + dt * d[k + 1] * G.Fb[i * Mx + x].im;
            }
    }
}

/* ------------------------------------------------------------------ */
/* the (kx,kz) = (0,0) mode: project0 (channel/project0.c:22-271),      */
/* increproject0 (channel/increproject0.c:14-312),                      */
/* adjproject0 (channel/adjproject0.c:8-310)                            */
/*   mode: 0 primal (flux target G.flux), 1 tangent (target 0 or 2),    */
/*         2 adjoint (target 0, correction through the LHS matrix)      */
/* ------------------------------------------------------------------ */
static void mean_rhs(const cx *Ua, int adj, double *fa, double *fb)
{
    int q = G.qpts, nR = G.dimR;
    for (int i = 0; i < nR; ++i) {
        double sa = 0, sb = 0;
        for (int j = 0; j < q; ++j) {
            if (adj) {
                sa += G.Rw[i * q + j] * Ua[ui(0, S_HX, j, 0)].re +
                      G.Rpw[i * q + j] * Ua[ui(0, S_DW, j, 0)].re;
                sb += G.Rw[i * q + j] * Ua[ui(0, S_HZ, j, 0)].re -
                      G.Rpw[i * q + j] * Ua[ui(0, S_DU, j, 0)].re;
            } else {
                sa += G.Rw[i * q + j] * Ua[ui(0, S_HX, j, 0)].re;
                sb += G.Rw[i * q + j] * Ua[ui(0, S_HZ, j, 0)].re;
            }
        }
        fa[i] = sa;
        fb[i] = sb;
    }
}

/* flux-correction trace (test hook): per mean-mode stage, the record
   {mode, stage k, target, flux error before the correction, after it} */
static double *g_ftrace = NULL;
static int g_ftrace_n = 0, g_ftrace_cap = 0;

int orc_flux_trace(double *buf, int cap_records)
{
    int n = g_ftrace_n;
    g_ftrace = buf;
    g_ftrace_cap = buf ? cap_records : 0;
    g_ftrace_n = 0;
    return n;
}

/* 1: the adjoint's mean mode is the exact transpose of the tangent's (the
   objective adjoint's mode); 0: the reference's (channel/adjproject0.c) */
static int g_exact_mean = 0;

/* The exact transpose of the tangent's flux correction x <- P x,
   P = I - add w1^T / 2 (add = Rs^{-1} w1), in the pairing a^T Rs b of the
   mean column: P is self-adjoint there, and in the reversed stage order of
   the adjoint it falls on the stage's right-hand side before its solve
   (rhs <- rhs - w1 (add^T rhs) / 2, at adjoint stages 0 and 1) and on the
   adjoint state at the start of every step, before the step's adjoint
   velocities are formed (lambda <- P lambda: the driver's part); the real
   parts only, as in the tangent. */
static void mean_flux_transpose(cx *al, long s, int on_rhs)
{
    const int nR = G.dimR, q = G.qpts;
    double M[5 * 1024], w1[1024];
    cx add[1024];
    memcpy(M, G.Rs, sizeof(double) * nR * 5);
    for (int i = 0; i < nR; ++i) {
        w1[i] = 0.0;
        for (int j = 0; j < q; ++j) w1[i] += G.Rw[i * q + j];
        add[i].re = w1[i];
        add[i].im = 0.0;
    }
    band_factor(M, nR, 2);
    band_subst(M, add, 1, nR, 2);
    double t = 0.0;
    if (on_rhs) {
        for (int i = 0; i < nR; ++i) t += add[i].re * al[i * s].re;
        for (int i = 0; i < nR; ++i) al[i * s].re -= 0.5 * t * w1[i];
    } else {
        for (int i = 0; i < nR; ++i) t += w1[i] * al[i * s].re;
        for (int i = 0; i < nR; ++i) al[i * s].re -= 0.5 * t * add[i].re;
    }
}

static void stage_mean(cx *Ca, const cx *Ua, const cx *LUa, cx *Uout,
                       cx *store, int k, int mode, double target, int wall)
{
    int nR = G.dimR, q = G.qpts;
    int adj = (mode == 2);
    const double *a = adj ? BA : FA, *b = adj ? BB : FB;
    const double *c = adj ? BC : FC, *d = adj ? BD : FD;
    double dt = G.dt, re = G.re;
    double M[5 * 1024], add[1024];
    double fa[1024], fb[1024];
    cx tm[1024];
    cx *al = Ca + ci(0, 0, 0, 0), *be = Ca + ci(0, 1, 0, 0);
    long s = G.Mx;
    const int exact = adj && g_exact_mean;
    if (k == 0) {
        op_mean(M, -(re * a[0] * dt));
        band_mult(M, al, s, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) al[i * s] = tm[i];
        band_mult(M, be, s, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) be[i * s] = tm[i];
    }
    /* add = (Rs or Rs + re b dt Rps)^{-1} (Rw 1) */
    for (int i = 0; i < nR; ++i) {
        double t = 0;
        for (int j = 0; j < q; ++j) t += G.Rw[i * q + j];
        add[i] = t;
    }
    if (adj) op_mean(M, re * b[k] * dt);
    else memcpy(M, G.Rs, sizeof(double) * nR * 5);
    {
        cx tmp[1024];
        for (int i = 0; i < nR; ++i) { tmp[i].re = add[i]; tmp[i].im = 0; }
        band_factor(M, nR, 2);
        band_subst(M, tmp, 1, nR, 2);
        for (int i = 0; i < nR; ++i) add[i] = tmp[i].re;
    }
    mean_rhs(Ua, adj, fa, fb);
    op_mean(M, re * b[k] * dt);
    band_factor(M, nR, 2);
    for (int i = 0; i < nR; ++i) al[i * s].re += dt * c[k] * fa[i];
    if (g_src)
        for (int i = 0; i < nR; ++i) {
            al[i * s].re += g_src[ci(0, 0, i, 0)].re;
            al[i * s].im += g_src[ci(0, 0, i, 0)].im;
        }
    if (exact && k < 2) mean_flux_transpose(al, s, 1);
    band_subst(M, al, s, nR, 2);
    for (int i = 0; i < nR; ++i) be[i * s].re += dt * c[k] * fb[i];
    if (g_src)
        for (int i = 0; i < nR; ++i) {
            be[i * s].re += g_src[ci(0, 1, i, 0)].re;
            be[i * s].im += g_src[ci(0, 1, i, 0)].im;
        }
    band_subst(M, be, s, nR, 2);
    double flux_t = 0;
    for (int i = 0; i < nR; ++i)
        for (int j = 0; j < q; ++j) flux_t += G.Rw[i * q + j] * al[i * s].re;
    flux_t = target - flux_t;
    const double err_before = flux_t;
    if (!exact)
        for (int j = 0; j < nR; ++j) {
            al[j * s].re = al[j * s].re + flux_t * add[j] * 0.5;
            al[j * s].im = al[j * s].im + flux_t * 0.0 * 0.5;
        }
    if (g_ftrace && g_ftrace_n < g_ftrace_cap) {
        /* flux error target - sum_ij Rw_ij Re alpha_i after the correction */
        double after = 0;
        for (int i = 0; i < nR; ++i)
            for (int j = 0; j < q; ++j) after += G.Rw[i * q + j] * al[i * s].re;
        double *r = g_ftrace + 5 * g_ftrace_n++;
        r[0] = mode; r[1] = k; r[2] = target; r[3] = err_before; r[4] = target - after;
    }
    rebuild_mean(Ca, Uout);
    if (wall) wall_mean(Ca);
    if (store)
        for (int v = 0; v < 2; ++v)
            for (int i = 0; i < nR; ++i) store[ci(0, v, i, 0)] = Ca[ci(0, v, i, 0)];
    if (k != 2) {
        if (adj) mean_rhs(LUa, 1, fa, fb);
        op_mean(M, -(re * a[k + 1] * dt));
        band_mult(M, al, s, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) {
            al[i * s].re = tm[i].re + dt * d[k + 1] * fa[i];
            al[i * s].im = tm[i].im + dt * d[k + 1] * 0.0;
        }
        band_mult(M, be, s, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) {
            be[i * s].re = tm[i].re + dt * d[k + 1] * fb[i];
            be[i * s].im = tm[i].im + dt * d[k + 1] * 0.0;
        }
    }
}

/* one full stage over all modes of an array (the loops of
   channel/channel.c:456-470, 546-559, 668-679) */
static void stage_all(cx *Ca, cx *Ua, const cx *LUa, cx *Uout, cx *storebase,
                      int k, int mode, double target)
{
    int Nz = G.Nz;
    int adj = (mode == 2), wall = (mode == 0);
    stage_mean(Ca, Ua, LUa, Uout, storebase, k, mode, target, wall);
    stage_modes(Ca, Ua, LUa, Uout, storebase, k, 0, 1, adj, wall);
    for (int z = 1; z < Nz; ++z) {
        if (z == Nz / 2) {
            for (int s = 0; s < 5; ++s)
                memset(Uout + ui(z, s, 0, 0), 0, sizeof(cx) * G.qpts * G.Mx);
            memset(Ca + ci(z, 0, 0, 0), 0, sizeof(cx) * 2 * G.dimR * G.Mx);
            continue;
        }
        stage_modes(Ca, Ua, LUa, Uout, storebase, k, z, 0, adj, wall);
    }
}

/* ------------------------------------------------------------------ */
/* initial projection from Gauss points (channel/initAlphaBeta.c:24-266)*/
/* ------------------------------------------------------------------ */
static void gauss_to_coef(void)
{
    int q = G.qpts, nQ = G.dimQ, nR = G.dimR, Mx = G.Mx;
    double M[9 * 1024];
    /* mean mode: alpha from u, beta from w, R-basis mass solves */
    for (int v = 0; v < 2; ++v) {
        int slot = v == 0 ? S_U : S_W;
        for (int i = 0; i < nR; ++i) {
            double sr = 0, si = 0;
            for (int j = 0; j < q; ++j) {
                sr += G.Rw[i * q + j] * G.U[ui(0, slot, j, 0)].re;
                si += G.Rw[i * q + j] * G.U[ui(0, slot, j, 0)].im;
            }
            G.C[ci(0, v, i, 0)].re = sr;
            G.C[ci(0, v, i, 0)].im = si;
        }
        memcpy(M, G.Rs, sizeof(double) * nR * 5);
        band_factor(M, nR, 2);
        band_subst(M, G.C + ci(0, v, 0, 0), Mx, nR, 2);
    }
    for (int z = 0; z < G.Nz; ++z) {
        int x0 = z == 0 ? 1 : 0;
        double kz = G.Kz[z];
        /* alpha from v with the Q mass matrix (rows dimQ.. untouched) */
        for (int i = 0; i < nQ; ++i)
            for (int x = x0; x < Mx; ++x) {
                double sr = 0, si = 0;
                for (int j = 0; j < q; ++j) {
                    sr += G.Qw[i * q + j] * G.U[ui(z, S_V, j, x)].re;
                    si += G.Qw[i * q + j] * G.U[ui(z, S_V, j, x)].im;
                }
                G.C[ci(z, 0, i, x)].re = sr;
                G.C[ci(z, 0, i, x)].im = si;
            }
        memcpy(M, G.Qs, sizeof(double) * nQ * 9);
        band_factor(M, nQ, 4);
        for (int x = x0; x < Mx; ++x) band_subst(M, G.C + ci(z, 0, 0, x), Mx, nQ, 4);
        /* beta from omega_y = i kz u - i kx w */
        for (int i = 0; i < nR; ++i)
            for (int x = x0; x < Mx; ++x) {
                double kx = G.Kx[x], sr = 0, si = 0;
                for (int j = 0; j < q; ++j) {
                    cx u = G.U[ui(z, S_U, j, x)], w = G.U[ui(z, S_W, j, x)];
                    sr += G.Rw[i * q + j] * (kx * w.im - kz * u.im);
                    si += G.Rw[i * q + j] * (kz * u.re - kx * w.re);
                }
                G.C[ci(z, 1, i, x)].re = sr;
                G.C[ci(z, 1, i, x)].im = si;
            }
        memcpy(M, G.Rs, sizeof(double) * nR * 5);
        band_factor(M, nR, 2);
        for (int x = x0; x < Mx; ++x) band_subst(M, G.C + ci(z, 1, 0, x), Mx, nR, 2);
    }
}

/* ------------------------------------------------------------------ */
/* public API: channel/channel.c                                        */
/* ------------------------------------------------------------------ */
static void freeall(void)
{
    void **ptrs[] = {
        (void **) &G.y, (void **) &G.w, (void **) &G.cfl2, (void **) &G.Kx,
        (void **) &G.Kz, (void **) &G.K2, (void **) &G.Q, (void **) &G.Qp,
        (void **) &G.Qpp, (void **) &G.R, (void **) &G.Rp, (void **) &G.Qw,
        (void **) &G.Qpw, (void **) &G.Qppw, (void **) &G.Rw, (void **) &G.Rpw,
        (void **) &G.Qs, (void **) &G.Qps, (void **) &G.Qpps, (void **) &G.Rs,
        (void **) &G.Rps, (void **) &G.Rp0, (void **) &G.Rpp0, (void **) &G.addm,
        (void **) &G.C, (void **) &G.U, (void **) &G.IC, (void **) &G.IU,
        (void **) &G.AC, (void **) &G.AU, (void **) &G.LU, (void **) &G.MC,
        (void **) &G.MIC, (void **) &G.Uxb, (void **) &G.Uzb, (void **) &G.Fa,
        (void **) &G.Fb, (void **) &G.TM, (void **) &G.band, (void **) &G.spec,
        (void **) &G.phys, (void **) &G.zcol, (void **) &G.xrow,
        (void **) &g_scratchU };
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) {
        free(*ptrs[i]);
        *ptrs[i] = NULL;
    }
    if (G.pz.n) plan_free(&G.pz);
    if (G.px.n) plan_free(&G.px);
    G.ok = 0;
}

void orc_destroy(void) { freeall(); }

int orc_init(int Nx, int Ny, int Nz, double Lx, double Lz, double Re,
             double flux, double dt, int nsteps)
{
    int faithful = G.faithful;
    if (G.ok) freeall();
    memset(&G, 0, sizeof(G));
    G.faithful = faithful;
    /* channel/channel.c:274-312 */
    if (Nx % 4 != 0 || Nz % 2 != 0 || Ny - 4 < 0 || nsteps < 0) return 1;
    G.Nx = Nx; G.Ny = Ny; G.Nz = Nz; G.Lx = Lx; G.Lz = Lz;
    G.dt = dt; G.nsteps = nsteps; G.flux = flux;
    G.re = 1. / Re;
    G.qpts = 3 * Ny / 2;
    G.dimR = Ny - 2;
    G.dimQ = Ny - 4;
    G.Mx = Nx / 2;
    G.Px = 3 * Nx / 2;
    G.Pz = 3 * Nz / 2;
    if (G.dimR > 1000) return 1;
    int q = G.qpts, nQ = G.dimQ, nR = G.dimR, Mx = G.Mx;
#define ALLOC(p, n) do { (p) = calloc((size_t) (n), sizeof(*(p))); if (!(p)) { freeall(); return 1; } } while (0)
    ALLOC(G.y, q); ALLOC(G.w, q); ALLOC(G.cfl2, q);
    ALLOC(G.Kx, Mx); ALLOC(G.Kz, Nz); ALLOC(G.K2, Nz * Mx);
    ALLOC(G.Q, q * nQ); ALLOC(G.Qp, q * nQ); ALLOC(G.Qpp, q * nQ);
    ALLOC(G.R, q * nR); ALLOC(G.Rp, q * nR);
    ALLOC(G.Qw, nQ * q); ALLOC(G.Qpw, nQ * q); ALLOC(G.Qppw, nQ * q);
    ALLOC(G.Rw, nR * q); ALLOC(G.Rpw, nR * q);
    ALLOC(G.Qs, nQ * 9); ALLOC(G.Qps, nQ * 9); ALLOC(G.Qpps, nQ * 9);
    ALLOC(G.Rs, nR * 5); ALLOC(G.Rps, nR * 5);
    ALLOC(G.Rp0, nR); ALLOC(G.Rpp0, nR); ALLOC(G.addm, nR);
    ALLOC(G.C, csize()); ALLOC(G.IC, csize()); ALLOC(G.AC, csize());
    ALLOC(G.U, usize()); ALLOC(G.IU, usize()); ALLOC(G.AU, usize());
    ALLOC(G.LU, usize()); ALLOC(g_scratchU, usize());
    ALLOC(G.MC, csize() * (3L * nsteps + 1));
    ALLOC(G.MIC, csize() * (3L * nsteps + 1));
    ALLOC(G.Uxb, Nz * Mx); ALLOC(G.Uzb, Nz * Mx);
    ALLOC(G.Fa, nR * Mx); ALLOC(G.Fb, nR * Mx); ALLOC(G.TM, nR * Mx);
    ALLOC(G.band, nR * 9 * Mx);
    ALLOC(G.spec, 16L * Nz * Mx);
    ALLOC(G.phys, 16L * G.Pz * G.Px);
    ALLOC(G.zcol, (long) G.Pz * Mx);
    ALLOC(G.xrow, G.Px);
#undef ALLOC
    plan_make(&G.pz, G.Pz);
    plan_make(&G.px, G.Px);

    /* quadrature, Legendre matrix, bases (channel/LegendreSetup.c:353-400) */
    gauss_nodes(q, G.y, G.w);
    double *P = (double *) malloc(sizeof(double) * q * nR);
    for (int m = 0; m < q; ++m) {
        P[m * nR + 0] = 1.0;
        if (nR > 1) P[m * nR + 1] = G.y[m];
        for (int n = 1; n < nR - 1; ++n)
            P[m * nR + n + 1] = ((2. * n + 1.) * (G.y[m] * P[m * nR + n]) -
                                 n * P[m * nR + n - 1]) / (double) (n + 1);
    }
    for (int i = 0; i < q; ++i) {
        double yy = G.y[i], om = 1.0 - yy * yy, y2 = yy * yy;
        for (int j = 0; j < nR; ++j) {
            G.R[i * nR + j] = om * P[i * nR + j];
            G.Rp[i * nR + j] = (j == 0) ? -2.0 * yy
                : j * P[i * nR + j - 1] - (j + 2.0) * (yy * P[i * nR + j]);
        }
        for (int j = 0; j < nQ; ++j) {
            G.Q[i * nQ + j] = om * G.R[i * nR + j];
            G.Qp[i * nQ + j] = (j == 0) ? -4.0 * yy * om
                : om * (j * P[i * nR + j - 1] - (j + 4.0) * (yy * P[i * nR + j]));
            G.Qpp[i * nQ + j] = (j == 0) ? 12.0 * y2 - 4.0
                : (12.0 + 7.0 * j + j * j) * (y2 * P[i * nR + j]) -
                  (4.0 + j * j + j) * P[i * nR + j] - 6.0 * j * (yy * P[i * nR + j - 1]);
        }
    }
    free(P);
    for (int j = 0; j < nR; ++j) {
        double sg = (j % 2 == 0) ? 1.0 : -1.0;     /* (-1)^j */
        G.Rp0[j] = 2.0 * sg;                       /* R_j'(-1)  */
        G.Rpp0[j] = (j == 0) ? -2.0 : -2.0 * sg * (1.0 + j + (double) j * j);
    }
    /* weighted transposes Mw[i][j] = w[j] M[j][i] (:756-765) */
    for (int i = 0; i < nQ; ++i)
        for (int j = 0; j < q; ++j) {
            G.Qw[i * q + j] = G.w[j] * G.Q[j * nQ + i];
            G.Qpw[i * q + j] = G.w[j] * G.Qp[j * nQ + i];
            G.Qppw[i * q + j] = G.w[j] * G.Qpp[j * nQ + i];
        }
    for (int i = 0; i < nR; ++i)
        for (int j = 0; j < q; ++j) {
            G.Rw[i * q + j] = G.w[j] * G.R[j * nR + i];
            G.Rpw[i * q + j] = G.w[j] * G.Rp[j * nR + i];
        }
    /* band Gram matrices: nonzero diagonals only (:387-396) */
    gram_band(G.Qs, nQ, 4, 5, G.Qw, G.Q, nQ);
    gram_band(G.Qps, nQ, 4, 4, G.Qpw, G.Qp, nQ);
    gram_band(G.Qpps, nQ, 4, 3, G.Qppw, G.Qpp, nQ);
    gram_band(G.Rs, nR, 2, 3, G.Rw, G.R, nR);
    gram_band(G.Rps, nR, 2, 2, G.Rpw, G.Rp, nR);
    /* wavenumbers (channel/waveNumsAndCFL.c:45-59) */
    for (int i = 0; i < Mx; ++i) G.Kx[i] = i * (2. * M_PI / Lx);
    for (int i = 0; i < Nz / 2 + 1; ++i) G.Kz[i] = i * (2. * M_PI / Lz);
    for (int i = Nz / 2 + 1; i < Nz; ++i) G.Kz[i] = (i - Nz) * (2. * M_PI / Lz);
    for (int i = 0; i < Nz; ++i)
        for (int j = 0; j < Mx; ++j) G.K2[i * Mx + j] = G.Kz[i] * G.Kz[i] + G.Kx[j] * G.Kx[j];
    /* CFL constants (:65-96) */
    G.cfl1 = M_PI / (Lx / Nx);
    G.cfl3 = M_PI / (Lz / Nz);
    for (int j = 1; j < q - 1; ++j) G.cfl2[j] = 2. * M_PI / (G.y[j + 1] - G.y[j - 1]);
    G.cfl2[0] = G.cfl2[q - 1] = M_PI / (G.y[0] + 1.0);
    G.ok = 1;
    return 0;
}

/* laminar Poiseuille + rand() perturbation start (channel/channel.c:409-422) */
static void laminar_start(void)
{
    memset(G.U, 0, sizeof(cx) * usize());
    for (int i = 0; i < G.qpts; ++i) {
        G.U[ui(0, S_U, i, 0)].re = (1.0 - G.y[i] * G.y[i]) * G.flux * 3. / 4;
        G.U[ui(0, S_U, i, 0)].im = 0.0;
    }
    if (G.Nz > 1 && G.Mx > 1)
        for (int i = 0; i < G.qpts; ++i) {
            G.U[ui(1, S_U, i, 1)].re = (orc_rand() / (double) 2147483647 - 0.5);
            G.U[ui(1, S_U, i, 1)].im = (orc_rand() / (double) 2147483647 - 0.5);
        }
    gauss_to_coef();
}

int orc_primal(int ru_steps, orc_cplx *Cg)
{
    if (!G.ok || ru_steps < 0) return -2;
    long cs = csize();
    if (Cg) {
        memcpy(G.C, Cg, sizeof(cx) * cs);
        coef_to_gauss(G.C, G.U, 1);
    } else {
        laminar_start();
    }
    memset(G.MC, 0, sizeof(cx) * cs * (3L * G.nsteps + 1));
    memcpy(G.MC, G.C, sizeof(cx) * cs);
    G.status = 0;
    for (int n = 0; n < G.nsteps + ru_steps; ++n) {
        for (int k = 0; k < 3; ++k) {
            if (pass1(G.U, G.faithful ? g_scratchU : NULL, k == 0) != 0) {
                printf("Pass1 failure\n");
                G.status = -1;
                goto done;
            }
            int count = (n - ru_steps) * 3 + k + 1;
            stage_all(G.C, G.U, NULL, G.U, count >= 0 ? G.MC + cs * count : NULL,
                      k, 0, G.flux);
        }
    }
done:
    if (Cg) memcpy(Cg, G.C, sizeof(cx) * cs);
    return 0;
}

int orc_tangent(int start, int end, orc_cplx *ICg, int inhomo)
{
    if (!G.ok) return -2;
    if (!(end > start) || start < 0 || end > G.nsteps || !ICg) return -3;
    long cs = csize();
    memcpy(G.C, G.MC + cs * (3L * start), sizeof(cx) * cs);
    memcpy(G.IC, ICg, sizeof(cx) * cs);
    coef_to_gauss(G.C, G.U, 1);
    coef_to_gauss(G.IC, G.IU, 0);
    memcpy(G.MIC + cs * (3L * start), G.IC, sizeof(cx) * cs);
    G.status = 0;
    for (int n = start; n < end; ++n) {
        for (int k = 0; k < 3; ++k) {
            if (pass1(G.U, G.IU, k == 0) != 0) {
                printf("Pass1 failure\n");
                G.status = -1;
                goto done;
            }
            int count = n * 3 + k + 1;
            stage_all(G.IC, G.IU, NULL, G.IU, G.MIC + cs * count, k, 1,
                      inhomo ? 2.0 : 0.0);
            stage_all(G.C, G.U, NULL, G.U, NULL, k, 0, G.flux);
        }
    }
done:
    memcpy(ICg, G.IC, sizeof(cx) * cs);
    return 0;
}

static void spec2phys_impl(const cx *Cin, double *flow)
{
    int Px = G.Px, Pz = G.Pz, q = G.qpts;
    long fs = (long) G.Nz * G.Mx, ps = (long) Pz * Px;
    memcpy(G.C, Cin, sizeof(cx) * csize());
    coef_to_gauss(G.C, G.U, 1);
    for (int y = 0; y < q; ++y) {
        for (int s = 0; s < 3; ++s) {
            for (int z = 0; z < G.Nz; ++z)
                for (int x = 0; x < G.Mx; ++x)
                    G.spec[s * fs + (long) z * G.Mx + x] = G.U[ui(z, s, y, x)];
            inv2d(G.spec + s * fs, G.Mx, G.phys + s * ps);
        }
        for (int s = 0; s < 3; ++s)
            for (int zp = 0; zp < Pz; ++zp)
                for (int x = 0; x < Px; ++x)
                    flow[(long) s * Px * q * Pz + ((long) x * q + y) * Pz + zp] =
                        G.phys[s * ps + (long) zp * Px + x];
    }
}

void orc_spec2phys(const orc_cplx *Cin, double *flow) { spec2phys_impl(Cin, flow); }

double orc_ddt_project(int i_step, orc_cplx *v)
{
    if (!G.ok || i_step < 0 || i_step > G.nsteps || G.nsteps < 1) return NAN;
    long cs = csize();
    long fl = 3L * G.Px * G.qpts * G.Pz;
    cx *dC = (cx *) calloc(cs, sizeof(cx));
    double *du = (double *) calloc(fl, sizeof(double));
    double *vp = (double *) calloc(fl, sizeof(double));
    int ip, im;
    if (i_step == 0) { ip = 3 * (i_step + 1); im = 3 * i_step; }
    else { ip = 3 * i_step; im = 3 * (i_step - 1); }
    const cx *Mp = G.MC + cs * ip, *Mm = G.MC + cs * im;
    for (long e = 0; e < cs; ++e) {
        dC[e].re = (1.0 / G.dt) * (Mp[e].re - Mm[e].re);
        dC[e].im = (1.0 / G.dt) * (Mp[e].im - Mm[e].im);
    }
    spec2phys_impl(dC, du);
    spec2phys_impl(v, vp);
    double num = 0.0, den = 0.0;
    long fX = 0, fY = (long) G.Px * G.qpts * G.Pz, fZ = 2 * fY;
    for (int y = 0; y < G.qpts; ++y)
        for (int z = 0; z < G.Pz; ++z)
            for (int x = 0; x < G.Px; ++x) {
                long i = ((long) x * G.qpts + y) * G.Pz + z;
                num += du[i + fX] * vp[i + fX] * G.w[y];
                den += du[i + fX] * du[i + fX] * G.w[y];
                num += du[i + fY] * vp[i + fY] * G.w[y];
                den += du[i + fY] * du[i + fY] * G.w[y];
                num += du[i + fZ] * vp[i + fZ] * G.w[y];
                den += du[i + fZ] * du[i + fZ] * G.w[y];
            }
    for (long e = 0; e < cs; ++e) {
        v[e].re -= (num / den) * dC[e].re;
        v[e].im -= (num / den) * dC[e].im;
    }
    free(dC); free(du); free(vp);
    return num / den;
}

/* The reference subtracts the trapezoid forcing from AC only: the adjoint
   velocities AU, rebuilt once before the loop (adj_initAlphaBeta2,
   channel/channel.c:617) and after every stage solve, are not rebuilt after
   a forcing (channel/channel.c:622-631, 682-689), so the forcing reaches the
   convective term of pass2 one stage late.  Test hook: with the refresh on,
   AU is rebuilt from AC after each forcing, which makes the forced adjoint
   exactly the superposition of unforced adjoints of the forcing terms. */
static int g_adj_refresh = 0;
void orc_set_adjoint_refresh(int on) { g_adj_refresh = on; }

int orc_adjoint(int start, int end, orc_cplx *ACg, int inhomo, double strength)
{
    (void) inhomo;
    if (!G.ok) return -2;
    if (!(end < start) || end < 0 || start > G.nsteps || !ACg) return -3;
    long cs = csize(), us = usize();
    cx *tmp = (cx *) malloc(sizeof(cx) * cs);
    memcpy(G.C, G.MC + cs * (3L * start), sizeof(cx) * cs);
    memcpy(G.AC, ACg, sizeof(cx) * cs);
    coef_to_gauss(G.AC, G.AU, 0);
    strength *= G.dt;
    for (int n = start; n > end; --n) {
        if (g_exact_mean) {
            mean_flux_transpose(G.AC, G.Mx, 0);
            coef_to_gauss(G.AC, G.AU, 0);
        }
        memcpy(tmp, G.MIC + cs * (3L * n), sizeof(cx) * cs);
        orc_ddt_project(n, tmp);
        for (long e = 0; e < cs; ++e) {
            G.AC[e].re -= 0.5 * strength * tmp[e].re;
            G.AC[e].im -= 0.5 * strength * tmp[e].im;
        }
        if (g_adj_refresh) coef_to_gauss(G.AC, G.AU, 0);
        for (int k = 0; k < 3; ++k) {
            int count = n * 3 - k - 1;
            if (k < 2) {
                memcpy(G.C, G.MC + cs * (count - 1), sizeof(cx) * cs);
                coef_to_gauss(G.C, G.U, 1);
                pass2(G.U, G.AU, G.faithful);
                memcpy(G.LU, G.U, sizeof(cx) * us);
            }
            memcpy(G.C, G.MC + cs * count, sizeof(cx) * cs);
            coef_to_gauss(G.C, G.U, 1);
            pass2(G.U, G.AU, G.faithful);
            stage_all(G.AC, G.U, G.LU, G.AU, NULL, k, 2, 0.0);
        }
        memcpy(tmp, G.MIC + cs * (3L * (n - 1)), sizeof(cx) * cs);
        orc_ddt_project(n - 1, tmp);
        for (long e = 0; e < cs; ++e) {
            G.AC[e].re -= 0.5 * strength * tmp[e].re;
            G.AC[e].im -= 0.5 * strength * tmp[e].im;
        }
        if (g_adj_refresh) coef_to_gauss(G.AC, G.AU, 0);
    }
    memcpy(ACg, G.AC, sizeof(cx) * cs);
    free(tmp);
    return 0;
}

/* Reynolds-number sensitivity forcing of stored state MC[count]
 * (channel/tangent_forcing.c:10-75, channel/tangent_forcing0.c:10-44), the
 * band products in the order of smMult / smMult0 (channel/mvOps.c:198-237):
 *   (kz,kx) != (0,0): f_a = -re^2 (K2^2 Qs + 2 K2 Qps + Qpps) alpha,
 *                     f_b =  re^2 (Rps + K2 Rs) beta
 *   mean column:      f0_a = re^2 Rps alpha(0,0), f0_b = re^2 Rps beta(0,0)
 * f_a [Nz][dimQ][Mx], f_b [Nz][dimR][Mx] (column (0,0) zero), f0 [dimR]. */
int orc_tangent_forcing(int count, orc_cplx *fa_, orc_cplx *fb_, orc_cplx *f0a_, orc_cplx *f0b_)
{
    if (!G.ok || count < 0 || count > 3 * G.nsteps) return -1;
    const int Nz = G.Nz, Mx = G.Mx, nQ = G.dimQ, nR = G.dimR;
    const cx *S = G.MC + (long) count * csize();
    cx *fa = (cx *) fa_, *fb = (cx *) fb_, *f0a = (cx *) f0a_, *f0b = (cx *) f0b_;
    const double re = G.re;
    memset(fa, 0, sizeof(cx) * (size_t) Nz * nQ * Mx);
    memset(fb, 0, sizeof(cx) * (size_t) Nz * nR * Mx);
    for (int z = 0; z < Nz; ++z) {
        const cx *al = S + (long) (z * 2 + 0) * nR * Mx, *be = S + (long) (z * 2 + 1) * nR * Mx;
        for (int x = (z == 0 ? 1 : 0); x < Mx; ++x) {
            const double k2 = G.K2[z * Mx + x], s = k2 * k2;
            for (int i = 0; i < nQ; ++i) {
                double sr = 0.0, si = 0.0;
                for (int m = -MIN(4, i / 2); m <= MIN(4, (nQ - i + 1) / 2 - 1); ++m) {
                    const int e = i * 9 + 4 + m;
                    const double a = -re * re * (s * G.Qs[e] + 2. * k2 * G.Qps[e] + G.Qpps[e]);
                    sr += a * al[(long) (i + 2 * m) * Mx + x].re;
                    si += a * al[(long) (i + 2 * m) * Mx + x].im;
                }
                fa[((long) z * nQ + i) * Mx + x].re = sr;
                fa[((long) z * nQ + i) * Mx + x].im = si;
            }
            for (int i = 0; i < nR; ++i) {
                double sr = 0.0, si = 0.0;
                for (int m = -MIN(2, i / 2); m <= MIN(2, (nR - i + 1) / 2 - 1); ++m) {
                    const int e = i * 5 + 2 + m;
                    const double a = re * re * (G.Rps[e] + k2 * G.Rs[e]);
                    sr += a * be[(long) (i + 2 * m) * Mx + x].re;
                    si += a * be[(long) (i + 2 * m) * Mx + x].im;
                }
                fb[((long) z * nR + i) * Mx + x].re = sr;
                fb[((long) z * nR + i) * Mx + x].im = si;
            }
        }
    }
    const cx *al0 = S, *be0 = S + (long) nR * Mx;
    for (int i = 0; i < nR; ++i) {
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        for (int m = -MIN(2, i / 2); m <= MIN(2, (nR - i + 1) / 2 - 1); ++m) {
            const double a = re * re * G.Rps[i * 5 + 2 + m];
            ar += a * al0[(long) (i + 2 * m) * Mx].re;
            ai += a * al0[(long) (i + 2 * m) * Mx].im;
            br += a * be0[(long) (i + 2 * m) * Mx].re;
            bi += a * be0[(long) (i + 2 * m) * Mx].im;
        }
        f0a[i].re = ar;
        f0a[i].im = ai;
        f0b[i].re = br;
        f0b[i].im = bi;
    }
    return 0;
}

const orc_cplx *orc_getsoln(int i)
{
    if (!G.ok || i < 0 || i > G.nsteps) return NULL;
    return G.MC + csize() * 3L * i;
}

const orc_cplx *orc_getincresoln(int i)
{
    if (!G.ok || i < 0 || i > G.nsteps) return NULL;
    return G.MIC + csize() * 3L * i;
}

const orc_cplx *orc_getincrestage(int count)
{
    if (!G.ok || count < 0 || count > 3 * G.nsteps) return NULL;
    return G.MIC + csize() * (long) count;
}

/* Reynolds-number source of the tangent at stage k of step n, per unit
   change of Re: the implicit and explicit operators M -+ re b/a dt D depend
   on re = 1/Re, so with TF(X) = (dre/dRe) D X = -re^2 D X (the forcing of
   channel/tangent_forcing.c, tangent_forcing0.c) the stage's right-hand side
   gains dt (a_k TF(Y_k) + b_k TF(X_k)), Y_k the state the stage's explicit
   part acts on (MC[3n+k]: the step's start at k = 0, the previous stage's
   solution after) and X_k its own solution (MC[3n+k+1]).  src: C layout. */
static void re_source(int n, int k, cx *src)
{
    const int Nz = G.Nz, Mx = G.Mx, nQ = G.dimQ, nR = G.dimR;
    const long na = (long) Nz * nQ * Mx, nb = (long) Nz * nR * Mx;
    cx *fa = (cx *) calloc(na, sizeof(cx)), *fb = (cx *) calloc(nb, sizeof(cx));
    cx *f0a = (cx *) calloc(nR, sizeof(cx)), *f0b = (cx *) calloc(nR, sizeof(cx));
    memset(src, 0, sizeof(cx) * csize());
    const double w[2] = {G.dt * FA[k], G.dt * FB[k]};
    for (int t = 0; t < 2; ++t) {
        orc_tangent_forcing(3 * n + k + t, (orc_cplx *) fa, (orc_cplx *) fb, (orc_cplx *) f0a,
                            (orc_cplx *) f0b);
        for (int z = 0; z < Nz; ++z)
            for (int x = 0; x < Mx; ++x) {
                if (z == 0 && x == 0) continue;
                for (int i = 0; i < nQ; ++i) {
                    src[ci(z, 0, i, x)].re += w[t] * fa[((long) z * nQ + i) * Mx + x].re;
                    src[ci(z, 0, i, x)].im += w[t] * fa[((long) z * nQ + i) * Mx + x].im;
                }
                for (int i = 0; i < nR; ++i) {
                    src[ci(z, 1, i, x)].re += w[t] * fb[((long) z * nR + i) * Mx + x].re;
                    src[ci(z, 1, i, x)].im += w[t] * fb[((long) z * nR + i) * Mx + x].im;
                }
            }
        if (t == 1) {
            /* The implicit operator acts on the mean solve's own result, before
               the flux correction alpha += e add / 2 (add = Rs^{-1} w1, w1_i =
               sum_j Rw_ij) that the stored state has had: X_pre = X - e add / 2
               with e = (flux - w1.X) / (1 - w1.add / 2) recovered from X.  TF is
               linear, so TF(X_pre) = TF(X) - e/2 re^2 Rps add. */
            const int q = G.qpts;
            const cx *X = G.MC + (long) (3 * n + k + 1) * csize();
            double M[5 * 1024], w1[1024], wx = 0.0, wa = 0.0;
            cx add[1024], rps_add[1024];
            memcpy(M, G.Rs, sizeof(double) * nR * 5);
            for (int i = 0; i < nR; ++i) {
                w1[i] = 0.0;
                for (int j = 0; j < q; ++j) w1[i] += G.Rw[i * q + j];
                add[i].re = w1[i];
                add[i].im = 0.0;
            }
            band_factor(M, nR, 2);
            band_subst(M, add, 1, nR, 2);
            for (int i = 0; i < nR; ++i) {
                wx += w1[i] * X[ci(0, 0, i, 0)].re;
                wa += w1[i] * add[i].re;
            }
            const double e = (G.flux - wx) / (1.0 - 0.5 * wa);
            band_mult(G.Rps, add, 1, rps_add, 1, nR, 2);
            for (int i = 0; i < nR; ++i) f0a[i].re -= 0.5 * e * G.re * G.re * rps_add[i].re;
        }
        for (int i = 0; i < nR; ++i) {
            src[ci(0, 0, i, 0)].re += w[t] * f0a[i].re;
            src[ci(0, 0, i, 0)].im += w[t] * f0a[i].im;
            src[ci(0, 1, i, 0)].re += w[t] * f0b[i].re;
            src[ci(0, 1, i, 0)].im += w[t] * f0b[i].im;
        }
    }
    free(fa); free(fb); free(f0a); free(f0b);
}

int orc_tangent_src(int start, int end, orc_cplx *ICg, int inhomo, int re_src)
{
    if (!G.ok) return -2;
    if (!(end > start) || start < 0 || end > G.nsteps || !ICg) return -3;
    long cs = csize();
    cx *src = re_src ? (cx *) malloc(sizeof(cx) * cs) : NULL;
    memcpy(G.C, G.MC + cs * (3L * start), sizeof(cx) * cs);
    memcpy(G.IC, ICg, sizeof(cx) * cs);
    coef_to_gauss(G.C, G.U, 1);
    coef_to_gauss(G.IC, G.IU, 0);
    memcpy(G.MIC + cs * (3L * start), G.IC, sizeof(cx) * cs);
    G.status = 0;
    for (int n = start; n < end; ++n) {
        for (int k = 0; k < 3; ++k) {
            if (pass1(G.U, G.IU, k == 0) != 0) {
                printf("Pass1 failure\n");
                G.status = -1;
                goto done;
            }
            int count = n * 3 + k + 1;
            if (src) re_source(n, k, src);
            g_src = src;
            stage_all(G.IC, G.IU, NULL, G.IU, G.MIC + cs * count, k, 1, inhomo ? 2.0 : 0.0);
            g_src = NULL;
            stage_all(G.C, G.U, NULL, G.U, NULL, k, 0, G.flux);
        }
    }
done:
    g_src = NULL;
    memcpy(ICg, G.IC, sizeof(cx) * cs);
    free(src);
    return 0;
}

/* The adjoint (no trapezoid forcing) of the objective
     J = weight sum_{m = end}^{start - 1} tau(C_m),  tau(C) = sum_j R_j'(-1) Re alpha_j(0,0)
   (the lower-wall shear of the mean profile, channel/project0.c:227-234):
   after each step n -> n - 1, AC += weight G with G the representer of tau in
   the Gauss-weighted physical inner product <a, b> of drivers/testAdj.py:55-56:
   only its mean streamwise column is nonzero, alpha = Rs^{-1} Rp0 / (Px Pz)
   (the plane sums of a mean column are Px Pz times its value, and the Gauss
   rule makes <a, b> of two mean columns Px Pz alpha_a^T Rs alpha_b).  The
   adjoint velocities are rebuilt after each addition.  On return,
   <AC, d>_phys = <AC_given, T d> + weight sum_m tau(T_m d) for a tangent d
   from step end. */
void orc_set_exact_mean_adjoint(int on) { g_exact_mean = on; }

int orc_adjoint_objective(int start, int end, orc_cplx *ACg, double weight)
{
    if (!G.ok) return -2;
    if (!(end < start) || end < 0 || start > G.nsteps || !ACg) return -3;
    const long cs = csize(), us = usize();
    const int nR = G.dimR;
    double gv[1024], M[5 * 1024];
    {
        cx tmp[1024];
        memcpy(M, G.Rs, sizeof(double) * nR * 5);
        for (int i = 0; i < nR; ++i) {
            tmp[i].re = G.Rp0[i] / ((double) G.Px * G.Pz);
            tmp[i].im = 0.0;
        }
        band_factor(M, nR, 2);
        band_subst(M, tmp, 1, nR, 2);
        for (int i = 0; i < nR; ++i) gv[i] = tmp[i].re;
    }
    memcpy(G.C, G.MC + cs * (3L * start), sizeof(cx) * cs);
    memcpy(G.AC, ACg, sizeof(cx) * cs);
    const int was_exact = g_exact_mean;
    g_exact_mean = 1;
    for (int n = start; n > end; --n) {
        mean_flux_transpose(G.AC, G.Mx, 0);
        coef_to_gauss(G.AC, G.AU, 0);
        for (int k = 0; k < 3; ++k) {
            int count = n * 3 - k - 1;
            if (k < 2) {
                memcpy(G.C, G.MC + cs * (count - 1), sizeof(cx) * cs);
                coef_to_gauss(G.C, G.U, 1);
                pass2(G.U, G.AU, 0);
                memcpy(G.LU, G.U, sizeof(cx) * us);
            }
            memcpy(G.C, G.MC + cs * count, sizeof(cx) * cs);
            coef_to_gauss(G.C, G.U, 1);
            pass2(G.U, G.AU, 0);
            stage_all(G.AC, G.U, G.LU, G.AU, NULL, k, 2, 0.0);
        }
        for (int i = 0; i < nR; ++i) G.AC[ci(0, 0, i, 0)].re += weight * gv[i];
    }
    g_exact_mean = was_exact;
    memcpy(ACg, G.AC, sizeof(cx) * cs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* incremental adjoint stage: increadjproject0                          */
/* (channel/increadjproject0.c:14-357) for (0,0), then increadjproject  */
/* (channel/increadjproject.c:10-645) for row 0 (x >= 1) and every row  */
/* z != Nz/2, the order of the reference's calling loop               */
/* (shan/increadjoint_test7.c:673-685).                                 */
/* ------------------------------------------------------------------ */
/* Galerkin projections of the wall lifts (channel/increadjproject.c:40-46,
   the reference's table): LG[i] = 2 (R_i, (1-y)/2), LX[i] = (Q_i, lv) and
   LR[i] = 2 (Q_i, lv'') with lv = (1-y)^2 (1+y)/4; the projections vanish
   beyond these entries. */
static const double LG[4] = { 4. / 3., -4. / 15., -4. / 15., 4. / 35. };
static const double LR[8] = { -16. / 15., 16. / 35., 32. / 105., -32. / 105.,
                              -16. / 315., 16. / 231., 0., 0. };
static const double LX[8] = { 8. / 35., -8. / 315., -8. / 105., 8. / 385.,
                              8. / 385., -8. / 1001., -8. / 3003., 8. / 6435. };

/* viscous parts alone: K2^2 Qs + 2 K2 Qps + Qpps and -(Rps + K2 Rs) */
static void visc_alpha(double *M, double K2)
{
    for (int e = 0; e < G.dimQ * 9; ++e)
        M[e] = K2 * K2 * G.Qs[e] + 2. * K2 * G.Qps[e] + G.Qpps[e];
}

static void visc_beta(double *M, double K2)
{
    for (int e = 0; e < G.dimR * 5; ++e) M[e] = -(G.Rps[e] + K2 * G.Rs[e]);
}

/* one (z, x) column: y += coef * V * v (V banded, rows N, half width dl) */
static void add_band(const double *V, const cx *v, long vs, cx *y, long ys, int N, int dl,
                     double coef)
{
    cx t[1024];
    band_mult(V, v, vs, t, 1, N, dl);
    for (int i = 0; i < N; ++i) {
        y[i * ys].re += t[i].re * coef;
        y[i * ys].im += t[i].im * coef;
    }
}

static inline cx wall_at(const cx *W, int f, int z, int x)
{
    return W[((long) f * G.Nz + z) * G.Mx + x];
}

/* forcing value: stage s (0: k, 1: k+1), side v (0 alpha, 1 beta) */
static inline cx force_at(const cx *F, int s, int v, int z, int x, int i)
{
    cx zero = { 0.0, 0.0 };
    if (!F) return zero;
    return F[((((long) s * 2 + v) * G.Nz + z) * G.Mx + x) * G.dimR + i];
}

static void iadj_modes(cx *Ca, const cx *Ua, const cx *LUa, const cx *W, const cx *F,
                       cx *Uout, cx *Ub, int k, int count, int z, int x0)
{
    int Mx = G.Mx, nQ = G.dimQ, nR = G.dimR, q = G.qpts;
    const double *a = BA, *b = BB, *c = BC, *d = BD;
    double dt = G.dt, re = G.re, kz = G.Kz[z];
    double *M = G.band;
    const cx *mic0 = G.MIC + csize() * (long) count;
    const cx *mic1 = count > 0 ? G.MIC + csize() * (long) (count - 1) : NULL;
    for (int x = x0; x < Mx; ++x) {
        double k2 = G.K2[z * Mx + x];
        cx *al = Ca + ci(z, 0, 0, x);
        cx tm[1024];
        if (k == 0) {
            op_alpha(M, k2, re * a[0] * dt);
            band_mult(M, al, Mx, tm, 1, nQ, 4);
            for (int i = 0; i < nQ; ++i) al[i * Mx] = tm[i];
            visc_alpha(M, k2);
            add_band(M, mic0 + ci(z, 0, 0, x), Mx, al, Mx, nQ, 4, re * dt * (a[0] + b[0]));
        }
    }
    forcing_modes(Ua, z, x0, 1);
    for (int x = x0; x < Mx; ++x) {
        double k2 = G.K2[z * Mx + x], k4 = k2 * k2;
        cx *al = Ca + ci(z, 0, 0, x);
        cx wb = wall_at(W, 1, z, x), wbt = wall_at(W, 3, z, x), awb = wall_at(W, 5, z, x);
        for (int i = 0; i < nQ; ++i) {
            cx f = G.Fa[i * Mx + x], g = force_at(F, 0, 0, z, x, i);
            al[i * Mx].re += dt * c[k] * (f.re + g.re);
            al[i * Mx].im += dt * c[k] * (f.im + g.im);
        }
        for (int i = 0; i < 8 && i < nQ; ++i) {
            double m = LR[i] / 2. - k2 * LX[i], v = -k2 * LR[i] + k4 * LX[i];
            al[i * Mx].re += m * (wbt.re - wb.re) + dt * re * (a[k] * wbt.re + b[k] * wb.re) * v +
                             dt * re * (a[k] + b[k]) * awb.re * v;
            al[i * Mx].im += m * (wbt.im - wb.im) + dt * re * (a[k] * wbt.im + b[k] * wb.im) * v +
                             dt * re * (a[k] + b[k]) * awb.im * v;
        }
        op_alpha(M, k2, -(re * b[k] * dt));
        band_factor(M, nQ, 4);
        band_subst(M, al, Mx, nQ, 4);
    }
    for (int x = x0; x < Mx; ++x) {
        double k2 = G.K2[z * Mx + x];
        cx *be = Ca + ci(z, 1, 0, x);
        cx ub = wall_at(W, 0, z, x), ubt = wall_at(W, 2, z, x), aub = wall_at(W, 4, z, x);
        cx tm[1024];
        if (k == 0) {
            op_beta(M, k2, -(re * a[0] * dt));
            band_mult(M, be, Mx, tm, 1, nR, 2);
            for (int i = 0; i < nR; ++i) be[i * Mx] = tm[i];
            visc_beta(M, k2);
            add_band(M, mic0 + ci(z, 1, 0, x), Mx, be, Mx, nR, 2, re * dt * (a[0] + b[0]));
        }
        for (int i = 0; i < nR; ++i) {
            cx f = G.Fb[i * Mx + x], g = force_at(F, 0, 1, z, x, i);
            be[i * Mx].re += dt * c[k] * (f.re + g.re);
            be[i * Mx].im += dt * c[k] * (f.im + g.im);
        }
        for (int i = 0; i < 4 && i < nR; ++i) {
            double h = LG[i] / 2.;
            be[i * Mx].re += h * (ubt.re - ub.re) + dt * re * (a[k] * ubt.re + b[k] * ub.re) * (-k2 * h) +
                             dt * re * (a[k] + b[k]) * aub.re * (-k2 * h);
            be[i * Mx].im += h * (ubt.im - ub.im) + dt * re * (a[k] * ubt.im + b[k] * ub.im) * (-k2 * h) +
                             dt * re * (a[k] + b[k]) * aub.im * (-k2 * h);
        }
        op_beta(M, k2, re * b[k] * dt);
        band_factor(M, nR, 2);
        band_subst(M, be, Mx, nR, 2);
    }
    /* Gauss-point values with the wall lifts; the omega_y sums run over the
       first dimQ coefficients (channel/increadjproject.c:371-395) */
    for (int x = x0; x < Mx; ++x) {
        double kx = G.Kx[x], k2 = G.K2[z * Mx + x];
        const cx *al = Ca + ci(z, 0, 0, x), *be = Ca + ci(z, 1, 0, x);
        cx ub = wall_at(W, 0, z, x), wb = wall_at(W, 1, z, x);
        for (int i = 0; i < q; ++i) {
            double y = G.y[i];
            double vr = 0, vi = 0, fr = 0, fi = 0, qr = 0, qi = 0, gr = 0, gi = 0, pr = 0, pi = 0;
            for (int j = 0; j < nQ; ++j) {
                cx aj = al[j * Mx], bj = be[j * Mx];
                double cq = G.Q[i * nQ + j], cp = G.Qp[i * nQ + j], c2 = G.Qpp[i * nQ + j];
                double cr = G.R[i * nR + j], crp = G.Rp[i * nR + j];
                vr += cq * aj.re; vi += cq * aj.im;
                fr -= cp * aj.re; fi -= cp * aj.im;
                qr += c2 * aj.re; qi += c2 * aj.im;
                gr += cr * bj.re; gi += cr * bj.im;
                pr += crp * bj.re; pi += crp * bj.im;
            }
            double lv = (1 - y) * (1 - y) * (1 + y) * 0.25, lv1 = (3 * y * y - 2 * y - 1) * 0.25;
            double lv2 = (3 * y - 1) * 0.5, lg = (1 - y) * 0.5;
            vr += wb.re * lv; vi += wb.im * lv;
            fr -= wb.re * lv1; fi -= wb.im * lv1;
            qr += wb.re * lv2; qi += wb.im * lv2;
            gr += ub.re * lg; gi += ub.im * lg;
            pr += -ub.re / 2.; pi += -ub.im / 2.;
            cx *o = Uout;
            o[ui(z, S_V, i, x)].re = vr;
            o[ui(z, S_V, i, x)].im = vi;
            o[ui(z, S_U, i, x)].re = (kx * fi + kz * gi) / k2;
            o[ui(z, S_U, i, x)].im = -(kx * fr + kz * gr) / k2;
            o[ui(z, S_W, i, x)].re = (-kx * gi + kz * fi) / k2;
            o[ui(z, S_W, i, x)].im = (kx * gr - kz * fr) / k2;
            o[ui(z, S_DU, i, x)].re = (-kx * qi + kz * pi) / k2;
            o[ui(z, S_DU, i, x)].im = (kx * qr - kz * pr) / k2;
            o[ui(z, S_DW, i, x)].re = -(kx * pi + kz * qi) / k2;
            o[ui(z, S_DW, i, x)].im = (kx * pr + kz * qr) / k2;
        }
        /* wall values IAUxb, IAUzb (channel/increadjproject.c:446-485) */
        cx *xb = Ub + z * Mx + x, *zb = Ub + ((long) G.Nz + z) * Mx + x;
        for (int j = 0; j < nR; ++j) {
            cx bj = be[j * Mx];
            xb->re += (G.Rp0[j] * bj.im) * kz / k2;
            xb->im += (-G.Rp0[j] * bj.re) * kz / k2;
            zb->re += -(G.Rp0[j] * bj.im) * kx / k2;
            zb->im += (G.Rp0[j] * bj.re) * kx / k2;
        }
        xb->re += -ub.im / 2. * kz / k2;
        xb->im += ub.re / 2. * kz / k2;
        zb->re += ub.im / 2. * kx / k2;
        zb->im += -ub.re / 2. * kx / k2;
        for (int j = 0; j < nQ; ++j) {
            double sg = (j & 1) ? -1.0 : 1.0;
            cx aj = al[j * Mx];
            xb->re -= (sg * 8. * aj.im) * kx / k2;
            xb->im -= (-sg * 8. * aj.re) * 2 * kx / k2;
            zb->re -= (sg * 8. * aj.im) * kz / k2;
            zb->im -= (-sg * 8. * aj.re) * 2 * kz / k2;
        }
        xb->re -= -wb.im * 2. * kx / k2;
        xb->im -= wb.re * 2. * kx / k2;
        zb->re -= -wb.im * 2. * kz / k2;
        zb->im -= wb.re * 2 * kz / k2;
    }
    if (k == 2) return;
    /* explicit part of the next stage: LIU, forcing k+1, MIC[count-1] */
    forcing_modes(LUa, z, x0, 1);
    for (int x = x0; x < Mx; ++x) {
        double k2 = G.K2[z * Mx + x];
        cx *al = Ca + ci(z, 0, 0, x), *be = Ca + ci(z, 1, 0, x);
        cx tm[1024];
        op_alpha(M, k2, re * a[k + 1] * dt);
        band_mult(M, al, Mx, tm, 1, nQ, 4);
        for (int i = 0; i < nQ; ++i) {
            cx f = G.Fa[i * Mx + x], g = force_at(F, 1, 0, z, x, i);
            al[i * Mx].re = tm[i].re + dt * d[k + 1] * (f.re + g.re);
            al[i * Mx].im = tm[i].im + dt * d[k + 1] * (f.im + g.im);
        }
        visc_alpha(M, k2);
        add_band(M, mic1 + ci(z, 0, 0, x), Mx, al, Mx, nQ, 4, re * dt * (a[k + 1] + b[k + 1]));
        op_beta(M, k2, -(re * a[k + 1] * dt));
        band_mult(M, be, Mx, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) {
            cx f = G.Fb[i * Mx + x], g = force_at(F, 1, 1, z, x, i);
            be[i * Mx].re = tm[i].re + dt * d[k + 1] * (f.re + g.re);
            be[i * Mx].im = tm[i].im + dt * d[k + 1] * (f.im + g.im);
        }
        visc_beta(M, k2);
        add_band(M, mic1 + ci(z, 1, 0, x), Mx, be, Mx, nR, 2, re * dt * (a[k + 1] + b[k + 1]));
    }
}

static void iadj_mean(cx *Ca, const cx *Ua, const cx *LUa, const cx *W, const cx *F, cx *Uout,
                      cx *Ub, int k, int count)
{
    int nR = G.dimR, q = G.qpts;
    const double *a = BA, *b = BB, *c = BC, *d = BD;
    double dt = G.dt, re = G.re;
    double M[5 * 1024], fa[1024], fb[1024];
    cx *al = Ca + ci(0, 0, 0, 0), *be = Ca + ci(0, 1, 0, 0);
    long s = G.Mx;
    const cx *mic0 = G.MIC + csize() * (long) count;
    const cx *mic1 = count > 0 ? G.MIC + csize() * (long) (count - 1) : NULL;
    cx ub = wall_at(W, 0, 0, 0), wb = wall_at(W, 1, 0, 0);
    cx ubt = wall_at(W, 2, 0, 0), wbt = wall_at(W, 3, 0, 0);
    if (k == 0) {
        cx tm[1024];
        op_mean(M, -(re * a[0] * dt));
        band_mult(M, al, s, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) al[i * s] = tm[i];
        add_band(G.Rps, mic0 + ci(0, 0, 0, 0), s, al, s, nR, 2, -(dt * re * (a[0] + b[0])));
        band_mult(M, be, s, tm, 1, nR, 2);
        for (int i = 0; i < nR; ++i) be[i * s] = tm[i];
        add_band(G.Rps, mic0 + ci(0, 1, 0, 0), s, be, s, nR, 2, -(dt * re * (a[0] + b[0])));
    }
    mean_rhs(Ua, 1, fa, fb);
    op_mean(M, re * b[k] * dt);
    band_factor(M, nR, 2);
    /* only the real part of the forcing reaches the stage's right-hand side
       (channel/increadjproject0.c:130-141) */
    for (int i = 0; i < nR; ++i) al[i * s].re += dt * c[k] * (fa[i] + force_at(F, 0, 0, 0, 0, i).re);
    for (int i = 0; i < nR; ++i) {
        double lr = 0, li = 0;
        for (int j = 0; j < q; ++j) {
            lr += G.Rw[i * q + j] * (1 - G.y[j]) * 0.5 * (ubt.re - ub.re);
            li += G.Rw[i * q + j] * (1 - G.y[j]) * 0.5 * (ubt.im - ub.im);
        }
        al[i * s].re += lr;
        al[i * s].im += li;
    }
    band_subst(M, al, s, nR, 2);
    for (int i = 0; i < nR; ++i) be[i * s].re += dt * c[k] * (fb[i] + force_at(F, 0, 1, 0, 0, i).re);
    for (int i = 0; i < nR; ++i) {
        double lr = 0, li = 0;
        for (int j = 0; j < q; ++j) {
            lr += G.Rw[i * q + j] * (1 - G.y[j]) * 0.5 * (wbt.re - wb.re);
            li += G.Rw[i * q + j] * (1 - G.y[j]) * 0.5 * (wbt.im - wb.im);
        }
        be[i * s].re += lr;
        be[i * s].im += li;
    }
    band_subst(M, be, s, nR, 2);
    for (int i = 0; i < q; ++i) {
        double ur = 0, uim = 0, dr = 0, di = 0, wr = 0, wi = 0, er = 0, ei = 0;
        for (int j = 0; j < nR; ++j) {
            cx aj = al[j * s], bj = be[j * s];
            double r = G.R[i * nR + j], rp = G.Rp[i * nR + j];
            ur += r * aj.re; uim += r * aj.im;
            dr += rp * aj.re; di += rp * aj.im;
            wr += r * bj.re; wi += r * bj.im;
            er += rp * bj.re; ei += rp * bj.im;
        }
        double lg = (1 - G.y[i]) * 0.5;
        Uout[ui(0, S_U, i, 0)].re = ur + ub.re * lg;
        Uout[ui(0, S_U, i, 0)].im = uim + ub.im * lg;
        Uout[ui(0, S_DU, i, 0)].re = dr - ub.re / 2.;
        Uout[ui(0, S_DU, i, 0)].im = di - ub.im / 2.;
        Uout[ui(0, S_V, i, 0)].re = 0;
        Uout[ui(0, S_V, i, 0)].im = 0;
        Uout[ui(0, S_W, i, 0)].re = wr + wb.re * lg;
        Uout[ui(0, S_W, i, 0)].im = wi + wb.im * lg;
        Uout[ui(0, S_DW, i, 0)].re = er - wb.re / 2.;
        Uout[ui(0, S_DW, i, 0)].im = ei - wb.im / 2.;
    }
    cx *xb = Ub, *zb = Ub + (long) G.Nz * G.Mx;
    for (int j = 0; j < nR; ++j) {
        xb->re += G.Rp0[j] * al[j * s].re;
        xb->im += G.Rp0[j] * al[j * s].im;
        zb->re += G.Rp0[j] * be[j * s].re;
        zb->im += G.Rp0[j] * be[j * s].im;
    }
    xb->re -= ub.re / 2.;
    xb->im -= ub.im / 2.;
    zb->re -= wb.re / 2.;
    zb->im -= wb.im / 2.;
    if (k == 2) return;
    mean_rhs(LUa, 1, fa, fb);
    cx tm[1024];
    op_mean(M, -(re * a[k + 1] * dt));
    band_mult(M, al, s, tm, 1, nR, 2);
    for (int i = 0; i < nR; ++i) {
        cx g = force_at(F, 1, 0, 0, 0, i);
        al[i * s].re = tm[i].re + dt * d[k + 1] * (fa[i] + g.re);
        al[i * s].im = tm[i].im + dt * d[k + 1] * g.im;
    }
    add_band(G.Rps, mic1 + ci(0, 0, 0, 0), s, al, s, nR, 2, -(dt * re * (a[k + 1] + b[k + 1])));
    band_mult(M, be, s, tm, 1, nR, 2);
    for (int i = 0; i < nR; ++i) {
        cx g = force_at(F, 1, 1, 0, 0, i);
        be[i * s].re = tm[i].re + dt * d[k + 1] * (fb[i] + g.re);
        be[i * s].im = tm[i].im + dt * d[k + 1] * g.im;
    }
    add_band(G.Rps, mic1 + ci(0, 1, 0, 0), s, be, s, nR, 2, -(dt * re * (a[k + 1] + b[k + 1])));
}

/* one adjoint stage over all modes (adjproject0 + adjproject, the loop of
   channel/channel.c:668-679) on caller arrays, for pinning tests */
int orc_adjproject_stage(int k, orc_cplx *AC, orc_cplx *U, const orc_cplx *LU, orc_cplx *AU)
{
    if (!G.ok || k < 0 || k > 2 || !AC || !U || (k < 2 && !LU) || !AU) return -1;
    stage_all((cx *) AC, (cx *) U, (const cx *) LU, (cx *) AU, NULL, k, 2, 0.0);
    return 0;
}

int orc_increadjproject(int k, int count, orc_cplx *IAC, const orc_cplx *IU,
                        const orc_cplx *LIU, const orc_cplx *W, const orc_cplx *F,
                        orc_cplx *IAU, orc_cplx *IAUb)
{
    if (!G.ok || k < 0 || k > 2 || count < (k < 2 ? 1 : 0) || count > 3 * G.nsteps) return -1;
    if (!IAC || !IU || (k < 2 && !LIU) || !W || !IAU || !IAUb) return -1;
    cx *Ca = (cx *) IAC, *Uout = (cx *) IAU, *Ub = (cx *) IAUb;
    const cx *Ua = (const cx *) IU, *LUa = (const cx *) LIU, *Wc = (const cx *) W;
    const cx *Fc = (const cx *) F;
    memset(Ub, 0, sizeof(cx) * 2 * G.Nz * G.Mx);
    iadj_mean(Ca, Ua, LUa, Wc, Fc, Uout, Ub, k, count);
    iadj_modes(Ca, Ua, LUa, Wc, Fc, Uout, Ub, k, count, 0, 1);
    for (int z = 1; z < G.Nz; ++z) {
        if (z == G.Nz / 2) {
            for (int sl = 0; sl < 5; ++sl)
                memset(Uout + ui(z, sl, 0, 0), 0, sizeof(cx) * G.qpts * G.Mx);
            continue;
        }
        iadj_modes(Ca, Ua, LUa, Wc, Fc, Uout, Ub, k, count, z, 0);
    }
    return 0;
}

void orc_quad(double *y, double *w)
{
    memcpy(y, G.y, sizeof(double) * G.qpts);
    memcpy(w, G.w, sizeof(double) * G.qpts);
}

int orc_qpts(void) { return G.qpts; }
int orc_dimR(void) { return G.dimR; }
int orc_dimQ(void) { return G.dimQ; }
int orc_Nx(void) { return G.Nx; }
int orc_Nz(void) { return G.Nz; }
int orc_nsteps(void) { return G.nsteps; }
int orc_last_status(void) { return G.status; }
void orc_set_faithful_cost(int on) { G.faithful = on; }

void orc_get_state(orc_cplx *Cout, orc_cplx *Uout)
{
    if (Cout) memcpy(Cout, G.C, sizeof(cx) * csize());
    if (Uout) memcpy(Uout, G.U, sizeof(cx) * usize());
}

int orc_get_array(const char *name, double *out, int cap)
{
    struct { const char *n; double *p; long len; } t[] = {
        { "Q", G.Q, (long) G.qpts * G.dimQ }, { "Qp", G.Qp, (long) G.qpts * G.dimQ },
        { "Qpp", G.Qpp, (long) G.qpts * G.dimQ }, { "R", G.R, (long) G.qpts * G.dimR },
        { "Rp", G.Rp, (long) G.qpts * G.dimR }, { "Qw", G.Qw, (long) G.qpts * G.dimQ },
        { "Qpw", G.Qpw, (long) G.qpts * G.dimQ }, { "Qppw", G.Qppw, (long) G.qpts * G.dimQ },
        { "Rw", G.Rw, (long) G.qpts * G.dimR }, { "Rpw", G.Rpw, (long) G.qpts * G.dimR },
        { "Qs", G.Qs, 9L * G.dimQ }, { "Qps", G.Qps, 9L * G.dimQ },
        { "Qpps", G.Qpps, 9L * G.dimQ }, { "Rs", G.Rs, 5L * G.dimR },
        { "Rps", G.Rps, 5L * G.dimR }, { "Rp0", G.Rp0, G.dimR },
        { "Rpp0", G.Rpp0, G.dimR }, { "Kx", G.Kx, G.Mx }, { "Kz", G.Kz, G.Nz },
        { "cfl2", G.cfl2, G.qpts }, { "y", G.y, G.qpts }, { "w", G.w, G.qpts },
    };
    if (!G.ok) return -1;
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
        if (strcmp(name, t[i].n) == 0) {
            if (t[i].len > cap) return -1;
            memcpy(out, t[i].p, sizeof(double) * t[i].len);
            return (int) t[i].len;
        }
    return -1;
}

double orc_get_scalar(const char *name)
{
    if (!strcmp(name, "cfl1")) return G.cfl1;
    if (!strcmp(name, "cfl3")) return G.cfl3;
    if (!strcmp(name, "re")) return G.re;
    if (!strcmp(name, "flux")) return G.flux;
    if (!strcmp(name, "dt")) return G.dt;
    return NAN;
}
