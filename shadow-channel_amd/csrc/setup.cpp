// Host-side setup of the channel solver: Gauss-Legendre quadrature, the
// Legendre-based bases and their Gram band matrices, wavenumbers, CFL
// constants, LU factors of every implicit stage operator and FFT plans.
//
// Follows channel/LegendreSetup.c:25-527 (quadrature, bases, weighted
// transposes, band Gram matrices with structurally-zero diagonals left zero),
// channel/RsQs.c:55-234 (analytic derivatives of R = (1-y^2) P_j and
// Q = (1-y^2)^2 P_j), channel/waveNumsAndCFL.c:22-96 and channel/mvOps.c:29-144
// (stride-2 banded LU without pivoting).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "sc_internal.h"

namespace sc {

static double legendre_p(int n, double x)
{
    if (n == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// ascending Gauss-Legendre nodes (mirror-symmetric) and weights
static void gauss_rule(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int m = 0; m < n / 2; ++m) {
        double t = -std::cos(M_PI * (m + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p = legendre_p(n, t), pm = legendre_p(n - 1, t);
            double dx = p / (n * (pm - t * p) / (1.0 - t * t));
            t -= dx;
            if (std::fabs(dx) <= 1e-16 * std::fabs(t) || dx == 0.0) break;
        }
        x[m] = t;
        x[n - 1 - m] = -t;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
    for (int k = 0; k < (n + 1) / 2; ++k) {
        double v = n * (legendre_p(n - 1, x[k]) - x[k] * legendre_p(n, x[k]));
        w[k] = 2.0 * (1.0 - x[k] * x[k]) / (v * v);
        w[n - 1 - k] = w[k];
    }
}

// D[i][dl+m] = sum_k A[i][k] B[k][i+2m] for |m| < nd, mirrored
static void gram_band(std::vector<double> &D, int n, int dl, int nd, const std::vector<double> &A,
                      const std::vector<double> &B, int dim, int q)
{
    int W = 2 * dl + 1;
    D.assign((size_t) n * W, 0.0);
    for (int m = 0; m < nd; ++m)
        for (int i = 0; i + 2 * m < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < q; ++k) s += A[(size_t) i * q + k] * B[(size_t) k * dim + i + 2 * m];
            D[(size_t) i * W + dl + m] = s;
            if (m > 0) D[(size_t) (i + 2 * m) * W + dl - m] = s;
        }
}

void band_factor(double *A, int N, int dl)
{
    int W = 2 * dl + 1;
    for (int i = 2; i < N; ++i) {
        int mmax = std::min(dl, i / 2);
        for (int m = mmax; m >= 1; --m) {
            double v = A[i * W + dl - m];
            for (int k = 1; m + k <= mmax; ++k)
                v -= A[i * W + dl - (m + k)] * A[(i - 2 * (m + k)) * W + dl + k];
            A[i * W + dl - m] = v / A[(i - 2 * m) * W + dl];
        }
        int umax = std::min(dl, (N - i + 1) / 2);
        for (int u = 0; u < umax; ++u) {
            double v = A[i * W + dl + u];
            int kmax = std::min(i / 2, std::min(dl, dl - u));
            for (int k = 1; k <= kmax; ++k) v -= A[i * W + dl - k] * A[(i - 2 * k) * W + dl + u + k];
            A[i * W + dl + u] = v;
        }
    }
}

// Device form of the factors: the pivot U(i, i) replaced by its reciprocal
// and the entries that fall outside the matrix set to zero, so the device
// substitution runs without bounds tests or divisions.
static void device_factors(double *A, int N, int dl)
{
    const int W = 2 * dl + 1;
    for (int i = 0; i < N; ++i) {
        A[i * W + dl] = 1.0 / A[i * W + dl];
        for (int m = 1; m <= dl; ++m) {
            if (i - 2 * m < 0) A[i * W + dl - m] = 0.0;
            if (i + 2 * m >= N) A[i * W + dl + m] = 0.0;
        }
    }
}

void band_subst_real(const double *A, double *b, int N, int dl)
{
    int W = 2 * dl + 1;
    for (int i = 2; i < N; ++i)
        for (int m = std::min(dl, i / 2); m >= 1; --m) b[i] -= A[i * W + dl - m] * b[i - 2 * m];
    for (int i = N - 1; i >= 0; --i) {
        int jmax = std::min(dl, (N - i + 1) / 2 - 1);
        for (int j = 1; j <= jmax; ++j) b[i] -= A[i * W + dl + j] * b[i + 2 * j];
        b[i] /= A[i * W + dl];
    }
}

static int plan_fft(FftPlan &p, int n)
{
    p.n = n;
    p.nf = 0;
    // fewest passes over the radices the device implements; among those, the
    // smallest per-thread register tile (ceil(FFT_TILE/(FFT_BLOCK R))
    // butterflies of R points)
    static const int RAD[] = {16, 12, 8, 6, 5, 4, 3, 2};
    auto tile = [](int r) { return (FFT_TILE / FFT_BLOCK + r - 1) / r * r; };
    std::vector<int> passes(n + 1, 1 << 20), worst(n + 1, 0), first(n + 1, 0);
    if (n >= 1) passes[1] = 0;
    for (int m = 2; m <= n; ++m) {
        if (n % m) continue;
        for (int r : RAD) {
            if (m % r || passes[m / r] >= (1 << 20)) continue;
            const int np = passes[m / r] + 1, nw = std::max(worst[m / r], tile(r));
            if (np < passes[m] || (np == passes[m] && nw < worst[m])) {
                passes[m] = np;
                worst[m] = nw;
                first[m] = r;
            }
        }
    }
    int m = n;
    while (m > 1 && first[m] && p.nf < MAX_FACTORS) {
        p.f[p.nf++] = first[m];
        m /= first[m];
    }
    if (m != 1 || n > 4096) {
        last_error() = "transform length " + std::to_string(n) +
                       " must be 2^a 3^b 5^c and at most 4096";
        return -1;
    }
    std::vector<double2> tw(n);
    for (int j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * (double) j / (double) n;
        tw[j] = make_double2(std::cos(a), std::sin(a));
    }
    if (hipMalloc(&p.tw, sizeof(double2) * n) != hipSuccess) return -1;
    if (hipMemcpy(p.tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &v)
{
    if (hipMalloc(dst, sizeof(T) * std::max<size_t>(v.size(), 1)) != hipSuccess) return -1;
    if (!v.empty() && hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    return 0;
}

static const double FWD_B[3] = {1. / 6, 2. / 3, 0.};
static const double ADJ_B[3] = {2. / 3, 1. / 6, 0.};

int setup_host(Context &c)
{
    Dims &d = c.d;
    const int q = d.qpts, nQ = d.dimQ, nR = d.dimR, Mx = d.Mx, Nz = d.Nz;
    gauss_rule(q, c.y, c.w);
    std::vector<double> P((size_t) q * nR);
    for (int m = 0; m < q; ++m) {
        P[(size_t) m * nR] = 1.0;
        if (nR > 1) P[(size_t) m * nR + 1] = c.y[m];
        for (int n = 1; n < nR - 1; ++n)
            P[(size_t) m * nR + n + 1] = ((2. * n + 1.) * (c.y[m] * P[(size_t) m * nR + n]) -
                                          n * P[(size_t) m * nR + n - 1]) / (double) (n + 1);
    }
    c.Q.assign((size_t) q * nQ, 0); c.Qp = c.Q; c.Qpp = c.Q;
    c.R.assign((size_t) q * nR, 0); c.Rp = c.R;
    for (int i = 0; i < q; ++i) {
        double yy = c.y[i], om = 1.0 - yy * yy, y2 = yy * yy;
        const double *Pi = &P[(size_t) i * nR];
        for (int j = 0; j < nR; ++j) {
            c.R[(size_t) i * nR + j] = om * Pi[j];
            c.Rp[(size_t) i * nR + j] = j == 0 ? -2.0 * yy : j * Pi[j - 1] - (j + 2.0) * (yy * Pi[j]);
        }
        for (int j = 0; j < nQ; ++j) {
            c.Q[(size_t) i * nQ + j] = om * c.R[(size_t) i * nR + j];
            c.Qp[(size_t) i * nQ + j] = j == 0 ? -4.0 * yy * om
                : om * (j * Pi[j - 1] - (j + 4.0) * (yy * Pi[j]));
            c.Qpp[(size_t) i * nQ + j] = j == 0 ? 12.0 * y2 - 4.0
                : (12.0 + 7.0 * j + j * j) * (y2 * Pi[j]) - (4.0 + j * j + j) * Pi[j] -
                  6.0 * j * (yy * Pi[j - 1]);
        }
    }
    c.Rp0.assign(nR, 0.0);
    for (int j = 0; j < nR; ++j) c.Rp0[j] = (j % 2 == 0) ? 2.0 : -2.0;
    auto wt = [&](const std::vector<double> &M, int n) {
        std::vector<double> Mw((size_t) n * q);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < q; ++j) Mw[(size_t) i * q + j] = c.w[j] * M[(size_t) j * n + i];
        return Mw;
    };
    c.Qw = wt(c.Q, nQ); c.Qpw = wt(c.Qp, nQ); c.Qppw = wt(c.Qpp, nQ);
    c.Rw = wt(c.R, nR); c.Rpw = wt(c.Rp, nR);
    gram_band(c.Qs, nQ, 4, 5, c.Qw, c.Q, nQ, q);
    gram_band(c.Qps, nQ, 4, 4, c.Qpw, c.Qp, nQ, q);
    gram_band(c.Qpps, nQ, 4, 3, c.Qppw, c.Qpp, nQ, q);
    gram_band(c.Rs, nR, 2, 3, c.Rw, c.R, nR, q);
    gram_band(c.Rps, nR, 2, 2, c.Rpw, c.Rp, nR, q);
    c.Kx.assign(Mx, 0.0);
    c.Kz.assign(Nz, 0.0);
    for (int i = 0; i < Mx; ++i) c.Kx[i] = i * (2. * M_PI / d.Lx);
    for (int i = 0; i < Nz / 2 + 1; ++i) c.Kz[i] = i * (2. * M_PI / d.Lz);
    for (int i = Nz / 2 + 1; i < Nz; ++i) c.Kz[i] = (i - Nz) * (2. * M_PI / d.Lz);
    d.cfl1 = M_PI / (d.Lx / d.Nx);
    d.cfl3 = M_PI / (d.Lz / d.Nz);
    c.cfl2.assign(q, 0.0);
    for (int j = 1; j < q - 1; ++j) c.cfl2[j] = 2. * M_PI / (c.y[j + 1] - c.y[j - 1]);
    c.cfl2[0] = c.cfl2[q - 1] = M_PI / (c.y[0] + 1.0);
    std::vector<double> rowsum(nR, 0.0);
    for (int i = 0; i < nR; ++i) {
        double t = 0.0;
        for (int j = 0; j < q; ++j) t += c.Rw[(size_t) i * q + j];
        rowsum[i] = t;
    }

    // upload constants
    if (upload(&c.dKx, c.Kx) || upload(&c.dKz, c.Kz) || upload(&c.dcfl2, c.cfl2) ||
        upload(&c.dW, c.w) || upload(&c.dQ, c.Q) || upload(&c.dQp, c.Qp) ||
        upload(&c.dQpp, c.Qpp) || upload(&c.dR, c.R) || upload(&c.dRp, c.Rp) ||
        upload(&c.dQw, c.Qw) || upload(&c.dQpw, c.Qpw) || upload(&c.dQppw, c.Qppw) ||
        upload(&c.dRw, c.Rw) || upload(&c.dRpw, c.Rpw) || upload(&c.dQs, c.Qs) ||
        upload(&c.dQps, c.Qps) || upload(&c.dQpps, c.Qpps) || upload(&c.dRs, c.Rs) ||
        upload(&c.dRps, c.Rps) || upload(&c.dRowsum, rowsum)) {
        last_error() = "device allocation of constants failed";
        return -1;
    }

    // y-transform operands in MFMA fragment order, zero outside the bases
    {
        const int nyt = (q + 15) / 16, nksR = ((nR + 3) / 4 + KC_REBUILD - 1) / KC_REBUILD * KC_REBUILD;
        std::vector<double> fr((size_t) nyt * nksR * 5 * 64, 0.0);
        const std::vector<double> *B[5] = {&c.Q, &c.Qp, &c.Qpp, &c.R, &c.Rp};
        for (int t = 0; t < nyt; ++t)
            for (int s = 0; s < nksR; ++s)
                for (int m = 0; m < 5; ++m)
                    for (int l = 0; l < 64; ++l) {
                        const int y = 16 * t + l % 16, j = 4 * s + l / 16, dim = m < 3 ? nQ : nR;
                        if (y < q && j < dim)
                            fr[(((size_t) t * nksR + s) * 5 + m) * 64 + l] = (*B[m])[(size_t) y * dim + j];
                    }
        const int nit = (nR + 15) / 16, nksF = ((q + 3) / 4 + KC_FORCE - 1) / KC_FORCE * KC_FORCE;
        std::vector<double> ff((size_t) nit * nksF * 5 * 64, 0.0);
        const std::vector<double> *Wt[5] = {&c.Qpw, &c.Qw, &c.Qppw, &c.Rw, &c.Rpw};
        for (int t = 0; t < nit; ++t)
            for (int s = 0; s < nksF; ++s)
                for (int m = 0; m < 5; ++m)
                    for (int l = 0; l < 64; ++l) {
                        const int i = 16 * t + l % 16, y = 4 * s + l / 16, dim = m < 3 ? nQ : nR;
                        if (i < dim && y < q)
                            ff[(((size_t) t * nksF + s) * 5 + m) * 64 + l] = (*Wt[m])[(size_t) i * q + y];
                    }
        c.nyt = nyt; c.nksR = nksR; c.nit = nit; c.nksF = nksF;
        // parity of the bases on the mirrored Gauss points, checked exactly:
        // Q, Qpp, R of column j have the parity of j, Qp and Rp the opposite
        const int hq = (q + 1) / 2;
        bool par = true;
        for (int i = 0; i < hq && par; ++i) {
            const int im = q - 1 - i;
            par = c.y[im] == -c.y[i];
            for (int m = 0; m < 5 && par; ++m) {
                const int dim = m < 3 ? nQ : nR;
                const bool deriv = m == 1 || m == 4;
                for (int j = 0; j < dim && par; ++j) {
                    const double b = (*B[m])[(size_t) i * dim + j];
                    const double bm = (*B[m])[(size_t) im * dim + j];
                    par = bm == (((j & 1) != 0) != deriv ? -b : b);
                }
            }
        }
        // k-steps per parity: whole groups of four (the kernel moves its
        // coefficients in quarters of two-k-step pairs)
        const int nkh = ((((nR + 1) / 2 + 3) / 4) + 3) / 4 * 4;
        c.ypar = par;
        c.nht = (hq + 15) / 16;
        c.nksP = 2 * nkh;
        std::vector<double> fp((size_t) c.nht * c.nksP * 5 * 64, 0.0);
        for (int t = 0; t < c.nht; ++t)
            for (int s = 0; s < c.nksP; ++s)
                for (int m = 0; m < 5; ++m)
                    for (int l = 0; l < 64; ++l) {
                        const int odd = s >= nkh, so = s - odd * nkh;
                        const int y = 16 * t + l % 16, j = 8 * so + 2 * (l / 16) + odd;
                        const int dim = m < 3 ? nQ : nR;
                        if (y < hq && j < dim)
                            fp[(((size_t) t * c.nksP + s) * 5 + m) * 64 + l] = (*B[m])[(size_t) y * dim + j];
                    }
        // weighted bases of the projections for the parity form: the same
        // parity rule (Qpw, Rpw opposite to their row), checked on the weights
        for (int i = 0; i < hq && par; ++i) par = c.w[q - 1 - i] == c.w[i];
        c.ypar = par;
        c.npair = ((nR + 1) / 2 + 15) / 16;
        // k-steps over the half domain, rounded up to a count the kernels
        // are built for (8, 16, 20, 28; larger counts use the direct form)
        {
            const int need = (hq + 3) / 4;
            c.nksQ = need;
            for (int n : {28, 20, 16, 8})
                if (need <= n) c.nksQ = n;
        }
        {
            const std::vector<double> *Wq[5] = {&c.Qpw, &c.Qw, &c.Qppw, &c.Rw, &c.Rpw};
            std::vector<double> fq((size_t) c.npair * c.nksQ * 2 * 5 * 64, 0.0);
            for (int t = 0; t < c.npair; ++t)
                for (int s = 0; s < c.nksQ; ++s)
                    for (int pr = 0; pr < 2; ++pr)
                        for (int m = 0; m < 5; ++m)
                            for (int l = 0; l < 64; ++l) {
                                const int i = 2 * (16 * t + l % 16) + pr, y = 4 * s + l / 16;
                                const int dim = m < 3 ? nQ : nR;
                                if (i < dim && y < hq)
                                    fq[((((size_t) t * c.nksQ + s) * 2 + pr) * 5 + m) * 64 + l] =
                                        (*Wq[m])[(size_t) i * q + y] * (y == q - 1 - y ? 0.5 : 1.0);
                            }
            if (upload(&c.dFragQ, fq)) {
                last_error() = "device allocation of transform operands failed";
                return -1;
            }
        }
        if (upload(&c.dFragP, fp)) {
            last_error() = "device allocation of transform operands failed";
            return -1;
        }
        const size_t qbytes = sizeof(unsigned) * 2 * 8 * (size_t) nyt;
        if (hipMalloc(&c.rb_queue, qbytes) != hipSuccess || hipMemset(c.rb_queue, 0, qbytes) != hipSuccess) {
            last_error() = "device allocation of the rebuild queues failed";
            return -1;
        }
        if (upload(&c.dFragR, fr) || upload(&c.dFragF, ff)) {
            last_error() = "device allocation of transform operands failed";
            return -1;
        }
    }

    // LU factors of the implicit operators per direction and stage
    // v:      -(K2 Qs + Qps) - re b dt (K2^2 Qs + 2 K2 Qps + Qpps)
    // omega:  Rs + re b dt (Rps + K2 Rs)
    // mean:   Rs + re b dt Rps
    std::vector<double> A9(9 * (size_t) nQ), A5(5 * (size_t) nR);
    const int sbt = solve_columns_per_block(nR);
    if (sbt <= 0) {
        last_error() = "no solve block shape for dimR " + std::to_string(nR);
        return -1;
    }
    c.sbt = sbt;
    // Columns kz and Nz - kz have the same k^2 (Kz[Nz - z] = -Kz[z] exactly),
    // so the same factors: when the blocks follow the kz rows (Mx % sbt ==
    // 0), only the rows z <= Nz/2 are stored and the solve reads the mirror
    // row's image for z > Nz/2.
    const bool mirror = Mx % sbt == 0;
    c.factor_mirror = mirror;
    const int zrows = mirror ? Nz / 2 + 1 : Nz;
    const size_t nblk = ((size_t) zrows * Mx + sbt - 1) / sbt;
    const size_t rqa = (nQ + 1) / 2, rqb = (nR + 1) / 2;
    // position of factor j of matrix row i of column col in the block images
    auto image = [&](size_t rmax, int W, int col, int i, int j) {
        const size_t b = col / sbt, cc = col % sbt, par = i & 1, r = i >> 1;
        return (((b * 2 + par) * rmax + r) * W + j) * sbt + cc;
    };
    for (int dir = 0; dir < 2; ++dir)
        for (int k = 0; k < 3; ++k) {
            const double bk = dir == 0 ? FWD_B[k] : ADJ_B[k];
            const double coef = d.re * bk * d.dt;
            std::vector<double> La(nblk * 2 * rqa * 9 * sbt, 0.0), Lb(nblk * 2 * rqb * 5 * sbt, 0.0);
            for (int z = 0; z < zrows; ++z)
                for (int x = 0; x < Mx; ++x) {
                    if ((z == 0 && x == 0) || z == Nz / 2) continue;
                    const int col = z * Mx + x;
                    double k2 = c.Kz[z] * c.Kz[z] + c.Kx[x] * c.Kx[x], s = k2 * k2;
                    for (int e = 0; e < 9 * nQ; ++e)
                        A9[e] = -(k2 * c.Qs[e] + c.Qps[e]) -
                                coef * (s * c.Qs[e] + 2. * k2 * c.Qps[e] + c.Qpps[e]);
                    band_factor(A9.data(), nQ, 4);
                    device_factors(A9.data(), nQ, 4);
                    for (int i = 0; i < nQ; ++i)
                        for (int j = 0; j < 9; ++j)
                            La[image(rqa, 9, col, i, j)] = A9[(size_t) i * 9 + j];
                    for (int e = 0; e < 5 * nR; ++e)
                        A5[e] = c.Rs[e] + coef * (c.Rps[e] + k2 * c.Rs[e]);
                    band_factor(A5.data(), nR, 2);
                    device_factors(A5.data(), nR, 2);
                    for (int i = 0; i < nR; ++i)
                        for (int j = 0; j < 5; ++j)
                            Lb[image(rqb, 5, col, i, j)] = A5[(size_t) i * 5 + j];
                }
            if (upload(&c.dLa[dir][k], La) || upload(&c.dLb[dir][k], Lb)) {
                last_error() = "device allocation of factors failed";
                return -1;
            }
            // mean mode: LHS factors and the flux-correction vector
            for (int e = 0; e < 5 * nR; ++e) A5[e] = c.Rs[e] + coef * c.Rps[e];
            std::vector<double> Lm(A5);
            band_factor(Lm.data(), nR, 2);
            device_factors(Lm.data(), nR, 2);
            // primal and tangent use the mass matrix (channel/project0.c:85-98),
            // the adjoint the stage matrix (channel/adjproject0.c:82-95)
            std::vector<double> M = dir == 0 ? c.Rs : A5;
            band_factor(M.data(), nR, 2);
            std::vector<double> add(rowsum);
            band_subst_real(M.data(), add.data(), nR, 2);
            if (upload(&c.dLm[dir][k], Lm) || upload(&c.dAdd[dir][k], add)) {
                last_error() = "device allocation of mean factors failed";
                return -1;
            }
        }
    if (plan_fft(c.pz, d.Pz) || plan_fft(c.px, d.Px)) return -1;
    return 0;
}

}  // namespace sc
