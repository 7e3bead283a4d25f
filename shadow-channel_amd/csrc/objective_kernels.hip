// Objectives and statistics of stored states as device reductions.
//
// The reference forms its objectives from full physical fields: velAvg /
// IvelAvg / uavgGrad (channel/__init__.py:217-299) run spec2phys on every
// stored step and average the x-z plane on the host, and comp_stat
// (channel/comp_stat.c:9-230) transforms U at every Gauss point to form plane
// means and second moments.  On the dealiased grid those plane averages are
// exact spectral sums, so they are computed here from the coefficients:
//  * the plane mean of u at Gauss point y is Re u_00(y) = sum_j R_j(y)
//    Re alpha_j(0,0) (the mean column holds the streamwise mean profile in
//    the R basis; the complex-to-real x transform reads the real part of the
//    DC bin), so a time average over stored steps is R applied to the time
//    sum of Re alpha(0,0): one pass over dimR values per step;
//  * the wall shear du_00/dy at y = -1 (the Uxb(0,0) of
//    channel/project0.c:227-234) and at y = +1 is the same sum with R_j'(-1)
//    and R_j'(+1);
//  * the plane mean of a product f g is sum_k c_k Re(f_k conj(g_k)) with
//    c = 2 for kx > 0 and c = 1 at kx = 0, where each coefficient is replaced
//    by its Hermitian part (f(kz) + conj f(-kz)) / 2, the only part the
//    complex-to-real x transform sees (Parseval; exact because the product
//    of two fields band-limited to |kx| < Nx/2 has no mode at the padded
//    grid's aliases +-3Nx/2, and likewise in z).
#include <algorithm>

#include "../../include/shadow_channel.h"
#include "sc_internal.h"

namespace sc {

constexpr int OBJ_CHUNK = 32;   // stored steps per partial-sum block
constexpr int OBJ_THREADS = 256;

// part[b][j] = sum over the steps of chunk b of Re alpha_j(0,0); stored step
// t of the window is S + (start + t) * stride (stride: one step, 3 stages).
__global__ __launch_bounds__(OBJ_THREADS) void k_mean_alpha_partial(const cplx *S, size_t stride,
                                                                    int start, int count,
                                                                    int dimR, int Mx,
                                                                    double *part)
{
    const int t0 = blockIdx.x * OBJ_CHUNK, t1 = min(t0 + OBJ_CHUNK, count);
    for (int j = threadIdx.x; j < dimR; j += OBJ_THREADS) {
        double s = 0.0;
        for (int t = t0; t < t1; ++t) s += S[(size_t) (start + t) * stride + (size_t) j * Mx].x;
        part[(size_t) blockIdx.x * dimR + j] = s;
    }
}

// out[y] = weight sum_j R[y][j] s_j for y < qpts, out[qpts + w] = weight
// sum_j Rwall[w][j] s_j (w = 0: y = -1, 1: y = +1), with s_j the sum of the
// nchunk partials (in chunk order, so the result does not depend on timing).
__global__ __launch_bounds__(OBJ_THREADS) void k_mean_profile_final(const double *part, int nchunk,
                                                                    int dimR, int qpts,
                                                                    const double *R,
                                                                    const double *Rwall,
                                                                    double weight, double *out)
{
    __shared__ double s[512];
    for (int j = threadIdx.x; j < dimR; j += OBJ_THREADS) {
        double v = 0.0;
        for (int b = 0; b < nchunk; ++b) v += part[(size_t) b * dimR + j];
        s[j] = weight * v;
    }
    __syncthreads();
    for (int y = threadIdx.x; y < qpts + 2; y += OBJ_THREADS) {
        const double *row = y < qpts ? R + (size_t) y * dimR : Rwall + (size_t) (y - qpts) * dimR;
        double v = 0.0;
        for (int j = 0; j < dimR; ++j) v += row[j] * s[j];
        out[y] = v;
    }
}

// Plane statistics of U ([Nz][5][qpts][Mx]: u, v, w, du/dy, dw/dy) at Gauss
// point y = blockIdx.x, into stats[q][y] for the quantities
// q = 0 u, 1 v, 2 w, 3 uv, 4 uw, 5 vw, 6 uu, 7 vv, 8 ww, 9 du/dy.
constexpr int NSTAT = 10;
__global__ __launch_bounds__(OBJ_THREADS) void k_plane_stats(const cplx *U, int Nz, int Mx,
                                                             int qpts, double *stats)
{
    const int y = blockIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};   // uv, uw, vw, uu, vv, ww
    auto at = [&](int z, int s, int x) { return U[(((size_t) z * 5 + s) * qpts + y) * Mx + x]; };
    for (int col = threadIdx.x; col < Nz * Mx; col += OBJ_THREADS) {
        const int z = col / Mx, x = col - z * Mx;
        if (z == Nz / 2) continue;                      // row Nz/2 is never transformed
        cplx f[3];
        for (int s = 0; s < 3; ++s) f[s] = at(z, s, x);
        double c = 2.0;
        if (x == 0) {
            // Hermitian part against the mirror row -kz
            const int zm = (Nz - z) % Nz;
            for (int s = 0; s < 3; ++s) {
                const cplx g = at(zm, s, 0);
                f[s] = make_double2(0.5 * (f[s].x + g.x), 0.5 * (f[s].y - g.y));
            }
            c = 1.0;
        }
        auto re = [](cplx a, cplx b) { return a.x * b.x + a.y * b.y; };
        acc[0] += c * re(f[0], f[1]);
        acc[1] += c * re(f[0], f[2]);
        acc[2] += c * re(f[1], f[2]);
        acc[3] += c * re(f[0], f[0]);
        acc[4] += c * re(f[1], f[1]);
        acc[5] += c * re(f[2], f[2]);
    }
    // block sums in a fixed order: lane exchanges, then the wave sums in order
    __shared__ double red[6][OBJ_THREADS / 64];
    for (int q = 0; q < 6; ++q) {
        double v = acc[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 6; ++q) {
            double v = 0.0;
            for (int w = 0; w < OBJ_THREADS / 64; ++w) v += red[q][w];
            stats[(size_t) (3 + q) * qpts + y] = v;
        }
        // means: the real part of the mean column (Hermitian part of (0,0))
        for (int s = 0; s < 3; ++s) stats[(size_t) s * qpts + y] = at(0, s, 0).x;
        stats[(size_t) 9 * qpts + y] = at(0, 3, 0).x;
    }
}

// ------------------------------------------------------------ sensitivities
// Reynolds-number source of the tangent's stage (per unit change of Re).  The
// stage operators M + re a_k dt D (explicit) and M - re b_k dt D (implicit)
// depend on re = 1/Re, so the right-hand side of the tangent's stage k gains
//   Src = TF(dt a_k Y + dt b_k X),  TF(S) = (dre/dRe) D S = -re^2 D S,
// TF being the forcing of channel/tangent_forcing.c and tangent_forcing0.c:
// Y = MC[3n+k] (what the stage's explicit part acts on), X = MC[3n+k+1] (its
// solution).  In the mean column the implicit operator acts on the solve's
// own result, before the flux correction alpha += e add / 2 the stored X
// carries: X_pre = X - e add / 2 with e = (flux - w1.Re X) / (1 - w1.add / 2).
// Non-mean columns: one thread per element of Src (C layout); the mean column
// is k_re_source_mean's.  Grid y: member (strides ms_y, ms_x, ms_s).
__global__ void k_re_source(const cplx *Y, const cplx *X, double ca, double cb, const double *Qs,
                            const double *Qps, const double *Qpps, const double *Rs,
                            const double *Rps, const double *Kx, const double *Kz, double re,
                            int Nz, int Mx, int nQ, int nR, cplx *Src, size_t ms_y, size_t ms_x,
                            size_t ms_s)
{
    const size_t n = (size_t) Nz * 2 * nR * Mx;
    const size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
    if (e >= n) return;
    Y += blockIdx.y * ms_y;
    X += blockIdx.y * ms_x;
    Src += blockIdx.y * ms_s;
    const int x = (int) (e % Mx), i = (int) ((e / Mx) % nR), v = (int) ((e / ((size_t) Mx * nR)) % 2);
    const int z = (int) (e / ((size_t) Mx * nR * 2));
    if ((z == 0 && x == 0) || (v == 0 && i >= nQ)) {
        if (!(z == 0 && x == 0)) Src[e] = make_double2(0.0, 0.0);
        return;
    }
    const double k2 = Kz[z] * Kz[z] + Kx[x] * Kx[x], s2 = k2 * k2;
    const size_t col = ((size_t) (z * 2 + v) * nR) * Mx + x;
    double sr = 0.0, si = 0.0;
    const int N = v == 0 ? nQ : nR, dl = v == 0 ? 4 : 2;
    for (int m = -min(dl, i / 2); m <= min(dl, (N - i + 1) / 2 - 1); ++m) {
        const size_t o = col + (size_t) (i + 2 * m) * Mx;
        const double yr = ca * Y[o].x + cb * X[o].x, yi = ca * Y[o].y + cb * X[o].y;
        double g;
        if (v == 0) {
            const int b = i * 9 + 4 + m;
            g = -re * re * (s2 * Qs[b] + 2. * k2 * Qps[b] + Qpps[b]);
        } else {
            const int b = i * 5 + 2 + m;
            g = re * re * (Rps[b] + k2 * Rs[b]);
        }
        sr += g * yr;
        si += g * yi;
    }
    Src[e] = make_double2(sr, si);
}

// the mean column of the source: one block per member (grid x)
__global__ __launch_bounds__(OBJ_THREADS) void k_re_source_mean(
    const cplx *Y, const cplx *X, double ca, double cb, const double *Rps, const double *w1,
    const double *add, double flux, double re, int nR, int Mx, cplx *Src, size_t ms_y,
    size_t ms_x, size_t ms_s)
{
    Y += blockIdx.x * ms_y;
    X += blockIdx.x * ms_x;
    Src += blockIdx.x * ms_s;
    __shared__ double red[2][OBJ_THREADS];
    __shared__ double2 comb[2][512];
    double wx = 0.0, wa = 0.0;
    for (int i = threadIdx.x; i < nR; i += OBJ_THREADS) {
        wx += w1[i] * X[(size_t) i * Mx].x;
        wa += w1[i] * add[i];
    }
    red[0][threadIdx.x] = wx;
    red[1][threadIdx.x] = wa;
    __syncthreads();
    for (int o = OBJ_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    const double e = (flux - red[0][0]) / (1.0 - 0.5 * red[1][0]);
    for (int i = threadIdx.x; i < nR; i += OBJ_THREADS) {
        const cplx ya = Y[(size_t) i * Mx], xa = X[(size_t) i * Mx];
        const cplx yb = Y[(size_t) (nR + i) * Mx], xb = X[(size_t) (nR + i) * Mx];
        comb[0][i] = make_double2(ca * ya.x + cb * (xa.x - 0.5 * e * add[i]), ca * ya.y + cb * xa.y);
        comb[1][i] = make_double2(ca * yb.x + cb * xb.x, ca * yb.y + cb * xb.y);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * nR; t += OBJ_THREADS) {
        const int v = t / nR, i = t - v * nR;
        double sr = 0.0, si = 0.0;
        for (int m = -min(2, i / 2); m <= min(2, (nR - i + 1) / 2 - 1); ++m) {
            const double g = re * re * Rps[i * 5 + 2 + m];
            sr += g * comb[v][i + 2 * m].x;
            si += g * comb[v][i + 2 * m].y;
        }
        Src[(size_t) (v * nR + i) * Mx] = make_double2(sr, si);
    }
}

// alpha(0,0) <- P alpha(0,0) on the real part, P = I - add w1^T / 2 (the
// tangent's flux correction, self-adjoint in the mean column's pairing): the
// exact adjoint's step start.  One block per member.
__global__ __launch_bounds__(OBJ_THREADS) void k_mean_project(cplx *A, const double *w1,
                                                              const double *add, int nR, int Mx,
                                                              size_t ms)
{
    A += blockIdx.x * ms;
    __shared__ double red[OBJ_THREADS];
    double t = 0.0;
    for (int i = threadIdx.x; i < nR; i += OBJ_THREADS) t += w1[i] * A[(size_t) i * Mx].x;
    red[threadIdx.x] = t;
    __syncthreads();
    for (int o = OBJ_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double h = 0.5 * red[0];
    for (int i = threadIdx.x; i < nR; i += OBJ_THREADS) A[(size_t) i * Mx].x -= h * add[i];
}

// alpha(0,0) += weight g on the real part (the objective's representer).
// One block per member.
__global__ void k_add_mean(cplx *A, const double *g, double weight, int nR, int Mx, size_t ms)
{
    A += blockIdx.x * ms;
    for (int i = threadIdx.x; i < nR; i += blockDim.x) A[(size_t) i * Mx].x += weight * g[i];
}

void launch_re_source(Context &c, const cplx *Y, const cplx *X, double ca, double cb, cplx *Src)
{
    const Dims &d = c.d;
    const size_t n = (size_t) d.Nz * 2 * d.dimR * d.Mx;
    const size_t my = mstride(c, Y), mx = mstride(c, X), ms = mstride(c, Src);
    hipLaunchKernelGGL(k_re_source, dim3((unsigned) ((n + 255) / 256), c.nrun), dim3(256), 0,
                       c.stream, Y, X, ca, cb, c.dQs, c.dQps, c.dQpps, c.dRs, c.dRps, c.dKx,
                       c.dKz, d.re, d.Nz, d.Mx, d.dimQ, d.dimR, Src, my, mx, ms);
    hipLaunchKernelGGL(k_re_source_mean, dim3(c.nrun), dim3(OBJ_THREADS), 0, c.stream, Y, X, ca,
                       cb, c.dRps, c.dRowsum, c.dAdd[0][0], d.flux, d.re, d.dimR, d.Mx, Src, my,
                       mx, ms);
}

void launch_mean_project(Context &c, cplx *A)
{
    hipLaunchKernelGGL(k_mean_project, dim3(c.nrun), dim3(OBJ_THREADS), 0, c.stream, A,
                       c.dRowsum, c.dAdd[0][0], c.d.dimR, c.d.Mx, mstride(c, A));
}

void launch_add_mean(Context &c, cplx *A, const double *g, double weight)
{
    hipLaunchKernelGGL(k_add_mean, dim3(c.nrun), dim3(256), 0, c.stream, A, g, weight, c.d.dimR,
                       c.d.Mx, mstride(c, A));
}

size_t objective_scratch_doubles(const Dims &d)
{
    const size_t nchunk = (size_t) std::max(1, (d.nsteps + 1 + OBJ_CHUNK - 1) / OBJ_CHUNK);
    return std::max(nchunk * d.dimR + d.qpts + 2, (size_t) NSTAT * d.qpts);
}

}  // namespace sc

using namespace sc;

static int obj_fail(const char *msg)
{
    last_error() = msg;
    return SC_EINVAL;
}

extern "C" {

// the profile and wall shear of the stored trajectory S (one member's MC or MIC)
static int mean_profile(Context &c, const cplx *S, int start_step, int end_step, double weight,
                        double *profile, double *shear)
{
    const Dims &d = c.d;
    const int count = end_step - start_step;
    const int nchunk = std::max(1, (count + OBJ_CHUNK - 1) / OBJ_CHUNK);
    // scratch (objective_scratch_doubles): the partials, then the profile and
    // the two wall values; the wall-derivative rows sit in c.dRwall
    double *scr = c.objscr, *out = scr + (size_t) nchunk * d.dimR;
    // (an empty window writes zero partials)
    hipLaunchKernelGGL(k_mean_alpha_partial, dim3(nchunk), dim3(OBJ_THREADS), 0, c.stream, S,
                       3 * c.csize, start_step, count, d.dimR, d.Mx, scr);
    hipLaunchKernelGGL(k_mean_profile_final, dim3(1), dim3(OBJ_THREADS), 0, c.stream, scr, nchunk,
                       d.dimR, d.qpts, c.dR, c.dRwall, weight, out);
    std::vector<double> h(d.qpts + 2);
    SC_HIP(hipMemcpyAsync(h.data(), out, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                          c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
    std::copy(h.begin(), h.begin() + d.qpts, profile);
    shear[0] = h[d.qpts];
    shear[1] = h[d.qpts + 1];
    return SC_OK;
}

int sc_mean_profile(int incre, int start_step, int end_step, double weight, double *profile,
                    double *shear)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!profile || !shear || start_step < 0 || end_step < start_step || end_step > d.nsteps + 1)
        return obj_fail("mean_profile requires 0 <= start_step <= end_step <= nsteps + 1 "
                        "and two output arrays");
    return mean_profile(c, incre ? c.MIC : c.MC, start_step, end_step, weight, profile, shear);
}

int sc_member_mean_profile(int member, int incre, int start_step, int end_step, double weight,
                           double *profile, double *shear)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!profile || !shear || member < 0 || member >= c.nmem || start_step < 0 ||
        end_step < start_step || end_step > d.nsteps + 1)
        return obj_fail("member_mean_profile requires 0 <= member < members, 0 <= start_step "
                        "<= end_step <= nsteps + 1 and two output arrays");
    const cplx *S = (const cplx *) ((const char *) (incre ? c.MIC : c.MC) +
                                    (size_t) member * c.mcbytes);
    return mean_profile(c, S, start_step, end_step, weight, profile, shear);
}

int sc_step_statistics(int incre, int i_step, double *stats)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!stats || i_step < 0 || i_step > d.nsteps)
        return obj_fail("step_statistics requires 0 <= i_step <= nsteps and an output array");
    const cplx *S = (incre ? c.MIC : c.MC) + 3 * c.csize * (size_t) i_step;
    double *dst = c.objscr;
    // the state's Gauss-point values in the spare field buffer (all five
    // slots: du/dy is one of the statistics)
    c.fp64_out = true;                 // the statistics read FP64 slots
    launch_rebuild(c, S, c.Ut2);
    c.fp64_out = false;
    hipLaunchKernelGGL(k_plane_stats, dim3(d.qpts), dim3(OBJ_THREADS), 0, c.stream, c.Ut2, d.Nz,
                       d.Mx, d.qpts, dst);
    SC_HIP(hipMemcpyAsync(stats, dst, sizeof(double) * NSTAT * d.qpts, hipMemcpyDeviceToHost,
                          c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
    return SC_OK;
}

}  // extern "C"
