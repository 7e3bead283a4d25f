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

int sc_mean_profile(int incre, int start_step, int end_step, double weight, double *profile,
                    double *shear)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!profile || !shear || start_step < 0 || end_step < start_step || end_step > d.nsteps + 1)
        return obj_fail("mean_profile requires 0 <= start_step <= end_step <= nsteps + 1 "
                        "and two output arrays");
    const int count = end_step - start_step;
    const int nchunk = std::max(1, (count + OBJ_CHUNK - 1) / OBJ_CHUNK);
    // scratch (objective_scratch_doubles): the partials, then the profile and
    // the two wall values; the wall-derivative rows sit in c.dRwall
    double *scr = c.objscr, *out = scr + (size_t) nchunk * d.dimR;
    const cplx *S = incre ? c.MIC : c.MC;
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
    launch_rebuild(c, S, c.Ut2);
    hipLaunchKernelGGL(k_plane_stats, dim3(d.qpts), dim3(OBJ_THREADS), 0, c.stream, c.Ut2, d.Nz,
                       d.Mx, d.qpts, dst);
    SC_HIP(hipMemcpyAsync(stats, dst, sizeof(double) * NSTAT * d.qpts, hipMemcpyDeviceToHost,
                          c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
    return SC_OK;
}

}  // extern "C"
