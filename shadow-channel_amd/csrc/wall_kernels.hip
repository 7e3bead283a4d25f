// Wall-plane kernels: the wall values of a coefficient array
// (channel/initAlphaBeta2.c:70-81, 210-223) and the incremental boundary
// condition built from them (channel/increBoundary.c:20-162).
//
// The planes are [Nz][Mx] complex (kx fastest), a few hundred KB at most, and
// the product is taken once per restart, so the transforms are direct sums
// over the stored modes (one output per thread, operands and the library's
// twiddle tables exp(+2 pi i m / P) in LDS) rather than FFT passes: three
// small launches with coalesced loads and stores.
#include "sc_internal.h"

namespace sc {

// One thread per (z, x).  R_j'(-1) = 2 (-1)^j (the Rp0 of the setup).
// (0, 0): Uxb = Rp0.alpha, Uzb = Rp0.beta; other columns, every kz row:
// Uxb = -i kz/k2 Rp0.beta, Uzb = i kx/k2 Rp0.beta.
__global__ void k_wall_plane(const cplx *C, const double *Kx, const double *Kz, int Nz, int Mx,
                             int dimR, cplx *Uxb, cplx *Uzb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Nz * Mx) return;
    const int z = e / Mx, x = e - z * Mx;
    const cplx *a = C + (size_t) z * 2 * dimR * Mx + x;
    const cplx *b = a + (size_t) dimR * Mx;
    double br = 0.0, bi = 0.0, ar = 0.0, ai = 0.0;
    for (int j = 0; j < dimR; ++j) {
        const double r0 = (j & 1) ? -2.0 : 2.0;
        const cplx bj = b[(size_t) j * Mx];
        br += r0 * bj.x;
        bi += r0 * bj.y;
        if (e == 0) {
            const cplx aj = a[(size_t) j * Mx];
            ar += r0 * aj.x;
            ai += r0 * aj.y;
        }
    }
    if (e == 0) {
        Uxb[0] = make_double2(ar, ai);
        Uzb[0] = make_double2(br, bi);
        return;
    }
    const double kx = Kx[x], kz = Kz[z], k2 = kx * kx + kz * kz;
    Uxb[e] = make_double2(bi * kz / k2, -br * kz / k2);
    Uzb[e] = make_double2(-bi * kx / k2, br * kx / k2);
}

// grid (Mx, 3): z inverse transform of column x of field f (0: Uxb, 1: Uzb,
// 2: the plane w.n, Re = 1 at kz row 1, kx 1) into T[f][Pz][Mx].  Stored kz
// row r < Nz/2 sits at padded row r, r > Nz/2 at r + Nz/2; row Nz/2 is not
// loaded.
__global__ void k_wall_zinv(const cplx *Uxb, const cplx *Uzb, const cplx *twz, int Nz, int Mx,
                            int Pz, cplx *T)
{
    extern __shared__ cplx sh[];
    cplx *col = sh, *tw = sh + Nz;
    const int x = blockIdx.x, f = blockIdx.y;
    for (int i = threadIdx.x; i < Pz; i += blockDim.x) tw[i] = twz[i];
    for (int r = threadIdx.x; r < Nz; r += blockDim.x) {
        cplx v = make_double2(0.0, 0.0);
        if (r != Nz / 2) {
            if (f == 0)
                v = Uxb[(size_t) r * Mx + x];
            else if (f == 1)
                v = Uzb[(size_t) r * Mx + x];
            else if (r == 1 && x == 1)
                v = make_double2(1.0, 0.0);
        }
        col[r] = v;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < Pz; p += blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < Nz; ++r) {
            if (r == Nz / 2) continue;
            const int pr = r < Nz / 2 ? r : r + Nz / 2;
            const cplx t = tw[(pr * p) % Pz];
            const cplx v = col[r];
            sr += v.x * t.x - v.y * t.y;
            si += v.x * t.y + v.y * t.x;
        }
        T[((size_t) f * Pz + p) * Mx + x] = make_double2(sr, si);
    }
}

// grid (Pz): row p of the three fields to physical x (complex-to-real with
// the Hermitian extension; the real part only at kx = 0), the products
// Uxb w.n and Uzb w.n, and their real-to-complex transform at kx < Mx into
// T2[2][Pz][Mx] (unscaled).
__global__ void k_wall_xmix(const cplx *T, const cplx *twx, int Mx, int Px, int Pz, cplx *T2)
{
    extern __shared__ cplx sh[];
    cplx *X = sh;                          // [3][Mx]
    cplx *tw = X + 3 * Mx;                 // [Px]
    double *q = (double *) (tw + Px);      // [2][Px]
    const int p = blockIdx.x;
    for (int i = threadIdx.x; i < Px; i += blockDim.x) tw[i] = twx[i];
    for (int i = threadIdx.x; i < 3 * Mx; i += blockDim.x) {
        const int f = i / Mx, k = i - f * Mx;
        X[i] = T[((size_t) f * Pz + p) * Mx + k];
    }
    __syncthreads();
    for (int n = threadIdx.x; n < Px; n += blockDim.x) {
        double v[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            double s = X[f * Mx].x;
            for (int k = 1; k < Mx; ++k) {
                const cplx t = tw[(k * n) % Px];
                const cplx a = X[f * Mx + k];
                s += 2.0 * (a.x * t.x - a.y * t.y);
            }
            v[f] = s;
        }
        q[n] = v[0] * v[2];
        q[Px + n] = v[1] * v[2];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * Mx; i += blockDim.x) {
        const int f = i / Mx, k = i - f * Mx;
        double sr = 0.0, si = 0.0;
        for (int n = 0; n < Px; ++n) {
            const cplx t = tw[(k * n) % Px];
            const double a = q[f * Px + n];
            sr += a * t.x;
            si -= a * t.y;
        }
        T2[((size_t) f * Pz + p) * Mx + k] = make_double2(sr, si);
    }
}

// grid (Mx): z forward transform of column x of both products, scaled by
// norm = 1/(Px Pz) and truncated to the stored kz rows (row Nz/2 zero), then
// the rewrite of every (z, x) != (0, 0):
//     Uxb <- i kz Uxb - i kx Uzb,   Uzb <- -i kx Uxb - i kz Uzb.
__global__ void k_wall_zfwd(const cplx *T2, const cplx *twz, const double *Kx, const double *Kz,
                            int Nz, int Mx, int Pz, double norm, cplx *Uxb, cplx *Uzb)
{
    extern __shared__ cplx sh[];
    cplx *col = sh;             // [2][Pz]
    cplx *tw = sh + 2 * Pz;     // [Pz]
    const int x = blockIdx.x;
    for (int i = threadIdx.x; i < Pz; i += blockDim.x) {
        tw[i] = twz[i];
        col[i] = T2[(size_t) i * Mx + x];
        col[Pz + i] = T2[((size_t) Pz + i) * Mx + x];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < Nz; r += blockDim.x) {
        double g[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        if (r != Nz / 2) {
            const int pr = r < Nz / 2 ? r : r + Nz / 2;
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                double sr = 0.0, si = 0.0;
                for (int p = 0; p < Pz; ++p) {
                    const cplx t = tw[(pr * p) % Pz];      // forward: the conjugate
                    const cplx v = col[f * Pz + p];
                    sr += v.x * t.x + v.y * t.y;
                    si += v.y * t.x - v.x * t.y;
                }
                g[f][0] = norm * sr;
                g[f][1] = norm * si;
            }
        }
        double ax = g[0][0], ay = g[0][1], bx = g[1][0], by = g[1][1];
        if (r != 0 || x != 0) {
            const double kx = Kx[x], kz = Kz[r];
            const double t1 = -kz * ay + kx * by;
            const double t2 = kz * ax - kx * bx;
            const double t3 = kx * ay + kz * by;
            by = -kx * ax - kz * bx;
            bx = t3;
            ax = t1;
            ay = t2;
        }
        Uxb[(size_t) r * Mx + x] = make_double2(ax, ay);
        Uzb[(size_t) r * Mx + x] = make_double2(bx, by);
    }
}

void launch_wall_plane(Context &c, const cplx *C, cplx *Uxb, cplx *Uzb)
{
    const Dims &d = c.d;
    const int n = d.Nz * d.Mx;
    hipLaunchKernelGGL(k_wall_plane, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, c.stream, C,
                       c.dKx, c.dKz, d.Nz, d.Mx, d.dimR, Uxb, Uzb);
}

void launch_incre_boundary(Context &c, cplx *Uxb, cplx *Uzb, cplx *scratch)
{
    const Dims &d = c.d;
    cplx *T = scratch;                                   // [3][Pz][Mx]
    cplx *T2 = scratch + (size_t) 3 * d.Pz * d.Mx;       // [2][Pz][Mx]
    hipLaunchKernelGGL(k_wall_zinv, dim3(d.Mx, 3), dim3(256), sizeof(cplx) * (d.Nz + d.Pz),
                       c.stream, Uxb, Uzb, c.pz.tw, d.Nz, d.Mx, d.Pz, T);
    hipLaunchKernelGGL(k_wall_xmix, dim3(d.Pz), dim3(256),
                       sizeof(cplx) * (3 * d.Mx + d.Px) + sizeof(double) * 2 * d.Px, c.stream, T,
                       c.px.tw, d.Mx, d.Px, d.Pz, T2);
    const double norm = 1.0 / ((3. * d.Nx / 2.) * (3. * d.Nz / 2.));
    hipLaunchKernelGGL(k_wall_zfwd, dim3(d.Mx), dim3(256), sizeof(cplx) * 3 * d.Pz, c.stream, T2,
                       c.pz.tw, c.dKx, c.dKz, d.Nz, d.Mx, d.Pz, norm, Uxb, Uzb);
}

}  // namespace sc
