// ddt_project on device-resident vectors without any allocation, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops).  The kernels are those of c_ddt_project; their partial sums go to
// a file-scope device array instead of the buffer c_init allocates.
#include "../../include/shadow_channel.h"
#include "sc_internal.h"

namespace sc {

// 65536 doubles (512 KB) of partial sums: two per block of the inner-product
// kernel, i.e. 2 Nz ceil(Nx/2 / ncol) with ncol the columns per block (4 up
// to dimR = 128); grids that need more are refused.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 1 << 16;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

// the array's device address per device (each GPU holds its own copy)
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES];

int noalloc_setup()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        last_error() = "no-allocation scratch: no current device below 64";
        return SC_EDEVICE;
    }
    if (g_scratch_ptr[dev]) return SC_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        last_error() = std::string("no-allocation scratch: hipGetSymbolAddress: ") +
                       hipGetErrorString(e);
        return SC_EDEVICE;
    }
    g_scratch_ptr[dev] = (double *) p;
    return SC_OK;
}

}  // namespace sc

using namespace sc;

extern "C" int sc_ddt_project_noalloc(int i_step, sc_complex *v_dev, double *r_dev)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!v_dev || !r_dev || i_step < 0 || i_step > d.nsteps || d.nsteps < 1) {
        last_error() = "ddt_project_noalloc requires device arrays v and r, "
                       "0 <= i_step <= nsteps and nsteps >= 1";
        return SC_EINVAL;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES ||
        !g_scratch_ptr[dev]) {
        last_error() = "ddt_project_noalloc: the scratch array of this device is not set up "
                       "(c_init on it first)";
        return SC_EDEVICE;
    }
    const size_t need = 2 * (size_t) ddt_gram_blocks(d);
    if (need > NOALLOC_SCRATCH_DOUBLES) {
        last_error() = "ddt_project_noalloc: this grid needs " + std::to_string(need) +
                       " doubles of scratch, more than the " +
                       std::to_string(NOALLOC_SCRATCH_DOUBLES) + " of the no-allocation entry";
        return SC_EINVAL;
    }
    double *scr = g_scratch_ptr[dev];
    // dudt = (MC[ip] - MC[im]) / dt as in c_ddt_project (forward difference
    // at step 0, backward otherwise; channel/channel.c:729-735)
    const long ip = i_step == 0 ? 3 : 3L * i_step, im = i_step == 0 ? 0 : 3L * (i_step - 1);
    const cplx *P = c.MC + c.csize * ip, *M = c.MC + c.csize * im;
    cplx *v = (cplx *) v_dev;
    // the inner-product kernel skips its work while the word it is given is
    // nonzero (a stopped run); abort_flag[3] is never written, so it always
    // runs, and no memset is queued
    const int nparts = launch_ddt_gram(c, P, M, 1.0 / d.dt, v, scr, NOALLOC_SCRATCH_DOUBLES,
                                       c.abort_flag + 3);
    if (nparts <= 0) return SC_EINVAL;
    launch_ddt_apply(c, v, v, P, M, 1.0 / d.dt, c.csize, nullptr, 0.0, nparts, 0, scr, r_dev);
    if (hipGetLastError() != hipSuccess) {
        last_error() = "ddt_project_noalloc: launch failed";
        return SC_EDEVICE;
    }
    return SC_OK;
}
