// The transform launchers the drivers call: the FP32 build of
// fft_kernels.hip for a mixed-precision context, the FP64 build otherwise.
#include "sc_internal.h"

namespace sc {

void launch_zinv(Context &c, const FieldSpec &fs, cplx *W1)
{
    c.mixed ? fp32::launch_zinv(c, fs, W1) : fp64::launch_zinv(c, fs, W1);
}

void launch_xmix(Context &c, int mode, const cplx *W1, cplx *W2, int check_cfl, double *flow,
                 const cplx *W1b, int nf1)
{
    c.mixed ? fp32::launch_xmix(c, mode, W1, W2, check_cfl, flow, W1b, nf1)
            : fp64::launch_xmix(c, mode, W1, W2, check_cfl, flow, W1b, nf1);
}

void launch_zfwd(Context &c, const cplx *W2, int nout, cplx *dst, int dst_kind, int cfl_slot)
{
    c.mixed ? fp32::launch_zfwd(c, W2, nout, dst, dst_kind, cfl_slot)
            : fp64::launch_zfwd(c, W2, nout, dst, dst_kind, cfl_slot);
}

bool xmix_adj2_available(Context &c)
{
    return c.mixed ? fp32::xmix_adj2_available(c) : fp64::xmix_adj2_available(c);
}

void launch_xmix_adj2(Context &c, const cplx *Wa, const cplx *Wb, const cplx *Wl, cplx *W2,
                      cplx *W2b)
{
    c.mixed ? fp32::launch_xmix_adj2(c, Wa, Wb, Wl, W2, W2b)
            : fp64::launch_xmix_adj2(c, Wa, Wb, Wl, W2, W2b);
}

void launch_assemble_adj(Context &c, const cplx *SP, cplx *Udst)
{
    fp64::launch_assemble_adj(c, SP, Udst);     // FP64 data in both builds
}

bool launch_zfwd_adj(Context &c, const cplx *W2, cplx *Udst)
{
    return c.mixed ? fp32::launch_zfwd_adj(c, W2, Udst) : fp64::launch_zfwd_adj(c, W2, Udst);
}

bool launch_zfwd_adj2(Context &c, const cplx *W2, cplx *Udst, const cplx *W2b, cplx *Udstb)
{
    return c.mixed ? fp32::launch_zfwd_adj2(c, W2, Udst, W2b, Udstb)
                   : fp64::launch_zfwd_adj2(c, W2, Udst, W2b, Udstb);
}

}  // namespace sc
