// Internal declarations of the MI355X channel-DNS library.
//
// Device layouts (complex = double2 {re, im}, kx fastest), identical to the
// reference's arrays so that C, MC and MIC can be exchanged without reshuffle
// (channel/main.h:81-200, channel/minChnl.h:46-72):
//   C  [Nz][2][dimR][Mx]     alpha (Q basis), beta (R basis)
//   U  [Nz][5][qpts][Mx]     u, v, w, du/dy, dw/dy at the Gauss points;
//                            after a nonlinear pass slots 0..2 hold Hz, Hx, Hy
//                            (adjoint: slots 3, 4 hold S_x, S_z)
//   W1 [qpts][nf][Pz][Mx]    fields after the z inverse transform
//   W2 [qpts][no][Pz][Mx]    products after the x forward transform
// with Mx = Nx/2, Pz = 3Nz/2, Px = 3Nx/2 (3/2-rule padding).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace sc {

typedef double2 cplx;

// kernel error bits in abort_flag[2]
constexpr int SC_KERR_SOLVE_WAIT = 1;   // band solve: factor staging never signalled

constexpr int MAXF = 12;        // fields in one z-inverse launch
constexpr int FFT_TILE = 2048;  // complex values (data + twiddles) in one transform block
constexpr int FFT_BLOCK = 256;  // threads of a transform block
// MFMA k-steps per staged chunk of the y-transforms (tuning constants)
#ifndef SC_KC_REBUILD
#define SC_KC_REBUILD 8
#endif
#ifndef SC_KC_FORCE
#define SC_KC_FORCE 4
#endif
// minimum waves per SIMD the y-transform kernels are compiled for
#ifndef SC_YT_WAVES
#define SC_YT_WAVES 2
#endif
constexpr int KC_REBUILD = SC_KC_REBUILD;   // coefficients -> Gauss points
constexpr int KC_FORCE = SC_KC_FORCE;       // Gauss points -> coefficients
constexpr int MAX_FACTORS = 24; // radix factors of an FFT length

struct FftPlan {
    int n = 0;
    int nf = 0;
    int f[MAX_FACTORS];
    double2 *tw = nullptr;      // device: exp(+2 pi i m / n), m < n
};

// What the z-inverse kernel loads into field f.
enum FieldKind : int {
    FK_SLOT = 0,     // copy slot `slot` of `src`
    FK_OMEGA_Z = 1,  // -du/dy + i kx v
    FK_OMEGA_X = 2,  // dw/dy - i kz v
    FK_OMEGA_Y = 3,  // i kz u - i kx w
};

struct FieldSpec {
    const cplx *src[MAXF];
    int kind[MAXF];
    int slot[MAXF];
    int nf;
    size_t ms[MAXF];    // member strides of the sources (set by launch_zinv)
};

// Pointwise stage of the x kernel.
enum XMode : int {
    XM_PRIMAL = 0,   // 6 in (u,v,w,oz,ox,oy) -> 3 out (Hz, Hx, Hy)
    XM_TANGENT = 1,  // 12 in (base, tangent) -> 3 out (IHz, IHx, IHy)
    XM_ADJOINT = 2,  // 9 in (u,v,w,oz,ox,oy,lx,ly,lz) -> 6 out (Sx,Sy,Sz,Tx,Ty,Tz)
    XM_PHYS = 3,     // 3 in -> physical flow[3][Px][qpts][Pz]
    XM_ADJ2 = 5,     // 15 in (two states' 6 fields, lambda) -> 2 x 6 out (the two
                     // adjoint product sets of one lambda, into W2 and W2b)
};

struct Dims {
    int Nx, Ny, Nz, Mx, qpts, dimR, dimQ, Px, Pz, nsteps;
    double Lx, Lz, dt, re, flux, cfl1, cfl3;
};

// One per-member allocation: nmem copies of per bytes, one after the other.
struct MemberArray {
    const char *lo = nullptr;
    size_t per = 0;
    bool trajectory = false;   // the base trajectory (MC, field cache): shared in a shared-base run
};

struct Context {
    bool ok = false;
    Dims d{};
    hipStream_t stream = nullptr;

    // Members: every per-member array holds nmem copies, member m at
    // m * (its size); the kernels take the member from their grid and each
    // pointer's member stride from mstride().  A run advances nrun members
    // (1 for the c_* entry points: member 0).  In a shared-base run the
    // members are tangent or adjoint solutions along member 0's trajectory.
    int nmem = 1;
    int nrun = 1;
    bool shared_base = false;
    std::vector<MemberArray> marrays;

    // HIP graphs of the runs' stepping loops (sc_set_graphs): one per run
    // signature, captured at its first call and replayed after; dropped when
    // the trajectory's cached fields change or the context is freed
    bool use_graphs = false;
    // mixed precision (sc_set_precision): the x/z transforms and the rebuild
    // of the time steps in FP32, everything else FP64; fixed at c_init
    bool mixed = false;
    // set while an output in physical space is formed (spec2phys): its
    // transforms run in FP64 whatever the precision of the steps
    bool fp64_out = false;
    struct GraphEntry {
        std::vector<double> key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;

    // host copies of the setup (also used for the laminar start)
    std::vector<double> y, w, cfl2, Kx, Kz;
    std::vector<double> Q, Qp, Qpp, R, Rp, Qw, Qpw, Qppw, Rw, Rpw;
    std::vector<double> Qs, Qps, Qpps, Rs, Rps, Rp0;

    // device constants
    double *dKx = nullptr, *dKz = nullptr, *dcfl2 = nullptr, *dW = nullptr;
    double *dQ = nullptr, *dQp = nullptr, *dQpp = nullptr, *dR = nullptr, *dRp = nullptr;
    double *dQw = nullptr, *dQpw = nullptr, *dQppw = nullptr, *dRw = nullptr, *dRpw = nullptr;
    double *dQs = nullptr, *dQps = nullptr, *dQpps = nullptr, *dRs = nullptr, *dRps = nullptr;
    double *dRowsum = nullptr;                  // sum_j Rw[i][j]
    // y-transform operands in the B-operand order of v_mfma_f64_16x16x4f64
    // (lane l of k-step s, tile t: row 4s + l/16, column 16t + l%16), zero-padded:
    //   fragR [nyt][nksR][5][64]: Q, Qp, Qpp, R, Rp at (j = 4s + l/16, y = 16t + l%16)
    //   fragF [nit][nksF][5][64]: Qpw, Qw, Qppw, Rw, Rpw at (y = 4s + l/16, i = 16t + l%16)
    //   fragP [nht][nksP][5][64]: the same bases on the half domain i < (q+1)/2
    //   (y = 16t + l%16), k-steps s < nksP/2 at the even rows j = 8s + 2(l/16),
    //   the others at the odd rows j = 8(s - nksP/2) + 2(l/16) + 1
    //   fragQ [npair][nksQ][2][5][64]: Qpw, Qw, Qppw, Rw, Rpw of the rows
    //   i = 2(16p + l%16) + parity at the half-domain points y = 4s + l/16
    //   (weights of the middle point halved)
    double *dFragR = nullptr, *dFragF = nullptr, *dFragP = nullptr, *dFragQ = nullptr;
    int nyt = 0, nksR = 0, nit = 0, nksF = 0, nht = 0, nksP = 0, npair = 0, nksQ = 0;
    // the Gauss points are mirror images (y[q-1-i] = -y[i]) and every basis
    // column has the parity of its index (derivatives the opposite), exactly
    bool ypar = false;
    unsigned *rb_queue = nullptr;               // rebuild task queues [8][nyt][2], zero between launches
    // LU factors of the implicit operators, per (direction, stage), as the
    // LDS images of the solve's blocks of sbt columns (col = z Mx + x):
    //   alpha: [nblk][2 parities][(dimQ+1)/2 chain rows][9][sbt]
    //   beta:  [nblk][2 parities][(dimR+1)/2 chain rows][5][sbt]
    // (j<DL: L(i,i-2(DL-j)); j=DL: 1/U(i,i); j>DL: U(i,i+2(j-DL)); chain row
    // r of parity p is matrix row p + 2r; zero outside the matrix and for the
    // columns without a chain)
    double *dLa[2][3] = {{nullptr}}, *dLb[2][3] = {{nullptr}};
    int sbt = 0;                                // columns per solve block
    bool factor_mirror = false;                 // images stored for kz rows <= Nz/2 only
    // mean mode: factors [dimR][5] and flux-correction vector [dimR]
    double *dLm[2][3] = {{nullptr}}, *dAdd[2][3] = {{nullptr}};

    // state
    cplx *C = nullptr, *IC = nullptr, *AC = nullptr;
    cplx *U = nullptr, *IU = nullptr, *AU = nullptr, *LU = nullptr, *Ut = nullptr, *Ut2 = nullptr;
    cplx *MC = nullptr, *MIC = nullptr;
    size_t mcbytes = 0;                         // bytes of one member's MC (and MIC)
    cplx *Csol = nullptr, *Ctmp = nullptr, *Ctmp2 = nullptr;
    // per member, the state one step before its window (sc_set_member_previous):
    // with use_prev, a member's time derivative at its step 0 is the backward
    // difference from it, except member 0's (the forward difference)
    cplx *Cprev = nullptr;
    bool use_prev = false;
    // per-member windows (sc_set_member_windows): member m of a member run
    // advances over steps < win[m] only (empty: every member over the whole
    // run); ACinit keeps the adjoint members' given states until their
    // windows' ends
    std::vector<int> win;
    cplx *ACinit = nullptr;
    cplx *Fa = nullptr, *Fb = nullptr;          // [Nz][dimQ][Mx], [Nz][dimR][Mx]
    cplx *FLa = nullptr, *FLb = nullptr;        // adjoint next-stage projection
    cplx *W1 = nullptr, *W2 = nullptr, *W2b = nullptr, *SP = nullptr;
    // Per stage state j = 3n + k of the stored trajectory: the six
    // velocity/vorticity fields after the z inverse transform, as written by
    // the primal's first nonlinear pass ([qpts][6][Pz][Mx] each).  The
    // tangent and adjoint read them instead of rebuilding and transforming
    // MC[j] again.  Allocated when it fits in half of the free HBM.
    cplx *wcache = nullptr;
    size_t wslot = 0;                           // complex values per state
    std::vector<char> wvalid;                   // entry j holds MC[j]'s fields
    double *flowdev = nullptr;                  // [3][Px][qpts][Pz]
    // [qpts + 1]: per-plane CFL maxima of the running check (zero between
    // checks), then the maximum of the plane that stopped a run
    double *cflmax = nullptr;
    double *dotpart = nullptr;                  // partial sums
    size_t ndotpart = 0;                        // doubles in dotpart
    double *scal = nullptr;                     // [8] device scalars
    // [0] CFL stop (1 + plane), [1] first unwritten slot, [2] kernel error
    // bits (SC_KERR_*, sticky until reported), [3] always zero (never written)
    int *abort_flag = nullptr;
    double *objscr = nullptr;                   // objectives and statistics (objective_scratch_doubles)
    double *dRwall = nullptr;                   // [2][dimR]: R_j'(-1), R_j'(+1)
    // incremental adjoint stage (sc_increadjproject): transposed bases and
    // lift values, and the staging of its host arrays; built at its first call
    double *dIadj = nullptr;
    cplx *iadjbuf = nullptr;
    cplx *p2ibuf = nullptr;     // sc_pass2_incremental: 4 inputs and 2 product sets
    double *dGwall = nullptr;                   // [dimR]: Rs^{-1} Rp0 / (Px Pz), the wall shear's representer
    FftPlan pz, px;
    size_t csize = 0, usize = 0;                // element counts
    int status = 0;                             // -1 after a CFL failure

    // host views handed out by c_getsoln / c_getincresoln, per step index
    std::vector<std::vector<cplx>> views, iviews;

    // kernel timing (benchmarks)
    int prof_class = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double prof_bytes = 0.0;        // algorithmic bytes of the recorded launches
    std::function<void()> replay;   // the last recorded launch, re-issuable
    double last_bytes = 0.0, replay_bytes = 0.0;
};

Context &ctx();

// member stride of p in bytes: the size of one copy of the per-member
// array p points into; 0 when p is in no per-member array (constants), when
// the run advances one member, or for the base trajectory of a shared-base
// run (every member reads member 0's)
size_t mstride_bytes(const Context &c, const void *p);
template <class T>
inline size_t mstride(const Context &c, const T *p)
{
    return mstride_bytes(c, (const void *) p) / sizeof(T);
}

// Launch shapes the tuner may set (sc_set_tuning; 0 = the built-in rule).
struct Tuning {
    int solve_mb = 0;        // members per band-solve block: 1, 2 or 4
    int rebuild_blocks = 0;  // blocks per XCD of the parity rebuild (at most 256)
    int force_blocks = 0;    // blocks per pair of i-tiles of the parity projection (multiple of 8)
};
Tuning &tuning();

// ---- setup (host) ----
int setup_host(Context &c);
void band_factor(double *A, int N, int dl);
// forward and back substitution of a real right-hand side with band_factor's LU
void band_subst_real(const double *A, double *b, int N, int dl);

// ---- launches ----
void launch_zinv(Context &c, const FieldSpec &fs, cplx *W1);
// Fields 0..nf1-1 come from W1 ([qpts][nf1][Pz][Mx]), the rest from W1b
// ([qpts][nf-nf1][Pz][Mx]); W1b null: all fields from W1.
void launch_xmix(Context &c, int mode, const cplx *W1, cplx *W2, int check_cfl,
                 double *flow, const cplx *W1b = nullptr, int nf1 = 0);
// cfl_slot >= 0: the first block runs the CFL check (cfl_check_block) with
// that slot: the first stored state of the step being checked
void launch_zfwd(Context &c, const cplx *W2, int nout, cplx *dst, int dst_kind, int cfl_slot = -1);
// the merged adjoint x-kernel exists for this grid (fixed transform length)
bool xmix_adj2_available(Context &c);
// both adjoint product sets of one lambda: states Wa, Wb ([qpts][6][Pz][Mx]
// each), lambda Wl ([qpts][3][Pz][Mx]) -> W2 (with Wa), W2b (with Wb)
void launch_xmix_adj2(Context &c, const cplx *Wa, const cplx *Wb, const cplx *Wl, cplx *W2,
                      cplx *W2b);
void launch_assemble_adj(Context &c, const cplx *SP, cplx *Udst);
// forward z transform of the six adjoint products fused with the assembly
// into the five state slots of Udst; false when the tile does not fit
bool launch_zfwd_adj(Context &c, const cplx *W2, cplx *Udst);
// the same for two product sets (W2 -> Udst, W2b -> Udstb) in one launch
// where the kernel allows it (two launches otherwise); false when the tile
// does not fit
bool launch_zfwd_adj2(Context &c, const cplx *W2, cplx *Udst, const cplx *W2b, cplx *Udstb);
// The transforms above dispatch on Context::mixed to one of two builds of
// fft_kernels.hip: fp64 (all FP64) and fp32 (transforms, products, W1, W2,
// W2b and the cached fields in FP32; U and the slots written stay FP64).
#define SC_FFT_LAUNCHERS                                                                        \
    void launch_zinv(Context &c, const FieldSpec &fs, cplx *W1);                                \
    void launch_xmix(Context &c, int mode, const cplx *W1, cplx *W2, int check_cfl,             \
                     double *flow, const cplx *W1b, int nf1);                                   \
    void launch_zfwd(Context &c, const cplx *W2, int nout, cplx *dst, int dst_kind,             \
                     int cfl_slot);                                                             \
    bool xmix_adj2_available(Context &c);                                                       \
    void launch_xmix_adj2(Context &c, const cplx *Wa, const cplx *Wb, const cplx *Wl, cplx *W2, \
                          cplx *W2b);                                                           \
    void launch_assemble_adj(Context &c, const cplx *SP, cplx *Udst);                           \
    bool launch_zfwd_adj(Context &c, const cplx *W2, cplx *Udst);                               \
    bool launch_zfwd_adj2(Context &c, const cplx *W2, cplx *Udst, const cplx *W2b, cplx *Udstb);
namespace fp64 {
SC_FFT_LAUNCHERS
}
namespace fp32 {
SC_FFT_LAUNCHERS
}
// vel_only: only the velocity slots u, v, w are needed (du/dy and dw/dy may
// be left unwritten where the kernel allows it)
// Whether the product slots the z-forward transforms write (and the
// Galerkin projections read) are FP32 complex: mixed-precision contexts whose
// projections run in parity form.  Same element index as the FP64 layout.
inline bool prods32(const Context &c)
{
    return c.mixed && c.ypar && (c.nksQ == 8 || c.nksQ == 16 || c.nksQ == 20 || c.nksQ == 28);
}

// Whether the Gauss-point slots the rebuild writes (and the z-inverse
// transform reads) are FP32 complex: in mixed-precision contexts whose
// rebuild runs in parity form on the FP32 matrix cores, except while an
// output in physical space is formed.  Same element index as the FP64 layout.
inline bool slots32(const Context &c)
{
    return c.mixed && !c.fp64_out && c.ypar && (c.nksP == 8 || c.nksP == 16 || c.nksP == 24);
}
void launch_rebuild(Context &c, const cplx *Csrc, cplx *Udst, bool vel_only = false);
// n complex values at p <- 0 (a kernel, so that graph capture records no memset)
void launch_zero(Context &c, cplx *p, size_t n);
// two rebuilds in one launch where the kernel allows it
void launch_rebuild2(Context &c, const cplx *C0, cplx *U0, const cplx *C1, cplx *U1,
                     bool vel_only = false);
void launch_force(Context &c, const cplx *Usrc, int adj);
void launch_force(Context &c, const cplx *Usrc, int adj, cplx *Fa, cplx *Fb);
// two projections in one launch where the kernel allows it
void launch_force2(Context &c, int adj, const cplx *U0, cplx *Fa0, cplx *Fb0, const cplx *U1,
                   cplx *Fa1, cplx *Fb1);
// target: the mean mode's flux target; members m with bit m of alt_mask set
// use target_alt instead
void launch_solve(Context &c, cplx *Ccarry, cplx *Csol, int k, int adj, const cplx *Fa2,
                  const cplx *Fb2, double target, double target_alt = 0.0,
                  unsigned alt_mask = 0, const cplx *Src = nullptr, int exact_mean = 0);
int solve_columns_per_block(int dimR);
// Re-sensitivity forcing of state S: fa [Nz][dimQ][Mx], fb [Nz][dimR][Mx], f0 [2][dimR]
void launch_tangent_forcing(Context &c, const cplx *S, cplx *fa, cplx *fb, cplx *f0);
// wall plane of C: Uxb, Uzb [Nz][Mx]
void launch_wall_plane(Context &c, const cplx *C, cplx *Uxb, cplx *Uzb);
// incremental boundary condition in place on Uxb, Uzb; scratch holds
// 5 Pz Mx complex values
void launch_incre_boundary(Context &c, cplx *Uxb, cplx *Uzb, cplx *scratch);
// partial inner products of ddt_project from the coefficients of
// dudt = s (P - M) and V; returns the number of partial pairs written to
// dotpart (or to part, of part_cap doubles), 0 if they do not fit
int launch_ddt_gram(Context &c, const cplx *P, const cplx *M, double s, const cplx *V);
// (stop: the word whose nonzero value skips the work, abort_flag by default)
int launch_ddt_gram(Context &c, const cplx *P, const cplx *M, double s, const cplx *V,
                    double *part, size_t part_cap, const int *stop, const cplx *P0 = nullptr,
                    const cplx *M0 = nullptr);
// blocks (partial pairs) of the inner-product launch for a grid
int ddt_gram_blocks(const Dims &d);
// vout = vin - (num/den) s (P - M); with acc, also acc -= a vout (twice);
// num, den, num/den into scal (c.scal unless given), partials from dotpart
// (or part)
void launch_ddt_apply(Context &c, cplx *vout, const cplx *vin, const cplx *P, const cplx *M,
                      double s, size_t n, cplx *acc, double a, int nparts, int twice);
void launch_ddt_apply(Context &c, cplx *vout, const cplx *vin, const cplx *P, const cplx *M,
                      double s, size_t n, cplx *acc, double a, int nparts, int twice,
                      const double *part, double *scal, const cplx *P0 = nullptr,
                      const cplx *M0 = nullptr);
// incremental adjoint stage k over every mode (iadj_kernels.hip): device
// arrays in the reference's layouts, the tangent states MIC[count] and
// MIC[count - 1]; iadj_prepare builds the constants and the staging buffer
// (IAC, IU, LIU, IAU, walls, forcing, IAUb, one after the other)
int iadj_prepare(Context &c);
int launch_iadj(Context &c, int k, int count, cplx *IAC, const cplx *IU, const cplx *LIU,
                const cplx *W, const cplx *F, cplx *IAU, cplx *IAUb);
// a += b over Gauss-point arrays [Nz][5][qpts][Mx], row Nz/2 of a cleared
void launch_gauss_sum(Context &c, cplx *a, const cplx *b);
// the no-allocation entry's scratch array: its device address for the
// current device, looked up once per device (0 on success)
int noalloc_setup();

// solution files in the reference's HDF5 layout (h5file.cpp); host arrays
// C, IC of [Nz][2][dimR][Nx/2]; SC_OK or SC_EIO with err set
int h5_write_solution(const std::string &path, const Dims &d, double mpg, const cplx *C,
                      const cplx *IC, std::string &err);
int h5_read_solution(const std::string &path, const Dims &d, cplx *C, cplx *IC, std::string &err);

// Reynolds-number source of a tangent stage: Src = TF(ca Y + cb X_pre)
// (k_re_source: TF the tangent forcing, X_pre the mean solve before its
// flux correction)
void launch_re_source(Context &c, const cplx *Y, const cplx *X, double ca, double cb, cplx *Src);
// the flux correction P on alpha(0,0) (the exact adjoint's step start)
void launch_mean_project(Context &c, cplx *A);
// alpha(0,0) += weight g (the wall-shear objective's representer)
void launch_add_mean(Context &c, cplx *A, const double *g, double weight);
// doubles of the objective/statistics scratch buffer for a grid and window
size_t objective_scratch_doubles(const Dims &d);

// dst_kind for launch_zfwd
enum { ZF_USLOTS = 0, ZF_SCRATCH = 1 };

// timing of one kernel class inside the timed region of a benchmark
void prof_begin(Context &c, hipStream_t s, int kernel_id, double bytes = 0.0);
// relaunch: the same launch again (kept for sc_prof_replay when kernel_id is
// the recorded class)
void prof_end(Context &c, hipStream_t s, int kernel_id,
              const std::function<void()> &relaunch = std::function<void()>());

std::string &last_error();
}  // namespace sc

// CFL check by one block (any size): the first plane whose CFL number
// reaches the limit sets abort_flag[0] = 1 + plane, abort_flag[1] = slot
// (the first stored state the stopped step would have written) and leaves
// its maximum in cflmax[qpts]; the per-plane maxima are cleared for the next
// check (channel/pass1.c CFL test)
__device__ __forceinline__ void cfl_check_block(double *cflmax, int qpts, double dt, int *abort_flag,
                                                int slot)
{
    __shared__ int first;
    if (threadIdx.x == 0) first = qpts;
    __syncthreads();
    const bool live = *abort_flag == 0;
    for (int y = threadIdx.x; y < qpts; y += blockDim.x) {
        if (live && cflmax[y] * dt >= 2.8) atomicMin(&first, y);
    }
    __syncthreads();
    const int y0 = first;
    if (y0 < qpts && threadIdx.x == 0) {
        cflmax[qpts] = cflmax[y0];
        abort_flag[1] = slot;
        abort_flag[0] = 1 + y0;
    }
    __syncthreads();
    for (int y = threadIdx.x; y < qpts; y += blockDim.x) cflmax[y] = 0.0;
}

// Workgroups are dispatched round-robin over the 8 XCDs in launch order, and
// each XCD has its own L2.  With a 1-D grid of 8*ceil(n/8) blocks, block b
// takes logical tile (b%8)*ceil(n/8) + b/8: consecutive logical tiles (tiles
// that read neighbouring data) share an XCD.  Returns -1 for the padding
// blocks.
__device__ __forceinline__ int xcd_tile(int n)
{
    const int per = (n + 7) / 8;
    const int t = (int) (blockIdx.x & 7) * per + (int) (blockIdx.x >> 3);
    return t < n ? t : -1;
}

static inline int xcd_grid(int n) { return 8 * ((n + 7) / 8); }

#define SC_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            sc::last_error() = std::string(#call) + ": " + hipGetErrorString(e_); \
            fprintf(stderr, "shadow-channel: %s\n", sc::last_error().c_str()); \
            return -10;                                                     \
        }                                                                   \
    } while (0)
