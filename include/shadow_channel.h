/*
 * shadow_channel.h -- C ABI of the MI355X channel-DNS library
 * (libshadowchannel.so).
 *
 * These are the functions the reference's SWIG module `c_channel` binds
 * (channel/channel.i), with the numpy typemaps expanded into plain pointers
 * and dimensions.  Arrays are host memory, C-contiguous, complex values laid
 * out as {double re, im} (fftw_complex, channel/mcomplex.h:19-22):
 *   C, IC, AC   [Nz][2][Ny-2][Nx/2]
 *   flow        [3][3Nx/2][qpts][3Nz/2]        (qpts = 3Ny/2)
 * One simulation per process, as in the reference (all state is global;
 * calls are not re-entrant).  Where the reference aborts on a failed
 * assert, these functions return SC_EINVAL and leave the state unchanged.
 */
#ifndef SHADOW_CHANNEL_H
#define SHADOW_CHANNEL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } sc_complex;

#define SC_OK 0
#define SC_EINIT 1        /* c_init: bad grid (EXIT_FAILURE in the reference) */
#define SC_ENOTREADY -2   /* no successful c_init */
#define SC_EINVAL -3      /* a dimension or step-range assert would fail */
#define SC_EIO -4         /* file could not be read or written */
#define SC_EDEVICE -10    /* HIP runtime error (see sc_last_error) */

/* channel/channel.i:17-23 -> init (channel/channel.c:207-391).
   Returns 0, or 1 if Nx % 4, Nz % 2 or Ny > 4 fail. */
int c_init(int Nx, int Ny, int Nz, double Lx, double Lz, double Re,
           double flux, double dt, int nstep);

/* channel/channel.i:26-31 -> destroy (channel/channel.c:52-199) */
void c_destroy(void);

/* channel/channel.i:34-53 -> primal (channel/channel.c:399-478).
   All four dims 0 (or C_given NULL): laminar profile + rand() start.
   Otherwise C_given is the initial state and receives the final state. */
int c_primal(int ru_steps, sc_complex *C_given,
             int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup);

/* channel/channel.i:56-71 -> tangent (channel/channel.c:486-587).
   Requires 0 <= start_step < end_step <= nsteps. */
int c_tangent(int start_step, int end_step, sc_complex *IC_given,
              int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup, int inhomo);

/* channel/channel.i:74-89 -> adjoint (channel/channel.c:596-696).
   Requires 0 <= end_step < start_step <= nsteps. */
int c_adjoint(int start_step, int end_step, sc_complex *AC_given,
              int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup,
              int inhomo, double strength);

/* channel/channel.i:92-106 -> ddt_project (channel/channel.c:704-780).
   Returns num/den (NaN on an invalid call). */
double c_ddt_project(int i_step, sc_complex *IC_given,
                     int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup);

/* channel/channel.i:109-143: state at step i (MC[3i] / MIC[3i]).  The
   pointer is library-owned host memory refreshed by each call for that
   step, valid until c_destroy. */
int c_getsoln(int i_step, sc_complex **MC_ptr,
              int *Nz_ptr, int *Nvar_ptr, int *Ny_ptr, int *Nx_ptr);
int c_getincresoln(int i_step, sc_complex **MIC_ptr,
                   int *Nz_ptr, int *Nvar_ptr, int *Ny_ptr, int *Nx_ptr);

/* channel/channel.i:146-169 -> spec2phys (channel/spec2phys.c:10-90) */
int c_spec2phys(sc_complex *C_ptr, int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup,
                double *flow_ptr, int Nd_flow, int Nx_flow, int Ny_flow, int Nz_flow);

/* channel/channel.i:172-203 -> read_solution / save_solution
   (channel/channel.c:791-1028).  HDF5 files in the reference's layout,
   written and read without libhdf5: /data_alpha, /data_beta, /data_ialpha,
   /data_ibeta as [dimR][Nz][Nx/2] compound {double real; double imaginary}
   and data_Nx, data_Ny (= 2 qpts / 3), data_Nz (int32), data_dt, data_Re
   (= 1/Re: the reference stores its global re), data_mpg (float64) of
   shape [1].  ialpha/ibeta hold the library's
   tangent state (the last tangent's final state), which read_solution
   also restores.  SC_EIO with sc_last_error() on an unreadable file or a
   grid other than c_init's. */
int c_read_solution(const char *filename, sc_complex *C_ptr,
                    int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup);
int c_save_solution(const char *filename, sc_complex *C_ptr,
                    int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup);

/* Host-side HDF5 access (no counterpart in the reference, which calls
   libhdf5; no device work and no c_init needed).  sc_h5_dataset_shape:
   rank of a root-group dataset (dims[0 .. rank-1] filled, kind 0 integer,
   1 floating point, 2 complex) or a negative SC_* code.  sc_h5_read_doubles:
   the dataset converted to double (complex as re, im pairs) into out of
   cap doubles; returns the count or a negative SC_* code.  Files of the
   libhdf5 family the reference writes: superblock 0 or 1 (after a
   userblock or not), symbol-table groups, version-1 object headers,
   contiguous or compact data; no compression or chunking. */
int sc_h5_dataset_shape(const char *path, const char *dataset, long long *dims, int max_rank,
                        int *kind);
long long sc_h5_read_doubles(const char *path, const char *dataset, double *out, long long cap);
/* save_solution's file from host arrays C and IC [Nz][2][Ny-2][Nx/2] of the
   grid (Nx, Ny, Nz) with dt and Re, without a c_init (same datasets as
   c_save_solution; qpts = 3 Ny / 2 as in c_init). */
int sc_h5_write_solution(const char *path, int Nx, int Ny, int Nz, double dt, double Re,
                         const sc_complex *C, const sc_complex *IC);

/* channel/channel.h:17-18 -> tangent_forcing0 / tangent_forcing
   (channel/tangent_forcing0.c:10-44, channel/tangent_forcing.c:10-75):
   Reynolds-number sensitivity forcing of the stored state MC[count]
   (0 <= count <= 3 nsteps).  f_a [Nz][dimQ][Nx/2] and f_b [Nz][dimR][Nx/2]
   for the (kz,kx) != (0,0) columns (column (0,0) set to zero); f0_a, f0_b
   [dimR] for the mean column.  The reference declares both but calls neither
   (channel/channel.c:519-522, channel/increproject0.c:46-90). */
int sc_tangent_forcing(int count, sc_complex *f_a, sc_complex *f_b, sc_complex *f0_a,
                       sc_complex *f0_b);

/* channel/initAlphaBeta2.c:70-81, 210-223 (the wall-plane part of
   initAlphaBeta2, which the reference's read_solution runs at
   channel/channel.c:868): Uxb, Uzb [Nz][Nx/2] of the coefficient array C
   [Nz][2][dimR][Nx/2].  (0,0): Uxb = Rp0.alpha, Uzb = Rp0.beta; other
   columns: Uxb = -i kz/k2 Rp0.beta, Uzb = i kx/k2 Rp0.beta, with
   Rp0_j = R_j'(-1).  The reference keeps them in globals; here the caller
   owns both planes. */
int sc_wall_plane(const sc_complex *C_ptr, sc_complex *Uxb, sc_complex *Uzb);

/* channel/increBoundary.c:20-162 (int increBoundary(void), called at
   channel/channel.c:872): in place on the wall planes Uxb, Uzb [Nz][Nx/2]:
   the dealiased products Uxb (w.n) and Uzb (w.n) with the fixed single-mode
   plane w.n (Re = 1 at kz row 1, kx 1: channel/increBoundary.c:66), then
   Uxb <- i kz Uxb - i kx Uzb and Uzb <- -i kx Uxb - i kz Uzb at every
   (z,x) != (0,0); row Nz/2 of the result is zero.  SC_EINVAL for Nx or
   Nz < 4. */
int sc_incre_boundary(sc_complex *Uxb, sc_complex *Uzb);

/* channel/pass2.c:120-566, the incremental side of int pass2(int dctr, int n)
   (the reference computes it inside pass2 with the adjoint side; the
   incremental adjoint's driver loop uses it, shan/increadjoint_test7.c:
   640-700): from the Gauss-point arrays [Nz][5][qpts][Nx/2] of the state U,
   the tangent IU, the adjoint AU and the incremental adjoint IAU (slots u, v,
   w, du/dy, dw/dy), IS = u x I lambda + Iu x lambda and IT = lambda x I omega
   + I lambda x omega on the dealiased grid, then IH <- -IT - the x,z curl of
   IS in slots 0..2 (H_z, H_x, H_y) and IS_x, IS_z in slots 3, 4: the
   products sc_increadjproject takes as IU.  Host arrays; row Nz/2 of IH is
   zero.  FP64 whatever the precision setting. */
int sc_pass2_incremental(const sc_complex *U, const sc_complex *IU, const sc_complex *AU,
                         const sc_complex *IAU, sc_complex *IH);

/* channel/increadjproject0.c:14-357 (void increadjproject0(int k, int n,
   int count, func_force_t force)) for the mean mode and
   channel/increadjproject.c:10-645 (void increadjproject(int n, int k, int z,
   int x0, int count, func_force_t force)) for every other mode, in the
   order of the reference's driver loop (shan/increadjoint_test7.c:673-685):
   one Runge-Kutta stage k (0..2) of the incremental adjoint, on host arrays
     IAC   [Nz][2][dimR][Nx/2]   in: the stage's right-hand side; out: the
                                 next stage's (k < 2) or the solution (k = 2);
                                 row Nz/2 is left as it is
     IU    [Nz][5][qpts][Nx/2]   the stage's incremental adjoint products
                                 (slots 0..2 H_z, H_x, H_y; slots 3, 4 S_x, S_z:
                                 the IU of channel/pass2.c)
     LIU   as IU                 the next stage's products (NULL allowed at k = 2)
     walls [6][Nz][Nx/2]         Uxb, Uzb, Uxbt, Uzbt, AUxb, AUzb
     force [2][2][Nz][Nx/2][dimR] or NULL: what the reference's force callback
                                 returns for stages k and k+1, the alpha
                                 equation's values (modes: its second array,
                                 the mean: its first) then the beta equation's
     IAU   [Nz][5][qpts][Nx/2]   out: u, v, w, du/dy, dw/dy at the Gauss points
     IAUb  [2][Nz][Nx/2]         out: IAUxb, IAUzb
   The tangent states MIC[count] and MIC[count-1] are those of the last
   tangent run (c_tangent, or member 0 of sc_tangent_members).  The callback
   argument n of the reference only reaches the callback and has no
   counterpart here.  SC_EINVAL unless 0 <= k <= 2, 1 <= count <= 3 nsteps
   (0 allowed at k = 2) and dimR <= 128. */
int sc_increadjproject(int k, int count, sc_complex *IAC, const sc_complex *IU,
                       const sc_complex *LIU, const sc_complex *walls, const sc_complex *force,
                       sc_complex *IAU, sc_complex *IAUb);

/* sc_increadjproject and sc_pass2_incremental on arrays in device memory
   (same layouts; no host transfer; IAC_dev in place; the output arrays must
   not overlap the inputs), for callers that chain stages on resident
   fields.  Results are bitwise those of the host-array calls. */
int sc_increadjproject_dev(int k, int count, sc_complex *IAC_dev, const sc_complex *IU_dev,
                           const sc_complex *LIU_dev, const sc_complex *walls_dev,
                           const sc_complex *force_dev, sc_complex *IAU_dev, sc_complex *IAUb_dev);
int sc_pass2_incremental_dev(const sc_complex *U_dev, const sc_complex *IU_dev,
                             const sc_complex *AU_dev, const sc_complex *IAU_dev,
                             sc_complex *IH_dev);

/* channel/__init__.py:217-299 (velAvg, IvelAvg and the profiles of
   uavgGrad) and the wall shear Uxb(0,0) of channel/project0.c:227-234, as
   device reductions over stored steps start_step .. end_step - 1 of the
   primal (incre 0, MC[3i]) or the tangent (incre 1, MIC[3i]):
     profile[y] = weight sum_i <u>(y_j, step i)   (y = 0 .. qpts-1; <u> is the
                  x-z plane mean of the streamwise velocity, Re u_00)
     shear[0]   = weight sum_i d<u>/dy at y = -1, shear[1] the same at y = +1.
   The reference forms the profile from a full spec2phys per step; here it is
   R (sum_i Re alpha_i(0,0)) weight, equal up to rounding.  Requires
   0 <= start_step <= end_step <= nsteps + 1. */
int sc_mean_profile(int incre, int start_step, int end_step, double weight, double *profile,
                    double *shear);

/* channel/comp_stat.c:9-230: plane statistics of the stored state at step
   i_step (primal MC[3i], or with incre the tangent MIC[3i]) at every Gauss
   point, stats[q][y] (q < 10, y < qpts): plane means of u, v, w, then the
   plane means of uv, uw, vw, uu, vv, ww, then of du/dy.  Formed from the
   spectra (Parseval on the dealiased grid) instead of transformed planes. */
int sc_step_statistics(int incre, int i_step, double *stats);

/* channel/channel.i:206-218: Gauss points and weights (library-owned) */
int c_quad(double **y_ptr, int *qpts_ptr1, double **w_ptr, int *qpts_ptr2);

/* channel/channel.i:221-229 */
int c_qpts(void);
int c_dimR(void);
int c_dimQ(void);
int c_Nx(void);
int c_Nz(void);
int c_nsteps(void);

/* ---- additions without a counterpart in the reference ---- */
/* Text of the last error. */
const char *sc_last_error(void);
/* 0 after a clean run; -1 when a CFL violation truncated the last primal or
   tangent run (the reference prints "Pass1 failure" and returns). */
int sc_last_status(void);
/* Reseed the process-global rand() stream used by the laminar start
   (glibc TYPE_3 generator; the reference calls the C library rand()). */
void sc_srand(unsigned seed);
/* Kernel timing for benchmarks: with kernel_class > 0, every launch of that
   class is bracketed by HIP events on its stream; sc_prof_read returns the
   number of launches and the summed milliseconds, and clears the record.
   Classes: 1 z-inverse FFT, 2 x transform/products, 3 z-forward FFT,
   4 y-transform to Gauss points, 5 Galerkin projection, 6 band solves,
   7 the merged adjoint x-kernel (two product sets of one lambda),
   8 the incremental adjoint stage's launches (sc_increadjproject). */
int sc_prof_enable(int kernel_class);
int sc_prof_read(double *total_ms);
/* Algorithmic bytes (compulsory reads + writes, each element once) of the
   launches recorded since sc_prof_enable. */
double sc_prof_bytes(void);
/* Re-issue the last recorded launch of the enabled class reps times back to
   back; returns the average milliseconds per launch (events around the whole
   batch, so launch latency overlaps the previous launch) and that launch's
   algorithmic bytes.  The replayed kernels overwrite their outputs: use after
   the results are read back. */
int sc_prof_replay(int reps, double *avg_ms, double *bytes);
/* The same replay with each launch preceded by a fill of flush_mb MB of
   scratch memory (evicting the last-level caches, as between the launches of
   a step); the fills' own time, measured alone, is subtracted.  No
   counterpart in the reference (benchmark hook). */
int sc_prof_replay_cold(int reps, double flush_mb, double *avg_ms, double *bytes);
/* Wait until all work queued by the library has finished. */
int sc_synchronize(void);
/* c_primal / c_adjoint for callers that keep the state in device memory:
   C_dev / AC_dev are device pointers to [Nz][2][dimR][Nx/2] complex arrays,
   read at the start and overwritten with the final state (device-to-device
   copies; no host transfer).  sc_primal_dev always starts from the given
   state (no laminar start). */
int sc_primal_dev(int ru_steps, sc_complex *C_dev);
int sc_adjoint_dev(int start_step, int end_step, sc_complex *AC_dev, double strength);
/* The primal keeps, per stored stage state, its velocity and vorticity after
   the z transform, so that the tangent and adjoint need not rebuild and
   transform that state again (capacity traded for recomputation; the results
   are bitwise the same).  c_init allocates it when it fits in half of the
   free device memory and it is enabled (default 1); the setting applies to
   the next c_init.  sc_field_cache_active reports whether it is allocated. */
int sc_set_field_cache(int enable);
int sc_field_cache_active(void);

/* HIP graphs (no counterpart in the reference, which launches from C loops
   on the CPU).  With sc_set_graphs(1), a stepping loop of c_primal,
   c_tangent, c_adjoint and their device and member forms is recorded as a
   HIP graph on its first call with a given signature (entry point, step
   range, member layout, sources, forcing strength, cached-field layout) and
   replayed on later calls with the same signature: the host no longer
   issues thousands of launches per window.  The CFL test still stops a
   primal run (its flag is read after the replay instead of every 64
   steps).  sc_set_graphs(0) (the default) launches kernel by kernel and
   drops the recorded graphs; kernel timing (sc_prof_enable) also bypasses
   them.  Results are identical either way. */
int sc_set_graphs(int enable);

/* Precision (no counterpart in the reference, which is FP64 throughout).
   sc_set_precision(1) makes the next c_init a mixed-precision context: the
   x/z transforms of the time steps, the nonlinear products, the arrays
   between the transforms (and the cached fields, at half the memory) and
   the y-transform to the Gauss points (FP32 matrix cores, FP32
   accumulation; where dimR <= 96) in FP32; the Galerkin projections, the
   band solves, the mean-mode flux correction, the trapezoid projection and
   every state array stay FP64.  0 (default): FP64 everywhere.  spec2phys
   is FP64 in both.  sc_precision() reports the current context's
   setting. */
int sc_set_precision(int mixed);
int sc_precision(void);

/* Launch shapes (no counterpart in the reference).  sc_set_tuning(name,
   value) sets, for all later launches of the process, one of
     "solve_members_per_block"  1, 2 or 4 members whose band solves share a
                                block and its staged factor image (0: 4, or
                                fewer where they do not fit);
     "rebuild_blocks"           blocks per XCD of the parity rebuild (<= 256);
     "force_blocks"             blocks per pair of i-tiles of the parity
                                projection (multiple of 8, <= 256);
   0 restores the built-in rule.  Results are the same bit for bit whatever
   the shapes; channel/tuning.py times candidates and keeps the fastest.
   sc_get_tuning(name) returns the current value. */
int sc_set_tuning(const char *name, int value);
int sc_get_tuning(const char *name);

/* Members (no counterpart in the reference: its drivers run one solution per
   process, drivers/wrapper.py:14-30 loops over them).  sc_set_members(n)
   makes the next c_init allocate n members (1 .. 32, default 1): n copies
   of every state array and stored trajectory.  The c_* entry points above
   act on member 0.  The *_members calls advance all n in every kernel
   launch, their arrays on the device, member m at m Nz 2 dimR Nx/2 complex
   values:
     sc_primal_members: n independent trajectories from C_dev;
     sc_tangent_members / sc_adjoint_members: with shared_base 0, member m
       along its own trajectory; with shared_base 1, every member along
       member 0's (the least-squares-shadowing segments' tangent directions);
       bit m of inhomo_mask gives member m the mean-velocity source.
   A CFL violation in any member stops all.  sc_member_state copies MC[3i]
   (incre 0) or MIC[3i] (incre 1) of one member to host memory. */
int sc_set_members(int n);
int sc_members(void);
int sc_primal_members(int ru_steps, sc_complex *C_dev);
int sc_tangent_members(int start_step, int end_step, sc_complex *IC_dev, unsigned inhomo_mask,
                       int shared_base);
int sc_adjoint_members(int start_step, int end_step, sc_complex *AC_dev, double strength,
                       int shared_base);
int sc_member_state(int incre, int member, int i_step, sc_complex *out);
/* c_ddt_project of the host array v [Nz][2][dimR][Nx/2] at step i_step of
   member `member`'s stored trajectory (the least-squares-shadowing segments
   run as members project at their own ends; at step 0 of a member with a
   previous state, sc_set_member_previous, the backward difference); returns
   num/den, NaN on an invalid call. */
double sc_member_ddt_project(int member, int i_step, sc_complex *v);

/* Members as consecutive windows of one trajectory (no counterpart in the
   reference; the least-squares-shadowing segments run as members): prev
   holds, for each member, the state one step before its window (members one
   after the other, [Nz][2][dimR][Nx/2] each; member 0's entry unused).  Then
   a member's time derivative at its step 0 (the forced adjoint's last
   forcing, sc_member_ddt_project at step 0) is the backward difference from
   that state, as c_ddt_project's at that step of the long trajectory;
   member 0 keeps the forward difference.  NULL switches it off; c_init
   clears it. */
int sc_set_member_previous(const sc_complex *prev);

/* Windows of uneven length (no counterpart in the reference, whose drivers
   run each least-squares-shadowing segment by a call of its own,
   drivers/wrapper.py:14-54): len[m] (0 .. nsteps, one per member) is the
   last step of member m's window.  sc_tangent_members then advances member
   m over steps < len[m] only: its result is its state after step len[m]
   (sc_member_state(1, m, len[m])); from there on the member is zero and
   gets no mean-velocity source, so its later stored states are zero.
   sc_adjoint_members starts member m at step min(len[m], start_step) from
   its given state, with the forcing c_adjoint applies at a run's start;
   before that the member is zero (len[m] = 0: zero throughout; a tangent
   member with len[m] <= start_step is zero from the start).  Each member's results equal a run of
   its own over its window bit for bit.  sc_primal_members is not affected.
   NULL: every member over the whole run; c_init clears it. */
int sc_set_member_windows(const int *len);

/* Sensitivities (no counterpart in the reference; the building blocks of the
   time-averaged wall-shear gradient):
   sc_tangent_sources: c_tangent with the mean-velocity source (inhomo, as
     c_tangent) and/or the Reynolds-number source (re_source): the
     derivative of the stage operators in Re from the forcing of
     channel/tangent_forcing.c and tangent_forcing0.c (declared but never
     wired into the tangent by the reference, channel/channel.c:519-522), so
     that a tangent from IC = 0 is dC/dRe.  IC: host array, in place.
   sc_adjoint_objective: the adjoint of J = weight sum_{m=end}^{start-1}
     tau(C_m), tau = sum_j R_j'(-1) Re alpha_j(0,0) the lower-wall shear
     (channel/project0.c:227-234), from AC (host, in place): the exact
     transpose of the tangent in the Gauss-weighted physical inner product
     (drivers/testAdj.py:55-56), including the mean column's flux
     correction, where the reference's c_adjoint is exact only to O(re dt).
   sc_gradient_euclid: the gradient in the real and imaginary parts of the
     coefficients, g, of d -> <lam, d>_phys (g = its derivative in Re d plus
     i times that in Im d). */
int sc_tangent_sources(int start_step, int end_step, sc_complex *IC, int inhomo, int re_source);
int sc_adjoint_objective(int start_step, int end_step, sc_complex *AC, double weight);
/* Members (no counterpart in the reference): sc_adjoint_objective of every
   member in each launch, AC_dev on the device (members one after the
   other), each member's own trajectory (shared_base 0) or member 0's; with
   member windows (sc_set_member_windows) member m starts at step len[m]
   from its given state and is zero before (len[m] = 0: zero throughout).
   sc_member_mean_profile: sc_mean_profile of member `member`'s stored
   trajectory. */
int sc_adjoint_objective_members(int start_step, int end_step, sc_complex *AC_dev, double weight,
                                 int shared_base);
int sc_member_mean_profile(int member, int incre, int start_step, int end_step, double weight,
                           double *profile, double *shear);
int sc_gradient_euclid(const sc_complex *lam, sc_complex *g);

/* No counterpart in the reference: c_ddt_project for callers that cannot
   allocate device memory (code under HIP graph capture, serving loops).
   v_dev: device array [Nz][2][dimR][Nx/2] projected in place; r_dev: device
   array of 3 doubles that receives num, den and num/den.  Queued on the
   library's stream without any allocation or host synchronisation (wait
   with sc_synchronize).  Its partial sums live in a fixed 65536-double
   device array per GPU (set up by c_init): grids needing more are refused
   with SC_EINVAL.  Calls that overlap on one GPU share that array and race. */
int sc_ddt_project_noalloc(int i_step, sc_complex *v_dev, double *r_dev);

#ifdef __cplusplus
}
#endif
#endif
