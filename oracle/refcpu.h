/*
 * refcpu -- CPU restatement of the shadow-channel DNS path (TEST INFRASTRUCTURE).
 *
 * This library is the parity oracle for the HIP implementation.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it; the product never links it.
 *
 * It restates, in FP64 and single-threaded, the reference's Fourier-Fourier-
 * Legendre-Galerkin channel solver: setup (channel/LegendreSetup.c,
 * channel/RsQs.c, channel/waveNumsAndCFL.c), the nonlinear terms
 * (channel/pass1.c, channel/pass2.c), the Galerkin/Helmholtz solves
 * (channel/project*.c, channel/increproject*.c, channel/adjproject*.c,
 * channel/increadjproject*.c, channel/mvOps.c), the coefficient<->Gauss-point transforms
 * (channel/initAlphaBeta*.c, channel/spec2phys.c) and the drivers
 * (channel/channel.c: init, primal, tangent, adjoint, ddt_project).
 *
 * Layouts (complex = {re, im} doubles, x fastest):
 *   C  [Nz][2][dimR][Nx/2]      alpha (v, Q basis) and beta (omega_y, R basis)
 *   U  [Nz][5][qpts][Nx/2]      u, v, w, du/dy, dw/dy at Gauss points
 *   flow [3][3Nx/2][qpts][3Nz/2] physical velocity (spec2phys)
 */
#ifndef REFCPU_H
#define REFCPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_cplx;

int    orc_init(int Nx, int Ny, int Nz, double Lx, double Lz, double Re,
                double flux, double dt, int nsteps);
void   orc_destroy(void);
/* C_given == NULL: laminar profile + rand() perturbation start. */
int    orc_primal(int ru_steps, orc_cplx *C_given);
int    orc_tangent(int start_step, int end_step, orc_cplx *IC, int inhomo);
int    orc_adjoint(int start_step, int end_step, orc_cplx *AC, int inhomo,
                   double strength);
double orc_ddt_project(int i_step, orc_cplx *IC);
/* Re-sensitivity forcing of MC[count] (channel/tangent_forcing.c,
   channel/tangent_forcing0.c): f_a [Nz][dimQ][Nx/2], f_b [Nz][dimR][Nx/2],
   f0_a, f0_b [dimR]. */
int orc_tangent_forcing(int count, orc_cplx *fa, orc_cplx *fb, orc_cplx *f0a, orc_cplx *f0b);
/* Tangent with optional sources: inhomo 1 the mean-velocity (flux) source as
   orc_tangent; re_src 1 adds the Reynolds-number source (the derivative of
   the stage operators in Re, from the forcing of tangent_forcing at the
   stage's stored states), so that the tangent from IC = 0 is dC/dRe. */
int orc_tangent_src(int start_step, int end_step, orc_cplx *IC, int inhomo, int re_src);
/* Adjoint of the time-averaged lower-wall shear objective: no trapezoid
   forcing; after each step n -> n-1, AC += weight G(tau), G the representer
   of tau(C) = sum_j R_j'(-1) Re alpha_j(0,0) in the physical inner product. */
int orc_adjoint_objective(int start_step, int end_step, orc_cplx *AC, double weight);
/* 1: orc_adjoint's mean column is the exact transpose of the tangent's (the
   flux correction as in orc_adjoint_objective); 0 (default): the
   reference's (channel/adjproject0.c:82-95, 153-167: corrected after each
   solve through the left-hand-side matrix), exact only to O(re dt). */
void   orc_set_exact_mean_adjoint(int on);
void   orc_spec2phys(const orc_cplx *C, double *flow);
const orc_cplx *orc_getsoln(int i_step);
const orc_cplx *orc_getincresoln(int i_step);
/* MIC[count], the tangent state stored at stage slot count (0 .. 3 nsteps). */
const orc_cplx *orc_getincrestage(int count);
void   orc_quad(double *y, double *w);
int    orc_qpts(void);
int    orc_dimR(void);
int    orc_dimQ(void);
int    orc_Nx(void);
int    orc_Nz(void);
int    orc_nsteps(void);
/* 0 after a clean run, -1 after a CFL violation truncated the last run. */
int    orc_last_status(void);
/* 1: also compute the terms the reference computes and discards (tangent
   side of pass1 in the primal, incremental side of pass2); 0: skip them.
   Results are identical either way; only the cost changes. */
void   orc_set_faithful_cost(int on);
/* Reset the process-global rand() stream (glibc TYPE_3 generator). */
void   orc_srand(unsigned seed);
int    orc_rand(void);

/* One Runge-Kutta stage k of the incremental adjoint over all modes
   (channel/increadjproject0.c:14-357, channel/increadjproject.c:10-645):
   IAC [Nz][2][dimR][Nx/2] in/out (the next stage's right-hand side after
   k < 2), IU / LIU [Nz][5][qpts][Nx/2] the incremental adjoint products of
   this stage and of the next (LIU unused at k = 2), W [6][Nz][Nx/2] the wall
   planes Uxb, Uzb, Uxbt, Uzbt, AUxb, AUzb, F [2][2][Nz][Nx/2][dimR] or NULL
   the forcing of stages k and k+1 on the alpha and beta equations, IAU
   [Nz][5][qpts][Nx/2] and IAUb [2][Nz][Nx/2] (IAUxb, IAUzb) out; the tangent
   states MIC[count] and MIC[count-1] of the last orc_tangent.  Requires
   1 <= count (0 at k = 2) <= 3 nsteps; -1 on bad arguments. */
int orc_increadjproject(int k, int count, orc_cplx *IAC, const orc_cplx *IU,
                        const orc_cplx *LIU, const orc_cplx *W, const orc_cplx *F,
                        orc_cplx *IAU, orc_cplx *IAUb);

/* ---- internals exposed for pinning tests ---- */
/* One adjoint stage k over all modes (channel/adjproject0.c,
   channel/adjproject.c) on caller arrays: AC in/out, U the adjoint products
   of the stage, LU those of the next (k < 2), AU out. */
int orc_adjproject_stage(int k, orc_cplx *AC, orc_cplx *U, const orc_cplx *LU, orc_cplx *AU);
/* name: "Q","Qp","Qpp","R","Rp","Qw","Qpw","Qppw","Rw","Rpw" (row-major),
   "Qs","Qps","Qpps" (dimQ x 9), "Rs","Rps" (dimR x 5), "Rp0","Rpp0",
   "Kx","Kz","cfl2","y","w".  Returns the element count, or -1. */
int    orc_get_array(const char *name, double *out, int cap);
double orc_get_scalar(const char *name);
/* One padded 2-D inverse transform of a spectral plane [Nz][Nx/2] into a
   physical plane [3Nz/2][3Nx/2] (z rows, x fastest), as pass1 does. */
void   orc_inv2d(const orc_cplx *spec, double *phys);
/* Forward transform of a physical plane, scaled and truncated to [Nz][Nx/2]
   (row Nz/2 set to zero). */
void   orc_fwd2d(const double *phys, orc_cplx *spec);
/* Solve a stride-2 banded system in place (no pivoting): A is N x (2dl+1),
   entry A[i][dl+m] is the matrix element (i, i+2m).  A is overwritten. */
void   orc_band_solve(double *A, orc_cplx *b, int N, int dl);
/* Run the state nonlinear term on a user U: U[z][0..2] <- H at all planes. */
int    orc_pass1_state(orc_cplx *U, int check_cfl);
/* Adjoint nonlinear term: U (state velocities) and AU (adjoint velocities)
   in; U[z][HZ,HX,HY] <- adjoint H, U[z][DX] <- S_x, U[z][DZ] <- S_z. */
void   orc_pass2_state(orc_cplx *U, const orc_cplx *AU);
/* The incremental side of the adjoint nonlinear term (channel/pass2.c:
   120-566): U (state), IU (tangent), AU (adjoint), IAU (incremental
   adjoint) Gauss-point arrays in; IH [Nz][5][qpts][Nx/2] <- slots 0..2 the
   incremental adjoint H, slots 3, 4 IS_x, IS_z (what pass2 leaves in IU).
   -1 on bad arguments. */
int orc_pass2_incremental(const orc_cplx *U, const orc_cplx *IU, const orc_cplx *AU,
                          const orc_cplx *IAU, orc_cplx *IH);
/* C -> U at Gauss points (initAlphaBeta2). */
void   orc_coef_to_gauss(const orc_cplx *C, orc_cplx *U);
/* Copy the current working state C / U of the solver. */
void   orc_get_state(orc_cplx *C, orc_cplx *U);
/* Wall plane of C (initAlphaBeta2's Uxb, Uzb): two [Nz][Nx/2] planes. */
void   orc_wall_plane(const orc_cplx *C, orc_cplx *Uxb, orc_cplx *Uzb);
/* increBoundary in place on two wall planes; -1 if Nz < 4 or Nx < 4. */
int    orc_incre_boundary(orc_cplx *Uxb, orc_cplx *Uzb);
/* Flux-correction trace of the mean-mode stages: with buf non-NULL, each
   later stage appends {mode (0 primal, 1 tangent, 2 adjoint), stage k,
   flux target, flux error before the correction, error after it} while
   fewer than cap_records are stored; NULL stops it.  Returns the number of
   records stored since the previous call. */
int    orc_flux_trace(double *buf, int cap_records);
/* Test hook: 1 rebuilds the adjoint velocities after each trapezoid forcing
   of orc_adjoint (the reference does not: the forcing reaches the
   convective term one stage late); 0 (default) follows the reference. */
void   orc_set_adjoint_refresh(int on);

#ifdef __cplusplus
}
#endif
#endif
