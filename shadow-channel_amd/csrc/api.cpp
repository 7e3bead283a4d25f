// C ABI of the MI355X channel-DNS library and its time-stepping drivers.
//
// The drivers follow channel/channel.c: primal (:399-478), tangent (:486-587),
// adjoint (:596-696), ddt_project (:704-780), with the stage loops over all
// Fourier modes replaced by batched kernels.  State stays on the device for
// the whole run; the host only launches.  Two deliberate changes of
// mechanism, with identical results:
//  * the tangent and the adjoint rebuild the base velocity of each stage from
//    the stored stage coefficients MC (the reference re-integrates the base
//    in the tangent; MC holds exactly the coefficients that integration
//    produces, so the rebuilt velocities are the same);
//  * the terms the reference computes and discards (the tangent side of the
//    nonlinear pass in the primal, the incremental-adjoint side of the
//    adjoint pass) are not computed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <string>

#include "../../include/shadow_channel.h"
#include "sc_internal.h"

namespace sc {

Context &ctx()
{
    static Context c;
    return c;
}

std::string &last_error()
{
    static std::string e;
    return e;
}

size_t mstride_bytes(const Context &c, const void *p)
{
    if (c.nrun <= 1 || !p) return 0;
    const char *q = (const char *) p;
    for (const MemberArray &m : c.marrays)
        if (q >= m.lo && q < m.lo + m.per * (size_t) c.nmem)
            return (m.trajectory && c.shared_base) ? 0 : m.per;
    return 0;
}

// The members a run advances: nrun members, sharing member 0's trajectory
// or not; restores the single-member setting when the run ends.
struct MemberRun {
    Context &c;
    MemberRun(Context &cc, int n, bool shared) : c(cc)
    {
        c.nrun = n;
        c.shared_base = shared;
    }
    ~MemberRun()
    {
        c.nrun = 1;
        c.shared_base = false;
    }
};

// ---------------------------------------------------------------- timing
void prof_begin(Context &c, hipStream_t s, int id, double bytes)
{
    if (c.prof_class != id) return;
    c.prof_bytes += bytes;
    c.last_bytes = bytes;
    if (c.ev_used + 2 > c.ev_pool.size()) {
        for (int i = 0; i < 256; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            c.ev_pool.push_back(e);
        }
    }
    (void) hipEventRecord(c.ev_pool[c.ev_used], s);
}

void prof_end(Context &c, hipStream_t s, int id, const std::function<void()> &relaunch)
{
    if (c.prof_class != id) return;
    if (relaunch) {
        c.replay = relaunch;
        c.replay_bytes = c.last_bytes;
    }
    if (c.ev_used + 2 > c.ev_pool.size()) return;
    (void) hipEventRecord(c.ev_pool[c.ev_used + 1], s);
    c.ev_used += 2;
}

static void drop_graphs(Context &c)
{
    for (auto &g : c.graphs) (void) hipGraphExecDestroy(g.exec);
    c.graphs.clear();
}

// The launches of a run's stepping loop (loop(capturing) queues them on the
// library's stream).  With graphs on (and no kernel timing), the loop is
// captured into a HIP graph at the first call of its signature `key` and
// the instantiated graph is replayed afterwards; the loop then must not
// synchronise while capturing.
static int stepping(Context &c, const std::vector<double> &key,
                    const std::function<int(bool)> &loop)
{
    if (!c.use_graphs || c.prof_class) return loop(false);
    for (auto &g : c.graphs)
        if (g.key == key) {
            SC_HIP(hipGraphLaunch(g.exec, c.stream));
            return SC_OK;
        }
    hipGraph_t graph = nullptr;
    SC_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    const int rc = loop(true);
    const hipError_t e = hipStreamEndCapture(c.stream, &graph);
    if (rc != SC_OK || e != hipSuccess) {
        if (graph) (void) hipGraphDestroy(graph);
        if (rc == SC_OK) last_error() = std::string("graph capture: ") + hipGetErrorString(e);
        return rc != SC_OK ? rc : SC_EDEVICE;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void) hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        last_error() = std::string("graph instantiation: ") + hipGetErrorString(ei);
        return SC_EDEVICE;
    }
    c.graphs.push_back({key, exec});
    SC_HIP(hipGraphLaunch(exec, c.stream));
    return SC_OK;
}

Tuning &tuning()
{
    static Tuning t;
    return t;
}

// whether c_init allocates the transformed-field cache (sc_set_field_cache)
static int g_field_cache = 1;
// members c_init allocates (sc_set_members)
static int g_members = 1;
// precision of the next c_init's time steps (sc_set_precision)
static int g_mixed = 0;

// ---------------------------------------------------------------- rand()
// glibc random() TYPE_3: r_i = r_{i-3} + r_{i-31} (mod 2^32) after the
// srand() seeding r_0 = seed, r_i = 16807 r_{i-1} mod (2^31 - 1) (i < 31),
// r_31..33 = r_0..2 and 310 discarded outputs; rand() = r_i >> 1.
static unsigned g_ring[31];
static bool g_seeded = false;

static int glibc_rand();

static void glibc_srand(unsigned seed)
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
    for (int i = 0; i < 31; ++i) g_ring[i] = r[i + 3];
    g_seeded = true;
    for (int i = 0; i < 310; ++i) glibc_rand();
}

static int glibc_rand()
{
    if (!g_seeded) glibc_srand(1);
    unsigned v = g_ring[0] + g_ring[28];
    memmove(g_ring, g_ring + 1, sizeof(unsigned) * 30);
    g_ring[30] = v;
    return (int) (v >> 1);
}

// ---------------------------------------------------------------- helpers
static void free_all(Context &c)
{
    void *ptrs[] = {c.dKx, c.dKz, c.dcfl2, c.dW, c.dQ, c.dQp, c.dQpp, c.dR, c.dRp, c.dQw,
                    c.dQpw, c.dQppw, c.dRw, c.dRpw, c.dQs, c.dQps, c.dQpps, c.dRs, c.dRps,
                    c.dRowsum, c.C, c.IC, c.AC, c.U, c.IU, c.AU, c.LU, c.Ut, c.Ut2, c.MC,
                    c.MIC, c.Csol, c.Ctmp, c.Ctmp2, c.Cprev, c.ACinit, c.Fa, c.Fb, c.FLa, c.FLb, c.W1, c.W2, c.W2b,
                    c.SP, c.flowdev, c.cflmax, c.dotpart, c.scal, c.abort_flag, c.pz.tw,
                    c.px.tw, c.dFragR, c.dFragF, c.dFragP, c.dFragQ, c.wcache, c.rb_queue,
                    c.objscr, c.dRwall, c.dGwall, c.dIadj, c.iadjbuf, c.p2ibuf};
    for (void *p : ptrs)
        if (p) (void) hipFree(p);
    for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 3; ++k) {
            if (c.dLa[a][k]) (void) hipFree(c.dLa[a][k]);
            if (c.dLb[a][k]) (void) hipFree(c.dLb[a][k]);
            if (c.dLm[a][k]) (void) hipFree(c.dLm[a][k]);
            if (c.dAdd[a][k]) (void) hipFree(c.dAdd[a][k]);
        }
    for (hipEvent_t e : c.ev_pool) (void) hipEventDestroy(e);
    for (auto &g : c.graphs) (void) hipGraphExecDestroy(g.exec);
    if (c.stream) (void) hipStreamDestroy(c.stream);
    const bool graphs = c.use_graphs;
    c = Context();
    c.use_graphs = graphs;
}

static int dims_ok(const Context &c, int Nz, int Nvar, int Ny, int Nx)
{
    if (Nx != c.d.Mx || Ny != c.d.dimR || Nz != c.d.Nz || Nvar != 2) {
        last_error() = "array shape must be (Nz, 2, dimR, Nx/2)";
        return 0;
    }
    return 1;
}

static cplx *mc(Context &c, long slot) { return c.MC + c.csize * slot; }
static cplx *mic(Context &c, long slot) { return c.MIC + c.csize * slot; }

static FieldSpec velocity_and_vorticity(const cplx *Ub)
{
    FieldSpec fs;
    fs.nf = 6;
    for (int s = 0; s < 3; ++s) {
        fs.src[s] = Ub;
        fs.kind[s] = FK_SLOT;
        fs.slot[s] = s;
    }
    fs.src[3] = fs.src[4] = fs.src[5] = Ub;
    fs.kind[3] = FK_OMEGA_Z;
    fs.kind[4] = FK_OMEGA_X;
    fs.kind[5] = FK_OMEGA_Y;
    fs.slot[3] = fs.slot[4] = fs.slot[5] = 0;
    return fs;
}

// cached transformed fields of stored state j, or null when not available
static cplx *wslot(Context &c, long j)
{
    if (!c.wcache || j < 0 || j >= (long) c.wvalid.size()) return nullptr;
    return c.wcache + (size_t) j * c.wslot;
}

static const cplx *wcached(Context &c, long j)
{
    cplx *p = wslot(c, j);
    return p && c.wvalid[j] ? p : nullptr;
}

// H = u x omega into the H slots of Ub (channel/pass1.c, state side).  With
// wdst the transformed fields are kept there (the cache slot of this state).
static void pass1_state(Context &c, cplx *Ub, int check_cfl, cplx *wdst = nullptr, int slot = 0)
{
    cplx *w = wdst ? wdst : c.W1;
    launch_zinv(c, velocity_and_vorticity(Ub), w);
    launch_xmix(c, XM_PRIMAL, w, c.W2, check_cfl, nullptr);
    launch_zfwd(c, c.W2, 3, Ub, ZF_USLOTS, check_cfl ? slot : -1);
}

// IH = u x I omega + I u x omega into the H slots of IUb (channel/pass1.c).
// Base fields from wbase (cached) when given, else transformed from Ub.
static void pass1_tangent(Context &c, const cplx *Ub, cplx *IUb, int check_cfl,
                          const cplx *wbase = nullptr)
{
    if (wbase) {
        launch_zinv(c, velocity_and_vorticity(IUb), c.W1);
        launch_xmix(c, XM_TANGENT, wbase, c.W2, check_cfl, nullptr, c.W1, 6);
    } else {
        FieldSpec a = velocity_and_vorticity(Ub), b = velocity_and_vorticity(IUb);
        for (int f = 0; f < 6; ++f) {
            a.src[6 + f] = b.src[f];
            a.kind[6 + f] = b.kind[f];
            a.slot[6 + f] = b.slot[f];
        }
        a.nf = 12;
        launch_zinv(c, a, c.W1);
        launch_xmix(c, XM_TANGENT, c.W1, c.W2, check_cfl, nullptr);
    }
    launch_zfwd(c, c.W2, 3, IUb, ZF_USLOTS, check_cfl ? 0 : -1);
}

// adjoint convective operator (channel/pass2.c, adjoint side) into Udst.
// State fields from wbase (cached) when given, else transformed from Ub.
// lam_ready: W1 already holds AUb's transformed fields (a previous call with
// a cached state and the same AUb).
static void pass2_adjoint(Context &c, const cplx *Ub, const cplx *AUb, cplx *Udst,
                          const cplx *wbase = nullptr, bool lam_ready = false)
{
    FieldSpec lam;
    lam.nf = 3;
    for (int s = 0; s < 3; ++s) {
        lam.src[s] = AUb;
        lam.kind[s] = FK_SLOT;
        lam.slot[s] = s;
    }
    if (wbase) {
        if (!lam_ready) launch_zinv(c, lam, c.W1);
        launch_xmix(c, XM_ADJOINT, wbase, c.W2, 0, nullptr, c.W1, 6);
    } else {
        FieldSpec a = velocity_and_vorticity(Ub);
        for (int s = 0; s < 3; ++s) {
            a.src[6 + s] = lam.src[s];
            a.kind[6 + s] = lam.kind[s];
            a.slot[6 + s] = lam.slot[s];
        }
        a.nf = 9;
        launch_zinv(c, a, c.W1);
        launch_xmix(c, XM_ADJOINT, c.W1, c.W2, 0, nullptr);
    }
    if (!launch_zfwd_adj(c, c.W2, Udst)) {
        launch_zfwd(c, c.W2, 6, c.SP, ZF_SCRATCH);
        launch_assemble_adj(c, c.SP, Udst);
    }
}

// vout = vin - (num/den) dC/dt with the Gauss-weighted physical inner
// products (channel/channel.c:704-780); device arrays, vout may be vin.
// Result in scal[2].  With acc, also acc -= a vout.
static int ddt_project_dev(Context &c, int i, const cplx *vin, cplx *vout,
                           cplx *acc = nullptr, double a = 0.0, int twice = 0)
{
    const Dims &d = c.d;
    long ip = i == 0 ? 3 : 3L * i, im = i == 0 ? 0 : 3L * (i - 1);
    // dudt = (MC[ip] - MC[im]) / dt; the inner products in closed form from
    // the coefficients (equal to the grid sums of the rebuilt fields)
    const cplx *P = mc(c, ip), *M = mc(c, im), *P0 = nullptr, *M0 = nullptr;
    if (i == 0 && c.use_prev) {
        // members whose windows continue a trajectory: the backward
        // difference from the state before the window, member 0 (the
        // trajectory's start) keeping the forward difference
        P = mc(c, 0);
        M = c.Cprev;
        P0 = mc(c, 3);
        M0 = mc(c, 0);
    }
    const int nparts = launch_ddt_gram(c, P, M, 1.0 / d.dt, vin, c.dotpart, c.ndotpart,
                                       c.abort_flag, P0, M0);
    if (nparts <= 0) return SC_EDEVICE;
    launch_ddt_apply(c, vout, vin, P, M, 1.0 / d.dt, c.csize, acc, a, nparts, twice, c.dotpart,
                     c.scal, P0, M0);
    return SC_OK;
}

// the two adjoint passes of one lambda (AUb) with the cached states wa
// (into Uda) and wb (into Udb): lambda transformed once, one x-kernel launch
static void pass2_adjoint_pair(Context &c, const cplx *AUb, const cplx *wa, const cplx *wb,
                               cplx *Uda, cplx *Udb)
{
    FieldSpec lam;
    lam.nf = 3;
    for (int s = 0; s < 3; ++s) {
        lam.src[s] = AUb;
        lam.kind[s] = FK_SLOT;
        lam.slot[s] = s;
    }
    launch_zinv(c, lam, c.W1);
    launch_xmix_adj2(c, wa, wb, c.W1, c.W2, c.W2b);
    if (!launch_zfwd_adj2(c, c.W2, Uda, c.W2b, Udb)) {
        const cplx *W2s[2] = {c.W2, c.W2b};
        cplx *Ud[2] = {Uda, Udb};
        for (int p = 0; p < 2; ++p) {
            launch_zfwd(c, W2s[p], 6, c.SP, ZF_SCRATCH);
            launch_assemble_adj(c, c.SP, Ud[p]);
        }
    }
}

// Kernel error bits (abort_flag[2]) left by the last run: reported once
// through last_error() and cleared.
static int check_kernel_error(Context &c)
{
    int e = 0;
    SC_HIP(hipMemcpy(&e, c.abort_flag + 2, sizeof(int), hipMemcpyDeviceToHost));
    if (!e) return SC_OK;
    SC_HIP(hipMemset(c.abort_flag + 2, 0, sizeof(int)));
    last_error() = "kernel error " + std::to_string(e) +
                   ((e & SC_KERR_SOLVE_WAIT) ? ": band solve factor staging never signalled" : "");
    fprintf(stderr, "shadow-channel: %s\n", last_error().c_str());
    return SC_EDEVICE;
}

static int check_abort(Context &c, const char *what)
{
    int flag = 0;
    SC_HIP(hipMemcpy(&flag, c.abort_flag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        int y = flag - 1;
        double m = 0.0;
        (void) hipMemcpy(&m, c.cflmax + c.d.qpts, sizeof(double), hipMemcpyDeviceToHost);
        printf("Pass1:  CFL[%d]:%e, Max:%e dt:%e.  CFL condition violated.\n", y, m * c.d.dt,
               2.8, c.d.dt);
        printf("%s failure\n", what);
        c.status = -1;
    }
    return flag;
}

// host restatement of the initial projection (channel/initAlphaBeta.c:24-266)
static void host_gauss_to_coef(Context &c, const std::vector<cplx> &U, std::vector<cplx> &C)
{
    const Dims &d = c.d;
    const int q = d.qpts, nQ = d.dimQ, nR = d.dimR, Mx = d.Mx;
    auto ui = [&](int z, int s, int j, int x) { return (((size_t) z * 5 + s) * q + j) * Mx + x; };
    auto ci = [&](int z, int v, int i, int x) { return (((size_t) z * 2 + v) * nR + i) * Mx + x; };
    std::vector<double> Ms(c.Rs), Mq(c.Qs);
    band_factor(Ms.data(), nR, 2);
    band_factor(Mq.data(), nQ, 4);
    auto subst = [&](const std::vector<double> &A, int N, int dl, size_t base) {
        int W = 2 * dl + 1;
        for (int i = 2; i < N; ++i)
            for (int m = std::min(dl, i / 2); m >= 1; --m) {
                double l = A[(size_t) i * W + dl - m];
                C[base + (size_t) i * Mx].x -= l * C[base + (size_t) (i - 2 * m) * Mx].x;
                C[base + (size_t) i * Mx].y -= l * C[base + (size_t) (i - 2 * m) * Mx].y;
            }
        for (int i = N - 1; i >= 0; --i) {
            int jmax = std::min(dl, (N - i + 1) / 2 - 1);
            for (int j = 1; j <= jmax; ++j) {
                double u = A[(size_t) i * W + dl + j];
                C[base + (size_t) i * Mx].x -= u * C[base + (size_t) (i + 2 * j) * Mx].x;
                C[base + (size_t) i * Mx].y -= u * C[base + (size_t) (i + 2 * j) * Mx].y;
            }
            C[base + (size_t) i * Mx].x /= A[(size_t) i * W + dl];
            C[base + (size_t) i * Mx].y /= A[(size_t) i * W + dl];
        }
    };
    for (int v = 0; v < 2; ++v) {
        int slot = v == 0 ? 0 : 2;
        for (int i = 0; i < nR; ++i) {
            double sr = 0, si = 0;
            for (int j = 0; j < q; ++j) {
                sr += c.Rw[(size_t) i * q + j] * U[ui(0, slot, j, 0)].x;
                si += c.Rw[(size_t) i * q + j] * U[ui(0, slot, j, 0)].y;
            }
            C[ci(0, v, i, 0)] = make_double2(sr, si);
        }
        subst(Ms, nR, 2, ci(0, v, 0, 0));
    }
    for (int z = 0; z < d.Nz; ++z) {
        int x0 = z == 0 ? 1 : 0;
        double kz = c.Kz[z];
        for (int x = x0; x < Mx; ++x) {
            double kx = c.Kx[x];
            for (int i = 0; i < nQ; ++i) {
                double sr = 0, si = 0;
                for (int j = 0; j < q; ++j) {
                    sr += c.Qw[(size_t) i * q + j] * U[ui(z, 1, j, x)].x;
                    si += c.Qw[(size_t) i * q + j] * U[ui(z, 1, j, x)].y;
                }
                C[ci(z, 0, i, x)] = make_double2(sr, si);
            }
            subst(Mq, nQ, 4, ci(z, 0, 0, x));
            for (int i = 0; i < nR; ++i) {
                double sr = 0, si = 0;
                for (int j = 0; j < q; ++j) {
                    cplx u = U[ui(z, 0, j, x)], w = U[ui(z, 2, j, x)];
                    sr += c.Rw[(size_t) i * q + j] * (kx * w.y - kz * u.y);
                    si += c.Rw[(size_t) i * q + j] * (kz * u.x - kx * w.x);
                }
                C[ci(z, 1, i, x)] = make_double2(sr, si);
            }
            subst(Ms, nR, 2, ci(z, 1, 0, x));
        }
    }
}

static int stream_sync(Context &c)
{
    SC_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

}  // namespace sc

using namespace sc;

extern "C" {

const char *sc_last_error(void) { return last_error().c_str(); }
int sc_last_status(void) { return ctx().status; }
void sc_srand(unsigned seed) { glibc_srand(seed); }

int sc_set_field_cache(int enable)
{
    g_field_cache = enable ? 1 : 0;
    return SC_OK;
}

int sc_field_cache_active(void) { return ctx().wcache != nullptr; }

int sc_set_precision(int mixed)
{
    if (mixed != 0 && mixed != 1) {
        last_error() = "precision: 0 (FP64) or 1 (mixed)";
        return SC_EINVAL;
    }
    g_mixed = mixed;
    return SC_OK;
}

int sc_precision(void) { return ctx().mixed ? 1 : 0; }

int sc_set_tuning(const char *name, int value)
{
    if (!name || value < 0) return SC_EINVAL;
    Tuning &t = tuning();
    const std::string n(name);
    if (n == "solve_members_per_block" && (value == 0 || value == 1 || value == 2 || value == 4))
        t.solve_mb = value;
    else if (n == "rebuild_blocks" && value <= 256)
        t.rebuild_blocks = value;
    else if (n == "force_blocks" && value % 8 == 0 && value <= 256)
        t.force_blocks = value;
    else {
        last_error() = "tuning: unknown name or value out of range";
        return SC_EINVAL;
    }
    // recorded graphs hold the old launch shapes
    drop_graphs(ctx());
    return SC_OK;
}

int sc_get_tuning(const char *name)
{
    const Tuning &t = tuning();
    const std::string n(name ? name : "");
    if (n == "solve_members_per_block") return t.solve_mb;
    if (n == "rebuild_blocks") return t.rebuild_blocks;
    if (n == "force_blocks") return t.force_blocks;
    last_error() = "tuning: unknown name";
    return SC_EINVAL;
}

int sc_prof_enable(int kernel_class)
{
    Context &c = ctx();
    c.prof_class = kernel_class;
    c.ev_used = 0;
    c.prof_bytes = 0.0;
    c.replay = nullptr;
    return 0;
}

int sc_prof_replay(int reps, double *avg_ms, double *bytes)
{
    Context &c = ctx();
    if (!c.stream) return SC_ENOTREADY;
    if (!c.replay || reps < 1 || !avg_ms || !bytes) return SC_EINVAL;
    *bytes = c.replay_bytes;
    hipEvent_t e0, e1;
    SC_HIP(hipEventCreate(&e0));
    SC_HIP(hipEventCreate(&e1));
    c.replay();   // warm the code and caches
    SC_HIP(hipEventRecord(e0, c.stream));
    for (int i = 0; i < reps; ++i) c.replay();
    SC_HIP(hipEventRecord(e1, c.stream));
    SC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return SC_OK;
}

int sc_prof_replay_cold(int reps, double flush_mb, double *avg_ms, double *bytes)
{
    Context &c = ctx();
    if (!c.stream) return SC_ENOTREADY;
    if (!c.replay || reps < 1 || flush_mb <= 0.0 || !avg_ms || !bytes) return SC_EINVAL;
    *bytes = c.replay_bytes;
    const size_t fb = (size_t) (flush_mb * 1048576.0);
    void *flush = nullptr;
    SC_HIP(hipMalloc(&flush, fb));
    hipEvent_t e[4];
    for (auto &x : e) SC_HIP(hipEventCreate(&x));
    c.replay();   // warm the code
    // the fills alone, then each fill followed by the launch: every launch
    // finds its data evicted from the caches, as inside a step
    SC_HIP(hipEventRecord(e[0], c.stream));
    for (int i = 0; i < reps; ++i) SC_HIP(hipMemsetAsync(flush, i & 0xff, fb, c.stream));
    SC_HIP(hipEventRecord(e[1], c.stream));
    SC_HIP(hipEventRecord(e[2], c.stream));
    for (int i = 0; i < reps; ++i) {
        SC_HIP(hipMemsetAsync(flush, i & 0xff, fb, c.stream));
        c.replay();
    }
    SC_HIP(hipEventRecord(e[3], c.stream));
    SC_HIP(hipEventSynchronize(e[3]));
    float m0 = 0.f, m1 = 0.f;
    SC_HIP(hipEventElapsedTime(&m0, e[0], e[1]));
    SC_HIP(hipEventElapsedTime(&m1, e[2], e[3]));
    for (auto &x : e) (void) hipEventDestroy(x);
    (void) hipFree(flush);
    *avg_ms = (m1 - m0) / reps;
    return SC_OK;
}

double sc_prof_bytes(void) { return ctx().prof_bytes; }

int sc_synchronize(void)
{
    Context &c = ctx();
    if (!c.stream) return SC_ENOTREADY;
    SC_HIP(hipStreamSynchronize(c.stream));
    return SC_OK;
}

int sc_prof_read(double *total_ms)
{
    Context &c = ctx();
    if (c.stream) (void) hipStreamSynchronize(c.stream);
    double tot = 0.0;
    int n = 0;
    for (size_t i = 0; i + 1 < c.ev_used; i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.ev_pool[i], c.ev_pool[i + 1]) == hipSuccess) {
            tot += ms;
            ++n;
        }
    }
    c.ev_used = 0;
    if (total_ms) *total_ms = tot;
    return n;
}

void c_destroy(void)
{
    Context &c = ctx();
    if (c.stream) (void) hipStreamSynchronize(c.stream);
    free_all(c);
}

int c_init(int Nx, int Ny, int Nz, double Lx, double Lz, double Re, double flux, double dt,
           int nstep)
{
    Context &c = ctx();
    if (c.ok || c.stream) c_destroy();
    printf("Nx,Ny,Nz,Lx  |  Lz,dt,nsteps  |  flux,Re\n%d %d %d %f  |  %f %f %d |  %f %f\n", Nx,
           Ny, Nz, Lx, Lz, dt, nstep, flux, Re);
    if (Nx % 4 != 0 || Nz % 2 != 0 || Ny - 4 < 0 || Nx <= 0 || Nz <= 0 || nstep < 0) {
        last_error() = "grid requires Nx % 4 == 0, Nz % 2 == 0, Ny >= 4, nsteps >= 0";
        return SC_EINIT;
    }
    c.mixed = g_mixed != 0;
    Dims &d = c.d;
    d.Nx = Nx; d.Ny = Ny; d.Nz = Nz; d.Lx = Lx; d.Lz = Lz; d.dt = dt; d.nsteps = nstep;
    d.flux = flux;
    d.re = 1. / Re;
    d.qpts = 3 * Ny / 2;
    d.dimR = Ny - 2;
    d.dimQ = Ny - 4;
    d.Mx = Nx / 2;
    d.Px = 3 * Nx / 2;
    d.Pz = 3 * Nz / 2;
    // transform tiles (data plus twiddles, FFT_TILE complex values): the x
    // kernel holds at least 6 padded rows of Px + 1 and a Px twiddle table,
    // the z kernels one column of Pz and a Pz twiddle table
    if (d.dimR > 512 || 6 * (d.Px + 1) + d.Px > FFT_TILE || 2 * d.Pz > FFT_TILE) {
        last_error() = "grid too large: need Ny <= 514, Nx <= 194 and Nz <= 682";
        return SC_EINIT;
    }
    if (hipStreamCreate(&c.stream) != hipSuccess) {
        last_error() = "no HIP device";
        return SC_EINIT;
    }
    if (setup_host(c) != 0 || noalloc_setup() != SC_OK) {
        free_all(c);
        return SC_EINIT;
    }
    c.csize = (size_t) d.Nz * 2 * d.dimR * d.Mx;
    c.usize = (size_t) d.Nz * 5 * d.qpts * d.Mx;
    const size_t cb = sizeof(cplx) * c.csize, ub = sizeof(cplx) * c.usize;
    const size_t fb = sizeof(cplx) * (size_t) d.Nz * d.dimR * d.Mx;
    const size_t mcb = cb * (3 * (size_t) nstep + 1);
    const int nm = g_members;
    c.nmem = nm;
    c.mcbytes = mcb;
    // ddt_project's partial sums: one pair per block of its inner-product
    // kernel, at most one block per (kz, kx) column, per member
    const size_t ndot = 2 * (size_t) d.Nz * d.Mx * nm;
    c.ndotpart = ndot;
    // per-member arrays (nm copies each; trajectory: the base trajectory)
    struct { void **p; size_t n; bool member, trajectory; } al[] = {
        {(void **) &c.C, cb, true, false}, {(void **) &c.IC, cb, true, false},
        {(void **) &c.AC, cb, true, false}, {(void **) &c.Csol, cb, true, false},
        {(void **) &c.Ctmp, cb, true, false}, {(void **) &c.Ctmp2, cb, true, false},
        {(void **) &c.Cprev, cb, true, false},
        {(void **) &c.ACinit, cb, true, false},
        {(void **) &c.U, ub, true, false}, {(void **) &c.IU, ub, true, false},
        {(void **) &c.AU, ub, true, false}, {(void **) &c.LU, ub, true, false},
        {(void **) &c.Ut, ub, true, false}, {(void **) &c.Ut2, ub, true, false},
        {(void **) &c.MC, mcb, true, true}, {(void **) &c.MIC, mcb, true, false},
        {(void **) &c.Fa, fb, true, false}, {(void **) &c.Fb, fb, true, false},
        {(void **) &c.FLa, fb, true, false}, {(void **) &c.FLb, fb, true, false},
        {(void **) &c.W1, sizeof(cplx) * (size_t) d.qpts * MAXF * d.Pz * d.Mx, true, false},
        {(void **) &c.W2, sizeof(cplx) * (size_t) d.qpts * 6 * d.Pz * d.Mx, true, false},
        {(void **) &c.W2b, sizeof(cplx) * (size_t) d.qpts * 6 * d.Pz * d.Mx, true, false},
        {(void **) &c.SP, sizeof(cplx) * 6 * (size_t) d.Nz * d.qpts * d.Mx, true, false},
        {(void **) &c.flowdev, sizeof(double) * 3 * (size_t) d.Px * d.qpts * d.Pz, false, false},
        {(void **) &c.cflmax, sizeof(double) * (d.qpts + 1), false, false},
        {(void **) &c.dotpart, sizeof(double) * ndot, false, false},
        {(void **) &c.scal, sizeof(double) * 8 * nm, false, false},
        {(void **) &c.abort_flag, 4 * sizeof(int), false, false},
        {(void **) &c.objscr, sizeof(double) * objective_scratch_doubles(d), false, false},
    };
    c.marrays.clear();
    for (auto &e : al) {
        const size_t n = e.member ? e.n * nm : e.n;
        if (hipMalloc(e.p, n) != hipSuccess || hipMemset(*e.p, 0, n) != hipSuccess) {
            last_error() = "device allocation failed";
            free_all(c);
            return SC_EINIT;
        }
        if (e.member) {
            MemberArray m;
            m.lo = (const char *) *e.p;
            m.per = e.n;
            m.trajectory = e.trajectory;
            c.marrays.push_back(m);
        }
    }
    {
        // wall derivatives of the R basis: R_j'(-1) = Rp0_j and, R_j having
        // the parity of j, R_j'(+1) = -(-1)^j Rp0_j
        std::vector<double> rw(2 * (size_t) d.dimR);
        for (int j = 0; j < d.dimR; ++j) {
            rw[j] = c.Rp0[j];
            rw[d.dimR + j] = (j % 2 == 0 ? -1.0 : 1.0) * c.Rp0[j];
        }
        if (hipMalloc(&c.dRwall, sizeof(double) * rw.size()) != hipSuccess ||
            hipMemcpy(c.dRwall, rw.data(), sizeof(double) * rw.size(), hipMemcpyHostToDevice) !=
                hipSuccess) {
            last_error() = "device allocation failed";
            free_all(c);
            return SC_EINIT;
        }
    }
    {
        // representer of the lower-wall shear in the physical inner product:
        // alpha(0,0) = Rs^{-1} Rp0 / (Px Pz)
        std::vector<double> M(c.Rs), g(c.Rp0);
        band_factor(M.data(), d.dimR, 2);
        for (double &v : g) v /= (double) d.Px * d.Pz;
        band_subst_real(M.data(), g.data(), d.dimR, 2);
        if (hipMalloc(&c.dGwall, sizeof(double) * d.dimR) != hipSuccess ||
            hipMemcpy(c.dGwall, g.data(), sizeof(double) * d.dimR, hipMemcpyHostToDevice) !=
                hipSuccess) {
            last_error() = "device allocation failed";
            free_all(c);
            return SC_EINIT;
        }
    }
    // transformed-field cache of the stored trajectory, if it fits in half of
    // the free memory (otherwise the tangent and adjoint recompute)
    {
        // six fields per stored stage: FP64 complex, or FP32 complex (half the
        // bytes) when the transforms run in FP32
        const size_t per = (size_t) d.qpts * 6 * d.Pz * d.Mx / (c.mixed ? 2 : 1);
        const size_t need = sizeof(cplx) * per * 3 * (size_t) nstep;   // one member
        size_t freeb = 0, totalb = 0;
        if (g_field_cache && need > 0 && hipMemGetInfo(&freeb, &totalb) == hipSuccess &&
            need * nm <= freeb / 2 &&
            hipMalloc(&c.wcache, need * nm) == hipSuccess) {
            c.wslot = per;
            c.wvalid.assign(3 * (size_t) nstep, 0);
            MemberArray m;
            m.lo = (const char *) c.wcache;
            m.per = need;
            m.trajectory = true;
            c.marrays.push_back(m);
        } else {
            c.wcache = nullptr;
            c.wvalid.clear();
        }
    }
    c.views.assign(nstep + 1, {});
    c.iviews.assign(nstep + 1, {});
    c.ok = true;
    return SC_OK;
}

}   // extern "C"

// slot `slot` of a trajectory array (MC or MIC) of every run member <- the
// member copies of a coefficient array src (one csize apart)
static int copy_to_slots(Context &c, cplx *traj, long slot, const cplx *src)
{
    const size_t cb = sizeof(cplx) * c.csize;
    cplx *dst = traj + c.csize * slot;
    if (c.nrun == 1) {
        SC_HIP(hipMemcpyAsync(dst, src, cb, hipMemcpyDeviceToDevice, c.stream));
    } else {
        SC_HIP(hipMemcpy2DAsync(dst, c.mcbytes, src, cb, cb, c.nrun, hipMemcpyDeviceToDevice,
                                c.stream));
    }
    return SC_OK;
}

// member m's copy of a member array in the current member run
static cplx *member_at(const Context &c, const cplx *p, int m)
{
    return (cplx *) ((const char *) p + (size_t) m * mstride_bytes(c, p));
}

// the member windows of sc_set_member_windows apply to this run
static bool windowed(const Context &c)
{
    return c.nrun > 1 && (int) c.win.size() == c.nrun;
}

// primal run from a given state (host or device memory, per `in`/`out`) or
// from the laminar start (C_given null); c.nrun members, C_given holding
// their states one after the other
static int primal_run(int ru_steps, sc_complex *C_given, hipMemcpyKind in, hipMemcpyKind out)
{
    Context &c = ctx();
    const Dims &d = c.d;
    const bool given = C_given != nullptr;
    const size_t cb = sizeof(cplx) * c.csize, cbm = cb * c.nrun;
    if (given) {
        SC_HIP(hipMemcpyAsync(c.C, C_given, cbm, in, c.stream));
        launch_rebuild(c, c.C, c.U);
    } else {
        // laminar Poiseuille + rand() perturbation at Gauss points; C by
        // projection, U left as set (channel/channel.c:409-422)
        std::vector<cplx> Uh(c.usize, make_double2(0.0, 0.0)), Ch(c.csize);
        SC_HIP(hipMemcpy(Ch.data(), c.C, cb, hipMemcpyDeviceToHost));
        const int q = d.qpts, Mx = d.Mx;
        for (int i = 0; i < q; ++i)
            Uh[(size_t) i * Mx] = make_double2((1.0 - c.y[i] * c.y[i]) * d.flux * 3. / 4, 0.0);
        for (int i = 0; i < q; ++i) {
            double re = glibc_rand() / (double) 2147483647 - 0.5;
            double im = glibc_rand() / (double) 2147483647 - 0.5;
            if (d.Nz > 1) Uh[(((size_t) 1 * 5 + 0) * q + i) * Mx + 1] = make_double2(re, im);
        }
        host_gauss_to_coef(c, Uh, Ch);
        SC_HIP(hipMemcpy(c.C, Ch.data(), cb, hipMemcpyHostToDevice));
        if (slots32(c)) {
            // FP32 slots, as the rebuild writes them in this context
            std::vector<float2> Uf(c.usize);
            for (size_t i = 0; i < c.usize; ++i)
                Uf[i] = make_float2((float) Uh[i].x, (float) Uh[i].y);
            SC_HIP(hipMemcpy(c.U, Uf.data(), sizeof(float2) * c.usize, hipMemcpyHostToDevice));
        } else {
            SC_HIP(hipMemcpy(c.U, Uh.data(), sizeof(cplx) * c.usize, hipMemcpyHostToDevice));
        }
    }
    // Every stored state is written by a completed run; after a CFL stop the
    // states from the stopped step on are cleared (below), so the trajectory
    // reads as the reference's zeroed array (channel/channel.c:426-432).
    if (int rc = copy_to_slots(c, c.MC, 0, c.C)) return rc;
    SC_HIP(hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream));
    // each CFL check clears the maxima it read; a run starts from zero
    SC_HIP(hipMemsetAsync(c.cflmax, 0, sizeof(double) * c.d.qpts, c.stream));
    c.status = 0;
    // the stage working on stored state j = 3(n - ru) + k writes the cached
    // fields of that state (its fields equal those of rebuild(MC[j])), except
    // for state 0 after a laminar start or a run-up (MC[0] is the state
    // before either)
    {
        std::vector<char> valid(c.wvalid.size(), 0);
        for (size_t j = 0; j < valid.size(); ++j)
            valid[j] = wslot(c, (long) j) && (j > 0 || (given && ru_steps == 0));
        if (valid != c.wvalid) drop_graphs(c);
        c.wvalid = valid;
    }
    const int total = d.nsteps + ru_steps;
    const std::vector<double> key = {0, (double) ru_steps, (double) given, (double) c.nrun};
    int rc = stepping(c, key, [&](bool capturing) {
        for (int n = 0; n < total; ++n) {
            for (int k = 0; k < 3; ++k) {
                const long j = (long) (n - ru_steps) * 3 + k;
                cplx *w = (j >= 0 && j < (long) c.wvalid.size() && c.wvalid[j]) ? wslot(c, j)
                                                                                 : nullptr;
                pass1_state(c, c.U, k == 0, w,
                            (int) std::max<long>(1, (long) (n - ru_steps) * 3 + 1));
                long count = (long) (n - ru_steps) * 3 + k + 1;
                cplx *Cs = count >= 0 ? mc(c, count) : c.Csol;
                launch_force(c, c.U, 0);
                launch_solve(c, c.C, Cs, k, 0, c.Fa, c.Fb, d.flux);
                launch_rebuild(c, Cs, c.U);
            }
            if ((n & 63) == 63 && !capturing) {
                int flag = 0;
                SC_HIP(hipMemcpy(&flag, c.abort_flag, sizeof(int), hipMemcpyDeviceToHost));
                if (flag) break;
            }
        }
        return SC_OK;
    });
    if (!rc) rc = stream_sync(c);
    if (!rc) rc = check_kernel_error(c);
    if (rc) return rc;
    // after an abort the queued launches did nothing: no cached field is
    // valid, and the states from the stopped step on were not written
    if (check_abort(c, "Pass1")) {
        std::fill(c.wvalid.begin(), c.wvalid.end(), 0);
        drop_graphs(c);
        int slot = 1;
        SC_HIP(hipMemcpy(&slot, c.abort_flag + 1, sizeof(int), hipMemcpyDeviceToHost));
        const long last = 3L * d.nsteps;
        if (slot >= 1 && slot <= last)
            for (int m = 0; m < c.nrun; ++m)
                SC_HIP(hipMemset((char *) mc(c, slot) + m * c.mcbytes, 0,
                                 cb * (size_t) (last - slot + 1)));
    }
    if (given) SC_HIP(hipMemcpy(C_given, c.C, cbm, out));
    return SC_OK;
}

extern "C" {

int c_primal(int ru_steps, sc_complex *C_given, int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    bool given = C_given && !(Nz_dup == 0 && Nvar_dup == 0 && Ny_dup == 0 && Nx_dup == 0);
    if (given && !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup)) return SC_EINVAL;
    if (ru_steps < 0) return SC_EINVAL;
    return primal_run(ru_steps, given ? C_given : nullptr, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}

int sc_primal_dev(int ru_steps, sc_complex *C_dev)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!C_dev || ru_steps < 0) return SC_EINVAL;
    return primal_run(ru_steps, C_dev, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice);
}

static int tangent_range_ok(const Context &c, int start, int end)
{
    if (!(end > start) || start < 0 || end > c.d.nsteps) {
        last_error() = "tangent requires 0 <= start_step < end_step <= nsteps";
        return 0;
    }
    return 1;
}

// tangent run of c.nrun members (IC_given: their states one after the
// other, host or device memory per in / out); member m has the mean-velocity
// source when bit m of inhomo_mask is set
static int tangent_run(int start, int end, sc_complex *IC_given, unsigned inhomo_mask,
                       hipMemcpyKind in, hipMemcpyKind out, int re_source = 0)
{
    Context &c = ctx();
    const size_t cbm = sizeof(cplx) * c.csize * c.nrun;
    SC_HIP(hipMemcpyAsync(c.IC, IC_given, cbm, in, c.stream));
    if (int rc = copy_to_slots(c, c.MIC, 3L * start, c.IC)) return rc;
    SC_HIP(hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream));
    // each CFL check clears the maxima it read; a run starts from zero
    SC_HIP(hipMemsetAsync(c.cflmax, 0, sizeof(double) * c.d.qpts, c.stream));
    // member windows: a member is zero from the end of its window on (a
    // zero tangent stays zero without the mean-velocity source), so the
    // stored states after its window are zero and its result stays at
    // MIC[3 win]
    const bool windows = windowed(c);
    if (windows)
        for (int m = 0; m < c.nrun; ++m)
            if (c.win[m] <= start) launch_zero(c, member_at(c, c.IC, m), c.csize);
    launch_rebuild(c, c.IC, c.IU);
    c.status = 0;
    const std::vector<double> key = {1, (double) start, (double) end, (double) c.nrun,
                                     (double) c.shared_base, (double) inhomo_mask,
                                     (double) re_source};
    int rc = stepping(c, key, [&](bool capturing) {
    for (int n = start; n < end; ++n) {
        unsigned mask = inhomo_mask;
        if (windows)
            for (int m = 0; m < c.nrun; ++m)
                if (c.win[m] <= n) mask &= ~(1u << m);
        for (int k = 0; k < 3; ++k) {
            const cplx *w = wcached(c, 3L * n + k);
            if (!w) launch_rebuild(c, mc(c, 3L * n + k), c.U);
            pass1_tangent(c, c.U, c.IU, k == 0, w);
            long count = 3L * n + k + 1;
            cplx *Cs = mic(c, count);
            launch_force(c, c.IU, 0);
            // the Reynolds-number source: forward RK3/CN coefficients a, b of
            // the stage (channel/project.c:44-47)
            static const double RA[3] = {1. / 3., -1. / 2, 1. / 3.}, RB[3] = {1. / 6, 2. / 3, 0.};
            if (re_source)
                launch_re_source(c, mc(c, 3L * n + k), mc(c, 3L * n + k + 1), c.d.dt * RA[k],
                                 c.d.dt * RB[k], c.Ctmp2);
            launch_solve(c, c.IC, Cs, k, 0, c.Fa, c.Fb, 0.0, 2.0, mask,
                         re_source ? c.Ctmp2 : nullptr);
            launch_rebuild(c, Cs, c.IU);
        }
        if (windows)
            for (int m = 0; m < c.nrun; ++m)
                if (c.win[m] == n + 1 && n + 1 < end) {
                    launch_zero(c, member_at(c, c.IC, m), c.csize);
                    launch_zero(c, member_at(c, c.IU, m), mstride_bytes(c, c.IU) / sizeof(cplx));
                }
        if (((n - start) & 63) == 63 && !capturing) {
            int flag = 0;
            SC_HIP(hipMemcpy(&flag, c.abort_flag, sizeof(int), hipMemcpyDeviceToHost));
            if (flag) break;
        }
    }
    return SC_OK;
    });
    if (!rc) rc = stream_sync(c);
    if (!rc) rc = check_kernel_error(c);
    if (rc) return rc;
    check_abort(c, "Pass1");
    SC_HIP(hipMemcpy(IC_given, c.IC, cbm, out));
    return SC_OK;
}

extern "C" int sc_tangent_sources(int start, int end, sc_complex *IC, int inhomo, int re_source)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!IC || !tangent_range_ok(c, start, end)) return SC_EINVAL;
    return tangent_run(start, end, IC, inhomo ? 1u : 0u, hipMemcpyHostToDevice,
                       hipMemcpyDeviceToHost, re_source != 0);
}

extern "C" int c_tangent(int start, int end, sc_complex *IC_given, int Nz_dup, int Nvar_dup,
                         int Ny_dup, int Nx_dup, int inhomo)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!IC_given || !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup)) return SC_EINVAL;
    if (!tangent_range_ok(c, start, end)) return SC_EINVAL;
    return tangent_run(start, end, IC_given, inhomo ? 1u : 0u, hipMemcpyHostToDevice,
                       hipMemcpyDeviceToHost);
}

}   // extern "C"

static int adjoint_range_ok(const Context &c, int start, int end)
{
    if (!(end < start) || end < 0 || start > c.d.nsteps) {
        last_error() = "adjoint requires 0 <= end_step < start_step <= nsteps";
        return 0;
    }
    return 1;
}

// adjoint run; AC_given in host or device memory per `in`/`out`
// the three stages of adjoint step n (channel/channel.c:633-679) on AC, AU;
// exact_mean: the mean column as the exact transpose of the tangent's
static void adjoint_stages(Context &c, int n, int exact_mean)
{
        for (int k = 0; k < 3; ++k) {
            long count = 3L * n - k - 1;
            // with both states cached, lambda is transformed once for the pair
            bool lam_ready = false, paired = false;
            if (k < 2) {
                const cplx *wa = wcached(c, count - 1), *wb = wcached(c, count);
                if (wa && wb && xmix_adj2_available(c)) {
                    // both states cached: one merged pass for the pair
                    pass2_adjoint_pair(c, c.AU, wa, wb, c.LU, c.U);
                    paired = true;
                } else {
                    if (!wa) launch_rebuild(c, mc(c, count - 1), c.U);
                    pass2_adjoint(c, c.U, c.AU, c.LU, wa);
                    lam_ready = wa != nullptr;
                }
            }
            if (!paired) {
                const cplx *w = wcached(c, count);
                if (!w) launch_rebuild(c, mc(c, count), c.U);
                pass2_adjoint(c, c.U, c.AU, c.U, w, lam_ready && w != nullptr);
            }
            if (k < 2)
                launch_force2(c, 1, c.U, c.Fa, c.Fb, c.LU, c.FLa, c.FLb);
            else
                launch_force(c, c.U, 1);
            launch_solve(c, c.AC, c.Csol, k, 1, c.FLa, c.FLb, 0.0, 0.0, 0, nullptr, exact_mean);
            launch_rebuild(c, c.Csol, c.AU, true);
        }
}

// Member m of a windowed adjoint member run starts at step n: its given
// state, lambda rebuilt from it, and the forcing c_adjoint applies at its
// start_step (acc -= a v with v the projected MIC[3n]), on that member alone
// (a < 0: the given state only).
static int start_member(Context &c, int m, int n, double a)
{
    cplx *ac = member_at(c, c.AC, m), *au = member_at(c, c.AU, m);
    const cplx *init = member_at(c, c.ACinit, m), *vin = member_at(c, mic(c, 3L * n), m);
    const cplx *P = member_at(c, mc(c, 3L * n), m), *M = member_at(c, mc(c, 3L * (n - 1)), m);
    SC_HIP(hipMemcpyAsync(ac, init, sizeof(cplx) * c.csize, hipMemcpyDeviceToDevice, c.stream));
    if (a < 0) return SC_OK;        // the objective adjoint: rebuilt at its step's start
    const int nrun = c.nrun;
    c.nrun = 1;                 // the launches below see one member at these addresses
    launch_rebuild(c, ac, au, true);
    const int nparts = launch_ddt_gram(c, P, M, 1.0 / c.d.dt, vin, c.dotpart, c.ndotpart,
                                       c.abort_flag);
    if (nparts > 0)
        launch_ddt_apply(c, c.Ctmp, vin, P, M, 1.0 / c.d.dt, c.csize, ac, a, nparts, 0,
                         c.dotpart, c.scal);
    c.nrun = nrun;
    return nparts > 0 ? SC_OK : SC_EDEVICE;
}

// Members of a windowed adjoint run whose windows end before step n are
// still waiting: held at zero (whatever forcing or source reached them is
// dropped), so that a member that never starts returns zero
static void zero_pending(Context &c, int n)
{
    for (int m = 0; m < c.nrun; ++m)
        if (c.win[m] < n) launch_zero(c, member_at(c, c.AC, m), c.csize);
}

// top of iteration n of a windowed adjoint run from `start`: the waiting
// members held at zero, the members whose windows end at n < start started
// (a >= 0: the forced adjoint's start forcing; a < 0: none, the objective
// adjoint); members whose windows reach `start` begin with the run itself
static int hold_pending(Context &c, int start, int n, double a)
{
    zero_pending(c, n);
    for (int m = 0; m < c.nrun; ++m)
        if (c.win[m] == n && n < start && start_member(c, m, n, a)) return SC_EDEVICE;
    return SC_OK;
}

static int adjoint_run(int start, int end, sc_complex *AC_given, double strength, hipMemcpyKind in,
                       hipMemcpyKind out)
{
    Context &c = ctx();
    const Dims &d = c.d;
    const size_t cbm = sizeof(cplx) * c.csize * c.nrun;
    SC_HIP(hipMemcpyAsync(c.AC, AC_given, cbm, in, c.stream));
    SC_HIP(hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream));
    // member windows: a member whose window ends before `start` waits at
    // zero and takes its given state at its window's end
    const bool windows = windowed(c);
    if (windows) {
        SC_HIP(hipMemcpyAsync(c.ACinit, c.AC, cbm, hipMemcpyDeviceToDevice, c.stream));
        zero_pending(c, start);
    }
    launch_rebuild(c, c.AC, c.AU, true);      // lambda: velocity slots only
    const double s = strength * d.dt;
    const std::vector<double> key = {2, (double) start, (double) end, (double) c.nrun,
                                     (double) c.shared_base, s};
    int rc = stepping(c, key, [&](bool) {
        for (int n = start; n > end; --n) {
            if (windows && hold_pending(c, start, n, 0.5 * s)) return SC_EDEVICE;
            // The projection of MIC[3n] at step n was made at the end of the
            // previous iteration (the reference computes it twice), which also
            // applied it for this step's start.
            if (n == start && ddt_project_dev(c, n, mic(c, 3L * n), c.Ctmp, c.AC, 0.5 * s))
                return SC_EDEVICE;
            adjoint_stages(c, n, 0);
            // forcing of MIC[3(n-1)]: end of this step and, when there is a
            // next step, its start (the two subtractions back to back)
            if (ddt_project_dev(c, n - 1, mic(c, 3L * (n - 1)), c.Ctmp, c.AC, 0.5 * s, n - 1 > end))
                return SC_EDEVICE;
        }
        if (windows) zero_pending(c, end + 1);   // never started: win <= end
        return SC_OK;
    });
    if (!rc) rc = stream_sync(c);
    if (!rc) rc = check_kernel_error(c);
    if (rc) return rc;
    SC_HIP(hipMemcpy(AC_given, c.AC, cbm, out));
    return SC_OK;
}

// The adjoint of J = weight sum_{m = end}^{start-1} tau(C_m), tau the lower
// wall shear (sum_j R_j'(-1) Re alpha_j(0,0), channel/project0.c:227-234):
// no trapezoid forcing; the mean column as the exact transpose of the
// tangent's (its flux correction P at each step's start, before the step's
// adjoint velocities); after each step the representer of tau in the
// physical inner product, weight G (alpha(0,0) = Rs^{-1} Rp0 / (Px Pz)).
static int adjoint_objective_run(int start, int end, sc_complex *AC_given, double weight,
                                 hipMemcpyKind in, hipMemcpyKind out)
{
    Context &c = ctx();
    const size_t cbm = sizeof(cplx) * c.csize * c.nrun;
    SC_HIP(hipMemcpyAsync(c.AC, AC_given, cbm, in, c.stream));
    SC_HIP(hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream));
    // member windows: a member waits at zero until its window's end
    const bool windows = windowed(c);
    if (windows) {
        SC_HIP(hipMemcpyAsync(c.ACinit, c.AC, cbm, hipMemcpyDeviceToDevice, c.stream));
        zero_pending(c, start);
    }
    for (int n = start; n > end; --n) {
        if (windows && hold_pending(c, start, n, -1.0)) return SC_EDEVICE;
        launch_mean_project(c, c.AC);
        launch_rebuild(c, c.AC, c.AU, true);
        adjoint_stages(c, n, 1);
        launch_add_mean(c, c.AC, c.dGwall, weight);
    }
    if (windows) zero_pending(c, end + 1);   // never started: win <= end
    int rc = stream_sync(c);
    if (!rc) rc = check_kernel_error(c);
    if (rc) return rc;
    SC_HIP(hipMemcpy(AC_given, c.AC, cbm, out));
    return SC_OK;
}

extern "C" {

int sc_adjoint_objective(int start, int end, sc_complex *AC, double weight)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!AC || !adjoint_range_ok(c, start, end)) return SC_EINVAL;
    return adjoint_objective_run(start, end, AC, weight, hipMemcpyHostToDevice,
                                 hipMemcpyDeviceToHost);
}

int c_adjoint(int start, int end, sc_complex *AC_given, int Nz_dup, int Nvar_dup, int Ny_dup,
              int Nx_dup, int inhomo, double strength)
{
    (void) inhomo;   // unused by the reference as well (channel/channel.c:596-696)
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!AC_given || !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup)) return SC_EINVAL;
    if (!adjoint_range_ok(c, start, end)) return SC_EINVAL;
    return adjoint_run(start, end, AC_given, strength, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}

int sc_adjoint_dev(int start, int end, sc_complex *AC_dev, double strength)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!AC_dev) return SC_EINVAL;
    if (!adjoint_range_ok(c, start, end)) return SC_EINVAL;
    return adjoint_run(start, end, AC_dev, strength, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice);
}

double c_ddt_project(int i_step, sc_complex *IC_given, int Nz_dup, int Nvar_dup, int Ny_dup,
                     int Nx_dup)
{
    Context &c = ctx();
    last_error().clear();
    if (!c.ok || !IC_given || !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup) || i_step < 0 ||
        i_step > c.d.nsteps || c.d.nsteps < 1) {
        last_error() = "ddt_project requires an initialised run, 0 <= i_step <= nsteps, "
                       "nsteps >= 1 and a (Nz, 2, dimR, Nx/2) array";
        return NAN;
    }
    const size_t cb = sizeof(cplx) * c.csize;
    if (hipMemcpyAsync(c.Ctmp, IC_given, cb, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream) != hipSuccess) {
        last_error() = "ddt_project: HIP transfer failed";
        return NAN;
    }
    if (ddt_project_dev(c, i_step, c.Ctmp, c.Ctmp) != SC_OK) return NAN;
    double scal[3] = {0, 0, 0};
    if (hipMemcpyAsync(IC_given, c.Ctmp, cb, hipMemcpyDeviceToHost, c.stream) != hipSuccess) return NAN;
    if (hipMemcpyAsync(scal, c.scal, sizeof(scal), hipMemcpyDeviceToHost, c.stream) != hipSuccess) return NAN;
    if (hipStreamSynchronize(c.stream) != hipSuccess) {
        last_error() = "ddt_project: HIP synchronisation failed";
        return NAN;
    }
    return scal[2];
}

static int get_view(int i_step, bool incre, sc_complex **ptr, int *Nz, int *Nvar, int *Ny, int *Nx)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (i_step < 0 || i_step > c.d.nsteps) {
        last_error() = "step index out of range";
        return SC_EINVAL;
    }
    auto &v = incre ? c.iviews[i_step] : c.views[i_step];
    v.resize(c.csize);
    const cplx *src = incre ? mic(c, 3L * i_step) : mc(c, 3L * i_step);
    SC_HIP(hipStreamSynchronize(c.stream));
    SC_HIP(hipMemcpy(v.data(), src, sizeof(cplx) * c.csize, hipMemcpyDeviceToHost));
    *ptr = (sc_complex *) v.data();
    *Nz = c.d.Nz;
    *Nvar = 2;
    *Ny = c.d.dimR;
    *Nx = c.d.Mx;
    return SC_OK;
}

int c_getsoln(int i_step, sc_complex **MC_ptr, int *Nz_ptr, int *Nvar_ptr, int *Ny_ptr, int *Nx_ptr)
{
    return get_view(i_step, false, MC_ptr, Nz_ptr, Nvar_ptr, Ny_ptr, Nx_ptr);
}

int c_getincresoln(int i_step, sc_complex **MIC_ptr, int *Nz_ptr, int *Nvar_ptr, int *Ny_ptr,
                   int *Nx_ptr)
{
    return get_view(i_step, true, MIC_ptr, Nz_ptr, Nvar_ptr, Ny_ptr, Nx_ptr);
}

int c_spec2phys(sc_complex *C_ptr, int Nz_dup, int Nvar_dup, int Ny_dup, int Nx_dup,
                double *flow_ptr, int Nd_flow, int Nx_flow, int Ny_flow, int Nz_flow)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!C_ptr || !flow_ptr || !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup) || Nd_flow != 3 ||
        Nx_flow != d.Px || Ny_flow != d.qpts || Nz_flow != d.Pz) {
        last_error() = "spec2phys: flow must be (3, 3Nx/2, qpts, 3Nz/2)";
        return SC_EINVAL;
    }
    SC_HIP(hipMemcpyAsync(c.Ctmp, C_ptr, sizeof(cplx) * c.csize, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream));
    c.fp64_out = true;                          // an output: FP64 transforms
    launch_rebuild(c, c.Ctmp, c.Ut, true);    // spec2phys reads u, v, w only
    c.fp64_out = false;
    FieldSpec fs;
    fs.nf = 3;
    for (int s = 0; s < 3; ++s) {
        fs.src[s] = c.Ut;
        fs.kind[s] = FK_SLOT;
        fs.slot[s] = s;
    }
    // physical fields in FP64 whatever the precision of the time steps
    fp64::launch_zinv(c, fs, c.W1);
    fp64::launch_xmix(c, XM_PHYS, c.W1, nullptr, 0, c.flowdev, nullptr, 0);
    SC_HIP(hipMemcpyAsync(flow_ptr, c.flowdev, sizeof(double) * 3 * (size_t) d.Px * d.qpts * d.Pz,
                          hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
    return SC_OK;
}

// save_solution / read_solution (channel/channel.c:791-1028) in the
// reference's HDF5 layout (h5file.cpp): C from the caller, IC from the
// library's tangent state (the reference's global IC: the last tangent's
// final state); read_solution fills the caller's C and the tangent state.
// data_mpg is 0 as in the reference's channel module, which never sets mpg.
int c_save_solution(const char *filename, sc_complex *C_ptr, int Nz_dup, int Nvar_dup, int Ny_dup,
                    int Nx_dup)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!filename || !C_ptr || !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup)) return SC_EINVAL;
    std::vector<cplx> IC(c.csize);
    SC_HIP(hipMemcpy(IC.data(), c.IC, sizeof(cplx) * c.csize, hipMemcpyDeviceToHost));
    std::string err;
    const int rc = h5_write_solution(filename, c.d, 0.0, (const cplx *) C_ptr, IC.data(), err);
    if (rc != SC_OK) last_error() = err;
    return rc;
}

int c_read_solution(const char *filename, sc_complex *C_ptr, int Nz_dup, int Nvar_dup, int Ny_dup,
                    int Nx_dup)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!filename || !C_ptr || !dims_ok(c, Nz_dup, Nvar_dup, Ny_dup, Nx_dup)) return SC_EINVAL;
    std::vector<cplx> C(c.csize), IC(c.csize);
    std::string err;
    if (h5_read_solution(filename, c.d, C.data(), IC.data(), err) != SC_OK) {
        last_error() = err;
        return SC_EIO;
    }
    memcpy(C_ptr, C.data(), sizeof(cplx) * c.csize);
    SC_HIP(hipMemcpy(c.C, C.data(), sizeof(cplx) * c.csize, hipMemcpyHostToDevice));
    SC_HIP(hipMemcpy(c.IC, IC.data(), sizeof(cplx) * c.csize, hipMemcpyHostToDevice));
    return SC_OK;
}

int c_quad(double **y_ptr, int *qpts_ptr1, double **w_ptr, int *qpts_ptr2)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    *y_ptr = c.y.data();
    *w_ptr = c.w.data();
    *qpts_ptr1 = *qpts_ptr2 = c.d.qpts;
    return SC_OK;
}

int sc_tangent_forcing(int count, sc_complex *f_a, sc_complex *f_b, sc_complex *f0_a,
                       sc_complex *f0_b)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!f_a || !f_b || !f0_a || !f0_b || count < 0 || count > 3 * d.nsteps) {
        last_error() = "tangent_forcing requires 0 <= count <= 3 nsteps and four output arrays";
        return SC_EINVAL;
    }
    // outputs staged in the projection buffers (same shapes) and Ctmp2
    launch_tangent_forcing(c, mc(c, count), c.Fa, c.Fb, c.Ctmp2);
    int rc = stream_sync(c);
    if (rc) return rc;
    const size_t na = (size_t) d.Nz * d.dimQ * d.Mx, nb = (size_t) d.Nz * d.dimR * d.Mx;
    SC_HIP(hipMemcpy(f_a, c.Fa, sizeof(cplx) * na, hipMemcpyDeviceToHost));
    SC_HIP(hipMemcpy(f_b, c.Fb, sizeof(cplx) * nb, hipMemcpyDeviceToHost));
    SC_HIP(hipMemcpy(f0_a, c.Ctmp2, sizeof(cplx) * d.dimR, hipMemcpyDeviceToHost));
    SC_HIP(hipMemcpy(f0_b, c.Ctmp2 + d.dimR, sizeof(cplx) * d.dimR, hipMemcpyDeviceToHost));
    return SC_OK;
}

int sc_wall_plane(const sc_complex *C_ptr, sc_complex *Uxb, sc_complex *Uzb)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!C_ptr || !Uxb || !Uzb) {
        last_error() = "wall_plane requires C and two output planes";
        return SC_EINVAL;
    }
    const Dims &d = c.d;
    const size_t np = (size_t) d.Nz * d.Mx;
    SC_HIP(hipMemcpyAsync(c.Ctmp, C_ptr, sizeof(cplx) * c.csize, hipMemcpyHostToDevice, c.stream));
    launch_wall_plane(c, c.Ctmp, c.Ctmp2, c.Ctmp2 + np);
    SC_HIP(hipMemcpyAsync(Uxb, c.Ctmp2, sizeof(cplx) * np, hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipMemcpyAsync(Uzb, c.Ctmp2 + np, sizeof(cplx) * np, hipMemcpyDeviceToHost, c.stream));
    return stream_sync(c);
}

int sc_incre_boundary(sc_complex *Uxb, sc_complex *Uzb)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (!Uxb || !Uzb || d.Nz < 4 || d.Mx < 2) {
        last_error() = "incre_boundary requires two wall planes and Nx, Nz >= 4";
        return SC_EINVAL;
    }
    const size_t np = (size_t) d.Nz * d.Mx;
    // planes in Ctmp2 (2 Nz Mx <= one coefficient array), transforms in W1
    SC_HIP(hipMemcpyAsync(c.Ctmp2, Uxb, sizeof(cplx) * np, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemcpyAsync(c.Ctmp2 + np, Uzb, sizeof(cplx) * np, hipMemcpyHostToDevice, c.stream));
    launch_incre_boundary(c, c.Ctmp2, c.Ctmp2 + np, c.W1);
    SC_HIP(hipMemcpyAsync(Uxb, c.Ctmp2, sizeof(cplx) * np, hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipMemcpyAsync(Uzb, c.Ctmp2 + np, sizeof(cplx) * np, hipMemcpyDeviceToHost, c.stream));
    return stream_sync(c);
}

int sc_pass2_incremental(const sc_complex *U, const sc_complex *IU, const sc_complex *AU,
                         const sc_complex *IAU, sc_complex *IH)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!U || !IU || !AU || !IAU || !IH) {
        last_error() = "pass2_incremental requires U, IU, AU, IAU and IH";
        return SC_EINVAL;
    }
    const size_t us = c.usize;
    if (!c.p2ibuf && hipMalloc(&c.p2ibuf, sizeof(cplx) * 6 * us) != hipSuccess) {
        c.p2ibuf = nullptr;
        last_error() = "pass2_incremental: device allocation failed";
        return SC_EINIT;
    }
    cplx *bU = c.p2ibuf, *bIU = bU + us, *bAU = bIU + us, *bIAU = bAU + us, *T1 = bIAU + us,
         *T2 = T1 + us;
    SC_HIP(hipMemcpyAsync(bU, U, sizeof(cplx) * us, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemcpyAsync(bIU, IU, sizeof(cplx) * us, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemcpyAsync(bAU, AU, sizeof(cplx) * us, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemcpyAsync(bIAU, IAU, sizeof(cplx) * us, hipMemcpyHostToDevice, c.stream));
    // bilinear in (state, lambda): the adjoint operator of (u, I lambda)
    // plus that of (Iu, lambda), in FP64 whatever the context's precision
    const bool mixed = c.mixed;
    c.mixed = false;
    pass2_adjoint(c, bU, bIAU, T1);
    pass2_adjoint(c, bIU, bAU, T2);
    launch_gauss_sum(c, T1, T2);
    c.mixed = mixed;
    SC_HIP(hipMemcpyAsync(IH, T1, sizeof(cplx) * us, hipMemcpyDeviceToHost, c.stream));
    return stream_sync(c);
}

int sc_increadjproject(int k, int count, sc_complex *IAC, const sc_complex *IU,
                       const sc_complex *LIU, const sc_complex *walls, const sc_complex *force,
                       sc_complex *IAU, sc_complex *IAUb)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (k < 0 || k > 2 || count < (k < 2 ? 1 : 0) || count > 3 * d.nsteps || !IAC || !IU ||
        (k < 2 && !LIU) || !walls || !IAU || !IAUb) {
        last_error() = "increadjproject requires 0 <= k <= 2, 1 <= count <= 3 nsteps (0 allowed "
                       "at k = 2), and IAC, IU, LIU (k < 2), walls, IAU, IAUb";
        return SC_EINVAL;
    }
    if (d.dimR > 128) {
        last_error() = "increadjproject supports dimR <= 128 (Ny <= 130)";
        return SC_EINVAL;
    }
    if (iadj_prepare(c)) return SC_EINIT;
    const size_t np = (size_t) d.Nz * d.Mx, nf = 4 * np * d.dimR;
    cplx *bIAC = c.iadjbuf, *bIU = bIAC + c.csize, *bLIU = bIU + c.usize, *bIAU = bLIU + c.usize;
    cplx *bW = bIAU + c.usize, *bF = bW + 6 * np, *bIAUb = bF + nf;
    SC_HIP(hipMemcpyAsync(bIAC, IAC, sizeof(cplx) * c.csize, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemcpyAsync(bIU, IU, sizeof(cplx) * c.usize, hipMemcpyHostToDevice, c.stream));
    if (k < 2)
        SC_HIP(hipMemcpyAsync(bLIU, LIU, sizeof(cplx) * c.usize, hipMemcpyHostToDevice, c.stream));
    SC_HIP(hipMemcpyAsync(bW, walls, sizeof(cplx) * 6 * np, hipMemcpyHostToDevice, c.stream));
    if (force)
        SC_HIP(hipMemcpyAsync(bF, force, sizeof(cplx) * nf, hipMemcpyHostToDevice, c.stream));
    if (launch_iadj(c, k, count, bIAC, bIU, bLIU, bW, force ? bF : nullptr, bIAU, bIAUb)) {
        last_error() = "increadjproject: launch failed";
        return SC_EINIT;
    }
    SC_HIP(hipMemcpyAsync(IAC, bIAC, sizeof(cplx) * c.csize, hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipMemcpyAsync(IAU, bIAU, sizeof(cplx) * c.usize, hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipMemcpyAsync(IAUb, bIAUb, sizeof(cplx) * 2 * np, hipMemcpyDeviceToHost, c.stream));
    return stream_sync(c);
}

int sc_increadjproject_dev(int k, int count, sc_complex *IAC_dev, const sc_complex *IU_dev,
                           const sc_complex *LIU_dev, const sc_complex *walls_dev,
                           const sc_complex *force_dev, sc_complex *IAU_dev, sc_complex *IAUb_dev)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    const Dims &d = c.d;
    if (k < 0 || k > 2 || count < (k < 2 ? 1 : 0) || count > 3 * d.nsteps || !IAC_dev ||
        !IU_dev || (k < 2 && !LIU_dev) || !walls_dev || !IAU_dev || !IAUb_dev) {
        last_error() = "increadjproject_dev requires 0 <= k <= 2, 1 <= count <= 3 nsteps (0 "
                       "allowed at k = 2), and IAC, IU, LIU (k < 2), walls, IAU, IAUb";
        return SC_EINVAL;
    }
    if (d.dimR > 128) {
        last_error() = "increadjproject supports dimR <= 128 (Ny <= 130)";
        return SC_EINVAL;
    }
    if (iadj_prepare(c)) return SC_EINIT;
    if (launch_iadj(c, k, count, (cplx *) IAC_dev, (const cplx *) IU_dev, (const cplx *) LIU_dev,
                    (const cplx *) walls_dev, (const cplx *) force_dev, (cplx *) IAU_dev,
                    (cplx *) IAUb_dev)) {
        last_error() = "increadjproject: launch failed";
        return SC_EINIT;
    }
    return stream_sync(c);
}

int sc_pass2_incremental_dev(const sc_complex *U_dev, const sc_complex *IU_dev,
                             const sc_complex *AU_dev, const sc_complex *IAU_dev, sc_complex *IH_dev)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!U_dev || !IU_dev || !AU_dev || !IAU_dev || !IH_dev) {
        last_error() = "pass2_incremental_dev requires U, IU, AU, IAU and IH";
        return SC_EINVAL;
    }
    const size_t us = c.usize;
    if (!c.p2ibuf && hipMalloc(&c.p2ibuf, sizeof(cplx) * 6 * us) != hipSuccess) {
        c.p2ibuf = nullptr;
        last_error() = "pass2_incremental: device allocation failed";
        return SC_EINIT;
    }
    cplx *T2 = c.p2ibuf + 5 * us;
    const bool mixed = c.mixed;
    c.mixed = false;
    pass2_adjoint(c, (const cplx *) U_dev, (const cplx *) IAU_dev, (cplx *) IH_dev);
    pass2_adjoint(c, (const cplx *) IU_dev, (const cplx *) AU_dev, T2);
    launch_gauss_sum(c, (cplx *) IH_dev, T2);
    c.mixed = mixed;
    return stream_sync(c);
}

int c_qpts(void) { return ctx().d.qpts; }
int c_dimR(void) { return ctx().d.dimR; }
int c_dimQ(void) { return ctx().d.dimQ; }
int c_Nx(void) { return ctx().d.Nx; }
int c_Nz(void) { return ctx().d.Nz; }
int c_nsteps(void) { return ctx().d.nsteps; }

}  // extern "C"

// ---------------------------------------------------------------- members
extern "C" {

int sc_set_graphs(int on)
{
    Context &c = ctx();
    c.use_graphs = on != 0;
    if (!c.use_graphs) drop_graphs(c);
    return SC_OK;
}

int sc_set_members(int n)
{
    if (n < 1 || n > 32) {
        last_error() = "members: 1 .. 32";
        return SC_EINVAL;
    }
    g_members = n;
    return SC_OK;
}

int sc_members(void) { return ctx().ok ? ctx().nmem : g_members; }

int sc_primal_members(int ru_steps, sc_complex *C_dev)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!C_dev || ru_steps < 0) return SC_EINVAL;
    MemberRun run(c, c.nmem, false);
    return primal_run(ru_steps, C_dev, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice);
}

int sc_tangent_members(int start_step, int end_step, sc_complex *IC_dev, unsigned inhomo_mask,
                       int shared_base)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!IC_dev || !tangent_range_ok(c, start_step, end_step)) return SC_EINVAL;
    MemberRun run(c, c.nmem, shared_base != 0);
    return tangent_run(start_step, end_step, IC_dev, inhomo_mask, hipMemcpyDeviceToDevice,
                       hipMemcpyDeviceToDevice);
}

int sc_adjoint_objective_members(int start_step, int end_step, sc_complex *AC_dev,
                                 double weight, int shared_base)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!AC_dev || !adjoint_range_ok(c, start_step, end_step)) return SC_EINVAL;
    MemberRun run(c, c.nmem, shared_base != 0);
    return adjoint_objective_run(start_step, end_step, AC_dev, weight, hipMemcpyDeviceToDevice,
                                 hipMemcpyDeviceToDevice);
}

int sc_adjoint_members(int start_step, int end_step, sc_complex *AC_dev, double strength,
                       int shared_base)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!AC_dev || !adjoint_range_ok(c, start_step, end_step)) return SC_EINVAL;
    MemberRun run(c, c.nmem, shared_base != 0);
    return adjoint_run(start_step, end_step, AC_dev, strength, hipMemcpyDeviceToDevice,
                       hipMemcpyDeviceToDevice);
}

int sc_member_state(int incre, int member, int i_step, sc_complex *out)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!out || member < 0 || member >= c.nmem || i_step < 0 || i_step > c.d.nsteps) {
        last_error() = "member_state requires 0 <= member < members and 0 <= i_step <= nsteps";
        return SC_EINVAL;
    }
    const char *src = (const char *) ((incre ? c.MIC : c.MC) + 3 * c.csize * (size_t) i_step) +
                      (size_t) member * c.mcbytes;
    SC_HIP(hipStreamSynchronize(c.stream));
    SC_HIP(hipMemcpy(out, src, sizeof(cplx) * c.csize, hipMemcpyDeviceToHost));
    return SC_OK;
}

// The state one step before each member's window (host, members one after
// the other; member 0's entry is not used): afterwards the members' time
// derivative at their step 0 (the adjoint's last forcing, sc_member_ddt_project
// at step 0) is the backward difference from it, as on one long trajectory
// cut into the members' windows.  NULL: the forward difference again.
int sc_set_member_previous(const sc_complex *prev)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!prev) {
        c.use_prev = false;
        drop_graphs(c);
        return SC_OK;
    }
    const size_t cb = sizeof(cplx) * c.csize;
    for (int m = 0; m < c.nmem; ++m)
        SC_HIP(hipMemcpy((char *) c.Cprev + (size_t) m * cb, prev + (size_t) m * c.csize, cb,
                         hipMemcpyHostToDevice));
    c.use_prev = true;
    drop_graphs(c);
    return SC_OK;
}

// Windows of their own for the members of a run: member m advances over
// steps < len[m] of the tangent member runs and starts at step len[m] of the
// adjoint member runs (drivers/wrapper.py:14-54 runs each segment by a call
// of its own; here segments of uneven length share the launches)
int sc_set_member_windows(const int *len)
{
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (len) {
        for (int m = 0; m < c.nmem; ++m)
            if (len[m] < 0 || len[m] > c.d.nsteps) {
                last_error() = "member windows must lie in 0 .. nsteps";
                return SC_EINVAL;
            }
        c.win.assign(len, len + c.nmem);
    } else {
        c.win.clear();
    }
    drop_graphs(c);
    return SC_OK;
}

// c_ddt_project on member `member`'s stored trajectory (its own states;
// the least-squares-shadowing segments run as members use it at their ends;
// at step 0 of a member with a previous state, the backward difference)
double sc_member_ddt_project(int member, int i_step, sc_complex *v)
{
    Context &c = ctx();
    last_error().clear();
    if (!c.ok || !v || member < 0 || member >= c.nmem || i_step < 0 || i_step > c.d.nsteps ||
        c.d.nsteps < 1) {
        last_error() = "member_ddt_project requires 0 <= member < members, 0 <= i_step <= "
                       "nsteps and nsteps >= 1";
        return NAN;
    }
    const size_t cb = sizeof(cplx) * c.csize;
    const long ip = i_step == 0 ? 3 : 3L * i_step, im = i_step == 0 ? 0 : 3L * (i_step - 1);
    const cplx *P = (const cplx *) ((const char *) mc(c, ip) + (size_t) member * c.mcbytes);
    const cplx *M = (const cplx *) ((const char *) mc(c, im) + (size_t) member * c.mcbytes);
    if (i_step == 0 && c.use_prev && member > 0) {
        // the backward difference from the state before the member's window
        P = (const cplx *) ((const char *) mc(c, 0) + (size_t) member * c.mcbytes);
        M = c.Cprev + (size_t) member * c.csize;
    }
    if (hipMemcpyAsync(c.Ctmp, v, cb, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        hipMemsetAsync(c.abort_flag, 0, sizeof(int), c.stream) != hipSuccess) {
        last_error() = "member_ddt_project: HIP transfer failed";
        return NAN;
    }
    const int nparts = launch_ddt_gram(c, P, M, 1.0 / c.d.dt, c.Ctmp);
    if (nparts <= 0) return NAN;
    launch_ddt_apply(c, c.Ctmp, c.Ctmp, P, M, 1.0 / c.d.dt, c.csize, nullptr, 0.0, nparts, 0);
    double scal[3] = {0, 0, 0};
    if (hipMemcpyAsync(v, c.Ctmp, cb, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipMemcpyAsync(scal, c.scal, sizeof(scal), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        last_error() = "member_ddt_project: HIP transfer failed";
        return NAN;
    }
    return scal[2];
}

}  // extern "C"

// ---------------------------------------------------------------- gradients
extern "C" int sc_gradient_euclid(const sc_complex *lam, sc_complex *g)
{
    // <lam, d> in the Gauss-weighted physical inner product of
    // drivers/testAdj.py:55-56 is, by Parseval on the dealiased grid and the
    // exactness of the Gauss rule (as in ddt_project's inner products),
    //   Px Pz sum over columns c Re[lam_h^H G d_h]
    // with G = Qs + Qps/k2 on alpha (rows < dimQ) and Rs/k2 on beta, c = 2
    // for kx > 0, c = 1 at kx = 0 where both enter through their Hermitian
    // parts (f(kz) + conj f(-kz)) / 2, and the mean column's real parts with
    // Rs; row Nz/2 counts nothing.  Its gradient in the real and imaginary
    // parts of d: 2 Px Pz G lam (kx > 0), Px Pz G lam_h (kx = 0), Px Pz Rs
    // Re lam (mean, real), 0 elsewhere.
    Context &c = ctx();
    if (!c.ok) return SC_ENOTREADY;
    if (!lam || !g) return SC_EINVAL;
    const Dims &d = c.d;
    const int Nz = d.Nz, Mx = d.Mx, nR = d.dimR, nQ = d.dimQ;
    const double pp = (double) d.Px * d.Pz;
    const cplx *L = (const cplx *) lam;
    cplx *G = (cplx *) g;
    auto at = [&](int z, int v, int i, int x) { return (((size_t) z * 2 + v) * nR + i) * Mx + x; };
    std::fill(G, G + c.csize, make_double2(0.0, 0.0));
    // band mat-vec y = A lam over rows of one column (A: N x (2dl+1), entry
    // (i, i + 2m) at A[i (2dl+1) + dl + m])
    auto band = [&](const double *A, int N, int dl, auto value, auto put) {
        for (int i = 0; i < N; ++i) {
            double sr = 0.0, si = 0.0;
            for (int m = -std::min(dl, i / 2); m <= std::min(dl, (N - i + 1) / 2 - 1); ++m) {
                const double a = A[(size_t) i * (2 * dl + 1) + dl + m];
                const cplx v = value(i + 2 * m);
                sr += a * v.x;
                si += a * v.y;
            }
            put(i, make_double2(sr, si));
        }
    };
    for (int z = 0; z < Nz; ++z) {
        if (z == Nz / 2) continue;
        for (int x = 0; x < Mx; ++x) {
            if (z == 0 && x == 0) {
                for (int v = 0; v < 2; ++v)
                    band(c.Rs.data(), nR, 2,
                         [&](int i) { return make_double2(L[at(0, v, i, 0)].x, 0.0); },
                         [&](int i, cplx y) { G[at(0, v, i, 0)] = make_double2(pp * y.x, 0.0); });
                continue;
            }
            const double k2 = c.Kz[z] * c.Kz[z] + c.Kx[x] * c.Kx[x];
            const int zm = (Nz - z) % Nz;
            auto lv = [&](int v, int i) {
                const cplx a = L[at(z, v, i, x)];
                if (x > 0) return a;
                const cplx b = L[at(zm, v, i, 0)];
                return make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
            };
            const double cw = (x > 0 ? 2.0 : 1.0) * pp;
            // alpha: (Qs + Qps / k2) on rows < dimQ
            for (int i = 0; i < nQ; ++i) {
                double sr = 0.0, si = 0.0;
                for (int m = -std::min(4, i / 2); m <= std::min(4, (nQ - i + 1) / 2 - 1); ++m) {
                    const size_t e = (size_t) i * 9 + 4 + m;
                    const double a = c.Qs[e] + c.Qps[e] / k2;
                    const cplx v = lv(0, i + 2 * m);
                    sr += a * v.x;
                    si += a * v.y;
                }
                G[at(z, 0, i, x)] = make_double2(cw * sr, cw * si);
            }
            band(c.Rs.data(), nR, 2, [&](int i) { return lv(1, i); },
                 [&](int i, cplx y) { G[at(z, 1, i, x)] = make_double2(cw * y.x / k2, cw * y.y / k2); });
        }
    }
    return SC_OK;
}
