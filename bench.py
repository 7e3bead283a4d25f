"""Benchmark: forward+adjoint steps per second of the channel DNS at
64x64 Fourier modes x 97 Legendre modes (96x145x96 dealiased grid) on one GPU.

One step of the workload is one primal RK3/CN step followed by one adjoint
step (strength 1, with the two trapezoid ddt_project forcings per step), the
unit the least-squares-shadowing drivers repeat per time step
(drivers/wrapper.py:14-30).  Timed: K primal steps then the K adjoint steps
back over the same window, with the states resident in device memory; the
host only launches.  Prints one JSON line.

The headline value is the batched mixed-precision rate: 32 independent
members (seeded states; --members) advanced together in every kernel launch, with the
x/z transforms, the nonlinear products and the y-transform to the Gauss
points in FP32 (Galerkin projections, solves, flux correction and the
trapezoid projection FP64), in member-steps per second.  Beside it:
the plain FP64 single-solution rate (the metric's denominator), the FP64
batched rate, the rate without the transformed-field cache, and the wall-
shear gradient of a 1000-step window (checkpoints every 50) in mixed
precision against FP64 and the FP64 gradient against fourth-order central
differences (both untimed).

Box: Lx = 2 pi, Lz = pi, Re = 3000, dt = 0.004, meanU = 2/3 (flux 4/3),
synthetic seeded state (laminar profile + k^-2 perturbation, amplitude 0.02).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shadow-channel_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "fwd+adjoint steps/s at 64x64x97 vs plain FP64 GPU; grad err vs FP64/FD"
CFG = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)
HBM_PEAK_GBS = 8000.0
CLASS_NAMES = {1: "z-inverse FFT (k_zinv)", 2: "x transform + products (k_xmix)",
               3: "z-forward FFT (k_zfwd)", 4: "y-transform to Gauss points (k_rebuild)",
               5: "Galerkin projection (k_force)", 6: "band solves (k_solve)",
               7: "merged adjoint x transform + products (k_xmix ADJ2)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fwd_adj(cc, C0, nsteps):
    """Host-array interface (c_primal / c_adjoint): includes the transfers of
    the state and the adjoint state between host and device."""
    C = C0.copy()
    cc.c_primal(0, C)
    AC = np.zeros_like(C0)
    cc.c_adjoint(nsteps, 0, AC, 0, 1.0)
    return C, AC


def fwd_adj_dev(cc, Cd, ACd, nsteps):
    """Device-resident interface: the states stay in HBM."""
    cc.sc_primal_dev(0, Cd)
    cc.sc_adjoint_dev(nsteps, 0, ACd, 1.0)


def unit_direction(shape, seed):
    """A seeded tangent direction of the state's spectral shape (unit norm)."""
    from channel.synthetic import synthetic_state
    c = CFG
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=seed,
                           unit=True)


def fwd_adj_members(cc, Cd, ACd, nsteps):
    """All members in every launch: their primal runs, then their forced
    adjoints along their own trajectories (states resident in HBM)."""
    cc.sc_primal_members(0, Cd)
    cc.sc_adjoint_members(nsteps, 0, ACd, 1.0, False)


RETUNE = False      # --retune: time the launch shapes instead of reading them


def member_rate(cc, states, K, W, mixed, roofline_class=0, replays=0):
    """fwd+adj member-steps per second of len(states) independent members
    advanced together (mixed: FP32 transforms): W untimed steps, then K
    timed.  With roofline_class, one more fwd+adj step of the same members
    records that kernel class's launches and the last one is replayed cold:
    (rate, elapsed, roofline tuple or None, the launch shapes the autotuner
    chose for the configuration before timing)."""
    from channel import tuning
    c = CFG
    nm = len(states)
    # untimed: the launch shapes of this configuration, as the autotuner
    # chose them on an MI355X (channel/shapes.json), or timed again
    shapes = tuning.autotune((c["Nx"], c["Ny"], c["Nz"]), (c["Lx"], c["Lz"]), c["Re"], c["flux"],
                             c["dt"], np.stack(states), mixed=bool(mixed),
                             cache=None if RETUNE else tuning.SHAPES_FILE)
    cc.sc_set_members(nm)
    cc.sc_set_precision(mixed)
    roof = None
    try:
        if W > 0:
            assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                             c["dt"], W) == 0
            Cd = cc.DeviceArray(np.stack(states))
            ACd = cc.DeviceArray(np.zeros((nm,) + states[0].shape, complex))
            fwd_adj_members(cc, Cd, ACd, W)
            Cd.free()
            ACd.free()
            cc.c_destroy()
        assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                         c["dt"], K) == 0
        # untimed: each member's trajectory and a tangent along it, so that
        # the timed adjoints' forcing (their stored tangents) is not zero; the
        # timed primal runs recompute the same trajectories
        Cd = cc.DeviceArray(np.stack(states))
        cc.sc_primal_members(0, Cd)
        Cd.upload(np.stack(states))
        Id = cc.DeviceArray(np.stack([unit_direction(s.shape, 100 + m)
                                      for m, s in enumerate(states)]))
        cc.sc_tangent_members(0, K, Id, 0, False)
        Id.free()
        ACd = cc.DeviceArray(np.zeros((nm,) + states[0].shape, complex))
        cc.sc_synchronize()
        t0 = time.perf_counter()
        fwd_adj_members(cc, Cd, ACd, K)
        cc.sc_synchronize()
        elapsed = time.perf_counter() - t0
        C, AC = Cd.download(), ACd.download()
        assert cc.last_status() == 0 and np.all(np.isfinite(C)) and np.all(np.isfinite(AC))
        Cd.free()
        ACd.free()
        cc.c_destroy()
        if roofline_class:
            # one fwd+adj step of the same members records the class's
            # launches; the last is re-issued `replays` times, each after a
            # cache-evicting fill whose own time is subtracted
            assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                             c["dt"], 1) == 0
            Cd = cc.DeviceArray(np.stack(states))
            ACd = cc.DeviceArray(np.zeros((nm,) + states[0].shape, complex))
            cc.sc_prof_enable(roofline_class)
            fwd_adj_members(cc, Cd, ACd, 1)
            n, ms, nbytes = cc.sc_prof_read()
            avg_ms, bytes1 = cc.sc_prof_replay_cold(replays, FLUSH_MB)
            cc.sc_prof_enable(0)
            Cd.free()
            ACd.free()
            cc.c_destroy()
            roof = (n, ms, nbytes, avg_ms, bytes1)
    finally:
        cc.sc_set_members(1)
        cc.sc_set_precision(0)
        tuning.reset()
    return nm * K / elapsed, elapsed, roof, shapes


def gradient_errors(cc):
    """Wall-shear gradient of a 1000-step window (checkpoints every 50) at the
    benchmark grid from the seed-0 state: mixed precision against FP64 (J,
    dJ/dmeanU, dJ/dRe, dJ/dC0), and the FP64 adjoint along a unit direction
    against a fourth-order central difference at 1e-3 of the state's size
    (the chaotic growth of rounding differences over 1000 steps makes
    smaller steps noisy)."""
    from channel.gradient import WallShearWindow
    from channel.synthetic import synthetic_state
    c = CFG
    N, meanU = 1000, c["flux"] / 2
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
    d = unit_direction(C0.shape, 9)
    out = {}
    for mixed in (0, 1):
        cc.sc_set_precision(mixed)
        win = WallShearWindow(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["dt"], N, every=50)
        try:
            J, cps, dmu, dre = win.forward(C0, meanU, c["Re"], params=True)
            g = win.state_gradient(cps, meanU, c["Re"])
            fd = None
            if not mixed:
                f = lambda t: win.forward(C0 + t * d, meanU, c["Re"], params=False)[0]
                h = 1e-3 * np.linalg.norm(C0)
                fd = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
            out[mixed] = (J, dmu, dre, g, fd)
        finally:
            win.release()
            cc.sc_set_precision(0)
    (J0, mu0, re0, g0, fd), (J1, mu1, re1, g1, _) = out[0], out[1]
    gd = float(np.sum(g0.real * d.real + g0.imag * d.imag))
    rel = lambda a, b: float(abs(a - b) / abs(b))
    return {"window_steps": N, "checkpoint_every": 50,
            "mixed_vs_fp64": {"J": rel(J1, J0), "dJ_dmeanU": rel(mu1, mu0),
                              "dJ_dRe": rel(re1, re0),
                              "dJ_dC0": float(np.linalg.norm(g1 - g0) / np.linalg.norm(g0))},
            "fp64_vs_fd": {"dJ_dC0_along_unit_direction": rel(gd, fd)}}


# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs
# (scaled by a copy of known size), per kernel class of the roofline replay and
# member count, mixed precision
FLUSH_MB = 320       # above the 256 MB Infinity Cache

TRAFFIC_PROFILES = {(6, 8): "profiles/solve_8members_mixed_64x64x97_pmc.md",
                    (6, 32): "profiles/solve_32members_mixed_64x64x97_pmc.md",
                    (8, 1): "profiles/iadj_stage_64x64x97.md"}


def measured_traffic(kernel_class, members):
    path = TRAFFIC_PROFILES.get((kernel_class, members))
    if not path or not os.path.exists(os.path.join(ROOT, path)):
        return None
    for ln in open(os.path.join(ROOT, path)):
        if ln.startswith("traffic_bytes_per_launch:"):
            return int(ln.split(":")[1])
    return None


def cpu_baseline():
    """The oracle (single-threaded FP64 restatement, computing the reference's
    discarded terms as well) on a bounded sample: one primal step and one
    adjoint step of the same 64x64x97 workload."""
    import pyorc
    from channel.synthetic import synthetic_state
    c = CFG
    L = pyorc.load()
    L.orc_set_faithful_cost(1)
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], 1)
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
    t0 = time.perf_counter()
    o.primal(0, C0.copy())
    AC = np.zeros_like(C0)
    o.adjoint(1, 0, AC, 0, 1.0)
    t = time.perf_counter() - t0
    L.orc_set_faithful_cost(0)
    o.destroy()
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": 1.0 / t, "unit": "steps/s", "cores": 1, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(),
            "sample": "1 primal + 1 adjoint step of the 64x64x97 workload, oracle (FP64 "
                      "restatement, single thread, incl. the reference's discarded "
                      "tangent/incremental-adjoint transforms), %.1f s" % t}


def iadj_flops(q, nQ, nR, k):
    """Algorithmic FP64 operations of one mode of the incremental adjoint
    stage: the Galerkin projections (alpha: Qpw, Qw, Qppw rows; beta: Rw,
    Rpw; a real-by-complex multiply-add is 4 operations) of this stage's
    products and, for k < 2, the next stage's; the five y-sums of the
    Gauss-point values; the band products (9- and 5-point) of the explicit
    parts and the tangent coupling."""
    proj = 4 * q * (3 * nQ + 2 * nR)
    band = 4 * 2 * (9 * nQ + 5 * nR)
    f = proj + 4 * 5 * q * nQ + (band if k == 0 else 0)
    if k < 2:
        f += proj + band
    return f


def iadj_cpu_baseline(IAC, IU, LIU, W):
    """The oracle's incremental adjoint stage (single thread) on the same
    inputs at the same grid, after its own one-step primal and tangent."""
    import pyorc
    from channel.synthetic import synthetic_state
    c = CFG
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], 1)
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
    o.primal(0, C0.copy())
    o.tangent(0, 1, unit_direction(C0.shape, 7), 0)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        o.increadjproject(0, 3, IAC, IU, LIU, W)
    t = (time.perf_counter() - t0) / reps
    o.destroy()
    return {"value": round(1.0 / t, 3), "unit": "stages/s", "cores": 1, "kind": "port",
            "sample": "%d stages (k = 0 with the next stage's explicit part) of the 64x64x97 "
                      "grid on the same inputs, oracle (FP64 restatement, single thread), "
                      "%.2f s each" % (reps, t)}


def incremental_adjoint_line(cc, calls, cpu=True):
    """Kernel time of the incremental adjoint stage (sc_increadjproject,
    channel/increadjproject.c): stage 0 with the next stage's explicit part
    at the benchmark grid, random products and wall planes, the tangent
    states of a one-step tangent run; `calls` calls, each on fresh inputs,
    the stage's launches (projections, right-hand sides and solves, the
    y-transform, the wall lifts) bracketed by HIP events on the library's
    stream."""
    from channel.synthetic import synthetic_state
    c = CFG
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     1) == 0
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
    cc.c_primal(0, C0.copy())
    cc.c_tangent(0, 1, unit_direction(C0.shape, 7), 0)
    Nz, Mx, q, nR, nQ = c["Nz"], c["Nx"] // 2, cc.c_qpts(), cc.c_dimR(), cc.c_dimQ()
    rng = np.random.default_rng(0)

    def rc(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    IAC, IU, LIU, W = rc((Nz, 2, nR, Mx)), rc((Nz, 5, q, Mx)), rc((Nz, 5, q, Mx)), rc((6, Nz, Mx))
    cc.sc_increadjproject(0, 3, IAC.copy(), IU, LIU, W)      # untimed: constants, code
    cc.sc_prof_enable(8)
    for _ in range(calls):
        cc.sc_increadjproject(0, 3, IAC.copy(), IU, LIU, W)
    n, ms, nbytes = cc.sc_prof_read()
    cc.sc_prof_enable(0)
    cc.c_destroy()
    base = iadj_cpu_baseline(IAC, IU, LIU, W) if cpu else None
    avg_s = ms / n / 1000.0
    modes = (Nz - 1) * Mx - 1
    flops = modes * iadj_flops(q, nQ, nR, 0)
    tf = flops / avg_s / 1e12
    gbs = nbytes / n / avg_s / 1e9
    return {"metric": "incremental adjoint stage (increadjproject) at 64x64x97",
            "value": round(1.0 / avg_s, 1), "unit": "stages/s", "n_gpus": 1, "steps": n,
            "warmup": 1, "ms_per_step": round(avg_s * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "one incremental adjoint stage (k = 0 with the next stage's "
                                   "explicit part) over every mode of the 64x64x97 grid, random "
                                   "products and wall planes, tangent states of a one-step run",
                       "modes": modes, "flops_per_launch": flops,
                       "bytes_per_launch": round(nbytes / n)},
            "roofline": {"kernel": "incremental adjoint stage (all launches)", "bound": "mfma", "achieved": round(tf, 3),
                         "peak": 78.6, "unit": "TFLOP/s", "frac": round(tf / 78.6, 4),
                         "traffic": measured_traffic(8, 1), "hbm_achieved_gbs": round(gbs, 1),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                         "avg_launch_us": round(avg_s * 1e6, 2)},
            "cpu_baseline": base}


def main():
    # the library prints the reference's grid banner on stdout at init: send
    # file descriptor 1 to stderr while it runs so that stdout carries only
    # the JSON line
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)
    try:
        line = run()
    finally:
        sys.stdout.flush()
        os.dup2(json_fd, 1)
        os.close(json_fd)
    if line is None:
        return 2
    print(json.dumps(line), flush=True)
    return 0


def run():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--roofline-class", type=int, default=6)
    ap.add_argument("--replays", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the untimed gradient-error check")
    ap.add_argument("--members", type=int, default=32,
                    help="independent members advanced together in the headline figure")
    ap.add_argument("--retune", action="store_true",
                    help="time the launch shapes instead of using channel/shapes.json")
    ap.add_argument("--incremental-adjoint", type=int, default=0, metavar="CALLS",
                    help="instead of the headline, time the incremental adjoint stage's "
                         "kernel over CALLS calls and print its own line")
    args = ap.parse_args()
    global RETUNE
    RETUNE = args.retune
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver runs one simulation per device and "
              "multi-GPU sharding is not implemented (--gpus must be 1)",
              file=sys.stderr)
        return None
    import c_channel as cc
    from channel.synthetic import synthetic_state
    if args.incremental_adjoint > 0:
        return incremental_adjoint_line(cc, args.incremental_adjoint, not args.no_cpu_baseline)
    c = CFG
    K, W = args.steps, args.warmup
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)

    def single_rate(field_cache):
        """plain FP64, one solution: device-resident fwd+adj rate (and the
        host-array rate with the cache on)"""
        cc.sc_set_field_cache(field_cache)
        try:
            if W > 0:
                assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"],
                                 c["flux"], c["dt"], W) == 0
                fwd_adj(cc, C0, W)
                cc.c_destroy()
            assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                             c["dt"], K) == 0
            assert cc.sc_field_cache_active() == field_cache
            # untimed: the trajectory and a tangent along it (the adjoint's
            # forcing); the timed primal recomputes the same trajectory
            Cd = cc.DeviceArray(C0)
            cc.sc_primal_dev(0, Cd)
            Cd.upload(C0)
            cc.c_tangent(0, K, unit_direction(C0.shape, 100), 0)
            ACd = cc.DeviceArray(np.zeros_like(C0))
            cc.sc_synchronize()
            t0 = time.perf_counter()
            fwd_adj_dev(cc, Cd, ACd, K)
            cc.sc_synchronize()
            elapsed = time.perf_counter() - t0
            C, AC = Cd.download(), ACd.download()
            Cd.free()
            ACd.free()
            assert cc.last_status() == 0 and np.all(np.isfinite(C)) and np.all(np.isfinite(AC))
            host = None
            if field_cache:
                # the same window through the host-array interface (PCIe
                # transfers of the states included)
                cc.sc_synchronize()
                t0 = time.perf_counter()
                Ch, ACh = fwd_adj(cc, C0, K)
                host = K / (time.perf_counter() - t0)
                assert np.array_equal(Ch, C) and np.array_equal(ACh, AC)
            cc.c_destroy()
        finally:
            cc.sc_set_field_cache(1)
        return K / elapsed, host

    log("timing plain FP64 fwd+adj, %d steps" % K)
    plain, host_rate = single_rate(1)
    log("timing plain FP64 without the field cache")
    nocache, _ = single_rate(0)
    nm = args.members
    states = [synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=s)
              for s in range(nm)]
    log("timing %d FP64 members x %d fwd+adj steps" % (nm, K))
    rate64, _, _, shapes64 = member_rate(cc, states, K, W, 0)
    # the headline: members in mixed precision, and the roofline of one
    # kernel class in that same configuration
    log("timing %d mixed-precision members x %d fwd+adj steps (+ roofline class %d)" % (
        nm, K, args.roofline_class))
    mrate, melapsed, rf, shapes = member_rate(cc, states, K, W, 1, args.roofline_class, args.replays)
    n, ms, nbytes, avg_ms, bytes1 = rf
    achieved = bytes1 / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.roofline_class, nm)
    roof = {"kernel": CLASS_NAMES.get(args.roofline_class, str(args.roofline_class)),
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "avg_launch_us": round(1e3 * avg_ms, 2),
            "bytes_per_launch": round(bytes1), "launches_per_step": n,
            "members_per_launch": nm,
            "timing": "HIP events on the library's stream around %d launches of the step's "
                      "last one, each after a %d MB cache-evicting fill (fill time measured "
                      "alone and subtracted)" % (args.replays, FLUSH_MB),
            "step_avg_launch_us": round(1e3 * ms / n, 2),
            "step_bytes_per_launch": round(nbytes / n),
            "step_avg_frac": round((nbytes / n) / (ms / n * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    grad = None
    if not args.no_gradient:
        log("gradient errors: 1000-step window, mixed vs FP64, FP64 vs finite differences")
        grad = gradient_errors(cc)
    line = {
        "metric": METRIC, "value": round(mrate, 3), "unit": "steps/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(1e3 * melapsed / K, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "fwd+adj step: 1 primal RK3/CN step + 1 adjoint step "
                               "(strength 1) per member, %d independent members advanced "
                               "together in every kernel launch, mixed precision (x/z "
                               "transforms, products and the y-transform to the Gauss "
                               "points FP32; Galerkin projections, solves, flux "
                               "correction FP64), states resident in HBM; "
                               "value = member-steps/s" % nm,
                   "grid": "64x64x97 (96x145x96 dealiased)", "Lx": "2pi", "Lz": "pi",
                   "Re": c["Re"], "dt": c["dt"], "meanU": c["flux"] / 2,
                   "members": nm, "window_steps": K, "precision": "mixed",
                   "plain_fp64_steps_per_s": round(plain, 3),
                   "ratio_vs_plain_fp64": round(mrate / plain, 3),
                   "fp64_members_steps_per_s": round(rate64, 3),
                   "plain_fp64_no_field_cache_steps_per_s": round(nocache, 3),
                   "launch_shapes": shapes, "launch_shapes_fp64_members": shapes64,
                   "host_array_steps_per_s": round(host_rate, 3)},
        "roofline": roof,
    }
    if grad is not None:
        line["config"]["gradient_errors"] = grad
    # the incremental adjoint stage at the same grid (untimed for the
    # headline; its own line: --incremental-adjoint)
    try:
        ia = incremental_adjoint_line(cc, 10, cpu=False)
        line["config"]["incremental_adjoint_stage"] = {
            "stages_per_s": ia["value"], "avg_us": ia["roofline"]["avg_launch_us"],
            "tflops": ia["roofline"]["achieved"], "frac_fp64_peak": ia["roofline"]["frac"],
            "workload": ia["config"]["workload"]}
    except Exception as e:   # reported, never fatal for the headline
        line["config"]["incremental_adjoint_stage"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        line["cpu_baseline"] = cpu_baseline()
    return line


if __name__ == "__main__":
    sys.exit(main())
