"""Benchmark of the DEM step on one MI355X.

Workload (N=1 GPU): a bed of polydisperse spheres (log-normal radii, median 1 mm,
sigma_ln 0.2, clipped to [0.5, 2] mm, 64 radius classes, rho 2500) in a drum
of 64 quasi-plates + 2 end plates (R = 0.15 m) rotating at 2 rad/s about a
horizontal axis; ISwBSDBF linear spring-dashpot kn = 1e4 N/m, ks = 0.8 kn,
e = 0.6, mu = 0.4; dt = 5e-6 s; LeapConsultant candidate lists rebuilt every
UpIact steps.  A "step" is one SimMediator::Run over all spheres.  The run is
the configuration's own: the spheres start as a packed bed (see
scenes.drum_scene) and settle into the drum, which turns at 2 rad/s from
t = 0.  The neighbour-list skin is tuned on the first steps of the run,
--warmup steps follow, the run goes on to the next rebuild boundary, and then
--steps steps (10,000 by default, the configuration's step count) are timed.
UpIact is the largest period not above --up-iact that divides --steps, so the
window is whole rebuild periods with their rebuilds.  The headline is FP64,
the reference's arithmetic; a second handle brought to the same state steps
the same window with the FP32/FP64-mixed pair solves (vg_set_precision),
reported under "mixed" with the error of one mixed step against one FP64 step
from the window's end state (contact histories included).  --prep > 0 first
relaxes the bed in the resting drum instead (see settle()).

Prints one JSON line (see the metric/roofline/cpu_baseline fields).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M spheres vs plain GPU; p50/p99 scene latency batched"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per step (fixed here, used by the roofline):
#   280 B per sphere  = FP64 x,v,w read+write (144) + FP64 Ox,Oz read+write (96)
#                       + radius, mass and model read (20) + the relist guard's
#                       path length read+write (16), rounded to 280
#   72 B per contact  = (lo,hi) key (8) + FP64 tangential history read+write (48)
#                       + per-contact scalar state read+write (16)
# (the neighbour-list words of listed partners not in contact are not counted)
BYTES_PER_SPHERE = 280
BYTES_PER_CONTACT = 72


def measured_traffic():
    """HBM bytes per step launch of the drum workload from the committed
    counter summary (two rocprofv3 --pmc passes calibrated with the copy
    kernel), or None."""
    path = os.path.join(ROOT, "profiles", "k_step_drum1m_hbm_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1]), "profiles/k_step_drum1m_hbm_counters.md"
    except OSError:
        pass
    return None, None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(state, wa, n_gpu, seconds):
    """Oracle (FP64 C restatement of the reference step, single thread) on a
    bounded sample of the same workload: the spheres of a slab of the settled
    drum along its axis, with the drum walls, stepped for about `seconds`."""
    from oracle.oracle import OracleSim
    from vedo_amd.world import WorldArrays
    n = n_gpu
    pos = state["pos"]
    y = pos[:n, 1]
    target = min(20000, n)
    y_mid = float(np.median(y))
    keep = np.sort(np.argsort(np.abs(y - y_mid), kind="stable")[:target])
    y_lo, y_hi = float(y[keep].min()), float(y[keep].max())
    walls = np.arange(n, wa.n_do)
    ids = np.concatenate([keep, walls])
    sub = WorldArrays(**{k: getattr(wa, k) for k in wa.__dataclass_fields__})
    sub.model_of = wa.model_of[ids]
    sub.model_behavior = np.where(wa.model_shape == 1, wa.model_behavior, 2).astype(np.int32)   # walls fixed
    for f in ("pos", "vel", "ox", "oz", "av"):
        setattr(sub, f, np.ascontiguousarray(state[f][ids]))
    sim = OracleSim(sub)
    # one core: the process is pinned to the first core it may use while the oracle runs
    nproc, model = cpu_info()
    old_aff = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
    core = min(old_aff) if old_aff else None
    if core is not None:
        os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        steps = 0
        while time.perf_counter() - t0 < seconds:
            sim.step(1)
            steps += 1
        el = time.perf_counter() - t0
    finally:
        if old_aff:
            os.sched_setaffinity(0, old_aff)
    out = {"value": len(keep) * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
           "nproc": nproc, "cpu_model": model, "pinned_core": core,
           "sample": f"oracle (FP64 C restatement, 1 thread pinned to core {core}) on the {len(keep)} spheres of "
                     f"the same bed in the axial slab {y_lo:.4f} <= y <= {y_hi:.4f} m (middle of the drum) with its "
                     f"{len(walls)} walls held fixed, {steps} steps in {el:.1f} s"}
    cal = cpu_calibration()
    if cal:
        out["calibration"] = cal
    return out


def settle(sim, n, total_mass, max_steps, chunk=100, v_rms_stop=2e-3):
    """Untimed preparation: the bed relaxes in the resting drum.  Each time
    the spheres' kinetic energy has passed a maximum, all sphere velocities
    are reset to zero through SimMediator::Run(map) (a quench): that removes
    the slow elastic "breathing" of the deep bed, which contact damping alone
    damps only weakly.  Stops when the rms speed is below v_rms_stop."""
    ids = np.arange(n, dtype=np.uint32)
    zeros = np.zeros((n, 3))
    last, rising, done, quenches = None, False, 0, 0
    while done < max_steps:
        sim.step(chunk)
        done += chunk
        e = sim.system_energy()
        ke = e["translation"] + e["rotation"]
        v_rms = (2.0 * e["translation"] / total_mass) ** 0.5
        log(f"[bench] settle {done} steps: v_rms {v_rms:.2e} m/s, v_max {e['velocity_max']:.2e} m/s, "
            f"active contacts {sim.count_active()}, quenches {quenches}")
        if v_rms < v_rms_stop:
            break
        if last is not None and rising and ke < last:
            st = sim.state()
            sim.step_assigned(ids, st["pos"][:n], zeros, st["ox"][:n], st["oz"][:n], zeros)
            done += 1
            quenches += 1
            last, rising = None, False
            continue
        rising = last is None or ke > last
        last = ke
    return done, quenches


def batched_boxes(count, steps, up_iact, seed=4, loop=True):
    """The batched configuration: `count` box-settling scenes (ISwH), sphere
    counts log-uniform in [1e3, 2e5], mu ~ U[0.2, 0.7], e ~ U[0.3, 0.9],
    E ~ logU[1e6, 1e8], r = 1 mm, box sized to its lattice, dt = 0.1 of the
    critical time step of the contacts' tangential-rotational mode.  All scenes run `steps` steps in one batch (admission
    to results on the host, timed), and, when `loop`, the plain per-scene GPU
    loop: each scene alone, one after the other, in its own handle."""
    from vedo_amd import scenes
    from vedo_amd.gpu import SCENE_DONE, SCENE_FAILED, GpuBatch, GpuError, GpuSimulation
    was = [scenes.batched_box(k, steps, up_iact, count=count, seed=seed) for k in range(count)]
    n_tot = sum(int((wa.model_shape[wa.model_of] == 1).sum()) for wa in was)
    ps = float(n_tot) * steps
    log(f"[bench] batched: {count} box scenes, {n_tot} spheres, {steps} steps each")
    t0 = time.perf_counter()
    b = GpuBatch(n_tot, max_walls=6 * count, max_scenes=count, up_iact=up_iact)
    ids = [b.admit(wa, steps) for wa in was]
    t_adm = time.perf_counter() - t0
    # neighbour-list skin tuned on the batch's first steps (part of the run)
    from vedo_amd.autotune import tune_skin
    skin, tune = tune_skin(b, 1e-3, up_iact, factors=(0.25, 0.5, 1.0, 2.0))
    log(f"[bench] batched: skin {skin * 1e3:.2f} mm ({({f'{k * 1e3:.2f}': round(v * 1e3, 2) for k, v in tune.items()})} "
        f"ms/step)")
    # from the next rebuild on, each scene's skin follows its speed (the tuned
    # skin as the floor): the spheres fall faster as the run goes on
    skin_cap = max(4.0 * skin, 1e-3)
    b.set_skin_auto(skin, skin_cap)
    from vedo_amd.gpu import SCENE_RUNNING
    done_steps = sum(up_iact + 2 for _ in tune)   # the tuning's steps
    while True:
        st = [b.status(s) for s in ids]
        states = [x[0] for x in st]
        if all(x in (SCENE_DONE, SCENE_FAILED) for x in states):
            break
        # on to the nearest scene's end in one call (whole periods, which the
        # batch replays as graphs): no host round trip between periods
        rem = [steps - x[1] for x in st if x[0] == SCENE_RUNNING]
        chunk = max(up_iact, (min(rem) // up_iact) * up_iact) if rem else up_iact
        b.step(chunk)
        done_steps += chunk
        log(f"[bench] batched: {done_steps} steps, {sum(x == SCENE_DONE for x in states)} done, "
            f"{sum(x == SCENE_FAILED for x in states)} failed")
    t_run = time.perf_counter() - t0
    ok = [s for s, x in zip(ids, states) if x == SCENE_DONE]
    res = [b.state(s) for s in ok]
    t_batch = time.perf_counter() - t0
    log(f"[bench] batched: admission {t_adm:.2f} s, stepping {t_run - t_adm:.2f} s, results {t_batch - t_run:.2f} s")
    stats = b.stats()
    graph_replays = b.graph_replays()
    b.close()
    del res
    failed = [k for k, x in enumerate(states) if x == SCENE_FAILED]
    n_ok = sum(int((was[k].model_shape[was[k].model_of] == 1).sum()) for k in range(count) if k not in failed)
    ps = float(n_ok) * steps
    n_of = [int((wa.model_shape[wa.model_of] == 1).sum()) for wa in was]
    out = {"scenes": count, "spheres": n_tot, "steps_per_scene": steps, "up_iact": up_iact,
           "failed_scenes": len(failed), "failed": [[k, n_of[k]] for k in failed],
           "batched_seconds": t_batch, "batched_particle_steps_per_s": ps / t_batch,
           "batch_steps": done_steps, "relists": stats["relists"], "skin_m": skin,
           "graph_replays": graph_replays}
    log(f"[bench] batched: {t_batch:.2f} s, {ps / t_batch:.3e} particle-steps/s ({len(failed)} scenes blew up: "
        f"soft ISwH contacts)")
    if loop:
        t0 = time.perf_counter()
        for k, wa in enumerate(was):
            if k in failed:    # the same scene blows up alone: skipped in both
                continue
            sim = GpuSimulation(wa)
            sim.set_skin_auto(skin, skin_cap)   # the batch's skin policy
            try:
                sim.step(steps)
                sim.state()
            except GpuError:
                log(f"[bench] per-scene loop: scene {k} failed alone")
            sim.close()
            if k % 64 == 63:
                log(f"[bench] per-scene loop: {k + 1} scenes, {time.perf_counter() - t0:.1f} s")
        t_loop = time.perf_counter() - t0
        out.update({"loop_seconds": t_loop, "loop_particle_steps_per_s": ps / t_loop,
                    "speedup_vs_loop": t_loop / t_batch})
        log(f"[bench] per-scene loop: {t_loop:.2f} s, {ps / t_loop:.3e} particle-steps/s, "
            f"batched/loop {t_loop / t_batch:.2f}x")
    return out


def serving_run(count, up_iact, seed=5, loop=True):
    """The serving configuration: `count` requests of a Poisson stream (20/s)
    of drum / hopper / box scenes, 5e3-5e5 spheres, 2e3-2e4 steps, ISwH or
    ISwBSDBF, served by continuous batching; p50/p99 time-to-result and the
    aggregate throughput, and the same stream served one scene at a time."""
    from vedo_amd import serving
    reqs = serving.poisson_requests(count, rate=20.0, seed=seed)
    for r in reqs:
        serving.build_scene(r, up_iact=up_iact)
    need = sum(r.n for r in reqs)
    # a device layout for half the stream's spheres (at least its largest
    # scene): admission by footprint binds, later requests wait for slots
    cap = max(max(r.n for r in reqs), need // 2)
    log(f"[bench] serving: {count} requests, {need} spheres in all, capacity {cap} spheres "
        f"({cap * serving.slot_bytes() / 2 ** 30:.1f} GiB)")
    srv = serving.Server(capacity=cap, up_iact=up_iact, max_walls=80 * count, max_scenes=max(count, 16),
                         keep_results=False)
    st = srv.run(reqs)
    srv.close()
    waits = np.array([r.t_admit - r.arrival for r in reqs])
    period = max(w for w in waits) if len(waits) else 0.0
    out = {"requests": count, "rate_per_s": 20.0, "spheres": need, "capacity_spheres": cap, "failed": st.failed,
           "seconds": st.wall, "particle_steps_per_s": st.throughput, "p50_s": st.pct(50), "p99_s": st.pct(99),
           "admission_wait_p50_s": float(np.percentile(waits, 50)), "admission_wait_max_s": float(period),
           "requests_waiting_over_1s": int((waits > 1.0).sum())}
    log(f"[bench] serving: {st.wall:.2f} s, {st.throughput:.3e} particle-steps/s, "
        f"time-to-result p50 {st.pct(50):.2f} s p99 {st.pct(99):.2f} s")
    if loop:
        st1 = serving.serve_one_by_one(reqs)
        out.update({"loop_seconds": st1.wall, "loop_particle_steps_per_s": st1.throughput,
                    "loop_p50_s": st1.pct(50), "loop_p99_s": st1.pct(99), "loop_failed": st1.failed})
        log(f"[bench] serving one scene at a time: {st1.wall:.2f} s, {st1.throughput:.3e} particle-steps/s, "
            f"p50 {st1.pct(50):.2f} s p99 {st1.pct(99):.2f} s")
    return out


def window_up_iact(steps, up_max):
    """The largest rebuild period (UpIact) not above `up_max` that divides
    `steps`: the timed window is then whole rebuild periods, each with its
    candidate-list rebuild."""
    for u in range(min(up_max, max(steps, 1)), 0, -1):
        if steps % u == 0:
            return u
    return 1


def drum_prepare(args, wa, wall_ids, c, a, skin=None, steps=None):
    """A fresh FP64 handle of the drum workload advanced to the start of the
    timed window: the optional resting prep, the drum turning at 2 rad/s from
    then on, the neighbour-list skin tuned on the live run, --warmup steps,
    then on to the next rebuild boundary.  With `skin` and `steps` given (a
    second handle for the other precision) the handle takes that skin and
    runs the same number of steps: results do not depend on the skin, so it
    arrives at the first handle's state bit for bit.  Returns
    (handle, skin, steps run)."""
    from vedo_amd.gpu import GpuSimulation
    t = time.perf_counter()
    sim = GpuSimulation(wa)
    log(f"[bench] created on GPU ({time.perf_counter() - t:.1f} s)")
    done = 0
    if args.prep > 0:
        t = time.perf_counter()
        sim.set_wall_rotation(wall_ids, c, a, 0.0, 0.0)
        r_sph = wa.model_attr[wa.model_of[:args.n], 0]
        total_mass = float(np.sum(wa.model_density[wa.model_of[:args.n]] * 4.0 / 3.0 * np.pi * r_sph ** 3))
        done, quenches = settle(sim, args.n, total_mass, args.prep)
        log(f"[bench] prep {done} steps, {quenches} quenches, {time.perf_counter() - t:.1f} s")
    sim.set_wall_rotation(wall_ids, c, a, 2.0, sim.time)
    up = int(wa.up_iact)
    if skin is not None:
        sim.set_skin(skin)
        sim.step(steps - done)
        return sim, skin, steps
    if args.skin is not None:
        sim.set_skin(args.skin)
    elif not args.no_tune:
        from vedo_amd.autotune import tune_skin
        r0 = done
        skin, table = tune_skin(sim, float(wa.model_attr[wa.model_shape == 1, 0].min()), up)
        done += 4 * (up + 2)
        assert r0 + 4 * (up + 2) == done
        log(f"[bench] skin autotune: {({f'{k * 1e3:.3f}mm': round(v * 1e6, 1) for k, v in table.items()})} "
            f"us/step -> {skin * 1e3:.3f} mm")
    sim.step(args.warmup)
    done += args.warmup
    # on to a rebuild boundary: the window's steps are whole rebuild periods
    align = (-done) % up
    sim.step(align)
    done += align
    return sim, sim.skin, done


def time_window(sim, n, steps):
    """Times `steps` steps (whole rebuild periods) bracketed by device
    synchronisation; the step kernel's launches and the rebuilds are also
    timed with HIP events on the library's stream."""
    s0 = sim.stats()
    sim.kernel_timing(steps)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps, sync=False)
    sim.sync()
    wall = time.perf_counter() - t0
    k_ms, k_launches = sim.timing()
    r_ms, r_count = sim.rebuild_timing()
    s1 = sim.stats()
    stats = {k: s1[k] - s0[k] for k in ("rebuilds", "relists", "contact_steps")}
    # average number of contact pairs each timed step solved (on-device count)
    c_act = stats["contact_steps"] / steps
    bytes_step = BYTES_PER_SPHERE * n + BYTES_PER_CONTACT * c_act
    k_avg_s = k_ms / 1e3 / max(k_launches, 1)
    log(f"[bench] timed {steps} steps in {wall:.3f} s: step kernel avg {k_avg_s * 1e6:.1f} us over {k_launches} "
        f"launches, {r_count} rebuilds {r_ms:.2f} ms in all, {c_act:.0f} contact pairs per step, {stats}")
    return {"value": n * steps / wall, "ms_per_step": wall * 1e3 / steps, "wall_s": wall,
            "kernel_avg_us": k_avg_s * 1e6, "kernel_launches": k_launches, "achieved_gbs": bytes_step / k_avg_s / 1e9,
            "bytes_per_step": bytes_step, "active_contacts": c_act, "rebuilds": r_count,
            "rebuild_ms_each": r_ms / max(r_count, 1), "rebuild_share": r_ms / 1e3 / wall,
            "step_kernel_share": k_ms / 1e3 / wall, "relists": stats["relists"]}


def cpu_info():
    """(nproc, CPU model) of this host."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count(), model


def cpu_calibration():
    """The oracle against the compiled reference on one core of the build
    host (oracle/cpu_calibration.json, written by oracle/calibrate_cpu.py)."""
    try:
        with open(os.path.join(ROOT, "oracle", "cpu_calibration.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--prep", type=int, default=0, help="untimed settling steps in the resting drum")
    ap.add_argument("--up-iact", type=int, default=100,
                    help="largest candidate-list rebuild period; the run uses the largest one that divides --steps")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--skin", type=float, default=None, help="neighbour-list skin in m (skips the tuning)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=256, help="batched configuration: scenes (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=5000, help="batched configuration: steps per scene")
    ap.add_argument("--serve-requests", type=int, default=24, help="serving configuration: requests (0: skip)")
    ap.add_argument("--no-loop", action="store_true", help="skip the one-scene-at-a-time comparisons")
    ap.add_argument("--batch-up-iact", type=int, default=20, help="UpIact of the batched and served scenes")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="arithmetic of the headline (fp64: the reference's); the other one is reported beside it")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32/FP64-mixed leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU holds the whole domain); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from vedo_amd import scenes
    from vedo_amd.gpu import PRECISION_MIXED, copy_bandwidth

    copy_gbps = copy_bandwidth(1 << 30, 10)
    log(f"[bench] streaming copy kernel: {copy_gbps:.0f} GB/s")

    up = window_up_iact(args.steps, args.up_iact)
    t = time.perf_counter()
    wa, wall_ids, c, a = scenes.drum_scene(args.n, arrays=True, up_iact=up)
    log(f"[bench] scene: {args.n} spheres + {len(wall_ids)} walls, drum length {wa.pos[:, 1].max():.3f} m, "
        f"UpIact {up} ({time.perf_counter() - t:.1f} s)")
    n = args.n
    # two handles at the same state at the start of the same window: one
    # steps in FP64, the other in the FP32/FP64-mixed precision
    sim, skin, done = drum_prepare(args, wa, wall_ids, c, a)
    other = None
    if not args.no_mixed:
        other, _, _ = drum_prepare(args, wa, wall_ids, c, a, skin=skin, steps=done)
    log(f"[bench] window starts after {done} steps (a rebuild boundary), skin {skin * 1e3:.3f} mm")
    run64 = time_window(sim, n, args.steps)
    runmix = None
    if other is not None:
        other.set_precision(PRECISION_MIXED)
        runmix = time_window(other, n, args.steps)
        other.close()
        del other
    headline_mixed = args.precision == "mixed" and runmix is not None
    run = runmix if headline_mixed else run64
    traffic, traffic_src = measured_traffic()
    out = {
        "metric": METRIC, "value": run["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": run["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32" if headline_mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "rotating-drum polydisperse spheres, ISwBSDBF, LeapConsultant", "spheres": n,
                   "walls": len(wall_ids), "radius_classes": 64, "dt": wa.dt, "up_iact": up,
                   "window": f"steps {done + 1}-{done + args.steps} of the run: {args.steps // up} whole rebuild "
                             f"periods, their {run['rebuilds']} rebuilds included",
                   "prep_steps": done, "active_contacts": run["active_contacts"], "neighbour_skin_m": skin,
                   "precision": "mixed" if headline_mixed else "fp64", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": run["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": run["achieved_gbs"] / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "copy_kernel_gbps": copy_gbps, "frac_of_copy": run["achieved_gbs"] / copy_gbps,
                     "kernel": "k_step", "kernel_avg_us": run["kernel_avg_us"],
                     "algorithmic_bytes_per_launch": run["bytes_per_step"]},
        "rebuild": {"count": run["rebuilds"], "ms_each": run["rebuild_ms_each"], "share_of_window": run["rebuild_share"],
                    "step_kernel_share_of_window": run["step_kernel_share"]},
        "relists": run["relists"],
    }
    st = sim.state()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, wa, n, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    if runmix is not None:
        # error of one mixed step against one FP64 step from the window's end
        # state (positions, velocities, orientations and the contacts' histories)
        import dataclasses
        from vedo_amd.precision import compare_precisions
        hist = sim.active()
        wa_end = dataclasses.replace(wa, pos=st["pos"], vel=st["vel"], ox=st["ox"], oz=st["oz"], av=st["av"],
                                     t_current=sim.time)
        sim.close()
        del sim
        err = compare_precisions(wa_end, steps=1, history=hist)
        log(f"[bench] mixed vs fp64 on the end state: {err}")
        side, name = (run64, "fp64") if headline_mixed else (runmix, "mixed")
        out[name] = {"value": side["value"], "unit": "particle-steps/s", "steps": args.steps,
                     "ms_per_step": side["ms_per_step"], "kernel_avg_us": side["kernel_avg_us"],
                     "achieved_gbs": side["achieved_gbs"], "frac": side["achieved_gbs"] / HBM_PEAK_GBS,
                     "rebuild_ms_each": side["rebuild_ms_each"], "window": "the headline's (same start state)",
                     "speedup_mixed_vs_fp64": runmix["value"] / run64["value"], "error_vs_fp64": err}
    else:
        sim.close()
        del sim
    if args.batch_scenes > 0:
        out["batched"] = batched_boxes(args.batch_scenes, args.batch_steps, args.batch_up_iact, loop=not args.no_loop)
    if args.serve_requests > 0:
        out["serving"] = serving_run(args.serve_requests, args.batch_up_iact, loop=not args.no_loop)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
