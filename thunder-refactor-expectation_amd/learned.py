"""Host-side driver of the learned-proposal expectation.

The proposal network is trained on particles the library's own projector
generates from the expectation's reference volume (random poses, defocus and
shifts, white noise over a range of SNRs), with the hand-written HIP training
kernels.  The fallback rule of thx_ctx_expect_learned (proposal spread, tail
depth of the best rotation, score drop below the batch median, ESS) is
calibrated on a held-out synthetic set that is scored both ways: the plain
whole-grid scan and the learned step.  The calibration picks, among candidate
selection sizes and thresholds, the one with the lowest predicted time whose
assignments agree with the plain scan's (rotation within one sampling step)
on at least the target fraction of the held-out particles.

There is no counterpart in the reference: it has no learned proposal.
"""
import time

import numpy as np

import thx
import workloads as wl


def train_proposal(net, ctx, workload, n_samples=31_457_280, chunk=8192, batch=512,
                   snr=(0.01, 0.1), seed=5000, schedule=((0.5, 1e-3), (0.25, 3e-4), (0.25, 1e-4)),
                   log=None):
    """Train `net` on synthetic particles of the reference volume held by `ctx`
    (class 0, pixel set set).  Returns (seconds, last losses)."""
    w = wl.WORKLOADS[workload]
    rng = np.random.default_rng(seed)
    n_chunks = max(1, n_samples // chunk)
    lrs = []
    for frac, lr in schedule:
        lrs += [lr] * int(round(frac * n_chunks))
    lrs = (lrs + [schedule[-1][1]] * n_chunks)[:n_chunks]
    t0 = time.perf_counter()
    losses = []
    for ch, lr in enumerate(lrs):
        q, t, _ = ctx.synth_particles(0, chunk, w["max_shift"], w["df_min"], w["df_max"],
                                      float(rng.uniform(*snr)), seed + ch, **wl.CTF_DEFAULTS)
        losses.append(net.train_ctx(ctx, q, t, batch, 1, lr))
        if log and (ch % 60 == 0 or ch == n_chunks - 1):
            log(f"train chunk {ch + 1}/{n_chunks} loss {losses[-1]:.3f} "
                f"{time.perf_counter() - t0:.1f} s")
    return time.perf_counter() - t0, losses[-8:]


# (rotations, translations) selected per particle.  The scorer pads
# translations to 32 / 64 / 128 rows and rotations to blocks of 128 / 256, so
# the 32 / 64-translation and 128 / 256 / 512-rotation sizes fill the tiles
# they pay for.
CANDIDATES = [(128, 32), (256, 32), (512, 32), (128, 64), (192, 64), (256, 64), (300, 49),
              (500, 49), (512, 64), (256, 121)]
SPREAD = [20.0, 30.0, 50.0, 80.0, 0.0]
TAIL = [3.0, 4.0, 5.0, 6.0, 0.0]
DROP = [100.0, 150.0, 0.0]
# fallbacks by a wider selection of the same proposal, (rotations, translations)
WIDE = [(1024, 64), (1024, 128), (2048, 64), (2048, 128)]   # clipped to the grid


def wide_name(nr, nt):
    return f"sel{nr}x{nt}"


def fallback_setup(ctx, name):
    """Put the fallback `name` on `ctx` ("bf16" / "fp16" / "fp8": the whole-grid
    scan in that precision; "sel<R>x<T>": the R x T most probable grid poses of
    the proposal) and return the precision argument of expect_learned."""
    if name.startswith("sel"):
        nr, nt = (int(x) for x in name[3:].split("x"))
        ctx.set_learned_fallback_selection(nr, nt)
        return thx.BF16
    ctx.set_learned_fallback_selection(0, 0)
    return thx.PRECISIONS[name]


def _held_out(ctx, net, step, candidates, fallbacks, wide):
    """One held-out set's statistics: per fallback name, (ms per particle,
    agreement with the plain FP32 scan per particle); per candidate selection
    size, (ms per particle of the learned step without fallback, agreement,
    spread, tail depth, best score) per particle."""
    quat = ctx.grid_quat
    P = ctx.P
    ctx.score_global(0, thx.FP32)
    gR, _, _ = ctx.get_best(0)

    def agree(bR):
        return wl.geodesic_angle(quat[bR], quat[gR]) <= step * (1 + 1e-6)

    full = {}
    for name in fallbacks:
        # timed on a second call: the first may grow buffers inside its timed stages
        for _ in range(2):
            ctx.score_global(0, thx.PRECISIONS[name])
        ctx.sync()
        t = sum(ctx.last_timing().values()) / P
        full[name] = (t, agree(ctx.get_best(0)[0]))
    ctx.set_learned_fallback_selection(0, 0)
    for nr, nt in sorted({(min(nr, ctx.R), min(nt, ctx.T)) for nr, nt in wide}):
        # a wider selection scored for every particle: its agreement per particle,
        # and its selection + scoring time per particle
        for _ in range(2):
            ctx.expect_learned(net, 0, nr, nt)
        tm = ctx.last_timing()
        t = (tm.get("select", 0.0) + tm.get("score_sel", 0.0)) / P
        full[wide_name(nr, nt)] = (t, agree(ctx.get_learned()[0]))
    cand = {}
    for nr, nt in candidates:
        for _ in range(2):   # the second call is the timed one, as above
            ctx.expect_learned(net, 0, nr, nt)
        tm = ctx.last_timing()
        t_sel = sum(v for k, v in tm.items() if not k.startswith("fallback")) / P
        bR, bT, sc, ess, spread, tail, fell = ctx.get_learned()
        cand[(nr, nt)] = (t_sel, agree(bR), spread, tail, sc)
    return full, cand


def calibrate(ctx, net, step, target=0.99, margin=0.003, candidates=CANDIDATES, log=None,
              fallbacks=("bf16", "fp8"), wide=WIDE, n_measure=6, robust=()):
    """Choose selection sizes, fallback thresholds and the fallback on the
    held-out set in `ctx` (particles, grid and class-0 volume set).  The
    fallback is the whole-grid scan in one of `fallbacks` precisions or a
    wider selection of the same proposal (`wide` sizes); a particle that falls
    back is counted correct when that fallback alone agrees with the plain
    FP32 scan on it.  A rule must reach target + margin on `ctx` and on every
    context in `robust` (held-out sets of other maps, e.g. the reference map
    as one refinement iteration would change it); times are those of `ctx`.
    Returns a dict with the chosen n_rot, n_trans, spread_max, tail_max,
    score_drop, fallback (a name for fallback_setup) and the calibration
    statistics; the contexts' fallback is left at the whole-grid scan."""
    P = ctx.P
    sets = [_held_out(c, net, step, candidates, fallbacks, wide) for c in (ctx,) + tuple(robust)]
    full0 = sets[0][0]
    ok_rules = []   # every rule that meets the target on every set, with its predicted time
    ok_rules0 = []  # those that meet it on ctx's set alone
    for nr, nt in candidates:
        t_sel = sets[0][1][(nr, nt)][0]
        for sm in SPREAD:
            for tmx in TAIL:
                for dz in DROP:
                    accs, fbs = {}, []
                    for full, cand in sets:
                        _, ok, spread, tail, sc = cand[(nr, nt)]
                        fb = np.zeros(len(ok), bool)
                        if sm > 0:
                            fb |= spread > sm
                        if tmx > 0:
                            fb |= tail > tmx
                        if dz > 0:
                            fb |= sc < np.median(sc) - dz
                        fbs.append(fb)
                        for name, (_, ok_full) in full.items():
                            accs.setdefault(name, []).append(
                                float(np.mean(np.where(fb, ok_full, ok))))
                    for name, acc in accs.items():
                        if acc[0] < target + margin:
                            continue
                        rule = dict(n_rot=nr, n_trans=nt, spread_max=sm, tail_max=tmx,
                                    score_drop=dz, fallback=name, cal_accuracy=acc[0],
                                    cal_accuracy_robust=acc[1:],
                                    cal_fallback=float(fbs[0].mean()),
                                    cal_accuracy_no_fallback=float(sets[0][1][(nr, nt)][1].mean()),
                                    pred_ms_per_particle=float(
                                        t_sel + fbs[0].mean() * full0[name][0]))
                        ok_rules0.append(rule)
                        if min(acc) >= target + margin:
                            ok_rules.append(rule)
        if log:
            log(f"calibrate {nr}x{nt}: no-fallback agreement "
                f"{[float(c[(nr, nt)][1].mean()) for _, c in sets]}")
    # the prediction prices a fallback particle at the fallback's per-particle
    # time over a whole set, which a small flagged subset exceeds (gathering its
    # rows, fixed launches): the few cheapest predicted rules are run as they
    # are on the held-out set and the fastest measured is kept
    def measure(rules):
        rules.sort(key=lambda r: r["pred_ms_per_particle"])
        top = None
        for r in rules[:n_measure]:
            if "measured_ms_per_particle" not in r:
                args = (net, 0, r["n_rot"], r["n_trans"], 0.0, r["spread_max"], r["tail_max"],
                        r["score_drop"], fallback_setup(ctx, r["fallback"]))
                ctx.expect_learned(*args)   # untimed: may grow buffers
                ctx.sync()
                ts = []
                for _ in range(3):   # the fastest of three calls
                    t0 = time.perf_counter()
                    ctx.expect_learned(*args)
                    ctx.sync()
                    ts.append(time.perf_counter() - t0)
                r["measured_ms_per_particle"] = 1e3 * min(ts) / P
            if top is None or r["measured_ms_per_particle"] < top["measured_ms_per_particle"]:
                top = r
        if top is not None:
            top = dict(top)
            top["rules_measured"] = [{k: r[k] for k in (
                "n_rot", "n_trans", "spread_max", "tail_max", "score_drop", "fallback",
                "pred_ms_per_particle", "measured_ms_per_particle")} for r in rules[:n_measure]]
        return top

    best = measure(ok_rules)
    if robust:
        # the fastest rule for ctx's map alone, for comparison
        best0 = measure(ok_rules0)
        if best is not None and best0 is not None:
            best["same_map_rule"] = {k: best0[k] for k in (
                "n_rot", "n_trans", "spread_max", "tail_max", "score_drop", "fallback",
                "cal_accuracy", "cal_accuracy_robust", "cal_fallback",
                "measured_ms_per_particle")}
    ctx.set_learned_fallback_selection(0, 0)
    if best is None:   # nothing meets the target: every particle falls back
        best = dict(n_rot=candidates[0][0], n_trans=candidates[0][1], spread_max=0.5,
                    tail_max=0.0, score_drop=0.0, fallback=fallbacks[0], cal_accuracy=float("nan"),
                    cal_fallback=1.0, cal_accuracy_no_fallback=float("nan"),
                    pred_ms_per_particle=full0[fallbacks[0]][0])
    best["fallback_agreement"] = [{k: float(np.mean(v[1])) for k, v in full.items()}
                                  for full, _ in sets]
    best["cal_particles"] = P
    best["cal_sets"] = 1 + len(robust)
    best["target"] = target
    return best
