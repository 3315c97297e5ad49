"""Benchmark: plain wavefront path tracer on one MI355X, 1M-triangle terrain + icosphere scene.

Workload (one step = one full frame): seeded fBm terrain (1000x500 grid, 997,002 triangles) plus
200 icospheres (subdivision 3, 256,000 triangles), 50% Lambert / 50% GGX conductor, 16 emissive
quads, 1920x1080 at 64 spp, max depth 8, Russian roulette after depth 3, FP32 film. Scene data is
resident in HBM before the timed region.

Prints one JSON line: Mpaths/s, the extension-traversal kernel's roofline (algorithmic bytes per
launch / mean launch time, HIP events on the library's stream), and the CPU oracle timed on this
host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_host():
    """This process's CPUs (as nproc counts them: its affinity set), the machine's logical CPUs and
    the CPU model string."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return nproc, os.cpu_count() or nproc, model


def cpu_baseline(sc, wide, c1, c1_wide, threads, budget_s=10.0):
    """CPU oracle (the C++ restatement of the reference CPU path, 8x8 tiles over std::thread
    workers with samples looped per pixel, as FixedSampler.cpp:73-94 and Parallel.cpp:475-491 run
    it) traversing the same wide BVH the GPU renders with (downloaded from the library), on one
    centred crop of the same frame with `threads` workers and with one, and the whole Cornell-box
    frame (256x256, 16 spp)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleScene

    nproc, ncpu, model = _cpu_host()
    t0 = time.perf_counter()
    o = OracleScene(sc)
    o.use_wide_bvh(*wide)
    log(f"cpu oracle: scene loaded in {time.perf_counter() - t0:.1f}s")
    w, h = sc["width"], sc["height"]

    def run(cw, ch, spp, nthreads):
        rect = ((w - cw) // 2, (h - ch) // 2, cw, ch)
        t = time.perf_counter()
        o.render(rect=rect, spp_range=(0, spp), threads=nthreads, film=True)
        return time.perf_counter() - t

    # one centred crop for both figures, sized so one thread renders it at 1 spp in budget / 3
    dt = run(96, 54, 1, 1)
    scale = min(1.0, (96 * 54 / max(dt, 1e-6) * budget_s / 3 / (w * h)) ** 0.5)
    cw, ch = min(w, max(64, int(w * scale) // 8 * 8)), min(h, max(36, int(h * scale) // 8 * 8))
    dt1 = run(cw, ch, 1, 1)
    dtn = run(cw, ch, 1, threads)
    spp = max(1, min(64, int(budget_s / max(dtn, 1e-6))))  # the multithreaded figure over about budget_s
    if spp > 1:
        dtn = run(cw, ch, spp, threads)
    mp1, mp = cw * ch / dt1 / 1e6, cw * ch * spp / dtn / 1e6
    sample1 = f"{cw}x{ch} crop at 1 spp ({cw * ch} paths, {dt1:.1f} s)"
    sample = f"{cw}x{ch} crop at {spp} spp ({cw * ch * spp} paths, {dtn:.1f} s)"
    oc = OracleScene(c1)
    oc.use_wide_bvh(*c1_wide)
    t = time.perf_counter()
    oc.render(threads=threads, film=True)
    dt_c1 = time.perf_counter() - t
    c1_paths = c1["width"] * c1["height"] * c1["spp"]
    width = wide[0]
    return {"value": round(mp, 3), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++ restatement, traversing the GPU's own BVH{width}) on a {sample} of the same "
                      f"1920x1080 frame, {threads} threads",
            "nproc": nproc, "logical_cpus": ncpu, "cpu_model": model,
            "threads_note": "worker threads capped at 16, the host-CPU share of one GPU on this machine",
            "one_thread": {"value": round(mp1, 3), "unit": "Mpaths/s", "sample": sample1},
            "cornell_256x256_16spp": {"value": round(c1_paths / dt_c1 / 1e6, 3), "unit": "Mpaths/s",
                                      "cores": threads, "seconds": round(dt_c1, 2)}}


def measured_traffic(bvh_width, wavefront_paths):
    """HBM bytes per kExtend launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
    this workload (profiles/), when they were taken at the same BVH width and pool size."""
    path = os.path.join(ROOT, "profiles", "kextend_hbm_traffic_terrain_spheres_1080p64.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    if t["bvh_width"] != bvh_width or t["wavefront_paths"] != wavefront_paths:
        return None
    return t["hbm_bytes_per_launch"]


def _img(f):
    import numpy as np
    return f[..., :3] / np.maximum(f[..., 3:], 1e-30)


def _rel_mse(a, b):
    import numpy as np
    return float(np.mean((a - b) ** 2 / (b * b + 0.01)))


def time_to_relmse(sc, film, ref, render, target=0.01, max_level=14):
    """Progressive render with doubling sample counts (1, 2, 4, ... 2^max_level spp accumulated in
    one film); returns (seconds to reach `target` relMSE, interpolated in log-log between levels,
    or None, and the whole trace of (spp, seconds, relMSE))."""
    import math
    film.clear()
    t, prev, trace, hit = 0.0, None, [], None
    for k in range(max_level + 1):
        s0, s1 = (0, 1) if k == 0 else (1 << (k - 1), 1 << k)
        st = render(pg_params(sc, (s0, s1)))
        t += st["ms_total"] / 1e3
        e = _rel_mse(_img(film.download()), ref)
        trace.append((s1, round(t, 4), round(e, 6)))
        if hit is None and e <= target:
            if prev is None or prev[1] <= e:
                hit = t
            else:
                f = (math.log(target) - math.log(prev[1])) / (math.log(e) - math.log(prev[1]))
                hit = math.exp(math.log(prev[0]) + f * (math.log(t) - math.log(prev[0])))
        prev = (t, e)
    return hit, trace


GUIDE_FLOP_PER_SAMPLE = 2 * (139 * 64 + 3 * 64 * 64 + 64 * 40)  # unpadded MLP forward: 47,488
MFMA_BF16_PEAK_TFLOPS = 2500.0


def guide_roofline(st):
    """Guide inference of one guided render against MI355X's dense BF16 matrix-core peak: one
    forward per extension-queue entry (= extension ray), algorithmic FLOPs over the event-timed
    inference stage."""
    if not st.get("ms_guide"):
        return None
    tflops = st["extension_rays"] * GUIDE_FLOP_PER_SAMPLE / (st["ms_guide"] / 1e3) / 1e12
    return {"bound": "mfma", "kernel": "kGuideInfer", "achieved": round(tflops, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4), "samples": st["extension_rays"],
            "ms": round(st["ms_guide"], 3), "flop_per_sample": GUIDE_FLOP_PER_SAMPLE}


def guided_profile():
    """Profiling workload of the guided path (for rocprofv3): the 256x256 hard room, 16 online
    training passes of 16 spp, then one 1024-spp render with the trained guide frozen and one plain
    1024-spp render; prints one JSON line with their stage times."""
    import piper2_amd as pg
    from piper2_amd import scenes
    dev = pg.Device(0)
    sc = scenes.hard_room(seed=0, width=256, height=256)
    scene = pg.Scene(dev, sc)
    film = pg.Film(dev, 256, 256)
    g = pg.Guide(dev, seed=1)
    base = 1 << 17  # training samples: disjoint from the measured renders' [0, 1024)
    for p in range(16):
        scene.render_guided(g, pg_params(sc, (base + 16 * p, base + 16 * (p + 1))), None)
    frozen = pg.Guide(dev, seed=1, train_records=0)
    frozen.set_params(g.get_params())
    guided = scene.render_guided(frozen, pg_params(sc, (0, 1024)), film)
    plain = scene.render(pg_params(sc, (0, 1024)), film)
    keys = ("ms_total", "ms_raygen", "ms_extend", "ms_guide", "ms_shade", "ms_shadow", "ms_film", "extension_rays")
    print(json.dumps({"workload": "hard_room_256x256_1024spp", "guided": {k: guided[k] for k in keys},
                      "plain": {k: plain[k] for k in keys}, "inference_roofline": guide_roofline(guided)}), flush=True)


def inference_profile(n=1 << 22, calls=4):
    """Profiling workload of the guide network alone (for rocprofv3): n samples of seeded random
    inputs (position in [0,1]^3, unit normal and direction) through pg_guide_infer, `calls` times,
    with a perturbed random network; prints the host time per call (with its copies)."""
    import numpy as np

    import piper2_amd as pg
    rng = np.random.default_rng(0)
    pos = rng.random((n, 3), dtype=np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3)).astype(np.float32)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    dev = pg.Device(0)
    g = pg.Guide(dev, seed=3)
    th = g.get_params()
    g.set_params(th + rng.normal(size=th.shape).astype(np.float32) * 0.05)
    times = []
    for _ in range(calls):
        t0 = time.perf_counter()
        mix = g.infer(pos, nrm, wo)
        times.append(time.perf_counter() - t0)
    assert np.isfinite(mix).all()
    print(json.dumps({"workload": f"guide_inference_{n}_samples", "calls": calls,
                      "host_ms_per_call": [round(t * 1e3, 2) for t in times],
                      "flop_per_sample": GUIDE_FLOP_PER_SAMPLE}), flush=True)


def equal_time(trace_a, trace_b):
    """relMSE of trace b at the times of trace a's levels (log-log interpolation between b's levels),
    as (seconds, relMSE a, relMSE b) for every time both traces cover."""
    import math
    out = []
    for (_, t, e) in trace_a:
        for (_, t0, e0), (_, t1, e1) in zip(trace_b, trace_b[1:]):
            if t0 <= t <= t1 and min(e0, e1, t0) > 0:
                f = (math.log(t) - math.log(t0)) / (math.log(t1) - math.log(t0))
                out.append((round(t, 4), e, round(math.exp(math.log(e0) + f * (math.log(e1) - math.log(e0))), 6)))
                break
    return out


SPP_DECLARED = 1 << 18  # sample index space: [0, 2^16) measured renders, [2^16, 2^17) reference, 2^17+ training


def pg_params(sc, spp_range, flags=0):
    import piper2_amd as pg
    return pg.render_params(sc, spp=SPP_DECLARED, spp_range=spp_range, flags=flags)


def guided_vs_plain(dev, sc, ref_spp, label, train_passes=64, train_spp=16, max_level=14, learned=False):
    """Guided (online-trained, one-sample MIS at 0.5) against the plain tracer on one scene.

    The reference R is an FP64-film render of ref_spp samples by the guided estimator with a guide
    trained beforehand on samples no measured render uses, then frozen (the estimator is unbiased
    for any guide; a trained one reaches a given error in fewer samples than the plain tracer). A
    plain FP64 render of the same samples is compared with it as a cross-check. Reported: time to
    1% relMSE from a fresh guide trained online during the progressive render (training included),
    the plain tracer's time, and both estimators' relMSE at equal samples with the trained guide.
    learned=True adds the same two measurements for a guide whose selection probability is learned
    per vertex (mis_fraction < 0) instead of fixed at 0.5."""
    import piper2_amd as pg
    scene = pg.Scene(dev, sc)
    film = pg.Film(dev, sc["width"], sc["height"])
    R = 1 << 16
    chunk = max(1, min(4096, (1 << 31) // (sc["width"] * sc["height"])))  # < 2^32 paths per call
    t0 = time.perf_counter()
    trained = pg.Guide(dev, seed=1)
    for p in range(train_passes):
        scene.render_guided(trained, pg_params(sc, (2 * R + train_spp * p, 2 * R + train_spp * (p + 1))), None)
    frozen = pg.Guide(dev, seed=1, train_records=0)
    frozen.set_params(trained.get_params())
    trained.close()
    for s0 in range(R, R + ref_spp, chunk):
        scene.render_guided(frozen, pg_params(sc, (s0, min(s0 + chunk, R + ref_spp)), pg.PG_FILM_FP64), film)
    ref = _img(film.download(fp64=True))
    film.clear()
    for s0 in range(R, R + ref_spp, chunk):
        scene.render(pg_params(sc, (s0, min(s0 + chunk, R + ref_spp)), pg.PG_FILM_FP64), film)
    ref_plain = _img(film.download(fp64=True))
    log(f"{label}: {ref_spp}-spp FP64 references in {time.perf_counter() - t0:.1f}s; "
        f"relMSE(plain, guided) = {_rel_mse(ref_plain, ref):.5f}, mean {ref.mean():.5f}")
    equal_spp = {}
    for spp in (256, 1024):
        film.clear()
        scene.render(pg_params(sc, (0, spp)), film)
        e_plain = _rel_mse(_img(film.download()), ref)
        film.clear()
        st = scene.render_guided(frozen, pg_params(sc, (0, spp)), film)
        equal_spp[str(spp)] = {"plain": round(e_plain, 5), "trained_guide": round(_rel_mse(_img(film.download()), ref), 5)}
    inference = guide_roofline(st)
    frozen.close()
    t_plain, tr_plain = time_to_relmse(sc, film, ref, lambda p: scene.render(p, film), max_level=max_level)
    guide = pg.Guide(dev, seed=1, lr=1e-2, mis_fraction=0.5, train_batch=1 << 14, train_records=1 << 18)
    t_guided, tr_guided = time_to_relmse(sc, film, ref, lambda p: scene.render_guided(guide, p, film),
                                         max_level=max_level)
    log(f"{label}: plain {t_plain} s {tr_plain}")
    log(f"{label}: guided {t_guided} s {tr_guided}")
    out = {"plain_s": t_plain and round(t_plain, 4), "guided_s": t_guided and round(t_guided, 4),
           "reference": f"{ref_spp}-spp FP64 guided render (trained guide, frozen)",
           "reference_vs_plain_fp64_relmse": round(_rel_mse(ref_plain, ref), 5),
           "reference_mean": round(float(ref.mean()), 5),
           "resolution": [sc["width"], sc["height"]], "relmse_at_equal_spp": equal_spp,
           "inference_roofline": inference,
           "plain_trace_spp_s_relmse": tr_plain, "guided_trace_spp_s_relmse": tr_guided}
    if t_plain and t_guided:
        out["guided_speedup"] = round(t_plain / t_guided, 3)
    if learned:  # the selection probability learned per vertex
        tl = pg.Guide(dev, seed=1, mis_fraction=-1.0)
        for p in range(train_passes):
            scene.render_guided(tl, pg_params(sc, (2 * R + train_spp * p, 2 * R + train_spp * (p + 1))), None)
        fl = pg.Guide(dev, seed=1, train_records=0, mis_fraction=-1.0)
        fl.set_params(tl.get_params())
        tl.close()
        eq_l = {}
        for spp in (256, 1024):
            film.clear()
            st_l = scene.render_guided(fl, pg_params(sc, (0, spp)), film)
            eq_l[str(spp)] = {"relmse": round(_rel_mse(_img(film.download()), ref), 5), "ms": round(st_l["ms_total"], 2)}
        fl.close()
        gl = pg.Guide(dev, seed=1, mis_fraction=-1.0)
        t_l, tr_l = time_to_relmse(sc, film, ref, lambda p: scene.render_guided(gl, p, film), max_level=max_level)
        gl.close()
        out["learned_selection"] = {"time_to_1pct_s": t_l and round(t_l, 4),
                                    "speedup_vs_plain": round(t_plain / t_l, 3) if t_plain and t_l else None,
                                    "relmse_at_equal_spp_trained": eq_l, "trace_spp_s_relmse": tr_l}
        log(f"{label}: learned selection {t_l} s, equal spp {eq_l}")
    eq = equal_time(tr_plain, tr_guided)
    if eq:
        t, e_plain, e_guided = eq[-1]
        out["equal_time"] = {"seconds": t, "plain_relmse": e_plain, "guided_relmse": e_guided,
                             "plain_over_guided": round(e_plain / e_guided, 3)}
    guide.close()
    film.close()
    scene.close()
    return out


def guided_batch(dev, n, bvh_width, crop=64, ref_spp=1 << 16, train_scale=8, train_passes=32, train_spp=16,
                 budget_s=12.0, max_level=16, pass_spp=16, which=("plain", "guided", "guided_shared", "guided_4x_training")):
    """Hard-indirect batch at its stated size: n seeded hard rooms at their own resolutions (512x512
    to 1920x1080, 499k triangles each, depth 12), each batched call one persistent launch.
    Throughput: one pass_spp pass of all rooms, plain batch vs guided batch (trained guides, frozen).
    Convergence: progressive renders of the whole images (one batch call per doubling level, all
    samples accumulated) by each estimator; after every level the relMSE of a fixed crop x crop
    window of each image (the window at the 90th percentile of window brightness in a 256-spp
    preview at 1/train_scale resolution: lit, but not the light) against a ref_spp-sample FP64
    reference of that window, rendered
    by frozen guides trained beforehand on 1/train_scale-resolution renders of the same rooms
    (disjoint samples; any guide leaves the estimator unbiased). Estimators: the plain tracer; fresh
    per-room guides trained online during the render; one fresh guide shared by all rooms (the rooms
    are seeded variants of one family, so it trains on all of them); per-room guides taking 4x the
    Adam steps per call. Levels stop once every room is below 1% or after budget_s seconds of render
    time (budget_s = 0: no limit). A room's time to 1% is measured (log-log between levels) or, if its
    last level is above 1%, extrapolated as t x relMSE / 0.01 (relMSE falling as 1/time) and marked.
    The error breakdown at the last level: median relMSE, and the share of squared error in the worst
    1% of crop pixels."""
    import math
    import numpy as np
    import piper2_amd as pg
    from piper2_amd import scenes
    t0 = time.perf_counter()
    full = [scenes.hard_room_variant(k) for k in range(n)]
    small = [scenes.hard_room_variant(k, scale=train_scale) for k in range(n)]
    objs = [pg.Scene(dev, sc, bvh_width=bvh_width) for sc in full]  # small differs only in resolution
    log(f"guided batch: {n} rooms, {sum(sc['indices'].shape[0] for sc in full)} triangles, built in "
        f"{time.perf_counter() - t0:.1f}s")
    R = 1 << 16

    def params(scs, rng, flags=0, rects=None):
        return [pg.render_params(sc, spp=1 << 18, spp_range=rng, flags=flags, rect=None if rects is None else rects[i])
                for i, sc in enumerate(scs)]

    guides = [pg.Guide(dev, seed=100 + k) for k in range(n)]
    t0 = time.perf_counter()
    for p in range(train_passes):
        pg.render_batch(dev, objs, params(small, (2 * R + train_spp * p, 2 * R + train_spp * (p + 1))), None,
                        guides=guides)
    frozen = []
    for g in guides:
        f = pg.Guide(dev, seed=0, train_records=0)
        f.set_params(g.get_params())
        frozen.append(f)
        g.close()
    fsmall = [pg.Film(dev, sc["width"], sc["height"]) for sc in small]
    pg.render_batch(dev, objs, params(small, (3 * R, 3 * R + 256)), fsmall, guides=frozen)
    rects, b = [], crop // train_scale
    for f, sc in zip(fsmall, full):
        img = _img(f.download()).mean(-1)
        h, w = img.shape
        sums = img[:h - h % b, :w - w % b].reshape(h // b, b, w // b, b).mean((1, 3))
        flat = sums.reshape(-1)
        k = int(np.argsort(flat)[int(0.9 * (flat.size - 1))])
        y, x = divmod(k, sums.shape[1])
        rects.append((min(x * crop, sc["width"] - crop), min(y * crop, sc["height"] - crop), crop, crop))
        f.close()
    ffull = [pg.Film(dev, sc["width"], sc["height"]) for sc in full]
    for s0 in range(R, R + ref_spp, 8192):
        pg.render_batch(dev, objs, params(full, (s0, min(s0 + 8192, R + ref_spp)), pg.PG_FILM_FP64, rects), ffull,
                        guides=frozen)
    refs = [_img(f.download(fp64=True))[y:y + crop, x:x + crop] for f, (x, y, _, _) in zip(ffull, rects)]
    log(f"guided batch: guides trained and {ref_spp}-spp {crop}x{crop} crop references in "
        f"{time.perf_counter() - t0:.1f}s")

    def progressive(gs, label):
        for f in ffull:
            f.clear()
        t, traces = 0.0, [[] for _ in range(n)]
        for k in range(max_level + 1):
            rng = (0, 1) if k == 0 else (1 << (k - 1), 1 << k)
            st, _ = pg.render_batch(dev, objs, params(full, rng), ffull, guides=gs)
            t += st["ms_total"] / 1e3
            errs = []
            for i, (f, ref, (x, y, _, _)) in enumerate(zip(ffull, refs, rects)):
                img = _img(f.download())[y:y + crop, x:x + crop]
                errs.append(_rel_mse(img, ref))
                traces[i].append((rng[1], t, errs[-1], img))
            log(f"guided batch: {label} level {k} ({rng[1]} spp) at {t:.3f} s, median crop relMSE "
                f"{float(np.median(errs)):.4f}")
            if max(errs) <= 0.01 or (budget_s and t >= budget_s):
                break
        hit, marked = [], []
        for tr in traces:
            h = None
            for (s_a, t_a, e_a, _), (s_b, t_b, e_b, _) in zip([(0, 0.0, None, None)] + tr, tr):
                if e_b <= 0.01:
                    if e_a is None or e_a <= 0.01:
                        h = t_b
                    else:
                        fr = (math.log(0.01) - math.log(e_a)) / (math.log(e_b) - math.log(e_a))
                        h = math.exp(math.log(t_a) + fr * (math.log(t_b) - math.log(t_a)))
                    break
            if h is None:  # extrapolated: relMSE ~ 1 / time beyond the last level
                h = tr[-1][1] * tr[-1][2] / 0.01
            hit.append(h)
            marked.append(tr[-1][2] > 0.01)
        last = []
        for tr, ref in zip(traces, refs):
            img = tr[-1][3]
            sq = ((img - ref) ** 2 / (ref * ref + 0.01)).sum(-1).reshape(-1)
            top = np.sort(sq)[::-1][:max(1, sq.size // 100)]
            last.append((tr[-1][2], float(top.sum() / max(sq.sum(), 1e-30))))
        out = {"s_to_1pct": [round(v, 3) for v in hit], "extrapolated": marked, "seconds": round(t, 3),
               "spp": traces[0][-1][0],
               "median_relmse_last": round(float(np.median([e for e, _ in last])), 5),
               "worst1pct_error_share_median": round(float(np.median([w for _, w in last])), 3),
               "trace_median": [[tr0[0], round(tr0[1], 3), round(float(np.median([tr[j][2] for tr in traces])), 5)]
                                for j, tr0 in enumerate(traces[0])]}
        return out, traces

    results, all_traces = {}, {}
    estimators = [("plain", lambda: None),
                  ("guided", lambda: [pg.Guide(dev, seed=200 + k) for k in range(n)]),
                  ("guided_shared", lambda: [pg.Guide(dev, seed=300)] * n),
                  ("guided_4x_training", lambda: [pg.Guide(dev, seed=400 + k, train_records=1 << 20) for k in range(n)])]
    for name, make in estimators:
        if name not in which:
            continue
        gs = make()
        results[name], all_traces[name] = progressive(gs, name)
        for g in set(gs or []):
            g.close()
    plain = results["plain"]["s_to_1pct"] if "plain" in results else None
    for name in results:
        if name != "plain" and plain is not None:
            sp = [a / b for a, b in zip(plain, results[name]["s_to_1pct"])]
            results[name]["speedup_per_room"] = [round(v, 3) for v in sp]
            results[name]["median_speedup"] = round(float(np.median(sp)), 3)
            # relMSE at equal render time: the plain trace's last level against this estimator's trace
            eq = []
            for tp, tg in zip(all_traces["plain"], all_traces[name]):
                a = equal_time([(s_, t_, e_) for s_, t_, e_, _ in tp], [(s_, t_, e_) for s_, t_, e_, _ in tg])
                if a:
                    eq.append(a[-1][1] / a[-1][2])
            if eq:
                results[name]["equal_time_plain_over_guided_median"] = round(float(np.median(eq)), 3)
    out = {"rooms": n, "resolutions": [[sc["width"], sc["height"]] for sc in full], "depth": 12,
           "crop": f"{crop}x{crop}, the window at the 90th percentile of brightness of a 1/{train_scale}-resolution preview",
           "crops": [list(r[:2]) for r in rects],
           "reference": f"{ref_spp}-spp FP64 of the crop, frozen guides trained at 1/{train_scale} resolution",
           "budget_s": budget_s, "estimators": results}
    # throughput at the rooms' own resolutions
    tp = {}
    for name, gs in (("plain", None), ("guided", frozen)):
        st, lat = pg.render_batch(dev, objs, params(full, (0, pass_spp)), ffull, guides=gs)  # warm
        st, lat = pg.render_batch(dev, objs, params(full, (pass_spp, 2 * pass_spp)), ffull, guides=gs)
        tp[name] = {"mpaths_per_s": round(st["paths"] / st["ms_total"] / 1e3, 1), "ms": round(st["ms_total"], 2),
                    "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
                    "latency_ms_p99": round(float(np.percentile(lat, 99)), 2)}
    out["pass_spp"] = pass_spp
    out["throughput"] = tp
    log(f"guided batch: { {k: {kk: v[kk] for kk in ('median_speedup', 'seconds', 'spp', 'median_relmse_last') if kk in v} for k, v in results.items()} }")
    for o in ffull + objs + frozen:
        o.close()
    return out


def batched(dev, n_scenes, steps, bvh_width, wavefront_paths):
    """Batched workload: n seeded scenes (log-uniform 10k-2M triangles, mixed resolutions, spp
    4/16/64, depth 8) in one pg_render_batch call; aggregate Mrays/s and per-scene latency
    (device time from the batch's start until the scene's film is complete). The same scenes
    rendered one pg_render call after another are timed beside it."""
    import numpy as np

    import piper2_amd as pg
    from piper2_amd import scenes
    t0 = time.perf_counter()
    scs = [scenes.batch_scene(i) for i in range(n_scenes)]
    log(f"batch: {n_scenes} scenes generated in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    objs = [pg.Scene(dev, sc, bvh_width=bvh_width) for sc in scs]
    log(f"batch: BVH{bvh_width}s built in {time.perf_counter() - t0:.1f}s")
    params = [pg.render_params(sc, flags=pg.PG_NO_STAGE_TIMING, wavefront_paths=wavefront_paths) for sc in scs]
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc in scs]
    st, lat = pg.render_batch(dev, objs, params, films)  # warmup
    log(f"batch warmup: {st['ms_total']:.0f} ms")
    rays = paths = 0
    lats = []
    t_start = time.perf_counter()
    for i in range(steps):
        for f in films:
            f.clear()
        st, lat = pg.render_batch(dev, objs, params, films)
        rays += st["extension_rays"] + st["shadow_rays"] + st["mis_rays"]
        paths += st["paths"]
        lats.append(lat)
        log(f"batch step {i}: {st['ms_total']:.0f} ms, {st['paths'] / st['ms_total'] / 1e3:.1f} Mpaths/s")
    elapsed = time.perf_counter() - t_start
    lats = np.concatenate(lats)
    t_seq = 0.0
    for s, p, f in zip(objs, params, films):  # the same work as one render call per scene
        f.clear()
        t = time.perf_counter()
        s.render(p, f)
        t_seq += time.perf_counter() - t
    out = {"scenes": n_scenes, "triangles_total": int(sum(sc["indices"].shape[0] for sc in scs)),
           "bvh_width": bvh_width, "wavefront_paths": wavefront_paths,
           "paths_per_batch": int(paths // steps), "steps": steps,
           "ms_per_batch": round(elapsed * 1e3 / steps, 1),
           "mrays_per_s": round(rays / elapsed / 1e6, 1), "mpaths_per_s": round(paths / elapsed / 1e6, 1),
           "latency_ms_p50": round(float(np.percentile(lats, 50)), 1),
           "latency_ms_p99": round(float(np.percentile(lats, 99)), 1),
           "sequential_ms": round(t_seq * 1e3, 1),
           "speedup_vs_sequential": round(t_seq / (elapsed / steps), 3)}
    for f in films:
        f.close()
    for s in objs:
        s.close()
    return out


def batched_small(dev, n_scenes, steps):
    """Small-scene batch: n seeded Cornell boxes (scenes.small_batch_scene: 64x64 to 256x256, 1 to
    16 spp, depth 5), each far below a path pool's capacity, in one pg_render_batch call against
    one pg_render call per scene (each call waits for its own film). Aggregate Mpaths/s and
    per-scene latency (device time from the batch's start until the scene's film is complete)."""
    import numpy as np

    import piper2_amd as pg
    from piper2_amd import scenes
    scs = [scenes.small_batch_scene(i) for i in range(n_scenes)]
    objs = [pg.Scene(dev, sc, bvh_width=4) for sc in scs]
    params = [pg.render_params(sc, flags=pg.PG_NO_STAGE_TIMING) for sc in scs]
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc in scs]
    pg.render_batch(dev, objs, params, films)  # warmup
    paths = rays = 0
    lats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st, lat = pg.render_batch(dev, objs, params, films)
        paths += st["paths"]
        rays += st["extension_rays"] + st["shadow_rays"] + st["mis_rays"]
        lats.append(lat)
    elapsed = (time.perf_counter() - t0) / steps
    for s, p, f in zip(objs, params, films):  # warm the per-scene path
        s.render(p, f)
    t0 = time.perf_counter()
    for s, p, f in zip(objs, params, films):
        s.render(p, f)
    t_seq = time.perf_counter() - t0
    lats = np.concatenate(lats)
    out = {"scenes": n_scenes, "paths_per_batch": int(paths // steps), "steps": steps,
           "ms_per_batch": round(elapsed * 1e3, 2), "mpaths_per_s": round(paths / steps / elapsed / 1e6, 1),
           "mrays_per_s": round(rays / steps / elapsed / 1e6, 1),
           "latency_ms_p50": round(float(np.percentile(lats, 50)), 2),
           "latency_ms_p99": round(float(np.percentile(lats, 99)), 2),
           "sequential_ms": round(t_seq * 1e3, 2), "sequential_mpaths_per_s": round(paths / steps / t_seq / 1e6, 1),
           "speedup_vs_sequential": round(t_seq / elapsed, 3)}
    for f in films:
        f.close()
    for s in objs:
        s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--wavefront-paths", type=int, default=0, help="path-pool capacity (0: autotuned)")
    ap.add_argument("--bvh-width", type=int, default=0, help="4 or 8 (0: autotuned)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-guided", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay timing")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--small-batch", type=int, default=256, help="scenes of the small-scene batch leg (0: skip it)")
    ap.add_argument("--guided-batch", type=int, default=16, help="hard rooms of the guided batch leg (0: skip it)")
    ap.add_argument("--guided-batch-budget", type=float, default=12.0,
                    help="render seconds per estimator in the guided batch convergence (0: until 1%%)")
    ap.add_argument("--guided-batch-estimators", default="plain,guided,guided_shared,guided_4x_training",
                    help="estimators of the guided batch convergence (comma-separated)")
    ap.add_argument("--only-guided-batch", action="store_true", help="run the guided batch leg alone")
    ap.add_argument("--guided-profile", action="store_true", help="only the guided profiling workload (for rocprofv3)")
    ap.add_argument("--inference-profile", action="store_true",
                    help="only the guide-network inference workload (for rocprofv3)")
    ap.add_argument("--ref-spp", type=int, default=1 << 16, help="reference samples per pixel of the guided comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a render would shard by film tiles "
              "or by scenes, with no exchange between GPUs). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.inference_profile:
        inference_profile()
        return
    if args.guided_profile:
        guided_profile()
        return
    import piper2_amd as pg
    from piper2_amd import scenes

    if args.only_guided_batch:  # the hard-room batch alone (long convergence runs)
        dev = pg.Device(0)
        res = guided_batch(dev, args.guided_batch, args.bvh_width or 4, budget_s=args.guided_batch_budget,
                           which=tuple(args.guided_batch_estimators.split(",")))
        print(json.dumps({"guided_batch_hard_rooms": res}), flush=True)
        return
    t0 = time.perf_counter()
    sc = scenes.terrain_spheres(spp=args.spp)
    log(f"scene generated in {time.perf_counter() - t0:.1f}s: {sc['indices'].shape[0]} triangles")
    dev = pg.Device(0)
    t0 = time.perf_counter()
    scene = pg.Scene(dev, sc, bvh_width=args.bvh_width or 8)
    log(f"device {dev.platform}; BVH built in {time.perf_counter() - t0:.1f}s")
    tune = None
    if not (args.bvh_width and args.wavefront_paths):  # autotuner: BVH width x path-pool capacity
        t0 = time.perf_counter()
        tune = scene.autotune(pg.render_params(sc))
        log(f"autotune in {time.perf_counter() - t0:.1f}s: {tune}")
        if args.bvh_width and args.bvh_width != tune["bvh_width"]:
            scene.build_bvh(width=args.bvh_width)
        args.wavefront_paths = args.wavefront_paths or tune["wavefront_paths"]
    info = scene.bvh_info()
    log(f"BVH{info['width']}: {info}; path pool {args.wavefront_paths}")
    film = pg.Film(dev, sc["width"], sc["height"])
    params = pg.render_params(sc, wavefront_paths=args.wavefront_paths)
    copy_gbps = dev.copy_bandwidth(1 << 30, 3)  # kCopy: 1 GiB read + 1 GiB written per launch
    log(f"plain copy kernel: {copy_gbps:.0f} GB/s")

    # traversal work per extension ray on a 4-sample slice (counting perturbs timing: untimed)
    cnt = pg.render_params(sc, spp_range=(0, 4), flags=pg.PG_COUNT_TRAVERSAL,
                           wavefront_paths=args.wavefront_paths)
    st = scene.render(cnt, None)
    nodes_per_ray = st["nodes_visited"] / st["extension_rays"]
    tris_per_ray = st["tris_tested"] / st["extension_rays"]
    log(f"traversal: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} triangles per extension ray")

    for i in range(args.warmup):
        film.clear()
        st = scene.render(params, film)
        log(f"warmup {i}: {st['ms_total']:.1f} ms")
    totals = {k: 0.0 for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_mis", "ms_film", "ms_raygen")}
    rays = {"extension_rays": 0, "shadow_rays": 0, "mis_rays": 0, "extend_launches": 0, "paths": 0}
    t_start = time.perf_counter()
    for i in range(args.steps):
        film.clear()
        st = scene.render(params, film)
        for k in totals:
            totals[k] += st[k]
        for k in rays:
            rays[k] += st[k]
        log(f"step {i}: {st['ms_total']:.1f} ms, {st['paths'] / st['ms_total'] / 1e3:.1f} Mpaths/s")
    elapsed = time.perf_counter() - t_start
    graph_ms = None
    if not args.no_graph:  # the same frames with the bounce loop replayed from a captured graph
        gparams = pg.render_params(sc, wavefront_paths=args.wavefront_paths, flags=pg.PG_USE_GRAPH)
        film.clear()
        scene.render(gparams, film)
        t_graph = time.perf_counter()
        for i in range(args.steps):
            film.clear()
            scene.render(gparams, film)
        graph_ms = round((time.perf_counter() - t_graph) * 1e3 / args.steps, 2)
        log(f"graph replay: {graph_ms} ms per frame")
    # the Cornell-box frame (256x256, 16 spp, depth 5): whole-frame throughput on the GPU
    c1 = scenes.cornell()
    c1_scene = pg.Scene(dev, c1, bvh_width=info["width"])
    c1_film = pg.Film(dev, c1["width"], c1["height"])
    c1_params = pg.render_params(c1, flags=pg.PG_NO_STAGE_TIMING)
    c1_scene.render(c1_params, c1_film)
    t_c1 = time.perf_counter()
    for _ in range(10):
        c1_film.clear()
        c1_scene.render(c1_params, c1_film)
    c1_ms = (time.perf_counter() - t_c1) * 1e3 / 10
    c1_gpu = {"mpaths_per_s": round(c1["width"] * c1["height"] * c1["spp"] / c1_ms / 1e3, 1), "ms": round(c1_ms, 3)}
    log(f"cornell 256x256x16: {c1_gpu}")
    c1_film.close()
    c1_scene.close()
    ms_per_step = elapsed * 1e3 / args.steps
    paths = sc["width"] * sc["height"] * args.spp * args.steps
    mpaths = paths / elapsed / 1e6
    total_rays = rays["extension_rays"] + rays["shadow_rays"] + rays["mis_rays"]

    bytes_per_ray = 32 + 20 + nodes_per_ray * info["node_bytes"] + tris_per_ray * info["tri_bytes"]
    launches = max(rays["extend_launches"], 1)
    bytes_per_launch = rays["extension_rays"] / launches * bytes_per_ray
    mean_launch_s = totals["ms_extend"] / launches / 1e3
    achieved = bytes_per_launch / mean_launch_s / 1e9

    result = {
        "metric": "Mpaths/s (plain GPU wavefront path tracer)",
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "terrain_spheres_1920x1080_64spp_depth8_rr3", "triangles": int(sc["indices"].shape[0]),
                   "lights": 16, "spp": args.spp, "bvh_width": info["width"],
                   "wavefront_paths": args.wavefront_paths},
        "autotune": tune,
        "graph_ms_per_step": graph_ms,
        "cornell_256x256_16spp": c1_gpu,
        "mrays_per_s": round(total_rays / elapsed / 1e6, 2),
        "copy_kernel_gbps": round(copy_gbps, 1),
        "rays_per_path": round(total_rays / paths, 3),
        "stage_ms_per_step": {k[3:]: round(v / args.steps, 2) for k, v in totals.items()},
        "roofline": {"bound": "hbm", "kernel": "kExtend", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(info["width"], args.wavefront_paths),
                     "algorithmic_bytes_per_launch": round(bytes_per_launch),
                     "bytes_per_ray": round(bytes_per_ray, 1), "nodes_per_ray": round(nodes_per_ray, 3),
                     "tris_per_ray": round(tris_per_ray, 3), "mean_launch_ms": round(mean_launch_s * 1e3, 4),
                     "launches": launches},
    }
    if args.batch_scenes > 0:
        result["batched"] = batched(dev, args.batch_scenes, args.batch_steps, info["width"], args.wavefront_paths)
        log(f"batched: {result['batched']}")
    if args.small_batch > 0:
        result["batched_small"] = batched_small(dev, args.small_batch, 3)
        log(f"batched small scenes: {result['batched_small']}")
    if not args.no_guided:
        # time to 1% relMSE, guided (online training, one-sample MIS at 0.5) vs plain
        result["time_to_1pct_relmse"] = {
            "hard_room_256x256_depth12": guided_vs_plain(dev, scenes.hard_room(seed=0, width=256, height=256),
                                                         args.ref_spp, "hard room", learned=True),
            "terrain_slit_1920x1080_depth8": guided_vs_plain(dev, scenes.terrain_slit(), args.ref_spp // 16,
                                                             "terrain slit", max_level=9),
        }
    if args.guided_batch > 0:
        result["guided_batch_hard_rooms"] = guided_batch(dev, args.guided_batch, info["width"],
                                                         budget_s=args.guided_batch_budget,
                                                         which=tuple(args.guided_batch_estimators.split(",")))
    if not args.no_cpu:
        threads = min(_cpu_host()[0], 16)
        c1_scene = pg.Scene(dev, c1, bvh_width=info["width"])
        result["cpu_baseline"] = cpu_baseline(sc, (info["width"], *scene.bvh_download()), c1,
                                              (info["width"], *c1_scene.bvh_download()), threads)
        c1_scene.close()
        log(f"cpu baseline: {result['cpu_baseline']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
