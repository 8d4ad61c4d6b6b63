"""Autotuner for the batched encode: picks the code-block size (64x64 or
32x32) and the tier-1 launch form (one workgroup per code-block, or a
persistent grid of N wavefronts per CU pulling from a device work queue) by
timing the actual batch, and caches the choice in a JSON file keyed by the
batch's shape signature (image count, total megapixels, share of 9/7 and RGB
images) and code-block style."""
import json
import os
import time

CACHE = os.environ.get("RG_AUTOTUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                          ".autotune.json"))
# (code-block size, waves per CU; 0 = per-block grid).  The context modeller
# runs at four waves per SIMD, i.e. up to 16 per CU.
CANDIDATES = [(64, 0), (64, 8), (64, 12), (64, 16), (32, 0), (32, 12), (32, 16)]


def signature(images):
    mpix = sum(m["w"] * m["h"] for m in images) / 1e6
    lossy = sum(1 for m in images if not m["reversible"]) / len(images)
    rgb = sum(1 for m in images if m["ncomp"] == 3) / len(images)
    return "n%d_mpix%d_lossy%.2f_rgb%.2f" % (len(images), round(mpix), lossy, rgb)


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(ctx, batch, images, candidates=CANDIDATES, reps=2, log=None, style=0):
    """Plans `batch` (code-block style `style`) with each candidate, times
    `reps` runs, leaves the batch planned with the best one and returns
    {cb, waves_per_cu, ms, table}."""
    key = signature(images) + "_style%d" % style
    cache = _load()
    if key in cache and "waves_per_cu" in cache[key]:
        best = cache[key]
        batch.plan(images, cbw=best["cb"], cbh=best["cb"], use_graph=True, style=style,
                   waves_per_cu=best["waves_per_cu"])
        return dict(best, cached=True)
    table = []
    for cb, wpc in candidates:
        batch.plan(images, cbw=cb, cbh=cb, use_graph=True, style=style, waves_per_cu=wpc)
        batch.run()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            batch.run()
        ctx.sync()
        ms = (time.perf_counter() - t) * 1e3 / reps
        table.append({"cb": cb, "waves_per_cu": wpc, "ms": round(ms, 3)})
        if log:
            log("autotune cb=%d waves/CU=%d: %.2f ms" % (cb, wpc, ms))
    best = min(table, key=lambda r: r["ms"])
    cache[key] = {"cb": best["cb"], "waves_per_cu": best["waves_per_cu"], "ms": best["ms"]}
    try:
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass
    batch.plan(images, cbw=best["cb"], cbh=best["cb"], use_graph=True, style=style,
               waves_per_cu=best["waves_per_cu"])
    return dict(best, cached=False, table=table)
