"""Launch-shape autotuner with a persistent cache.

The ray tile (how a wavefront's 64 rays cover the screen), the samples per
thread (voxel fetches in flight per ray), the resident wavefronts per SIMD
and, for 8/16-bit volumes, the classification mode (tabulated per raw value
or computed per sample) change only the speed of a frame, never its pixels.
`Tuner` picks them per request class -- voxel type, viewport size (log2 of
the pixel count), sampling rate (samples per voxel of the finest level, log2)
and filter -- by timing every candidate on the first frame of that class
(`lvr_autotune`), and keeps the choices in a JSON file so that later runs
start tuned. The package ships the table measured on MI355X (tune_gfx950.json
beside this module); the user cache (LIVRE_AMD_TUNE_CACHE, default
~/.cache/livre_amd/tune_gfx950.json) takes precedence over it and receives
every class tuned later. Entries without a wavefronts-per-SIMD choice (older
caches) are tuned again.

A batch renders with one samples-per-thread value (it selects the kernel)
and one workgroup size for all of its requests, so `tune_batch` gives every
request its class's ray tile and the batch the samples-per-thread and
wavefronts-per-SIMD values that the classes chose most often, weighted by
pixels.
"""
import json
import math
import os

from . import DT_UINT8, DT_UINT16, RenderParams, lib

ARCH = "gfx950"
SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_%s.json" % ARCH)


def default_cache_path():
    return os.environ.get("LIVRE_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "livre_amd",
                                       "tune_%s.json" % ARCH))


def request_class(vi, width, height, params):
    """Key of the class a frame belongs to."""
    vmax = max(vi.voxels[0], vi.voxels[1], vi.voxels[2])
    rate = params.n_samples_per_ray / float(vmax)
    return "%s/dt%d/px%d/rate%+d/%s" % (
        ARCH, vi.data_type, int(round(math.log2(max(1, width * height)))),
        int(round(math.log2(max(rate, 1e-6)))), "linear" if params.linear_filter else "nearest")


def _type_class(data_type):
    """The batch kernel's voxel-type branch: 8-bit, 16-bit or float."""
    return {DT_UINT8: 0, DT_UINT16: 1}.get(int(data_type), 2)


class Tuner:
    def __init__(self, path=None, reps=3):
        self.path = path or default_cache_path()
        self.reps = reps
        self.tuned = 0  # classes tuned by this instance (cache misses)
        self.table = self._load(SHIPPED)
        self.table.update(self._load(self.path))

    @staticmethod
    def _load(path):
        try:
            with open(path) as f:
                table = json.load(f)
            return table if isinstance(table, dict) else {}
        except (OSError, ValueError):
            return {}

    def save(self):
        os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def choice(self, key):
        return self.table.get(key)

    def params_for(self, ctx, fr, width, height, ordered, params):
        """params with the tuned tile and samples per thread of the frame's
        class; tunes the class on this frame if it is not cached."""
        lib()
        key = request_class(ctx.vi, width, height, params)
        c = self.table.get(key)
        if c is None or "waves_per_simd" not in c:
            tuned, cands = ctx.autotune(fr, width, height, ordered, params, self.reps)
            c = {"tile": [tuned.tile_w, tuned.tile_h],
                 "samples_per_thread": tuned.samples_per_thread,
                 "classify": tuned.classify,
                 "waves_per_simd": tuned.waves_per_simd,
                 "candidates": [list(x) for x in cands]}
            self.table[key] = c
            self.tuned += 1
            self.save()
        out = RenderParams.from_buffer_copy(params)
        out.tile_w, out.tile_h = c["tile"]
        out.samples_per_thread = c["samples_per_thread"]
        out.classify = c.get("classify", 0)
        out.waves_per_simd = c.get("waves_per_simd", 0)
        return out

    def cached(self, vi, width, height, params):
        """params with the cached choice of the frame's class applied, without
        tuning; params unchanged when the class has no entry."""
        c = self.table.get(request_class(vi, width, height, params))
        out = RenderParams.from_buffer_copy(params)
        if c is not None:
            out.tile_w, out.tile_h = c["tile"]
            out.samples_per_thread = c["samples_per_thread"]
            out.classify = c.get("classify", 0)
            out.waves_per_simd = c.get("waves_per_simd", 0)
        return out

    def tune_batch(self, requests):
        """requests: list of (ctx, fr, width, height, ordered, params).
        Returns one RenderParams per request: the class's tile, samples per
        thread by pixel-weighted vote per voxel type (8-bit, 16-bit, float:
        the batch kernel's branches) and wavefronts per SIMD by vote over
        the batch."""
        tuned = [self.params_for(*r) for r in requests]

        def vote(field, group):
            votes = {}
            for (ctx, _fr, w, h, _o, _p), t in zip(requests, tuned):
                g = group(ctx)
                if g is not None and not (field == "samples_per_thread" and t.linear_filter):
                    v = getattr(t, field)
                    votes.setdefault(g, {})
                    votes[g][v] = votes[g].get(v, 0) + w * h
            for (ctx, *_r), t in zip(requests, tuned):
                g = group(ctx)
                if g in votes:
                    setattr(t, field, max(votes[g].items(), key=lambda kv: kv[1])[0])

        # samples per thread per voxel type (the batch kernel's branches),
        # wavefronts per SIMD for the whole launch
        vote("samples_per_thread", lambda ctx: _type_class(ctx.vi.data_type))
        vote("waves_per_simd", lambda ctx: 0)
        return tuned
