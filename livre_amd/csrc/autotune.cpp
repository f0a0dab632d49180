// Autotuner of the ray-cast launch shape: the ray tile (how a wavefront's 64
// rays cover the screen), the resident wavefronts per SIMD, the samples per
// thread (voxel fetches in flight per ray) and, for 8/16-bit volumes,
// tabulated or per-sample classification.
// Every candidate renders the caller's frame; the ray-cast kernel is timed
// with HIP events on the context stream, median of `reps` launches, and
// the fastest candidate is written back into the parameters. Results are
// bit-identical for every candidate, so the choice only affects speed.
#include <algorithm>
#include <cstring>
#include <vector>

#include "lvr_internal.h"

namespace
{
const int kTiles[][2] = {{8, 8}, {16, 4}, {4, 16}, {32, 2}, {2, 32}};
const int kWavesPerSimd[] = {4, 5, 6, 8};
} // namespace

extern "C" int lvr_autotune(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                            const uint64_t* ordered_ids, int n, int reps,
                            lvr_render_params* params, lvr_tune_entry* results, int cap,
                            int* n_results)
{
    if (!ctx || !f || !viewport || !params || reps < 1 || reps > 64 || cap < 0 ||
        (cap > 0 && !results))
        return LVR_EINVAL;
    const bool wasTiming = ctx->timing;
    int rc = lvr_set_timing(ctx, 1);
    if (rc)
        return rc;
    // NEAREST: 1 to 3 fetches in flight; GL_LINEAR: brick by brick (1) or
    // through the two-brick window (2)
    const int depths = params->linear_filter ? 2 : 3;
    // 8/16-bit NEAREST: tabulated or per-sample classification
    const int dt = ctx->vi.data_type;
    const bool tabulable =
        !params->linear_filter && (dt == LVR_DT_UINT8 || dt == LVR_DT_UINT16);
    const int modes = tabulable ? 2 : 1;
    std::vector<lvr_tune_entry> all;
    const int nTiles = int(sizeof kTiles / sizeof kTiles[0]);
    const int nOcc = int(sizeof kWavesPerSimd / sizeof kWavesPerSimd[0]);
    for (int c = 0; c < nTiles * nOcc * depths * modes; ++c)
    {
        const auto& t = kTiles[c / (nOcc * depths * modes)];
        const int w = kWavesPerSimd[(c / (depths * modes)) % nOcc];
        const int d = 1 + (c / modes) % depths;
        const int m = 1 + c % modes;
        lvr_render_params p = *params;
        p.tile_w = t[0];
        p.tile_h = t[1];
        p.samples_per_thread = d;
        p.classify = tabulable ? m : 0;
        p.waves_per_simd = w;
        std::vector<float> ms;
        for (int r = 0; r < reps && rc == LVR_OK; ++r)
        {
            rc = lvr_render(ctx, f, viewport, ordered_ids, n, &p);
            lvr_frame_stats st;
            if (rc == LVR_OK)
                rc = lvr_frame_stats_get(ctx, &st);
            if (rc == LVR_OK)
                ms.push_back(st.raycast_ms);
        }
        if (rc)
            break;
        std::nth_element(ms.begin(), ms.begin() + ms.size() / 2, ms.end());
        lvr_tune_entry e;
        std::memset(&e, 0, sizeof e);
        e.tile_w = p.tile_w;
        e.tile_h = p.tile_h;
        e.samples_per_thread = p.samples_per_thread;
        e.classify = p.classify;
        e.waves_per_simd = p.waves_per_simd;
        e.raycast_ms = ms[ms.size() / 2];
        all.push_back(e);
    }
    if (!wasTiming)
        (void)lvr_set_timing(ctx, 0);
    if (rc)
        return rc;
    const auto best = std::min_element(all.begin(), all.end(),
                                       [](const lvr_tune_entry& a, const lvr_tune_entry& b) {
                                           return a.raycast_ms < b.raycast_ms;
                                       });
    params->tile_w = best->tile_w;
    params->tile_h = best->tile_h;
    params->samples_per_thread = best->samples_per_thread;
    params->classify = best->classify;
    params->waves_per_simd = best->waves_per_simd;
    const int k = std::min(cap, int(all.size()));
    for (int i = 0; i < k; ++i)
        results[i] = all[size_t(i)];
    if (n_results)
        *n_results = int(all.size());
    return LVR_OK;
}
