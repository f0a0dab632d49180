// Serving loop: concurrent camera sessions rendered by continuous batching.
//
// The reference serves one session per process with an event-driven redraw
// loop (Client.cpp:89-122, Config.cpp:298-342) and no batching. Here one
// process serves many sessions on one device:
//  - sessions are admitted against a byte budget of device memory (brick
//    pools of the volumes + per-session frame state); a session that does not
//    fit waits (FIFO) until another closes;
//  - frame requests queue as they arrive; whenever one of the two batch
//    slots is free, every queued request of an admitted session (up to the
//    batch limit) goes into one persistent batched launch, so requests that
//    arrive while a batch is on the GPU form the next one and the host
//    prepares it while the first renders;
//  - a session's newer request supersedes its older one still queued
//    (interactive sessions only need their latest view);
//  - bricks a request needs and the pool lacks are produced on the device by
//    the volume's procedural source, evicting least-recently-used bricks; the
//    bricks of every queued-for-launch or in-flight request are pinned, so
//    eviction never touches what a launch reads;
//  - a finished frame is copied into its session's device framebuffer (the
//    session state stays on the device) and reported with its arrival,
//    launch and completion times.
// Every frame is rendered by the same kernels as a frame rendered alone, so
// its pixels and sample count equal the standalone result.
#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "lvr_internal.h"

namespace
{
// Launches in flight at once: while one is retired and the next staged on
// the host, the others keep the GPU busy. Each session holds one
// framebuffer more than this: its finished frame is never one being written.
constexpr int kFlights = 4;

struct Volume
{
    lvr_ctx* ctx = nullptr;
    int source = 0; // 0 none (caller uploads), 1 noise, 2 field
    lvr_noise_params noise;
    lvr_field_params field;
    size_t bytes = 0;
    // a selection queued ahead on the context for a pending request
    // (lvr_select_visibles_async), and the finished ones held for their
    // requests
    int64_t selQueued = -1;
    std::vector<std::pair<int64_t, std::vector<uint64_t>>> ready;
    std::vector<uint64_t> selBuf; // what lvr_select_visibles_wait writes
};

struct Session
{
    int state = 0; // 0 closed, 1 waiting for admission, 2 admitted
    int volume = -1;
    int width = 0, height = 0;
    std::vector<float> tf;
    float r0 = 0.f, r1 = 0.f;
    lvr_render_params params{};
    size_t bytes = 0;
    uint32_t* d_frames = nullptr; // kFlights + 1 frames: finished + those in flight
    int nextBuf = 0, shownBuf = -1;
    int64_t shownRequest = -1; // launches on two streams may finish out of order
    bool hasFrame = false;
    uint64_t lastSamples = 0;
    uint64_t frames = 0;
};

struct Request
{
    int64_t id = 0;
    int session = 0;
    int buf = 0; // session framebuffer the launch writes
    lvr_frustum f{};
    double arrive = 0.0;
};

struct Flight
{
    lvr_batch* batch = nullptr;
    std::vector<Request> reqs;
    std::vector<std::pair<int, uint64_t>> pins; // (volume, brick)
    double launch = 0.0;
    bool busy = false;
};

double nowMs()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}
} // namespace

struct lvr_server
{
    int device = 0;
    size_t budget = 0, used = 0;
    int maxBatch = 64;
    bool coalesce = true;
    std::string error;
    std::vector<Volume> volumes;
    std::vector<Session> sessions;
    std::deque<int> waiting;
    std::deque<Request> pending;
    Flight flights[kFlights];
    hipStream_t copyStream = nullptr;
    std::vector<lvr_server_frame> done;
    std::vector<uint64_t> vis, ord;
    int64_t nextId = 0;
    lvr_server_stats stats{};
};

namespace
{
int sfail(lvr_server* s, int code, const std::string& msg)
{
    s->error = msg;
    return code;
}

int ctxFail(lvr_server* s, lvr_ctx* ctx, int code)
{
    s->error = lvr_last_error(ctx);
    return code;
}

size_t frameBytes(int w, int h)
{
    return size_t(w) * size_t(h) * 4u;
}

size_t sessionBytes(int w, int h)
{
    // the finished frame + one per batch slot in flight
    return frameBytes(w, h) * size_t(kFlights + 1);
}

// Kept free of the batch kernel for selection and brick-source kernels.
constexpr int kServerReservedCUs = 8;

void admitWaiting(lvr_server* s)
{
    while (!s->waiting.empty())
    {
        Session& ss = s->sessions[size_t(s->waiting.front())];
        if (s->used + ss.bytes > s->budget)
            return;
        if (hipMalloc((void**)&ss.d_frames, sessionBytes(ss.width, ss.height)) != hipSuccess)
            return; // stays waiting; retried at the next close
        s->used += ss.bytes;
        ss.state = 2;
        ++s->stats.admitted;
        s->waiting.pop_front();
    }
}

void report(lvr_server* s, const Request& r, int status, uint64_t samples, double launch,
            double end)
{
    lvr_server_frame f;
    std::memset(&f, 0, sizeof f);
    f.request = r.id;
    f.session = r.session;
    f.status = status;
    f.samples = samples;
    f.arrive_ms = r.arrive;
    f.launch_ms = launch;
    f.done_ms = end;
    s->done.push_back(f);
}

void unpin(lvr_server* s, Flight& fl)
{
    for (const auto& p : fl.pins)
        (void)lvr::pinBrick(s->volumes[size_t(p.first)].ctx, p.second, -1);
    fl.pins.clear();
}

// Retire a finished launch: frames into the sessions' framebuffers, pins off.
int retire(lvr_server* s, Flight& fl)
{
    const double t0 = nowMs();
    struct Acc
    {
        double t0;
        double* acc;
        ~Acc() { *acc += nowMs() - t0; }
    } acc{t0, &s->stats.retire_ms};
    lvr_batch_stats st;
    int rc = lvr_batch_wait(fl.batch, &st);
    if (rc)
        return sfail(s, rc, lvr_batch_last_error(fl.batch));
    s->stats.gpu_ms += st.gpu_ms;
    const double end = nowMs();
    for (size_t i = 0; i < fl.reqs.size(); ++i)
    {
        const Request& r = fl.reqs[i];
        uint64_t samples = 0;
        rc = lvr_batch_read(fl.batch, int(i), nullptr, 0, &samples);
        if (rc)
            return sfail(s, rc, lvr_batch_last_error(fl.batch));
        Session& ss = s->sessions[size_t(r.session)];
        if (ss.state == 2 && r.id > ss.shownRequest)
        {
            ss.shownBuf = r.buf;
            ss.shownRequest = r.id;
            ss.hasFrame = true;
            ss.lastSamples = samples;
        }
        if (ss.state == 2)
            ++ss.frames;
        s->stats.samples += samples;
        report(s, r, 0, samples, fl.launch, end);
    }
    s->stats.frames += fl.reqs.size();
    unpin(s, fl);
    fl.reqs.clear();
    fl.busy = false;
    return LVR_OK;
}

// The visible-set capacity of a context: every node of its octree.
uint64_t visCap(lvr_ctx* ctx)
{
    return ctx->pt.levelOffset.back();
}

// Finish a volume's queued selection and hold its result for its request.
void settle(Volume& v)
{
    if (v.selQueued < 0)
        return;
    const uint64_t cap = visCap(v.ctx);
    if (v.selBuf.size() < cap)
        v.selBuf.resize(cap);
    int n = 0;
    if (lvr_select_visibles_wait(v.ctx, v.selBuf.data(), int(cap), &n) == LVR_OK)
        v.ready.emplace_back(v.selQueued,
                             std::vector<uint64_t>(v.selBuf.begin(), v.selBuf.begin() + n));
    v.selQueued = -1;
}

bool isPending(const lvr_server* s, int64_t id)
{
    for (const Request& r : s->pending)
        if (r.id == id)
            return true;
    return false;
}

// Keep each volume's context selecting ahead: a finished selection is held
// for its request, and the oldest pending request of the volume without a
// visible set is queued next (lvr_select_visibles_async), so that it runs
// on the GPU while launches are in flight and the request's launch finds
// its visible set ready. Results of requests that are gone (superseded, or
// their session closed) are dropped.
void preselect(lvr_server* s)
{
    for (size_t vi = 0; vi < s->volumes.size(); ++vi)
    {
        Volume& v = s->volumes[vi];
        if (v.selQueued >= 0 && hipEventQuery(v.ctx->selDone) == hipSuccess)
            settle(v);
        v.ready.erase(std::remove_if(v.ready.begin(), v.ready.end(),
                                     [s](const std::pair<int64_t, std::vector<uint64_t>>& e) {
                                         return !isPending(s, e.first);
                                     }),
                      v.ready.end());
        if (v.selQueued >= 0)
            continue;
        for (const Request& r : s->pending)
        {
            const Session& ss = s->sessions[size_t(r.session)];
            if (ss.state != 2 || size_t(ss.volume) != vi)
                continue;
            bool have = false;
            for (const auto& e : v.ready)
                have = have || e.first == r.id;
            if (have)
                continue;
            if (lvr_select_visibles_async(v.ctx, &r.f, uint32_t(ss.height), 1.0f, 0,
                                          v.ctx->vi.depth - 1) == LVR_OK)
                v.selQueued = r.id;
            break;
        }
    }
}

// Visible set, order and residency of one request; pins its bricks into the
// flight and adds it to the batch. Returns LVR_OK, LVR_ENOMEM when the pool
// cannot hold the request next to what is pinned (try again later), or an
// error.
int stage(lvr_server* s, Flight& fl, Request& r)
{
    Session& ss = s->sessions[size_t(r.session)];
    Volume& v = s->volumes[size_t(ss.volume)];
    lvr_ctx* ctx = v.ctx;
    const uint64_t cap = visCap(ctx);
    if (s->vis.size() < cap)
    {
        s->vis.resize(cap);
        s->ord.resize(cap);
    }
    int n = 0;
    double t = nowMs();
    int rc = LVR_OK;
    if (v.selQueued == r.id)
        settle(v);
    auto held = std::find_if(v.ready.begin(), v.ready.end(),
                             [&r](const std::pair<int64_t, std::vector<uint64_t>>& e) {
                                 return e.first == r.id;
                             });
    if (held != v.ready.end())
    {
        // selected ahead (preselect)
        n = int(held->second.size());
        std::copy(held->second.begin(), held->second.end(), s->vis.begin());
        v.ready.erase(held);
    }
    else
    {
        // the context takes one selection at a time: another request's
        // queued selection is finished and held for it first
        settle(v);
        rc = lvr_select_visibles(ctx, &r.f, uint32_t(ss.height), 1.0f, 0, ctx->vi.depth - 1,
                                 s->vis.data(), int(cap), &n);
    }
    s->stats.select_ms += nowMs() - t;
    if (rc)
        return ctxFail(s, ctx, rc);
    t = nowMs();
    rc = lvr_order(&ctx->vi, &r.f, s->vis.data(), n, s->ord.data());
    s->stats.order_ms += nowMs() - t;
    if (rc)
        return sfail(s, rc, "order");
    t = nowMs();
    struct Acc
    {
        double t0;
        double* acc;
        ~Acc() { *acc += nowMs() - t0; }
    };
    std::vector<uint64_t> missing;
    for (int i = 0; i < n; ++i)
        if (!ctx->resident.count(s->ord[size_t(i)]))
            missing.push_back(s->ord[size_t(i)]);
    if (!missing.empty())
    {
        if (v.source == 0)
            return sfail(s, LVR_ENOENT, "bricks missing and the volume has no source");
        // the pool must take the misses without touching pinned bricks,
        // nor the resident bricks of this request (pinned below)
        uint64_t residentHere = 0;
        for (int i = 0; i < n; ++i)
        {
            auto it = ctx->resident.find(s->ord[size_t(i)]);
            residentHere += it != ctx->resident.end() && it->second.pinned == 0;
        }
        if (lvr::evictableSlots(ctx) < missing.size() + residentHere)
            return LVR_ENOMEM;
        // protect this request's resident bricks from its own misses
        for (int i = 0; i < n; ++i)
            if (ctx->resident.count(s->ord[size_t(i)]))
            {
                (void)lvr::pinBrick(ctx, s->ord[size_t(i)], 1);
                fl.pins.emplace_back(ss.volume, s->ord[size_t(i)]);
            }
        rc = v.source == 1
                 ? lvr_noise_bricks_device(ctx, &v.noise, missing.data(), int(missing.size()))
                 : lvr_field_bricks_device(ctx, &v.field, missing.data(), int(missing.size()));
        if (rc)
            return ctxFail(s, ctx, rc);
        s->stats.bricks_generated += missing.size();
        for (uint64_t id : missing)
        {
            (void)lvr::pinBrick(ctx, id, 1);
            fl.pins.emplace_back(ss.volume, id);
        }
    }
    else
        for (int i = 0; i < n; ++i)
        {
            (void)lvr::pinBrick(ctx, s->ord[size_t(i)], 1);
            fl.pins.emplace_back(ss.volume, s->ord[size_t(i)]);
        }
    s->stats.residency_ms += nowMs() - t;
    Acc addTime{nowMs(), &s->stats.add_ms};
    const int vp[4] = {0, 0, ss.width, ss.height};
    int req = 0;
    rc = lvr_batch_add(fl.batch, ctx, &r.f, vp, s->ord.data(), n, &ss.params, ss.tf.data(),
                       int(ss.tf.size() / 4), ss.r0, ss.r1, &req);
    if (rc)
        return sfail(s, rc, lvr_batch_last_error(fl.batch));
    // the next of the kFlights + 1 framebuffers, neither shown nor in flight
    r.buf = ss.nextBuf;
    ss.nextBuf = (ss.nextBuf + 1) % (kFlights + 1);
    rc = lvr_batch_set_output(fl.batch, req, ss.d_frames + size_t(r.buf) * ss.width * ss.height);
    if (rc)
        return sfail(s, rc, lvr_batch_last_error(fl.batch));
    return LVR_OK;
}

// Form and launch the next batch from the queue on a free slot.
int launch(lvr_server* s, int* launched)
{
    Flight* fl = nullptr;
    for (Flight& f : s->flights)
        if (!f.busy)
        {
            fl = &f;
            break;
        }
    if (!fl || s->pending.empty())
        return LVR_OK;
    bool otherBusy = false;
    for (const Flight& f : s->flights)
        otherBusy = otherBusy || f.busy;
    int rc = lvr_batch_clear(fl->batch);
    if (rc)
        return sfail(s, rc, lvr_batch_last_error(fl->batch));
    std::deque<Request> keep;
    std::vector<int> inFlight; // sessions with a request in this launch
    while (!s->pending.empty())
    {
        Request r = s->pending.front();
        s->pending.pop_front();
        const Session& ss = s->sessions[size_t(r.session)];
        // one request per session per launch keeps at most two of its
        // frames in flight, next to the finished one
        const bool again =
            std::find(inFlight.begin(), inFlight.end(), r.session) != inFlight.end();
        if (ss.state != 2 || int(fl->reqs.size()) >= s->maxBatch || again)
        {
            if (ss.state == 0)
                report(s, r, LVR_EINVAL, 0, 0.0, nowMs()); // session closed
            else
                keep.push_back(r);
            continue;
        }
        const size_t pinsBefore = fl->pins.size();
        rc = stage(s, *fl, r);
        if (rc == LVR_OK)
        {
            fl->reqs.push_back(r);
            inFlight.push_back(r.session);
            continue;
        }
        // undo this request's pins
        for (size_t i = pinsBefore; i < fl->pins.size(); ++i)
            (void)lvr::pinBrick(s->volumes[size_t(fl->pins[i].first)].ctx, fl->pins[i].second,
                                -1);
        fl->pins.resize(pinsBefore);
        if (rc == LVR_ENOMEM && (otherBusy || !fl->reqs.empty()))
        {
            keep.push_back(r); // fits once pinned bricks are released
            continue;
        }
        ++s->stats.failed;
        report(s, r, rc, 0, 0.0, nowMs());
    }
    s->pending.swap(keep);
    if (fl->reqs.empty())
        return LVR_OK;
    const double tRun = nowMs();
    rc = lvr_batch_run(fl->batch);
    s->stats.run_ms += nowMs() - tRun;
    if (rc)
    {
        unpin(s, *fl);
        for (const Request& r : fl->reqs)
            report(s, r, rc, 0, 0.0, nowMs());
        fl->reqs.clear();
        return sfail(s, rc, lvr_batch_last_error(fl->batch));
    }
    fl->launch = nowMs();
    fl->busy = true;
    ++s->stats.batches;
    if (launched)
        *launched = int(fl->reqs.size());
    return LVR_OK;
}
} // namespace

extern "C" {

double lvr_server_now_ms(void)
{
    return nowMs();
}

int lvr_server_create(int device, size_t budget_bytes, int max_batch, lvr_server** out)
{
    if (!out || max_batch < 1 || max_batch > 4096 || budget_bytes == 0)
        return LVR_EINVAL;
    *out = nullptr;
    lvr_server* s = new lvr_server();
    s->device = device;
    s->budget = budget_bytes;
    s->maxBatch = max_batch;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copyStream, hipStreamNonBlocking) != hipSuccess)
    {
        lvr_server_destroy(s);
        return LVR_EHIP;
    }
    for (Flight& f : s->flights)
        if (lvr_batch_create(device, &f.batch) != LVR_OK ||
            lvr_batch_set_graph(f.batch, 0) != LVR_OK ||
            lvr_batch_reserve_cus(f.batch, kServerReservedCUs) != LVR_OK)
        {
            lvr_server_destroy(s);
            return LVR_EHIP;
        }
    *out = s;
    return LVR_OK;
}

void lvr_server_destroy(lvr_server* s)
{
    if (!s)
        return;
    for (Flight& f : s->flights)
    {
        if (f.batch)
        {
            if (f.busy)
                (void)lvr_batch_wait(f.batch, nullptr);
            unpin(s, f);
            lvr_batch_destroy(f.batch);
        }
    }
    for (Volume& v : s->volumes)
        if (v.ctx)
        {
            settle(v); // the context is the caller's: no selection left queued
            v.ctx->pinsGuardReaders = false;
        }
    for (Session& ss : s->sessions)
        (void)hipFree(ss.d_frames);
    if (s->copyStream)
        (void)hipStreamDestroy(s->copyStream);
    delete s;
}

const char* lvr_server_last_error(const lvr_server* s)
{
    return s ? s->error.c_str() : "null server";
}

int lvr_server_set_coalesce(lvr_server* s, int enable)
{
    if (!s)
        return LVR_EINVAL;
    s->coalesce = enable != 0;
    return LVR_OK;
}

int lvr_server_add_volume(lvr_server* s, lvr_ctx* ctx, const lvr_noise_params* noise,
                          const lvr_field_params* field, int* volume)
{
    if (!s || !ctx || (noise && field) || ctx->device != s->device)
        return LVR_EINVAL;
    for (const Volume& v : s->volumes)
        if (v.ctx == ctx)
            return sfail(s, LVR_EINVAL, "volume already added");
    Volume v;
    v.ctx = ctx;
    if (noise)
    {
        v.source = 1;
        v.noise = *noise;
    }
    else if (field)
    {
        v.source = 2;
        v.field = *field;
    }
    v.bytes = size_t(ctx->nSlots) * ctx->slotStride;
    if (s->used + v.bytes > s->budget)
        return sfail(s, LVR_ENOMEM, "brick pool exceeds the server budget");
    s->used += v.bytes;
    // every launch of this server pins the bricks it reads
    ctx->pinsGuardReaders = true;
    s->volumes.push_back(v);
    if (volume)
        *volume = int(s->volumes.size()) - 1;
    return LVR_OK;
}

int lvr_server_open(lvr_server* s, int volume, int width, int height, const float* tf_rgba,
                    int n_tf, float range0, float range1, const lvr_render_params* params,
                    int* session, int* admitted)
{
    if (!s || volume < 0 || size_t(volume) >= s->volumes.size() || width <= 0 ||
        height <= 0 || width > 16384 || height > 16384 || !tf_rgba || n_tf < 1 ||
        n_tf > 8192 || !params || params->linear_filter || !session)
        return LVR_EINVAL;
    Session ss;
    ss.volume = volume;
    ss.width = width;
    ss.height = height;
    ss.tf.assign(tf_rgba, tf_rgba + 4 * n_tf);
    ss.r0 = range0;
    ss.r1 = range1;
    ss.params = *params;
    ss.bytes = sessionBytes(width, height);
    ss.state = 1;
    s->sessions.push_back(ss);
    const int id = int(s->sessions.size()) - 1;
    s->waiting.push_back(id);
    admitWaiting(s);
    *session = id;
    if (admitted)
        *admitted = s->sessions[size_t(id)].state == 2;
    return LVR_OK;
}

int lvr_server_close(lvr_server* s, int session)
{
    if (!s || session < 0 || size_t(session) >= s->sessions.size())
        return LVR_EINVAL;
    Session& ss = s->sessions[size_t(session)];
    if (ss.state == 0)
        return LVR_OK;
    // frames of this session still in flight finish first
    for (Flight& f : s->flights)
        if (f.busy)
            for (const Request& r : f.reqs)
                if (r.session == session)
                {
                    const int rc = retire(s, f);
                    if (rc)
                        return rc;
                    break;
                }
    if (ss.state == 2)
    {
        s->used -= ss.bytes;
        (void)hipFree(ss.d_frames);
        ss.d_frames = nullptr;
    }
    else
        s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), session),
                         s->waiting.end());
    ss.state = 0;
    admitWaiting(s);
    return LVR_OK;
}

int lvr_server_submit(lvr_server* s, int session, const lvr_frustum* f, double arrive_ms,
                      int64_t* request)
{
    if (!s || !f || session < 0 || size_t(session) >= s->sessions.size() ||
        s->sessions[size_t(session)].state == 0)
        return LVR_EINVAL;
    Request r;
    r.id = s->nextId++;
    r.session = session;
    r.f = *f;
    r.arrive = arrive_ms;
    if (s->coalesce)
        for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
            if (it->session == session)
            {
                report(s, *it, 1, 0, 0.0, arrive_ms); // superseded
                ++s->stats.superseded;
                s->pending.erase(it);
                break;
            }
    s->pending.push_back(r);
    if (request)
        *request = r.id;
    if (hipSetDevice(s->device) == hipSuccess)
        preselect(s);
    return LVR_OK;
}

int lvr_server_step(lvr_server* s, int* launched, int* retired)
{
    if (!s)
        return LVR_EINVAL;
    if (launched)
        *launched = 0;
    if (retired)
        *retired = 0;
    if (hipSetDevice(s->device) != hipSuccess)
        return sfail(s, LVR_EHIP, "hipSetDevice");
    for (Flight& f : s->flights)
    {
        if (!f.busy)
            continue;
        const int q = lvr_batch_query(f.batch);
        if (q < 0)
            return sfail(s, q, lvr_batch_last_error(f.batch));
        if (q == 1)
        {
            if (retired)
                *retired += int(f.reqs.size());
            const int rc = retire(s, f);
            if (rc)
                return rc;
        }
    }
    const int rc = launch(s, launched);
    if (rc == LVR_OK)
        preselect(s);
    return rc;
}

int lvr_server_drain(lvr_server* s)
{
    if (!s)
        return LVR_EINVAL;
    for (;;)
    {
        int rc = lvr_server_step(s, nullptr, nullptr);
        if (rc)
            return rc;
        Flight* oldest = nullptr;
        for (Flight& f : s->flights)
            if (f.busy && (!oldest || f.launch < oldest->launch))
                oldest = &f;
        if (oldest)
        {
            // block on the earliest launch instead of polling
            if ((rc = retire(s, *oldest)))
                return rc;
            continue;
        }
        // nothing in flight after a step: the queue holds only requests of
        // sessions still waiting for admission (or nothing)
        bool any = false;
        for (const Request& r : s->pending)
            any = any || s->sessions[size_t(r.session)].state == 2;
        if (!any)
        {
            // the contexts are the caller's again: no selection left queued
            for (Volume& v : s->volumes)
                settle(v);
            return LVR_OK;
        }
    }
}

int lvr_server_poll(lvr_server* s, lvr_server_frame* out, int cap, int* n)
{
    if (!s || cap < 0 || (cap > 0 && !out) || !n)
        return LVR_EINVAL;
    const int k = std::min(cap, int(s->done.size()));
    std::copy(s->done.begin(), s->done.begin() + k, out);
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return LVR_OK;
}

int lvr_server_read(lvr_server* s, int session, uint8_t* out_rgba8, size_t bytes,
                    uint64_t* samples)
{
    if (!s || session < 0 || size_t(session) >= s->sessions.size())
        return LVR_EINVAL;
    const Session& ss = s->sessions[size_t(session)];
    if (ss.state != 2 || !ss.hasFrame)
        return sfail(s, LVR_ENOENT, "session has no finished frame");
    if (out_rgba8)
    {
        const size_t need = size_t(ss.width) * ss.height * 4u;
        if (bytes < need)
            return sfail(s, LVR_ERANGE, "frame buffer too small");
        // on the server's own stream (a null-stream copy would wait for the
        // launches in flight)
        if (hipMemcpyAsync(out_rgba8, ss.d_frames + size_t(ss.shownBuf) * ss.width * ss.height,
                           need, hipMemcpyDeviceToHost, s->copyStream) != hipSuccess ||
            hipStreamSynchronize(s->copyStream) != hipSuccess)
            return sfail(s, LVR_EHIP, "session frame read");
    }
    if (samples)
        *samples = ss.lastSamples;
    return LVR_OK;
}

int lvr_server_set_tf(lvr_server* s, int session, const float* tf_rgba, int n_tf,
                      float range0, float range1)
{
    if (!s || session < 0 || size_t(session) >= s->sessions.size() || !tf_rgba || n_tf < 1 ||
        n_tf > 8192 || !std::isfinite(range0) || !std::isfinite(range1))
        return LVR_EINVAL;
    Session& ss = s->sessions[size_t(session)];
    if (ss.state == 0)
        return sfail(s, LVR_EINVAL, "session closed");
    ss.tf.assign(tf_rgba, tf_rgba + 4 * size_t(n_tf));
    ss.r0 = range0;
    ss.r1 = range1;
    return LVR_OK;
}

int lvr_server_set_params(lvr_server* s, int session, const lvr_render_params* params)
{
    if (!s || session < 0 || size_t(session) >= s->sessions.size() || !params ||
        params->linear_filter || params->n_clip_planes < 0 || params->n_clip_planes > 6 ||
        params->n_samples_per_ray < 0)
        return LVR_EINVAL;
    Session& ss = s->sessions[size_t(session)];
    if (ss.state == 0)
        return sfail(s, LVR_EINVAL, "session closed");
    ss.params = *params;
    return LVR_OK;
}

int lvr_server_grab_jpeg(lvr_server* s, int session, lvr_jpeg_encoder* enc, int quality,
                         uint8_t* out, size_t cap, size_t* size)
{
    if (!s || !enc || session < 0 || size_t(session) >= s->sessions.size() || !size)
        return LVR_EINVAL;
    const Session& ss = s->sessions[size_t(session)];
    if (ss.state != 2 || !ss.hasFrame)
        return sfail(s, LVR_ENOENT, "session has no finished frame");
    // the finished frame is never the one a launch writes; its launch has
    // been retired, so its pixels are complete
    const uint32_t* px = ss.d_frames + size_t(ss.shownBuf) * ss.width * ss.height;
    const int rc = lvr_jpeg_encode(enc, px, ss.width, ss.height, ss.width * 4, LVR_PIXEL_RGBA,
                                   1, quality, out, cap, size);
    if (rc)
        return sfail(s, rc, "session frame grab");
    return LVR_OK;
}

int lvr_server_get_stats(lvr_server* s, lvr_server_stats* out)
{
    if (!s || !out)
        return LVR_EINVAL;
    *out = s->stats;
    out->budget_bytes = s->budget;
    out->used_bytes = s->used;
    out->pending = s->pending.size();
    out->waiting_sessions = s->waiting.size();
    out->in_flight = 0;
    for (const Flight& f : s->flights)
        out->in_flight += f.busy ? f.reqs.size() : 0;
    return LVR_OK;
}
}
