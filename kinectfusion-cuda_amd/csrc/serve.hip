// Serving: a request loop over a kf_batch.
//
// Sessions are admitted against an HBM budget (FIFO of waiting requests) and
// stay resident until closed.  Frames are copied into a per-session ring of
// pinned host memory as they arrive; each tick takes the oldest waiting frame
// of every admitted session, copies those frames to the sessions' device
// buffers and runs one batch tick (one launch per stage for all of them).  At
// most two ticks are in flight, so frames that arrive while the GPU works on
// one tick are batched into the next: continuous batching.  Latency of a
// frame = the loop clock when its tick is seen complete - its arrival stamp.
// A tick is launched once enough sessions have a frame waiting, or, with the
// GPU idle, once the oldest waiting frame has waited long enough
// (kf_server_config::launch_fill, max_wait_ms).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "kf_internal.hpp"

namespace {

constexpr int kInFlight = 2;

struct Ticket {
    kf_session_config cfg{};
    size_t bytes = 0;
    size_t frame_bytes = 0;
    int state = 0;  // 0 waiting, 1 admitted, -1 released
    int sid = -1;
    bool closing = false;
    uint16_t* raw_dev = nullptr;  // kInFlight device frame buffers, one per tick slot
    uint16_t* ring = nullptr;  // pinned, cap frames
    int cap = 0, head = 0, count = 0, inflight = 0;
    std::vector<double> arrival;
    float last_pose[16];
    int status = 0;
};

struct Tick {
    hipEvent_t ev = nullptr;
    hipEvent_t copied = nullptr;  // the tick's frame uploads (copy stream) are done
    bool busy = false;
    std::vector<int> tickets;
    std::vector<double> arrivals;
};

double clock_s()
{
    using namespace std::chrono;
    return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

// SplitMix64 (Steele, Lea, Flood 2014) and a uniform double in (0, 1).
struct SplitMix64 {
    uint64_t x;
    uint64_t next()
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct kf_server {
    kf_server_config cfg{};
    hipStream_t s = nullptr;
    hipStream_t cs = nullptr;  // frame uploads, overlapping the running tick
    kf_batch* b = nullptr;
    std::vector<Ticket> t;
    std::deque<int> waiting;
    size_t resident = 0, peak = 0;
    Tick fl[kInFlight];
    int next = 0;
    std::vector<double> lat;
    int64_t ticks = 0, frames_in_ticks = 0, admitted = 0, queued_admissions = 0, failed = 0;
    double t_first = -1.0, t_last = 0.0, t_last_arrival = 0.0;
    int64_t submitted = 0;
};

namespace {

int release(kf_server* s, Ticket& q)
{
    int st = kf_batch_pose(s->b, q.sid, q.last_pose);
    if (!st) st = kf_batch_session_status(s->b, q.sid, &q.status);
    if (!st && q.status) s->failed += 1;
    if (!st) st = kf_batch_remove_session(s->b, q.sid);
    (void)hipFree(q.raw_dev);
    if (q.ring) (void)hipHostFree(q.ring);
    q.raw_dev = nullptr;
    q.ring = nullptr;
    q.state = -1;
    q.sid = -1;
    s->resident -= q.bytes;
    return st;
}

int admit(kf_server* s)
{
    while (!s->waiting.empty()) {
        Ticket& q = s->t[s->waiting.front()];
        if (s->resident + q.bytes > s->cfg.hbm_budget_bytes) return KF_OK;
        int free_slots = 0;
        for (const Ticket& o : s->t) free_slots += (o.state == 1) ? 1 : 0;
        if (free_slots >= s->cfg.batch.max_sessions) return KF_OK;
        // the frame buffer first: a failure then leaves nothing half admitted
        hipError_t e = hipMalloc(&q.raw_dev, q.frame_bytes * kInFlight);
        if (e != hipSuccess) {
            q.raw_dev = nullptr;
            return kf::fail_hip(e, "server session buffer");
        }
        int st = kf_batch_add_session(s->b, &q.cfg, &q.sid);
        if (st) {
            (void)hipFree(q.raw_dev);
            q.raw_dev = nullptr;
            q.sid = -1;
            return st;
        }
        q.state = 1;
        s->resident += q.bytes;
        s->peak = std::max(s->peak, s->resident);
        s->admitted += 1;
        s->waiting.pop_front();
    }
    return KF_OK;
}

int retire(kf_server* s, Tick& k, bool wait)
{
    if (!k.busy) return KF_OK;
    if (wait) {
        KF_HIP(hipEventSynchronize(k.ev));
    } else {
        hipError_t e = hipEventQuery(k.ev);
        if (e == hipErrorNotReady) return KF_OK;
        if (e != hipSuccess) return kf::fail_hip(e, "server tick");
    }
    const double now = clock_s();
    for (size_t q = 0; q < k.tickets.size(); ++q) {
        Ticket& tk = s->t[k.tickets[q]];
        tk.inflight -= 1;
        tk.head = (tk.head + 1) % tk.cap;  // the copied frame's ring slot is free again
        tk.count -= 1;
        s->lat.push_back((now - k.arrivals[q]) * 1e3);
    }
    s->t_last = now;
    k.busy = false;
    k.tickets.clear();
    k.arrivals.clear();
    return KF_OK;
}

int launch(kf_server* s, int* launched, bool flush)
{
    Tick& k = s->fl[s->next];
    if (k.busy) return KF_OK;
    // A tick costs nearly as much with a few frames as with many, so ticks are
    // launched full: once frames of launch_fill of the resident sessions wait
    // (then it may also queue behind a running tick), or, with the GPU idle,
    // once the oldest waiting frame has waited max_wait_ms.
    bool running = false;
    for (const Tick& t : s->fl) running = running || t.busy;
    int resident = 0, waiting = 0;
    double oldest = 0.0;
    for (const Ticket& q : s->t) {
        if (q.state != 1) continue;
        resident += 1;
        if (q.count - q.inflight > 0) {
            const double a = q.arrival[(q.head + q.inflight) % q.cap];
            oldest = waiting ? std::min(oldest, a) : a;
            waiting += 1;
        }
    }
    if (waiting == 0) return KF_OK;
    const int need = std::max(1, (int)ceil((double)s->cfg.launch_fill * resident - 1e-9));
    if (waiting < need && !flush) {
        if (running) return KF_OK;
        if (s->cfg.max_wait_ms > 0.0f && (clock_s() - oldest) * 1e3 < (double)s->cfg.max_wait_ms) return KF_OK;
    }
    std::vector<int> sids;
    std::vector<const uint16_t*> raws;
    for (int id = 0; id < (int)s->t.size(); ++id) {
        Ticket& q = s->t[id];
        // the oldest frame not yet in a tick: ring position head + inflight
        if (q.state != 1 || q.count - q.inflight <= 0) continue;
        const int slot = (q.head + q.inflight) % q.cap;
        // this tick slot's own device buffer: the tick that used it before has
        // retired, and the running tick reads the other one
        uint16_t* dst = q.raw_dev + (size_t)s->next * (q.frame_bytes / 2);
        KF_HIP(hipMemcpyAsync(dst, q.ring + (size_t)slot * (q.frame_bytes / 2), q.frame_bytes, hipMemcpyHostToDevice,
                              s->cs));
        sids.push_back(q.sid);
        raws.push_back(dst);
        k.tickets.push_back(id);
        k.arrivals.push_back(q.arrival[slot]);
        q.inflight += 1;
    }
    if (sids.empty()) return KF_OK;
    KF_HIP(hipEventRecord(k.copied, s->cs));
    KF_HIP(hipStreamWaitEvent(s->s, k.copied, 0));
    int st = kf_batch_tick(s->b, (int)sids.size(), sids.data(), raws.data());
    if (st) return st;
    KF_HIP(hipEventRecord(k.ev, s->s));
    k.busy = true;
    s->next = (s->next + 1) % kInFlight;
    s->ticks += 1;
    s->frames_in_ticks += (int64_t)sids.size();
    if (launched) *launched = 1;
    return KF_OK;
}

Ticket* ticket_of(kf_server* s, int id)
{
    if (!s || id < 0 || id >= (int)s->t.size()) return nullptr;
    return &s->t[id];
}

// One turn of the loop; flush launches whatever waits (draining: nothing more arrives).
int loop_turn(kf_server* s, int* launched, bool flush)
{
    if (launched) *launched = 0;
    int st = KF_OK;
    for (Tick& k : s->fl)
        if ((st = retire(s, k, false))) return st;
    for (Ticket& q : s->t)
        if (q.state == 1 && q.closing && q.count == 0 && q.inflight == 0)
            if ((st = release(s, q))) return st;
    if ((st = admit(s))) return st;
    return launch(s, launched, flush);
}

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest rank
    size_t r = (size_t)ceil(p / 100.0 * (double)v.size());
    if (r < 1) r = 1;
    return v[r - 1];
}

}  // namespace

extern "C" {

double kf_server_now(void) { return clock_s(); }

int kf_server_destroy(kf_server* s)
{
    if (!s) return KF_OK;
    if (s->s) (void)hipStreamSynchronize(s->s);
    if (s->cs) (void)hipStreamSynchronize(s->cs);
    for (Ticket& q : s->t) {
        (void)hipFree(q.raw_dev);
        if (q.ring) (void)hipHostFree(q.ring);
    }
    kf_batch_destroy(s->b);
    for (Tick& k : s->fl) {
        if (k.ev) (void)hipEventDestroy(k.ev);
        if (k.copied) (void)hipEventDestroy(k.copied);
    }
    if (s->s) (void)hipStreamDestroy(s->s);
    if (s->cs) (void)hipStreamDestroy(s->cs);
    delete s;
    return KF_OK;
}

int kf_server_create(const kf_server_config* cfg, kf_server** out)
{
    if (!cfg || !out) return kf::fail(KF_ERR_ARG, "kf_server_create: null argument");
    *out = nullptr;
    kf_server* s = new kf_server();
    s->cfg = *cfg;
    if (s->cfg.queue_depth <= 0) s->cfg.queue_depth = 8;
    if (!(s->cfg.launch_fill > 0.0f)) s->cfg.launch_fill = 0.5f;
    if (s->cfg.launch_fill > 1.0f) s->cfg.launch_fill = 1.0f;
    if (s->cfg.max_wait_ms == 0.0f) s->cfg.max_wait_ms = 2.0f;
    hipError_t e = hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking);
    int st = (e == hipSuccess) ? KF_OK : kf::fail_hip(e, "kf_server_create stream");
    for (Tick& k : s->fl)
        if (!st && (hipEventCreateWithFlags(&k.ev, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&k.copied, hipEventDisableTiming) != hipSuccess))
            st = kf::fail(KF_ERR_HIP, "kf_server_create events");
    if (!st) st = kf_batch_create(&cfg->batch, reinterpret_cast<kf_stream_t>(s->s), &s->b);
    if (st) {
        kf_server_destroy(s);
        return st;
    }
    *out = s;
    return KF_OK;
}

int kf_server_open(kf_server* s, const kf_session_config* c, int* ticket)
{
    if (!s || !c || !ticket) return kf::fail(KF_ERR_ARG, "kf_server_open: null argument");
    Ticket q;
    q.cfg = *c;
    int st = kf_session_bytes(c, &q.bytes);
    if (st) return st;
    // charged against the budget: what the batch holds for the session and the
    // session's device frame buffers (one per tick slot)
    q.frame_bytes = sizeof(uint16_t) * (size_t)c->frame_width * c->frame_height;
    q.bytes += q.frame_bytes * kInFlight;
    if (q.bytes > s->cfg.hbm_budget_bytes)
        return kf::fail(KF_ERR_BUDGET, "kf_server_open: the session alone exceeds the HBM budget");
    q.cap = s->cfg.queue_depth;
    q.arrival.assign((size_t)q.cap, 0.0);
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&q.ring), q.frame_bytes * (size_t)q.cap,
                                 hipHostMallocDefault);
    if (e != hipSuccess) return kf::fail_hip(e, "kf_server_open frame queue");
    const int id = (int)s->t.size();
    s->t.push_back(q);
    s->waiting.push_back(id);
    st = admit(s);
    if (!st && s->t[id].state == 0) s->queued_admissions += 1;  // could not be admitted at once
    *ticket = id;
    return st;
}

int kf_server_close(kf_server* s, int id)
{
    Ticket* q = ticket_of(s, id);
    if (!q || q->state == -1) return kf::fail(KF_ERR_ARG, "kf_server_close: unknown or released session");
    q->closing = true;
    if (q->state == 0) {
        // never admitted: drop the request
        s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), id), s->waiting.end());
        if (q->ring) (void)hipHostFree(q->ring);
        q->ring = nullptr;
        q->state = -1;
    }
    return KF_OK;
}

int kf_server_submit(kf_server* s, int id, const uint16_t* frame, double arrival_s)
{
    Ticket* q = ticket_of(s, id);
    if (!q || !frame || q->state == -1 || q->closing)
        return kf::fail(KF_ERR_ARG, "kf_server_submit: unknown, closing or released session");
    if (q->count >= q->cap) return kf::fail(KF_ERR_BUDGET, "kf_server_submit: the session's frame queue is full");
    const int slot = (q->head + q->count) % q->cap;
    memcpy(q->ring + (size_t)slot * (q->frame_bytes / 2), frame, q->frame_bytes);
    q->arrival[slot] = arrival_s;
    q->count += 1;
    if (s->t_first < 0.0 || arrival_s < s->t_first) s->t_first = arrival_s;
    if (arrival_s > s->t_last_arrival) s->t_last_arrival = arrival_s;
    s->submitted += 1;
    return KF_OK;
}

int kf_server_poll(kf_server* s, int* launched)
{
    if (!s) return kf::fail(KF_ERR_ARG, "kf_server_poll: null server");
    return loop_turn(s, launched, false);
}

int kf_server_drain(kf_server* s)
{
    if (!s) return kf::fail(KF_ERR_ARG, "kf_server_drain: null server");
    for (;;) {
        int st = loop_turn(s, nullptr, true);
        if (st) return st;
        bool busy = false, admitted_work = false, waiting_work = false;
        for (const Tick& k : s->fl) busy = busy || k.busy;
        for (const Ticket& q : s->t) {
            admitted_work = admitted_work || (q.state == 1 && (q.count > 0 || q.closing));
            waiting_work = waiting_work || (q.state == 0 && q.count > 0);
        }
        if (!busy && !admitted_work && !waiting_work) return KF_OK;
        if (busy) {
            // nothing new can start until the older tick finishes
            Tick& k = s->fl[s->next].busy ? s->fl[s->next] : s->fl[(s->next + 1) % kInFlight];
            if ((st = retire(s, k, true))) return st;
            continue;
        }
        // an idle GPU with frames of admitted sessions launches (or releases) on the next poll
        if (admitted_work) continue;
        return kf::fail(KF_ERR_BUDGET, "kf_server_drain: waiting sessions can never be admitted");
    }
}

int kf_server_stats_get(kf_server* s, kf_server_stats* o)
{
    if (!s || !o) return kf::fail(KF_ERR_ARG, "kf_server_stats_get: null argument");
    memset(o, 0, sizeof(*o));
    o->frames = (int64_t)s->lat.size();
    o->ticks = s->ticks;
    o->admitted = s->admitted;
    o->queued_admissions = s->queued_admissions;
    o->failed_frames = s->failed;
    o->wall_s = (s->t_first >= 0.0 && s->t_last > s->t_first) ? s->t_last - s->t_first : 0.0;
    o->frames_per_s = o->wall_s > 0 ? (double)o->frames / o->wall_s : 0.0;
    o->mean_frames_per_tick = s->ticks ? (double)s->frames_in_ticks / (double)s->ticks : 0.0;
    const double span = s->t_last_arrival - s->t_first;
    o->offered_frames_per_s = (s->t_first >= 0.0 && span > 0.0) ? (double)s->submitted / span : 0.0;
    o->latency_p50_ms = percentile(s->lat, 50.0);
    o->latency_p99_ms = percentile(s->lat, 99.0);
    double mx = 0.0, sum = 0.0;
    for (double v : s->lat) {
        mx = std::max(mx, v);
        sum += v;
    }
    o->latency_max_ms = mx;
    o->latency_mean_ms = s->lat.empty() ? 0.0 : sum / (double)s->lat.size();
    o->peak_resident_bytes = s->peak;
    return KF_OK;
}

int kf_server_latencies(kf_server* s, double* out, size_t cap, size_t* count)
{
    if (!s || !count) return kf::fail(KF_ERR_ARG, "kf_server_latencies: null argument");
    *count = s->lat.size();
    if (out)
        for (size_t q = 0; q < std::min(cap, s->lat.size()); ++q) out[q] = s->lat[q];
    return KF_OK;
}

int kf_server_state(kf_server* s, int id, int* state)
{
    Ticket* q = ticket_of(s, id);
    if (!q || !state) return kf::fail(KF_ERR_ARG, "kf_server_state: bad argument");
    *state = q->state;
    return KF_OK;
}

int kf_server_pose(kf_server* s, int id, float pose[16])
{
    Ticket* q = ticket_of(s, id);
    if (!q || !pose) return kf::fail(KF_ERR_ARG, "kf_server_pose: bad argument");
    if (q->state == 1) return kf_batch_pose(s->b, q->sid, pose);
    if (q->state == -1 && q->sid == -1) {
        memcpy(pose, q->last_pose, sizeof(q->last_pose));
        return KF_OK;
    }
    return kf::fail(KF_ERR_ARG, "kf_server_pose: the session was never admitted");
}

int kf_server_volume(kf_server* s, int id, kf_volume** v)
{
    Ticket* q = ticket_of(s, id);
    if (!q || !v || q->state != 1) return kf::fail(KF_ERR_ARG, "kf_server_volume: the session is not resident");
    return kf_batch_volume(s->b, q->sid, v);
}

int kf_server_checkpoint(kf_server* s, int id, const char* path)
{
    Ticket* q = ticket_of(s, id);
    if (!q || q->state != 1) return kf::fail(KF_ERR_ARG, "kf_server_checkpoint: the session is not resident");
    return kf_batch_session_save(s->b, q->sid, path);
}

int kf_server_restore(kf_server* s, int id, const char* path)
{
    Ticket* q = ticket_of(s, id);
    if (!q || q->state != 1) return kf::fail(KF_ERR_ARG, "kf_server_restore: the session is not resident");
    if (q->count > 0 || q->inflight > 0)
        return kf::fail(KF_ERR_ARG, "kf_server_restore: restore before the session's first frame");
    return kf_batch_session_restore(s->b, q->sid, path);
}

int kf_server_run_load(kf_server* s, const kf_poisson_load* L, int* tickets_out, kf_server_stats* out)
{
    if (!s || !L || L->sessions <= 0 || !L->configs || !L->banks || !L->bank_len || !L->frames_per_session ||
        !L->opens_s || !(L->rate_hz > 0.0) || L->jitter < 0.0 || L->jitter >= 1.0)
        return kf::fail(KF_ERR_ARG, "kf_server_run_load: bad load");
    const int S = L->sessions;
    // arrival schedule (seconds after the start)
    SplitMix64 rng{L->seed};
    std::vector<std::vector<double>> when((size_t)S);
    for (int k = 0; k < S; ++k) {
        if (L->bank_len[k] < 1 || L->frames_per_session[k] < 0 || !L->banks[k])
            return kf::fail(KF_ERR_ARG, "kf_server_run_load: empty frame bank");
        double t = L->opens_s[k];
        for (int f = 0; f < L->frames_per_session[k]; ++f) {
            when[k].push_back(t);
            const double gap = -log(rng.uniform()) / L->rate_hz;
            t += gap * (1.0 + L->jitter * (2.0 * rng.uniform() - 1.0));
        }
    }
    std::vector<int> ticket((size_t)S, -1), sent((size_t)S, 0);
    const double t0 = clock_s();
    int remaining = 0;
    for (int k = 0; k < S; ++k) remaining += L->frames_per_session[k];
    for (;;) {
        const double now = clock_s() - t0;
        for (int k = 0; k < S; ++k) {
            if (ticket[k] < 0) {
                if (now < L->opens_s[k]) continue;
                int st = kf_server_open(s, &L->configs[k], &ticket[k]);
                if (st) return st;
            }
            while (sent[k] < L->frames_per_session[k] && when[k][sent[k]] <= now) {
                const int n = L->bank_len[k];
                const int period = n > 1 ? 2 * (n - 1) : 1;
                const int ph = sent[k] % period;
                const int idx = ph < n ? ph : period - ph;
                const size_t px = (size_t)L->configs[k].frame_width * L->configs[k].frame_height;
                int st = kf_server_submit(s, ticket[k], L->banks[k] + (size_t)idx * px, t0 + when[k][sent[k]]);
                if (st == KF_ERR_BUDGET) break;  // queue full: retry on the next turn, stamp unchanged
                if (st) return st;
                sent[k] += 1;
                remaining -= 1;
                if (sent[k] == L->frames_per_session[k] && !L->keep_open) kf_server_close(s, ticket[k]);
            }
        }
        int launched = 0;
        int st = kf_server_poll(s, &launched);
        if (st) return st;
        if (remaining == 0) break;
        if (!launched) std::this_thread::yield();
    }
    int st = kf_server_drain(s);
    if (st) return st;
    if (tickets_out)
        for (int k = 0; k < S; ++k) tickets_out[k] = ticket[k];
    if (out) return kf_server_stats_get(s, out);
    return KF_OK;
}

}  // extern "C"
