// Serving loop: continuous batching of template requests against the
// resident waveform (tmx_server_start / tmx_submit / tmx_poll / tmx_wait /
// tmx_server_stop in tmx.h).
//
// Requests queue in arrival order.  Each lane is a worker thread with its own
// HIP stream and work buffers; when free it takes the waiting requests in
// order, as many as fit the batch size and the memory budget, and runs them
// as one batch (run_batch: one launch sequence for all of their templates).
// Requests that arrive while batches run wait only for the next free lane, so
// lanes' batches overlap on the device and the batch size follows the load.
// A batch that fails is retried request by request, so one bad request (for
// example a threshold with more than 2^31 hits) fails alone.  A stopped loop
// hands its lanes back to the context, and the next start reuses them with
// the buffers and FFT plans they have grown.
#include "tremor_amd/tmx.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "ctx.hpp"

namespace tmx {

struct Request {
    int64_t ticket = 0;
    tmx_request req{};
    std::vector<float> data;
    size_t est_bytes = 0;
    double submit_s = 0.0, start_s = 0.0, done_s = 0.0;
    bool done = false;
    int status = 0;
    std::string error;
    QueryOut out;
    int64_t batch = -1;
    int batch_size = 0;
    int lane = -1;
};

struct Server {
    tmx_server_cfg cfg{};
    tmx_cfg run_cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> waiting;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<std::thread> threads;
    bool stopping = false;
    bool stopped = false;
    int64_t next_ticket = 1, next_batch = 0;
    size_t budget = 0, in_use = 0;
};

namespace {

// Device memory a request holds while its batch runs: template copies and
// block order (~16 B per sample), matrix-core operands (1 KB per 16 samples
// of padded length), its FFT spectrum, and its share of candidates and
// results.
size_t estimate_bytes(const tmx_request &r)
{
    const size_t m = static_cast<size_t>(r.m);
    size_t B = 1u << 15;
    while (B < 4 * m)
        B *= 2;
    const size_t tmpl = 16 * m + 1024 * ((m + 31 + 15) / 16);
    const size_t spec = 8 * (B / 2 + 1);
    const size_t res = r.kind == TMX_REQ_KNN ? 64 * 24 * static_cast<size_t>(r.k) + (1u << 16) : (16u << 20);
    return tmpl + spec + res;
}

void finish(Request *r, const QueryOut *o, int rc, const std::string &err, double t0, double t1, int64_t bid, int bsize,
            int lane)
{
    r->start_s = t0;
    r->done_s = t1;
    r->batch = bid;
    r->batch_size = bsize;
    r->lane = lane;
    if (rc) {
        r->status = rc < 0 ? rc : -rc;
        r->error = err.empty() ? "search failed" : err;
    } else {
        r->status = 0;
        r->out = *o;
    }
    r->done = true;
}

void lane_main(tmx_ctx *ctx, Server *s, int li)
{
    Lane &L = *s->lanes[li];
    (void)hipSetDevice(L.device);
    for (;;) {
        std::vector<Request *> batch;
        size_t bytes = 0;
        int64_t bid = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                while (!s->waiting.empty() && static_cast<int>(batch.size()) < s->cfg.max_batch) {
                    Request *r = s->waiting.front();
                    // A request larger than the whole budget still runs, alone.
                    const bool alone = batch.empty() && s->in_use == 0;
                    if (!alone && s->in_use + bytes + r->est_bytes > s->budget)
                        break;
                    s->waiting.pop_front();
                    batch.push_back(r);
                    bytes += r->est_bytes;
                }
                if (!batch.empty())
                    break;
                if (s->waiting.empty() && s->stopping)
                    return;
                s->cv_work.wait(lk);
            }
            s->in_use += bytes;
            bid = s->next_batch++;
        }
        const int nb = static_cast<int>(batch.size());
        std::vector<Query> q(nb);
        for (int i = 0; i < nb; i++) {
            const tmx_request &rq = batch[i]->req;
            q[i] = Query{batch[i]->data.data(), rq.m, rq.kind == TMX_REQ_KNN ? rq.k : 0, rq.merge_offset, rq.rho};
        }
        const double t0 = now_s();
        std::vector<QueryOut> out;
        int rc = run_batch(ctx, L, s->run_cfg, q.data(), nb, out, nullptr);
        double t1 = now_s();
        if (rc == 0) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (int i = 0; i < nb; i++)
                finish(batch[i], &out[i], 0, "", t0, t1, bid, nb, li);
        } else if (nb == 1) {
            std::lock_guard<std::mutex> lk(s->mu);
            finish(batch[0], nullptr, rc, L.err, t0, t1, bid, 1, li);
        } else {
            for (int i = 0; i < nb; i++) {
                std::vector<QueryOut> one;
                const int r1 = run_batch(ctx, L, s->run_cfg, &q[i], 1, one, nullptr);
                t1 = now_s();
                std::lock_guard<std::mutex> lk(s->mu);
                finish(batch[i], r1 ? nullptr : &one[0], r1, L.err, t0, t1, bid, nb, li);
            }
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use -= bytes;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

// Ends the lane threads once the queue has drained; the lanes go back to the
// context's spares with their buffers and plans.
void stop_server(tmx_ctx *ctx, Server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopped)
            return;
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads)
        if (t.joinable())
            t.join();
    s->threads.clear();
    {
        std::lock_guard<std::mutex> lk(ctx->spare_mu);
        for (auto &L : s->lanes)
            ctx->spare_lanes.push_back(std::move(L));
    }
    s->lanes.clear();
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopped = true;
}

void fill_result(const Request &r, tmx_result *out)
{
    std::memset(out, 0, sizeof(*out));
    out->status = r.status;
    out->kind = r.req.kind;
    out->count = static_cast<int64_t>(r.out.pos.size());
    const size_t n = r.out.pos.size() ? r.out.pos.size() : 1;
    out->pos = static_cast<uint64_t *>(std::malloc(sizeof(uint64_t) * n));
    out->d = static_cast<float *>(std::malloc(sizeof(float) * n));
    if (!r.out.pos.empty()) {
        std::memcpy(out->pos, r.out.pos.data(), sizeof(uint64_t) * r.out.pos.size());
        std::memcpy(out->d, r.out.d.data(), sizeof(float) * r.out.d.size());
    }
    out->submit_s = r.submit_s;
    out->start_s = r.start_s;
    out->done_s = r.done_s;
    out->batch = r.batch;
    out->batch_size = r.batch_size;
    out->lane = r.lane;
    std::strncpy(out->error, r.error.c_str(), sizeof(out->error) - 1);
}

}  // namespace

void server_shutdown(tmx_ctx *ctx)
{
    if (ctx->server) {
        stop_server(ctx, ctx->server.get());
        ctx->server.reset();
    }
}

bool server_running(const tmx_ctx *ctx) { return ctx->server && !ctx->server->stopped; }

void ServerDelete::operator()(Server *s) const { delete s; }

}  // namespace tmx

using namespace tmx;

extern "C" {

int tmx_server_cfg_default(tmx_server_cfg *cfg)
{
    if (!cfg)
        return -1;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->lanes = 2;
    cfg->max_batch = 64;
    cfg->mem_budget = 0;
    cfg->max_template = 65536;
    return 0;
}

int tmx_server_start(tmx_ctx *ctx, const tmx_server_cfg *cfg)
{
    if (!ctx)
        return -1;
    if (server_running(ctx))
        return ctx->fail("the serving loop is already running");
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    tmx_server_cfg c;
    tmx_server_cfg_default(&c);
    if (cfg)
        c = *cfg;
    if (c.lanes < 1 || c.lanes > 8)
        return ctx->fail("lanes must be in [1, 8]");
    if (c.max_batch < 1 || c.max_batch > 65535)
        return ctx->fail("max_batch must be in [1, 65535]");
    if (c.max_template < 1 || c.max_template > 65536)
        return ctx->fail("max_template must be in [1, 65536]");
    if (c.mem_budget < 0)
        return ctx->fail("mem_budget must be non-negative");
    ctx->server.reset();
    std::unique_ptr<Server, ServerDelete> s(new Server());
    s->cfg = c;
    {
        std::lock_guard<std::mutex> call(ctx->call_mu);
        s->run_cfg = ctx->cfg;
    }
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
        return ctx->fail("cannot query device memory", -2);
    s->budget = c.mem_budget > 0 ? static_cast<size_t>(c.mem_budget) : free_b / 2;
    {
        // Spare lanes of an earlier loop on this device first.
        std::lock_guard<std::mutex> lk(ctx->spare_mu);
        auto &sp = ctx->spare_lanes;
        for (size_t i = 0; i < sp.size() && static_cast<int>(s->lanes.size()) < c.lanes;) {
            if (sp[i]->device == ctx->cfg.device) {
                s->lanes.push_back(std::move(sp[i]));
                sp.erase(sp.begin() + static_cast<std::ptrdiff_t>(i));
            } else {
                i++;
            }
        }
    }
    while (static_cast<int>(s->lanes.size()) < c.lanes) {
        auto L = std::make_unique<Lane>();
        if (L->init(ctx->cfg.device) != hipSuccess)
            return ctx->fail("cannot create a lane stream", -2);
        s->lanes.push_back(std::move(L));
    }
    Server *sp = s.get();
    ctx->server = std::move(s);
    for (int i = 0; i < c.lanes; i++)
        sp->threads.emplace_back(lane_main, ctx, sp, i);
    return 0;
}

int tmx_server_stop(tmx_ctx *ctx)
{
    if (!ctx)
        return -1;
    if (!ctx->server)
        return ctx->fail("the serving loop is not running");
    stop_server(ctx, ctx->server.get());
    return 0;
}

int64_t tmx_submit(tmx_ctx *ctx, const tmx_request *req)
{
    if (!ctx)
        return -1;
    Server *s = ctx->server.get();
    if (!s || s->stopped)
        return ctx->fail("the serving loop is not running", -4);
    if (!req || !req->tmpl || req->m <= 0)
        return ctx->fail("empty template");
    if (req->m > s->cfg.max_template)
        return ctx->fail("template longer than the serving loop's max_template");
    if (req->kind == TMX_REQ_KNN) {
        if (req->k < 1 || req->k > 4096)
            return ctx->fail("k must be in [1, 4096]");
    } else if (req->kind == TMX_REQ_THRESHOLD) {
        if (!std::isfinite(req->rho))
            return ctx->fail("correlation threshold must be finite");
    } else {
        return ctx->fail("unknown request kind");
    }
    if (req->merge_offset < 0)
        return ctx->fail("merge offset must be non-negative");
    auto r = std::make_unique<Request>();
    r->req = *req;
    r->data.assign(req->tmpl, req->tmpl + req->m);
    r->req.tmpl = nullptr;
    r->est_bytes = estimate_bytes(*req);
    int64_t ticket;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping)
            return ctx->fail("the serving loop is stopping", -4);
        ticket = s->next_ticket++;
        r->ticket = ticket;
        r->submit_s = now_s();
        s->waiting.push_back(r.get());
        s->reqs.emplace(ticket, std::move(r));
    }
    s->cv_work.notify_one();
    return ticket;
}

int tmx_poll(tmx_ctx *ctx, int64_t ticket, tmx_result *out)
{
    return tmx_wait(ctx, ticket, 0.0, out);
}

int tmx_wait(tmx_ctx *ctx, int64_t ticket, double timeout_s, tmx_result *out)
{
    if (!ctx)
        return -1;
    Server *s = ctx->server.get();
    if (!s || !out)
        return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end())
        return -1;
    Request *r = it->second.get();
    if (!r->done) {
        if (timeout_s == 0.0)
            return 1;
        if (timeout_s < 0.0) {
            s->cv_done.wait(lk, [&] { return r->done; });
        } else {
            const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
            if (!s->cv_done.wait_until(lk, until, [&] { return r->done; }))
                return 1;
        }
    }
    fill_result(*r, out);
    s->reqs.erase(it);
    return 0;
}

void tmx_result_free(tmx_result *res)
{
    if (!res)
        return;
    std::free(res->pos);
    std::free(res->d);
    res->pos = nullptr;
    res->d = nullptr;
    res->count = 0;
}

}  // extern "C"
