/*
 * server.hip - the serving loop: solves on cached composites with
 * continuous batching (an extension: the reference solves one right-hand
 * side per call, src/SOLVER/solver.cu:20-129).
 *
 * A registered composite (operator) gets a lane while it has work: a
 * stream of its own and a slot engine (host_multi.hip) of column slots.
 * A step of the loop visits the lanes in turn and, for each,
 *   1. waits for the iteration it queued in the previous step,
 *   2. collects the slots whose stopping test ended them and retires the
 *      done ones (x copied to the request's output, the result recorded),
 *   3. admits queued requests into free slots - requests that arrived while
 *      others were on the device join those lanes at an iteration boundary,
 *      and a lane is opened for an operator only while the device memory it
 *      needs fits the budget (idle lanes are closed, and their memory freed,
 *      to make room; requests of an operator whose one-slot lane exceeds the
 *      whole budget fail with BCMG_ENOMEM),
 *   4. queues its next iteration (a captured graph) at once; the lanes'
 *      streams run concurrently on the device while the host turns to the
 *      next lane.
 * A slot computes exactly what a single solve of its right-hand side
 * computes, so a result does not depend on what else is being served.
 * A lane reads a request's b only after the work queued on the context's
 * stream before the step (an event the lane streams wait on), so a caller
 * may write b asynchronously on that stream right before submitting.
 */
#include "bcmg_internal.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>

namespace bcmg {

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    long long id = 0;
    int op = 0;
    const double *b = nullptr;
    double *x = nullptr;
    double t_submit = 0.0;
};

struct Lane {
    Ctx c;
    Boot *B = nullptr;
    MultiWork *M = nullptr;
    std::unique_ptr<SlotEngine> E;
    std::vector<Request> slot_req;
    size_t bytes = 0;
    double last_used = 0.0;
    ~Lane()
    {
        if (c.stream) (void)hipStreamSynchronize(c.stream);
        E.reset();
        destroy_multi(M);
    }
};

} // namespace

struct ServerResult {
    bool done = false;
    int iters = 0, status = BCMG_OK;
    double resid = 0.0, latency_ms = 0.0;
};

struct Server {
    /* one lane per operator, max_slots wide; its launches span only the
       slots in use (host_multi.hip), so a lightly used lane stays cheap */
    Ctx &main;
    size_t budget;
    int max_slots;
    hipEvent_t ev_main = nullptr;   /* end of the context stream's work at a step */
    std::vector<Boot *> ops;
    std::vector<std::deque<Request>> queue;
    std::vector<std::vector<std::unique_ptr<Lane>>> lanes; /* per operator */
    std::unordered_map<long long, ServerResult> results;
    std::vector<long long> completed_ids; /* finished since the last drain */
    size_t in_use = 0;              /* device bytes held by open lanes (pool classes) */
    long long next_id = 1;
    Server(Ctx &c, size_t b, int k) : main(c), budget(b), max_slots(k)
    {
        BCMG_HIP(hipEventCreateWithFlags(&ev_main, hipEventDisableTiming));
    }
    ~Server()
    {
        lanes.clear();
        if (ev_main) (void)hipEventDestroy(ev_main);
    }

    /* an upper bound of the device bytes a k-slot lane of B takes: its
       buffers' sizes, each rounded up to the pool's size class (less than a
       quarter more, and at most 255 bytes more for blocks up to 4 KiB) */
    static size_t lane_bound(const Boot &B, int k)
    {
        return multi_work_bytes(B, k) / 4 * 5 + ((size_t)64 << 10);
    }

    int add(Boot *B)
    {
        ops.push_back(B);
        queue.emplace_back();
        lanes.emplace_back();
        return (int)ops.size() - 1;
    }

    long long submit(int op, const double *b, double *x)
    {
        if (op < 0 || op >= (int)ops.size()) throw Error(BCMG_EINVAL, "unknown operator id");
        Request r;
        r.id = next_id++;
        r.op = op;
        r.b = b;
        r.x = x;
        r.t_submit = now_ms();
        queue[op].push_back(r);
        results[r.id] = ServerResult();
        return r.id;
    }

    static bool lane_idle(const Lane &L)
    {
        for (auto &s : L.E->slots)
            if (s.state != SLOT_FREE) return false;
        return true;
    }

    int slots_of(int op) const
    {
        int k = 0;
        for (auto &L : lanes[op]) k += L->M->kmax;
        return k;
    }

    /* close the least recently used idle lane of another operator (an idle
       lane of op itself would have free slots); false when none is idle */
    bool evict_one(int op)
    {
        int vo = -1;
        size_t vl = 0;
        double oldest = 0.0;
        for (size_t o = 0; o < lanes.size(); o++) {
            for (size_t l = 0; l < lanes[o].size(); l++) {
                const Lane &L = *lanes[o][l];
                if (!lane_idle(L)) continue;
                if ((int)o == op) continue;
                if (vo < 0 || L.last_used < oldest) {
                    vo = (int)o;
                    vl = l;
                    oldest = L.last_used;
                }
            }
        }
        if (vo < 0) return false;
        in_use -= lanes[vo][vl]->bytes;
        /* the lane's blocks go back to the device, not to the pool's cache
           (the pool reuses a block only for its own size class, so cached
           blocks of closed lanes would add up past the budget) */
        Pool *pool = current_pool();
        const bool prev = pool ? pool->free_released : false;
        if (pool) pool->free_released = true;
        try {
            lanes[vo].erase(lanes[vo].begin() + vl);
        } catch (...) {
            if (pool) pool->free_released = prev;
            throw;
        }
        if (pool) pool->free_released = prev;
        return true;
    }

    /* device bytes the calling thread's pool holds in live blocks */
    static size_t pool_live()
    {
        const Pool *pool = current_pool();
        return pool ? pool->live : 0;
    }

    /* open a lane for op (up to its slot limit, within the memory budget,
       closing idle lanes of other operators, least recently used first,
       when it does not fit) */
    Lane *open_lane(int op)
    {
        Boot &B = *ops[op];
        const int want = max_slots - slots_of(op);
        if (want < 1) return nullptr;
        auto fits = [&](int k) { return in_use + lane_bound(B, k) <= budget; };
        while (!fits(1))
            if (!evict_one(op)) return nullptr;
        int k = want;
        while (k > 1 && !fits(k)) k--;
        /* the pool hands blocks out in the order of one stream: the lane's
           stream may only take blocks the main stream is done with */
        BCMG_HIP(hipStreamSynchronize(main.stream));
        auto L = std::make_unique<Lane>();
        L->c.device = main.device;
        L->c.use_graphs = main.use_graphs;
        BCMG_HIP(hipStreamCreateWithFlags(&L->c.stream, hipStreamNonBlocking));
        L->B = &B;
        const size_t live0 = pool_live();
        multi_work(L->c, B, L->M, k);
        L->E = std::make_unique<SlotEngine>(L->c, B, *L->M);
        L->slot_req.assign(k, Request());
        L->bytes = pool_live() - live0;
        in_use += L->bytes;
        lanes[op].push_back(std::move(L));
        return lanes[op].back().get();
    }

    /* Visit the lanes in turn: a lane's previous iteration is waited for,
       its stopped slots collected and retired, queued requests admitted into
       its free slots, and its next iteration queued at once, so that it runs
       while the host turns to the next lane. Requests left over open new
       lanes while the operator's slot limit and the budget allow. */
    int step()
    {
        int completed = 0;
        /* b of a request may come from work the caller queued on the
           context's stream before this step: lanes admit after it */
        BCMG_HIP(hipEventRecord(ev_main, main.stream));
        for (size_t op = 0; op < ops.size(); op++) {
            auto &q = queue[op];
            for (auto &L : lanes[op]) {
                completed += finish(*L);
                if (!q.empty()) {
                    BCMG_HIP(hipStreamWaitEvent(L->c.stream, ev_main, 0));
                    admit(*L, q);
                }
                L->E->launch();
            }
            while (!q.empty()) {
                Lane *L = open_lane((int)op);
                if (!L) {
                    /* a one-slot lane larger than the whole budget never fits */
                    if (lane_bound(*ops[op], 1) > budget) completed += fail_queued((int)op);
                    break;
                }
                BCMG_HIP(hipStreamWaitEvent(L->c.stream, ev_main, 0));
                admit(*L, q);
                L->E->launch();
            }
        }
        return completed;
    }

    /* every queued request of op ends with BCMG_ENOMEM */
    int fail_queued(int op)
    {
        const double t = now_ms();
        int k = 0;
        for (const Request &r : queue[op]) {
            ServerResult &res = results[r.id];
            res.done = true;
            res.status = BCMG_ENOMEM;
            res.latency_ms = t - r.t_submit;
            completed_ids.push_back(r.id);
            k++;
        }
        queue[op].clear();
        return k;
    }

    /* wait, collect, retire (results recorded after the copies of x) */
    int finish(Lane &L)
    {
        SlotEngine &E = *L.E;
        if (E.launched) E.wait();
        E.collect();
        bool retired = false;
        for (size_t j = 0; j < E.slots.size(); j++) {
            if (E.slots[j].state != SLOT_DONE) continue;
            const SlotInfo S = E.slots[j];
            const Request &r = L.slot_req[j];
            E.retire((int)j, r.x);
            ServerResult &res = results[r.id];
            res.iters = S.zero_rhs ? 0 : S.iters;
            res.resid = S.zero_rhs ? 0.0 : S.l2;
            res.status = S.status;
            retired = true;
        }
        if (!retired) return 0;
        BCMG_HIP(hipStreamSynchronize(L.c.stream));
        const double t = now_ms();
        int completed = 0;
        for (size_t j = 0; j < E.slots.size(); j++) {
            Request &r = L.slot_req[j];
            if (r.id && E.slots[j].state == SLOT_FREE) {
                ServerResult &res = results[r.id];
                if (!res.done) {
                    res.done = true;
                    res.latency_ms = t - r.t_submit;
                    completed_ids.push_back(r.id);
                    completed++;
                }
                r = Request();
            }
        }
        L.last_used = t;
        return completed;
    }

    /* queued requests of the lane's operator into its free slots (zero
       right-hand sides are done at admission and retired next step) */
    void admit(Lane &L, std::deque<Request> &q)
    {
        if (L.B->prec != BCMG_PREC_FP64 && L.M->prec != L.B->prec) {
            /* reduced-precision buffers are allocated now (see open_lane) */
            BCMG_HIP(hipStreamSynchronize(main.stream));
            const size_t live0 = pool_live();
            ensure_multi_into(L.c, *L.B, *L.M, L.M->kmax);
            const size_t grown = pool_live() - live0;
            L.bytes += grown;
            in_use += grown;
        }
        std::vector<int> js;
        std::vector<const double *> bs;
        for (size_t j = 0; j < L.E->slots.size() && !q.empty(); j++) {
            if (L.E->slots[j].state != SLOT_FREE) continue;
            L.slot_req[j] = q.front();
            q.pop_front();
            js.push_back((int)j);
            bs.push_back(L.slot_req[j].b);
        }
        if (js.empty()) return;
        L.E->admit(js, bs);
        L.last_used = now_ms();
    }

    void info(int *queued, int *running, size_t *bytes, int *nlanes) const
    {
        int qn = 0, rn = 0, ln = 0;
        for (auto &q : queue) qn += (int)q.size();
        for (auto &v : lanes)
            for (auto &L : v) {
                ln++;
                for (auto &s : L->E->slots) rn += s.state != SLOT_FREE;
            }
        if (queued) *queued = qn;
        if (running) *running = rn;
        if (bytes) *bytes = in_use;
        if (nlanes) *nlanes = ln;
    }
};

Server *server_create(Ctx &c, size_t budget, int slots)
{
    if (slots < 1 || slots > MULTI_MAX_COLS)
        throw Error(BCMG_EINVAL, "slots per operator must be 1..32");
    return new Server(c, budget, slots);
}
void server_destroy(Server *S) { delete S; }
int server_add(Server &S, Boot &B) { return S.add(&B); }
long long server_submit(Server &S, int op, const double *b, double *x)
{
    return S.submit(op, b, x);
}
int server_step(Server &S) { return S.step(); }
bool server_result(Server &S, long long id, int *done, int *iters, double *resid, int *status,
                   double *latency_ms)
{
    auto it = S.results.find(id);
    if (it == S.results.end()) return false;
    const ServerResult &r = it->second;
    if (done) *done = r.done ? 1 : 0;
    if (iters) *iters = r.iters;
    if (resid) *resid = r.resid;
    if (status) *status = r.status;
    if (latency_ms) *latency_ms = r.latency_ms;
    if (r.done) S.results.erase(it);
    return true;
}
int server_completed(Server &S, long long *ids, int cap)
{
    const int k = std::min<int>(cap, (int)S.completed_ids.size());
    for (int i = 0; i < k; i++) ids[i] = S.completed_ids[i];
    S.completed_ids.erase(S.completed_ids.begin(), S.completed_ids.begin() + k);
    return k;
}
void server_info(const Server &S, int *queued, int *running, size_t *bytes, int *lanes)
{
    S.info(queued, running, bytes, lanes);
}

} // namespace bcmg
