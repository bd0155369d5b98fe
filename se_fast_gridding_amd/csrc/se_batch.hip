// Batched and served solves: an executor that runs many independent SE3P/SE2P systems
// concurrently on one GPU.
//
// The executor owns `slots`: one HIP stream (default priority) each, with a cache of
// device-resident plans keyed by geometry and precision.  A job (one system, device
// arrays) goes to a slot, reuses that slot's plan for its geometry when the capacity
// suffices and otherwise builds one, and is launched asynchronously on the slot's
// stream; its completion is a HIP event behind a ticket.  Jobs on different slots
// overlap on the device; jobs on one slot run in submission order, so a plan is never
// used by two jobs at once.  Admission is by HBM: the plans of all slots together stay
// within the executor's byte budget (least recently used plans of idle slots are freed
// to make room), and se_batch_try_submit refuses a job when no slot is idle.
// se_batch_submit runs a job through its plan's captured graph (se_solve_graph), so a
// job that is submitted again costs one launch; se_batch_try_submit replays the job's
// graph when an idle slot already holds one (one launch instead of one per kernel) and
// otherwise runs the job eagerly, so that no capture lands on a request's latency (a
// serving loop captures its jobs ahead of traffic, e.g. by submitting each job once per
// slot).  se_batch_run takes a whole set of
// jobs, groups them by geometry and output mode, and solves each group with one group
// plan (se_solve_group: one launch per stage for all systems of the group); the groups
// run on different slots and overlap.  Group plans are cached by geometry across calls
// and ordered behind their previous use with events, so any slot can run them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <list>
#include <vector>

#include "se_gpu.h"
#include "se_internal.h"

namespace {

struct PlanEntry {
    se_job_shape key{};
    int cap = 0;
    se_plan* plan = nullptr;
    size_t bytes = 0;
    unsigned long last_use = 0;
};

struct GroupEntry {
    se_job_shape key{};
    int s_cap = 0, n_cap = 0;
    se_plan* plan = nullptr;
    size_t bytes = 0;
    unsigned long last_use = 0;
    hipEvent_t done = nullptr;   // completion of the plan's latest group solve
};

struct Slot {
    hipStream_t stream = nullptr;
    std::list<PlanEntry> plans;
    hipEvent_t last = nullptr;   // completion of the slot's latest job
    long pending = 0;            // jobs submitted and not yet seen complete
};

bool same_shape(const se_job_shape& a, const se_job_shape& b)
{
    return a.periodicity == b.periodicity && a.M == b.M && a.P == b.P && a.box == b.box && a.xi == b.xi &&
           a.m == b.m && a.slab == b.slab && a.precision == b.precision;
}

}  // namespace

struct se_batch {
    std::vector<Slot> slots;
    std::list<GroupEntry> groups;
    hipStream_t join = nullptr;   // completion of se_batch_run tickets
    size_t budget = 0, used = 0;
    unsigned long clock = 0;
    int next = 0;   // round-robin cursor
    std::deque<hipEvent_t> free_events;
    std::vector<hipEvent_t> tickets;   // ticket -> event (nullptr once collected)
    std::vector<int> ticket_slot;

    ~se_batch()
    {
        if (join) (void)hipStreamSynchronize(join);
        for (auto& s : slots)
            if (s.stream) (void)hipStreamSynchronize(s.stream);
        for (auto& e : groups) {
            se_plan_destroy(e.plan);
            if (e.done) (void)hipEventDestroy(e.done);
        }
        if (join) (void)hipStreamDestroy(join);
        for (auto& s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            for (auto& e : s.plans) se_plan_destroy(e.plan);
            if (s.last) (void)hipEventDestroy(s.last);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        for (hipEvent_t e : tickets)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    }

    bool idle(int i)
    {
        Slot& s = slots[i];
        if (s.pending == 0) return true;
        if (hipEventQuery(s.last) == hipSuccess) {
            s.pending = 0;
            return true;
        }
        return false;
    }

    // make room for `need` bytes of new plans: free least recently used plans of idle
    // slots and group plans whose last solve has completed
    void reserve(size_t need)
    {
        while (used + need > budget) {
            Slot* best_s = nullptr;
            std::list<PlanEntry>::iterator best;
            unsigned long best_use = ~0ul;
            for (int i = 0; i < (int)slots.size(); ++i) {
                if (!idle(i)) continue;
                for (auto it = slots[i].plans.begin(); it != slots[i].plans.end(); ++it)
                    if (it->last_use < best_use) {
                        best_s = &slots[i];
                        best = it;
                        best_use = it->last_use;
                    }
            }
            auto gbest = groups.end();
            for (auto it = groups.begin(); it != groups.end(); ++it)
                if (it->last_use < best_use && hipEventQuery(it->done) == hipSuccess) {
                    gbest = it;
                    best_use = it->last_use;
                }
            if (gbest != groups.end()) {
                used -= gbest->bytes;
                se_plan_destroy(gbest->plan);
                (void)hipEventDestroy(gbest->done);
                groups.erase(gbest);
                continue;
            }
            SE_REQUIRE(best_s, SE_ERR_CAPACITY, "se_batch: job does not fit the HBM budget");
            used -= best->bytes;
            se_plan_destroy(best->plan);
            best_s->plans.erase(best);
        }
    }

    // group plan of a geometry with room for S systems and N particles in total; an
    // existing one that is too small is replaced once its last solve has completed
    GroupEntry& group_for(const se_job_shape& k, int S, long N)
    {
        for (auto it = groups.begin(); it != groups.end(); ++it) {
            if (!same_shape(it->key, k)) continue;
            if (it->s_cap >= S && it->n_cap >= N) {
                it->last_use = ++clock;
                return *it;
            }
            SE_HIP(hipEventSynchronize(it->done));
            used -= it->bytes;
            se_plan_destroy(it->plan);
            (void)hipEventDestroy(it->done);
            groups.erase(it);
            break;
        }
        se_plan_desc d{};
        d.periodicity = k.periodicity;
        d.M = k.M;
        d.P = k.P;
        d.box = k.box;
        d.xi = k.xi;
        d.m = k.m;
        d.precision = k.precision;
        d.slab = k.slab;
        // capacities rounded up so that nearby batches share a plan
        const int scap = (S + 7) / 8 * 8;
        long cap = 1 << 16;
        while (cap < N) cap *= 2;
        SE_REQUIRE(cap < (1l << 31), SE_ERR_CAPACITY, "se_batch: too many particles in one group");
        d.n_max = (int)cap;
        const size_t need = se_plan_bytes_for_group(&d, scap);
        SE_REQUIRE(need > 0, SE_ERR_ARG, se_gpu_last_error());
        reserve(need);
        se_plan* p = nullptr;
        const int rc = se_plan_create_group(&p, &d, scap, slots[0].stream);
        SE_REQUIRE(rc == SE_OK, rc, se_gpu_last_error());
        GroupEntry e;
        e.key = k;
        e.s_cap = scap;
        e.n_cap = (int)cap;
        e.plan = p;
        e.bytes = se_plan_bytes(p);
        e.last_use = ++clock;
        SE_HIP(hipEventCreateWithFlags(&e.done, hipEventDisableTiming));
        used += e.bytes;
        groups.push_front(e);
        return groups.front();
    }

    bool has_plan(int si, const se_job_shape& k, int N)
    {
        for (auto& e : slots[si].plans)
            if (same_shape(e.key, k) && e.cap >= N) return true;
        return false;
    }

    // the slot's plan for the job holds the job's captured graph
    bool has_graph(int si, const se_job* job)
    {
        for (auto& e : slots[si].plans)
            if (same_shape(e.key, job->shape) && e.cap >= job->N)
                return se::plan_has_graph(e.plan, job->N, job->d_x, job->d_q, job->d_phi, job->d_force);
        return false;
    }

    se_plan* plan_for(int si, const se_job_shape& k, int N)
    {
        Slot& s = slots[si];
        for (auto it = s.plans.begin(); it != s.plans.end(); ++it)
            if (same_shape(it->key, k) && it->cap >= N) {
                it->last_use = ++clock;
                return it->plan;
            }
        // a plan of this shape but too small is replaced (after the slot's work drains)
        for (auto it = s.plans.begin(); it != s.plans.end(); ++it)
            if (same_shape(it->key, k)) {
                SE_HIP(hipStreamSynchronize(s.stream));
                used -= it->bytes;
                se_plan_destroy(it->plan);
                s.plans.erase(it);
                break;
            }
        se_plan_desc d{};
        d.periodicity = k.periodicity;
        d.M = k.M;
        d.P = k.P;
        d.box = k.box;
        d.xi = k.xi;
        d.m = k.m;
        d.precision = k.precision;
        d.slab = k.slab;
        // capacity rounded up so that nearby sizes share a plan
        int cap = 1024;
        while (cap < N) cap *= 2;
        d.n_max = cap;
        const size_t need = se_plan_bytes_for(&d);
        SE_REQUIRE(need > 0, SE_ERR_ARG, se_gpu_last_error());
        reserve(need);
        se_plan* p = nullptr;
        const int rc = se_plan_create(&p, &d, s.stream);
        SE_REQUIRE(rc == SE_OK, rc, se_gpu_last_error());
        PlanEntry e;
        e.key = k;
        e.cap = cap;
        e.plan = p;
        e.bytes = se_plan_bytes(p);
        e.last_use = ++clock;
        used += e.bytes;
        s.plans.push_front(e);
        return p;
    }

    hipEvent_t new_event()
    {
        if (!free_events.empty()) {
            hipEvent_t e = free_events.front();
            free_events.pop_front();
            return e;
        }
        hipEvent_t e;
        SE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }

    int launch(int si, const se_job* job, bool graph)
    {
        Slot& s = slots[si];
        se_plan* p = plan_for(si, job->shape, job->N);
        const int rc = graph ? se_solve_graph(p, job->N, job->d_x, job->d_q, job->d_phi, job->d_force)
                             : se_solve(p, job->N, job->d_x, job->d_q, job->d_phi, job->d_force);
        SE_REQUIRE(rc == SE_OK, rc, se_gpu_last_error());
        hipEvent_t ev = new_event();
        SE_HIP(hipEventRecord(ev, s.stream));
        SE_HIP(hipEventRecord(s.last, s.stream));
        ++s.pending;
        tickets.push_back(ev);
        ticket_slot.push_back(si);
        return (int)tickets.size() - 1;
    }
};

namespace {

template <typename F>
int guarded_b(F&& f)
{
    try {
        f();
        return SE_OK;
    } catch (const se::Error& e) {
        se::set_error(e.msg);
        return e.code;
    }
}

}  // namespace

extern "C" {

int se_batch_create(se_batch** out, int n_slots, size_t hbm_budget)
{
    if (!out || n_slots < 1 || n_slots > 64) return SE_ERR_ARG;
    *out = nullptr;
    se_batch* b = new se_batch();
    const int rc = guarded_b([&] {
        b->budget = hbm_budget ? hbm_budget : (size_t)64 << 30;
        b->slots.resize(n_slots);
        for (auto& s : b->slots) {
            SE_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            SE_HIP(hipEventCreateWithFlags(&s.last, hipEventDisableTiming));
        }
        SE_HIP(hipStreamCreateWithFlags(&b->join, hipStreamNonBlocking));
    });
    if (rc != SE_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return SE_OK;
}

int se_batch_destroy(se_batch* b)
{
    delete b;
    return SE_OK;
}

int se_batch_submit(se_batch* b, const se_job* job, int* ticket)
{
    if (!b || !job || !ticket) return SE_ERR_ARG;
    return guarded_b([&] {
        // round robin: queueing needs no completion queries (each costs a driver call)
        const int si = b->next;
        b->next = (b->next + 1) % (int)b->slots.size();
        *ticket = b->launch(si, job, true);
    });
}

int se_batch_try_submit(se_batch* b, const se_job* job, int* ticket)
{
    if (!b || !job || !ticket) return SE_ERR_ARG;
    *ticket = -1;
    return guarded_b([&] {
        // an idle slot that holds the job's captured graph (replayed: one launch), else an
        // idle slot that holds a plan for the job, else any idle slot (both eager)
        const int n = (int)b->slots.size();
        int with_plan = -1, fallback = -1;
        for (int r = 0; r < n; ++r) {
            const int i = (b->next + r) % n;
            if (!b->idle(i)) continue;
            if (b->has_graph(i, job)) {
                b->next = (i + 1) % n;
                *ticket = b->launch(i, job, true);
                return;
            }
            if (with_plan < 0 && b->has_plan(i, job->shape, job->N)) with_plan = i;
            if (fallback < 0) fallback = i;
        }
        const int si = with_plan >= 0 ? with_plan : fallback;
        if (si < 0) throw se::Error{SE_BUSY, "se_batch: no idle slot"};
        b->next = (si + 1) % n;
        *ticket = b->launch(si, job, false);
    });
}

int se_batch_run(se_batch* b, const se_job* jobs, int n_jobs, int* ticket)
{
    if (!b || (!jobs && n_jobs > 0) || n_jobs < 0 || !ticket) return SE_ERR_ARG;
    *ticket = -1;
    return guarded_b([&] {
        // groups: same geometry and the same requested outputs, in first-seen order
        std::vector<std::vector<int>> members;
        std::vector<int> lead;
        for (int j = 0; j < n_jobs; ++j) {
            const se_job& a = jobs[j];
            SE_REQUIRE(a.N >= 0, SE_ERR_ARG, "se_batch_run: negative job size");
            if (a.N == 0) continue;   // nothing to compute
            int gi = -1;
            for (int g = 0; g < (int)lead.size(); ++g) {
                const se_job& l = jobs[lead[g]];
                if (same_shape(l.shape, a.shape) && (l.d_phi != nullptr) == (a.d_phi != nullptr) &&
                    (l.d_force != nullptr) == (a.d_force != nullptr)) {
                    gi = g;
                    break;
                }
            }
            if (gi < 0) {
                gi = (int)lead.size();
                lead.push_back(j);
                members.emplace_back();
            }
            members[gi].push_back(j);
        }
        std::vector<se_system> sys;
        std::vector<hipEvent_t> ends;
        for (size_t g = 0; g < members.size(); ++g) {
            const auto& mem = members[g];
            long total = 0;
            for (int j : mem) total += jobs[j].N;
            GroupEntry& e = b->group_for(jobs[lead[g]].shape, (int)mem.size(), total);
            const int si = b->next;
            b->next = (b->next + 1) % (int)b->slots.size();
            Slot& s = b->slots[si];
            // the plan's previous solve (on whichever slot) completes first
            SE_HIP(hipStreamWaitEvent(s.stream, e.done, 0));
            int rc = se_plan_set_stream(e.plan, s.stream);
            SE_REQUIRE(rc == SE_OK, rc, se_gpu_last_error());
            sys.resize(mem.size());
            for (size_t i = 0; i < mem.size(); ++i) {
                const se_job& a = jobs[mem[i]];
                sys[i] = se_system{a.N, a.d_x, a.d_q, a.d_phi, a.d_force};
            }
            rc = se_solve_group(e.plan, (int)sys.size(), sys.data());
            SE_REQUIRE(rc == SE_OK, rc, se_gpu_last_error());
            SE_HIP(hipEventRecord(e.done, s.stream));
            SE_HIP(hipEventRecord(s.last, s.stream));
            ++s.pending;
            SE_HIP(hipStreamWaitEvent(b->join, e.done, 0));
        }
        hipEvent_t ev = b->new_event();
        SE_HIP(hipEventRecord(ev, b->join));
        b->tickets.push_back(ev);
        b->ticket_slot.push_back(-1);
        *ticket = (int)b->tickets.size() - 1;
    });
}

int se_batch_prepare(se_batch* b, const se_job_shape* shape, int n_max)
{
    if (!b || !shape || n_max < 0) return SE_ERR_ARG;
    return guarded_b([&] {
        for (int i = 0; i < (int)b->slots.size(); ++i) b->plan_for(i, *shape, std::max(n_max, 1));
        for (auto& s : b->slots) SE_HIP(hipStreamSynchronize(s.stream));
    });
}

int se_batch_poll(se_batch* b, int ticket)
{
    if (!b || ticket < 0 || ticket >= (int)b->tickets.size()) return SE_ERR_ARG;
    hipEvent_t e = b->tickets[ticket];
    if (!e) return 1;
    const hipError_t q = hipEventQuery(e);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) {
        se::set_error(std::string("se_batch_poll: ") + hipGetErrorString(q));
        return SE_ERR_HIP;
    }
    b->free_events.push_back(e);
    b->tickets[ticket] = nullptr;
    return 1;
}

int se_batch_wait(se_batch* b, int ticket)
{
    if (!b || ticket < 0 || ticket >= (int)b->tickets.size()) return SE_ERR_ARG;
    hipEvent_t e = b->tickets[ticket];
    if (!e) return SE_OK;
    return guarded_b([&] {
        SE_HIP(hipEventSynchronize(e));
        b->free_events.push_back(e);
        b->tickets[ticket] = nullptr;
    });
}

int se_batch_wait_all(se_batch* b)
{
    if (!b) return SE_ERR_ARG;
    return guarded_b([&] {
        for (auto& s : b->slots) {
            SE_HIP(hipStreamSynchronize(s.stream));
            s.pending = 0;
        }
        SE_HIP(hipStreamSynchronize(b->join));
        for (auto& e : b->tickets)
            if (e) {
                b->free_events.push_back(e);
                e = nullptr;
            }
    });
}

int se_batch_stats(se_batch* b, size_t* plan_bytes, int* n_plans)
{
    if (!b) return SE_ERR_ARG;
    int n = (int)b->groups.size();
    for (auto& s : b->slots) n += (int)s.plans.size();
    if (plan_bytes) *plan_bytes = b->used;
    if (n_plans) *n_plans = n;
    return SE_OK;
}

}  // extern "C"
