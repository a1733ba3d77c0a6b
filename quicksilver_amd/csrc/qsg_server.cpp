// Serving loop over the batched form: solve requests (input decks) are
// queued, admitted in arrival order while their device footprint fits the
// HBM budget, kept resident on the device across cycles, advanced together
// one cycle per qsg_server_step, and retired (device memory released) when
// they have run their nSteps cycles.  Continuous batching: a request joins
// or leaves the running batch at a cycle boundary by taking or freeing one
// of its slots, so work arriving while others run shares their launches and
// the batch's captured graphs stay valid; the batch is rebuilt only when it
// runs out of slots.  No reference counterpart: the reference runs one
// problem per process.
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "qsg_internal.h"

using qsgi::hipCheck;
using qsgi::QsgError;

namespace {

struct Request {
    uint64_t id = 0;
    qs::Parameters params;
    qs::ProblemTables tables;
    uint64_t bytes = 0;
    uint64_t capacity = 0;  // particle-buffer entries asked for (0: the default)
    int nSteps = 0;
    int state = 0;  // 0 queued, 1 active, 2 done, 3 failed, 4 cancelled
    qsg_problem_s* prob = nullptr;
    std::vector<qsg_balance> bal;
    std::vector<double> flux;
    std::string error;
    int errorCode = 0;  // the QSG_ERR_* a failed request reports
    int slot = -1;  // batch slot while running
};

}  // namespace

struct qsg_server_s {
    int device = 0;
    uint64_t budget = 0;
    uint64_t used = 0;
    int maxActive = 0;
    uint64_t nextId = 1;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    qsg_batch_s* batch = nullptr;  // slots for the running requests, kept across steps
    qs::TrackTuning tune = qs::kDefaultTuning;
    bool haveTune = false;
    qsgi::StackPool stacks;  // shared by the successive batches
    qsg_server_stats stats = {};
};

namespace {

void dropBatch(qsg_server_s* s) {
    if (s->batch) qsgiDestroyBatch(s->batch);
    s->batch = nullptr;
}

void retire(qsg_server_s* s, Request* r, int state, int code = 0, const std::string& error = std::string()) {
    if (state == 3) {
        r->errorCode = code ? code : QSG_ERR_DEVICE;
        r->error = error;
    }
    if (r->slot >= 0 && s->batch) {
        try {
            qsgiBatchSetMember(s->batch, r->slot, nullptr);
        } catch (const std::exception&) {
            dropBatch(s);  // a device error: the next admission builds a new batch
        }
    }
    r->slot = -1;
    if (r->prob) {
        qsgi::destroy(r->prob);
        r->prob = nullptr;
        s->used -= r->bytes;
    }
    r->state = state;
}

// A batch slot for a newly admitted request.  The batch keeps its slots
// (and captured graphs) across steps; when all are taken it is rebuilt with
// twice as many, the running requests keeping their state.
int freeSlot(qsg_server_s* s) {
    if (s->batch) {
        std::vector<bool> used(qsgiBatchCapacity(s->batch), false);
        for (Request* r : s->active)
            if (r->slot >= 0) used[r->slot] = true;
        for (size_t i = 0; i < used.size(); i++)
            if (!used[i]) return (int)i;
    }
    const int old = s->batch ? qsgiBatchCapacity(s->batch) : 0;
    const int cap = old ? 2 * old : (s->maxActive > 0 ? s->maxActive : 16);
    std::vector<qsg_problem> hs;
    for (Request* r : s->active) hs.push_back(r->prob);
    dropBatch(s);
    s->batch = qsgiCreateBatch(hs.data(), (int)hs.size(), cap, s->device, &s->stacks);
    if (s->haveTune) qsgiBatchSetTuning(s->batch, s->tune);
    for (size_t i = 0; i < s->active.size(); i++) s->active[i]->slot = (int)i;
    return (int)s->active.size();
}

// Admit queued requests in arrival order while they fit the budget (and the
// active-count cap); head-of-line requests that never fit are failed.
void admit(qsg_server_s* s) {
    while (!s->queue.empty()) {
        Request* r = s->queue.front();
        if (r->bytes > s->budget) {
            s->queue.pop_front();
            r->error = "request needs more device memory than the whole budget";
            r->errorCode = QSG_ERR_CAPACITY;
            r->state = 3;
            s->stats.rejected++;
            continue;
        }
        if (s->used + r->bytes > s->budget) break;
        if (s->maxActive > 0 && (int)s->active.size() >= s->maxActive) break;
        s->queue.pop_front();
        try {
            r->prob = qsgi::createFromTables(r->params, std::move(r->tables), s->device, false, r->capacity);
            r->tables = qs::ProblemTables();
            s->used += r->bytes;
            const int slot = freeSlot(s);
            qsgiBatchSetMember(s->batch, slot, r->prob);
            r->slot = slot;
            r->state = 1;
            s->active.push_back(r);
            s->stats.admitted++;
        } catch (const std::exception& e) {
            if (r->prob) {
                qsgi::destroy(r->prob);
                r->prob = nullptr;
                s->used -= r->bytes;
            }
            const qsgi::QsgError* q = dynamic_cast<const qsgi::QsgError*>(&e);
            r->error = e.what();
            r->errorCode = q ? q->code : QSG_ERR_DEVICE;
            r->state = 3;
            s->stats.rejected++;
        }
    }
}

// The running batch holds every active request.  It can be missing while
// requests run: retiring one may drop it after a device error, and a
// rebuild for more slots may fail after the old batch is gone.  The active
// problems are then still at a cycle boundary, so a new batch takes them.
void ensureBatch(qsg_server_s* s) {
    if (s->batch || s->active.empty()) return;
    std::vector<qsg_problem> hs;
    for (Request* r : s->active) hs.push_back(r->prob);
    int cap = s->maxActive > 0 ? s->maxActive : 16;
    while (cap < (int)hs.size()) cap *= 2;
    s->batch = qsgiCreateBatch(hs.data(), (int)hs.size(), cap, s->device, &s->stacks);
    if (s->haveTune) qsgiBatchSetTuning(s->batch, s->tune);
    for (size_t i = 0; i < s->active.size(); i++) s->active[i]->slot = (int)i;
}

void failAllActive(qsg_server_s* s, int code, const std::string& msg) {
    for (Request* r : s->active) retire(s, r, 3, code, msg);
    s->active.clear();
    dropBatch(s);
}

void serverStep(qsg_server_s* s) {
    admit(s);
    if (s->active.empty()) return;
    try {
        ensureBatch(s);
    } catch (const std::exception& e) {
        const qsgi::QsgError* q = dynamic_cast<const qsgi::QsgError*>(&e);
        failAllActive(s, q ? q->code : QSG_ERR_DEVICE, std::string("no batch for the running requests: ") + e.what());
        throw;
    }
    const int cap = qsgiBatchCapacity(s->batch);
    std::vector<qsg_balance> bal(cap);
    std::vector<double> flux(cap);
    try {
        qsgiBatchStep(s->batch, bal.data(), flux.data());
    } catch (const qsgi::QsgError& e) {
        if (qsgiBatchBroken(s->batch) || (e.code != QSG_ERR_CAPACITY && e.code != QSG_ERR_NUMERIC)) {
            // a failed step leaves every problem's device state undefined
            failAllActive(s, e.code, std::string("batch step failed: ") + e.what());
            throw;
        }
        // some problems were taken out of the batch; the others advanced
    } catch (const std::exception& e) {
        failAllActive(s, QSG_ERR_DEVICE, std::string("batch step failed: ") + e.what());
        throw;
    }
    s->stats.steps++;
    std::vector<Request*> still;
    for (Request* r : s->active) {
        if (const int code = qsgiBatchSlotError(s->batch, r->slot)) {
            int c = 0;
            char msg[512];
            qsg_problem_error(r->prob, &c, msg, sizeof msg);
            retire(s, r, 3, code, msg);
            continue;
        }
        r->bal.push_back(bal[r->slot]);
        r->flux.push_back(flux[r->slot]);
        s->stats.segments += bal[r->slot].num_segments;
        if ((int)r->bal.size() >= r->nSteps) {
            retire(s, r, 2);
            s->stats.completed++;
        } else {
            still.push_back(r);
        }
    }
    s->active.swap(still);
}

// Stops a request between steps: a queued one leaves the queue, a running
// one frees its slot and device memory (it stands at a cycle boundary).
void cancel(qsg_server_s* s, Request* r) {
    if (r->state == 0) {
        for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
            if (*it == r) {
                s->queue.erase(it);
                break;
            }
        r->state = 4;
    } else if (r->state == 1) {
        for (auto it = s->active.begin(); it != s->active.end(); ++it)
            if (*it == r) {
                s->active.erase(it);
                break;
            }
        retire(s, r, 4);
    }
}

Request* find(qsg_server_s* s, uint64_t id) {
    auto it = s->all.find(id);
    if (it == s->all.end()) throw QsgError(QSG_ERR_ARG, "unknown request id");
    return it->second.get();
}

}  // namespace

extern "C" {

int qsg_server_create(int device, uint64_t hbm_budget_bytes, int max_active, qsg_server* out) {
    QSG_GUARD({
        if (!out) throw QsgError(QSG_ERR_ARG, "out is NULL");
        if (hbm_budget_bytes == 0) throw QsgError(QSG_ERR_ARG, "budget must be positive");
        hipCheck(hipSetDevice(device), "hipSetDevice");
        qsg_server_s* s = new qsg_server_s();
        s->device = device;
        s->budget = hbm_budget_bytes;
        s->maxActive = max_active;
        *out = s;
    })
}

int qsg_server_destroy(qsg_server s) {
    QSG_GUARD({
        if (!s) return QSG_OK;
        for (Request* r : s->active) retire(s, r, 4);
        dropBatch(s);
        (void)hipSetDevice(s->device);
        s->stacks.release();
        delete s;
    })
}

int qsg_server_submit(qsg_server s, const char* deck_text, uint64_t* request_id) {
    return qsg_server_submit_opts(s, deck_text, nullptr, request_id);
}

int qsg_server_submit_opts(qsg_server s, const char* deck_text, const qsg_options* opts, uint64_t* request_id) {
    QSG_GUARD({
        if (!s || !deck_text || !request_id) throw QsgError(QSG_ERR_ARG, "NULL argument");
        std::unique_ptr<Request> r(new Request());
        const char* argv0[] = {"qs"};
        r->params = qs::getParameters(1, argv0);
        r->params.geometryParams.clear();
        r->params.materialParams.clear();
        r->params.crossSectionParams.clear();
        qs::parseDeckText(deck_text, r->params);
        qs::supplyDefaults(r->params);
        r->tables = qs::buildProblem(r->params);
        r->capacity = opts ? opts->particle_capacity : 0;
        r->bytes = qsgi::deviceBytesEstimate(r->tables, r->capacity);
        r->nSteps = r->tables.nSteps;
        r->id = s->nextId++;
        *request_id = r->id;
        s->queue.push_back(r.get());
        s->all[r->id] = std::move(r);
        s->stats.submitted++;
    })
}

int qsg_server_step(qsg_server s, int* n_active, int* n_queued) {
    QSG_GUARD({
        if (!s) throw QsgError(QSG_ERR_ARG, "NULL server");
        serverStep(s);
        admit(s);
        if (n_active) *n_active = (int)s->active.size();
        if (n_queued) *n_queued = (int)s->queue.size();
    })
}

int qsg_server_status(qsg_server s, uint64_t id, int* state, int* cycles_done) {
    QSG_GUARD({
        if (!s) throw QsgError(QSG_ERR_ARG, "NULL server");
        Request* r = find(s, id);
        if (state) *state = r->state;
        if (cycles_done) *cycles_done = (int)r->bal.size();
    })
}

int qsg_server_result(qsg_server s, uint64_t id, qsg_balance* balances, double* flux_sums, int n) {
    QSG_GUARD({
        if (!s) throw QsgError(QSG_ERR_ARG, "NULL server");
        Request* r = find(s, id);
        if (r->state == 3) throw QsgError(r->errorCode ? r->errorCode : QSG_ERR_STATE, "request failed: " + r->error);
        if (n < 0 || n > (int)r->bal.size()) throw QsgError(QSG_ERR_ARG, "more cycles requested than completed");
        for (int i = 0; i < n; i++) {
            if (balances) balances[i] = r->bal[i];
            if (flux_sums) flux_sums[i] = r->flux[i];
        }
    })
}

int qsg_server_cancel(qsg_server s, uint64_t id) {
    QSG_GUARD({
        if (!s) throw QsgError(QSG_ERR_ARG, "NULL server");
        Request* r = find(s, id);
        if (r->state != 0 && r->state != 1) throw QsgError(QSG_ERR_STATE, "request is not queued or running");
        cancel(s, r);
    })
}

int qsg_server_release(qsg_server s, uint64_t id) {
    QSG_GUARD({
        if (!s) throw QsgError(QSG_ERR_ARG, "NULL server");
        Request* r = find(s, id);
        cancel(s, r);
        s->all.erase(id);
    })
}

int qsg_server_set_tuning(qsg_server s, const qsg_tuning* t) {
    QSG_GUARD({
        if (!s || !t) throw QsgError(QSG_ERR_ARG, "NULL argument");
        s->tune = qsgi::tuningFromC(t);
        s->haveTune = true;
        if (s->batch) qsgiBatchSetTuning(s->batch, s->tune);
    })
}

int qsg_server_get_stats(qsg_server s, qsg_server_stats* stats) {
    QSG_GUARD({
        if (!s || !stats) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *stats = s->stats;
        stats->bytes_in_use = s->used;
        stats->active = (uint64_t)s->active.size();
        stats->queued = (uint64_t)s->queue.size();
    })
}

}  // extern "C"
