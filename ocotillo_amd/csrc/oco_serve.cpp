/*
 * oco_serve.cpp -- the serving loop of include/ocotillo.h (oco_server_*).
 *
 * One scheduler thread owns the launches, one copy thread the transfers:
 *   admit    queued jobs in arrival order while a job slot is free and the
 *            job's buffers fit the arena (one device allocation of the HBM
 *            budget, first-fit, so admission never calls hipMalloc); the copy
 *            thread uploads its inputs and builds its cell records on the copy
 *            stream (a host copy from pageable memory blocks only that thread);
 *   iterate  form one launch from the tiles (column, 64 wavelengths) of every
 *            admitted job whose records are ready, shared round-robin so each
 *            active job progresses every iteration, at most two launches in
 *            flight on the compute stream;
 *   retire   when a launch completes, count its tiles against their jobs; a
 *            job with all tiles done has its emergent flux copied out and its
 *            arena blocks freed, which admits the next queued job.
 * Wavelength tables stay resident in the arena, one per distinct grid, while
 * a job uses them (and after, until the room is needed); a new grid's table
 * is built on the host at admission and uploaded by the copy thread ahead of
 * its first job's inputs.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/ocotillo.h"
#include "oco_internal.h"
#include "oco_physics.h"

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point t0, clk::time_point t1) { return std::chrono::duration<double>(t1 - t0).count(); }

constexpr int MAX_SLOTS = 256;
constexpr double ITER_POINTS = 1.5e9;   /* points per launch: ~30 ms of work, bounds a newcomer's wait */

/* first-fit allocator over one device allocation */
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;   /* offset -> bytes */
    void init(void* p, size_t n)
    {
        base = (char*)p;
        size = n;
        free_.clear();
        free_[0] = n;
    }
    static size_t round(size_t n) { return (n + 255) & ~(size_t)255; }
    bool alloc(size_t n, size_t& off)
    {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                used += n;
                peak = std::max(peak, used);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = round(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

enum JobState { QUEUED, LOADING, RUNNING, COPYING, DONE };

struct Job {
    long long ticket = 0;
    oco_job desc{};
    long ncol = 0, nsys = 0, ntiles = 0, issued = 0, done = 0;
    int nblk = 0;   /* wavelength blocks of 64 */
    int slot = -1;
    size_t off = 0, bytes = 0;
    const double* d_wtab = nullptr;
    int table = -1;   /* index of its wavelength table in oco_server::tables while admitted */
    std::shared_ptr<std::vector<double>> wt_host;   /* the host table it uploaded, kept until it is done */
    double* d_em = nullptr;
    hipEvent_t ready = nullptr, copied = nullptr;
    JobState state = QUEUED;
    bool loaded = false, copy_issued = false;   /* events recorded by the copy thread */
    clk::time_point t_submit, t_admit, t_done;
    oco_job_result res{};
};

struct Iter {
    std::vector<std::pair<Job*, long>> tiles;   /* tiles of each job in this launch */
    hipEvent_t done = nullptr;
    int buf = 0;
};

struct Table {
    std::vector<double> waves;
    size_t off = 0, bytes = 0;
    int users = 0;    /* admitted jobs on this grid */
    bool live = false;
    /* the host table until the copy thread has uploaded it (with the
     * inputs of the first job on the grid, ahead of that job's records) */
    std::shared_ptr<std::vector<double>> pending;
};

/* Arena bytes of a job's own buffers: rho and T, cell records (mixed:
 * the serving kernel's chunk-major layout), emergent flux. */
size_t job_bytes(bool mixed, long ncol, int nz, long nsys)
{
    const size_t ncells = (size_t)ncol * nz;
    const size_t b_in = ncells * sizeof(double);
    const size_t b_cells = (mixed ? oco_cells_chunked_doubles(ncol, nz) : ncells * OCO_CELL_STRIDE) * sizeof(double);
    const size_t b_em = (size_t)nsys * 2 * sizeof(double);
    return Arena::round(b_in) * 2 + Arena::round(b_cells) + Arena::round(b_em);
}

size_t table_bytes(int nw) { return Arena::round((size_t)nw * OCO_WT_STRIDE * sizeof(double)); }

}  // namespace

struct oco_server {
    oco_ctx* ctx = nullptr;
    int device = 0, mixed = 0, max_nz = 0, nblocks = 0;
    double fully_ionized_T = 2e4;
    hipStream_t cs = nullptr, xs = nullptr;   /* compute, copy */
    void* mem = nullptr;
    Arena arena;
    double* scratch = nullptr;
    unsigned* counter = nullptr;
    unsigned long long* fail = nullptr;        /* per slot, device */
    unsigned long long* h_fail = nullptr;      /* pinned */
    oco_serve_job* d_jobs = nullptr;
    oco_serve_job* h_jobs[2] = {nullptr, nullptr};   /* pinned, double-buffered */
    int* d_tiles[2] = {nullptr, nullptr};
    int* h_tiles[2] = {nullptr, nullptr};
    size_t tile_cap = 0;
    std::vector<oco_serve_job> jobs;            /* slot descriptors, host copy */
    std::vector<int> free_slots;
    std::vector<Table> tables;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::vector<Job*> active;                   /* admitted, in admission order */
    std::map<long long, Job*> all;
    std::deque<Iter> inflight;
    std::thread th, copier;
    std::deque<std::pair<Job*, int>> copy_tasks;   /* (job, 0 upload / 1 download) */
    std::condition_variable cv_copy;
    bool stop = false, copier_stop = false;
    int fatal = OCO_OK;
    long long next_ticket = 1, submitted = 0, completed = 0, iterations = 0;
    std::vector<double> lat_ms;
    double points_done = 0.0;
    bool any_submit = false;
    clk::time_point t_first, t_last;
    int buf = 0;

    bool copier_busy = false;                  /* the copy thread holds a task outside the lock */
    std::condition_variable cv_copier_idle;

    void loop(std::unique_lock<std::mutex>& lk);
    bool admit_one(Job* j);
    int table_for(const double* waves, int nw);
    void drop_unused_tables();
    bool launch_iteration();
    void retire();
    void finish_copies();
    void fail_all(int st, std::unique_lock<std::mutex>& lk);
    void copy_loop();
};

/* index of the resident table of this wavelength grid, built in the arena
 * if it is not there yet (on the host here; the copy thread uploads it with
 * the job's inputs, so the scheduler never waits on a transfer); -1 if it
 * does not fit */
int oco_server::table_for(const double* waves, int nw)
{
    for (size_t i = 0; i < tables.size(); i++) {
        const Table& t = tables[i];
        if (t.live && (int)t.waves.size() == nw && std::equal(t.waves.begin(), t.waves.end(), waves)) return (int)i;
    }
    Table t;
    t.waves.assign(waves, waves + nw);
    auto wt = std::make_shared<std::vector<double>>();
    oco_build_wtab(t.waves, *wt);
    t.bytes = wt->size() * sizeof(double);
    if (!arena.alloc(t.bytes, t.off)) return -1;
    t.pending = std::move(wt);
    t.live = true;
    for (size_t i = 0; i < tables.size(); i++)
        if (!tables[i].live) {
            tables[i] = std::move(t);
            return (int)i;
        }
    tables.push_back(std::move(t));
    return (int)tables.size() - 1;
}

/* free the resident tables no admitted job uses (their grids are built
 * again when a later job needs them); mu held */
void oco_server::drop_unused_tables()
{
    for (Table& t : tables)
        if (t.live && t.users == 0) {
            arena.release(t.off, t.bytes);
            t.live = false;
            t.waves.clear();
            t.pending.reset();
        }
}

/* called with mu held; false if the job does not fit yet (its table and
 * buffers: first with the resident tables, then with the unused ones freed) */
bool oco_server::admit_one(Job* j)
{
    if (free_slots.empty()) return false;
    const oco_job& d = j->desc;
    const long ncells = j->ncol * d.nz;
    const size_t b_in = (size_t)ncells * sizeof(double);
    const size_t b_cells = (mixed ? oco_cells_chunked_doubles(j->ncol, d.nz) : (size_t)ncells * OCO_CELL_STRIDE) *
                           sizeof(double);
    j->bytes = job_bytes(mixed, j->ncol, d.nz, j->nsys);
    int ti = table_for(d.waves, d.nw);
    if (ti < 0 || !arena.alloc(j->bytes, j->off)) {
        /* free the tables of grids no admitted job uses (this job's own if it
         * was just built: it is built again) and try once more */
        drop_unused_tables();
        ti = table_for(d.waves, d.nw);
        if (ti < 0 || !arena.alloc(j->bytes, j->off)) return false;
    }
    tables[ti].users++;
    j->table = ti;
    const double* wt = (const double*)(arena.base + tables[ti].off);
    j->slot = free_slots.back();
    free_slots.pop_back();
    j->d_wtab = wt;
    char* p = arena.base + j->off;
    double* d_cells = (double*)(p + 2 * Arena::round(b_in));
    j->d_em = (double*)(p + 2 * Arena::round(b_in) + Arena::round(b_cells));
    copy_tasks.push_back({j, 0});
    cv_copy.notify_one();
    oco_serve_job& s = jobs[j->slot];
    s.cells = d_cells;
    s.wtab = wt;
    s.em = j->d_em;
    s.ncol = j->ncol;
    s.nz = d.nz;
    s.nw = d.nw;
    s.dz = d.dz;
    s.dz1 = 1.0 / d.dz;
    s.dz2 = d.dz * d.dz;
    j->state = LOADING;
    j->t_admit = clk::now();
    return true;
}

/* one launch over the tiles of every job whose records are ready; mu held */
bool oco_server::launch_iteration()
{
    std::vector<Job*> ready;
    for (Job* j : active) {
        if (j->state == LOADING && j->loaded && hipEventQuery(j->ready) == hipSuccess) j->state = RUNNING;
        if (j->state == RUNNING && j->issued < j->ntiles) ready.push_back(j);
    }
    if (ready.empty()) return false;
    const int b = buf;
    buf ^= 1;
    Iter it;
    it.buf = b;
    int* ht = h_tiles[b];
    size_t nt = 0;
    /* round-robin shares of the iteration's points, then the rest in order */
    double budget = ITER_POINTS;
    std::vector<long> take(ready.size(), 0);
    long planned = 0;   /* tiles taken so far; the staging buffers hold tile_cap */
    for (int pass = 0; pass < 2 && budget > 0; pass++) {
        const double share = pass == 0 ? budget / (double)ready.size() : budget;
        for (size_t r = 0; r < ready.size() && budget > 0; r++) {
            Job* j = ready[r];
            const double pts_tile = 64.0 * j->desc.nz;
            long avail = j->ntiles - j->issued - take[r];
            long want = std::max(1L, (long)(std::min(share, budget) / pts_tile));
            long n = std::min(std::min(avail, want), (long)tile_cap - planned);
            if (n <= 0) continue;
            take[r] += n;
            planned += n;
            budget -= n * pts_tile;
        }
    }
    for (size_t r = 0; r < ready.size(); r++) {
        Job* j = ready[r];
        for (long k = 0; k < take[r]; k++) {
            const long t = j->issued + k;
            ht[3 * nt] = j->slot;
            ht[3 * nt + 1] = (int)(t / j->nblk);
            ht[3 * nt + 2] = (int)(t % j->nblk) * 64;
            nt++;
        }
        if (take[r]) {
            j->issued += take[r];
            it.tiles.push_back({j, take[r]});
        }
    }
    if (nt == 0) return false;
    memcpy(h_jobs[b], jobs.data(), jobs.size() * sizeof(oco_serve_job));
    oco_serve_args a;
    a.ntiles = (int)nt;
    a.tiles = d_tiles[b];
    a.jobs = d_jobs;
    a.scratch = scratch;
    a.scr_stride = (long)max_nz * 3 * 64;
    a.counter = counter;
    a.fail = fail;
    a.win_rows = oco_serve_win_rows(max_nz);
    /* the tiles' records were built on the copy stream */
    hipError_t e = hipSuccess;
    for (auto& pr : it.tiles)
        if (e == hipSuccess) e = hipStreamWaitEvent(cs, pr.first->ready, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tiles[b], ht, nt * 3 * sizeof(int), hipMemcpyHostToDevice, cs);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_jobs, h_jobs[b], jobs.size() * sizeof(oco_serve_job), hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = oco_launch_serve(a, mixed, nblocks, cs);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&it.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(it.done, cs);
    if (e != hipSuccess) {
        fatal = OCO_E_HIP;
        return false;
    }
    iterations++;
    inflight.push_back(std::move(it));
    return true;
}

/* completed launches: count tiles, start output copies; mu held */
void oco_server::retire()
{
    while (!inflight.empty() && hipEventQuery(inflight.front().done) == hipSuccess) {
        Iter& it = inflight.front();
        for (auto& pr : it.tiles) {
            Job* j = pr.first;
            j->done += pr.second;
            if (j->done == j->ntiles) {
                j->state = COPYING;
                copy_tasks.push_back({j, 1});
                cv_copy.notify_one();
            }
        }
        (void)hipEventDestroy(it.done);
        inflight.pop_front();
    }
}

/* jobs whose outputs have landed; mu held */
void oco_server::finish_copies()
{
    for (size_t i = 0; i < active.size();) {
        Job* j = active[i];
        if (j->state == COPYING && j->copy_issued && hipEventQuery(j->copied) == hipSuccess) {
            j->t_done = clk::now();
            unsigned long long key = h_fail[j->slot];
            j->res.status = OCO_OK;
            if (key != OCO_FAIL_NONE) {
                const long sys = (long)(key >> 32);   /* column * nw + iw */
                const long col = sys / j->desc.nw;
                j->res.status = (int)((key >> 28) & 0xf);
                j->res.fail_j = (int)(key & 0x0fffffffULL);
                j->res.fail_ix = (int)(col / j->desc.ny) + 1;
                j->res.fail_iy = (int)(col % j->desc.ny) + 1;
                j->res.fail_iw = (int)(sys % j->desc.nw) + 1;
            }
            j->res.queue_ms = 1e3 * secs(j->t_submit, j->t_admit);
            j->res.latency_ms = 1e3 * secs(j->t_submit, j->t_done);
            j->res.points = (double)j->nsys * j->desc.nz;
            lat_ms.push_back(j->res.latency_ms);
            points_done += j->res.points;
            t_last = j->t_done;
            completed++;
            arena.release(j->off, j->bytes);
            if (j->table >= 0) tables[j->table].users--;
            j->table = -1;
            free_slots.push_back(j->slot);
            j->slot = -1;
            j->state = DONE;
            active.erase(active.begin() + i);
            cv_done.notify_all();
        } else {
            i++;
        }
    }
}

/* A HIP error: every unfinished job fails with st.  Before any is marked
 * done (after which its waiter deletes it), the copy thread lets go of the
 * task it holds, the tasks not started are dropped, and both streams are
 * drained, so no launch, copy or thread still reads a job. */
void oco_server::fail_all(int st, std::unique_lock<std::mutex>& lk)
{
    copy_tasks.clear();
    cv_copier_idle.wait(lk, [&] { return !copier_busy; });
    (void)hipStreamSynchronize(cs);
    (void)hipStreamSynchronize(xs);
    (void)hipGetLastError();
    for (Iter& it : inflight) (void)hipEventDestroy(it.done);
    inflight.clear();
    for (auto& kv : all)
        if (kv.second->state != DONE) {
            kv.second->res.status = st;
            kv.second->state = DONE;
        }
    for (Job* j : active) {
        if (j->table >= 0) tables[j->table].users--;
        j->table = -1;
        if (j->slot >= 0) arena.release(j->off, j->bytes), free_slots.push_back(j->slot);
        j->slot = -1;
    }
    queue.clear();
    active.clear();
    cv_done.notify_all();
}

/* the copy thread: uploads (inputs, cell records) and downloads (emergent
 * flux, failure key) on the copy stream, in submission order */
void oco_server::copy_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_copy.wait(lk, [&] { return copier_stop || !copy_tasks.empty(); });
        if (copy_tasks.empty()) return;
        auto task = copy_tasks.front();
        copy_tasks.pop_front();
        copier_busy = true;
        Job* j = task.first;
        const oco_job d = j->desc;
        const long ncells = j->ncol * d.nz;
        const size_t b_in = (size_t)ncells * sizeof(double);
        char* p = arena.base + j->off;
        const int slot = j->slot;
        /* the job's wavelength table, if this is its grid's first upload */
        std::shared_ptr<std::vector<double>> wt;
        void* d_wt = nullptr;
        if (task.second == 0 && j->table >= 0 && tables[j->table].pending) {
            wt = std::move(tables[j->table].pending);
            d_wt = arena.base + tables[j->table].off;
        }
        lk.unlock();
        hipError_t e = hipSuccess;
        if (task.second == 0) {
            double* d_rho = (double*)p;
            double* d_T = (double*)(p + Arena::round(b_in));
            double* d_cells = (double*)(p + 2 * Arena::round(b_in));
            if (wt) e = hipMemcpyAsync(d_wt, wt->data(), wt->size() * sizeof(double), hipMemcpyHostToDevice, xs);
            if (e == hipSuccess) e = hipMemcpyAsync(d_rho, d.rho3d, b_in, hipMemcpyHostToDevice, xs);
            if (e == hipSuccess) e = hipMemcpyAsync(d_T, d.T3d, b_in, hipMemcpyHostToDevice, xs);
            if (e == hipSuccess)
                e = mixed ? oco_launch_cells_chunked(d_rho, d_T, ncells, d.nz, fully_ionized_T, d_cells, xs)
                          : oco_launch_cells(d_rho, d_T, ncells, fully_ionized_T, d_cells, 0, xs);
            if (e == hipSuccess) e = oco_launch_fail_reset(fail + slot, 1, xs);
            if (e == hipSuccess) e = hipEventRecord(j->ready, xs);
        } else {
            e = hipMemcpyAsync(h_fail + slot, fail + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost, xs);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d.emergent, j->d_em, (size_t)j->nsys * 2 * sizeof(double), hipMemcpyDeviceToHost,
                                   xs);
            if (e == hipSuccess) e = hipEventRecord(j->copied, xs);
        }
        lk.lock();
        if (e != hipSuccess) fatal = OCO_E_HIP;
        if (wt) j->wt_host = std::move(wt);
        if (task.second == 0) j->loaded = true;
        else j->copy_issued = true;
        copier_busy = false;
        cv_copier_idle.notify_all();
        cv_work.notify_one();
    }
}

void oco_server::loop(std::unique_lock<std::mutex>& lk)
{
    for (;;) {
        if (fatal) {
            fail_all(fatal, lk);
            if (stop) return;
            cv_work.wait(lk);
            continue;
        }
        retire();
        finish_copies();
        while (!queue.empty()) {
            Job* h = queue.front();
            if (admit_one(h)) {
                active.push_back(h);
                queue.pop_front();
                continue;
            }
            if (active.empty() && inflight.empty() && free_slots.size() == MAX_SLOTS) {
                /* the arena holds no job and no used table, and the job still
                 * does not fit: it never will */
                h->res.status = OCO_E_OOM;
                h->t_admit = h->t_done = clk::now();
                h->res.queue_ms = h->res.latency_ms = 1e3 * secs(h->t_submit, h->t_done);
                h->state = DONE;
                queue.pop_front();
                cv_done.notify_all();
                continue;
            }
            break;
        }
        bool launched = false;
        if (inflight.size() < 2) launched = launch_iteration();
        if (stop && queue.empty() && active.empty() && inflight.empty()) return;
        if (!launched) {
            /* wait for a submission or for device progress */
            cv_work.wait_for(lk, std::chrono::microseconds(200));
        }
    }
}

extern "C" {

int oco_server_create(oco_ctx* ctx, size_t hbm_budget, int max_nz, int mode, oco_server** out)
{
    if (!ctx || !out || hbm_budget < ((size_t)1 << 20) || max_nz < 4) return OCO_E_ARG;
    if (mode != OCO_MODE_FP64 && mode != OCO_MODE_MIXED) return OCO_E_ARG;
    *out = nullptr;
    oco_server* s = new (std::nothrow) oco_server;
    if (!s) return OCO_E_OOM;
    s->ctx = ctx;
    s->device = oco_ctx_device(ctx);
    s->fully_ionized_T = oco_ctx_fully_ionized_T(ctx);
    s->mixed = mode == OCO_MODE_MIXED;
    s->max_nz = max_nz;
    int st = OCO_OK;
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && st == OCO_OK) {
            (void)hipGetLastError();
            st = e == hipErrorOutOfMemory ? OCO_E_OOM : OCO_E_HIP;
        }
    };
    chk(hipSetDevice(s->device));
    chk(hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking));
    chk(hipStreamCreateWithFlags(&s->xs, hipStreamNonBlocking));
    s->nblocks = oco_serve_grid(s->mixed, max_nz, s->device);
    s->tile_cap = (size_t)(ITER_POINTS / (64.0 * 4)) + 4 * MAX_SLOTS;
    if (st == OCO_OK) chk(hipMalloc(&s->mem, hbm_budget));
    /* elimination scratch: FP64 mode only (the mixed kernel keeps its state in registers) */
    if (st == OCO_OK && !s->mixed)
        chk(hipMalloc((void**)&s->scratch, (size_t)s->nblocks * max_nz * 3 * 64 * sizeof(double)));
    if (st == OCO_OK) chk(hipMalloc((void**)&s->counter, sizeof(unsigned)));
    if (st == OCO_OK) chk(hipMalloc((void**)&s->fail, MAX_SLOTS * sizeof(unsigned long long)));
    if (st == OCO_OK) chk(hipMalloc((void**)&s->d_jobs, MAX_SLOTS * sizeof(oco_serve_job)));
    if (st == OCO_OK) chk(hipHostMalloc((void**)&s->h_fail, MAX_SLOTS * sizeof(unsigned long long), 0));
    for (int b = 0; b < 2 && st == OCO_OK; b++) {
        chk(hipMalloc((void**)&s->d_tiles[b], s->tile_cap * 3 * sizeof(int)));
        if (st == OCO_OK) chk(hipHostMalloc((void**)&s->h_tiles[b], s->tile_cap * 3 * sizeof(int), 0));
        if (st == OCO_OK) chk(hipHostMalloc((void**)&s->h_jobs[b], MAX_SLOTS * sizeof(oco_serve_job), 0));
    }
    if (st != OCO_OK) {
        oco_server_destroy(s);
        return st;
    }
    s->arena.init(s->mem, hbm_budget);
    s->jobs.assign(MAX_SLOTS, oco_serve_job{});
    for (int i = MAX_SLOTS - 1; i >= 0; i--) s->free_slots.push_back(i);
    s->copier = std::thread([s] { s->copy_loop(); });
    s->th = std::thread([s] {
        (void)hipSetDevice(s->device);
        std::unique_lock<std::mutex> lk(s->mu);
        s->loop(lk);
    });
    *out = s;
    return OCO_OK;
}

int oco_server_submit(oco_server* s, const oco_job* job, long long* ticket)
{
    if (!s || !job || !ticket) return OCO_E_ARG;
    const oco_job& d = *job;
    if (d.nx < 1 || d.ny < 1 || d.nz < 4 || d.nz > s->max_nz || d.nw < 1 || !(d.dz > 0.0) || !d.rho3d ||
        !d.T3d || !d.waves || !d.emergent)
        return OCO_E_ARG;
    for (int i = 0; i < d.nw; i++)
        if (!(d.waves[i] > 0.0) || !std::isfinite(d.waves[i])) return OCO_E_ARG;
    Job* j = new (std::nothrow) Job;
    if (!j) return OCO_E_OOM;
    j->desc = d;
    j->ncol = (long)d.nx * d.ny;
    j->nsys = j->ncol * d.nw;
    j->nblk = (d.nw + 63) / 64;
    j->ntiles = j->ncol * j->nblk;
    /* what admission needs: the job's buffers and its grid's table (a job
     * that passes this can still be refused at admission by the first-fit
     * arena's fragmentation, with OCO_E_OOM through oco_server_wait) */
    const size_t need = job_bytes(s->mixed, j->ncol, d.nz, j->nsys) + table_bytes(d.nw);
    if (need > s->arena.size) {
        delete j;
        return OCO_E_OOM;
    }
    if (hipEventCreateWithFlags(&j->ready, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&j->copied, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        delete j;
        return OCO_E_HIP;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) {
        (void)hipEventDestroy(j->ready);
        (void)hipEventDestroy(j->copied);
        delete j;
        return s->fatal;
    }
    j->ticket = s->next_ticket++;
    j->t_submit = clk::now();
    if (!s->any_submit) {
        s->t_first = j->t_submit;
        s->any_submit = true;
    }
    s->all[j->ticket] = j;
    s->queue.push_back(j);
    s->submitted++;
    *ticket = j->ticket;
    s->cv_work.notify_one();
    return OCO_OK;
}

static int take_result(oco_server* s, std::map<long long, Job*>::iterator it, oco_job_result* res)
{
    Job* j = it->second;
    if (res) *res = j->res;
    int st = j->res.status;
    (void)hipEventDestroy(j->ready);
    (void)hipEventDestroy(j->copied);
    s->all.erase(it);
    delete j;
    return st;
}

int oco_server_wait(oco_server* s, long long ticket, oco_job_result* res)
{
    if (!s) return OCO_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return OCO_E_ARG;
    Job* j = it->second;
    s->cv_done.wait(lk, [&] { return j->state == DONE; });
    return take_result(s, s->all.find(ticket), res);
}

int oco_server_poll(oco_server* s, long long ticket, oco_job_result* res)
{
    if (!s) return OCO_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return OCO_E_ARG;
    if (it->second->state != DONE) return OCO_E_BUSY;
    return take_result(s, it, res);
}

int oco_server_stats(oco_server* s, oco_serve_stats* st)
{
    if (!s || !st) return OCO_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->queued = (long long)s->queue.size();
    st->active = (long long)s->active.size();
    st->iterations = s->iterations;
    st->hbm_budget = s->arena.size;
    st->hbm_peak = s->arena.peak;
    st->points = s->points_done;
    if (!s->lat_ms.empty()) {
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            double x = p * (double)(v.size() - 1);
            size_t i = (size_t)x;
            double f = x - (double)i;
            return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = v.back();
        double sum = 0.0;
        for (double x : v) sum += x;
        st->mean_ms = sum / (double)v.size();
        st->wall_s = secs(s->t_first, s->t_last);
        st->points_per_s = st->wall_s > 0 ? st->points / st->wall_s : 0.0;
    }
    return OCO_OK;
}

int oco_server_destroy(oco_server* s)
{
    if (!s) return OCO_OK;
    if (s->th.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->cv_work.notify_all();
        }
        s->th.join();
    }
    if (s->copier.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->copier_stop = true;
            s->cv_copy.notify_all();
        }
        s->copier.join();
    }
    (void)hipSetDevice(s->device);
    if (s->cs) (void)hipStreamSynchronize(s->cs);
    if (s->xs) (void)hipStreamSynchronize(s->xs);
    for (auto& kv : s->all) {
        (void)hipEventDestroy(kv.second->ready);
        (void)hipEventDestroy(kv.second->copied);
        delete kv.second;
    }
    for (int b = 0; b < 2; b++) {
        if (s->d_tiles[b]) (void)hipFree(s->d_tiles[b]);
        if (s->h_tiles[b]) (void)hipHostFree(s->h_tiles[b]);
        if (s->h_jobs[b]) (void)hipHostFree(s->h_jobs[b]);
    }
    if (s->h_fail) (void)hipHostFree(s->h_fail);
    if (s->d_jobs) (void)hipFree(s->d_jobs);
    if (s->fail) (void)hipFree(s->fail);
    if (s->counter) (void)hipFree(s->counter);
    if (s->scratch) (void)hipFree(s->scratch);
    if (s->mem) (void)hipFree(s->mem);
    if (s->cs) (void)hipStreamDestroy(s->cs);
    if (s->xs) (void)hipStreamDestroy(s->xs);
    delete s;
    return OCO_OK;
}

}  /* extern "C" */
