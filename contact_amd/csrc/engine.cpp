// engine.cpp -- host orchestration of the batched HIP contact solve.
//
// Each patch's case runs as a C++20 coroutine whose control flow follows the reference solve
// (contac/panprc -> snorm/normcg -> stang/gdsteady|tangcg).  Device work is requested through
// the scheduler: requests that need no result are queued and the coroutine continues; requests
// that return reduction values suspend it.  When every coroutine is suspended the scheduler
// launches the queued requests level by level -- one kernel per operation kind for all patches
// that asked for it -- then synchronises once, copies the reductions back and resumes the
// coroutines.  A single patch and a batch of thousands run through the same code.
#include "engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <coroutine>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <initializer_list>
#include <map>
#include <stdexcept>

#include "kernels.hpp"
#include "ops.hpp"

namespace cntc {

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                                     __FILE__ + ":" + std::to_string(__LINE__));                  \
    } while (0)

static const double PI = 3.14159265358979323846264338327950288;
static const double TINY = 1e-20;   // m_globals.f90:37
static double sp(float x) { return (double)x; }

enum { ERR_NORM = -27, ERR_TANG = -28, ERR_OTHER = -99 };

DeviceArena::~DeviceArena()
{
    if (base) (void)hipFree(base);
}

void DeviceArena::ensure(size_t n)
{
    if (n <= bytes) return;
    if (base) HIPCHK(hipFree(base));
    base = nullptr;
    HIPCHK(hipMalloc(&base, n));
    bytes = n;
}

// opt_fft_size, src/m_aijpj.f90:1022-1119
int opt_fft_size(int fft_size)
{
    static const int kpat[9][8] = {
        {-2, 1, 0, 0, 0, 4, 3, 3},   {-5, 3, 0, 0, 0, 32, 27, 3}, {-1, -1, 1, 0, 0, 6, 5, 5},
        {-4, 1, 1, 0, 0, 16, 15, 5}, {0, -3, 2, 0, 0, 27, 25, 5}, {-3, 0, 0, 1, 0, 8, 7, 7},
        {1, -1, -1, 1, 0, 15, 14, 7}, {-1, 2, -1, 0, 0, 10, 9, 5}, {-2, -1, 0, 0, 1, 12, 11, 11}};
    int k2 = (int)std::ceil(std::log(1.0 * fft_size) / std::log(2.0));
    long prod = 1L << k2;
    int kf[5] = {k2, 0, 0, 0, 0};
    bool changed = true;
    while (changed) {
        changed = false;
        for (int ip = 0; ip < 9; ip++) {
            if (kpat[ip][5] <= 0 || kpat[ip][7] > 7) continue;
            for (;;) {
                bool ok = true;
                for (int f = 0; f < 5; f++)
                    if (kf[f] + kpat[ip][f] < 0) ok = false;
                if (!ok || (long)kpat[ip][6] * prod < (long)kpat[ip][5] * fft_size) break;
                for (int f = 0; f < 5; f++) kf[f] += kpat[ip][f];
                prod = (long)kpat[ip][6] * prod / kpat[ip][5];
                changed = true;
            }
        }
    }
    return (int)prod;
}

// ------------------------------------------------------------------------------------------
// coroutine task
// ------------------------------------------------------------------------------------------

struct Task {
    struct promise_type {
        std::coroutine_handle<> cont;
        int rc = 0;
        Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
        std::suspend_always initial_suspend() noexcept { return {}; }
        struct Final {
            bool await_ready() noexcept { return false; }
            std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept
            {
                auto c = h.promise().cont;
                return c ? c : std::noop_coroutine();
            }
            void await_resume() noexcept {}
        };
        Final final_suspend() noexcept { return {}; }
        void return_value(int v) { rc = v; }
        void unhandled_exception() { std::terminate(); }
    };
    std::coroutine_handle<promise_type> h;
    explicit Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Task(Task&& o) noexcept : h(o.h) { o.h = nullptr; }
    Task& operator=(Task&& o) noexcept
    {
        if (this != &o) {
            if (h) h.destroy();
            h = o.h;
            o.h = nullptr;
        }
        return *this;
    }
    Task(const Task&) = delete;
    ~Task()
    {
        if (h) h.destroy();
    }
    bool await_ready() { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> parent)
    {
        h.promise().cont = parent;
        return h;
    }
    int await_resume() { return h.promise().rc; }
};

// ------------------------------------------------------------------------------------------
// scheduler
// ------------------------------------------------------------------------------------------

union AnyItem {
    VecItem vec;
    NcgItem ncg;
    AijItem aij;
    TcgItem tcg;
    StItem st;
    GdItem gd;
    ConvItem conv;
    CoefItem coef;
    EmbedItem emb;
    ExtractItem ext;
    SetupItem setup;
    GsItem gs;
    CgsItem cgs;
    DenseItem dense;
    SensItem sens;
    LedgeItem ledge;
    AnyItem() { std::memset(this, 0, sizeof(*this)); }
};

static size_t item_size(int kind)
{
    switch (kind) {
    case OP_VEC: return sizeof(VecItem);
    case OP_NCG_NEGPN: case OP_NCG_DD: case OP_NCG_EXPAND: case OP_NCG_HMIN:
    case OP_SN_CONTRACT: case OP_SN_EXPAND: case OP_SN_FINAL: return sizeof(NcgItem);
    case OP_AIJ_ONES: return sizeof(AijItem);
    case OP_TCG_NT: case OP_TCG_RSET: case OP_TCG_ZPREP: case OP_TCG_QPOST: case OP_TCG_PUPD:
    case OP_TCG_AS: case OP_TCG_SA: case OP_TCG_VUPD: return sizeof(TcgItem);
    case OP_ST_AS: case OP_ST_SA: return sizeof(StItem);
    case OP_GD_DP: case OP_GD_PROJECT: case OP_GD_PROBE: case OP_GD_UPDATE: case OP_GD_MOVES:
    case OP_GD_DIAGRES: return sizeof(GdItem);
    case OP_CONV: case OP_CONV32: return sizeof(ConvItem);
    case OP_ELASCF: return sizeof(CoefItem);
    case OP_EMBED: return sizeof(EmbedItem);
    case OP_EXTRACT: return sizeof(ExtractItem);
    case OP_SETUP: case OP_FORCES: return sizeof(SetupItem);
    case OP_STDYGS: case OP_GUARD: return sizeof(GsItem);
    case OP_CNVXGS: return sizeof(CgsItem);
    case OP_DENSE: return sizeof(DenseItem);
    case OP_SENS: case OP_SENS_EXT: return sizeof(SensItem);
    case OP_LEDGE: return sizeof(LedgeItem);
    }
    return 0;
}

struct Req {
    int kind;
    int slot = -1;
    AnyItem item;
    double* res = nullptr;
};

// operations that run for a long time in one launch (SteadyGS: all sweeps of a patch) go to a
// side stream, so that the batched launches of the other patches are not held behind them
static bool is_async_kind(int kind) { return kind == OP_STDYGS; }

// Device buffers of the scheduler's flushes: item staging, group partials, results and the
// per-item arrival counters of the fused reductions
struct FlushBufs {
    char* stage = nullptr;
    size_t stage_cap = 0;
    double* part = nullptr;
    size_t part_cap = 0;      // doubles
    double* res = nullptr;
    size_t res_cap = 0;       // doubles
    unsigned* cnt = nullptr;
    size_t cnt_cap = 0;       // counters
};

class Scheduler {
public:
    explicit Scheduler(hipStream_t s) : s_(s)
    {
        // one hardware queue stays with the main stream: side streams that share it would hold
        // the main stream's work behind their long launches
        int hwq = 4;
        if (const char* v = std::getenv("GPU_MAX_HW_QUEUES")) hwq = std::max(2, std::atoi(v));
        int nside = std::min(7, hwq - 1);
        if (const char* v = std::getenv("CNTC_SIDE_STREAMS")) nside = std::max(1, std::min(hwq - 1, std::atoi(v)));
        side_.assign(nside, nullptr);
        if (const char* v = std::getenv("CNTC_PRIO_EXTRA")) PRIO_EXTRA = std::max(0, std::min(64, std::atoi(v)));
        if (const char* v = std::getenv("CNTC_PRIO_DIV")) PRIO_DIV = std::max(1, std::min(64, std::atoi(v)));
        for (auto& st : side_) HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        cap_ = 64ull << 20;
        HIPCHK(hipHostMalloc((void**)&hstage_, cap_, hipHostMallocDefault));
        HIPCHK(hipMalloc((void**)&dstage_, cap_));
        HIPCHK(hipHostMalloc((void**)&hres_, rescap_ * sizeof(double), hipHostMallocDefault));
        HIPCHK(hipMalloc((void**)&dres_, rescap_ * sizeof(double)));
        HIPCHK(hipMalloc((void**)&dpart_, partcap_ * sizeof(double)));
        HIPCHK(hipMalloc((void**)&dcnt_, cntcap_ * sizeof(unsigned)));
        HIPCHK(hipMemset(dcnt_, 0, cntcap_ * sizeof(unsigned)));
    }
    ~Scheduler()
    {
        drop_async();
        for (auto& op : async_free_) free_async(op);
        for (auto& st : side_) (void)hipStreamDestroy(st);
        for (auto ev : ev_pool) (void)hipEventDestroy(ev);
        (void)hipHostFree(hstage_);
        (void)hipFree(dstage_);
        (void)hipHostFree(hres_);
        (void)hipFree(dres_);
        (void)hipFree(dpart_);
        (void)hipFree(dcnt_);
    }
    void resize(int nslots)
    {
        drop_async();
        q_.assign(nslots, {});
        dirty_.clear();
        in_dirty_.assign(nslots, 0);
        waiter_.assign(nslots, nullptr);
        async_wait_.assign(nslots, 0);
        prio_.assign(nslots, 0);
    }
    // patches marked here get flushes of their own between the full ones (PRIO_EXTRA per full
    // flush), so that the largest patches of a batch, which set its length, are not held to
    // the pace of every other patch's work; each patch's own sequence of operations is unchanged
    void set_priority(int slot, bool p) { prio_[slot] = p ? 1 : 0; }
    // Fixed buffers: flushes use `b` (no growth: a flush that needs more throws) and every
    // operation, SteadyGS included, runs on the main stream (no side-stream buffers).  Used by
    // the entry point that allocates nothing; restore() returns to the scheduler's own buffers.
    void use_fixed(const FlushBufs& b)
    {
        saved_ = FlushBufs{dstage_, cap_, dpart_, partcap_, dres_, rescap_, dcnt_, cntcap_};
        dstage_ = b.stage;
        dpart_ = b.part;
        dres_ = b.res;
        dcnt_ = b.cnt;
        fixed_stage_cap_ = std::min(b.stage_cap, cap_);
        fixed_part_cap_ = b.part_cap;
        fixed_res_cap_ = std::min(b.res_cap, rescap_);
        fixed_cnt_cap_ = b.cnt_cap;
        fixed_ = true;
    }
    void restore()
    {
        if (!fixed_) return;
        dstage_ = saved_.stage;
        dpart_ = saved_.part;
        dres_ = saved_.res;
        dcnt_ = saved_.cnt;
        fixed_ = false;
    }
    bool fixed() const { return fixed_; }
    // the end of the pinned staging buffer, which fixed-mode flushes never reach (they use
    // at most the fixed staging capacity from the start)
    char* host_tail(size_t bytes) { return hstage_ + ((cap_ - bytes) & ~(size_t)255); }
    int prio_div() const { return PRIO_DIV; }
    void enqueue(int slot, int kind, const AnyItem& it)
    {
        Req r;
        r.kind = kind;
        r.slot = slot;
        r.item = it;
        q_[slot].push_back(r);
        if (!in_dirty_[slot]) {
            in_dirty_[slot] = 1;
            dirty_.push_back(slot);
        }
    }
    // dirty_ lists exactly the slots with queued operations, so that a flush visits those
    // and not every slot of the serving loop's 16384
    bool any_queued() const { return !dirty_.empty(); }
    bool any_async() const { return !async_.empty(); }
    void set_result_dest(int slot, double* dst) { q_[slot].back().res = dst; }

    // per-kernel timing with HIP events on the scheduler's stream (enabled on request)
    struct ProfEntry {
        double seconds = 0, bytes = 0;
        long launches = 0;
    };
    // 0: off; 1: every kernel; 2: only the kernels whose rooflines are reported (SteadyGS, the
    // GD line-search probe and the FFT product stages), so that timing the rest costs nothing
    int profile = 0;
    bool timed_kind(int kind) const
    {
        return profile == 1 || (profile == 2 && (kind == OP_STDYGS || kind == OP_GD_PROBE ||
                                                 kind == OP_CONV || kind == OP_CONV32 ||
                                                 kind == OP_DENSE));
    }
    std::mutex prof_mtx;             // prof and nflush are read by the caller's thread
    std::map<int, ProfEntry> prof;   // key kind*4 + stage
    std::vector<hipEvent_t> ev_pool;
    void set_waiter(int slot, std::coroutine_handle<> h) { waiter_[slot] = h; }

    // one flush of every pending request and resumption of the waiting tasks of `slots`;
    // false when nothing was pending
    bool step(const std::vector<int>& slots)
    {
        bool any = false, anyp = false;
        for (int sl : slots)
            if (!q_[sl].empty()) {
                any = true;
                if (prio_[sl]) anyp = true;
            }
        if (any) {
            const bool only = anyp && (prio_turn_++ % (PRIO_EXTRA + 1)) != 0;
            flush(only);
            resume_ready(slots, only);
        }
        launch_pending(!any_queued());
        if (!async_.empty()) any = complete_async(false) > 0 || any;
        return any;
    }
    int nslots() const { return (int)q_.size(); }

    // run all tasks to completion
    void run(std::vector<Task>& tasks, const std::vector<int>& slots)
    {
        for (size_t i = 0; i < tasks.size(); i++) tasks[i].h.resume();
        for (;;) {
            if (any_queued()) {
                bool anyp = false;
                for (int sl : dirty_)
                    if (prio_[sl]) anyp = true;
                const bool only = anyp && (prio_turn_++ % (PRIO_EXTRA + 1)) != 0;
                flush(only);
                resume_ready(slots, only);
            }
            launch_pending(!any_queued());
            // side-stream operations: wait for one only when nothing else can proceed
            if (!async_.empty()) complete_async(!any_queued() && pending_.empty());
            if (!any_queued() && async_.empty() && pending_.empty()) break;
            // side-stream work waits for a stream: poll gently
            if (!any_queued() && !pending_.empty()) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        for (auto& t : tasks)
            if (!t.h.done()) throw std::runtime_error("scheduler: a patch solve did not finish");
    }

private:
    // resume the tasks of `slots` whose results arrived (not those waiting on a side stream)
    void resume_ready(const std::vector<int>& slots, bool prio_only = false)
    {
        for (int sl : slots)
            if (waiter_[sl] && !async_wait_[sl] && (!prio_only || prio_[sl])) {
                auto h = waiter_[sl];
                waiter_[sl] = nullptr;
                h.resume();
            }
    }

    struct GsbGroup;
    // one launch of side-stream operations (all items of one flush)
    struct AsyncOp {
        hipStream_t stream = nullptr;
        hipEvent_t ev_start = nullptr, ev_done = nullptr;
        hipEvent_t t0 = nullptr, t1 = nullptr;    // kernel timing (profiling)
        int kind = 0, variant = 0;
        bool timed = false;
        char* host = nullptr;        // pinned: items | group table | results
        char* dev = nullptr;         // device: items | group table | partials
        size_t cap = 0;
        int nitems = 0, K = 0;
        size_t res_off = 0;
        std::vector<int> slots;
        std::vector<double*> dest;
        std::vector<long> ns;        // element counts of the items (CNTC_GS_TIMING)
        double t_launch = 0.0;       // host clock at launch (CNTC_GS_TIMING)
        long max_n = 0;
        std::shared_ptr<GsbGroup> gsb;   // block SteadyGS: an operation in steps
    };
    void free_async(AsyncOp& op)
    {
        if (op.ev_start) (void)hipEventDestroy(op.ev_start);
        if (op.ev_done) (void)hipEventDestroy(op.ev_done);
        if (op.t0) (void)hipEventDestroy(op.t0);
        if (op.t1) (void)hipEventDestroy(op.t1);
        if (op.host) (void)hipHostFree(op.host);
        if (op.dev) (void)hipFree(op.dev);
        op = AsyncOp();
    }
    AsyncOp take_async(size_t bytes)
    {
        for (size_t k = 0; k < async_free_.size(); k++)
            if (async_free_[k].cap >= bytes) {
                AsyncOp op = async_free_[k];
                async_free_.erase(async_free_.begin() + k);
                op.slots.clear();
                op.dest.clear();
                op.ns.clear();
                op.gsb.reset();
                return op;
            }
        AsyncOp op;
        op.cap = std::max(bytes, (size_t)1 << 16);
        HIPCHK(hipHostMalloc((void**)&op.host, op.cap, hipHostMallocDefault));
        HIPCHK(hipMalloc((void**)&op.dev, op.cap));
        HIPCHK(hipEventCreateWithFlags(&op.ev_start, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&op.ev_done, hipEventDisableTiming));
        HIPCHK(hipEventCreate(&op.t0));
        HIPCHK(hipEventCreate(&op.t1));
        return op;
    }
    // wait for every side-stream operation and forget it (its tasks are being discarded)
    void drop_async()
    {
        pending_.clear();
        for (auto& op : async_) {
            (void)hipEventSynchronize(op.ev_done);
            async_free_.push_back(op);
        }
        async_.clear();
    }
    // hand the results of finished side-stream operations to their tasks and resume them;
    // with `block`, wait for the oldest one first
    int complete_async(bool block)
    {
        int n = 0;
        for (size_t k = 0; k < async_.size();) {
            AsyncOp& op = async_[k];
            if (block && k == 0) {
                HIPCHK(hipEventSynchronize(op.ev_done));
            } else {
                hipError_t e = hipEventQuery(op.ev_done);
                if (e == hipErrorNotReady) { k++; continue; }
                HIPCHK(e);
            }
            if (op.gsb && !gsb_advance(op)) {
                // the next step is queued on the operation's stream
                k++;
                block = false;
                continue;
            }
            const double* r = (const double*)(op.host + op.res_off);
            if (op.timed) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, op.t0, op.t1));
                std::lock_guard<std::mutex> lk(prof_mtx);
                auto& pe = prof[op.kind * 4 + op.variant];
                pe.seconds += ms * 1e-3;
                for (int i = 0; i < op.nitems; i++) pe.bytes += op.K >= 3 ? r[(size_t)i * op.K + 2] : 0.0;
                pe.launches++;
            }
            if (const char* tv = std::getenv("CNTC_GS_TIMING"); tv && tv[0] == '2')
                std::fprintf(stderr, "steadygs launch of %d items (largest %ld elements): launched at %.2f s, "
                             "done by %.2f s\n", op.nitems, op.max_n, op.t_launch,
                             std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t_origin);
            if (const char* tv = std::getenv("CNTC_GS_TIMING"); tv && tv[0] == '1' && op.K >= 10)
                for (int i = 0; i < op.nitems; i++) {
                    const double* q = r + (size_t)i * op.K;
                    std::fprintf(stderr, "steadygs item %d (%ld elements): sweeps %.0f, row sums %.3f s, sequential %.3f s "
                                 "(own row %.3f, search %.3f, element %.3f, chain %.3f)\n", i,
                                 i < (int)op.ns.size() ? op.ns[i] : 0L, q[0],
                                 q[3] * 1e-8, q[4] * 1e-8, q[6] * 1e-8, q[7] * 1e-8, q[8] * 1e-8, q[9] * 1e-8);
                }
            std::vector<int> sl = op.slots;
            for (int i = 0; i < op.nitems; i++) {
                if (op.dest[i])
                    for (int j = 0; j < op.K; j++) op.dest[i][j] = r[(size_t)i * op.K + j];
                async_wait_[op.slots[i]] = 0;
            }
            async_free_.push_back(op);
            async_.erase(async_.begin() + k);
            for (int s2 : sl)
                if (waiter_[s2]) {
                    auto h = waiter_[s2];
                    waiter_[s2] = nullptr;
                    h.resume();
                }
            n++;
            block = false;
        }
        return n;
    }
    // launch the pending side-stream operations: when the main stream has nothing queued, or
    // when enough have gathered that one launch fills a good part of the GPU (each item is
    // one work-group)
    bool side_idle(hipStream_t st) const
    {
        for (auto& op : async_)
            if (op.stream == st) return false;
        return true;
    }
    void launch_pending(bool main_idle)
    {
        (void)main_idle;
        if (pending_.empty()) return;
        // variants: 3 blocks of rows with FFT products, 2 row sums overlapped with the walk,
        // 1 staged row sums, 0 direct sums; the highest whose LDS fits.  CNTC_GS_KERNEL=1..4
        // caps the variant at 0..3 (testing)
        const char* forced = std::getenv("CNTC_GS_KERNEL");
        const int maxvar = forced && forced[0] >= '1' && forced[0] <= '4' ? forced[0] - '1' : 3;
        // largest patches first; one launch per idle side stream, each of patches of similar
        // size (down to a quarter of the launch's largest, at most 64), because a launch
        // returns when its longest item ends and a launch queued behind another on the same
        // stream waits for all of it
        std::vector<Req*> order;
        for (auto& r : pending_) order.push_back(&r);
        auto bigger = [](const Req* a, const Req* b) { return a->item.gs.n > b->item.gs.n; };
        std::stable_sort(order.begin(), order.end(), bigger);
        size_t k = 0;
        std::vector<Req> rest;
        for (auto st : side_) {
            if (k >= order.size()) break;
            if (!side_idle(st)) continue;
            const long top = order[k]->item.gs.n;
            std::vector<Req*> byvar[4];
            for (size_t c = 0; k < order.size() && c < 64 && 4L * order[k]->item.gs.n >= top; k++, c++) {
                Req* r = order[k];
                int v = maxvar;
                // (2 KB left for the kernels' static LDS)
                while (v > 0 && op_lds(r->kind, &r->item, v) + 2048 > MAX_LDS) v--;
                byvar[v].push_back(r);
            }
            if (!byvar[3].empty()) launch_gsb(byvar[3], st);
            for (int v = 2; v >= 0; v--)
                if (!byvar[v].empty()) launch_async(byvar[v], v, st);
        }
        for (; k < order.size(); k++) rest.push_back(*order[k]);
        pending_.swap(rest);
    }
    void launch_async(const std::vector<Req*>& reqs, int variant, hipStream_t stream)
    {
        int kind = reqs[0]->kind, nitems = (int)reqs.size();
        size_t isz = item_size(kind);
        int K = op_nred(kind, &reqs[0]->item);
        auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
        size_t items_b = align(isz * nitems), gs_b = align(sizeof(int) * (nitems + 1));
        size_t res_b = align(sizeof(double) * K * nitems);
        AsyncOp op = take_async(items_b + gs_b + res_b);
        op.stream = stream;
        op.nitems = nitems;
        op.K = K;
        op.res_off = items_b + gs_b;
        op.t_launch = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t_origin;
        op.max_n = 0;
        for (auto* r : reqs) op.max_n = std::max(op.max_n, (long)r->item.gs.n);
        size_t lds = 0;
        int* gs = (int*)(op.host + items_b);
        int tot = 0;
        for (int i = 0; i < nitems; i++) {
            std::memcpy(op.host + isz * i, &reqs[i]->item, isz);
            gs[i] = tot;
            tot += op_groups(kind, &reqs[i]->item, variant);
            lds = std::max(lds, op_lds(kind, &reqs[i]->item, variant));
            op.slots.push_back(reqs[i]->slot);
            op.dest.push_back(reqs[i]->res);
            op.ns.push_back((long)reqs[i]->item.gs.n);
            async_wait_[reqs[i]->slot] = 1;
        }
        gs[nitems] = tot;
        // start after everything issued so far on the main stream (the inputs of these items)
        HIPCHK(hipEventRecord(op.ev_start, s_));
        HIPCHK(hipStreamWaitEvent(op.stream, op.ev_start, 0));
        HIPCHK(hipMemcpyAsync(op.dev, op.host, items_b + gs_b, hipMemcpyHostToDevice, op.stream));
        op.kind = kind;
        op.variant = variant;
        op.timed = timed_kind(kind);
        if (op.timed) HIPCHK(hipEventRecord(op.t0, op.stream));
        // one work-group per item: its partials are the item's results
        launch_op(kind, op.dev, (const int*)(op.dev + items_b), nitems, tot,
                  RedOut{(double*)(op.dev + op.res_off), nullptr, nullptr, nullptr}, variant, lds, op.stream);
        HIPCHK(hipGetLastError());
        if (op.timed) HIPCHK(hipEventRecord(op.t1, op.stream));
        HIPCHK(hipMemcpyAsync(op.host + op.res_off, op.dev + op.res_off, res_b, hipMemcpyDeviceToHost,
                              op.stream));
        HIPCHK(hipEventRecord(op.ev_done, op.stream));
        async_.push_back(op);
    }


    // SteadyGS by blocks of rows (variant 3) on a side stream, as one side-stream operation in
    // steps: the set-up (differences, row ranges), then one step per sweep -- the product
    // z = A dp, for each block the walk and the product of its change into z (none after
    // the last block), and the convergence test.  complete_async advances it each time the
    // step's event has completed; patches that have converged leave the later sweeps.
    struct GsbGroup {
        int B = 8;                            // rows per block
        std::vector<GsItem> gs;
        std::vector<int> first, last, nb;     // interior rows and blocks per patch
        std::vector<char> active;
        std::vector<std::array<double, 10>> fin;
        std::vector<int> act;                 // patches of the sweep in progress
        int nbmax = 0, sweeps = 0;
        size_t off_st = 0, off_sw = 0, cap_sw = 0;
        bool filled = false;                  // the sweep area holds the launches of `act`
        struct Launch {
            int kind, n, ng[3];
            size_t items, gst[3], lds[3];
        };
        std::vector<Launch> seq;
        double seconds = 0.0;
        bool timed = false;
        // profiling: events around each walk launch of the sweep in flight, their summed time
        std::vector<hipEvent_t> wev;
        int nwev = 0, nwalk = 0;
        double walk_seconds = 0.0;
        ~GsbGroup()
        {
            for (auto e : wev) (void)hipEventDestroy(e);
        }
    };
    static GsbItem gsb_item(const GsItem& g, int mode, double* st)
    {
        GsbItem b{};
        b.n = g.n; b.mx = g.mx; b.my = g.my; b.k = g.k;
        b.mode = mode; b.maxgs = g.maxgs;
        b.eps = g.eps; b.omegah = g.omegah; b.omegas = g.omegas; b.ga_inv = g.ga_inv; b.mu = g.mu;
        b.facnel = g.facnel;
        b.el = g.el; b.px = g.px; b.py = g.py; b.pn = g.pn; b.sx = g.sx; b.sy = g.sy;
        b.wsx = g.wsx; b.wsy = g.wsy; b.dpx = g.dpx; b.dpy = g.dpy;
        b.zx = g.zx; b.zy = g.zy; b.ddx = g.ddx; b.ddy = g.ddy;
        b.rng = g.list;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) b.cf[i][j] = g.cf[i][j];
        b.p0x = g.p0x; b.p0y = g.p0y;
        b.st = st;
        return b;
    }
    static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
    void launch_gsb(const std::vector<Req*>& reqs, hipStream_t stream)
    {
        const int nitems = (int)reqs.size(), K = 10;
        auto G = std::make_shared<GsbGroup>();
        if (const char* e = std::getenv("CNTC_GSB_ROWS"); e && std::atoi(e) > 0) G->B = std::atoi(e);
        int mymax = 0;
        for (auto* r : reqs) {
            G->gs.push_back(r->item.gs);
            mymax = std::max(mymax, r->item.gs.my);
        }
        const int nbb = (mymax + G->B - 1) / G->B;
        const size_t isz = std::max(sizeof(ConvItem), sizeof(GsbItem));
        const size_t one = align256(isz * nitems) + 3 * align256(sizeof(int) * (nitems + 1));
        G->cap_sw = (size_t)(2 * nbb + 2) * one;
        const size_t items_b = align256(sizeof(GsbItem) * nitems), gs_b = align256(sizeof(int) * (nitems + 1));
        const size_t res_b = align256(sizeof(double) * K * nitems);
        G->off_st = items_b + gs_b + res_b;
        G->off_sw = G->off_st + align256(sizeof(double) * 4 * nitems);
        AsyncOp op = take_async(G->off_sw + G->cap_sw);
        op.stream = stream;
        op.nitems = nitems;
        op.K = K;
        op.res_off = items_b + gs_b;
        op.kind = OP_STDYGS;
        op.variant = 3;
        op.timed = false;   // timed per sweep by the group itself
        op.t_launch = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t_origin;
        op.max_n = 0;
        int* gst = (int*)(op.host + items_b);
        for (int i = 0; i < nitems; i++) {
            double* st = (double*)(op.dev + G->off_st) + 4 * i;
            GsbItem b = gsb_item(G->gs[i], 0, st);
            std::memcpy(op.host + sizeof(GsbItem) * i, &b, sizeof(GsbItem));
            gst[i] = i;
            op.slots.push_back(reqs[i]->slot);
            op.dest.push_back(reqs[i]->res);
            op.ns.push_back((long)G->gs[i].n);
            op.max_n = std::max(op.max_n, (long)G->gs[i].n);
            async_wait_[reqs[i]->slot] = 1;
        }
        gst[nitems] = nitems;
        G->timed = timed_kind(OP_STDYGS);
        op.gsb = G;
        HIPCHK(hipEventRecord(op.ev_start, s_));
        HIPCHK(hipStreamWaitEvent(op.stream, op.ev_start, 0));
        HIPCHK(hipMemcpyAsync(op.dev, op.host, items_b + gs_b, hipMemcpyHostToDevice, op.stream));
        launch_op(OP_GSB_SETUP, op.dev, (const int*)(op.dev + items_b), nitems, nitems,
                  RedOut{(double*)(op.dev + op.res_off), nullptr, nullptr, nullptr}, 0, 0, op.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(op.host + op.res_off, op.dev + op.res_off, res_b, hipMemcpyDeviceToHost,
                              op.stream));
        HIPCHK(hipEventRecord(op.ev_done, op.stream));
        async_.push_back(op);
    }
    // the launches of a sweep over G.act, written to the sweep area of op (host and device)
    void gsb_fill(AsyncOp& op, GsbGroup& G)
    {
        G.seq.clear();
        size_t off = 0;
        char* h = op.host + G.off_sw;
        const int na = (int)G.act.size();
        auto add_conv = [&](const std::vector<int>& who, bool change) {
            if (who.empty()) return;
            GsbGroup::Launch L{};
            L.kind = OP_CONV;
            L.n = (int)who.size();
            L.items = off;
            ConvItem* ci = (ConvItem*)(h + off);
            for (int a = 0; a < L.n; a++) {
                const GsItem& g = G.gs[who[a]];
                ConvItem c = g.conv;
                c.in[0] = change ? g.ddx : g.dpx;
                c.in[1] = change ? g.ddy : g.dpy;
                c.addend[0] = change ? g.zx : nullptr;
                c.addend[1] = change ? g.zy : nullptr;
                ci[a] = c;
            }
            off += align256(sizeof(ConvItem) * L.n);
            for (int st = 0; st < 3; st++) {
                L.gst[st] = off;
                int* gp = (int*)(h + off);
                int tot = 0;
                size_t lds = 0;
                for (int a = 0; a < L.n; a++) {
                    gp[a] = tot;
                    tot += op_groups(OP_CONV, &ci[a], st);
                    lds = std::max(lds, op_lds(OP_CONV, &ci[a], st));
                }
                gp[L.n] = tot;
                L.ng[st] = tot;
                L.lds[st] = lds;
                off += align256(sizeof(int) * (L.n + 1));
            }
            G.seq.push_back(L);
        };
        auto add_gsb = [&](int kind, const std::vector<int>& who, const std::vector<GsbItem>& its) {
            if (who.empty()) return;
            GsbGroup::Launch L{};
            L.kind = kind;
            L.n = (int)who.size();
            L.items = off;
            std::memcpy(h + off, its.data(), sizeof(GsbItem) * L.n);
            off += align256(sizeof(GsbItem) * L.n);
            L.gst[0] = off;
            int* gp = (int*)(h + off);
            size_t lds = 0;
            for (int a = 0; a <= L.n; a++) gp[a] = a;
            for (auto& b : its) lds = std::max(lds, op_lds(kind, &b, 0));
            L.ng[0] = L.n;
            L.lds[0] = lds;
            off += align256(sizeof(int) * (L.n + 1));
            G.seq.push_back(L);
        };
        auto st_of = [&](int i) { return (double*)(op.dev + G.off_st) + 4 * i; };
        // z = A dp for every patch of the sweep
        add_conv(G.act, false);
        for (int b = 0; b < G.nbmax; b++) {
            std::vector<int> who, whoc;
            std::vector<GsbItem> its;
            for (int a = 0; a < na; a++) {
                const int i = G.act[a];
                if (G.nb[i] <= b) continue;
                GsbItem w = gsb_item(G.gs[i], 1, st_of(i));
                w.y0 = G.first[i] + b * G.B;
                w.y1 = std::min(G.last[i] + 1, w.y0 + G.B);
                // d of the previous block (for the first block: the last block of the sweep
                // before, whose change no product used)
                const int pb = b > 0 ? b - 1 : G.nb[i] - 1;
                w.z0 = G.first[i] + pb * G.B;
                w.z1 = std::min(G.last[i] + 1, w.z0 + G.B);
                who.push_back(i);
                its.push_back(w);
                if (G.nb[i] > b + 1) whoc.push_back(i);
            }
            add_gsb(OP_GSB_WALK, who, its);
            add_conv(whoc, true);
        }
        std::vector<GsbItem> chk;
        for (int a = 0; a < na; a++) chk.push_back(gsb_item(G.gs[G.act[a]], 2, st_of(G.act[a])));
        add_gsb(OP_GSB_CHECK, G.act, chk);
        if (off > G.cap_sw) throw std::runtime_error("steadygs blocks: sweep area too small");
        HIPCHK(hipMemcpyAsync(op.dev + G.off_sw, h, off, hipMemcpyHostToDevice, op.stream));
        G.filled = true;
    }
    // after a step's event: take its results; queue the next sweep and return false, or write
    // the final results (op order, 10 per patch) and return true
    bool gsb_advance(AsyncOp& op)
    {
        GsbGroup& G = *op.gsb;
        const int nitems = op.nitems, K = op.K;
        double* r = (double*)(op.host + op.res_off);
        if (G.timed && G.sweeps > 0) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, op.t0, op.t1));
            G.seconds += ms * 1e-3;
            for (int j = 0; j < G.nwev; j += 2) {
                HIPCHK(hipEventElapsedTime(&ms, G.wev[j], G.wev[j + 1]));
                G.walk_seconds += ms * 1e-3;
                G.nwalk++;
            }
        }
        G.nwev = 0;
        if (G.first.empty()) {
            // set-up done: interior rows and blocks per patch
            G.first.assign(nitems, 0);
            G.last.assign(nitems, -1);
            G.nb.assign(nitems, 0);
            G.active.assign(nitems, 1);
            G.fin.assign(nitems, std::array<double, 10>{});
            for (int i = 0; i < nitems; i++) {
                const double f = r[(size_t)i * K], l = r[(size_t)i * K + 1];
                if (std::isfinite(f) && std::isfinite(l) && l >= f) {
                    G.first[i] = (int)f;
                    G.last[i] = (int)l;
                    G.nb[i] = (G.last[i] - G.first[i]) / G.B + 1;
                }
                G.nbmax = std::max(G.nbmax, G.nb[i]);
                G.act.push_back(i);
            }
        } else {
            // a sweep done: patches that stop leave
            std::vector<int> keep;
            for (int a = 0; a < (int)G.act.size(); a++) {
                const int i = G.act[a];
                const double* q = r + (size_t)a * K;
                if (q[6] != 0.0) {
                    keep.push_back(i);
                } else {
                    for (int j = 0; j < 10; j++) G.fin[i][j] = q[j];
                    G.fin[i][6] = 0.0;
                    G.fin[i][7] = 0.0;
                    G.active[i] = 0;
                }
            }
            if (keep.size() != G.act.size()) G.filled = false;
            G.act.swap(keep);
        }
        if (G.act.empty()) {
            for (int i = 0; i < nitems; i++)
                for (int j = 0; j < K; j++) r[(size_t)i * K + j] = G.fin[i][j];
            if (G.timed) {
                // the sweeps as a whole (walks, products and tests), and the walk kernel alone
                // with the flops of its direct sums
                std::lock_guard<std::mutex> lk(prof_mtx);
                auto& pe = prof[OP_STDYGS * 4 + 3];
                pe.seconds += G.seconds;
                pe.launches++;
                auto& pw = prof[OP_GSB_WALK * 4];
                pw.seconds += G.walk_seconds;
                for (int i = 0; i < nitems; i++) pw.bytes += G.fin[i][2];
                pw.launches += G.nwalk;
            }
            return true;
        }
        if (!G.filled) gsb_fill(op, G);
        if (G.timed) HIPCHK(hipEventRecord(op.t0, op.stream));
        for (const auto& L : G.seq) {
            const char* d = op.dev + G.off_sw;
            if (L.kind == OP_CONV) {
                for (int st = 0; st < 3; st++)
                    launch_op(OP_CONV, d + L.items, (const int*)(d + L.gst[st]), L.n, L.ng[st], RedOut{},
                              st, L.lds[st], op.stream);
            } else {
                RedOut ro{};
                if (L.kind == OP_GSB_CHECK) ro.partials = (double*)(op.dev + op.res_off);
                const bool tw = G.timed && L.kind == OP_GSB_WALK;
                if (tw) {
                    while ((int)G.wev.size() < G.nwev + 2) {
                        hipEvent_t e;
                        HIPCHK(hipEventCreate(&e));
                        G.wev.push_back(e);
                    }
                    HIPCHK(hipEventRecord(G.wev[G.nwev], op.stream));
                }
                launch_op(L.kind, d + L.items, (const int*)(d + L.gst[0]), L.n, L.ng[0], ro, 0, L.lds[0],
                          op.stream);
                if (tw) {
                    HIPCHK(hipEventRecord(G.wev[G.nwev + 1], op.stream));
                    G.nwev += 2;
                }
            }
            HIPCHK(hipGetLastError());
        }
        if (G.timed) HIPCHK(hipEventRecord(op.t1, op.stream));
        HIPCHK(hipMemcpyAsync(op.host + op.res_off, op.dev + op.res_off, sizeof(double) * K * G.act.size(),
                              hipMemcpyDeviceToHost, op.stream));
        HIPCHK(hipEventRecord(op.ev_done, op.stream));
        G.sweeps++;
        return false;
    }

    struct Group {
        int kind;
        std::vector<int> reqs;   // indices into flat list
    };
    // prio_only: the queues of the priority patches only; the others stay queued
    void flush(bool prio_only = false)
    {
        const auto tf0 = std::chrono::steady_clock::now();
        // collect requests level by level; side-stream operations are set aside (a task
        // waiting on one has nothing queued behind it)
        std::vector<Req*> flat, async_reqs;
        std::vector<std::vector<Group>> levels;
        size_t maxlen = 0;
        auto taken = [&](size_t sl) { return !prio_only || prio_[sl]; };
        // the slots with queued work, in slot order (the order of a launch's items)
        std::vector<int> work(dirty_);
        std::sort(work.begin(), work.end());
        for (int sl : work)
            if (taken(sl)) maxlen = std::max(maxlen, q_[sl].size());
        for (size_t lev = 0; lev < maxlen; lev++) {
            std::map<int, Group> byk;
            for (int sl : work) {
                auto& q = q_[sl];
                if (taken(sl) && lev < q.size()) {
                    if (is_async_kind(q[lev].kind) && !fixed_) {
                        async_reqs.push_back(&q[lev]);
                        continue;
                    }
                    int idx = (int)flat.size();
                    flat.push_back(&q[lev]);
                    auto& g = byk[q[lev].kind];
                    g.kind = q[lev].kind;
                    g.reqs.push_back(idx);
                }
            }
            std::vector<Group> gs;
            for (auto& kv : byk) gs.push_back(kv.second);
            levels.push_back(gs);
        }
        // stage items, group tables and result slots
        struct Launch {
            int kind, nitems, K, stage;
            size_t items_off, gstart_off, modes_off, part_off, res_off, cnt_off;
            int ngroups;
            size_t lds;
        };
        std::vector<Launch> launches;
        size_t off = 0;
        size_t part = 0, res = 0, cnt = 0;
        auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
        std::vector<std::pair<Req*, size_t>> resmap;
        for (auto& lev : levels)
            for (auto& g : lev) {
                int kind = g.kind;
                size_t isz = item_size(kind);
                // an operation without an item record would launch on whatever the staging
                // buffer holds
                if (isz == 0) throw std::runtime_error("scheduler: operation kind without an item size");
                int nitems = (int)g.reqs.size();
                size_t items_off = off;
                if (off + isz * nitems > stage_cap()) throw std::runtime_error("scheduler: staging overflow");
                for (int i = 0; i < nitems; i++)
                    std::memcpy(hstage_ + off + isz * i, &flat[g.reqs[i]]->item, isz);
                off = align(off + isz * nitems);
                int K = op_nred(kind, &flat[g.reqs[0]]->item);
                int nst = (kind == OP_CONV || kind == OP_CONV32) ? 3 : 1;
                int st0 = 0;
                if (kind == OP_STDYGS) {
                    // SteadyGS on the main stream (fixed-buffer mode): the overlapped variant
                    // where every item's LDS fits, else the staged one, else the direct one
                    st0 = 2;
                    for (int i = 0; i < nitems; i++)
                        while (st0 > 0 && op_lds(kind, &flat[g.reqs[i]]->item, st0) + 2048 > MAX_LDS) st0--;
                }
                for (int st = st0; st < st0 + nst; st++) {
                    Launch L;
                    L.kind = kind;
                    L.nitems = nitems;
                    L.K = K;
                    L.stage = st;
                    L.items_off = items_off;
                    L.gstart_off = off;
                    int* gs = (int*)(hstage_ + off);
                    int tot = 0;
                    size_t lds = 0;
                    for (int i = 0; i < nitems; i++) {
                        gs[i] = tot;
                        const int ng = op_groups(kind, &flat[g.reqs[i]]->item, st);
                        // every item of a reduction must have a group to form its result
                        if (K > 0 && ng <= 0) throw std::runtime_error("scheduler: reduction item without work-groups");
                        tot += ng;
                        lds = std::max(lds, op_lds(kind, &flat[g.reqs[i]]->item, st));
                    }
                    gs[nitems] = tot;
                    off = align(off + sizeof(int) * (nitems + 1));
                    L.ngroups = tot;
                    L.lds = lds;
                    L.modes_off = off;
                    int* md = (int*)(hstage_ + off);
                    for (int i = 0; i < nitems; i++) md[i] = op_combine_mode(kind, &flat[g.reqs[i]]->item);
                    off = align(off + sizeof(int) * nitems);
                    L.part_off = part;
                    L.res_off = res;
                    L.cnt_off = cnt;
                    if (K > 0) {
                        cnt += nitems;
                        part += (size_t)tot * K;
                        for (int i = 0; i < nitems; i++)
                            resmap.push_back({flat[g.reqs[i]], res + (size_t)i * K});
                        res += (size_t)nitems * K;
                    }
                    if (off > stage_cap()) throw std::runtime_error("scheduler: staging overflow");
                    launches.push_back(L);
                }
            }
        if (fixed_ && (part > fixed_part_cap_ || cnt > fixed_cnt_cap_ || res > fixed_res_cap_))
            throw std::runtime_error("scheduler: a flush needs more than the fixed buffers hold");
        if (!fixed_ && part > partcap_) {
            HIPCHK(hipStreamSynchronize(s_));
            HIPCHK(hipFree(dpart_));
            partcap_ = part * 2;
            HIPCHK(hipMalloc((void**)&dpart_, partcap_ * sizeof(double)));
        }
        if (!fixed_ && cnt > cntcap_) {
            HIPCHK(hipStreamSynchronize(s_));
            HIPCHK(hipFree(dcnt_));
            cntcap_ = cnt * 2;
            HIPCHK(hipMalloc((void**)&dcnt_, cntcap_ * sizeof(unsigned)));
            HIPCHK(hipMemset(dcnt_, 0, cntcap_ * sizeof(unsigned)));
        }
        if (!fixed_ && res > rescap_) {
            HIPCHK(hipStreamSynchronize(s_));
            HIPCHK(hipHostFree(hres_));
            HIPCHK(hipFree(dres_));
            rescap_ = res * 2;
            HIPCHK(hipHostMalloc((void**)&hres_, rescap_ * sizeof(double), hipHostMallocDefault));
            HIPCHK(hipMalloc((void**)&dres_, rescap_ * sizeof(double)));
        }
        HIPCHK(hipMemcpyAsync(dstage_, hstage_, off, hipMemcpyHostToDevice, s_));
        std::vector<std::pair<int, double>> pkeys;
        size_t nev = 0;
        if (profile) {
            size_t need = 0;
            for (auto& L : launches) need += timed_kind(L.kind) && L.ngroups > 0 ? 2 : 0;
            while (ev_pool.size() < need) {
                hipEvent_t ev;
                HIPCHK(hipEventCreate(&ev));
                ev_pool.push_back(ev);
            }
        }
        for (auto& L : launches) {
            const bool timed = timed_kind(L.kind) && L.ngroups > 0;
            if (timed) {
                double bytes = 0;
                for (int i = 0; i < L.nitems; i++)
                    bytes += op_bytes(L.kind, hstage_ + L.items_off + item_size(L.kind) * i, L.stage);
                pkeys.push_back({L.kind * 4 + L.stage, bytes});
                HIPCHK(hipEventRecord(ev_pool[nev++], s_));
            }
            RedOut ro{dpart_ + L.part_off, L.K > 0 ? dcnt_ + L.cnt_off : nullptr, dres_ + L.res_off,
                      (const int*)(dstage_ + L.modes_off)};
            // SteadyGS and ConvexGS write their one group's values as the item's results
            if (L.kind == OP_STDYGS || L.kind == OP_CNVXGS)
                ro = RedOut{dres_ + L.res_off, nullptr, nullptr, nullptr};
            launch_op(L.kind, dstage_ + L.items_off, (const int*)(dstage_ + L.gstart_off), L.nitems,
                      L.ngroups, ro, L.stage, L.lds, s_);
            if (hipError_t le = hipGetLastError(); le != hipSuccess) {
                char msg[256];
                std::snprintf(msg, sizeof msg, "launch of operation %d stage %d failed (%s): %d groups, %zu B LDS",
                              L.kind, L.stage, hipGetErrorString(le), L.ngroups, L.lds);
                throw std::runtime_error(msg);
            }
            if (timed) HIPCHK(hipEventRecord(ev_pool[nev++], s_));
        }
        HIPCHK(hipGetLastError());
        for (Req* r : async_reqs) {
            pending_.push_back(*r);
            async_wait_[r->slot] = 1;
        }
        if (res > 0) HIPCHK(hipMemcpyAsync(hres_, dres_, res * sizeof(double), hipMemcpyDeviceToHost, s_));
        const auto tf1 = std::chrono::steady_clock::now();
        HIPCHK(hipStreamSynchronize(s_));
        const auto tf2 = std::chrono::steady_clock::now();
        flush_issue_s += std::chrono::duration<double>(tf1 - tf0).count();
        flush_wait_s += std::chrono::duration<double>(tf2 - tf1).count();
        flush_launches += (long)launches.size();
        std::lock_guard<std::mutex> lk(prof_mtx);
        for (size_t k = 0; k < pkeys.size(); k++) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev_pool[2 * k], ev_pool[2 * k + 1]));
            auto& pe = prof[pkeys[k].first];
            pe.seconds += ms * 1e-3;
            pe.bytes += pkeys[k].second;
            pe.launches++;
        }
        for (auto& rm : resmap)
            if (rm.first->res) {
                int K = op_nred(rm.first->kind, &rm.first->item);
                for (int k = 0; k < K; k++) rm.first->res[k] = hres_[rm.second + k];
            }
        dirty_.clear();
        for (int sl : work) {
            if (taken(sl)) {
                q_[sl].clear();
                in_dirty_[sl] = 0;
            } else {
                dirty_.push_back(sl);
            }
        }
        nflush++;
    }

    hipStream_t s_;
    // side streams (CNTC_SIDE_STREAMS, default and at most GPU_MAX_HW_QUEUES - 1, at most 7 by
    // default): operations on one stream run one after another, so more streams let more
    // SteadyGS launches run at once
    std::vector<hipStream_t> side_;
    std::vector<AsyncOp> async_, async_free_;
    std::vector<Req> pending_;
    std::vector<char> async_wait_;
    std::vector<char> prio_;
    int PRIO_EXTRA = 2;        // CNTC_PRIO_EXTRA
    // patches of at least 1/PRIO_DIV of the batch's largest element count get priority
    int PRIO_DIV = 4;          // CNTC_PRIO_DIV
    long prio_turn_ = 0;
    std::vector<std::vector<Req>> q_;
    std::vector<int> dirty_;        // slots whose queue is not empty
    std::vector<char> in_dirty_;
    std::vector<std::coroutine_handle<>> waiter_;
    char* hstage_ = nullptr;
    char* dstage_ = nullptr;
    size_t cap_ = 0;
    double* hres_ = nullptr;
    double* dres_ = nullptr;
    size_t rescap_ = 1 << 20;
    double* dpart_ = nullptr;
    size_t partcap_ = 4 << 20;
    // one arrival counter per reduction item of a flush; each item's last group resets its own
    unsigned* dcnt_ = nullptr;
    size_t cntcap_ = 1 << 16;
    // fixed-buffer mode (use_fixed): caps of the caller's buffers, the scheduler's own saved
    bool fixed_ = false;
    size_t fixed_stage_cap_ = 0, fixed_part_cap_ = 0, fixed_res_cap_ = 0, fixed_cnt_cap_ = 0;
    FlushBufs saved_;
    size_t stage_cap() const { return fixed_ ? fixed_stage_cap_ : cap_; }

public:
    long nflush = 0;
    double t_origin = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    // host seconds spent issuing flushes and waiting for them, kernel launches issued
    double flush_issue_s = 0.0, flush_wait_s = 0.0;
    long flush_launches = 0;
};

// awaiter of a request with results
struct Res {
    double v[12];
};

struct ResAwait {
    Scheduler* S;
    int slot;
    Res r;
    bool await_ready() { return false; }
    void await_suspend(std::coroutine_handle<> h)
    {
        S->set_result_dest(slot, r.v);
        S->set_waiter(slot, h);
    }
    Res await_resume() { return r; }
};

// per-patch issuing context
// the item of the FFT product u[ik] (mask) = sum_jk C[ik][jk] * p[jk] on patch P,
// src/m_aijpj.f90:258-451.  which: 0 = cs, 1 = cv, 2 = ms; fp32 selects the single-precision
// spectra.
// the dense product of the same blocks as conv_item (which 0: cs, 1: cv, 2: preconditioner)
static DenseItem dense_item(const Patch* P, int which, int mask, int nout, const int* ik,
                           double* const* out, int nin, const int* jk, double* const* in,
                           const double* const* addend)
{
    DenseItem d{};
    d.mx = P->mx;
    d.my = P->my;
    d.nin = nin;
    d.nout = nout;
    d.mask = mask;
    const size_t nb = (size_t)(2 * P->mx) * (2 * P->my);
    const double* base = which == 0 ? P->cf : (which == 1 ? P->cfv : P->cfm);
    for (int j = 0; j < nin; j++) d.in[j] = in[j];
    for (int i = 0; i < nout; i++) {
        d.out[i] = out[i];
        d.addend[i] = addend ? addend[i] : nullptr;
        for (int j = 0; j < nin; j++) {
            const int a = ik[i], b = jk[j];
            const double* c = nullptr;
            if (which == 2) {
                if (a == b) c = base + (size_t)a * nb;
            } else if (P->nt_cpl || !((a == 2) != (b == 2))) {
                c = base + (size_t)(a * 3 + b) * nb;
            }
            d.cf[i][j] = c;
        }
    }
    d.el = P->el;
    d.ga_inv = 1.0 / P->ga;   // every product, the preconditioner's included, is (1/G) cf * in
    return d;
}

// Where the dense FP64-MFMA product replaces the FFT product on patches of at most
// DENSE_MAX_N elements (CNTC_DENSE): 1 (default) in the sensitivity solves, 2 in every
// product, 0 nowhere.
static int dense_mode()
{
    const char* v = std::getenv("CNTC_DENSE");
    return v ? std::max(0, std::min(2, std::atoi(v))) : 1;
}

static ConvItem conv_item(const Patch* P, int which, int mask, int nout, const int* ik,
                          double* const* out, int nin, const int* jk, double* const* in,
                          const double* const* addend, bool fp32)
{
    ConvItem c{};
    c.mx = P->mx;
    c.my = P->my;
    c.Lx = P->Lx;
    c.Ly = P->Ly;
    c.nin = nin;
    c.nout = nout;
    c.mask = mask;
    c.mode = 0;
    int M = P->Lx / 2, H = M + 1;
    size_t bsz = (size_t)H * P->Ly;
    const char* base;
    size_t esz = fp32 ? 8 : 16;
    if (which == 0) base = (const char*)(fp32 ? P->spec32 : P->spec);
    else if (which == 1) base = (const char*)P->specv;
    else base = (const char*)(fp32 ? P->specm32 : P->specm);
    for (int j = 0; j < nin; j++) c.in[j] = in[j];
    for (int i = 0; i < nout; i++) {
        c.out[i] = out[i];
        c.addend[i] = addend ? addend[i] : nullptr;
        for (int j = 0; j < nin; j++) {
            int a = ik[i], b = jk[j];   // 0-based directions
            const void* sp = nullptr;
            if (which == 2) {
                if (a == b) sp = base + (size_t)a * bsz * esz;
            } else if (P->nt_cpl || !((a == 2) != (b == 2))) {
                sp = base + (size_t)(a * 3 + b) * bsz * esz;
            }
            c.spec[i][j] = sp;
        }
    }
    c.el = P->el;
    c.wsS = P->wsS;
    c.wsT = P->wsT;
    c.twM = fp32 ? P->twM32 : P->twM;
    c.twLx = fp32 ? P->twLx32 : P->twLx;
    c.twLy = fp32 ? P->twLy32 : P->twLy;
    c.scale = 1.0;
    return c;
}

struct Ctx {
    Scheduler* S;
    Patch* P;
    int slot;
    Engine* E;

    void op(int kind, const AnyItem& it) { S->enqueue(slot, kind, it); }
    ResAwait opr(int kind, const AnyItem& it)
    {
        S->enqueue(slot, kind, it);
        return ResAwait{S, slot, {}};
    }
    // an operation whose results land in dst when the flush that runs it ends; the task goes
    // on issuing (a later co_await of the same flush makes them readable)
    void opr_into(int kind, const AnyItem& it, double* dst)
    {
        S->enqueue(slot, kind, it);
        S->set_result_dest(slot, dst);
    }
    // masked BLAS-1 helpers (src/m_gridfunc.f90)
    AnyItem vitem(int sub, int mask, double* x0, double* x1, double* y0, double* y1, double a, double b)
    {
        AnyItem it;
        it.vec.n = P->n;
        it.vec.sub = sub;
        it.vec.mask = mask;
        it.vec.ncomp = x1 ? 2 : 1;
        it.vec.el = P->el;
        it.vec.x[0] = x0;
        it.vec.x[1] = x1;
        it.vec.y[0] = y0;
        it.vec.y[1] = y1;
        it.vec.a = a;
        it.vec.b = b;
        return it;
    }
    // device-scalar forms: a reduction whose results stay on the device (sout, 4 values) for
    // later launches, and an update whose scalar the device forms from such results
    void vdev(int sub, int mask, double* x, double* y, double* sout)
    {
        AnyItem it = vitem(sub, mask, x, nullptr, y, nullptr, 0, 0);
        it.vec.sout = sout;
        op(OP_VEC, it);
    }
    void vfn(int sub, int mask, double* x, double* y, int sfn, const double* s0, const double* s1 = nullptr,
             const double* s2 = nullptr, int flag = 0)
    {
        AnyItem it = vitem(sub, mask, x, nullptr, y, nullptr, 0, 0);
        it.vec.sfn = sfn;
        it.vec.sin[0] = s0;
        it.vec.sin[1] = s1;
        it.vec.sin[2] = s2;
        it.vec.flag = flag;
        op(OP_VEC, it);
    }
    // a sum of squares that also returns the device scalars e0[0], e1[0] as results 2 and 3
    ResAwait sumsq_echo(int mask, double* x, const double* e0, const double* e1)
    {
        AnyItem it = vitem(V_SUMSQ, mask, x, nullptr, nullptr, nullptr, 0, 0);
        it.vec.echo[0] = e0;
        it.vec.echo[1] = e1;
        return opr(OP_VEC, it);
    }
    void set(int mask, double* x, double a, double* x1 = nullptr) { op(OP_VEC, vitem(V_SET, mask, x, x1, nullptr, nullptr, a, 0)); }
    void copy(int mask, double* x, double* y, double* x1 = nullptr, double* y1 = nullptr) { op(OP_VEC, vitem(V_COPY, mask, x, x1, y, y1, 0, 0)); }
    void scal(int mask, double* x, double a, double* x1 = nullptr) { op(OP_VEC, vitem(V_SCAL, mask, x, x1, nullptr, nullptr, a, 0)); }
    void axpy(int mask, double a, double* x, double* y, double* x1 = nullptr, double* y1 = nullptr) { op(OP_VEC, vitem(V_AXPY, mask, x, x1, y, y1, a, 0)); }
    void axpby(int mask, double a, double* x, double b, double* y, double* x1 = nullptr, double* y1 = nullptr) { op(OP_VEC, vitem(V_AXPBY, mask, x, x1, y, y1, a, b)); }
    void addc(int mask, double* x, double a) { op(OP_VEC, vitem(V_ADDC, mask, x, nullptr, nullptr, nullptr, a, 0)); }
    ResAwait dot(int mask, double* x, double* y, double* x1 = nullptr, double* y1 = nullptr) { return opr(OP_VEC, vitem(V_DOT, mask, x, x1, y, y1, 0, 0)); }
    ResAwait sum(int mask, double* x) { return opr(OP_VEC, vitem(V_SUM, mask, x, nullptr, nullptr, nullptr, 0, 0)); }
    ResAwait sumsq(int mask, double* x, double* x1 = nullptr) { return opr(OP_VEC, vitem(V_SUMSQ, mask, x, x1, nullptr, nullptr, 0, 0)); }
    ResAwait maxabs(int mask, double* x, double* x1 = nullptr) { return opr(OP_VEC, vitem(V_MAXABS, mask, x, x1, nullptr, nullptr, 0, 0)); }
    ResAwait vmax(int mask, double* x) { return opr(OP_VEC, vitem(V_MAX, mask, x, nullptr, nullptr, nullptr, 0, 0)); }
    ResAwait count(int mask) { return opr(OP_VEC, vitem(V_COUNT, mask, P->pn, nullptr, nullptr, nullptr, 0, 0)); }

    // FFT product u[ik] (mask) = sum_jk C[ik][jk] * p[jk] (conv_item)
    void conv(int which, int mask, int nout, const int* ik, double* const* out, int nin, const int* jk,
              double* const* in, const double* const* addend = nullptr, bool fp32 = false)
    {
        AnyItem it;
        if (P->dense) {
            // small patch: the dense product on FP64 MFMA from the coefficient tables (FP64
            // also where the FFT product would run in FP32)
            it.dense = dense_item(P, which, mask, nout, ik, out, nin, jk, in, addend);
            op(OP_DENSE, it);
            return;
        }
        it.conv = conv_item(P, which, mask, nout, ik, out, nin, jk, in, addend, fp32);
        const auto sh = E->conv_shape(fp32, P->Lx, P->Ly, nin, nout);
        it.conv.rpg = sh.first;
        it.conv.cpg = sh.second;
        op(fp32 ? OP_CONV32 : OP_CONV, it);
    }
    // normal-normal product
    void conv_n(int which, int mask, double* out, double* in, bool fp32 = false)
    {
        int ik[1] = {2}, jk[1] = {2};
        double* o[1] = {out};
        double* i[1] = {in};
        conv(which, mask, 1, ik, o, 1, jk, i, nullptr, fp32);
    }
    // tangential-tangential product with optional addend
    void conv_t(int which, int mask, double* ox, double* oy, double* ix, double* iy,
                const double* ax = nullptr, const double* ay = nullptr, bool fp32 = false)
    {
        int ik[2] = {0, 1}, jk[2] = {0, 1};
        double* o[2] = {ox, oy};
        double* i[2] = {ix, iy};
        const double* ad[2] = {ax, ay};
        conv(which, mask, 2, ik, o, 2, jk, i, (ax || ay) ? ad : nullptr, fp32);
    }
};

// ------------------------------------------------------------------------------------------
// engine basics
// ------------------------------------------------------------------------------------------

// The fixed device array of the solve that allocates nothing (cntc_calculatenoalloc): 64 MiB
// per GPU, held until the process ends.  Host code never uses the name for its memory: the
// engine looks up the array's device address on its GPU at set-up and uses that address.
// Layout per call: the patch arena and its twiddle tables, then the scheduler's staging,
// partials, results and arrival counters.
constexpr size_t NOALLOC_BYTES = 64ull << 20;
constexpr size_t NOALLOC_STAGE = 4ull << 20, NOALLOC_PART = 2ull << 20, NOALLOC_RES = 1ull << 20,
                 NOALLOC_CNT = 256ull << 10;
constexpr size_t NOALLOC_ARENA = NOALLOC_BYTES - NOALLOC_STAGE - NOALLOC_PART - NOALLOC_RES - NOALLOC_CNT;
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

size_t Engine::noalloc_bytes() { return NOALLOC_BYTES; }
size_t Engine::noalloc_patch_limit() { return NOALLOC_ARENA; }

Engine::Engine()
{
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    sched_.reset(new Scheduler(stream_));
    // set-up of the no-allocation path on this GPU (the current device): the fixed array's
    // device address
    hipError_t e = hipGetSymbolAddress((void**)&noalloc_base_, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        noalloc_base_ = nullptr;
        (void)hipGetLastError();
    }
}

Engine::~Engine()
{
    serve_stop();
    if (const char* v = std::getenv("CNTC_FLUSH_STATS"); v && v[0] == '1' && sched_) {
        std::fprintf(stderr, "flushes %ld, launches %ld, host issue %.3f s, wait %.3f s\n", sched_->nflush,
                     sched_->flush_launches, sched_->flush_issue_s, sched_->flush_wait_s);
        int tuned = 0;
        for (auto& kv : shapes_) tuned += kv.second.first != 0 || kv.second.second != 0;
        std::fprintf(stderr, "FFT size classes %zu, %d with tuned tile shapes\n", shapes_.size(), tuned);
    }
    sched_.reset();
    for (auto& t : tw_) (void)hipFree(t.second);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// twiddle table exp(-2 pi i e / L), e < L, into host memory h (2L values)
template <typename T>
static void twiddle_table(int L, T* h)
{
    for (int e = 0; e < L; e++) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * e / L;
        h[2 * e] = (T)cosl(a);
        h[2 * e + 1] = (T)sinl(a);
    }
}

// FFT tile-shape tuner.  For one size class (precision, Lx, Ly, nin, nout) it runs the
// product on synthetic zero data of that size (mx = Lx/2, my = Ly/2, every (out, in) block)
// for the default shapes and for half and twice the rows per row-stage work-group and the
// columns per column-stage work-group (within the LDS), three times each, timed by HIP
// events on the engine's stream, and keeps the fastest.  Every shape computes the same bits.
std::pair<int, int> Engine::conv_shape(bool fp32, int Lx, int Ly, int nin, int nout)
{
    const long key = ((((long)(fp32 ? 1 : 0) * 8192 + Lx) * 8192 + Ly) * 4 + nin) * 4 + nout;
    auto f = shapes_.find(key);
    if (f != shapes_.end()) return f->second;
    if (noalloc_) return {0, 0};
    if (const char* v = std::getenv("CNTC_FFT_TUNE"); v && v[0] == '0') return shapes_[key] = {0, 0};
    const int kind = fp32 ? OP_CONV32 : OP_CONV;
    const int M = Lx / 2, H = M + 1, mx = M, my = std::max(1, Ly / 2);
    const size_t es = fp32 ? 8 : 16, npt = (size_t)mx * my, bsz = (size_t)H * Ly;
    // device buffers: inputs, outputs, spectra of every block, the two row-spectrum
    // workspaces (3 components of Ly rows of H bins, as a patch arena has), item and group table
    const size_t b_in = 8 * npt * nin, b_out = 8 * npt * nout, b_spec = es * bsz * nin * nout,
                 b_ws = 16 * 3 * bsz, b_item = 4096;
    char* base = nullptr;
    const size_t total = b_in + b_out + b_spec + 2 * b_ws + b_item + 5 * 256;
    if (hipMalloc(&base, total) != hipSuccess) {
        (void)hipGetLastError();
        return shapes_[key] = {0, 0};
    }
    HIPCHK(hipMemsetAsync(base, 0, total, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char* q = base;
    auto take = [&](size_t b) { char* r = q; q += al(b); return r; };
    ConvItem c{};
    c.mx = mx; c.my = my; c.Lx = Lx; c.Ly = Ly;
    c.nin = nin; c.nout = nout; c.mask = M_ALL; c.mode = 0;
    double* in = (double*)take(b_in);
    double* out = (double*)take(b_out);
    char* spec = take(b_spec);
    for (int j = 0; j < nin; j++) c.in[j] = in + (size_t)j * npt;
    for (int i = 0; i < nout; i++) {
        c.out[i] = out + (size_t)i * npt;
        for (int j = 0; j < nin; j++) c.spec[i][j] = spec + (size_t)(i * nin + j) * bsz * es;
    }
    c.wsS = take(b_ws);
    c.wsT = take(b_ws);
    c.twM = twiddles(M, fp32);
    c.twLx = twiddles(Lx, fp32);
    c.twLy = twiddles(Ly, fp32);
    c.scale = 1.0;
    char* d_item = take(b_item);   // the item at 0, its group table at 2048
    static_assert(sizeof(ConvItem) <= 2048, "the tuner's item slot");
    int r0 = 0, c0 = 0;
    conv_default_shape(&c, &r0, &c0);
    const int nc = std::max(nin, nout);
    // the stages of one shape, best of three, in milliseconds (stages: {0, 2} rows, {1} columns)
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto timed = [&](int rpg, int cpg, std::initializer_list<int> stages) {
        c.rpg = rpg;
        c.cpg = cpg;
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            float tot = 0.f;
            for (int st : stages) {
                const int ng = op_groups(kind, &c, st);
                const int gs[2] = {0, ng};
                // (the previous launch has ended: its event was waited for)
                HIPCHK(hipMemcpy(d_item, &c, sizeof c, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(d_item + 2048, gs, sizeof gs, hipMemcpyHostToDevice));
                HIPCHK(hipEventRecord(e0, stream_));
                launch_op(kind, d_item, (const int*)(d_item + 2048), 1, ng, RedOut{}, st, op_lds(kind, &c, st),
                          stream_);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(e1, stream_));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                tot += ms;
            }
            best = std::min(best, tot);
        }
        return best;
    };
    int br = 0, bc = 0;
    {
        float tbest = timed(0, 0, {0, 2});
        for (int r : {r0 / 2, r0 * 2}) {
            if (r < 1 || r > 32 || r == r0 || 2 * (size_t)r * M * es + 2048 > MAX_LDS) continue;
            const float t = timed(r, 0, {0, 2});
            if (t < tbest) { tbest = t; br = r; }
        }
    }
    {
        float tbest = timed(0, 0, {1});
        for (int cc : {c0 / 2, c0 * 2}) {
            if (cc < 1 || cc > 16 || cc == c0 || 2 * (size_t)nc * cc * Ly * es + 2048 > MAX_LDS) continue;
            const float t = timed(0, cc, {1});
            if (t < tbest) { tbest = t; bc = cc; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipStreamSynchronize(stream_));
    HIPCHK(hipFree(base));
    return shapes_[key] = {br, bc};
}

const void* Engine::twiddles(int L, bool fp32)
{
    long key = (long)L * 2 + (fp32 ? 1 : 0);
    for (auto& t : tw_)
        if (t.first == key) return t.second;
    void* d = nullptr;
    if (fp32) {
        std::vector<float> h(2 * (size_t)L);
        for (int e = 0; e < L; e++) {
            long double a = -2.0L * 3.141592653589793238462643383279502884L * e / L;
            h[2 * e] = (float)cosl(a);
            h[2 * e + 1] = (float)sinl(a);
        }
        HIPCHK(hipMalloc(&d, h.size() * sizeof(float)));
        HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
        std::vector<double> h(2 * (size_t)L);
        for (int e = 0; e < L; e++) {
            long double a = -2.0L * 3.141592653589793238462643383279502884L * e / L;
            h[2 * e] = (double)cosl(a);
            h[2 * e + 1] = (double)sinl(a);
        }
        HIPCHK(hipMalloc(&d, h.size() * sizeof(double)));
        HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    tw_.push_back({key, d});
    return d;
}

// ------------------------------------------------------------------------------------------
// host set-up of a case: material, Hertz, grid, undeformed distance, rigid slip
// ------------------------------------------------------------------------------------------

// complete elliptic integrals B(m), D(m) (mc = 1-m) by Carlson's duplication algorithms for
// R_F and R_D (B.C. Carlson, Numer. Algorithms 10 (1995) 13-26): K = R_F(0,mc,1),
// D = R_D(0,mc,1)/3, B = K - D; used in place of ellip_bd (src/m_hertz.f90:503-1004)
static double rf(double x, double y, double z)
{
    double a0 = (x + y + z) / 3.0;
    double q = std::pow(3e-17, -1.0 / 6.0) * std::max({std::fabs(a0 - x), std::fabs(a0 - y), std::fabs(a0 - z)});
    double a = a0, f = 1.0;
    for (int it = 0; it < 60 && q * f >= std::fabs(a); it++) {
        double sx = std::sqrt(x), sy = std::sqrt(y), sz = std::sqrt(z), lam = sx * sy + sx * sz + sy * sz;
        x = 0.25 * (x + lam); y = 0.25 * (y + lam); z = 0.25 * (z + lam); a = 0.25 * (a + lam); f *= 0.25;
    }
    double X = (a - x) / a, Y = (a - y) / a, Z = -(X + Y), e2 = X * Y - Z * Z, e3 = X * Y * Z;
    return (1.0 - e2 / 10.0 + e3 / 14.0 + e2 * e2 / 24.0 - 3.0 * e2 * e3 / 44.0) / std::sqrt(a);
}
static double rd(double x, double y, double z)
{
    double a0 = (x + y + 3.0 * z) / 5.0;
    double q = std::pow(0.25e-17, -1.0 / 6.0) * std::max({std::fabs(a0 - x), std::fabs(a0 - y), std::fabs(a0 - z)});
    double a = a0, f = 1.0, s = 0.0;
    for (int it = 0; it < 60 && q * f >= std::fabs(a); it++) {
        double sx = std::sqrt(x), sy = std::sqrt(y), sz = std::sqrt(z), lam = sx * sy + sx * sz + sy * sz;
        s += f / (sz * (z + lam));
        x = 0.25 * (x + lam); y = 0.25 * (y + lam); z = 0.25 * (z + lam); a = 0.25 * (a + lam); f *= 0.25;
    }
    double X = (a - x) / a, Y = (a - y) / a, Z = -(X + Y) / 3.0;
    double e2 = X * Y - 6.0 * Z * Z, e3 = (3.0 * X * Y - 8.0 * Z * Z) * Z, e4 = 3.0 * (X * Y - Z * Z) * Z * Z,
           e5 = X * Y * Z * Z * Z;
    double ser = 1.0 - 3.0 * e2 / 14.0 + e3 / 6.0 + 9.0 * e2 * e2 / 88.0 - 3.0 * e4 / 22.0 - 9.0 * e2 * e3 / 52.0 + 3.0 * e5 / 26.0;
    return f * ser / (a * std::sqrt(a)) + 3.0 * s;
}
static double eli(double elmodk, double& K, double& E)
{
    double mc = 1.0 - elmodk * elmodk, b, d;
    if (1.0 - mc < 1.11e-16) { b = PI / 4; d = PI / 4; }
    else {
        double mcc = mc < 1.11e-32 ? 1.11e-32 : mc;
        double kk = rf(0.0, mcc, 1.0);
        d = rd(0.0, mcc, 1.0) / 3.0;
        b = kk - d;
    }
    K = b + d;
    E = b + mc * d;
    return (E - b) / b;
}

// hzcalc3d, src/m_hertz.f90:267-384
static bool hertz3d(double e_star, int ipotcn, int norm, double& a1, double& b1, double& aa, double& bb,
                    double& pen, double& fn, double& cp, double& rho)
{
    double K, E, y;
    if (ipotcn == -1) {
        if (std::min(a1, b1) < TINY) return false;
        bool zbla = b1 <= a1;
        y = zbla ? b1 / a1 : a1 / b1;
        // bisection on the modulus, src/m_hertz.f90:386-442
        const double epshz = sp(1e-9f);
        double xl = 0.0, xr = 1.0, x = 0.5;
        double elr = eli(xr, K, E), ell = eli(xl, K, E);
        while (std::fabs(xr - xl) > epshz) {
            if ((y - elr) * (y - ell) > 0.0) return false;
            x = 0.5 * (xl + xr);
            double elx = eli(x, K, E);
            if ((y - elr) * (y - elx) <= 0.0) { xl = x; ell = elx; }
            else { xr = x; elr = elx; }
        }
        eli(x, K, E);
        double g = std::sqrt(std::max(1e-40, 1.0 - x * x)), sg = std::sqrt(g);
        rho = 2.0 / (a1 + b1);
        if (norm == 1) {
            cp = std::pow((3.0 * std::max(0.0, fn) * rho * E) / (4.0 * PI * e_star * sg), 1.0 / 3.0);
            pen = 2.0 * (cp * sg) * (cp * sg) * K / (rho * E);
        } else {
            cp = std::sqrt((std::max(0.0, pen) * rho * E) / (2.0 * K * sg * sg));
            fn = 4.0 * PI * cp * cp * cp * e_star * sg / (3.0 * rho * E);
        }
        if (zbla) { aa = cp * sg; bb = cp / sg; } else { aa = cp / sg; bb = cp * sg; }
        return true;
    }
    bool zbla = bb <= aa;
    double g = zbla ? bb / aa : aa / bb, sg = std::sqrt(g);
    y = eli(std::sqrt(1.0 - g * g), K, E);
    cp = std::sqrt(aa * bb);
    if (norm == 1) {
        rho = 4.0 * PI * cp * cp * cp * e_star * sg / (3.0 * fn * E);
        pen = 2.0 * (cp * sg) * (cp * sg) * K / (rho * E);
    } else {
        pen = std::max(pen, 1e-9);
        rho = 2.0 * (cp * sg) * (cp * sg) * K / (pen * E);
        fn = 4.0 * PI * cp * cp * cp * e_star * sg / (3.0 * rho * E);
    }
    double apb = 2.0 / rho, ama = apb / (y + 1.0), ami = y * ama;
    if (zbla) { a1 = ami; b1 = ama; } else { a1 = ama; b1 = ami; }
    return true;
}

// elliptic_to_curv, src/m_hertz.f90:95-213: the curvature b1 for which the Hertz ellipse of
// curvature a1 has semi-axis ratio aob -- b1 halved or doubled from b1 = a1 until aob is
// bracketed, then regula falsi to a relative tolerance of 10 epshz
static bool elliptic_to_curv(double e_star, int norm, double a1, double aob, double& b1, double& aa,
                             double& bb, double& pen, double& fn, double& cp, double& rho)
{
    const double epshz = sp(1e-9f), epsaob = 10.0 * epshz;
    int iter = 0;
    double a1w = a1;
    b1 = a1;
    auto solve = [&]() { return hertz3d(e_star, -1, norm, a1w, b1, aa, bb, pen, fn, cp, rho); };
    if (!solve()) return false;
    if (aa < TINY && bb < TINY) return true;
    double b1low, b1high, aoblow, aobhigh;
    if (aa / bb > aob) {
        b1high = b1;
        aobhigh = aa / bb;
        while (iter <= 100 && aa / bb > aob) {
            iter++;
            b1high = b1;
            aobhigh = aa / bb;
            b1 *= 0.5;
            if (!solve()) return false;
        }
        b1low = b1;
        aoblow = aa / bb;
    } else {
        b1low = b1;
        aoblow = aa / bb;
        while (iter <= 100 && aa / bb < aob) {
            iter++;
            b1low = b1;
            aoblow = aa / bb;
            b1 *= 2.0;
            if (!solve()) return false;
        }
        b1high = b1;
        aobhigh = aa / bb;
    }
    while (iter <= 100 && std::fabs(aa / bb - aob) > epsaob * aob) {
        iter++;
        double rico = (aobhigh - aoblow) / (b1high - b1low);
        b1 = b1low + (aob - aoblow) / rico;
        if (!solve()) return false;
        if (aa / bb > aob) {
            b1high = b1;
            aobhigh = aa / bb;
        } else {
            b1low = b1;
            aoblow = aa / bb;
        }
    }
    return true;
}

// Host part of contac's set-up (src/m_scontc.f90:37-176): checks, Hertz, grid, h and w.
// device bytes of one patch's arena (grid functions, coefficients, spectra, workspaces)
size_t patch_bytes(int mx, int my)
{
    int Lx = 2 * opt_fft_size(mx), Ly = 2 * opt_fft_size(my);
    int M = Lx / 2, H = M + 1;
    size_t n = (size_t)mx * my;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t nb = (size_t)(2 * mx) * (2 * my);
    size_t bsz = (size_t)H * Ly;
    size_t pH = (size_t)mx + 1, pL = 2 * (size_t)my;
    size_t ws = std::max((size_t)3 * Ly * H, pH * pL) * 16;
    size_t emb = std::max((size_t)Lx * Ly, nb) * 8;
    return al(n) + al(46 * n * 8) + al(9 * nb * 8) * 2 + al(3 * nb * 8) + al(9 * bsz * 16) * 2 +
           al(3 * bsz * 16) + al(9 * bsz * 8) + al(3 * bsz * 8) + al(3 * emb) + 2 * al(ws) + al(64 * 8) +
           4096;
}

bool Engine::prepare(Patch& P)
{
    CaseInput& in = P.in;
    CaseResult& R = P.out;
    R = CaseResult();
    if (in.unsupported) {
        R.ierror = ERR_OTHER;
        last_error = in.unsupported;
        return false;
    }
    bool hertzian = in.ipotcn <= -1 && in.ipotcn >= -5;
    const bool hertz2d = in.ipotcn == -4 || in.ipotcn == -5;
    if (in.bound != 0 && in.bound != 2) {
        R.ierror = ERR_OTHER;
        last_error = "the B digit supports 0 and 2";
        return false;
    }
    if (in.bound == 2 && !hertzian) {
        // src/m_sinput.f90:176-181: B=2 needs a Hertzian potential contact
        R.ierror = ERR_OTHER;
        last_error = "B=2 needs a Hertzian potential contact";
        return false;
    }
    if (hertz2d && in.norm != 1) {
        R.ierror = ERR_OTHER;
        last_error = "2D Hertz needs N=1";
        return false;
    }
    if (in.tang < 0 || in.tang > 3 || in.ipotcn == 0 || (!hertzian && in.ibase == 0)) {
        R.ierror = ERR_OTHER;
        last_error = "potential contact or undeformed distance not specified";
        return false;
    }
    if (in.pvtime != 0 && in.pvtime != 2) { R.ierror = ERR_OTHER; return false; }
    // combin_mater, src/m_hierarch_data.f90:1543-1553
    P.ga = 2.0 / (1.0 / in.gg1 + 1.0 / in.gg2);
    P.nu = P.ga * (in.poiss1 / in.gg1 + in.poiss2 / in.gg2) / 2.0;
    P.ak = (P.ga / 4.0) * ((1.0 - 2.0 * in.poiss1) / in.gg1 - (1.0 - 2.0 * in.poiss2) / in.gg2);
    P.nt_cpl = std::fabs(P.ak) >= 1e-6;
    P.muscal = in.fstat;
    P.fntrue = in.fn;
    P.pen = in.pen;
    // potential contact and grid (potcon_fill / potcon_hertz, src/m_hierarch_data.f90:1801-1959)
    int mx = in.mx, my = in.my;
    double dx = in.dx, dy = in.dy, xc1 = in.xc1, yc1 = in.yc1;
    int ibase = in.ibase;
    double prm[6];
    std::memcpy(prm, in.prmudf, sizeof prm);
    if (hertzian) {
        // hzsol, src/m_hertz.f90:29-90
        double e_star = P.ga / (1.0 - P.nu);
        R.hz_a1 = in.hz_a1; R.hz_b1 = in.hz_b1; R.hz_aa = in.hz_aa; R.hz_bb = in.hz_bb;
        bool ok = true;
        if (hertz2d) {
            // hzcalc2d, src/m_hertz.f90:217-263: line contact of half-width bb
            const double fn_2d = P.fntrue / (2.0 * R.hz_bb);
            R.hz_b1 = 0.0;
            if (in.ipotcn == -4) {
                const double rx = 1.0 / (2.0 * R.hz_a1);
                R.hz_aa = std::sqrt(4.0 * rx * fn_2d / (PI * e_star));
            } else {
                const double rx = PI * R.hz_aa * R.hz_aa * e_star / (4.0 * fn_2d);
                R.hz_a1 = 1.0 / (2.0 * rx);
            }
            R.hz_rho = 2.0 / (R.hz_a1 + R.hz_b1);
            R.hz_cp = std::sqrt(4.0 * R.hz_aa * R.hz_bb / PI);
        } else if (in.ipotcn == -2) {
            ok = elliptic_to_curv(e_star, in.norm, in.hz_a1, in.hz_aob, R.hz_b1, R.hz_aa, R.hz_bb, P.pen,
                                  P.fntrue, R.hz_cp, R.hz_rho);
        } else {
            ok = hertz3d(e_star, in.ipotcn, in.norm, R.hz_a1, R.hz_b1, R.hz_aa, R.hz_bb, P.pen, P.fntrue,
                         R.hz_cp, R.hz_rho);
        }
        if (!ok) { R.ierror = ERR_OTHER; last_error = "Hertz solution failed"; return false; }
        ibase = 1;
        std::memset(prm, 0, sizeof prm);
        prm[0] = R.hz_a1;
        prm[2] = R.hz_b1;
        // potcon_hertz, src/m_hierarch_data.f90:1889-1959
        double xl = -in.hz_scale * std::max(1e-9, R.hz_aa), yl = -in.hz_scale * std::max(1e-9, R.hz_bb);
        dx = (-xl - xl) / mx;
        dy = (-yl - yl) / my;
        if (hertz2d) {
            // n_int interior rows of the strip: dy such that n_int dy = 2 bb
            int n_int = my;
            if (in.hz_scale > 1.0) {
                n_int = std::max(1, (int)std::lround(my / in.hz_scale));
                if ((my - n_int) % 2 == 1) n_int = n_int + 1;
            }
            dy = 2.0 * std::max(1e-9, R.hz_bb) / n_int;
            yl = -my * dy / 2.0;
        }
        xc1 = xl + 0.5 * dx;
        yc1 = yl + 0.5 * dy;
    } else if (in.ipotcn == 1) {
        xc1 = in.xl + 0.5 * dx;
        yc1 = in.yl + 0.5 * dy;
    } else if (in.ipotcn == 2) {
        dx = (in.xh - in.xl) / mx;
        dy = (in.yh - in.yl) / my;
        xc1 = in.xl + 0.5 * dx;
        yc1 = in.yl + 0.5 * dy;
    } else if (in.ipotcn == 4) {
        dx = (in.xcm - in.xc1) / std::max(1, mx - 1);
        dy = (in.ycm - in.yc1) / std::max(1, my - 1);
    } else if (in.ipotcn != 3) {
        R.ierror = ERR_OTHER;
        return false;
    }
    // a single row (2-D problems) only with ConvexGS: the row-walk solvers are not set up for it
    if (mx < 2 || my < 1 || (my < 2 && in.gausei != 2)) {
        R.ierror = ERR_OTHER;
        last_error = "grids need mx >= 2 and my >= 2 (my = 1 with ConvexGS, G=2)";
        return false;
    }
    bool newgrid = mx != P.mx || my != P.my;
    if (newgrid && P.have_solution && (in.pvtime != 2 || in.iestim != 0)) {
        R.ierror = ERR_OTHER;
        last_error = "continuation needs an unchanged grid";
        return false;
    }
    // K != 0 couples the normal and tangential problems.  Transient rolling after a previous
    // case at dq != dx would need A'_tn pv_n on the interior but not at the leading edge
    // (form 1b), which is not built: that combination is refused before the patch changes.
    if (P.nt_cpl && in.tang == 2 && in.pvtime == 0 && (in.dq > 0 ? in.dq : dx) > dx * (1.0 + 1e-12)) {
        R.ierror = ERR_OTHER;
        last_error = "dissimilar materials (K != 0) in transient rolling at dq != dx after a previous case are not supported";
        return false;
    }
    P.mx = mx; P.my = my; P.n = mx * my;
    P.dx = dx; P.dy = dy; P.dxdy = dx * dy; P.xc1 = xc1; P.yc1 = yc1;
    // check_roll_stepsize, src/m_sdis.f90:125-204
    bool is_roll = in.tang == 2 || in.tang == 3;
    P.chi = in.chi;
    P.dq = in.dq > 0 ? in.dq : dx;
    if (in.tang == 3 && in.gausei != 2) {
        // steady rolling with SteadyGS or GDsteady: dq = dx; ConvexGS (G=2) keeps dq
        if (std::fabs(P.chi) > 0.01 && std::fabs(P.chi - PI) > 0.01) P.chi = 0.0;
        P.dq = dx;
    }
    if (!is_roll) { P.chi = 0.0; P.dq = 1.0; }
    if (std::fabs(P.chi) > 0.01) { R.ierror = ERR_OTHER; last_error = "rolling direction chi=0 only"; return false; }
    // transient rolling with dq > dx: elements entering the contact area near the leading edge
    // spend part of the step in it (sxbnd/subnd, src/m_leadedge.f90); with dq <= dx every
    // interior element has facdt = 1 and no leading-edge position is used
    P.ledge = (in.tang == 2 || (in.tang == 3 && in.gausei == 2)) && P.dq > dx * (1.0 + 1e-12);
    // set_norm_rhs, src/m_sdis.f90:321-494: the quadratic gap is formed on the GPU; the host
    // keeps a copy only where it is needed (element-wise input, or eldiv0's initial estimate)
    int ihertz = in.ipotcn < 0 ? -in.ipotcn : 0;
    P.ibase_eff = ibase;
    std::memcpy(P.prm, prm, sizeof prm);
    bool need_host_h = ibase == 9 || (ihertz == 0 && in.iestim == 0);
    P.h_host.clear();
    if (need_host_h) {
        P.h_host.resize(P.n);
        for (int iy = 0; iy < my; iy++)
            for (int ix = 0; ix < mx; ix++) {
                int i = ix + iy * mx;
                double x = xc1 + ix * dx, y = yc1 + iy * dy;
                if (ibase == 1)
                    P.h_host[i] = prm[0] * x * x + prm[1] * x * y + prm[2] * y * y + prm[3] * x + prm[4] * y + prm[5];
                else if (ibase == 9) {
                    if ((int)in.h_elem.size() < P.n) { R.ierror = ERR_OTHER; return false; }
                    P.h_host[i] = in.h_elem[i];
                } else { R.ierror = ERR_OTHER; return false; }
            }
    } else if (ibase != 1) {
        R.ierror = ERR_OTHER;
        return false;
    }
    // padded sizes of the products
    P.Lx = 2 * opt_fft_size(mx);
    P.Ly = 2 * opt_fft_size(my);
    if (P.Ly > 2048 || P.Lx > 4096) {   // column stage of the products must fit the LDS
        R.ierror = ERR_OTHER;
        last_error = "grid too large for the FFT products (my <= 1024, mx <= 2048)";
        return false;
    }
    // device memory
    int M = P.Lx / 2, H = M + 1;
    size_t n = P.n;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t nb = (size_t)(2 * mx) * (2 * my);
    size_t bsz = (size_t)H * P.Ly;
    size_t pH = (size_t)mx + 1, pL = 2 * (size_t)my;   // preconditioner transform (2mx x 2my)
    size_t ws = std::max((size_t)3 * P.Ly * H, pH * pL) * 16;
    size_t emb = std::max((size_t)P.Lx * P.Ly, (size_t)(2 * mx) * (2 * my)) * 8;
    size_t need = patch_bytes(mx, my);
    // twiddle tables of the no-allocation path, after the arena: FP64 and FP32 at M, Lx, Ly,
    // FP64 at mx, 2mx, 2my (the preconditioner's transform)
    const int twl[9] = {M, P.Lx, P.Ly, M, P.Lx, P.Ly, mx, 2 * mx, 2 * my};
    const int twsz[9] = {16, 16, 16, 8, 8, 8, 16, 16, 16};
    size_t twbytes = 0;
    for (int k = 0; k < 9; k++) twbytes += al((size_t)twl[k] * twsz[k]);
    char* b;
    if (noalloc_) {
        if (need + twbytes > NOALLOC_ARENA) {
            R.ierror = ERR_OTHER;
            char msg[160];
            std::snprintf(msg, sizeof msg, "a %dx%d patch needs %zu bytes, more than the %zu of the "
                          "no-allocation array", mx, my, need + twbytes, NOALLOC_ARENA);
            last_error = msg;
            return false;
        }
        b = noalloc_base_;
    } else {
        P.mem.ensure(need);
        b = P.mem.base;
    }
    char* const arena = b;
    auto take = [&](size_t bytes) { char* r = b; b += al(bytes); return r; };
    P.el = (int8_t*)take(n);
    double* d = (double*)take(46 * n * 8);
    P.pn = d; P.px = d + n; P.py = d + 2 * n; P.sx = d + 3 * n; P.sy = d + 4 * n;
    P.pvx = d + 5 * n; P.pvy = d + 6 * n; P.pvn = d + 7 * n;
    P.h = d + 8 * n; P.hx = d + 9 * n; P.hy = d + 10 * n; P.wsx = d + 11 * n; P.wsy = d + 12 * n;
    for (int k = 0; k < 24; k++) P.w[k] = d + (13 + k) * n;
    P.un = d + 37 * n; P.ux = d + 38 * n; P.uy = d + 39 * n;
    P.facdt = d + 40 * n; P.lix = d + 41 * n;
    P.hn = P.nt_cpl ? d + 42 * n : P.h;
    for (int k = 0; k < 3; k++) P.po[k] = d + (43 + k) * n;
    P.cf = (double*)take(9 * nb * 8);
    P.cfv = (double*)take(9 * nb * 8);
    P.cfm = (double*)take(3 * nb * 8);
    P.spec = take(9 * bsz * 16);
    P.specv = take(9 * bsz * 16);
    P.specm = take(3 * bsz * 16);
    P.spec32 = take(9 * bsz * 8);
    P.specm32 = take(3 * bsz * 8);
    P.embed = (double*)take(3 * emb);
    P.wsS = take(ws);
    P.wsT = take(ws);
    P.sc = (double*)take(64 * 8);
    if (noalloc_) {
        // tables computed into the pinned staging tail and copied into the array
        char* host = sched_->host_tail(twbytes);
        char* dev = b;
        const void** dst[9] = {&P.twM, &P.twLx, &P.twLy, &P.twM32, &P.twLx32, &P.twLy32,
                               &P.ptwM, &P.ptwLx, &P.ptwLy};
        size_t o = 0;
        for (int k = 0; k < 9; k++) {
            if (twsz[k] == 16) twiddle_table(twl[k], (double*)(host + o));
            else twiddle_table(twl[k], (float*)(host + o));
            *dst[k] = dev + o;
            o += al((size_t)twl[k] * twsz[k]);
        }
        HIPCHK(hipMemcpyAsync(dev, host, twbytes, hipMemcpyHostToDevice, stream_));
    } else {
        P.twM = twiddles(M, false); P.twLx = twiddles(P.Lx, false); P.twLy = twiddles(P.Ly, false);
        P.twM32 = twiddles(M, true); P.twLx32 = twiddles(P.Lx, true); P.twLy32 = twiddles(P.Ly, true);
        P.ptwM = twiddles(mx, false); P.ptwLx = twiddles(2 * mx, false); P.ptwLy = twiddles(2 * my, false);
    }
    // The element states and the first 8 grid functions (tractions, slip, previous tractions)
    // carry a problem's state to its next case.  They are in this arena unless the last solve
    // allocated nothing (then on the host) or used the other memory (arena vs fixed array).
    const size_t lead = al(n) + 8 * n * 8;   // the arena's leading block: el, pn .. pvn
    if (newgrid) {
        HIPCHK(hipMemsetAsync(P.el, 0, n, stream_));
        HIPCHK(hipMemsetAsync(P.pn, 0, 8 * n * 8, stream_));
        P.have_solution = false;
        P.state_on_host = false;
    } else if (P.state_on_host && P.host_state.size() >= lead) {
        HIPCHK(hipMemcpyAsync(arena, P.host_state.data(), lead, hipMemcpyHostToDevice, stream_));
    } else if (noalloc_ && P.mem.base && P.mem.bytes >= lead) {
        HIPCHK(hipMemcpyAsync(arena, P.mem.base, lead, hipMemcpyDeviceToDevice, stream_));
    } else if (noalloc_) {
        HIPCHK(hipMemsetAsync(P.el, 0, n, stream_));
        HIPCHK(hipMemsetAsync(P.pn, 0, 8 * n * 8, stream_));
        P.have_solution = false;
    }
    if (!noalloc_) P.state_on_host = false;
    if (ibase == 9)
        HIPCHK(hipMemcpyAsync(P.h, P.h_host.data(), n * 8, hipMemcpyHostToDevice, stream_));
    if (in.rztang == 1) {
        // extra rigid slip of cntc_setextrarigidslip, sized for the grid it was given for
        if (in.exrhs.size() != 2 * n) {
            R.ierror = ERR_OTHER;
            last_error = "extra rigid slip was set for another grid size";
            return false;
        }
        HIPCHK(hipMemcpyAsync(P.hx, in.exrhs.data(), n * 8, hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(P.hy, in.exrhs.data() + n, n * 8, hipMemcpyHostToDevice, stream_));
    }
    // init_curr_data, src/m_sdis.f90:587-768: eldiv0 (src/m_sdis.f90:818-1012) on the host for
    // a non-Hertzian zero initial estimate; the Hertzian estimate is formed on the GPU
    if (in.iestim == 0 && ihertz == 0) {
        const double facpen = sp(0.60f), reltol = sp(0.01f);
        double fnscal = P.fntrue / P.ga;
        size_t imin = 0;
        for (size_t i = 1; i < n; i++)
            if (P.h_host[i] < P.h_host[imin]) imin = i;
        double hsmin = P.h_host[imin];
        double pentru = P.pen - hsmin;
        if (in.norm == 1) {
            double fac = 1.0 * 0.75 * std::sqrt(PI) * (1.0 - P.nu);
            double penmin = 0.0, fnmin = 0.0, penmax = fac * fnscal / std::sqrt(P.dxdy);
            long ncnmax = 0;
            for (size_t i = 0; i < n; i++)
                if (P.h_host[i] - hsmin < facpen * penmax) ncnmax++;
            double fnmax = penmax * std::sqrt(P.dxdy * ncnmax) / fac;
            int iter = 0;
            while (iter < 50 && std::fabs(fnmax - fnmin) > reltol * fnscal) {
                iter++;
                double penmid = 0.5 * (penmax + penmin);
                long ncnmid = 0;
                for (size_t i = 0; i < n; i++)
                    if (P.h_host[i] - hsmin < facpen * penmid) ncnmid++;
                double fnmid = penmid * std::sqrt(P.dxdy * ncnmid) / fac;
                if (fnmid < fnscal) { penmin = penmid; fnmin = fnmid; }
                else { penmax = penmid; fnmax = fnmid; }
            }
            pentru = penmax;
            P.pen = pentru + hsmin;
        }
        P.el_host.resize(n);
        for (size_t i = 0; i < n; i++) P.el_host[i] = (P.h_host[i] - hsmin < facpen * pentru) ? ADHES : EXTER;
        HIPCHK(hipMemcpyAsync(P.el, P.el_host.data(), n, hipMemcpyHostToDevice, stream_));
    }
    // Hertzian pressure: prescribed (B=2) or the initial estimate of I=0 (src/m_sdis.f90:641-670)
    P.hertz_init = ihertz >= 1 && (in.bound == 2 || (in.iestim == 0 && in.bound == 0));
    P.hz_pnmax = !P.hertz_init ? 0.0
                 : ihertz <= 3 ? 3.0 * P.fntrue / (2.0 * PI * R.hz_aa * R.hz_bb)
                               : P.fntrue / (PI * R.hz_aa * R.hz_bb);
    P.xofs = is_roll ? std::cos(P.chi) * P.dq * (1.0 / 6.0) : 0.0;
    P.yofs = is_roll ? std::sin(P.chi) * P.dq * (1.0 / 6.0) : 0.0;
    return true;
}

// ------------------------------------------------------------------------------------------
// device set-up: coefficients and spectra (sgencr, src/m_visc.f90:127-376; fft_makePrec,
// src/m_aijpj.f90:457-708)
// ------------------------------------------------------------------------------------------

// need_cv: the coefficients cv of the previous time (transient rolling) with their spectra;
// cv_table: cv without spectra (steady rolling, for ConvexGS's cs - cv).  The tables and spectra
// of the previous case are kept when grid, material, rolling step and precision are unchanged
// (the reference keeps its spectra c%fft_cf the same way, src/m_aijpj.f90:873-914); the
// solve that allocates nothing shares its memory between problems and always recomputes.
static Task setup_coefficients(Ctx C, bool need_cv, bool cv_table = false)
{
    Patch& P = *C.P;
    Patch::CoefKey key;
    key.cf = P.cf; key.mx = P.mx; key.my = P.my; key.Lx = P.Lx; key.Ly = P.Ly;
    key.cvmode = (need_cv ? 1 : 0) + (cv_table ? 2 : 0);
    key.precision = P.in.precision;
    key.dx = P.dx; key.dy = P.dy; key.ak = P.ak; key.nu = P.nu; key.ga = P.ga;
    key.chi = P.chi; key.dq = P.dq;
    if (P.coef_valid && key == P.coef_key) co_return 0;
    P.coef_valid = false;
    P.prec_valid = 0;
    int mx = P.mx, my = P.my;
    size_t nb = (size_t)(2 * mx) * (2 * my);
    int M = P.Lx / 2, H = M + 1;
    size_t bsz = (size_t)H * P.Ly;
    AnyItem it;
    it.coef.mx = mx; it.coef.my = my;
    it.coef.dx = P.dx; it.coef.dy = P.dy;
    it.coef.xshft = 0.0; it.coef.yshft = 0.0;
    it.coef.akv = P.ak; it.coef.nuv = P.nu;
    it.coef.cf = P.cf;
    C.op(OP_ELASCF, it);
    if (need_cv || cv_table) {
        AnyItem iv = it;
        iv.coef.xshft = std::cos(P.chi) * P.dq;
        iv.coef.yshft = std::sin(P.chi) * P.dq;
        iv.coef.cf = P.cfv;
        C.op(OP_ELASCF, iv);
    }
    co_await C.count(M_ALL);   // completes the coefficient generation before the roll step
    if (need_cv || cv_table) launch_rollcoef((int)nb, P.cf, P.cfv, C.E->stream());

    // spectra of the blocks: embed (wrap, scaled by 1/(G M Ly)) then forward transform
    double scale = (1.0 / P.ga) / ((double)M * (double)P.Ly);
    std::vector<std::pair<const double*, void*>> blocks;
    // xx, xy, yx, yy, zz; the normal-tangential blocks xz, yz, zx, zy are zero unless K != 0
    std::vector<int> bl = {0, 1, 3, 4, 8};
    if (P.nt_cpl) bl = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    for (int b : bl) blocks.push_back({P.cf + (size_t)b * nb, (char*)P.spec + (size_t)b * bsz * 16});
    if (need_cv)
        for (int b : bl) blocks.push_back({P.cfv + (size_t)b * nb, (char*)P.specv + (size_t)b * bsz * 16});
    // three blocks per transform while the column stage fits the LDS (two at Ly = 2048)
    const size_t per = 2 * (size_t)P.Ly * 16;   // LDS per block and column: two buffers, FP64
    const size_t bper = per * 3 <= MAX_LDS ? 3 : (per * 2 <= MAX_LDS ? 2 : 1);
    for (size_t k0 = 0; k0 < blocks.size(); k0 += bper) {
        int nin = (int)std::min<size_t>(bper, blocks.size() - k0);
        AnyItem cv;
        ConvItem& c = cv.conv;
        c.mx = mx; c.my = my; c.Lx = P.Lx; c.Ly = P.Ly;
        c.nin = nin; c.nout = nin; c.mask = M_ALL; c.mode = 1;
        for (int j = 0; j < nin; j++) {
            AnyItem e;
            e.emb.Lx = P.Lx; e.emb.Ly = P.Ly; e.emb.rx = mx; e.emb.ry = my;
            e.emb.cf_mx = mx; e.emb.cf_my = my;
            e.emb.cf = blocks[k0 + j].first;
            e.emb.out = P.embed + (size_t)j * P.Lx * P.Ly;
            e.emb.scale = scale;
            C.op(OP_EMBED, e);
            c.in[j] = e.emb.out;
            c.out[j] = (double*)blocks[k0 + j].second;
        }
        c.wsS = P.wsS; c.wsT = P.wsT;
        c.twM = P.twM; c.twLx = P.twLx; c.twLy = P.twLy;
        c.scale = 1.0;
        C.op(OP_CONV, cv);
    }
    if (P.in.precision == 1) {
        co_await C.count(M_ALL);
        for (int b : bl)
            launch_to_float(2 * bsz, (const double*)((char*)P.spec + (size_t)b * bsz * 16),
                            (float*)((char*)P.spec32 + (size_t)b * bsz * 8), C.E->stream());
    }
    P.coef_key = key;
    P.coef_valid = !C.E->noalloc_active();
    co_return 0;
}

// approximate inverse of diagonal block kk: m = G^2 IFFT(1/FFT(c)) on the 2mx x 2my circulant,
// then its spectrum at the product size
static Task setup_prec(Ctx C, int kk)
{
    Patch& P = *C.P;
    if (P.coef_valid && (P.prec_valid >> kk & 1u)) co_return 0;
    int mx = P.mx, my = P.my;
    size_t nb = (size_t)(2 * mx) * (2 * my);
    int M = P.Lx / 2, H = M + 1;
    size_t bsz = (size_t)H * P.Ly;
    int bidx = kk * 3 + kk;
    int pLx = 2 * mx, pLy = 2 * my;
    AnyItem e;
    e.emb.Lx = pLx; e.emb.Ly = pLy; e.emb.rx = mx; e.emb.ry = my;
    e.emb.cf_mx = mx; e.emb.cf_my = my;
    e.emb.cf = P.cf + (size_t)bidx * nb;
    e.emb.out = P.embed;
    e.emb.scale = 1.0;
    C.op(OP_EMBED, e);
    AnyItem cv;
    ConvItem& c = cv.conv;
    c.mx = mx; c.my = my; c.Lx = pLx; c.Ly = pLy;
    c.nin = 1; c.nout = 1; c.mask = M_ALL; c.mode = 2;
    c.in[0] = P.embed;
    c.out[0] = P.embed + (size_t)pLx * pLy;
    c.wsS = P.wsS; c.wsT = P.wsT;
    c.twM = P.ptwM; c.twLx = P.ptwLx; c.twLy = P.ptwLy;
    c.scale = P.ga * P.ga / ((double)mx * (double)pLy);
    C.op(OP_CONV, cv);
    AnyItem x;
    x.ext.Lx = pLx; x.ext.Ly = pLy; x.ext.cf_mx = mx; x.ext.cf_my = my;
    x.ext.in = P.embed + (size_t)pLx * pLy;
    x.ext.cf = P.cfm + (size_t)kk * nb;
    x.ext.scale = 1.0;
    C.op(OP_EXTRACT, x);
    // spectrum at the product size
    AnyItem e2;
    e2.emb.Lx = P.Lx; e2.emb.Ly = P.Ly; e2.emb.rx = mx; e2.emb.ry = my;
    e2.emb.cf_mx = mx; e2.emb.cf_my = my;
    e2.emb.cf = P.cfm + (size_t)kk * nb;
    e2.emb.out = P.embed;
    e2.emb.scale = (1.0 / P.ga) / ((double)M * (double)P.Ly);
    C.op(OP_EMBED, e2);
    AnyItem cs;
    ConvItem& c2 = cs.conv;
    c2.mx = mx; c2.my = my; c2.Lx = P.Lx; c2.Ly = P.Ly;
    c2.nin = 1; c2.nout = 1; c2.mask = M_ALL; c2.mode = 1;
    c2.in[0] = P.embed;
    c2.out[0] = (double*)((char*)P.specm + (size_t)kk * bsz * 16);
    c2.wsS = P.wsS; c2.wsT = P.wsT;
    c2.twM = P.twM; c2.twLx = P.twLx; c2.twLy = P.twLy;
    c2.scale = 1.0;
    C.op(OP_CONV, cs);
    co_await C.count(M_ALL);
    if (P.in.precision == 1)
        launch_to_float(2 * bsz, (const double*)((char*)P.specm + (size_t)kk * bsz * 16),
                        (float*)((char*)P.specm32 + (size_t)kk * bsz * 8), C.E->stream());
    if (P.coef_valid) P.prec_valid |= 1u << kk;
    co_return 0;
}

// row-extent sum of one influence row with unit tractions, times 1/G (gf3_AijPj)
static ResAwait aij_ones(Ctx& C, int bidx)
{
    Patch& P = *C.P;
    AnyItem it;
    it.aij.mx = P.mx; it.aij.my = P.my;
    it.aij.ix = std::max(1, P.mx / 2);
    it.aij.iy = std::max(1, P.my / 2);
    it.aij.cf_mx = P.mx; it.aij.cf_my = P.my;
    it.aij.el = P.el;
    it.aij.cf = P.cf + (size_t)bidx * (2 * P.mx) * (2 * P.my);
    it.aij.ga_inv = 1.0 / P.ga;
    return C.opr(OP_AIJ_ONES, it);
}

// ------------------------------------------------------------------------------------------
// NormCG, src/m_solvpn.f90:24-461
// ------------------------------------------------------------------------------------------

static NcgItem ncg_item(Patch& P)
{
    NcgItem I{};
    I.n = P.n;
    I.el = P.el;
    I.p = P.pn;
    I.rhs = P.w[0];
    I.res = P.w[1];
    I.dd = P.w[3];
    I.v = P.w[5];
    I.h = P.hn;
    I.unn = P.w[7];
    I.hmin = P.hmin;
    return I;
}

// ps0: initial estimate of a sensitivity solve (is_sens), whose stopping test measures the
// change from it (src/m_solvpn.f90:94-97, 281-289); null otherwise
static Task normcg(Ctx C, int norm, int maxcg, double eps, double* pen, int* itcg_out, double* err_out,
                   double* ps0 = nullptr)
{
    Patch& P = *C.P;
    const bool fp32 = P.in.precision == 1;
    const int npot = P.n;
    double dxdy = P.dxdy, fntrue = P.fntrue;
    double *rhs = P.w[0], *res = P.w[1], *r_prv = P.w[2], *dd = P.w[3], *z = P.w[4], *v = P.w[5], *q = P.w[6];
    double* ps = P.pn;
    int numinn = npot <= 150 ? 3 : (npot <= 400 ? 2 : 1);
    int itcg = 0, ret = 0;
    double davg = 0.0;
    Res r;

    if (norm == 1) *pen = 0.0;
    C.set(M_ALL, rhs, *pen);
    C.axpy(M_ALL, -1.0, P.hn, rhs);
    C.set(M_ALL, res, 0.0);
    C.set(M_ALL, v, 0.0);
    C.set(M_ALL, r_prv, 0.0);
    r = co_await C.count(M_INT);
    long ncon = (long)r.v[0];
    if (norm == 0) {
        if (P.hmin - *pen >= 0.0) {
            // no interpenetration: zero solution, empty contact area
            C.set(M_ALL, ps, 0.0);
            co_await C.count(M_ALL);
            HIPCHK(hipMemsetAsync(P.el, 0, npot, C.E->stream()));
            *itcg_out = 0;
            *err_out = 0.0;
            co_return 0;
        }
    } else {
        if (ncon <= 0) {
            double htrsh = P.hmin + 0.1 * std::max(P.hmax - P.hmin, 1e-10);
            NcgItem I = ncg_item(P);
            I.mode = 1;          // strict h < threshold, pressure kept
            I.hmin = htrsh;
            AnyItem it;
            it.ncg = I;
            r = co_await C.opr(OP_NCG_HMIN, it);
            ncon += (long)r.v[0];
        }
        r = co_await C.sum(M_ALL, ps);
        double fk = dxdy * r.v[0];
        if (std::fabs(fk) < sp(1e-3f) * fntrue) {
            double pn = fntrue / (dxdy * (double)ncon);
            C.set(M_INT, ps, pn);
        } else {
            C.scal(M_INT, ps, fntrue / fk);
        }
    }
    // res = rhs - A ps on the interior
    C.conv_n(0, M_INT, res, ps);
    C.axpby(M_INT, 1.0, rhs, -1.0, res);
    if (norm == 1) {
        r = co_await C.sum(M_INT, res);
        C.addc(M_INT, res, -r.v[0] / std::max(1.0, r.v[1]));
    }
    double rms_xk = 1.0, rms_upd = 2.0 * eps, rms_upd1 = 0.0, alpha = 0.0;
    int itinn = 0;
    bool lchanged = false;
    // The iteration's scalars stay on the device (slots of 4 in the patch's scalar area): the
    // means that project z, v, q on zero sum (N=1), rz2 = z.res (two slots by the parity of
    // the iteration, so that the previous one is rz1), z.r_prv, r.v and v.A v.  The updates
    // form beta and alpha from them as the host formed them (src/m_solvpn.f90:130-250); the
    // iteration's one host round trip is the sum of squares of v, which also returns r.v and
    // v.A v so that the host forms the same alpha.
    double* sc = P.sc;
    double *s_zm = sc + 16, *s_rzr = sc + 28, *s_vm = sc + 32, *s_qm = sc + 36, *s_rv = sc + 40,
           *s_vav = sc + 44;
    while ((lchanged || rms_upd > eps * rms_xk) && itcg < maxcg) {
        itcg++;
        itinn++;
        double* s_rz2 = sc + 20 + 4 * (itcg & 1);
        double* s_rz1 = sc + 20 + 4 * ((itcg + 1) & 1);
        C.conv_n(2, M_INT, z, res, fp32);
        if (norm == 1) {
            C.vdev(V_SUM, M_INT, z, nullptr, s_zm);
            C.vfn(V_ADDC, M_INT, z, nullptr, SF_NEG_MEAN, s_zm);
        }
        C.vdev(V_DOT, M_INT, z, res, s_rz2);
        if (itcg <= 1) {
            C.copy(M_INT, z, v);
        } else {
            C.vdev(V_DOT, M_INT, z, r_prv, s_rzr);
            C.vfn(V_AXPBY, M_INT, z, v, SF_NCG_BETA, s_rz2, s_rzr, s_rz1);
        }
        if (norm == 1) {
            C.vdev(V_SUM, M_INT, v, nullptr, s_vm);
            C.vfn(V_ADDC, M_INT, v, nullptr, SF_NEG_MEAN, s_vm);
        }
        C.conv_n(0, M_INT, q, v, fp32);
        if (norm == 1) {
            C.vdev(V_SUM, M_INT, q, nullptr, s_qm);
            C.vfn(V_ADDC, M_INT, q, nullptr, SF_NEG_MEAN, s_qm);
        }
        C.vdev(V_DOT, M_INT, res, v, s_rv);
        C.vdev(V_DOT, M_INT, q, v, s_vav);
        C.vfn(V_AXPY, M_INT, v, ps, SF_NCG_ALPHA, s_rv, s_vav, nullptr, ncon == 1 ? 1 : 0);
        r = co_await C.sumsq_echo(M_INT, v, s_rv, s_vav);
        const double rv = r.v[2], vav = r.v[3];
        if (std::fabs(vav) > 1e-32 && ncon == 1) alpha = rv / vav;
        else alpha = rv / std::max(TINY, vav);
        rms_upd = std::fabs(alpha) * std::sqrt(r.v[0] / std::max(1.0, r.v[1]));
        if (itcg == 1) rms_upd1 = rms_upd;
        if (itcg <= 3 || itcg % 10 == 0) {
            if (ps0) {
                double* tmp = P.w[8];
                C.copy(M_ALL, ps, tmp);
                C.axpy(M_ALL, -1.0, ps0, tmp);
                r = co_await C.sumsq(M_INT, tmp);
            } else {
                r = co_await C.sumsq(M_INT, ps);
            }
            rms_xk = std::sqrt(r.v[0] / std::max(1.0, r.v[1]));
        }
        C.copy(M_ALL, res, r_prv);
        if (itinn < numinn && rms_upd >= eps * rms_xk) {
            C.axpy(M_INT, -alpha, q, res);
        } else {
            AnyItem it;
            it.ncg = ncg_item(P);
            r = co_await C.opr(OP_NCG_NEGPN, it);
            bool lchg_negpn = r.v[0] > 0;
            ncon = (long)r.v[1];
            double fk = dxdy * r.v[2];
            if (ncon <= 0) {
                NcgItem I = ncg_item(P);
                I.mode = 2;      // h <= hmin + 1e-5, pressure set to zero
                I.hmin = P.hmin + 1e-5;
                AnyItem ih;
                ih.ncg = I;
                r = co_await C.opr(OP_NCG_HMIN, ih);
                ncon += (long)r.v[0];
                lchg_negpn = true;
                Res s2 = co_await C.sum(M_ALL, ps);
                fk = dxdy * s2.v[0];
            }
            if (norm == 1 && lchg_negpn) {
                if (std::fabs(fk) < sp(1e-3f) * fntrue) {
                    C.set(M_INT, ps, 1.0);
                    r = co_await C.sum(M_ALL, ps);
                    fk = r.v[0];
                }
                C.scal(M_INT, ps, fntrue / fk);
            }
            C.conv_n(0, M_ALL, dd, ps);
            AnyItem idd;
            idd.ncg = ncg_item(P);
            r = co_await C.opr(OP_NCG_DD, idd);
            double avg = 0.0;
            if (norm == 1) {
                davg = r.v[0] / (double)ncon;
                avg = r.v[0] / std::max(1.0, r.v[1]);
            }
            AnyItem iex;
            iex.ncg = ncg_item(P);
            iex.ncg.avg = avg;
            iex.ncg.davg = davg;
            r = co_await C.opr(OP_NCG_EXPAND, iex);
            bool lchg_intpen = r.v[0] > 0;
            ncon += (long)r.v[0];
            itinn = 0;
            lchanged = lchg_intpen || lchg_negpn;
        }
    }
    if (norm == 1) *pen = davg;
    double conv = 1.0;
    if (rms_upd * rms_upd1 > 0.0 && itcg > 1) conv = std::exp(std::log(rms_upd / rms_upd1) / (itcg - 1));
    if (rms_upd > rms_xk && conv > 1.0 && itcg >= maxcg) ret = ERR_NORM;
    *itcg_out = itcg;
    *err_out = rms_upd;
    co_return ret;
}

// ------------------------------------------------------------------------------------------
// sensitivities (src/m_solvpn.f90:531-650, src/m_solvpt.f90:497-796)
// ------------------------------------------------------------------------------------------

enum { IOUT_FN = 0, IOUT_FX = 1, IOUT_FY = 2, IOUT_MZ = 3 };
enum { IIN_PEN = 0, IIN_KSI = 1, IIN_ETA = 2, IIN_PHI = 3 };
static double& sens_at(Patch& P, int iout, int iin) { return P.sens[iout + 4 * iin]; }

// copy of the element division and of up to 6 grid functions of a patch, as one launch
static void snap(Ctx& C, const int8_t* el_src, int8_t* el_dst, std::initializer_list<double*> src,
                 std::initializer_list<double*> dst)
{
    AnyItem it;
    SensItem& q = it.sens;
    q.n = C.P->n;
    q.mx = C.P->mx;
    q.mode = 0;
    q.el_src = el_src;
    q.el_dst = el_dst;
    int k = 0;
    for (double* a : src) q.src[k++] = a;
    q.nvec = k;
    k = 0;
    for (double* a : dst) q.dst[k++] = a;
    C.op(OP_SENS, it);
}

// sens_norm, src/m_solvpn.f90:531-650: dFn/dpen by one NormCG (N=0, maxit mxsens, eps
// epsens) at pen + 0.0067 max(un), started from the converged pressures and element division,
// which are restored afterwards.  unn holds A pn of the converged state.
static Task sens_norm(Ctx C, double* unn)
{
    Patch& P = *C.P;
    Res r = co_await C.vmax(M_INT, unn);
    double pentru = std::max(r.v[0], -1e20);        // gf3_max starts at -1e20
    double dpen = 0.0067 * pentru;
    double penloc = P.pen + dpen;
    int8_t* el0 = (int8_t*)P.w[20];
    double* ps0 = P.w[15];
    snap(C, P.el, el0, {P.pn}, {ps0});
    int it = 0;
    double err = 0.0;
    const bool dense0 = P.dense;
    P.dense = dense0 || (dense_mode() >= 1 && P.n <= DENSE_MAX_N);
    int rc = co_await normcg(C, 0, P.in.mxsens, P.in.epsens, &penloc, &it, &err, ps0);
    P.dense = dense0;
    r = co_await C.sum(M_ALL, P.pn);
    double fnloc = P.dxdy * r.v[0];
    sens_at(P, IOUT_FN, IIN_PEN) = (fnloc - P.fntrue) / dpen;
    snap(C, el0, P.el, {ps0}, {P.pn});
    co_return rc;
}

// snorm, src/m_snorm.f90:31-378
static Task snorm(Ctx C, int* itnorm_out)
{
    Patch& P = *C.P;
    const int npot = P.n;
    double* ps = P.pn;
    double* unn = P.w[7];
    bool use_vecfft = true;   // the product of the whole area serves the expansion test
    co_await setup_prec(C, 2);
    if (P.nt_cpl) {
        // K != 0: the normal displacement of the current tangential tractions joins the gap,
        // hn = h + A_nt p_t on every element (src/m_snorm.f90:112-119)
        int ik[1] = {2}, jk[2] = {0, 1};
        double* o[1] = {P.hn};
        double* in[2] = {P.px, P.py};
        const double* ad[1] = {P.h};
        C.conv(0, M_ALL, 1, ik, o, 2, jk, in, ad);
    }
    int itnorm = 0, itcg = 0, ret = 0;
    bool zready;
    Res r;
    do {
        itnorm++;
        zready = true;
        C.set(M_EXT, ps, 0.0);
        int it = 0;
        double errpn = 0.0;
        int rc = co_await normcg(C, P.in.norm, P.in.maxgs, P.in.eps, &P.pen, &it, &errpn);
        itcg += it;
        if (rc) { ret = rc; break; }
        AnyItem ic;
        ic.ncg = ncg_item(P);
        ic.ncg.errpn = errpn;
        r = co_await C.opr(OP_SN_CONTRACT, ic);
        if (r.v[0] > 0) zready = false;
        if (zready) {
            if (use_vecfft) C.conv_n(0, M_ALL, unn, ps);
            r = co_await aij_ones(C, 8);
            double tol = std::fabs(errpn * r.v[0] * (1.0 / P.ga));
            AnyItem ie;
            ie.ncg = ncg_item(P);
            ie.ncg.pen = P.pen;
            ie.ncg.tol = tol;
            r = co_await C.opr(OP_SN_EXPAND, ie);
            if (r.v[0] != 0) zready = false;
        }
        if (it >= P.in.maxgs) zready = false;
    } while (!zready && itnorm < P.in.maxin);
    if (!zready && !ret) itnorm = -1;
    AnyItem ifin;
    ifin.ncg = ncg_item(P);
    co_await C.opr(OP_SN_FINAL, ifin);
    P.out.itcg = itcg;
    if (P.in.norm == 0) {
        r = co_await C.sum(M_ALL, ps);
        P.fntrue = P.dxdy * r.v[0];
    }
    // sensitivity dFn/dpen (src/m_snorm.f90:357-368); below 700 elements the reference forms
    // unn after the final clean-up of the pressures
    if (!ret && P.in.sens >= 2) {
        if (npot < 700) C.conv_n(0, M_ALL, unn, ps);
        ret = co_await sens_norm(C, unn);
    } else {
        sens_at(P, IOUT_FN, IIN_PEN) = 0.0;
    }
    *itnorm_out = itnorm;
    co_return ret;
}

// ------------------------------------------------------------------------------------------
// GDsteady, src/gdsteady.f90:9-606
// ------------------------------------------------------------------------------------------

struct GdState {
    int gd_meth, kdown;
};

static GdItem gd_item(Patch& P)
{
    GdItem I{};
    I.n = P.n; I.mx = P.mx; I.my = P.my;
    I.el = P.el; I.px = P.px; I.py = P.py; I.pn = P.pn;
    I.dpx = P.w[0]; I.dpy = P.w[1];
    I.sx = P.sx; I.sy = P.sy;
    I.wsx = P.wsx; I.wsy = P.wsy;
    I.nx = P.w[2]; I.ny = P.w[3];
    I.rx = P.w[4]; I.ry = P.w[5];
    I.dvx = P.w[6]; I.dvy = P.w[7];
    I.vx = P.w[8]; I.vy = P.w[9];
    I.qx = P.w[10]; I.qy = P.w[11];
    I.dscl = P.w[12];
    I.poldx = P.w[13]; I.poldy = P.w[14];
    I.mu = P.in.fstat;
    I.fac_v = 1000.0;
    I.fdecay = P.in.fdecay;
    I.d_ifc = P.in.d_ifc; I.d_lin = P.in.d_lin; I.d_cns = P.in.d_cns; I.d_slp = P.in.d_slp; I.pow_s = P.in.pow_s;
    return I;
}

// cf00: the tangential coefficients of element (0,0), unscaled (cs blocks xx, xy, yx, yy)
static Task gdsteady(Ctx C, const double cf00[4], int maxgd, double eps, int* itgd_out,
                     double* err_out, int* lstagn)
{
    Patch& P = *C.P;
    const double tiny = 1e-12;
    const int npot = P.n;
    const bool fp32 = P.in.precision == 1;
    GdItem G = gd_item(P);
    G.c11 = cf00[0] / P.ga; G.c12 = cf00[1] / P.ga; G.c21 = cf00[2] / P.ga; G.c22 = cf00[3] / P.ga;
    // gd_meth / kdown (src/contact_addon.f90:1240-1247)
    int gd_meth = 1, kdown = 3;
    if (P.in.fdecay >= 0.999) gd_meth = 1;
    else if (P.in.fdecay <= 0.001) { gd_meth = 2; kdown = std::max(1, (int)std::lround(-P.in.fdecay)); }
    else gd_meth = 3;
    {
        // the row tiles of the GD kernels must fit the LDS of one work-group
        AnyItem t;
        t.gd = G;
        if (op_lds(OP_GD_PROJECT, &t.gd, 0) > MAX_LDS || op_lds(OP_GD_PROBE, &t.gd, 0) > MAX_LDS)
            co_return ERR_OTHER;
    }
    Res r;
    r = co_await C.count(M_INT);
    double ncon = r.v[0];
    double facnel = (double)std::sqrt((float)npot / (float)ncon);
    C.set(M_ALL, G.dscl, 1.0);
    C.set(M_ALL, G.poldx, 0.0, G.poldy);
    C.set(M_ALL, G.vx, 0.0, G.vy);
    C.set(M_ALL, G.dvx, 0.0, G.dvy);
    C.set(M_ALL, G.qx, 0.0, G.qy);
    int itgd = 0, it_fb = -99;
    bool lchanged = false;
    double dif = 2.0, difid = 1.0, dif1 = 0.0, beta = 1.0, alpha = 0.0, alpha0 = 0.0;
    *lstagn = 0;
    AnyItem it;
    it.gd = G;
    C.op(OP_TCG_NT, [&] { AnyItem t; TcgItem& T = t.tcg; T.n = npot; T.px = P.px; T.py = P.py; T.nx = G.nx; T.ny = G.ny; return t; }());
    it.gd.n = P.n;
    C.op(OP_GD_DP, it);
    C.set(M_ALL, P.sx, 0.0, P.sy);
    C.conv_t(0, M_INT, P.sx, P.sy, G.dpx, G.dpy, G.wsx, G.wsy);
    AnyItem id = it;
    id.gd.first = 1;
    co_await C.opr(OP_GD_DIAGRES, id);
    r = co_await C.maxabs(M_ALL, G.rx, G.ry);
    if (r.v[0] < tiny) dif = 0.0;

    double resrms2 = 0.0;
    while ((lchanged || dif > difid) && itgd < maxgd) {
        itgd++;
        // project_searchdir
        AnyItem ip = it;
        if (beta > P.in.betath || itgd - it_fb <= 1) { ip.gd.imeth = gd_meth; ip.gd.kdown = kdown; }
        else { ip.gd.imeth = 2; ip.gd.kdown = P.in.kdowfb; it_fb = itgd; }
        C.op(OP_GD_PROJECT, ip);
        C.conv_t(0, M_INT, G.qx, G.qy, G.dvx, G.dvy, nullptr, nullptr, fp32);

        // perform_linesearch, src/gdsteady.f90:1046-1462
        const int MAX_J = 99;
        const double eps_j = 0.001;
        double tbl_alpha[MAX_J + 2], tbl_rho[MAX_J + 2], tbl_drhoda[MAX_J + 2];
        int j = 0, itb_a = 0, itb_b = 0, itb_x = 0, itb_w = 0, itb_v = 0, itb_u;
        bool stop_j = false, has_bracket = false, prev_bracket;
        double alpha_j = 0.0, alpha_prv = -1.0, dst_cur = 0.0, dst_prv1 = 0.0, dst_prv2 = 0.0, ddrho;
        double sum_dq2 = 0.0, sum_r2 = 0.0;
        // After the first iteration the first two trial steps are known before any probe
        // result: 0, then 0.6 alpha0 with alpha0 = (alpha0 + alpha) / 2 (set at j = 0 below,
        // taken at j = 1).  Both probes go into one flush; probes only reduce, so the kernels
        // and their order are those of one probe per flush.
        double spec[2][sizeof(Res::v) / sizeof(double)];
        double spec_alpha = 0.0;
        int nspec = 0;
        if (itgd > 1) {
            spec_alpha = 0.0 + 0.6 * (0.5 * (alpha0 + alpha));
            AnyItem p0 = it;
            p0.gd.alpha = 0.0;
            p0.gd.first = 1;
            C.opr_into(OP_GD_PROBE, p0, spec[0]);
            AnyItem p1 = it;
            p1.gd.alpha = spec_alpha;
            p1.gd.first = 0;
            Res rs = co_await C.opr(OP_GD_PROBE, p1);
            std::memcpy(spec[1], rs.v, sizeof(rs.v));
            nspec = 2;
        }
        while (!stop_j) {
            AnyItem pr = it;
            pr.gd.alpha = alpha_j;
            pr.gd.first = j == 0 ? 1 : 0;
            if (j < nspec && alpha_j == (j == 0 ? 0.0 : spec_alpha)) std::memcpy(r.v, spec[j], sizeof(r.v));
            else r = co_await C.opr(OP_GD_PROBE, pr);
            double rhonew = r.v[0];
            double drho_da = r.v[1] + r.v[3] + alpha_j * (r.v[2] + r.v[4]) + alpha_j * alpha_j * r.v[5];
            if (j == 0) {
                if (itgd <= 1) alpha0 = r.v[6] / r.v[7];
                else alpha0 = 0.5 * (alpha0 + alpha);
                sum_dq2 = r.v[8];
                sum_r2 = r.v[9];
            }
            int itb = 1;
            while (itb <= j && alpha_j > tbl_alpha[itb]) itb++;
            itb_u = itb;
            for (int k = j; k >= itb; k--) {
                tbl_alpha[k + 1] = tbl_alpha[k];
                tbl_rho[k + 1] = tbl_rho[k];
                tbl_drhoda[k + 1] = tbl_drhoda[k];
            }
            if (has_bracket) {
                if (itb_a >= itb_u) itb_a++;
                if (itb_b >= itb_u) itb_b++;
                if (itb_x >= itb_u) itb_x++;
                if (itb_w >= itb_u) itb_w++;
                if (itb_v >= itb_u) itb_v++;
            }
            j++;
            tbl_alpha[itb] = alpha_j;
            tbl_rho[itb] = rhonew;
            tbl_drhoda[itb] = drho_da;
            prev_bracket = has_bracket;
            has_bracket = false;
            int ibrack = 1;
            while (ibrack < j - 1 && !has_bracket) {
                has_bracket = tbl_rho[ibrack + 1] <= std::min(tbl_rho[ibrack], tbl_rho[ibrack + 2]);
                if (!has_bracket) ibrack++;
            }
            if (has_bracket && !prev_bracket) {
                itb_a = ibrack;
                itb_x = ibrack + 1;
                itb_b = ibrack + 2;
                if (tbl_rho[itb_a] < tbl_rho[itb_b]) { itb_w = itb_a; itb_v = itb_b; }
                else { itb_w = itb_b; itb_v = itb_a; }
                dst_prv1 = tbl_alpha[itb_x] - tbl_alpha[itb_w];
                dst_prv2 = tbl_alpha[itb_w] - tbl_alpha[itb_v];
            } else if (has_bracket) {
                if (tbl_rho[itb_u] <= tbl_rho[itb_x]) {
                    if (tbl_alpha[itb_u] >= tbl_alpha[itb_x]) itb_a = itb_x;
                    else itb_b = itb_x;
                    itb_v = itb_w;
                    itb_w = itb_x;
                    itb_x = itb_u;
                } else {
                    if (tbl_alpha[itb_u] < tbl_alpha[itb_x]) itb_a = itb_u;
                    else itb_b = itb_u;
                    if (tbl_rho[itb_u] <= tbl_rho[itb_w] || tbl_alpha[itb_w] == tbl_alpha[itb_x]) {
                        itb_v = itb_w;
                        itb_w = itb_u;
                    } else if (tbl_rho[itb_u] <= tbl_rho[itb_v] || tbl_alpha[itb_v] == tbl_alpha[itb_x] ||
                               tbl_alpha[itb_v] == tbl_alpha[itb_w]) {
                        itb_v = itb_u;
                    }
                }
                dst_prv2 = dst_prv1;
                dst_prv1 = dst_cur;
            }
            alpha_prv = alpha_j;
            if (has_bracket) {
                double xmid = 0.5 * (tbl_alpha[itb_a] + tbl_alpha[itb_b]);
                double tolx = eps_j * std::fabs(tbl_alpha[itb_x]) + tiny;
                bool ldone = std::fabs(tbl_alpha[itb_x] - xmid) <= 2.0 * tolx - 0.5 * (tbl_alpha[itb_b] - tbl_alpha[itb_a]);
                if (!ldone) {
                    bool use_parab;
                    if (std::fabs(dst_prv2) <= tolx) {
                        use_parab = false;
                    } else {
                        double br = (tbl_alpha[itb_x] - tbl_alpha[itb_w]) * (tbl_rho[itb_x] - tbl_rho[itb_v]);
                        double bq = (tbl_alpha[itb_x] - tbl_alpha[itb_v]) * (tbl_rho[itb_x] - tbl_rho[itb_w]);
                        double bp = (tbl_alpha[itb_x] - tbl_alpha[itb_v]) * bq - (tbl_alpha[itb_x] - tbl_alpha[itb_w]) * br;
                        bq = 2.0 * (bq - br);
                        if (bq > 0.0) bp = -bp;
                        bq = std::fabs(bq);
                        if (std::fabs(bp) >= std::fabs(0.5 * bq * dst_prv2) || bp <= bq * (tbl_alpha[itb_a] - tbl_alpha[itb_x]) ||
                            bp >= bq * (tbl_alpha[itb_b] - tbl_alpha[itb_x])) {
                            use_parab = false;
                        } else {
                            use_parab = true;
                            dst_cur = bp / bq;
                            alpha_j = tbl_alpha[itb_x] + dst_cur;
                            if (alpha_j - tbl_alpha[itb_a] < 2.0 * tolx || tbl_alpha[itb_b] - alpha_j < 2.0 * tolx)
                                dst_cur = tolx * (xmid - tbl_alpha[itb_x] >= 0 ? 1.0 : -1.0);
                        }
                    }
                    if (!use_parab) {
                        if (tbl_alpha[itb_x] >= xmid) dst_prv1 = tbl_alpha[itb_a] - tbl_alpha[itb_x];
                        else dst_prv1 = tbl_alpha[itb_b] - tbl_alpha[itb_x];
                        dst_cur = 0.381966 * dst_prv1;
                    }
                    alpha_j = tbl_alpha[itb_x] + dst_cur;
                }
            } else if (j == 1) {
                alpha_j = alpha_j + 0.6 * alpha0;
                ddrho = 10.0;
            } else if (tbl_rho[j] < tbl_rho[1]) {
                double dalph = tbl_alpha[j] - tbl_alpha[j - 1];
                if (j < 3) {
                    ddrho = (tbl_drhoda[j] - tbl_drhoda[j - 1]) / dalph;
                    alpha_j = tbl_alpha[j] + std::min(3.0 * dalph, -tbl_drhoda[j - 1] / ddrho);
                } else {
                    alpha_j = tbl_alpha[j] + 3.0 * dalph;
                }
            } else {
                double dalph = tbl_alpha[2] - tbl_alpha[1];
                if (j < 3) {
                    ddrho = (tbl_drhoda[2] - tbl_drhoda[1]) / dalph;
                    alpha_j = tbl_alpha[1] + std::min(-3.0 * dalph, tbl_drhoda[1] / ddrho);
                } else {
                    alpha_j = tbl_alpha[1] - 3.0 * dalph;
                }
            }
            stop_j = (j >= MAX_J || std::fabs(alpha_j - alpha_prv) < eps_j * std::max(std::fabs(alpha_j), std::fabs(alpha_prv)));
        }
        (void)ddrho;
        if (!has_bracket) {   // the reference aborts the run here
            *itgd_out = itgd;
            *err_out = dif;
            co_return ERR_TANG;
        }
        if (j >= MAX_J) {
            int best = 1;
            for (int k = 2; k <= MAX_J; k++)
                if (tbl_rho[k] < tbl_rho[best]) best = k;
            alpha_j = tbl_alpha[best];
        }
        alpha = alpha_j;
        beta = alpha * std::sqrt(sum_dq2) / std::sqrt(sum_r2);
        // update p and the slip
        AnyItem iu = it;
        iu.gd.alpha = alpha;
        C.op(OP_GD_UPDATE, iu);
        C.set(M_ALL, P.sx, 0.0, P.sy);
        C.conv_t(0, M_INT, P.sx, P.sy, G.dpx, G.dpy, G.wsx, G.wsy);
        // the moves count and the residual terms in one flush (nothing between them reads the
        // count)
        double moves[sizeof(Res::v) / sizeof(double)];
        C.opr_into(OP_GD_MOVES, it, moves);
        AnyItem idr = it;
        idr.gd.first = 0;
        r = co_await C.opr(OP_GD_DIAGRES, idr);
        lchanged = moves[0] > 0;
        resrms2 = r.v[2];
        double facdif;
        if (beta >= 0.1) facdif = 1.0;
        else if (beta >= 0.001) facdif = 0.1 / beta;
        else facdif = 100.0;
        dif = facnel * std::sqrt(r.v[0] / (2.0 * npot)) * facdif;
        difid = eps * std::max(1e-6, facnel * std::sqrt(r.v[1] / (2.0 * npot)));
        if (itgd == 1) dif1 = dif;
    }
    {
        Res c2 = co_await C.count(M_INT);
        double resrms = std::sqrt(resrms2 / std::max(1.0, 2.0 * c2.v[0]));
        double res_dp = resrms / std::max(G.c11, G.c22);
        double conv = 1.0;
        if (dif * dif1 > 0.0 && itgd > 1) conv = std::exp(std::log(dif / dif1) / (itgd - 1));
        if (lchanged && itgd >= maxgd) *lstagn = 1;
        else if (dif > difid && conv > 1.0 && itgd >= maxgd) *lstagn = 1;
        else if (itgd > 0 && conv < 1.0 - tiny && res_dp > 5.0 * difid / (1.0 - conv)) { itgd = -itgd; *lstagn = 1; }
    }
    *itgd_out = itgd;
    *err_out = dif;
    co_return 0;
}

// ------------------------------------------------------------------------------------------
// TangCG, src/m_solvpt.f90:1841-2442
// ------------------------------------------------------------------------------------------

// cf00: the tangential coefficients of element (0,0), unscaled (cs blocks xx, xy, yx, yy)
// p0x, p0y: initial estimate of a sensitivity solve (is_sens, src/m_solvpt.f90:1910-1914,
// 2209-2216); null otherwise
static Task tangcg(Ctx C, const double cf00[4], int maxcg, double eps, int* itcg_out, double* err_out,
                   const double* p0x = nullptr, const double* p0y = nullptr)
{
    Patch& P = *C.P;
    const int num_inn = 4;
    const double small = 1e-6;
    const int npot = P.n;
    const bool fp32 = P.in.precision == 1;
    const double c11 = cf00[0], c22 = cf00[3];
    co_await setup_prec(C, 0);
    co_await setup_prec(C, 1);
    TcgItem T{};
    T.n = npot; T.el = P.el; T.px = P.px; T.py = P.py; T.pn = P.pn; T.sx = P.sx; T.sy = P.sy;
    T.nx = P.w[0]; T.ny = P.w[1]; T.rx = P.w[2]; T.ry = P.w[3]; T.zx = P.w[4]; T.zy = P.w[5];
    T.vx = P.w[6]; T.vy = P.w[7]; T.qx = P.w[8]; T.qy = P.w[9]; T.poldx = P.w[10]; T.poldy = P.w[11];
    T.mu = P.in.fstat; T.ga = P.ga; T.c11 = c11; T.c22 = c22;
    T.p0x = p0x; T.p0y = p0y;
    T.sc = P.sc;
    AnyItem t;
    t.tcg = T;
    Res r;
    r = co_await C.count(M_ADH);
    double nadh = r.v[0];
    r = co_await C.count(M_INT);
    double ncon = r.v[0];
    double nslip = ncon - nadh;
    double facnel = (double)std::sqrt((float)npot / (float)ncon);
    bool use_fftprec = true;
    int itcg = 0, it_inn = 0;
    bool lchanged = false;
    double dif = 2.0, difid = 1.0, difinn = 0.0, trsinn = 0.0, alpha = 0.0;
    C.set(M_ALL, T.vx, 0.0, T.vy);
    C.set(M_ALL, T.qx, 0.0, T.qy);
    C.set(M_ALL, T.zx, 0.0, T.zy);
    C.set(M_ALL, T.poldx, 0.0, T.poldy);
    C.op(OP_TCG_NT, t);
    C.set(M_ALL, P.sx, 0.0, P.sy);
    C.conv_t(0, M_INT, P.sx, P.sy, P.px, P.py, P.wsx, P.wsy);
    C.op(OP_TCG_RSET, t);
    while ((lchanged || dif > difid) && itcg < maxcg) {
        itcg++;
        it_inn++;
        if (use_fftprec) {
            if (nslip > 3 * nadh && P.my > 1) use_fftprec = false;
            if (itcg >= maxcg / 2) use_fftprec = false;
            if (!use_fftprec) lchanged = true;
        }
        if (use_fftprec) {
            int ik[2] = {0, 1}, jk[2] = {0, 1};
            double* o[2] = {T.zx, T.zy};
            double* i2[2] = {T.rx, T.ry};
            C.conv(2, M_INT, 2, ik, o, 2, jk, i2, nullptr, fp32);
        } else {
            C.copy(M_ALL, T.rx, T.zx, T.ry, T.zy);
        }
        // zprep leaves (z.q, v.q) in the patch's device scalars, from which vupd forms beta
        // (src/m_solvpt.f90:2150-2160); qpost leaves (r.v, v.q), from which pupd forms alpha
        // (:2190-2200).  The iteration's only host round trip is pupd's result, which also
        // returns (r.v, v.q) so that the host forms the same alpha.
        C.op(OP_TCG_ZPREP, t);
        if (itcg <= 1 || lchanged) C.copy(M_ALL, T.zx, T.vx, T.zy, T.vy);
        else C.op(OP_TCG_VUPD, t);
        C.conv_t(0, M_INT, T.qx, T.qy, T.vx, T.vy, nullptr, nullptr, fp32);
        C.op(OP_TCG_QPOST, t);
        r = co_await C.opr(OP_TCG_PUPD, t);
        const double rv = r.v[2], vq = r.v[3];
        if (std::fabs(rv) < 1e-60) alpha = 0.0;
        else if (std::fabs(vq) < small * std::fabs(rv)) alpha = 1.0;
        else alpha = rv / vq;
        dif = alpha * facnel * std::sqrt(r.v[0] / (2.0 * npot));
        difinn += dif;
        if (itcg <= 5 || itcg % 5 == 1) {
            double ptang = facnel * std::sqrt(r.v[1] / (2.0 * npot));
            difid = eps * std::max(1e-6, ptang);
            trsinn = sp(0.01f) * ptang;
        }
        if (it_inn >= num_inn || dif <= difid || difinn > trsinn) {
            lchanged = false;
            r = co_await C.opr(OP_TCG_AS, t);
            if (r.v[0] > 0) lchanged = true;
            C.conv_t(0, M_INT, P.sx, P.sy, P.px, P.py, P.wsx, P.wsy);
            r = co_await C.opr(OP_TCG_SA, t);
            if (r.v[0] > 0) lchanged = true;
            dif = facnel * std::sqrt(r.v[1] / (2.0 * npot));
            it_inn = 0;
            difinn = 0.0;
            if (lchanged) {
                Res a = co_await C.count(M_ADH);
                Res b = co_await C.count(M_INT);
                nadh = a.v[0];
                nslip = b.v[0] - a.v[0];
            }
        }
        if ((lchanged || dif > difid) && itcg < maxcg) {
            if (it_inn <= 0) {
                C.op(OP_TCG_NT, t);
                C.op(OP_TCG_RSET, t);
            } else {
                C.axpy(M_INT, -alpha, T.qx, T.rx, T.qy, T.ry);
            }
        }
    }
    *itcg_out = itcg;
    *err_out = dif;
    co_return 0;
}

// ------------------------------------------------------------------------------------------
// stang, src/m_stang.f90:28-951 (Coulomb friction, creepages prescribed)
// ------------------------------------------------------------------------------------------

static GsItem gs_item(Patch& P, int k, double omegah, double omegas, int maxgs, double eps)
{
    GsItem g{};
    g.n = P.n; g.mx = P.mx; g.my = P.my; g.k = k;
    g.maxgs = maxgs;
    g.eps = eps; g.omegah = omegah; g.omegas = omegas;
    g.ga_inv = 1.0 / P.ga;
    g.mu = P.in.fstat;
    g.facnel = (double)std::sqrt((float)P.n / (float)std::max(1, k));
    g.el = P.el;
    g.px = P.px; g.py = P.py; g.pn = P.pn;
    g.sx = P.sx; g.sy = P.sy;
    g.wsx = P.wsx; g.wsy = P.wsy;
    g.dpx = P.w[0]; g.dpy = P.w[1];
    g.list = (int*)P.w[2];
    size_t blk = (size_t)(2 * P.mx) * (2 * P.my);
    g.cf[0][0] = P.cf;
    g.cf[0][1] = P.cf + blk;
    g.cf[1][0] = P.cf + 3 * blk;
    g.cf[1][1] = P.cf + 4 * blk;
    // the block variant: z = A dp and the block change d, and the item of their product
    g.zx = P.w[3]; g.zy = P.w[4];
    g.ddx = P.w[5]; g.ddy = P.w[6];
    const int ik[2] = {0, 1}, jk[2] = {0, 1};
    double* out[2] = {g.zx, g.zy};
    double* in[2] = {g.ddx, g.ddy};
    const double* add[2] = {g.zx, g.zy};
    g.conv = conv_item(&P, 0, M_ALL, 2, ik, out, 2, jk, in, add, false);
    return g;
}

// SteadyGS, src/m_solvpt.f90:2825-3254: all sweeps in one launch per patch.  p0x, p0y: initial
// estimate of a sensitivity solve (is_sens, :2895-2899, 3138-3144); null otherwise.
// info: 0, or 1 when maxgs was reached, 2 when also the mean rate of convergence exceeds 0.997,
// 3 when it exceeds 1 (:3205-3234).
static Task stdygs(Ctx C, int k, double omegah, double omegas, int maxgs, double eps, int* itgs_out,
                   double* err_out, int* info, const double* p0x = nullptr, const double* p0y = nullptr)
{
    Patch& P = *C.P;
    AnyItem it;
    it.gs = gs_item(P, k, omegah, omegas, maxgs, eps);
    it.gs.p0x = p0x;
    it.gs.p0y = p0y;
    if (op_lds(OP_STDYGS, &it.gs, 0) + 2048 > MAX_LDS && op_lds(OP_STDYGS, &it.gs, 1) + 2048 > MAX_LDS)
        co_return ERR_OTHER;
    Res r = co_await C.opr(OP_STDYGS, it);
    int itgs = (int)r.v[0];
    double dif = r.v[1], dif1 = r.v[5];
    double conv = 1.0;
    if (dif * dif1 != 0.0 && itgs > 1) conv = std::exp(std::log(dif / dif1) / (itgs - 1));
    *info = 0;
    if (itgs >= maxgs) *info = 1;
    if (itgs >= maxgs && conv > sp(0.997f)) *info = 2;
    if (itgs >= maxgs && conv > 1.0) *info = 3;
    *itgs_out = itgs;
    *err_out = dif;
    P.out.used_stdygs = 1;
    co_return 0;
}

// ConvexGS, src/m_solvpt.f90:2446-2821: all sweeps in one launch per patch (on the main
// stream: the solver of G=2 and of steady rolling with contact in the trailing-edge column).
// p0x, p0y: initial estimate of a sensitivity solve; info as for SteadyGS.
static Task cnvxgs(Ctx C, int k, double omegah, double omegas, int maxgs, double eps, int* itgs_out,
                   double* err_out, int* info, const double* p0x = nullptr, const double* p0y = nullptr)
{
    Patch& P = *C.P;
    AnyItem it;
    CgsItem& g = it.cgs;
    g.n = P.n; g.mx = P.mx; g.my = P.my; g.k = k;
    g.maxgs = maxgs; g.steady = P.in.tang == 3 ? 1 : 0;
    g.eps = eps; g.omegah = omegah; g.omegas = omegas;
    g.ga_inv = 1.0 / P.ga;
    g.mu = P.in.fstat;
    g.facnel = (double)std::sqrt((float)P.n / (float)std::max(1, k));
    g.el = P.el;
    g.px = P.px; g.py = P.py; g.pn = P.pn;
    g.sx = P.sx; g.sy = P.sy;
    g.wsx = P.wsx; g.wsy = P.wsy;
    g.lix = P.lix;
    g.ubx = P.w[3]; g.uby = P.w[4];
    g.list = (int*)P.w[2];
    const size_t blk = (size_t)(2 * P.mx) * (2 * P.my);
    const size_t off[2][2] = {{0, blk}, {3 * blk, 4 * blk}};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            g.cf[a][b] = P.cf + off[a][b];
            g.cv[a][b] = P.cfv + off[a][b];
        }
    g.p0x = p0x; g.p0y = p0y;
    Res r = co_await C.opr(OP_CNVXGS, it);
    const int itgs = (int)r.v[0];
    const double dif = r.v[1], dif1 = r.v[5];
    double conv = 1.0;
    if (dif * dif1 != 0.0 && itgs > 1) conv = std::exp(std::log(dif / dif1) / (itgs - 1));
    *info = 0;
    if (itgs >= maxgs) *info = 1;
    if (itgs >= maxgs && conv > sp(0.997f)) *info = 2;
    if (itgs >= maxgs && conv > 1.0) *info = 3;
    *itgs_out = itgs;
    *err_out = dif;
    co_return 0;
}

// tang_solver, src/m_solvpt.f90:382-493: GDsteady for T=3 with G=5, restarted by SteadyGS when it
// stagnates (:466-485); SteadyGS for other G in steady rolling; TangCG for T=1, 2.  *it counts
// the iterations of both solvers of a fall-back; *info is set by SteadyGS only.  p0x, p0y: the
// initial estimate of a sensitivity solve (is_sens), null otherwise.
static Task tang_solver(Ctx C, int k, const double cf00[4], double omegah, double omegas, int maxit,
                        double eps, int* it_out, double* err_out, int* info,
                        const double* p0x = nullptr, const double* p0y = nullptr)
{
    Patch& P = *C.P;
    int it = 0, rc;
    if (P.cnvx) {
        rc = co_await cnvxgs(C, k, omegah, omegas, maxit, eps, &it, err_out, info, p0x, p0y);
        *it_out = std::abs(it);
    } else if (P.in.tang == 3 && P.in.gausei == 5) {
        int lstagn = 0;
        rc = co_await gdsteady(C, cf00, maxit, eps, &it, err_out, &lstagn);
        P.out.gd_stagnated = lstagn;
        *it_out = std::abs(it);
        if (!rc && lstagn) {
            rc = co_await stdygs(C, k, omegah, omegas, maxit, eps, &it, err_out, info, p0x, p0y);
            *it_out += std::abs(it);
        }
    } else if (P.in.tang == 3) {
        rc = co_await stdygs(C, k, omegah, omegas, maxit, eps, &it, err_out, info, p0x, p0y);
        *it_out = std::abs(it);
    } else {
        rc = co_await tangcg(C, cf00, maxit, eps, &it, err_out, p0x, p0y);
        *it_out = std::abs(it);
    }
    co_return rc;
}

// k_sens mode 1 on the interior: ws = wb + d dq along x (which 0) or y (which 1), or the spin
// change -d y dq, +d x dq (which 2)
static void add_rigid_slip(Ctx& C, double* wbx, double* wby, double* wsx, double* wsy, int which, double d)
{
    Patch& P = *C.P;
    AnyItem wi;
    SensItem& q = wi.sens;
    q.n = P.n; q.mx = P.mx; q.mode = 1; q.which = which;
    q.el = P.el; q.wbx = wbx; q.wby = wby; q.wsx = wsx; q.wsy = wsy;
    q.xc1 = P.xc1; q.yc1 = P.yc1; q.dx = P.dx; q.dy = P.dy; q.dq = P.dq; q.d = d;
    q.facdt = P.ledge ? P.facdt : nullptr;
    C.op(OP_SENS, wi);
}

static SetupItem setup_item(Patch& P)
{
    SetupItem S{};
    S.n = P.n; S.mx = P.mx; S.my = P.my; S.ibase = P.ibase_eff;
    S.el = P.el; S.h = P.h; S.hx = P.hx; S.hy = P.hy;
    S.pn = P.pn; S.px = P.px; S.py = P.py; S.pvn = P.pvn; S.pvx = P.pvx; S.pvy = P.pvy;
    S.xc1 = P.xc1; S.yc1 = P.yc1; S.dx = P.dx; S.dy = P.dy;
    for (int k = 0; k < 6; k++) S.prm[k] = P.prm[k];
    S.facx = S.facy = 1.0;
    return S;
}

// Kalker's creep coefficients C11, C22, C23 (the table of src/m_hertz.f90:1158-1179, single-
// precision values there) at g = 0.1, ..., 1.0 and poisson ratio 0, 1/4, 1/2; case 0 holds
// g = a/b <= 1, case 1 holds g = b/a < 1.
static const float kalker_c[2][10][3][3] = {   // [case][g row][C11, C22, C23][nu]
    {{{2.51f, 3.31f, 4.85f}, {2.51f, 2.52f, 2.53f}, {.334f, .473f, .731f}},
     {{2.59f, 3.37f, 4.81f}, {2.59f, 2.63f, 2.66f}, {.483f, .603f, .809f}},
     {{2.68f, 3.44f, 4.80f}, {2.68f, 2.75f, 2.81f}, {.607f, .715f, .889f}},
     {{2.78f, 3.53f, 4.82f}, {2.78f, 2.88f, 2.98f}, {.720f, .823f, .977f}},
     {{2.88f, 3.62f, 4.83f}, {2.88f, 3.01f, 3.14f}, {.827f, .929f, 1.07f}},
     {{2.98f, 3.72f, 4.91f}, {2.98f, 3.14f, 3.31f}, {.930f, 1.03f, 1.18f}},
     {{3.09f, 3.81f, 4.97f}, {3.09f, 3.28f, 3.48f}, {1.03f, 1.14f, 1.29f}},
     {{3.19f, 3.91f, 5.05f}, {3.19f, 3.41f, 3.65f}, {1.13f, 1.25f, 1.40f}},
     {{3.29f, 4.01f, 5.12f}, {3.29f, 3.54f, 3.82f}, {1.23f, 1.36f, 1.51f}},
     {{3.40f, 4.12f, 5.20f}, {3.40f, 3.67f, 3.98f}, {1.33f, 1.47f, 1.63f}}},
    {{{10.7f, 11.7f, 12.9f}, {10.7f, 12.8f, 16.0f}, {12.2f, 14.6f, 18.0f}},
     {{6.96f, 7.78f, 8.82f}, {6.96f, 8.14f, 9.79f}, {5.72f, 6.63f, 7.89f}},
     {{5.57f, 6.34f, 7.34f}, {5.57f, 6.40f, 7.51f}, {3.79f, 4.32f, 5.01f}},
     {{4.84f, 5.57f, 6.57f}, {4.84f, 5.48f, 6.31f}, {2.88f, 3.24f, 3.70f}},
     {{4.37f, 5.10f, 6.11f}, {4.37f, 4.90f, 5.56f}, {2.35f, 2.62f, 2.96f}},
     {{4.06f, 4.78f, 5.80f}, {4.06f, 4.50f, 5.04f}, {2.01f, 2.23f, 2.50f}},
     {{3.82f, 4.54f, 5.58f}, {3.82f, 4.21f, 4.67f}, {1.76f, 1.95f, 2.18f}},
     {{3.65f, 4.36f, 5.42f}, {3.65f, 3.99f, 4.39f}, {1.58f, 1.75f, 1.94f}},
     {{3.51f, 4.22f, 5.30f}, {3.51f, 3.81f, 4.16f}, {1.44f, 1.59f, 1.77f}},
     {{3.40f, 4.12f, 5.20f}, {3.40f, 3.67f, 3.98f}, {1.33f, 1.47f, 1.63f}}}};

// linrol, src/m_hertz.f90:1147-1247: linear-theory creep coefficients at a/b and poisson ratio
// nu: asymptotic forms for slender ellipses (g < 0.101), otherwise the table, linear in g and
// quadratic (Lagrange) in nu for 1/C
static void linrol(double nu, double aob, double& cxx, double& cyy, double& cyz)
{
    int icase = aob <= 1.0 ? 0 : 1;
    double g = aob <= 1.0 ? aob : 1.0 / aob;
    if (icase == 0 && g < 0.101) {
        cxx = PI * PI / 4.0 / (1.0 - nu);
        cyy = PI * PI / 4.0;
        cyz = PI * std::sqrt(g) / 3.0 / (1.0 - nu) * (1.0 + nu * (std::log(16.0 / g) - 5.0));
    } else if (icase == 1 && g < 0.101) {
        double al = std::log(16.0 / (g * g));
        double l4 = std::log(4.0);
        cxx = 2.0 * PI / (al - 2.0 * nu) / g * (1.0 + (3.0 - l4) / (al - 2.0 * nu));
        cyy = 2.0 * PI / g * (1.0 + (1.0 - nu) * (3.0 - l4) / ((1.0 - nu) * al + 2.0 * nu)) /
              ((1.0 - nu) * al + 2.0 * nu);
        cyz = 2.0 * PI / 3.0 / std::pow(g, 1.5) / ((1.0 - nu) * al - 2.0 + 4.0 * nu);
    } else {
        int irow = (int)(10.0 * g - 0.005);
        double fac = 10.0 * g - (double)irow;
        double c[3][3], cc[3];
        for (int inu = 0; inu < 3; inu++)
            for (int ic = 0; ic < 3; ic++)
                c[inu][ic] = fac * (double)kalker_c[icase][irow][ic][inu] +
                             (1.0 - fac) * (double)kalker_c[icase][irow - 1][ic][inu];
        for (int ic = 0; ic < 3; ic++)
            cc[ic] = (nu - 0.25) * (nu - 0.50) * 8.0 / c[0][ic] - (nu - 0.50) * nu * 16.0 / c[1][ic] +
                     nu * (nu - 0.25) * 8.0 / c[2][ic];
        cxx = 1.0 / cc[0];
        cyy = 1.0 / cc[1];
        cyz = 1.0 / cc[2];
    }
}

// sens_tang, src/m_solvpt.f90:544-796: d(Fx, Fy)/(mu Fn) and dMz per creepage.  Each creepage
// is raised by 1e-4 of its characteristic size (char_creep, :497-540) and one tangential solve
// (tang_solver with is_sens: maxit mxsens, eps epsens, :382-493) runs from the converged
// state on the fixed right-hand side plus that change; the state is restored after each.
static Task sens_tang(Ctx C, int k, const double cf00[4], double omegah, double omegas)
{
    Patch& P = *C.P;
    AnyItem fi;
    fi.setup = setup_item(P);
    Res s = co_await C.opr(OP_FORCES, fi);
    double den = P.muscal * P.fntrue;
    double fxrel = P.dxdy * s.v[1] / den, fyrel = P.dxdy * s.v[2] / den;
    double mztrue = P.dxdy * (-s.v[3] + s.v[4]);
    // char_creep: effective radius from the contact area, half-axes from its extents
    Res cn = co_await C.count(M_INT);
    AnyItem ei;
    ei.sens.n = P.n;
    ei.sens.mx = P.mx;
    ei.sens.mode = 2;
    ei.sens.el = P.el;
    Res ex = co_await C.opr(OP_SENS_EXT, ei);
    double ncon = cn.v[0];
    double c = std::sqrt((double)(float)ncon * P.dx * P.dy / PI);
    int ixmin = (int)-ex.v[0], ixmax = (int)ex.v[1], iymin = (int)-ex.v[2], iymax = (int)ex.v[3];
    double a = (double)std::max(1, ixmax - ixmin + 1) * P.dx * 0.5;
    double b = (double)std::max(1, iymax - iymin + 1) * P.dy * 0.5;
    double c11, c22, c23;
    linrol(P.nu, a / b, c11, c22, c23);
    double ch[3];
    ch[0] = P.in.fstat * P.fntrue / (c * c * P.ga * c11);
    ch[1] = P.in.fstat * P.fntrue / (c * c * P.ga * c22);
    ch[2] = P.in.fstat * P.fntrue / (c * c * c * P.ga * c23);
    // saved state: element division, tractions and slip.  The right-hand side of the solves
    // is wstot = wsfix1 plus, when forces are prescribed, the creepages solvpt adds (:600-614).
    int8_t* el0 = (int8_t*)P.w[20];
    double *p0x = P.w[15], *p0y = P.w[16], *s0x = P.w[17], *s0y = P.w[18];
    double *wbx = P.w[19], *wby = P.w[21];
    const int F = P.in.force3;
    snap(C, P.el, el0, {P.px, P.py, P.sx, P.sy}, {p0x, p0y, s0x, s0y});
    if (F == 0) {
        snap(C, nullptr, nullptr, {P.wsx, P.wsy}, {wbx, wby});
    } else {
        add_rigid_slip(C, P.w[22], P.w[23], wbx, wby, 0, P.in.cksi);
        if (F >= 2) add_rigid_slip(C, wbx, wby, wbx, wby, 1, P.in.ceta);
    }
    int stagn0 = P.out.gd_stagnated, used0 = P.out.used_stdygs, ret = 0;
    const bool dense0 = P.dense;
    P.dense = dense0 || (dense_mode() >= 1 && P.n <= DENSE_MAX_N);
    for (int icrp = 0; icrp < 3 && !ret; icrp++) {
        double dcrp = 0.0001 * ch[icrp];
        add_rigid_slip(C, wbx, wby, P.wsx, P.wsy, icrp, dcrp);
        int it = 0, info = 0, rc;
        double err = 0.0;
        rc = co_await tang_solver(C, k, cf00, omegah, omegas, P.in.mxsens, P.in.epsens, &it, &err, &info,
                                  p0x, p0y);
        if (rc) {
            ret = rc;
        } else {
            Res f = co_await C.opr(OP_FORCES, fi);
            double fxloc = P.dxdy * f.v[1] / den, fyloc = P.dxdy * f.v[2] / den;
            double mzloc = P.dxdy * (-f.v[3] + f.v[4]);
            sens_at(P, IOUT_FX, IIN_KSI + icrp) = (fxloc - fxrel) / dcrp;
            sens_at(P, IOUT_FY, IIN_KSI + icrp) = (fyloc - fyrel) / dcrp;
            sens_at(P, IOUT_MZ, IIN_KSI + icrp) = (mzloc - mztrue) / dcrp;
        }
        snap(C, el0, P.el, {p0x, p0y, s0x, s0y}, {P.px, P.py, P.sx, P.sy});
    }
    P.dense = dense0;
    snap(C, nullptr, nullptr, {wbx, wby}, {P.wsx, P.wsy});
    P.out.gd_stagnated = stagn0;
    P.out.used_stdygs = used0;
    co_return ret;
}

// solvpt, src/m_solvpt.f90:51-378.  F=0: one tang_solver call on wsfix1.  F=1 (Fx prescribed)
// or F=2 (Fx and Fy): wstot = wsfix1 + facdt (cksi, ceta) dq, then Newton-Raphson on cksi (and
// ceta) with the sensitivities d(Fx, Fy)/(mu Fn) per creepage, refreshed from each half step.
// The creepages are the problem's kinematic state and keep their final values.
static Task solvpt(Ctx C, int k, const double cf00[4], double omegah, double omegas, int* itgs_out,
                   double* err_out)
{
    Patch& P = *C.P;
    const int F = P.in.force3;
    double* wfx = P.w[22];
    double* wfy = P.w[23];
    if (F >= 1) {
        add_rigid_slip(C, wfx, wfy, P.wsx, P.wsy, 0, P.in.cksi);
        if (F >= 2) add_rigid_slip(C, P.wsx, P.wsy, P.wsx, P.wsy, 1, P.in.ceta);
    }
    const double den = P.muscal * P.fntrue;
    AnyItem fi;
    fi.setup = setup_item(P);
    int itgs = 0, info1 = 0, it = 0, rc;
    double err = 0.0;
    rc = co_await tang_solver(C, k, cf00, omegah, omegas, P.in.maxgs, P.in.eps, &it, &err, &info1);
    itgs += it;
    if (!rc && F >= 1 && info1 <= 1) {
        Res f = co_await C.opr(OP_FORCES, fi);
        double fxkp1 = P.dxdy * f.v[1] / den, fykp1 = P.dxdy * f.v[2] / den;
        Res ca = co_await C.count(M_ADH);
        Res ci = co_await C.count(M_INT);
        double nadh = ca.v[0], ncon = ci.v[0];
        int itnr = 0;
        double df = std::fabs(P.in.fxrel - fxkp1);
        if (F >= 2) df = df + std::fabs(P.in.fyrel - fykp1);
        double& s_xk = sens_at(P, IOUT_FX, IIN_KSI);
        double& s_yk = sens_at(P, IOUT_FY, IIN_KSI);
        double& s_xe = sens_at(P, IOUT_FX, IIN_ETA);
        double& s_ye = sens_at(P, IOUT_FY, IIN_ETA);
        while (!rc && df > P.in.eps && itnr < P.in.maxnr && info1 <= 1) {
            itnr++;
            double dfx = P.in.fxrel - fxkp1, dfy = P.in.fyrel - fykp1, dcksi = 0.0, dceta = 0.0;
            if (F == 1) {
                dcksi = std::fabs(s_xk) > sp(1e-6f) ? dfx / s_xk : sp(0.00003f);
            } else {
                double det = s_xk * s_ye - s_yk * s_xe;
                if (det > P.in.eps) {
                    dcksi = (s_ye * dfx - s_yk * dfy) / det;
                    dceta = (-s_xe * dfx + s_xk * dfy) / det;
                } else {
                    dcksi = 0.000003;
                    dceta = 0.000003;
                }
            }
            for (int ifxy = 1; ifxy <= F && !rc; ifxy++) {
                double fxk = fxkp1, fyk = fykp1;
                if (ifxy == 1) {
                    P.in.cksi = P.in.cksi + dcksi;
                    add_rigid_slip(C, P.wsx, P.wsy, P.wsx, P.wsy, 0, dcksi);
                } else {
                    P.in.ceta = P.in.ceta + dceta;
                    add_rigid_slip(C, P.wsx, P.wsy, P.wsx, P.wsy, 1, dceta);
                }
                rc = co_await tang_solver(C, k, cf00, omegah, omegas, P.in.maxgs, P.in.eps, &it, &err, &info1);
                f = co_await C.opr(OP_FORCES, fi);
                fxkp1 = P.dxdy * f.v[1] / den;
                fykp1 = P.dxdy * f.v[2] / den;
                itgs += it;
                ca = co_await C.count(M_ADH);
                ci = co_await C.count(M_INT);
                nadh = ca.v[0];
                ncon = ci.v[0];
                double dfxk = fxkp1 - fxk, dfyk = fykp1 - fyk;
                double area = ncon * P.dxdy;
                if (ifxy == 1) {
                    if (nadh > 0) {
                        double dpxavg = dfxk * P.muscal * P.fntrue / area;
                        if (s_xk == 0.0 || std::fabs(dpxavg) > 10.0 * err) s_xk = dfxk / dcksi;
                        double dpyavg = dfyk * P.muscal * P.fntrue / area;
                        if (std::fabs(dpyavg) > 10.0 * err) s_yk = dfyk / dcksi;
                    } else if (std::fabs(s_xk) < sp(1e-6f)) {
                        s_xk = fxkp1 / P.in.cksi;
                    }
                } else {
                    if (nadh > 0) {
                        double dpxavg = dfxk * P.muscal * P.fntrue / area;
                        if (std::fabs(dpxavg) > 10.0 * err) s_xe = dfxk / dceta;
                        double dpyavg = dfyk * P.muscal * P.fntrue / area;
                        if (s_ye == 0.0 || std::fabs(dpyavg) > 10.0 * err) s_ye = dfyk / dceta;
                    } else if (std::fabs(s_ye) < sp(1e-6f)) {
                        s_ye = fykp1 / P.in.ceta;
                    }
                }
                df = std::fabs(P.in.fxrel - fxkp1);
                if (F >= 2) df = df + std::fabs(P.in.fyrel - fykp1);
            }
        }
        // the reference's error estimate adds the slip elements as the state code Slip = 2
        // (`slip` in src/m_solvpt.f90:366 names that constant); no plastic elements here
        err = err + 2.0 * df * P.muscal * P.fntrue / ((nadh + (double)SLIP) * P.dxdy);
    } else {
        sens_at(P, IOUT_FX, IIN_KSI) = 0.0;
        sens_at(P, IOUT_FY, IIN_KSI) = 0.0;
        sens_at(P, IOUT_FX, IIN_ETA) = 0.0;
        sens_at(P, IOUT_FY, IIN_ETA) = 0.0;
    }
    *itgs_out = itgs;
    *err_out = err;
    co_return rc;
}

static Task stang(Ctx C, int* ittang_out)
{
    Patch& P = *C.P;
    Res r;
    r = co_await C.count(M_INT);
    if (r.v[0] <= 0) co_return ERR_OTHER;
    const int k = (int)r.v[0];
    double cf00[4] = {0.0, 0.0, 0.0, 0.0};
    {
        // the solver (src/m_stang.f90:145-190): ConvexGS for G=2, and for steady rolling with
        // contact in the trailing-edge column (no exterior element to start a row's walk).  The
        // same launch returns the coefficients of element (0,0).
        AnyItem gi;
        gi.gs = gs_item(P, k, 1.0, 1.0, P.in.maxgs, P.in.eps);
        Res g = co_await C.opr(OP_GUARD, gi);
        P.cnvx = P.in.gausei == 2 || (P.in.tang == 3 && g.v[0] > 0);
        for (int q = 0; q < 4; q++) cf00[q] = g.v[1 + q];
    }
    // relaxation (src/m_stang.f90:192-223)
    double omegah = P.in.omegah, omegas = P.in.omegas;
    if (P.in.gausei == 0 || P.in.gausei == 4 || P.in.gausei == 5) {
        if (k <= 25) { omegah = 1.0; omegas = 1.0; }
        else if (P.in.tang == 3 && P.cnvx) { omegah = 0.5; omegas = 0.5; }
        else if (P.in.tang == 3) {
            double ar = P.dx / P.dy;
            if (ar <= 5.0) { omegah = 0.9; omegas = 1.0; }
            else if (ar <= 15.0) { omegah = 0.8; omegas = 0.8; }
            else { omegah = 0.8; omegas = 0.6; }
        } else { omegah = 0.5; omegas = 0.5; }
    }
    // stang_rhs: ws = -facdt h_t (- A'_tt(cv) pv_t for T=2) on the interior; facdt = 1 for
    // interior elements when dq = dx (src/m_leadedge.f90:290-318)
    if (P.ledge && P.in.tang == 2) {
        // transient rolling with dq > dx: facdt and the leading-edge positions from the element
        // division (sxbnd), then (1a) in the interior and (1b) at the leading edge, where the
        // previous-time displacement is replaced by ubnd, interpolated from A pv (cs)
        AnyItem li;
        LedgeItem& L = li.ledge;
        L.n = P.n; L.mx = P.mx; L.my = P.my; L.mode = 0;
        L.el = P.el; L.facdt = P.facdt; L.lix = P.lix;
        L.xc1 = P.xc1; L.dx = P.dx; L.dq = P.dq; L.fxdfac = 1.0;   // TangCG: edge 1 dx beyond ix0
        C.op(OP_LEDGE, li);
        double *ux = P.w[20], *uy = P.w[21], *bx = P.w[18], *by = P.w[19];
        C.conv_t(1, M_INT, ux, uy, P.pvx, P.pvy);
        C.conv_t(0, M_ALL, bx, by, P.pvx, P.pvy);
        L.mode = 1;
        L.hx = P.hx; L.hy = P.hy; L.ux = ux; L.uy = uy; L.bx = bx; L.by = by;
        L.wsx = P.wsx; L.wsy = P.wsy;
        C.op(OP_LEDGE, li);
    } else if (P.in.tang == 3 && P.cnvx) {
        // steady rolling with ConvexGS: facdt and the leading-edge positions (sxbnd, the edge
        // 1 dx beyond ix0), ws = -facdt h_t; ubnd enters ConvexGS's shift at each sweep
        AnyItem li;
        LedgeItem& L = li.ledge;
        L.n = P.n; L.mx = P.mx; L.my = P.my; L.mode = 0;
        L.el = P.el; L.facdt = P.facdt; L.lix = P.lix;
        L.xc1 = P.xc1; L.dx = P.dx; L.dq = P.dq; L.fxdfac = 1.0;
        C.op(OP_LEDGE, li);
        L.mode = 1;
        L.steady = 1;
        L.hx = P.hx; L.hy = P.hy;
        L.wsx = P.wsx; L.wsy = P.wsy;
        C.op(OP_LEDGE, li);
    } else {
        C.set(M_ALL, P.wsx, 0.0, P.wsy);
        C.axpy(M_INT, -1.0, P.hx, P.wsx, P.hy, P.wsy);
        if (P.in.tang == 2 || (P.in.tang == 1 && P.in.pvtime == 0)) {
            // the previous tractions' displacements A' pv_t (src/m_stang.f90:873): shifted
            // coefficients in transient rolling, cs itself in a shift (cv = cs)
            double *ux = P.w[20], *uy = P.w[21];
            C.conv_t(P.in.tang == 2 ? 1 : 0, M_INT, ux, uy, P.pvx, P.pvy);
            C.axpy(M_INT, -1.0, ux, P.wsx, uy, P.wsy);
        }
    }
    if (P.nt_cpl) {
        // K != 0: the tangential displacements of the normal tractions, ws += A_tn ps_n - A'_tn
        // pv_n on the interior, with pv = ps in steady rolling; pv = 0 in shift and transient
        // rolling without a previous case (stang_rhs, src/m_stang.f90:855-861)
        double *ux = P.w[20], *uy = P.w[21];
        int ik[2] = {0, 1}, jk[1] = {2};
        double* o[2] = {ux, uy};
        double* in[1] = {P.pn};
        C.conv(0, M_INT, 2, ik, o, 1, jk, in);
        C.axpy(M_INT, 1.0, ux, P.wsx, uy, P.wsy);
        if (P.in.tang == 3 || P.in.pvtime == 0) {
            double* prev[1] = {P.in.tang == 3 ? P.pn : P.pvn};
            C.conv(P.in.tang == 1 ? 0 : 1, M_INT, 2, ik, o, 1, jk, prev);
            C.axpy(M_INT, -1.0, ux, P.wsx, uy, P.wsy);
        }
    }
    // with forces prescribed, solvpt rebuilds its right-hand side from wsfix1 (kept in w[22],
    // w[23]) and the current creepages at every call
    if (P.in.force3 >= 1) snap(C, nullptr, nullptr, {P.wsx, P.wsy}, {P.w[22], P.w[23]});
    int ittang = 0, itgs = 0, ret = 0;
    bool zready = false;
    while (!zready && ittang < P.in.maxin) {
        ittang++;
        zready = true;
        int it = 0;
        double errpt = 0.0;
        int rc = co_await solvpt(C, k, cf00, omegah, omegas, &it, &errpt);
        itgs += it;
        if (rc) { ret = rc; break; }
        StItem S{};
        S.n = P.n; S.el = P.el; S.px = P.px; S.py = P.py; S.pn = P.pn; S.sx = P.sx; S.sy = P.sy;
        S.mu = P.in.fstat;
        S.tol = std::sqrt(2.0) * errpt;
        S.errpt = errpt;
        AnyItem sa;
        sa.st = S;
        r = co_await C.opr(OP_ST_AS, sa);
        if (r.v[0] != 0) zready = false;
        if (zready) {
            Res a = co_await aij_ones(C, 0);
            Res b = co_await aij_ones(C, 4);
            sa.st.tol1 = errpt * 2.0 * std::fabs(a.v[0] / P.ga);
            sa.st.tol2 = errpt * 2.0 * std::fabs(b.v[0] / P.ga);
            r = co_await C.opr(OP_ST_SA, sa);
            if (r.v[0] != 0) zready = false;
        }
    }
    if (!zready && !ret) ittang = -1;
    P.out.itgs = itgs;
    *ittang_out = ittang;
    // sensitivities to the creepages, src/m_stang.f90:724-731
    if (!ret && P.in.sens >= 3) ret = co_await sens_tang(C, k, cf00, omegah, omegas);
    co_return ret;
}

// ------------------------------------------------------------------------------------------
// one case: panprc (src/m_scontc.f90:356-553) and the force integration of soutpt
// (src/m_soutpt.f90:424-450)
// ------------------------------------------------------------------------------------------


static Task calculate(Ctx C)
{
    Patch& P = *C.P;
    CaseResult& R = P.out;
    P.dense = dense_mode() == 2 && P.n <= DENSE_MAX_N;
    {
        // set_prev_data, set_norm_rhs, set_tang_rhs, init_curr_data (src/m_sdis.f90:208-768)
        AnyItem it;
        it.setup = setup_item(P);
        SetupItem& S = it.setup;
        S.prev = P.in.pvtime == 0;
        S.zero_pn = P.in.iestim == 0;
        S.zero_pt = P.in.iestim == 0 || P.in.tang == 0;
        S.hertz = !P.hertz_init ? 0 : ((P.in.ipotcn == -4 || P.in.ipotcn == -5) ? 2 : 1);
        S.regularize = P.in.iestim == 1;
        S.mu = P.in.fstat;
        // set_tang_rhs (src/m_sdis.f90:566-575): a creepage enters the rigid slip only when it
        // is prescribed (F=0: both; F=1: ceta); init_curr_data starts the force-controlled ones
        // from cksi = 1e-6, ceta = 0 (:759-763)
        if (P.in.iestim == 0 || P.in.iestim == 2) {
            if (P.in.force3 >= 1) P.in.cksi = 1e-6;
            if (P.in.force3 == 2) P.in.ceta = 0.0;
        }
        S.cksi = P.in.force3 == 0 ? P.in.cksi : 0.0;
        S.ceta = P.in.force3 <= 1 ? P.in.ceta : 0.0;
        S.cphi = P.in.cphi; S.dq = P.dq;
        S.xofs = P.xofs; S.yofs = P.yofs;
        S.hz_aa = R.hz_aa; S.hz_bb = R.hz_bb; S.pnmax = P.hz_pnmax;
        // E digit (src/m_sdis.f90:517-553): 1 extra rigid slip (copied into hx, hy by prepare),
        // 2 no spin in wy, 3 no spin in wx
        S.exrhs = P.in.rztang == 1 ? 1 : 0;
        S.facy = P.in.rztang == 2 ? 0.0 : 1.0;
        S.facx = P.in.rztang == 3 ? 0.0 : 1.0;
        Res r = co_await C.opr(OP_SETUP, it);
        P.hmax = r.v[0];
        P.hmin = -r.v[1];
    }
    // initial sensitivities (init_curr_data, src/m_sdis.f90:753-764); eldiv0's estimate of
    // dFn/dpen for N=1 is always replaced in snorm and is not formed
    {
        const int ihertz = P.in.ipotcn < 0 ? -P.in.ipotcn : 0;
        if (P.in.iestim == 0)
            for (int io = IOUT_FX; io <= IOUT_MZ; io++) sens_at(P, io, IIN_PEN) = 0.0;
        if (P.in.iestim == 0 && (P.in.norm != 1 || ihertz != 0)) sens_at(P, IOUT_FN, IIN_PEN) = 0.0;
        if (P.in.iestim == 0 || P.in.iestim == 2)
            for (int ii = IIN_KSI; ii <= IIN_PHI; ii++)
                for (int io = IOUT_FN; io <= IOUT_MZ; io++) sens_at(P, io, ii) = 0.0;
    }
    // steady rolling with K != 0 also needs the spectra of the shifted coefficients (A'_tn)
    co_await setup_coefficients(C, P.in.tang == 2 || (P.in.tang == 3 && P.nt_cpl), P.in.tang == 3);
    // panprc (src/m_scontc.f90:356-553): NORM and TANG alternate until the tractions change by
    // at most 5 eps of their RMS on the interior, up to maxout times (1 unless K != 0)
    const int maxout = P.nt_cpl ? std::max(1, P.in.maxout) : 1;
    int itnorm_tot = 0, ittang_tot = 0, itout = 0, rc = 0;
    double difid = 1.0, dif = 200.0 * difid;
    if (maxout > 1) C.copy(M_ALL, P.px, P.po[0], P.py, P.po[1]), C.copy(M_ALL, P.pn, P.po[2]);
    while (dif > difid && itout < maxout && itnorm_tot >= 0 && ittang_tot >= 0) {
        itout++;
        int itnorm = 0, ittang = 0;
        // B=2: the Hertzian pressure is the normal solution (panprc skips NORM, :409-411)
        rc = P.in.bound == 2 ? 0 : co_await snorm(C, &itnorm);
        if (rc) break;
        itnorm_tot = itnorm >= 0 ? itnorm_tot + itnorm : -1;
        C.set(M_EXT, P.px, 0.0, P.py);
        C.set(M_EXT, P.pn, 0.0);
        Res r = co_await C.count(M_INT);
        if (P.in.tang == 0 || r.v[0] <= 0) {
            C.set(M_ALL, P.sx, 0.0, P.sy);
            dif = 0.0;
        } else {
            rc = co_await stang(C, &ittang);
            if (rc) break;
            ittang_tot = ittang >= 0 ? ittang_tot + ittang : -1;
            if (maxout > 1) {
                // dif = RMS of the change of (px, py, pn) over the interior, difid = 5 eps RMS(p)
                C.axpy(M_ALL, -1.0, P.px, P.po[0], P.py, P.po[1]);
                C.axpy(M_ALL, -1.0, P.pn, P.po[2]);
                Res d2 = co_await C.sumsq(M_INT, P.po[0], P.po[1]);
                Res d1 = co_await C.sumsq(M_INT, P.po[2]);
                Res p2 = co_await C.sumsq(M_INT, P.px, P.py);
                Res p1 = co_await C.sumsq(M_INT, P.pn);
                const double cnt = 3.0 * std::max(1.0, d1.v[1]);
                dif = std::sqrt((d2.v[0] + d1.v[0]) / cnt);
                difid = 5.0 * P.in.eps * std::sqrt((p2.v[0] + p1.v[0]) / cnt);
                C.copy(M_ALL, P.px, P.po[0], P.py, P.po[1]);
                C.copy(M_ALL, P.pn, P.po[2]);
            }
        }
    }
    R.itnorm = itnorm_tot;
    R.ittang = ittang_tot;
    // displacement differences of the solution (soutpt, src/m_soutpt.f90:371-397): us = A ps on
    // the interior, 0 elsewhere (init_curr_data clears us for every case, src/m_sdis.f90:624);
    // with K = 0 the normal and tangential parts decouple.  Read by cntc_getdisplacements.
    C.set(M_ALL, P.un, 0.0);
    C.set(M_ALL, P.ux, 0.0, P.uy);
    if (P.nt_cpl && P.in.bound != 2 && P.in.tang != 0) {
        int ik[3] = {0, 1, 2}, jk[3] = {0, 1, 2};
        double* o[3] = {P.ux, P.uy, P.un};
        double* in[3] = {P.px, P.py, P.pn};
        C.conv(0, M_INT, 3, ik, o, 3, jk, in);
    } else {
        if (P.in.bound != 2) C.conv_n(0, M_INT, P.un, P.pn);   // B=2: normal part 0 (:379-380)
        if (P.in.tang != 0) C.conv_t(0, M_INT, P.ux, P.uy, P.px, P.py);
    }
    // forces (soutpt, src/m_soutpt.f90:424-450)
    AnyItem fi;
    fi.setup = setup_item(P);
    Res s = co_await C.opr(OP_FORCES, fi);
    if (P.in.norm == 0) P.fntrue = P.dxdy * s.v[0];
    double den = P.fntrue * P.muscal + TINY;
    double fxrel = P.dxdy * s.v[1] / den, fyrel = P.dxdy * s.v[2] / den;
    R.fx = fxrel * den;
    R.fy = fyrel * den;
    R.fn = P.fntrue;
    R.mz = -P.dxdy * s.v[3] + P.dxdy * s.v[4];
    if (std::fabs(R.mz) < 0.5 * P.in.eps * (P.muscal * P.fntrue + TINY)) R.mz = 0.0;
    R.pen = P.pen;
    P.have_solution = true;
    R.t_done = std::chrono::duration<double>(std::chrono::steady_clock::now() - C.E->t_solve0).count();
    if (rc) R.ierror = rc;
    else if (R.itnorm < 0) R.ierror = ERR_NORM;
    else if (R.ittang < 0) R.ierror = ERR_TANG;
    co_return R.ierror;
}

int Engine::solve_noalloc(Patch* p)
{
    last_error.clear();
    if (!noalloc_base_) {
        last_error = "the no-allocation array is not available on this GPU";
        return -1;
    }
    if (serving_) {
        last_error = "the serving loop is running";
        return -1;
    }
    char* base = noalloc_base_ + NOALLOC_ARENA;
    FlushBufs fb;
    fb.stage = base;
    fb.stage_cap = NOALLOC_STAGE;
    fb.part = (double*)(base + NOALLOC_STAGE);
    fb.part_cap = NOALLOC_PART / sizeof(double);
    fb.res = (double*)(base + NOALLOC_STAGE + NOALLOC_PART);
    fb.res_cap = NOALLOC_RES / sizeof(double);
    fb.cnt = (unsigned*)(base + NOALLOC_STAGE + NOALLOC_PART + NOALLOC_RES);
    fb.cnt_cap = NOALLOC_CNT / sizeof(unsigned);
    // the array keeps the last call's contents: the arrival counters must start at zero
    launch_fill_u32(fb.cnt, fb.cnt_cap, 0u, stream_);
    HIPCHK(hipGetLastError());
    sched_->use_fixed(fb);
    noalloc_ = true;
    int rc = 0;
    try {
        rc = solve({p});
    } catch (...) {
        noalloc_ = false;
        sched_->restore();
        throw;
    }
    noalloc_ = false;
    sched_->restore();
    if (p->out.ierror == 0 || p->have_solution) {
        // keep the solution with the problem: the array is reused by the next call
        const size_t n = p->n, lead = ((n + 255) & ~(size_t)255) + 40 * n * 8;
        p->host_state.resize(lead);
        HIPCHK(hipMemcpyAsync(p->host_state.data(), p->el, lead, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        p->state_on_host = true;
    }
    return rc;
}

int Engine::solve(const std::vector<Patch*>& patches)
{
    last_error.clear();
    t_solve0 = std::chrono::steady_clock::now();
    std::vector<Patch*> ok;
    for (Patch* p : patches)
        if (prepare(*p)) ok.push_back(p);
    HIPCHK(hipStreamSynchronize(stream_));
    sched_->resize((int)ok.size());
    std::vector<Task> tasks;
    std::vector<int> slots;
    // the largest patches of a big batch set its length: give them flushes of their own
    long maxn = 0;
    for (Patch* p : ok) maxn = std::max(maxn, (long)p->n);
    for (size_t i = 0; i < ok.size(); i++) {
        const bool pr = ok.size() >= 16 && (long)sched_->prio_div() * ok[i]->n >= maxn;
        sched_->set_priority((int)i, pr);
        if (pr) ok[i]->out.prio = 1;
    }
    for (size_t i = 0; i < ok.size(); i++) {
        ok[i]->slot = (int)i;
        tasks.push_back(calculate(Ctx{sched_.get(), ok[i], (int)i, this}));
        slots.push_back((int)i);
    }
    sched_->run(tasks, slots);
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

// ------------------------------------------------------------------------------------------
// continuous batching: requests are admitted between flushes while others are in flight
// ------------------------------------------------------------------------------------------

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool Engine::serve_start(size_t budget_bytes)
{
    std::lock_guard<std::mutex> lk(smtx_);
    if (serving_) return true;
    budget_ = budget_bytes;
    in_use_ = 0;
    stop_ = false;
    serving_ = true;
    server_ = std::thread([this] { serve_loop(); });
    return true;
}

void Engine::serve_stop()
{
    {
        std::lock_guard<std::mutex> lk(smtx_);
        if (!serving_) return;
        stop_ = true;
    }
    scv_.notify_all();
    server_.join();
    std::lock_guard<std::mutex> lk(smtx_);
    serving_ = false;
}

int Engine::submit(Patch* p)
{
    std::lock_guard<std::mutex> lk(smtx_);
    if (!serving_ || stop_) return -1;
    int id = next_ticket_++;
    Ticket& t = tickets_[id];
    t.p = p;
    t.t_submit = now_s();
    queue_.push_back(id);
    scv_.notify_all();
    return id;
}

int Engine::poll(int ticket, int* rc, double* latency, bool wait)
{
    std::unique_lock<std::mutex> lk(smtx_);
    auto it = tickets_.find(ticket);
    if (it == tickets_.end()) return -1;
    if (wait) dcv_.wait(lk, [&] { return tickets_[ticket].state == 2; });
    Ticket& t = tickets_[ticket];
    if (t.state != 2) return 0;
    if (rc) *rc = t.rc;
    if (latency) *latency = t.t_done - t.t_submit;
    return 1;
}

void Engine::forget(int ticket)
{
    std::lock_guard<std::mutex> lk(smtx_);
    auto it = tickets_.find(ticket);
    if (it != tickets_.end() && it->second.state == 2) tickets_.erase(it);
}

void Engine::serve_loop()
{
    struct Act {
        int ticket, slot;
        size_t bytes;
        Task task;
    };
    const int cap = 16384;
    sched_->resize(cap);
    std::vector<int> free_slots;
    for (int k = cap - 1; k >= 0; k--) free_slots.push_back(k);
    std::vector<Act> active;
    auto finish = [&](int id, int rc, size_t bytes) {
        std::lock_guard<std::mutex> lk(smtx_);
        Ticket& t = tickets_[id];
        t.rc = rc;
        t.state = 2;
        t.t_done = now_s();
        in_use_ -= bytes;
        dcv_.notify_all();
    };
    for (;;) {
        std::vector<std::pair<int, size_t>> admit;
        {
            std::unique_lock<std::mutex> lk(smtx_);
            if (active.empty()) scv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (stop_ && queue_.empty() && active.empty()) break;
            // admission by device-memory footprint of the active patches
            while (!queue_.empty() && (int)admit.size() < (int)free_slots.size()) {
                Ticket& t = tickets_[queue_.front()];
                size_t b = patch_bytes(std::max(1, t.p->in.mx), std::max(1, t.p->in.my));
                if ((!active.empty() || !admit.empty()) && in_use_ + b > budget_) break;
                in_use_ += b;
                t.state = 1;
                admit.push_back({queue_.front(), b});
                queue_.pop_front();
            }
        }
        for (auto& ab : admit) {
            Patch* p = tickets_at(ab.first);
            bool ok = false;
            try {
                t_solve0 = std::chrono::steady_clock::now();
                ok = prepare(*p);
            } catch (const std::exception& ex) {
                last_error = ex.what();
                p->out.ierror = ERR_OTHER;
            }
            if (!ok) {
                finish(ab.first, p->out.ierror ? p->out.ierror : ERR_OTHER, ab.second);
                continue;
            }
            int slot = free_slots.back();
            free_slots.pop_back();
            p->slot = slot;
            active.push_back(Act{ab.first, slot, ab.second, calculate(Ctx{sched_.get(), p, slot, this})});
            active.back().task.h.resume();
        }
        if (active.empty()) continue;
        if (!admit.empty()) {
            // the largest patches of a big active set set its length: give them flushes of their
            // own.  Decided over everything active, so that patches admitted over several
            // rounds (a large batch arriving while the first requests start) are ranked together.
            long maxn = 0;
            for (auto& a : active) maxn = std::max(maxn, (long)tickets_at(a.ticket)->n);
            const bool many = active.size() >= 16;
            for (auto& a : active) {
                Patch* p = tickets_at(a.ticket);
                const bool pr = many && (long)sched_->prio_div() * p->n >= maxn;
                sched_->set_priority(a.slot, pr);
                if (pr) p->out.prio = 1;
            }
        }
        std::vector<int> slots;
        for (auto& a : active) slots.push_back(a.slot);
        try {
            // nothing to flush while only side-stream work is pending: poll it gently
            if (!sched_->step(slots)) std::this_thread::sleep_for(std::chrono::microseconds(50));
        } catch (const std::exception& ex) {
            last_error = ex.what();
            for (auto& a : active) finish(a.ticket, ERR_OTHER, a.bytes);
            active.clear();
            sched_->resize(cap);
            free_slots.clear();
            for (int k = cap - 1; k >= 0; k--) free_slots.push_back(k);
            continue;
        }
        for (size_t k = 0; k < active.size();) {
            if (active[k].task.h.done()) {
                Patch* p = tickets_at(active[k].ticket);
                finish(active[k].ticket, p->out.ierror, active[k].bytes);
                free_slots.push_back(active[k].slot);
                active.erase(active.begin() + k);
            } else {
                k++;
            }
        }
    }
}

Patch* Engine::tickets_at(int id)
{
    std::lock_guard<std::mutex> lk(smtx_);
    return tickets_[id].p;
}

void Engine::set_profile(int mode)
{
    std::lock_guard<std::mutex> lk(sched_->prof_mtx);
    sched_->profile = mode == 1 || mode == 2 ? mode : 0;
    sched_->prof.clear();
    sched_->nflush = 0;
}

int Engine::get_profile(int len, int* kinds, double* seconds, double* bytes, long* launches) const
{
    std::lock_guard<std::mutex> lk(sched_->prof_mtx);
    int k = 0;
    for (auto& kv : sched_->prof) {
        if (k < len) {
            kinds[k] = kv.first;
            seconds[k] = kv.second.seconds;
            bytes[k] = kv.second.bytes;
            launches[k] = kv.second.launches;
        }
        k++;
    }
    return k;
}

long Engine::rounds() const
{
    std::lock_guard<std::mutex> lk(sched_->prof_mtx);
    return sched_->nflush;
}

void Engine::get_field(const Patch& p, int what, double* out) const
{
    if (p.state_on_host && p.n > 0) {
        // grid functions in arena order: pn px py sx sy pvx pvy pvn h hx hy ws.. w[24] un ux uy
        static const int slot[11] = {0, 1, 2, 3, 4, 8, 37, 38, 39, 9, 10};
        if (what < 0 || what > 10) return;
        const size_t n = p.n, off = ((n + 255) & ~(size_t)255) + (size_t)slot[what] * n * 8;
        if (p.host_state.size() >= off + n * 8) std::memcpy(out, p.host_state.data() + off, n * 8);
        return;
    }
    const double* src = nullptr;
    switch (what) {
    case 0: src = p.pn; break;
    case 1: src = p.px; break;
    case 2: src = p.py; break;
    case 3: src = p.sx; break;
    case 4: src = p.sy; break;
    case 5: src = p.h; break;
    case 6: src = p.un; break;
    case 7: src = p.ux; break;
    case 8: src = p.uy; break;
    case 9: src = p.hx; break;
    case 10: src = p.hy; break;
    }
    if (!src || p.n <= 0) return;
    HIPCHK(hipMemcpyAsync(out, src, (size_t)p.n * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
}

void Engine::get_eldiv(const Patch& p, int* out) const
{
    if (p.n <= 0) return;
    if (!p.el) {   // never laid out: no solution yet
        for (int i = 0; i < p.n; i++) out[i] = 0;
        return;
    }
    if (p.state_on_host && p.host_state.size() >= (size_t)p.n) {
        for (int i = 0; i < p.n; i++) out[i] = (int8_t)p.host_state[i];
        return;
    }
    std::vector<int8_t> e(p.n);
    HIPCHK(hipMemcpyAsync(e.data(), p.el, p.n, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int i = 0; i < p.n; i++) out[i] = e[i];
}

}  // namespace cntc
