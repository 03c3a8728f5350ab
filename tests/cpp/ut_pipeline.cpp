// CPU unit tests of the pipe and agent host (no device code).  Modelled on the
// reference's UnitTests/UTPipeline (UTPipe.cpp:261-432 concurrent writers and
// readers with overlap and shots; UTAgentBase.cpp:22-85 a 7-thread sum agent
// and agents that throw).  Prints one line per test; exit status 0 iff all pass.
#include <atomic>
#include <cstdio>
#include <functional>
#include <numeric>
#include <random>
#include <thread>

#include "jh/Agent.h"

using namespace jh;

static int g_failed = 0;

#define CHECK(cond)                                                                      \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            std::fprintf(stderr, "  CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            throw std::runtime_error("check failed");                                    \
        }                                                                                \
    } while (0)

static void Run(const char* name, const std::function<void()>& f)
{
    try {
        f();
        std::printf("PASS %s\n", name);
    } catch (std::exception& e) {
        ++g_failed;
        std::printf("FAIL %s: %s\n", name, e.what());
    }
    std::fflush(stdout);
}

// --------------------------------------------------------------------------

static void TokenSizesAndShots()
{
    Pipe<size_t> p("p");
    p.SetTemplate(size_t(0), {1});
    p.SetProducer("w", 1, 3);
    p.SetConsumer("r", 1, 2, 0);
    CHECK(p.GetBufferSize() == 3 + 2);
    {
        auto w = p.GetWriteToken(3, false);
        CHECK(w.IsShotStart() && !w.IsShotEnd() && w.GetStartIndex() == 0 && w.GetSize() == 3);
        for (size_t i = 0; i < 3; ++i) w.GetBuffer(i) = i;
    }
    {
        auto w = p.GetWriteToken(2, true);
        CHECK(!w.IsShotStart() && w.IsShotEnd() && w.GetStartIndex() == 3);
        for (size_t i = 0; i < 2; ++i) w.GetBuffer(i) = 3 + i;
    }
    size_t expect = 0;
    for (int t = 0; t < 3; ++t) {
        auto r = p.GetReadToken();
        CHECK(r.GetStartIndex() == expect);
        CHECK(r.IsShotStart() == (t == 0));
        CHECK(r.IsShotEnd() == (t == 2));
        CHECK(r.GetSize() == (t == 2 ? 1u : 2u));
        for (size_t i = 0; i < r.GetSize(); ++i) CHECK(r.GetBuffer(i) == expect + i);
        expect += r.GetSize();
    }
    p.Close();
    bool drained = false;
    try {
        p.GetReadToken();
    } catch (PipeClosedAndEmptySignal&) {
        drained = true;
    }
    CHECK(drained);
    bool threw = false;
    try {
        p.GetWriteToken(1, false);
    } catch (Exception&) {
        threw = true;
    }
    CHECK(threw);
}

static void OverlapWithinShots()
{
    // reads of 5 with overlap 2 never cross a shot and repeat 2 elements
    Pipe<size_t> p("p");
    p.SetTemplate(size_t(0), {1});
    p.SetProducer("w", 1, 1);
    p.SetConsumer("r", 1, 5, 2);
    std::vector<size_t> shots = {7, 3, 12};
    std::thread writer([&] {
        size_t v = 0;
        for (size_t s : shots)
            for (size_t i = 0; i < s; ++i) {
                auto w = p.GetWriteToken(1, i + 1 == s);
                w.GetBuffer(0) = v++;
            }
        p.Close();
    });
    size_t shotStart = 0, shot = 0, next = 0;
    try {
        for (;;) {
            auto r = p.GetReadToken();
            CHECK(r.GetStartIndex() >= shotStart);
            if (r.IsShotStart()) {
                CHECK(r.GetOverlapSize() == 0);
                CHECK(r.GetStartIndex() == shotStart);
            } else {
                CHECK(r.GetOverlapSize() == 2);
            }
            for (size_t i = 0; i < r.GetSize(); ++i) CHECK(r.GetBuffer(i) == r.GetStartIndex() + i);
            CHECK(r.GetStartIndex() + r.GetOverlapSize() == next);
            next = r.GetStartIndex() + r.GetSize();
            if (r.IsShotEnd()) {
                CHECK(next == shotStart + shots[shot]);
                shotStart = next;
                ++shot;
            }
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    writer.join();
    CHECK(shot == shots.size());
}

static void ConcurrentWritersReaders()
{
    // 3 writers, 11 readers, read 5 with overlap 2, shots of 3..99
    Pipe<size_t> p("p");
    p.SetTemplate(size_t(0), {1});
    p.SetProducer("w", 3, 1);
    p.SetConsumer("r", 11, 5, 2);
    std::mt19937 rng(7);
    std::vector<size_t> shots;
    size_t total = 0;
    for (int i = 0; i < 40; ++i) {
        shots.push_back(3 + rng() % 97);
        total += shots.back();
    }
    std::mutex wm;
    size_t wshot = 0, wi = 0;
    std::vector<std::thread> ws;
    for (int t = 0; t < 3; ++t)
        ws.emplace_back([&] {
            for (;;) {
                PipeWriteToken<size_t> w;
                {
                    std::lock_guard<std::mutex> lk(wm);
                    if (wshot == shots.size()) return;
                    const bool end = wi + 1 == shots[wshot];
                    w = p.GetWriteToken(1, end);
                    if (end) { ++wshot; wi = 0; } else ++wi;
                }
                std::this_thread::yield();
                w.GetBuffer(0) = w.GetStartIndex() * 3 + 1;
            }
        });
    std::atomic<size_t> seen{0};
    std::atomic<int> bad{0};
    std::vector<std::thread> rs;
    for (int t = 0; t < 11; ++t)
        rs.emplace_back([&] {
            try {
                for (;;) {
                    auto r = p.GetReadToken();
                    for (size_t i = 0; i < r.GetSize(); ++i)
                        if (r.GetBuffer(i) != (r.GetStartIndex() + i) * 3 + 1) ++bad;
                    seen += r.GetSize() - r.GetOverlapSize();
                }
            } catch (PipeClosedAndEmptySignal&) {
            }
        });
    for (auto& t : ws) t.join();
    p.Close();
    for (auto& t : rs) t.join();
    CHECK(bad == 0);
    CHECK(seen == total);
}

static void DeadlockDetected()
{
    // a read needs 4 elements but the ring holds 3: the writer can never
    // progress and nobody holds a token
    Pipe<int> p("p");
    p.SetTemplate(0, {1});
    p.SetBufferSize(3);
    p.SetProducer("w", 1, 1);
    p.SetConsumer("r", 1, 4, 0);
    bool threw = false;
    try {
        for (int i = 0; i < 10; ++i) p.GetWriteToken(1, false);
    } catch (Exception& e) {
        threw = std::string(e.what()).find("deadlock") != std::string::npos;
    }
    CHECK(threw);
    CHECK(p.Closed());
}

// ----- agents ---------------------------------------------------------------

class SourceAgent : public SequentialAgentBase {
public:
    SourceAgent(size_t n, bool fail = false) : SequentialAgentBase("Source", 1), n_(n), fail_(fail) {}

private:
    void SetPipesImpl() override
    {
        out_ = CreateNewPipe<size_t>("Num");
        out_->SetTemplate(size_t(0), {1});
        out_->SetProducer(GetAgentName(), 1, 1);
        MergePipeToTrunk(out_);
    }
    void WorkFlow0() override
    {
        for (size_t i = 0; i < n_; ++i) {
            if (fail_ && i == n_ / 2) throw std::runtime_error("source fails on purpose");
            auto w = out_->GetWriteToken(1, i + 1 == n_);
            w.GetBuffer(0) = i + 1;
        }
        out_->Close();
    }
    size_t n_;
    bool fail_;
    std::shared_ptr<Pipe<size_t>> out_;
};

// 7 workers add up chunks of the input (UnitTests/UTPipeline/SumAgent.cpp:21-55)
class SumAgent : public ConcurrentAgentBase {
public:
    explicit SumAgent(bool failWorker = false) : ConcurrentAgentBase("Sum", 7), failWorker_(failWorker) {}
    std::atomic<unsigned long long> sum{0};

private:
    struct Task : TaskBase {
        PipeReadToken<size_t> token;
    };
    void SetPipesImpl() override
    {
        in_ = BranchPipeFromTrunk<size_t>("Num");
        in_->SetConsumer(GetAgentName(), 8, 1000, 0);
    }
    void ManagerWorkFlow() override
    {
        for (;;) {
            auto r = in_->GetReadToken();
            WaitIdleWorkerIndex();
            auto t = std::make_unique<Task>();
            t->token = r;
            SubmitTask(std::move(t));
        }
    }
    void ProcessTask(size_t, std::unique_ptr<TaskBase> base) override
    {
        auto* t = static_cast<Task*>(base.get());
        if (failWorker_ && t->token.GetStartIndex() >= 500000) throw Exception("worker fails on purpose");
        unsigned long long s = 0;
        for (size_t i = 0; i < t->token.GetSize(); ++i) s += t->token.GetBuffer(i);
        sum += s;
    }
    bool failWorker_;
    std::shared_ptr<Pipe<size_t>> in_;
};

static void RunGraph(std::vector<std::shared_ptr<AgentBase>> agents)
{
    std::vector<std::shared_ptr<PipeBase>> trunk;
    for (auto& a : agents) a->SetPipes(trunk);
    for (auto& a : agents) a->GetReady();
    for (auto& a : agents) a->Start();
    for (auto& a : agents) a->Join();
}

static void ConcurrentSum()
{
    const size_t n = 1000000;
    auto src = std::make_shared<SourceAgent>(n);
    auto sum = std::make_shared<SumAgent>();
    RunGraph({src, sum});
    CHECK(sum->sum == (unsigned long long)n * (n + 1) / 2);
    CHECK(!src->Failed() && !sum->Failed());
}

static void FailingSourceShutsDownGraph()
{
    auto src = std::make_shared<SourceAgent>(100000, true);
    auto sum = std::make_shared<SumAgent>();
    RunGraph({src, sum});  // must return
    CHECK(src->Failed());
    CHECK(sum->sum < 100000ull * 100001 / 2);
}

static void FailingWorkerShutsDownGraph()
{
    auto src = std::make_shared<SourceAgent>(1000000);
    auto sum = std::make_shared<SumAgent>(true);
    RunGraph({src, sum});  // must return
    CHECK(sum->Failed());
}

static void MissingPipeIsReported()
{
    auto sum = std::make_shared<SumAgent>();
    std::vector<std::shared_ptr<PipeBase>> trunk;
    bool threw = false;
    try {
        sum->SetPipes(trunk);
    } catch (Exception& e) {
        threw = std::string(e.what()).find("does not contain") != std::string::npos;
    }
    CHECK(threw);
}

int main()
{
    Run("TokenSizesAndShots", TokenSizesAndShots);
    Run("OverlapWithinShots", OverlapWithinShots);
    Run("ConcurrentWritersReaders", ConcurrentWritersReaders);
    Run("DeadlockDetected", DeadlockDetected);
    Run("ConcurrentSum", ConcurrentSum);
    Run("FailingSourceShutsDownGraph", FailingSourceShutsDownGraph);
    Run("FailingWorkerShutsDownGraph", FailingWorkerShutsDownGraph);
    Run("MissingPipeIsReported", MissingPipeIsReported);
    std::printf("%d failed\n", g_failed);
    return g_failed ? 1 : 0;
}
