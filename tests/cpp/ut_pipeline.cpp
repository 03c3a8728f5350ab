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

#include <cstdlib>
#include <fstream>
#include <unistd.h>

#include "jh/Agent.h"
#include "jh/FileAgents.h"

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

// ----- dump / load and the on-disk formats ----------------------------------

static std::string TempDir()
{
    char tmpl[] = "/tmp/jh_ut_XXXXXX";
    char* d = mkdtemp(tmpl);
    CHECK(d != nullptr);
    return d;
}

static std::vector<char> Slurp(const std::string& path)
{
    std::ifstream is(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
}

// PipeBase.cpp:263-268,310-313: a dump holds every finished element in
// stream order; a load replaces each read token's new elements, in order,
// also when read tokens overlap.
static void PipeDumpAndLoad()
{
    const std::string dir = TempDir();
    const std::string path = dir + "/proj.dump";
    const size_t n = 7, len = 5;
    {
        Pipe<FloatVec> p("Proj");
        p.SetTemplate(FloatVec(len), {len});
        p.SetProducer("w", 1, 1);
        p.SetConsumer("r", 1, 3, 1);
        p.SetDump(path);
        std::thread reader([&] {
            try {
                for (;;) auto r = p.GetReadToken();
            } catch (PipeClosedAndEmptySignal&) {
            }
        });
        for (size_t i = 0; i < n; ++i) {
            auto w = p.GetWriteToken(1, i + 1 == n);
            for (size_t k = 0; k < len; ++k) w.GetBuffer(0)[k] = (float)(10 * i + k);
        }
        p.Close();
        reader.join();
    }
    const std::vector<char> bytes = Slurp(path);
    CHECK(bytes.size() == n * len * sizeof(float));
    const float* f = reinterpret_cast<const float*>(bytes.data());
    for (size_t i = 0; i < n * len; ++i) CHECK(f[i] == (float)(10 * (i / len) + i % len));
    // replay: the producer writes zeros, the reader sees the dumped stream
    Pipe<FloatVec> q("Proj");
    q.SetTemplate(FloatVec(len), {len});
    q.SetProducer("w", 1, 1);
    q.SetConsumer("r", 1, 3, 1);
    q.SetLoad(path);
    std::vector<float> seen;
    std::thread reader([&] {
        try {
            for (;;) {
                auto r = q.GetReadToken();
                for (size_t i = r.GetOverlapSize(); i < r.GetSize(); ++i) seen.push_back(r.GetBuffer(i)[0]);
                if (r.GetOverlapSize()) CHECK(r.GetBuffer(0)[0] == seen[seen.size() - r.GetSize()]);
            }
        } catch (PipeClosedAndEmptySignal&) {
        }
    });
    for (size_t i = 0; i < n; ++i) {
        auto w = q.GetWriteToken(1, i + 1 == n);
        std::fill(w.GetBuffer(0).begin(), w.GetBuffer(0).end(), 0.0f);
    }
    q.Close();
    reader.join();
    CHECK(seen.size() == n);
    for (size_t i = 0; i < n; ++i) CHECK(seen[i] == (float)(10 * i));
    unlink(path.c_str());
    rmdir(dir.c_str());
}

static void LoadFileTooShortFailsTheReader()
{
    const std::string dir = TempDir();
    const std::string path = dir + "/short.dump";
    {
        std::ofstream os(path, std::ios::binary);
        const float v[3] = {1, 2, 3};
        os.write(reinterpret_cast<const char*>(v), sizeof v);
    }
    Pipe<FloatVec> q("Proj");
    q.SetTemplate(FloatVec(2), {2});
    q.SetProducer("w", 1, 1);
    q.SetConsumer("r", 1, 1, 0);
    q.SetLoad(path);
    for (int i = 0; i < 2; ++i) q.GetWriteToken(1, i == 1);
    { auto r = q.GetReadToken(); CHECK(r.GetBuffer(0)[0] == 1.0f && r.GetBuffer(0)[1] == 2.0f); }
    bool threw = false;
    try {
        auto r = q.GetReadToken();
    } catch (Exception& e) {
        threw = std::string(e.what()).find("ended early") != std::string::npos;
    }
    CHECK(threw);
    unlink(path.c_str());
    rmdir(dir.c_str());
}

// JEngine/PTMReader.h:17-23: 12 whitespace-separated numbers per view.
static void PTMFileRoundTrip()
{
    const std::string dir = TempDir();
    const std::string path = dir + "/PTM.dat";
    std::vector<double> ptm(24);
    for (size_t i = 0; i < ptm.size(); ++i) ptm[i] = (i % 5 ? 1.0 : -1.0) * (0.1 + i) / 3.0 * (i == 11 ? 1e-7 : 1.0);
    WritePTMFile(path, ptm);
    CHECK(ReadPTMFile(path, 2) == ptm);
    bool threw = false;
    try {
        ReadPTMFile(path, 3);
    } catch (Exception& e) {
        threw = std::string(e.what()).find("numbers") != std::string::npos;
    }
    CHECK(threw);
    {
        std::ofstream os(path);
        os << "1 2 3 4\n5 6 7 8 9 10 11 x12\n";
    }
    threw = false;
    try {
        ReadPTMFile(path, 1);
    } catch (Exception& e) {
        threw = std::string(e.what()).find("bad number") != std::string::npos;
    }
    CHECK(threw);
    {
        // the reference's test matrices are written as plain decimals
        std::ofstream os(path);
        os << "-1.33108 -7.78814 -0.01779 573.81219\n-0.96645 -0.02576 7.50237 344.55353\n-0.00225 0 0 1\n";
    }
    const std::vector<double> r = ReadPTMFile(path, 1);
    CHECK(r[0] == -1.33108 && r[3] == 573.81219 && r[8] == -0.00225 && r[11] == 1.0);
    unlink(path.c_str());
    rmdir(dir.c_str());
}

static void IndexedFileNames()
{
    CHECK(IndexedFileName("/data/scan", "Proj_%d.prj", 7) == "/data/scan/Proj_7.prj");
    CHECK(IndexedFileName("/data/scan/", "slice_%04d.slc", 12) == "/data/scan/slice_0012.slc");
    CHECK(IndexedFileName("", "a%%b_%u", 3) == "a%b_3");
    for (const char* bad : {"Proj.prj", "Proj_%s.prj", "P_%d_%d", "P_%-3d", "P_%"}) {
        bool threw = false;
        try {
            IndexedFileName("d", bad, 1);
        } catch (Exception&) {
            threw = true;
        }
        CHECK(threw);
    }
}

class U16Collector : public SequentialAgentBase {
public:
    U16Collector() : SequentialAgentBase("Collector", 1) {}
    std::vector<unsigned> first;
    std::vector<size_t> shotEnds;

private:
    void SetPipesImpl() override
    {
        in_ = BranchPipeFromTrunk<UINT16Vec>("Proj");
        in_->SetConsumer(GetAgentName(), 1, 1, 0);
    }
    void WorkFlow0() override
    {
        try {
            for (;;) {
                auto r = in_->GetReadToken();
                first.push_back(r.GetBuffer(0)[0]);
                if (r.IsShotEnd()) shotEnds.push_back(first.size());
            }
        } catch (PipeClosedAndEmptySignal&) {
        }
    }
    std::shared_ptr<Pipe<UINT16Vec>> in_;
};

// InputAgent.cpp:47-85: view v, then its conjugate (v + N/2) % N; one shot
// per pass, the pipe closed after the last pass.
static void InputAgentReadsFramesAndPairsConjugates()
{
    const std::string dir = TempDir();
    const size_t N = 6, U = 4, V = 3;
    for (size_t v = 0; v < N; ++v) {
        std::vector<uint16_t> fr(U * V, (uint16_t)(100 + v));
        std::ofstream os(IndexedFileName(dir, "Proj_%d.prj", v), std::ios::binary);
        os.write(reinterpret_cast<const char*>(fr.data()), (std::streamsize)(fr.size() * 2));
    }
    for (bool pairs : {false, true}) {
        auto in = std::make_shared<InputAgent>(U, V, N, 2, dir, "Proj_%d.prj", pairs);
        auto col = std::make_shared<U16Collector>();
        RunGraph({in, col});
        CHECK(!in->Failed() && !col->Failed());
        std::vector<unsigned> expect;
        for (int pass = 0; pass < 2; ++pass)
            for (size_t v = 0; v < N; ++v) {
                expect.push_back(100 + (unsigned)v);
                if (pairs) expect.push_back(100 + (unsigned)((v + N / 2) % N));
            }
        CHECK(col->first == expect);
        const size_t per = pairs ? 2 * N : N;
        CHECK(col->shotEnds == std::vector<size_t>({per, 2 * per}));
    }
    // a missing frame fails the agent and shuts the graph down
    unlink(IndexedFileName(dir, "Proj_%d.prj", 4).c_str());
    auto in = std::make_shared<InputAgent>(U, V, N, 1, dir, "Proj_%d.prj", false);
    auto col = std::make_shared<U16Collector>();
    RunGraph({in, col});
    CHECK(in->Failed());
    for (size_t v = 0; v < N; ++v) unlink(IndexedFileName(dir, "Proj_%d.prj", v).c_str());
    rmdir(dir.c_str());
}

class SliceSource : public SequentialAgentBase {
public:
    SliceSource(size_t n, size_t len) : SequentialAgentBase("SliceSource", 1), n_(n), len_(len) {}

private:
    void SetPipesImpl() override
    {
        out_ = CreateNewPipe<FloatVec>("Slice");
        out_->SetTemplate(FloatVec(len_), {len_});
        out_->SetProducer(GetAgentName(), 1, 1);
        MergePipeToTrunk(out_);
    }
    void WorkFlow0() override
    {
        for (size_t i = 0; i < n_; ++i) {
            auto w = out_->GetWriteToken(1, i + 1 == n_);
            for (size_t k = 0; k < len_; ++k) w.GetBuffer(0)[k] = (float)(i * 1000 + k);
        }
        out_->Close();
    }
    size_t n_, len_;
    std::shared_ptr<Pipe<FloatVec>> out_;
};

// OutputAgent.cpp:53-76: slice i -> <folder>/<template % i>, X * Y floats.
static void OutputAgentWritesSliceFiles()
{
    const std::string dir = TempDir() + "/out";
    auto src = std::make_shared<SliceSource>(5, 6);
    auto out = std::make_shared<OutputAgent>(3, 2, dir, "slice_%d.slc");
    RunGraph({src, out});
    CHECK(!out->Failed() && out->SlicesWritten() == 5);
    for (size_t i = 0; i < 5; ++i) {
        const std::string name = IndexedFileName(dir, "slice_%d.slc", i);
        const std::vector<char> b = Slurp(name);
        CHECK(b.size() == 6 * sizeof(float));
        const float* f = reinterpret_cast<const float*>(b.data());
        for (size_t k = 0; k < 6; ++k) CHECK(f[k] == (float)(i * 1000 + k));
        unlink(name.c_str());
    }
    rmdir(dir.c_str());
    rmdir(dir.substr(0, dir.size() - 4).c_str());
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
    Run("PipeDumpAndLoad", PipeDumpAndLoad);
    Run("LoadFileTooShortFailsTheReader", LoadFileTooShortFailsTheReader);
    Run("PTMFileRoundTrip", PTMFileRoundTrip);
    Run("IndexedFileNames", IndexedFileNames);
    Run("InputAgentReadsFramesAndPairsConjugates", InputAgentReadsFramesAndPairsConjugates);
    Run("OutputAgentWritesSliceFiles", OutputAgentWritesSliceFiles);
    std::printf("%d failed\n", g_failed);
    return g_failed ? 1 : 0;
}
