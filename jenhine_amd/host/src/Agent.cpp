// Agent bases and common utilities of the host.
#include "jh/Agent.h"

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>

namespace jh {

// ---------------------------------------------------------------------------
// logging and exceptions
// ---------------------------------------------------------------------------

static std::mutex g_logMutex;

static void WriteLog(const std::string& line)
{
    std::lock_guard<std::mutex> lk(g_logMutex);
    std::time_t t = std::time(nullptr);
    char ts[32];
    std::strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", std::localtime(&t));
    std::cerr << "[" << ts << "] " << line << std::endl;
    if (const char* f = std::getenv("JH_LOG_FILE")) {
        std::ofstream os(f, std::ios::app);
        os << "[" << ts << "] " << line << "\n";
    }
}

void GLog(const std::string& msg) { WriteLog(msg); }

void GLogDebug(const std::string& msg)
{
    if (std::getenv("JH_LOG_DEBUG")) WriteLog("[DEBUG] " + msg);
}

void ThrowException(const std::string& msg) { throw Exception(msg); }

void ThrowExceptionAndLog(const std::string& msg)
{
    GLog(msg);
    throw Exception(msg);
}

// ---------------------------------------------------------------------------
// AgentBase
// ---------------------------------------------------------------------------

void AgentBase::CloseAllConnectedPipes()
{
    std::lock_guard<std::mutex> lk(closeMutex_);
    for (auto& p : connected_) {
        if (!p->Closed()) {
            GLogDebug(name_ + ": " + p->GetName() + " is force-closed");
            p->Close();
        }
    }
}

void AgentBase::AddOccupied(double span)
{
    if (span <= 1e-5) return;
    std::lock_guard<std::mutex> lk(statMutex_);
    occupied_ = occupied_.load() + span;
    if (span > maxOccupied_.load()) maxOccupied_ = span;
    ++count_;
}

// ---------------------------------------------------------------------------
// SequentialAgentBase
// ---------------------------------------------------------------------------

SequentialAgentBase::SequentialAgentBase(const std::string& name, int numSteps) : AgentBase(name), numSteps_(numSteps)
{
    if (numSteps < 1 || numSteps > 10) ThrowException("SequentialAgentBase holds 1 to 10 steps");
}

SequentialAgentBase::~SequentialAgentBase()
{
    for (auto& t : threads_)
        if (t.joinable()) t.join();
}

void SequentialAgentBase::Start()
{
    for (int i = 0; i < numSteps_; ++i) threads_.emplace_back(&SequentialAgentBase::Run, this, i);
}

void SequentialAgentBase::Join()
{
    for (auto& t : threads_)
        if (t.joinable()) t.join();
    CloseAllConnectedPipes();
}

void SequentialAgentBase::Run(int index)
{
    try {
        switch (index) {
        case 0: WorkFlow0(); break;
        case 1: WorkFlow1(); break;
        case 2: WorkFlow2(); break;
        case 3: WorkFlow3(); break;
        case 4: WorkFlow4(); break;
        case 5: WorkFlow5(); break;
        case 6: WorkFlow6(); break;
        case 7: WorkFlow7(); break;
        case 8: WorkFlow8(); break;
        default: WorkFlow9(); break;
        }
        return;
    } catch (PipeClosedAndEmptySignal&) {
        // input drained: a normal end
    } catch (std::exception& e) {
        MarkFailed();
        GLog(GetAgentName() + " step #" + std::to_string(index) + " died: " + e.what());
    } catch (...) {
        MarkFailed();
        GLog(GetAgentName() + " step #" + std::to_string(index) + " died: unknown exception");
    }
    CloseAllConnectedPipes();
}

// ---------------------------------------------------------------------------
// ConcurrentAgentBase
// ---------------------------------------------------------------------------

ConcurrentAgentBase::ConcurrentAgentBase(const std::string& name, size_t numThreads)
    : AgentBase(name), numWorkers_(numThreads)
{
    if (numThreads == 0) ThrowException("ConcurrentAgentBase needs at least one worker");
}

ConcurrentAgentBase::~ConcurrentAgentBase()
{
    tasks_.Close();
    idle_.Close();
    if (manager_.joinable()) manager_.join();
    for (auto& t : workers_)
        if (t.joinable()) t.join();
}

void ConcurrentAgentBase::Start()
{
    manager_ = std::thread(&ConcurrentAgentBase::RunManager, this);
    for (size_t i = 0; i < numWorkers_; ++i) workers_.emplace_back(&ConcurrentAgentBase::RunWorker, this, i);
}

void ConcurrentAgentBase::Join()
{
    if (manager_.joinable()) manager_.join();
    tasks_.Close();
    for (auto& t : workers_)
        if (t.joinable()) t.join();
    CloseAllConnectedPipes();
}

void ConcurrentAgentBase::RunManager()
{
    try {
        ManagerWorkFlow();
    } catch (PipeClosedAndEmptySignal&) {
    } catch (SyncQueueClosedAndEmptySignal&) {
        GLog(GetAgentName() + ": workers have quit");
    } catch (std::exception& e) {
        MarkFailed();
        GLog(GetAgentName() + " manager died: " + e.what());
    } catch (...) {
        MarkFailed();
        GLog(GetAgentName() + " manager died: unknown exception");
    }
    idle_.Close();
    tasks_.Close();
    CloseAllConnectedPipes();
}

void ConcurrentAgentBase::RunWorker(size_t index)
{
    try {
        for (;;) {
            idle_.Push(index);
            std::unique_ptr<TaskBase> t = tasks_.Pop();
            ProcessTask(index, std::move(t));
        }
    } catch (SyncQueueClosedAndEmptySignal&) {
        return;  // normal end
    } catch (std::exception& e) {
        MarkFailed();
        GLog(GetAgentName() + " #" + std::to_string(index) + " died: " + e.what());
    } catch (...) {
        MarkFailed();
        GLog(GetAgentName() + " #" + std::to_string(index) + " died: unknown exception");
    }
    idle_.Close();
    tasks_.Close();
    CloseAllConnectedPipes();
}

}  // namespace jh
