// Agents: the working units of a pipeline.  Interface of the reference's
// Pipeline/AgentBase.h:22-168, SequentialAgentBase.h:12-48 and
// ConcurrentAgentBase.h:12-66:
//   AgentBase::SetPipes(trunk) -> SetPipesImpl(); GetReady() (allocation,
//   main thread); Start(); Join(); BranchPipeFromTrunk<T>(name),
//   CreateNewPipe<T>(name), MergePipeToTrunk(pipe); ThreadOccupiedScope /
//   WaitAsyncScope time accounting.
//   SequentialAgentBase: WorkFlow0..WorkFlow9, one thread each.
//   ConcurrentAgentBase: ManagerWorkFlow() + ProcessTask(thread, task) on N
//   workers, WaitIdleWorkerIndex() / SubmitTask().
// Any exception leaving a work flow is logged and closes every pipe the agent
// is connected to, which shuts the graph down (AgentBase.cpp:25-36).
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "Common.h"
#include "Pipe.h"

namespace jh {

class AgentBase {
public:
    explicit AgentBase(const std::string& name) : name_(name) {}
    virtual ~AgentBase() = default;
    AgentBase(const AgentBase&) = delete;
    AgentBase& operator=(const AgentBase&) = delete;

    // Takes the pipes this agent reads or passes by; leaves in `pipes` the
    // ones it writes or passes by.
    void SetPipes(std::vector<std::shared_ptr<PipeBase>>& pipes)
    {
        trunk_ = pipes;
        SetPipesImpl();
        pipes = trunk_;
    }

    virtual void GetReady() {}
    virtual void Start() = 0;
    virtual void Join() = 0;
    virtual size_t GetNumThreads() const = 0;

    std::string GetAgentName() const { return name_; }
    // Every pipe this agent reads, writes or creates (the main program
    // collects them to list the graph and to set dump/load files,
    // JEngine.cpp:320-400).
    const std::set<std::shared_ptr<PipeBase>>& GetConnectedPipes() const { return connected_; }
    // True once a work flow of this agent ended with an exception.
    bool Failed() const { return failed_.load(); }
    double GetOccupiedTime() const { return occupied_.load(); }
    double GetAsyncTime() const { return async_.load(); }
    double GetMaxOccupiedTime() const { return maxOccupied_.load(); }
    double GetAverageOccupiedSpan() const { return count_ ? occupied_.load() / (double)count_ : 0.0; }

    class ThreadOccupiedScope;
    class WaitAsyncScope;

protected:
    template <typename T>
    std::shared_ptr<Pipe<T>> BranchPipeFromTrunk(const std::string& name)
    {
        for (auto it = trunk_.begin(); it != trunk_.end(); ++it) {
            if ((*it)->GetName() == name) {
                auto p = std::dynamic_pointer_cast<Pipe<T>>(*it);
                if (!p) ThrowExceptionAndLog(name_ + ": pipe " + name + " has another element type");
                connected_.insert(*it);
                trunk_.erase(it);
                return p;
            }
        }
        ThrowExceptionAndLog(name_ + ": Pipe trunk does not contain a pipe named " + name);
    }

    template <typename T>
    std::shared_ptr<Pipe<T>> CreateNewPipe(const std::string& name)
    {
        auto p = std::make_shared<Pipe<T>>(name);
        connected_.insert(p);
        return p;
    }

    template <typename T>
    void MergePipeToTrunk(std::shared_ptr<Pipe<T>> branch)
    {
        if (!connected_.count(branch)) ThrowExceptionAndLog("Please use CreateNewPipe<T>() to create a pipe");
        trunk_.push_back(branch);
    }

    void CloseAllConnectedPipes();
    void MarkFailed() { failed_ = true; }
    void AddOccupied(double span);
    void AddAsync(double span) { async_ = async_.load() + span; }

    std::set<std::shared_ptr<PipeBase>> connected_;

private:
    virtual void SetPipesImpl() = 0;

    const std::string name_;
    std::vector<std::shared_ptr<PipeBase>> trunk_;
    std::mutex closeMutex_, statMutex_;
    std::atomic<double> occupied_{0.0}, maxOccupied_{0.0}, async_{0.0};
    std::atomic<size_t> count_{0};
    std::atomic<bool> failed_{false};
};

// Claims a host thread for its lifetime; the span is added to the agent's
// occupied time.  A WaitAsyncScope inside it splits off time spent waiting on
// asynchronous work (device, file) (AgentBase.cpp:38-78).
class AgentBase::ThreadOccupiedScope {
public:
    explicit ThreadOccupiedScope(AgentBase* owner) : owner_(owner) {}
    ~ThreadOccupiedScope() { owner_->AddOccupied(t_.Toc()); }

private:
    friend class WaitAsyncScope;
    AgentBase* owner_;
    Timer t_;
};

class AgentBase::WaitAsyncScope {
public:
    explicit WaitAsyncScope(ThreadOccupiedScope* occ) : occ_(occ)
    {
        occ_->owner_->AddOccupied(occ_->t_.Toc());
        t_.Tic();
    }
    ~WaitAsyncScope()
    {
        occ_->owner_->AddAsync(t_.Toc());
        occ_->t_.Tic();
    }

private:
    ThreadOccupiedScope* occ_;
    Timer t_;
};

class SequentialAgentBase : public AgentBase {
public:
    SequentialAgentBase(const std::string& name, int numSteps);
    ~SequentialAgentBase() override;
    void Start() final;
    void Join() final;
    size_t GetNumThreads() const final { return (size_t)numSteps_; }

protected:
    virtual void WorkFlow0() {}
    virtual void WorkFlow1() {}
    virtual void WorkFlow2() {}
    virtual void WorkFlow3() {}
    virtual void WorkFlow4() {}
    virtual void WorkFlow5() {}
    virtual void WorkFlow6() {}
    virtual void WorkFlow7() {}
    virtual void WorkFlow8() {}
    virtual void WorkFlow9() {}

private:
    void Run(int index);
    const int numSteps_;
    std::vector<std::thread> threads_;
};

class ConcurrentAgentBase : public AgentBase {
public:
    ConcurrentAgentBase(const std::string& name, size_t numThreads);
    ~ConcurrentAgentBase() override;
    void Start() final;
    void Join() final;
    size_t GetNumThreads() const final { return numWorkers_; }

    struct TaskBase {
        virtual ~TaskBase() = default;
    };

protected:
    virtual void ManagerWorkFlow() = 0;
    size_t WaitIdleWorkerIndex() { return idle_.Pop(); }
    void SubmitTask(std::unique_ptr<TaskBase>&& t) { tasks_.Push(std::move(t)); }

private:
    virtual void ProcessTask(size_t threadIndex, std::unique_ptr<TaskBase> task) = 0;
    void RunManager();
    void RunWorker(size_t index);

    const size_t numWorkers_;
    std::thread manager_;
    std::vector<std::thread> workers_;
    SyncQueue<std::unique_ptr<TaskBase>> tasks_;
    SyncQueue<size_t> idle_;
};

}  // namespace jh
