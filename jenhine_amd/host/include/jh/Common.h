// Common utilities of the agent/pipe host: exception type, logging, timer,
// blocking queue.  Interfaces follow the reference's Common/ module
// (Common/Exception.h:12-41, GLog.h:8-12, Timer.h:12-26, SyncQueue.h:15-113).
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

namespace jh {

typedef std::vector<float> FloatVec;
typedef std::vector<unsigned short> UINT16Vec;

// Library exception; what() returns the message.
class Exception : public std::exception {
public:
    explicit Exception(std::string msg) : msg_(std::move(msg)) {}
    const char* what() const noexcept override { return msg_.c_str(); }

private:
    std::string msg_;
};

// Thread-safe log line to stderr (and to a log file when JH_LOG_FILE is set).
void GLog(const std::string& msg);
// Debug line, printed only when JH_LOG_DEBUG is set.
void GLogDebug(const std::string& msg);

[[noreturn]] void ThrowException(const std::string& msg);
[[noreturn]] void ThrowExceptionAndLog(const std::string& msg);

// Wall-clock stopwatch in seconds (Common/Timer.h:12-26).
class Timer {
public:
    Timer() { Tic(); }
    void Tic() { t0_ = std::chrono::steady_clock::now(); }
    double Toc() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

private:
    std::chrono::steady_clock::time_point t0_;
};

struct SyncQueueClosedAndEmptySignal {};

// Blocking FIFO: Pop waits while empty, Push waits while full; after Close,
// Pop drains what is left then throws SyncQueueClosedAndEmptySignal and Push
// throws at once.
template <typename T>
class SyncQueue {
public:
    explicit SyncQueue(size_t maxSize = SIZE_MAX) : maxSize_(maxSize) {}

    T Pop()
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) throw SyncQueueClosedAndEmptySignal();
        T v = std::move(q_.front());
        q_.pop_front();
        cv_.notify_all();
        return v;
    }

    void Push(T v)
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return q_.size() < maxSize_ || closed_; });
        if (closed_) throw SyncQueueClosedAndEmptySignal();
        q_.push_back(std::move(v));
        cv_.notify_all();
    }

    size_t Size()
    {
        std::lock_guard<std::mutex> lk(m_);
        return q_.size();
    }

    bool Empty() { return Size() == 0; }

    void Close()
    {
        std::lock_guard<std::mutex> lk(m_);
        closed_ = true;
        cv_.notify_all();
    }

private:
    const size_t maxSize_;
    bool closed_ = false;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<T> q_;
};

// Counting semaphore (Common/Semaphore.h:21-61): Signal adds, Wait(n) blocks
// until the count reaches n and consumes it.
class Semaphore {
public:
    void Signal(size_t n = 1)
    {
        std::lock_guard<std::mutex> lk(m_);
        count_ += n;
        cv_.notify_all();
    }
    void Wait(size_t n = 1)
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return count_ >= n; });
        count_ -= n;
    }

private:
    size_t count_ = 0;
    std::mutex m_;
    std::condition_variable cv_;
};

}  // namespace jh
