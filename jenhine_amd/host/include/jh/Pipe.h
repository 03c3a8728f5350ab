// Pipe<T>: bounded multi-producer / multi-consumer ring of elements between
// agents, with shot framing and read overlap, handed out as RAII tokens.
//
// Interface of the reference's Pipeline/Pipe.h:19-66, PipeBase.h:19-154 and
// the tokens of PipeReadToken.h / PipeWriteToken.h:
//   SetTemplate(T&&, dims), SetProducer(name, nWriters, writeSize),
//   SetConsumer(name, nReaders, readSize incl. overlap, overlap),
//   SetBufferSize, GetWriteToken(size, isShotEnd), GetReadToken(), Close().
// Semantics kept from the reference (PipeBase.cpp:147-331): positions are
// global element indices; a write token waits until the ring has room both
// behind unfinished writes and behind the oldest element a reader may still
// need; a read token covers readSize elements (fewer at a shot end) and
// repeats the last `overlap` elements of the previous token of the same shot;
// tokens are released in order, when the last copy of a token is destroyed;
// GetReadToken throws PipeClosedAndEmptySignal once the pipe is closed and
// drained; a writer that can never make progress gets "Pipe deadlock".
// Differences: tokens notify the pipe from their shared control block
// instead of being polled by use count; a closed pipe whose last shot was
// never finished drains to PipeClosedAndEmptySignal instead of blocking.
//
// Dump and load (Pipeline/PipeBase.h:54-56, PipeBase.cpp:263-268,310-313,
// PipeDumpLoadImpl.cpp:15-57): SetDump streams every element to a file as
// its write finishes, in stream order; SetLoad overwrites the new (non-
// overlap) elements of every read token with the next elements of a file
// before the reader sees them.  The file holds the elements' payload bytes
// back to back (PipeElementIO<T>), so a stream dumped from one run replays in
// another.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "Common.h"

namespace jh {

struct PipeClosedAndEmptySignal {};

// Payload I/O of one pipe element for dump and load files.  Provided for
// vectors of trivially copyable values (FloatVec, UINT16Vec, ...); other
// element types specialise it (the device buffers do, in HipAgents.h).
template <typename T, typename = void>
struct PipeElementIO {
    static constexpr bool kSupported = false;
    static void Write(std::ostream&, const T&) {}
    static void Read(std::istream&, T&) {}
};

template <typename U>
struct PipeElementIO<std::vector<U>, std::enable_if_t<std::is_trivially_copyable<U>::value>> {
    static constexpr bool kSupported = true;
    static void Write(std::ostream& os, const std::vector<U>& v)
    {
        os.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(U)));
    }
    static void Read(std::istream& is, std::vector<U>& v)
    {
        is.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(U)));
    }
};

// How a pipe fills its ring from the element template: copies by default.
// Element types may specialise it (the device buffers allocate the ring's
// slots from one contiguous allocation, HipAgents.h).
template <typename T, typename = void>
struct PipeRing {
    static void Fill(std::vector<T>& buffers, size_t n, const T& tmpl) { buffers.assign(n, tmpl); }
};

class PipeBase {
public:
    explicit PipeBase(std::string name);
    virtual ~PipeBase();
    PipeBase(const PipeBase&) = delete;
    PipeBase& operator=(const PipeBase&) = delete;

    void SetConsumer(const std::string& name, size_t number, size_t sizeIncludingOverlap, size_t overlapSize);
    void SetProducer(const std::string& name, size_t number, size_t size);
    void SetBufferSize(size_t size);
    // The default ring size is rounded up to a multiple of `granule` (an
    // extension: consumers that process whole chunks of a contiguous ring
    // need chunk-aligned tokens never to wrap).
    void SetBufferGranule(size_t granule);

    const std::string& GetName() const { return name_; }
    const std::string& GetConsumer() const { return consumer_; }
    const std::string& GetProducer() const { return producer_; }
    size_t GetNumReaders() const { return numReaders_; }
    size_t GetReadSize() const { return readSize_; }
    size_t GetOverlapSize() const { return overlap_; }
    size_t GetNumWriters() const { return numWriters_; }
    size_t GetWriteSize() const { return writeSize_; }
    size_t GetBufferSize() const { return bufferSize_; }
    std::vector<size_t> GetElementShape() const { return shape_; }

    size_t GetWritingPos();
    size_t GetWriteDonePos();
    size_t GetReadingPos();
    size_t GetReadDonePos();

    void Close();
    bool Closed();
    double GetProducerBlockedTime() const { return blocked_; }

    // Dump every finished write to / load every read from a file (see above).
    // Set before the graph starts; each at most once.
    virtual void SetDump(const std::string& path) = 0;
    virtual void SetLoad(const std::string& path) = 0;

    // Token bookkeeping shared by the typed tokens.
    struct Span {
        size_t start = 0, size = 0, overlap = 0;
        bool shotStart = false, shotEnd = false;
    };
    struct TokenState {
        PipeBase* pipe;
        bool write;
        uint64_t id;
        Span span;
        ~TokenState() { pipe->Release(write, id); }
    };

protected:
    void SetElementShape(const std::vector<size_t>& d) { shape_ = d; }
    std::shared_ptr<TokenState> AcquireWrite(size_t size, bool isShotEnd);
    std::shared_ptr<TokenState> AcquireRead();
    // Called (under the lock) once producer, consumer and template are known.
    virtual void InitBuffers() = 0;
    virtual bool HasTemplate() const = 0;
    // Element payload of global positions [start, end) to the dump file /
    // from the load file; called under the lock.
    virtual void DumpElements(size_t start, size_t end) = 0;
    virtual void LoadElements(size_t start, size_t end) = 0;
    void OpenDump(const std::string& path);
    void OpenLoad(const std::string& path);
    void TryInit();
    size_t bufferSize_ = 0;
    size_t granule_ = 1;
    std::unique_ptr<std::ofstream> dump_;
    std::unique_ptr<std::ifstream> load_;

private:
    friend struct TokenState;
    void Release(bool write, uint64_t id);
    size_t ReadStart() const;
    size_t ReadEnd() const;
    bool DataReady() const;
    bool RoomToWrite(size_t size);

    struct Slot {
        uint64_t id;
        size_t start, size;
        bool released;
    };
    struct Shot {
        size_t start, end;
    };

    const std::string name_;
    std::string consumer_, producer_;
    size_t numWriters_ = 0, writeSize_ = 0, numReaders_ = 0, readSize_ = 0, overlap_ = 0;
    std::vector<size_t> shape_;
    bool ready_ = false;
    bool closed_ = false;
    size_t writingPos_ = 0, writeDonePos_ = 0, readingPos_ = 0, readDonePos_ = 0;
    uint64_t nextId_ = 1;
    std::deque<Slot> writes_, reads_;
    std::deque<Shot> shots_;
    double blocked_ = 0.0;
    std::string ioError_;  // a failed dump (raised by the next token request)
    std::mutex m_;
    std::condition_variable cv_;
};

template <typename T>
class PipeWriteToken {
public:
    PipeWriteToken() = default;
    PipeWriteToken(std::shared_ptr<PipeBase::TokenState> st, T* const* ptrs) : st_(std::move(st)), ptrs_(ptrs) {}
    PipeWriteToken(const PipeWriteToken&) = default;
    PipeWriteToken(PipeWriteToken&&) = default;
    PipeWriteToken& operator=(const PipeWriteToken& o)
    {
        if (st_) ThrowException("WriteToken is already assigned.");
        st_ = o.st_;
        ptrs_ = o.ptrs_;
        return *this;
    }
    bool Null() const { return !st_; }
    size_t GetStartIndex() const { return st_->span.start; }
    size_t GetSize() const { return st_->span.size; }
    bool IsShotStart() const { return st_->span.shotStart; }
    bool IsShotEnd() const { return st_->span.shotEnd; }
    T* const* GetBufferPtrs() const { return ptrs_; }
    T& GetBuffer(size_t i) const { return *ptrs_[i]; }

private:
    std::shared_ptr<PipeBase::TokenState> st_;
    T* const* ptrs_ = nullptr;
};

template <typename T>
class PipeReadToken {
public:
    PipeReadToken() = default;
    PipeReadToken(std::shared_ptr<PipeBase::TokenState> st, T* const* ptrs) : st_(std::move(st)), ptrs_(ptrs) {}
    PipeReadToken(const PipeReadToken&) = default;
    PipeReadToken(PipeReadToken&&) = default;
    PipeReadToken& operator=(const PipeReadToken& o)
    {
        if (st_) ThrowException("ReadToken is already assigned.");
        st_ = o.st_;
        ptrs_ = o.ptrs_;
        return *this;
    }
    void Reset() { st_.reset(); ptrs_ = nullptr; }
    bool Null() const { return !st_; }
    size_t GetStartIndex() const { return st_->span.start; }
    size_t GetSize() const { return st_->span.size; }
    size_t GetOverlapSize() const { return st_->span.overlap; }
    bool IsShotStart() const { return st_->span.shotStart; }
    bool IsShotEnd() const { return st_->span.shotEnd; }
    const T* const* GetBufferPtrs() const { return ptrs_; }
    const T& GetBuffer(size_t i) const { return *ptrs_[i]; }
    T& GetMutableBuffer(size_t i) const { return *ptrs_[i]; }

private:
    std::shared_ptr<PipeBase::TokenState> st_;
    T* const* ptrs_ = nullptr;
};

template <typename T>
class Pipe : public PipeBase {
public:
    typedef PipeWriteToken<T> WriteToken;
    typedef PipeReadToken<T> ReadToken;

    explicit Pipe(const std::string& name) : PipeBase(name) {}

    void SetTemplate(T&& elementTemplate, std::vector<size_t>&& dims)
    {
        if (!buffers_.empty() || hasTemplate_) ThrowExceptionAndLog("Pipe: buffer is already initialized.");
        template_.push_back(std::move(elementTemplate));
        hasTemplate_ = true;
        SetElementShape(dims);
        TryInit();
    }

    T GetTemplate() const
    {
        if (template_.empty()) ThrowException("Pipe: pipe is not initialized.");
        return template_[0];
    }

    WriteToken GetWriteToken(size_t size, bool isShotEnd)
    {
        auto st = AcquireWrite(size, isShotEnd);
        return WriteToken(st, ptrs_.data() + st->span.start % bufferSize_);
    }

    ReadToken GetReadToken()
    {
        auto st = AcquireRead();
        return ReadToken(st, ptrs_.data() + st->span.start % bufferSize_);
    }

    void SetDump(const std::string& path) override
    {
        if (!PipeElementIO<T>::kSupported) ThrowExceptionAndLog(GetName() + ": this element type cannot be dumped");
        OpenDump(path);
    }

    void SetLoad(const std::string& path) override
    {
        if (!PipeElementIO<T>::kSupported) ThrowExceptionAndLog(GetName() + ": this element type cannot be loaded");
        OpenLoad(path);
    }

protected:
    bool HasTemplate() const override { return hasTemplate_; }
    void DumpElements(size_t start, size_t end) override
    {
        for (; start != end; ++start) PipeElementIO<T>::Write(*dump_, buffers_[start % bufferSize_]);
        dump_->flush();
        if (!*dump_) ThrowException(GetName() + ": writing the dump file failed");
    }
    void LoadElements(size_t start, size_t end) override
    {
        for (; start != end; ++start) {
            PipeElementIO<T>::Read(*load_, buffers_[start % bufferSize_]);
            if (!*load_) ThrowException(GetName() + ": the load file ended early");
        }
    }
    void InitBuffers() override
    {
        PipeRing<T>::Fill(buffers_, bufferSize_, template_[0]);
        // pointers repeated once so that a token crossing the end of the
        // ring still sees contiguous element pointers
        ptrs_.resize(2 * bufferSize_ - 1);
        for (size_t i = 0; i < ptrs_.size(); ++i) ptrs_[i] = &buffers_[i % bufferSize_];
    }

private:
    bool hasTemplate_ = false;
    std::vector<T> template_;
    std::vector<T> buffers_;
    std::vector<T*> ptrs_;
};

}  // namespace jh
