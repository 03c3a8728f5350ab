// Ring bookkeeping of PipeBase (see Pipe.h for the semantics and their
// reference: Pipeline/PipeBase.cpp:147-331).
#include "jh/Pipe.h"

#include <algorithm>

namespace jh {

PipeBase::PipeBase(std::string name) : name_(std::move(name)) {}

PipeBase::~PipeBase()
{
    std::unique_lock<std::mutex> lk(m_);
    cv_.wait(lk, [&] { return writes_.empty() && reads_.empty(); });
    GLogDebug(name_ + " pipe destructed; " + producer_ + " was blocked by " + consumer_ + " for " +
              std::to_string(blocked_) + " s.");
}

void PipeBase::SetConsumer(const std::string& name, size_t number, size_t sizeIncludingOverlap, size_t overlapSize)
{
    {
        std::lock_guard<std::mutex> lk(m_);
        if (numReaders_ * readSize_ != 0) ThrowExceptionAndLog(name_ + ": consumer is already set.");
        if (number == 0) ThrowExceptionAndLog(name_ + ": number of readers can not be 0.");
        if (sizeIncludingOverlap == 0) ThrowExceptionAndLog(name_ + ": read size can not be 0.");
        if (overlapSize >= sizeIncludingOverlap) ThrowExceptionAndLog(name_ + ": overlap must be smaller than the read size.");
        consumer_ = name;
        numReaders_ = number;
        readSize_ = sizeIncludingOverlap;
        overlap_ = overlapSize;
    }
    TryInit();
}

void PipeBase::SetProducer(const std::string& name, size_t number, size_t size)
{
    {
        std::lock_guard<std::mutex> lk(m_);
        if (numWriters_ * writeSize_ != 0) ThrowExceptionAndLog(name_ + ": producer is already set.");
        if (number == 0) ThrowExceptionAndLog(name_ + ": number of writers can not be 0.");
        if (size == 0) ThrowExceptionAndLog(name_ + ": write size can not be 0.");
        producer_ = name;
        numWriters_ = number;
        writeSize_ = size;
    }
    TryInit();
}

void PipeBase::SetBufferSize(size_t size)
{
    {
        std::lock_guard<std::mutex> lk(m_);
        if (bufferSize_ != 0) ThrowExceptionAndLog(name_ + ": pipe size is already set.");
        if (size == 0) ThrowExceptionAndLog(name_ + ": pipe size can not be 0.");
        bufferSize_ = size;
    }
    TryInit();
}

void PipeBase::SetBufferGranule(size_t granule)
{
    {
        std::lock_guard<std::mutex> lk(m_);
        if (ready_) ThrowExceptionAndLog(name_ + ": pipe is already initialized.");
        if (granule == 0) ThrowExceptionAndLog(name_ + ": granule can not be 0.");
        granule_ = granule;
    }
    TryInit();
}

void PipeBase::TryInit()
{
    std::lock_guard<std::mutex> lk(m_);
    if (ready_ || numReaders_ == 0 || numWriters_ == 0 || !HasTemplate()) return;
    // default: every writer and reader can hold a token at once
    // (Pipeline/PipeImpl.h:84-89)
    if (bufferSize_ == 0) {
        bufferSize_ = numWriters_ * writeSize_ + numReaders_ * (readSize_ - overlap_) + overlap_;
        bufferSize_ = (bufferSize_ + granule_ - 1) / granule_ * granule_;
    }
    InitBuffers();
    ready_ = true;
    cv_.notify_all();
}

void PipeBase::OpenDump(const std::string& path)
{
    std::lock_guard<std::mutex> lk(m_);
    if (dump_) ThrowExceptionAndLog(name_ + ": dump file is already set.");
    dump_.reset(new std::ofstream(path, std::ios::binary));
    if (!*dump_) {
        dump_.reset();
        ThrowExceptionAndLog(name_ + ": can not open dump file " + path);
    }
}

void PipeBase::OpenLoad(const std::string& path)
{
    std::lock_guard<std::mutex> lk(m_);
    if (load_) ThrowExceptionAndLog(name_ + ": load file is already set.");
    load_.reset(new std::ifstream(path, std::ios::binary));
    if (!*load_) {
        load_.reset();
        ThrowExceptionAndLog(name_ + ": can not open file: " + path);
    }
}

size_t PipeBase::GetWritingPos() { std::lock_guard<std::mutex> lk(m_); return writingPos_; }
size_t PipeBase::GetWriteDonePos() { std::lock_guard<std::mutex> lk(m_); return writeDonePos_; }
size_t PipeBase::GetReadingPos() { std::lock_guard<std::mutex> lk(m_); return readingPos_; }
size_t PipeBase::GetReadDonePos() { std::lock_guard<std::mutex> lk(m_); return readDonePos_; }

void PipeBase::Close()
{
    std::lock_guard<std::mutex> lk(m_);
    if (!closed_) {
        closed_ = true;
        GLogDebug(name_ + " pipe is closed.");
        cv_.notify_all();
    }
}

bool PipeBase::Closed()
{
    std::lock_guard<std::mutex> lk(m_);
    return closed_;
}

size_t PipeBase::ReadStart() const
{
    return std::max(shots_.front().start, readingPos_ > overlap_ ? readingPos_ - overlap_ : (size_t)0);
}

size_t PipeBase::ReadEnd() const { return std::min(ReadStart() + readSize_, shots_.front().end); }

bool PipeBase::DataReady() const
{
    return !shots_.empty() && ReadEnd() <= std::min(writeDonePos_, readDonePos_ + bufferSize_);
}

bool PipeBase::RoomToWrite(size_t size)
{
    const size_t oldestNeeded = !reads_.empty() ? reads_.front().start : (shots_.empty() ? readingPos_ : ReadStart());
    const bool room = writingPos_ + size <= oldestNeeded + bufferSize_;
    if (!room && reads_.empty() && writes_.empty() && !DataReady()) {
        closed_ = true;
        cv_.notify_all();
        ThrowExceptionAndLog(name_ + ": Pipe deadlock");
    }
    return room;
}

std::shared_ptr<PipeBase::TokenState> PipeBase::AcquireWrite(size_t size, bool isShotEnd)
{
    std::unique_lock<std::mutex> lk(m_);
    cv_.wait(lk, [&] { return ready_ || closed_; });
    if (closed_) ThrowExceptionAndLog(name_ + ": trying to get a write token when the pipe is closed");
    if (!ioError_.empty()) ThrowExceptionAndLog(ioError_);
    if (size == 0 || size > bufferSize_) ThrowExceptionAndLog(name_ + ": write size is 0 or larger than the buffer");
    cv_.wait(lk, [&] { return closed_ || writingPos_ + size <= writeDonePos_ + bufferSize_; });
    if (!closed_ && !RoomToWrite(size)) {
        Timer t;
        cv_.wait(lk, [&] { return closed_ || RoomToWrite(size); });
        blocked_ += t.Toc();
    }
    if (closed_) ThrowExceptionAndLog(name_ + ": trying to get a write token when the pipe is closed");
    const bool shotStart = shots_.empty() || writingPos_ == shots_.back().end;
    if (shotStart) shots_.push_back(Shot{writingPos_, SIZE_MAX});
    auto st = std::make_shared<TokenState>();
    st->pipe = this;
    st->write = true;
    st->id = nextId_++;
    st->span.start = writingPos_;
    st->span.size = size;
    st->span.shotStart = shotStart;
    st->span.shotEnd = isShotEnd;
    writes_.push_back(Slot{st->id, writingPos_, size, false});
    writingPos_ += size;
    if (isShotEnd) {
        shots_.back().end = writingPos_;
        if (shots_.back().end - shots_.back().start < overlap_ + 1)
            ThrowExceptionAndLog(name_ + ": shot size is smaller than or equal to the overlap size.");
    }
    cv_.notify_all();
    return st;
}

std::shared_ptr<PipeBase::TokenState> PipeBase::AcquireRead()
{
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
        if (ready_ && DataReady()) break;
        if (closed_ && writes_.empty() && readDonePos_ == writeDonePos_ && writeDonePos_ == writingPos_)
            throw PipeClosedAndEmptySignal();
        // closed with an unfinished shot and no writer left: nothing more can come
        if (closed_ && writes_.empty() && reads_.empty()) throw PipeClosedAndEmptySignal();
        cv_.wait(lk);
    }
    if (!ioError_.empty()) ThrowExceptionAndLog(ioError_);
    const size_t start = ReadStart();
    const size_t end = ReadEnd();
    // replace the token's new elements before anyone can see them (the
    // overlap was loaded with the previous token)
    if (load_) LoadElements(readingPos_, end);
    auto st = std::make_shared<TokenState>();
    st->pipe = this;
    st->write = false;
    st->id = nextId_++;
    st->span.start = start;
    st->span.size = end - start;
    st->span.overlap = readingPos_ - start;
    st->span.shotStart = start == shots_.front().start;
    st->span.shotEnd = end == shots_.front().end;
    if (st->span.shotEnd) shots_.pop_front();
    reads_.push_back(Slot{st->id, start, end - start, false});
    readingPos_ += st->span.size - st->span.overlap;
    cv_.notify_all();
    return st;
}

void PipeBase::Release(bool write, uint64_t id)
{
    std::lock_guard<std::mutex> lk(m_);
    std::deque<Slot>& q = write ? writes_ : reads_;
    for (Slot& s : q)
        if (s.id == id) { s.released = true; break; }
    const size_t doneBefore = writeDonePos_;
    while (!q.empty() && q.front().released) {
        if (write) writeDonePos_ += q.front().size;
        else readDonePos_ = q.front().start + q.front().size;
        q.pop_front();
    }
    if (write && dump_ && writeDonePos_ != doneBefore && ioError_.empty()) {
        // runs from a token's destructor: a failure is kept and raised by the
        // next token request, and the pipe is closed
        try {
            DumpElements(doneBefore, writeDonePos_);
        } catch (std::exception& e) {
            ioError_ = e.what();
            closed_ = true;
            GLog(ioError_);
        }
    }
    cv_.notify_all();
}

}  // namespace jh
