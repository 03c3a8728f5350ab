// Device-side agents of the reconstruction pipeline on MI355X, built on the
// C-ABI of include/jhip.h.  They take the places of the reference's CUDA
// module and CPU cores in the agent graph (JEngine/JEngine.cpp:185-313):
//
//   ProjectionUploadAgent      Pipe<UINT16Vec> "Proj" -> Pipe<DeviceMemory<u16>> "Proj"
//                              (replaces CUDA/Tex2DUploadAgent.cpp: raw counts
//                              go up once, so every later stage stays on the device)
//   PreProcessingHIPAgent      DeviceMemory<u16> -> DeviceMemory<float>
//                              (replaces PreProcessingAgent/PreProcessingAgent.cpp)
//   CompositeHIPAgent          DeviceMemory<float> pairs (view, conjugate) ->
//                              DeviceMemory<float> of 2W columns
//                              (replaces CompositeAgent/CompositeAgent.cpp)
//   FilterHIPAgent             DeviceMemory<float> -> DeviceMemory<float>
//                              (replaces FilterAgent/FilterAgent.cpp)
//   BPHIPAgent                 DeviceMemory<float> "Proj" -> DeviceMemory<float> "Slice"
//                              (replaces BPCUDAAgent/BPCUDAAgent.cpp)
//   DeviceMemoryDownloadAgent  DeviceMemory<float> "Slice" -> FloatVec "Slice"
//                              (replaces CUDA/DeviceMemoryDownloadAgent.cpp)
//
// Each agent thread owns one HIP stream created on that thread and
// synchronises it before releasing a write token, so the next agent may use
// the data (the reference's ordering rule, BPCUDAAgent.cpp:240-261).  The
// device agents take chunks of views (ReconParams::ViewsPerChunk): device
// frame pipes allocate their ring as one buffer (PipeRing), so an aligned
// chunk token is one contiguous run of frames and the chunk is one launch and
// one stream synchronisation.  The copy agents (upload, download) run
// kCopyThreads threads, each with its own stream; a thread takes its read and write tokens
// together under the agent's lock, so the stream order is kept.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "jhip.h"
#include "Agent.h"
#include "Params.h"

namespace jh {

// Message prefix of a failed device allocation; the host's part-retry loop
// keys on it (the reference keys on "CUDA failed to malloc ",
// JEngine/JEngine.cpp:413-421).
extern const char* const kDeviceOOMPrefix;

// Throws jh::Exception carrying jh_last_error() when s != JH_OK.
void CheckJh(jh_status s, const char* what);
void CheckHip(hipError_t e, const char* what);

// JEngine's adjustment of the matrices read from PTM.dat, applied in FP64 on
// `device` through jh_ptm_modify (JEngine/JEngine.cpp:139-150): per view
// MoveUAxis(BorderSizeLeft), MoveVAxis(BorderSizeUp), ReverseXAxis, then
// ReverseYAxis when MirroringY, and MoveZAxis(-CenterZ).  The detector-side
// decision (JEngine.cpp:153-161) is taken on the adjusted matrices.
void AdjustProjectionMatrices(std::vector<double>& ptm, const ScanParams& scan, const ReconParams& recon,
                              int device);

namespace hip {

// Throws with kDeviceOOMPrefix when the allocation fails.  (The library's
// plans, which hold the large buffers, also honour the JH_MAX_DEVICE_BYTES
// test cap; pipe frames allocated here do not.)
void* DeviceMalloc(size_t bytes);
void DeviceFree(void* p, size_t bytes);

// One device allocation shared by several DeviceMemory views (a pipe's ring).
std::shared_ptr<void> DeviceArena(size_t bytes);

// Owned device buffer of n elements, or a view of n elements of an arena.
// Copyable (a pipe copies its template into every ring slot): a copy
// allocates and copies the contents device to device.  Replaces
// CUDA/DeviceMemory.cpp:19-79.
template <typename T>
class DeviceMemory {
public:
    DeviceMemory() = default;
    explicit DeviceMemory(size_t n) { Malloc(n); }
    // n elements at element offset `offset` of an arena (kept alive by the view)
    static DeviceMemory View(const std::shared_ptr<void>& arena, size_t offset, size_t n)
    {
        DeviceMemory d;
        d.arena_ = arena;
        d.p_ = static_cast<T*>(arena.get()) + offset;
        d.n_ = n;
        return d;
    }
    DeviceMemory(const DeviceMemory& o)
    {
        Malloc(o.n_);
        if (n_) CheckHip(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice), "DeviceMemory copy");
    }
    DeviceMemory(DeviceMemory&& o) noexcept : p_(o.p_), n_(o.n_), arena_(std::move(o.arena_))
    {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceMemory& operator=(const DeviceMemory& o)
    {
        if (this != &o) {
            Free();
            Malloc(o.n_);
            if (n_) CheckHip(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice), "DeviceMemory copy");
        }
        return *this;
    }
    DeviceMemory& operator=(DeviceMemory&& o) noexcept
    {
        if (this != &o) {
            Free();
            p_ = o.p_;
            n_ = o.n_;
            arena_ = std::move(o.arena_);
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DeviceMemory() { Free(); }

    T* Data() const { return p_; }
    size_t Size() const { return n_; }
    size_t Bytes() const { return n_ * sizeof(T); }

private:
    void Malloc(size_t n);
    void Free()
    {
        if (arena_) arena_.reset();
        else DeviceFree(p_, n_ * sizeof(T));
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0;
    std::shared_ptr<void> arena_;  // set for a view: p_ points into it
};

template <typename T>
void DeviceMemory<T>::Malloc(size_t n)
{
    n_ = n;
    p_ = n ? static_cast<T*>(DeviceMalloc(n * sizeof(T))) : nullptr;
}

}  // namespace hip

// A ring of device frames is one allocation: slot i follows slot i-1, so the
// frames of a token that does not wrap around the ring are contiguous and a
// chunk of views can be processed by one launch.
template <typename U>
struct PipeRing<hip::DeviceMemory<U>> {
    static void Fill(std::vector<hip::DeviceMemory<U>>& buffers, size_t n, const hip::DeviceMemory<U>& tmpl)
    {
        const size_t e = tmpl.Size();
        auto arena = hip::DeviceArena(n * e * sizeof(U));
        buffers.clear();
        buffers.reserve(n);
        for (size_t i = 0; i < n; ++i) buffers.push_back(hip::DeviceMemory<U>::View(arena, i * e, e));
    }
};

// Whether the frames of a token lie back to back (one contiguous buffer).
template <typename U>
bool Contiguous(const hip::DeviceMemory<U>* const* frames, size_t n)
{
    for (size_t i = 1; i < n; ++i)
        if (frames[i]->Data() != frames[0]->Data() + i * frames[0]->Size()) return false;
    return true;
}

// Dump / load payload of a device buffer: copied through host memory
// (the reference's CUDA specialisation, CUDA/PipeDumpLoadImplCUDA.cpp:15-65).
template <typename U>
struct PipeElementIO<hip::DeviceMemory<U>, void> {
    static constexpr bool kSupported = true;
    static void Write(std::ostream& os, const hip::DeviceMemory<U>& d)
    {
        std::vector<U> h(d.Size());
        if (!h.empty()) CheckHip(hipMemcpy(h.data(), d.Data(), d.Bytes(), hipMemcpyDeviceToHost), "dump copy");
        os.write(reinterpret_cast<const char*>(h.data()), (std::streamsize)d.Bytes());
    }
    static void Read(std::istream& is, hip::DeviceMemory<U>& d)
    {
        std::vector<U> h(d.Size());
        is.read(reinterpret_cast<char*>(h.data()), (std::streamsize)d.Bytes());
        if (is && !h.empty()) CheckHip(hipMemcpy(d.Data(), h.data(), d.Bytes(), hipMemcpyHostToDevice), "load copy");
    }
};

// Per-agent HIP stream, created by the agent's GetReady (set-up time, on the
// main thread, with the agent's device current) or else on first use.
class StreamHolder {
public:
    ~StreamHolder();
    hipStream_t Get();
    void Sync();

private:
    hipStream_t s_ = nullptr;
};

// Threads of the copy agents.  Each thread stages its frames through a pinned
// host buffer of its own (hipHostMalloc): the host memcpy between the pipe's
// pageable frame and that buffer runs on the copy threads in parallel, and
// the device DMA runs from / to pinned memory.
constexpr int kCopyThreads = 4;

// A pinned host buffer, allocated at set-up or on first use.
class PinnedBuffer {
public:
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer();
    void* Get(size_t bytes);

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Copies the raw frames up; viewsPerChunk aligns the device ring to the
// chunks its consumer reads.  GetReady creates the copy threads' streams and
// pinned buffers and moves one frame through each stream, so that the HIP
// runtime's one-time set-up of the copy path is not paid by the first frames.
class ProjectionUploadAgent : public SequentialAgentBase {
public:
    ProjectionUploadAgent(size_t width, size_t height, size_t viewsPerChunk = 1, int device = 0);
    void GetReady() override;

private:
    void SetPipesImpl() override;
    void WorkFlow0() override { Copy(0); }
    void WorkFlow1() override { Copy(1); }
    void WorkFlow2() override { Copy(2); }
    void WorkFlow3() override { Copy(3); }
    void Copy(int thread);
    const size_t width_, height_, chunk_;
    const int device_;
    std::shared_ptr<Pipe<UINT16Vec>> in_;
    std::shared_ptr<Pipe<hip::DeviceMemory<uint16_t>>> out_;
    std::mutex tokens_;
    std::atomic<int> running_{kCopyThreads};
    StreamHolder stream_[kCopyThreads];
    PinnedBuffer pinned_[kCopyThreads];
};

class PreProcessingHIPAgent : public SequentialAgentBase {
public:
    // dark (inH x inW) and gain (outH x outW) are optional extensions
    PreProcessingHIPAgent(const ScanParams& scan, const std::vector<uint16_t>& dark = {},
                          const std::vector<float>& gain = {}, int device = 0, size_t viewsPerChunk = 1);
    ~PreProcessingHIPAgent() override;
    void GetReady() override;

private:
    void SetPipesImpl() override;
    void WorkFlow0() override;
    ScanParams scan_;
    std::vector<uint16_t> darkHost_;
    std::vector<float> gainHost_;
    int device_;
    size_t chunk_;
    jh_pre_plan* plan_ = nullptr;
    hip::DeviceMemory<uint16_t> dark_;
    hip::DeviceMemory<float> gain_;
    std::shared_ptr<Pipe<hip::DeviceMemory<uint16_t>>> in_;
    std::shared_ptr<Pipe<hip::DeviceMemory<float>>> out_;
    StreamHolder stream_;
};

// Composes every (view, conjugate) pair of the "Proj" pipe into one frame of
// 2 NumUsedDetsU columns (read size 2, no overlap: CompositeAgent.cpp:22-34).
// The view of a pair is its start index / 2 modulo the view count.
class CompositeHIPAgent : public SequentialAgentBase {
public:
    CompositeHIPAgent(const ScanParams& scan, const std::vector<double>& ptms, bool detectorOnRight, int device = 0);
    ~CompositeHIPAgent() override;
    void GetReady() override;

private:
    void SetPipesImpl() override;
    void WorkFlow0() override;
    ScanParams scan_;
    std::vector<double> ptms_;
    bool right_;
    int device_;
    jh_composite_plan* plan_ = nullptr;
    std::shared_ptr<Pipe<hip::DeviceMemory<float>>> in_, out_;
    StreamHolder stream_;
};

// With composite = true the input frames are the composite's 2W columns and
// the filter runs with widthOut = W and yMirroring = detectorOnRight (the
// FilterAgent arguments of JEngine.cpp:223-238).
class FilterHIPAgent : public SequentialAgentBase {
public:
    FilterHIPAgent(const ScanParams& scan, const ReconParams& recon, const std::vector<double>& ptms, int device = 0,
                   bool composite = false, bool detectorOnRight = false);
    ~FilterHIPAgent() override;
    void GetReady() override;

private:
    void SetPipesImpl() override;
    void WorkFlow0() override;
    ScanParams scan_;
    ReconParams recon_;
    std::vector<double> ptms_;
    int device_;
    bool composite_, right_;
    size_t chunk_;
    jh_filter_plan* plan_ = nullptr;
    std::shared_ptr<Pipe<hip::DeviceMemory<float>>> in_, out_;
    StreamHolder stream_;
};

// Backprojection of the views of each shot into the volume, written out as
// slices at the shot end.  With numParts > 1 the volume is split into
// z-slabs (the reference's "recon parts", BPCUDACore.cpp:32-38): shot p of
// the input stream is back-projected into slab p % numParts.
//
// The first half of a shot's views (whole chunks) is accumulated over the
// whole volume as it arrives; the rest is staged and held.  At the shot end
// the held views are back-projected z-slab by z-slab (jh_bp_accumulate_slab),
// and each slab's slices are finalised and handed on before the next slab
// runs, so the download of the volume overlaps the last half of the
// backprojection.  Every voxel still takes its views in order: the volume is
// the one-launch volume to the bit.
class BPHIPAgent : public SequentialAgentBase {
public:
    BPHIPAgent(const ScanParams& scan, const ReconParams& recon, const std::vector<double>& ptms,
               size_t numParts = 1, int device = 0);
    ~BPHIPAgent() override;
    void GetReady() override;

    // first slice and slice count of slab p
    size_t PartStart(size_t p) const;
    size_t PartSlices(size_t p) const;

private:
    void SetPipesImpl() override;
    void WorkFlow0() override;
    jh_bp_plan* PlanFor(size_t part);
    // back-projects the held views slab by slab and writes the part's slices
    void FinishShot(jh_bp_plan* plan, size_t part, size_t held);
    ScanParams scan_;
    ReconParams recon_;
    std::vector<double> ptms_;
    size_t numParts_;
    int device_;
    size_t part_ = SIZE_MAX;
    jh_bp_plan* plan_ = nullptr;
    hip::DeviceMemory<uint8_t> q_;
    size_t tail_ = 0;                   // first held view of a shot
    size_t frameBytes_ = 0;             // staged bytes per view
    hip::DeviceMemory<uint8_t> tailQ_;  // the held views, staged
    std::shared_ptr<Pipe<hip::DeviceMemory<float>>> in_, out_;
    StreamHolder stream_;
};

// Copies the slices down; GetReady sets up its streams and pinned buffers as
// ProjectionUploadAgent's does.
class DeviceMemoryDownloadAgent : public SequentialAgentBase {
public:
    explicit DeviceMemoryDownloadAgent(size_t elements, int device = 0);
    void GetReady() override;

private:
    void SetPipesImpl() override;
    void WorkFlow0() override { Copy(0); }
    void WorkFlow1() override { Copy(1); }
    void WorkFlow2() override { Copy(2); }
    void WorkFlow3() override { Copy(3); }
    void Copy(int thread);
    const size_t n_;
    const int device_;
    std::shared_ptr<Pipe<hip::DeviceMemory<float>>> in_;
    std::shared_ptr<Pipe<FloatVec>> out_;
    std::mutex tokens_;
    std::atomic<int> running_{kCopyThreads};
    StreamHolder stream_[kCopyThreads];
    PinnedBuffer pinned_[kCopyThreads];
};

}  // namespace jh
