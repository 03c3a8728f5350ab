// Device-side agents (see HipAgents.h).
#include "jh/HipAgents.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace jh {

const char* const kDeviceOOMPrefix = "HIP failed to malloc ";

void CheckJh(jh_status s, const char* what)
{
    if (s != JH_OK) throw Exception(std::string(what) + ": " + jh_last_error());
}

void CheckHip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Exception(std::string(what) + ": " + hipGetErrorString(e));
}

void AdjustProjectionMatrices(std::vector<double>& ptm, const ScanParams& scan, const ReconParams& recon,
                              int device)
{
    if (ptm.size() % 12) ThrowException("AdjustProjectionMatrices: not a whole number of 3x4 matrices");
    const int nv = (int)(ptm.size() / 12);
    if (nv == 0) return;
    CheckHip(hipSetDevice(device), "hipSetDevice");
    double* d = nullptr;
    CheckHip(hipMalloc(&d, ptm.size() * sizeof(double)), "hipMalloc ptm");
    jh_status s = JH_OK;
    const char* what = "ptm upload";
    if (hipMemcpy(d, ptm.data(), ptm.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        CheckHip(hipErrorInvalidValue, what);
    }
    // the reference's order (JEngine.cpp:141-149); values as the reference
    // forms them (float(BorderSize*), -CenterZ)
    struct Op {
        int op;
        double value;
        bool on;
    };
    const Op ops[] = {
        {JH_PTM_MOVE_U, (double)(float)scan.BorderSizeLeft, true},
        {JH_PTM_MOVE_V, (double)(float)scan.BorderSizeUp, true},
        {JH_PTM_REVERSE_X, 0.0, true},
        {JH_PTM_REVERSE_Y, 0.0, recon.MirroringY},
        {JH_PTM_MOVE_Z, -(double)recon.CenterZ, true},
    };
    for (const Op& o : ops) {
        if (!o.on) continue;
        s = jh_ptm_modify(d, nv, o.op, o.value, nullptr);
        if (s != JH_OK) {
            what = "jh_ptm_modify";
            break;
        }
    }
    const hipError_t e = s == JH_OK ? hipMemcpy(ptm.data(), d, ptm.size() * sizeof(double), hipMemcpyDeviceToHost)
                                    : hipSuccess;
    (void)hipFree(d);
    CheckJh(s, what);
    CheckHip(e, "ptm download");
}

namespace hip {

static std::atomic<size_t> g_allocated{0};

void* DeviceMalloc(size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess || !p) {
        (void)hipGetLastError();
        throw Exception(std::string(kDeviceOOMPrefix) + std::to_string(bytes) + " bytes");
    }
    g_allocated += bytes;
    return p;
}

void DeviceFree(void* p, size_t bytes)
{
    if (p) {
        (void)hipFree(p);
        g_allocated -= bytes;
    }
}

std::shared_ptr<void> DeviceArena(size_t bytes)
{
    void* p = bytes ? DeviceMalloc(bytes) : nullptr;
    return std::shared_ptr<void>(p, [bytes](void* q) { DeviceFree(q, bytes); });
}

}  // namespace hip

StreamHolder::~StreamHolder()
{
    if (s_) (void)hipStreamDestroy(s_);
}

hipStream_t StreamHolder::Get()
{
    if (!s_) CheckHip(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate");
    return s_;
}

void StreamHolder::Sync() { CheckHip(hipStreamSynchronize(Get()), "hipStreamSynchronize"); }

PinnedBuffer::~PinnedBuffer()
{
    if (p_) (void)hipHostFree(p_);
}

void* PinnedBuffer::Get(size_t bytes)
{
    if (bytes > n_) {
        if (p_) CheckHip(hipHostFree(p_), "hipHostFree");
        p_ = nullptr;
        n_ = 0;
        if (hipHostMalloc(&p_, bytes, hipHostMallocDefault) != hipSuccess || !p_) {
            (void)hipGetLastError();
            p_ = nullptr;
            throw Exception(std::string("hipHostMalloc of ") + std::to_string(bytes) + " bytes failed");
        }
        n_ = bytes;
    }
    return p_;
}

// ---------------------------------------------------------------------------
// ProjectionUploadAgent
// ---------------------------------------------------------------------------

ProjectionUploadAgent::ProjectionUploadAgent(size_t width, size_t height, size_t viewsPerChunk, int device)
    : SequentialAgentBase("ProjectionUploadAgent", kCopyThreads), width_(width), height_(height),
      chunk_(std::max<size_t>(1, viewsPerChunk)), device_(device)
{
}

// Streams and pinned buffers of the copy threads, and one frame through each
// stream: stream creation and the first transfers take milliseconds each.
void ProjectionUploadAgent::GetReady()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const size_t bytes = width_ * height_ * sizeof(uint16_t);
    hip::DeviceMemory<uint16_t> frame(width_ * height_);
    for (int t = 0; t < kCopyThreads; ++t) {
        void* stage = pinned_[t].Get(bytes);
        std::memset(stage, 0, bytes);
        CheckHip(hipMemcpyAsync(frame.Data(), stage, bytes, hipMemcpyHostToDevice, stream_[t].Get()), "upload set-up");
    }
    for (int t = 0; t < kCopyThreads; ++t) stream_[t].Sync();
}

void ProjectionUploadAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<UINT16Vec>("Proj");
    in_->SetConsumer(GetAgentName(), kCopyThreads, 1, 0);
    out_ = CreateNewPipe<hip::DeviceMemory<uint16_t>>("Proj");
    out_->SetTemplate(hip::DeviceMemory<uint16_t>(width_ * height_), {width_, height_});
    out_->SetBufferGranule(chunk_);
    out_->SetProducer(GetAgentName(), kCopyThreads, 1);
    MergePipeToTrunk(out_);
}

// Several threads, each with its own stream; a thread takes a read token and the
// matching write token together, so frames keep their stream order.
void ProjectionUploadAgent::Copy(int thread)
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    try {
        for (;;) {
            Pipe<UINT16Vec>::ReadToken rt;
            Pipe<hip::DeviceMemory<uint16_t>>::WriteToken wt;
            {
                std::lock_guard<std::mutex> lock(tokens_);
                rt = in_->GetReadToken();
                wt = out_->GetWriteToken(1, rt.IsShotEnd());
            }
            ThreadOccupiedScope occ(this);
            const UINT16Vec& src = rt.GetBuffer(0);
            if (src.size() != width_ * height_) ThrowExceptionAndLog(GetAgentName() + ": frame size mismatch");
            const size_t bytes = src.size() * sizeof(uint16_t);
            void* stage = pinned_[thread].Get(bytes);
            std::memcpy(stage, src.data(), bytes);
            CheckHip(hipMemcpyAsync(wt.GetBuffer(0).Data(), stage, bytes, hipMemcpyHostToDevice, stream_[thread].Get()),
                     "upload");
            WaitAsyncScope wait(&occ);
            stream_[thread].Sync();
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    if (--running_ == 0) out_->Close();
}

// ---------------------------------------------------------------------------
// PreProcessingHIPAgent
// ---------------------------------------------------------------------------

PreProcessingHIPAgent::PreProcessingHIPAgent(const ScanParams& scan, const std::vector<uint16_t>& dark,
                                             const std::vector<float>& gain, int device, size_t viewsPerChunk)
    : SequentialAgentBase("PreProcessingHIPAgent", 1), scan_(scan), darkHost_(dark), gainHost_(gain), device_(device),
      chunk_(std::max<size_t>(1, viewsPerChunk))
{
    scan_.Derive();
}

PreProcessingHIPAgent::~PreProcessingHIPAgent()
{
    if (plan_) jh_pre_destroy(plan_);
}

void PreProcessingHIPAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<hip::DeviceMemory<uint16_t>>("Proj");
    in_->SetBufferGranule(chunk_);
    in_->SetConsumer(GetAgentName(), 1, chunk_, 0);
    out_ = CreateNewPipe<hip::DeviceMemory<float>>("Proj");
    out_->SetTemplate(hip::DeviceMemory<float>(scan_.NumUsedDetsU * scan_.NumUsedDetsV),
                      {scan_.NumUsedDetsU, scan_.NumUsedDetsV});
    out_->SetBufferGranule(chunk_);
    out_->SetProducer(GetAgentName(), 1, chunk_);
    MergePipeToTrunk(out_);
}

void PreProcessingHIPAgent::GetReady()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const float* bh = scan_.BeamHardeningParams.data();
    if (scan_.BeamHardeningParams.size() != 3) ThrowException("The number of BeamHardeningParams should be 3.");
    CheckJh(jh_pre_create(&plan_, device_, (int)scan_.NumDetsU, (int)scan_.NumDetsV, (int)scan_.BorderSizeUp,
                          (int)scan_.BorderSizeDown, (int)scan_.BorderSizeLeft, (int)scan_.BorderSizeRight,
                          (int)scan_.NumUsedDetsU, (int)scan_.NumUsedDetsV, scan_.BrightField, bh),
            "jh_pre_create");
    if (!darkHost_.empty()) {
        dark_ = hip::DeviceMemory<uint16_t>(darkHost_.size());
        CheckHip(hipMemcpy(dark_.Data(), darkHost_.data(), dark_.Bytes(), hipMemcpyHostToDevice), "dark upload");
    }
    if (!gainHost_.empty()) {
        gain_ = hip::DeviceMemory<float>(gainHost_.size());
        CheckHip(hipMemcpy(gain_.Data(), gainHost_.data(), gain_.Bytes(), hipMemcpyHostToDevice), "gain upload");
    }
    stream_.Get();
}

void PreProcessingHIPAgent::WorkFlow0()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    try {
        for (;;) {
            auto rt = in_->GetReadToken();
            const size_t n = rt.GetSize();
            auto wt = out_->GetWriteToken(n, rt.IsShotEnd());
            ThreadOccupiedScope occ(this);
            // one launch over the chunk when both runs of frames are contiguous
            if (Contiguous(rt.GetBufferPtrs(), n) && Contiguous(wt.GetBufferPtrs(), n)) {
                CheckJh(jh_pre_run(plan_, rt.GetBuffer(0).Data(), dark_.Data(), gain_.Data(), wt.GetBuffer(0).Data(),
                                   (int)n, stream_.Get()),
                        "jh_pre_run");
            } else {
                for (size_t i = 0; i < n; ++i)
                    CheckJh(jh_pre_run(plan_, rt.GetBuffer(i).Data(), dark_.Data(), gain_.Data(),
                                       wt.GetBuffer(i).Data(), 1, stream_.Get()),
                            "jh_pre_run");
            }
            WaitAsyncScope wait(&occ);
            stream_.Sync();
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    out_->Close();
}

// ---------------------------------------------------------------------------
// CompositeHIPAgent
// ---------------------------------------------------------------------------

CompositeHIPAgent::CompositeHIPAgent(const ScanParams& scan, const std::vector<double>& ptms, bool detectorOnRight,
                                     int device)
    : SequentialAgentBase("CompositeAgent", 1), scan_(scan), ptms_(ptms), right_(detectorOnRight), device_(device)
{
    scan_.Derive();
}

CompositeHIPAgent::~CompositeHIPAgent()
{
    if (plan_) jh_composite_destroy(plan_);
}

void CompositeHIPAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<hip::DeviceMemory<float>>("Proj");
    in_->SetConsumer(GetAgentName(), 1, 2, 0);
    out_ = CreateNewPipe<hip::DeviceMemory<float>>("Proj");
    out_->SetProducer(GetAgentName(), 1, 1);
    out_->SetTemplate(hip::DeviceMemory<float>(2 * scan_.NumUsedDetsU * scan_.NumUsedDetsV),
                      {2 * scan_.NumUsedDetsU, scan_.NumUsedDetsV});
    MergePipeToTrunk(out_);
}

void CompositeHIPAgent::GetReady()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    CheckJh(jh_composite_create(&plan_, device_, (int)scan_.NumUsedDetsU, (int)scan_.NumUsedDetsV,
                                (int)(ptms_.size() / 12), ptms_.data(), scan_.DSO, right_ ? 1 : 0),
            "jh_composite_create");
    stream_.Get();
}

void CompositeHIPAgent::WorkFlow0()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const size_t nv = ptms_.size() / 12;
    try {
        for (;;) {
            auto rt = in_->GetReadToken();
            if (rt.GetSize() != 2) ThrowExceptionAndLog(GetAgentName() + ": a shot holds an odd number of frames");
            auto wt = out_->GetWriteToken(1, rt.IsShotEnd());
            ThreadOccupiedScope occ(this);
            const int view = (int)((rt.GetStartIndex() / 2) % nv);
            CheckJh(jh_composite_run(plan_, rt.GetBuffer(0).Data(), rt.GetBuffer(1).Data(), 0, wt.GetBuffer(0).Data(),
                                     view, 1, stream_.Get()),
                    "jh_composite_run");
            WaitAsyncScope wait(&occ);
            stream_.Sync();
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    out_->Close();
}

// ---------------------------------------------------------------------------
// FilterHIPAgent
// ---------------------------------------------------------------------------

FilterHIPAgent::FilterHIPAgent(const ScanParams& scan, const ReconParams& recon, const std::vector<double>& ptms,
                               int device, bool composite, bool detectorOnRight)
    : SequentialAgentBase("FilterHIPAgent", 1), scan_(scan), recon_(recon), ptms_(ptms), device_(device),
      composite_(composite), right_(composite && detectorOnRight), chunk_(std::max<size_t>(1, recon.ViewsPerChunk))
{
    scan_.Derive();
}

FilterHIPAgent::~FilterHIPAgent()
{
    if (plan_) jh_filter_destroy(plan_);
}

void FilterHIPAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<hip::DeviceMemory<float>>("Proj");
    in_->SetBufferGranule(chunk_);
    in_->SetConsumer(GetAgentName(), 1, chunk_, 0);
    out_ = CreateNewPipe<hip::DeviceMemory<float>>("Proj");
    out_->SetTemplate(hip::DeviceMemory<float>(scan_.NumUsedDetsU * scan_.NumUsedDetsV),
                      {scan_.NumUsedDetsU, scan_.NumUsedDetsV});
    out_->SetBufferGranule(chunk_);
    out_->SetProducer(GetAgentName(), 1, chunk_);
    MergePipeToTrunk(out_);
}

void FilterHIPAgent::GetReady()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const int nv = (int)(ptms_.size() / 12);
    const int nAdj = (int)recon_.FilterAdjustPoints.size();
    if (recon_.FilterAdjustLevelInDB.size() != (size_t)nAdj) ThrowException("filter adjust points and levels differ");
    const int W = (int)scan_.NumUsedDetsU;
    CheckJh(jh_filter_create_ex(&plan_, device_, composite_ ? 2 * W : W, (int)scan_.NumUsedDetsV, W, right_ ? 1 : 0,
                                nv, ptms_.data(), scan_.DSO, scan_.HalfSampleRate, recon_.FilterCutOffStart,
                                recon_.FilterCutOffEnd, nAdj ? recon_.FilterAdjustPoints.data() : nullptr,
                                nAdj ? recon_.FilterAdjustLevelInDB.data() : nullptr, nAdj, recon_.FilterWindow,
                                recon_.FilterMode),
            "jh_filter_create_ex");
    stream_.Get();
}

void FilterHIPAgent::WorkFlow0()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const size_t nv = ptms_.size() / 12;
    try {
        for (;;) {
            auto rt = in_->GetReadToken();
            const size_t n = rt.GetSize();
            auto wt = out_->GetWriteToken(n, rt.IsShotEnd());
            ThreadOccupiedScope occ(this);
            const int view = (int)(rt.GetStartIndex() % nv);
            // one launch over the chunk when its frames are contiguous and its
            // views consecutive (a chunk never spans two shots)
            if (view + n <= nv && Contiguous(rt.GetBufferPtrs(), n) && Contiguous(wt.GetBufferPtrs(), n)) {
                CheckJh(jh_filter_run(plan_, rt.GetBuffer(0).Data(), wt.GetBuffer(0).Data(), view, (int)n,
                                      stream_.Get()),
                        "jh_filter_run");
            } else {
                for (size_t i = 0; i < n; ++i)
                    CheckJh(jh_filter_run(plan_, rt.GetBuffer(i).Data(), wt.GetBuffer(i).Data(),
                                          (int)((rt.GetStartIndex() + i) % nv), 1, stream_.Get()),
                            "jh_filter_run");
            }
            WaitAsyncScope wait(&occ);
            stream_.Sync();
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    out_->Close();
}

// ---------------------------------------------------------------------------
// BPHIPAgent
// ---------------------------------------------------------------------------

BPHIPAgent::BPHIPAgent(const ScanParams& scan, const ReconParams& recon, const std::vector<double>& ptms,
                       size_t numParts, int device)
    : SequentialAgentBase("BPHIPAgent", 1), scan_(scan), recon_(recon), ptms_(ptms), numParts_(numParts),
      device_(device)
{
    scan_.Derive();
    if (numParts_ < 1 || numParts_ > recon_.NumPixelsZ) ThrowException("BPHIPAgent: bad number of parts");
    if (recon_.ViewsPerChunk < 1) ThrowException("BPHIPAgent: ViewsPerChunk must be positive");
}

BPHIPAgent::~BPHIPAgent()
{
    if (plan_) jh_bp_destroy(plan_);
}

size_t BPHIPAgent::PartStart(size_t p) const
{
    const size_t per = (recon_.NumPixelsZ + numParts_ - 1) / numParts_;
    return std::min(p * per, recon_.NumPixelsZ);
}

size_t BPHIPAgent::PartSlices(size_t p) const
{
    const size_t per = (recon_.NumPixelsZ + numParts_ - 1) / numParts_;
    return std::min(per, recon_.NumPixelsZ - PartStart(p));
}

// slices finalised per launch (one contiguous write token)
constexpr size_t kSliceChunk = 16;
// slices per z-slab of the held views (rounded to the z tile)
constexpr size_t kSlabSlices = 64;

void BPHIPAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<hip::DeviceMemory<float>>("Proj");
    in_->SetBufferGranule(recon_.ViewsPerChunk);
    in_->SetConsumer(GetAgentName(), 1, recon_.ViewsPerChunk, 0);
    out_ = CreateNewPipe<hip::DeviceMemory<float>>("Slice");
    out_->SetTemplate(hip::DeviceMemory<float>(recon_.NumPixelsX * recon_.NumPixelsY),
                      {recon_.NumPixelsX, recon_.NumPixelsY});
    out_->SetBufferGranule(kSliceChunk);
    out_->SetProducer(GetAgentName(), 1, kSliceChunk);
    // room for four slabs of slices, so that the next slab is back-projected
    // while the download drains the last ones (the default ring, one token
    // per writer and reader, would hold the agent to the download's pace)
    const size_t slices = (PartSlices(0) + kSliceChunk - 1) / kSliceChunk * kSliceChunk;
    out_->SetBufferSize(std::max<size_t>(2 * kSliceChunk, std::min<size_t>(slices, 4 * kSlabSlices)));
    MergePipeToTrunk(out_);
}

// Plan of slab p: its slices, with the z offset of the slab folded into the
// matrices (ProjectionMatrixModifier::MoveZAxis semantics, in FP64 on the
// device).
jh_bp_plan* BPHIPAgent::PlanFor(size_t p)
{
    if (part_ == p) return plan_;
    if (plan_) {
        CheckJh(jh_bp_destroy(plan_), "jh_bp_destroy");
        plan_ = nullptr;
    }
    const size_t nv = ptms_.size() / 12;
    const size_t z0 = PartStart(p), nz = PartSlices(p);
    const double Z = (double)recon_.NumPixelsZ;
    const double dz = ((double)z0 + ((double)nz - 1.0) / 2.0 - (Z - 1.0) / 2.0) * (double)recon_.PitchZ;
    std::vector<double> ptm(ptms_);
    if (dz != 0.0) {
        double* d = nullptr;
        CheckHip(hipMalloc(&d, ptm.size() * sizeof(double)), "hipMalloc ptm");
        CheckHip(hipMemcpy(d, ptm.data(), ptm.size() * sizeof(double), hipMemcpyHostToDevice), "ptm upload");
        const jh_status s = jh_ptm_modify(d, (int)nv, JH_PTM_MOVE_Z, dz, nullptr);
        if (s == JH_OK) (void)hipMemcpy(ptm.data(), d, ptm.size() * sizeof(double), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        CheckJh(s, "jh_ptm_modify");
    }
    CheckJh(jh_bp_create(&plan_, device_, ptm.data(), (int)nv, (int)scan_.NumUsedDetsU, (int)scan_.NumUsedDetsV,
                         (int)recon_.NumPixelsX, (int)recon_.NumPixelsY, (int)nz, recon_.PitchXY, recon_.PitchZ),
            "jh_bp_create");
    if (recon_.ZFilter) CheckJh(jh_bp_set_zfilter(plan_, 1), "jh_bp_set_zfilter");
    if (recon_.CTNumber)
        CheckJh(jh_bp_set_ctnum(plan_, 1, recon_.CTNumNorm0, recon_.CTNumNorm1, recon_.MuWater), "jh_bp_set_ctnum");
    // the accumulators are allocated here so that a shortage shows up now
    CheckJh(jh_bp_reset(plan_, stream_.Get()), "jh_bp_reset");
    stream_.Sync();
    part_ = p;
    return plan_;
}

void BPHIPAgent::GetReady()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    PlanFor(0);
    const size_t nv = ptms_.size() / 12;
    size_t bytes = 0;
    CheckJh(jh_bp_stage_bytes(plan_, recon_.StorageType, (int)recon_.ViewsPerChunk, &bytes), "jh_bp_stage_bytes");
    q_ = hip::DeviceMemory<uint8_t>(bytes);
    CheckJh(jh_bp_stage_bytes(plan_, recon_.StorageType, 1, &frameBytes_), "jh_bp_stage_bytes");
    // the views from the chunk boundary nearest below half the shot on are held
    tail_ = nv / 2 / recon_.ViewsPerChunk * recon_.ViewsPerChunk;
    tailQ_ = hip::DeviceMemory<uint8_t>((nv - tail_) * frameBytes_);
}

void BPHIPAgent::FinishShot(jh_bp_plan* plan, size_t part, size_t held)
{
    const size_t nz = PartSlices(part);
    int zc = 0, ty = 0;
    CheckJh(jh_bp_get_tile(plan, &zc, &ty), "jh_bp_get_tile");
    const size_t slab = held ? std::max<size_t>((size_t)zc, kSlabSlices / (size_t)zc * (size_t)zc) : nz;
    for (size_t z0 = 0; z0 < nz; z0 += slab) {
        const size_t m = std::min(slab, nz - z0);
        if (held)
            CheckJh(jh_bp_accumulate_slab(plan, tailQ_.Data(), recon_.StorageType, (int)tail_, (int)held, (int)z0,
                                          (int)m, stream_.Get()),
                    "jh_bp_accumulate_slab");
        for (size_t z = z0; z < z0 + m; z += kSliceChunk) {
            const size_t k = std::min(kSliceChunk, z0 + m - z);
            auto wt = out_->GetWriteToken(k, z + k == nz);
            if (Contiguous(wt.GetBufferPtrs(), k)) {
                CheckJh(jh_bp_finalize(plan, wt.GetBuffer(0).Data(), (int)z, (int)k, stream_.Get()), "jh_bp_finalize");
            } else {
                for (size_t i = 0; i < k; ++i)
                    CheckJh(jh_bp_finalize(plan, wt.GetBuffer(i).Data(), (int)(z + i), 1, stream_.Get()),
                            "jh_bp_finalize");
            }
            stream_.Sync();
        }
    }
}

void BPHIPAgent::WorkFlow0()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const size_t nv = ptms_.size() / 12;
    const int dt = recon_.StorageType;
    size_t shot = 0, held = 0;
    try {
        for (;;) {
            auto rt = in_->GetReadToken();
            ThreadOccupiedScope occ(this);
            jh_bp_plan* plan = PlanFor(shot % numParts_);
            if (rt.IsShotStart()) {
                CheckJh(jh_bp_reset(plan, stream_.Get()), "jh_bp_reset");
                held = 0;
            }
            const size_t n = rt.GetSize();
            const size_t first = rt.GetStartIndex() % nv;
            if (first + n > nv) ThrowException("BPHIPAgent: a chunk spans two shots");
            // stages the token's frames [i0, i0 + m) (views first + i0 ..)
            // to dst: one launch when the frames are contiguous
            auto stage = [&](size_t i0, size_t m, uint8_t* dst) {
                const int v = (int)(first + i0);
                if (Contiguous(rt.GetBufferPtrs() + i0, m)) {
                    CheckJh(jh_bp_stage(plan, rt.GetBuffer(i0).Data(), v, (int)m, dt, dst, stream_.Get()),
                            "jh_bp_stage");
                } else {
                    for (size_t i = 0; i < m; ++i)
                        CheckJh(jh_bp_stage(plan, rt.GetBuffer(i0 + i).Data(), v + (int)i, 1, dt,
                                            dst + i * frameBytes_, stream_.Get()),
                                "jh_bp_stage");
                }
            };
            // views before tail_ go into the whole volume now, the others are held
            const size_t head = first < tail_ ? std::min(n, tail_ - first) : 0;
            if (head) {
                stage(0, head, q_.Data());
                CheckJh(jh_bp_accumulate(plan, q_.Data(), dt, (int)first, (int)head, stream_.Get()),
                        "jh_bp_accumulate");
            }
            if (head < n) {
                if (first + head != tail_ + held) ThrowException("BPHIPAgent: held views out of order");
                stage(head, n - head, tailQ_.Data() + held * frameBytes_);
                held += n - head;
            }
            {
                WaitAsyncScope wait(&occ);
                stream_.Sync();
            }
            if (rt.IsShotEnd()) {
                rt.Reset();  // views are no longer needed
                FinishShot(plan, shot % numParts_, held);
                held = 0;
                ++shot;
            }
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    out_->Close();
}

// ---------------------------------------------------------------------------
// DeviceMemoryDownloadAgent
// ---------------------------------------------------------------------------

DeviceMemoryDownloadAgent::DeviceMemoryDownloadAgent(size_t elements, int device)
    : SequentialAgentBase("DeviceMemoryDownloadAgent", kCopyThreads), n_(elements), device_(device)
{
}

// As ProjectionUploadAgent::GetReady, one slice down through each stream.
void DeviceMemoryDownloadAgent::GetReady()
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    const size_t bytes = n_ * sizeof(float);
    hip::DeviceMemory<float> slice(n_);
    CheckHip(hipMemset(slice.Data(), 0, bytes), "download set-up");
    for (int t = 0; t < kCopyThreads; ++t)
        CheckHip(hipMemcpyAsync(pinned_[t].Get(bytes), slice.Data(), bytes, hipMemcpyDeviceToHost, stream_[t].Get()),
                 "download set-up");
    for (int t = 0; t < kCopyThreads; ++t) stream_[t].Sync();
}

void DeviceMemoryDownloadAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<hip::DeviceMemory<float>>("Slice");
    in_->SetConsumer(GetAgentName(), kCopyThreads, 1, 0);
    out_ = CreateNewPipe<FloatVec>("Slice");
    out_->SetTemplate(FloatVec(n_), {n_});
    out_->SetProducer(GetAgentName(), kCopyThreads, 1);
    MergePipeToTrunk(out_);
}

// Several threads, each with its own stream; read and write tokens are taken
// together, so slices keep their order.
void DeviceMemoryDownloadAgent::Copy(int thread)
{
    CheckHip(hipSetDevice(device_), "hipSetDevice");
    try {
        for (;;) {
            Pipe<hip::DeviceMemory<float>>::ReadToken rt;
            Pipe<FloatVec>::WriteToken wt;
            {
                std::lock_guard<std::mutex> lock(tokens_);
                rt = in_->GetReadToken();
                wt = out_->GetWriteToken(1, rt.IsShotEnd());
            }
            ThreadOccupiedScope occ(this);
            void* stage = pinned_[thread].Get(n_ * sizeof(float));
            CheckHip(hipMemcpyAsync(stage, rt.GetBuffer(0).Data(), n_ * sizeof(float), hipMemcpyDeviceToHost,
                                    stream_[thread].Get()),
                     "download");
            {
                WaitAsyncScope wait(&occ);
                stream_[thread].Sync();
            }
            std::memcpy(wt.GetBuffer(0).data(), stage, n_ * sizeof(float));
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
    if (--running_ == 0) out_->Close();
}

}  // namespace jh
