// jh_fdk_pipeline: FDK reconstruction of one scan through the agent graph
//
//   InputAgent | MemoryInputAgent -> ProjectionUploadAgent -> PreProcessingHIPAgent
//   [-> CompositeHIPAgent] -> FilterHIPAgent -> BPHIPAgent -> DeviceMemoryDownloadAgent
//   -> OutputAgent | SliceOutputAgent
//
// wired by pipe name ("Proj", "Slice") the way JEngine/JEngine.cpp:185-346
// wires its agents, with the reference's retry on device shortage
// (JEngine.cpp:167-176, 413-426): when set-up fails with the device
// out-of-memory prefix the graph is rebuilt with one more z-slab ("part") and
// the input is streamed once per part.
//
// Input: either the reference's scan folder (input_folder=, input_template=
// e.g. Proj_%d.prj, u16 frames; ptm_file= PTM.dat text) or raw files
// (raw= N x NumDetsV x NumDetsU u16, ptm= N x 12 f64); optional dark (u16)
// and gain (f32) frames.  Matrices from PTM.dat get JEngine's adjustment
// (JEngine.cpp:139-150: borders, ReverseX, ReverseY with mirror_y=1,
// MoveZ(-PitchZ * translation_z)); adjust_ptm=0/1 overrides the default
// (1 for ptm_file=, 0 for ptm=).  composite=1 runs the offset-detector chain (each
// view paired with its conjugate, JEngine.cpp:152-238).  Output: the volume
// Z x Y x X f32 (out=) and/or slice files (output_folder=, output_template=
// e.g. slice_%d.slc).  dump= / load= take comma-separated
// <pipe>_<producer>_<consumer>:<path> items (the keys of the reference's
// pipe_dump_load.xml, Config/PipeDumpLoadReader.cpp:14-55).  Parameters are
// key=value arguments.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>

#include "jh/FileAgents.h"
#include "jh/HipAgents.h"

using namespace jh;

namespace {

template <typename T>
std::vector<T> ReadFile(const std::string& path, size_t count)
{
    std::vector<T> v(count);
    std::ifstream is(path, std::ios::binary);
    if (!is) ThrowException("cannot open " + path);
    is.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(count * sizeof(T)));
    if ((size_t)is.gcount() != count * sizeof(T)) ThrowException("short read from " + path);
    return v;
}

// Streams the scan from memory, once per part, one shot per pass
// (JEngine/InputAgent.cpp:39-91 streams files the same way), with each view
// followed by its conjugate when pairs is set.  Four threads copy frames;
// each takes the next frame and its write token together, so the order is
// kept.
constexpr int kHostThreads = 4;

class MemoryInputAgent : public SequentialAgentBase {
public:
    MemoryInputAgent(const std::vector<uint16_t>& raw, size_t frame, size_t views, size_t passes, bool pairs)
        : SequentialAgentBase("MemoryInputAgent", kHostThreads), raw_(raw), frame_(frame), views_(views),
          passes_(passes), pairs_(pairs)
    {
    }

private:
    void SetPipesImpl() override
    {
        out_ = CreateNewPipe<UINT16Vec>("Proj");
        out_->SetTemplate(UINT16Vec(frame_), {frame_});
        out_->SetProducer(GetAgentName(), kHostThreads, 1);
        MergePipeToTrunk(out_);
    }
    void WorkFlow0() override { Copy(); }
    void WorkFlow1() override { Copy(); }
    void WorkFlow2() override { Copy(); }
    void WorkFlow3() override { Copy(); }
    void Copy()
    {
        const size_t per = pairs_ ? 2u : 1u, total = passes_ * views_ * per;
        for (;;) {
            size_t view;
            Pipe<UINT16Vec>::WriteToken wt;
            {
                std::lock_guard<std::mutex> lock(m_);
                if (next_ == total) break;
                const size_t e = next_++;
                const size_t k = e % per, v = (e / per) % views_;
                view = k ? (v + views_ / 2) % views_ : v;
                wt = out_->GetWriteToken(1, v + 1 == views_ && k + 1 == per);
            }
            ThreadOccupiedScope occ(this);
            std::memcpy(wt.GetBuffer(0).data(), raw_.data() + view * frame_, frame_ * sizeof(uint16_t));
        }
        if (--running_ == 0) out_->Close();
    }
    const std::vector<uint16_t>& raw_;
    const size_t frame_, views_, passes_;
    const bool pairs_;
    std::mutex m_;
    size_t next_ = 0;
    std::atomic<int> running_{kHostThreads};
    std::shared_ptr<Pipe<UINT16Vec>> out_;
};

// Collects slices into the output volume (JEngine/OutputAgent.cpp:37-82
// writes them to files instead).  Four threads; each takes a slice token and
// its index together.
class SliceOutputAgent : public SequentialAgentBase {
public:
    SliceOutputAgent(std::vector<float>& vol, size_t plane, const BPHIPAgent& bp, size_t parts)
        : SequentialAgentBase("SliceOutputAgent", kHostThreads), vol_(vol), plane_(plane), bp_(bp), parts_(parts)
    {
    }

private:
    void SetPipesImpl() override
    {
        in_ = BranchPipeFromTrunk<FloatVec>("Slice");
        in_->SetConsumer(GetAgentName(), kHostThreads, 1, 0);
    }
    void WorkFlow0() override { Collect(); }
    void WorkFlow1() override { Collect(); }
    void WorkFlow2() override { Collect(); }
    void WorkFlow3() override { Collect(); }
    void Collect()
    {
        try {
            for (;;) {
                Pipe<FloatVec>::ReadToken rt;
                size_t slice;
                {
                    std::lock_guard<std::mutex> lock(m_);
                    rt = in_->GetReadToken();
                    slice = bp_.PartStart(part_) + z_;
                    ++z_;
                    if (rt.IsShotEnd()) {
                        ++part_;
                        z_ = 0;
                    }
                }
                ThreadOccupiedScope occ(this);
                std::memcpy(vol_.data() + slice * plane_, rt.GetBuffer(0).data(), plane_ * sizeof(float));
            }
        } catch (PipeClosedAndEmptySignal&) {
        }
        if (--running_ == 0 && part_ != parts_)
            ThrowExceptionAndLog("SliceOutputAgent: received " + std::to_string(part_) + " of " +
                                 std::to_string(parts_) + " parts");
    }
    std::vector<float>& vol_;
    const size_t plane_;
    const BPHIPAgent& bp_;
    const size_t parts_;
    std::mutex m_;
    size_t part_ = 0, z_ = 0;
    std::atomic<int> running_{kHostThreads};
    std::shared_ptr<Pipe<FloatVec>> in_;
};

struct Args {
    std::map<std::string, std::string> kv;
    std::string S(const std::string& k, const std::string& d = "") const
    {
        auto it = kv.find(k);
        if (it == kv.end()) {
            if (d.empty()) ThrowException("missing argument " + k);
            return d;
        }
        return it->second;
    }
    double D(const std::string& k, const std::string& d = "") const { return std::atof(S(k, d).c_str()); }
    long I(const std::string& k, const std::string& d = "") const { return std::atol(S(k, d).c_str()); }
};

// "key:path,key:path" -> map
std::map<std::string, std::string> ParsePathMap(const std::string& s)
{
    std::map<std::string, std::string> m;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        const size_t c = item.find(':');
        if (c == std::string::npos || c == 0) ThrowException("dump/load item needs <pipe>_<producer>_<consumer>:<path>");
        m[item.substr(0, c)] = item.substr(c + 1);
    }
    return m;
}

std::vector<float> ParseList(const std::string& s)
{
    std::vector<float> v;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, '|'))
        if (!item.empty()) v.push_back((float)std::atof(item.c_str()));
    return v;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        Args a;
        for (int i = 1; i < argc; ++i) {
            const char* eq = std::strchr(argv[i], '=');
            if (!eq) ThrowException(std::string("argument without '=': ") + argv[i]);
            a.kv[std::string(argv[i], (size_t)(eq - argv[i]))] = eq + 1;
        }
        ScanParams scan;
        scan.NumViews = (size_t)a.I("views");
        scan.DSO = (float)a.D("dso");
        scan.DSD = (float)a.D("dsd");
        scan.NumDetsU = (size_t)a.I("detu");
        scan.NumDetsV = (size_t)a.I("detv");
        scan.DetectorPixelSize = (float)a.D("pitch");
        scan.BorderSizeUp = (size_t)a.I("border_up", "0");
        scan.BorderSizeDown = (size_t)a.I("border_down", "0");
        scan.BorderSizeLeft = (size_t)a.I("border_left", "0");
        scan.BorderSizeRight = (size_t)a.I("border_right", "0");
        scan.BrightField = (float)a.D("bright_field");
        scan.BeamHardeningParams = ParseList(a.S("bh", "1|0|0"));
        scan.Derive();
        ReconParams recon;
        recon.NumPixelsX = (size_t)a.I("x");
        recon.NumPixelsY = (size_t)a.I("y");
        recon.NumPixelsZ = (size_t)a.I("z");
        recon.PitchXY = (float)a.D("pitch_xy");
        recon.PitchZ = (float)a.D("pitch_z");
        recon.FilterCutOffStart = (float)a.D("cut_start", "0");
        recon.FilterCutOffEnd = (float)a.D("cut_end", "0");
        recon.FilterAdjustPoints = ParseList(a.S("adj_pts", "|"));
        recon.FilterAdjustLevelInDB = ParseList(a.S("adj_db", "|"));
        recon.FilterWindow = (int)a.I("window", "0");
        recon.FilterMode = (int)a.I("filter", "0");
        recon.StorageType = (int)a.I("storage", "0");
        recon.ViewsPerChunk = (size_t)a.I("chunk", "8");
        recon.ZFilter = a.I("zfilter", "0") != 0;
        if (a.kv.count("ct")) {  // ct=norm0|norm1|muWater
            const std::vector<float> ct = ParseList(a.S("ct"));
            if (ct.size() != 3) ThrowException("ct needs norm0|norm1|muWater");
            recon.CTNumber = true;
            recon.CTNumNorm0 = ct[0];
            recon.CTNumNorm1 = ct[1];
            recon.MuWater = ct[2];
        }
        const int device = (int)a.I("device", "0");
        const bool composite = a.I("composite", "0") != 0;
        const size_t frame = scan.NumDetsU * scan.NumDetsV;
        const bool fromFolder = a.kv.count("input_folder") != 0;
        std::vector<uint16_t> raw;
        if (!fromFolder) raw = ReadFile<uint16_t>(a.S("raw"), frame * scan.NumViews);
        recon.MirroringY = a.I("mirror_y", "0") != 0;
        recon.CenterZ = recon.PitchZ * (float)a.D("translation_z", "0");  // ConfigHelper.cpp:85-86
        const bool ptmFile = a.kv.count("ptm_file") != 0;
        std::vector<double> ptm = ptmFile ? ReadPTMFile(a.S("ptm_file"), scan.NumViews)
                                          : ReadFile<double>(a.S("ptm"), 12 * scan.NumViews);
        // PTM.dat holds the scanner's matrices: JEngine adjusts them for the
        // borders, the axis directions and the z translation before use
        // (JEngine.cpp:139-150); raw FP64 matrices (ptm=) are taken as the
        // adjusted ones unless adjust_ptm=1
        if (a.I("adjust_ptm", ptmFile ? "1" : "0") != 0) AdjustProjectionMatrices(ptm, scan, recon, device);
        // JEngine.cpp:153-161, on the adjusted matrices
        const bool detectorOnRight = jh_detector_on_right(ptm.data(), (int)scan.NumUsedDetsU) != 0;
        const auto dumps = ParsePathMap(a.S("dump", ","));
        const auto loads = ParsePathMap(a.S("load", ","));
        const bool toFolder = a.kv.count("output_folder") != 0;
        std::vector<uint16_t> dark;
        std::vector<float> gain;
        if (a.kv.count("dark")) dark = ReadFile<uint16_t>(a.S("dark"), frame);
        if (a.kv.count("gain")) gain = ReadFile<float>(a.S("gain"), scan.NumUsedDetsU * scan.NumUsedDetsV);
        const size_t plane = recon.NumPixelsX * recon.NumPixelsY;
        std::vector<float> vol(plane * recon.NumPixelsZ, 0.0f);

        for (size_t parts = (size_t)a.I("parts", "1"); parts <= recon.NumPixelsZ; ++parts) {
            std::vector<std::shared_ptr<AgentBase>> agents;
            auto bp = std::make_shared<BPHIPAgent>(scan, recon, ptm, parts, device);
            if (fromFolder)
                agents.push_back(std::make_shared<InputAgent>(scan.NumDetsU, scan.NumDetsV, scan.NumViews, parts,
                                                              a.S("input_folder"), a.S("input_template"), composite));
            else
                agents.push_back(std::make_shared<MemoryInputAgent>(raw, frame, scan.NumViews, parts, composite));
            agents.push_back(
                std::make_shared<ProjectionUploadAgent>(scan.NumDetsU, scan.NumDetsV, recon.ViewsPerChunk, device));
            agents.push_back(std::make_shared<PreProcessingHIPAgent>(scan, dark, gain, device, recon.ViewsPerChunk));
            if (composite) agents.push_back(std::make_shared<CompositeHIPAgent>(scan, ptm, detectorOnRight, device));
            agents.push_back(std::make_shared<FilterHIPAgent>(scan, recon, ptm, device, composite, detectorOnRight));
            agents.push_back(bp);
            agents.push_back(std::make_shared<DeviceMemoryDownloadAgent>(plane, device));
            if (toFolder)
                agents.push_back(std::make_shared<OutputAgent>(recon.NumPixelsX, recon.NumPixelsY,
                                                               a.S("output_folder"), a.S("output_template")));
            else
                agents.push_back(std::make_shared<SliceOutputAgent>(vol, plane, *bp, parts));
            try {
                std::vector<std::shared_ptr<PipeBase>> trunk;
                for (auto& ag : agents) ag->SetPipes(trunk);
                // dump / load files by <pipe>_<producer>_<consumer> (JEngine.cpp:349-400)
                std::set<std::shared_ptr<PipeBase>> pipes;
                for (auto& ag : agents) pipes.insert(ag->GetConnectedPipes().begin(), ag->GetConnectedPipes().end());
                for (auto& p : pipes) {
                    const std::string key = p->GetName() + "_" + p->GetProducer() + "_" + p->GetConsumer();
                    if (dumps.count(key)) p->SetDump(dumps.at(key));
                    if (loads.count(key)) p->SetLoad(loads.at(key));
                }
                for (auto& ag : agents) ag->GetReady();
            } catch (Exception& e) {
                if (std::strncmp(e.what(), kDeviceOOMPrefix, std::strlen(kDeviceOOMPrefix)) == 0 ||
                    std::strstr(e.what(), kDeviceOOMPrefix)) {
                    GLog("device memory short with " + std::to_string(parts) + " part(s): " + e.what());
                    continue;
                }
                throw;
            }
            Timer t;
            for (auto& ag : agents) ag->Start();
            for (auto& ag : agents) ag->Join();
            const double span = t.Toc();
            for (auto& ag : agents)
                if (ag->Failed()) ThrowException(ag->GetAgentName() + " failed; see the log");
            for (auto& ag : agents)
                std::cerr << ag->GetAgentName() << ": occupied " << ag->GetOccupiedTime() << " s, async "
                          << ag->GetAsyncTime() << " s\n";
            std::cout << "{\"parts\": " << parts << ", \"seconds\": " << span << ", \"composite\": "
                      << (composite ? 1 : 0) << ", \"detector_on_right\": " << (detectorOnRight ? 1 : 0) << "}"
                      << std::endl;
            if (!toFolder) {
                std::ofstream os(a.S("out"), std::ios::binary);
                os.write(reinterpret_cast<const char*>(vol.data()), (std::streamsize)(vol.size() * sizeof(float)));
                if (!os) ThrowException("cannot write " + a.S("out"));
            }
            return 0;
        }
        ThrowException("device memory is short even with one slice per part");
    } catch (std::exception& e) {
        std::cerr << "jh_fdk_pipeline: " << e.what() << std::endl;
        return 1;
    }
}
