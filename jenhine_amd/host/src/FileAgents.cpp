// On-disk formats and file agents (see FileAgents.h).
#include "jh/FileAgents.h"

#include <cctype>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sys/stat.h>
#include <thread>

namespace jh {

std::vector<double> ReadPTMFile(const std::string& path, size_t numViews)
{
    std::ifstream is(path);
    if (!is) ThrowExceptionAndLog("can not open " + path);
    std::vector<double> ptm(12 * numViews);
    std::string tok;
    for (size_t i = 0; i < ptm.size(); ++i) {
        if (!(is >> tok)) ThrowExceptionAndLog(path + " holds " + std::to_string(i) + " numbers, " +
                                               std::to_string(ptm.size()) + " needed");
        char* end = nullptr;
        errno = 0;
        ptm[i] = std::strtod(tok.c_str(), &end);
        if (errno != 0 || end == tok.c_str() || *end != '\0') ThrowExceptionAndLog(path + ": bad number '" + tok + "'");
    }
    return ptm;
}

void WritePTMFile(const std::string& path, const std::vector<double>& ptm)
{
    if (ptm.size() % 12) ThrowException("WritePTMFile: not a whole number of 3x4 matrices");
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) ThrowExceptionAndLog("can not create " + path);
    for (size_t i = 0; i < ptm.size(); ++i) std::fprintf(f, "%.17g%c", ptm[i], (i % 12 == 11) ? '\n' : ' ');
    const bool ok = std::fclose(f) == 0;
    if (!ok) ThrowExceptionAndLog("writing " + path + " failed");
}

std::string IndexedFileName(const std::string& folder, const std::string& nameTemplate, size_t index)
{
    // exactly one %[0][width]{d,i,u}; "%%" stands for a literal percent
    int conversions = 0;
    for (size_t i = 0; i < nameTemplate.size(); ++i) {
        if (nameTemplate[i] != '%') continue;
        size_t j = i + 1;
        if (j < nameTemplate.size() && nameTemplate[j] == '%') { i = j; continue; }
        while (j < nameTemplate.size() && (nameTemplate[j] == '0' || std::isdigit((unsigned char)nameTemplate[j]))) ++j;
        if (j >= nameTemplate.size() || !std::strchr("diu", nameTemplate[j]))
            ThrowException("file name template '" + nameTemplate + "' has a non-integer conversion");
        ++conversions;
        i = j;
    }
    if (conversions != 1) ThrowException("file name template '" + nameTemplate + "' needs exactly one integer field");
    // normalise the conversion to %zu-compatible unsigned long long formatting
    std::string fmt;
    for (size_t i = 0; i < nameTemplate.size(); ++i) {
        fmt += nameTemplate[i];
        if (nameTemplate[i] != '%') continue;
        if (i + 1 < nameTemplate.size() && nameTemplate[i + 1] == '%') { fmt += '%'; ++i; continue; }
        size_t j = i + 1;
        while (std::isdigit((unsigned char)nameTemplate[j])) fmt += nameTemplate[j++];
        fmt += "llu";
        i = j;
    }
    char buf[4096];
    const int n = std::snprintf(buf, sizeof buf, fmt.c_str(), (unsigned long long)index);
    if (n < 0 || (size_t)n >= sizeof buf) ThrowException("file name too long");
    std::string name(buf);
    if (folder.empty()) return name;
    return folder.back() == '/' ? folder + name : folder + "/" + name;
}

// ---------------------------------------------------------------------------

InputAgent::InputAgent(size_t numDetsU, size_t numDetsV, size_t numViews, size_t numRoundsPerIteration,
                       const std::string& folder, const std::string& nameTemplate, bool pairConjugates)
    : SequentialAgentBase("InputAgent", 1), u_(numDetsU), v_(numDetsV), views_(numViews),
      passes_(numRoundsPerIteration), folder_(folder), template_(nameTemplate), pairs_(pairConjugates)
{
    struct stat st;
    if (stat(folder.c_str(), &st) != 0) ThrowExceptionAndLog(folder + " does not exist.");
    if (!S_ISDIR(st.st_mode)) ThrowExceptionAndLog(folder + " is not a folder");
    if (views_ == 0 || passes_ == 0) ThrowExceptionAndLog("InputAgent: no views or no passes");
    if (pairs_ && views_ % 2) ThrowExceptionAndLog("InputAgent: conjugate pairs need an even view count");
    (void)IndexedFileName(folder_, template_, 0);  // validates the template
}

void InputAgent::SetPipesImpl()
{
    out_ = CreateNewPipe<UINT16Vec>("Proj");
    out_->SetTemplate(UINT16Vec(u_ * v_), {u_, v_});
    out_->SetProducer(GetAgentName(), 1, 1);
    MergePipeToTrunk(out_);
}

void InputAgent::ReadFrame(size_t view, UINT16Vec& frame) const
{
    const std::string name = IndexedFileName(folder_, template_, view);
    std::ifstream is(name, std::ios::binary);
    if (!is) ThrowExceptionAndLog("can not open " + name);
    is.read(reinterpret_cast<char*>(frame.data()), (std::streamsize)(frame.size() * sizeof(uint16_t)));
    if ((size_t)is.gcount() != frame.size() * sizeof(uint16_t)) ThrowExceptionAndLog("short read from " + name);
}

void InputAgent::WorkFlow0()
{
    for (size_t pass = 0; pass < passes_; ++pass)
        for (size_t v = 0; v < views_; ++v) {
            {
                auto wt = out_->GetWriteToken(1, !pairs_ && v + 1 == views_);
                ThreadOccupiedScope occ(this);
                WaitAsyncScope wait(&occ);
                ReadFrame(v, wt.GetBuffer(0));
            }
            if (pairs_) {
                auto wt = out_->GetWriteToken(1, v + 1 == views_);
                ThreadOccupiedScope occ(this);
                WaitAsyncScope wait(&occ);
                ReadFrame((v + views_ / 2) % views_, wt.GetBuffer(0));
            }
        }
    out_->Close();
}

// ---------------------------------------------------------------------------

OutputAgent::OutputAgent(size_t sizeX, size_t sizeY, const std::string& folder, const std::string& nameTemplate)
    : SequentialAgentBase("OutputAgent", 1), n_(sizeX * sizeY), folder_(folder), template_(nameTemplate)
{
    // create the folder (one level; the reference uses create_directories)
    struct stat st;
    if (stat(folder.c_str(), &st) != 0) (void)mkdir(folder.c_str(), 0777);
    if (stat(folder.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) ThrowExceptionAndLog("can not create folder: " + folder);
    (void)IndexedFileName(folder_, template_, 0);
}

void OutputAgent::SetPipesImpl()
{
    in_ = BranchPipeFromTrunk<FloatVec>("Slice");
    in_->SetConsumer(GetAgentName(), 1, 1, 0);
}

void OutputAgent::WorkFlow0()
{
    try {
        for (;;) {
            auto rt = in_->GetReadToken();
            ThreadOccupiedScope occ(this);
            const std::string name = IndexedFileName(folder_, template_, written_);
            WaitAsyncScope wait(&occ);
            for (int attempt = 0;; ++attempt) {
                std::ofstream os(name, std::ios::binary);
                if (os) os.write(reinterpret_cast<const char*>(rt.GetBuffer(0).data()), (std::streamsize)(n_ * sizeof(float)));
                if (os) {
                    os.close();
                    if (os) break;
                }
                GLog("Fail to write file " + name);
                if (attempt == 2) ThrowExceptionAndLog("Error writing output file #" + std::to_string(written_));
                std::this_thread::sleep_for(std::chrono::seconds(1));
            }
            ++written_;
        }
    } catch (PipeClosedAndEmptySignal&) {
    }
}

}  // namespace jh
