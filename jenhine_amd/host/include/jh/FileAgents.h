// The reference's on-disk formats and the agents that read and write them
// (host code, no device calls):
//
//   ReadPTMFile / WritePTMFile  PTM.dat: 12 numbers per view, whitespace
//                               separated, row-major 3x4 (JEngine/PTMReader.h:17-23)
//   InputAgent                  raw frames <folder>/<template % view>, NumDetsU x
//                               NumDetsV u16 each (JEngine/InputAgent.cpp:39-91)
//                               -> Pipe<UINT16Vec> "Proj"
//   OutputAgent                 Pipe<FloatVec> "Slice" -> <folder>/<template % index>,
//                               X x Y f32 each (JEngine/OutputAgent.cpp:37-82)
//
// Templates hold one integer conversion (e.g. "Proj_%d.prj", "slice_%04d.slc"),
// the way the reference formats them with swprintf.
#pragma once

#include <string>
#include <vector>

#include "Agent.h"

namespace jh {

std::vector<double> ReadPTMFile(const std::string& path, size_t numViews);
void WritePTMFile(const std::string& path, const std::vector<double>& ptm);

// folder/template with the integer conversion replaced by index.  Throws when
// the template does not hold exactly one integer conversion.
std::string IndexedFileName(const std::string& folder, const std::string& nameTemplate, size_t index);

// Reads the scan's raw frames, numRoundsPerIteration times over (one shot per
// pass: the z-slab parts of the device-shortage retry).  With
// pairConjugates each view v is followed by its conjugate (v + N/2) % N, the
// pairing the offset-detector composite reads (InputAgent.cpp:47-85); the
// shot then holds 2N frames.  Unlike the reference, the pipe is closed once,
// after the last pass (the reference closes it inside the pass loop, which
// breaks more than one pass).
class InputAgent : public SequentialAgentBase {
public:
    InputAgent(size_t numDetsU, size_t numDetsV, size_t numViews, size_t numRoundsPerIteration,
               const std::string& folder, const std::string& nameTemplate, bool pairConjugates = false);

private:
    void SetPipesImpl() override;
    void WorkFlow0() override;
    void ReadFrame(size_t view, UINT16Vec& frame) const;
    const size_t u_, v_, views_, passes_;
    const std::string folder_, template_;
    const bool pairs_;
    std::shared_ptr<Pipe<UINT16Vec>> out_;
};

// Writes every slice of the "Slice" pipe to its own file, the index running
// over all slices received; a failed write is retried twice, one second
// apart (OutputAgent.cpp:53-76).
class OutputAgent : public SequentialAgentBase {
public:
    OutputAgent(size_t sizeX, size_t sizeY, const std::string& folder, const std::string& nameTemplate);
    size_t SlicesWritten() const { return written_; }

private:
    void SetPipesImpl() override;
    void WorkFlow0() override;
    const size_t n_;
    const std::string folder_, template_;
    size_t written_ = 0;
    std::shared_ptr<Pipe<FloatVec>> in_;
};

}  // namespace jh
