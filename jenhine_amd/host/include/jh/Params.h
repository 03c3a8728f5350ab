// Scan and reconstruction parameters of the reconstruction path, with the
// field names of the reference (Config/ScanParams.h:13-40,
// Config/ReconParams.h:14-68) for the fields this path reads, and the derived
// fields computed as Config/ConfigHelper.cpp:52-63 does.
#pragma once

#include <cstddef>
#include <vector>

namespace jh {

struct ScanParams {
    size_t NumViews = 0;
    float DSO = 0.0f;
    float DSD = 0.0f;
    size_t NumDetsU = 0;
    size_t NumDetsV = 0;
    float DetectorPixelSize = 0.0f;
    size_t BorderSizeUp = 0;
    size_t BorderSizeDown = 0;
    size_t BorderSizeLeft = 0;
    size_t BorderSizeRight = 0;
    float BrightField = 0.0f;
    std::vector<float> BeamHardeningParams{1.0f, 0.0f, 0.0f};

    // derived
    float HalfSampleRate = 0.0f;
    size_t NumUsedDetsU = 0;
    size_t NumUsedDetsV = 0;

    void Derive()
    {
        const float detectorSizeAtISO = DetectorPixelSize / DSD * DSO;
        HalfSampleRate = 0.5f * (1.0f / detectorSizeAtISO);
        NumUsedDetsU = NumDetsU - BorderSizeLeft - BorderSizeRight;
        NumUsedDetsV = NumDetsV - BorderSizeUp - BorderSizeDown;
    }
};

struct ReconParams {
    size_t NumPixelsX = 0;
    size_t NumPixelsY = 0;
    size_t NumPixelsZ = 0;
    float PitchXY = 0.0f;
    float PitchZ = 0.0f;
    float FilterCutOffStart = 0.0f;
    float FilterCutOffEnd = 0.0f;
    std::vector<float> FilterAdjustPoints;
    std::vector<float> FilterAdjustLevelInDB;
    // extensions of this back end
    int FilterWindow = 0;    // JH_WINDOW_REFERENCE or JH_WINDOW_SHEPP_LOGAN
    int FilterMode = 0;      // JH_FILTER_FP32, _MFMA_BF16X3, _MFMA_BF16X6, _FP32_VALU or _MFMA_F16X3
    int StorageType = 0;     // JH_F32, JH_BF16 or JH_F16
    size_t ViewsPerChunk = 8;
    // the production chain's two stages around the backprojection, fused
    // into it (ZFilterAgent between filter and upload, CTNumAgent after the
    // download: JEngine.cpp:240-245,297-305)
    bool ZFilter = false;
    bool CTNumber = false;
    float CTNumNorm0 = 1.0f;  // ReconParams::CTNumNorm0 (ConfigHelper.cpp:91)
    float CTNumNorm1 = 1.0f;  // ReconParams::CTNumNorm1 (ConfigHelper.cpp:92)
    float MuWater = 0.02f;    // ReconParams::MuWater (ConfigHelper.cpp:70)
    // the projection-matrix adjustment of JEngine.cpp:139-150
    bool MirroringY = false;  // ReconParams::MirroringY (ConfigHelper.cpp:82)
    float CenterZ = 0.0f;     // ReconParams::CenterZ = PitchZ * Volume_Translation_Z (ConfigHelper.cpp:85-86)
};

}  // namespace jh
