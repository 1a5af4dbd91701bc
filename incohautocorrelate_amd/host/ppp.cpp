// ppp.cpp -- PPP (IAC/PPP.h, IAC/PPP.cpp): the per-frame operations of pattern pre-processing,
// run on the device through iac_ppp_lap.  The HDF5 / XML handling of PPP::ProcessData_PF_LAP
// (file reads, hyperslab writes, the event list) stays with the caller.
#include <stdexcept>
#include <vector>

#include "host_util.hpp"
#include "iac_host.hpp"

namespace PPP {

namespace {

// The reference's PhotonFinder_LargestAdjacentPixel takes no Settings; its device comes from
// this pool, created on first use and kept for the life of the process (no context is torn
// down after the HIP runtime during exit).
Settings &default_pool()
{
    static Settings *pool = [] {
        auto *s = new Settings();
        s->echo = false;
        return s;
    }();
    return *pool;
}

}  // namespace

void PhotonFinder_LargestAdjacentPixel(float *Intensity, std::vector<DetectorPanel> DetectorPanels, int FullDetSize,
                                       float ADU_perPhoton, float SeedThershold, float CombinedThershold)
{
    PhotonFinder_LargestAdjacentPixel(Intensity, DetectorPanels, FullDetSize, ADU_perPhoton, SeedThershold,
                                      CombinedThershold, default_pool());
}

void PhotonFinder_LargestAdjacentPixel(float *Intensity, const std::vector<DetectorPanel> &DetectorPanels,
                                       int FullDetSize, float ADU_perPhoton, float SeedThershold,
                                       float CombinedThershold, Settings &Options)
{
    if (!Intensity || FullDetSize <= 0)
        throw std::invalid_argument("PPP::PhotonFinder_LargestAdjacentPixel: no intensity");
    std::vector<iac_panel> pan;
    pan.reserve(DetectorPanels.size());
    for (const DetectorPanel &p : DetectorPanels) pan.push_back({p.FirstInd, p.Scans[0], p.Scans[1]});
    iac_host::Reserved r(Options, "PPP::PhotonFinder_LargestAdjacentPixel");
    iac_ctx *ctx = r.ctx;
    // the reference's arrays span the panels' FullDetSize pixels; the rest of the frame is untouched
    const size_t bytes = (size_t)FullDetSize * sizeof(float);
    iac_host::DevBuf buf(ctx, bytes);
    buf.upload(Intensity, bytes);
    Options.checkErr(iac_ppp_lap(ctx, buf.as<float>(), buf.as<float>(), 1, (uint32_t)FullDetSize, nullptr, nullptr,
                                 nullptr, pan.data(), (uint32_t)pan.size(), ADU_perPhoton, SeedThershold,
                                 CombinedThershold, 0, nullptr, nullptr),
                     "iac_ppp_lap in PPP::PhotonFinder_LargestAdjacentPixel()");
    buf.download(Intensity, bytes);
}

}  // namespace PPP
