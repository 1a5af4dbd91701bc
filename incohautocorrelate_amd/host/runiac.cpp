// runiac.cpp -- RunIAC::Run_AutoCorr_DataEval (IAC/RunIAC.cpp:185-632) and
// Merge_ACandCQ (:120-158): the evaluation sequence with every pass on the device.
// 3D: averaged intensity -> C(q) small -> rotation merge into the big C(q) ->
// AC_uw over all events (sparsified in device stacks, one batched pair pass per
// stack into the resident mesh) -> AC = AC_uw / C(q).  1D: C(q) vector, AC_uw
// vector, ratio.  Outputs are raw binary arrays as SafeArrayToFile(Binary) writes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "host_util.hpp"
#include "iac_host.hpp"

using iac_host::check;

namespace RunIAC {

// IAC/RunIAC.cpp:120-158: Output = IntToFloat(AC.Mesh) / CQ.CQMesh, 0 where CQ <= 0
// or NaN; both meshes must be resident on the same device.
void Merge_ACandCQ(double *&Output, ACMesh &AC, ACMesh &CQ, Settings &)
{
    if (!AC.d_Mesh || !CQ.d_CQMesh) throw std::runtime_error("Merge_ACandCQ: AC or C(q) mesh not calculated");
    if (AC.ctx != CQ.ctx) throw std::runtime_error("Merge_ACandCQ: meshes on different devices");
    AC.Flush();  // events queued by Detector::AutoCorrelateSparseList
    if (AC.NumVoxels() != CQ.NumVoxels()) {
        std::cerr << "ERROR: AC and C(q) meshes differ in size\n    -> in Merge_ACandCQ()\n";
        throw std::runtime_error("Merge_ACandCQ: mesh size mismatch");
    }
    const size_t n = AC.NumVoxels();
    iac_host::DevBuf out(AC.ctx, n * 8);
    check(iac_merge_ac_cq(AC.ctx, AC.d_Mesh, CQ.d_CQMesh, n, out.as<double>()), "iac_merge_ac_cq in Merge_ACandCQ()");
    delete[] Output;
    Output = new double[n];
    out.download(Output, n * 8);
}

static void save(const std::string &path, const void *p, size_t bytes, const char *what, int echo)
{
    if (path.empty()) return;
    iac_host::write_raw(path, p, bytes);
    if (echo > 0) std::cout << " -> Saved " << what << " as \"" << path << "\"\n";
}

IAC_Report Run_AutoCorr_DataEval(Settings &PrgSettings, CreateDataEval_Settings EvalSettings)
{
    if (EvalSettings.PixelMap_Path.empty())
        throw std::runtime_error("Run_AutoCorr_DataEval: no pixel map path set => can't load Hit Events");
    if (!EvalSettings.DetectorSize[0] || !EvalSettings.DetectorSize[1]) {
        // the raw pixel map holds [ss][fs][3] floats; without a stated size, one row of all pixels
        std::ifstream f(EvalSettings.PixelMap_Path, std::ios::binary | std::ios::ate);
        if (!f) throw std::runtime_error("Run_AutoCorr_DataEval: cannot open pixel map " + EvalSettings.PixelMap_Path);
        const long long bytes = (long long)f.tellg();
        if (bytes <= 0 || bytes % 12) throw std::runtime_error("Run_AutoCorr_DataEval: pixel map is not [n][3] floats");
        EvalSettings.DetectorSize[0] = 1;
        EvalSettings.DetectorSize[1] = (unsigned int)(bytes / 12);
    }
    Detector Det(PrgSettings);
    Det.LoadPixelMap(EvalSettings.PixelMap_Path, EvalSettings.DetectorSize[0], EvalSettings.DetectorSize[1]);
    if (EvalSettings.DetDisturb)
        throw std::runtime_error("Run_AutoCorr_DataEval: pixel-map disturbance (simulated data) is not supported");
    Det.Calc_kMap();
    Det.LoadPixelMask(EvalSettings.PixelMask_Path);
    return Run_AutoCorr_DataEval(PrgSettings, EvalSettings, Det);
}

IAC_Report Run_AutoCorr_DataEval(Settings &PrgSettings, CreateDataEval_Settings EvalSettings, Detector &Det)
{
    IAC_Report Report;
    iac_host::ingest_stats() = iac_host::IngestStats();
    const int echo = EvalSettings.EchoLevel;
    if (!Det.Checklist.KMap) throw std::runtime_error("Run_AutoCorr_DataEval: detector k-map not calculated");
    if (!Det.Checklist.PixelMask) Det.GenerateFlatOnesPixelMask();
    if (!EvalSettings.XML_Path.empty()) {  // IAC/RunIAC.cpp:227-240
        PrgSettings.Echo("Load Hit Event List");
        PrgSettings.LoadHitEventListFromFile(EvalSettings.XML_Path);
    }
    if (PrgSettings.HitEvents.empty()) throw std::runtime_error("Run_AutoCorr_DataEval: no hit events");
    if (EvalSettings.InvertRotMatrix) PrgSettings.InvertRotationMatrices();
    unsigned int lowerBound = 0, upperBound = (unsigned int)PrgSettings.HitEvents.size();
    if (EvalSettings.RestrictStackToBoundaries) {
        if (EvalSettings.UpperBoundary > PrgSettings.HitEvents.size())
            throw std::runtime_error("Run_AutoCorr_DataEval: upper hit boundary exceeds the number of hits");
        lowerBound = EvalSettings.LowerBoundary;
        upperBound = EvalSettings.UpperBoundary;
    }
    if (upperBound <= lowerBound) throw std::runtime_error("Run_AutoCorr_DataEval: empty event range");
    Report.LowerBound = lowerBound;
    Report.UpperBound = upperBound;
    if (EvalSettings.IgnoreOrientation) {
        const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (auto &ev : PrgSettings.HitEvents) std::copy(id, id + 9, ev.RotMatrix);
    }
    if (EvalSettings.FractionalCq) throw std::runtime_error("Fractional C(q) is not implemented");
    const size_t P = Det.NumPixels();

    // averaged intensity (flat field)
    if (EvalSettings.UseExistingAvInt && !EvalSettings.UsePixelMask_as_Flatfield) {
        PrgSettings.Echo("Load existing averaged intensity");
        delete[] Det.Intensity;
        Det.Intensity = new float[P];
        iac_host::read_raw(EvalSettings.Out_AvIntensity_Path, Det.Intensity, P * 4);
        Det.ApplyPixelMask();
        Det.Checklist.Intensity = true;
    } else if (!EvalSettings.UsePixelMask_as_Flatfield) {
        PrgSettings.Echo("Load and average all intensities");
        Det.LoadAndAverageIntensity(PrgSettings.HitEvents, EvalSettings.PhotonOffset, EvalSettings.PhotonStep,
                                    (int)lowerBound, (int)upperBound, true);
        save(EvalSettings.Out_AvIntensity_Path, Det.Intensity, P * 4, "averaged intensity", echo);
        Det.ApplyPixelMask();
    } else {
        PrgSettings.Echo("Use pixel-mask as flat-field (instead of averaged intensity)");
        delete[] Det.Intensity;
        Det.Intensity = new float[P];
        for (size_t i = 0; i < P; i++) Det.Intensity[i] = (float)Det.PixelMask[i];
        Det.Checklist.Intensity = true;
    }

    if (EvalSettings.AngularAveraged) {
        const Settings::Interpolation AvInterpol = Settings::Linear;  // IAC/RunIAC.cpp:344
        AC1D Vector_AC;
        Vector_AC.Initialize(Det, EvalSettings.MeshSize, EvalSettings.QZoom);
        Vector_AC.EchoLevel = echo;
        Vector_AC.Calculate_CQ(Det, PrgSettings, AvInterpol);
        save(EvalSettings.Out_Cq_Path, Vector_AC.CQ, Vector_AC.Shape.Size * 8, "angular averaged C(q)", echo);
        std::array<float, 2> Photonisation = {EvalSettings.PhotonOffset, EvalSettings.PhotonStep};
        Vector_AC.Calculate_AC_UW_MR(PrgSettings, Det, AvInterpol, Photonisation, lowerBound, upperBound);
        save(EvalSettings.Out_ACuw_Path, Vector_AC.AC_UW, Vector_AC.Shape.Size * 8, "angular averaged AC_uw", echo);
        Vector_AC.CalcAC();
        save(EvalSettings.Out_Final_AC_Path, Vector_AC.AC, Vector_AC.Shape.Size * 8, "angular averaged AC", echo);
        if (!EvalSettings.Out_Q_Vector.empty()) {
            Vector_AC.CreateQVector();
            save(EvalSettings.Out_Q_Vector, Vector_AC.Q, Vector_AC.Shape.Size * 8, "q vector", echo);
        }
        Report.dQperVox = Vector_AC.Shape.dq_per_Step;
        Report.FinalMeshSize = Vector_AC.Shape.Size;
        return Report;
    }

    // 3D: C(q) small, merged into the lab-frame C(q)
    ACMesh smallMesh(&PrgSettings);
    smallMesh.CreateSmallMesh_CofQ_ForDetector(Det, EvalSettings.MeshSize, EvalSettings.QZoom);
    Detector::AutoCorrFlags Flags;
    Flags.InterpolationMode = EvalSettings.AC_FirstMap_Flags.InterpolationMode;
    Det.AutoCorrelate_CofQ_SmallMesh(smallMesh, Flags, PrgSettings);
    if (!EvalSettings.Out_Cq_small_Path.empty()) {
        smallMesh.Download();
        save(EvalSettings.Out_Cq_small_Path, smallMesh.CQMesh, smallMesh.NumVoxels() * 8, "small C(q)", echo);
    }
    ACMesh Mesh_CQ(&PrgSettings);
    Mesh_CQ.CreateBigMesh_CofQ_ForDetector(Det, EvalSettings.MeshSize, EvalSettings.QZoom);
    Det.Merge_smallCofQ(Mesh_CQ, smallMesh, PrgSettings.HitEvents, lowerBound, upperBound, PrgSettings,
                        EvalSettings.AC_SecondMap_Flags);
    if (!EvalSettings.Out_Cq_Path.empty()) {
        Mesh_CQ.Download();
        save(EvalSettings.Out_Cq_Path, Mesh_CQ.CQMesh, Mesh_CQ.NumVoxels() * 8, "C(q)", echo);
    }

    // AC_uw: the frame loop of IAC/RunIAC.cpp:524-550 in device stacks
    ACMesh AC_uw(&PrgSettings);
    AC_uw.CreateBigMeshForDetector(Det, EvalSettings.MeshSize, EvalSettings.QZoom);
    iac_ctx *ctx = nullptr;
    {
    iac_host::Reserved res(PrgSettings, "Run_AutoCorr_DataEval");
    ctx = res.ctx;
    if (Mesh_CQ.ctx != ctx) throw std::runtime_error("Run_AutoCorr_DataEval: C(q) on another device");
    AC_uw.ctx = ctx;
    const uint32_t N = upperBound - lowerBound;
    const uint32_t B = std::max<uint32_t>(
        1, std::min<uint32_t>(PrgSettings.BatchFrames, (uint32_t)((1ull << 31) / (P * 20))));
    iac_host::DevBuf frames(ctx, (size_t)B * P * 4), rec(ctx, (size_t)B * P * 16), off(ctx, (size_t)(B + 1) * 8),
        rots(ctx, (size_t)B * 36);
    std::vector<float> hrot((size_t)B * 9);
    const float *kmap = Det.DeviceKMap(ctx);
    const int *mask = Det.DeviceMask(ctx);
    iac_mesh_shape shape = AC_uw.CShape();
    uint64_t *mesh = AC_uw.DeviceMesh();
    // frames read one stack ahead on a reader thread into page-locked stacks
    iac_host::FrameStream stream(PrgSettings, PrgSettings.HitEvents, lowerBound, N, P, B);
    const auto t_loop = std::chrono::steady_clock::now();
    double up_s = 0;
    uint32_t b0 = 0, nb = 0;
    while (const float *stack = stream.next(nb)) {
        bool identity = true;
        for (uint32_t f = 0; f < nb; f++) {
            const Settings::HitEvent &ev = PrgSettings.HitEvents[lowerBound + b0 + f];
            std::copy(ev.RotMatrix, ev.RotMatrix + 9, hrot.begin() + 9 * f);
            static const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            identity = identity && std::equal(ev.RotMatrix, ev.RotMatrix + 9, id);
        }
        const auto t_up = std::chrono::steady_clock::now();
        frames.upload(stack, (size_t)nb * P * 4);
        up_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_up).count();
        stream.release();
        if (!identity) rots.upload(hrot.data(), (size_t)nb * 36);
        uint64_t n = 0;
        check(iac_sparse_hit_list(ctx, kmap, (uint32_t)P, frames.as<float>(), nb, mask, EvalSettings.PhotonOffset,
                                  EvalSettings.PhotonStep, rec.as<float>(), off.as<uint64_t>(), &n),
              "iac_sparse_hit_list in Run_AutoCorr_DataEval()");
        check(iac_ac3d_records_batch(ctx, kmap, (uint32_t)P, rec.as<float>(), off.as<uint64_t>(), nb, n,
                                     identity ? nullptr : rots.as<float>(), &shape,
                                     EvalSettings.AC_FirstMap_Flags.InterpolationMode,
                                     EvalSettings.AC_SecondMap_Flags.InterpolationMode, EvalSettings.DoubleMap ? 1 : 0,
                                     mesh),
              "iac_ac3d_records_batch in Run_AutoCorr_DataEval()");
        if (echo > 2) std::cout << "AC_uw: " << (b0 + nb) << "/" << N << " events\n";
        b0 += nb;
    }
    check(iac_ctx_sync(ctx), "iac_ctx_sync");
    iac_host::IngestStats &st = iac_host::ingest_stats();
    st.frames += N;
    st.read_seconds += stream.read_seconds();
    st.wait_seconds += stream.wait_seconds();
    st.upload_seconds += up_s;
    st.loop_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
    }
    // Report.MeanPhotonCount += PhotonCount per event, then / N (IAC/RunIAC.cpp:528, 553)
    {
        float m = 0.0f;
        for (uint32_t i = lowerBound; i < upperBound; i++) m += PrgSettings.HitEvents[i].PhotonCount;
        Report.MeanPhotonCount = m / ((float)(upperBound - lowerBound));
    }
    if (!EvalSettings.Out_ACuw_Path.empty()) {
        AC_uw.Download();
        std::vector<double> acd(AC_uw.NumVoxels());
        for (size_t i = 0; i < acd.size(); i++) acd[i] = PrgSettings.IntToFloat(AC_uw.Mesh[i]);
        save(EvalSettings.Out_ACuw_Path, acd.data(), acd.size() * 8, "AC_uw", echo);
    }
    double *FinalAC = nullptr;
    Merge_ACandCQ(FinalAC, AC_uw, Mesh_CQ, PrgSettings);
    save(EvalSettings.Out_Final_AC_Path, FinalAC, AC_uw.NumVoxels() * 8, "final AC", echo);
    delete[] FinalAC;
    Report.dQperVox = AC_uw.Shape.dq_per_Voxel;
    Report.FinalMeshSize = AC_uw.Shape.Size_AB;
    Report.SmallCqMeshSize[0] = Mesh_CQ.Shape.Size_AB;
    Report.SmallCqMeshSize[1] = Mesh_CQ.Shape.Size_AB;
    Report.SmallCqMeshSize[2] = Mesh_CQ.Shape.Size_C;
    return Report;
}

}  // namespace RunIAC
