// iac_eval_main.cpp -- command-line driver of RunIAC::Run_AutoCorr_DataEval: the evaluate
// run mode of IAC/MainRunModes.cpp:1254-1330 over a settings file (iac_eval -e Settings.xml,
// the hit-event list named in it; arrays are raw binary files), or flags over raw inputs:
//   iac_eval --pixelmap PM --size SS FS --frames STACK --nframes N [options]
// PM: float32 [SS][FS][3]; STACK: float32 [N][SS][FS]; --mask: int32 [SS][FS];
// --rotations: float32 [N][9] row major (default identity).  Outputs are raw
// float64 arrays.  Prints the IAC_Report as one JSON line.
#include <array>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_util.hpp"
#include "iac_host.hpp"

static void usage()
{
    std::cerr << "usage: iac_eval --pixelmap PM --size SS FS --frames STACK --nframes N\n"
                 "  [--mask M] [--rotations R] [--mesh-size S] [--qzoom Z] [--1d]\n"
                 "  [--photon-offset T] [--photon-step S] [--interp1 0|1] [--interp2 0|1] [--no-double-map]\n"
                 "  [--bounds LO HI] [--invert-rot] [--ignore-orientation] [--mask-flatfield] [--batch B]\n"
                 "  [--out-avint F] [--out-acuw F] [--out-cq F] [--out-cq-small F] [--out-final F] [--out-q F]\n"
                 "  [--echo L] [--version]\n"
                 "       iac_eval -e Settings.xml      (evaluation settings file, hit-event list from it)\n"
                 "       iac_eval -eec Example.xml     (write an example evaluation settings file)\n"
                 "       iac_eval --lap STACK N NPIX PANELS ADU SEED COMB OUT\n"
                 "       iac_eval --queue PIXELMAP SS FS PIX CNT OFF NFRAMES EDGE OUT\n"
                 "  (PPP::PhotonFinder_LargestAdjacentPixel frame by frame; PANELS = first:fs:ss,...)\n";
}

// the LAP photon finder over a raw stack, one frame at a time as IAC/PPP.cpp:469-479 loops
static int lap_main(int argc, char **argv)
{
    if (argc != 10) {
        usage();
        return 2;
    }
    const unsigned long n = std::strtoul(argv[3], nullptr, 10), npix = std::strtoul(argv[4], nullptr, 10);
    std::vector<PPP::DetectorPanel> panels;
    unsigned long full = 0;
    for (const char *p = argv[5]; *p;) {
        PPP::DetectorPanel d;
        char *e = nullptr;
        d.FirstInd = (unsigned)std::strtoul(p, &e, 10);
        if (*e != ':') break;
        d.Scans[0] = (unsigned)std::strtoul(e + 1, &e, 10);
        if (*e != ':') break;
        d.Scans[1] = (unsigned)std::strtoul(e + 1, &e, 10);
        panels.push_back(d);
        full += (unsigned long)d.Scans[0] * d.Scans[1];
        p = *e == ',' ? e + 1 : e;
    }
    if (n == 0 || npix == 0 || panels.empty() || full > npix) {
        usage();
        return 2;
    }
    std::vector<float> I(n * npix);
    iac_host::read_raw(argv[2], I.data(), I.size() * sizeof(float));
    for (unsigned long f = 0; f < n; f++)
        PPP::PhotonFinder_LargestAdjacentPixel(I.data() + f * npix, panels, (int)full, std::strtof(argv[6], nullptr),
                                               std::strtof(argv[7], nullptr), std::strtof(argv[8], nullptr));
    iac_host::write_raw(argv[9], I.data(), I.size() * sizeof(float));
    std::cout << "lap ok " << n << std::endl;
    return 0;
}

// Detector::AutoCorrelateSparseList once per event, as RunIAC's frame loop calls it
// (IAC/RunIAC.cpp:524-550), on hit lists {kx, ky, kz, count} built on the host from photon
// lists (PIX uint32 and CNT uint16 per photon, OFF uint64 frame offsets); NN/NN double map
// into the big mesh of edge EDGE.  After one untimed pass, times the calls and the queue's
// flush to the device's completion, then downloads the mesh to OUT (uint64).
static int queue_main(int argc, char **argv)
{
    if (argc != 11) {
        usage();
        return 2;
    }
    const unsigned ss = (unsigned)std::strtoul(argv[3], nullptr, 10), fs = (unsigned)std::strtoul(argv[4], nullptr, 10);
    const unsigned long nframes = std::strtoul(argv[8], nullptr, 10);
    const int edge = (int)std::strtol(argv[9], nullptr, 10);
    if (ss == 0 || fs == 0 || nframes == 0 || edge <= 0) {
        usage();
        return 2;
    }
    Settings S;
    S.echo = false;
    Detector D(S);
    D.LoadPixelMap(argv[2], ss, fs);
    D.Calc_kMap();
    ACMesh M(&S);
    M.CreateBigMeshForDetector(D, edge);
    std::vector<uint64_t> off(nframes + 1);
    iac_host::read_raw(argv[7], off.data(), off.size() * sizeof(uint64_t));
    const uint64_t n = off[nframes];
    std::vector<uint32_t> pix(n);
    std::vector<uint16_t> cnt(n);
    iac_host::read_raw(argv[5], pix.data(), n * sizeof(uint32_t));
    iac_host::read_raw(argv[6], cnt.data(), n * sizeof(uint16_t));
    const size_t npix = D.NumPixels();
    std::vector<std::vector<std::array<float, 4>>> lists(nframes);
    for (unsigned long f = 0; f < nframes; f++) {
        for (uint64_t i = off[f]; i < off[f + 1]; i++) {
            if (pix[i] >= npix) throw std::runtime_error("--queue: pixel index outside the detector");
            const float *k = D.kMap + 3 * (size_t)pix[i];
            lists[f].push_back({k[0], k[1], k[2], (float)cnt[i]});
        }
    }
    Detector::AutoCorrFlags nn;
    // the queue and the device k-map exist before the timed calls (the first event of a run
    // makes them); the device is given back before the calls reserve it one by one
    iac_ctx *ctx = nullptr;
    {
        iac_host::Reserved res(S, "iac_eval --queue");
        ctx = res.ctx;
        M.ctx = ctx;
        M.DeviceMesh();
        M.Queue(D.DeviceKMap(ctx), (uint32_t)npix, nn.InterpolationMode, nn.InterpolationMode, true);
        iac_host::check(iac_ctx_sync(ctx), "iac_ctx_sync");
    }
    auto pass = [&]() {
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned long f = 0; f < nframes; f++) {
            D.SparseHitList.swap(lists[f]);  // the event's list (swapped back below)
            D.Checklist.SparseHitList = true;
            D.AutoCorrelateSparseList(M, nn, nn, true, S, 1);
            D.SparseHitList.swap(lists[f]);
        }
        M.Flush();
        iac_host::check(iac_ctx_sync(ctx), "iac_ctx_sync");
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    // one untimed pass (the process's first launches load their kernels), a cleared mesh, the timed pass
    pass();
    iac_host::check(iac_memset_zero(ctx, M.d_Mesh, M.NumVoxels() * sizeof(uint64_t)), "iac_memset_zero");
    const double ms = pass();
    M.Download();
    iac_host::write_raw(argv[10], M.Mesh, M.NumVoxels() * sizeof(uint64_t));
    std::cout << "queue ms " << ms << " frames " << nframes << std::endl;
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && std::string(argv[1]) == "--lap") return lap_main(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "--queue") {
        try {
            return queue_main(argc, argv);
        } catch (const std::exception &ex) {
            std::cerr << "ERROR: " << ex.what() << "\n";
            return 1;
        }
    }
    if (argc >= 2) {
        // the reference's run modes over settings files (IAC/main.cpp:560-600)
        const std::string a1 = argv[1];
        const bool eval = a1 == "evaluate" || a1 == "-evaluate" || a1 == "e" || a1 == "-e";
        const bool example = a1 == "exampleevaluationconfig" || a1 == "-exampleevaluationconfig" || a1 == "eec" ||
                             a1 == "-eec";
        if (eval || example) {
            if (argc < 3) {
                std::cerr << (eval ? "-Evaluate" : "-ExampleEvaluateConfig")
                          << " requires one additional argument (\"Settings.xml\")\n";
                return -1;
            }
            try {
                Settings S;
                if (example) return MainRunModes::Create_Example_Evaluation_Config_File(argv[2], S);
                std::cout << "Run ICA in evaluation mode - autocorrelation\n";
                return MainRunModes::AutoCorrelateData(argv[2], S);
            } catch (const std::exception &ex) {
                std::cerr << "ERROR: " << ex.what() << "\n";
                return 1;
            }
        }
    }
    RunIAC::CreateDataEval_Settings E;
    E.EchoLevel = 0;
    std::string frames, rotations;
    unsigned int nframes = 0, batch = 0;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                usage();
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--version") {
            std::cout << "iac_eval abi " << iac_abi_version() << "\n";
            return 0;
        } else if (a == "--pixelmap") E.PixelMap_Path = next();
        else if (a == "--size") {
            E.DetectorSize[0] = (unsigned)std::stoul(next());
            E.DetectorSize[1] = (unsigned)std::stoul(next());
        } else if (a == "--frames") frames = next();
        else if (a == "--nframes") nframes = (unsigned)std::stoul(next());
        else if (a == "--mask") E.PixelMask_Path = next();
        else if (a == "--rotations") rotations = next();
        else if (a == "--mesh-size") E.MeshSize = std::stoi(next());
        else if (a == "--qzoom") E.QZoom = std::stof(next());
        else if (a == "--1d") E.AngularAveraged = true;
        else if (a == "--photon-offset") E.PhotonOffset = std::stof(next());
        else if (a == "--photon-step") E.PhotonStep = std::stof(next());
        else if (a == "--interp1") E.AC_FirstMap_Flags.InterpolationMode = (Settings::Interpolation)std::stoi(next());
        else if (a == "--interp2") E.AC_SecondMap_Flags.InterpolationMode = (Settings::Interpolation)std::stoi(next());
        else if (a == "--no-double-map") E.DoubleMap = false;
        else if (a == "--bounds") {
            E.RestrictStackToBoundaries = true;
            E.LowerBoundary = (unsigned)std::stoul(next());
            E.UpperBoundary = (unsigned)std::stoul(next());
        } else if (a == "--invert-rot") E.InvertRotMatrix = true;
        else if (a == "--ignore-orientation") E.IgnoreOrientation = true;
        else if (a == "--mask-flatfield") E.UsePixelMask_as_Flatfield = true;
        else if (a == "--batch") batch = (unsigned)std::stoul(next());
        else if (a == "--out-avint") E.Out_AvIntensity_Path = next();
        else if (a == "--out-acuw") E.Out_ACuw_Path = next();
        else if (a == "--out-cq") E.Out_Cq_Path = next();
        else if (a == "--out-cq-small") E.Out_Cq_small_Path = next();
        else if (a == "--out-final") E.Out_Final_AC_Path = next();
        else if (a == "--out-q") E.Out_Q_Vector = next();
        else if (a == "--echo") E.EchoLevel = std::stoi(next());
        else {
            usage();
            return 2;
        }
    }
    if (E.PixelMap_Path.empty() || frames.empty() || nframes == 0 || !E.DetectorSize[0]) {
        usage();
        return 2;
    }
    try {
        Settings S;
        S.echo = E.EchoLevel > 0;
        if (batch) S.BatchFrames = batch;
        std::vector<float> rot;
        if (!rotations.empty()) {
            rot.resize((size_t)nframes * 9);
            iac_host::read_raw(rotations, rot.data(), rot.size() * 4);
        }
        for (unsigned int i = 0; i < nframes; i++) {
            Settings::HitEvent ev;
            ev.Filename = frames;
            ev.Event = (int)i;
            ev.SerialNumber = (int)i;
            if (!rot.empty()) std::memcpy(ev.RotMatrix, rot.data() + 9 * (size_t)i, 36);
            S.HitEvents.push_back(ev);
        }
        RunIAC::IAC_Report R = RunIAC::Run_AutoCorr_DataEval(S, E);
        const iac_host::IngestStats &st = iac_host::ingest_stats();
        std::printf("{\"dQperVox\": %.9g, \"FinalMeshSize\": %u, \"SmallCqMeshSize\": [%u, %u, %u], "
                    "\"MeanPhotonCount\": %.9g, \"LowerBound\": %u, \"UpperBound\": %u, \"ingest\": "
                    "{\"frames\": %llu, \"read_s\": %.6f, \"wait_s\": %.6f, \"upload_s\": %.6f, "
                    "\"loop_s\": %.6f}}\n",
                    R.dQperVox, R.FinalMeshSize, R.SmallCqMeshSize[0], R.SmallCqMeshSize[1], R.SmallCqMeshSize[2],
                    R.MeanPhotonCount, R.LowerBound, R.UpperBound, (unsigned long long)st.frames, st.read_seconds,
                    st.wait_seconds, st.upload_seconds, st.loop_seconds);
    } catch (const std::exception &ex) {
        std::cerr << "ERROR: " << ex.what() << "\n";
        return 1;
    }
    return 0;
}
