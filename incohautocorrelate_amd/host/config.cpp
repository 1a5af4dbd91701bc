// config.cpp -- the reference's configuration files without Boost: the hit-event list
// (Settings::LoadHitEventListFromFile / SafeHitEventListToFile, IAC/Settings.cpp:473-676)
// and the evaluation settings (MainRunModes::LoadEvaluationSettings, IAC/MainRunModes.cpp:32-91,
// the example file of :403-500 and the evaluate run mode of :1254-1330), read and written
// through the property tree of ptree.hpp in the same key layout, so files written by the
// reference load here and the other way round.
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "host_util.hpp"
#include "iac_host.hpp"
#include "ptree.hpp"

using iac_host::PTree;

// IAC/Settings.cpp:563-676.  The list lives under root.Content.N<i> (version >= 2) with a
// "a1; a2; ...; c3" Rotation string; version 1 files have R0..R8 keys; the oldest files
// root.content.<i> or root.content.N<i>.  A missing Info.Size reads as 2^32 - 1 as in the
// reference (so the first missing entry throws); Size 0 leaves HitEvents untouched.
void Settings::LoadHitEventListFromFile(std::string Filename)
{
    const PTree pt = iac_host::load_property_tree(Filename);
    const unsigned int Size = pt.get("root.Info.Size", (unsigned int)-1);
    const unsigned int Version = pt.get("root.Info.Version", 1u);
    if (Size == 0) return;
    HitEvents.clear();
    std::string pathbase;
    if (!pt.find("root.content.N0.Filename") && !pt.find("root.Content.N0.Filename")) {
        pathbase = "root.content.";
    } else {
        pathbase = "root.content.N";
        if (Version >= 2) pathbase = "root.Content.N";
    }
    for (unsigned int i = 0; i < Size; i++) {
        const std::string path = pathbase + std::to_string(i);
        HitEvent tmp;
        tmp.Filename = pt.get_required(path + ".Filename");
        tmp.Dataset = pt.get_required(path + ".Dataset");
        tmp.Event = pt.get_required_int(path + ".Event");
        tmp.SerialNumber = pt.get_required_int(path + ".SerialNumber");
        tmp.MeanIntensity = pt.get_required_float(path + ".MeanIntensity");
        tmp.PhotonCount = pt.get_required_int(path + ".PhotonCount");
        tmp.SupplementInfo = pt.get(path + ".SupplementInfo", std::string());
        if (Version == 1) {
            static const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            for (int j = 0; j < 9; j++) tmp.RotMatrix[j] = pt.get(path + ".R" + std::to_string(j), id[j]);
        } else {
            // "a1; a2; a3; b1; b2; b3; c1; c2; c3", each entry read like atof; more than nine
            // entries: a warning and the identity
            std::string in = pt.get(path + ".Rotation", std::string("1.0; 0.0; 0.0; 0.0; 1.0; 0.0; 0.0; 0.0; 1.0"));
            float val[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            int count = 0;
            bool ok = true;
            size_t pos;
            while ((pos = in.find(';')) != std::string::npos) {
                if (count > 7) {
                    ok = false;
                    break;
                }
                val[count++] = (float)std::atof(in.substr(0, pos).c_str());
                in.erase(0, pos + 1);
            }
            if (ok) {
                val[count] = (float)std::atof(in.c_str());
                std::copy(val, val + 9, tmp.RotMatrix);
            } else {
                std::cerr << "WARNING: Invalid rotation matrix format! use \"a1; a2; a3; b1; b2; b3; c1; c2; c3\" \n"
                             "    -> Rotation matrix is set to \"1.0; 0.0; 0.0; 0.0; 1.0; 0.0; 0.0; 0.0; 1.0 \"\n";
                static const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
                std::copy(id, id + 9, tmp.RotMatrix);
            }
        }
        HitEvents.push_back(tmp);
    }
}

void Settings::SafeHitEventListToFile(std::string Filename) { SafeHitEventListToFile(Filename, HitEvents); }

// IAC/Settings.cpp:478-536: Info.Size, Info.Version (XML_HITLIST_VERSION), optional extra
// Info keys, and per event N<i> with the rotation as "r0; r1; ...; r8" (std::to_string).
void Settings::SafeHitEventListToFile(std::string Filename, std::vector<HitEvent> &HitEventList,
                                      bool AdditionalInformations,
                                      std::unordered_map<std::string, std::string> AdditioInfoMap)
{
    PTree pt;
    pt.put("root.Info.Size", std::to_string(HitEventList.size()));
    pt.put("root.Info.Version", std::to_string(XML_HITLIST_VERSION));
    if (AdditionalInformations)
        for (const auto &kv : AdditioInfoMap) pt.put("root.Info." + kv.first, kv.second);
    for (size_t i = 0; i < HitEventList.size(); i++) {
        const std::string path = "root.Content.N" + std::to_string(i);
        const HitEvent &e = HitEventList[i];
        pt.put(path + ".Filename", e.Filename);
        pt.put(path + ".Dataset", e.Dataset);
        pt.put(path + ".Event", std::to_string(e.Event));
        pt.put(path + ".SerialNumber", std::to_string(e.SerialNumber));
        pt.put(path + ".MeanIntensity", iac_host::format_float(e.MeanIntensity));
        pt.put(path + ".PhotonCount", std::to_string(e.PhotonCount));
        pt.put(path + ".SupplementInfo", e.SupplementInfo);
        std::string rm = std::to_string(e.RotMatrix[0]);
        for (int j = 1; j < 9; j++) rm += "; " + std::to_string(e.RotMatrix[j]);
        pt.put(path + ".Rotation", rm);
    }
    iac_host::save_property_tree(pt, Filename);
}

namespace MainRunModes {

// IAC/MainRunModes.cpp:32-91, key for key with the reference's defaults.  One addition for
// raw pixel maps: root.EvalSettings.InputFiles.PixelMap_Size ("SS; FS"); without it the
// size comes from the pixel map file (ss = 1, fs = bytes / 12).
RunIAC::CreateDataEval_Settings LoadEvaluationSettings(std::string Filename, Settings &Options)
{
    RunIAC::CreateDataEval_Settings EVS;
    const PTree pt = iac_host::load_property_tree(Filename);
    const std::string in = "root.EvalSettings.InputFiles.", out = "root.EvalSettings.OutputFiles.",
                      mesh = "root.EvalSettings.OutputMesh.", ph = "root.EvalSettings.PhotonThreshold.",
                      misc = "root.EvalSettings.Misc.";
    EVS.XML_Path = pt.get(in + "EventList_Path", std::string());
    EVS.PixelMap_Path = pt.get(in + "PixelMap_Path", std::string());
    EVS.PixelMap_DataSet = pt.get(in + "PixelMap_Dataset", std::string());
    EVS.PixelMask_Path = pt.get(in + "PixelMask_Path", std::string());
    EVS.PixelMask_Dataset = pt.get(in + "PixelMask_Dataset", std::string());
    EVS.UseExistingAvInt = pt.get(in + "UseExistingAvInt", false);
    const std::string size = pt.get(in + "PixelMap_Size", std::string());
    if (!size.empty()) {
        const size_t s = size.find_first_of(";, x");
        if (s == std::string::npos) throw std::runtime_error("PixelMap_Size must read \"SS; FS\"");
        EVS.DetectorSize[0] = (unsigned int)std::stoul(size.substr(0, s));
        EVS.DetectorSize[1] = (unsigned int)std::stoul(size.substr(size.find_first_not_of(";, x", s)));
    }
    EVS.Out_Final_AC_Path = pt.get(out + "AC_Final_Path", std::string());
    EVS.Out_AvIntensity_Path = pt.get(out + "AvIntensity_Path", std::string());
    EVS.Out_Cq_Path = pt.get(out + "Cq_Merged_Path", std::string());
    EVS.Out_Cq_small_Path = pt.get(out + "Cq_small_Path", std::string());
    EVS.Out_ACuw_Path = pt.get(out + "ACuw_Path", std::string());
    EVS.Out_Q_Vector = pt.get(out + "Q_Vector", std::string());
    EVS.MeshSize = pt.get(mesh + "MeshSize", 201);
    EVS.QZoom = pt.get(mesh + "QZoom", 1.0f);
    EVS.AngularAveraged = pt.get(mesh + "AngularAveraged", false);
    EVS.DoubleMap = pt.get(mesh + "DoubleMap", true);
    EVS.AC_FirstMap_Flags.InterpolationMode =
        (Settings::Interpolation)pt.get(mesh + "AC_FirstMap_Flags.InterpolationMode", (int)Settings::NearestNeighbour);
    EVS.AC_SecondMap_Flags.InterpolationMode =
        (Settings::Interpolation)pt.get(mesh + "AC_SecondMap_Flags.InterpolationMode", (int)Settings::NearestNeighbour);
    EVS.PhotonOffset = pt.get(ph + "PhotonOffset", 0.0f);
    EVS.PhotonStep = pt.get(ph + "PhotonStep", 1.0f);
    EVS.RestrictStackToBoundaries = pt.get(misc + "RestrictStackToBoundaries", false);
    EVS.LowerBoundary = pt.get(misc + "LowerBoundary", 0u);
    EVS.UpperBoundary = pt.get(misc + "UpperBoundary", 10u);
    EVS.UsePixelMask_as_Flatfield = pt.get(misc + "UsePixelmaskAsFlatField", false);
    EVS.FractionalCq = pt.get(misc + "FractionalCq", false);
    EVS.SizeOfCqFraction = pt.get(misc + "SizeOfCqFraction", 10u);
    EVS.InvertRotMatrix = pt.get(misc + "InvertRotMatrix", false);
    EVS.EchoLevel = pt.get(misc + "EchoLevel", 3);
    EVS.Out_Report = pt.get(misc + "Report_Path", std::string());
    if (EVS.EchoLevel == 0) Options.echo = false;
    return EVS;
}

// IAC/MainRunModes.cpp:354-500: an example evaluation settings file (the reference's values)
int Create_Example_Evaluation_Config_File(std::string Filename, Settings &Options)
{
    PTree pt;
    auto b = [](bool v) { return std::string(v ? "true" : "false"); };
    pt.put("root.Info.Info_Evaluation",
           "In this file are all Settings for data-evaluation (autocorrelation) stored. This is an example file, "
           "just edit it as you need. Input and output arrays are raw binary files here.");
    pt.put("root.Info.Version", std::to_string(Settings::INTERNAL_VERSION));
    pt.put("root.Info.VersionRevision", std::to_string(Settings::INTERNAL_VERSION_Revision));
    const std::string in = "root.EvalSettings.InputFiles.", out = "root.EvalSettings.OutputFiles.",
                      mesh = "root.EvalSettings.OutputMesh.", ph = "root.EvalSettings.PhotonThreshold.",
                      misc = "root.EvalSettings.Misc.";
    pt.put(in + "EventList_Path", "EventList.xml");
    pt.put(in + "PixelMap_Path", "Pixelmap.bin");
    pt.put(in + "PixelMap_Dataset", "");
    pt.put(in + "PixelMap_Size", "1024; 1024");
    pt.put(in + "PixelMask_Path", "Pixelmask.bin");
    pt.put(in + "PixelMask_Dataset", "");
    pt.put(in + "UseExistingAvInt", b(false));
    pt.put(out + "AC_Final_Path", "AC_Final.bin");
    pt.put(out + "AvIntensity_Path", "AvIntensity.bin");
    pt.put(out + "Cq_Merged_Path", "CQ_Merged.bin");
    pt.put(out + "Cq_small_Path", "CQ_small.bin");
    pt.put(out + "ACuw_Path", "AC_unweighted.bin");
    pt.put(out + "Q_Vector", "QVector.bin");
    pt.put(mesh + "MeshSize", "201");
    pt.put(mesh + "QZoom", "1");
    pt.put(mesh + "AngularAveraged", b(false));
    pt.put(mesh + "DoubleMap", b(true));
    pt.put(mesh + "AC_FirstMap_Flags.InterpolationMode", "0");
    pt.put(mesh + "AC_SecondMap_Flags.InterpolationMode", "0");
    pt.put(ph + "PhotonOffset", "0");
    pt.put(ph + "PhotonStep", "1");
    pt.put(misc + "RestrictStackToBoundaries", b(false));
    pt.put(misc + "LowerBoundary", "0");
    pt.put(misc + "UpperBoundary", "10");
    pt.put(misc + "UsePixelmaskAsFlatField", b(false));
    pt.put(misc + "FractionalCq", b(false));
    pt.put(misc + "SizeOfCqFraction", "10");
    pt.put(misc + "InvertRotMatrix", b(false));
    pt.put(misc + "EchoLevel", "3");
    pt.put(misc + "Report_Path", "Report.txt");
    iac_host::save_property_tree(pt, Filename);
    if (Options.echo) std::cout << "Example Evaluation Setting File have been created and saved as \"" << Filename << "\".\n";
    return 0;
}

// IAC/MainRunModes.cpp:1254-1330: load the settings, run the evaluation, write the report.
int AutoCorrelateData(std::string ConfigFile, Settings &Options)
{
    if (Options.echo) std::cout << "load Settings\n";
    RunIAC::CreateDataEval_Settings EVS = LoadEvaluationSettings(ConfigFile, Options);
    const auto t0 = std::chrono::steady_clock::now();
    if (Options.echo) std::cout << "start evaluation\n";
    const RunIAC::IAC_Report Report = RunIAC::Run_AutoCorr_DataEval(Options, EVS);
    const double TimeNeeded = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!EVS.Out_Report.empty()) {
        std::ofstream file(EVS.Out_Report);
        file << " *** IAC - Report (autocorrelation mode) *** \n";
        file << "Ver: " << Settings::GetVersion() << "\n\n";
        file << "From Settings file: \"" << ConfigFile << "\"\n\n";
        if (EVS.AngularAveraged) {
            file << "Angular averaged (1D - Mode)\n";
            file << "Vector Length: " << Report.FinalMeshSize << "\n";
            file << "dq/dx = " << Report.dQperVox << "  (dq/dx is normalized to be 2 for 180 deg between k1 and k2)\n";
        } else {
            file << "AC-Volume (3D - Mode)\n";
            file << "Final Mesh Size: " << Report.FinalMeshSize << " x " << Report.FinalMeshSize << " x "
                 << Report.FinalMeshSize << "\n";
            file << "dq/dVox = " << Report.dQperVox
                 << "  (dq/dVox is normalized to be 2 for 180 deg between k1 and k2)\n";
            file << "small C(q) Mesh Size: " << Report.SmallCqMeshSize[0] << " x " << Report.SmallCqMeshSize[1]
                 << " x " << Report.SmallCqMeshSize[2] << "\n";
            file << "Mean photon count per pattern: " << Report.MeanPhotonCount << "\n";
            file << "\n Further Information (from Settings file):\n";
            file << "Input paths:\n";
            file << "Event-XML-file path: \"" << EVS.XML_Path << "\"\n";
            file << "Pixel-map path: \"" << EVS.PixelMap_Path << "\"; Dataset: \"" << EVS.PixelMap_DataSet << "\"\n";
            file << "Pixel-mask path: \"" << EVS.PixelMask_Path << "\"; Dataset: \"" << EVS.PixelMask_Dataset
                 << "\"\n";
            file << "Output paths:\n";
            file << "AC (merged) Output path: \"" << EVS.Out_Final_AC_Path << "\"\n";
            if (!EVS.Out_Cq_Path.empty()) file << "CQ (merged) path: \"" << EVS.Out_Cq_Path << "\"\n";
            if (!EVS.Out_ACuw_Path.empty()) file << "AC (uw) Output path: \"" << EVS.Out_ACuw_Path << "\"\n";
            if (!EVS.Out_Cq_small_Path.empty())
                file << "CQ (small) Output path: \"" << EVS.Out_Cq_small_Path << "\"\n";
            if (!EVS.Out_AvIntensity_Path.empty())
                file << "Averaged intensity (flat) path: \"" << EVS.Out_AvIntensity_Path << "\"\n";
            file << "Other Settings:\n";
            file << "Evaluated Stack: Events: " << Report.LowerBound << " - " << Report.UpperBound << "\n";
        }
        file << "\nEvaluation performed within " << TimeNeeded << "s (" << TimeNeeded / 3600 << "h)\n";
    }
    return 0;
}

}  // namespace MainRunModes
