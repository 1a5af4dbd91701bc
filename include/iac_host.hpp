// iac_host.hpp -- the reference's host call surface (Settings, Detector, ACMesh,
// AC1D, RunIAC, Simulator) kept for its callers, with every device step going
// through the C ABI of iac_hip.h (libiac_hip.so).  Link with -liac_host -liac_hip.
//
// Class, member and method names, argument order and meaning follow the
// reference headers (IAC/Settings.h, IAC/Detector.h, IAC/ACMesh.h, IAC/AC1D.h,
// IAC/RunIAC.h, IAC/Simulator.h).  Differences, all at the edges of the path:
//  * no OpenCL: Settings::SetUp_GPU / GPU_ReserveDevice / GPU_FreeDevice hold one
//    iac_ctx per device in place of SetUp_OpenCL / OCL_ReserveDevice / OCL_FreeDevice;
//  * no HDF5 / XML / Boost: frames come from Settings::FrameLoader (default: the
//    event's in-memory Data pointer, else a raw float32 stack file named by
//    Filename, frame index Event); pixel maps, masks and outputs are raw binary
//    arrays (ArrayOperators::SafeArrayToFile / LoadArrayFromFile, Binary);
//  * meshes stay resident on the device (ACMesh::d_Mesh / d_CQMesh) across the
//    frame loop; ACMesh::Download() fills the host arrays;
//  * Detector keeps a Settings pointer (like ACMesh) for the device it runs on;
//  * the CPU branch of AutoCorrelateSparseList (CpuGpu = 0) is not part of this
//    library and returns an error; the GPU result equals that branch bit for bit.
// Errors follow Settings::checkErr (IAC/Settings.cpp:322-329): print and
// exit(EXIT_FAILURE); precondition failures throw std::runtime_error.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "iac_hip.h"

// IAC/Settings.h:64-187
class Settings {
public:
    Settings();
    ~Settings();
    Settings(const Settings &) = delete;
    Settings &operator=(const Settings &) = delete;

    // IAC/Settings.h:40-42
    const static int XML_HITLIST_VERSION = 9;
    const static int INTERNAL_VERSION = 12;
    const static int INTERNAL_VERSION_Revision = 2;
    static std::string GetVersion();

    // IAC/Settings.h:105-119
    struct HitEvent {
        float RotMatrix[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // row major, q = M q_local
        std::string Filename;
        std::string Dataset;
        int Event = 0;
        int SerialNumber = 0;
        float MeanIntensity = -1.0f;
        int PhotonCount = -1;
        float HitsPixelRatio = -1.0f;
        std::string SupplementInfo = "";
        const float *Data = nullptr;  // in-memory frame (DetectorSize[0] * DetectorSize[1] floats)
    };
    std::vector<HitEvent> HitEvents;
    // IAC/Settings.cpp:473-676: the hit-event list file (XML, or JSON / INFO by extension)
    void LoadHitEventListFromFile(std::string Filename);
    void SafeHitEventListToFile(std::string Filename);
    void SafeHitEventListToFile(std::string Filename, std::vector<HitEvent> &HitEventList,
                                bool AdditionalInformations = false,
                                std::unordered_map<std::string, std::string> AdditioInfoMap = {});
    bool echo = true;
    void Echo(const std::string &output);

    // IAC/Settings.h:158
    enum Interpolation { NearestNeighbour = 0, Linear = 1 };

    // IAC/Settings.h:160-187; the device path uses the default Offset 0, Step 0.001
    struct FloatIntConversion {
        float Offset = 0.0f;
        float Step = 0.001f;
    };
    FloatIntConversion F_I_Conversion;
    long FloatToInt(float f) const;
    long FloatToInt(double f) const;
    float IntToFloat(int i) const;
    float IntToFloat(long i) const;
    float IntToFloat(unsigned long i) const;

    // IAC/Settings.cpp:539-560 (3x3 inverse in double, stored back as float)
    void InvertRotationMatrices();

    // device pool in place of IAC/Settings.cpp:331-470
    void SetUp_GPU(int MaxDevices = -1);
    // A pool of contexts the caller made and keeps (e.g. ones bound to its own streams):
    // the Settings reserves and frees them like its own but never destroys them.
    void UseContexts(const std::vector<iac_ctx *> &Contexts);
    int GPU_ReserveDevice();  // -1 if every device is taken
    void GPU_FreeDevice(int DeviceIndex);
    iac_ctx *Context(int DeviceIndex);
    int NumDevices() const { return (int)ctxs_.size(); }
    unsigned int ThreadSleepForGPUDev = 60000;  // microseconds between reservation attempts
    void checkErr(int err, const char *name);  // IAC/Settings.cpp:322-329

    // frame source (GetSliceOutOfHDFCuboid in the reference, IAC/Detector.cpp:800-850)
    std::function<void(const HitEvent &, float *dst, size_t npix)> FrameLoader;
    void LoadFrame(const HitEvent &ev, float *dst, size_t npix) const;

    // frames per device batch of the evaluation frame loop
    unsigned int BatchFrames = 512;

private:
    std::vector<iac_ctx *> ctxs_;
    std::vector<bool> available_;
    bool owns_ctxs_ = true;
    std::mutex pool_mutex_;
};

class ACMesh;

// IAC/Detector.h
class Detector {
public:
    // IAC/Detector.h:46
    struct AutoCorrFlags {
        Settings::Interpolation InterpolationMode = Settings::NearestNeighbour;
    };
    struct DetChecklist {
        bool SparseHitList = false;
        bool Event = false;
        bool PixelMap = false;
        bool KMap = false;
        bool PixelMask = false;
        bool Intensity = false;
    };
    DetChecklist Checklist;

    Detector();
    explicit Detector(Settings &Options);
    Detector(const Detector &RefDet, bool DeepCopy);
    ~Detector();
    Detector &operator=(const Detector &) = delete;

    Settings *Options = nullptr;

    void Calc_kMap();                                                               // IAC/Detector.cpp:310-361
    void CreateEmptyPixelMap(int FsSize, int SsSize);                               // :297-308
    void LoadPixelMap(const float *PixelMap3, unsigned int SsSize, unsigned int FsSize);
    void LoadPixelMap(const std::string &RawPath, unsigned int SsSize, unsigned int FsSize);
    void LoadPixelMap(unsigned int fs, unsigned int ss);                            // size only
    void LoadPixelMask(const int *Mask);
    void LoadPixelMask(const std::string &Path = "");                               // raw int32, "" = flat ones
    void GenerateFlatOnesPixelMask();
    void ApplyPixelMask();
    void LoadIntensityData(Settings::HitEvent *Event);                              // :792-800
    void LoadIntensityData();
    // IAC/Detector.cpp:687-790 (events [LowerBound, UpperBound), on the device)
    void LoadAndAverageIntensity(std::vector<Settings::HitEvent> &Events, float Threshold, float PhotonSamplingStep,
                                 int LowerBound, int UpperBound, bool Pixelmask = true);
    void LoadAndAverageIntensity(std::vector<Settings::HitEvent> &Events, float Threshold, float PhotonSamplingStep);
    // IAC/Detector.cpp:926-977 (on the device; fills SparseHitList)
    void CreateSparseHitList(float Threshold);
    void CreateSparseHitList(float Threshold, float PhotonSamplingStep);
    void CreateSparseHitList(float Threshold, float PhotonSamplingStep, bool Par);
    float CalculateMeanIntensity(bool FromSparse);
    // IAC/Detector.cpp:1010-1181: returns 1 (GPU path)
    int AutoCorrelateSparseList(ACMesh &BigMesh, AutoCorrFlags FlagsFirstMap, AutoCorrFlags FlagsSecondMap,
                                bool DoubleMapping, Settings &Options, int CpuGpu = -1);
    // IAC/Detector.cpp:1184-1315
    void AutoCorrelate_CofQ_SmallMesh(ACMesh &SmallMesh, AutoCorrFlags Flags, Settings &Options);
    // IAC/Detector.cpp:1317-1573
    void Merge_smallCofQ(ACMesh &BigMesh, ACMesh &SmallMesh, std::vector<Settings::HitEvent> &Events,
                         unsigned int LowerBound, unsigned int UpperBound, Settings &Options, AutoCorrFlags Flags);
    void Merge_smallCofQ(ACMesh &BigMesh, ACMesh &SmallMesh, std::vector<Settings::HitEvent> &Events,
                         Settings &Options, AutoCorrFlags Flags);

    unsigned int DetectorSize[2] = {0, 0};  // [slow scan][fast scan]
    float *Intensity = nullptr;              // [n_y][n_x]
    int *PixelMask = nullptr;
    float *PixelMap = nullptr;               // [n_y][n_x][3]
    float *kMap = nullptr;                   // [n_y][n_x][3]
    float Max_k[3] = {0, 0, 0};
    float Min_k[3] = {0, 0, 0};
    float Max_q[3] = {0, 0, 0};
    std::vector<std::array<float, 4>> SparseHitList;
    Settings::HitEvent *DetectorEvent = nullptr;

    size_t NumPixels() const { return (size_t)DetectorSize[0] * DetectorSize[1]; }
    // device copies made on demand (k-map, mask) for the device passes
    const float *DeviceKMap(iac_ctx *ctx);
    const int *DeviceMask(iac_ctx *ctx);

private:
    Settings &opts(const char *where);
    void free_device();
    iac_ctx *dev_ctx_ = nullptr;
    float *d_kmap_ = nullptr;
    int *d_mask_ = nullptr;
    bool d_kmap_valid_ = false, d_mask_valid_ = false;
};

// IAC/ACMesh.h
class ACMesh {
public:
    explicit ACMesh(Settings *PrgSettings);
    ~ACMesh();
    ACMesh(const ACMesh &) = delete;
    ACMesh &operator=(const ACMesh &) = delete;

    double *CQMesh = nullptr;       // host copy (Download)
    unsigned long *Mesh = nullptr;  // host copy (Download); 64-bit on LP64
    // IAC/ACMesh.h:21-29; the same layout as iac_mesh_shape
    struct MeshShape {
        unsigned int k_A = 0, k_B = 1, k_C = 2;
        unsigned int Size_AB = 0, Size_C = 0;
        unsigned int Center[3] = {0, 0, 0};
        float dq_per_Voxel = 0.0f;
        unsigned int Voxel_per_dq = 1;
        float Max_Q = 0.0f;
    };
    Settings *Options = nullptr;
    MeshShape Shape;
    struct ACMesh_Checklist {
        bool SmallMesh = false;
        bool BigMesh = false;
        bool CofQBigMesh = false;
        bool CofQSmallMesh = false;
    };
    ACMesh_Checklist Checklist;

    void CreateBigMeshForDetector(Detector &Det, int EdgeSize, float q_Zoom = 1.0f);        // :74-106
    void CreateBigMesh_CofQ_ForDetector(Detector &Det, int EdgeSize, float q_Zoom = 1.0f);  // :107-139
    void CreateSmallMesh_CofQ_ForDetector(Detector &Det, int PerpSize, float q_Zoom = 1.0f);  // :140-206

    // device residency
    size_t NumVoxels() const { return (size_t)Shape.Size_AB * Shape.Size_AB * Shape.Size_C; }
    uint64_t *d_Mesh = nullptr;  // uint64 accumulation mesh (zeroed at creation)
    double *d_CQMesh = nullptr;  // double C(q) mesh
    iac_ctx *ctx = nullptr;      // the context the device meshes live on
    uint64_t *DeviceMesh();      // allocates and zeroes d_Mesh on first use
    double *DeviceCQMesh();
    void Download();             // Flush(), then d_Mesh -> Mesh, d_CQMesh -> CQMesh
    iac_mesh_shape CShape() const;
    // The frame queue of d_Mesh that Detector::AutoCorrelateSparseList appends events to
    // (made on first use for the detector's device k-map and the binning flags; other
    // flags or another k-map flush it and make a new one).  Flush() enqueues the binning
    // of the queued events (no synchronisation); ~ACMesh and Download() flush as well.
    iac_ac3d_queue *Queue(const float *d_kmap, uint32_t npix, int interp1, int interp2, bool double_map);
    void Flush();

private:
    void reset_device();
    void take_shape(const iac_mesh_shape &s);
    void drop_queue();
    iac_ac3d_queue *queue_ = nullptr;
    const float *queue_kmap_ = nullptr;
    int queue_flags_ = -1;
};

// IAC/AC1D.h
class AC1D {
public:
    AC1D();
    ~AC1D();
    AC1D(const AC1D &) = delete;
    AC1D &operator=(const AC1D &) = delete;
    double *CQ = nullptr;
    double *AC_UW = nullptr;
    double *AC = nullptr;
    double *Q = nullptr;
    int EchoLevel = 1;
    struct MeshShape {
        unsigned int Size = 1000;
        float dq_per_Step = 0.0f;
        float Max_Q = 0.0f;
    };
    MeshShape Shape;
    void Initialize(Detector &Det, unsigned int ArraySize);
    void Initialize(Detector &Det, unsigned int ArraySize, float QZoom);  // IAC/AC1D.cpp:29-59
    void CreateQVector();                                                // :472-480
    void CalcAC();                                                       // :482-493
    void Calculate_CQ(Detector &Det, Settings &Options, Settings::Interpolation IterpolMode);  // :61-195
    // :393-469; the thread count of the reference's map-reduce has no meaning here
    void Calculate_AC_UW_MR(Settings &Options, Detector &RefDet, Settings::Interpolation IterpolMode,
                            std::array<float, 2> Photonisation, unsigned int LowerBoundary,
                            unsigned int UpperBoundary, unsigned int Threads = 200);
};

// IAC/RunIAC.h
namespace RunIAC {
struct IAC_Report {
    float dQperVox = 0.0f;
    unsigned int SmallCqMeshSize[3] = {0, 0, 0};
    unsigned int FinalMeshSize = 101;
    float MeanPhotonCount = 0.0f;
    unsigned int LowerBound = 0;
    unsigned int UpperBound = 0;
};

class CreateDataEval_Settings {
public:
    int MeshSize = 501;
    float QZoom = 1.0f;
    bool AngularAveraged = false;
    bool RestrictStackToBoundaries = false;
    unsigned int LowerBoundary = 0;
    unsigned int UpperBoundary = 1;
    bool FractionalCq = false;
    unsigned int SizeOfCqFraction = 100;
    bool InvertRotMatrix = false;
    bool UsePixelMask_as_Flatfield = false;
    std::string XML_Path = "";
    std::string PixelMap_Path = "";  // raw float32 [ss][fs][3]
    std::string PixelMap_DataSet = "";
    std::string PixelMask_Path = "";  // raw int32
    std::string PixelMask_Dataset = "";
    std::string Out_AvIntensity_Path = "";
    std::string Out_ACuw_Path = "";
    std::string Out_Cq_Path = "";
    std::string Out_Cq_small_Path = "";
    std::string Out_Final_AC_Path = "";
    std::string Out_Q_Vector = "";
    std::string Out_Report = "";
    float PhotonOffset = 0.0f;
    float PhotonStep = 1.0f;
    int EchoLevel = 1;
    bool DoubleMap = true;
    Detector::AutoCorrFlags AC_SecondMap_Flags;
    Detector::AutoCorrFlags AC_FirstMap_Flags;
    bool UseExistingAvInt = false;
    bool DetDisturb = false;
    double DetDisturb_Shift = 0.0;
    double DetDisturb_Rot = 0.0;
    bool IgnoreOrientation = false;
    // size of the pixel map when it is read from a raw file
    unsigned int DetectorSize[2] = {0, 0};
};

// IAC/RunIAC.cpp:120-158
void Merge_ACandCQ(double *&Output, ACMesh &AC, ACMesh &CQ, Settings &PrgSettings);
// IAC/RunIAC.cpp:185-632; the event list is read from XML_Path when it is set, else
// Settings::HitEvents is used as it stands
IAC_Report Run_AutoCorr_DataEval(Settings &PrgSettings, CreateDataEval_Settings EvalSettings);
// An evaluation whose detector the caller prepared (pixel map, k-map, mask loaded).
IAC_Report Run_AutoCorr_DataEval(Settings &PrgSettings, CreateDataEval_Settings EvalSettings, Detector &Det);
}  // namespace RunIAC

// IAC/MainRunModes.h: the evaluation settings file and the evaluate run mode
namespace MainRunModes {
// IAC/MainRunModes.cpp:32-91 (plus root.EvalSettings.InputFiles.PixelMap_Size "SS; FS" for raw
// pixel maps; without it the size is taken from the file)
RunIAC::CreateDataEval_Settings LoadEvaluationSettings(std::string Filename, Settings &Options);
// IAC/MainRunModes.cpp:354-500
int Create_Example_Evaluation_Config_File(std::string Filename, Settings &Options);
// IAC/MainRunModes.cpp:1254-1330: settings -> Run_AutoCorr_DataEval -> report file
int AutoCorrelateData(std::string ConfigFile, Settings &Options);
}  // namespace MainRunModes

// IAC/PPP.h:11-101: the per-frame operations of pattern pre-processing (the HDF5 / XML handling
// of ProcessData_PF_LAP stays with the caller; iac_ppp_lap runs a whole stack with mask, gain
// maps and photon counts)
namespace PPP {
// IAC/PPP.h:14-18
struct DetectorPanel {
    unsigned int FirstInd = 0;
    unsigned int Scans[2] = {2, 2};  // fs, ss
};
// IAC/PPP.h:20-37 (the finder's part)
struct Create_LAPSettings {
    unsigned int DetPanels_Num = 1;
    std::vector<DetectorPanel> DetectorPanels;
    float ADU_perPhoton = 1.0f;
    float SeedThershold = 0.5f;
    float CombinedThershold = 0.9f;
};
// IAC/PPP.cpp:18-95, in place on one host frame of FullDetSize pixels (the panels' total), on a
// device of a process-wide pool
void PhotonFinder_LargestAdjacentPixel(float *Intensity, std::vector<DetectorPanel> DetectorPanels, int FullDetSize,
                                       float ADU_perPhoton = 1.0f, float SeedThershold = 0.5f,
                                       float CombinedThershold = 0.9f);
// the same on a device of Options
void PhotonFinder_LargestAdjacentPixel(float *Intensity, const std::vector<DetectorPanel> &DetectorPanels,
                                       int FullDetSize, float ADU_perPhoton, float SeedThershold,
                                       float CombinedThershold, Settings &Options);
}  // namespace PPP

// IAC/Simulator.h:26-56 (+ the device simulator); emitter positions are supplied by the
// caller (the reference's Crystal generator is not part of this library)
class Simulator {
public:
    struct SimulationSettings {
        unsigned int NumberOfSimulations = 1;
        unsigned int Modes = 1;
        unsigned int SubSampling = 0;
        float AveragePhotonesPerEmitterOnDetector = 1.0f;
        bool PoissonSample = true;
        float Value_per_Photon = 1.0f;
        float ADUNoise = 0.0f;
        bool AutoPixelOrientation = true;
        bool AutoPixelSize = true;
        double PixelOrientationVectors[6] = {1, 0, 0, 0, 1, 0};
        double PixelSize[2] = {0, 0};
        double Wavelength = 1.94;
        float FlYield = 1.0f;          // Crystal::EmittingCrystSettings::FlYield
        int Precision = IAC_SIM_BF16X3;  // IAC_SIM_* of the field computation
        uint64_t Seed = 0;              // phases and Poisson draws
    };
    struct SimulationOutput {
        std::vector<std::vector<float>> Intensities;
        std::vector<Settings::HitEvent> HitEvents;
        unsigned int DetectorSize[2] = {0, 0};
    };
    // Simulator::SimulatePart, IAC/Simulator.cpp:662-944, for the emitter positions
    // Emitters ([nem][3], Angstrom) with phases drawn uniformly in [0, 2 pi) per frame
    // and mode (or taken from Phases, [frames * modes][nem], when non-null).
    void Simulate(Detector &Det, const std::vector<std::array<double, 3>> &Emitters, SimulationSettings &SimSettings,
                  SimulationOutput &Output, Settings &Options, const double *Phases = nullptr);
};
