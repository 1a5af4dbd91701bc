// detector.cpp -- Detector (IAC/Detector.h, IAC/Detector.cpp): geometry, masks,
// frames, and the device passes of the evaluation (averaged intensity,
// sparsification, AC_uw, C(q) small, rotation merge).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "host_util.hpp"
#include "iac_host.hpp"

namespace iac_host {

void check(int rc, const char *where)
{
    if (rc != IAC_OK) {
        std::cerr << "ERROR: " << where << " (" << rc << ": " << iac_strerror(rc) << ") " << iac_last_error()
                  << std::endl;
        exit(EXIT_FAILURE);
    }
}

void DevBuf::reset(iac_ctx *ctx, size_t bytes)
{
    release();
    ctx_ = ctx;
    bytes_ = bytes;
    if (bytes) check(iac_dev_alloc(ctx, bytes, &ptr_), "device allocation");
}

void DevBuf::release()
{
    if (ptr_) iac_dev_free(ctx_, ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

void DevBuf::zero()
{
    if (ptr_) check(iac_memset_zero(ctx_, ptr_, bytes_), "device memset");
}

void DevBuf::upload(const void *src, size_t bytes)
{
    if (bytes) check(iac_memcpy_h2d(ctx_, ptr_, src, bytes), "host to device copy");
}

void DevBuf::download(void *dst, size_t bytes) const
{
    if (bytes) check(iac_memcpy_d2h(ctx_, dst, ptr_, bytes), "device to host copy");
}

}  // namespace iac_host

using iac_host::check;

Detector::Detector() {}

Detector::Detector(Settings &Options_) : Options(&Options_) {}

// IAC/Detector.cpp:28-85 (deep copy of the arrays)
Detector::Detector(const Detector &RefDet, bool DeepCopy)
{
    Checklist = RefDet.Checklist;
    Options = RefDet.Options;
    DetectorSize[0] = RefDet.DetectorSize[0];
    DetectorSize[1] = RefDet.DetectorSize[1];
    std::memcpy(Max_k, RefDet.Max_k, sizeof Max_k);
    std::memcpy(Min_k, RefDet.Min_k, sizeof Min_k);
    std::memcpy(Max_q, RefDet.Max_q, sizeof Max_q);
    DetectorEvent = RefDet.DetectorEvent;
    const size_t n = NumPixels();
    auto dup = [&](auto *src, size_t count) -> decltype(src) {
        if (!src) return nullptr;
        using T = std::remove_pointer_t<decltype(src)>;
        T *d = new T[count];
        std::copy(src, src + count, d);
        return d;
    };
    if (DeepCopy) {
        Intensity = dup(RefDet.Intensity, n);
        PixelMask = dup(RefDet.PixelMask, n);
        PixelMap = dup(RefDet.PixelMap, 3 * n);
        kMap = dup(RefDet.kMap, 3 * n);
        SparseHitList = RefDet.SparseHitList;
    } else {
        PixelMap = dup(RefDet.PixelMap, 3 * n);
        kMap = dup(RefDet.kMap, 3 * n);
        PixelMask = dup(RefDet.PixelMask, n);
        Checklist.Intensity = false;
        Checklist.SparseHitList = false;
    }
}

Detector::~Detector()
{
    free_device();
    delete[] Intensity;
    delete[] PixelMask;
    delete[] PixelMap;
    delete[] kMap;
}

void Detector::free_device()
{
    if (d_kmap_) iac_dev_free(dev_ctx_, d_kmap_);
    if (d_mask_) iac_dev_free(dev_ctx_, d_mask_);
    d_kmap_ = nullptr;
    d_mask_ = nullptr;
    d_kmap_valid_ = d_mask_valid_ = false;
    dev_ctx_ = nullptr;
}

Settings &Detector::opts(const char *where)
{
    if (!Options) throw std::runtime_error(std::string("Detector::") + where + ": no Settings bound (Detector(Settings&))");
    return *Options;
}

const float *Detector::DeviceKMap(iac_ctx *ctx)
{
    if (!Checklist.KMap || !kMap) throw std::runtime_error("Detector: k-map not calculated (Calc_kMap)");
    if (dev_ctx_ != ctx) free_device();
    dev_ctx_ = ctx;
    if (!d_kmap_) check(iac_dev_alloc(ctx, NumPixels() * 12, (void **)&d_kmap_), "k-map allocation");
    if (!d_kmap_valid_) {
        check(iac_memcpy_h2d(ctx, d_kmap_, kMap, NumPixels() * 12), "k-map upload");
        d_kmap_valid_ = true;
    }
    return d_kmap_;
}

const int *Detector::DeviceMask(iac_ctx *ctx)
{
    if (!PixelMask) return nullptr;
    if (dev_ctx_ != ctx) free_device();
    dev_ctx_ = ctx;
    if (!d_mask_) check(iac_dev_alloc(ctx, NumPixels() * 4, (void **)&d_mask_), "mask allocation");
    if (!d_mask_valid_) {
        check(iac_memcpy_h2d(ctx, d_mask_, PixelMask, NumPixels() * 4), "mask upload");
        d_mask_valid_ = true;
    }
    return d_mask_;
}

// IAC/Detector.cpp:310-361
void Detector::Calc_kMap()
{
    if (!Checklist.PixelMap) throw std::runtime_error("Detector::Calc_kMap: no pixel map loaded");
    delete[] kMap;
    kMap = new float[3 * NumPixels()];
    check(iac_calc_kmap(PixelMap, (uint32_t)NumPixels(), kMap, Max_k, Min_k, Max_q), "iac_calc_kmap");
    Checklist.KMap = true;
    d_kmap_valid_ = false;
}

// IAC/Detector.cpp:363-375
void Detector::CreateEmptyPixelMap(int FsSize, int SsSize)
{
    delete[] PixelMap;
    PixelMap = new float[(size_t)FsSize * SsSize * 3]();
    DetectorSize[0] = SsSize;
    DetectorSize[1] = FsSize;
    Checklist.PixelMap = true;
}

void Detector::LoadPixelMap(const float *PixelMap3, unsigned int SsSize, unsigned int FsSize)
{
    DetectorSize[0] = SsSize;
    DetectorSize[1] = FsSize;
    delete[] PixelMap;
    PixelMap = new float[3 * NumPixels()];
    std::copy(PixelMap3, PixelMap3 + 3 * NumPixels(), PixelMap);
    Checklist.PixelMap = true;
    Checklist.KMap = false;
}

void Detector::LoadPixelMap(const std::string &RawPath, unsigned int SsSize, unsigned int FsSize)
{
    std::vector<float> pm(3 * (size_t)SsSize * FsSize);
    iac_host::read_raw(RawPath, pm.data(), pm.size() * sizeof(float));
    LoadPixelMap(pm.data(), SsSize, FsSize);
}

// IAC/Detector.cpp:514-535
void Detector::LoadPixelMap(unsigned int fs, unsigned int ss)
{
    DetectorSize[0] = ss;
    DetectorSize[1] = fs;
    delete[] PixelMap;
    PixelMap = new float[3 * (size_t)fs * ss]();
    for (unsigned int i_x = 0; i_x < DetectorSize[0]; i_x++) {
        for (unsigned int i_y = 0; i_y < DetectorSize[1]; i_y++) {
            PixelMap[0 + 3 * i_y + 3 * DetectorSize[1] * i_x] = (float)i_x - 0.5f * (DetectorSize[0]) + 0.5f;
            PixelMap[1 + 3 * i_y + 3 * DetectorSize[1] * i_x] = (float)i_y - 0.5f * (DetectorSize[1]) + 0.5f;
            PixelMap[2 + 3 * i_y + 3 * DetectorSize[1] * i_x] = 0;
        }
    }
    Checklist.PixelMap = true;
    Checklist.KMap = false;
}

void Detector::LoadPixelMask(const int *Mask)
{
    delete[] PixelMask;
    PixelMask = new int[NumPixels()];
    std::copy(Mask, Mask + NumPixels(), PixelMask);
    Checklist.PixelMask = true;
    d_mask_valid_ = false;
}

// IAC/Detector.cpp:566-650 (binary int mask; "" = flat ones)
void Detector::LoadPixelMask(const std::string &Path)
{
    if (Path.empty()) {
        GenerateFlatOnesPixelMask();
        return;
    }
    std::vector<int> m(NumPixels());
    iac_host::read_raw(Path, m.data(), m.size() * sizeof(int));
    LoadPixelMask(m.data());
}

// IAC/Detector.cpp:651-660
void Detector::GenerateFlatOnesPixelMask()
{
    delete[] PixelMask;
    PixelMask = new int[NumPixels()];
    std::fill(PixelMask, PixelMask + NumPixels(), 1);
    Checklist.PixelMask = true;
    d_mask_valid_ = false;
}

// IAC/Detector.cpp:662-665 (ParMultiplyElementwise(float*, int*))
void Detector::ApplyPixelMask()
{
    if (!Intensity || !PixelMask) return;
    for (size_t i = 0; i < NumPixels(); i++) Intensity[i] *= PixelMask[i];
}

void Detector::LoadIntensityData(Settings::HitEvent *Event)
{
    DetectorEvent = Event;
    Checklist.Event = true;
    LoadIntensityData();
}

// IAC/Detector.cpp:799-830
void Detector::LoadIntensityData()
{
    if (!Checklist.Event || !DetectorEvent) throw std::runtime_error("Detector::LoadIntensityData: no Event set");
    delete[] Intensity;
    Intensity = new float[NumPixels()];
    opts("LoadIntensityData").LoadFrame(*DetectorEvent, Intensity, NumPixels());
    Checklist.Intensity = true;
}

void Detector::LoadAndAverageIntensity(std::vector<Settings::HitEvent> &Events, float Threshold,
                                       float PhotonSamplingStep)
{
    LoadAndAverageIntensity(Events, Threshold, PhotonSamplingStep, 0, (int)Events.size(), true);
}

// IAC/Detector.cpp:687-790: frames streamed to the device in stacks; the per-pixel
// sums continue across stacks in event order, so the result is the serial one.
void Detector::LoadAndAverageIntensity(std::vector<Settings::HitEvent> &Events, float Threshold,
                                       float PhotonSamplingStep, int LowerBound, int UpperBound, bool)
{
    if (Events.empty()) {
        std::cerr << "WARNING: Event list is empty, no intensity integration/averaging can be performed.\n";
        return;
    }
    if (LowerBound < 0 || UpperBound > (int)Events.size() || UpperBound <= LowerBound)
        throw std::runtime_error("Detector::LoadAndAverageIntensity: bad event bounds");
    if (!Checklist.PixelMask) GenerateFlatOnesPixelMask();
    Settings &O = opts("LoadAndAverageIntensity");
    iac_host::Reserved res(O, "Detector::LoadAndAverageIntensity");
    iac_ctx *ctx = res.ctx;
    const size_t P = NumPixels();
    const uint32_t N = (uint32_t)(UpperBound - LowerBound);
    const uint32_t B = std::max<uint32_t>(1, std::min<uint32_t>(O.BatchFrames, (uint32_t)((1ull << 30) / (P * 4))));
    iac_host::DevBuf acc(ctx, P * 8), frames(ctx, (size_t)B * P * 4), mi(ctx, (size_t)B * 4), pc(ctx, (size_t)B * 4);
    acc.zero();
    std::vector<float> mean(B);
    std::vector<int> count(B);
    const int *d_mask = DeviceMask(ctx);
    float *sum = PhotonSamplingStep > 0 ? nullptr : acc.as<float>();
    int64_t *tally = PhotonSamplingStep > 0 ? acc.as<int64_t>() : nullptr;
    // frames read one stack ahead on a reader thread into page-locked stacks
    iac_host::FrameStream stream(O, Events, (size_t)LowerBound, N, P, B);
    const auto t_loop = std::chrono::steady_clock::now();
    double up_s = 0;
    uint32_t b0 = 0, nb = 0;
    while (const float *stack = stream.next(nb)) {
        const auto t_up = std::chrono::steady_clock::now();
        frames.upload(stack, (size_t)nb * P * 4);
        up_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_up).count();
        stream.release();
        check(iac_intensity_accumulate(ctx, frames.as<float>(), nb, (uint32_t)P, d_mask, Threshold,
                                       PhotonSamplingStep, sum, tally, mi.as<float>(), pc.as<int>()),
              "iac_intensity_accumulate in Detector::LoadAndAverageIntensity()");
        mi.download(mean.data(), (size_t)nb * 4);
        if (PhotonSamplingStep > 0) pc.download(count.data(), (size_t)nb * 4);
        for (uint32_t f = 0; f < nb; f++) {
            Settings::HitEvent &ev = Events[LowerBound + b0 + f];
            ev.MeanIntensity = mean[f];
            if (PhotonSamplingStep > 0) ev.PhotonCount = count[f];
        }
        b0 += nb;
    }
    iac_host::IngestStats &st = iac_host::ingest_stats();
    st.frames += N;
    st.read_seconds += stream.read_seconds();
    st.wait_seconds += stream.wait_seconds();
    st.upload_seconds += up_s;
    st.loop_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
    iac_host::DevBuf avg(ctx, P * 4);
    check(iac_intensity_finish(ctx, (uint32_t)P, N, PhotonSamplingStep, sum, tally, avg.as<float>()),
          "iac_intensity_finish in Detector::LoadAndAverageIntensity()");
    delete[] Intensity;
    Intensity = new float[P];
    avg.download(Intensity, P * 4);
    Checklist.Intensity = true;
}

void Detector::CreateSparseHitList(float Threshold) { CreateSparseHitList(Threshold, 0.0f, true); }

void Detector::CreateSparseHitList(float Threshold, float PhotonSamplingStep)
{
    CreateSparseHitList(Threshold, PhotonSamplingStep, true);
}

// IAC/Detector.cpp:926-977 on the device (one frame: the loaded Intensity); a step
// <= 0 keeps the intensity (the one-argument form).
void Detector::CreateSparseHitList(float Threshold, float PhotonSamplingStep, bool)
{
    if (!Intensity) throw std::runtime_error("Detector::CreateSparseHitList: no intensity loaded");
    Settings &O = opts("CreateSparseHitList");
    iac_host::Reserved res(O, "Detector::CreateSparseHitList");
    iac_ctx *ctx = res.ctx;
    const size_t P = NumPixels();
    iac_host::DevBuf fr(ctx, P * 4), rec(ctx, P * 16), off(ctx, 16);
    fr.upload(Intensity, P * 4);
    uint64_t n = 0;
    check(iac_sparse_hit_list(ctx, DeviceKMap(ctx), (uint32_t)P, fr.as<float>(), 1, nullptr, Threshold,
                              PhotonSamplingStep, rec.as<float>(), off.as<uint64_t>(), &n),
          "iac_sparse_hit_list in Detector::CreateSparseHitList()");
    SparseHitList.resize(n);
    rec.download(SparseHitList.data(), n * 16);
    Checklist.SparseHitList = true;
}

// IAC/Detector.cpp:980-1000
float Detector::CalculateMeanIntensity(bool FromSparse)
{
    double IntInt = 0;
    if (FromSparse) {
        for (const auto &e : SparseHitList) IntInt += e[3];
    } else {
        for (size_t i = 0; i < NumPixels(); i++) IntInt += Intensity[i];
    }
    IntInt = IntInt / (DetectorSize[0] * DetectorSize[1]);
    if (DetectorEvent) DetectorEvent->MeanIntensity = (float)IntInt;
    return (float)IntInt;
}

// IAC/Detector.cpp:1010-1181: the current SparseHitList, with the event's rotation,
// appended to the mesh's frame queue (pinned staging, no allocation, no device
// synchronisation); the LDS engine bins the queued events in one pass when the queue
// fills and on ACMesh::Flush / Download.
int Detector::AutoCorrelateSparseList(ACMesh &BigMesh, AutoCorrFlags FlagsFirstMap, AutoCorrFlags FlagsSecondMap,
                                      bool DoubleMapping, Settings &Options_, int CpuGpu)
{
    if (!Checklist.SparseHitList) throw std::runtime_error("Detector::AutoCorrelateSparseList: no sparse hit list");
    if (!BigMesh.Checklist.BigMesh) throw std::runtime_error("Detector::AutoCorrelateSparseList: no big mesh");
    if (CpuGpu == 0) {
        std::cerr << "ERROR: the CPU branch of AutoCorrelateSparseList is not part of this library; "
                     "the device path gives the same mesh (CpuGpu = -1 or 1)\n";
        exit(EXIT_FAILURE);
    }
    iac_host::Reserved res(Options_, "Detector::AutoCorrelateSparseList");
    iac_ctx *ctx = res.ctx;
    if (BigMesh.ctx && BigMesh.ctx != ctx) throw std::runtime_error("AutoCorrelateSparseList: mesh lives on another device");
    BigMesh.ctx = ctx;
    const float identity[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const float *rot = DetectorEvent ? DetectorEvent->RotMatrix : identity;
    const size_t n = SparseHitList.size();
    if (n >= (1ull << 32)) throw std::runtime_error("Detector::AutoCorrelateSparseList: hit list too long");
    iac_ac3d_queue *q = BigMesh.Queue(DeviceKMap(ctx), (uint32_t)NumPixels(), FlagsFirstMap.InterpolationMode,
                                      FlagsSecondMap.InterpolationMode, DoubleMapping);
    check(iac_ac3d_queue_push(q, n ? SparseHitList.data()->data() : nullptr, (uint32_t)n, rot),
          "iac_ac3d_queue_push in Detector::AutoCorrelateSparseList()");
    return 1;
}

// IAC/Detector.cpp:1184-1315
void Detector::AutoCorrelate_CofQ_SmallMesh(ACMesh &SmallMesh, AutoCorrFlags Flags, Settings &Options_)
{
    if (!Checklist.Intensity || !Intensity) throw std::runtime_error("AutoCorrelate_CofQ_SmallMesh: no intensity");
    if (!SmallMesh.Checklist.CofQSmallMesh) throw std::runtime_error("AutoCorrelate_CofQ_SmallMesh: no small mesh");
    iac_host::Reserved res(Options_, "AutoCorrelate_CofQ_SmallMesh");
    iac_ctx *ctx = res.ctx;
    SmallMesh.ctx = ctx;
    const size_t P = NumPixels();
    double mult = 0;
    check(iac_cq_small_multiplicator(Intensity, (uint32_t)P, &mult), "iac_cq_small_multiplicator");
    iac_host::DevBuf I(ctx, P * 4);
    I.upload(Intensity, P * 4);
    iac_mesh_shape s = SmallMesh.CShape();
    uint64_t *m = SmallMesh.DeviceMesh();
    check(iac_cq_small(ctx, DeviceKMap(ctx), (uint32_t)P, I.as<float>(), &s, Flags.InterpolationMode, mult, m),
          "iac_cq_small in Detector::AutoCorrelate_CofQ_SmallMesh()");
    check(iac_u64_to_f64(ctx, m, SmallMesh.NumVoxels(), mult, SmallMesh.DeviceCQMesh()), "iac_u64_to_f64");
    check(iac_ctx_sync(ctx), "iac_ctx_sync");
}

void Detector::Merge_smallCofQ(ACMesh &BigMesh, ACMesh &SmallMesh, std::vector<Settings::HitEvent> &Events,
                               Settings &Options_, AutoCorrFlags Flags)
{
    Merge_smallCofQ(BigMesh, SmallMesh, Events, 0, (unsigned int)Events.size(), Options_, Flags);
}

// IAC/Detector.cpp:1317-1573: weights and Multiplicator on the host (the small
// mesh is downloaded for them), the merge on the device into BigMesh's C(q).
void Detector::Merge_smallCofQ(ACMesh &BigMesh, ACMesh &SmallMesh, std::vector<Settings::HitEvent> &Events,
                               unsigned int LowerBound, unsigned int UpperBound, Settings &Options_,
                               AutoCorrFlags Flags)
{
    if (!SmallMesh.Checklist.CofQSmallMesh || !SmallMesh.d_CQMesh)
        throw std::runtime_error("Detector::Merge_smallCofQ: small C(q) not calculated");
    if (!BigMesh.Checklist.CofQBigMesh) throw std::runtime_error("Detector::Merge_smallCofQ: no big C(q) mesh");
    if (SmallMesh.Shape.Size_AB != BigMesh.Shape.Size_AB)  // IAC/Detector.cpp:1324-1329
        throw std::runtime_error("Detector::Merge_smallCofQ: perpendicular mesh size does not fit");
    if (UpperBound > Events.size() || UpperBound <= LowerBound)
        throw std::runtime_error("Detector::Merge_smallCofQ: bad event bounds");
    iac_host::Reserved res(Options_, "Detector::Merge_smallCofQ");
    iac_ctx *ctx = res.ctx;
    if (SmallMesh.ctx != ctx) throw std::runtime_error("Detector::Merge_smallCofQ: small mesh on another device");
    BigMesh.ctx = ctx;
    const uint32_t nev = UpperBound - LowerBound;
    std::vector<float> rot((size_t)nev * 9), mi(nev), rw((size_t)nev * 10);
    for (uint32_t i = 0; i < nev; i++) {
        std::copy(Events[LowerBound + i].RotMatrix, Events[LowerBound + i].RotMatrix + 9, rot.begin() + 9 * i);
        mi[i] = Events[LowerBound + i].MeanIntensity;
    }
    const size_t nsmall = SmallMesh.NumVoxels();
    std::vector<double> small(nsmall);
    check(iac_memcpy_d2h(ctx, small.data(), SmallMesh.d_CQMesh, nsmall * 8), "small C(q) download");
    double mult = 0;
    check(iac_merge_prepare(rot.data(), mi.data(), nev, small.data(), nsmall, rw.data(), &mult), "iac_merge_prepare");
    iac_host::DevBuf drw(ctx, rw.size() * 4);
    drw.upload(rw.data(), rw.size() * 4);
    iac_mesh_shape ss = SmallMesh.CShape();
    uint64_t *big = BigMesh.DeviceMesh();
    check(iac_merge_cq(ctx, SmallMesh.d_CQMesh, &ss, drw.as<float>(), nev, Flags.InterpolationMode, mult, big),
          "iac_merge_cq in Detector::Merge_smallCofQ()");
    check(iac_u64_to_f64(ctx, big, BigMesh.NumVoxels(), mult, BigMesh.DeviceCQMesh()), "iac_u64_to_f64");
    check(iac_ctx_sync(ctx), "iac_ctx_sync");
}
