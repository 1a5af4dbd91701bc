// ac1d.cpp -- AC1D (IAC/AC1D.h, IAC/AC1D.cpp): radial C(q), unweighted AC and
// their ratio, with the histograms built on the device.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "host_util.hpp"
#include "iac_host.hpp"

using iac_host::check;

AC1D::AC1D() {}

AC1D::~AC1D()
{
    delete[] CQ;
    delete[] AC_UW;
    delete[] AC;
    delete[] Q;
}

static void fresh(double *&a, unsigned int n)
{
    delete[] a;
    a = new double[n]();
}

void AC1D::Initialize(Detector &Det, unsigned int ArraySize) { Initialize(Det, ArraySize, 1.0f); }

// IAC/AC1D.cpp:29-59
void AC1D::Initialize(Detector &Det, unsigned int ArraySize, float QZoom)
{
    Shape.Size = ArraySize;
    check(iac_ac1d_shape(Det.Max_q, ArraySize, QZoom, &Shape.Max_Q, &Shape.dq_per_Step), "AC1D::Initialize()");
    fresh(CQ, Shape.Size);
    fresh(AC_UW, Shape.Size);
    fresh(AC, Shape.Size);
    fresh(Q, Shape.Size);
    for (unsigned int i = 0; i < Shape.Size; i++) Q[i] = (float)i * Shape.dq_per_Step;
}

// IAC/AC1D.cpp:472-480
void AC1D::CreateQVector()
{
    fresh(Q, Shape.Size);
    for (unsigned int i = 0; i < Shape.Size; i++) Q[i] = i * Shape.dq_per_Step;
}

// IAC/AC1D.cpp:482-493
void AC1D::CalcAC()
{
    fresh(AC, Shape.Size);
    for (unsigned int i = 0; i < Shape.Size; i++) {
        AC[i] = AC_UW[i] / CQ[i];
        if (std::isnan(AC[i])) AC[i] = 0;
    }
}

// IAC/AC1D.cpp:61-195: C(q) = histogram / Multiplicator
void AC1D::Calculate_CQ(Detector &Det, Settings &Options, Settings::Interpolation IterpolMode)
{
    if (!Det.Intensity) throw std::runtime_error("AC1D::Calculate_CQ: averaged intensity not loaded");
    iac_host::Reserved res(Options, "AC1D::Calculate_CQ");
    iac_ctx *ctx = res.ctx;
    const size_t P = Det.NumPixels();
    double mult = 0;
    check(iac_ac1d_cq_multiplicator(Det.Intensity, (uint32_t)P, &mult), "iac_ac1d_cq_multiplicator");
    iac_host::DevBuf I(ctx, P * 4), hist(ctx, (size_t)Shape.Size * 8), cq(ctx, (size_t)Shape.Size * 8);
    I.upload(Det.Intensity, P * 4);
    hist.zero();
    check(iac_ac1d_cq(ctx, Det.DeviceKMap(ctx), (uint32_t)P, I.as<float>(), Shape.Size, Shape.Max_Q,
                      Shape.dq_per_Step, IterpolMode, mult, hist.as<uint64_t>()),
          "iac_ac1d_cq in AC1D::Calculate_CQ()");
    check(iac_u64_to_f64(ctx, hist.as<uint64_t>(), Shape.Size, mult, cq.as<double>()), "iac_u64_to_f64");
    fresh(CQ, Shape.Size);
    cq.download(CQ, (size_t)Shape.Size * 8);
}

// IAC/AC1D.cpp:393-469 with the per-event work of Calculate_AC_UW_Mapped (:201-293):
// load, mask, CreateSparseHitList(Photonisation), pair histogram -- all events of
// [LowerBoundary, UpperBoundary) in device stacks.
void AC1D::Calculate_AC_UW_MR(Settings &Options, Detector &RefDet, Settings::Interpolation IterpolMode,
                              std::array<float, 2> Photonisation, unsigned int LowerBoundary,
                              unsigned int UpperBoundary, unsigned int)
{
    if (Options.HitEvents.empty() || UpperBoundary <= LowerBoundary || UpperBoundary > Options.HitEvents.size())
        throw std::runtime_error("AC1D::Calculate_AC_UW_MR: no entries in HitEvents");
    iac_host::Reserved res(Options, "AC1D::Calculate_AC_UW_MR");
    iac_ctx *ctx = res.ctx;
    const size_t P = RefDet.NumPixels();
    const uint32_t N = UpperBoundary - LowerBoundary;
    const uint32_t B = std::max<uint32_t>(1, std::min<uint32_t>(Options.BatchFrames, (uint32_t)((1ull << 31) / (P * 20))));
    iac_host::DevBuf frames(ctx, (size_t)B * P * 4), rec(ctx, (size_t)B * P * 16), off(ctx, (size_t)(B + 1) * 8);
    iac_host::DevBuf hist(ctx, (size_t)Shape.Size * 8);
    hist.zero();
    const float *kmap = RefDet.DeviceKMap(ctx);
    const int *mask = RefDet.DeviceMask(ctx);
    // frames read one stack ahead on a reader thread into page-locked stacks
    iac_host::FrameStream stream(Options, Options.HitEvents, LowerBoundary, N, P, B);
    uint32_t nb = 0;
    while (const float *stack = stream.next(nb)) {
        frames.upload(stack, (size_t)nb * P * 4);
        stream.release();
        uint64_t n = 0;
        check(iac_sparse_hit_list(ctx, kmap, (uint32_t)P, frames.as<float>(), nb, mask, Photonisation[0],
                                  Photonisation[1], rec.as<float>(), off.as<uint64_t>(), &n),
              "iac_sparse_hit_list in AC1D::Calculate_AC_UW_MR()");
        check(iac_ac1d_acuw_records(ctx, rec.as<float>(), off.as<uint64_t>(), nb, n, Shape.Size, Shape.Max_Q,
                                    Shape.dq_per_Step, IterpolMode, 2, hist.as<double>()),
              "iac_ac1d_acuw_records in AC1D::Calculate_AC_UW_MR()");
    }
    fresh(AC_UW, Shape.Size);
    hist.download(AC_UW, (size_t)Shape.Size * 8);
}
