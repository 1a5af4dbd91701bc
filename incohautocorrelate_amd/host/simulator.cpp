// simulator.cpp -- Simulator (IAC/Simulator.h, IAC/Simulator.cpp:55-130, 662-944):
// pixel orientation / size guess, subpixel steps, phases per frame and mode, the
// device field (split-precision complex GEMM or the FP32 direct sum), rescale,
// Poisson, noise and the event bookkeeping of SimulatePart.
#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <stdexcept>
#include <vector>

#include "host_util.hpp"
#include "iac_host.hpp"

using iac_host::check;

namespace {

// IAC/Simulator.cpp:69-120
void guess_pixel_geometry(const Detector &Det, Simulator::SimulationSettings &S)
{
    const float *pm = Det.PixelMap;
    const size_t row = 3 * (size_t)Det.DetectorSize[1];
    double v[6] = {pm[3] - pm[0], pm[4] - pm[1], pm[5] - pm[2], pm[row] - pm[0], pm[row + 1] - pm[1], pm[row + 2] - pm[2]};
    const double n0 = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double n1 = sqrt(v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    if (S.AutoPixelSize) {
        S.PixelSize[0] = n0;
        S.PixelSize[1] = n1;
    }
    if (S.AutoPixelOrientation) {
        for (int i = 0; i < 3; i++) S.PixelOrientationVectors[i] = v[i] / n0;
        for (int i = 3; i < 6; i++) S.PixelOrientationVectors[i] = v[i] / n1;
    }
}

// IAC/Simulator.cpp:742-760
void subpixel_steps(const Simulator::SimulationSettings &S, double u[3], double v[3])
{
    const double *o = S.PixelOrientationVectors;
    double t = 1.0 / sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    t = (t / ((double)(2 * S.SubSampling + 1))) * S.PixelSize[0];
    for (int i = 0; i < 3; i++) u[i] = o[i] * t;
    t = 1.0 / sqrt(o[3] * o[3] + o[4] * o[4] + o[5] * o[5]);
    t = (t / ((double)(2 * S.SubSampling + 1))) * S.PixelSize[1];
    for (int i = 0; i < 3; i++) v[i] = o[3 + i] * t;
}

}  // namespace

void Simulator::Simulate(Detector &Det, const std::vector<std::array<double, 3>> &Emitters,
                         SimulationSettings &SimSettings, SimulationOutput &Output, Settings &Options,
                         const double *Phases)
{
    if (!Det.Checklist.PixelMap || !Det.PixelMap)
        throw std::runtime_error("Simulator::Simulate: detector needs a pixel map");
    if (Emitters.empty()) throw std::runtime_error("Simulator::Simulate: no emitters");
    guess_pixel_geometry(Det, SimSettings);
    double us[3], vs[3];
    subpixel_steps(SimSettings, us, vs);
    const size_t P = Det.NumPixels();
    const uint32_t nem = (uint32_t)Emitters.size();
    const uint32_t N = SimSettings.NumberOfSimulations, M = std::max(1u, SimSettings.Modes);
    Output.Intensities.clear();
    Output.HitEvents.clear();
    Output.DetectorSize[0] = Det.DetectorSize[0];
    Output.DetectorSize[1] = Det.DetectorSize[1];
    if (N == 0) return;

    iac_host::Reserved res(Options, "Simulator::Simulate");
    iac_ctx *ctx = res.ctx;
    // positions and phases as the reference hands them to the kernel: float
    // (IAC/Simulator.cpp:727-737), carried in double to the device
    std::vector<double> em(3 * (size_t)nem);
    for (uint32_t e = 0; e < nem; e++)
        for (int d = 0; d < 3; d++) em[3 * e + d] = (double)(float)Emitters[e][d];
    iac_host::DevBuf d_pm(ctx, P * 12), d_em(ctx, em.size() * 8);
    d_pm.upload(Det.PixelMap, P * 12);
    d_em.upload(em.data(), em.size() * 8);
    const bool direct = SimSettings.Precision == IAC_SIM_FP32_DIRECT;
    iac_sim_plan *plan = nullptr;
    // the plan is destroyed on every exit, exceptions included
    std::unique_ptr<iac_sim_plan, int (*)(iac_sim_plan *)> plan_guard(nullptr, iac_sim_plan_destroy);
    iac_host::DevBuf d_em32;
    if (direct) {
        std::vector<float> e32(em.begin(), em.end());
        d_em32.reset(ctx, e32.size() * 4);
        d_em32.upload(e32.data(), e32.size() * 4);
        if (SimSettings.SubSampling > 8) check(IAC_ERR_LIMIT, "Simulator::Simulate: subsampling is 0..8");
    } else {
        check(iac_sim_plan_create(ctx, d_pm.as<float>(), (uint32_t)P, (int)SimSettings.SubSampling, us, vs,
                                  SimSettings.Wavelength, d_em.as<double>(), nem, SimSettings.Precision, &plan),
              "iac_sim_plan_create in Simulator::Simulate()");
        plan_guard.reset(plan);
    }
    std::mt19937_64 rng(SimSettings.Seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    const double two_pi = 2.0 * 3.14159265358979323846;
    const uint32_t B = std::max<uint32_t>(1, std::min<uint32_t>(
        N, (uint32_t)std::min<uint64_t>(4096, (1ull << 31) / (std::max<uint64_t>(1, (uint64_t)M * nem) * 8))));
    const double n_exp = (double)SimSettings.AveragePhotonesPerEmitterOnDetector * (double)SimSettings.FlYield *
                         (double)nem;  // IAC/Simulator.cpp:871
    iac_host::DevBuf d_ph(ctx, (size_t)B * M * nem * 8), d_out(ctx, (size_t)B * P * 4), d_ph32;
    if (direct) d_ph32.reset(ctx, (size_t)B * M * nem * 4);
    std::vector<double> ph((size_t)B * M * nem);
    std::vector<float> out((size_t)B * P);
    for (uint32_t f0 = 0; f0 < N; f0 += B) {
        const uint32_t nb = std::min(B, N - f0);
        const size_t nph = (size_t)nb * M * nem;
        for (size_t i = 0; i < nph; i++)
            ph[i] = Phases ? Phases[(size_t)f0 * M * nem + i] : (double)(float)(uni(rng) * two_pi);
        if (direct) {
            std::vector<float> p32(ph.begin(), ph.begin() + nph);
            d_ph32.upload(p32.data(), nph * 4);
            const float u32[3] = {(float)us[0], (float)us[1], (float)us[2]};
            const float v32[3] = {(float)vs[0], (float)vs[1], (float)vs[2]};
            check(iac_sim_direct_fp32(ctx, d_pm.as<float>(), (uint32_t)P, (int)SimSettings.SubSampling, u32, v32,
                                      (float)SimSettings.Wavelength, d_em32.as<float>(), nem, d_ph32.as<float>(), nb,
                                      M, IAC_SIM_AMPLITUDE, d_out.as<float>()),
                  "iac_sim_direct_fp32 in Simulator::Simulate()");
        } else {
            d_ph.upload(ph.data(), nph * 8);
            check(iac_sim_run(ctx, plan, d_ph.as<double>(), nb, M, IAC_SIM_AMPLITUDE, d_out.as<float>()),
                  "iac_sim_run in Simulator::Simulate()");
        }
        check(iac_sim_postprocess(ctx, d_out.as<float>(), (uint32_t)P, nb, n_exp, SimSettings.PoissonSample ? 1 : 0,
                                  SimSettings.ADUNoise, SimSettings.Value_per_Photon, SimSettings.Seed, f0),
              "iac_sim_postprocess in Simulator::Simulate()");
        d_out.download(out.data(), (size_t)nb * P * 4);
        for (uint32_t f = 0; f < nb; f++) {
            const float *I = out.data() + (size_t)f * P;
            Output.Intensities.emplace_back(I, I + P);
            Settings::HitEvent ev;  // identity orientation: emitters are given in the lab frame
            ev.SerialNumber = ev.Event = (int)(f0 + f);
            // IAC/Simulator.cpp:913-922
            float s = 0.0f;
            for (size_t j = 0; j < P; j++) s += I[j];
            double mean = s;
            ev.PhotonCount = (int)floor((mean / ((double)SimSettings.Value_per_Photon)) + 0.5);
            mean = mean / ((double)P);
            ev.MeanIntensity = (float)mean;
            Output.HitEvents.push_back(ev);
        }
    }
}
