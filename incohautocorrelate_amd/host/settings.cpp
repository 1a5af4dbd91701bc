// settings.cpp -- Settings (IAC/Settings.h, IAC/Settings.cpp): hit events,
// fixed point, rotation inversion, the device pool and the frame source.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <thread>

#include "host_util.hpp"
#include "iac_host.hpp"

Settings::Settings() {}

// IAC/Settings.h:86-99: "M.mm.rr" from INTERNAL_VERSION and its revision
std::string Settings::GetVersion()
{
    std::string ret = std::to_string(INTERNAL_VERSION / 100) + ".";
    if (INTERNAL_VERSION % 100 < 10) ret += "0";
    ret += std::to_string(INTERNAL_VERSION % 100) + ".";
    if (INTERNAL_VERSION_Revision / 10 == 0) ret += "0";
    return ret + std::to_string(INTERNAL_VERSION_Revision);
}

Settings::~Settings()
{
    if (owns_ctxs_)
        for (iac_ctx *c : ctxs_) iac_ctx_destroy(c);
}

void Settings::Echo(const std::string &output)
{
    if (echo) std::cout << output << "\n";
}

// IAC/Settings.h:168-187
long Settings::FloatToInt(float f) const
{
    return (long)floorf(((f + F_I_Conversion.Offset) / F_I_Conversion.Step) + 0.5f);
}
long Settings::FloatToInt(double f) const
{
    return (long)floor(((f + F_I_Conversion.Offset) / F_I_Conversion.Step) + 0.5);
}
float Settings::IntToFloat(int i) const { return ((float)i * F_I_Conversion.Step) - (float)F_I_Conversion.Offset; }
float Settings::IntToFloat(long i) const { return ((float)i * F_I_Conversion.Step) - (float)F_I_Conversion.Offset; }
float Settings::IntToFloat(unsigned long i) const
{
    return ((float)i * F_I_Conversion.Step) - (float)F_I_Conversion.Offset;
}

// IAC/Settings.cpp:539-560: the inverse of each event's matrix in double
// (adjugate times 1 / determinant), stored back as float.
void Settings::InvertRotationMatrices()
{
    for (HitEvent &ev : HitEvents) {
        double a[9];
        for (int i = 0; i < 9; i++) a[i] = ev.RotMatrix[i];
        const double adj[9] = {a[4] * a[8] - a[5] * a[7], a[2] * a[7] - a[1] * a[8], a[1] * a[5] - a[2] * a[4],
                               a[5] * a[6] - a[3] * a[8], a[0] * a[8] - a[2] * a[6], a[2] * a[3] - a[0] * a[5],
                               a[3] * a[7] - a[4] * a[6], a[1] * a[6] - a[0] * a[7], a[0] * a[4] - a[1] * a[3]};
        const double det = a[0] * adj[0] + a[1] * adj[3] + a[2] * adj[6];
        if (det == 0.0) throw std::runtime_error("Settings::InvertRotationMatrices: singular rotation matrix");
        const double inv_det = 1.0 / det;
        for (int i = 0; i < 9; i++) ev.RotMatrix[i] = (float)(adj[i] * inv_det);
    }
}

void Settings::checkErr(int err, const char *name)
{
    if (err != IAC_OK) {
        std::cerr << "ERROR: " << name << " (" << err << ": " << iac_strerror(err) << ") " << iac_last_error()
                  << std::endl;
        exit(EXIT_FAILURE);
    }
}

void Settings::SetUp_GPU(int MaxDevices)
{
    if (!ctxs_.empty()) return;
    int n = 0;
    checkErr(iac_device_count(&n), "iac_device_count in Settings::SetUp_GPU()");
    if (n <= 0) checkErr(IAC_ERR_NO_DEVICE, "Settings::SetUp_GPU(): no HIP device visible");
    if (MaxDevices > 0 && MaxDevices < n) n = MaxDevices;
    for (int d = 0; d < n; d++) {
        iac_ctx *c = nullptr;
        checkErr(iac_ctx_create(d, &c), "iac_ctx_create in Settings::SetUp_GPU()");
        ctxs_.push_back(c);
        available_.push_back(true);
    }
    if (echo) std::cout << "GPU pool: " << n << " device(s)\n";
}

void Settings::UseContexts(const std::vector<iac_ctx *> &Contexts)
{
    std::lock_guard<std::mutex> g(pool_mutex_);
    if (!ctxs_.empty()) throw std::runtime_error("Settings::UseContexts: the device pool is already set up");
    if (Contexts.empty()) throw std::runtime_error("Settings::UseContexts: no contexts");
    ctxs_ = Contexts;
    available_.assign(Contexts.size(), true);
    owns_ctxs_ = false;
}

// OCL_ReserveDevice / OCL_FreeDevice, IAC/Settings.cpp:432-470
int Settings::GPU_ReserveDevice()
{
    if (ctxs_.empty()) SetUp_GPU();
    std::lock_guard<std::mutex> g(pool_mutex_);
    for (size_t i = 0; i < available_.size(); i++) {
        if (available_[i]) {
            available_[i] = false;
            return (int)i;
        }
    }
    return -1;
}

void Settings::GPU_FreeDevice(int DeviceIndex)
{
    std::lock_guard<std::mutex> g(pool_mutex_);
    if (DeviceIndex >= 0 && (size_t)DeviceIndex < available_.size()) available_[DeviceIndex] = true;
}

iac_ctx *Settings::Context(int DeviceIndex)
{
    if (ctxs_.empty()) SetUp_GPU();
    if (DeviceIndex < 0 || (size_t)DeviceIndex >= ctxs_.size())
        throw std::runtime_error("Settings::Context: no such device");
    return ctxs_[DeviceIndex];
}

void Settings::LoadFrame(const HitEvent &ev, float *dst, size_t npix) const
{
    if (FrameLoader) {
        FrameLoader(ev, dst, npix);
        return;
    }
    if (ev.Data) {
        std::copy(ev.Data, ev.Data + npix, dst);
        return;
    }
    if (ev.Filename.empty()) throw std::runtime_error("Settings::LoadFrame: event has neither data nor a file");
    iac_host::read_raw(ev.Filename, dst, npix * sizeof(float), (size_t)ev.Event * npix * sizeof(float));
}

namespace iac_host {

Reserved::Reserved(Settings &pool, const char *who) : pool_(pool)
{
    dev = pool.GPU_ReserveDevice();
    if (dev < 0) throw std::runtime_error(std::string(who) + ": no free device");
    try {
        ctx = pool.Context(dev);
    } catch (...) {
        pool.GPU_FreeDevice(dev);
        throw;
    }
}

Reserved::~Reserved() { pool_.GPU_FreeDevice(dev); }

// ArrayOperators::LoadArrayFromFile / SafeArrayToFile (Binary), IAC/ArrayOperators.h
void read_raw(const std::string &path, void *dst, size_t bytes, size_t offset)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    f.seekg((std::streamoff)offset);
    f.read(static_cast<char *>(dst), (std::streamsize)bytes);
    if ((size_t)f.gcount() != bytes) throw std::runtime_error("short read from " + path);
}

void write_raw(const std::string &path, const void *src, size_t bytes)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write(static_cast<const char *>(src), (std::streamsize)bytes);
    if (!f) throw std::runtime_error("short write to " + path);
}

}  // namespace iac_host
