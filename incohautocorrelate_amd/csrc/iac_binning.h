// iac_binning.h -- exact per-pair binning of the 3D AC_uw path on gfx950.
//
// Every function reproduces the reference CPU path's float/double operation
// sequence (IAC/ACMesh.cpp:209-385, IAC/ArrayOperators.h:228-238,
// IAC/Settings.h:168-175) so that each pair's voxel index and fixed-point value
// are bit-identical; uint64 accumulation then makes the mesh order independent.
// Translation units including this header are built with -ffp-contract=off and
// correctly rounded fp32 division; the pragma below keeps contraction off even
// if a unit is built without the flag.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace iac {

struct BinParams {
    float cut_sq;        // s >= cut_sq  <=>  sqrtf(s) > Max_Q   (ACMesh.cpp:216)
    float vox;           // (float)Voxel_per_dq                   (ACMesh.cpp:221)
    int c0, c1, c2;      // Shape.Center
    int sab, sc;         // Size_AB, Size_C
    int interp1, interp2, double_map;
    int rot_identity;    // rotation is exactly the identity matrix
    // symmetric binning valued by FloatToInt(double): a linear second level over
    // the integral first-level voxel of a double map under the identity rotation
    // puts the whole pair on one corner, weight 1 (the LDS engine's pair spec only)
    int dval;
};

// 1 / (double)0.001f, correctly rounded (the compiler divides in double)
constexpr double kInvStep = 1.0 / (double)0.001f;

// FloatToInt(float) then the implicit long -> unsigned int of ACMesh.cpp:390-391.
// For integer-valued p in [0, 21054) the result is exactly 1000*p (checked
// exhaustively against the formula); other values take the formula itself, its
// float quotient p / 0.001f as (float)((double)p * RN(1 / 0.001f)).  The exact
// quotient of two floats is never within 2^-49 (relative) of a float rounding
// midpoint and the double product errs by at most 2^-52, so the two round alike
// whenever the quotient is a normal float; a smaller one (|p| < 1.2e-41) gives 0
// after the + 0.5f either way.
__device__ __forceinline__ uint32_t pair_val(float p)
{
    if (p >= 0.0f && p < 21054.0f && p == truncf(p)) return (uint32_t)p * 1000u;
    float t = (float)((double)p * kInvStep) + 0.5f;
    return (uint32_t)(int64_t)floorf(t);
}

// FloatToInt(double) -> unsigned int, Settings.h:172-175 / ACMesh.cpp:412
__device__ __forceinline__ uint32_t pair_val_d(double f)
{
    return (uint32_t)(int64_t)floor(((f + 0.0) / (double)0.001f) + 0.5);
}

// ArrayOperators::Rotate (row major, q = M q_local), evaluated left to right.
__device__ __forceinline__ void rotate3(float &x, float &y, float &z, const float *R)
{
    float r0 = x * R[0] + y * R[1] + z * R[2];
    float r1 = x * R[3] + y * R[4] + z * R[5];
    float r2 = x * R[6] + y * R[7] + z * R[8];
    x = r0;
    y = r1;
    z = r2;
}

// Nearest-neighbour second level with nearest-neighbour (or no) first level:
// returns false when the pair is cut by Max_Q, otherwise the rounded voxel
// offsets (dx,dy,dz) from the centre.  The mapping is odd in q: the mirrored
// pair (b,a) lands exactly on the point-reflected voxel (negation is exact in
// every float operation used and round() is half-away-from-zero).
__device__ __forceinline__ bool bin_nn(float ax, float ay, float az, float bx, float by, float bz,
                                       const BinParams &P, const float *R, int &dx, int &dy,
                                       int &dz)
{
    float q0 = ax - bx, q1 = ay - by, q2 = az - bz;
    float s = (q0 * q0 + q1 * q1) + q2 * q2;
    if (!(s < P.cut_sq)) return false;
    q0 = q0 * P.vox;
    q1 = q1 * P.vox;
    q2 = q2 * P.vox;
    if (P.double_map) {  // (float)((int)round(q) + C - C) == round(q) for |q| < 2^24
        q0 = roundf(q0);
        q1 = roundf(q1);
        q2 = roundf(q2);
    }
    if (!P.rot_identity) rotate3(q0, q1, q2, R);
    dx = (int)roundf(q0);
    dy = (int)roundf(q1);
    dz = (int)roundf(q2);
    return true;
}

// Atomic_Add_rotated_q_Entry, ACMesh.cpp:319-385.  `1 - (q - floor(q))` is
// float arithmetic in the linear branch (int literal, float floor overload).
template <class Sink>
__device__ __forceinline__ void add_rotated(Sink &sink, float q0, float q1, float q2, float value,
                                            const BinParams &P)
{
    if (P.interp2 == 0) {
        sink.add((int)roundf(q0) + P.c0, (int)roundf(q1) + P.c1, (int)roundf(q2) + P.c2,
                 pair_val(value));
    } else {
        double sfs = (double)(1.0f - (q0 - floorf(q0)));
        double sms = (double)(1.0f - (q1 - floorf(q1)));
        double sss = (double)(1.0f - (q2 - floorf(q2)));
        int fs = (int)floorf(q0) + P.c0;
        int ms = (int)floorf(q1) + P.c1;
        int ss = (int)floorf(q2) + P.c2;
        double v = (double)value;
        // corner (ex, ey, ez) gets v * ((wx * wy) * wz), w = s or 1 - s per axis.  A corner of
        // weight 0 adds FloatToInt(0) = 0 (finite v) and is skipped: an integral q -- the first
        // level's voxel under the identity rotation -- leaves one corner of eight
#pragma unroll 1
        for (int c = 0; c < 8; c++) {
            const int ex = c & 1, ey = (c >> 1) & 1, ez = (c >> 2) & 1;
            const double wx = ex ? (1 - sfs) : sfs;
            const double wy = ey ? (1 - sms) : sms;
            const double wz = ez ? (1 - sss) : sss;
            const double w = (wx * wy) * wz;
            if (w == 0.0) continue;
            sink.add(fs + ex, ms + ey, ss + ez, pair_val_d(v * w));
        }
    }
}

// Atomic_Add_q_Entry, ACMesh.cpp:213-317, for one ordered pair: q = k_a - k_b,
// value = v_a * v_b.  Used for the linear interpolation modes.
template <class Sink>
__device__ __forceinline__ void add_q_entry(Sink &sink, float ax, float ay, float az, float bx,
                                            float by, float bz, float value, const BinParams &P,
                                            const float *R)
{
    float q0 = ax - bx, q1 = ay - by, q2 = az - bz;
    float s = (q0 * q0 + q1 * q1) + q2 * q2;
    if (!(s < P.cut_sq)) return;
    q0 = q0 * P.vox;
    q1 = q1 * P.vox;
    q2 = q2 * P.vox;
    if (P.double_map) {
        if (P.interp1 == 0) {
            q0 = roundf(q0);
            q1 = roundf(q1);
            q2 = roundf(q2);
            if (!P.rot_identity) rotate3(q0, q1, q2, R);
            add_rotated(sink, q0, q1, q2, value, P);
        } else {
            double sfs = 1.0 - (double)(q0 - floorf(q0));
            double sms = 1.0 - (double)(q1 - floorf(q1));
            double sss = 1.0 - (double)(q2 - floorf(q2));
            float f0 = floorf(q0), f1 = floorf(q1), f2 = floorf(q2);
            double v = (double)value;
#pragma unroll 1  // eight corners, each up to eight second-level adds: keep the code (and registers) small
            for (int c = 0; c < 8; c++) {
                const int ex = c & 1, ey = (c >> 2) & 1, ez = (c >> 1) & 1;
                // (float)((fs_l + e) - C) == floor(q) + e exactly for small q
                float x = f0 + (float)ex, y = f1 + (float)ey, z = f2 + (float)ez;
                if (!P.rot_identity) rotate3(x, y, z, R);
                double wx = ex ? (1 - sfs) : sfs;
                double wy = ey ? (1 - sms) : sms;
                double wz = ez ? (1 - sss) : sss;
                if (wx == 0.0 || wy == 0.0 || wz == 0.0) continue;  // its value is 0: every add is 0
                add_rotated(sink, x, y, z, (float)(((v * wx) * wy) * wz), P);
            }
        }
    } else {
        if (!P.rot_identity) rotate3(q0, q1, q2, R);
        add_rotated(sink, q0, q1, q2, value, P);
    }
}

}  // namespace iac
