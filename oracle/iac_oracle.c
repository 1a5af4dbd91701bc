/*
 * iac_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see iac_oracle.h).
 *
 * Restates the reference's arithmetic operation by operation.  Each function
 * cites the reference lines it follows (paths relative to the reference root,
 * IAC = IncohAutoCorrelate/).  Built with gcc -O3 -fopenmp -ffp-contract=off
 * and no -march, i.e. the reference's own flags (IAC/IncohAutoCorrelate.vcxproj:115)
 * plus an explicit ban on FMA contraction.
 */
#include "iac_oracle.h"

/* (int)x on the reference's x86-64 build: cvttss2si truncates toward zero and returns
 * 0x80000000 for NaN and out-of-range values (C leaves those undefined, so the rule is
 * written out; tests compare it with the host's own conversion). */
int ora_cvtt_i32(float x)
{
    return (x >= -0x1p31f && x < 0x1p31f) ? (int)x : INT32_MIN;
}
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted mistakes for the pin tests (tests/test_oracle_slips.py).  The shipped
 * oracle is built with ORA_SLIP = 0, where every SLIP(n) is a constant false and
 * the compiler drops the branch.  A test build with -DORA_SLIP=n plants mistake n:
 *   1  AC1D AC_uw: self pairs weighted 1 instead of 2
 *   2  AC1D AC_uw nearest neighbour: floor(q) instead of floor(q + 0.5)
 *   3  AC1D AC_uw linear: Sep and the split values in double, no float rounding
 *   4  AC3D linear second level: the x weights of the two corners swapped
 *   5  AC3D linear first level: the x weights of the two corners swapped
 *   6  C(q) small linear: the x weights of the two corners swapped
 *   7  Merge_CQ: the transposed rotation matrix
 *   8  Merge_CQ linear: the x weights of the two corners swapped
 *   9  AC1D C(q) linear: Sep given to the lower bin
 *  10  AC1D C(q) nearest neighbour: floor(q) instead of floor(q + 0.5) */
#ifndef ORA_SLIP
#define ORA_SLIP 0
#endif
#define SLIP(n) (ORA_SLIP == (n))

/* ------------------------------------------------------------------------- */
/* Fixed point: Settings::FloatToInt / IntToFloat, IAC/Settings.h:162-187.
 * F_I_Conversion is {Offset = 0.0f, Step = 0.001f} and never configured. */
static const float kOffset = 0.0f;
static const float kStep = 0.001f;

int64_t ora_float_to_int_f(float f) { return (int64_t)floorf(((f + kOffset) / kStep) + 0.5f); }
int64_t ora_float_to_int_d(double f)
{
    return (int64_t)floor(((f + (double)kOffset) / (double)kStep) + 0.5);
}
float ora_int_to_float_u64(uint64_t i) { return ((float)i * kStep) - (float)kOffset; }

/* `unsigned int val = Options->FloatToInt(Value);` IAC/ACMesh.cpp:388-391 */
uint32_t ora_pair_val(float p) { return (uint32_t)ora_float_to_int_f(p); }

/* ------------------------------------------------------------------------- */
/* Geometry */

/* Simulator::GeneratePixelMap, IAC/Simulator.cpp:495-519.  The element
 * expression is Center + (VecA*PS)*(a - 0.5*SizeA) + (VecB*PS)*(b - 0.5*SizeB):
 * float products, promoted to double by the 0.5 literal, stored as float. */
void ora_generate_pixel_map(int size_a, int size_b, float pixel_size, const float center[3],
                            const float vec_a_in[3], const float vec_b_in[3], float *pixmap)
{
    float va[3] = {vec_a_in[0], vec_a_in[1], vec_a_in[2]};
    float vb[3] = {vec_b_in[0], vec_b_in[1], vec_b_in[2]};
    float norm = sqrtf(va[0] * va[0] + va[1] * va[1] + va[2] * va[2]);
    va[0] = va[0] / norm;
    va[1] = va[1] / norm;
    va[2] = va[2] / norm;
    norm = sqrtf(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
    vb[0] = vb[0] / norm;
    vb[1] = vb[1] / norm;
    vb[2] = vb[2] / norm;
    for (int a = 0; a < size_a; a++) {
        for (int b = 0; b < size_b; b++) {
            double da = (double)a - 0.5 * size_a;
            double db = (double)b - 0.5 * size_b;
            for (int d = 0; d < 3; d++) {
                float ca = va[d] * pixel_size;
                float cb = vb[d] * pixel_size;
                pixmap[3 * (size_t)a * size_b + 3 * (size_t)b + d] =
                    (float)((double)center[d] + (double)ca * da + (double)cb * db);
            }
        }
    }
}

/* Detector::Calc_kMap, IAC/Detector.cpp:310-361 */
void ora_calc_kmap(const float *pixmap, uint32_t npix, float *kmap, float max_k[3],
                   float min_k[3], float max_q[3])
{
    for (int d = 0; d < 3; d++) {
        max_k[d] = -100.0f;
        min_k[d] = 100.0f;
    }
    for (uint32_t i = 0; i < npix; i++) {
        float x = pixmap[3 * (size_t)i + 0];
        float y = pixmap[3 * (size_t)i + 1];
        float z = pixmap[3 * (size_t)i + 2];
        float r = sqrtf(x * x + y * y + z * z);
        x = x / r;
        y = y / r;
        z = z / r;
        kmap[3 * (size_t)i + 0] = x;
        kmap[3 * (size_t)i + 1] = y;
        kmap[3 * (size_t)i + 2] = z;
        if (x > max_k[0]) max_k[0] = x;
        if (y > max_k[1]) max_k[1] = y;
        if (z > max_k[2]) max_k[2] = z;
        if (x < min_k[0]) min_k[0] = x;
        if (y < min_k[1]) min_k[1] = y;
        if (z < min_k[2]) min_k[2] = z;
    }
    for (int d = 0; d < 3; d++) max_q[d] = fabsf(max_k[d] - min_k[d]);
}

/* ACMesh::CreateBigMeshForDetector, IAC/ACMesh.cpp:74-106 (CreateBigMesh_CofQ_ForDetector
 * :107-139 computes the same shape).  Returns -1 for an even edge size (the reference
 * throws std::invalid_argument, :76-81). */
int ora_mesh_shape_big(const float max_q[3], int edge_size, float q_zoom, ora_mesh_shape *s)
{
    if (edge_size % 2 == 0) return -1;
    memset(s, 0, sizeof(*s));
    s->k_a = 0;
    s->k_b = 1;
    s->k_c = 2;
    s->size_ab = (uint32_t)(edge_size + 2);
    s->size_c = s->size_ab;
    s->center[0] = (s->size_ab - 1) / 2;
    s->center[1] = (s->size_ab - 1) / 2;
    s->center[2] = (s->size_c - 1) / 2;
    float mq = fmaxf(fmaxf(max_q[0], max_q[1]), max_q[2]);
    float maxq = mq / q_zoom;
    s->max_q = maxq;
    uint32_t den = ((s->size_ab - 1) / 2) - 2;
    float dq = maxq / (float)den;
    s->voxel_per_dq = (uint32_t)floorf(1 / dq);
    s->dq_per_voxel = 1.0f / ((float)s->voxel_per_dq);
    return 0;
}

/* ACMesh::CreateSmallMesh_CofQ_ForDetector, IAC/ACMesh.cpp:140-206 */
int ora_mesh_shape_small_cofq(const float max_q[3], int perp_size, float q_zoom,
                              ora_mesh_shape *s)
{
    if (perp_size % 2 == 0) return -1;
    memset(s, 0, sizeof(*s));
    uint32_t ka = 0, kb = 1, kc = 2, t;
    if (max_q[kb] > max_q[ka]) { t = ka; ka = kb; kb = t; }
    if (max_q[kc] > max_q[kb]) { t = kc; kc = kb; kb = t; }
    if (max_q[kb] > max_q[ka]) { t = ka; ka = kb; kb = t; }
    s->k_a = ka;
    s->k_b = kb;
    s->k_c = kc;
    s->size_ab = (uint32_t)(perp_size + 2);
    if (q_zoom == 1) {
        float ratio = (max_q[kc] / max_q[ka]) * (float)perp_size;
        s->size_c = (uint32_t)(int)floor((double)ratio + 2.5);
    } else {
        s->size_c = s->size_ab;
    }
    if (s->size_c % 2 == 0) s->size_c += 1;
    s->center[0] = (s->size_ab - 1) / 2;
    s->center[1] = (s->size_ab - 1) / 2;
    s->center[2] = (s->size_c - 1) / 2;
    float maxq = max_q[ka] / q_zoom;
    s->max_q = maxq;
    uint32_t den = ((s->size_ab - 1) / 2) - 2;
    float dq = maxq / (float)den;
    s->voxel_per_dq = (uint32_t)floorf(1 / dq);
    s->dq_per_voxel = 1.0f / ((float)s->voxel_per_dq);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Sparse hit list */

/* Detector::CreateSparseHitList(Threshold) IAC/Detector.cpp:926-948, then the
 * DiscretizeToPhotones pass :950-958 -> ArrayOperators::DiscretizeToPhotons
 * IAC/ArrayOperators.h:138-148 (ceilf((I-T)/S) * (I >= T)); the frame was
 * multiplied by the pixel mask first (IAC/RunIAC.cpp:532, ArrayOperators.h:120-127). */
uint32_t ora_sparse_hit_list(const float *intensity, const int *mask, const float *kmap,
                             uint32_t npix, float threshold, float step, float *out4)
{
    uint32_t n = 0;
    for (uint32_t i = 0; i < npix; i++) {
        float I = intensity[i];
        if (mask) I = I * (float)mask[i];
        if (I >= threshold && I > 0) {
            float v = I;
            if (step > 0) v = ceilf((v - threshold) / step) * (float)(v >= threshold);
            out4[4 * (size_t)n + 0] = kmap[3 * (size_t)i + 0];
            out4[4 * (size_t)n + 1] = kmap[3 * (size_t)i + 1];
            out4[4 * (size_t)n + 2] = kmap[3 * (size_t)i + 2];
            out4[4 * (size_t)n + 3] = v;
            n++;
        }
    }
    return n;
}

void ora_hit_list_from_photons(const float *kmap, const uint32_t *pix, const uint16_t *cnt,
                               uint32_t n, float *out4)
{
    for (uint32_t i = 0; i < n; i++) {
        uint32_t p = pix[i];
        out4[4 * (size_t)i + 0] = kmap[3 * (size_t)p + 0];
        out4[4 * (size_t)i + 1] = kmap[3 * (size_t)p + 1];
        out4[4 * (size_t)i + 2] = kmap[3 * (size_t)p + 2];
        out4[4 * (size_t)i + 3] = (float)cnt[i];
    }
}

/* ------------------------------------------------------------------------- */
/* 3D binning: ACMesh::Atomic_Add_q_Entry / Atomic_Add_rotated_q_Entry */

/* ArrayOperators::Rotate, IAC/ArrayOperators.h:228-238 (row major, q = M q_local) */
static inline void rotate(float v[3], const float m[9])
{
    float r0 = v[0] * m[0] + v[1] * m[1] + v[2] * m[2];
    float r1 = v[0] * m[3] + v[1] * m[4] + v[2] * m[5];
    float r2 = v[0] * m[6] + v[1] * m[7] + v[2] * m[8];
    v[0] = r0;
    v[1] = r1;
    v[2] = r2;
}

static inline void mesh_add(uint64_t *mesh, const ora_mesh_shape *s, int fs, int ms, int ss,
                            uint64_t v, uint64_t *dropped)
{
    int sab = (int)s->size_ab, sc = (int)s->size_c;
    if (fs < 0 || ms < 0 || ss < 0 || fs >= sab || ms >= sab || ss >= sc) {
#pragma omp atomic
        (*dropped)++;
        return;
    }
    size_t idx = (size_t)fs + (size_t)ms * sab + (size_t)ss * sab * sab;
#pragma omp atomic
    mesh[idx] += v;
}

/* ACMesh::Atomic_Add_rotated_q_Entry, IAC/ACMesh.cpp:319-385.  In the linear
 * branch `1 - (q - floor(q))` is evaluated in float (int literal, float
 * overload of floor) and only then widened to double. */
static void add_rotated(uint64_t *mesh, const ora_mesh_shape *s, const float q[3], float value,
                        int interp2, uint64_t *dropped)
{
    int c0 = (int)s->center[0], c1 = (int)s->center[1], c2 = (int)s->center[2];
    if (interp2 == 0) {
        int fs = (int)roundf(q[0]) + c0;
        int ms = (int)roundf(q[1]) + c1;
        int ss = (int)roundf(q[2]) + c2;
        uint32_t val = (uint32_t)ora_float_to_int_f(value);
        mesh_add(mesh, s, fs, ms, ss, val, dropped);
    }
    if (interp2 == 1) {
        double sfs = (double)(1.0f - (q[0] - floorf(q[0])));
        double sms = (double)(1.0f - (q[1] - floorf(q[1])));
        double sss = (double)(1.0f - (q[2] - floorf(q[2])));
        if (SLIP(4)) sfs = 1 - sfs;
        int fs = (int)floorf(q[0]) + c0;
        int ms = (int)floorf(q[1]) + c1;
        int ss = (int)floorf(q[2]) + c2;
        double v = (double)value;
        mesh_add(mesh, s, fs + 0, ms + 0, ss + 0,
                 (uint32_t)ora_float_to_int_d(v * ((sfs * sms) * sss)), dropped);
        mesh_add(mesh, s, fs + 1, ms + 0, ss + 0,
                 (uint32_t)ora_float_to_int_d(v * (((1 - sfs) * sms) * sss)), dropped);
        mesh_add(mesh, s, fs + 0, ms + 0, ss + 1,
                 (uint32_t)ora_float_to_int_d(v * ((sfs * sms) * (1 - sss))), dropped);
        mesh_add(mesh, s, fs + 1, ms + 0, ss + 1,
                 (uint32_t)ora_float_to_int_d(v * (((1 - sfs) * sms) * (1 - sss))), dropped);
        mesh_add(mesh, s, fs + 0, ms + 1, ss + 0,
                 (uint32_t)ora_float_to_int_d(v * ((sfs * (1 - sms)) * sss)), dropped);
        mesh_add(mesh, s, fs + 1, ms + 1, ss + 0,
                 (uint32_t)ora_float_to_int_d(v * (((1 - sfs) * (1 - sms)) * sss)), dropped);
        mesh_add(mesh, s, fs + 0, ms + 1, ss + 1,
                 (uint32_t)ora_float_to_int_d(v * ((sfs * (1 - sms)) * (1 - sss))), dropped);
        mesh_add(mesh, s, fs + 1, ms + 1, ss + 1,
                 (uint32_t)ora_float_to_int_d(v * (((1 - sfs) * (1 - sms)) * (1 - sss))), dropped);
    }
}

/* ACMesh::Atomic_Add_q_Entry, IAC/ACMesh.cpp:213-317.  q is modified in place
 * like the reference's q_local. */
static void add_q_entry(uint64_t *mesh, const ora_mesh_shape *s, float q[3], const float R[9],
                        float value, int interp1, int interp2, int double_map, uint64_t *dropped)
{
    if (sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) > s->max_q) return;
    float f = (float)s->voxel_per_dq;
    q[0] *= f;
    q[1] *= f;
    q[2] *= f;
    int c0 = (int)s->center[0], c1 = (int)s->center[1], c2 = (int)s->center[2];
    if (double_map) {
        if (interp1 == 0) {
            int fs_l = (int)roundf(q[0]) + c0;
            int ms_l = (int)roundf(q[1]) + c1;
            int ss_l = (int)roundf(q[2]) + c2;
            q[0] = (float)(fs_l - c0);
            q[1] = (float)(ms_l - c1);
            q[2] = (float)(ss_l - c2);
            rotate(q, R);
            add_rotated(mesh, s, q, value, interp2, dropped);
        }
        if (interp1 == 1) {
            double sfs = 1.0 - (double)(q[0] - floorf(q[0]));
            double sms = 1.0 - (double)(q[1] - floorf(q[1]));
            double sss = 1.0 - (double)(q[2] - floorf(q[2]));
            int fs_l = (int)floorf(q[0]) + c0;
            int ms_l = (int)floorf(q[1]) + c1;
            int ss_l = (int)floorf(q[2]) + c2;
            double v = (double)value;
            /* corner order and weight expressions of :311-365 */
            const int dx[8] = {0, 1, 0, 1, 0, 1, 0, 1};
            const int dy[8] = {0, 0, 0, 0, 1, 1, 1, 1};
            const int dz[8] = {0, 0, 1, 1, 0, 0, 1, 1};
            for (int c = 0; c < 8; c++) {
                float qq[3];
                qq[0] = (float)((fs_l + dx[c]) - c0);
                qq[1] = (float)((ms_l + dy[c]) - c1);
                qq[2] = (float)((ss_l + dz[c]) - c2);
                rotate(qq, R);
                double wx = dx[c] ? (1 - sfs) : sfs;
                if (SLIP(5)) wx = 1 - wx;
                double wy = dy[c] ? (1 - sms) : sms;
                double wz = dz[c] ? (1 - sss) : sss;
                float w = (float)(((v * wx) * wy) * wz);
                add_rotated(mesh, s, qq, w, interp2, dropped);
            }
        }
    } else {
        rotate(q, R);
        add_rotated(mesh, s, q, value, interp2, dropped);
    }
}

/* Detector::AutoCorrelateSparseList CPU branch, IAC/Detector.cpp:1025-1051:
 * `omp parallel for` over i, j from i, the (i,j) update and the mirrored (j,i)
 * update when i != j, value SparseHitList[i][3]*SparseHitList[j][3]. */
void ora_ac3d_sparse(const float *hl4, uint32_t n, const float rot[9], const ora_mesh_shape *s,
                     int interp1, int interp2, int double_map, uint64_t *mesh,
                     uint64_t *dropped, int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
#pragma omp parallel for num_threads(nt)
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t j = i; j < n; j++) {
            float q[3];
            float value = hl4[4 * (size_t)i + 3] * hl4[4 * (size_t)j + 3];
            q[0] = hl4[4 * (size_t)i + 0] - hl4[4 * (size_t)j + 0];
            q[1] = hl4[4 * (size_t)i + 1] - hl4[4 * (size_t)j + 1];
            q[2] = hl4[4 * (size_t)i + 2] - hl4[4 * (size_t)j + 2];
            add_q_entry(mesh, s, q, rot, value, interp1, interp2, double_map, dropped);
            q[0] = hl4[4 * (size_t)j + 0] - hl4[4 * (size_t)i + 0];
            q[1] = hl4[4 * (size_t)j + 1] - hl4[4 * (size_t)i + 1];
            q[2] = hl4[4 * (size_t)j + 2] - hl4[4 * (size_t)i + 2];
            if (i != j) add_q_entry(mesh, s, q, rot, value, interp1, interp2, double_map, dropped);
        }
    }
}

static const float kIdentity[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

void ora_ac3d_sparse_frames(const float *kmap, const uint32_t *pix, const uint16_t *cnt,
                            const uint64_t *off, uint32_t nframes, const float *rots,
                            const ora_mesh_shape *s, int interp1, int interp2, int double_map,
                            uint64_t *mesh, uint64_t *dropped, int nthreads)
{
    uint64_t maxn = 0;
    for (uint32_t f = 0; f < nframes; f++)
        if (off[f + 1] - off[f] > maxn) maxn = off[f + 1] - off[f];
    float *hl = (float *)malloc(sizeof(float) * 4 * (maxn ? maxn : 1));
    for (uint32_t f = 0; f < nframes; f++) {
        uint32_t n = (uint32_t)(off[f + 1] - off[f]);
        ora_hit_list_from_photons(kmap, pix + off[f], cnt + off[f], n, hl);
        const float *R = rots ? rots + 9 * (size_t)f : kIdentity;
        ora_ac3d_sparse(hl, n, R, s, interp1, interp2, double_map, mesh, dropped, nthreads);
    }
    free(hl);
}

void ora_ac3d_bruteforce(const float *hl4, uint32_t n, const float rot[9], const ora_mesh_shape *s,
                         int interp1, int interp2, int double_map, uint64_t *mesh, uint64_t *dropped)
{
    for (uint32_t a = 0; a < n; a++) {
        for (uint32_t b = 0; b < n; b++) {
            float q[3];
            q[0] = hl4[4 * (size_t)a + 0] - hl4[4 * (size_t)b + 0];
            q[1] = hl4[4 * (size_t)a + 1] - hl4[4 * (size_t)b + 1];
            q[2] = hl4[4 * (size_t)a + 2] - hl4[4 * (size_t)b + 2];
            float value = hl4[4 * (size_t)a + 3] * hl4[4 * (size_t)b + 3];
            add_q_entry(mesh, s, q, rot, value, interp1, interp2, double_map, dropped);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* C(q) small */

/* ArrayOperators::Min_Max_Mean_Value(float*), IAC/ArrayOperators.h:160-175 */
static float mean_value_f(const float *a, uint32_t n)
{
    double t = 0;
    for (uint32_t i = 0; i < n; i++) t += a[i];
    t = t / n;
    return (float)t;
}

/* Detector::AutoCorrelate_CofQ_SmallMesh, IAC/Detector.cpp:1215-1227 */
double ora_cq_small_multiplicator(const float *intensity, uint32_t npix)
{
    float mean = mean_value_f(intensity, npix);
    if (mean <= 0) return -1.0;
    double m = 1e-10;
    for (; 1 > mean * mean * m;) m *= 10;
    return m * 100000;
}

static inline void small_add(uint64_t *mesh, const ora_mesh_shape *s, int fs, int ms, int ss,
                             uint64_t v, uint64_t *dropped)
{
    mesh_add(mesh, s, fs, ms, ss, v, dropped);
}

/* AutoCorr_CQ_small, IAC/kernel.cl:27-184: per pixel ind over all pixels i.
 * The OpenCL `1 - (q - floor(q))` in the linear branch is float arithmetic. */
void ora_cq_small(const float *intensity, const float *kmap, uint32_t npix, const ora_mesh_shape *s,
                  int interp, double mult, uint64_t *mesh, uint64_t *dropped, int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    const float vox = (float)s->voxel_per_dq;
    const float maxq = s->max_q;
    const int ai = (int)s->k_a, bi = (int)s->k_b, ci = (int)s->k_c;
    const int cab = ((int)s->size_ab - 1) / 2;
    const int cc = ((int)s->size_c - 1) / 2;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (uint32_t ind = 0; ind < npix; ind++) {
        float k1[3] = {kmap[3 * (size_t)ind], kmap[3 * (size_t)ind + 1], kmap[3 * (size_t)ind + 2]};
        for (uint32_t i = 0; i < npix; i++) {
            double val = (double)(intensity[ind] * intensity[i]);
            if (val < (double)1e-37f) continue;
            uint64_t vconv = (uint64_t)(val * mult);
            float q[3];
            q[0] = k1[0] - kmap[3 * (size_t)i + 0];
            q[1] = k1[1] - kmap[3 * (size_t)i + 1];
            q[2] = k1[2] - kmap[3 * (size_t)i + 2];
            if (sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) > maxq) continue;
            q[0] = q[0] * vox;
            q[1] = q[1] * vox;
            q[2] = q[2] * vox;
            if (interp == 0) {
                int fs = (int)(roundf(q[ai]) + (float)cab);
                int ms = (int)(roundf(q[bi]) + (float)cab);
                int ss = (int)(roundf(q[ci]) + (float)cc);
                small_add(mesh, s, fs, ms, ss, vconv, dropped);
            }
            if (interp == 1) {
                double sfs = (double)(1.0f - (q[ai] - floorf(q[ai])));
                double sms = (double)(1.0f - (q[bi] - floorf(q[bi])));
                double sss = (double)(1.0f - (q[ci] - floorf(q[ci])));
                if (SLIP(6)) sfs = 1 - sfs;
                int fs = (int)(floorf(q[ai]) + (float)cab);
                int ms = (int)(floorf(q[bi]) + (float)cab);
                int ss = (int)(floorf(q[ci]) + (float)cc);
                double vm = val * mult;
                small_add(mesh, s, fs + 0, ms + 0, ss + 0, (uint64_t)(vm * ((sfs * sms) * sss)), dropped);
                small_add(mesh, s, fs + 0, ms + 0, ss + 1, (uint64_t)(vm * ((sfs * sms) * (1 - sss))), dropped);
                small_add(mesh, s, fs + 1, ms + 0, ss + 0, (uint64_t)(vm * (((1 - sfs) * sms) * sss)), dropped);
                small_add(mesh, s, fs + 1, ms + 0, ss + 1, (uint64_t)(vm * (((1 - sfs) * sms) * (1 - sss))), dropped);
                small_add(mesh, s, fs + 0, ms + 1, ss + 0, (uint64_t)(vm * ((sfs * (1 - sms)) * sss)), dropped);
                small_add(mesh, s, fs + 0, ms + 1, ss + 1, (uint64_t)(vm * ((sfs * (1 - sms)) * (1 - sss))), dropped);
                small_add(mesh, s, fs + 1, ms + 1, ss + 0, (uint64_t)(vm * (((1 - sfs) * (1 - sms)) * sss)), dropped);
                small_add(mesh, s, fs + 1, ms + 1, ss + 1, (uint64_t)(vm * (((1 - sfs) * (1 - sms)) * (1 - sss))), dropped);
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Merge of the small C(q) into the lab-frame mesh */

/* ArrayOperators::Min_Max_Mean_Value(double*), IAC/ArrayOperators.h:177-192 */
static double mean_value_d(const double *a, size_t n)
{
    double t = 0;
    for (size_t i = 0; i < n; i++) t += a[i];
    return t / n;
}

/* Detector::Merge_smallCofQ host preparation, IAC/Detector.cpp:1339-1427 */
double ora_merge_prepare(const float *rot9, const float *mean_intensity, uint32_t nev,
                         const double *small_cq, size_t nsmall, float *rw10)
{
    double max_weight = 0.0;
    for (uint32_t i = 0; i < nev; i++) {
        for (int k = 0; k < 9; k++) rw10[10 * (size_t)i + k] = rot9[9 * (size_t)i + k];
        rw10[10 * (size_t)i + 9] =
            (float)((double)mean_intensity[i] * (double)mean_intensity[i]);
        if (mean_intensity[i] > max_weight) max_weight = mean_intensity[i];
    }
    double mean_cq = mean_value_d(small_cq, nsmall);
    double m = 1e-12;
    for (; 1 > max_weight * max_weight * mean_cq * m;) m *= 10;
    uint64_t oom = 1;
    for (; oom < (uint64_t)nev;) oom *= 10;
    m = m / oom;
    m *= 1e10;
    if (m > 0.6) m = round(m);
    if (m < 1 && m > 0.1) m = 1;
    if (m < 1e-12) m = 1e-12;
    for (uint32_t i = 0; i < nev; i++) {
        float red = (float)m;
        if (m > 0.6) red = roundf((float)m);
        float w = rw10[10 * (size_t)i + 9];
        w *= red;
        w = floorf((float)((double)w + 0.5));
        rw10[10 * (size_t)i + 9] = w / red;
    }
    return m;
}

/* Merge_CQ, IAC/kernel.cl:189-344.  The big mesh is cubic Size_AB^3 and its
 * slow-scan index uses the AB centre (:287).  `1 - (q - floor(q))` is float. */
void ora_merge_cq(const double *small_cq, const ora_mesh_shape *ss_, const float *rw10,
                  uint32_t nev, int interp, double mult, uint64_t *big, uint64_t *dropped,
                  int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    const int sab = (int)ss_->size_ab, sc = (int)ss_->size_c;
    const int cab = (sab - 1) / 2, cc = (sc - 1) / 2;
    const unsigned ai = ss_->k_a, bi = ss_->k_b, ci = ss_->k_c;
    ora_mesh_shape bigs = *ss_;
    bigs.size_c = (uint32_t)sab;
    const size_t nsmall = (size_t)sab * sab * sc;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (size_t ind = 0; ind < nsmall; ind++) {
        int ss_in = (int)(ind / ((size_t)sab * sab));
        int ms_in = (int)((ind - (size_t)ss_in * sab * sab) / sab);
        int fs_in = (int)(ind - (size_t)ss_in * sab * sab - (size_t)ms_in * sab);
        float q_in[3];
        q_in[ai] = (float)(fs_in - cab);
        q_in[bi] = (float)(ms_in - cab);
        q_in[ci] = (float)(ss_in - cc);
        double val_in = small_cq[ind];
        if (val_in < (double)1e-37f) continue;
        for (uint32_t i = 0; i < nev; i++) {
            const float *R = rw10 + 10 * (size_t)i;
            float q[3];
            if (SLIP(7)) {
                q[0] = R[0] * q_in[0] + R[3] * q_in[1] + R[6] * q_in[2];
                q[1] = R[1] * q_in[0] + R[4] * q_in[1] + R[7] * q_in[2];
                q[2] = R[2] * q_in[0] + R[5] * q_in[1] + R[8] * q_in[2];
            } else {
                q[0] = R[0] * q_in[0] + R[1] * q_in[1] + R[2] * q_in[2];
                q[1] = R[3] * q_in[0] + R[4] * q_in[1] + R[5] * q_in[2];
                q[2] = R[6] * q_in[0] + R[7] * q_in[1] + R[8] * q_in[2];
            }
            double w = (double)R[9];
            if (interp == 0) {
                uint64_t vout = (uint64_t)((val_in * w) * mult);
                int fs = (int)(roundf(q[0]) + (float)cab);
                int ms = (int)(roundf(q[1]) + (float)cab);
                int ss = (int)(roundf(q[2]) + (float)cab);
                mesh_add(big, &bigs, fs, ms, ss, vout, dropped);
            }
            if (interp == 1) {
                double sfs = (double)(1.0f - (q[0] - floorf(q[0])));
                double sms = (double)(1.0f - (q[1] - floorf(q[1])));
                double sss = (double)(1.0f - (q[2] - floorf(q[2])));
                if (SLIP(8)) sfs = 1 - sfs;
                int fs = (int)(floorf(q[0]) + (float)cab);
                int ms = (int)(floorf(q[1]) + (float)cab);
                int ss = (int)(floorf(q[2]) + (float)cab);
                double vmw = (val_in * mult) * w;
                mesh_add(big, &bigs, fs + 0, ms + 0, ss + 0, (uint64_t)(vmw * ((sfs * sms) * sss)), dropped);
                mesh_add(big, &bigs, fs + 0, ms + 0, ss + 1, (uint64_t)(vmw * ((sfs * sms) * (1 - sss))), dropped);
                mesh_add(big, &bigs, fs + 1, ms + 0, ss + 0, (uint64_t)(vmw * (((1 - sfs) * sms) * sss)), dropped);
                mesh_add(big, &bigs, fs + 1, ms + 0, ss + 1, (uint64_t)(vmw * (((1 - sfs) * sms) * (1 - sss))), dropped);
                mesh_add(big, &bigs, fs + 0, ms + 1, ss + 0, (uint64_t)(vmw * ((sfs * (1 - sms)) * sss)), dropped);
                mesh_add(big, &bigs, fs + 0, ms + 1, ss + 1, (uint64_t)(vmw * ((sfs * (1 - sms)) * (1 - sss))), dropped);
                mesh_add(big, &bigs, fs + 1, ms + 1, ss + 0, (uint64_t)(vmw * (((1 - sfs) * (1 - sms)) * sss)), dropped);
                mesh_add(big, &bigs, fs + 1, ms + 1, ss + 1, (uint64_t)(vmw * (((1 - sfs) * (1 - sms)) * (1 - sss))), dropped);
            }
        }
    }
}

void ora_cq_small_f64(const float *intensity, const float *kmap, uint32_t npix,
                      const ora_mesh_shape *s, double *mesh, int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    const float vox = (float)s->voxel_per_dq;
    const int ai = (int)s->k_a, bi = (int)s->k_b, ci = (int)s->k_c;
    const int sab = (int)s->size_ab, sc = (int)s->size_c;
    const int cab = (sab - 1) / 2, cc = (sc - 1) / 2;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (uint32_t ind = 0; ind < npix; ind++) {
        for (uint32_t i = 0; i < npix; i++) {
            const double val = (double)intensity[ind] * (double)intensity[i];
            if (val == 0.0) continue;
            float q[3];
            for (int d = 0; d < 3; d++) q[d] = kmap[3 * (size_t)ind + d] - kmap[3 * (size_t)i + d];
            if (sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) > s->max_q) continue;
            for (int d = 0; d < 3; d++) q[d] = q[d] * vox;
            const int fs = (int)(roundf(q[ai]) + (float)cab);
            const int ms = (int)(roundf(q[bi]) + (float)cab);
            const int ss = (int)(roundf(q[ci]) + (float)cc);
            if (fs < 0 || ms < 0 || ss < 0 || fs >= sab || ms >= sab || ss >= sc) continue;
#pragma omp atomic
            mesh[(size_t)fs + (size_t)ms * sab + (size_t)ss * sab * sab] += val;
        }
    }
}

void ora_merge_f64(const double *small_cq, const ora_mesh_shape *s, const float *R, double *big)
{
    const int sab = (int)s->size_ab, sc = (int)s->size_c;
    const int cab = (sab - 1) / 2, cc = (sc - 1) / 2;
    for (size_t ind = 0; ind < (size_t)sab * sab * sc; ind++) {
        const double v = small_cq[ind];
        if (v == 0.0) continue;
        const int ss_in = (int)(ind / ((size_t)sab * sab));
        const int ms_in = (int)((ind - (size_t)ss_in * sab * sab) / sab);
        const int fs_in = (int)(ind - (size_t)ss_in * sab * sab - (size_t)ms_in * sab);
        float q_in[3];
        q_in[s->k_a] = (float)(fs_in - cab);
        q_in[s->k_b] = (float)(ms_in - cab);
        q_in[s->k_c] = (float)(ss_in - cc);
        const float q0 = R[0] * q_in[0] + R[1] * q_in[1] + R[2] * q_in[2];
        const float q1 = R[3] * q_in[0] + R[4] * q_in[1] + R[5] * q_in[2];
        const float q2 = R[6] * q_in[0] + R[7] * q_in[1] + R[8] * q_in[2];
        const int fs = (int)(roundf(q0) + (float)cab), ms = (int)(roundf(q1) + (float)cab),
                  ss = (int)(roundf(q2) + (float)cab);
        if (fs < 0 || ms < 0 || ss < 0 || fs >= sab || ms >= sab || ss >= sab) continue;
        big[(size_t)fs + (size_t)ms * sab + (size_t)ss * sab * sab] += v * (double)R[9];
    }
}

/* ------------------------------------------------------------------------- */
/* AC1D */

/* AC1D::Initialize(Det, ArraySize, QZoom), IAC/AC1D.cpp:53-58 */
void ora_ac1d_shape(const float mq[3], uint32_t size, float q_zoom, float *max_q_out,
                    float *dq_out)
{
    float m = sqrtf(mq[0] * mq[0] + mq[1] * mq[1] + mq[2] * mq[2]) / q_zoom;
    *max_q_out = m;
    *dq_out = m / ((float)(size - 2));
}

/* AC1D::Calculate_CQ, IAC/AC1D.cpp:78-90 */
double ora_ac1d_cq_multiplicator(const float *intensity, uint32_t npix)
{
    float mean = mean_value_f(intensity, npix);
    double m = 1e-10;
    for (; 1 > mean * mean * m;) m *= 10;
    m = m * 1e5;
    if (m > 1e18) m = 1e18;
    return m;
}

static inline void hist_add(uint64_t *h, uint32_t size, uint32_t sc, uint64_t v, uint64_t *dropped)
{
    if (sc >= size) {
#pragma omp atomic
        (*dropped)++;
        return;
    }
#pragma omp atomic
    h[sc] += v;
}

/* AutoCorr_CQ_AV, IAC/kernel.cl:656-764 (bins summed over the Map-and-Reduce copies) */
void ora_ac1d_cq(const float *intensity, const float *kmap, uint32_t npix, uint32_t size,
                 float max_q, float dq, int interp, double mult, uint64_t *hist,
                 uint64_t *dropped, int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (uint32_t ind = 0; ind < npix; ind++) {
        const float *k1 = kmap + 3 * (size_t)ind;
        float iind = intensity[ind];
        for (uint32_t i = 0; i < npix; i++) {
            double val = (double)(iind * intensity[i]);
            if (val < (double)1e-37f) continue;
            const float *k2 = kmap + 3 * (size_t)i;
            float q = sqrtf((k1[0] - k2[0]) * (k1[0] - k2[0]) + (k1[1] - k2[1]) * (k1[1] - k2[1]) +
                            (k1[2] - k2[2]) * (k1[2] - k2[2]));
            if (q > max_q) continue;
            q = q / dq;
            if (interp == 0) {
                int64_t vc = (int64_t)(val * mult);
                uint32_t sc = (uint32_t)(floor((double)q + 0.5));
                if (SLIP(10)) sc = (uint32_t)floorf(q);
                hist_add(hist, size, sc, (uint64_t)vc, dropped);
            }
            if (interp == 1) {
                uint32_t sc1 = (uint32_t)(floorf(q));
                uint32_t sc2 = sc1 + 1;
                double sep = (double)(q - floorf(q));
                if (SLIP(9)) sep = 1 - sep;
                int64_t v1 = (int64_t)((val * (1 - sep)) * mult);
                int64_t v2 = (int64_t)((val * sep) * mult);
                hist_add(hist, size, sc1, (uint64_t)v1, dropped);
                hist_add(hist, size, sc2, (uint64_t)v2, dropped);
            }
        }
    }
}

/* Calculate_AC_UW_Mapped CPU branch, IAC/AC1D.cpp:256-293 */
void ora_ac1d_acuw(const float *hl4, uint32_t n, uint32_t size, float max_q, float dqdx,
                   int interp, int self_pair_factor, double *ac, uint64_t *dropped)
{
    for (uint32_t j = 0; j < n; j++) {
        for (uint32_t k = j; k < n; k++) {
            const float *a = hl4 + 4 * (size_t)j;
            const float *b = hl4 + 4 * (size_t)k;
            float q = sqrtf((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                            (a[2] - b[2]) * (a[2] - b[2]));
            if (q > max_q) continue;
            q = q / dqdx;
            float f2 = (j == k && self_pair_factor == 1) ? 1.0f : 2.0f;
            if (SLIP(1) && j == k) f2 = 1.0f;
            if (interp == 0) {
                uint32_t sc = (uint32_t)(floor((double)q + 0.5));
                if (SLIP(2)) sc = (uint32_t)floorf(q);
                if (sc >= size) { (*dropped)++; continue; }
                ac[sc] += (double)(f2 * (a[3] * b[3]));
            } else if (interp == 1) {
                uint32_t sc1 = (uint32_t)(floorf(q));
                uint32_t sc2 = sc1 + 1;
                float sep = q - (float)sc1;
                float v1 = f2 * ((a[3] * b[3]) * (1 - sep));
                float v2 = f2 * ((a[3] * b[3]) * (sep));
                if (SLIP(3)) {
                    double dsep = (double)q - floor((double)q);
                    double d1 = (double)f2 * ((double)a[3] * (double)b[3] * (1 - dsep));
                    double d2 = (double)f2 * ((double)a[3] * (double)b[3] * dsep);
                    if (sc1 < size) ac[sc1] += d1; else (*dropped)++;
                    if (sc2 < size) ac[sc2] += d2; else (*dropped)++;
                    continue;
                }
                if (sc1 < size) ac[sc1] += (double)v1; else (*dropped)++;
                if (sc2 < size) ac[sc2] += (double)v2; else (*dropped)++;
            }
        }
    }
}

/* AC1D::CalcAC, IAC/AC1D.cpp:482-493 */
void ora_ac1d_calc_ac(const double *ac_uw, const double *cq, uint32_t size, double *ac)
{
    for (uint32_t i = 0; i < size; i++) {
        ac[i] = ac_uw[i] / cq[i];
        if (isnan(ac[i])) ac[i] = 0;
    }
}

/* RunIAC::Merge_ACandCQ, IAC/RunIAC.cpp:120-158: IntToFloat(Mesh) (float), then
 * divide by CQ in double; 0 where CQ <= 0 or the quotient is NaN. */
void ora_merge_ac_cq(const uint64_t *ac, const double *cq, size_t n, double *out)
{
    for (size_t i = 0; i < n; i++) {
        double a = (double)ora_int_to_float_u64(ac[i]);
        if (cq[i] <= 0)
            out[i] = 0.0;
        else
            out[i] = a / cq[i];
        if (isnan(out[i])) out[i] = 0;
    }
}

/* ------------------------------------------------------------------------- */
/* Simulator, FP64 */

/* SimulateCrystal, IAC/kernel.cl:768-859, evaluated in double: subpixel
 * positions p + v*v_step + u*u_step (:823-825), k = 2π/λ p/|p| (:828-832),
 * arg = k·r + φ (:834-838), |ψ| summed over subpixels and averaged (:844,856). */
void ora_simulate_fp64(const float *pixmap, uint32_t npix, uint32_t p0, uint32_t np,
                       const double *em, uint32_t nem, const double *phases, uint32_t f0,
                       uint32_t nf, double wavelength, int susa, const double u_step[3],
                       const double v_step[3], int out_kind, double *out, int nthreads)
{
    (void)npix;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    const double kk = (2.0 * 3.14159265358979323846) / wavelength;
    const int nsub = (2 * susa + 1) * (2 * susa + 1);
#pragma omp parallel for collapse(2) schedule(dynamic, 4) num_threads(nt)
    for (uint32_t fi = 0; fi < nf; fi++) {
        for (uint32_t pi = 0; pi < np; pi++) {
            uint32_t p = p0 + pi;
            const double *ph = phases + (size_t)(f0 + fi) * nem;
            double acc = 0;
            for (int v = -susa; v <= susa; v++) {
                for (int u = -susa; u <= susa; u++) {
                    double pos[3];
                    for (int d = 0; d < 3; d++)
                        pos[d] = (double)pixmap[3 * (size_t)p + d] + v * v_step[d] + u * u_step[d];
                    double nrm = kk / sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
                    double k0 = pos[0] * nrm, k1 = pos[1] * nrm, k2 = pos[2] * nrm;
                    double re = 0, im = 0;
                    for (uint32_t e = 0; e < nem; e++) {
                        double arg = k0 * em[3 * (size_t)e] + k1 * em[3 * (size_t)e + 1] +
                                     k2 * em[3 * (size_t)e + 2] + ph[e];
                        re += cos(arg);
                        im += sin(arg);
                    }
                    double a2 = re * re + im * im;
                    acc += out_kind == 0 ? sqrt(a2) : a2;
                }
            }
            out[(size_t)fi * np + pi] = acc / nsub;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Poisson sampling on a counter-based stream */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

void ora_uniform2(uint64_t seed, uint32_t frame, uint32_t pixel, uint32_t stream, double u[2])
{
    uint32_t ctr[4] = {pixel, frame, 0x49414350u /* tag */, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    ora_philox4x32_10(ctr, key, r);
    u[0] = ((double)(r[0] >> 5) * 67108864.0 + (double)(r[1] >> 6)) * (1.0 / 9007199254740992.0);
    u[1] = ((double)(r[2] >> 5) * 67108864.0 + (double)(r[3] >> 6)) * (1.0 / 9007199254740992.0);
}

/* exp(x): x = k ln2 + r with ln2 split in a 32-bit head and a tail (Cody-Waite),
 * |r| <= ln2/2, Taylor series of e^r to degree 13 (truncation < 1e-17). */
double ora_exp(double x)
{
    if (x < -708.0) return 0.0;
    if (x > 709.0) return INFINITY;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double log2e = 1.44269504088896338700e+00;
    double k = floor(x * log2e + 0.5);
    double r = (x - k * ln2_hi) - k * ln2_lo;
    double p = 1.0 / 6227020800.0; /* 1/13! */
    static const double inv_fact[13] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                        1.0 / 362880.0,    1.0 / 40320.0,    1.0 / 5040.0,
                                        1.0 / 720.0,       1.0 / 120.0,      1.0 / 24.0,
                                        1.0 / 6.0,         1.0 / 2.0,        1.0,
                                        1.0};
    for (int i = 0; i < 13; i++) p = p * r + inv_fact[i];
    return ldexp(p, (int)k);
}

/* log(x), x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^23. */
double ora_log(double x)
{
    if (!(x > 0)) return -INFINITY;
    int e;
    double m = frexp(x, &e); /* m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e -= 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double t = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2) t = t * s2 + 1.0 / k;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return (e * ln2_hi + 2.0 * s * t) + e * ln2_lo;
}

/* cos(x) for moderate |x|: x = n π/2 + r (π/2 split in head and tail),
 * |r| <= π/4, Taylor series of cos/sin to degree 16/17. */
double ora_cos(double x)
{
    const double pio2_hi = 1.57079632673412561417e+00;
    const double pio2_lo = 6.07710050650619224932e-11;
    const double twoopi = 6.36619772367581382433e-01;
    double n = floor(x * twoopi + 0.5);
    double r = (x - n * pio2_hi) - n * pio2_lo;
    double r2 = r * r;
    double c = 1.0 / 20922789888000.0;   /* 1/16! */
    double sn = 1.0 / 355687428096000.0; /* 1/17! */
    static const double cf[8] = {-1.0 / 87178291200.0, 1.0 / 479001600.0, -1.0 / 3628800.0,
                                 1.0 / 40320.0,        -1.0 / 720.0,      1.0 / 24.0,
                                 -1.0 / 2.0,           1.0};
    static const double sf[8] = {-1.0 / 1307674368000.0, 1.0 / 6227020800.0, -1.0 / 39916800.0,
                                 1.0 / 362880.0,         -1.0 / 5040.0,      1.0 / 120.0,
                                 -1.0 / 6.0,             1.0};
    for (int i = 0; i < 8; i++) c = c * r2 + cf[i];
    for (int i = 0; i < 8; i++) sn = sn * r2 + sf[i];
    sn = sn * r;
    long q = ((long)n) & 3;
    if (q == 0) return c;
    if (q == 1) return -sn;
    if (q == 2) return -c;
    return sn;
}

/* ArrayOperators::ScalarPoissonSampling, IAC/ArrayOperators.h:25-55: inversion
 * by sequential search for mean < 25, Box-Muller otherwise (clamped at 0,
 * truncated to unsigned).  u[0] is the first Drand() draw, u[1] the second.
 * The search is capped at 4096 steps so that a draw above the accumulated
 * (rounded) CDF cannot loop forever. */
uint32_t ora_poisson_sample(double mean, const double u[2])
{
    double ret = 0;
    if (mean < 25) {
        int i = 0;
        double p, q, r;
        r = ora_exp(-mean);
        p = r;
        q = u[0];
        while (p < q && i < 4096) {
            i++;
            r *= mean / i;
            p += r;
        }
        ret = i;
    } else {
        double a = 1 - u[0];
        double b = 1 - u[1];
        ret = (mean + sqrt(mean) * sqrt(-2. * ora_log(a)) * ora_cos(2. * 3.14159265358979323846 * b));
        if (ret < 0) ret = 0;
    }
    return (uint32_t)ret;
}

/* Simulator::SimulatePart post-processing, IAC/Simulator.cpp:868-888, one frame
 * in place: the integrated intensity in double (the reference sums serially;
 * here in 1024-pixel chunks summed in order, then the chunk sums in order --
 * the same value up to double rounding, and the order the GPU reduction uses);
 * the array scaled by (float)(n_expected / sum) (ParMultiplyScalar(float*,
 * float, int), ArrayOperators.h:94-101); if do_poisson, each value replaced by
 * a Poisson draw (ParPoissonSampling(float*), ArrayOperators.h:203-210) from
 * stream 0; if adu_noise != 0, (2u - 1) * adu_noise added in double from
 * stream 1 (ParAddWhiteNoise(float*, float, ...), ArrayOperators.h:219-226);
 * finally a float multiply by value_per_photon (MultiplyScalar<float>). */
void ora_sim_postprocess(float *I, uint32_t npix, double n_expected, int do_poisson, float adu_noise,
                         float value_per_photon, uint64_t seed, uint32_t frame)
{
    double integ = 0.0;
    for (uint32_t c0 = 0; c0 < npix; c0 += 1024) {
        double part = 0.0;
        for (uint32_t l = c0; l < npix && l < c0 + 1024; l++) part += I[l];
        integ += part;
    }
    float fac = (float)(n_expected / integ);
    for (uint32_t l = 0; l < npix; l++) {
        float v = I[l] * fac;
        double u[2];
        if (do_poisson) {
            ora_uniform2(seed, frame, l, 0, u);
            v = (float)ora_poisson_sample((double)v, u);
        }
        if (adu_noise != 0.0f) {
            ora_uniform2(seed, frame, l, 1, u);
            v = (float)((double)v + (2.0 * u[0] - 1.0) * (double)adu_noise);
        }
        I[l] = v * value_per_photon;
    }
}

/* ------------------------------------------------------------------------- */
/* Detector::LoadAndAverageIntensity(Events, Threshold, PhotonSamplingStep, Lower,
 * Upper, Pixelmask), IAC/Detector.cpp:687-790, over frames already loaded
 * ([nframes][npix]): tmpIntensity *= mask (ParMultiplyElementwise); the three
 * branches (no discretisation: ParAdd with threshold; threshold only: zero below,
 * ParAdd; discretisation: DiscretizeToPhotones, long tally of floor(v + 0.5));
 * MeanIntensity = Sum<float>(tmp) / (float)P, PhotonCount = (int)Sum (third branch
 * only); Intensity = tally (third branch) x (1.0f / (float)N) (ParMultiplyScalar). */
void ora_average_intensity(const float *frames, uint32_t nframes, uint32_t npix, const int *mask,
                           float thr, float step, float *avg, float *mean_intensity, int *photon_count)
{
    float *tmp = (float *)malloc(sizeof(float) * npix);
    long *discr = (long *)calloc(npix, sizeof(long));
    for (uint32_t j = 0; j < npix; j++) avg[j] = 0.0f;
    for (uint32_t f = 0; f < nframes; f++) {
        for (uint32_t j = 0; j < npix; j++) {
            tmp[j] = frames[(size_t)f * npix + j];
            if (mask) tmp[j] = tmp[j] * (float)mask[j];
        }
        if (step <= 0) {
            if (thr > 0)
                for (uint32_t j = 0; j < npix; j++)
                    if (tmp[j] < thr) tmp[j] = 0.0f;
            for (uint32_t j = 0; j < npix; j++)
                if (tmp[j] >= thr) avg[j] += tmp[j];
        } else {
            for (uint32_t j = 0; j < npix; j++) {
                if (tmp[j] >= thr) {
                    tmp[j] = ceilf((tmp[j] - thr) / step) * (float)(tmp[j] >= thr);
                    discr[j] += (long)floor(tmp[j] + 0.5);
                } else {
                    tmp[j] = 0.0f;
                }
            }
        }
        float s = 0.0f;
        for (uint32_t j = 0; j < npix; j++) s += tmp[j];
        if (step > 0 && photon_count) photon_count[f] = ora_cvtt_i32(s);
        if (mean_intensity) mean_intensity[f] = s / ((float)npix);
    }
    if (step > 0)
        for (uint32_t j = 0; j < npix; j++) avg[j] = (float)discr[j];
    float inv = 1.0f / ((float)nframes);
    for (uint32_t j = 0; j < npix; j++) avg[j] *= inv;
    free(tmp);
    free(discr);
}

/* ---------------- pattern pre-processing: PPP.cpp ---------------- */
/* PPP::PhotonFinder_LargestAdjacentPixel, IAC/PPP.cpp:18-95: whole photons and the
 * remaining fraction of every pixel of the panels' pixel range first, then per panel (in
 * list order) one more photon where the fraction reaches the seed threshold and, with the
 * largest fraction among the left / right / upper / lower neighbours inside the panel, the
 * combined threshold.  The upper neighbour is looked at only when i - fs > 0. */
static void ora_lap(float *I, const ora_panel *panels, uint32_t npanels, float adu, float seed, float comb)
{
    size_t full = 0;
    for (uint32_t p = 0; p < npanels; p++) full += (size_t)panels[p].fs * panels[p].ss;
    float *whole = (float *)calloc(full ? full : 1, sizeof(float));
    float *frag = (float *)calloc(full ? full : 1, sizeof(float));
    for (size_t j = 0; j < full; j++) {
        if (I[j] <= 0) continue;
        whole[j] = floorf(I[j] / adu);
        frag[j] = (I[j] / adu) - whole[j];
    }
    for (uint32_t p = 0; p < npanels; p++) {
        const int fs = (int)panels[p].fs, n = (int)(panels[p].fs * panels[p].ss);
        const size_t base = panels[p].first_ind;
        for (int i = 0; i < n; i++) {
            const size_t j = base + (size_t)i;
            if (!(frag[j] >= seed)) {
                I[j] = whole[j];
                continue;
            }
            float lap = 0.0f;
            if (i % fs != 0 && frag[j - 1] > lap) lap = frag[j - 1];
            if ((i + 1) % fs != 0 && frag[j + 1] > lap) lap = frag[j + 1];
            if (i - fs > 0 && frag[j - fs] > lap) lap = frag[j - fs];
            if (i + fs < n && frag[j + fs] > lap) lap = frag[j + fs];
            I[j] = (frag[j] + lap >= comb) ? whole[j] + 1 : whole[j];
        }
    }
    free(whole);
    free(frag);
}

float ora_ppp_lap_frame(float *I, uint32_t npix, const int *mask, const float *offset, const float *gain,
                        const ora_panel *panels, uint32_t npanels, float adu_per_photon, float seed_threshold,
                        float combined_threshold, int gain_only)
{
    if (mask)
        for (uint32_t j = 0; j < npix; j++) I[j] = I[j] * (float)mask[j];
    if (offset && gain) {
        /* IAC/PPP.cpp:781-803: the offset map is negated and added, then the division by
         * the gain where allowed, 0 elsewhere */
        for (uint32_t j = 0; j < npix; j++) I[j] = I[j] + (offset[j] * -1.0f);
        for (uint32_t j = 0; j < npix; j++) {
            const int m = mask ? mask[j] : 1;
            if ((I[j] >= 0 || gain_only) && gain[j] > 0 && m == 1)
                I[j] = I[j] / gain[j];
            else
                I[j] = 0;
        }
    }
    if (!gain_only) ora_lap(I, panels, npanels, adu_per_photon, seed_threshold, combined_threshold);
    float s = 0.0f;
    for (uint32_t j = 0; j < npix; j++) s += I[j];
    return s;
}
