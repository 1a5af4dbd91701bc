/*
 * iac_oracle.h -- CPU oracle for the pixel-pair autocorrelation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (FabianTr/IncohAutoCorrelate, CPU path and kernel.cl) that the
 * tests, the smoke check and the benchmark's CPU-baseline leg use as the
 * checker.  The product (incohautocorrelate_amd/, include/iac_hip.h) never
 * links, imports or calls anything in this directory.
 *
 * Arithmetic follows the reference's x86-64 build: IEEE single for float
 * expressions, double where the reference promotes, no FMA contraction
 * (built with -ffp-contract=off, like IncohAutoCorrelate.vcxproj:115 has no
 * -march and therefore no FMA), correctly rounded sqrtf and division.
 *
 * Parity pins (see tests/): brute force on tiny inputs, closed forms,
 * invariants (point symmetry, mass), exhaustive fixed-point tables, and
 * rotation special cases.  Functions whose output bits depend on a vendor
 * math library (the FP32 SimulateCrystal) are pinned to FP64 instead.
 */
#ifndef IAC_ORACLE_H
#define IAC_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ACMesh::MeshShape, IncohAutoCorrelate/ACMesh.h:21-29 */
typedef struct {
    uint32_t k_a, k_b, k_c;
    uint32_t size_ab, size_c;
    uint32_t center[3];
    float dq_per_voxel;
    uint32_t voxel_per_dq;
    float max_q;
} ora_mesh_shape;

/* ---------------- fixed point: Settings.h:168-187 ---------------- */
int64_t ora_float_to_int_f(float f);
int64_t ora_float_to_int_d(double f);
float ora_int_to_float_u64(uint64_t i);
/* val(p) of one NN pair update, ACMesh.cpp:331-335 (long -> unsigned int) */
uint32_t ora_pair_val(float p);

/* ---------------- geometry ---------------- */
/* Simulator::GeneratePixelMap, Simulator.cpp:495-519 (layout [SizeA][SizeB][3]) */
void ora_generate_pixel_map(int size_a, int size_b, float pixel_size, const float center[3],
                            const float vec_a[3], const float vec_b[3], float *pixmap);
/* Detector::Calc_kMap, Detector.cpp:310-361 */
void ora_calc_kmap(const float *pixmap, uint32_t npix, float *kmap,
                   float max_k[3], float min_k[3], float max_q[3]);
/* ACMesh::CreateBigMeshForDetector / CreateBigMesh_CofQ_ForDetector, ACMesh.cpp:74-139 */
int ora_mesh_shape_big(const float max_q[3], int edge_size, float q_zoom, ora_mesh_shape *s);
/* ACMesh::CreateSmallMesh_CofQ_ForDetector, ACMesh.cpp:140-206 */
int ora_mesh_shape_small_cofq(const float max_q[3], int perp_size, float q_zoom, ora_mesh_shape *s);

/* ---------------- sparse hit list ---------------- */
/* Detector::CreateSparseHitList(T[,S]) Detector.cpp:926-958 with the mask multiply
 * ArrayOperators.h:120-127 and DiscretizeToPhotons ArrayOperators.h:138-148.
 * mask may be NULL (flat ones).  step <= 0 means no discretisation.
 * out4 receives {kx,ky,kz,value} per entry; returns the number of entries. */
uint32_t ora_sparse_hit_list(const float *intensity, const int *mask, const float *kmap,
                             uint32_t npix, float threshold, float step, float *out4);
/* hit list from a (pixel, count) photon list: the same list CreateSparseHitList
 * builds for a frame whose only nonzero pixels carry integer counts (T=0, S=1). */
void ora_hit_list_from_photons(const float *kmap, const uint32_t *pix, const uint16_t *cnt,
                               uint32_t n, float *out4);

/* ---------------- 3D pair accumulation, CPU path ---------------- */
/* Detector::AutoCorrelateSparseList CPU branch, Detector.cpp:1010-1051, binning by
 * ACMesh::Atomic_Add_q_Entry / Atomic_Add_rotated_q_Entry ACMesh.cpp:209-385.
 * interp: 0 nearest neighbour, 1 linear.  Updates falling outside the mesh (the
 * reference writes them out of bounds) are dropped and counted in *dropped.
 * nthreads <= 0 uses the OpenMP default. */
void ora_ac3d_sparse(const float *hl4, uint32_t n, const float rot[9], const ora_mesh_shape *s,
                     int interp1, int interp2, int double_map, uint64_t *mesh,
                     uint64_t *dropped, int nthreads);
/* The RunIAC frame loop (RunIAC.cpp:524-550) over photon-list frames:
 * frame f holds entries [off[f], off[f+1]); rots is 9 floats per frame or NULL
 * for the identity. */
void ora_ac3d_sparse_frames(const float *kmap, const uint32_t *pix, const uint16_t *cnt,
                            const uint64_t *off, uint32_t nframes, const float *rots,
                            const ora_mesh_shape *s, int interp1, int interp2, int double_map,
                            uint64_t *mesh, uint64_t *dropped, int nthreads);
/* The same frame loop, the serial plain-loop form: ordered pairs enumerated
 * one by one without the i<=j mirror.  Used as a brute-force cross-check. */
void ora_ac3d_bruteforce(const float *hl4, uint32_t n, const float rot[9], const ora_mesh_shape *s,
                         int interp1, int interp2, int double_map, uint64_t *mesh, uint64_t *dropped);

/* ---------------- C(q) small: kernel.cl:27-184 + Detector.cpp:1184-1315 ---------------- */
double ora_cq_small_multiplicator(const float *intensity, uint32_t npix);
void ora_cq_small(const float *intensity, const float *kmap, uint32_t npix, const ora_mesh_shape *s,
                  int interp, double mult, uint64_t *mesh, uint64_t *dropped, int nthreads);

/* ---------------- merge: kernel.cl:189-344 + Detector.cpp:1317-1534 ---------------- */
/* Builds Rot_and_Weight (10 floats per event, weights quantised) and the
 * Multiplicator from the events' rotations and mean intensities and the small
 * C(q) mesh, Detector.cpp:1339-1427. */
double ora_merge_prepare(const float *rot9, const float *mean_intensity, uint32_t nev,
                         const double *small_cq, size_t nsmall, float *rw10);
void ora_merge_cq(const double *small_cq, const ora_mesh_shape *small_shape, const float *rw10,
                  uint32_t nev, int interp, double mult, uint64_t *big, uint64_t *dropped,
                  int nthreads);

/* FP64 oracle of the dense float path: the C(q) small binning of every ordered
 * pixel pair (nearest neighbour, AutoCorr_CQ_small index rule) with the value
 * (double)I_a * (double)I_b accumulated in double -- no fixed point. */
void ora_cq_small_f64(const float *intensity, const float *kmap, uint32_t npix,
                      const ora_mesh_shape *s, double *mesh, int nthreads);
/* FP64 merge: small-mesh voxel v rotated by R (nearest neighbour, Merge_CQ index
 * rule) adds w * small[v] to the cubic big mesh in double. */
void ora_merge_f64(const double *small_cq, const ora_mesh_shape *small_shape, const float *rw10,
                   double *big);

/* ---------------- AC1D: AC1D.cpp ---------------- */
/* AC1D::Initialize(Det, Size, QZoom), AC1D.cpp:53-58 */
void ora_ac1d_shape(const float max_q[3], uint32_t size, float q_zoom, float *max_q_out,
                    float *dq_out);
/* AC1D::Calculate_CQ multiplicator, AC1D.cpp:78-90 */
double ora_ac1d_cq_multiplicator(const float *intensity, uint32_t npix);
/* AutoCorr_CQ_AV kernel.cl:656-764: per-bin uint64 totals over all copies */
void ora_ac1d_cq(const float *intensity, const float *kmap, uint32_t npix, uint32_t size,
                 float max_q, float dq, int interp, double mult, uint64_t *hist,
                 uint64_t *dropped, int nthreads);
/* Calculate_AC_UW_Mapped CPU branch AC1D.cpp:256-293 (self pairs weighted 2,
 * like the reference; self_pair_factor=1 counts them once). */
void ora_ac1d_acuw(const float *hl4, uint32_t n, uint32_t size, float max_q, float dqdx,
                   int interp, int self_pair_factor, double *ac, uint64_t *dropped);
/* AC1D::CalcAC AC1D.cpp:482-493 */
void ora_ac1d_calc_ac(const double *ac_uw, const double *cq, uint32_t size, double *ac);

/* ---------------- normalisation: RunIAC.cpp:120-158 ---------------- */
void ora_merge_ac_cq(const uint64_t *ac, const double *cq, size_t n, double *out);

/* ---------------- simulator (FP64): kernel.cl:768-859 ---------------- */
/* ψ(p,f) = Σ_e exp(i(k_p·r_e + φ_{e,f})), k_p = 2π/λ · p/|p| evaluated in double.
 * emitters: 3 doubles per emitter; phases: [nframes][nem] doubles.
 * out_kind 0 = |ψ| (reference SimulateCrystal, kernel.cl:844), 1 = |ψ|².
 * Pixels [p0, p0+np) and frames [f0, f0+nf) only; out is [nf][np]. */
void ora_simulate_fp64(const float *pixmap, uint32_t npix, uint32_t p0, uint32_t np,
                       const double *emitters, uint32_t nem, const double *phases,
                       uint32_t f0, uint32_t nf, double wavelength, int subsampling,
                       const double u_step[3], const double v_step[3], int out_kind,
                       double *out, int nthreads);

/* ---------------- Poisson: ArrayOperators.h:25-55 on a Philox4x32-10 stream ---------------- */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Uniform pair for (seed, frame, pixel, stream): two 53-bit doubles in [0,1). */
void ora_uniform2(uint64_t seed, uint32_t frame, uint32_t pixel, uint32_t stream, double u[2]);
/* Own exp / log / cos (range reduction + polynomial) used by the sampler. */
double ora_exp(double x);
double ora_log(double x);
double ora_cos(double x);
uint32_t ora_poisson_sample(double mean, const double u[2]);
/* Rescale to n_expected photons (Simulator.cpp:870-878), Poisson sample
 * (ArrayOperators.h:203-210), white noise, value per photon; one frame in place. */
void ora_sim_postprocess(float *intensity, uint32_t npix, double n_expected, int do_poisson, float adu_noise,
                         float value_per_photon, uint64_t seed, uint32_t frame);

/* ---------------- averaged intensity: Detector.cpp:687-790 ---------------- */
void ora_average_intensity(const float *frames, uint32_t nframes, uint32_t npix, const int *mask,
                           float thr, float step, float *avg, float *mean_intensity, int *photon_count);

/* (int)x as x86-64 cvttss2si converts it (0x80000000 for NaN and out-of-range values) */
int ora_cvtt_i32(float x);

/* ---------------- pattern pre-processing: PPP.cpp ---------------- */
/* PPP::DetectorPanel, IAC/PPP.h:14-18 (Scans = {fs, ss}) */
typedef struct {
    uint32_t first_ind, fs, ss;
} ora_panel;
/* One frame of PPP::ProcessData_PF_LAP (IAC/PPP.cpp:469-495) in place, in the reference's
 * order: Detector::ApplyPixelMask when mask is given, PPP::GainCorrection (IAC/PPP.cpp:
 * 672-808, AllowNegativeValues = gain_only) when offset and gain are given, then unless
 * gain_only PPP::PhotonFinder_LargestAdjacentPixel (IAC/PPP.cpp:18-95) over the panels in
 * list order.  Returns the serial float sum of the frame (ArrayOperators::Sum). */
float ora_ppp_lap_frame(float *intensity, uint32_t npix, const int *mask, const float *offset, const float *gain,
                        const ora_panel *panels, uint32_t npanels, float adu_per_photon, float seed_threshold,
                        float combined_threshold, int gain_only);

#ifdef __cplusplus
}
#endif
#endif
