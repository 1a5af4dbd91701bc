// acmesh.cpp -- ACMesh (IAC/ACMesh.h, IAC/ACMesh.cpp:74-206): mesh shapes from the
// detector geometry, device-resident uint64 / double meshes, download.
#include <stdexcept>

#include "host_util.hpp"
#include "iac_host.hpp"

using iac_host::check;

ACMesh::ACMesh(Settings *PrgSettings) : Options(PrgSettings) {}

ACMesh::~ACMesh()
{
    reset_device();
    delete[] Mesh;
    delete[] CQMesh;
}

void ACMesh::drop_queue()
{
    if (queue_) {
        iac_ac3d_queue *q = queue_;
        queue_ = nullptr;
        queue_kmap_ = nullptr;
        queue_flags_ = -1;
        check(iac_ac3d_queue_destroy(q), "iac_ac3d_queue_destroy");
    }
}

iac_ac3d_queue *ACMesh::Queue(const float *d_kmap, uint32_t npix, int interp1, int interp2, bool double_map)
{
    const int flags = interp1 + 2 * interp2 + 4 * (double_map ? 1 : 0);
    if (queue_ && (queue_kmap_ != d_kmap || queue_flags_ != flags)) drop_queue();
    if (!queue_) {
        iac_mesh_shape s = CShape();
        check(iac_ac3d_queue_create(ctx, d_kmap, npix, &s, interp1, interp2, double_map ? 1 : 0, DeviceMesh(), 0, 0,
                                    &queue_),
              "iac_ac3d_queue_create in ACMesh::Queue()");
        queue_kmap_ = d_kmap;
        queue_flags_ = flags;
    }
    return queue_;
}

void ACMesh::Flush()
{
    if (queue_) check(iac_ac3d_queue_flush(queue_), "iac_ac3d_queue_flush in ACMesh::Flush()");
}

void ACMesh::reset_device()
{
    drop_queue();
    if (d_Mesh) iac_dev_free(ctx, d_Mesh);
    if (d_CQMesh) iac_dev_free(ctx, d_CQMesh);
    d_Mesh = nullptr;
    d_CQMesh = nullptr;
}

void ACMesh::take_shape(const iac_mesh_shape &s)
{
    reset_device();
    delete[] Mesh;
    delete[] CQMesh;
    Mesh = nullptr;
    CQMesh = nullptr;
    Shape.k_A = s.k_a;
    Shape.k_B = s.k_b;
    Shape.k_C = s.k_c;
    Shape.Size_AB = s.size_ab;
    Shape.Size_C = s.size_c;
    for (int i = 0; i < 3; i++) Shape.Center[i] = s.center[i];
    Shape.dq_per_Voxel = s.dq_per_voxel;
    Shape.Voxel_per_dq = s.voxel_per_dq;
    Shape.Max_Q = s.max_q;
}

iac_mesh_shape ACMesh::CShape() const
{
    iac_mesh_shape s;
    s.k_a = Shape.k_A;
    s.k_b = Shape.k_B;
    s.k_c = Shape.k_C;
    s.size_ab = Shape.Size_AB;
    s.size_c = Shape.Size_C;
    for (int i = 0; i < 3; i++) s.center[i] = Shape.Center[i];
    s.dq_per_voxel = Shape.dq_per_Voxel;
    s.voxel_per_dq = Shape.Voxel_per_dq;
    s.max_q = Shape.Max_Q;
    return s;
}

// IAC/ACMesh.cpp:74-106
void ACMesh::CreateBigMeshForDetector(Detector &Det, int EdgeSize, float q_Zoom)
{
    iac_mesh_shape s;
    check(iac_mesh_shape_big(Det.Max_q, EdgeSize, q_Zoom, &s), "ACMesh::CreateBigMeshForDetector()");
    take_shape(s);
    Checklist.BigMesh = true;
}

// IAC/ACMesh.cpp:107-139
void ACMesh::CreateBigMesh_CofQ_ForDetector(Detector &Det, int EdgeSize, float q_Zoom)
{
    iac_mesh_shape s;
    check(iac_mesh_shape_big(Det.Max_q, EdgeSize, q_Zoom, &s), "ACMesh::CreateBigMesh_CofQ_ForDetector()");
    take_shape(s);
    Checklist.CofQBigMesh = true;
}

// IAC/ACMesh.cpp:140-206
void ACMesh::CreateSmallMesh_CofQ_ForDetector(Detector &Det, int PerpSize, float q_Zoom)
{
    iac_mesh_shape s;
    check(iac_mesh_shape_small_cofq(Det.Max_q, PerpSize, q_Zoom, &s), "ACMesh::CreateSmallMesh_CofQ_ForDetector()");
    take_shape(s);
    Checklist.CofQSmallMesh = true;
}

uint64_t *ACMesh::DeviceMesh()
{
    if (!ctx) {
        if (!Options) throw std::runtime_error("ACMesh: no Settings to place the mesh on a device");
        ctx = Options->Context(0);
    }
    if (!d_Mesh) {
        check(iac_dev_alloc(ctx, NumVoxels() * 8, (void **)&d_Mesh), "ACMesh mesh allocation");
        check(iac_memset_zero(ctx, d_Mesh, NumVoxels() * 8), "ACMesh mesh zero");
    }
    return d_Mesh;
}

double *ACMesh::DeviceCQMesh()
{
    if (!ctx) {
        if (!Options) throw std::runtime_error("ACMesh: no Settings to place the mesh on a device");
        ctx = Options->Context(0);
    }
    if (!d_CQMesh) {
        check(iac_dev_alloc(ctx, NumVoxels() * 8, (void **)&d_CQMesh), "ACMesh C(q) allocation");
        check(iac_memset_zero(ctx, d_CQMesh, NumVoxels() * 8), "ACMesh C(q) zero");
    }
    return d_CQMesh;
}

void ACMesh::Download()
{
    Flush();
    const size_t n = NumVoxels();
    if (d_Mesh) {
        if (!Mesh) Mesh = new unsigned long[n];
        check(iac_memcpy_d2h(ctx, Mesh, d_Mesh, n * 8), "ACMesh download");
    }
    if (d_CQMesh) {
        if (!CQMesh) CQMesh = new double[n];
        check(iac_memcpy_d2h(ctx, CQMesh, d_CQMesh, n * 8), "ACMesh C(q) download");
    }
}
