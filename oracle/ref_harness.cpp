// Test-infrastructure driver for the reference's own CPU cloth class.
//
// This file is ours; it is compiled together with the reference's CPU
// translation units (sources left where they lie under the reference tree, see
// oracle/Makefile.ref) into oracle/_ref/ref_cloth.  It never ships in the
// product and never leaves this container: tests/golden/make_golden.py runs it
// to write the golden fixtures, and the oracle is timed against it here; the
// product path never loads or runs it.
//
// It follows the reference's own call sequence for one cloth
// (src/sim/SimObjectBuilder.cpp:40-54 -> cBaraffCloth::Init,
//  src/sim/cloth/BaseCloth.cpp:97-109 SetPos/SetPrePos,
//  src/sim/cloth/BaraffCloth.cpp:90-163 UpdatePos)
// and dumps state into a small tagged binary file:
//   repeated records:  "<name> <dtype> <count>\n" + count raw little-endian values
//   dtype is "f8" (double) or "i4" (int32).
//
// usage: ref_cloth <config.json> <x0.bin|-> <nsteps> <out.bin>
//                  [--dump-steps a,b,c] [--stages] [--time]
//                  [--contacts contacts.bin [--contact-steps a,b,c]]
// contacts.bin (float64 little-endian): n_pt, n_ee, then n_pt records
//   {cloth_is_point, vertex, triangle, depth, normal[3], bary[3]} and n_ee
//   records {cloth_is_edge0, edge0, edge1, depth, bary0, bary1, normal[3]};
//   the same contacts are handed to the cloth before each listed step
//   (every step by default) through cBaseObject::SetPointTriangleCollisionInfo
//   and SetEdgeEdgeCollisionInfo, as cSimScene::PerformCollisionDetection
//   does (src/scenes/SimScene.cpp:174-190).
//                  [--perturb perturb.bin] [--ydrag F] [--xpre xpre.bin]
//                  [--qbend] [--normals]
// --qbend dumps a cQBendingMaterial built on the cloth at its rest positions
//   with unit stiffness (QBendingMaterial.cpp:47-118, the constant bending
//   matrix of the reference's GPU form before its per-edge stiffness): the
//   assembled matrix (qbend_K) and each interior edge's 12x12 element matrix
//   (qbend_ele, in edge order) with the edge ids (qbend_edges).
// --normals dumps the vertex and triangle normals after the last step, as
//   cBaseCloth::Update refreshes them (BaseCloth.cpp:585-590:
//   cBaseObject::UpdateTriangleNormal, UpdateVertexNormalFromTriangleNormal,
//   BaseObject.cpp:119-160).
// --xpre gives the previous positions separately (SetPrePos), so that a step
//   starts from any state of a trajectory (x0.bin is then only SetPos).
// perturb.bin (float64 little-endian): records {step, kind, triangle, v[3]};
//   before that step, kind 0 hands a tPerturb with force v on the triangle to
//   cBaseCloth::ApplyPerturb (BaseCloth.cpp:76-87), kind 1 one with goal
//   position v to cBaraffCloth::ApplyUserPerturbForceOnce
//   (BaraffCloth.cpp:165-178), in file order, as cSimScene::UpdateObjects
//   calls the latter before each object's update (SimScene.cpp:148-169).
//   --ydrag sets cBaseCloth::mYDragForce (the ImGui slider,
//   BaraffCloth.cpp:384), which every step's ApplyYDragForce reads.
#include "sim/cloth/BaraffCloth.h"
#include "sim/collision/CollisionInfo.h"
#include "sim/cloth/BaraffMaterial.h"
#include "sim/cloth/BaseBendingMaterial.h"
#include "sim/Perturb.h"
#include "sim/cloth/QBendingMaterial.h"
#include "geometries/Primitives.h"
#include "utils/JsonUtil.h"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>
#include <unistd.h>
#include <fcntl.h>

namespace
{
// Exposes protected state of the reference class for dumping only.
class cProbeCloth : public cBaraffCloth
{
public:
    explicit cProbeCloth(int id) : cBaraffCloth(id) {}
    const tSparseMatd &K() const { return mStiffnessMatrix; }
    const tVectorXd &IntForce() const { return mIntForce; }
    const tVectorXd &GravityForce() const { return mGravityForce; }
    const tVectorXd &MassDiag() const { return mMassMatrixDiag; }
    const tVectorXd &Dx() const { return mDx_buf; }
    const tVectorXd &Xpre() const { return mXpre; }
    cBaraffMaterialPtr Stretch() const { return mMaterial; }
    cBaseBendingMaterialPtr Bending() const { return mBendingMaterial; }
    const tEigenArr<tVector4i> &Stencils() const { return mEdgeAffectVertexId; }
    void SetYDragForce(float f) { mYDragForce = f; }
    // the normal refresh of cBaseCloth::Update (BaseCloth.cpp:585-590)
    void RefreshNormals()
    {
        UpdateTriangleNormal();
        UpdateVertexNormalFromTriangleNormal();
    }
};

// A user perturbation whose force and goal are given directly (the scene
// derives them from a mouse ray, tPerturb::UpdatePerturbPos, Perturb.cpp).
struct tProbePerturb : public tPerturb
{
    void Set(const tVector &force, const tVector &goal)
    {
        mPerturbForce = force;
        mGoalPos = goal;
    }
};

struct PerturbRecord
{
    int step, kind, tri;
    double v[3];
};

FILE *gOut = nullptr;

void WriteF8(const std::string &name, const double *p, size_t n)
{
    fprintf(gOut, "%s f8 %zu\n", name.c_str(), n);
    fwrite(p, sizeof(double), n, gOut);
}
void WriteI4(const std::string &name, const int *p, size_t n)
{
    fprintf(gOut, "%s i4 %zu\n", name.c_str(), n);
    fwrite(p, sizeof(int), n, gOut);
}
void WriteVec(const std::string &name, const tVectorXd &v)
{
    WriteF8(name, v.data(), v.size());
}
// Sparse matrix as COO over its stored entries (row-major CSR order).
void WriteSparse(const std::string &name, const tSparseMatd &m)
{
    std::vector<int> rows, cols;
    std::vector<double> vals;
    for (int k = 0; k < m.outerSize(); ++k)
        for (tSparseMatd::InnerIterator it(m, k); it; ++it)
        {
            rows.push_back(it.row());
            cols.push_back(it.col());
            vals.push_back(it.value());
        }
    WriteI4(name + "_row", rows.data(), rows.size());
    WriteI4(name + "_col", cols.data(), cols.size());
    WriteF8(name + "_val", vals.data(), vals.size());
}

std::vector<int> ParseList(const char *s)
{
    std::vector<int> out;
    std::string cur;
    for (const char *c = s;; ++c)
    {
        if (*c == ',' || *c == 0)
        {
            if (!cur.empty())
                out.push_back(atoi(cur.c_str()));
            cur.clear();
            if (*c == 0)
                break;
        }
        else
            cur.push_back(*c);
    }
    return out;
}
} // namespace

int main(int argc, char **argv)
{
    if (argc < 5)
    {
        fprintf(stderr, "usage: %s config.json x0.bin|- nsteps out.bin "
                        "[--dump-steps a,b] [--stages] [--time]\n",
                argv[0]);
        return 2;
    }
    std::string conf_path = argv[1], x0_path = argv[2], out_path = argv[4];
    int nsteps = atoi(argv[3]);
    std::set<int> dump_steps, contact_steps;
    bool stages = false, timing = false, qbend = false, normals = false;
    std::string contacts_path, perturb_path, xpre_path;
    double ydrag = 0.0;
    bool has_ydrag = false;
    for (int i = 5; i < argc; i++)
    {
        if (!strcmp(argv[i], "--dump-steps") && i + 1 < argc)
        {
            for (int s : ParseList(argv[++i]))
                dump_steps.insert(s);
        }
        else if (!strcmp(argv[i], "--stages"))
            stages = true;
        else if (!strcmp(argv[i], "--time"))
            timing = true;
        else if (!strcmp(argv[i], "--contacts") && i + 1 < argc)
            contacts_path = argv[++i];
        else if (!strcmp(argv[i], "--qbend"))
            qbend = true;
        else if (!strcmp(argv[i], "--normals"))
            normals = true;
        else if (!strcmp(argv[i], "--xpre") && i + 1 < argc)
            xpre_path = argv[++i];
        else if (!strcmp(argv[i], "--perturb") && i + 1 < argc)
            perturb_path = argv[++i];
        else if (!strcmp(argv[i], "--ydrag") && i + 1 < argc)
        {
            ydrag = atof(argv[++i]);
            has_ydrag = true;
        }
        else if (!strcmp(argv[i], "--contact-steps") && i + 1 < argc)
        {
            for (int s : ParseList(argv[++i]))
                contact_steps.insert(s);
        }
    }
    std::vector<tPointTriangleCollisionInfoPtr> pt_contacts;
    std::vector<tEdgeEdgeCollisionInfoPtr> ee_contacts;
    if (!contacts_path.empty())
    {
        FILE *f = fopen(contacts_path.c_str(), "rb");
        if (!f)
            return 7;
        std::vector<double> buf;
        double v;
        while (fread(&v, sizeof(double), 1, f) == 1)
            buf.push_back(v);
        fclose(f);
        if (buf.size() < 2)
            return 8;
        const int n_pt = (int)buf[0], n_ee = (int)buf[1];
        if (buf.size() != size_t(2 + 10 * n_pt + 9 * n_ee))
            return 8;
        const double *r = buf.data() + 2;
        for (int k = 0; k < n_pt; k++, r += 10)
        {
            auto info = std::make_shared<tPointTriangleCollisionInfo>();
            info->obj0_is_cloth_obj1_is_rb = r[0] != 0;
            info->mVertexId0 = (int)r[1];
            info->mTriangleId1 = (int)r[2];
            info->mDepth = (float)r[3];
            info->mOuterNormal = tVector3d(r[4], r[5], r[6]);
            info->mBary = tVector3d(r[7], r[8], r[9]);
            pt_contacts.push_back(info);
        }
        for (int k = 0; k < n_ee; k++, r += 9)
        {
            auto info = std::make_shared<tEdgeEdgeCollisionInfo>();
            info->obj0_is_cloth_obj1_is_rb = r[0] != 0;
            info->mEdgeId0 = (int)r[1];
            info->mEdgeId1 = (int)r[2];
            info->mDepth = (float)r[3];
            info->mBary0 = (float)r[4];
            info->mBary1 = (float)r[5];
            info->mOuterNormal = tVector3d(r[6], r[7], r[8]);
            ee_contacts.push_back(info);
        }
    }
    std::vector<PerturbRecord> perturbs;
    if (!perturb_path.empty())
    {
        FILE *f = fopen(perturb_path.c_str(), "rb");
        if (!f)
            return 9;
        double r[6];
        while (fread(r, sizeof(double), 6, f) == 6)
            perturbs.push_back({(int)r[0], (int)r[1], (int)r[2], {r[3], r[4], r[5]}});
        fclose(f);
    }
    dump_steps.insert(nsteps);

    Json::Value conf;
    if (!cJsonUtil::LoadJson(conf_path, conf))
        return 3;

    // the reference prints on every step; keep stdout quiet
    fflush(stdout);
    int devnull = open("/dev/null", O_WRONLY);
    int saved_stdout = dup(1);
    dup2(devnull, 1);

    auto cloth = std::make_shared<cProbeCloth>(0);
    cloth->Init(conf);
    if (has_ydrag)
        cloth->SetYDragForce((float)ydrag);
    int nv = cloth->GetNumOfVertices();
    tVectorXd x0 = cloth->GetPos();
    if (x0_path != "-")
    {
        FILE *f = fopen(x0_path.c_str(), "rb");
        if (!f)
            return 4;
        size_t got = fread(x0.data(), sizeof(double), 3 * nv, f);
        fclose(f);
        if (got != size_t(3 * nv))
            return 5;
        cloth->SetPos(x0);
        cloth->SetPrePos(x0);
    }
    if (!xpre_path.empty())
    {
        tVectorXd xpre = x0;
        FILE *f = fopen(xpre_path.c_str(), "rb");
        if (!f)
            return 4;
        size_t got = fread(xpre.data(), sizeof(double), 3 * nv, f);
        fclose(f);
        if (got != size_t(3 * nv))
            return 5;
        cloth->SetPrePos(xpre);
    }

    gOut = fopen(out_path.c_str(), "wb");
    if (!gOut)
        return 6;

    // topology and rest data (reference numbering)
    {
        std::vector<int> tri, edge, stencil;
        for (auto &t : cloth->GetTriangleArray())
        {
            tri.push_back(t->mId0);
            tri.push_back(t->mId1);
            tri.push_back(t->mId2);
        }
        for (auto &e : cloth->GetEdgeArray())
        {
            edge.push_back(e->mId0);
            edge.push_back(e->mId1);
            edge.push_back(e->mTriangleId0);
            edge.push_back(e->mTriangleId1);
            edge.push_back(e->mIsBoundary ? 1 : 0);
        }
        for (auto &s : cloth->Stencils())
            for (int k = 0; k < 4; k++)
                stencil.push_back(s[k]);
        WriteI4("triangles", tri.data(), tri.size());
        WriteI4("edges", edge.data(), edge.size());
        WriteI4("stencils", stencil.data(), stencil.size());
        const std::vector<int> &pins = cloth->GetConstraintStaticVertices();
        WriteI4("pins", pins.data(), pins.size());
        std::vector<double> uv, rest;
        for (auto &v : cloth->GetVertexArray())
        {
            uv.push_back(v->muv[0]);
            uv.push_back(v->muv[1]);
        }
        WriteF8("uv", uv.data(), uv.size());
        WriteVec("x_init", cloth->GetInitPos());
        WriteVec("x0", x0);
        WriteVec("mass_diag", cloth->MassDiag());
        WriteVec("gravity", cloth->GravityForce());
    }

    if (qbend)
    {
        // at the initial positions, where the GPU form builds its matrix
        // (cBaraffClothGpu::Init, before any SetPos)
        tVectorXd xnow = cloth->GetPos();
        cloth->SetPos(cloth->GetInitPos());
        auto qb = std::make_shared<cQBendingMaterial>();
        qb->Init(cloth->GetVertexArray(), cloth->GetEdgeArray(), cloth->GetTriangleArray(),
                 tVector3d::Ones());
        WriteSparse("qbend_K", qb->GetStiffnessMatrix());
        std::vector<double> ele;
        for (const tMatrix12f &k : qb->GetEleStiffnessMatrixLst())
            for (int r = 0; r < 12; r++)
                for (int c = 0; c < 12; c++)
                    ele.push_back((double)k(r, c));
        WriteF8("qbend_ele", ele.data(), ele.size());
        std::vector<int> ids;
        const auto &edges = cloth->GetEdgeArray();
        for (int e = 0; e < (int)edges.size(); e++)
            if (!edges[e]->mIsBoundary)
                ids.push_back(e);
        WriteI4("qbend_edges", ids.data(), ids.size());
        cloth->SetPos(xnow);
    }

    std::vector<double> step_seconds;
    for (int step = 1; step <= nsteps; step++)
    {
        if (!contacts_path.empty() && (contact_steps.empty() || contact_steps.count(step)))
        {
            cloth->SetPointTriangleCollisionInfo(pt_contacts);
            cloth->SetEdgeEdgeCollisionInfo(ee_contacts);
        }
        for (const PerturbRecord &p : perturbs)
        {
            if (p.step != step)
                continue;
            tProbePerturb pert;
            pert.mObject = cloth;
            pert.mAffectedTriId = p.tri;
            tVector v(p.v[0], p.v[1], p.v[2], 0.0);
            if (p.kind == 0)
            {
                pert.Set(v, tVector::Zero());
                cloth->ApplyPerturb(&pert);
            }
            else
            {
                pert.Set(tVector::Zero(), v);
                cloth->ApplyUserPerturbForceOnce(&pert);
            }
        }
        tVectorXd xbefore = cloth->GetPos();
        auto t0 = std::chrono::steady_clock::now();
        cloth->UpdatePos(cloth->GetDefaultTimestep());
        auto t1 = std::chrono::steady_clock::now();
        step_seconds.push_back(std::chrono::duration<double>(t1 - t0).count());
        if (stages && step == 1)
        {
            // state of the first step: K and f were formed at xbefore
            WriteVec("s1_x", xbefore);
            WriteSparse("s1_K", cloth->K());
            WriteVec("s1_fint", cloth->IntForce());
            WriteVec("s1_dx", cloth->Dx());
            WriteVec("s1_f_stretch", cloth->Stretch()->CalcTotalForce());
            WriteSparse("s1_K_stretch",
                        cloth->Stretch()->CalcTotalStiffnessMatrix());
            WriteSparse("s1_K_bend", cloth->Bending()->GetStiffnessMatrix());
            WriteVec("s1_f_bend", cloth->Bending()->CalcForce(xbefore));
        }
        if (dump_steps.count(step))
            WriteVec("x_step" + std::to_string(step), cloth->GetPos());
    }
    if (timing)
        WriteF8("step_seconds", step_seconds.data(), step_seconds.size());
    if (normals)
    {
        cloth->RefreshNormals();
        std::vector<double> vn, tn;
        for (auto &v : cloth->GetVertexArray())
            for (int a = 0; a < 3; a++)
                vn.push_back(v->mNormal[a]);
        for (auto &t : cloth->GetTriangleArray())
            for (int a = 0; a < 3; a++)
                tn.push_back(t->mNormal[a]);
        WriteF8("vertex_normals", vn.data(), vn.size());
        WriteF8("triangle_normals", tn.data(), tn.size());
    }
    fclose(gOut);
    fflush(stdout);
    dup2(saved_stdout, 1);
    return 0;
}
