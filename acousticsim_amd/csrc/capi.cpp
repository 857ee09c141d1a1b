// C ABI (include/bw_cloth.h) over the device batch runtime.  Every entry
// point catches exceptions and returns a status; nothing exits the process.
#include "../../include/bw_cloth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch.h"
#include "json_lite.h"
#include "mesh_gen.h"
#include "model.h"
#include "server.h"
#include "status.h"

namespace
{
thread_local std::string g_error;

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

template <typename F> int guard(F &&f)
{
    try
    {
        return f();
    }
    catch (...)
    {
        std::string msg;
        const int code = bw::status_of_current_exception(msg);
        return fail(code, msg);
    }
}

bw::ClothModel model_from_json(const char *json)
{
    if (!json)
        throw std::runtime_error("config: NULL JSON text");
    return bw::ClothModel::from_json_text(json);
}

bool all_finite(const double *x, size_t n)
{
    for (size_t k = 0; k < n; k++)
        if (!std::isfinite(x[k]))
            return false;
    return true;
}
} // namespace

struct bw_cloth
{
    std::unique_ptr<bw::Batch> batch;
    std::vector<double> init_pos;
};

struct bw_batch
{
    std::unique_ptr<bw::Batch> batch;
};

struct bw_body
{
    std::unique_ptr<bw::Body> body;
};

namespace
{
template <typename F> int with_member(bw_batch *b, int i, F &&f)
{
    return guard([&] {
        if (!b || i < 0 || i >= b->batch->size())
            return fail(BW_ERR_ARG, "bad batch or member index");
        f();
        return BW_OK;
    });
}
} // namespace

extern "C" {

const char *bw_last_error(void) { return g_error.c_str(); }

namespace
{
int create_from_model(bw::ClothModel model, const double *x0, bw_cloth **out)
{
    std::vector<bw::ClothModel> ms;
    ms.push_back(std::move(model));
    auto c = std::make_unique<bw_cloth>();
    c->init_pos = ms[0].xrest;
    const size_t n3 = 3 * (size_t)ms[0].nv;
    if (x0 && !all_finite(x0, n3))
        return fail(BW_ERR_ARG, "x0 has NaN/Inf");
    c->batch = std::make_unique<bw::Batch>(std::move(ms));
    if (x0)
    {
        c->batch->set_pos(0, x0);
        c->batch->set_pre_pos(0, x0);
    }
    *out = c.release();
    return BW_OK;
}
} // namespace

int bw_create_mesh(const char *json, const double *rest_pos, int nv, const int *tri, int nt,
                   const double *x0, bw_cloth **out)
{
    return guard([&] {
        if (!out || !json || !rest_pos || !tri)
            return fail(BW_ERR_ARG, "NULL argument");
        *out = nullptr;
        if (nv < 3 || nt < 1)
            return fail(BW_ERR_ARG, "a mesh needs >= 3 vertices and >= 1 triangle");
        if (!all_finite(rest_pos, 3 * (size_t)nv))
            return fail(BW_ERR_ARG, "rest positions hold NaN/Inf");
        bw::Json j = bw::Json::parse(json);
        if (j.kind != bw::Json::Object)
            return fail(BW_ERR_ARG, "config: top level must be an object");
        bw::ClothConfig cfg = bw::ClothConfig::from_json(j);
        if (cfg.geometry_type != "mesh")
            return fail(BW_ERR_ARG, "config: bw_create_mesh needs geometry_type \"mesh\"");
        return create_from_model(bw::ClothModel::build_mesh(cfg, rest_pos, nv, tri, nt), x0, out);
    });
}

int bw_delaunay_mesh(double height, double width, int n, unsigned long long seed,
                     double min_spacing, double *rest_pos, int cap_vertices, int *tri,
                     int cap_triangles, int *nv, int *nt)
{
    return guard([&] {
        if (!nv || !nt)
            return fail(BW_ERR_ARG, "NULL argument");
        if (n < 4 || !(height > 0) || !(width > 0) || !(min_spacing >= 0))
            return fail(BW_ERR_ARG, "a delaunay cloth needs n >= 4, a positive size and "
                                    "min_spacing >= 0");
        std::vector<double> xyz;
        std::vector<std::array<int, 3>> t;
        bw::delaunay_cloth((float)width, (float)height, n, seed, (float)min_spacing, xyz, t);
        *nv = (int)(xyz.size() / 3);
        *nt = (int)t.size();
        if ((rest_pos && cap_vertices < *nv) || (tri && cap_triangles < *nt))
            return fail(BW_ERR_CAPACITY, "delaunay mesh larger than the buffers");
        if (rest_pos)
            std::memcpy(rest_pos, xyz.data(), xyz.size() * sizeof(double));
        if (tri)
            std::memcpy(tri, t.data(), t.size() * sizeof(t[0]));
        return BW_OK;
    });
}

int bw_create(const char *json, const double *x0, bw_cloth **out)
{
    return guard([&] {
        if (!out)
            return fail(BW_ERR_ARG, "out is NULL");
        *out = nullptr;
        std::vector<bw::ClothModel> ms;
        ms.push_back(model_from_json(json));
        auto c = std::make_unique<bw_cloth>();
        c->init_pos = ms[0].xrest;
        const size_t n3 = 3 * (size_t)ms[0].nv;
        c->batch = std::make_unique<bw::Batch>(std::move(ms));
        if (x0)
        {
            if (!all_finite(x0, n3))
                return fail(BW_ERR_ARG, "x0 has NaN/Inf");
            c->batch->set_pos(0, x0);
            c->batch->set_pre_pos(0, x0);
        }
        *out = c.release();
        return BW_OK;
    });
}

void bw_destroy(bw_cloth *c) { delete c; }

int bw_update_pos(bw_cloth *c, double)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->step(1);
        return BW_OK;
    });
}

int bw_reset(bw_cloth *c)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->set_pos(0, c->init_pos.data());
        c->batch->set_pre_pos(0, c->init_pos.data());
        c->batch->clear_warm_start(0);
        c->batch->clear_user_force(0);
        return BW_OK;
    });
}

int bw_apply_perturb(bw_cloth *c, int triangle, const double force[3])
{
    return guard([&] {
        if (!c || !force)
            return fail(BW_ERR_ARG, "NULL argument");
        c->batch->apply_perturb(0, triangle, force);
        return BW_OK;
    });
}

int bw_apply_user_perturb_force_once(bw_cloth *c, int triangle, const double goal[3])
{
    return guard([&] {
        if (!c || !goal)
            return fail(BW_ERR_ARG, "NULL argument");
        c->batch->apply_user_perturb_force_once(0, triangle, goal);
        return BW_OK;
    });
}

int bw_set_y_drag_force(bw_cloth *c, float force)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->set_y_drag_force(0, force);
        return BW_OK;
    });
}

int bw_get_pos(const bw_cloth *c, double *x)
{
    return guard([&] {
        if (!c || !x)
            return fail(BW_ERR_ARG, "NULL argument");
        c->batch->get_pos(0, x);
        return BW_OK;
    });
}

int bw_get_pre_pos(const bw_cloth *c, double *x)
{
    return guard([&] {
        if (!c || !x)
            return fail(BW_ERR_ARG, "NULL argument");
        c->batch->get_pre_pos(0, x);
        return BW_OK;
    });
}

int bw_get_velocity(const bw_cloth *c, double *v)
{
    return guard([&] {
        if (!c || !v)
            return fail(BW_ERR_ARG, "NULL argument");
        const size_t n3 = 3 * (size_t)c->batch->model(0).nv;
        std::vector<double> xp(n3);
        c->batch->get_pos(0, v);
        c->batch->get_pre_pos(0, xp.data());
        const double dt = c->batch->model(0).cfg.dt;
        for (size_t k = 0; k < n3; k++)
            v[k] = (v[k] - xp[k]) / dt;
        return BW_OK;
    });
}

int bw_set_pos(bw_cloth *c, const double *x)
{
    return guard([&] {
        if (!c || !x)
            return fail(BW_ERR_ARG, "NULL argument");
        if (!all_finite(x, 3 * (size_t)c->batch->model(0).nv))
            return fail(BW_ERR_ARG, "positions have NaN/Inf");
        c->batch->set_pos(0, x);
        return BW_OK;
    });
}

int bw_set_pre_pos(bw_cloth *c, const double *x)
{
    return guard([&] {
        if (!c || !x)
            return fail(BW_ERR_ARG, "NULL argument");
        if (!all_finite(x, 3 * (size_t)c->batch->model(0).nv))
            return fail(BW_ERR_ARG, "positions have NaN/Inf");
        c->batch->set_pre_pos(0, x);
        return BW_OK;
    });
}

int bw_get_normals(const bw_cloth *c, double *vn, double *tn)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->normals(0, vn, tn);
        return BW_OK;
    });
}

int bw_set_point_triangle_contacts(bw_cloth *c, int n, const int *ids, const double *normal,
                                   const double *bary, const float *depth)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->set_point_triangle_contacts(0, n, ids, normal, bary, depth);
        return BW_OK;
    });
}

int bw_set_edge_edge_contacts(bw_cloth *c, int n, const int *ids, const double *normal,
                              const float *bary, const float *depth)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->set_edge_edge_contacts(0, n, ids, normal, bary, depth);
        return BW_OK;
    });
}

int bw_get_point_triangle_contacts(const bw_cloth *c, int cap, int *ids, double *normal,
                                   double *bary, float *depth)
{
    return guard([&] {
        if (!c || cap < 0)
            return fail(BW_ERR_ARG, "bad arguments");
        const bw::Contacts &ct = c->batch->contacts(0);
        const int n = (int)ct.pt_depth.size(), m = std::min(n, cap);
        if (m > 0 && (!ids || !normal || !bary || !depth))
            return fail(BW_ERR_ARG, "NULL output array");
        std::copy(ct.pt_ids.begin(), ct.pt_ids.begin() + 3 * (size_t)m, ids);
        std::copy(ct.pt_normal.begin(), ct.pt_normal.begin() + 3 * (size_t)m, normal);
        std::copy(ct.pt_bary.begin(), ct.pt_bary.begin() + 3 * (size_t)m, bary);
        std::copy(ct.pt_depth.begin(), ct.pt_depth.begin() + m, depth);
        return n;
    });
}

int bw_get_edge_edge_contacts(const bw_cloth *c, int cap, int *ids, double *normal, float *bary,
                              float *depth)
{
    return guard([&] {
        if (!c || cap < 0)
            return fail(BW_ERR_ARG, "bad arguments");
        const bw::Contacts &ct = c->batch->contacts(0);
        const int n = (int)ct.ee_depth.size(), m = std::min(n, cap);
        if (m > 0 && (!ids || !normal || !bary || !depth))
            return fail(BW_ERR_ARG, "NULL output array");
        std::copy(ct.ee_ids.begin(), ct.ee_ids.begin() + 3 * (size_t)m, ids);
        std::copy(ct.ee_normal.begin(), ct.ee_normal.begin() + 3 * (size_t)m, normal);
        std::copy(ct.ee_bary.begin(), ct.ee_bary.begin() + 2 * (size_t)m, bary);
        std::copy(ct.ee_depth.begin(), ct.ee_depth.begin() + m, depth);
        return n;
    });
}

int bw_body_create(const double *pos, int nv, const int *tri, int nt, bw_body **out)
{
    return guard([&] {
        if (!out)
            return fail(BW_ERR_ARG, "NULL output handle");
        *out = nullptr;
        auto b = std::make_unique<bw_body>();
        b->body = std::make_unique<bw::Body>(pos, nv, tri, nt);
        *out = b.release();
        return BW_OK;
    });
}

void bw_body_destroy(bw_body *b) { delete b; }

int bw_body_set_pos(bw_body *b, const double *pos)
{
    return guard([&] {
        if (!b)
            return fail(BW_ERR_ARG, "NULL body");
        b->body->set_pos(pos);
        return BW_OK;
    });
}

int bw_body_num_edges(const bw_body *b)
{
    if (!b)
        return fail(BW_ERR_ARG, "NULL body");
    return b->body->ne;
}

int bw_body_get_edges(const bw_body *b, int *edge4)
{
    if (!b || !edge4)
        return fail(BW_ERR_ARG, "NULL argument");
    std::copy(b->body->edge.begin(), b->body->edge.end(), edge4);
    return BW_OK;
}

int bw_detect_contacts(bw_cloth *c, const bw_body *b, int cloth_first, int *n_pt, int *n_ee)
{
    return guard([&] {
        if (!c || !b)
            return fail(BW_ERR_ARG, "NULL cloth or body");
        c->batch->detect_contacts(0, *b->body, cloth_first != 0, n_pt, n_ee);
        return BW_OK;
    });
}

int bw_get_init_pos(const bw_cloth *c, double *x)
{
    return guard([&] {
        if (!c || !x)
            return fail(BW_ERR_ARG, "NULL argument");
        std::memcpy(x, c->init_pos.data(), c->init_pos.size() * sizeof(double));
        return BW_OK;
    });
}

double bw_get_default_timestep(const bw_cloth *c) { return c ? c->batch->model(0).cfg.dt : 0.0; }
int bw_num_vertices(const bw_cloth *c) { return c ? c->batch->model(0).nv : BW_ERR_ARG; }
int bw_num_triangles(const bw_cloth *c) { return c ? c->batch->model(0).nt : BW_ERR_ARG; }
int bw_num_edges(const bw_cloth *c) { return c ? c->batch->model(0).ne : BW_ERR_ARG; }

int bw_get_constraint_static_vertices(const bw_cloth *c, int *ids, int cap)
{
    if (!c)
        return fail(BW_ERR_ARG, "NULL cloth");
    const auto &p = c->batch->model(0).pins;
    for (int k = 0; k < (int)p.size() && k < cap && ids; k++)
        ids[k] = p[k];
    return (int)p.size();
}

int bw_add_constraint_static_vertices(bw_cloth *c, const int *ids, int n)
{
    return guard([&] {
        if (!c || (n > 0 && !ids) || n < 0)
            return fail(BW_ERR_ARG, "bad arguments");
        std::vector<int> pins = c->batch->model(0).pins;
        for (int k = 0; k < n; k++)
        {
            if (ids[k] < 0 || ids[k] >= c->batch->model(0).nv)
                return fail(BW_ERR_ARG, "vertex id out of range");
            pins.push_back(ids[k]);
        }
        c->batch->refresh_pins(0, pins);
        return BW_OK;
    });
}

int bw_clear_constraint_static_vertices(bw_cloth *c)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        c->batch->refresh_pins(0, {});
        return BW_OK;
    });
}

int bw_get_triangles(const bw_cloth *c, int *tri3)
{
    if (!c || !tri3)
        return fail(BW_ERR_ARG, "NULL argument");
    const auto &m = c->batch->model(0);
    for (int t = 0; t < m.nt; t++)
        for (int k = 0; k < 3; k++)
            tri3[3 * t + k] = m.tri[t][k];
    return BW_OK;
}

int bw_get_edges(const bw_cloth *c, int *e5)
{
    if (!c || !e5)
        return fail(BW_ERR_ARG, "NULL argument");
    const auto &m = c->batch->model(0);
    for (int e = 0; e < m.ne; e++)
    {
        const auto &ed = m.edges[e];
        e5[5 * e] = ed.v0;
        e5[5 * e + 1] = ed.v1;
        e5[5 * e + 2] = ed.t0;
        e5[5 * e + 3] = ed.t1;
        e5[5 * e + 4] = ed.boundary ? 1 : 0;
    }
    return BW_OK;
}

int bw_get_ell_map(const bw_cloth *c, int *row_ptr, int *cols)
{
    return guard([&] {
        if (!c)
            return fail(BW_ERR_ARG, "NULL cloth");
        std::vector<int> rp, cl;
        c->batch->model(0).reference_ell(rp, cl);
        if (row_ptr && cols)
        {
            std::memcpy(row_ptr, rp.data(), rp.size() * sizeof(int));
            std::memcpy(cols, cl.data(), cl.size() * sizeof(int));
        }
        return (int)cl.size();
    });
}

int bw_assemble_system(bw_cloth *c, double *W, double *rhs)
{
    return guard([&] {
        if (!c || !W || !rhs)
            return fail(BW_ERR_ARG, "NULL argument");
        c->batch->assemble_system(0, W, rhs);
        return c->batch->model(0).nnzb();
    });
}

int bw_get_last_stats(const bw_cloth *c, bw_step_stats *st)
{
    if (!c || !st)
        return fail(BW_ERR_ARG, "NULL argument");
    const auto &s = c->batch->stats(0);
    st->pcg_iterations = s.pcg_iterations;
    st->outer_iterations = s.outer_iterations;
    st->matrix_fp64 = s.matrix_fp64;
    st->reserved = 0;
    st->rel_residual = s.rel_residual;
    st->step_us = s.step_us;
    return BW_OK;
}

int bw_batch_create(const char *const *jsons, const double *const *x0s, int n, bw_batch **out)
{
    return guard([&] {
        if (!out || !jsons || n <= 0)
            return fail(BW_ERR_ARG, "bad arguments");
        *out = nullptr;
        std::vector<bw::ClothModel> ms;
        for (int i = 0; i < n; i++)
            ms.push_back(model_from_json(jsons[i]));
        std::vector<size_t> n3;
        for (auto &m : ms)
            n3.push_back(3 * (size_t)m.nv);
        auto b = std::make_unique<bw_batch>();
        b->batch = std::make_unique<bw::Batch>(std::move(ms));
        for (int i = 0; i < n; i++)
            if (x0s && x0s[i])
            {
                if (!all_finite(x0s[i], n3[i]))
                    return fail(BW_ERR_ARG, "x0 has NaN/Inf");
                b->batch->set_pos(i, x0s[i]);
                b->batch->set_pre_pos(i, x0s[i]);
            }
        *out = b.release();
        return BW_OK;
    });
}

void bw_batch_destroy(bw_batch *b) { delete b; }
int bw_batch_size(const bw_batch *b) { return b ? b->batch->size() : BW_ERR_ARG; }

int bw_batch_step(bw_batch *b, int nsteps)
{
    return guard([&] {
        if (!b || nsteps < 0)
            return fail(BW_ERR_ARG, "bad arguments");
        b->batch->step(nsteps);
        return BW_OK;
    });
}

int bw_batch_num_vertices(const bw_batch *b, int i)
{
    if (!b || i < 0 || i >= b->batch->size())
        return fail(BW_ERR_ARG, "bad index");
    return b->batch->model(i).nv;
}

int bw_batch_num_triangles(const bw_batch *b, int i)
{
    if (!b || i < 0 || i >= b->batch->size())
        return fail(BW_ERR_ARG, "bad index");
    return b->batch->model(i).nt;
}

int bw_batch_get_pos(const bw_batch *b, int i, double *x)
{
    return guard([&] {
        if (!b || !x || i < 0 || i >= b->batch->size())
            return fail(BW_ERR_ARG, "bad arguments");
        b->batch->get_pos(i, x);
        return BW_OK;
    });
}

int bw_batch_set_state(bw_batch *b, int i, const double *x, const double *xpre)
{
    return guard([&] {
        if (!b || !x || !xpre || i < 0 || i >= b->batch->size())
            return fail(BW_ERR_ARG, "bad arguments");
        const size_t n3 = 3 * (size_t)b->batch->model(i).nv;
        if (!all_finite(x, n3) || !all_finite(xpre, n3))
            return fail(BW_ERR_ARG, "state has NaN/Inf");
        b->batch->set_pos(i, x);
        b->batch->set_pre_pos(i, xpre);
        return BW_OK;
    });
}

int bw_batch_get_normals(bw_batch *b, int i, double *vn, double *tn)
{
    return guard([&] {
        if (!b || i < 0 || i >= b->batch->size())
            return fail(BW_ERR_ARG, "bad arguments");
        b->batch->normals(i, vn, tn);
        return BW_OK;
    });
}

int bw_batch_set_point_triangle_contacts(bw_batch *b, int i, int n, const int *ids,
                                         const double *normal, const double *bary,
                                         const float *depth)
{
    return guard([&] {
        if (!b)
            return fail(BW_ERR_ARG, "NULL batch");
        b->batch->set_point_triangle_contacts(i, n, ids, normal, bary, depth);
        return BW_OK;
    });
}

int bw_batch_set_edge_edge_contacts(bw_batch *b, int i, int n, const int *ids,
                                    const double *normal, const float *bary, const float *depth)
{
    return guard([&] {
        if (!b)
            return fail(BW_ERR_ARG, "NULL batch");
        b->batch->set_edge_edge_contacts(i, n, ids, normal, bary, depth);
        return BW_OK;
    });
}

int bw_batch_step_noalloc(bw_batch *b, int nsteps)
{
    return guard([&] {
        if (!b || nsteps < 0)
            return fail(BW_ERR_ARG, "bad batch or step count");
        b->batch->step_noalloc(nsteps);
        return BW_OK;
    });
}

int bw_batch_apply_perturb(bw_batch *b, int i, int triangle, const double force[3])
{
    if (!force)
        return fail(BW_ERR_ARG, "NULL force");
    return with_member(b, i, [&] { b->batch->apply_perturb(i, triangle, force); });
}

int bw_batch_apply_user_perturb_force_once(bw_batch *b, int i, int triangle,
                                           const double goal[3])
{
    if (!goal)
        return fail(BW_ERR_ARG, "NULL goal");
    return with_member(b, i,
                       [&] { b->batch->apply_user_perturb_force_once(i, triangle, goal); });
}

int bw_batch_set_y_drag_force(bw_batch *b, int i, float force)
{
    return with_member(b, i, [&] { b->batch->set_y_drag_force(i, force); });
}

int bw_batch_detect_contacts(bw_batch *bt, int i, const bw_body *b, int cloth_first, int *n_pt,
                             int *n_ee)
{
    return guard([&] {
        if (!bt || !b)
            return fail(BW_ERR_ARG, "NULL batch or body");
        bt->batch->detect_contacts(i, *b->body, cloth_first != 0, n_pt, n_ee);
        return BW_OK;
    });
}

int bw_batch_num_blocks(const bw_batch *b, int i)
{
    if (!b || i < 0 || i >= b->batch->size())
        return fail(BW_ERR_ARG, "bad arguments");
    return b->batch->model(i).nnzb();
}

int bw_batch_get_stats(const bw_batch *b, int i, bw_step_stats *st)
{
    if (!b || !st || i < 0 || i >= b->batch->size())
        return fail(BW_ERR_ARG, "bad arguments");
    const auto &s = b->batch->stats(i);
    st->pcg_iterations = s.pcg_iterations;
    st->outer_iterations = s.outer_iterations;
    st->matrix_fp64 = s.matrix_fp64;
    st->reserved = 0;
    st->rel_residual = s.rel_residual;
    st->step_us = s.step_us;
    return BW_OK;
}

int bw_batch_time_solver_kernel(bw_batch *b, int reps, double *avg_us, double *alg_bytes)
{
    return guard([&] {
        if (!b || !avg_us)
            return fail(BW_ERR_ARG, "NULL argument");
        *avg_us = b->batch->time_solver_kernel(reps, alg_bytes);
        return BW_OK;
    });
}

int bw_batch_time_launch_floor(bw_batch *b, int reps, double *avg_us)
{
    return guard([&] {
        if (!b || !avg_us)
            return fail(BW_ERR_ARG, "NULL argument");
        *avg_us = b->batch->time_launch_floor(reps);
        return BW_OK;
    });
}

int bw_batch_autotune(bw_batch *b, int *lanes, int *staged, int *nontemporal, double *us)
{
    return guard([&] {
        if (!b || !lanes)
            return fail(BW_ERR_ARG, "NULL argument");
        *lanes = b->batch->autotune();
        if (staged)
            *staged = b->batch->spmv_stage() ? 1 : 0;
        if (nontemporal)
            *nontemporal = b->batch->spmv_nt() ? 1 : 0;
        if (us)
            for (int k = 0; k < bw::Batch::TUNE_VARIANTS; k++)
                us[k] = b->batch->tune_us()[k];
        return BW_OK;
    });
}

int bw_batch_set_graphs(bw_batch *b, int enable)
{
    if (!b)
        return fail(BW_ERR_ARG, "NULL batch");
    b->batch->set_graphs(enable != 0);
    return BW_OK;
}

} // extern "C"

// ---- serving ----------------------------------------------------------

struct bw_server
{
    std::unique_ptr<bw::Server> server;
};

extern "C" int bw_server_create(double budget_bytes, int max_resident, const char *precision,
                                bw_server **out)
{
    return guard([&] {
        if (!out)
            return fail(BW_ERR_ARG, "NULL output pointer");
        *out = nullptr;
        bw::ServerOptions o;
        if (!(budget_bytes > 0))
            return fail(BW_ERR_ARG, "budget_bytes must be positive");
        o.budget_bytes = (size_t)budget_bytes;
        o.max_resident = max_resident;
        const std::string p = precision ? precision : "mixed";
        if (p != "mixed" && p != "fp64")
            return fail(BW_ERR_ARG, "precision must be \"mixed\" or \"fp64\"");
        o.fp32 = p == "mixed";
        auto s = std::make_unique<bw_server>();
        s->server = std::make_unique<bw::Server>(o);
        *out = s.release();
        return BW_OK;
    });
}

extern "C" void bw_server_destroy(bw_server *s) { delete s; }

extern "C" int bw_server_submit(bw_server *s, const char *json, const double *x0, int nx0,
                                int nsteps, long long *id)
{
    return guard([&] {
        if (!s || !json || !id)
            return fail(BW_ERR_ARG, "NULL argument");
        if (x0 && nx0 > 0 && !all_finite(x0, (size_t)nx0))
            return fail(BW_ERR_ARG, "x0 holds NaN/Inf");
        *id = s->server->submit(json, x0, x0 ? nx0 : 0, nsteps);
        return BW_OK;
    });
}

extern "C" int bw_server_wait(bw_server *s, long long id, double timeout_s)
{
    return guard([&] {
        if (!s)
            return fail(BW_ERR_ARG, "NULL server");
        if (!s->server->wait(id, timeout_s))
            return fail(BW_ERR_STATE, "request not done within the timeout");
        return BW_OK;
    });
}

extern "C" int bw_server_get_info(bw_server *s, long long id, bw_request_info *info)
{
    return guard([&] {
        if (!s || !info)
            return fail(BW_ERR_ARG, "NULL argument");
        const bw::RequestInfo r = s->server->info(id);
        info->status = r.status;
        info->done = r.done ? 1 : 0;
        info->num_vertices = r.nv;
        info->steps = r.steps;
        info->steps_done = r.steps_done;
        info->pcg_iterations = r.iterations;
        info->max_rel_residual = r.max_residual;
        info->t_submit = r.t_submit;
        info->t_admit = r.t_admit;
        info->t_first_step = r.t_first;
        info->t_done = r.t_done;
        info->device_bytes = (double)r.device_bytes;
        info->admitted_bytes = (double)r.admitted_bytes;
        if (r.status != BW_OK)
            g_error = r.error;
        return BW_OK;
    });
}

extern "C" int bw_server_get_pos(bw_server *s, long long id, double *x)
{
    return guard([&] {
        if (!s || !x)
            return fail(BW_ERR_ARG, "NULL argument");
        const bw::RequestInfo r = s->server->info(id);
        if (!r.done)
            return fail(BW_ERR_STATE, "request still running");
        if (r.status != BW_OK)
            return fail(r.status, r.error);
        std::memcpy(x, r.pos.data(), r.pos.size() * sizeof(double));
        return BW_OK;
    });
}

extern "C" int bw_server_release(bw_server *s, long long id)
{
    return guard([&] {
        if (!s)
            return fail(BW_ERR_ARG, "NULL server");
        s->server->release(id);
        return BW_OK;
    });
}

extern "C" int bw_server_get_counters(bw_server *s, bw_server_counters *c)
{
    return guard([&] {
        if (!s || !c)
            return fail(BW_ERR_ARG, "NULL argument");
        const bw::ServerCounters k = s->server->counters();
        c->queued = k.queued;
        c->resident = k.resident;
        c->completed = k.completed;
        c->max_resident_seen = k.max_resident_seen;
        c->bytes_in_use = (double)k.bytes_in_use;
        c->chunks = k.chunks;
        c->launches = k.launches;
        return BW_OK;
    });
}

extern "C" double bw_server_now(bw_server *s) { return s ? s->server->now() : 0.0; }
