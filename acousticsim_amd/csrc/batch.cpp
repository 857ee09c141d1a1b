#include "batch.h"
#include "mesh_gen.h"
#include "noalloc.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace bw
{

namespace
{
void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
} // namespace

// Slices above which a cloth's solver sums go through group sums
// (BW_GROUP_MIN overrides the choice for tuning runs).  The group-sum
// launches cost every iteration a little latency and save each large cloth
// a redundant pass over its partials that grows with the square of its
// slice count.  A batch's cloth may be solved alone and pay the latency in
// full: 2048 (a 257 x 257 cloth, 1,032 slices, runs 23.7 steps/s without
// group sums and 21.0 with).  The server's cloths always share their rounds'
// launches: 512 (the serving mix runs 42.6 cloth-steps/s at 512 against
// 32.0 at 2048).
static int group_min(bool shared_launches)
{
    if (const char *e = std::getenv("BW_GROUP_MIN"))
        return std::max(0, std::atoi(e));
    return shared_launches ? 512 : 2048;
}

template <typename T> T *Batch::dalloc(size_t n)
{
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    check(hipMalloc(&p, bytes), "hipMalloc");
    allocations_.push_back(p);
    device_bytes_ += bytes;
    check(hipMemset(p, 0, bytes), "hipMemset");
    // the fill runs on the null stream, which does not order work on this
    // batch's non-blocking stream: finish it before a kernel can write here
    check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    return static_cast<T *>(p);
}

namespace
{
template <typename T> void upload(T *dst, const std::vector<T> &src)
{
    if (!src.empty())
        check(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice),
              "hipMemcpy H2D");
}
} // namespace

// Row order of one cloth.  Grids: vertices go to "wide" or "narrow" rows by their
// block-row width and are packed into 64-row slices in vertex order, one open
// slice per class.  Regular grids alternate 13- and 9-wide rows in a
// checkerboard, so vertex order would pad every slice to 13; this order pads
// under 1% while keeping both classes' slices on the same stretch of the
// cloth (the gathered columns stay cache-local).
std::vector<int> slice_row_order(const ClothModel &m)
{
    if (m.cfg.geometry_type != "regular_triangle")
    {
        // Irregular meshes: vertex ids carry no locality, so rows follow a
        // Hilbert curve over the rest positions; inside every window of
        // 65536 rows they are sorted by width class (widths in classes of
        // 2 slots, widest first, curve order kept within a class), so that
        // a 64-row slice holds rows of similar width (ELL padding 9% with
        // classes of 4 on Delaunay meshes with widths 4..25, 0.2% with exact
        // widths, 39% in vertex order) whose columns stay close along the
        // curve.  SpMV on a 1.5M-point cloth, staged directions (the
        // autotuner's pick on such meshes): 165 us (classes of 2, 65536-row
        // windows), 168 us (2, 16384), 170 us (4, 16384), 172 us (1, 16384
        // or 2, 131072), 186 us (4, 256).  With gathered directions the
        // locality is worth more than the padding: 226 us (4, 16384),
        // 237 us (exact widths), 255 us (curve order only), 355 us (global
        // width sort).
        std::vector<int> order = hilbert_order(m.flat_xy);
        int window = 65536; // BW_SORT_WINDOW overrides it for tuning runs
        if (const char *e = std::getenv("BW_SORT_WINDOW"))
            window = std::max(64, std::atoi(e));
        int bucket = 2; // widths compared in classes of this many slots
        if (const char *e = std::getenv("BW_WIDTH_BUCKET"))
            bucket = std::max(1, std::atoi(e));
        for (size_t s = 0; s < order.size(); s += window)
        {
            auto b = order.begin() + s,
                 e = order.begin() + std::min(order.size(), s + (size_t)window);
            std::stable_sort(b, e, [&](int x, int y) {
                return (m.row_ptr[x + 1] - m.row_ptr[x]) / bucket >
                       (m.row_ptr[y + 1] - m.row_ptr[y]) / bucket;
            });
        }
        return order;
    }
    int wmin = 1 << 30, wmax = 0;
    for (int v = 0; v < m.nv; v++)
    {
        const int w = m.row_ptr[v + 1] - m.row_ptr[v];
        wmin = std::min(wmin, w);
        wmax = std::max(wmax, w);
    }
    std::vector<int> order, wide, narrow;
    order.reserve(m.nv);
    for (int v = 0; v < m.nv; v++)
    {
        const int w = m.row_ptr[v + 1] - m.row_ptr[v];
        std::vector<int> &cls = 2 * w > wmin + wmax ? wide : narrow;
        cls.push_back(v);
        if ((int)cls.size() == SLICE)
        {
            order.insert(order.end(), cls.begin(), cls.end());
            cls.clear();
        }
    }
    order.insert(order.end(), wide.begin(), wide.end());
    order.insert(order.end(), narrow.begin(), narrow.end());
    return order;
}

Batch::Batch(std::vector<ClothModel> models, const BatchOptions &opt)
    : models_(std::move(models)), opt_(opt)
{
    if (models_.empty())
        throw std::runtime_error("batch: no cloths");
    auto kind_of = [](const ClothModel &m) {
        if (m.cfg.gpu_form)
            return SolverKind::GpuForm;
        if (m.cfg.solve_mode == SolveMode::Reference)
            return SolverKind::RefPCG64;
        return m.cfg.precision == Precision::Mixed ? SolverKind::CG32 : SolverKind::CG64;
    };
    kind_ = kind_of(models_[0]);
    for (auto &m : models_)
        if (kind_of(m) != kind_)
            throw std::runtime_error(
                "batch: members must share one precision and solver_mode");
    stats_.resize(models_.size());
    if (opt_.stream)
        stream_ = opt_.stream;
    else
    {
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        own_stream_ = true;
    }
    // BW_GRAPHS=0 launches the iteration chains directly (profilers that do
    // not follow graph launches)
    if (const char *g = std::getenv("BW_GRAPHS"))
        graphs_disabled_ = std::string(g) == "0";
    use_graphs_ = !graphs_disabled_;
    check(hipHostMalloc((void **)&host_state_, sizeof(InstState) * models_.size()),
          "hipHostMalloc");
    check(hipHostMalloc((void **)&host_cg_, 2 * sizeof(CgState) * models_.size()),
          "hipHostMalloc");
    check(hipHostMalloc((void **)&host_gf_, 2 * sizeof(GfState) * models_.size()),
          "hipHostMalloc");
    if (kind_ == SolverKind::GpuForm)
        tuned_ = true; // one SpMV form
    build_layout();
    // lane 0: the batch's own stream, list and state buffers
    lanes_.resize(1);
    Lane &L0 = lanes_[0];
    L0.stream = stream_;
    L0.spmv_list = d_.spmv_list;
    L0.upd_list = d_.upd_list;
    L0.host_spmv = host_spmv_list_;
    L0.host_upd = host_upd_list_;
    L0.host_cg = host_cg_;
    check(hipEventCreateWithFlags(&L0.ev, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&lanes_start_, hipEventDisableTiming), "hipEventCreate");
}

Batch::~Batch()
{
    for (auto &L : lanes_)
        close_lane(L);
    if (stream_)
        (void)hipStreamSynchronize(stream_);
    if (!lanes_.empty())
    {
        for (auto &g : lanes_[0].graphs)
            (void)hipGraphExecDestroy(g.exec);
        if (lanes_[0].ev)
            (void)hipEventDestroy(lanes_[0].ev);
    }
    if (lanes_start_)
        (void)hipEventDestroy(lanes_start_);
    for (void *p : allocations_)
        (void)hipFree(p);
    if (host_state_)
        (void)hipHostFree(host_state_);
    if (host_cg_)
        (void)hipHostFree(host_cg_);
    if (host_gf_)
        (void)hipHostFree(host_gf_);
    if (host_spmv_list_)
        (void)hipHostFree(host_spmv_list_);
    if (host_upd_list_)
        (void)hipHostFree(host_upd_list_);
    if (contact_dev_)
        (void)hipFree(contact_dev_);
    if (detect_count_dev_)
        (void)hipFree(detect_count_dev_);
    if (detect_hit_dev_)
        (void)hipFree(detect_hit_dev_);
    if (stream_ && own_stream_)
        (void)hipStreamDestroy(stream_);
}

void Batch::build_layout()
{
    const int n = (int)models_.size();
    row_base_.resize(n);
    row_of_.resize(n);
    slice_base_.resize(n);
    nslices_.resize(n);
    int slices = 0;
    int nvert = 0, ntri = 0, nie = 0, nblk = 0;
    std::vector<int> vert_base(n), tri_base(n), ie_base(n), blk_base(n);
    for (int i = 0; i < n; i++)
    {
        const ClothModel &m = models_[i];
        const std::vector<int> order = slice_row_order(m);
        row_of_[i].assign(m.nv, 0);
        for (int r = 0; r < m.nv; r++)
            row_of_[i][order[r]] = r;
        slice_base_[i] = slices;
        nslices_[i] = ((m.nv + SLICE - 1) / SLICE + BW_SB - 1) / BW_SB * BW_SB;
        row_base_[i] = slices * SLICE;
        slices += nslices_[i];
        vert_base[i] = nvert;
        tri_base[i] = ntri;
        ie_base[i] = nie;
        blk_base[i] = nblk;
        nvert += m.nv;
        ntri += m.nt;
        nie += (int)m.interior.size();
        nblk += m.nnzb();
    }
    if ((long long)ntri >= (1LL << 27) || (long long)nie >= (1LL << 27))
        throw std::runtime_error("batch: too many elements");
    vert_base_ = vert_base;
    tri_base_ = tri_base;
    d_.n_inst = n;
    if (const char *e = std::getenv("BW_XCD_MAP"))
        d_.xcd_map = std::atoi(e) != 0;
    d_.n_slices = slices;
    d_.n_rows = slices * SLICE;
    d_.n_vert = nvert;
    d_.n_tri = ntri;
    d_.n_ie = nie;
    d_.n_blocks = nblk;

    std::vector<int> slice_off(slices), slice_w(slices), slice_inst(slices);
    std::vector<int> row_vertex(d_.n_rows, -1), row_w(d_.n_rows, 0);
    std::vector<uint8_t> row_pinned(d_.n_rows, 0);
    std::vector<double> row_mass(d_.n_rows, 0.0);
    long long entries = 0;
    for (int i = 0; i < n; i++)
    {
        const ClothModel &m = models_[i];
        for (int v = 0; v < m.nv; v++)
        {
            const int p = row_base_[i] + row_of_[i][v];
            row_vertex[p] = vert_base[i] + v;
            row_w[p] = m.row_ptr[v + 1] - m.row_ptr[v];
            row_pinned[p] = m.pinned[v];
            row_mass[p] = m.mass[v];
        }
        for (int s = 0; s < nslices_[i]; s++)
        {
            const int gs = slice_base_[i] + s;
            int w = 0;
            for (int l = 0; l < SLICE; l++)
                w = std::max(w, row_w[gs * SLICE + l]);
            slice_w[gs] = w;
            slice_inst[gs] = i;
            slice_off[gs] = (int)entries;
            entries += (long long)w * SLICE;
        }
    }
    if (entries >= (1LL << 31) / 9)
        throw std::runtime_error("batch: matrix too large for 32-bit entry offsets");
    d_.n_entries = entries;
    std::vector<int> col(entries, 0);
    for (int s = 0; s < slices; s++)
        for (int k = 0; k < slice_w[s]; k++)
            for (int l = 0; l < SLICE; l++)
                col[slice_off[s] + (size_t)k * SLICE + l] = s * SLICE + l; // padding: self
    std::vector<int> vert_row(nvert), blk_ptr(nvert + 1, 0), blk_col(nblk);
    std::vector<int> s_ptr(nblk + 1, 0), b_ptr(nblk + 1, 0), s_list, b_list;
    std::vector<int> vt_ptr(nvert + 1, 0), ve_ptr(nvert + 1, 0), vt_list, ve_list;
    std::vector<int> tri_rows(3 * (size_t)ntri), tri_inst(ntri), ie_rows(4 * (size_t)nie);
    std::vector<float> tri_coef(12 * (size_t)ntri), tri_area(ntri);
    std::vector<double> ie_factor(nie);
    std::vector<InstParams> params(n);
    std::vector<int> group_inst;
    // a batch's own rule does not depend on its other members, so that a
    // cloth gives the same result alone and in a mix
    const int gmin = group_min(opt_.shared_launches);
    for (int i = 0; i < n; i++)
    {
        const ClothModel &m = models_[i];
        const int rb = row_base_[i];
        const std::vector<int> &ro = row_of_[i];
        for (int v = 0; v < m.nv; v++)
        {
            const int vg = vert_base[i] + v;
            vert_row[vg] = rb + ro[v];
            blk_ptr[vg + 1] = blk_base[i] + m.row_ptr[v + 1];
            vt_ptr[vg + 1] = (int)vt_list.size() + m.vt_ptr[v + 1];
            ve_ptr[vg + 1] = (int)ve_list.size() + m.ve_ptr[v + 1];
            for (int k = m.row_ptr[v]; k < m.row_ptr[v + 1]; k++)
            {
                const int p = rb + ro[v], ks = k - m.row_ptr[v];
                const int s = p / SLICE, l = p % SLICE;
                col[slice_off[s] + (size_t)ks * SLICE + l] = rb + ro[m.cols[k]];
            }
        }
        if (m.nv > 0)
        {
            blk_ptr[vert_base[i]] = blk_base[i];
            vt_ptr[vert_base[i]] = (int)vt_list.size();
            ve_ptr[vert_base[i]] = (int)ve_list.size();
        }
        for (int code : m.vt_list)
            vt_list.push_back(((code >> 2) + tri_base[i]) << 2 | (code & 3));
        for (int code : m.ve_list)
            ve_list.push_back(((code >> 2) + ie_base[i]) << 2 | (code & 3));
        const int sb = (int)s_list.size(), bb = (int)b_list.size();
        for (int k = 0; k < m.nnzb(); k++)
        {
            const int gb = blk_base[i] + k;
            blk_col[gb] = rb + ro[m.cols[k]];
            s_ptr[gb + 1] = sb + m.s_ptr[k + 1];
            b_ptr[gb + 1] = bb + m.b_ptr[k + 1];
        }
        if (m.nnzb() > 0)
        {
            s_ptr[blk_base[i]] = sb;
            b_ptr[blk_base[i]] = bb;
        }
        for (int code : m.s_list)
            s_list.push_back(((code >> 4) + tri_base[i]) << 4 | (code & 15));
        for (int code : m.b_list)
            b_list.push_back(((code >> 4) + ie_base[i]) << 4 | (code & 15));
        for (int t = 0; t < m.nt; t++)
        {
            const int tg = tri_base[i] + t;
            for (int k = 0; k < 3; k++)
                tri_rows[3 * (size_t)tg + k] = rb + ro[m.tri[t][k]];
            tri_inst[tg] = i;
            for (int k = 0; k < 12; k++)
                tri_coef[12 * (size_t)tg + k] = m.coef[12 * (size_t)t + k];
            tri_area[tg] = m.area[t];
        }
        for (size_t ie = 0; ie < m.interior.size(); ie++)
        {
            const size_t g = ie_base[i] + ie;
            for (int k = 0; k < 4; k++)
                ie_rows[4 * g + k] = rb + ro[m.stencil[m.interior[ie]][k]];
            ie_factor[g] = m.bend_factor[ie];
        }
        InstParams &P = params[i];
        P.dt = m.cfg.dt;
        P.sk = m.sk;
        P.sm = m.sm;
        P.k0 = m.k[0];
        P.k1 = m.k[1];
        P.k2 = m.k[2];
        P.grav = m.cfg.gravity ? 1.0 : 0.0;
        P.ref_mode = m.cfg.solve_mode == SolveMode::Reference ? 1 : 0;
        P.tol = P.ref_mode ? 0.0 : m.cfg.tolerance;
        P.max_iter = P.ref_mode ? m.cfg.max_iters - 1 : m.cfg.solver_max_iters;
        // reliable updates: FP32 recurrences are replaced by the FP64 true
        // residual each time r.r drops by rel2, or after max_since
        // iterations; FP64 recurrences are checked only at convergence.
        // BW_REL2 and BW_MAX_SINCE override them for tuning runs.
        // The forced check's window grows with the cloth's linear size: CG
        // needs iterations in proportion to the mesh's resolution to reduce
        // the residual by a given factor, so a fixed window takes a large
        // cloth's ordinary slow stretches for stalls.  2000 iterations per
        // 128 vertices of side (sqrt(V / 16384)), at least 2000: the
        // batch64 cloths (up to 129 x 129) keep 2000-2016, a 257 x 257
        // cloth gets 4016.
        const bool mixed = kind_ == SolverKind::CG32;
        P.rel2 = mixed ? 1e-2 : 0.0;
        const double side = std::sqrt(std::max(1.0, (double)m.nv / 16384.0));
        P.max_since = mixed ? (int)std::min(2000.0 * side, 1e9) : (1 << 30);
        if (const char *e = std::getenv("BW_REL2"))
            P.rel2 = std::atof(e);
        if (const char *e = std::getenv("BW_MAX_SINCE"))
            P.max_since = std::atoi(e);
        // a forced check that finds r.r above 1e-4 of its value at the last
        // replacement (less than 100x down in 2000 iterations) moves the
        // cloth to the FP64 matrix for the rest of the solve (BW_STALL
        // overrides the ratio; 0 keeps the FP32 matrix)
        P.stall = mixed ? 1e-4 : 0.0;
        if (const char *e = std::getenv("BW_STALL"))
            P.stall = mixed ? std::atof(e) : 0.0;
        // the switch carries over to the cloth's next solves once `carry`
        // solves in a row have switched (BW_STALL_CARRY overrides the
        // count): 1 keeps the FP64 matrix after the first stall, which the
        // stiffest batch members need (batch64 515 cloth-steps/s against 480
        // with 2); a cloth whose solves stall only now and then does better
        // with 2 (single257 under a fixed 2000-iteration window: 19.6
        // steps/s against 16.1; under the size-scaled window it does not
        // switch at all)
        P.carry = 1;
        if (const char *e = std::getenv("BW_STALL_CARRY"))
            P.carry = std::max(1, std::atoi(e));
        P.gf_a = m.cfg.rayleigh_a;
        P.gf_b = m.cfg.rayleigh_b;
        P.blk_begin = slice_base_[i] / BW_SB;
        P.blk_count = nslices_[i] / BW_SB;
        P.slice_begin = slice_base_[i];
        P.slice_count = nslices_[i];
        // group sums for cloths whose per-workgroup pass over all slice
        // partials would cost more than one extra small launch
        P.group_begin = (int)group_inst.size();
        P.group_count = nslices_[i] > gmin ? (nslices_[i] + GROUP - 1) / GROUP : 0;
        group_inst.insert(group_inst.end(), P.group_count, i);
    }
    d_.n_groups = (int)group_inst.size();
    // per slice: cloth, solver-sum range (negative count: group sums) and
    // first ELL entry, so that a solver kernel needs one load to find them
    std::vector<int4> slice_info(slices);
    for (int i = 0; i < n; i++)
    {
        const InstParams &P = params[i];
        for (int s = 0; s < nslices_[i]; s++)
        {
            const int gs = slice_base_[i] + s;
            slice_info[gs] = P.group_count > 0
                                 ? make_int4(i, P.group_begin, -P.group_count, slice_off[gs])
                                 : make_int4(i, P.slice_begin, P.slice_count, slice_off[gs]);
        }
    }
    d_.slice_info = dalloc<int4>(slices);
    upload(d_.slice_info, slice_info);
    slice_info_ = slice_info;

    d_.slice_off = dalloc<int>(slices);
    d_.slice_w = dalloc<int>(slices);
    d_.slice_inst = dalloc<int>(slices);
    d_.n_blocks_wg = slices / BW_SB;
    d_.blk_inst = dalloc<int>(d_.n_blocks_wg);
    std::vector<int> blk_inst(d_.n_blocks_wg);
    for (int s = 0; s < slices; s += BW_SB)
        blk_inst[s / BW_SB] = slice_inst[s];
    d_.row_vertex = dalloc<int>(d_.n_rows);
    d_.row_w = dalloc<int>(d_.n_rows);
    d_.row_pinned = dalloc<uint8_t>(d_.n_rows);
    d_.row_mass = dalloc<double>(d_.n_rows);
    d_.col = dalloc<int>(entries + COL_PAD);
    // the staged SpMV's lists: the distinct columns of each slice, ascending,
    // and each entry's position in its slice's list
    std::vector<int> halo_off(slices + 1, 0), halo_row;
    std::vector<uint16_t> lcol(entries);
    {
        int hmax = 0;
        std::vector<int> u;
        for (int s = 0; s < slices; s++)
        {
            const size_t e0 = slice_off[s], e1 = e0 + (size_t)slice_w[s] * SLICE;
            u.assign(col.begin() + e0, col.begin() + e1);
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            hmax = std::max(hmax, (int)u.size());
            if (hmax > HALO_CAP)
                break;
            for (size_t e = e0; e < e1; e++)
                lcol[e] = (uint16_t)(std::lower_bound(u.begin(), u.end(), col[e]) - u.begin());
            halo_row.insert(halo_row.end(), u.begin(), u.end());
            halo_off[s + 1] = (int)halo_row.size();
        }
        if (hmax <= HALO_CAP && slices > 0)
        {
            // LDS for every list, or, for batches that share launches with
            // others (the server sizes the launch's LDS by the longest of
            // its batches), for the 99th percentile of the list lengths (at
            // least 256 entries, a multiple of 64); the rest of the slices
            // take the gathered product.  Serving mix: 47.5 against 36.1
            // cloth-steps/s; one 1.5M-point cloth: 165 us with every list
            // staged against 172 us with the percentile.  BW_STAGE_LDS=max
            // or p99 fixes the rule (a batch held to the server's form), a
            // number of entries (>= 64) fixes the LDS itself
            d_.halo_lds = hmax;
            bool p99 = opt_.shared_launches;
            int fixed = 0;
            if (const char *e = std::getenv("BW_STAGE_LDS"))
            {
                p99 = std::string(e) == "p99";
                fixed = std::atoi(e);
            }
            if (fixed >= 64)
            {
                p99 = false;
                d_.halo_lds = std::min(hmax, fixed);
            }
            if (p99)
            {
                std::vector<int> len(slices);
                for (int s = 0; s < slices; s++)
                    len[s] = halo_off[s + 1] - halo_off[s];
                const int k = (slices - 1) * 99 / 100;
                std::nth_element(len.begin(), len.begin() + k, len.end());
                d_.halo_lds = std::min(hmax, std::max(256, (len[k] + 63) / 64 * 64));
            }
            d_.halo_max = hmax;
            n_halo_ = (long long)halo_row.size();
            d_.halo_off = dalloc<int>(slices + 1);
            d_.halo_row = dalloc<int>(halo_row.size());
            d_.lcol = dalloc<uint16_t>(entries);
            upload(d_.halo_off, halo_off);
            upload(d_.halo_row, halo_row);
            upload(d_.lcol, lcol);
        }
    }
    d_.W64 = dalloc<double>(9 * (size_t)entries);
    d_.W32 = dalloc<float>(9 * (size_t)entries);
    d_.vert_row = dalloc<int>(nvert);
    d_.blk_ptr = dalloc<int>(nvert + 1);
    d_.blk_col = dalloc<int>(nblk);
    d_.s_ptr = dalloc<int>(nblk + 1);
    d_.s_list = dalloc<int>(s_list.size());
    d_.b_ptr = dalloc<int>(nblk + 1);
    d_.b_list = dalloc<int>(b_list.size());
    d_.vt_ptr = dalloc<int>(nvert + 1);
    d_.vt_list = dalloc<int>(vt_list.size());
    d_.ve_ptr = dalloc<int>(nvert + 1);
    d_.ve_list = dalloc<int>(ve_list.size());
    d_.tri_rows = dalloc<int>(3 * (size_t)ntri);
    d_.tri_inst = dalloc<int>(ntri);
    d_.tri_coef = dalloc<float>(12 * (size_t)ntri);
    d_.tri_area = dalloc<float>(ntri);
    d_.ie_rows = dalloc<int>(4 * (size_t)nie);
    d_.ie_factor = dalloc<double>(nie);
    d_.tri_ing = dalloc<double>(25 * (size_t)ntri);
    d_.ie_ing = dalloc<double>(25 * (size_t)nie);
    const size_t nv3 = 3 * (size_t)d_.n_rows;
    d_.x = dalloc<double>(nv3);
    d_.xpre = dalloc<double>(nv3);
    d_.dx = dalloc<double>(nv3);
    d_.b = dalloc<double>(nv3);
    d_.P64 = dalloc<double>(nv3);
    d_.P32 = dalloc<float>(nv3);
    d_.params = dalloc<InstParams>(n);
    d_.cg_fresh = dalloc<unsigned char>(n);
    upload(d_.cg_fresh, std::vector<unsigned char>(n, 1));
    if (kind_ == SolverKind::RefPCG64)
    {
        d_.r64 = dalloc<double>(nv3);
        d_.d64[0] = dalloc<double>(nv3);
        d_.d64[1] = dalloc<double>(nv3);
        d_.q64 = dalloc<double>(nv3);
        d_.partial = dalloc<double>(PARTS * (size_t)d_.n_blocks_wg);
        d_.state = dalloc<InstState>(n);
    }
    else if (kind_ == SolverKind::GpuForm)
    {
        d_.r32 = dalloc<float>(nv3);
        d_.q32 = dalloc<float>(nv3);
        d_.d32[0] = dalloc<float>(nv3);
        d_.d32[1] = dalloc<float>(nv3);
        d_.cg_partial = dalloc<double>(2 * PARTS * (size_t)d_.n_slices);
        d_.cg_group = dalloc<double>(2 * PARTS * (size_t)d_.n_groups);
        d_.group_inst = dalloc<int>(d_.n_groups);
        upload(d_.group_inst, group_inst);
        d_.gf_qb = dalloc<float>(entries);
        d_.gf_tri = dalloc<float>(90 * (size_t)ntri);
        d_.gf_pinv = dalloc<float>(9 * (size_t)d_.n_rows);
        d_.gf_dx = dalloc<float>(nv3);
        d_.gf_target = dalloc<float>(nv3);
        d_.gf_mass = dalloc<float>(d_.n_rows);
        d_.gf_grav = dalloc<float>(d_.n_rows);
        d_.gf_state = dalloc<GfState>(2 * (size_t)n);
        std::vector<float> qb(entries, 0.0f), target(nv3, 0.0f), mass(d_.n_rows, 0.0f),
            grav(d_.n_rows, 0.0f);
        for (int i = 0; i < n; i++)
        {
            const ClothModel &m = models_[i];
            for (int v = 0; v < m.nv; v++)
            {
                const int p = row_base_[i] + row_of_[i][v], s = p / SLICE, l = p % SLICE;
                for (int k = m.row_ptr[v]; k < m.row_ptr[v + 1]; k++)
                    qb[slice_off[s] + (size_t)(k - m.row_ptr[v]) * SLICE + l] = m.qb[k];
                for (int a = 0; a < 3; a++)
                    target[3 * (size_t)p + a] = (float)m.xrest[3 * (size_t)v + a];
                mass[p] = m.mass32[v];
                grav[p] = m.grav32[v];
            }
        }
        upload(d_.gf_qb, qb);
        upload(d_.gf_target, target);
        upload(d_.gf_mass, mass);
        upload(d_.gf_grav, grav);
    }
    else
    {
        if (kind_ == SolverKind::CG32)
        {
            d_.dx32 = dalloc<float>(nv3);
            d_.r32 = dalloc<float>(nv3);
            d_.z32 = dalloc<float>(nv3);
            d_.q32 = dalloc<float>(nv3);
            d_.d32[0] = dalloc<float>(nv3);
            d_.d32[1] = dalloc<float>(nv3);
        }
        else
        {
            d_.r64 = dalloc<double>(nv3);
            d_.z64 = dalloc<double>(nv3);
            d_.q64 = dalloc<double>(nv3);
            d_.d64[0] = dalloc<double>(nv3);
            d_.d64[1] = dalloc<double>(nv3);
        }
        d_.cg = opt_.cg_states ? opt_.cg_states : dalloc<CgState>(2 * (size_t)n);
        d_.spmv_list = dalloc<SolverEntry>(slices);
        d_.upd_list = dalloc<SolverEntry>(d_.n_blocks_wg);
        check(hipHostMalloc((void **)&host_spmv_list_, sizeof(SolverEntry) * (size_t)slices),
              "hipHostMalloc");
        check(hipHostMalloc((void **)&host_upd_list_,
                            sizeof(SolverEntry) * (size_t)d_.n_blocks_wg),
              "hipHostMalloc");
        d_.cg_partial = dalloc<double>(2 * PARTS * (size_t)d_.n_slices);
        d_.cg_group = dalloc<double>(2 * PARTS * (size_t)d_.n_groups);
        d_.group_inst = dalloc<int>(d_.n_groups);
        upload(d_.group_inst, group_inst);
    }

    upload(d_.slice_off, slice_off);
    upload(d_.slice_w, slice_w);
    host_slice_off_ = slice_off;
    contacts_.assign(models_.size(), Contacts());
    user_force_.assign(models_.size(), std::vector<double>());
    sched_pairs_.reserve(models_.size());
    lane_order_.reserve(models_.size());
    pred_launches_.reserve(models_.size());
    drag_vertex_.assign(models_.size(), -1);
    ydrag_.assign(models_.size(), 0.0f);
    fuser_dirty_.assign(models_.size(), 0);
    drag_pinned_.assign(models_.size(), 0);
    upload(d_.slice_inst, slice_inst);
    upload(d_.blk_inst, blk_inst);
    upload(d_.row_vertex, row_vertex);
    upload(d_.row_w, row_w);
    upload(d_.row_pinned, row_pinned);
    upload(d_.row_mass, row_mass);
    upload(d_.col, col);
    check(hipMemset(d_.col + entries, 0, COL_PAD * sizeof(int)), "hipMemset");
    upload(d_.vert_row, vert_row);
    upload(d_.blk_ptr, blk_ptr);
    upload(d_.blk_col, blk_col);
    upload(d_.s_ptr, s_ptr);
    upload(d_.s_list, s_list);
    upload(d_.b_ptr, b_ptr);
    upload(d_.b_list, b_list);
    upload(d_.vt_ptr, vt_ptr);
    upload(d_.vt_list, vt_list);
    upload(d_.ve_ptr, ve_ptr);
    upload(d_.ve_list, ve_list);
    upload(d_.tri_rows, tri_rows);
    upload(d_.tri_inst, tri_inst);
    upload(d_.tri_coef, tri_coef);
    upload(d_.tri_area, tri_area);
    upload(d_.ie_rows, ie_rows);
    upload(d_.ie_factor, ie_factor);
    upload(d_.params, params);
    // initial state: rest positions, at rest
    for (int i = 0; i < n; i++)
    {
        upload_state(i, models_[i].xrest.data(), d_.x);
        upload_state(i, models_[i].xrest.data(), d_.xpre);
    }
}

void Batch::refresh_pins(int i, const std::vector<int> &pins)
{
    ClothModel &m = models_[i];
    m.pins = pins;
    m.rebuild_pins_from_list();
    mark_fresh(i);
    std::vector<uint8_t> rows((size_t)nslices_[i] * SLICE, 0);
    for (int v = 0; v < m.nv; v++)
        rows[row_of_[i][v]] = m.pinned[v];
    check(hipMemcpy(d_.row_pinned + row_base_[i], rows.data(), rows.size(), hipMemcpyHostToDevice),
          "hipMemcpy pins");
}

void Batch::mark_fresh(int i)
{
    if (!d_.cg_fresh)
        return;
    const unsigned char one[1] = {1};
    if (i >= 0)
        check(hipMemcpy(d_.cg_fresh + i, one, 1, hipMemcpyHostToDevice), "hipMemcpy fresh");
    else
        upload(d_.cg_fresh, std::vector<unsigned char>(size(), 1));
}

void Batch::upload_state(int i, const double *src, double *dst_dev)
{
    mark_fresh(i);
    std::vector<double> rows(3 * (size_t)nslices_[i] * SLICE, 0.0);
    for (int v = 0; v < models_[i].nv; v++)
        for (int a = 0; a < 3; a++)
            rows[3 * (size_t)row_of_[i][v] + a] = src[3 * (size_t)v + a];
    check(hipMemcpy(dst_dev + 3 * (size_t)row_base_[i], rows.data(), rows.size() * sizeof(double),
                    hipMemcpyHostToDevice),
          "hipMemcpy state");
}

void Batch::download(int i, const double *src_dev, double *dst) const
{
    std::vector<double> rows(3 * (size_t)nslices_[i] * SLICE);
    check(hipMemcpy(rows.data(), src_dev + 3 * (size_t)row_base_[i], rows.size() * sizeof(double),
                    hipMemcpyDeviceToHost),
          "hipMemcpy D2H");
    for (int v = 0; v < models_[i].nv; v++)
        for (int a = 0; a < 3; a++)
            dst[3 * (size_t)v + a] = rows[3 * (size_t)row_of_[i][v] + a];
}

void Batch::normals(int i, double *vn, double *tn)
{
    if (!normals_dev_)
        normals_dev_ = dalloc<double>(3 * ((size_t)d_.n_tri + (size_t)d_.n_vert));
    double *tn_dev = normals_dev_, *vn_dev = normals_dev_ + 3 * (size_t)d_.n_tri;
    launch_normals(d_, tn_dev, vn_dev, stream_);
    check(hipGetLastError(), "normals launch");
    check(hipStreamSynchronize(stream_), "normals sync");
    const ClothModel &m = models_[i];
    if (vn)
        check(hipMemcpy(vn, vn_dev + 3 * (size_t)vert_base_[i], 3 * (size_t)m.nv * sizeof(double),
                        hipMemcpyDeviceToHost),
              "hipMemcpy normals");
    if (tn)
        check(hipMemcpy(tn, tn_dev + 3 * (size_t)tri_base_[i], 3 * (size_t)m.nt * sizeof(double),
                        hipMemcpyDeviceToHost),
              "hipMemcpy normals");
}

void Batch::get_pos(int i, double *x) const { download(i, d_.x, x); }
void Batch::get_pre_pos(int i, double *x) const { download(i, d_.xpre, x); }
void Batch::set_pos(int i, const double *x) { upload_state(i, x, d_.x); }
void Batch::set_pre_pos(int i, const double *x) { upload_state(i, x, d_.xpre); }

void Batch::assemble_system(int i, double *W, double *rhs)
{
    const bool ct = kind_ != SolverKind::GpuForm && upload_contacts();
    launch_elements(d_, stream_);
    launch_assemble_rows(d_, stream_);
    if (ct)
        launch_contact_hessian(d_, contact_blocks_, c_widx_, c_diag_, c_h9_, stream_);
    check(hipGetLastError(), "assemble launch");
    check(hipStreamSynchronize(stream_), "assemble sync");
    const ClothModel &m = models_[i];
    const int s0 = slice_base_[i], s1 = s0 + nslices_[i];
    std::vector<int> off(nslices_[i]);
    check(hipMemcpy(off.data(), d_.slice_off + s0, off.size() * sizeof(int),
                    hipMemcpyDeviceToHost),
          "hipMemcpy slice_off");
    long long e0 = off.empty() ? 0 : off[0];
    long long e1 = e0;
    {
        int last_w = 0;
        check(hipMemcpy(&last_w, d_.slice_w + s1 - 1, sizeof(int), hipMemcpyDeviceToHost),
              "hipMemcpy slice_w");
        e1 = (long long)off.back() + (long long)last_w * SLICE;
    }
    std::vector<double> vals(9 * (size_t)(e1 - e0));
    check(hipMemcpy(vals.data(), d_.W64 + 9 * (size_t)e0, vals.size() * sizeof(double),
                    hipMemcpyDeviceToHost),
          "hipMemcpy W");
    for (int v = 0; v < m.nv; v++)
    {
        const int p = row_base_[i] + row_of_[i][v], s = p / SLICE, l = p % SLICE;
        const size_t base = 9 * (size_t)(off[s - s0] - e0);
        for (int k = m.row_ptr[v]; k < m.row_ptr[v + 1]; k++)
        {
            const int ks = k - m.row_ptr[v];
            for (int q = 0; q < 9; q++)
                W[9 * (size_t)k + q] = vals[base + (size_t)(9 * ks + q) * SLICE + l];
        }
    }
    download(i, d_.b, rhs);
}

double Batch::time_solver_kernel(int reps, double *alg_bytes)
{
    if (reps < 2)
        throw std::runtime_error("batch: reps must be >= 2");
    if (kind_ == SolverKind::GpuForm)
        throw std::runtime_error("batch: kernel timing covers the plain method only");
    // the system at the current state (W and the Jacobi diagonal)
    launch_elements(d_, stream_);
    launch_assemble_rows(d_, stream_);
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    double nnzb = 0, nv = 0;
    for (auto &m : models_)
    {
        nnzb += m.nnzb();
        nv += m.nv;
    }
    double bytes = 0;
    if (kind_ == SolverKind::RefPCG64)
    {
        // a nonzero residual everywhere keeps every cloth active (d.Wd > 0)
        launch_reset_states(d_, stream_);
        launch_fill_residual(d_, false, stream_);
        launch_spmv_only(d_, false, stream_); // warm
        check(hipEventRecord(e0, stream_), "hipEventRecord");
        for (int k = 0; k < reps; k++)
            launch_spmv_only(d_, false, stream_);
        check(hipEventRecord(e1, stream_), "hipEventRecord");
        // W (72 B) + column (4 B) per block; r, P, d_prev read and d, W d written
        bytes = 76.0 * nnzb + 120.0 * nv;
    }
    else
    {
        // k_cg_spmv alone: it reads state slot 0 and writes slot 1, so
        // repeated launches redo the same SpMV with every cloth active
        const bool fp32 = kind_ == SolverKind::CG32;
        DeviceBatch all = d_; // every slice, no lists
        all.spmv_list = nullptr;
        all.upd_list = nullptr;
        launch_cg_timing_setup(all, fp32, stream_);
        launch_cg_spmv_only(all, fp32, stream_); // warm
        check(hipEventRecord(e0, stream_), "hipEventRecord");
        for (int k = 0; k < reps; k++)
            launch_cg_spmv_only(all, fp32, stream_);
        check(hipEventRecord(e1, stream_), "hipEventRecord");
        // per block: W + column (4 B index, or the staged variant's 2 B
        // position plus 4 B per entry of the slices' column lists); per row:
        // z, d_prev read, d, W d written
        const double wb = fp32 ? 36.0 : 72.0, vb = fp32 ? 48.0 : 96.0;
        bytes = d_.spmv_stage ? (wb + 2.0) * nnzb + 4.0 * (double)n_halo_ + vb * nv
                              : (wb + 4.0) * nnzb + vb * nv;
    }
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (alg_bytes)
        *alg_bytes = bytes;
    return 1000.0 * ms / reps;
}

double Batch::time_launch_floor(int reps)
{
    if (reps < 2)
        throw std::runtime_error("batch: reps must be >= 2");
    if (kind_ == SolverKind::GpuForm)
        throw std::runtime_error("batch: kernel timing covers the plain method only");
    // the grid and workgroup of the dominant kernel's launch with every
    // cloth active: one workgroup of spmv_lpr wavefronts per slice (the
    // reference mode's k_pcg_spmv: BW_SB slices per workgroup)
    const bool ref = kind_ == SolverKind::RefPCG64;
    const int grid = ref ? d_.n_blocks_wg : d_.n_slices;
    const int block = ref ? 64 * BW_SB : 64 * d_.spmv_lpr;
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    launch_empty(grid, block, stream_); // warm
    check(hipEventRecord(e0, stream_), "hipEventRecord");
    for (int k = 0; k < reps; k++)
        launch_empty(grid, block, stream_);
    check(hipEventRecord(e1, stream_), "hipEventRecord");
    check(hipGetLastError(), "empty launch");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0 * ms / reps;
}

void Batch::read_states(Lane &L, bool sync)
{
    if (kind_ == SolverKind::RefPCG64)
        check(hipMemcpyAsync(host_state_, d_.state, sizeof(InstState) * models_.size(),
                             hipMemcpyDeviceToHost, L.stream),
              "hipMemcpyAsync state");
    else
        check(hipMemcpyAsync(L.host_cg, d_.cg, 2 * sizeof(CgState) * models_.size(),
                             hipMemcpyDeviceToHost, L.stream),
              "hipMemcpyAsync state");
    if (sync)
        check(hipStreamSynchronize(L.stream), "hipStreamSynchronize");
    else
        check(hipEventRecord(L.ev, L.stream), "hipEventRecord");
}

void Batch::launch_chain(Lane &L, int length, int slices)
{
    if (kind_ == SolverKind::RefPCG64)
        for (int k = 0; k < length; k++)
            launch_pcg_iter64(d_, L.stream);
    else
    {
        DeviceBatch dl = d_;
        dl.spmv_list = L.spmv_list;
        dl.upd_list = L.upd_list;
        dl.lds_pad = L.lds_reserve;
        if (L.launch_first >= 0)
        {
            dl.spmv_list = nullptr;
            dl.upd_list = nullptr;
            dl.slice_first = L.launch_first;
        }
        for (int k = 0; k < length; k++)
            launch_cg_step(dl, kind_ == SolverKind::CG32, k & 1, slices, L.stream);
    }
}

void Batch::run_chunk(Lane &L, int length, int slices)
{
    if (!use_graphs_)
    {
        launch_chain(L, length, slices);
        check(hipGetLastError(), "solver launch");
        return;
    }
    // replayed as graphs of power-of-two lengths (>= 32): chunk lengths vary
    // from step to step, and a graph per exact length would be instantiated
    // over and over
    for (int piece = 1 << 30; piece >= 32; piece >>= 1)
        if (length & piece)
            run_graph(L, piece, slices);
    if (length & 31)
        launch_chain(L, length & 31, slices);
}

void Batch::run_graph(Lane &L, int length, int slices)
{
    ++graph_clock_;
    hipGraphExec_t exec = nullptr;
    for (auto &g : L.graphs)
        if (g.length == length && g.slices == slices && g.first == L.launch_first)
        {
            exec = g.exec;
            g.last_use = graph_clock_;
        }
    if (!exec)
    {
        // bounded cache: drop the least recently used graph
        constexpr size_t max_graphs = 128;
        if (L.graphs.size() >= max_graphs)
        {
            auto lru = std::min_element(L.graphs.begin(), L.graphs.end(),
                                        [](const GraphEntry &a, const GraphEntry &b) {
                                            return a.last_use < b.last_use;
                                        });
            check(hipStreamSynchronize(L.stream), "hipStreamSynchronize");
            (void)hipGraphExecDestroy(lru->exec);
            L.graphs.erase(lru);
        }
        hipGraph_t graph;
        check(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
        launch_chain(L, length, slices);
        check(hipStreamEndCapture(L.stream, &graph), "EndCapture");
        check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "GraphInstantiate");
        (void)hipGraphDestroy(graph);
        L.graphs.push_back({length, slices, L.launch_first, exec, graph_clock_});
    }
    check(hipGraphLaunch(exec, L.stream), "GraphLaunch");
}

void Batch::clear_graphs()
{
    for (auto &L : lanes_)
    {
        if (!L.graphs.empty())
            check(hipStreamSynchronize(L.stream), "hipStreamSynchronize");
        for (auto &g : L.graphs)
            (void)hipGraphExecDestroy(g.exec);
        L.graphs.clear();
    }
}

namespace
{
// BW_MAX_CHUNK caps the launches queued between host checks (default 1024)
int max_chunk()
{
    static const int v = [] {
        const char *e = std::getenv("BW_MAX_CHUNK");
        const int c = e ? std::atoi(e) : 1024;
        return std::max(32, std::min(c, 4096));
    }();
    return v;
}

// chunk lengths are multiples of 32 (even, so that a CG chunk ends with an
// update launch) and few distinct graphs get built
int chunk_for(long long estimate)
{
    const long long c = std::max<long long>(32, std::min<long long>(estimate, max_chunk()));
    return (int)(((c + 31) / 32) * 32);
}

// grid sizes of list launches, rounded up to 64 * {1, 1.5, 2, 3, 4, 6, ...}
// so that few distinct graphs get built; at most `all`
int slice_bucket(int n, int all)
{
    int b = 64;
    while (b < n)
        b = (b & (b - 1)) == 0 ? b / 2 * 3 : b / 3 * 4;
    return std::min(b, all);
}
} // namespace

void Batch::solve_gpu_form()
{
    launch_gf_assemble(d_, stream_);
    launch_gf_solve(d_, stream_);
    check(hipGetLastError(), "gpu-form launch");
    check(hipMemcpyAsync(host_gf_, d_.gf_state, 2 * sizeof(GfState) * models_.size(),
                         hipMemcpyDeviceToHost, stream_),
          "hipMemcpyAsync state");
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (int i = 0; i < size(); i++)
    {
        const GfState &s = host_gf_[2 * i + 1];
        StepStats &st = stats_[i];
        st.pcg_iterations = s.it;
        st.outer_iterations = 0;
        // the recursive residual (the reference checks it, absolute, every
        // 10 iterations)
        st.rel_residual = s.bb > 0 ? std::sqrt(s.rr / s.bb) : std::sqrt(s.rr);
        st.matrix_fp64 = 0;
    }
}

void Batch::clear_warm_start(int i)
{
    if (kind_ != SolverKind::GpuForm)
        return;
    check(hipMemsetAsync(d_.gf_dx + 3 * (size_t)row_base_[i], 0,
                         3 * (size_t)nslices_[i] * SLICE * sizeof(float), stream_),
          "hipMemsetAsync");
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Batch::solve_ref()
{
    launch_reset_states(d_, stream_);
    launch_pcg_init64(d_, stream_);
    int chunk = chunk_for(last_launches_ > 0 ? last_launches_ : 32);
    for (;;)
    {
        read_states(lanes_[0]);
        bool all_done = true;
        for (size_t i = 0; i < models_.size(); i++)
            all_done &= host_state_[i].done != 0;
        if (all_done)
            break;
        run_chunk(lanes_[0], chunk, d_.n_slices);
        chunk = chunk_for(32);
    }
    launch_residual64(d_, 2, stream_);
    read_states(lanes_[0]);
    int most = 0;
    for (size_t i = 0; i < models_.size(); i++)
    {
        const InstState &s = host_state_[i];
        StepStats &st = stats_[i];
        st.pcg_iterations = s.total_iter;
        st.outer_iterations = 1;
        st.rel_residual = s.bb > 0 ? std::sqrt(s.rr64 / s.bb) : std::sqrt(s.rr64);
        most = std::max(most, s.total_iter);
    }
    last_launches_ = most;
}

// Chunk boundaries (launch counts) for this solve of a lane's cloths from
// the launches every one of them needed in the previous solve: a boundary
// wherever the cloths expected to be still solving own at most half the
// slices of the previous boundary.  (The host buffers are the batch's own,
// sized at construction, so a solve allocates no host memory either.)
void Batch::chunk_schedule(Lane &L)
{
    std::vector<long long> &bounds = L.bounds;
    bounds.clear();
    if (pred_launches_.empty())
        return;
    std::vector<std::pair<long long, int>> &v = sched_pairs_;
    v.clear();
    long long remaining = 0;
    for (int i : L.members)
    {
        v.push_back({pred_launches_[i], nslices_[i]});
        remaining += nslices_[i];
    }
    std::sort(v.begin(), v.end());
    long long level = remaining / 2;
    for (size_t k = 0; k < v.size(); k++)
    {
        remaining -= v[k].second;
        if (remaining <= level || k + 1 == v.size())
        {
            bounds.push_back(v[k].first);
            while (level > 0 && remaining <= level)
                level /= 2;
        }
    }
}

// The slices the next chunk's launches cover.  With every cloth active the
// launches cover all slices directly (launch_first_ = 0): no list, so no
// dependent load of a list entry before a workgroup can start.  Otherwise
// lists of the active cloths' slices and solver workgroups (padded with -1
// to the bucketed grid) are uploaded on the stream: a replayed graph reads
// whatever list the buffer holds, so one graph serves every active set of a
// bucket, where a range baked into the launches would need a graph per
// range.  BW_RANGE=2 also runs any contiguous active range directly, 0 never
// (measurement runs).  Returns the k_cg_spmv grid.
int Batch::upload_lists(Lane &L)
{
    const std::vector<char> &active = L.active;
    static const int range_rule = [] {
        const char *e = std::getenv("BW_RANGE");
        return e ? std::atoi(e) : 1;
    }();
    int first = -1, last = -1;
    bool contiguous = true;
    for (int i = 0; i < size() && contiguous; i++)
        if (active[i])
        {
            if (first < 0)
                first = slice_base_[i];
            else if (slice_base_[i] != last)
                contiguous = false;
            last = slice_base_[i] + nslices_[i];
        }
    const bool whole = first == 0 && last == d_.n_slices;
    if (first >= 0 && contiguous && (range_rule == 2 || (range_rule == 1 && whole)))
    {
        L.launch_first = first;
        return last - first;
    }
    L.launch_first = -1;
    int n = 0, nb = 0;
    for (int i = 0; i < size(); i++)
        if (active[i])
        {
            for (int s = 0; s < nslices_[i]; s++)
            {
                const int sl = slice_base_[i] + s;
                L.host_spmv[n++] = {slice_info_[sl], sl, 0, 0, 0};
            }
            for (int b = 0; b < nslices_[i] / BW_SB; b++)
            {
                const int blk = slice_base_[i] / BW_SB + b;
                L.host_upd[nb++] = {slice_info_[(size_t)blk * BW_SB], blk, 0, 0, 0};
            }
        }
    const int grid = slice_bucket(n, d_.n_slices);
    const int ugrid = (grid + BW_SB - 1) / BW_SB;
    const SolverEntry pad = {make_int4(-1, -1, INT_MIN, -1), -1, 0, 0, 0};
    for (int k = n; k < grid; k++)
        L.host_spmv[k] = pad;
    for (int k = nb; k < ugrid; k++)
        L.host_upd[k] = pad;
    check(hipMemcpyAsync(L.spmv_list, L.host_spmv, grid * sizeof(SolverEntry),
                         hipMemcpyHostToDevice, L.stream),
          "hipMemcpyAsync list");
    check(hipMemcpyAsync(L.upd_list, L.host_upd, ugrid * sizeof(SolverEntry),
                         hipMemcpyHostToDevice, L.stream),
          "hipMemcpyAsync list");
    return grid;
}

// The cloths that solve in lanes of their own this step: with the previous
// solve's launch counts known and enough cloths, the ~1/10 of them with the
// most launches (at least one) form lane 0 on the batch's stream, the next
// ~1/3 lane 1 and the rest lane 2, each lane a chain of launches on its own
// stream (BW_LANE_SPLIT="a,b" sets the first two lanes' sizes; on batch64,
// 6 and 20 cloths ran 543 cloth-steps/s, against 532-539 with 1-3, 2-6,
// 3-10, 4-16, 6-30 or 10-20, 518-527 with 8-24 or 16-24, and 516 in one
// chain).  In
// one chain the stiffest cloths' iterations would wait on every launch of
// the others while all of them solve; in lanes of their own they run at the
// pace of their own launches, alongside the others' bandwidth-bound ones,
// and a step ends when its slowest lane does.  Each cloth's arithmetic is
// its own, so results are bit-identical to one chain.  One chain when
// BW_LANES=0, on the first solve, for fewer than 8 cloths, with group sums
// (whose launch covers every cloth's groups) and in bw_batch_step_noalloc
// (lanes 1 and 2 allocate streams and buffers on first use).  Returns the
// number of lanes.
int Batch::plan_lanes()
{
    const char *env = std::getenv("BW_LANES");
    const bool lanes_rule = env ? std::atoi(env) != 0 : true;
    const int n = size();
    lanes_[0].members.clear();
    if (!lanes_rule || !lanes_allowed_ || n < 8 || d_.n_groups > 0 ||
        (int)pred_launches_.size() != n)
    {
        for (int i = 0; i < n; i++)
            lanes_[0].members.push_back(i);
        return 1;
    }
    std::vector<int> &order = lane_order_;
    order.resize(n);
    for (int i = 0; i < n; i++)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return pred_launches_[a] > pred_launches_[b]; });
    int a = std::max(1, n / 10), b = std::max(1, n / 3);
    if (const char *e = std::getenv("BW_LANE_SPLIT")) // "a,b": lane sizes (tuning)
    {
        int ea = 0, eb = 0;
        if (std::sscanf(e, "%d,%d", &ea, &eb) == 2 && ea >= 1 && eb >= 1 && ea + eb < n)
        {
            a = ea;
            b = eb;
        }
    }
    while ((int)lanes_.size() < 3)
    {
        lanes_.emplace_back();
        open_lane(lanes_.back());
    }
    for (int k = 1; k < 3; k++)
        lanes_[k].members.clear();
    for (int r = 0; r < n; r++)
        lanes_[r < a ? 0 : r < a + b ? 1 : 2].members.push_back(order[r]);
    for (auto &L : lanes_)
        std::sort(L.members.begin(), L.members.end());
    return 3;
}

// stream, event and buffers of a lane beyond the first.  Its solver
// workgroups reserve LDS they do not use (BW_LANE_LDS bytes, default 24 KiB,
// 0 for none): fewer of them then fit on a CU at once, and the wave slots
// left over go to lane 0, whose few long-running cloths are bound by the
// latency of each launch rather than by bandwidth.
void Batch::open_lane(Lane &L)
{
    L.lds_reserve = 24576;
    if (const char *e = std::getenv("BW_LANE_LDS"))
        L.lds_reserve = std::min(std::max(std::atoi(e), 0), 32768);
    check(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming), "hipEventCreate");
    L.owned = true;
    L.spmv_list = dalloc<SolverEntry>(d_.n_slices);
    L.upd_list = dalloc<SolverEntry>(d_.n_blocks_wg);
    check(hipHostMalloc((void **)&L.host_spmv, sizeof(SolverEntry) * (size_t)d_.n_slices),
          "hipHostMalloc");
    check(hipHostMalloc((void **)&L.host_upd, sizeof(SolverEntry) * (size_t)d_.n_blocks_wg),
          "hipHostMalloc");
    check(hipHostMalloc((void **)&L.host_cg, 2 * sizeof(CgState) * models_.size()),
          "hipHostMalloc");
}

void Batch::close_lane(Lane &L)
{
    if (!L.owned)
        return;
    if (L.stream)
        (void)hipStreamSynchronize(L.stream);
    for (auto &g : L.graphs)
        (void)hipGraphExecDestroy(g.exec);
    L.graphs.clear();
    if (L.host_spmv)
        (void)hipHostFree(L.host_spmv);
    if (L.host_upd)
        (void)hipHostFree(L.host_upd);
    if (L.host_cg)
        (void)hipHostFree(L.host_cg);
    if (L.ev)
        (void)hipEventDestroy(L.ev);
    if (L.stream)
        (void)hipStreamDestroy(L.stream);
    L = Lane();
}

// the next chunk of a lane: its length from the lane's schedule, then the
// launches and the copy of the states after them
void Batch::lane_chunk(Lane &L, bool sync)
{
    while (L.next < L.bounds.size() && L.bounds[L.next] < L.launched + 32)
        L.next++;
    long long want;
    if (L.next < L.bounds.size())
        want = L.bounds[L.next] - L.launched;
    else
        want = 64LL << std::min(L.blind++, 4);
    const int length = chunk_for(want);
    run_chunk(L, length, upload_lists(L));
    L.launched += length;
    read_states(L, sync);
}

// after a lane's chunk: which of its cloths still solve; false when none
bool Batch::lane_update(Lane &L, long long cap)
{
    bool any = false;
    for (int i : L.members)
    {
        L.active[i] = L.host_cg[2 * i].done == 0;
        any |= L.active[i] != 0;
    }
    if (any && L.launched > cap)
        throw std::runtime_error("step: solver did not terminate");
    return any;
}

void Batch::solve_cg()
{
    launch_cg_reset(d_, stream_);
    long long cap = 16;
    for (auto &m : models_)
        cap = std::max<long long>(cap, 4LL * m.cfg.solver_max_iters + 64);
    const int nl = plan_lanes();
    for (int k = 0; k < nl; k++)
    {
        Lane &L = lanes_[k];
        chunk_schedule(L);
        L.active.assign(size(), 0);
        for (int i : L.members)
            L.active[i] = 1;
        L.next = 0;
        L.launched = 0;
        L.blind = 0;
        L.busy = false;
        L.done = L.members.empty();
    }
    if (nl == 1)
    {
        Lane &L = lanes_[0];
        for (;;)
        {
            lane_chunk(L, true);
            if (!lane_update(L, cap))
                break;
        }
    }
    else
    {
        // the other lanes start after the reset (and all work before it) on
        // the batch's stream; the batch's stream continues after all lanes
        check(hipEventRecord(lanes_start_, stream_), "hipEventRecord");
        for (int k = 1; k < nl; k++)
            check(hipStreamWaitEvent(lanes_[k].stream, lanes_start_, 0), "hipStreamWaitEvent");
        // on an error, no lane's queued work may outlive this call
        struct Drain
        {
            Batch *b;
            int nl;
            bool armed = true;
            ~Drain()
            {
                if (armed)
                    for (int k = 0; k < nl; k++)
                        (void)hipStreamSynchronize(b->lanes_[k].stream);
            }
        } drain{this, nl};
        for (;;)
        {
            bool all_done = true, launched = false;
            for (int k = 0; k < nl; k++)
            {
                Lane &L = lanes_[k];
                if (L.done)
                    continue;
                all_done = false;
                if (L.busy)
                {
                    const hipError_t q = hipEventQuery(L.ev);
                    if (q == hipErrorNotReady)
                        continue;
                    check(q, "hipEventQuery");
                    L.busy = false;
                    if (!lane_update(L, cap))
                    {
                        L.done = true;
                        continue;
                    }
                }
                lane_chunk(L, false);
                L.busy = true;
                launched = true;
            }
            if (all_done)
                break;
            if (!launched)
                std::this_thread::yield();
        }
        for (int k = 1; k < nl; k++)
            check(hipStreamWaitEvent(stream_, lanes_[k].ev, 0), "hipStreamWaitEvent");
        drain.armed = false;
    }
    pred_launches_.assign(size(), 0);
    for (int k = 0; k < nl; k++)
        for (int i : lanes_[k].members)
        {
            const CgState &s = lanes_[k].host_cg[2 * i];
            StepStats &st = stats_[i];
            st.pcg_iterations = s.total;
            st.outer_iterations = s.updates;
            st.matrix_fp64 = s.w64;
            st.rel_residual = s.bb > 0 ? std::sqrt(s.rr_true / s.bb) : std::sqrt(s.rr_true);
            // launches this cloth needed: two per iteration and per true residual
            pred_launches_[i] = 2LL * (s.total + s.updates) + 2;
        }
}

void Batch::begin_step()
{
    if (kind_ == SolverKind::RefPCG64 || kind_ == SolverKind::GpuForm)
        throw std::runtime_error("batch: reference-semantics and fem_gpu batches step through step()");
    launch_elements(d_, stream_);
    launch_assemble_rows(d_, stream_);
    launch_cg_reset(d_, stream_);
    check(hipGetLastError(), "step launch");
}

void Batch::end_step(const CgState *states, double step_us)
{
    bool bad = false;
    for (int i = 0; i < size(); i++)
    {
        const CgState &s = states[2 * i];
        StepStats &st = stats_[i];
        st.pcg_iterations = s.total;
        st.outer_iterations = s.updates;
        st.matrix_fp64 = s.w64;
        st.rel_residual = s.bb > 0 ? std::sqrt(s.rr_true / s.bb) : std::sqrt(s.rr_true);
        st.step_us = step_us;
        bad |= !std::isfinite(st.rel_residual);
    }
    launch_update_positions(d_, stream_);
    check(hipGetLastError(), "step launch");
    if (bad)
        throw std::runtime_error("step: non-finite residual (NaN/Inf in the state)");
}

void Batch::fixed_variant(int &lpr, int &stage, int &nt) const
{
    lpr = 0;
    stage = -1;
    nt = -1;
    if (const char *e = std::getenv("BW_SPMV_LPR"))
    {
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4)
            lpr = v;
    }
    if (const char *e = std::getenv("BW_SPMV_STAGE"))
        stage = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("BW_SPMV_NT"))
        nt = std::atoi(e) != 0 ? 1 : 0;
    if (d_.halo_max == 0)
        stage = 0;
}

// Picks the SpMV variant (lanes per row; directions gathered from memory or
// staged in LDS; matrix loads with or without the nontemporal hint; see
// cg.hip) that runs this batch's matrix fastest (the first in a fixed order
// within 3% of the fastest), timing each on the system assembled at the
// current state.  BW_SPMV_LPR=1|2|4, BW_SPMV_STAGE=0|1 and
// BW_SPMV_NT=0|1 in the environment fix those choices instead.
int Batch::autotune()
{
    if (kind_ == SolverKind::RefPCG64 || kind_ == SolverKind::GpuForm)
        return d_.spmv_lpr;
    clear_graphs(); // graphs hold the previous choice
    tuned_ = true;
    int fixed_lpr = 0, fixed_stage = -1, fixed_nt = -1;
    fixed_variant(fixed_lpr, fixed_stage, fixed_nt);
    double best = 1e300;
    int pick = 0;
    const int choices[3] = {1, 2, 4};
    // variant v: lanes choices[v % 3], staged (v / 3) % 2, nontemporal v / 6
    auto set = [&](int v) {
        d_.spmv_lpr = choices[v % 3];
        d_.spmv_stage = (v / 3) % 2;
        d_.spmv_nt = v / 6;
    };
    int candidates = 0;
    for (int v = 0; v < TUNE_VARIANTS; v++)
    {
        tune_us_[v] = 0.0;
        if ((fixed_lpr && choices[v % 3] != fixed_lpr) ||
            (fixed_stage >= 0 && (v / 3) % 2 != fixed_stage) || (fixed_nt >= 0 && v / 6 != fixed_nt))
            continue;
        pick = v;
        candidates++;
    }
    if (candidates > 1)
        for (int v = 0; v < TUNE_VARIANTS; v++)
        {
            if ((fixed_lpr && choices[v % 3] != fixed_lpr) ||
                (fixed_stage >= 0 && (v / 3) % 2 != fixed_stage) ||
                (fixed_nt >= 0 && v / 6 != fixed_nt))
                continue;
            set(v);
            tune_us_[v] = time_solver_kernel(20, nullptr);
            best = std::min(best, tune_us_[v]);
        }
    // the first variant, in the order above (fewer lanes, gathered, cached
    // first), within 3% of the fastest: variants that time alike would
    // otherwise swap from run to run on timing noise, and the pick sets the
    // rounding of the product, which a stiff cloth's iteration count follows
    if (candidates > 1)
        for (int v = TUNE_VARIANTS - 1; v >= 0; v--)
            if (tune_us_[v] > 0.0 && tune_us_[v] <= 1.03 * best)
                pick = v;
    set(pick);
    return d_.spmv_lpr;
}

// ---------------------------------------------------------------------------
// contacts

void Batch::set_point_triangle_contacts(int i, int n, const int *ids, const double *normal,
                                        const double *bary, const float *depth)
{
    if (i < 0 || i >= size())
        throw std::invalid_argument("contacts: bad cloth index");
    if (n < 0 || (n > 0 && (!ids || !normal || !bary || !depth)))
        throw std::invalid_argument("contacts: null array or negative count");
    const ClothModel &m = models_[i];
    for (int k = 0; k < n; k++)
    {
        const int flag = ids[3 * k], v = ids[3 * k + 1], t = ids[3 * k + 2];
        if (flag != 0 && flag != 1)
            throw std::invalid_argument("contacts: cloth_is_point must be 0 or 1");
        if (flag ? (v < 0 || v >= m.nv) : (t < 0 || t >= m.nt))
            throw std::invalid_argument("contacts: cloth vertex or triangle out of range");
        for (int a = 0; a < 3; a++)
            if (!std::isfinite(normal[3 * k + a]) || !std::isfinite(bary[3 * k + a]))
                throw std::invalid_argument("contacts: non-finite normal or barycentric coordinate");
        // the reference asserts 3e-3 - depth > 0 (BaraffCloth.cpp:832)
        if (!std::isfinite(depth[k]) || !(3e-3 - (double)depth[k] > 0))
            throw std::invalid_argument("contacts: point-triangle depth must be below 3e-3");
    }
    Contacts &c = contacts_[i];
    contact_total_ -= c.size();
    c.pt_ids.assign(ids, ids + 3 * (size_t)n);
    c.pt_normal.assign(normal, normal + 3 * (size_t)n);
    c.pt_bary.assign(bary, bary + 3 * (size_t)n);
    c.pt_depth.assign(depth, depth + n);
    contact_total_ += c.size();
}

void Batch::set_edge_edge_contacts(int i, int n, const int *ids, const double *normal,
                                   const float *bary, const float *depth)
{
    if (i < 0 || i >= size())
        throw std::invalid_argument("contacts: bad cloth index");
    if (n < 0 || (n > 0 && (!ids || !normal || !bary || !depth)))
        throw std::invalid_argument("contacts: null array or negative count");
    const ClothModel &m = models_[i];
    for (int k = 0; k < n; k++)
    {
        const int flag = ids[3 * k];
        if (flag != 0 && flag != 1)
            throw std::invalid_argument("contacts: cloth_is_edge0 must be 0 or 1");
        const int e = flag ? ids[3 * k + 1] : ids[3 * k + 2];
        if (e < 0 || e >= m.ne)
            throw std::invalid_argument("contacts: cloth edge out of range");
        for (int a = 0; a < 3; a++)
            if (!std::isfinite(normal[3 * k + a]))
                throw std::invalid_argument("contacts: non-finite normal");
        if (!std::isfinite(bary[2 * k]) || !std::isfinite(bary[2 * k + 1]) ||
            !std::isfinite(depth[k]))
            throw std::invalid_argument("contacts: non-finite barycentric coordinate or depth");
    }
    Contacts &c = contacts_[i];
    contact_total_ -= c.size();
    c.ee_ids.assign(ids, ids + 3 * (size_t)n);
    c.ee_normal.assign(normal, normal + 3 * (size_t)n);
    c.ee_bary.assign(bary, bary + 2 * (size_t)n);
    c.ee_depth.assign(depth, depth + n);
    contact_total_ += c.size();
}

// ---------------------------------------------------------------------------
// user inputs (cBaseCloth::ApplyPerturb, cBaraffCloth::ApplyUserPerturbForceOnce,
// cBaseCloth::mYDragForce)

namespace
{
void check_plain(SolverKind k)
{
    if (k == SolverKind::GpuForm)
        throw std::runtime_error("state: user perturbations of a fem_gpu cloth are not supported "
                                 "(its drag point follows the barycentric argmax, "
                                 "BaraffClothGpu.cpp:266-279)");
}
} // namespace

void Batch::apply_perturb(int i, int tri, const double force[3])
{
    check_plain(kind_);
    const ClothModel &m = models_[i];
    if (tri < 0 || tri >= m.nt)
        throw std::invalid_argument("perturb: triangle out of range");
    for (int a = 0; a < 3; a++)
        if (!std::isfinite(force[a]))
            throw std::invalid_argument("perturb: force has NaN/Inf");
    std::vector<double> &u = user_force_[i];
    if (u.empty())
        u.assign(3 * (size_t)m.nv, 0.0);
    // mUserForce.segment(3 id, 3) += force / 3 for mId0, mId1, mId2
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++)
        {
            double &f = u[3 * (size_t)m.tri[tri][k] + a];
            f = f + force[a] / 3;
        }
}

void Batch::apply_user_perturb_force_once(int i, int tri, const double goal[3])
{
    check_plain(kind_);
    const ClothModel &m = models_[i];
    if (tri < 0 || tri >= m.nt)
        throw std::invalid_argument("perturb: triangle out of range");
    for (int a = 0; a < 3; a++)
        if (!std::isfinite(goal[a]))
            throw std::invalid_argument("perturb: goal has NaN/Inf");
    const int v0 = m.tri[tri][0];
    const size_t at = 3 * ((size_t)row_base_[i] + row_of_[i][v0]);
    check(hipMemcpy(d_.x + at, goal, 3 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy goal");
    check(hipMemcpy(d_.xpre + at, goal, 3 * sizeof(double), hipMemcpyHostToDevice),
          "hipMemcpy goal");
    drag_vertex_[i] = v0;
}

void Batch::set_y_drag_force(int i, float f)
{
    check_plain(kind_);
    if (!std::isfinite(f))
        throw std::invalid_argument("perturb: drag force has NaN/Inf");
    ydrag_[i] = f;
}

void Batch::clear_user_force(int i) { user_force_[i].clear(); }

// The step's user force per row: the pending perturbation forces, then the y
// drag force (ApplyYDragForce runs inside UpdatePos, after them): +F/n_up on
// every vertex whose float v is within 1e-5 of the largest v, -F/n_low on the
// smallest, F/n in float.  Rows of cloths without inputs are zero.  The drag
// vertex's row is marked pinned for the step (its row and column leave the
// solve; its rhs and warm start become zero).
void Batch::prepare_user_inputs()
{
    const uint8_t one = 1;
    for (int i = 0; i < size(); i++)
    {
        const ClothModel &m = models_[i];
        const bool has = !user_force_[i].empty() || ydrag_[i] != 0.0f;
        const size_t rows = (size_t)nslices_[i] * SLICE;
        if (has)
        {
            std::vector<double> u = user_force_[i];
            if (u.empty())
                u.assign(3 * (size_t)m.nv, 0.0);
            if (ydrag_[i] != 0.0f)
            {
                double vmax = -1e5, vmin = 1e5;
                for (int v = 0; v < m.nv; v++)
                {
                    vmax = std::max(vmax, (double)m.uv[2 * v + 1]);
                    vmin = std::min(vmin, (double)m.uv[2 * v + 1]);
                }
                int n_up = 0, n_low = 0;
                for (int v = 0; v < m.nv; v++)
                {
                    n_up += std::fabs(vmax - (double)m.uv[2 * v + 1]) < 1e-5;
                    n_low += std::fabs(vmin - (double)m.uv[2 * v + 1]) < 1e-5;
                }
                const float up = ydrag_[i] / (float)n_up, low = -ydrag_[i] / (float)n_low;
                for (int v = 0; v < m.nv; v++)
                    if (std::fabs(vmax - (double)m.uv[2 * v + 1]) < 1e-5)
                        u[3 * (size_t)v + 1] = u[3 * (size_t)v + 1] + (double)up;
                for (int v = 0; v < m.nv; v++)
                    if (std::fabs(vmin - (double)m.uv[2 * v + 1]) < 1e-5)
                        u[3 * (size_t)v + 1] = u[3 * (size_t)v + 1] + (double)low;
            }
            if (!d_.fuser)
                d_.fuser = dalloc<double>(3 * (size_t)d_.n_rows);
            std::vector<double> r(3 * rows, 0.0);
            for (int v = 0; v < m.nv; v++)
                for (int a = 0; a < 3; a++)
                    r[3 * (size_t)row_of_[i][v] + a] = u[3 * (size_t)v + a];
            check(hipMemcpy(d_.fuser + 3 * (size_t)row_base_[i], r.data(), r.size() * sizeof(double),
                            hipMemcpyHostToDevice),
                  "hipMemcpy user force");
            fuser_dirty_[i] = 1;
        }
        else if (fuser_dirty_[i])
        {
            check(hipMemset(d_.fuser + 3 * (size_t)row_base_[i], 0, 3 * rows * sizeof(double)),
                  "hipMemset user force");
            check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            fuser_dirty_[i] = 0;
        }
        const int dv = drag_vertex_[i];
        if (dv >= 0 && !m.pinned[dv])
        {
            check(hipMemcpy(d_.row_pinned + row_base_[i] + row_of_[i][dv], &one, 1,
                            hipMemcpyHostToDevice),
                  "hipMemcpy drag");
            drag_pinned_[i] = 1;
        }
    }
}

void Batch::finish_user_inputs()
{
    const uint8_t zero = 0;
    for (int i = 0; i < size(); i++)
    {
        if (drag_pinned_[i])
        {
            check(hipMemcpy(d_.row_pinned + row_base_[i] + row_of_[i][drag_vertex_[i]], &zero, 1,
                            hipMemcpyHostToDevice),
                  "hipMemcpy drag");
            drag_pinned_[i] = 0;
        }
        drag_vertex_[i] = -1;
        user_force_[i].clear();
    }
}

void Batch::clear_contacts()
{
    if (contact_total_ == 0)
        return;
    for (auto &c : contacts_)
        c = Contacts();
    contact_total_ = 0;
}

namespace
{
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
} // namespace

// Host side of the collision response (the arithmetic of the records is the
// reference's, see contact.hip):
//  - collision Hessian blocks 1e-2 dC_i dC_j^T per point-triangle contact
//    (BaraffCloth.cpp:928-990), summed per entry in contact order as
//    setFromTriplets does; blocks in pinned rows or columns are left out
//    (the solve skips them, :469-506);
//  - one repulsion record per contact, edge-edge contacts first, then
//    point-triangle ones (:626-926), with its level for k_repulsion.
bool Batch::upload_contacts()
{
    if (contact_total_ == 0)
        return false;
    std::vector<long long> widx;
    std::vector<int> diag;
    std::vector<double> h9;
    std::vector<int> wg_chunk(1, 0), chunk_level(1, 0), level_ptr(1, 0), wg_rows(1, 0),
        touched_rows;
    std::vector<int4> ids;
    std::vector<double> rec;
    std::vector<int> last_level;
    for (int i = 0; i < size(); i++)
    {
        const Contacts &c = contacts_[i];
        if (c.size() == 0)
            continue;
        const ClothModel &m = models_[i];
        auto row = [&](int v) { return row_base_[i] + row_of_[i][v]; };
        // collision Hessian
        const int n_pt = (int)c.pt_depth.size();
        if (n_pt > 0)
        {
            std::unordered_map<long long, int> at;
            std::vector<double> hb;        // 9 per block
            std::vector<uint8_t> touched; // 9 per block
            std::vector<std::pair<int, int>> key;
            for (int k = 0; k < n_pt; k++)
            {
                const double *N = &c.pt_normal[3 * (size_t)k];
                double dc[3][3];
                int vid[3], nloc;
                if (c.pt_ids[3 * k])
                {
                    nloc = 1;
                    vid[0] = c.pt_ids[3 * k + 1];
                    for (int a = 0; a < 3; a++)
                        dc[0][a] = N[a];
                }
                else
                {
                    nloc = 3;
                    const auto &t = m.tri[c.pt_ids[3 * k + 2]];
                    for (int l = 0; l < 3; l++)
                    {
                        vid[l] = t[l];
                        for (int a = 0; a < 3; a++)
                            dc[l][a] = -c.pt_bary[3 * (size_t)k + l] * N[a];
                    }
                }
                for (int li = 0; li < nloc; li++)
                    for (int lj = 0; lj < nloc; lj++)
                    {
                        const long long kk = (long long)vid[li] * m.nv + vid[lj];
                        auto it = at.find(kk);
                        int blk;
                        if (it == at.end())
                        {
                            blk = (int)key.size();
                            at.emplace(kk, blk);
                            key.emplace_back(vid[li], vid[lj]);
                            hb.resize(hb.size() + 9, 0.0);
                            touched.resize(touched.size() + 9, 0);
                        }
                        else
                            blk = it->second;
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++)
                            {
                                const double h = 1e-2 * (dc[li][a] * dc[lj][b]);
                                const size_t q = 9 * (size_t)blk + 3 * a + b;
                                hb[q] = touched[q] ? hb[q] + h : h;
                                touched[q] = 1;
                            }
                    }
            }
            for (size_t blk = 0; blk < key.size(); blk++)
            {
                const int v = key[blk].first, col = key[blk].second;
                // the drag vertex is fixed in this step's solve like a pin
                const int dv = drag_vertex_[i];
                if (m.pinned[v] || m.pinned[col] || v == dv || col == dv)
                    continue;
                int ks = -1;
                for (int k = m.row_ptr[v]; k < m.row_ptr[v + 1]; k++)
                    if (m.cols[k] == col)
                        ks = k - m.row_ptr[v];
                if (ks < 0)
                    throw std::runtime_error("contacts: collision Hessian block outside the pattern");
                const int p = row(v), sl = p / SLICE, lane = p % SLICE;
                widx.push_back(9LL * host_slice_off_[sl] + 9LL * ks * SLICE + lane);
                diag.push_back(v == col ? p : -1);
                h9.insert(h9.end(), hb.begin() + 9 * blk, hb.begin() + 9 * blk + 9);
            }
        }
        // repulsion records and levels
        const double dt = m.cfg.dt;
        const float kf = (m.cfg.stretch_k[0] + (m.cfg.stretch_k[1] + m.cfg.stretch_k[2])) / 3.0f;
        const double dtK = dt * (double)(kf * 1);
        std::vector<int> lvl;
        std::vector<int4> cid;
        std::vector<double> crec;
        last_level.assign(m.nv, -1);
        // slots: this cloth's touched rows in order of first use
        std::unordered_map<int, int> slot_of;
        const int row0 = (int)touched_rows.size();
        auto slot = [&](int v) {
            auto it = slot_of.find(v);
            if (it != slot_of.end())
                return it->second;
            const int sl = (int)touched_rows.size() - row0;
            slot_of.emplace(v, sl);
            touched_rows.push_back(row(v));
            return sl;
        };
        auto add = [&](int type, const int *verts, int nvert, const double *r) {
            int L = 0;
            for (int l = 0; l < nvert; l++)
                L = std::max(L, last_level[verts[l]] + 1);
            for (int l = 0; l < nvert; l++)
                last_level[verts[l]] = L;
            lvl.push_back(L);
            int sl[3] = {0, 0, 0};
            for (int l = 0; l < nvert; l++)
                sl[l] = slot(verts[l]);
            cid.push_back(make_int4(type, sl[0], sl[1], sl[2]));
            crec.insert(crec.end(), r, r + CONTACT_REC);
        };
        const int n_ee = (int)c.ee_depth.size();
        for (int k = 0; k < n_ee; k++)
        {
            const bool cloth0 = c.ee_ids[3 * k] != 0;
            const int e = cloth0 ? c.ee_ids[3 * k + 1] : c.ee_ids[3 * k + 2];
            const float b0 = c.ee_bary[2 * k], b1 = c.ee_bary[2 * k + 1];
            const float bc = cloth0 ? b0 : b1;
            const int v[2] = {m.edges[e].v0, m.edges[e].v1};
            const float den = 1 + b0 * b0 + (1 - b0) * (1 - b0) + b1 * b1 + (1 - b1) * (1 - b1);
            const double r[CONTACT_REC] = {
                c.ee_normal[3 * k], c.ee_normal[3 * k + 1], c.ee_normal[3 * k + 2],
                (double)(1 - bc), (double)bc, 0.0,
                m.mass[v[0]] + m.mass[v[1]], (double)den, 3e-3 - (double)c.ee_depth[k],
                (double)0.1f, dtK, dt};
            add(0, v, 2, r);
        }
        for (int k = 0; k < n_pt; k++)
        {
            const double *bary = &c.pt_bary[3 * (size_t)k];
            const double den = 1 + ((bary[0] * bary[0] + bary[1] * bary[1]) + bary[2] * bary[2]);
            const double d = 3e-3 - (double)c.pt_depth[k];
            if (c.pt_ids[3 * k])
            {
                const int v = c.pt_ids[3 * k + 1];
                const double r[CONTACT_REC] = {c.pt_normal[3 * k], c.pt_normal[3 * k + 1],
                                               c.pt_normal[3 * k + 2], 0.0, 0.0, 0.0,
                                               m.mass[v], den, d, 0.1, dtK, dt};
                add(1, &v, 1, r);
            }
            else
            {
                const auto &t = m.tri[c.pt_ids[3 * k + 2]];
                const int v[3] = {t[0], t[1], t[2]};
                const double r[CONTACT_REC] = {c.pt_normal[3 * k], c.pt_normal[3 * k + 1],
                                               c.pt_normal[3 * k + 2], bary[0], bary[1], bary[2],
                                               m.mass[v[0]] + m.mass[v[1]] + m.mass[v[2]], den, d,
                                               0.1, dtK, dt};
                add(2, v, 3, r);
            }
        }
        // group by level, contact order kept within a level
        const int nl = lvl.empty() ? 0 : *std::max_element(lvl.begin(), lvl.end()) + 1;
        std::vector<int> count(nl + 1, 0);
        for (int L : lvl)
            count[L + 1]++;
        for (int L = 0; L < nl; L++)
            count[L + 1] += count[L];
        const int base = (int)ids.size();
        ids.resize(base + cid.size());
        rec.resize(rec.size() + crec.size());
        std::vector<int> fill(count.begin(), count.end() - 1);
        for (size_t k = 0; k < cid.size(); k++)
        {
            const int at = base + fill[lvl[k]]++;
            ids[at] = cid[k];
            std::copy(crec.begin() + CONTACT_REC * k, crec.begin() + CONTACT_REC * (k + 1),
                      rec.begin() + CONTACT_REC * (size_t)at);
        }
        // levels, then chunks of whole levels with at most REPULSION_CHUNK
        // contacts (a wider level alone)
        const int lbase = (int)level_ptr.size() - 1;
        for (int L = 0; L < nl; L++)
            level_ptr.push_back(base + count[L + 1]);
        for (int L = 0; L < nl;)
        {
            int L1 = L + 1;
            while (L1 < nl && count[L1 + 1] - count[L] <= REPULSION_CHUNK)
                L1++;
            chunk_level.push_back(lbase + L1);
            L = L1;
        }
        wg_chunk.push_back((int)chunk_level.size() - 1);
        wg_rows.push_back((int)touched_rows.size());
    }
    contact_blocks_ = (int)widx.size();
    contact_groups_ = (int)wg_chunk.size() - 1;
    // one staging buffer, one device buffer, one copy
    size_t off = 0;
    const size_t o_widx = off;
    off = align256(off + widx.size() * sizeof(long long));
    const size_t o_diag = off;
    off = align256(off + diag.size() * sizeof(int));
    const size_t o_h9 = off;
    off = align256(off + h9.size() * sizeof(double));
    const size_t o_wgl = off;
    off = align256(off + wg_chunk.size() * sizeof(int));
    const size_t o_chl = off;
    off = align256(off + chunk_level.size() * sizeof(int));
    const size_t o_lvl = off;
    off = align256(off + level_ptr.size() * sizeof(int));
    const size_t o_wgr = off;
    off = align256(off + wg_rows.size() * sizeof(int));
    const size_t o_rows = off;
    off = align256(off + touched_rows.size() * sizeof(int));
    const size_t o_ids = off;
    off = align256(off + ids.size() * sizeof(int4));
    const size_t o_rec = off;
    off = align256(off + rec.size() * sizeof(double));
    contact_host_.assign(off, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes)
            std::memcpy(contact_host_.data() + o, src, bytes);
    };
    put(o_widx, widx.data(), widx.size() * sizeof(long long));
    put(o_diag, diag.data(), diag.size() * sizeof(int));
    put(o_h9, h9.data(), h9.size() * sizeof(double));
    put(o_wgl, wg_chunk.data(), wg_chunk.size() * sizeof(int));
    put(o_chl, chunk_level.data(), chunk_level.size() * sizeof(int));
    put(o_lvl, level_ptr.data(), level_ptr.size() * sizeof(int));
    put(o_wgr, wg_rows.data(), wg_rows.size() * sizeof(int));
    put(o_rows, touched_rows.data(), touched_rows.size() * sizeof(int));
    put(o_ids, ids.data(), ids.size() * sizeof(int4));
    put(o_rec, rec.data(), rec.size() * sizeof(double));
    if (off > contact_cap_)
    {
        // the previous step's kernels have finished (step() synchronizes)
        check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        if (contact_dev_)
            check(hipFree(contact_dev_), "hipFree");
        contact_dev_ = nullptr;
        contact_cap_ = 0;
        check(hipMalloc(&contact_dev_, off), "hipMalloc contacts");
        contact_cap_ = off;
    }
    check(hipMemcpyAsync(contact_dev_, contact_host_.data(), off, hipMemcpyHostToDevice, stream_),
          "hipMemcpyAsync contacts");
    char *b = static_cast<char *>(contact_dev_);
    c_widx_ = reinterpret_cast<const long long *>(b + o_widx);
    c_diag_ = reinterpret_cast<const int *>(b + o_diag);
    c_h9_ = reinterpret_cast<const double *>(b + o_h9);
    c_wgl_ = reinterpret_cast<const int *>(b + o_wgl);
    c_chl_ = reinterpret_cast<const int *>(b + o_chl);
    c_lvl_ = reinterpret_cast<const int *>(b + o_lvl);
    c_wgr_ = reinterpret_cast<const int *>(b + o_wgr);
    c_rows_ = reinterpret_cast<const int *>(b + o_rows);
    c_ids_ = reinterpret_cast<const int4 *>(b + o_ids);
    c_rec_ = reinterpret_cast<const double *>(b + o_rec);
    return true;
}

namespace
{
// grow-only device scratch (the caller has synchronized the stream)
void *ensure(void *&p, size_t &cap, size_t bytes)
{
    if (bytes > cap)
    {
        if (p)
            check(hipFree(p), "hipFree");
        p = nullptr;
        cap = 0;
        check(hipMalloc(&p, bytes), "hipMalloc detection");
        cap = bytes;
    }
    return p;
}
} // namespace

void Batch::detect_contacts(int i, const Body &body, bool cloth_first, int *n_pt, int *n_ee)
{
    if (i < 0 || i >= size())
        throw std::invalid_argument("detect: bad cloth index");
    const ClothModel &m = models_[i];
    if (!vrow_dev_)
    {
        std::vector<int> vrow(std::max(1, d_.n_vert)), erow;
        edge_base_.assign(size(), 0);
        for (int k = 0; k < size(); k++)
        {
            for (int v = 0; v < models_[k].nv; v++)
                vrow[vert_base_[k] + v] = row_base_[k] + row_of_[k][v];
            edge_base_[k] = (int)erow.size() / 2;
            for (const auto &e : models_[k].edges)
            {
                erow.push_back(row_base_[k] + row_of_[k][e.v0]);
                erow.push_back(row_base_[k] + row_of_[k][e.v1]);
            }
        }
        vrow_dev_ = dalloc<int>(vrow.size());
        upload(vrow_dev_, vrow);
        erow_dev_ = dalloc<int>(std::max<size_t>(erow.size(), 1));
        upload(erow_dev_, erow);
    }
    if (!normals_dev_)
        normals_dev_ = dalloc<double>(3 * ((size_t)d_.n_tri + (size_t)d_.n_vert));
    double *tn_dev = normals_dev_, *vn_dev = normals_dev_ + 3 * (size_t)d_.n_tri;
    // the cloth's triangle normals at its current positions
    launch_normals(d_, tn_dev, vn_dev, stream_);
    ClothView cv;
    cv.nv = m.nv;
    cv.nt = m.nt;
    cv.ne = m.ne;
    cv.vbase = vert_base_[i];
    cv.tbase = tri_base_[i];
    cv.vrow = vrow_dev_ + vert_base_[i];
    cv.vt_ptr = d_.vt_ptr;
    cv.tri_rows = d_.tri_rows;
    cv.edge_rows = erow_dev_ + 2 * (size_t)edge_base_[i];
    cv.X = d_.x;
    cv.tn = tn_dev;
    // pass 1: counts per cloth vertex, triangle and edge
    const size_t N = (size_t)m.nv + m.nt + m.ne;
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    int *cnt = static_cast<int *>(ensure(detect_count_dev_, detect_count_cap_, 2 * N * sizeof(int)));
    int *c1 = cnt, *c2 = cnt + m.nv, *c3 = cnt + m.nv + m.nt;
    launch_detect_points(cv, body, c1, nullptr, nullptr, stream_);
    launch_detect_body_points(cv, body, c2, nullptr, nullptr, stream_);
    launch_detect_edges(cv, body, cloth_first ? 1 : 0, c3, nullptr, nullptr, stream_);
    check(hipGetLastError(), "detect launch");
    check(hipStreamSynchronize(stream_), "detect sync");
    std::vector<int> counts(N), offs(N);
    check(hipMemcpy(counts.data(), cnt, N * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy counts");
    // offsets: one exclusive scan per kernel, kernels back to back
    long long run = 0;
    size_t tot[3] = {0, 0, 0};
    const size_t seg[4] = {0, (size_t)m.nv, (size_t)m.nv + m.nt, N};
    for (int k = 0; k < 3; k++)
    {
        const long long start = run;
        for (size_t q = seg[k]; q < seg[k + 1]; q++)
        {
            offs[q] = (int)(run - start);
            run += counts[q];
        }
        tot[k] = (size_t)(run - start);
    }
    const size_t total = tot[0] + tot[1] + tot[2];
    std::vector<Hit> hits(total);
    if (total > 0)
    {
        if (total > (size_t)INT32_MAX)
            throw std::runtime_error("capacity: too many contacts");
        int *off_dev = cnt + N;
        check(hipMemcpy(off_dev, offs.data(), N * sizeof(int), hipMemcpyHostToDevice),
              "hipMemcpy offsets");
        Hit *h = static_cast<Hit *>(ensure(detect_hit_dev_, detect_hit_cap_, total * sizeof(Hit)));
        launch_detect_points(cv, body, nullptr, off_dev, h, stream_);
        launch_detect_body_points(cv, body, nullptr, off_dev + m.nv, h + tot[0], stream_);
        launch_detect_edges(cv, body, cloth_first ? 1 : 0, nullptr, off_dev + m.nv + m.nt,
                            h + tot[0] + tot[1], stream_);
        check(hipGetLastError(), "detect launch");
        check(hipStreamSynchronize(stream_), "detect sync");
        check(hipMemcpy(hits.data(), h, total * sizeof(Hit), hipMemcpyDeviceToHost),
              "hipMemcpy hits");
    }
    auto by_key = [](const Hit &a, const Hit &b) {
        return a.key0 != b.key0 ? a.key0 < b.key0 : a.key1 < b.key1;
    };
    // cloth-vertex hits come in (vertex, triangle) order; body-vertex hits in
    // (cloth triangle, body vertex) order, so sort those by (vertex, triangle)
    std::vector<Hit> cv_hits(hits.begin(), hits.begin() + tot[0]);
    std::vector<Hit> bv_hits(hits.begin() + tot[0], hits.begin() + tot[0] + tot[1]);
    std::vector<Hit> ee_hits(hits.begin() + tot[0] + tot[1], hits.end());
    std::sort(bv_hits.begin(), bv_hits.end(), by_key);
    if (!cloth_first)
        std::sort(ee_hits.begin(), ee_hits.end(), by_key);
    // obj0's points against obj1's triangles first (BVHCollisionDetecter.cpp:411-413)
    std::vector<Hit> pt;
    pt.reserve(tot[0] + tot[1]);
    const std::vector<Hit> &first = cloth_first ? cv_hits : bv_hits;
    const std::vector<Hit> &second = cloth_first ? bv_hits : cv_hits;
    pt.insert(pt.end(), first.begin(), first.end());
    pt.insert(pt.end(), second.begin(), second.end());
    std::vector<int> pids(3 * pt.size()), eids(3 * ee_hits.size());
    std::vector<double> pn(3 * pt.size()), pb(3 * pt.size()), en(3 * ee_hits.size());
    std::vector<float> pd(pt.size()), eb(2 * ee_hits.size()), ed(ee_hits.size());
    for (size_t k = 0; k < pt.size(); k++)
    {
        pids[3 * k] = pt[k].flag;
        pids[3 * k + 1] = pt[k].key0;
        pids[3 * k + 2] = pt[k].key1;
        for (int a = 0; a < 3; a++)
        {
            pn[3 * k + a] = pt[k].n[a];
            pb[3 * k + a] = (double)pt[k].bary[a];
        }
        pd[k] = pt[k].depth;
    }
    for (size_t k = 0; k < ee_hits.size(); k++)
    {
        eids[3 * k] = ee_hits[k].flag;
        eids[3 * k + 1] = ee_hits[k].key0;
        eids[3 * k + 2] = ee_hits[k].key1;
        for (int a = 0; a < 3; a++)
            en[3 * k + a] = ee_hits[k].n[a];
        eb[2 * k] = ee_hits[k].bary[0];
        eb[2 * k + 1] = ee_hits[k].bary[1];
        ed[k] = ee_hits[k].depth;
    }
    set_point_triangle_contacts(i, (int)pt.size(), pids.data(), pn.data(), pb.data(), pd.data());
    set_edge_edge_contacts(i, (int)ee_hits.size(), eids.data(), en.data(), eb.data(), ed.data());
    if (n_pt)
        *n_pt = (int)pt.size();
    if (n_ee)
        *n_ee = (int)ee_hits.size();
}

// bw_batch_step_noalloc.  The solver states are copied into the scratch
// before the steps (they carry each cloth's FP64-matrix choice from step to
// step) and back after them; the partial and group sums are zeroed first.
// Scratch layout: 2 CgState per cloth | 2 x PARTS doubles per slice | 2 x
// PARTS doubles per group, each part 256-byte aligned.
void Batch::step_noalloc(int nsteps)
{
    if (kind_ != SolverKind::CG32 && kind_ != SolverKind::CG64)
        throw std::runtime_error("state: the no-allocation step covers the tolerance-mode "
                                 "solvers of \"fem\" cloths only");
    if (contact_total_ > 0)
        throw std::runtime_error("state: contacts are pending; they need bw_batch_step, "
                                 "whose contact lists are allocated per step");
    for (int i = 0; i < size(); i++)
        if (!user_force_[i].empty() || drag_vertex_[i] >= 0 || ydrag_[i] != 0.0f)
            throw std::runtime_error("state: user perturbations are set; they need "
                                     "bw_batch_step");
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t st_bytes = align(2 * (size_t)size() * sizeof(CgState));
    const size_t part_bytes = align(2 * (size_t)PARTS * d_.n_slices * sizeof(double));
    const size_t grp_bytes = 2 * (size_t)PARTS * d_.n_groups * sizeof(double);
    if (st_bytes + part_bytes + grp_bytes > NOALLOC_SCRATCH_BYTES)
        throw std::runtime_error("capacity: this batch's solver scratch (" +
                                 std::to_string(st_bytes + part_bytes + grp_bytes) +
                                 " bytes) exceeds the " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                 " bytes of the no-allocation entry point");
    if (!tuned_)
    {
        int lpr, stage, nt;
        fixed_variant(lpr, stage, nt);
        if (lpr)
            d_.spmv_lpr = lpr;
        if (stage >= 0)
            d_.spmv_stage = stage;
        if (nt >= 0)
            d_.spmv_nt = nt;
    }
    char *base = static_cast<char *>(noalloc_scratch());
    CgState *own_cg = d_.cg;
    double *own_partial = d_.cg_partial, *own_group = d_.cg_group;
    const bool own_graphs = use_graphs_;
    struct Restore
    {
        Batch *b;
        CgState *cg;
        double *partial, *group;
        bool graphs;
        ~Restore()
        {
            b->d_.cg = cg;
            b->d_.cg_partial = partial;
            b->d_.cg_group = group;
            b->use_graphs_ = graphs;
            b->lanes_allowed_ = true;
        }
    } restore{this, own_cg, own_partial, own_group, own_graphs};
    const size_t st_copy = 2 * (size_t)size() * sizeof(CgState);
    check(hipMemcpyAsync(base, own_cg, st_copy, hipMemcpyDeviceToDevice, stream_),
          "hipMemcpyAsync states");
    noalloc_clear(base + st_bytes, part_bytes + grp_bytes, stream_);
    d_.cg = reinterpret_cast<CgState *>(base);
    d_.cg_partial = reinterpret_cast<double *>(base + st_bytes);
    d_.cg_group = d_.n_groups > 0 ? reinterpret_cast<double *>(base + st_bytes + part_bytes)
                                  : nullptr;
    use_graphs_ = false;
    lanes_allowed_ = false; // lanes 1 and 2 would allocate on first use
    for (int n = 0; n < nsteps; n++)
    {
        const auto t0 = std::chrono::steady_clock::now();
        launch_elements(d_, stream_);
        launch_assemble_rows(d_, stream_);
        solve_cg();
        launch_update_positions(d_, stream_);
        check(hipGetLastError(), "step launch");
        check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        const double us =
            std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        bool bad = false;
        for (auto &st : stats_)
        {
            st.step_us = us;
            bad |= !std::isfinite(st.rel_residual);
        }
        if (bad)
            throw std::runtime_error("step: non-finite residual (NaN/Inf in the state)");
    }
    check(hipMemcpyAsync(own_cg, base, st_copy, hipMemcpyDeviceToDevice, stream_),
          "hipMemcpyAsync states");
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Batch::step(int nsteps)
{
    if (!tuned_)
        autotune();
    for (int n = 0; n < nsteps; n++)
    {
        const auto t0 = std::chrono::steady_clock::now();
        if (kind_ == SolverKind::GpuForm)
            solve_gpu_form(); // assembly, solve and position update
        else
        {
            prepare_user_inputs();
            const bool ct = upload_contacts();
            launch_elements(d_, stream_);
            launch_assemble_rows(d_, stream_);
            if (ct) // W += collision Hessian (BaraffCloth.cpp:293)
                launch_contact_hessian(d_, contact_blocks_, c_widx_, c_diag_, c_h9_, stream_);
            if (kind_ == SolverKind::RefPCG64)
                solve_ref();
            else
                solve_cg();
            if (ct) // repulsion on dx (BaraffCloth.cpp:147)
                launch_repulsion(d_, contact_groups_, c_wgl_, c_chl_, c_lvl_, c_wgr_, c_rows_,
                                 c_ids_, c_rec_, stream_);
            launch_update_positions(d_, stream_);
        }
        check(hipGetLastError(), "step launch");
        check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        clear_contacts(); // contacts last one substep (BaraffCloth.cpp:156-157)
        if (kind_ != SolverKind::GpuForm)
            finish_user_inputs(); // mUserForce and mDragVertexIdx too (:154-155)
        const double us =
            std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        bool bad = false;
        for (auto &st : stats_)
        {
            st.step_us = us;
            bad |= !std::isfinite(st.rel_residual);
        }
        if (bad)
            throw std::runtime_error("step: non-finite residual (NaN/Inf in the state)");
    }
}

} // namespace bw
