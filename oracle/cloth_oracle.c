/*
 * cloth_oracle.c -- CPU restatement of the reference's plain implicit cloth
 * substep, cBaraffCloth::UpdatePos (src/sim/cloth/BaraffCloth.cpp:90-163).
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle).  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg load this library, as the
 * checker or the timed CPU baseline; the product never does.
 *
 * Arithmetic follows the reference expression by expression, FP64 with the
 * reference's FP32 roundings:
 *   - stretch coefficients a,b,c,d cast to float   BaraffMaterial.cpp:18-32
 *   - triangle areas returned as float             MathUtil2.cpp:344-357
 *   - UV coordinates stored as float               Primitives.h:11, Triangulator.cpp:621-658
 *   - stiffness, density, Rayleigh a/b as float    BaraffCloth.h:25-27, BaseCloth.h:53
 *   - dihedral rest data float                     DihedralBending.h:23-25
 *   - dihedral 12x12 Hessian stored as float       DihedralBending.cpp:437
 * and keeps the reference's behaviour where it is a bug:
 *   - Dv of the stretch Hessian scales with |Fu|   BaraffMaterial.cpp:142
 *   - dihedral force carries /2, Hessian does not  DihedralBending.cpp:425-426
 * Reductions in the solver follow Eigen 3.3.9's SSE2 redux order (four
 * partial sums), see oc_dot().
 *
 * The solver has two modes: the reference's own (exactly max_iters-1 PCG
 * iterations, BaraffCloth.cpp:529) and a tolerance mode (stop when the
 * recursive residual on free dofs is <= tol * ||b_free||), which is the
 * stopping rule of the HIP product.
 */
#include "cloth_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* small vector helpers, written in the reference's evaluation order   */
/* ------------------------------------------------------------------ */

static void v3sub(const double *a, const double *b, double *o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
/* Eigen cross(): (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0) */
static void v3cross(const double *a, const double *b, double *o)
{
    double t0 = a[1] * b[2] - a[2] * b[1];
    double t1 = a[2] * b[0] - a[0] * b[2];
    double t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}
/* fixed-size 3-vector redux: ((x0 + x1) + x2) */
static double v3dot(const double *a, const double *b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static double v3sqnorm(const double *a) { return v3dot(a, a); }
static double v3norm(const double *a) { return sqrt(v3sqnorm(a)); }
/* Eigen normalized(): n / sqrt(|n|^2) if |n|^2 > 0 else n (Dot.h:124-134) */
static void v3normalized(const double *a, double *o)
{
    double z = v3sqnorm(a);
    if (z > 0)
    {
        double s = sqrt(z);
        o[0] = a[0] / s;
        o[1] = a[1] / s;
        o[2] = a[2] / s;
    }
    else
    {
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2];
    }
}

/* Eigen 3.3.9 dot()/sum() of two aligned dynamic double vectors under SSE2:
   redux_impl<..., LinearVectorizedTraversal, NoUnrolling> (Redux.h) keeps two
   2-wide packet accumulators, folds them, then adds the scalar tail. */
double oc_dot(const double *a, const double *b, long n)
{
    const long ps = 2;
    long aligned2 = (n / (2 * ps)) * (2 * ps);
    long aligned = (n / ps) * ps;
    if (aligned == 0)
    {
        if (n == 0)
            return 0.0;
        double r = a[0] * b[0];
        for (long i = 1; i < n; i++)
            r = r + a[i] * b[i];
        return r;
    }
    double p00 = a[0] * b[0], p01 = a[1] * b[1];
    if (aligned > ps)
    {
        double p10 = a[2] * b[2], p11 = a[3] * b[3];
        for (long i = 2 * ps; i < aligned2; i += 2 * ps)
        {
            p00 = p00 + a[i] * b[i];
            p01 = p01 + a[i + 1] * b[i + 1];
            p10 = p10 + a[i + 2] * b[i + 2];
            p11 = p11 + a[i + 3] * b[i + 3];
        }
        p00 = p00 + p10;
        p01 = p01 + p11;
        if (aligned > aligned2)
        {
            p00 = p00 + a[aligned2] * b[aligned2];
            p01 = p01 + a[aligned2 + 1] * b[aligned2 + 1];
        }
    }
    double r = p00 + p01;
    for (long i = aligned; i < n; i++)
        r = r + a[i] * b[i];
    return r;
}

/* ------------------------------------------------------------------ */
/* state                                                               */
/* ------------------------------------------------------------------ */

struct oc_cloth
{
    int nv, nt, ne, npin;
    double *xrest; /* 3nv, after the init transform */
    double *x, *xpre;
    float *uv;     /* 2nv */
    int *tri;      /* 3nt */
    int *tri_e;    /* 3nt edge ids (mEId0..2) */
    int *edge;     /* 5ne: id0, id1, tri0, tri1, boundary */
    int *stencil;  /* 4ne, -1 for boundary edges */
    int *pins;
    unsigned char *pinned;
    float *flat_y; /* nv, untransformed rest y as float */
    /* GPU form (built on first use) */
    float *gf_qb, *gf_mass, *gf_grav, *gf_dx;
    float Ku, Kv, gf_a, gf_b, density_f;
    /* rest data */
    double *mass;    /* nv */
    double *gravity; /* 3nv */
    double *coef;    /* 12nt: cu_ww, cv_ww, cu_bias, cv_bias */
    double *area;    /* nt, float values */
    float *edge_len; /* ne */
    float *heron;    /* nt */
    float *height;   /* 2ne */
    double k_ww[2], k_bias, edge_k;
    float ray_a, ray_b;
    double dt;
    int max_iters;
    /* block pattern */
    int nnzb;
    int *row_ptr, *cols, *diag;
    /* per-step */
    double *Ks, *Kb, *K, *W; /* 9 nnzb each */
    double *fs, *fb, *f, *b, *dx;
    /* contacts for the next step only (cleared by oc_step) */
    int n_pt, n_ee;
    int *pt_ids;              /* 3 n_pt: cloth_is_point, vertex, triangle */
    double *pt_normal, *pt_bary; /* 3 n_pt each */
    float *pt_depth;
    int *ee_ids;              /* 3 n_ee: cloth_is_edge0, edge0, edge1 */
    double *ee_normal;        /* 3 n_ee */
    float *ee_bary, *ee_depth; /* 2 n_ee, n_ee */
    /* user inputs: cBaseCloth::mUserForce of the next step (cleared by
       oc_step, BaraffCloth.cpp:154), cBaraffCloth::mDragVertexIdx (-1: none;
       cleared by oc_step, :155) and cBaseCloth::mYDragForce (kept) */
    double *fuser; /* 3nv */
    int drag;
    float ydrag;
};

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    return p;
}

/* ------------------------------------------------------------------ */
/* geometry: Triangulator.cpp:141-390 (cells), 395-457 (vertices),     */
/* 11-31 (triangle edge ids), 116-131 (transform, UV)                  */
/* ------------------------------------------------------------------ */

typedef struct
{
    int a, b, id;
} epair;
static int epair_cmp(const void *p, const void *q)
{
    const epair *x = p, *y = q;
    if (x->a != y->a)
        return x->a < y->a ? -1 : 1;
    if (x->b != y->b)
        return x->b < y->b ? -1 : 1;
    return 0;
}
static int find_edge(const epair *e, int ne, int a, int b)
{
    epair key;
    key.a = a < b ? a : b;
    key.b = a < b ? b : a;
    const epair *f = bsearch(&key, e, ne, sizeof(epair), epair_cmp);
    return f ? f->id : -1;
}

/* rotation used by cRotUtil::TransformMat with XYZ order (RotUtil.cpp:1014-1046,
   1252-1258): R = Rz * Ry * Rx */
static void euler_xyz(const double *e, double R[3][3])
{
    double cx = cos(e[0]), sx = sin(e[0]);
    double cy = cos(e[1]), sy = sin(e[1]);
    double cz = cos(e[2]), sz = sin(e[2]);
    double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
    double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
    double Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
    double T[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T[i][j] = (Rz[i][0] * Ry[0][j] + Rz[i][1] * Ry[1][j]) + Rz[i][2] * Ry[2][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            R[i][j] = (T[i][0] * Rx[0][j] + T[i][1] * Rx[1][j]) + T[i][2] * Rx[2][j];
}

static void finish_geometry(oc_cloth *c, const oc_params *p, const double *uv_given);

static int build_grid(oc_cloth *c, const oc_params *p)
{
    double height = p->cloth_size[0], width = p->cloth_size[1];
    int hd = p->subdivision[0], wd = p->subdivision[1];
    if (hd < 1 || wd < 1)
        return -1;
    double uh = height / hd, uw = width / wd;
    int hl = hd + 1, wl = wd + 1;
    c->nv = wl * hl;
    c->nt = 2 * hd * wd;
    c->ne = wl * hd + hl * wd + hd * wd;
    c->xrest = xcalloc(3 * c->nv, sizeof(double));
    c->uv = xcalloc(2 * c->nv, sizeof(float));
    c->tri = xcalloc(3 * c->nt, sizeof(int));
    c->tri_e = xcalloc(3 * c->nt, sizeof(int));
    c->edge = xcalloc(5 * c->ne, sizeof(int));
    for (int k = 0; k < 5 * c->ne; k++)
        c->edge[k] = -2; /* unvisited marker */

    /* vertices: id = i * (hd + 1) + j at (i*uw, j*uh) minus the centre */
    for (int i = 0; i < wl; i++)
        for (int j = 0; j < hl; j++)
        {
            int v = i * hl + j;
            c->xrest[3 * v + 0] = i * uw - width / 2;
            c->xrest[3 * v + 1] = j * uh - height / 2;
            c->xrest[3 * v + 2] = 0.0 - 0.0;
        }

    int per_row = 2 * hd + hl;
    int tcount = 0;
    for (int row = 0; row < wd; row++)
        for (int col = 0; col < hd; col++)
        {
            int lu = row * hl + col, ru = lu + 1, ld = lu + hl, rd = ld + 1;
            int even = ((row + col) % 2) == 0;
            int top = row * per_row + col;
            int left = per_row * row + hd + col * 2;
            int skew = left + 1, right = skew + 1, bottom = top + per_row;
            int *t = c->tri + 3 * tcount;
            if (even)
            {
                t[0] = lu; t[1] = ld; t[2] = rd;
                t[3] = lu; t[4] = rd; t[5] = ru;
            }
            else
            {
                t[0] = lu; t[1] = ld; t[2] = ru;
                t[3] = ld; t[4] = rd; t[5] = ru;
            }
            tcount += 2;
            int *e;
            /* skew edge */
            e = c->edge + 5 * skew;
            e[0] = even ? lu : ru;
            e[1] = even ? rd : ld;
            e[2] = tcount - 2;
            e[3] = tcount - 1;
            e[4] = 0;
            /* bottom edge */
            e = c->edge + 5 * bottom;
            e[0] = ld;
            e[1] = rd;
            e[2] = even ? tcount - 2 : tcount - 1;
            if (row == wd - 1)
            {
                e[3] = -1;
                e[4] = 1;
            }
            else
            {
                e[3] = e[2] + hd * 2;
                e[4] = 0;
            }
            /* right edge */
            e = c->edge + 5 * right;
            e[0] = ru;
            e[1] = rd;
            e[2] = tcount - 1;
            e[4] = (col == hd - 1);
            e[3] = (col == hd - 1) ? -1 : e[2] + 1;
            if (row == 0)
            { /* top edge */
                e = c->edge + 5 * top;
                e[0] = lu;
                e[1] = ru;
                e[2] = even ? tcount - 1 : tcount - 2;
                e[3] = -1;
                e[4] = 1;
            }
            if (col == 0)
            { /* left edge */
                e = c->edge + 5 * left;
                e[0] = lu;
                e[1] = ld;
                e[2] = tcount - 2;
                e[3] = -1;
                e[4] = 1;
            }
        }
    for (int k = 0; k < c->ne; k++)
        if (c->edge[5 * k] == -2)
            return -2;
    finish_geometry(c, p, NULL);
    return 0;
}

/* triangle -> edge ids, init transform and UV, for any triangle mesh whose
   untransformed rest positions, triangles and edges are set */
static void finish_geometry(oc_cloth *c, const oc_params *p, const double *uv_given)
{
    c->flat_y = xcalloc(c->nv, sizeof(float));
    /* BuildTriangleEdgeId: (id0,id1), (id1,id2), (id2,id0) */
    epair *ep = xcalloc(c->ne, sizeof(epair));
    for (int k = 0; k < c->ne; k++)
    {
        int a = c->edge[5 * k], b = c->edge[5 * k + 1];
        ep[k].a = a < b ? a : b;
        ep[k].b = a < b ? b : a;
        ep[k].id = k;
    }
    qsort(ep, c->ne, sizeof(epair), epair_cmp);
    for (int t = 0; t < c->nt; t++)
    {
        int *v = c->tri + 3 * t;
        c->tri_e[3 * t + 0] = find_edge(ep, c->ne, v[0], v[1]);
        c->tri_e[3 * t + 1] = find_edge(ep, c->ne, v[1], v[2]);
        c->tri_e[3 * t + 2] = find_edge(ep, c->ne, v[2], v[0]);
    }
    free(ep);
    if (uv_given)
    {
        /* OBJ cloth: no init transform, UV from the file */
        for (int v = 0; v < c->nv; v++)
        {
            c->uv[2 * v] = (float)uv_given[2 * v];
            c->uv[2 * v + 1] = (float)uv_given[2 * v + 1];
            c->flat_y[v] = (float)c->xrest[3 * v + 1];
        }
        return;
    }

    /* init transform: mPos = TransformMat(pos, euler) * mPos */
    int identity = p->init_pos[0] == 0 && p->init_pos[1] == 0 && p->init_pos[2] == 0 &&
                   p->init_euler_xyz[0] == 0 && p->init_euler_xyz[1] == 0 &&
                   p->init_euler_xyz[2] == 0;
    double R[3][3];
    euler_xyz(p->init_euler_xyz, R);
    double flatxy[2];
    for (int v = 0; v < c->nv; v++)
    {
        double *q = c->xrest + 3 * v;
        flatxy[0] = q[0];
        flatxy[1] = q[1];
        if (!identity)
        {
            double o[3];
            for (int a = 0; a < 3; a++)
                o[a] = ((R[a][0] * q[0] + R[a][1] * q[1]) + R[a][2] * q[2]) + p->init_pos[a];
            q[0] = o[0];
            q[1] = o[1];
            q[2] = o[2];
        }
        /* UV = rot(uv_rotation) * float(inverse-transformed xy); the inverse
           transform returns the untransformed grid coordinates. */
        float ang = (float)(p->uv_rotation_deg / 180.0 * M_PI);
        float cr = (float)cos((double)ang), sr = (float)sin((double)ang);
        float px = (float)flatxy[0], py = (float)flatxy[1];
        float m01 = -sr;
        c->uv[2 * v + 0] = cr * px + m01 * py;
        c->uv[2 * v + 1] = sr * px + cr * py;
        c->flat_y[v] = py;
    }
}

/* SelectAnotherVertex (BaseBendingMaterial.cpp:65-74): the vertex of tri
   other than a and b */
static int other_vertex(const int *t, int a, int b)
{
    for (int k = 0; k < 3; k++)
        if (t[k] != a && t[k] != b)
            return t[k];
    return -1;
}

/* ------------------------------------------------------------------ */
/* block pattern: per row, sorted union of {v}, triangle pairs and     */
/* bending stencils                                                    */
/* ------------------------------------------------------------------ */

static int int_cmp(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* rows[v] gets candidates; returns sorted unique CSR */
static void build_csr(int nv, int ncand, const int *cand_row, const int *cand_col,
                      int **row_ptr_out, int **cols_out, int *nnz_out)
{
    int *cnt = xcalloc(nv + 1, sizeof(int));
    for (int k = 0; k < ncand; k++)
        cnt[cand_row[k] + 1]++;
    for (int v = 0; v < nv; v++)
        cnt[v + 1] += cnt[v];
    int *tmp = xcalloc(ncand, sizeof(int));
    int *fill = xcalloc(nv, sizeof(int));
    for (int k = 0; k < ncand; k++)
    {
        int r = cand_row[k];
        tmp[cnt[r] + fill[r]++] = cand_col[k];
    }
    int *row_ptr = xcalloc(nv + 1, sizeof(int));
    int *cols = xcalloc(ncand, sizeof(int));
    int nnz = 0;
    for (int v = 0; v < nv; v++)
    {
        int *s = tmp + cnt[v];
        int n = cnt[v + 1] - cnt[v];
        qsort(s, n, sizeof(int), int_cmp);
        row_ptr[v] = nnz;
        for (int k = 0; k < n; k++)
            if (k == 0 || s[k] != s[k - 1])
                cols[nnz++] = s[k];
    }
    row_ptr[nv] = nnz;
    free(cnt);
    free(tmp);
    free(fill);
    *row_ptr_out = row_ptr;
    *cols_out = cols;
    *nnz_out = nnz;
}

static int slot_of(const oc_cloth *c, int r, int col)
{
    int lo = c->row_ptr[r], hi = c->row_ptr[r + 1] - 1;
    while (lo <= hi)
    {
        int mid = (lo + hi) / 2;
        if (c->cols[mid] == col)
            return mid;
        if (c->cols[mid] < col)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

static void build_pattern(oc_cloth *c)
{
    int ncand = c->nv + 9 * c->nt + 16 * c->ne;
    int *cr = xcalloc(ncand, sizeof(int)), *cc = xcalloc(ncand, sizeof(int));
    int n = 0;
    for (int v = 0; v < c->nv; v++)
    {
        cr[n] = v;
        cc[n++] = v;
    }
    for (int t = 0; t < c->nt; t++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
            {
                cr[n] = c->tri[3 * t + i];
                cc[n++] = c->tri[3 * t + j];
            }
    for (int e = 0; e < c->ne; e++)
    {
        if (c->edge[5 * e + 4])
            continue;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
            {
                cr[n] = c->stencil[4 * e + i];
                cc[n++] = c->stencil[4 * e + j];
            }
    }
    build_csr(c->nv, n, cr, cc, &c->row_ptr, &c->cols, &c->nnzb);
    free(cr);
    free(cc);
    c->diag = xcalloc(c->nv, sizeof(int));
    for (int v = 0; v < c->nv; v++)
        c->diag[v] = slot_of(c, v, v);
}

int oc_get_ref_ell(const oc_cloth *c, int *row_ptr, int *cols)
{
    /* BaraffClothGpu.cpp:438-500: std::set per vertex of every interior
       stencil containing it, plus the vertex itself (triangles add only
       the diagonal, :460-469) */
    int ncand = c->nv + 16 * c->ne;
    int *cr = xcalloc(ncand, sizeof(int)), *cc = xcalloc(ncand, sizeof(int));
    int n = 0;
    for (int e = 0; e < c->ne; e++)
    {
        if (c->edge[5 * e + 4])
            continue;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
            {
                cr[n] = c->stencil[4 * e + i];
                cc[n++] = c->stencil[4 * e + j];
            }
    }
    for (int v = 0; v < c->nv; v++)
    {
        cr[n] = v;
        cc[n++] = v;
    }
    int *rp, *cl, nnz;
    build_csr(c->nv, n, cr, cc, &rp, &cl, &nnz);
    if (row_ptr)
        memcpy(row_ptr, rp, (c->nv + 1) * sizeof(int));
    if (cols)
        memcpy(cols, cl, nnz * sizeof(int));
    free(rp);
    free(cl);
    free(cr);
    free(cc);
    return nnz;
}

/* ------------------------------------------------------------------ */
/* rest precompute                                                     */
/* ------------------------------------------------------------------ */

/* cMathUtil::CalcTriangleArea3d: float(0.5 * |(p1-p0) x (p2-p0)|) */
static float tri_area_f(const double *p0, const double *p1, const double *p2)
{
    double a[3], b[3], cr[3];
    v3sub(p1, p0, a);
    v3sub(p2, p0, b);
    v3cross(a, b, cr);
    return (float)(0.5 * v3norm(cr));
}

/* CalcBaraffStretchCoef (BaraffMaterial.cpp:18-32) */
static void stretch_coef(const double m[2][2], double *cu, double *cv)
{
    float a = (float)m[0][0], b = (float)m[0][1], c = (float)m[1][0], d = (float)m[1][1];
    cu[0] = (double)(-a - c);
    cu[1] = (double)a;
    cu[2] = (double)c;
    cv[0] = (double)(-b - d);
    cv[1] = (double)b;
    cv[2] = (double)d;
}

static void rest_precompute(oc_cloth *c, const oc_params *p)
{
    int nv = c->nv, nt = c->nt, ne = c->ne;
    float rho = (float)p->density;
    /* cBaraffCloth::InitMass (BaraffCloth.cpp:209-236) */
    double *mdiag = xcalloc(3 * nv, sizeof(double));
    c->area = xcalloc(nt, sizeof(double));
    for (int t = 0; t < nt; t++)
    {
        const int *v = c->tri + 3 * t;
        double A = (double)tri_area_f(c->xrest + 3 * v[0], c->xrest + 3 * v[1],
                                      c->xrest + 3 * v[2]);
        c->area[t] = A;
        double add = A / 3 * (double)rho;
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++)
                mdiag[3 * v[k] + a] = mdiag[3 * v[k] + a] + add;
    }
    c->mass = xcalloc(nv, sizeof(double));
    c->gravity = xcalloc(3 * nv, sizeof(double));
    for (int v = 0; v < nv; v++)
    {
        c->mass[v] = mdiag[3 * v];
        /* BaraffCloth.cpp:73-77 */
        c->gravity[3 * v + 1] = -c->mass[v] * 9.8 * (p->enable_gravity ? 1 : 0);
    }
    free(mdiag);

    /* cBaraffMaterial::Init (BaraffMaterial.cpp:33-76) */
    c->coef = xcalloc(12 * nt, sizeof(double));
    double cs = cos(M_PI / 4), sn = sin(M_PI / 4);
    /* R45 = RotMat2D(pi/4) = [[cs, -sn], [sn, cs]]; R45^T = [[cs, sn], [-sn, cs]] */
    double RT[2][2] = {{cs, sn}, {-sn, cs}};
    for (int t = 0; t < nt; t++)
    {
        const int *v = c->tri + 3 * t;
        const float *u0 = c->uv + 2 * v[0], *u1 = c->uv + 2 * v[1], *u2 = c->uv + 2 * v[2];
        double m[2][2];
        m[0][0] = (double)(u1[0] - u0[0]);
        m[1][0] = (double)(u1[1] - u0[1]);
        m[0][1] = (double)(u2[0] - u0[0]);
        m[1][1] = (double)(u2[1] - u0[1]);
        double det = m[0][0] * m[1][1] - m[1][0] * m[0][1];
        double invdet = 1.0 / det;
        double inv[2][2];
        inv[0][0] = m[1][1] * invdet;
        inv[1][0] = -m[1][0] * invdet;
        inv[0][1] = -m[0][1] * invdet;
        inv[1][1] = m[0][0] * invdet;
        double *cf = c->coef + 12 * t;
        stretch_coef(inv, cf + 0, cf + 3);
        double r[2][2];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                r[i][j] = inv[i][0] * RT[0][j] + inv[i][1] * RT[1][j];
        stretch_coef(r, cf + 6, cf + 9);
    }
    c->k_ww[0] = (double)(float)p->stretch_k[0];
    c->k_ww[1] = (double)(float)p->stretch_k[1];
    c->k_bias = (double)(float)p->stretch_k[2];

    /* cDihedralMaterial::Init (DihedralBending.cpp:23-89) */
    c->edge_len = xcalloc(ne, sizeof(float));
    c->heron = xcalloc(nt, sizeof(float));
    c->height = xcalloc(2 * ne, sizeof(float));
    for (int e = 0; e < ne; e++)
    {
        const double *a = c->xrest + 3 * c->edge[5 * e], *b = c->xrest + 3 * c->edge[5 * e + 1];
        /* 4-vector norm with w = 1 - 1; SSE2 order (x^2 + z^2) + (y^2 + w^2) */
        double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2], dw = 0.0;
        c->edge_len[e] = (float)sqrt((dx * dx + dz * dz) + (dy * dy + dw * dw));
    }
    for (int t = 0; t < nt; t++)
    {
        double e0 = c->edge_len[c->tri_e[3 * t]], e1 = c->edge_len[c->tri_e[3 * t + 1]],
               e2 = c->edge_len[c->tri_e[3 * t + 2]];
        double s = (e0 + e1 + e2) / 2;
        c->heron[t] = (float)sqrt(s * (s - e0) * (s - e1) * (s - e2));
    }
    for (int e = 0; e < ne; e++)
    {
        float h0 = 0, h1 = 0;
        if (!c->edge[5 * e + 4])
        {
            /* height = 2 * tVector2f(a0, a1) / e_length: float arithmetic,
               the divisor converted to float (Eigen scalar promotion) */
            float a0 = c->heron[c->edge[5 * e + 2]], a1 = c->heron[c->edge[5 * e + 3]];
            float el = (float)(double)c->edge_len[e];
            h0 = (2.0f * a0) / el;
            h1 = (2.0f * a1) / el;
        }
        c->height[2 * e] = h0;
        c->height[2 * e + 1] = h1;
    }
    /* mEdgeK = mean of the 3-vector filled with float(k) (BaseBendingMaterial.cpp:84) */
    double kb = (double)(float)p->bending_k;
    c->edge_k = ((kb + kb) + kb) / 3;

    c->ray_a = (float)p->rayleigh_a;
    c->Ku = (float)p->bending_k;
    c->Kv = (float)(p->bending_k2 != 0 ? p->bending_k2 : p->bending_k);
    c->ray_b = (float)p->rayleigh_b;
    c->gf_a = c->ray_a;
    c->gf_b = c->ray_b;
    c->density_f = (float)p->density;
    c->dt = p->dt;
    c->max_iters = p->max_iters;
}

/* FindVerticesIdFromUV (BaseCloth.cpp:151-189) */
static void find_pins(oc_cloth *c, const oc_params *p)
{
    c->npin = p->num_pins;
    c->pins = xcalloc(c->npin, sizeof(int));
    c->pinned = xcalloc(c->nv, 1);
    for (int j = 0; j < c->npin; j++)
    {
        float tu = (float)p->pin_uv[2 * j], tv = (float)p->pin_uv[2 * j + 1];
        double best = NAN;
        int id = -1;
        for (int v = 0; v < c->nv; v++)
        {
            float du = c->uv[2 * v] - tu, dv = c->uv[2 * v + 1] - tv;
            double dist = (double)sqrtf(du * du + dv * dv);
            if (isnan(best) || dist < best)
            {
                id = v;
                best = dist;
            }
        }
        c->pins[j] = id;
        if (id >= 0)
            c->pinned[id] = 1;
    }
    if (p->pin_top_edge)
    {
        /* every vertex on the top side (largest untransformed y) */
        float top = -3.4e38f;
        int extra = 0;
        for (int v = 0; v < c->nv; v++)
            top = c->flat_y[v] > top ? c->flat_y[v] : top;
        for (int v = 0; v < c->nv; v++)
            extra += c->flat_y[v] == top;
        c->pins = realloc(c->pins, (c->npin + extra) * sizeof(int));
        for (int v = 0; v < c->nv; v++)
            if (c->flat_y[v] == top)
            {
                c->pins[c->npin++] = v;
                c->pinned[v] = 1;
            }
    }
}

static oc_cloth *complete(oc_cloth *c, const oc_params *p);

oc_cloth *oc_create_grid(const oc_params *p)
{
    oc_cloth *c = xcalloc(1, sizeof(oc_cloth));
    if (build_grid(c, p) != 0)
    {
        oc_destroy(c);
        return NULL;
    }
    return complete(c, p);
}

/* Edge list of a triangle list as cTriangulator::DelaunayTriangulation
   files it (add_edge_into_edge_array, TriangulatorDelaunay.cpp:9-43,
   118-142): each triangle's sides (id0,id1), (id1,id2), (id2,id0) in order;
   an unseen side becomes edge {min, max} with the next id and tri0 = this
   triangle, a seen one gets tri1 = this triangle; boundary when no tri1.
   Restated with a sort over the sides (the reference searches its edge list
   linearly); the first-appearance numbering is the same. */
typedef struct
{
    int a, b, side;
} oc_side;
static int side_cmp(const void *p, const void *q)
{
    const oc_side *x = p, *y = q;
    if (x->a != y->a)
        return x->a < y->a ? -1 : 1;
    if (x->b != y->b)
        return x->b < y->b ? -1 : 1;
    return x->side < y->side ? -1 : (x->side > y->side);
}
static int first_cmp(const void *p, const void *q)
{
    const int *x = p, *y = q;
    return x[0] < y[0] ? -1 : (x[0] > y[0]);
}

oc_cloth *oc_create_mesh(const oc_params *p, const double *xyz, int nv, const int *tri, int nt)
{
    return oc_create_mesh2(p, xyz, nv, tri, nt, NULL, 0);
}

/* ObjUtil.cpp:141-224 (cObjUtil::BuildEdge): edges keyed by {min, max} in a
   std::map, so numbered in ascending key order; tri0/tri1 = the first and
   second triangle (in triangle order) holding the edge */
oc_cloth *oc_create_mesh2(const oc_params *p, const double *xyz, int nv, const int *tri, int nt,
                          const double *uv, int sorted_edges)
{
    if (nv < 3 || nt < 1)
        return NULL;
    for (int k = 0; k < 3 * nt; k++)
        if (tri[k] < 0 || tri[k] >= nv)
            return NULL;
    oc_cloth *c = xcalloc(1, sizeof(oc_cloth));
    c->nv = nv;
    c->nt = nt;
    c->xrest = xcalloc(3 * (size_t)nv, sizeof(double));
    memcpy(c->xrest, xyz, 3 * (size_t)nv * sizeof(double));
    c->uv = xcalloc(2 * (size_t)nv, sizeof(float));
    c->tri = xcalloc(3 * (size_t)nt, sizeof(int));
    memcpy(c->tri, tri, 3 * (size_t)nt * sizeof(int));
    c->tri_e = xcalloc(3 * (size_t)nt, sizeof(int));
    int ns = 3 * nt;
    oc_side *sd = xcalloc(ns, sizeof(oc_side));
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
        {
            int a = tri[3 * t + k], b = tri[3 * t + (k + 1) % 3];
            sd[3 * t + k].a = a < b ? a : b;
            sd[3 * t + k].b = a < b ? b : a;
            sd[3 * t + k].side = 3 * t + k;
        }
    qsort(sd, ns, sizeof(oc_side), side_cmp);
    /* groups of equal {a, b} as (first side, group start), then ordered by
       first side = first appearance */
    int *grp = xcalloc(2 * (size_t)ns, sizeof(int));
    int ng = 0, bad = 0;
    for (int i = 0; i < ns;)
    {
        int j = i + 1;
        while (j < ns && sd[j].a == sd[i].a && sd[j].b == sd[i].b)
            j++;
        bad |= j - i > 2;
        grp[2 * ng] = sd[i].side;
        grp[2 * ng + 1] = i;
        ng++;
        i = j;
    }
    if (bad)
    {
        free(sd);
        free(grp);
        oc_destroy(c);
        return NULL;
    }
    if (!sorted_edges) /* groups are in key order already */
        qsort(grp, ng, 2 * sizeof(int), first_cmp);
    c->ne = ng;
    c->edge = xcalloc(5 * (size_t)ng, sizeof(int));
    for (int e = 0; e < ng; e++)
    {
        int i = grp[2 * e + 1];
        int two = i + 1 < ns && sd[i + 1].a == sd[i].a && sd[i + 1].b == sd[i].b;
        int *ed = c->edge + 5 * e;
        ed[0] = sd[i].a;
        ed[1] = sd[i].b;
        ed[2] = sd[i].side / 3;
        ed[3] = two ? sd[i + 1].side / 3 : -1;
        ed[4] = !two;
    }
    free(sd);
    free(grp);
    finish_geometry(c, p, uv);
    return complete(c, p);
}

static oc_cloth *complete(oc_cloth *c, const oc_params *p)
{
    /* stencils (BaseCloth.cpp:383-404) */
    c->stencil = xcalloc(4 * c->ne, sizeof(int));
    for (int e = 0; e < c->ne; e++)
    {
        int *s = c->stencil + 4 * e;
        if (c->edge[5 * e + 4])
        {
            s[0] = s[1] = s[2] = s[3] = -1;
            continue;
        }
        int v0 = c->edge[5 * e], v1 = c->edge[5 * e + 1];
        s[0] = v0;
        s[1] = v1;
        s[2] = other_vertex(c->tri + 3 * c->edge[5 * e + 2], v0, v1);
        s[3] = other_vertex(c->tri + 3 * c->edge[5 * e + 3], v0, v1);
    }
    rest_precompute(c, p);
    find_pins(c, p);
    build_pattern(c);
    size_t nb = 9 * (size_t)c->nnzb;
    c->Ks = xcalloc(nb, sizeof(double));
    c->Kb = xcalloc(nb, sizeof(double));
    c->K = xcalloc(nb, sizeof(double));
    c->W = xcalloc(nb, sizeof(double));
    c->fs = xcalloc(3 * c->nv, sizeof(double));
    c->fb = xcalloc(3 * c->nv, sizeof(double));
    c->f = xcalloc(3 * c->nv, sizeof(double));
    c->b = xcalloc(3 * c->nv, sizeof(double));
    c->dx = xcalloc(3 * c->nv, sizeof(double));
    c->x = xcalloc(3 * c->nv, sizeof(double));
    c->xpre = xcalloc(3 * c->nv, sizeof(double));
    c->fuser = xcalloc(3 * c->nv, sizeof(double));
    c->drag = -1;
    c->ydrag = 0.0f;
    memcpy(c->x, c->xrest, 3 * c->nv * sizeof(double));
    memcpy(c->xpre, c->xrest, 3 * c->nv * sizeof(double));
    return c;
}

void oc_destroy(oc_cloth *c)
{
    if (!c)
        return;
    free(c->xrest); free(c->x); free(c->xpre); free(c->uv); free(c->tri);
    free(c->tri_e); free(c->edge); free(c->stencil); free(c->pins); free(c->pinned);
    free(c->flat_y);
    free(c->gf_qb); free(c->gf_mass); free(c->gf_grav); free(c->gf_dx);
    free(c->mass); free(c->gravity); free(c->coef); free(c->area);
    free(c->edge_len); free(c->heron); free(c->height);
    free(c->row_ptr); free(c->cols); free(c->diag);
    free(c->Ks); free(c->Kb); free(c->K); free(c->W);
    free(c->fs); free(c->fb); free(c->f); free(c->b); free(c->dx); free(c->fuser);
    oc_set_contacts(c, 0, NULL, NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL);
    free(c);
}

/* ------------------------------------------------------------------ */
/* elements                                                            */
/* ------------------------------------------------------------------ */

/* one direction set of cBaraffMaterial::Update (BaraffMaterial.cpp:205-247):
   CalcFuAndFv :89-95, CalcFint :113-125, CalcHessian :135-172 */
static void stretch_pair(const double *x0, const double *x1, const double *x2,
                         const double *cu, const double *cv, double A, double Ku,
                         double Kv, double f[3][3], double H[9][3][3])
{
    double Fu[3], Fv[3];
    for (int a = 0; a < 3; a++)
    {
        Fu[a] = (cu[0] * x0[a] + cu[1] * x1[a]) + cu[2] * x2[a];
        Fv[a] = (cv[0] * x0[a] + cv[1] * x1[a]) + cv[2] * x2[a];
    }
    double Cu = v3norm(Fu) - 1;
    double Cv = v3norm(Fv) - 1;
    double nu[3], nv[3];
    v3normalized(Fu, nu);
    v3normalized(Fv, nv);
    /* CalcFint */
    double su = ((-A) * Ku) * Cu, sv = ((-A) * Kv) * Cv;
    double ub[3] = {su * nu[0], su * nu[1], su * nu[2]};
    double vb[3] = {sv * nv[0], sv * nv[1], sv * nv[2]};
    for (int i = 0; i < 3; i++)
        for (int a = 0; a < 3; a++)
            f[i][a] = ub[a] * cu[i] + vb[a] * cv[i];
    /* CalcHessian; Fv_norm = Fu.norm() is the reference's (bug :142) */
    double Fu_norm = v3norm(Fu);
    double Fv_norm = v3norm(Fu);
    double Du[3][3], Dv[3][3];
    double iu = 1.0 / Fu_norm, iv = 1.0 / Fv_norm;
    double lu[3] = {iu * nu[0], iu * nu[1], iu * nu[2]};
    double lv[3] = {iv * nv[0], iv * nv[1], iv * nv[2]};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
        {
            Du[a][b] = nu[b] * lu[a];
            Dv[a][b] = nv[b] * lv[a];
        }
    if (Fu_norm > 1)
    {
        double add = 1 - 1.0 / Fu_norm;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                Du[a][b] = Du[a][b] + (a == b ? add : add * 0.0);
    }
    if (Fv_norm > 1)
    {
        double add = 1 - 1.0 / Fv_norm;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                Dv[a][b] = Dv[a][b] + (a == b ? add : add * 0.0);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            double su2 = (Ku * cu[i]) * cu[j];
            double sv2 = (Kv * cv[i]) * cv[j];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    H[3 * i + j][a][b] = (-A) * (su2 * Du[a][b] + sv2 * Dv[a][b]);
        }
}

void oc_stretch_element(const double *x0, const double *x1, const double *x2,
                        const double *coef12, double area, const double *k3,
                        double *f9, double *H81)
{
    double f1[3][3], f2[3][3], H1[9][3][3], H2[9][3][3];
    stretch_pair(x0, x1, x2, coef12 + 0, coef12 + 3, area, k3[0], k3[1], f1, H1);
    stretch_pair(x0, x1, x2, coef12 + 6, coef12 + 9, area, k3[2], k3[2], f2, H2);
    for (int i = 0; i < 3; i++)
        for (int a = 0; a < 3; a++)
            f9[3 * i + a] = (0.0 + f1[i][a]) + f2[i][a];
    for (int k = 0; k < 9; k++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                H81[9 * k + 3 * a + b] = (0.0 + H1[k][a][b]) + H2[k][a][b];
}

/* CalcIprime (DihedralBending.cpp:115-119): (I - n n^T) / |v| */
static void iprime(const double *vec, double M[3][3])
{
    double n[3];
    v3normalized(vec, n);
    double len = v3norm(vec);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            M[a][b] = ((a == b ? 1.0 : 0.0) - n[b] * n[a]) / len;
}

/* CalcBeta + CalcDBetaDx (DihedralBending.cpp:120-271) */
void oc_dbeta_dx(const double *v0, const double *v1, const double *v2,
                 const double *v3, double *beta_out, double *g)
{
    double t0[3], t1[3], n0[3], n1[3], e[3];
    v3sub(v1, v0, t0);
    v3sub(v2, v1, t1);
    v3cross(t0, t1, n0);
    v3sub(v0, v1, t0);
    v3sub(v3, v0, t1);
    v3cross(t0, t1, n1);
    v3sub(v1, v0, e);
    double n0b[3], n1b[3], eb[3], c01[3];
    v3normalized(n0, n0b);
    v3normalized(n1, n1b);
    v3normalized(e, eb);
    double x = v3dot(n0b, n1b);
    v3cross(n0b, n1b, c01);
    double y = v3dot(eb, c01);
    double beta = atan2(y, x);
    /* CalcDBetaDnormal_and_De :163-223 */
    double sign_sinb = sin(beta) > 0 ? 1.0 : -1.0;
    double I0[3][3], I1[3][3];
    iprime(n0, I0);
    iprime(n1, I1);
    double m = 1 - x * x;
    double deno = sqrt(m > 0 ? m : 0);
    if (deno < 1e-10)
        deno = 1e-10;
    double factor = -sign_sinb * 1.0 / deno;
    double dn0[3], dn1[3];
    for (int a = 0; a < 3; a++)
    {
        /* (factor * I^T) * n_bar, coefficient-wise lazy product */
        dn0[a] = ((factor * I0[0][a]) * n1b[0] + (factor * I0[1][a]) * n1b[1]) +
                 (factor * I0[2][a]) * n1b[2];
        dn1[a] = ((factor * I1[0][a]) * n0b[0] + (factor * I1[1][a]) * n0b[1]) +
                 (factor * I1[2][a]) * n0b[2];
    }
    double a0[3], a1[3], c0[3], c1[3];
    /* dbeta/dv0 = (v1 - v2) x dn0 + (v3 - v1) x dn1 */
    v3sub(v1, v2, a0);
    v3sub(v3, v1, a1);
    v3cross(a0, dn0, c0);
    v3cross(a1, dn1, c1);
    for (int k = 0; k < 3; k++)
        g[k] = c0[k] + c1[k];
    /* dbeta/dv1 = (v2 - v0) x dn0 + (v0 - v3) x dn1 */
    v3sub(v2, v0, a0);
    v3sub(v0, v3, a1);
    v3cross(a0, dn0, c0);
    v3cross(a1, dn1, c1);
    for (int k = 0; k < 3; k++)
        g[3 + k] = c0[k] + c1[k];
    /* dbeta/dv2 = (v0 - v1) x dn0 ; dbeta/dv3 = (v1 - v0) x dn1 */
    v3sub(v0, v1, a0);
    v3cross(a0, dn0, g + 6);
    v3sub(v1, v0, a1);
    v3cross(a1, dn1, g + 9);
    *beta_out = beta;
}

/* ------------------------------------------------------------------ */
/* assembly                                                            */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* collision response                                                  */
/* ------------------------------------------------------------------ */

int oc_set_contacts(oc_cloth *c, int n_pt, const int *pt_ids, const double *pt_normal,
                    const double *pt_bary, const float *pt_depth, int n_ee, const int *ee_ids,
                    const double *ee_normal, const float *ee_bary, const float *ee_depth)
{
    free(c->pt_ids); free(c->pt_normal); free(c->pt_bary); free(c->pt_depth);
    free(c->ee_ids); free(c->ee_normal); free(c->ee_bary); free(c->ee_depth);
    c->pt_ids = NULL; c->pt_normal = NULL; c->pt_bary = NULL; c->pt_depth = NULL;
    c->ee_ids = NULL; c->ee_normal = NULL; c->ee_bary = NULL; c->ee_depth = NULL;
    c->n_pt = c->n_ee = 0;
    for (int k = 0; k < n_pt; k++)
    {
        int v = pt_ids[3 * k + 1], t = pt_ids[3 * k + 2];
        if (pt_ids[3 * k] ? (v < 0 || v >= c->nv) : (t < 0 || t >= c->nt))
            return -1;
    }
    for (int k = 0; k < n_ee; k++)
    {
        int e = ee_ids[3 * k] ? ee_ids[3 * k + 1] : ee_ids[3 * k + 2];
        if (e < 0 || e >= c->ne)
            return -1;
    }
    if (n_pt > 0)
    {
        c->pt_ids = xcalloc(3 * (size_t)n_pt, sizeof(int));
        c->pt_normal = xcalloc(3 * (size_t)n_pt, sizeof(double));
        c->pt_bary = xcalloc(3 * (size_t)n_pt, sizeof(double));
        c->pt_depth = xcalloc(n_pt, sizeof(float));
        memcpy(c->pt_ids, pt_ids, 3 * (size_t)n_pt * sizeof(int));
        memcpy(c->pt_normal, pt_normal, 3 * (size_t)n_pt * sizeof(double));
        memcpy(c->pt_bary, pt_bary, 3 * (size_t)n_pt * sizeof(double));
        memcpy(c->pt_depth, pt_depth, (size_t)n_pt * sizeof(float));
        c->n_pt = n_pt;
    }
    if (n_ee > 0)
    {
        c->ee_ids = xcalloc(3 * (size_t)n_ee, sizeof(int));
        c->ee_normal = xcalloc(3 * (size_t)n_ee, sizeof(double));
        c->ee_bary = xcalloc(2 * (size_t)n_ee, sizeof(float));
        c->ee_depth = xcalloc(n_ee, sizeof(float));
        memcpy(c->ee_ids, ee_ids, 3 * (size_t)n_ee * sizeof(int));
        memcpy(c->ee_normal, ee_normal, 3 * (size_t)n_ee * sizeof(double));
        memcpy(c->ee_bary, ee_bary, 2 * (size_t)n_ee * sizeof(float));
        memcpy(c->ee_depth, ee_depth, (size_t)n_ee * sizeof(float));
        c->n_ee = n_ee;
    }
    return 0;
}

/* cBaraffCloth::PrepareCollisionHessian (BaraffCloth.cpp:928-990) added to
   W (SolveForDx :293).  Per point-triangle contact: dC/dx = N on the cloth
   vertex, or -bary_i N on the cloth triangle's vertex i; blocks
   1e-2 * dC_i dC_j^T, triplets in contact order (i, j, a, b), summed per
   entry in that order by setFromTriplets, then W + H. */
static void add_contact_hessian(oc_cloth *c)
{
    if (c->n_pt == 0)
        return;
    size_t nb = 9 * (size_t)c->nnzb;
    double *H = xcalloc(nb, sizeof(double));
    unsigned char *touched = xcalloc(nb, 1);
    for (int k = 0; k < c->n_pt; k++)
    {
        const double *N = c->pt_normal + 3 * k;
        double dc[4][3];
        int id[4] = {-1, -1, -1, -1};
        if (c->pt_ids[3 * k])
        {
            for (int a = 0; a < 3; a++)
                dc[0][a] = N[a];
            id[0] = c->pt_ids[3 * k + 1];
        }
        else
        {
            const int *t = c->tri + 3 * c->pt_ids[3 * k + 2];
            for (int i = 0; i < 3; i++)
            {
                for (int a = 0; a < 3; a++)
                    dc[1 + i][a] = -c->pt_bary[3 * k + i] * N[a];
                id[1 + i] = t[i];
            }
        }
        for (int i = 0; i < 4; i++)
        {
            if (id[i] < 0)
                continue;
            for (int j = 0; j < 4; j++)
            {
                if (id[j] < 0)
                    continue;
                int s = slot_of(c, id[i], id[j]);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                    {
                        double h = 1e-2 * (dc[i][a] * dc[j][b]);
                        size_t q = 9 * (size_t)s + 3 * a + b;
                        H[q] = touched[q] ? H[q] + h : h;
                        touched[q] = 1;
                    }
            }
        }
    }
    for (size_t q = 0; q < nb; q++)
        if (touched[q])
            c->W[q] = c->W[q] + H[q];
    free(H);
    free(touched);
}

/* cBaraffCloth::Repulsion (BaraffCloth.cpp:626-926), one pass: impulses from
   edge-edge contacts, then point-triangle contacts, each applied to dx at
   once (later contacts see earlier impulses).  K = mStretchK.mean() in float
   (Eigen's unrolled sum k0 + (k1 + k2), then / 3). */
static void repulsion(oc_cloth *c)
{
    const double dt = c->dt;
    const float scale = 0.1f;
    const double v_thre = 0.01;
    const float kf = ((float)c->k_ww[0] + ((float)c->k_ww[1] + (float)c->k_bias)) / 3.0f;
    const double K = (double)(kf * 1);
    double *dx = c->dx;
    for (int k = 0; k < c->n_ee; k++)
    {
        const double *N = c->ee_normal + 3 * k;
        const int cloth0 = c->ee_ids[3 * k] != 0;
        const int e = cloth0 ? c->ee_ids[3 * k + 1] : c->ee_ids[3 * k + 2];
        const float b0 = c->ee_bary[2 * k], b1 = c->ee_bary[2 * k + 1];
        const float bc = cloth0 ? b0 : b1;
        const int v0 = c->edge[5 * e], v1 = c->edge[5 * e + 1];
        const double w0 = (double)(1 - bc), w1 = (double)bc;
        double vrel[3];
        for (int a = 0; a < 3; a++)
            vrel[a] = (w0 * dx[3 * v0 + a] + w1 * dx[3 * v1 + a]) / dt;
        const double mass = c->mass[v0] + c->mass[v1];
        const double vn = (vrel[0] * N[0] + vrel[1] * N[1]) + vrel[2] * N[2];
        const double Ic = vn > v_thre ? 0 : mass * fabs(vn) / 2;
        const double d = (3e-3 - (double)c->ee_depth[k]);
        const float inhibition_coef = 0.1f;
        double Ir = 0;
        if (!(vn > inhibition_coef * d / dt))
        {
            const double Ir0 = dt * K * d;
            const double Ir1 = mass * (inhibition_coef * d / dt - vn);
            Ir = Ir0 < Ir1 ? Ir0 : Ir1;
        }
        const double I = Ic + Ir;
        const float den = 1 + b0 * b0 + (1 - b0) * (1 - b0) + b1 * b1 + (1 - b1) * (1 - b1);
        const double I_tilde = 2 * I / den;
        const double s0 = scale * dt * (1 - bc) * I_tilde / mass;
        const double s1 = scale * dt * bc * I_tilde / mass;
        for (int a = 0; a < 3; a++)
            dx[3 * v0 + a] = dx[3 * v0 + a] + s0 * N[a];
        for (int a = 0; a < 3; a++)
            dx[3 * v1 + a] = dx[3 * v1 + a] + s1 * N[a];
    }
    for (int k = 0; k < c->n_pt; k++)
    {
        const double *N = c->pt_normal + 3 * k;
        const double *bary = c->pt_bary + 3 * k;
        const int cloth_point = c->pt_ids[3 * k] != 0;
        int vs[3];
        double vrel[3], mass;
        if (cloth_point)
        {
            vs[0] = c->pt_ids[3 * k + 1];
            for (int a = 0; a < 3; a++)
                vrel[a] = dx[3 * vs[0] + a] / dt;
            mass = c->mass[vs[0]];
        }
        else
        {
            const int *t = c->tri + 3 * c->pt_ids[3 * k + 2];
            for (int i = 0; i < 3; i++)
                vs[i] = t[i];
            for (int a = 0; a < 3; a++)
                vrel[a] = (dx[3 * vs[0] + a] * bary[0] + dx[3 * vs[1] + a] * bary[1] +
                           dx[3 * vs[2] + a] * bary[2]) /
                          dt;
            mass = c->mass[vs[0]] + c->mass[vs[1]] + c->mass[vs[2]];
        }
        const double vn = (vrel[0] * N[0] + vrel[1] * N[1]) + vrel[2] * N[2];
        const double Ic = vn > v_thre ? 0 : mass * fabs(vn) / 2;
        const double d = (3e-3 - (double)c->pt_depth[k]);
        double Ir = 0;
        if (!(vn > 0.1 * d / dt))
        {
            const double Ir0 = dt * K * d;
            const double Ir1 = mass * (0.1 * d / dt - vn);
            Ir = Ir0 < Ir1 ? Ir0 : Ir1;
        }
        const double I = Ic + Ir;
        const double I_tilde = 2 * I / (1 + ((bary[0] * bary[0] + bary[1] * bary[1]) + bary[2] * bary[2]));
        if (cloth_point)
        {
            const double f = dt * I_tilde / mass;
            for (int a = 0; a < 3; a++)
                dx[3 * vs[0] + a] = dx[3 * vs[0] + a] + (double)scale * (f * N[a]);
        }
        else
            for (int i = 0; i < 3; i++)
            {
                const double f = dt * bary[i] * I_tilde / mass;
                for (int a = 0; a < 3; a++)
                    dx[3 * vs[i] + a] = dx[3 * vs[i] + a] + (double)scale * (f * N[a]);
            }
    }
}

void oc_assemble(oc_cloth *c)
{
    int nv = c->nv;
    size_t nb = 9 * (size_t)c->nnzb;
    memset(c->Ks, 0, nb * sizeof(double));
    memset(c->Kb, 0, nb * sizeof(double));
    memset(c->fs, 0, 3 * nv * sizeof(double));
    memset(c->fb, 0, 3 * nv * sizeof(double));
    double k3[3] = {c->k_ww[0], c->k_ww[1], c->k_bias};
    /* stretch: triplets in triangle order, summed by setFromTriplets
       (BaraffMaterial.cpp:297-351); force CalcTotalForce :267-282 */
    for (int t = 0; t < c->nt; t++)
    {
        const int *v = c->tri + 3 * t;
        double f9[9], H[81];
        oc_stretch_element(c->x + 3 * v[0], c->x + 3 * v[1], c->x + 3 * v[2],
                           c->coef + 12 * t, c->area[t], k3, f9, H);
        for (int i = 0; i < 3; i++)
        {
            for (int a = 0; a < 3; a++)
                c->fs[3 * v[i] + a] = c->fs[3 * v[i] + a] + f9[3 * i + a];
            for (int j = 0; j < 3; j++)
            {
                int s = slot_of(c, v[i], v[j]);
                double *blk = c->Ks + 9 * (size_t)s;
                for (int k = 0; k < 9; k++)
                    blk[k] = blk[k] + H[9 * (3 * i + j) + k];
            }
        }
    }
    /* dihedral bending (DihedralBending.cpp:378-489) */
    for (int e = 0; e < c->ne; e++)
    {
        if (c->edge[5 * e + 4])
            continue;
        const int *s4 = c->stencil + 4 * e;
        double beta, g[12];
        oc_dbeta_dx(c->x + 3 * s4[0], c->x + 3 * s4[1], c->x + 3 * s4[2], c->x + 3 * s4[3],
                    &beta, g);
        double el = (double)c->edge_len[e];
        float hs = c->height[2 * e] + c->height[2 * e + 1];
        if (hs < 1e-7f)
            hs = 1e-7f;
        double factor = -c->edge_k * el / (double)hs;
        double fb = factor * beta;
        for (int i = 0; i < 4; i++)
        {
            int vi = s4[i];
            for (int a = 0; a < 3; a++)
                c->fb[3 * vi + a] = c->fb[3 * vi + a] + (fb * g[3 * i + a]) / 2;
            for (int j = 0; j < 4; j++)
            {
                int s = slot_of(c, vi, s4[j]);
                double *blk = c->Kb + 9 * (size_t)s;
                for (int a = 0; a < 3; a++)
                {
                    double l = factor * g[3 * i + a];
                    for (int b = 0; b < 3; b++)
                    {
                        /* mEleKLst = float(-H); K_b = -1 * sum(double(float(-H))) */
                        float hf = (float)(-(l * g[3 * j + b]));
                        blk[3 * a + b] = blk[3 * a + b] + (double)hf;
                    }
                }
            }
        }
    }
    for (size_t k = 0; k < nb; k++)
    {
        c->Kb[k] = -1 * c->Kb[k];
        c->K[k] = c->Ks[k] + c->Kb[k];
    }
    /* f = f_stretch + f_bend (BaraffCloth.cpp:202-207) */
    for (int k = 0; k < 3 * nv; k++)
        c->f[k] = c->fs[k] + c->fb[k];
    /* W = (1 + dt a) M + dt (b - dt) K ; rhs = dt^2 (g + u + f) + dt M (x - xpre) / dt
       (BaraffCloth.cpp:274-283) */
    double dt = c->dt;
    double sk = dt * ((double)c->ray_b - dt);
    double sm = 1 + dt * (double)c->ray_a;
    for (int v = 0; v < nv; v++)
        for (int s = c->row_ptr[v]; s < c->row_ptr[v + 1]; s++)
        {
            double *w = c->W + 9 * (size_t)s;
            const double *k = c->K + 9 * (size_t)s;
            for (int q = 0; q < 9; q++)
                w[q] = sk * k[q];
            if (c->cols[s] == v)
            {
                double mm = sm * c->mass[v];
                w[0] = mm + w[0];
                w[4] = mm + w[4];
                w[8] = mm + w[8];
            }
        }
    /* cBaseCloth::ApplyYDragForce (BaseCloth.cpp:592-623), called by
       UpdatePos after the internal force (BaraffCloth.cpp:129): the vertices
       whose float v coordinate is within 1e-5 of the largest (smallest) v
       get +F/n (-F/n) in y, F/n in float; added to mUserForce, which the
       step then clears */
    double *u = xcalloc(3 * nv, sizeof(double));
    memcpy(u, c->fuser, 3 * nv * sizeof(double));
    {
        double vmax = -1e5, vmin = 1e5;
        for (int v = 0; v < nv; v++)
        {
            double cv = (double)c->uv[2 * v + 1];
            vmax = cv > vmax ? cv : vmax;
            vmin = cv < vmin ? cv : vmin;
        }
        int n_up = 0, n_low = 0;
        for (int v = 0; v < nv; v++)
        {
            float cv = c->uv[2 * v + 1];
            n_up += fabs(vmax - (double)cv) < 1e-5;
            n_low += fabs(vmin - (double)cv) < 1e-5;
        }
        for (int v = 0; v < nv; v++)
            if (fabs(vmax - (double)c->uv[2 * v + 1]) < 1e-5)
                u[3 * v + 1] = u[3 * v + 1] + (double)(c->ydrag / (float)n_up);
        for (int v = 0; v < nv; v++)
            if (fabs(vmin - (double)c->uv[2 * v + 1]) < 1e-5)
                u[3 * v + 1] = u[3 * v + 1] + (double)(-c->ydrag / (float)n_low);
    }
    double dt2 = dt * dt;
    for (int v = 0; v < nv; v++)
        for (int a = 0; a < 3; a++)
        {
            int q = 3 * v + a;
            double t1 = dt2 * ((c->gravity[q] + u[q]) + c->f[q]);
            double t2 = ((dt * c->mass[v]) * (c->x[q] - c->xpre[q])) / dt;
            c->b[q] = t1 + t2;
        }
    free(u);
    add_contact_hessian(c);
}

/* ApplyMatmul (BaraffCloth.cpp:469-506): rows and columns of pinned vertices
   skipped, row sums in ascending column order */
static void apply_matmul(const oc_cloth *c, const double *x, double *res)
{
    for (int v = 0; v < c->nv; v++)
        for (int a = 0; a < 3; a++)
        {
            double acc = 0.0;
            if (!c->pinned[v])
                for (int s = c->row_ptr[v]; s < c->row_ptr[v + 1]; s++)
                {
                    int col = c->cols[s];
                    if (c->pinned[col])
                        continue;
                    const double *w = c->W + 9 * (size_t)s + 3 * a;
                    acc = acc + x[3 * col + 0] * w[0];
                    acc = acc + x[3 * col + 1] * w[1];
                    acc = acc + x[3 * col + 2] * w[2];
                }
            res[3 * v + a] = acc;
        }
}

/* cBaraffCloth::Solve (BaraffCloth.cpp:507-611) and the tolerance variant */
static int solve(oc_cloth *c, int mode, double tol, int max_iters, oc_stats *st)
{
    int n = 3 * c->nv;
    double *x = c->dx;
    double *r = xcalloc(n, sizeof(double)), *Ad = xcalloc(n, sizeof(double));
    double *d = xcalloc(n, sizeof(double)), *P = xcalloc(n, sizeof(double));
    double *Pr = xcalloc(n, sizeof(double)), *rn = xcalloc(n, sizeof(double));
    double *Prn = xcalloc(n, sizeof(double)), *bf = xcalloc(n, sizeof(double));
    /* warm start x = xcur - xpre (:284) */
    for (int k = 0; k < n; k++)
        x[k] = c->x[k] - c->xpre[k];
    apply_matmul(c, x, Ad);
    for (int k = 0; k < n; k++)
        r[k] = c->b[k] - Ad[k];
    for (int j = 0; j < c->npin; j++)
        for (int a = 0; a < 3; a++)
        {
            r[3 * c->pins[j] + a] = 0.0;
            x[3 * c->pins[j] + a] = 0.0;
        }
    if (c->drag >= 0) /* fix_array.push_back(mDragVertexIdx) (BaraffCloth.cpp:291-292) */
        for (int a = 0; a < 3; a++)
        {
            r[3 * c->drag + a] = 0.0;
            x[3 * c->drag + a] = 0.0;
        }
    for (int v = 0; v < c->nv; v++)
        for (int a = 0; a < 3; a++)
        {
            const double *w = c->W + 9 * (size_t)c->diag[v];
            P[3 * v + a] = 1.0 / w[4 * a];
            bf[3 * v + a] = c->pinned[v] ? 0.0 : c->b[3 * v + a];
        }
    double bnorm = sqrt(oc_dot(bf, bf, n));
    for (int k = 0; k < n; k++)
        d[k] = P[k] * r[k];
    int iters = 1;
    int done = 0;
    if (mode == OC_SOLVE_TOLERANCE)
    {
        double rn0 = sqrt(oc_dot(r, r, n));
        if (rn0 <= tol * bnorm)
            done = 1;
    }
    int cap = mode == OC_SOLVE_REFERENCE ? c->max_iters : max_iters + 1;
    while (!done && iters < cap)
    {
        apply_matmul(c, d, Ad);
        for (int k = 0; k < n; k++)
            Pr[k] = P[k] * r[k];
        double deno = oc_dot(d, Ad, n);
        if (deno == 0)
            break;
        double alpha = oc_dot(r, Pr, n) / deno;
        for (int k = 0; k < n; k++)
            x[k] = x[k] + alpha * d[k];
        for (int k = 0; k < n; k++)
            rn[k] = r[k] - alpha * Ad[k];
        for (int k = 0; k < n; k++)
            Prn[k] = P[k] * rn[k];
        double beta = oc_dot(rn, Prn, n) / oc_dot(r, Pr, n);
        if (mode == OC_SOLVE_TOLERANCE)
        {
            double rr = sqrt(oc_dot(rn, rn, n));
            if (rr <= tol * bnorm)
            {
                memcpy(r, rn, n * sizeof(double));
                iters += 1;
                break;
            }
        }
        if (beta == 0)
            break;
        for (int k = 0; k < n; k++)
            d[k] = Prn[k] + beta * d[k];
        memcpy(r, rn, n * sizeof(double));
        iters += 1;
    }
    if (st)
    {
        st->iterations = iters - 1;
        double rr = sqrt(oc_dot(r, r, n));
        st->rel_residual = bnorm > 0 ? rr / bnorm : rr;
        apply_matmul(c, x, Ad);
        for (int v = 0; v < c->nv; v++)
            for (int a = 0; a < 3; a++)
                Ad[3 * v + a] = c->pinned[v] ? 0.0 : c->b[3 * v + a] - Ad[3 * v + a];
        double tr = sqrt(oc_dot(Ad, Ad, n));
        st->true_rel_residual = bnorm > 0 ? tr / bnorm : tr;
    }
    free(r); free(Ad); free(d); free(P); free(Pr); free(rn); free(Prn); free(bf);
    return 0;
}

int oc_step(oc_cloth *c, int mode, double tol, int max_iters, oc_stats *st)
{
    oc_assemble(c);
    /* the drag vertex is fixed for this solve only: its row and column are
       filtered like a static vertex's */
    const int drag_added = c->drag >= 0 && !c->pinned[c->drag];
    if (drag_added)
        c->pinned[c->drag] = 1;
    solve(c, mode, tol, max_iters, st);
    if (drag_added)
        c->pinned[c->drag] = 0;
    /* cleared after the step (BaraffCloth.cpp:154-155) */
    c->drag = -1;
    memset(c->fuser, 0, 3 * c->nv * sizeof(double));
    repulsion(c);
    oc_set_contacts(c, 0, NULL, NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL);
    int n = 3 * c->nv;
    /* mXpre = mXcur; mXcur = mXpre + dx (BaraffCloth.cpp:150-153) */
    for (int k = 0; k < n; k++)
    {
        c->xpre[k] = c->x[k];
        c->x[k] = c->xpre[k] + c->dx[k];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* accessors                                                           */
/* ------------------------------------------------------------------ */

int oc_num_vertices(const oc_cloth *c) { return c->nv; }
int oc_num_triangles(const oc_cloth *c) { return c->nt; }
int oc_num_edges(const oc_cloth *c) { return c->ne; }
int oc_num_pins(const oc_cloth *c) { return c->npin; }
int oc_ell_nnzb(const oc_cloth *c) { return c->nnzb; }
void oc_get_triangles(const oc_cloth *c, int *o) { memcpy(o, c->tri, 3 * c->nt * sizeof(int)); }
void oc_get_edges(const oc_cloth *c, int *o) { memcpy(o, c->edge, 5 * c->ne * sizeof(int)); }
void oc_get_stencils(const oc_cloth *c, int *o) { memcpy(o, c->stencil, 4 * c->ne * sizeof(int)); }
void oc_get_pins(const oc_cloth *c, int *o) { memcpy(o, c->pins, c->npin * sizeof(int)); }
void oc_get_uv(const oc_cloth *c, double *o)
{
    for (int k = 0; k < 2 * c->nv; k++)
        o[k] = c->uv[k];
}
void oc_get_rest_pos(const oc_cloth *c, double *o) { memcpy(o, c->xrest, 3 * c->nv * sizeof(double)); }
void oc_get_mass(const oc_cloth *c, double *o) { memcpy(o, c->mass, c->nv * sizeof(double)); }
void oc_get_gravity(const oc_cloth *c, double *o) { memcpy(o, c->gravity, 3 * c->nv * sizeof(double)); }
void oc_get_stretch_coef(const oc_cloth *c, double *coef12, double *area)
{
    memcpy(coef12, c->coef, 12 * c->nt * sizeof(double));
    memcpy(area, c->area, c->nt * sizeof(double));
}
void oc_get_bending_rest(const oc_cloth *c, double *el, double *h2, double *ek)
{
    for (int e = 0; e < c->ne; e++)
    {
        el[e] = c->edge_len[e];
        h2[2 * e] = c->height[2 * e];
        h2[2 * e + 1] = c->height[2 * e + 1];
    }
    *ek = c->edge_k;
}
void oc_get_pattern(const oc_cloth *c, int *rp, int *cols)
{
    memcpy(rp, c->row_ptr, (c->nv + 1) * sizeof(int));
    memcpy(cols, c->cols, c->nnzb * sizeof(int));
}
/* cBaseCloth::ApplyPerturb (BaseCloth.cpp:76-87): a third of the force on
   each vertex of the triangle, into the next step's user force */
int oc_apply_perturb(oc_cloth *c, int tri, const double *force)
{
    if (tri < 0 || tri >= c->nt)
        return -1;
    for (int k = 0; k < 3; k++)
    {
        int v = c->tri[3 * tri + k];
        for (int a = 0; a < 3; a++)
            c->fuser[3 * v + a] = c->fuser[3 * v + a] + force[a] / 3;
    }
    return 0;
}

/* cBaraffCloth::ApplyUserPerturbForceOnce (BaraffCloth.cpp:165-178): the
   triangle's first vertex moves to the goal (current and previous position)
   and is fixed in the next step's solve */
int oc_apply_user_perturb_force_once(oc_cloth *c, int tri, const double *goal)
{
    if (tri < 0 || tri >= c->nt)
        return -1;
    int v0 = c->tri[3 * tri];
    for (int a = 0; a < 3; a++)
    {
        c->x[3 * v0 + a] = goal[a];
        c->xpre[3 * v0 + a] = goal[a];
    }
    c->drag = v0;
    return 0;
}

/* cBaseCloth::mYDragForce (BaseCloth.h:66) */
void oc_set_y_drag_force(oc_cloth *c, float f) { c->ydrag = f; }

/* cBaseCloth::ClearForce (BaseCloth.cpp:66-73) for the user force */
void oc_clear_user_force(oc_cloth *c) { memset(c->fuser, 0, 3 * c->nv * sizeof(double)); }

void oc_set_state(oc_cloth *c, const double *x, const double *xpre)
{
    memcpy(c->x, x, 3 * c->nv * sizeof(double));
    memcpy(c->xpre, xpre, 3 * c->nv * sizeof(double));
}
void oc_get_pos(const oc_cloth *c, double *x) { memcpy(x, c->x, 3 * c->nv * sizeof(double)); }
void oc_get_pre_pos(const oc_cloth *c, double *x) { memcpy(x, c->xpre, 3 * c->nv * sizeof(double)); }
void oc_get_K(const oc_cloth *c, double *k9, int which)
{
    const double *src = which == 1 ? c->Ks : which == 2 ? c->Kb : c->K;
    memcpy(k9, src, 9 * (size_t)c->nnzb * sizeof(double));
}
void oc_get_W(const oc_cloth *c, double *w9) { memcpy(w9, c->W, 9 * (size_t)c->nnzb * sizeof(double)); }
void oc_get_force(const oc_cloth *c, double *f3, int which)
{
    const double *src = which == 1 ? c->fs : which == 2 ? c->fb : c->f;
    memcpy(f3, src, 3 * c->nv * sizeof(double));
}
void oc_get_rhs(const oc_cloth *c, double *b3) { memcpy(b3, c->b, 3 * c->nv * sizeof(double)); }
void oc_get_dx(const oc_cloth *c, double *dx3) { memcpy(dx3, c->dx, 3 * c->nv * sizeof(double)); }

/* cBaseObject::UpdateTriangleNormal (BaseObject.cpp:119-131): n_t =
   normalized((x1 - x0).cross3(x2 - x1)); UpdateVertexNormalFromTriangleNormal
   (:138-160): n_v = normalized(sum over t in ascending order of n_t * A_t),
   A_t the float rest area (CalcTriangleInitArea, :291-302).  Eigen's
   normalized() on the 4-wide vector (w = 0) divides by sqrt of the squared
   norm summed pairwise, (x^2 + z^2) + (y^2 + w^2), and leaves a zero vector
   as it is. */
static void oc_normalize3(double *u)
{
    double n2 = (u[0] * u[0] + u[2] * u[2]) + u[1] * u[1];
    if (n2 > 0)
    {
        double n = sqrt(n2);
        u[0] = u[0] / n;
        u[1] = u[1] / n;
        u[2] = u[2] / n;
    }
}

void oc_normals(const oc_cloth *c, double *vn3, double *tn3)
{
    double *tn = xcalloc(3 * (size_t)c->nt, sizeof(double));
    for (int t = 0; t < c->nt; t++)
    {
        const double *p0 = c->x + 3 * c->tri[3 * t], *p1 = c->x + 3 * c->tri[3 * t + 1];
        const double *p2 = c->x + 3 * c->tri[3 * t + 2];
        double a[3], b[3], *n = tn + 3 * t;
        v3sub(p1, p0, a);
        v3sub(p2, p1, b);
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        oc_normalize3(n);
    }
    if (vn3)
    {
        memset(vn3, 0, 3 * (size_t)c->nv * sizeof(double));
        for (int t = 0; t < c->nt; t++)
            for (int k = 0; k < 3; k++)
            {
                double *v = vn3 + 3 * c->tri[3 * t + k];
                for (int a = 0; a < 3; a++)
                    v[a] = v[a] + tn[3 * t + a] * (double)c->area[t];
            }
        for (int v = 0; v < c->nv; v++)
            oc_normalize3(vn3 + 3 * v);
    }
    if (tn3)
        memcpy(tn3, tn, 3 * (size_t)c->nt * sizeof(double));
    free(tn);
}

/* ------------------------------------------------------------------ */
/* the reference's GPU form                                            */
/* ------------------------------------------------------------------ */

/* CalcCotangent (QBendingMaterial.cpp:11-19) */
static double gf_cot(const double *A, const double *B, const double *C)
{
    double ab[3], ac[3];
    v3sub(B, A, ab);
    v3sub(C, A, ac);
    float abac = (float)(ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2]);
    float abab = (float)(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    float acac = (float)(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
    return abac / sqrtf(abab * acac - abac * abac);
}

/* cBaraffClothGpu::InitBendingMatrix / UpdateBendingMatrix
   (BaraffClothGpu.cpp:566-679) and cBaseCloth::InitMass (BaseCloth.cpp:418-447),
   InitGravity (BaraffClothGpu.cpp:556-567) */
static void gf_init(oc_cloth *c)
{
    if (c->gf_qb)
        return;
    c->gf_qb = xcalloc(c->nnzb, sizeof(float));
    c->gf_mass = xcalloc(c->nv, sizeof(float));
    c->gf_grav = xcalloc(c->nv, sizeof(float));
    c->gf_dx = xcalloc(3 * (size_t)c->nv, sizeof(float));
    double *shared = xcalloc(c->nv, sizeof(double));
    for (int t = 0; t < c->nt; t++)
        for (int k = 0; k < 3; k++)
            shared[c->tri[3 * t + k]] += c->area[t] / 3;
    for (int v = 0; v < c->nv; v++)
    {
        double m = shared[v] * (double)c->density_f;
        c->gf_mass[v] = (float)m;
        c->gf_grav[v] = (float)(-9.8 * m);
    }
    free(shared);
    float threshold = (float)(1.0 / sin(1.0 / 180 * M_PI));
    for (int e = 0; e < c->ne; e++)
    {
        const int *ed = c->edge + 5 * e;
        if (ed[4])
            continue;
        const int *sv = c->stencil + 4 * e;
        const double *x[4];
        for (int k = 0; k < 4; k++)
            x[k] = c->xrest + 3 * sv[k];
        double cc[4] = {gf_cot(x[0], x[1], x[2]), gf_cot(x[0], x[1], x[3]),
                        gf_cot(x[1], x[0], x[2]), gf_cot(x[1], x[0], x[3])};
        for (int k = 0; k < 4; k++)
            cc[k] = cc[k] < -threshold ? -threshold : (cc[k] > threshold ? threshold : cc[k]);
        double coef[4] = {cc[2] + cc[3], cc[0] + cc[1], -cc[0] - cc[2], -cc[1] - cc[3]};
        float a0 = (float)c->area[ed[2]], a1 = (float)c->area[ed[3]];
        float prefix = (float)(3.0 / (a0 + a1));
        float du = c->uv[2 * ed[1]] - c->uv[2 * ed[0]];
        float dv = c->uv[2 * ed[1] + 1] - c->uv[2 * ed[0] + 1];
        float n2 = du * du + dv * dv;
        if (n2 > 0)
        {
            float n = sqrtf(n2);
            du = du / n;
            dv = dv / n;
        }
        float theta = acosf(du * 1.0f + dv * 0.0f);
        if (theta > (float)(M_PI / 2.0f))
            theta = (float)(M_PI - theta);
        float ct = cosf(theta), st = sinf(theta);
        float Ke = c->Ku * ct * ct + c->Kv * st * st;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
            {
                float base = (float)(coef[i] * coef[j] * prefix);
                c->gf_qb[slot_of(c, sv[i], sv[j])] += base * Ke;
            }
    }
}

void oc_gpu_form_bending(oc_cloth *c, double *qb)
{
    gf_init(c);
    for (int k = 0; k < c->nnzb; k++)
        qb[k] = c->gf_qb[k];
}

/* one stretch part in FP32 (BaraffClothGpu.cu:7-105; Dv scaled by |Fu|) */
static void gf_part(float Ku, float Kv, const float *cu, const float *cv, float x[3][3],
                    float area, float f[9], float H[81])
{
    float Fu[3], Fv[3], Un[3], Vn[3], Du[9], Dv[9];
    for (int a = 0; a < 3; a++)
    {
        Fu[a] = cu[0] * x[0][a] + cu[1] * x[1][a] + cu[2] * x[2][a];
        Fv[a] = cv[0] * x[0][a] + cv[1] * x[1][a] + cv[2] * x[2][a];
    }
    float nu = sqrtf(Fu[0] * Fu[0] + Fu[1] * Fu[1] + Fu[2] * Fu[2]);
    float nv = sqrtf(Fv[0] * Fv[0] + Fv[1] * Fv[1] + Fv[2] * Fv[2]);
    for (int a = 0; a < 3; a++)
    {
        Un[a] = nu > 0 ? Fu[a] / nu : Fu[a];
        Vn[a] = nv > 0 ? Fv[a] / nv : Fv[a];
    }
    float Cu = nu - 1, Cv = nv - 1;
    for (int a = 0; a < 3; a++)
    {
        float ub = -area * Ku * Cu * Un[a], vb = -area * Kv * Cv * Vn[a];
        for (int i = 0; i < 3; i++)
            f[3 * i + a] += ub * cu[i] + vb * cv[i];
    }
    float iu = 1.0f / nu;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
        {
            Du[3 * a + b] = iu * Un[a] * Un[b] + ((a == b && nu > 1) ? 1 - iu : 0.0f);
            Dv[3 * a + b] = iu * Vn[a] * Vn[b] + ((a == b && nu > 1) ? 1 - iu : 0.0f);
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            float su = Ku * cu[i] * cu[j], sv = Kv * cv[i] * cv[j];
            for (int q = 0; q < 9; q++)
                H[9 * (3 * i + j) + q] += -area * (su * Du[q] + sv * Dv[q]);
        }
}

static void gf_mv(const float *M, const float *v, float *o)
{
    for (int a = 0; a < 3; a++)
        o[a] = M[3 * a] * v[0] + M[3 * a + 1] * v[1] + M[3 * a + 2] * v[2];
}

void oc_gpu_form_step(oc_cloth *c, oc_stats *st)
{
    gf_init(c);
    const int nv = c->nv, n3 = 3 * nv;
    const float dt = (float)c->dt;
    /* pins back to their targets (UpdateConstraintVertexPos) */
    for (int j = 0; j < c->npin; j++)
        for (int a = 0; a < 3; a++)
        {
            int q = 3 * c->pins[j] + a;
            c->x[q] = c->xpre[q] = (double)(float)c->xrest[q];
        }
    /* stretch per triangle */
    float *tri = xcalloc(90 * (size_t)c->nt, sizeof(float));
    for (int t = 0; t < c->nt; t++)
    {
        float x[3][3], cf[12];
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++)
                x[k][a] = (float)c->x[3 * c->tri[3 * t + k] + a];
        for (int k = 0; k < 12; k++)
            cf[k] = (float)c->coef[12 * t + k];
        float f[9] = {0}, H[81] = {0};
        float area = (float)c->area[t];
        gf_part((float)c->k_ww[0], (float)c->k_ww[1], cf, cf + 3, x, area, f, H);
        gf_part((float)c->k_bias, (float)c->k_bias, cf + 6, cf + 9, x, area, f, H);
        memcpy(tri + 90 * t, f, 9 * sizeof(float));
        memcpy(tri + 90 * t + 9, H, 81 * sizeof(float));
    }
    /* K blocks (triangle order), W, rhs, block-Jacobi inverse */
    float *K = xcalloc(9 * (size_t)c->nnzb, sizeof(float));
    float *fs = xcalloc(n3, sizeof(float));
    for (int t = 0; t < c->nt; t++)
        for (int i = 0; i < 3; i++)
        {
            int vi = c->tri[3 * t + i];
            for (int a = 0; a < 3; a++)
                fs[3 * vi + a] += tri[90 * t + 3 * i + a];
            for (int j = 0; j < 3; j++)
            {
                float *blk = K + 9 * (size_t)slot_of(c, vi, c->tri[3 * t + j]);
                for (int q = 0; q < 9; q++)
                    blk[q] += tri[90 * t + 9 + 9 * (3 * i + j) + q];
            }
        }
    float *W = xcalloc(9 * (size_t)c->nnzb, sizeof(float));
    float *b = xcalloc(n3, sizeof(float)), *Pinv = xcalloc(9 * (size_t)nv, sizeof(float));
    const float ks = dt * c->gf_b - dt * dt;
    for (int v = 0; v < nv; v++)
    {
        float fb[3] = {0, 0, 0};
        for (int k = c->row_ptr[v]; k < c->row_ptr[v + 1]; k++)
        {
            float *Kb = K + 9 * (size_t)k, qb = c->gf_qb[k];
            Kb[0] -= qb;
            Kb[4] -= qb;
            Kb[8] -= qb;
            for (int a = 0; a < 3; a++)
                fb[a] += -1.0f * qb * (float)c->x[3 * c->cols[k] + a];
            float *Wb = W + 9 * (size_t)k;
            for (int q = 0; q < 9; q++)
                Wb[q] = ks * Kb[q];
            if (c->cols[k] == v)
            {
                float md = c->gf_mass[v] * (1 + dt * c->gf_a);
                Wb[0] += md;
                Wb[4] += md;
                Wb[8] += md;
                float A = Wb[0], B = Wb[1], C = Wb[2], D = Wb[3], E = Wb[4], F = Wb[5],
                      G2 = Wb[6], H2 = Wb[7], I = Wb[8];
                float inv[9] = {E * I - F * H2, C * H2 - B * I, B * F - C * E,
                                F * G2 - D * I, A * I - C * G2, C * D - A * F,
                                D * H2 - E * G2, B * G2 - A * H2, A * E - B * D};
                float pre = 1.0f / (A * (E * I - F * H2) - B * (D * I - F * G2) + C * (D * H2 - E * G2));
                for (int q = 0; q < 9; q++)
                    Pinv[9 * v + q] = inv[q] * pre;
            }
        }
        float g[3] = {0.0f, c->gf_grav[v], 0.0f};
        for (int a = 0; a < 3; a++)
        {
            float f = fs[3 * v + a] + fb[a];
            b[3 * v + a] = dt * dt * (g[a] + f) +
                           ((float)c->x[3 * v + a] - (float)c->xpre[3 * v + a]) * c->gf_mass[v];
        }
    }
    /* PCG (PCGSolver.cu:356-481) */
    float *r = xcalloc(n3, sizeof(float)), *d = xcalloc(n3, sizeof(float));
    float *z = xcalloc(n3, sizeof(float)), *dx = c->gf_dx;
    double rmr = 0, rr = 0, bb = 0;
    for (int v = 0; v < nv; v++)
    {
        float y[3] = {0, 0, 0};
        for (int k = c->row_ptr[v]; k < c->row_ptr[v + 1]; k++)
        {
            float t[3];
            gf_mv(W + 9 * (size_t)k, dx + 3 * c->cols[k], t);
            for (int a = 0; a < 3; a++)
                y[a] += t[a];
        }
        float mr[3];
        for (int a = 0; a < 3; a++)
            r[3 * v + a] = c->pinned[v] ? 0.0f : b[3 * v + a] - y[a];
        gf_mv(Pinv + 9 * v, r + 3 * v, mr);
        for (int a = 0; a < 3; a++)
        {
            d[3 * v + a] = mr[a];
            rmr += (double)r[3 * v + a] * mr[a];
            rr += (double)r[3 * v + a] * r[3 * v + a];
            if (!c->pinned[v])
                bb += (double)b[3 * v + a] * b[3 * v + a];
        }
    }
    int it = 0;
    for (; it < 100; it++)
    {
        double dz = 0;
        for (int v = 0; v < nv; v++)
        {
            float y[3] = {0, 0, 0};
            if (!c->pinned[v])
                for (int k = c->row_ptr[v]; k < c->row_ptr[v + 1]; k++)
                {
                    if (c->pinned[c->cols[k]])
                        continue;
                    float t[3];
                    gf_mv(W + 9 * (size_t)k, d + 3 * c->cols[k], t);
                    for (int a = 0; a < 3; a++)
                        y[a] += t[a];
                }
            for (int a = 0; a < 3; a++)
            {
                z[3 * v + a] = y[a];
                dz += (double)d[3 * v + a] * y[a];
            }
        }
        float alpha = dz == 0.0 ? 0.0f : (float)(rmr / dz);
        double rmr_new = 0;
        rr = 0;
        for (int q = 0; q < n3; q++)
        {
            r[q] = r[q] - alpha * z[q];
            dx[q] = dx[q] + alpha * d[q];
        }
        for (int v = 0; v < nv; v++)
        {
            float mr[3];
            gf_mv(Pinv + 9 * v, r + 3 * v, mr);
            for (int a = 0; a < 3; a++)
            {
                rmr_new += (double)r[3 * v + a] * mr[a];
                rr += (double)r[3 * v + a] * r[3 * v + a];
            }
        }
        float beta = rmr == 0.0 ? 0.0f : (float)(rmr_new / rmr);
        for (int v = 0; v < nv; v++)
        {
            float mr[3];
            gf_mv(Pinv + 9 * v, r + 3 * v, mr);
            for (int a = 0; a < 3; a++)
                d[3 * v + a] = mr[a] + beta * d[3 * v + a];
        }
        rmr = rmr_new;
        if ((it % 10 == 0 || it == 99) && sqrt(rr) < 1e-10)
        {
            it++;
            break;
        }
    }
    for (int q = 0; q < n3; q++)
    {
        c->xpre[q] = c->x[q];
        c->x[q] = c->x[q] + (double)dx[q];
    }
    if (st)
    {
        st->iterations = it;
        st->rel_residual = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
        st->true_rel_residual = st->rel_residual;
    }
    free(tri); free(K); free(fs); free(W); free(b); free(Pinv); free(r); free(d); free(z);
}

/* ------------------------------------------------------------------ */
/* collision detection against a kinematic body                        */
/* ------------------------------------------------------------------ */

struct oc_body
{
    int nv, nt, ne;
    double *x;   /* 3 nv, current positions */
    int *tri;    /* 3 nt */
    int *edge;   /* 4 ne: id0, id1, tri0, tri1 (-1) */
    double *tn;  /* 3 nt, normals at creation */
    double *vn;  /* 3 nv, normals at creation */
};

/* Vector3d dot and norm: Eigen's vectorised sum (a0 b0 + a1 b1) + a2 b2 */
static double d3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

oc_body *oc_body_create(const double *xyz, int nv, const int *tri, int nt)
{
    if (nv < 3 || nt < 1)
        return NULL;
    for (int k = 0; k < 3 * nt; k++)
        if (tri[k] < 0 || tri[k] >= nv)
            return NULL;
    oc_body *b = xcalloc(1, sizeof(oc_body));
    b->nv = nv;
    b->nt = nt;
    b->x = xcalloc(3 * (size_t)nv, sizeof(double));
    memcpy(b->x, xyz, 3 * (size_t)nv * sizeof(double));
    b->tri = xcalloc(3 * (size_t)nt, sizeof(int));
    memcpy(b->tri, tri, 3 * (size_t)nt * sizeof(int));
    /* edges as cObjUtil::BuildEdge numbers them (ObjUtil.cpp:141-224):
       ascending {min, max}, tri0 the first triangle with the side */
    int ns = 3 * nt;
    oc_side *sd = xcalloc(ns, sizeof(oc_side));
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
        {
            int a = tri[3 * t + k], c = tri[3 * t + (k + 1) % 3];
            sd[3 * t + k].a = a < c ? a : c;
            sd[3 * t + k].b = a < c ? c : a;
            sd[3 * t + k].side = 3 * t + k;
        }
    qsort(sd, ns, sizeof(oc_side), side_cmp);
    b->edge = xcalloc(4 * (size_t)ns, sizeof(int));
    for (int i = 0; i < ns;)
    {
        int j = i + 1;
        while (j < ns && sd[j].a == sd[i].a && sd[j].b == sd[i].b)
            j++;
        if (j - i > 2)
        {
            free(sd);
            oc_body_destroy(b);
            return NULL;
        }
        int *e = b->edge + 4 * b->ne++;
        e[0] = sd[i].a;
        e[1] = sd[i].b;
        e[2] = sd[i].side / 3;
        e[3] = j - i == 2 ? sd[i + 1].side / 3 : -1;
        i = j;
    }
    free(sd);
    /* cKinematicBody::Init (KinematicBody.cpp:114-116): CalcTriangleInitArea
       (BaseObject.cpp:291-302, float 0.5 |(p1 - p0) x (p2 - p0)|),
       UpdateTriangleNormal, UpdateVertexNormalFromTriangleNormal */
    b->tn = xcalloc(3 * (size_t)nt, sizeof(double));
    b->vn = xcalloc(3 * (size_t)nv, sizeof(double));
    float *area = xcalloc(nt, sizeof(float));
    for (int t = 0; t < nt; t++)
    {
        const double *p0 = b->x + 3 * tri[3 * t], *p1 = b->x + 3 * tri[3 * t + 1];
        const double *p2 = b->x + 3 * tri[3 * t + 2];
        double u[3], v[3], w[3], *n = b->tn + 3 * t;
        v3sub(p1, p0, u);
        v3sub(p2, p1, v);
        n[0] = u[1] * v[2] - u[2] * v[1];
        n[1] = u[2] * v[0] - u[0] * v[2];
        n[2] = u[0] * v[1] - u[1] * v[0];
        oc_normalize3(n);
        v3sub(p2, p0, v);
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
        area[t] = (float)(0.5 * sqrt(d3(w, w)));
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
        {
            double *v = b->vn + 3 * tri[3 * t + k];
            for (int a = 0; a < 3; a++)
                v[a] = v[a] + b->tn[3 * t + a] * (double)area[t];
        }
    for (int v = 0; v < nv; v++)
        oc_normalize3(b->vn + 3 * v);
    free(area);
    return b;
}

void oc_body_destroy(oc_body *b)
{
    if (!b)
        return;
    free(b->x); free(b->tri); free(b->edge); free(b->tn); free(b->vn);
    free(b);
}

void oc_body_set_pos(oc_body *b, const double *xyz) { memcpy(b->x, xyz, 3 * (size_t)b->nv * sizeof(double)); }
int oc_body_num_edges(const oc_body *b) { return b->ne; }
void oc_body_get_edges(const oc_body *b, int *e4) { memcpy(e4, b->edge, 4 * (size_t)b->ne * sizeof(int)); }
void oc_body_normals(const oc_body *b, double *vn3, double *tn3)
{
    if (vn3)
        memcpy(vn3, b->vn, 3 * (size_t)b->nv * sizeof(double));
    if (tn3)
        memcpy(tn3, b->tn, 3 * (size_t)b->nt * sizeof(double));
}

/* ClosestPtPointTriangle (BVHCollisionDetecter.cpp:17-83, Ericson's
   region test): dot products in double rounded to float, region
   arithmetic in float, the point in double */
void oc_closest_pt_triangle(const double *p, const double *a, const double *b, const double *c,
                            float *bary, double *out)
{
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    v3sub(b, a, ab);
    v3sub(c, a, ac);
    v3sub(p, a, ap);
    float d1 = (float)d3(ab, ap), d2 = (float)d3(ac, ap);
    bary[0] = bary[1] = bary[2] = 0.0f;
    if (d1 <= 0.0f && d2 <= 0.0f)
    {
        bary[0] = 1;
        memcpy(out, a, 3 * sizeof(double));
        return;
    }
    v3sub(p, b, bp);
    float d3f = (float)d3(ab, bp), d4 = (float)d3(ac, bp);
    if (d3f >= 0.0f && d4 <= d3f)
    {
        bary[1] = 1;
        memcpy(out, b, 3 * sizeof(double));
        return;
    }
    float vc = d1 * d4 - d3f * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3f <= 0.0f)
    {
        float v = d1 / (d1 - d3f);
        bary[0] = 1 - v;
        bary[1] = v;
        for (int k = 0; k < 3; k++)
            out[k] = a[k] + v * ab[k];
        return;
    }
    v3sub(p, c, cp);
    float d5 = (float)d3(ab, cp), d6 = (float)d3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6)
    {
        bary[2] = 1;
        memcpy(out, c, 3 * sizeof(double));
        return;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f)
    {
        float w = d2 / (d2 - d6);
        bary[0] = 1 - w;
        bary[2] = w;
        for (int k = 0; k < 3; k++)
            out[k] = a[k] + w * ac[k];
        return;
    }
    float va = d3f * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3f) >= 0.0f && (d5 - d6) >= 0.0f)
    {
        float w = (d4 - d3f) / ((d4 - d3f) + (d5 - d6));
        bary[1] = 1 - w;
        bary[2] = w;
        for (int k = 0; k < 3; k++)
            out[k] = b[k] + w * (c[k] - b[k]);
        return;
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    bary[0] = 1 - v - w;
    bary[1] = v;
    bary[2] = w;
    for (int k = 0; k < 3; k++)
        out[k] = a[k] + ab[k] * v + ac[k] * w;
}

/* cMathUtil::Clamp (MathUtil.cpp:31-34): max(lo, min(v, hi)) in double */
static float clampf01(float v)
{
    double m = (1.0 < (double)v) ? 1.0 : (double)v;
    return (float)((0.0 < m) ? m : 0.0);
}

/* EdgeEdgeDist (BVHCollisionDetecter.cpp:104-177, Ericson's closest points
   of two segments) with its float parameters: s on p1q1, t on p2q2 */
void oc_edge_edge_params(const double *p1, const double *q1, const double *p2, const double *q2,
                         float *s_out, float *t_out)
{
    double dd1[3], dd2[3], r[3];
    v3sub(q1, p1, dd1);
    v3sub(q2, p2, dd2);
    v3sub(p1, p2, r);
    float a = (float)d3(dd1, dd1), e = (float)d3(dd2, dd2), f = (float)d3(dd2, r);
    const float EPSILON = 1e-10;
    float s = 0, t = 0;
    if (a <= EPSILON && e <= EPSILON)
        s = t = 0.0f;
    if (a <= EPSILON)
    {
        s = 0.0f;
        t = f / e;
        t = clampf01(t);
    }
    else
    {
        float c = (float)d3(dd1, r);
        if (e <= EPSILON)
        {
            t = 0.0f;
            s = clampf01(-c / a);
        }
        else
        {
            float bb = (float)d3(dd1, dd2);
            float denom = a * e - bb * bb;
            if (denom != 0.0f)
                s = clampf01((bb * f - c * e) / denom);
            else
                s = 0.0f;
            t = (bb * s + f) / e;
            if (t < 0.0f)
            {
                t = 0.0f;
                s = clampf01(-c / a);
            }
            else if (t > 1.0f)
            {
                t = 1.0f;
                s = clampf01((bb - c) / a);
            }
        }
    }
    *s_out = s;
    *t_out = t;
}

static void normalized3d(const double *u, double *o)
{
    double z = d3(u, u);
    if (z > 0)
    {
        double n = sqrt(z);
        for (int k = 0; k < 3; k++)
            o[k] = u[k] / n;
    }
    else
        memcpy(o, u, 3 * sizeof(double));
}

typedef struct
{
    int key0, key1;
    int ids[3];
    double normal[3], bary[3];
    float depth, eb[2];
} oc_hit;

static int hit_cmp(const void *p, const void *q)
{
    const oc_hit *x = p, *y = q;
    if (x->key0 != y->key0)
        return x->key0 < y->key0 ? -1 : 1;
    return x->key1 < y->key1 ? -1 : (x->key1 > y->key1);
}

/* cBVHCollisionDetecter::PerformCD for the pair {cloth, body}
   (BVHCollisionDetecter.cpp:360-651).  The broadphase (leaf boxes of each
   triangle grown by 5e-3, IntersectTwoBVHs :187-240) keeps every pair
   closer than 1e-2, so the contacts are all pairs whose computed distance
   is below 3e-3: (point of obj0, triangle of obj1), then (point of obj1,
   triangle of obj0), each in ascending (point, triangle) order (std::set,
   :291-331), then edge pairs in ascending (edge of obj0, edge of obj1).
   obj0 is the object with the lower id: the cloth when cloth_first.
   Vertices in no triangle are never candidates.  The cloth's triangle
   normals are those at its current positions (cBaseCloth::Update refreshes
   them after every substep); the body's are those of its creation
   (cKinematicBody::Init).  The contacts replace the cloth's for its next
   step; returns 0, -1 for a non-manifold body. */
int oc_detect_contacts(oc_cloth *c, const oc_body *b, int cloth_first)
{
    double *ctn = xcalloc(3 * (size_t)c->nt, sizeof(double));
    oc_normals(c, NULL, ctn);
    unsigned char *c_in = xcalloc(c->nv, 1), *b_in = xcalloc(b->nv, 1);
    for (int k = 0; k < 3 * c->nt; k++)
        c_in[c->tri[k]] = 1;
    for (int k = 0; k < 3 * b->nt; k++)
        b_in[b->tri[k]] = 1;
    size_t cap = 64, n01 = 0, n10 = 0, nee = 0;
    oc_hit *h01 = xcalloc(cap, sizeof(oc_hit)), *h10 = xcalloc(cap, sizeof(oc_hit));
    size_t cap01 = cap, cap10 = cap, capee = cap;
    oc_hit *hee = xcalloc(cap, sizeof(oc_hit));
#define PUSH(arr, n, capv, hit)                                      \
    do                                                               \
    {                                                                \
        if (n == capv)                                               \
        {                                                            \
            capv *= 2;                                               \
            arr = realloc(arr, capv * sizeof(oc_hit));               \
        }                                                            \
        arr[n++] = hit;                                              \
    } while (0)
    /* cloth vertex against body triangle: obj0_is_cloth_obj1_is_rb = true */
    for (int p = 0; p < c->nv; p++)
    {
        if (!c_in[p])
            continue;
        const double *pt = c->x + 3 * p;
        for (int t = 0; t < b->nt; t++)
        {
            const int *tv = b->tri + 3 * t;
            float bary[3];
            double cl[3], dv[3];
            oc_closest_pt_triangle(pt, b->x + 3 * tv[0], b->x + 3 * tv[1], b->x + 3 * tv[2], bary, cl);
            v3sub(pt, cl, dv);
            if (sqrt(d3(dv, dv)) < 3e-3)
            {
                oc_hit hh;
                memset(&hh, 0, sizeof hh);
                hh.key0 = p;
                hh.key1 = t;
                hh.ids[0] = 1;
                hh.ids[1] = p;
                hh.ids[2] = t;
                memcpy(hh.normal, b->tn + 3 * t, 3 * sizeof(double));
                hh.depth = (float)d3(dv, hh.normal);
                for (int k = 0; k < 3; k++)
                    hh.bary[k] = (double)bary[k];
                if (cloth_first)
                    PUSH(h01, n01, cap01, hh);
                else
                    PUSH(h10, n10, cap10, hh);
            }
        }
    }
    /* body vertex against cloth triangle: obj0_is_cloth_obj1_is_rb = false */
    for (int p = 0; p < b->nv; p++)
    {
        if (!b_in[p])
            continue;
        const double *pt = b->x + 3 * p;
        for (int t = 0; t < c->nt; t++)
        {
            const int *tv = c->tri + 3 * t;
            float bary[3];
            double cl[3], dv[3], n[3];
            oc_closest_pt_triangle(pt, c->x + 3 * tv[0], c->x + 3 * tv[1], c->x + 3 * tv[2], bary, cl);
            memcpy(n, ctn + 3 * t, sizeof n);
            if (d3(n, b->vn + 3 * p) < 0)
                for (int k = 0; k < 3; k++)
                    n[k] = n[k] * -1;
            v3sub(cl, pt, dv);
            if (sqrt(d3(dv, dv)) < 3e-3)
            {
                oc_hit hh;
                memset(&hh, 0, sizeof hh);
                hh.key0 = p;
                hh.key1 = t;
                hh.ids[0] = 0;
                hh.ids[1] = p;
                hh.ids[2] = t;
                memcpy(hh.normal, n, sizeof n);
                hh.depth = (float)d3(dv, n);
                for (int k = 0; k < 3; k++)
                    hh.bary[k] = (double)bary[k];
                if (cloth_first)
                    PUSH(h10, n10, cap10, hh);
                else
                    PUSH(h01, n01, cap01, hh);
            }
        }
    }
    /* edge pairs: e0 of obj0, e1 of obj1 */
    for (int ce = 0; ce < c->ne; ce++)
        for (int be = 0; be < b->ne; be++)
        {
            const int *cev = c->edge + 5 * ce, *bev = b->edge + 4 * be;
            /* p0 p1 = obj0's edge, p2 p3 = obj1's */
            const double *p0, *p1, *p2, *p3;
            if (cloth_first)
            {
                p0 = c->x + 3 * cev[0]; p1 = c->x + 3 * cev[1];
                p2 = b->x + 3 * bev[0]; p3 = b->x + 3 * bev[1];
            }
            else
            {
                p0 = b->x + 3 * bev[0]; p1 = b->x + 3 * bev[1];
                p2 = c->x + 3 * cev[0]; p3 = c->x + 3 * cev[1];
            }
            float s, t;
            oc_edge_edge_params(p0, p1, p2, p3, &s, &t);
            double e0p[3], e1p[3], u[3], v[3], cr[3], outer[3], on[3], dv[3];
            for (int k = 0; k < 3; k++)
            {
                e0p[k] = (double)(1 - s) * p0[k] + (double)s * p1[k];
                e1p[k] = (double)(1 - t) * p2[k] + (double)t * p3[k];
            }
            v3sub(p1, p0, u);
            v3sub(p3, p2, v);
            cr[0] = u[1] * v[2] - u[2] * v[1];
            cr[1] = u[2] * v[0] - u[0] * v[2];
            cr[2] = u[0] * v[1] - u[1] * v[0];
            normalized3d(cr, outer);
            /* the body edge's triangle normals set the side */
            memcpy(on, b->tn + 3 * bev[2], sizeof on);
            if (bev[3] != -1)
                for (int k = 0; k < 3; k++)
                    on[k] = on[k] + b->tn[3 * bev[3] + k];
            normalized3d(on, on);
            if (d3(on, outer) < 0)
                for (int k = 0; k < 3; k++)
                    outer[k] = outer[k] * -1;
            /* from the body to the cloth */
            if (cloth_first)
                v3sub(e0p, e1p, dv);
            else
                v3sub(e1p, e0p, dv);
            if (sqrt(d3(dv, dv)) < 3e-3)
            {
                oc_hit hh;
                memset(&hh, 0, sizeof hh);
                hh.key0 = cloth_first ? ce : be;
                hh.key1 = cloth_first ? be : ce;
                hh.ids[0] = cloth_first ? 1 : 0;
                hh.ids[1] = hh.key0;
                hh.ids[2] = hh.key1;
                memcpy(hh.normal, outer, sizeof outer);
                hh.depth = (float)d3(dv, outer);
                hh.eb[0] = s;
                hh.eb[1] = t;
                PUSH(hee, nee, capee, hh);
            }
        }
#undef PUSH
    qsort(h01, n01, sizeof(oc_hit), hit_cmp);
    qsort(h10, n10, sizeof(oc_hit), hit_cmp);
    qsort(hee, nee, sizeof(oc_hit), hit_cmp);
    int npt = (int)(n01 + n10), ne = (int)nee;
    int *pid = xcalloc(3 * (size_t)npt + 1, sizeof(int)), *eid = xcalloc(3 * (size_t)ne + 1, sizeof(int));
    double *pn = xcalloc(3 * (size_t)npt + 1, sizeof(double)), *pb = xcalloc(3 * (size_t)npt + 1, sizeof(double));
    double *en = xcalloc(3 * (size_t)ne + 1, sizeof(double));
    float *pd = xcalloc(npt + 1, sizeof(float)), *eb = xcalloc(2 * (size_t)ne + 1, sizeof(float));
    float *ed = xcalloc(ne + 1, sizeof(float));
    for (int k = 0; k < npt; k++)
    {
        const oc_hit *hh = k < (int)n01 ? &h01[k] : &h10[k - n01];
        memcpy(pid + 3 * k, hh->ids, 3 * sizeof(int));
        memcpy(pn + 3 * k, hh->normal, 3 * sizeof(double));
        memcpy(pb + 3 * k, hh->bary, 3 * sizeof(double));
        pd[k] = hh->depth;
    }
    for (int k = 0; k < ne; k++)
    {
        memcpy(eid + 3 * k, hee[k].ids, 3 * sizeof(int));
        memcpy(en + 3 * k, hee[k].normal, 3 * sizeof(double));
        eb[2 * k] = hee[k].eb[0];
        eb[2 * k + 1] = hee[k].eb[1];
        ed[k] = hee[k].depth;
    }
    /* the repulsion asserts 3e-3 - depth > 0 for point-triangle contacts;
       the detector only records distances below 3e-3, and |depth| <= distance */
    oc_set_contacts(c, npt, pid, pn, pb, pd, ne, eid, en, eb, ed);
    free(pid); free(eid); free(pn); free(pb); free(en); free(pd); free(eb); free(ed);
    free(h01); free(h10); free(hee); free(ctn); free(c_in); free(b_in);
    return 0;
}

void oc_num_contacts(const oc_cloth *c, int *n_pt, int *n_ee)
{
    *n_pt = c->n_pt;
    *n_ee = c->n_ee;
}

void oc_get_contacts(const oc_cloth *c, int *pt_ids, double *pt_normal, double *pt_bary,
                     float *pt_depth, int *ee_ids, double *ee_normal, float *ee_bary, float *ee_depth)
{
    if (c->n_pt)
    {
        memcpy(pt_ids, c->pt_ids, 3 * (size_t)c->n_pt * sizeof(int));
        memcpy(pt_normal, c->pt_normal, 3 * (size_t)c->n_pt * sizeof(double));
        memcpy(pt_bary, c->pt_bary, 3 * (size_t)c->n_pt * sizeof(double));
        memcpy(pt_depth, c->pt_depth, (size_t)c->n_pt * sizeof(float));
    }
    if (c->n_ee)
    {
        memcpy(ee_ids, c->ee_ids, 3 * (size_t)c->n_ee * sizeof(int));
        memcpy(ee_normal, c->ee_normal, 3 * (size_t)c->n_ee * sizeof(double));
        memcpy(ee_bary, c->ee_bary, 2 * (size_t)c->n_ee * sizeof(float));
        memcpy(ee_depth, c->ee_depth, (size_t)c->n_ee * sizeof(float));
    }
}
