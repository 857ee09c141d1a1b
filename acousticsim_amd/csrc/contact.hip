// Collision response of the plain method for contacts handed in before a
// step (the output of the scene's collision detection):
//
//   - the contact Hessian (cBaraffCloth::PrepareCollisionHessian,
//     BaraffCloth.cpp:928-990) added to W before the solve (:293).  The
//     blocks depend on the contact records only, so the host sums them per
//     block in the reference's triplet order; k_contact_hessian adds each
//     summed block to W (FP64, and its FP32 copy) and refreshes the Jacobi
//     diagonal of diagonal blocks.
//   - the repulsion impulses (cBaraffCloth::Repulsion, :626-926) added to
//     the solution dx after the solve (:147).  The reference applies the
//     contacts one after another, each reading the dx its predecessors left,
//     so the host assigns each contact a level one above the last earlier
//     contact that shares a cloth vertex; contacts of one level touch
//     disjoint vertices and run in parallel, levels in order.  The result is
//     the serial one, bit for bit.  The touched rows of dx stay in LDS for
//     the pass when they fit, and levels are staged through LDS in chunks
//     (chains can be thousands of levels deep).
//
// Contraction is off for this file so that every product and sum rounds as
// in the reference.
#include "device.h"

namespace bw
{
namespace
{

__global__ void __launch_bounds__(256)
k_contact_hessian(int n, const long long *__restrict__ w_index, const int *__restrict__ diag_row,
                  const double *__restrict__ h9, double *__restrict__ W64, float *__restrict__ W32,
                  double *__restrict__ P64, float *__restrict__ P32)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const long long base = w_index[k];
    double w[9];
    for (int q = 0; q < 9; q++)
    {
        const long long at = base + (long long)q * SLICE;
        w[q] = W64[at] + h9[9 * (size_t)k + q];
        W64[at] = w[q];
        if (W32)
            W32[at] = (float)w[q];
    }
    const int p = diag_row[k];
    if (p >= 0)
        for (int a = 0; a < 3; a++)
        {
            const double pinv = 1.0 / w[4 * a];
            P64[3 * (size_t)p + a] = pinv;
            if (P32)
                P32[3 * (size_t)p + a] = (float)pinv;
        }
}

// one contact's impulse (BaraffCloth.cpp:640-780 edge-edge, :781-905
// point-triangle) on the solution rows x0, x1, x2 (the contact's vertices);
// rec = {N[3], w[3], mass, den, d, coef, dt*K, dt}
__device__ void apply_contact(double *x0, double *x1, double *x2, int type,
                              const double *__restrict__ rec)
{
    const double N[3] = {rec[0], rec[1], rec[2]};
    const double w0 = rec[3], w1 = rec[4], w2 = rec[5];
    const double mass = rec[6], den = rec[7], d = rec[8], coef = rec[9], dtK = rec[10],
                 dt = rec[11];
    double vrel[3];
    for (int a = 0; a < 3; a++)
    {
        if (type == 0) // edge-edge: the cloth edge's point
            vrel[a] = (w0 * x0[a] + w1 * x1[a]) / dt;
        else if (type == 1) // cloth vertex against a body triangle
            vrel[a] = x0[a] / dt;
        else // body vertex against a cloth triangle
            vrel[a] = (x0[a] * w0 + x1[a] * w1 + x2[a] * w2) / dt;
    }
    const double vn = (vrel[0] * N[0] + vrel[1] * N[1]) + vrel[2] * N[2];
    const double Ic = vn > 0.01 ? 0.0 : mass * fabs(vn) / 2;
    double Ir = 0.0;
    if (!(vn > coef * d / dt))
    {
        const double Ir0 = dtK * d;
        const double Ir1 = mass * (coef * d / dt - vn);
        Ir = Ir0 < Ir1 ? Ir0 : Ir1;
    }
    const double It = 2 * (Ic + Ir) / den;
    const double scale = (double)0.1f;
    if (type == 0)
    {
        const double s0 = scale * dt * w0 * It / mass;
        const double s1 = scale * dt * w1 * It / mass;
        for (int a = 0; a < 3; a++)
            x0[a] = x0[a] + s0 * N[a];
        for (int a = 0; a < 3; a++)
            x1[a] = x1[a] + s1 * N[a];
    }
    else if (type == 1)
    {
        const double f = dt * It / mass;
        for (int a = 0; a < 3; a++)
            x0[a] = x0[a] + scale * (f * N[a]);
    }
    else
    {
        const double f0 = dt * w0 * It / mass, f1 = dt * w1 * It / mass, f2 = dt * w2 * It / mass;
        for (int a = 0; a < 3; a++)
            x0[a] = x0[a] + scale * (f0 * N[a]);
        for (int a = 0; a < 3; a++)
            x1[a] = x1[a] + scale * (f1 * N[a]);
        for (int a = 0; a < 3; a++)
            x2[a] = x2[a] + scale * (f2 * N[a]);
    }
}

// One single-wave workgroup per cloth with contacts: its levels in order,
// the contacts of a level spread over the wave's lanes.  A contact names its
// vertices by slot in the cloth's list of touched rows; when that list fits,
// the touched rows of dx live in LDS for the whole pass, else they are used
// in place.  Levels come in chunks of whole levels holding at most
// REPULSION_CHUNK contacts, whose records are staged in LDS with one
// coalesced load, so that a level costs its arithmetic and a barrier
// (contact chains through shared vertices can be thousands of levels deep);
// a level wider than a chunk runs from memory.
__global__ void __launch_bounds__(64)
k_repulsion(double *__restrict__ dx, const int *__restrict__ wg_chunk,
            const int *__restrict__ chunk_level, const int *__restrict__ level_ptr,
            const int *__restrict__ wg_rows, const int *__restrict__ rows,
            const int4 *__restrict__ ids, const double *__restrict__ rec)
{
    __shared__ double s_dx[3 * REPULSION_LDS_ROWS];
    __shared__ double s_rec[REPULSION_CHUNK * CONTACT_REC];
    __shared__ int4 s_id[REPULSION_CHUNK];
    __shared__ int s_lp[REPULSION_CHUNK + 1];
    const int g = blockIdx.x, lane = threadIdx.x, nl = blockDim.x;
    const int *my_rows = rows + wg_rows[g];
    const int nr = wg_rows[g + 1] - wg_rows[g];
    const bool lds = nr <= REPULSION_LDS_ROWS;
    if (lds)
    {
        for (int k = lane; k < 3 * nr; k += nl)
            s_dx[k] = dx[3 * (size_t)my_rows[k / 3] + k % 3];
        __syncthreads();
    }
    auto row = [&](int sl) { return lds ? s_dx + 3 * sl : dx + 3 * (size_t)my_rows[sl]; };
    for (int c = wg_chunk[g]; c < wg_chunk[g + 1]; c++)
    {
        const int L0 = chunk_level[c], L1 = chunk_level[c + 1];
        const int k0 = level_ptr[L0], k1 = level_ptr[L1];
        if (k1 - k0 <= REPULSION_CHUNK)
        {
            for (int k = lane; k < k1 - k0; k += nl)
                s_id[k] = ids[k0 + k];
            for (int k = lane; k < (k1 - k0) * CONTACT_REC; k += nl)
                s_rec[k] = rec[CONTACT_REC * (size_t)k0 + k];
            for (int L = L0 + lane; L <= L1; L += nl)
                s_lp[L - L0] = level_ptr[L] - k0;
            __syncthreads();
            for (int L = 0; L < L1 - L0; L++)
            {
                for (int k = s_lp[L] + lane; k < s_lp[L + 1]; k += nl)
                {
                    const int4 id = s_id[k];
                    const int s1 = id.x == 1 ? id.y : id.z, s2 = id.x == 2 ? id.w : id.y;
                    apply_contact(row(id.y), row(s1), row(s2), id.x, s_rec + CONTACT_REC * k);
                }
                __syncthreads(); // this level's writes before the next level reads
            }
        }
        else // one level wider than a chunk
        {
            for (int k = k0 + lane; k < k1; k += nl)
            {
                const int4 id = ids[k];
                const int s1 = id.x == 1 ? id.y : id.z, s2 = id.x == 2 ? id.w : id.y;
                apply_contact(row(id.y), row(s1), row(s2), id.x, rec + CONTACT_REC * (size_t)k);
            }
            __syncthreads();
        }
    }
    if (lds)
        for (int k = lane; k < 3 * nr; k += nl)
            dx[3 * (size_t)my_rows[k / 3] + k % 3] = s_dx[k];
}

} // namespace

void launch_contact_hessian(const DeviceBatch &d, int n, const long long *w_index,
                            const int *diag_row, const double *h9, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_contact_hessian, dim3((n + 255) / 256), dim3(256), 0, s, n, w_index,
                           diag_row, h9, d.W64, d.W32, d.P64, d.P32);
}

void launch_repulsion(const DeviceBatch &d, int n_groups, const int *wg_chunk,
                      const int *chunk_level, const int *level_ptr, const int *wg_rows,
                      const int *rows, const int4 *ids, const double *rec, hipStream_t s)
{
    if (n_groups > 0)
        hipLaunchKernelGGL(k_repulsion, dim3(n_groups), dim3(64), 0, s, d.dx, wg_chunk,
                           chunk_level, level_ptr, wg_rows, rows, ids, rec);
}

} // namespace bw
