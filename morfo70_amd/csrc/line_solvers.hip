// line_solvers.hip -- tridiagonal line solvers of the Hydro ADI step for gfx950.
//
// Both kernels run the reference's recurrences unchanged, one line per lane, so each
// line's result is bitwise the reference's:
//   * k_thomas_cols: Math::TridiagonalSolver (src/Math.cpp:15-65). Lines are grid columns
//     (stride ld), so lane j reads element (p, j) and a wavefront reads 512 contiguous
//     bytes per row. Rows are fetched in chunks, the next chunk's loads issued before the
//     current chunk's recurrence runs (register double buffering).
//   * k_periodic_rows: Math::TridiagonalSolverPeriodic (El-Mikkawy 2005,
//     src/Math.cpp:68-118). Lines are grid rows (contiguous), so a wavefront owning 64
//     rows stages each chunk of PC positions through LDS: the global loads cover 128-byte
//     row segments and the LDS tile is read transposed, one row per lane. The forward
//     recurrences (k, inv_k, v, h, r and the rh/vh sums) are fused into one pass in the
//     reference's order; r, v, inv_k go to line-major scratch (coalesced) for the
//     backward pass.
#include "hydro_device.h"
#include "hydro_kernels.h"

namespace m70 {

// ------------------------------------------------------------------------------------
// Thomas along columns
// ------------------------------------------------------------------------------------
constexpr int TCH = 8;   // rows per prefetch chunk

struct TStage {
    double a[TCH], b[TCH], c[TCH], f[TCH];
};

__device__ __forceinline__ void t_load(TStage &s, const double *__restrict__ A, const double *__restrict__ B,
                                       const double *__restrict__ C, const double *__restrict__ F, int p0, int n,
                                       long ld, int j)
{
#pragma unroll
    for (int t = 0; t < TCH; ++t) {
        const int p = p0 + t;
        if (p < n) {
            const long o = (long)p * ld + j;
            s.a[t] = A[o];
            s.b[t] = B[o];
            s.c[t] = C[o - ld];   // c of row p-1
            s.f[t] = F[o];
        }
    }
}

__device__ __forceinline__ void t_fwd(const TStage &s, double &inv_w, double &y, double *__restrict__ Y,
                                      double *__restrict__ V, int p0, int n, long ld, int j)
{
#pragma unroll
    for (int t = 0; t < TCH; ++t) {
        const int p = p0 + t;
        if (p < n) {
            const long o = (long)p * ld + j;
            const double v = s.c[t] * inv_w;
            V[o - ld] = v;
            inv_w = 1 / (s.a[t] - s.b[t] * v);
            y = (s.f[t] - s.b[t] * y) * inv_w;
            Y[o] = y;
        }
    }
}

struct TBack {
    double y[TCH], v[TCH];
};

__device__ __forceinline__ void tb_load(TBack &s, const double *__restrict__ Y, const double *__restrict__ V,
                                        int p0, long ld, int j)
{   // rows p0, p0-1, ..., p0-TCH+1 (descending)
#pragma unroll
    for (int t = 0; t < TCH; ++t) {
        const int p = p0 - t;
        if (p >= 0) {
            const long o = (long)p * ld + j;
            s.y[t] = Y[o];
            s.v[t] = V[o];
        }
    }
}

__device__ __forceinline__ void tb_step(const TBack &s, double &yn, double *__restrict__ Y, int p0, long ld,
                                        int j)
{
#pragma unroll
    for (int t = 0; t < TCH; ++t) {
        const int p = p0 - t;
        if (p >= 0) {
            yn = s.y[t] - s.v[t] * yn;
            Y[(long)p * ld + j] = yn;
        }
    }
}

// Solves system j of n unknowns with diagonal A[p*ld+j], lower B, upper C and
// right-hand side F, p = 0..n-1. Solution to Y[p*ld+j]; V is n*ld scratch.
__device__ __forceinline__ void thomas_col(const double *__restrict__ A, const double *__restrict__ B,
                                           const double *__restrict__ C, const double *__restrict__ F,
                                           double *__restrict__ Y, double *__restrict__ V, int n, long ld, int j)
{
    // forward elimination, src/Math.cpp:20-48
    double inv_w = 1 / A[j];
    double y = F[j] * inv_w;
    Y[j] = y;
    TStage s0, s1;
    int p0 = 1;
    t_load(s0, A, B, C, F, p0, n, ld, j);
    while (true) {
        if (p0 >= n) break;
        t_load(s1, A, B, C, F, p0 + TCH, n, ld, j);
        t_fwd(s0, inv_w, y, Y, V, p0, n, ld, j);
        p0 += TCH;
        if (p0 >= n) break;
        t_load(s0, A, B, C, F, p0 + TCH, n, ld, j);
        t_fwd(s1, inv_w, y, Y, V, p0, n, ld, j);
        p0 += TCH;
    }
    // back substitution, src/Math.cpp:49-63: Y[p] -= V[p]*Y[p+1], p = n-2..0
    double yn = y;
    TBack b0, b1;
    int q0 = n - 2;
    tb_load(b0, Y, V, q0, ld, j);
    while (true) {
        if (q0 < 0) break;
        tb_load(b1, Y, V, q0 - TCH, ld, j);
        tb_step(b0, yn, Y, q0, ld, j);
        q0 -= TCH;
        if (q0 < 0) break;
        tb_load(b0, Y, V, q0 - TCH, ld, j);
        tb_step(b1, yn, Y, q0, ld, j);
        q0 -= TCH;
    }
}

__global__ void __launch_bounds__(64) k_thomas_cols(const double *__restrict__ A, const double *__restrict__ B,
                                                    const double *__restrict__ C, const double *__restrict__ F,
                                                    double *__restrict__ Y, double *__restrict__ V, int n, int m,
                                                    long ld)
{
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= m) return;
    thomas_col(A, B, C, F, Y, V, n, ld, j);
}


// ------------------------------------------------------------------------------------
// Periodic (El-Mikkawy) along rows
// ------------------------------------------------------------------------------------
constexpr int PC = 16;          // positions per chunk
constexpr int LW = 64;          // lines per workgroup (one wavefront)
constexpr int LP = LW + 1;      // LDS pitch
constexpr int LPT = LW / PC;    // lines covered by one staged load instruction

struct PStage {
    double x[4][PC];
};

// Row-major source (line L at L*ld): element t of lane = (line L0 + t*LPT + lane/PC,
// position p0 + lane%PC).
__device__ __forceinline__ void p_load_rows(double (&dst)[PC], const double *__restrict__ X, int L0, int nL,
                                            int n, long ld, int p0, int lane)
{
    const int po = lane % PC, ls = lane / PC;
#pragma unroll
    for (int t = 0; t < PC; ++t) {
        const int LL = L0 + t * LPT + ls;
        const int p = p0 + po;
        if (LL < nL && p < n) dst[t] = X[(long)LL * ld + p];
    }
}

__device__ __forceinline__ void p_store_rows_lds(double (*sh)[LP], const double (&src)[PC], int lane)
{
    const int po = lane % PC, ls = lane / PC;
#pragma unroll
    for (int t = 0; t < PC; ++t) sh[po][t * LPT + ls] = src[t];
}

// Line-major source (element (p, L) at p*nLp + L): element t of lane = (position p0+t,
// line L0+lane).
__device__ __forceinline__ void p_load_lm(double (&dst)[PC], const double *__restrict__ X, int L, int nL, int n,
                                          long nLp, int p0)
{
#pragma unroll
    for (int t = 0; t < PC; ++t) {
        const int p = p0 + t;
        if (L < nL && p < n) dst[t] = X[(long)p * nLp + L];
    }
}

__device__ __forceinline__ void p_store_lm_lds(double (*sh)[LP], const double (&src)[PC], int lane)
{
#pragma unroll
    for (int t = 0; t < PC; ++t) sh[t][lane] = src[t];
}

struct PFwd {
    double inv_k_prev, v_prev, h_prev, r_prev, c_prev, rh, vh, y_last;
};

// Solves nL periodic systems of n >= 3 unknowns; system L has diagonal A[L*ld+p], lower
// B (B[L*ld] is the corner A(0,n-1)), upper C (C[L*ld+n-1] is the corner A(n-1,0)) and
// right-hand side F. Solution to Y[L*ldy+p]. Sr, Sv, Sk are n*nLp line-major scratch.
__global__ void __launch_bounds__(64) k_periodic_rows(const double *__restrict__ A, const double *__restrict__ B,
                                                      const double *__restrict__ C, const double *__restrict__ F,
                                                      long ld, double *__restrict__ Y, long ldy,
                                                      double *__restrict__ Sr, double *__restrict__ Sv,
                                                      double *__restrict__ Sk, int nL, int n, int nLp)
{
    __shared__ double sh[4][PC][LP];
    const int lane = threadIdx.x;
    const int L0 = blockIdx.x * LW;
    const int L = L0 + lane;
    const bool act = L < nL;
    const long lrow = (long)(act ? L : 0) * ld;
    const double c_last = act ? C[lrow + n - 1] : 0.0;   // s = c[n-1], src/Math.cpp:70
    const double b_last = act ? B[lrow + n - 1] : 0.0;   // b[n-1], src/Math.cpp:91
    const int nch = (n + PC - 1) / PC;

    // ---------------- forward: src/Math.cpp:69-106 ----------------
    PFwd st = {0, 0, 0, 0, 0, 0, 0, 0};
    PStage rs;
    p_load_rows(rs.x[0], A, L0, nL, n, ld, 0, lane);
    p_load_rows(rs.x[1], B, L0, nL, n, ld, 0, lane);
    p_load_rows(rs.x[2], C, L0, nL, n, ld, 0, lane);
    p_load_rows(rs.x[3], F, L0, nL, n, ld, 0, lane);
    for (int ch = 0; ch < nch; ++ch) {
        const int p0 = ch * PC;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) p_store_rows_lds(sh[q], rs.x[q], lane);
        __syncthreads();
        if (ch + 1 < nch) {
            p_load_rows(rs.x[0], A, L0, nL, n, ld, p0 + PC, lane);
            p_load_rows(rs.x[1], B, L0, nL, n, ld, p0 + PC, lane);
            p_load_rows(rs.x[2], C, L0, nL, n, ld, p0 + PC, lane);
            p_load_rows(rs.x[3], F, L0, nL, n, ld, p0 + PC, lane);
        }
        if (act) {
#pragma unroll
            for (int q = 0; q < PC; ++q) {
                const int p = p0 + q;
                if (p < n) {
                    const double a = sh[0][q][lane], b = sh[1][q][lane], c = sh[2][q][lane], f = sh[3][q][lane];
                    const long so = (long)p * nLp + L;
                    if (p == 0) {
                        const double k0 = a;
                        const double v0 = b;                 // v[0] = t = b[0]
                        const double ik0 = 1 / k0;
                        const double h0 = c_last * ik0;      // h[0] = s*inv_k[0]
                        const double r0 = f;
                        st.rh = r0 * h0;
                        st.vh = v0 * h0;
                        Sk[so] = ik0; Sv[so] = v0; Sr[so] = r0;
                        st.inv_k_prev = ik0; st.v_prev = v0; st.h_prev = h0; st.r_prev = r0; st.c_prev = c;
                    } else if (p <= n - 3) {
                        const double kk = a - b * st.c_prev * st.inv_k_prev;
                        const double ik = 1 / kk;
                        const double vv = -b * st.v_prev * st.inv_k_prev;
                        const double hh = -st.c_prev * st.h_prev * ik;
                        const double rr = f - b * st.r_prev * st.inv_k_prev;
                        st.rh += hh * rr;
                        st.vh += hh * vv;
                        Sk[so] = ik; Sv[so] = vv; Sr[so] = rr;
                        st.inv_k_prev = ik; st.v_prev = vv; st.h_prev = hh; st.r_prev = rr; st.c_prev = c;
                    } else if (p == n - 2) {
                        const double kk = a - b * st.c_prev * st.inv_k_prev;
                        const double ik = 1 / kk;
                        const double vv = c - b * st.v_prev * st.inv_k_prev;
                        const double hh = (b_last - st.h_prev * st.c_prev) * ik;
                        const double rr = f - b * st.r_prev * st.inv_k_prev;
                        st.rh += hh * rr;
                        st.vh += hh * vv;
                        Sk[so] = ik; Sv[so] = vv; Sr[so] = rr;
                    } else {   // p == n-1: src/Math.cpp:105-108
                        const double k_last = a - st.vh;
                        const double r_last = f - st.rh;
                        st.y_last = r_last / k_last;
                    }
                }
            }
        }
    }

    // ---------------- backward: src/Math.cpp:108-112 ----------------
    // stage: c (row-major) and r, v, inv_k (line-major) of one chunk
    {
        const int pl = (nch - 1) * PC;
        p_load_rows(rs.x[0], C, L0, nL, n, ld, pl, lane);
        p_load_lm(rs.x[1], Sr, L, nL, n, nLp, pl);
        p_load_lm(rs.x[2], Sv, L, nL, n, nLp, pl);
        p_load_lm(rs.x[3], Sk, L, nL, n, nLp, pl);
    }
    double y1 = 0.0;   // y[p+1]
    const double yn = st.y_last;
    const int po = lane % PC, ls = lane / PC;
    for (int ch = nch - 1; ch >= 0; --ch) {
        const int p0 = ch * PC;
        __syncthreads();
        p_store_rows_lds(sh[0], rs.x[0], lane);
#pragma unroll
        for (int q = 1; q < 4; ++q) p_store_lm_lds(sh[q], rs.x[q], lane);
        __syncthreads();
        if (ch > 0) {
            const int pl = p0 - PC;
            p_load_rows(rs.x[0], C, L0, nL, n, ld, pl, lane);
            p_load_lm(rs.x[1], Sr, L, nL, n, nLp, pl);
            p_load_lm(rs.x[2], Sv, L, nL, n, nLp, pl);
            p_load_lm(rs.x[3], Sk, L, nL, n, nLp, pl);
        }
        double yv[PC];
#pragma unroll
        for (int q = PC - 1; q >= 0; --q) {
            const int p = p0 + q;
            double y = 0.0;
            if (p < n && act) {
                if (p == n - 1) {
                    y = yn;
                } else if (p == n - 2) {
                    y = (sh[1][q][lane] - sh[2][q][lane] * yn) * sh[3][q][lane];
                } else {
                    y = (sh[1][q][lane] - sh[0][q][lane] * y1 - sh[2][q][lane] * yn) * sh[3][q][lane];
                }
                y1 = y;
            }
            yv[q] = y;
        }
        __syncthreads();
        // reuse sh[1] (r) for the output tile: all reads of this chunk are done
#pragma unroll
        for (int q = 0; q < PC; ++q) sh[1][q][lane] = yv[q];
        __syncthreads();
#pragma unroll
        for (int t = 0; t < PC; ++t) {
            const int LL = L0 + t * LPT + ls;
            const int p = p0 + po;
            if (LL < nL && p < n) Y[(long)LL * ldy + p] = sh[1][po][t * LPT + ls];
        }
    }
}

hipError_t launch_thomas_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                              double *Y, double *V, int n, int m, long ld)
{
    hipLaunchKernelGGL(k_thomas_cols, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, st, A, B, C, F, Y, V, n, m,
                       ld);
    return hipGetLastError();
}


int periodic_rows_padded_lines(int nL) { return ((nL + LW - 1) / LW) * LW; }

hipError_t launch_periodic_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                long ld, double *Y, long ldy, double *Sr, double *Sv, double *Sk, int nL, int n)
{
    const int nLp = periodic_rows_padded_lines(nL);
    hipLaunchKernelGGL(k_periodic_rows, dim3((unsigned)(nLp / LW)), dim3(LW), 0, st, A, B, C, F, ld, Y, ldy, Sr, Sv,
                       Sk, nL, n, nLp);
    return hipGetLastError();
}

}  // namespace m70
